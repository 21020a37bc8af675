// Host-code checks of the catalog reader and .npy writer (csrc/catalog.cpp)
// for a build with AddressSanitizer and UndefinedBehaviorSanitizer: ordinary,
// empty, comment-only, CRLF, tab-separated, unterminated, malformed, short,
// over-long and binary inputs, bad arguments and small capacities.  Every
// call must return a status (never crash or touch memory it does not own);
// the parsed values of the well-formed cases are checked.  Exit code 0 when
// every expectation holds.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "baorec.h"

static int g_fail = 0;
#define EXPECT(c)                                                              \
    do {                                                                       \
        if (!(c)) {                                                            \
            fprintf(stderr, "expectation failed at line %d: %s\n", __LINE__, #c); \
            g_fail = 1;                                                        \
        }                                                                      \
    } while (0)

static std::string write_file(const std::string &dir, const char *name, const std::string &text) {
    std::string path = dir + "/" + name;
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) {
        perror("fopen");
        exit(2);
    }
    fwrite(text.data(), 1, text.size(), f);
    fclose(f);
    return path;
}

static int read_all(const std::string &path, int nsel, const int32_t *cols, std::vector<std::vector<float>> &out,
                    int64_t cap, int64_t *rows) {
    out.assign(nsel, std::vector<float>((size_t)(cap > 0 ? cap : 1), -12345.f));
    std::vector<float *> ptr;
    for (auto &v : out) ptr.push_back(v.data());
    return baorec_catalog_read(path.c_str(), nsel, cols, ptr.data(), cap, rows);
}

int main(int argc, char **argv) {
    const std::string dir = argc > 1 ? argv[1] : "/tmp";
    int64_t rows = 0;
    int32_t ncols = 0;
    const int32_t c012[3] = {0, 1, 2};
    std::vector<std::vector<float>> out;

    // well formed: comments, blank lines, tabs, runs of blanks, CRLF, no final newline
    std::string p = write_file(dir, "ok.txt", "# header\n1 2 3\n\n4\t5   6\r\n  # indented comment\n7.5 -8e2 9e-3");
    EXPECT(baorec_catalog_rows(p.c_str(), &rows, &ncols) == BAOREC_OK && rows == 3 && ncols == 3);
    EXPECT(read_all(p, 3, c012, out, 3, &rows) == BAOREC_OK && rows == 3);
    EXPECT(out[0][0] == 1.f && out[1][1] == 5.f && out[2][1] == 6.f && out[0][2] == 7.5f && out[1][2] == -800.f &&
           out[2][2] == 9e-3f);
    // capacity smaller than the rows: an error, nothing written past capacity
    EXPECT(read_all(p, 3, c012, out, 2, &rows) != BAOREC_OK);
    // empty and comment-only files
    p = write_file(dir, "empty.txt", "");
    EXPECT(baorec_catalog_rows(p.c_str(), &rows, &ncols) == BAOREC_OK && rows == 0);
    EXPECT(read_all(p, 3, c012, out, 1, &rows) == BAOREC_OK && rows == 0);
    p = write_file(dir, "comments.txt", "# a\n#b\n   \n");
    EXPECT(baorec_catalog_rows(p.c_str(), &rows, &ncols) == BAOREC_OK && rows == 0);
    // a short line and a malformed value are errors
    p = write_file(dir, "short.txt", "1 2 3\n4 5\n");
    EXPECT(baorec_catalog_rows(p.c_str(), &rows, &ncols) == BAOREC_OK && rows == 2);
    EXPECT(read_all(p, 3, c012, out, 2, &rows) != BAOREC_OK);
    p = write_file(dir, "bad.txt", "1 2 3\n4 x5 6\n");
    EXPECT(read_all(p, 3, c012, out, 2, &rows) != BAOREC_OK);
    p = write_file(dir, "bad2.txt", "1 2 3\n4 5e 6\n");
    EXPECT(read_all(p, 3, c012, out, 2, &rows) != BAOREC_OK);
    // out-of-range values round to 0 / inf like parse(Float32, .)
    p = write_file(dir, "range.txt", "1e-400 -1e400 3.4028236e38\n");
    EXPECT(read_all(p, 3, c012, out, 1, &rows) == BAOREC_OK && rows == 1);
    EXPECT(out[0][0] == 0.f && isinf(out[1][0]) && out[1][0] < 0 && isinf(out[2][0]));
    // a very long line and binary garbage
    std::string longline(1 << 20, '1');
    p = write_file(dir, "long.txt", longline + " 2 3\n");
    EXPECT(read_all(p, 3, c012, out, 1, &rows) == BAOREC_OK && rows == 1);
    std::string garbage;
    for (int i = 0; i < 4096; i++) garbage.push_back((char)((i * 131 + 7) & 0xff));
    p = write_file(dir, "garbage.bin", garbage);
    int64_t grows = -1;
    (void)baorec_catalog_rows(p.c_str(), &grows, &ncols);   // any status, no crash
    if (grows > 0 && grows < 100000) (void)read_all(p, 3, c012, out, grows, &rows);
    // bad arguments
    const int32_t cbad[1] = {7};
    EXPECT(read_all(write_file(dir, "one.txt", "1 2 3\n"), 1, cbad, out, 1, &rows) != BAOREC_OK);
    const int32_t cneg[1] = {-1};
    EXPECT(read_all(write_file(dir, "one2.txt", "1 2 3\n"), 1, cneg, out, 1, &rows) != BAOREC_OK);
    EXPECT(baorec_catalog_rows(nullptr, &rows, &ncols) != BAOREC_OK);
    EXPECT(baorec_catalog_rows((dir + "/does_not_exist").c_str(), &rows, &ncols) != BAOREC_OK);
    EXPECT(baorec_catalog_read(p.c_str(), 3, c012, nullptr, 1, &rows) != BAOREC_OK);
    EXPECT(strlen(baorec_io_last_error()) > 0);
    // a large file parsed by several threads (chunk boundaries inside lines)
    std::string big;
    for (int i = 0; i < 400000; i++) big += std::to_string(i) + " " + std::to_string(i % 97) + ".25 -1\n";
    p = write_file(dir, "big.txt", big);
    EXPECT(read_all(p, 3, c012, out, 400000, &rows) == BAOREC_OK && rows == 400000);
    EXPECT(out[0][399999] == 399999.f && out[1][12345] == (12345 % 97) + 0.25f && out[2][0] == -1.f);
    // .npy writer: header and Fortran-order payload
    std::vector<float> a = {1, 2, 3}, b = {4, 5, 6};
    const float *cols[2] = {a.data(), b.data()};
    std::string np = dir + "/out.npy";
    EXPECT(baorec_write_npy(np.c_str(), 2, cols, 3) == BAOREC_OK);
    FILE *f = fopen(np.c_str(), "rb");
    std::vector<char> buf(4096);
    size_t nread = f ? fread(buf.data(), 1, buf.size(), f) : 0;
    if (f) fclose(f);
    EXPECT(nread > 10 && memcmp(buf.data(), "\x93NUMPY", 6) == 0);
    float tail[6];
    if (nread >= 24) memcpy(tail, buf.data() + nread - 24, 24);
    EXPECT(nread >= 24 && tail[0] == 1 && tail[2] == 3 && tail[3] == 4 && tail[5] == 6);
    EXPECT(baorec_write_npy(np.c_str(), 0, cols, 3) != BAOREC_OK);
    EXPECT(baorec_write_npy(np.c_str(), 2, nullptr, 3) != BAOREC_OK);
    EXPECT(baorec_write_npy(nullptr, 2, cols, 3) != BAOREC_OK);
    printf(g_fail ? "catalog checks FAILED\n" : "catalog checks passed\n");
    return g_fail;
}
