// Catalog I/O at the boundary of the reconstruction path (host code):
//  - ASCII catalogs of whitespace-separated numeric columns, one galaxy per
//    line, as the reference's drivers read them with CSV.File(...; delim=" ",
//    types = Float32) (test/benchmarks_cpu.jl:22-23, examples/lightcone.jl:
//    20-21): every value is parsed to the nearest Float32.  The file is
//    memory-mapped and parsed by up to 16 threads over line-aligned chunks.
//  - .npy output of float32 columns, shape (n, ncols) in Fortran order: the
//    byte layout NPZ.jl's npzwrite gives the drivers' hcat(pos', nz)
//    (examples/lightcone.jl:125-127).
#include <errno.h>
#include <math.h>
#include <fcntl.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <charconv>
#include <string>
#include <thread>
#include <vector>

#include "../../include/baorec.h"

namespace {

thread_local std::string g_io_err;

int io_fail(int code, const std::string &msg) {
    g_io_err = msg;
    return code;
}

struct Mapped {
    const char *p = nullptr;
    size_t n = 0;
    int fd = -1;
    ~Mapped() {
        if (p && n) munmap(const_cast<char *>(p), n);
        if (fd >= 0) close(fd);
    }
};

int map_file(const char *path, Mapped &m) {
    if (!path) return io_fail(BAOREC_ERR_INVALID_ARGUMENT, "path is NULL");
    m.fd = open(path, O_RDONLY);
    if (m.fd < 0) return io_fail(BAOREC_ERR_INVALID_ARGUMENT, std::string("cannot open ") + path + ": " + strerror(errno));
    struct stat st;
    if (fstat(m.fd, &st) != 0) return io_fail(BAOREC_ERR_INVALID_ARGUMENT, std::string("cannot stat ") + path);
    m.n = (size_t)st.st_size;
    if (m.n == 0) return BAOREC_OK;
    void *p = mmap(nullptr, m.n, PROT_READ, MAP_PRIVATE, m.fd, 0);
    if (p == MAP_FAILED) return io_fail(BAOREC_ERR_INVALID_ARGUMENT, std::string("cannot map ") + path);
    m.p = static_cast<const char *>(p);
    madvise(p, m.n, MADV_SEQUENTIAL);
    return BAOREC_OK;
}

inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r'; }

// A data line has a non-space character and does not start with '#'.
inline bool data_line(const char *b, const char *e) {
    while (b < e && is_space(*b)) b++;
    return b < e && *b != '#';
}

int nthreads_for(size_t bytes) {
    unsigned hc = std::thread::hardware_concurrency();
    int t = (int)std::min<unsigned>(hc ? hc : 1, 16);
    const size_t per = 4u << 20;   // at least 4 MB per thread
    return (int)std::max<size_t>(1, std::min<size_t>((size_t)t, bytes / per + 1));
}

// Line-aligned chunk starts: chunk k covers [cut[k], cut[k+1]).
std::vector<size_t> chunks(const Mapped &m, int nt) {
    std::vector<size_t> cut(nt + 1, m.n);
    cut[0] = 0;
    for (int k = 1; k < nt; k++) {
        size_t pos = m.n / nt * k;
        while (pos < m.n && m.p[pos - 1] != '\n') pos++;
        cut[k] = std::max(pos, cut[k - 1]);
    }
    return cut;
}

template <class F>
void for_lines(const char *b, const char *e, F &&fn) {
    while (b < e) {
        const char *nl = static_cast<const char *>(memchr(b, '\n', (size_t)(e - b)));
        const char *le = nl ? nl : e;
        fn(b, le);
        b = nl ? nl + 1 : e;
    }
}

// Nearest Float32 of the decimal [s, e), as Julia's parse(Float32, .):
// values beyond the float range round to 0 or +-Inf (from_chars reports them
// as out of range; they are re-read in double and rounded).
bool parse_f32(const char *s, const char *e, float &v) {
    auto r = std::from_chars(s, e, v);
    if (r.ptr != e) return false;
    if (r.ec == std::errc()) return true;
    if (r.ec != std::errc::result_out_of_range) return false;
    double d;
    auto r2 = std::from_chars(s, e, d);
    if (r2.ptr != e) return false;
    if (r2.ec == std::errc::result_out_of_range) {
        // beyond double: the magnitude of the exponent decides
        const char *x = s;
        while (x < e && *x != 'e' && *x != 'E') x++;
        const bool tiny = x + 1 < e && x[1] == '-';
        const bool neg = *s == '-';
        v = tiny ? (neg ? -0.0f : 0.0f) : (neg ? -HUGE_VALF : HUGE_VALF);
        return true;
    }
    if (r2.ec != std::errc()) return false;
    v = (float)d;
    return true;
}

int64_t line_number(const Mapped &m, const char *at) {
    int64_t ln = 1;
    for (const char *q = m.p; q < at; q++) ln += *q == '\n';
    return ln;
}

}  // namespace

extern "C" {

const char *baorec_io_last_error(void) { return g_io_err.c_str(); }

int baorec_catalog_rows(const char *path, int64_t *nrows, int32_t *ncols) {
    if (!nrows) return io_fail(BAOREC_ERR_INVALID_ARGUMENT, "nrows is NULL");
    Mapped m;
    int rc = map_file(path, m);
    if (rc != BAOREC_OK) return rc;
    const int nt = nthreads_for(m.n);
    std::vector<size_t> cut = chunks(m, nt);
    std::vector<int64_t> cnt(nt, 0);
    std::vector<std::thread> th;
    for (int k = 0; k < nt; k++)
        th.emplace_back([&, k] {
            int64_t c = 0;
            for_lines(m.p + cut[k], m.p + cut[k + 1], [&](const char *b, const char *e) { c += data_line(b, e); });
            cnt[k] = c;
        });
    for (auto &t : th) t.join();
    int64_t tot = 0;
    for (int64_t c : cnt) tot += c;
    *nrows = tot;
    if (ncols) {
        *ncols = 0;
        for_lines(m.p, m.p + m.n, [&](const char *b, const char *e) {
            if (*ncols || !data_line(b, e)) return;
            int c = 0;
            while (b < e) {
                while (b < e && is_space(*b)) b++;
                if (b >= e) break;
                c++;
                while (b < e && !is_space(*b)) b++;
            }
            *ncols = c;
        });
    }
    g_io_err.clear();
    return BAOREC_OK;
}

int baorec_catalog_read(const char *path, int32_t nsel, const int32_t *cols, float *const *out,
                        int64_t capacity, int64_t *nrows) {
    if (nsel <= 0 || !cols || !out || !nrows) return io_fail(BAOREC_ERR_INVALID_ARGUMENT, "bad column selection");
    int need = 0;
    for (int k = 0; k < nsel; k++) {
        if (cols[k] < 0 || !out[k]) return io_fail(BAOREC_ERR_INVALID_ARGUMENT, "bad column index or output");
        need = std::max(need, cols[k] + 1);
    }
    Mapped m;
    int rc = map_file(path, m);
    if (rc != BAOREC_OK) return rc;
    const int nt = nthreads_for(m.n);
    std::vector<size_t> cut = chunks(m, nt);
    // pass 1: data lines per chunk -> row offsets
    std::vector<int64_t> cnt(nt, 0), off(nt + 1, 0);
    {
        std::vector<std::thread> th;
        for (int k = 0; k < nt; k++)
            th.emplace_back([&, k] {
                int64_t c = 0;
                for_lines(m.p + cut[k], m.p + cut[k + 1], [&](const char *b, const char *e) { c += data_line(b, e); });
                cnt[k] = c;
            });
        for (auto &t : th) t.join();
    }
    for (int k = 0; k < nt; k++) off[k + 1] = off[k] + cnt[k];
    if (off[nt] > capacity)
        return io_fail(BAOREC_ERR_INVALID_ARGUMENT, "catalog has " + std::to_string(off[nt]) +
                                                        " rows, more than the output capacity");
    // column k of a row -> output slot (or -1)
    std::vector<int> slot(need, -1);
    for (int k = 0; k < nsel; k++) slot[cols[k]] = k;
    // pass 2: parse
    std::vector<const char *> bad(nt, nullptr);
    {
        std::vector<std::thread> th;
        for (int k = 0; k < nt; k++)
            th.emplace_back([&, k] {
                int64_t row = off[k];
                for_lines(m.p + cut[k], m.p + cut[k + 1], [&](const char *b, const char *e) {
                    if (bad[k] || !data_line(b, e)) return;
                    const char *line = b;
                    int col = 0;
                    while (col < need) {
                        while (b < e && is_space(*b)) b++;
                        if (b >= e) break;
                        const char *tok = b;
                        while (b < e && !is_space(*b)) b++;
                        if (slot[col] >= 0) {
                            const char *s = tok;
                            if (s < b && *s == '+') s++;
                            float v;
                            if (!parse_f32(s, b, v)) {
                                bad[k] = line;
                                return;
                            }
                            out[slot[col]][row] = v;
                        }
                        col++;
                    }
                    if (col < need) {
                        bad[k] = line;
                        return;
                    }
                    row++;
                });
            });
        for (auto &t : th) t.join();
    }
    for (int k = 0; k < nt; k++)
        if (bad[k])
            return io_fail(BAOREC_ERR_INVALID_ARGUMENT, std::string(path) + ": line " +
                                                            std::to_string(line_number(m, bad[k])) +
                                                            ": missing or malformed value (need " +
                                                            std::to_string(need) + " numeric columns)");
    *nrows = off[nt];
    g_io_err.clear();
    return BAOREC_OK;
}

int baorec_write_npy(const char *path, int32_t ncols, const float *const *cols, int64_t n) {
    if (!path || ncols <= 0 || !cols || n < 0) return io_fail(BAOREC_ERR_INVALID_ARGUMENT, "bad arguments");
    for (int k = 0; k < ncols; k++)
        if (!cols[k] && n > 0) return io_fail(BAOREC_ERR_INVALID_ARGUMENT, "column is NULL");
    std::string dict = "{'descr': '<f4', 'fortran_order': True, 'shape': (" + std::to_string(n) + ", " +
                       std::to_string(ncols) + "), }";
    // magic (6) + version (2) + length (2) + dict + padding + '\n' -> multiple of 64
    size_t hl = dict.size() + 1;
    const size_t pad = (64 - (10 + hl) % 64) % 64;
    dict.append(pad, ' ');
    dict.push_back('\n');
    if (dict.size() > 65535) return io_fail(BAOREC_ERR_INVALID_ARGUMENT, "npy header too long");
    FILE *f = fopen(path, "wb");
    if (!f) return io_fail(BAOREC_ERR_INVALID_ARGUMENT, std::string("cannot create ") + path + ": " + strerror(errno));
    const unsigned char magic[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
    const uint16_t len = (uint16_t)dict.size();
    const unsigned char lb[2] = {(unsigned char)(len & 0xff), (unsigned char)(len >> 8)};
    bool ok = fwrite(magic, 1, 8, f) == 8 && fwrite(lb, 1, 2, f) == 2 &&
              fwrite(dict.data(), 1, dict.size(), f) == dict.size();
    for (int k = 0; ok && k < ncols; k++) ok = fwrite(cols[k], sizeof(float), (size_t)n, f) == (size_t)n;
    ok = (fclose(f) == 0) && ok;
    if (!ok) return io_fail(BAOREC_ERR_INVALID_ARGUMENT, std::string("write failed: ") + path);
    g_io_err.clear();
    return BAOREC_OK;
}

}  // extern "C"
