// C ABI of libbaorec_hip: context, device buffers and the reconstruction
// pipelines built from the kernels in fft.hip, mas.hip and mg.hip.
#include <math.h>
#include <cmath>
#include <stdio.h>
#include <string.h>

#include <array>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/baorec.h"
#include "kernels.hpp"

using namespace baorec;


struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

struct baorec_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::map<std::string, DevBuf> bufs;
    // recon.result_cache
    int has_result = 0;
    int result_kind = 0;
    int32_t result_grid[3] = {0, 0, 0};
    uint64_t result_gen = 0;
    // displacement grid cached for the result (Psi_x, Psi_y, Psi_z per cell)
    uint64_t disp_gen = 0;
    // block tables of the data + randoms sort on the device, for these sizes
    int64_t pair_n[2] = {-1, -1};
    int pair_nblk = 0;
    const void *pair_tab = nullptr;
    // generation of the result whose R2C spectrum is still in cwork[0]
    uint64_t spec_gen = 0;
    float disp_box[6] = {0, 0, 0, 0, 0, 0};
    int disp_valid = 0;
    // optional per-launch timing (HIP events on `stream`)
    int prof_on = 0;
    struct Pending {
        const char *name;
        double bytes;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
    struct Agg {
        int64_t count = 0;
        double ms = 0.0, bytes = 0.0;
    };
    std::map<std::string, Agg> prof;
    // autotuned choices (baorec_autotune, baorec_tune_load): scatter tile per
    // (nx, ny, nz, tsc, floor(log2 n)); strided kernel per (axis, line length)
    std::map<std::array<int64_t, 5>, std::array<int, 3>> tune_tile;
    std::map<std::array<int, 2>, int> tune_pair;
    int force_tile[3] = {0, 0, 0};   // set while tuning
    int force_pair = -1;
    // learned warm start: the network of baorec_warmstart_init / _load
    baorec::UNet *net = nullptr;
    // HIP graphs of baorec_run_positions_graph, keyed by its arguments; valid
    // while no library buffer has been (re)allocated since the capture
    struct GraphEntry {
        std::string key;
        uint64_t buf_gen;
        hipGraphExec_t exec;
    };
    std::vector<GraphEntry> graphs;
    uint64_t buf_gen = 0;   // counts buffer (re)allocations
    int capturing = 0;      // inside a stream capture: no host synchronisation
    // comoving-distance table of the last cosmology used
    int cosmo_valid = 0;
    baorec_cosmology cosmo{};
    double cosmo_r0 = 0.0, cosmo_dr_inv = 0.0;
};

static thread_local std::string g_create_err;

namespace {

struct Fail {
    int code;
};

[[noreturn]] void fail(baorec_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    throw Fail{code};
}

void hipchk(baorec_ctx *c, hipError_t e, const char *what) {
    if (e == hipSuccess) return;
    const int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
                         ? BAOREC_ERR_OUT_OF_MEMORY
                         : BAOREC_ERR_HIP;
    fail(c, code, std::string(what) + ": " + hipGetErrorString(e));
}

void *buf(baorec_ctx *c, const std::string &name, size_t bytes) {
    DevBuf &b = c->bufs[name];
    if (b.bytes < bytes) {
        if (c->capturing) fail(c, BAOREC_ERR_HIP, "buffer allocation inside a graph capture");
        c->buf_gen++;
        if (b.p) {
            hipchk(c, hipStreamSynchronize(c->stream), "sync before realloc");
            (void)hipFree(b.p);
            b.p = nullptr;
            b.bytes = 0;
        }
        void *p = nullptr;
        hipchk(c, hipMalloc(&p, bytes < 256 ? 256 : bytes), ("hipMalloc " + name).c_str());
        b.p = p;
        b.bytes = bytes;
    }
    return b.p;
}

template <typename T>
T *bufT(baorec_ctx *c, const std::string &name, size_t count) {
    return static_cast<T *>(buf(c, name, sizeof(T) * count));
}

// P: valid kx columns of the half spectrum (nx/2 + 1); PP: row pitch of the
// work spectra, rounded up to 16 complex (128 B) so that every column
// segment a strided pass touches starts on a cache line.
struct Grid {
    int nx, ny, nz, P;
    int PP() const { return (P + 15) / 16 * 16; }
    // kz planes are one row longer than ny rows, so that the z passes'
    // column loads are not all a multiple of 64 KB apart
    int64_t PZ() const { return (int64_t)PP() * (ny + 1); }
    int64_t R() const { return (int64_t)nx * ny * nz; }
    int64_t C() const { return PZ() * nz; }
    int64_t Cdense() const { return (int64_t)P * ny * nz; }
};

Grid check_grid(baorec_ctx *c, const int32_t *g, bool fft) {
    if (!g) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "grid is NULL");
    Grid G{g[0], g[1], g[2], g[0] / 2 + 1};
    for (int a = 0; a < 3; a++) {
        if (g[a] <= 0) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "grid sides must be positive");
        if (fft && !fft_side_ok(g[a]))
            fail(c, BAOREC_ERR_UNSUPPORTED,
                 "grid sides must be powers of two in [16, 1024] (got " + std::to_string(g[a]) + ")");
    }
    return G;
}

void check_params(baorec_ctx *c, const baorec_params *p) {
    if (!p) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "params is NULL");
    if (p->mas != BAOREC_MAS_CIC && p->mas != BAOREC_MAS_TSC)
        fail(c, BAOREC_ERR_INVALID_ARGUMENT, "unknown mass-assignment scheme");
}

void check_box(baorec_ctx *c, const baorec_params *p) {
    for (int a = 0; a < 3; a++)
        if (!(p->box_size[a] > 0.f) || !isfinite(p->box_min[a]))
            fail(c, BAOREC_ERR_INVALID_ARGUMENT, "box_size/box_min not set");
}

// Input array: device pointer used as is, host pointer staged into a named
// device buffer.
const float *stage_in(baorec_ctx *c, int mem, const float *src, int64_t n, const char *name) {
    if (n == 0) return nullptr;
    if (!src) fail(c, BAOREC_ERR_INVALID_ARGUMENT, std::string(name) + " is NULL");
    if (mem == BAOREC_MEM_DEVICE) return src;
    float *d = bufT<float>(c, name, n);
    hipchk(c, hipMemcpyAsync(d, src, sizeof(float) * n, hipMemcpyHostToDevice, c->stream), "H2D");
    return d;
}

float *stage_out(baorec_ctx *c, int mem, float *dst, int64_t n, const char *name) {
    if (n == 0) return nullptr;
    if (!dst) fail(c, BAOREC_ERR_INVALID_ARGUMENT, std::string(name) + " is NULL");
    if (mem == BAOREC_MEM_DEVICE) return dst;
    return bufT<float>(c, name, n);
}

void copy_out(baorec_ctx *c, int mem, float *host, const float *dev, int64_t n) {
    if (mem == BAOREC_MEM_DEVICE || n == 0) return;
    hipchk(c, hipMemcpyAsync(host, dev, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream), "D2H");
}

hipEvent_t take_event(baorec_ctx *c) {
    if (!c->event_pool.empty()) {
        hipEvent_t e = c->event_pool.back();
        c->event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    hipchk(c, hipEventCreate(&e), "hipEventCreate");
    return e;
}

void prof_resolve(baorec_ctx *c) {
    for (auto &p : c->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            auto &g = c->prof[p.name];
            g.count++;
            g.ms += ms;
            g.bytes += p.bytes;
        }
        c->event_pool.push_back(p.a);
        c->event_pool.push_back(p.b);
    }
    c->pending.clear();
}

void sync(baorec_ctx *c) {
    if (c->capturing) return;   // the graph's launcher synchronises after the replay
    hipchk(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    if (!c->pending.empty()) prof_resolve(c);
}

// Brackets one launch with events when profiling is on; `bytes` is the
// launch's algorithmic HBM traffic (compulsory reads + writes).
struct ProfScope {
    baorec_ctx *c;
    hipEvent_t a = nullptr, b = nullptr;
    const char *name;
    double bytes;
    ProfScope(baorec_ctx *c_, const char *n, double by) : c(c_), name(n), bytes(by) {
        if (c->prof_on) {
            a = take_event(c);
            b = take_event(c);
            hipchk(c, hipEventRecord(a, c->stream), "hipEventRecord");
        }
    }
    ~ProfScope() {
        if (a) {
            (void)hipEventRecord(b, c->stream);
            c->pending.push_back({name, bytes, a, b});
        }
    }
};

// ----------------------------------------------------------- device scalars

__global__ void k_slot_set(Slot *s, float b0, float b1, float b2, float m0, float m1, float m2,
                           double count_r) {
    if (threadIdx.x != 0) return;
    s->box_size[0] = b0;
    s->box_size[1] = b1;
    s->box_size[2] = b2;
    s->box_min[0] = m0;
    s->box_min[1] = m1;
    s->box_min[2] = m2;
    s->sum[0] = 0.0;
    s->sum[1] = 0.0;
    s->count_r = count_r;
    s->err = 0;
}

__global__ void k_slot_reset(Slot *s, double count_r) {
    if (threadIdx.x != 0) return;
    s->sum[0] = 0.0;
    s->sum[1] = 0.0;
    s->count_r = count_r;
    s->err = 0;
}

// n slots with the same box (the catalogs of a batch)
__global__ void k_slots_set(Slot *s, int n, float b0, float b1, float b2, float m0, float m1, float m2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Slot &t = s[i];
    t.box_size[0] = b0;
    t.box_size[1] = b1;
    t.box_size[2] = b2;
    t.box_min[0] = m0;
    t.box_min[1] = m1;
    t.box_min[2] = m2;
    t.sum[0] = 0.0;
    t.sum[1] = 0.0;
    t.count_r = 0.0;
    t.err = 0;
}

Slot *slot_buf(baorec_ctx *c) { return bufT<Slot>(c, "slot", 4); }

void slot_from_params(baorec_ctx *c, Slot *s, const baorec_params *p, double count_r) {
    hipLaunchKernelGGL(k_slot_set, dim3(1), dim3(64), 0, c->stream, s, p->box_size[0],
                       p->box_size[1], p->box_size[2], p->box_min[0], p->box_min[1], p->box_min[2],
                       count_r);
    hipchk(c, hipGetLastError(), "slot_set");
}

Slot read_slot(baorec_ctx *c, const Slot *s) {
    if (c->capturing) fail(c, BAOREC_ERR_UNSUPPORTED, "a host read of device scalars inside a graph capture");
    Slot h;
    hipchk(c, hipMemcpyAsync(&h, s, sizeof(Slot), hipMemcpyDeviceToHost, c->stream), "D2H slot");
    sync(c);
    return h;
}

// ----------------------------------------------------------------- scatter

int ilog2(int64_t n) {
    int k = 0;
    while (n > 1) {
        n >>= 1;
        k++;
    }
    return k;
}

std::array<int64_t, 5> tile_key(const Grid &G, int tsc, int64_t n) {
    return {G.nx, G.ny, G.nz, tsc, ilog2(n)};
}

// Scatter tile shape: the one being timed while tuning, a tuned one for this
// grid and catalog size, or 32 x 16 x 16.
void tiles_for(baorec_ctx *c, const Grid &G, ScatterArgs &a, int64_t n) {
    a.tx = G.nx < 32 ? G.nx : 32;
    a.ty = G.ny < 16 ? G.ny : 16;
    a.tz = G.nz < 16 ? G.nz : 16;
    if (c->force_tile[0]) {
        a.tx = c->force_tile[0];
        a.ty = c->force_tile[1];
        a.tz = c->force_tile[2];
    } else {
        auto it = c->tune_tile.find(tile_key(G, a.tsc, n));
        if (it != c->tune_tile.end()) {
            a.tx = it->second[0];
            a.ty = it->second[1];
            a.tz = it->second[2];
        }
    }
    a.ntx = (G.nx + a.tx - 1) / a.tx;
    a.nty = (G.ny + a.ty - 1) / a.ty;
    a.ntz = (G.nz + a.tz - 1) / a.tz;
}

// cic! into rho (overwrite unless accumulate) for one catalog.
void scatter(baorec_ctx *c, const Grid &G, const baorec_params *p, Slot *slot, float *rho, float *x,
             float *y, float *z, const float *w, int64_t n, int wrap, int accumulate,
             const char *tag, double *sums = nullptr) {
    // placements are indexed with 32 bits; the bound is the worst case of a
    // particle touching 8 (CIC) or 27 (TSC) tiles
    const int64_t per = p->mas == BAOREC_MAS_TSC ? 27 : 8;
    if (n * per >= ((int64_t)1 << 32))
        fail(c, BAOREC_ERR_UNSUPPORTED,
             "catalog too large for one scatter (at most " + std::to_string((((int64_t)1 << 32) - 1) / per) +
                 " particles)");
    ScatterArgs a{};
    a.x = x;
    a.y = y;
    a.z = z;
    a.w = w;
    a.offsets = nullptr;
    a.ncat = 1;
    a.ntotal = n;
    a.slot = slot;
    a.slotStride = 0;
    a.nx = G.nx;
    a.ny = G.ny;
    a.nz = G.nz;
    a.wrap = wrap;
    a.mutate = p->mutate_inputs;
    a.tsc = p->mas == BAOREC_MAS_TSC;
    tiles_for(c, G, a, n);
    const int64_t ntiles = (int64_t)a.ntx * a.nty * a.ntz;
    a.count = bufT<uint32_t>(c, "sc_count", ntiles);
    a.cursor = bufT<uint32_t>(c, "sc_cursor", ntiles);
    a.sorted = bufT<float4>(c, "sc_sorted", scatter_capacity(n > 0 ? n : 1, a.tsc));
    a.err = &slot->err;
    a.rho = rho;
    a.rstride = G.R();
    a.wmax = bufT<uint32_t>(c, "sc_wmax", 64);
    hipchk(c, launch_zero_words(a.wmax, 64, c->stream), "zero");
    (void)tag;
    if (sort_fits(a)) {
        const int nblk = sort_nblk(n);
        const int T = a.ntx * a.nty * a.ntz;
        uint32_t *H = bufT<uint32_t>(c, "sc_hist", (size_t)(nblk > 0 ? nblk : 1) * T);
        ProfScope ps(c, "scatter_bin", (double)n * (a.mutate && wrap ? 40.0 : 28.0) + 16.0 * n * 1.2);
        hipchk(c, launch_sort(a, nblk, H, nullptr, nullptr, nullptr, c->stream), "scatter sort");
    } else {
        ProfScope ps(c, "scatter_bin", (double)n * (a.mutate && wrap ? 40.0 : 28.0) + 16.0 * n * 1.2);
        hipchk(c, launch_scatter(a, nullptr, c->stream), "scatter");
    }
    // FP64 total of the grid (overwrite mode), from the tiles as they are written
    if (sums && !accumulate) a.part = bufT<double>(c, "sc_part", (size_t)ntiles);
    {
        // overwrite mode writes every cell once (zeros in empty tiles)
        ProfScope ps(c, "scatter_accum", 16.0 * n * 1.2 + 4.0 * G.R() * (accumulate ? 2 : 1));
        hipchk(c, launch_scatter_accum(a, accumulate ? 1 : 0, c->stream), "scatter accumulate");
    }
    if (a.part) hipchk(c, launch_sum_partials(a.part, ntiles, 1, sums, 1, c->stream), "grid total");
}

// cic! of the data and of the randoms into rho and rho + R (overwrite, no
// wrap) as one two-catalog sort and accumulation: the same tiles, sums and
// placements as two scatter() calls, in half the launches.  Falls back to two
// calls when the tiles do not fit the sort.
void scatter_pair(baorec_ctx *c, const Grid &G, const baorec_params *p, Slot *slot, float *rho, float *x,
                  float *y, float *z, const float *w, int64_t n, float *rx, float *ry, float *rz, const float *rw,
                  int64_t nr) {
    const int64_t per = p->mas == BAOREC_MAS_TSC ? 27 : 8;
    ScatterArgs a{};
    a.tsc = p->mas == BAOREC_MAS_TSC;
    tiles_for(c, G, a, nr);
    const int64_t ntot = n + nr;
    if (!sort_fits(a) || ntot * per >= ((int64_t)1 << 32) || n == 0 || nr == 0) {
        scatter(c, G, p, slot, rho, x, y, z, w, n, 0, 0, "data", &slot->sum[0]);
        scatter(c, G, p, slot, rho + G.R(), rx, ry, rz, rw, nr, 0, 0, "randoms", &slot->sum[1]);
        return;
    }
    a.x = x;
    a.y = y;
    a.z = z;
    a.w = w;
    a.x2 = rx;
    a.y2 = ry;
    a.z2 = rz;
    a.w2 = rw;
    a.n1 = n;
    a.ncat = 2;
    a.ntotal = ntot;
    a.slot = slot;
    a.slotStride = 0;
    a.nx = G.nx;
    a.ny = G.ny;
    a.nz = G.nz;
    a.wrap = 0;
    a.mutate = 0;
    const int64_t T = (int64_t)a.ntx * a.nty * a.ntz;
    // block tables (int64: offsets 3 | block starts nblk + 1; int32: block
    // catalogs nblk | first block per catalog 3), uploaded when the sizes change
    int64_t *d64 = bufT<int64_t>(c, "pair_t64", 3 + kSortTableBlocks + 1);
    int *d32 = bufT<int>(c, "pair_t32", kSortTableBlocks + 3);
    if (c->pair_n[0] != n || c->pair_n[1] != nr || c->pair_tab != d64) {
        const int64_t offs[3] = {0, n, ntot};
        std::vector<int> blk_cat, cat_blk0;
        std::vector<int64_t> blk_beg;
        const int nblk = sort_blocks(ntot, offs, 2, blk_cat, blk_beg, cat_blk0);
        if (nblk > kSortTableBlocks) fail(c, BAOREC_ERR_HIP, "sort block table overflow");
        std::vector<int64_t> t64(offs, offs + 3);
        t64.insert(t64.end(), blk_beg.begin(), blk_beg.end());
        std::vector<int> t32(blk_cat);
        t32.insert(t32.end(), cat_blk0.begin(), cat_blk0.end());
        hipchk(c, hipMemcpy(d64, t64.data(), sizeof(int64_t) * t64.size(), hipMemcpyHostToDevice), "copy");
        hipchk(c, hipMemcpy(d32, t32.data(), sizeof(int) * t32.size(), hipMemcpyHostToDevice), "copy");
        c->pair_n[0] = n;
        c->pair_n[1] = nr;
        c->pair_nblk = nblk;
        c->pair_tab = d64;
    }
    const int nblk = c->pair_nblk;
    a.offsets = d64;
    a.count = bufT<uint32_t>(c, "sc_count", (size_t)2 * T);
    a.cursor = bufT<uint32_t>(c, "sc_cursor", (size_t)2 * T);
    a.sorted = bufT<float4>(c, "sc_sorted", scatter_capacity(ntot, a.tsc));
    a.err = &slot->err;
    a.rho = rho;
    a.rstride = G.R();
    a.wmax = bufT<uint32_t>(c, "sc_wmax", 64);
    hipchk(c, launch_zero_words(a.wmax, 64, c->stream), "zero");
    uint32_t *H = bufT<uint32_t>(c, "sc_hist", (size_t)nblk * T);
    {
        ProfScope ps(c, "scatter_bin", 28.0 * ntot + 16.0 * ntot * 1.2);
        hipchk(c, launch_sort(a, nblk, H, d32, d64 + 3, d32 + nblk, c->stream), "scatter sort");
    }
    a.part = bufT<double>(c, "sc_part", (size_t)2 * T);
    {
        ProfScope ps(c, "scatter_accum", 16.0 * ntot * 1.2 + 4.0 * G.R() * 2);
        hipchk(c, launch_scatter_accum(a, 0, c->stream), "scatter accumulate");
    }
    // FP64 totals of the data and randoms grids into slot->sum[0], [1]
    hipchk(c, launch_sum_partials(a.part, T, 2, slot->sum, 1, c->stream), "grid totals");
}

// ------------------------------------------------------------------ FFT

struct FFTBufs {
    float2 *cw;   // 11 complex grids
    float *ktab, *xtab;
};

FFTBufs fft_bufs(baorec_ctx *c, const Grid &G) {
    c->spec_gen = 0;  // any new use of the work spectra invalidates the kept one
    FFTBufs f;
    f.cw = bufT<float2>(c, "cwork", (size_t)11 * G.C());
    f.ktab = bufT<float>(c, "ktab", (size_t)4 * (G.P + G.ny + G.nz));
    f.xtab = bufT<float>(c, "xtab", (size_t)4 * (G.nx + G.ny + G.nz));
    return f;
}

void tables(baorec_ctx *c, const Grid &G, const Slot *slot, FFTBufs &f) {
    hipchk(c, launch_tables(slot, f.ktab, f.xtab, G.nx, G.ny, G.nz, 1, c->stream), "tables");
}

void strided(baorec_ctx *c, const Grid &G, const StridedArgs &a0, int njobs, int nbatch, const char *what) {
    StridedArgs a = a0;
    // kernel choice of single-output passes: being timed, tuned, or the rule
    a.pair = c->force_pair;
    if (a.pair < 0) {
        auto it = c->tune_pair.find({a.axis, a.axis == 1 ? G.ny : G.nz});
        if (it != c->tune_pair.end()) a.pair = it->second;
    }
    double bytes = 0.0;
    for (int j = 0; j < njobs; j++) bytes += 8.0 * G.Cdense() * (1 + a.job[j].nout);
    const bool pair = strided_uses_pair(a, njobs);
    ProfScope ps(c, a.axis == 1 ? (pair ? "fft_y_pair" : "fft_y") : (pair ? "fft_z_pair" : "fft_z"), bytes * nbatch);
    hipchk(c, launch_strided(a, njobs, nbatch, c->stream), what);
}

void x_inv(baorec_ctx *c, const Grid &G, const XInvArgs &a, int epi, int nbatch, const char *what) {
    double out = a.out ? (epi == XE_DISP ? 12.0 : 4.0) : 0.0;
    double aux = (epi == XE_ITER_LOCAL || epi == XE_ITER_FIXED) ? 4.0 : 0.0;
    double fwd = a.fwd_out ? 8.0 * G.Cdense() : 0.0;
    static const char *const label[] = {"fft_x_c2r", "fft_x_c2r_od", "fft_x_c2r_od", "fft_x_c2r_iter",
                                        "fft_x_c2r_iter", "fft_x_c2r_disp"};
    const double in = a.rin[0] ? 4.0 * G.R() * a.nin : 8.0 * G.Cdense() * a.nin;
    ProfScope ps(c, label[epi], (double)nbatch * (in + (out + aux) * G.R() + fwd));
    hipchk(c, launch_x_c2r(a, epi, nbatch, c->stream), what);
}

StridedArgs sargs(const Grid &G, const FFTBufs &f, int axis, int mode, int mul) {
    StridedArgs a{};
    a.cstride = G.C();
    a.ktab = f.ktab;
    a.ktabStride = 0;
    a.P = G.PP();
        a.PZ = G.PZ();
    a.Pv = G.P;
    a.kyo = G.P;
    a.ny = G.ny;
    a.nz = G.nz;
    a.axis = axis;
    a.mode = mode;
    a.mul = mul;
    return a;
}

void x_fwd(baorec_ctx *c, const Grid &G, const float *in, float2 *out, int nbatch) {
    XFwdArgs a{};
    a.in = in;
    a.out = out;
    a.rstride = G.R();
    a.cstride = G.C();
    a.P = G.PP();
    a.PZ = G.PZ();
    a.nx = G.nx;
    a.ny = G.ny;
    a.nz = G.nz;
    ProfScope ps(c, "fft_x_r2c", (double)nbatch * (4.0 * G.R() + 8.0 * G.Cdense()));
    hipchk(c, launch_x_r2c(a, nbatch, c->stream), "x r2c");
}

void strided_plain(baorec_ctx *c, const Grid &G, const FFTBufs &f, int axis, bool inverse,
                   float2 *data, int nbatch) {
    StridedArgs a = sargs(G, f, axis, inverse ? 1 : 0, MUL_NONE);
    a.job[0].in = data;
    a.job[0].out[0] = data;
    a.job[0].nout = 1;
    a.job[0].pw[0] = 0;
    strided(c, G, a, 1, nbatch, inverse ? "strided inverse" : "strided fwd");
}

XInvArgs xargs(const Grid &G, const FFTBufs &f, const Slot *slot) {
    XInvArgs a{};
    a.cstride = G.C();
    a.rstride = G.R();
    a.P = G.PP();
    a.PZ = G.PZ();
    a.nx = G.nx;
    a.ny = G.ny;
    a.nz = G.nz;
    a.scale = (float)(1.0 / ((double)G.nx * G.ny * G.nz));
    a.ktab = f.ktab;
    a.ktabStride = 0;
    a.xtab = f.xtab;
    a.xtabStride = 0;
    a.slot = slot;
    return a;
}

// smooth! (src/utils.jl:45-56) of nb contiguous real grids, y and z axes,
// in place.  The Gaussian exp(-R^2 k^2 / 2) is the product of one factor per
// axis, so the smoothing runs one axis at a time on the real grid; here each
// grid is read as complex lines of two adjacent x cells, whose transform
// along y or z is that of the two real columns (the factor is real and even
// in k).  The x factor is applied by the x pass that reads the grids
// (XInvArgs::rin).
void smooth_yz(baorec_ctx *c, const Grid &G, FFTBufs &f, float *rin, int nb, float R) {
    for (int axis = 1; axis <= 2; axis++) {
        StridedArgs a = sargs(G, f, axis, 2, MUL_GAUSS1);
        a.P = G.nx / 2;
        a.Pv = G.nx / 2;
        a.PZ = (int64_t)(G.nx / 2) * G.ny;
        a.cstride = G.R() / 2;
        a.R2 = R * R;
        a.job[0].in = reinterpret_cast<float2 *>(rin);
        a.job[0].out[0] = reinterpret_cast<float2 *>(rin);
        a.job[0].pw[0] = 0;
        a.job[0].nout = 1;
        strided(c, G, a, 1, nb, axis == 1 ? "y smooth" : "z smooth");
    }
}

// setup_overdensity! (src/recon.jl:42-91) -> delta (device).
void overdensity(baorec_ctx *c, const Grid &G, const baorec_params *p, Slot *slot, float *delta,
                 float *x, float *y, float *z, const float *w, int64_t n, float *rx, float *ry,
                 float *rz, const float *rw, int64_t nr, bool fwd_spectrum = false) {
    FFTBufs f = fft_bufs(c, G);
    tables(c, G, slot, f);
    float *rho = bufT<float>(c, "rho2", (size_t)2 * G.R());
    const bool rand = nr > 0;
    if (rand) scatter_pair(c, G, p, slot, rho, x, y, z, w, n, rx, ry, rz, rw, nr);
    else scatter(c, G, p, slot, rho, x, y, z, w, n, 1, 0, "data", &slot->sum[0]);
    smooth_yz(c, G, f, rho, rand ? 2 : 1, p->smoothing_radius);
    // the x smoothing, the overdensity (FP64 totals from the scatter) and,
    // for the solvers, R2C(delta) along x
    XInvArgs a = xargs(G, f, slot);
    a.rin[0] = rho;
    a.kxpow[0] = 0;
    a.nin = 1;
    a.R2 = p->smoothing_radius * p->smoothing_radius;
    a.bias = p->bias;
    a.ran_min = p->ran_min;
    a.out = delta;
    if (fwd_spectrum) a.fwd_out = f.cw;
    if (rand) {
        a.rin[1] = rho + G.R();
        a.kxpow[1] = 0;
        a.nin = 2;
    }
    x_inv(c, G, a, rand ? XE_OD_RAND : XE_OD_BOX, 1, "x overdensity");
}

// Work spectra of the iterative solve and the displacement read-out for
// nb catalogs: quantity m of catalog b at base_m + b * C, quantities q
// complex elements apart.
struct Work {
    float2 *C0;   // R2C of the field being iterated / read out
    float2 *Gb;   // three quantities (z-pass outputs)
    float2 *Hb;   // three quantities (y-pass outputs)
    int64_t q;
};

// Layout of the single-catalog path inside the shared work spectra.
Work work_single(const FFTBufs &f, const Grid &G) {
    const int64_t C = G.C();
    return Work{f.cw, f.cw + 2 * C, f.cw + 5 * C, C};
}

// n_iter x iterate! (src/iterative.jl:151-211): delta_s -> delta_r for nb
// catalogs (grids R apart, spectra C apart, per-catalog slots).
// c0_ready: W.C0 already holds R2C(delta_s); keep_spectrum: leave
// R2C(delta_r) in W.C0 at the end (for the displacement read-out); dr may be
// NULL when only that spectrum is wanted.
void iterate_all(baorec_ctx *c, const Grid &G, const baorec_params *p, Slot *slot, const float *ds,
                 float *dr, bool c0_ready = false, bool keep_spectrum = false, int nb = 1,
                 const Work *Wp = nullptr) {
    FFTBufs f = fft_bufs(c, G);
    const Work W = Wp ? *Wp : work_single(f, G);
    float2 *C0 = W.C0;
    if (p->n_iter <= 0) {
        if (dr)
            hipchk(c, hipMemcpyAsync(dr, ds, sizeof(float) * G.R() * nb, hipMemcpyDeviceToDevice, c->stream),
                   "copy");
        return;
    }
    const float beta = p->beta;
    for (int it = 1; it <= p->n_iter; it++) {
        const bool last = it == p->n_iter;
        // later iterations get R2C(delta_r) from the previous x pass (fwd_out)
        if (it == 1 && !c0_ready) x_fwd(c, G, ds, C0, nb);
        strided_plain(c, G, f, 1, false, C0, nb);
        if (!p->has_los) {
            StridedArgs z = sargs(G, f, 2, 2, MUL_INVK2);
            z.job[0].in = C0;
            for (int m = 0; m < 3; m++) {
                z.job[0].out[m] = W.Gb + m * W.q;
                z.job[0].pw[m] = m;
            }
            z.job[0].nout = 3;
            strided(c, G, z, 1, nb, "z iterate");
            // y inverse: G0,G1,G2 -> A, B, C (the six Hessian terms grouped by
            // their power of x along an x line)
            YIterArgs y{};
            for (int m = 0; m < 3; m++) {
                y.g[m] = W.Gb + m * W.q;
                y.out[m] = W.Hb + m * W.q;
            }
            y.cstride = G.C();
            y.ktab = f.ktab;
            y.xtab = f.xtab;
            y.P = G.PP();
            y.PZ = G.PZ();
            y.Pv = G.P;
            y.nx = G.nx;
            y.ny = G.ny;
            y.nz = G.nz;
            const int pi[6] = {0, 0, 0, 1, 1, 2}, pj[6] = {0, 1, 2, 1, 2, 2};
            for (int q = 0; q < 6; q++) {
                // factor::T = (1. + T(i != j)) * beta  (src/iterative.jl:23-24)
                float fac = (float)((1.0 + (double)(float)(pi[q] != pj[q])) * (double)beta);
                if (it == 1) fac = fac / (1.0f + beta);
                y.fac[q] = fac;
            }
            {
                ProfScope ps(c, "fft_y_iter", 8.0 * G.Cdense() * 6 * nb);
                hipchk(c, launch_y_iter(y, nb, c->stream), "y iterate");
            }
            XInvArgs a = xargs(G, f, slot);
            const int kxp[3] = {2, 1, 0};
            for (int q = 0; q < 3; q++) {
                a.in[q] = W.Hb + q * W.q;
                a.kxpow[q] = kxp[q];
            }
            a.nin = 3;
            a.aux = ds;
            a.out = last ? dr : nullptr;
            if (!last || keep_spectrum) a.fwd_out = C0;
            x_inv(c, G, a, XE_ITER_LOCAL, nb, "x iterate");
        } else {
            StridedArgs z = sargs(G, f, 2, 2, MUL_FIXEDLOS);
            z.los[0] = p->los[0];
            z.los[1] = p->los[1];
            z.los[2] = p->los[2];
            z.job[0].in = C0;
            z.job[0].out[0] = C0;
            z.job[0].pw[0] = 0;
            z.job[0].nout = 1;
            strided(c, G, z, 1, nb, "z iterate fixed");
            strided_plain(c, G, f, 1, true, C0, nb);
            XInvArgs a = xargs(G, f, slot);
            a.in[0] = C0;
            a.kxpow[0] = 0;
            a.nin = 1;
            float fac = beta;
            if (it == 1) fac = fac / (1.0f + beta);
            a.fac[0] = fac;
            a.aux = ds;
            a.out = last ? dr : nullptr;
            if (!last || keep_spectrum) a.fwd_out = C0;
            x_inv(c, G, a, XE_ITER_FIXED, nb, "x iterate fixed");
        }
    }
}

// compute_displacements grids of nb catalogs: 3 floats per cell (Psi_x,
// Psi_y, Psi_z), catalog b at disp + 3 b R.  The k and x tables come from
// slot[0] (the catalogs share one box).
void displacement_grid(baorec_ctx *c, const Grid &G, int kind, Slot *slot, const float *field, float *disp,
                       bool c0_ready = false, int nb = 1, const Work *Wp = nullptr) {
    FFTBufs f = fft_bufs(c, G);
    tables(c, G, slot, f);
    const Work W = Wp ? *Wp : work_single(f, G);
    float2 *C0 = W.C0, *Gb = W.Gb, *Hb = W.Hb;
    if (!c0_ready) x_fwd(c, G, field, C0, nb);
    strided_plain(c, G, f, 1, false, C0, nb);
    StridedArgs z = sargs(G, f, 2, 2, kind == BAOREC_MULTIGRID ? MUL_I : MUL_I_INVK2);
    z.job[0].in = C0;
    z.job[0].out[0] = Gb;
    z.job[0].pw[0] = 0;
    z.job[0].out[1] = Gb + W.q;
    z.job[0].pw[1] = 1;
    z.job[0].nout = 2;
    strided(c, G, z, 1, nb, "z displacement");
    StridedArgs y = sargs(G, f, 1, 1, MUL_NONE);
    y.job[0].in = Gb;
    y.job[0].nout = 2;
    y.job[0].out[0] = Hb; y.job[0].pw[0] = 0;
    y.job[0].out[1] = Hb + W.q; y.job[0].pw[1] = 1;
    y.job[1].in = Gb + W.q;
    y.job[1].nout = 1;
    y.job[1].out[0] = Hb + 2 * W.q; y.job[1].pw[0] = 0;
    strided(c, G, y, 2, nb, "y displacement");
    XInvArgs a = xargs(G, f, slot);
    a.in[0] = Hb; a.kxpow[0] = 1;
    a.in[1] = Hb + W.q; a.kxpow[1] = 0;
    a.in[2] = Hb + 2 * W.q; a.kxpow[2] = 0;
    a.nin = 3;
    a.out = disp;
    x_inv(c, G, a, XE_DISP, nb, "x displacement");
}

// ---------------------------------------------------------------- multigrid

struct MGPyr {
    std::vector<Grid> g;
    std::vector<float *> A, B, F;      // FP32 levels
    std::vector<double *> Ad, Bd, Fd;  // FP64 levels (RBGS coarse)
    std::vector<MGLevel> L;
    std::vector<int> cur;              // 0: A holds v, 1: B holds v
    std::vector<int> dbl;
    std::vector<int> zero;             // 1: v is zero and not stored (its next fused sweep reads none)
    int lc = 0;                        // levels >= lc run in one k_mg_coarse launch
};

// Levels of at most this many cells are latency-bound: they run inside one
// workgroup (k_mg_coarse) instead of one launch per sweep.
constexpr int64_t kMGCoarseCells = 512;

bool mg_recurse(const Grid &G) {
    return G.nx > 4 && G.nx % 2 == 0 && G.ny > 4 && G.ny % 2 == 0 && G.nz > 4 && G.nz % 2 == 0;
}

MGPyr mg_build(baorec_ctx *c, const Grid &G0, const baorec_params *p, Slot *slot, float *v0,
               const float *f0, int fp64_max) {
    MGPyr P;
    Grid g = G0;
    while (true) {
        P.g.push_back(g);
        if (!mg_recurse(g)) break;
        g = Grid{g.nx / 2, g.ny / 2, g.nz / 2, 0};
    }
    const int nl = (int)P.g.size();
    size_t xvtot = 0;
    for (auto &q : P.g) xvtot += 2 * (q.nx + q.ny + q.nz);
    float *xv = bufT<float>(c, "mg_xv", xvtot);
    size_t off = 0;
    for (int i = 0; i < nl; i++) {
        const Grid &q = P.g[i];
        const bool d = i > 0 && q.nx <= fp64_max && q.ny <= fp64_max && q.nz <= fp64_max;
        P.dbl.push_back(d);
        std::string s = std::to_string(i);
        if (i == 0) {
            P.A.push_back(v0);
            P.B.push_back(bufT<float>(c, "mg_B0", q.R()));
            P.F.push_back(const_cast<float *>(f0));
        } else if (!d) {
            P.A.push_back(bufT<float>(c, "mg_A" + s, q.R()));
            P.B.push_back(bufT<float>(c, "mg_B" + s, q.R()));
            P.F.push_back(bufT<float>(c, "mg_F" + s, q.R()));
        } else {
            P.A.push_back(nullptr);
            P.B.push_back(nullptr);
            P.F.push_back(nullptr);
        }
        if (d) {
            P.Ad.push_back(bufT<double>(c, "mg_Ad" + s, q.R()));
            P.Bd.push_back(bufT<double>(c, "mg_Bd" + s, q.R()));
            P.Fd.push_back(bufT<double>(c, "mg_Fd" + s, q.R()));
        } else {
            P.Ad.push_back(nullptr);
            P.Bd.push_back(nullptr);
            P.Fd.push_back(nullptr);
        }
        MGLevel L{};
        L.n[0] = q.nx;
        L.n[1] = q.ny;
        L.n[2] = q.nz;
        for (int a = 0; a < 3; a++) {
            // src/multigrid.jl:196-199 (FP32 arithmetic)
            const float cell = p->box_size[a] / (float)L.n[a];
            L.cell[a] = cell;
            L.cell2[a] = cell * cell;
            L.icell2[a] = 1.0f / L.cell2[a];
            L.losn[a] = p->has_los ? p->los[a] / cell : 0.f;
        }
        L.xv = xv + off;
        L.pv = L.xv + (q.nx + q.ny + q.nz);
        hipchk(c, launch_mg_xvec(xv + off, slot, q.nx, q.ny, q.nz, c->stream), "mg xvec");
        off += 2 * (q.nx + q.ny + q.nz);
        P.L.push_back(L);
        P.cur.push_back(0);
        P.zero.push_back(0);
    }
    // first level from which every level is small and of one precision
    P.lc = nl;
    for (int i = nl - 1; i >= 0; i--) {
        if (P.g[i].R() > kMGCoarseCells || P.dbl[i] != P.dbl[nl - 1] || nl - i > kMGCoarseLevels) break;
        P.lc = i;
    }
    return P;
}

float *vcur(MGPyr &P, int l) { return P.cur[l] ? P.B[l] : P.A[l]; }
float *vother(MGPyr &P, int l) { return P.cur[l] ? P.A[l] : P.B[l]; }
double *vcurd(MGPyr &P, int l) { return P.cur[l] ? P.Bd[l] : P.Ad[l]; }
double *votherd(MGPyr &P, int l) { return P.cur[l] ? P.Ad[l] : P.Bd[l]; }

// n damped Jacobi sweeps (jacobi!, src/multigrid.jl:193-210) ping-ponging
// between a (the input) and b; returns the one holding the result.  On
// grids of at least 128^3 cells whose sides are multiples of 64, 16 and 8,
// sweeps run 3 (or 2) per launch, temporally blocked (the same bits).
static float *jacobi_sweeps(baorec_ctx *c, const MGLevel &L, int64_t R, float *a, float *b, const float *f,
                            const baorec_params *p, int n) {
    const bool tb = mg_jacobi_tb_ok(L) && R >= ((int64_t)1 << 21);
    for (int left = n; left > 0;) {
        const int k = !tb || left == 1 ? 1 : (left == 2 || left == 4 ? 2 : 3);
        ProfScope ps(c, k == 1 ? "mg_jacobi" : "mg_jacobi_tb", 12.0 * R);
        if (k == 1)
            hipchk(c, launch_mg_jacobi(b, a, f, L, p->beta, p->jacobi_damping_factor, p->has_los, c->stream),
                   "jacobi");
        else
            hipchk(c, launch_mg_jacobi_tb(b, a, f, L, p->beta, p->jacobi_damping_factor, p->has_los, k, c->stream),
                   "jacobi (blocked)");
        std::swap(a, b);
        left -= k;
    }
    return a;
}

void mg_jacobi_n(baorec_ctx *c, MGPyr &P, int l, const baorec_params *p, int n) {
    float *res = jacobi_sweeps(c, P.L[l], P.g[l].R(), vcur(P, l), vother(P, l), P.F[l], p, n);
    if (res != vcur(P, l)) P.cur[l] ^= 1;
}

void zero_level(baorec_ctx *c, MGPyr &P, int l) {
    const int64_t n = P.g[l].R();
    if (P.dbl[l]) {
        hipchk(c, launch_zero_words(vcurd(P, l), 2 * n, c->stream), "zero");
    } else {
        hipchk(c, launch_fill(vcur(P, l), n, 0.f, c->stream), "fill");
    }
}

// Levels [P.lc, nl) as one k_mg_coarse launch: a V-cycle from level lc, or
// the whole fmg from lc.
template <typename T>
void mg_coarse(baorec_ctx *c, MGPyr &P, const baorec_params *p, int fmg, int gs) {
    MGCoarse<T> a{};
    a.nl = (int)P.g.size() - P.lc;
    for (int i = 0; i < a.nl; i++) {
        const int l = P.lc + i;
        a.L[i] = P.L[l];
        if constexpr (sizeof(T) == sizeof(double)) {
            a.A[i] = P.Ad[l];
            a.B[i] = P.Bd[l];
            a.F[i] = P.Fd[l];
        } else {
            a.A[i] = P.A[l];
            a.B[i] = P.B[l];
            a.F[i] = P.F[l];
        }
        a.cur[i] = P.cur[l];
    }
    a.fmg = fmg;
    a.gs = gs;
    a.beta = p->beta;
    a.omega = p->jacobi_damping_factor;
    a.has_los = p->has_los;
    a.jacobi_n = p->jacobi_niterations;
    a.rb_sweeps = p->rb_sweeps;
    a.vcycles = p->vcycle_niterations;
    ProfScope ps(c, sizeof(T) == sizeof(double) ? "mg_coarse64" : "mg_coarse", 0.0);
    hipchk(c, launch_mg_coarse<T>(a, c->stream), "mg coarse levels");
    // damped Jacobi flips A/B once per sweep; every level visit does an even
    // number (pre + post), so P.cur is unchanged
}

// The FP32 residual of level l restricted to the right-hand side of level
// l + 1 (residual! + reduce!, src/multigrid.jl:378-395, 641-651): on grids of
// >= 128^3 cells one fused pass with the same bits (the residual is not
// stored), else the two kernels through vother(P, l).  coarse = NULL: level
// l + 1 is FP64 (P.Fd[l + 1]).
// fast: the fused pass computes the coefficient with the hardware
// reciprocal (the RBGS solve; the reference schedule keeps its division)
static void mg_residual_reduce(baorec_ctx *c, MGPyr &P, int l, const baorec_params *p, float *coarse,
                               int fast = 0) {
    const double R = (double)P.g[l].R();
    if (coarse && R >= (double)(1 << 21) && mg_residual_restrict_ok(P.L[l])) {
        ProfScope ps(c, "mg_residual_restrict", 8.5 * R);
        hipchk(c, launch_mg_residual_restrict(coarse, vcur(P, l), P.F[l], P.L[l], p->beta, p->has_los, c->stream,
                                              fast),
               "residual + restriction");
        return;
    }
    float *r = vother(P, l);
    {
        ProfScope ps(c, "mg_residual", 12.0 * R);
        hipchk(c, launch_mg_residual(r, vcur(P, l), P.F[l], P.L[l], p->beta, p->has_los, c->stream), "residual");
    }
    const int fn[3] = {P.g[l].nx, P.g[l].ny, P.g[l].nz};
    ProfScope ps(c, "mg_reduce", 4.5 * R);
    if (coarse)
        hipchk(c, launch_mg_reduce(coarse, r, fn, c->stream), "reduce");
    else
        hipchk(c, launch_mg_reduce_t<double, float>(P.Fd[l + 1], r, fn, c->stream), "reduce");
}

// vcycle! (src/multigrid.jl:689-719), reference schedule, FP32
void mg_vcycle_ref(baorec_ctx *c, MGPyr &P, int l, const baorec_params *p) {
    if (l == P.lc) {
        mg_coarse<float>(c, P, p, 0, 0);
        return;
    }
    mg_jacobi_n(c, P, l, p, p->jacobi_niterations);
    if (l + 1 < (int)P.g.size()) {
        mg_residual_reduce(c, P, l, p, P.F[l + 1]);
        zero_level(c, P, l + 1);
        mg_vcycle_ref(c, P, l + 1, p);
        const int cn[3] = {P.g[l + 1].nx, P.g[l + 1].ny, P.g[l + 1].nz};
        {
            ProfScope ps(c, "mg_prolong", 8.5 * P.g[l].R());
            hipchk(c, launch_mg_prolong(vcur(P, l), vcur(P, l + 1), cn, 1, c->stream), "prolong");
        }
    }
    mg_jacobi_n(c, P, l, p, p->jacobi_niterations);
}

// fmg (src/multigrid.jl:722-752), reference schedule
void mg_fmg_ref(baorec_ctx *c, MGPyr &P, int l, const baorec_params *p) {
    if (l == P.lc) {
        mg_coarse<float>(c, P, p, 1, 0);
        return;
    }
    if (l + 1 < (int)P.g.size()) {
        const int fn[3] = {P.g[l].nx, P.g[l].ny, P.g[l].nz};
        hipchk(c, launch_mg_reduce(P.F[l + 1], P.F[l], fn, c->stream), "reduce");
        mg_fmg_ref(c, P, l + 1, p);
        const int cn[3] = {P.g[l + 1].nx, P.g[l + 1].ny, P.g[l + 1].nz};
        hipchk(c, launch_mg_prolong(vcur(P, l), vcur(P, l + 1), cn, 0, c->stream), "prolong");
    } else {
        zero_level(c, P, l);
    }
    for (int i = 0; i < p->vcycle_niterations; i++) mg_vcycle_ref(c, P, l, p);
}

// Multicolour Gauss-Seidel V-cycle (RBGS schedule) with FP64 coarse levels.
// whole sweeps in one launch each (ping-pong between A and B) on levels of
// at least 128^3 cells; below, the per-colour launches are as fast
static bool mg_use_gs4(const MGPyr &P, int l) {
    return !P.dbl[l] && mg_gs4_ok(P.L[l]) && P.g[l].R() >= ((int64_t)1 << 21);
}

// v = 0 on level l before an RBGS V-cycle: on levels of fused sweeps no
// fill, the first sweep reads no v (mg_smooth_rb); otherwise zero_level.
void zero_level_rb(baorec_ctx *c, MGPyr &P, int l) {
    if (l != P.lc && mg_use_gs4(P, l))
        P.zero[l] = 1;
    else
        zero_level(c, P, l);
}

// sum_part (may be NULL; one-launch sweeps only): the last sweep leaves the
// per-workgroup sums of the result there (mg_gs4_blocks partials).  Leaves
// v stored (P.zero[l] cleared).
void mg_smooth_rb(baorec_ctx *c, MGPyr &P, int l, const baorec_params *p, int sweeps, double *sum_part = nullptr) {
    if (mg_use_gs4(P, l)) {
        for (int s = 0; s < sweeps; s++) {
            ProfScope ps(c, "mg_gs4", (P.zero[l] ? 8.0 : 12.0) * P.g[l].R());
            hipchk(c, launch_mg_gs4(vother(P, l), P.zero[l] ? nullptr : vcur(P, l), P.F[l], P.L[l], p->beta,
                                    p->has_los, c->stream, s + 1 == sweeps ? sum_part : nullptr),
                   "gs4");
            P.cur[l] ^= 1;
            P.zero[l] = 0;
        }
        if (P.zero[l]) {   // no sweep: store the zero
            zero_level(c, P, l);
            P.zero[l] = 0;
        }
        return;
    }
    for (int s = 0; s < sweeps; s++)
        for (int col = 0; col < 4; col++) {
            ProfScope ps(c, P.dbl[l] ? "mg_mcgs64" : "mg_mcgs", (P.dbl[l] ? 8.0 : 4.0) * 2.5 * P.g[l].R());
            if (P.dbl[l])
                hipchk(c, launch_mg_rbgs<double>(vcurd(P, l), P.Fd[l], P.L[l], p->beta, p->has_los, col,
                                                 c->stream),
                       "mcgs64");
            else
                hipchk(c, launch_mg_rbgs<float>(vcur(P, l), P.F[l], P.L[l], p->beta, p->has_los, col,
                                                c->stream),
                       "mcgs");
        }
}

// sum_part: as mg_smooth_rb's, for the post-smoothing of level l
void mg_vcycle_rb(baorec_ctx *c, MGPyr &P, int l, const baorec_params *p, double *sum_part = nullptr) {
    if (l == P.lc) {
        if (P.dbl[l])
            mg_coarse<double>(c, P, p, 0, 1);
        else
            mg_coarse<float>(c, P, p, 0, 1);
        return;
    }
    const int nl = (int)P.g.size();
    const bool coarsest = l + 1 >= nl;
    mg_smooth_rb(c, P, l, p, coarsest ? 16 * p->rb_sweeps : p->rb_sweeps);
    if (!coarsest) {
        const int fn[3] = {P.g[l].nx, P.g[l].ny, P.g[l].nz};
        const int cn[3] = {P.g[l + 1].nx, P.g[l + 1].ny, P.g[l + 1].nz};
        const double R = (double)P.g[l].R(), w = P.dbl[l] ? 8.0 : 4.0;
        if (P.dbl[l]) {
            double *r = votherd(P, l);
            {
                ProfScope ps(c, "mg_residual64", 3.0 * w * R);
                hipchk(c, launch_mg_residual_t<double>(r, vcurd(P, l), P.Fd[l], P.L[l], p->beta, p->has_los,
                                                       c->stream), "residual64");
            }
            ProfScope ps(c, "mg_reduce64", 1.125 * w * R);
            hipchk(c, launch_mg_reduce_t<double, double>(P.Fd[l + 1], r, fn, c->stream), "reduce64");
        } else {
            mg_residual_reduce(c, P, l, p, P.dbl[l + 1] ? nullptr : P.F[l + 1], 1);
        }
        zero_level_rb(c, P, l + 1);
        mg_vcycle_rb(c, P, l + 1, p);
        ProfScope ps(c, P.dbl[l] ? "mg_prolong64" : "mg_prolong", (2.0 + 0.125) * w * R);
        if (P.dbl[l])
            hipchk(c, launch_mg_prolong_t<double, double>(vcurd(P, l), vcurd(P, l + 1), cn, 1, c->stream),
                   "prolong64");
        else if (P.dbl[l + 1])
            hipchk(c, launch_mg_prolong_t<float, double>(vcur(P, l), vcurd(P, l + 1), cn, 1, c->stream),
                   "prolong");
        else
            hipchk(c, launch_mg_prolong(vcur(P, l), vcur(P, l + 1), cn, 1, c->stream), "prolong");
    }
    mg_smooth_rb(c, P, l, p, coarsest ? 16 * p->rb_sweeps : p->rb_sweeps, sum_part);
}

// Spectral seed of the RBGS solve: the exact inverse of the discrete
// operator with the LOS fixed, by one forward and inverse FFT.  The LOS is
// the given one, or for a local LOS the direction of the box centre seen
// from the origin (the operator's first-derivative term, which falls off as
// 1/|x|, is left out).  The k = 0 mode is zero.
void mg_spectral_seed(baorec_ctx *c, const Grid &G, const baorec_params *p, Slot *slot, const MGLevel &L0,
                      const float *f, float *out, bool accumulate) {
    double d[3];
    if (p->has_los) {
        for (int a = 0; a < 3; a++) d[a] = p->los[a];
    } else {
        for (int a = 0; a < 3; a++) d[a] = (double)p->box_min[a] + 0.5 * (double)p->box_size[a];
        const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if (r > 0) {
            for (int a = 0; a < 3; a++) d[a] /= r;
        } else {
            d[0] = d[1] = 0.0;
            d[2] = 1.0;
        }
    }
    // the coefficients of the per-cell operator (mg.hip coef()) for this LOS
    double pp[3], s2 = 0.0;
    for (int a = 0; a < 3; a++) {
        pp[a] = d[a] / (double)L0.cell[a];
        s2 += (double)L0.cell2[a] * pp[a] * pp[a];
    }
    const double g = s2 > 0 ? (double)p->beta / s2 : 0.0;
    FFTBufs fb = fft_bufs(c, G);
    tables(c, G, slot, fb);
    x_fwd(c, G, f, fb.cw, 1);
    strided_plain(c, G, fb, 1, false, fb.cw, 1);
    StridedArgs z = sargs(G, fb, 2, 2, MUL_MG_INV);
    for (int a = 0; a < 3; a++) {
        z.mg_gp[a] = (float)((double)L0.icell2[a] + g * pp[a] * pp[a]);
        z.mg_cell[a] = L0.cell[a];
    }
    z.mg_cross[0] = (float)(2.0 * g * pp[0] * pp[1]);
    z.mg_cross[1] = (float)(2.0 * g * pp[0] * pp[2]);
    z.mg_cross[2] = (float)(2.0 * g * pp[1] * pp[2]);
    z.job[0].in = fb.cw;
    z.job[0].out[0] = fb.cw;
    z.job[0].nout = 1;
    z.job[0].pw[0] = 0;
    strided(c, G, z, 1, 1, "z multigrid inverse");
    strided_plain(c, G, fb, 1, true, fb.cw, 1);
    XInvArgs a = xargs(G, fb, slot);
    a.in[0] = fb.cw;
    a.kxpow[0] = 0;
    a.nin = 1;
    a.out = out;
    a.accumulate = accumulate ? 1 : 0;
    x_inv(c, G, a, XE_PLAIN, 1, "x multigrid seed");
}

// ------------------------------------------------------------ learned warm start

// Direction of the spectral seed's fixed LOS (mg_spectral_seed): the given
// LOS, or the box centre seen from the origin.
static void seed_los(const baorec_params *p, float nc[3]) {
    double d[3];
    if (p->has_los) {
        for (int a = 0; a < 3; a++) d[a] = p->los[a];
    } else {
        for (int a = 0; a < 3; a++) d[a] = (double)p->box_min[a] + 0.5 * (double)p->box_size[a];
    }
    const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (int a = 0; a < 3; a++) nc[a] = r > 0 ? (float)(d[a] / r) : (a == 2 ? 1.f : 0.f);
}

static MGLevel cells_of(const Grid &G, const baorec_params *p) {
    MGLevel L{};
    const int n[3] = {G.nx, G.ny, G.nz};
    for (int a = 0; a < 3; a++) {
        L.n[a] = n[a];
        const float cell = p->box_size[a] / (float)n[a];
        L.cell[a] = cell;
        L.cell2[a] = cell * cell;
        L.icell2[a] = 1.0f / L.cell2[a];
        L.losn[a] = p->has_los ? p->los[a] / cell : 0.f;
    }
    return L;
}

// cells_of plus this grid's cell centres (for the local-LOS residual)
static MGLevel level_with_xv(baorec_ctx *c, const Grid &G, const baorec_params *p, const Slot *slot,
                             const std::string &key) {
    MGLevel L = cells_of(G, p);
    float *xv = bufT<float>(c, key, 2 * ((size_t)G.nx + G.ny + G.nz));
    hipchk(c, launch_mg_xvec(xv, slot, G.nx, G.ny, G.nz, c->stream), "xvec");
    L.xv = xv;
    L.pv = xv + (G.nx + G.ny + G.nz);
    return L;
}

// The spectral seed refined by one preconditioned Richardson step:
// phi_s = A_c^-1 f, r_s = f - A phi_s (the local-LOS operator), then
// out = phi_s + A_c^-1 r_s.  phis and rs (may be NULL) receive phi_s and
// psi = A_c^-1 r_s separately.
static void spectral_refined(baorec_ctx *c, const Grid &G, const baorec_params *p, Slot *slot, const MGLevel &L,
                             const float *f, float *phis, float *rs, float *psi) {
    mg_spectral_seed(c, G, p, slot, L, f, phis, false);
    hipchk(c, launch_mg_residual(rs, phis, f, L, p->beta, p->has_los, c->stream), "seed residual");
    mg_spectral_seed(c, G, p, slot, L, rs, psi, false);
}

static int ilog2i(int n) {
    int k = 0;
    while ((1 << k) < n) k++;
    return k;
}

// The network's input for nb grids f (each G.R() cells, device): per item the
// rms s of f (FP64), the spectral seed phis, its local-LOS residual rs and
// psi = A_c^-1 rs, and the 8 feature channels in the network's input
// tensor.  Returns the input tensor.  The slot must hold the box (its sums
// and error flag are left as they are).
static uint16_t *warm_inputs(baorec_ctx *c, const Grid &G, const baorec_params *p, Slot *slot, const float *f,
                             int nb, float *phis, float *rs, float *psi, double *ss) {
    const int64_t N = G.R();
    const MGLevel L = level_with_xv(c, G, p, slot, "warm_xv");
    std::vector<double> h(nb);
    double *tmp = bufT<double>(c, "warm_sum", 1);
    for (int b = 0; b < nb; b++) {
        hipchk(c, launch_sumsq<float>(f + b * N, N, tmp, bufT<double>(c, "red_part", 1024), c->stream), "rms f");
        hipchk(c, hipMemcpyAsync(&h[b], tmp, sizeof(double), hipMemcpyDeviceToHost, c->stream), "D2H");
        spectral_refined(c, G, p, slot, L, f + b * N, phis + b * N, rs + b * N, psi + b * N);
    }
    sync(c);
    for (int b = 0; b < nb; b++) h[b] = sqrt(h[b] / (double)N);
    hipchk(c, hipMemcpyAsync(ss, h.data(), sizeof(double) * nb, hipMemcpyHostToDevice, c->stream), "H2D");
    uint16_t *X = nullptr;
    hipchk(c, unet_input(c->net, nb, ilog2i(G.nx), ilog2i(G.ny), ilog2i(G.nz), &X), "network buffers");
    FeatArgs a{};
    a.lx = ilog2i(G.nx);
    a.ly = ilog2i(G.ny);
    a.lz = ilog2i(G.nz);
    a.batch = nb;
    for (int k = 0; k < 3; k++) {
        a.bmin[k] = p->box_min[k];
        a.h[k] = L.cell[k];
    }
    seed_los(p, a.nc);
    a.ih2 = L.icell2[0];
    a.href = 100.f * L.cell[0];
    hipchk(c, launch_unet_features(f, phis, rs, psi, ss, a, X, c->stream), "features");
    return X;
}

static void need_net(baorec_ctx *c) {
    if (!unet_ready(c->net))
        fail(c, BAOREC_ERR_NO_RESULT, "no warm-start network (baorec_warmstart_init or baorec_warmstart_load first)");
}

// The learned seed of the multigrid solve on level 0 of P: the spectral seed
// of f plus the network's correction.  The network runs on the pyramid level
// whose cell is closest to the one it was trained at (restricting f there),
// and its correction is prolonged back to level 0.
static void learned_seed(baorec_ctx *c, const Grid &G, const baorec_params *p, Slot *slot, const MGLevel &L0,
                         const float *f0, float *seed) {
    need_net(c);
    mg_spectral_seed(c, G, p, slot, L0, f0, seed, false);
    // the level of the network
    const float ht = unet_h_train(c->net);
    int lev = 0;
    Grid g = G;
    std::vector<Grid> gs{G};
    while (mg_recurse(g) && g.nx / 2 >= 16 && g.ny / 2 >= 16 && g.nz / 2 >= 16) {
        const float hc = p->box_size[0] / (float)g.nx, hn = 2.f * hc;
        if (!(ht > 0.f) || fabsf(logf(hn / ht)) >= fabsf(logf(hc / ht))) break;
        g = Grid{g.nx / 2, g.ny / 2, g.nz / 2, g.nx / 4 + 1};
        gs.push_back(g);
        lev++;
    }
    const Grid Gc = gs[lev];
    // f on that level
    const float *fc = f0;
    for (int l = 1; l <= lev; l++) {
        float *dst = bufT<float>(c, "warm_f" + std::to_string(l), gs[l].R());
        const int fn[3] = {gs[l - 1].nx, gs[l - 1].ny, gs[l - 1].nz};
        hipchk(c, launch_mg_reduce(dst, fc, fn, c->stream), "restrict f");
        fc = dst;
    }
    float *phis = bufT<float>(c, "warm_phis", Gc.R());
    float *rs = bufT<float>(c, "warm_rs", Gc.R());
    float *psi = bufT<float>(c, "warm_psi", Gc.R());
    double *ss = bufT<double>(c, "warm_ss", 1);
    warm_inputs(c, Gc, p, slot, fc, 1, phis, rs, psi, ss);
    {
        ProfScope ps(c, "unet_forward", 0.0);
        hipchk(c, unet_forward(c->net, c->stream), "network");
    }
    const float hc = p->box_size[0] / (float)Gc.nx;
    if (lev == 0) {
        hipchk(c, launch_unet_apply_seed(unet_output(c->net), ss, hc * hc, Gc.R(), 1, seed, c->stream), "correction");
        return;
    }
    float *e = bufT<float>(c, "warm_e" + std::to_string(lev), Gc.R());
    hipchk(c, launch_unet_apply_seed(unet_output(c->net), ss, hc * hc, Gc.R(), 0, e, c->stream), "correction");
    // prolong the correction to level 0 and add it to the spectral seed
    for (int l = lev; l >= 1; l--) {
        const int cn[3] = {gs[l].nx, gs[l].ny, gs[l].nz};
        float *dst = l == 1 ? seed : bufT<float>(c, "warm_e" + std::to_string(l - 1), gs[l - 1].R());
        hipchk(c, launch_mg_prolong(dst, e, cn, l == 1 ? 1 : 0, c->stream), "prolong correction");
        e = dst;
    }
}

// Solves A phi = f on level 0.  v0 (device) receives the potential.
// Partials of the fixed-order FP64 grid reductions (at most 1024 blocks).
static double *red_part(baorec_ctx *c) { return bufT<double>(c, "red_part", 1024); }

void multigrid_solve(baorec_ctx *c, const Grid &G, baorec_params *p, Slot *slot, float *v0,
                     const float *f0) {
    if (p->mg_schedule == BAOREC_MG_REFERENCE) {
        MGPyr P = mg_build(c, G, p, slot, v0, f0, 0);
        mg_fmg_ref(c, P, 0, p);
        if (P.cur[0])
            hipchk(c, hipMemcpyAsync(v0, P.B[0], sizeof(float) * G.R(), hipMemcpyDeviceToDevice, c->stream),
                   "copy");
        return;
    }
    if (p->mg_schedule != BAOREC_MG_RBGS) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "unknown mg_schedule");
    // RBGS in correction form: each V-cycle (FP32 fine levels, FP64 coarse
    // ones) solves A e = r from e = 0 for the residual r = f - A v of the
    // solution so far, and v += e.  v is kept in FP64 and its fine residual
    // computed in FP64 (iterative refinement): in FP32 alone the residual of
    // a converged v is rounding in A v, about 1e-4 of ||f|| at 512^3.
    // The periodic operator annihilates constants, so its range is the
    // orthogonal complement of the left null vector w (A^T w = 0): only
    // f - alpha with alpha = (w . f) / (w . 1) is consistent.  With a fixed LOS
    // A is symmetric, w = 1 and alpha is the mean of f.  With a local LOS the
    // first-derivative term makes A non-symmetric and w != 1; a V-cycle leaves
    // the residual's constant component unreduced, so after each residual its
    // mean is moved from the residual into the shift of f (alpha += mean(r),
    // r -= mean(r)).  alpha converges to (w . f) / (w . 1) and the residual to
    // zero (without it, an arbitrary f stalls at a floor: 1e-2 of ||f|| on the
    // 512^3 lightcone).  The solution is that of f minus this constant.
    const int64_t N = G.R();
    // FP64 scalars: [0] N alpha, [1] sum f^2, [2] |r|^2, [3] |dr|^2, [4] sum e,
    // [5] |r0|^2, [6] sum r, [7] sum f
    double *red = bufT<double>(c, "mg_red", 8);
    hipchk(c, launch_zero_words(red, 16, c->stream), "zero");
    hipchk(c, launch_sum<float>(f0, N, red + 7, red_part(c), c->stream), "sum f");
    hipchk(c, hipMemcpyAsync(red, red + 7, sizeof(double), hipMemcpyDeviceToDevice, c->stream), "copy");
    hipchk(c, launch_sumsq<float>(f0, N, red + 1, red_part(c), c->stream), "norm f");
    float *rb[2] = {bufT<float>(c, "mg_r0", N), bufT<float>(c, "mg_r1", N)};
    double *phi = bufT<double>(c, "mg_phi64", N);   // FP64 solution (set by the first residual call)
    MGPyr P = mg_build(c, G, p, slot, v0, rb[0], p->fp64_coarse_max);
    // With sides that are multiples of 64, 16 and 8 the refinement step
    // (phi += e - mean(e), r = f - alpha - A phi, sum r) is one z-marching
    // pass (k_mg_refine) and the shift with both norms another; otherwise
    // one launch per operation.
    const bool fused = mg_refine_ok(P.L[0]);
    double *phib = fused ? bufT<double>(c, "mg_phi64b", N) : nullptr;
    double *rpart = fused ? bufT<double>(c, "mg_refine_part", mg_refine_blocks(P.L[0])) : nullptr;
    double *npart = bufT<double>(c, "mg_norm_part", 2048);
    // the V-cycle's last sweep on level 0 sums the correction e (unless
    // level 0 runs inside mg_coarse or by per-colour launches)
    double *epart = P.lc != 0 && mg_use_gs4(P, 0) ? bufT<double>(c, "mg_gs4_part", mg_gs4_blocks(P.L[0])) : nullptr;
    // phi += e - mean(e) (e != NULL), r = residual, alpha += mean(r) and
    // r -= mean(r) (local LOS), norms[0] = |r|^2 and, with rprev, norms[1] =
    // |r - rprev|^2.  The first call (start) sets phi to the FP32 seed, or to
    // zero without one.
    auto residual = [&](float *r, const float *e, const float *rprev, double *norms, bool start = false,
                        const float *seed = nullptr) {
        if (fused) {
            {
                ProfScope ps(c, "mg_refine", (e ? 28.0 : (start ? (seed ? 24.0 : 16.0) : 16.0)) * N);
                hipchk(c, launch_mg_refine(phib, start ? nullptr : phi, seed, e, red + 4, f0, red, r, rpart, P.L[0],
                                           p->beta, p->has_los, c->stream), "refine");
            }
            if (e || start) std::swap(phi, phib);
            ProfScope ps(c, "mg_shift_norms", (p->has_los ? 4.0 : 8.0) * N + (rprev ? 4.0 : 0.0) * N);
            if (!p->has_los) {
                hipchk(c, launch_sum_partials(rpart, mg_refine_blocks(P.L[0]), 1, red + 6, 1, c->stream), "sum r");
                hipchk(c, launch_accumulate(red, red + 6, c->stream), "shift");
            }
            hipchk(c, launch_shift_norms(r, rprev, N, p->has_los ? nullptr : red + 6, norms, npart, c->stream),
                   "shift, norms");
            return;
        }
        if (start) {
            if (seed)
                hipchk(c, launch_convert<double, float>(phi, seed, N, c->stream), "seed to FP64");
            else
                hipchk(c, launch_zero_words(phi, 2 * N, c->stream), "zero");
        }
        if (e) {
            ProfScope ps(c, "mg_correct", 20.0 * N);
            hipchk(c, launch_add_correction<double>(phi, e, N, red + 4, c->stream), "v += e");
        }
        {
            ProfScope ps(c, "mg_residual_fp64", 16.0 * N);
            hipchk(c, launch_mg_residual_mixed(r, phi, f0, red, (double)N, P.L[0], p->beta, p->has_los,
                                               c->stream), "residual");
        }
        if (!p->has_los) {   // alpha += mean(r), r -= mean(r)
            ProfScope ps(c, "mg_shift", 8.0 * N);
            hipchk(c, launch_sum<float>(r, N, red + 6, red_part(c), c->stream), "sum r");
            hipchk(c, launch_accumulate(red, red + 6, c->stream), "shift");
            hipchk(c, launch_add_const<float>(r, N, red + 6, (double)N, -1.0, c->stream), "r - mean");
        }
        hipchk(c, launch_sumsq<float>(r, N, norms, red_part(c), c->stream), "norm r");
        if (rprev) hipchk(c, launch_sumsq_diff<float>(r, rprev, N, norms + 1, red_part(c), c->stream), "norm dr");
    };
    float *seed = nullptr;
    if (p->mg_warm_start) {
        if (p->mg_warm_start < 1 || p->mg_warm_start > 3)
            fail(c, BAOREC_ERR_INVALID_ARGUMENT, "unknown mg_warm_start");
        if (!fft_side_ok(G.nx) || !fft_side_ok(G.ny) || !fft_side_ok(G.nz))
            fail(c, BAOREC_ERR_UNSUPPORTED, "the warm starts need grid sides that are powers of two in [16, 1024]");
        seed = bufT<float>(c, "mg_seed", N);
        if (p->mg_warm_start == 1) {
            mg_spectral_seed(c, G, p, slot, P.L[0], f0, seed, false);
        } else {
            if (p->mg_seed_steps < 0 || p->mg_seed_steps > 16)
                fail(c, BAOREC_ERR_INVALID_ARGUMENT, "mg_seed_steps must be in [0, 16]");
            if (p->mg_warm_start == 3)
                mg_spectral_seed(c, G, p, slot, P.L[0], f0, seed, false);
            else
                learned_seed(c, G, p, slot, P.L[0], f0, seed);
            // spectral Richardson steps: seed += A_c^-1 (f - A seed)
            float *rs = bufT<float>(c, "mg_seed_r", N);
            for (int k = 0; k < p->mg_seed_steps; k++) {
                hipchk(c, launch_mg_residual_fast(rs, seed, f0, P.L[0], p->beta, p->has_los, c->stream),
                       "seed residual");
                mg_spectral_seed(c, G, p, slot, P.L[0], rs, seed, true);   // seed += A_c^-1 rs
            }
        }
    }
    residual(rb[0], nullptr, nullptr, red + 5, true, seed);   // of the start (f minus its mean for v = 0)
    double h[8];
    hipchk(c, hipMemcpyAsync(h, red, sizeof(double) * 8, hipMemcpyDeviceToHost, c->stream), "D2H");
    sync(c);
    const double f2 = h[1] - h[7] * h[7] / (double)N;   // ||f - mean(f)||^2
    const double fn = f2 > 0 ? sqrt(f2) : 0.0;
    p->seed_residual = fn > 0 ? sqrt(h[5]) / fn : 0.0;
    // Stop at ||f - alpha - A v|| <= tol ||f - mean(f)||; or once a V-cycle
    // changes the residual by <= tol ||f||, or two V-cycles in a row reduce
    // it by less than 10 % (a safeguard: the consistent system converges).
    int cyc = 0, stall = 0;
    double rel = p->seed_residual, chg = rel;
    while (rel > p->tolerance && chg > p->tolerance && stall < 2 && cyc < p->max_vcycles) {
        const double rel_prev = rel;
        float *r = rb[cyc & 1], *rn = rb[(cyc + 1) & 1];
        P.F[0] = r;
        zero_level_rb(c, P, 0);
        mg_vcycle_rb(c, P, 0, p, epart);
        // v += e - mean(e): v stays zero-mean (A annihilates constants), so an
        // inconsistent f does not make it drift by a constant every cycle
        if (epart)   // the last sweep's partial sums
            hipchk(c, launch_sum_partials(epart, mg_gs4_blocks(P.L[0]), 1, red + 4, 1, c->stream), "sum e");
        else
            hipchk(c, launch_sum<float>(vcur(P, 0), N, red + 4, red_part(c), c->stream), "sum e");
        cyc++;
        residual(rn, vcur(P, 0), r, red + 2);
        hipchk(c, hipMemcpyAsync(h + 2, red + 2, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream), "D2H");
        sync(c);
        rel = fn > 0 ? sqrt(h[2]) / fn : 0.0;
        chg = fn > 0 ? sqrt(h[3]) / fn : 0.0;
        stall = rel > 0.9 * rel_prev ? stall + 1 : 0;
    }
    p->vcycles_done = cyc;
    p->final_residual = rel;
    p->final_change = chg;
    hipchk(c, hipMemcpyAsync(h, red, sizeof(double), hipMemcpyDeviceToHost, c->stream), "D2H");
    sync(c);
    p->rhs_shift = h[0] / (double)N;
    hipchk(c, launch_convert<float, double>(v0, phi, N, c->stream), "phi to FP32");
}

// ------------------------------------------------------------ entry helpers

struct Inputs {
    float *x, *y, *z;
    const float *w;
    float *rx, *ry, *rz;
    const float *rw;
};

Inputs stage_catalogs(baorec_ctx *c, int mem, float *x, float *y, float *z, const float *w, int64_t n,
                      const float *rx, const float *ry, const float *rz, const float *rw, int64_t nr) {
    Inputs in{};
    in.x = const_cast<float *>(stage_in(c, mem, x, n, "in_x"));
    in.y = const_cast<float *>(stage_in(c, mem, y, n, "in_y"));
    in.z = const_cast<float *>(stage_in(c, mem, z, n, "in_z"));
    in.w = stage_in(c, mem, w, n, "in_w");
    in.rx = const_cast<float *>(stage_in(c, mem, rx, nr, "in_rx"));
    in.ry = const_cast<float *>(stage_in(c, mem, ry, nr, "in_ry"));
    in.rz = const_cast<float *>(stage_in(c, mem, rz, nr, "in_rz"));
    in.rw = stage_in(c, mem, rw, nr, "in_rw");
    return in;
}

void check_err_flag(baorec_ctx *c, const Slot *slot) {
    if (c->capturing) return;   // checked after the replay
    Slot h = read_slot(c, slot);
    if (h.err) fail(c, BAOREC_ERR_OUT_OF_BOX, "a particle indexes outside the grid (BoundsError)");
}

void mutate_back(baorec_ctx *c, int mem, const baorec_params *p, int wrap, float *x, float *y, float *z,
                 const Inputs &in, int64_t n) {
    if (mem == BAOREC_MEM_DEVICE || !wrap || !p->mutate_inputs) return;
    copy_out(c, mem, x, in.x, n);
    copy_out(c, mem, y, in.y, n);
    copy_out(c, mem, z, in.z, n);
}

// run! pieces shared by both recon kinds: box + overdensity into delta
void run_common(baorec_ctx *c, baorec_params *p, const Grid &G, int mem, float *x, float *y, float *z,
                const float *w, int64_t n, const float *rx, const float *ry, const float *rz,
                const float *rw, int64_t nr, float *delta, bool do_setup_box, Inputs &in, Slot *slot,
                bool fwd_spectrum = false) {
    if (n < 0 || nr < 0) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "negative particle count");
    in = stage_catalogs(c, mem, x, y, z, w, n, rx, ry, rz, rw, nr);
    if (nr > 0 && do_setup_box) {
        if (!(p->box_pad >= 0.f)) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "box_pad");
        slot_from_params(c, slot, p, (double)nr);
        hipchk(c, launch_setup_box(in.rx, in.ry, in.rz, nr, slot, p->box_pad, c->stream), "setup_box");
        Slot h = read_slot(c, slot);
        for (int a = 0; a < 3; a++) {
            p->box_size[a] = h.box_size[a];
            p->box_min[a] = h.box_min[a];
        }
    } else {
        check_box(c, p);
        slot_from_params(c, slot, p, (double)nr);
    }
    overdensity(c, G, p, slot, delta, in.x, in.y, in.z, in.w, n, in.rx, in.ry, in.rz, in.rw, nr,
                fwd_spectrum);
}

template <class F>
int guarded(baorec_ctx *c, F &&fn) {
    if (!c) return BAOREC_ERR_INVALID_ARGUMENT;
    try {
        hipError_t e = hipSetDevice(c->device);
        if (e != hipSuccess) {
            c->err = std::string("hipSetDevice: ") + hipGetErrorString(e);
            return BAOREC_ERR_HIP;
        }
        fn();
        c->err.clear();
        return BAOREC_OK;
    } catch (const Fail &f) {
        (void)hipStreamSynchronize(c->stream);
        return f.code;
    } catch (const std::bad_alloc &) {
        c->err = "host allocation failed";
        return BAOREC_ERR_OUT_OF_MEMORY;
    }
}

// ------------------------------------------------------------- batching

// A batch runs in scatter groups of consecutive catalogs: one multi-catalog
// sort and accumulation per group, into density grids laid out
// [catalog][cell] (at most kBatchScatterCats catalogs and kBatchPlacements
// sorted placements of 16 B, uint32 cursors).  The rest of the pipeline runs
// over FFT subgroups of a scatter group, whose other grid and spectrum
// buffers stay within kBatchFFTBytes.  Small FFT subgroups keep what one pass
// reads and writes near the size of the 256 MB Infinity Cache, so the next
// pass finds part of it there: at 256^3 (0.86 GB of FFT buffers per catalog)
// groups of 2 ran the 64-catalog batch in 72.9 ms, against 77.9 (1), 75.3
// (4), 77.3 (8) and 76.5 ms (16) with scatter and FFT grouped alike.
constexpr int kBatchScatterCats = 16;
constexpr double kBatchScatterBytes = 8.0 * (1ull << 30);   // density grids of a scatter group
constexpr double kBatchFFTBytes = 2.0 * (1ull << 30);
constexpr int64_t kBatchPlacements = (int64_t)1 << 29;

struct BatchGroup {
    int c0, c1;          // catalogs [c0, c1)
    int nblk;            // sort blocks
    size_t o64, o32;     // offsets of the group's tables in the packed arrays
};

// FFT buffers per catalog: delta_s (FP32), displacement (3 x FP32) and seven
// spectra.
double batch_fft_bytes_per_catalog(const Grid &G) { return 16.0 * G.R() + 7.0 * 8.0 * G.C(); }

// Catalogs per FFT subgroup.
int batch_fft_group(const Grid &G) {
    return (int)std::max(1.0, std::floor(kBatchFFTBytes / batch_fft_bytes_per_catalog(G)));
}

std::vector<BatchGroup> batch_groups(const Grid &G, const baorec_params *p, int ncat, const int64_t *off,
                                     std::vector<int64_t> &t64, std::vector<int> &t32) {
    const int per = p->mas == BAOREC_MAS_TSC ? 27 : 8;
    int maxnb = (int)std::max(1.0, std::min((double)kBatchScatterCats, std::floor(kBatchScatterBytes / (4.0 * G.R()))));
    maxnb = std::min(maxnb, ncat);
    const int ng = (ncat + maxnb - 1) / maxnb;
    maxnb = (ncat + ng - 1) / ng;   // balanced group sizes
    std::vector<BatchGroup> gs;
    int c = 0;
    while (c < ncat) {
        int e = c + 1;
        while (e < ncat && e - c < maxnb && (off[e + 1] - off[c]) * per <= kBatchPlacements) e++;
        BatchGroup g{c, e, 0, t64.size(), t32.size()};
        std::vector<int64_t> rel(e - c + 1);
        for (int k = c; k <= e; k++) rel[k - c] = off[k] - off[c];
        std::vector<int> blk_cat, cat_blk0;
        std::vector<int64_t> blk_beg;
        g.nblk = sort_blocks(rel[e - c], rel.data(), e - c, blk_cat, blk_beg, cat_blk0);
        // int64: relative offsets (nb + 1) | block starts (nblk + 1)
        t64.insert(t64.end(), rel.begin(), rel.end());
        t64.insert(t64.end(), blk_beg.begin(), blk_beg.end());
        // int32: block catalogs (nblk, at least one entry) | first block per catalog (nb + 1)
        if (blk_cat.empty()) blk_cat.push_back(0);
        t32.insert(t32.end(), blk_cat.begin(), blk_cat.end());
        t32.insert(t32.end(), cat_blk0.begin(), cat_blk0.end());
        gs.push_back(g);
        c = e;
    }
    return gs;
}

void batch_fft_group_run(baorec_ctx *c, const Grid &G, const baorec_params *p, const int64_t *off, int c0, int nb,
                         Slot *slots, float *rho, float *ds, float *disp, float2 *cw, FFTBufs &f, float *x,
                         float *y, float *z, float *ox, float *oy, float *oz, int32_t *err);

// One group: run!(IterativeRecon, grid, x, y, z, w) (periodic box, no randoms)
// followed by reconstructed_positions(recon, x, y, z; field = :sum) for each
// catalog, every step batched over the group's catalogs.
void batch_group(baorec_ctx *c, const Grid &G, const baorec_params *p, const int64_t *off,
                 const BatchGroup &g, const int64_t *d64, const int *d32, float *x, float *y, float *z,
                 const float *w, float *ox, float *oy, float *oz, int32_t *err, int maxnb) {
    const int nb = g.c1 - g.c0;
    const int64_t base = off[g.c0], ntot = off[g.c1] - base;
    const int64_t R = G.R(), C = G.C();
    Slot *slots = bufT<Slot>(c, "bt_slots", maxnb);
    hipLaunchKernelGGL(k_slots_set, dim3((nb + 63) / 64), dim3(64), 0, c->stream, slots, nb, p->box_size[0],
                       p->box_size[1], p->box_size[2], p->box_min[0], p->box_min[1], p->box_min[2]);
    hipchk(c, hipGetLastError(), "slots_set");
    float *rho = bufT<float>(c, "bt_rho", (size_t)maxnb * R);
    const int64_t *d_off = d64 + g.o64, *d_beg = d_off + nb + 1;
    const int *d_cat = d32 + g.o32, *d_blk0 = d_cat + std::max(g.nblk, 1);

    // cic! of every catalog (one-sided wrap into the box, src/recon.jl:48-50)
    ScatterArgs a{};
    a.x = x + base;
    a.y = y + base;
    a.z = z + base;
    a.w = w + base;
    a.offsets = d_off;
    a.ncat = nb;
    a.ntotal = ntot;
    a.slot = slots;
    a.slotStride = 1;
    a.nx = G.nx;
    a.ny = G.ny;
    a.nz = G.nz;
    a.wrap = 1;
    a.mutate = p->mutate_inputs;
    a.tsc = p->mas == BAOREC_MAS_TSC;
    tiles_for(c, G, a, nb > 0 ? ntot / nb : 0);
    const int64_t T = (int64_t)a.ntx * a.nty * a.ntz;
    a.count = bufT<uint32_t>(c, "bt_count", (size_t)maxnb * T);
    a.cursor = bufT<uint32_t>(c, "bt_cursor", (size_t)maxnb * T);
    a.sorted = bufT<float4>(c, "bt_sorted", scatter_capacity(ntot > 0 ? ntot : 1, a.tsc));
    a.err = err;
    a.rho = rho;
    a.rstride = R;
    a.wmax = bufT<uint32_t>(c, "bt_wmax", maxnb);
    hipchk(c, launch_zero_words(a.wmax, nb, c->stream), "zero");
    const double pb = (double)ntot * (a.mutate ? 40.0 : 28.0) + 16.0 * ntot * 1.2;
    if (sort_fits(a)) {
        uint32_t *H = bufT<uint32_t>(c, "bt_hist", (size_t)std::max(g.nblk, 1) * T);
        ProfScope ps(c, "scatter_bin", pb);
        hipchk(c, launch_sort(a, g.nblk, H, d_cat, d_beg, d_blk0, c->stream), "batch sort");
    } else {
        ProfScope ps(c, "scatter_bin", pb);
        hipchk(c, launch_scatter(a, nullptr, c->stream), "batch scatter");
    }
    a.part = bufT<double>(c, "bt_part", (size_t)maxnb * T);
    {
        ProfScope ps(c, "scatter_accum", 16.0 * ntot * 1.2 + 4.0 * R * nb);
        hipchk(c, launch_scatter_accum(a, 0, c->stream), "batch accumulate");
    }
    // per-catalog FP64 totals of the density grids into slots[k].sum[0]
    hipchk(c, launch_sum_partials(a.part, T, nb, &slots[0].sum[0], (int64_t)(sizeof(Slot) / sizeof(double)),
                                  c->stream),
           "grid totals");

    // the rest in FFT subgroups [s0, s0 + nf) of the scatter group
    const int nfmax = std::min(batch_fft_group(G), maxnb);
    float *ds = bufT<float>(c, "bt_delta_s", (size_t)nfmax * R);
    float *disp = bufT<float>(c, "bt_disp", (size_t)3 * nfmax * R);
    float2 *cw = bufT<float2>(c, "bt_cwork", (size_t)7 * nfmax * C);
    FFTBufs f = fft_bufs(c, G);
    tables(c, G, slots, f);
    for (int s0 = 0; s0 < nb; s0 += nfmax) {
        const int nf = std::min(nfmax, nb - s0);
        batch_fft_group_run(c, G, p, off, g.c0 + s0, nf, slots + s0, rho + (int64_t)s0 * R, ds, disp, cw, f, x, y,
                            z, ox, oy, oz, err);
    }
}

// Smoothing, overdensity, iterations, displacements and positions of nb
// catalogs (from catalog c0) whose density grids are rho (R apart).
void batch_fft_group_run(baorec_ctx *c, const Grid &G, const baorec_params *p, const int64_t *off, int c0, int nb,
                         Slot *slots, float *rho, float *ds, float *disp, float2 *cw, FFTBufs &f, float *x,
                         float *y, float *z, float *ox, float *oy, float *oz, int32_t *err) {
    const int64_t R = G.R(), C = G.C();
    const Work W{cw, cw + (int64_t)nb * C, cw + (int64_t)4 * nb * C, (int64_t)nb * C};
    // smooth! and the overdensity (src/recon.jl:42-58); the per-catalog FP64
    // totals came from the scatter
    smooth_yz(c, G, f, rho, nb, p->smoothing_radius);
    {
        XInvArgs xa = xargs(G, f, slots);
        xa.rin[0] = rho;
        xa.kxpow[0] = 0;
        xa.nin = 1;
        xa.R2 = p->smoothing_radius * p->smoothing_radius;
        xa.bias = p->bias;
        xa.ran_min = p->ran_min;
        xa.out = ds;
        xa.fwd_out = W.C0;   // R2C(delta_s) for the first iteration
        x_inv(c, G, xa, XE_OD_BOX, nb, "x overdensity");
    }
    // iterations; only the spectrum of delta_r is kept
    iterate_all(c, G, p, slots, ds, nullptr, true, true, nb, &W);
    displacement_grid(c, G, BAOREC_ITERATIVE, slots, nullptr, disp, true, nb, &W);

    // reconstructed_positions(...; field = :sum), src/recon.jl:366-381
    for (int b = 0; b < nb; b++) {
        const int64_t o = off[c0 + b], n = off[c0 + b + 1] - o;
        if (n == 0) continue;
        GatherArgs ga{};
        ga.x = x + o;
        ga.y = y + o;
        ga.z = z + o;
        ga.n = n;
        ga.disp = disp + (int64_t)3 * b * R;
        ga.slot = slots + b;
        ga.nx = G.nx;
        ga.ny = G.ny;
        ga.nz = G.nz;
        ga.field = BAOREC_FIELD_SUM;
        ga.positions = 1;
        ga.has_los = p->has_los;
        ga.los[0] = p->los[0];
        ga.los[1] = p->los[1];
        ga.los[2] = p->los[2];
        ga.f = p->f;
        ga.ox = ox + o;
        ga.oy = oy + o;
        ga.oz = oz + o;
        ga.err = err;
        ProfScope ps(c, "gather", (double)n * (24.0 + 96.0));
        hipchk(c, launch_gather(ga, c->stream), "batch gather");
    }
}

// ------------------------------------------------------------- cosmology

// Derived densities of BAOrec.Cosmology (src/cosmo.jl:22-58) in the Julia
// constructor's precision: Float32 fields, with the photon and neutrino
// densities computed in Float64 from physical constants and rounded to
// Float32, and Omega_Lambda closing the budget in Float32.
CosmoArgs cosmo_args(const baorec_cosmology *c) {
    const double G = 6.67408e-11, hbar = 6.582119514e-16, mpc = 3.085677581491367e22;
    const double evc2 = 1.782661907e-36, kb = 8.6173303e-5, ckms = kSpeedOfLightKms;
    const double pi = 3.14159265358979323846;
    const float h2 = c->h * c->h;
    const double hbar3 = hbar * hbar * hbar, impc = 1.0 / (1e-3 * mpc), c3 = (ckms * 1e3) * (ckms * 1e3) * (ckms * 1e3);
    const float rho_g0 = (float)(30000.0 / (8.0 * pi * G) * hbar3 * (impc * impc) * (1.0 / evc2) * c3);
    const double tk = 2.725 * kb, tr = (double)c->T_cmb / 2.725;
    const float omega_g0 = (float)((pi * pi / 15.0) * (tk * tk * tk * tk) / (double)rho_g0 * (tr * tr * tr * tr) /
                                   (double)h2);
    const float omega_nu0 = (float)((double)c->Neff * (7.0 / 8.0 * pow(4.0 / 11.0, 4.0 / 3.0)) * (double)omega_g0);
    const float omega_b0 = c->omega_b / h2, omega_c0 = c->omega_c / h2;
    const float omega_l0 = 1.0f - c->Omega_k0 - omega_b0 - omega_c0 - omega_nu0 - omega_g0;
    CosmoArgs a;
    a.omega_nu0 = omega_nu0;
    a.omega_g0 = omega_g0;
    a.omega_b0 = omega_b0;
    a.omega_c0 = omega_c0;
    a.omega_k0 = c->Omega_k0;
    a.omega_l0 = omega_l0;
    a.w0 = c->w0;
    a.wa = c->wa;
    a.h100 = (double)(c->h * 100.0f);
    return a;
}

// Knots of range(z_tab_min, z_tab_max, length = z_tab_num) (src/cosmo.jl:88).
void cosmo_knots(const baorec_cosmology *cp, double *h) {
    const int num = cp->z_tab_num;
    const long double a = cp->z_tab_min, b = cp->z_tab_max;
    for (int i = 0; i < num; i++) h[i] = (double)(a + (b - a) * (long double)i / (long double)(num - 1));
    h[num - 1] = cp->z_tab_max;
}

void check_cosmology(baorec_ctx *c, const baorec_cosmology *cp) {
    if (!cp) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "cosmology is NULL");
    if (cp->z_tab_num < 2 || !(cp->z_tab_max > cp->z_tab_min) || cp->z_tab_min < 0.0 || !(cp->h > 0.f))
        fail(c, BAOREC_ERR_INVALID_ARGUMENT, "bad cosmology table (need z_tab_num >= 2, 0 <= z_tab_min < z_tab_max, h > 0)");
}

// The table for cosmology c on the device (rebuilt when the parameters change).
CosmoTable cosmo_table(baorec_ctx *c, const baorec_cosmology *cp) {
    check_cosmology(c, cp);
    const int num = cp->z_tab_num;
    double *zk = bufT<double>(c, "cosmo_z", num);
    double *rk = bufT<double>(c, "cosmo_r", num);
    int *guide = bufT<int>(c, "cosmo_guide", num);
    CosmoTable t{zk, rk, num, cp->z_tab_min, (double)(num - 1) / (cp->z_tab_max - cp->z_tab_min), guide, num,
                 0.0, 0.0};
    if (!c->cosmo_valid || memcmp(&c->cosmo, cp, sizeof(*cp)) != 0) {
        std::vector<double> h(num);
        cosmo_knots(cp, h.data());
        hipchk(c, hipMemcpyAsync(zk, h.data(), sizeof(double) * num, hipMemcpyHostToDevice, c->stream), "H2D");
        double *seg = bufT<double>(c, "cosmo_seg", num);
        hipchk(c, launch_cosmo_table(cosmo_args(cp), zk, seg, rk, num, c->stream), "cosmology table");
        double ends[2];
        hipchk(c, hipMemcpyAsync(&ends[0], rk, sizeof(double), hipMemcpyDeviceToHost, c->stream), "D2H");
        hipchk(c, hipMemcpyAsync(&ends[1], rk + num - 1, sizeof(double), hipMemcpyDeviceToHost, c->stream), "D2H");
        hipchk(c, hipStreamSynchronize(c->stream), "cosmology table");
        c->cosmo_r0 = ends[0];
        c->cosmo_dr_inv = (double)num / (ends[1] - ends[0]);
        t.r0 = c->cosmo_r0;
        t.dr_inv = c->cosmo_dr_inv;
        hipchk(c, launch_cosmo_guide(t, guide, c->stream), "cosmology guide");
        c->cosmo = *cp;
        c->cosmo_valid = 1;
    }
    t.r0 = c->cosmo_r0;
    t.dr_inv = c->cosmo_dr_inv;
    return t;
}

int *cosmo_err(baorec_ctx *c) {
    int *e = bufT<int>(c, "cosmo_err", 1);
    hipchk(c, launch_zero_words(e, 1, c->stream), "zero");
    return e;
}

void cosmo_check(baorec_ctx *c, int *e, const char *what) {
    int h = 0;
    hipchk(c, hipMemcpyAsync(&h, e, sizeof(int), hipMemcpyDeviceToHost, c->stream), "D2H");
    sync(c);
    if (h) fail(c, BAOREC_ERR_OUT_OF_BOX, std::string(what) + ": value outside the cosmology table "
                                                          "(the reference raises BoundsError)");
}

}  // namespace

// ====================================================================== ABI

// Milliseconds of the fastest of `reps` runs of fn on the context's stream
// (after one untimed run).
template <class F>
double time_min(baorec_ctx *c, int reps, F &&fn) {
    fn();
    hipEvent_t a = take_event(c), b = take_event(c);
    double best = 1e30;
    for (int r = 0; r < reps; r++) {
        hipchk(c, hipEventRecord(a, c->stream), "hipEventRecord");
        fn();
        hipchk(c, hipEventRecord(b, c->stream), "hipEventRecord");
        hipchk(c, hipEventSynchronize(b), "hipEventSynchronize");
        float ms = 0.f;
        hipchk(c, hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
        best = ms < best ? ms : best;
    }
    c->event_pool.push_back(a);
    c->event_pool.push_back(b);
    return best;
}

// ----------------------------------------- allocation-free sky conversion
//
// baorec_sky_to_cartesian_noalloc keeps its comoving-distance table and error
// flag in `scratch`, a file-scope device array (one copy per GPU, held until
// the process ends), instead of context buffers.  Host code never uses the
// array's name as an address: baorec_create looks up each device's copy with
// hipGetSymbolAddress once, with that device current, and the entry point
// passes that address to the kernels.
namespace {
constexpr int kNoallocMaxKnots = 131072;
// zk, rk, seg (FP64, kNoallocMaxKnots each), guide (int32), error flag (int32)
constexpr size_t kNoallocZk = 0;
constexpr size_t kNoallocRk = kNoallocZk + sizeof(double) * kNoallocMaxKnots;
constexpr size_t kNoallocSeg = kNoallocRk + sizeof(double) * kNoallocMaxKnots;
constexpr size_t kNoallocGuide = kNoallocSeg + sizeof(double) * kNoallocMaxKnots;
constexpr size_t kNoallocErr = kNoallocGuide + sizeof(int) * kNoallocMaxKnots;
constexpr size_t kNoallocBytes = kNoallocErr + 256;
constexpr int kNoallocMaxDevices = 64;

struct NoallocDevice {
    unsigned char *ptr = nullptr;   // this device's copy of `scratch`
    int table_valid = 0;            // the table in it is for `cosmo`
    baorec_cosmology cosmo{};
    double r0 = 0.0, dr_inv = 0.0;
};
NoallocDevice g_noalloc[kNoallocMaxDevices];
double g_noalloc_knots[kNoallocMaxKnots];   // host staging of the knots
std::mutex g_noalloc_lock;                  // guards the two arrays above

__global__ void k_zero_flag(int *e) {
    if (threadIdx.x == 0) *e = 0;
}
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoallocBytes];

// Once per device, with that device current (called by baorec_create).
static hipError_t noalloc_setup(int device) {
    if (device < 0 || device >= kNoallocMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_noalloc_lock);
    NoallocDevice &d = g_noalloc[device];
    if (d.ptr) return hipSuccess;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    d.ptr = static_cast<unsigned char *>(p);
    return hipSuccess;
}

extern "C" {

void baorec_params_default(baorec_params *p, float bias, float f, float smoothing_radius) {
    if (!p) return;
    memset(p, 0, sizeof(*p));
    p->bias = bias;
    p->f = f;
    p->smoothing_radius = smoothing_radius;
    p->n_iter = 3;
    p->beta = f / bias;
    p->jacobi_damping_factor = 0.4f;
    p->jacobi_niterations = 5;
    p->vcycle_niterations = 6;
    p->mas = BAOREC_MAS_CIC;
    p->mg_schedule = BAOREC_MG_REFERENCE;
    p->tolerance = 1e-6;
    p->max_vcycles = 200;
    p->rb_sweeps = 2;
    p->fp64_coarse_max = 8;
    p->box_pad = 500.0f;
    p->ran_min = 0.01f;
    p->mutate_inputs = 1;
    p->mg_seed_steps = 3;
}

const char *baorec_version(void) { return "baorec_hip 0.1.0 (gfx950)"; }

int baorec_create(baorec_ctx **out, int device) {
    if (!out) return BAOREC_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        g_create_err = "no HIP device available";
        return BAOREC_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= n) {
        g_create_err = "device index out of range";
        return BAOREC_ERR_INVALID_ARGUMENT;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        g_create_err = "hipGetDeviceProperties failed";
        return BAOREC_ERR_HIP;
    }
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        g_create_err = std::string("device is ") + prop.gcnArchName + ", this library is built for gfx950";
        return BAOREC_ERR_NO_DEVICE;
    }
    baorec_ctx *c = new (std::nothrow) baorec_ctx();
    if (!c) return BAOREC_ERR_OUT_OF_MEMORY;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        g_create_err = "stream creation failed";
        delete c;
        return BAOREC_ERR_HIP;
    }
    std::vector<float2> tw(4096);
    fill_twiddles(tw.data());
    set_twiddle_table(tw.data());
    if (hipDeviceSynchronize() != hipSuccess) {
        g_create_err = "twiddle upload failed";
        (void)hipStreamDestroy(c->stream);
        delete c;
        return BAOREC_ERR_HIP;
    }
    if (noalloc_setup(device) != hipSuccess) {
        g_create_err = "hipGetSymbolAddress of the allocation-free scratch array failed";
        (void)hipStreamDestroy(c->stream);
        delete c;
        return BAOREC_ERR_HIP;
    }
    *out = c;
    return BAOREC_OK;
}

int baorec_destroy(baorec_ctx *c) {
    if (!c) return BAOREC_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto &g : c->graphs) (void)hipGraphExecDestroy(g.exec);
    for (auto &kv : c->bufs)
        if (kv.second.p) (void)hipFree(kv.second.p);
    unet_destroy(c->net);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return BAOREC_OK;
}

const char *baorec_last_error(const baorec_ctx *c) {
    if (!c) return g_create_err.c_str();
    return c->err.c_str();
}

static void run_body(baorec_ctx *c, int kind, baorec_params *p, const int32_t grid[3], int mem, float *x, float *y,
                     float *z, const float *w, int64_t n, const float *rx, const float *ry, const float *rz,
                     const float *rw, int64_t nr) {
        check_params(c, p);
        Grid G = check_grid(c, grid, true);
        if (mem != BAOREC_MEM_HOST && mem != BAOREC_MEM_DEVICE)
            fail(c, BAOREC_ERR_INVALID_ARGUMENT, "mem");
        Slot *slot = slot_buf(c);
        float *result = bufT<float>(c, "result", G.R());
        float *ds = bufT<float>(c, "delta_s", G.R());
        Inputs in;
        c->has_result = 0;
        c->disp_valid = 0;
        const bool iter = kind == BAOREC_ITERATIVE;
        run_common(c, p, G, mem, x, y, z, w, n, rx, ry, rz, rw, nr, ds, true, in, slot,
                   iter && p->n_iter > 0);
        if (iter) {
            iterate_all(c, G, p, slot, ds, result, p->n_iter > 0, true);
        } else {
            multigrid_solve(c, G, p, slot, result, ds);
        }
        check_err_flag(c, slot);
        mutate_back(c, mem, p, nr == 0, x, y, z, in, n);
        sync(c);
        c->has_result = 1;
        c->result_kind = kind;
        for (int a = 0; a < 3; a++) c->result_grid[a] = grid[a];
        c->result_gen++;
        // the spectrum of the iterative result is left in cwork[0]
        if (iter && p->n_iter > 0) c->spec_gen = c->result_gen;
}

static int run_impl(baorec_ctx *c, int kind, baorec_params *p, const int32_t grid[3], int mem,
                    float *x, float *y, float *z, const float *w, int64_t n, const float *rx,
                    const float *ry, const float *rz, const float *rw, int64_t nr) {
    return guarded(c, [&] { run_body(c, kind, p, grid, mem, x, y, z, w, n, rx, ry, rz, rw, nr); });
}

int baorec_run_iterative(baorec_ctx *c, baorec_params *p, const int32_t grid[3], int mem, float *x,
                         float *y, float *z, const float *w, int64_t n, const float *rx,
                         const float *ry, const float *rz, const float *rw, int64_t nr) {
    return run_impl(c, BAOREC_ITERATIVE, p, grid, mem, x, y, z, w, n, rx, ry, rz, rw, nr);
}

int baorec_run_multigrid(baorec_ctx *c, baorec_params *p, const int32_t grid[3], int mem, float *x,
                         float *y, float *z, const float *w, int64_t n, const float *rx,
                         const float *ry, const float *rz, const float *rw, int64_t nr) {
    return run_impl(c, BAOREC_MULTIGRID, p, grid, mem, x, y, z, w, n, rx, ry, rz, rw, nr);
}

int baorec_result_info(const baorec_ctx *c, int32_t grid[3], int32_t *kind, uint64_t *generation) {
    if (!c || !grid || !kind) return BAOREC_ERR_INVALID_ARGUMENT;
    if (!c->has_result) return BAOREC_ERR_NO_RESULT;
    for (int a = 0; a < 3; a++) grid[a] = c->result_grid[a];
    *kind = c->result_kind;
    if (generation) *generation = c->result_gen;
    return BAOREC_OK;
}

int baorec_result_copy(baorec_ctx *c, int mem, float *dst, uint64_t generation) {
    return guarded(c, [&] {
        if (!c->has_result) fail(c, BAOREC_ERR_NO_RESULT, "no cached result (run! first)");
        if (generation != c->result_gen)
            fail(c, BAOREC_ERR_NO_RESULT,
                 "stale result: a later run! on this context replaced the cached result");
        if (!dst) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "dst is NULL");
        const int64_t N = (int64_t)c->result_grid[0] * c->result_grid[1] * c->result_grid[2];
        float *r = bufT<float>(c, "result", N);
        hipchk(c, hipMemcpyAsync(dst, r, sizeof(float) * N,
                                 mem == BAOREC_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                                 c->stream), "copy result");
        sync(c);
    });
}

int baorec_setup_overdensity(baorec_ctx *c, const baorec_params *p, const int32_t grid[3], int mem,
                             float *delta, float *x, float *y, float *z, const float *w, int64_t n,
                             const float *rx, const float *ry, const float *rz, const float *rw,
                             int64_t nr) {
    return guarded(c, [&] {
        check_params(c, p);
        Grid G = check_grid(c, grid, true);
        baorec_params q = *p;
        Slot *slot = slot_buf(c) + 1;
        float *d = stage_out(c, mem, delta, G.R(), "out_grid");
        Inputs in;
        run_common(c, &q, G, mem, x, y, z, w, n, rx, ry, rz, rw, nr, d, false, in, slot);
        check_err_flag(c, slot);
        copy_out(c, mem, delta, d, G.R());
        mutate_back(c, mem, p, nr == 0, x, y, z, in, n);
        sync(c);
    });
}

int baorec_reconstructed_overdensity(baorec_ctx *c, const baorec_params *p, const int32_t grid[3],
                                     int mem, float *delta, float *x, float *y, float *z,
                                     const float *w, int64_t n, const float *rx, const float *ry,
                                     const float *rz, const float *rw, int64_t nr) {
    return guarded(c, [&] {
        check_params(c, p);
        Grid G = check_grid(c, grid, true);
        baorec_params q = *p;
        Slot *slot = slot_buf(c) + 1;
        float *d = stage_out(c, mem, delta, G.R(), "out_grid");
        float *ds = bufT<float>(c, "delta_s2", G.R());
        Inputs in;
        run_common(c, &q, G, mem, x, y, z, w, n, rx, ry, rz, rw, nr, ds, false, in, slot);
        iterate_all(c, G, &q, slot, ds, d);
        check_err_flag(c, slot);
        copy_out(c, mem, delta, d, G.R());
        mutate_back(c, mem, p, nr == 0, x, y, z, in, n);
        sync(c);
    });
}

int baorec_reconstructed_potential(baorec_ctx *c, baorec_params *p, const int32_t grid[3], int mem,
                                   float *phi, float *x, float *y, float *z, const float *w,
                                   int64_t n, const float *rx, const float *ry, const float *rz,
                                   const float *rw, int64_t nr) {
    return guarded(c, [&] {
        check_params(c, p);
        Grid G = check_grid(c, grid, true);
        Slot *slot = slot_buf(c) + 1;
        float *d = stage_out(c, mem, phi, G.R(), "out_grid");
        float *ds = bufT<float>(c, "delta_s2", G.R());
        Inputs in;
        run_common(c, p, G, mem, x, y, z, w, n, rx, ry, rz, rw, nr, ds, false, in, slot);
        multigrid_solve(c, G, p, slot, d, ds);
        check_err_flag(c, slot);
        copy_out(c, mem, phi, d, G.R());
        mutate_back(c, mem, p, nr == 0, x, y, z, in, n);
        sync(c);
    });
}

static void shifts_body(baorec_ctx *c, const baorec_params *p, int mem, const float *x, const float *y,
                        const float *z, int64_t n, const float *field3d, const int32_t grid[3], int kind,
                        int field, float *ox, float *oy, float *oz, int positions) {
        check_params(c, p);
        check_box(c, p);
        if (field < 0 || field > 2) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "field must be disp, rsd or sum");
        if (n < 0) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "negative particle count");
        Grid G;
        float *disp;
        Slot *slot;
        if (!field3d) {
            if (!c->has_result) fail(c, BAOREC_ERR_NO_RESULT, "no cached result (run! first)");
            G = check_grid(c, c->result_grid, true);
            kind = c->result_kind;
            slot = slot_buf(c);
            disp = bufT<float>(c, "disp_cache", (size_t)3 * G.R());
            const float key[6] = {p->box_size[0], p->box_size[1], p->box_size[2],
                                  p->box_min[0], p->box_min[1], p->box_min[2]};
            if (!c->disp_valid || c->disp_gen != c->result_gen || memcmp(key, c->disp_box, sizeof key)) {
                slot_from_params(c, slot, p, 0.0);
                const bool spec = c->spec_gen == c->result_gen;  // R2C(result) kept by run!
                displacement_grid(c, G, kind, slot, bufT<float>(c, "result", G.R()), disp, spec);
                c->disp_valid = 1;
                c->disp_gen = c->result_gen;
                memcpy(c->disp_box, key, sizeof key);
            } else {
                slot_from_params(c, slot, p, 0.0);
            }
        } else {
            G = check_grid(c, grid, true);
            if (kind != BAOREC_ITERATIVE && kind != BAOREC_MULTIGRID)
                fail(c, BAOREC_ERR_INVALID_ARGUMENT, "kind");
            slot = slot_buf(c) + 1;
            slot_from_params(c, slot, p, 0.0);
            const float *fd = stage_in(c, mem, field3d, G.R(), "in_field");
            disp = bufT<float>(c, "disp_tmp", (size_t)3 * G.R());
            displacement_grid(c, G, kind, slot, fd, disp);
        }
        const float *dx = stage_in(c, mem, x, n, "sh_x");
        const float *dy = stage_in(c, mem, y, n, "sh_y");
        const float *dz = stage_in(c, mem, z, n, "sh_z");
        float *o1 = stage_out(c, mem, ox, n, "sh_ox");
        float *o2 = stage_out(c, mem, oy, n, "sh_oy");
        float *o3 = stage_out(c, mem, oz, n, "sh_oz");
        GatherArgs a{};
        a.x = dx;
        a.y = dy;
        a.z = dz;
        a.n = n;
        a.disp = disp;
        a.slot = slot;
        a.nx = G.nx;
        a.ny = G.ny;
        a.nz = G.nz;
        a.field = field;
        a.positions = positions;
        a.has_los = p->has_los;
        a.los[0] = p->los[0];
        a.los[1] = p->los[1];
        a.los[2] = p->los[2];
        a.f = p->f;
        a.ox = o1;
        a.oy = o2;
        a.oz = o3;
        a.err = &slot->err;
        {
            ProfScope ps(c, "gather", (double)n * (24.0 + 96.0));
            hipchk(c, launch_gather(a, c->stream), "gather");
        }
        check_err_flag(c, slot);
        copy_out(c, mem, ox, o1, n);
        copy_out(c, mem, oy, o2, n);
        copy_out(c, mem, oz, o3, n);
        sync(c);
}

static int shifts_impl(baorec_ctx *c, const baorec_params *p, int mem, const float *x, const float *y,
                       const float *z, int64_t n, const float *field3d, const int32_t grid[3], int kind,
                       int field, float *ox, float *oy, float *oz, int positions) {
    return guarded(c, [&] {
        shifts_body(c, p, mem, x, y, z, n, field3d, grid, kind, field, ox, oy, oz, positions);
    });
}

int baorec_read_shifts(baorec_ctx *c, const baorec_params *p, int mem, const float *x, const float *y,
                       const float *z, int64_t n, const float *field3d, const int32_t grid[3], int kind,
                       int field, float *sx, float *sy, float *sz) {
    return shifts_impl(c, p, mem, x, y, z, n, field3d, grid, kind, field, sx, sy, sz, 0);
}

int baorec_reconstructed_positions(baorec_ctx *c, const baorec_params *p, int mem, const float *x,
                                   const float *y, const float *z, int64_t n, const float *field3d,
                                   const int32_t grid[3], int kind, int field, float *ox, float *oy,
                                   float *oz) {
    return shifts_impl(c, p, mem, x, y, z, n, field3d, grid, kind, field, ox, oy, oz, 1);
}

// The arguments that fix the launch sequence of baorec_run_positions_graph.
static std::string graph_key(const baorec_params *p, const int32_t grid[3], const void *const ptr[7], int64_t n,
                             const baorec_ctx *c) {
    std::string k;
    auto put = [&](const void *v, size_t b) { k.append(static_cast<const char *>(v), b); };
    put(grid, 3 * sizeof(int32_t));
    put(ptr, 7 * sizeof(void *));
    put(&n, sizeof n);
    put(&p->bias, sizeof(float));
    put(&p->f, sizeof(float));
    put(&p->smoothing_radius, sizeof(float));
    put(p->box_size, sizeof p->box_size);
    put(p->box_min, sizeof p->box_min);
    put(&p->has_los, sizeof(int32_t));
    put(p->los, sizeof p->los);
    put(&p->n_iter, sizeof(int32_t));
    put(&p->beta, sizeof(float));
    put(&p->mas, sizeof(int32_t));
    put(&p->ran_min, sizeof(float));
    put(&p->mutate_inputs, sizeof(int32_t));
    // autotuned choices change the kernels the sequence launches
    for (const auto &kv : c->tune_tile) {
        put(kv.first.data(), sizeof(int64_t) * kv.first.size());
        put(kv.second.data(), sizeof(int) * kv.second.size());
    }
    for (const auto &kv : c->tune_pair) {
        put(kv.first.data(), sizeof(int) * kv.first.size());
        put(&kv.second, sizeof(int));
    }
    return k;
}

int baorec_run_positions_graph(baorec_ctx *c, baorec_params *p, const int32_t grid[3], float *x, float *y, float *z,
                               const float *w, int64_t n, float *ox, float *oy, float *oz) {
    return guarded(c, [&] {
        if (!p || !grid) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "NULL argument");
        if (n <= 0 || !x || !y || !z || !w || !ox || !oy || !oz)
            fail(c, BAOREC_ERR_INVALID_ARGUMENT, "positions, weights and outputs (device) are required");
        check_params(c, p);
        check_box(c, p);
        const void *const ptr[7] = {x, y, z, w, ox, oy, oz};
        const std::string key = graph_key(p, grid, ptr, n, c);
        auto run = [&] {
            run_body(c, BAOREC_ITERATIVE, p, grid, BAOREC_MEM_DEVICE, x, y, z, w, n, nullptr, nullptr, nullptr,
                     nullptr, 0);
            shifts_body(c, p, BAOREC_MEM_DEVICE, x, y, z, n, nullptr, grid, BAOREC_ITERATIVE, 2, ox, oy, oz, 1);
        };
        for (size_t i = 0; i < c->graphs.size(); i++) {
            auto &g = c->graphs[i];
            if (g.key != key) continue;
            if (g.buf_gen != c->buf_gen) {   // a buffer it points into was reallocated
                (void)hipGraphExecDestroy(g.exec);
                c->graphs.erase(c->graphs.begin() + (long)i);
                break;
            }
            hipchk(c, hipGraphLaunch(g.exec, c->stream), "hipGraphLaunch");
            // the bookkeeping run_body and shifts_body do on the host
            check_grid(c, grid, true);
            Slot *slot = slot_buf(c);
            check_err_flag(c, slot);   // synchronises
            c->has_result = 1;
            c->result_kind = BAOREC_ITERATIVE;
            for (int a = 0; a < 3; a++) c->result_grid[a] = grid[a];
            c->result_gen++;
            if (p->n_iter > 0) c->spec_gen = c->result_gen;
            c->disp_valid = 1;
            c->disp_gen = c->result_gen;
            const float kb[6] = {p->box_size[0], p->box_size[1], p->box_size[2],
                                 p->box_min[0], p->box_min[1], p->box_min[2]};
            memcpy(c->disp_box, kb, sizeof kb);
            return;
        }
        // first call with these arguments: run eagerly (this allocates every
        // buffer the sequence uses), then capture the same sequence
        run();
        const int prof = c->prof_on;
        c->prof_on = 0;
        c->capturing = 1;
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed);
        if (e == hipSuccess) {
            try {
                run();
            } catch (const Fail &) {
                (void)hipStreamEndCapture(c->stream, &graph);
                if (graph) (void)hipGraphDestroy(graph);
                c->capturing = 0;
                c->prof_on = prof;
                throw;
            }
            e = hipStreamEndCapture(c->stream, &graph);
        }
        c->capturing = 0;
        c->prof_on = prof;
        hipchk(c, e, "stream capture");
        hipGraphExec_t exec = nullptr;
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        hipchk(c, e, "hipGraphInstantiate");
        c->graphs.push_back({key, c->buf_gen, exec});
    });
}

int baorec_setup_box(baorec_ctx *c, int mem, const float *x, const float *y, const float *z, int64_t n,
                     float pad, float box_size[3], float box_min[3]) {
    return guarded(c, [&] {
        if (n <= 0) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "setup_box needs at least one particle");
        if (!box_size || !box_min) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "output is NULL");
        const float *dx = stage_in(c, mem, x, n, "in_rx");
        const float *dy = stage_in(c, mem, y, n, "in_ry");
        const float *dz = stage_in(c, mem, z, n, "in_rz");
        Slot *slot = slot_buf(c) + 1;
        hipchk(c, launch_setup_box(dx, dy, dz, n, slot, pad, c->stream), "setup_box");
        Slot h = read_slot(c, slot);
        for (int a = 0; a < 3; a++) {
            box_size[a] = h.box_size[a];
            box_min[a] = h.box_min[a];
        }
    });
}

int baorec_cic(baorec_ctx *c, const baorec_params *p, const int32_t grid[3], int mem, float *rho,
               float *x, float *y, float *z, const float *w, int64_t n, int wrap) {
    return guarded(c, [&] {
        check_params(c, p);
        check_box(c, p);
        Grid G = check_grid(c, grid, false);
        if (n < 0) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "negative particle count");
        Slot *slot = slot_buf(c) + 1;
        slot_from_params(c, slot, p, 0.0);
        Inputs in = stage_catalogs(c, mem, x, y, z, w, n, nullptr, nullptr, nullptr, nullptr, 0);
        float *d = mem == BAOREC_MEM_DEVICE ? rho : const_cast<float *>(stage_in(c, mem, rho, G.R(), "out_grid"));
        scatter(c, G, p, slot, d, in.x, in.y, in.z, in.w, n, wrap, 1, "cic");
        check_err_flag(c, slot);
        copy_out(c, mem, rho, d, G.R());
        mutate_back(c, mem, p, wrap, x, y, z, in, n);
        sync(c);
    });
}

int baorec_read_cic(baorec_ctx *c, const baorec_params *p, const int32_t grid[3], int mem, float *out,
                    const float *field, const float *x, const float *y, const float *z, int64_t n) {
    return guarded(c, [&] {
        check_params(c, p);
        check_box(c, p);
        Grid G = check_grid(c, grid, false);
        Slot *slot = slot_buf(c) + 1;
        slot_from_params(c, slot, p, 0.0);
        const float *fd = stage_in(c, mem, field, G.R(), "in_field");
        const float *dx = stage_in(c, mem, x, n, "sh_x");
        const float *dy = stage_in(c, mem, y, n, "sh_y");
        const float *dz = stage_in(c, mem, z, n, "sh_z");
        float *o = stage_out(c, mem, out, n, "sh_ox");
        hipchk(c, launch_read_cic(o, fd, dx, dy, dz, n, slot, G.nx, G.ny, G.nz, &slot->err, c->stream),
               "read_cic");
        check_err_flag(c, slot);
        copy_out(c, mem, out, o, n);
        sync(c);
    });
}

int baorec_smooth(baorec_ctx *c, const baorec_params *p, const int32_t grid[3], int mem, float *field) {
    return guarded(c, [&] {
        check_params(c, p);
        check_box(c, p);
        Grid G = check_grid(c, grid, true);
        Slot *slot = slot_buf(c) + 1;
        slot_from_params(c, slot, p, 0.0);
        float *d = const_cast<float *>(stage_in(c, mem, field, G.R(), "in_field"));
        if (mem == BAOREC_MEM_DEVICE) d = field;
        FFTBufs f = fft_bufs(c, G);
        tables(c, G, slot, f);
        smooth_yz(c, G, f, d, 1, p->smoothing_radius);
        XInvArgs a = xargs(G, f, slot);
        a.rin[0] = d;   // line-local in place: each line is read before it is written
        a.kxpow[0] = 0;
        a.nin = 1;
        a.R2 = p->smoothing_radius * p->smoothing_radius;
        a.out = d;
        x_inv(c, G, a, XE_PLAIN, 1, "x smooth");
        copy_out(c, mem, field, d, G.R());
        sync(c);
    });
}

int baorec_rfft3d(baorec_ctx *c, const int32_t grid[3], int mem, const float *in, float *out_complex) {
    return guarded(c, [&] {
        Grid G = check_grid(c, grid, true);
        if (!out_complex) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "out is NULL");
        const float *d = stage_in(c, mem, in, G.R(), "in_field");
        FFTBufs f = fft_bufs(c, G);
        float2 *o = f.cw;
        x_fwd(c, G, d, o, 1);
        strided_plain(c, G, f, 1, false, o, 1);
        strided_plain(c, G, f, 2, false, o, 1);
        const int64_t nc = (int64_t)G.P * G.ny * G.nz;   // compact complex elements
        float *dout = stage_out(c, mem, out_complex, 2 * nc, "out_spectrum");
        hipchk(c, launch_spectrum_pack(reinterpret_cast<float2 *>(dout), o, G.P, G.PP(), G.PZ(), G.ny, G.nz, 1,
                                       c->stream), "pack spectrum");
        copy_out(c, mem, out_complex, dout, 2 * nc);
        sync(c);
    });
}

int baorec_irfft3d(baorec_ctx *c, const int32_t grid[3], int mem, const float *in_complex, float *out) {
    return guarded(c, [&] {
        Grid G = check_grid(c, grid, true);
        if (!in_complex || !out) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "NULL argument");
        FFTBufs f = fft_bufs(c, G);
        float2 *w = f.cw;  // work copy: the caller's spectrum is not modified
        const int64_t nc = (int64_t)G.P * G.ny * G.nz;
        const float *din = stage_in(c, mem, in_complex, 2 * nc, "in_spectrum");
        hipchk(c, launch_spectrum_pack(w, reinterpret_cast<const float2 *>(din), G.P, G.PP(), G.PZ(), G.ny, G.nz, 0,
                                       c->stream), "unpack spectrum");
        Slot *slot = slot_buf(c) + 1;
        strided_plain(c, G, f, 2, true, w, 1);
        strided_plain(c, G, f, 1, true, w, 1);
        float *o = stage_out(c, mem, out, G.R(), "out_grid");
        XInvArgs a = xargs(G, f, slot);
        a.in[0] = w;
        a.kxpow[0] = 0;
        a.nin = 1;
        a.out = o;
        x_inv(c, G, a, XE_PLAIN, 1, "x inverse");
        copy_out(c, mem, out, o, G.R());
        sync(c);
    });
}

int baorec_k_vec(const int32_t grid[3], const float box_size[3], float *kx, float *ky, float *kz) {
    if (!grid || !box_size || !kx || !ky || !kz) return BAOREC_ERR_INVALID_ARGUMENT;
    float *out[3] = {kx, ky, kz};
    for (int a = 0; a < 3; a++) {
        const int n = grid[a];
        if (n <= 0 || !(box_size[a] > 0.f)) return BAOREC_ERR_INVALID_ARGUMENT;
        const float fs = (float)((2.0 * M_PI * (double)n) / (double)box_size[a]);
        const float mult = fs / (float)n;
        const int m = a == 0 ? n / 2 + 1 : n;
        for (int i = 0; i < m; i++) {
            const int ii = a == 0 ? i : (i < (n + 1) / 2 ? i : i - n);
            out[a][i] = (float)ii * mult;
        }
    }
    return BAOREC_OK;
}

int baorec_x_vec(const int32_t grid[3], const float box_size[3], const float box_min[3], float *xx,
                 float *xy, float *xz) {
    if (!grid || !box_size || !box_min || !xx || !xy || !xz) return BAOREC_ERR_INVALID_ARGUMENT;
    float *out[3] = {xx, xy, xz};
    for (int a = 0; a < 3; a++) {
        if (grid[a] <= 0) return BAOREC_ERR_INVALID_ARGUMENT;
        const float h = box_size[a] / (float)grid[a];
        for (int i = 0; i < grid[a]; i++)
            out[a][i] = (float)((double)box_min[a] + ((double)i + 0.5) * (double)h);
    }
    return BAOREC_OK;
}

// ---------------------------------------------------------------- MG pieces

static MGLevel level_for(baorec_ctx *c, const Grid &G, const baorec_params *p, Slot *slot) {
    MGLevel L{};
    L.n[0] = G.nx;
    L.n[1] = G.ny;
    L.n[2] = G.nz;
    for (int a = 0; a < 3; a++) {
        const float cell = p->box_size[a] / (float)L.n[a];
        L.cell[a] = cell;
        L.cell2[a] = cell * cell;
        L.icell2[a] = 1.0f / L.cell2[a];
        L.losn[a] = p->has_los ? p->los[a] / cell : 0.f;
    }
    float *xv = bufT<float>(c, "mgp_xv", 2 * ((size_t)G.nx + G.ny + G.nz));
    hipchk(c, launch_mg_xvec(xv, slot, G.nx, G.ny, G.nz, c->stream), "xvec");
    L.xv = xv;
    L.pv = xv + (G.nx + G.ny + G.nz);
    return L;
}

int baorec_mg_jacobi(baorec_ctx *c, const baorec_params *p, const int32_t grid[3], int mem, float *v,
                     const float *f) {
    return guarded(c, [&] {
        check_params(c, p);
        check_box(c, p);
        Grid G = check_grid(c, grid, false);
        Slot *slot = slot_buf(c) + 1;
        slot_from_params(c, slot, p, 0.0);
        float *dv = const_cast<float *>(stage_in(c, mem, v, G.R(), "mgp_v"));
        if (mem == BAOREC_MEM_DEVICE) dv = v;
        const float *df = stage_in(c, mem, f, G.R(), "mgp_f");
        float *tmp = bufT<float>(c, "mgp_t", G.R());
        MGLevel L = level_for(c, G, p, slot);
        float *a = jacobi_sweeps(c, L, G.R(), dv, tmp, df, p, p->jacobi_niterations);
        if (a != dv)
            hipchk(c, hipMemcpyAsync(dv, a, sizeof(float) * G.R(), hipMemcpyDeviceToDevice, c->stream), "copy");
        copy_out(c, mem, v, dv, G.R());
        sync(c);
    });
}

int baorec_mg_gauss_seidel(baorec_ctx *c, const baorec_params *p, const int32_t grid[3], int mem, float *v,
                           const float *f, int sweeps, int fused) {
    return guarded(c, [&] {
        check_params(c, p);
        check_box(c, p);
        Grid G = check_grid(c, grid, false);
        if (sweeps < 0) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "sweeps < 0");
        if (G.nx % 2 || G.ny % 2 || G.nz % 2)
            fail(c, BAOREC_ERR_UNSUPPORTED, "the 4-colour Gauss-Seidel sweep needs even grid sides");
        Slot *slot = slot_buf(c) + 1;
        slot_from_params(c, slot, p, 0.0);
        float *dv = const_cast<float *>(stage_in(c, mem, v, G.R(), "mgp_v"));
        if (mem == BAOREC_MEM_DEVICE) dv = v;
        const float *df = stage_in(c, mem, f, G.R(), "mgp_f");
        MGLevel L = level_for(c, G, p, slot);
        if (fused) {
            if (!mg_gs4_ok(L))
                fail(c, BAOREC_ERR_UNSUPPORTED, "the one-launch sweep needs sides that are multiples of 64, 16, 8");
            float *tmp = bufT<float>(c, "mgp_t", G.R());
            float *a = dv, *b = tmp;
            for (int i = 0; i < sweeps; i++) {
                hipchk(c, launch_mg_gs4(b, a, df, L, p->beta, p->has_los, c->stream), "gs4");
                std::swap(a, b);
            }
            if (a != dv)
                hipchk(c, hipMemcpyAsync(dv, a, sizeof(float) * G.R(), hipMemcpyDeviceToDevice, c->stream),
                       "copy");
        } else {
            for (int i = 0; i < sweeps; i++)
                for (int col = 0; col < 4; col++)
                    hipchk(c, launch_mg_rbgs<float>(dv, df, L, p->beta, p->has_los, col, c->stream), "mcgs");
        }
        copy_out(c, mem, v, dv, G.R());
        sync(c);
    });
}

int baorec_mg_residual(baorec_ctx *c, const baorec_params *p, const int32_t grid[3], int mem, float *r,
                       const float *v, const float *f) {
    return guarded(c, [&] {
        check_params(c, p);
        check_box(c, p);
        Grid G = check_grid(c, grid, false);
        Slot *slot = slot_buf(c) + 1;
        slot_from_params(c, slot, p, 0.0);
        const float *dv = stage_in(c, mem, v, G.R(), "mgp_v");
        const float *df = stage_in(c, mem, f, G.R(), "mgp_f");
        float *dr = stage_out(c, mem, r, G.R(), "mgp_r");
        MGLevel L = level_for(c, G, p, slot);
        hipchk(c, launch_mg_residual(dr, dv, df, L, p->beta, p->has_los, c->stream), "residual");
        copy_out(c, mem, r, dr, G.R());
        sync(c);
    });
}

int baorec_mg_residual_restrict(baorec_ctx *c, const baorec_params *p, const int32_t grid[3], int mem,
                                float *coarse, const float *v, const float *f, int fused) {
    return guarded(c, [&] {
        check_params(c, p);
        check_box(c, p);
        Grid G = check_grid(c, grid, false);
        if (G.nx % 2 || G.ny % 2 || G.nz % 2) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "grid must be even");
        Slot *slot = slot_buf(c) + 1;
        slot_from_params(c, slot, p, 0.0);
        const int64_t nc = G.R() / 8;
        const float *dv = stage_in(c, mem, v, G.R(), "mgp_v");
        const float *df = stage_in(c, mem, f, G.R(), "mgp_f");
        float *dc = stage_out(c, mem, coarse, nc, "mgp_c");
        MGLevel L = level_for(c, G, p, slot);
        if (fused) {
            if (!mg_residual_restrict_ok(L))
                fail(c, BAOREC_ERR_UNSUPPORTED, "fused residual + restriction needs sides that are multiples of "
                                                "64, 16 and 2 and fewer than 2^31 cells");
            hipchk(c, launch_mg_residual_restrict(dc, dv, df, L, p->beta, p->has_los, c->stream),
                   "residual + restriction");
        } else {
            float *dr = bufT<float>(c, "mgp_r", G.R());
            hipchk(c, launch_mg_residual(dr, dv, df, L, p->beta, p->has_los, c->stream), "residual");
            const int fn[3] = {G.nx, G.ny, G.nz};
            hipchk(c, launch_mg_reduce(dc, dr, fn, c->stream), "reduce");
        }
        copy_out(c, mem, coarse, dc, nc);
        sync(c);
    });
}

int baorec_mg_reduce(baorec_ctx *c, const int32_t fine_grid[3], int mem, float *coarse, const float *fine) {
    return guarded(c, [&] {
        Grid G = check_grid(c, fine_grid, false);
        if (G.nx % 2 || G.ny % 2 || G.nz % 2) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "fine grid must be even");
        const int64_t nc = G.R() / 8;
        const float *df = stage_in(c, mem, fine, G.R(), "mgp_v");
        float *dc = stage_out(c, mem, coarse, nc, "mgp_r");
        const int fn[3] = {G.nx, G.ny, G.nz};
        hipchk(c, launch_mg_reduce(dc, df, fn, c->stream), "reduce");
        copy_out(c, mem, coarse, dc, nc);
        sync(c);
    });
}

int baorec_mg_prolong(baorec_ctx *c, const int32_t coarse_grid[3], int mem, float *fine,
                      const float *coarse, int add) {
    return guarded(c, [&] {
        Grid G = check_grid(c, coarse_grid, false);
        const int64_t nf = G.R() * 8;
        const float *dc = stage_in(c, mem, coarse, G.R(), "mgp_v");
        float *df = const_cast<float *>(add ? stage_in(c, mem, fine, nf, "mgp_r") : stage_out(c, mem, fine, nf, "mgp_r"));
        const int cn[3] = {G.nx, G.ny, G.nz};
        hipchk(c, launch_mg_prolong(df, dc, cn, add, c->stream), "prolong");
        copy_out(c, mem, fine, df, nf);
        sync(c);
    });
}

int baorec_mg_vcycle(baorec_ctx *c, const baorec_params *p, const int32_t grid[3], int mem, float *v,
                     const float *f) {
    return guarded(c, [&] {
        check_params(c, p);
        check_box(c, p);
        Grid G = check_grid(c, grid, false);
        Slot *slot = slot_buf(c) + 1;
        slot_from_params(c, slot, p, 0.0);
        float *dv = mem == BAOREC_MEM_DEVICE ? v : const_cast<float *>(stage_in(c, mem, v, G.R(), "mgp_vv"));
        const float *df = stage_in(c, mem, f, G.R(), "mgp_ff");
        MGPyr P = mg_build(c, G, p, slot, dv, df, 0);
        mg_vcycle_ref(c, P, 0, p);
        if (P.cur[0])
            hipchk(c, hipMemcpyAsync(dv, P.B[0], sizeof(float) * G.R(), hipMemcpyDeviceToDevice, c->stream),
                   "copy");
        copy_out(c, mem, v, dv, G.R());
        sync(c);
    });
}

int baorec_mg_fmg(baorec_ctx *c, baorec_params *p, const int32_t grid[3], int mem, float *v, const float *f) {
    return guarded(c, [&] {
        check_params(c, p);
        check_box(c, p);
        Grid G = check_grid(c, grid, false);
        Slot *slot = slot_buf(c) + 1;
        slot_from_params(c, slot, p, 0.0);
        float *dv = mem == BAOREC_MEM_DEVICE ? v : bufT<float>(c, "mgp_vv", G.R());
        const float *df = stage_in(c, mem, f, G.R(), "mgp_ff");
        multigrid_solve(c, G, p, slot, dv, df);
        copy_out(c, mem, v, dv, G.R());
        sync(c);
    });
}

void baorec_cosmology_default(baorec_cosmology *c) {
    if (!c) return;
    memset(c, 0, sizeof(*c));
    c->omega_b = 0.0225f;
    c->omega_c = 0.12f;
    c->h = 0.67f;
    c->Neff = 3.044f;
    c->Omega_k0 = 0.f;
    c->T_cmb = 2.725f;
    c->w0 = -1.f;
    c->wa = 0.f;
    c->z_tab_min = 0.0;
    c->z_tab_max = 3.0;
    c->z_tab_num = 100000;
}

void baorec_cosmology_desi(baorec_cosmology *c) {
    if (!c) return;
    baorec_cosmology_default(c);
    c->omega_b = 0.02237f;
    c->omega_c = 0.1200f;
    c->h = 0.6736f;
    c->Neff = 2.0328f + 1.f;
}

int baorec_comoving_distance(baorec_ctx *c, const baorec_cosmology *cp, int mem, const double *z, int64_t n,
                             double *r) {
    return guarded(c, [&] {
        if (n < 0 || (n > 0 && (!z || !r))) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "bad arrays");
        CosmoTable t = cosmo_table(c, cp);
        if (n == 0) return;
        const double *dz = z;
        double *dr = r;
        if (mem != BAOREC_MEM_DEVICE) {
            double *bz = bufT<double>(c, "cosmo_in", n);
            hipchk(c, hipMemcpyAsync(bz, z, sizeof(double) * n, hipMemcpyHostToDevice, c->stream), "H2D");
            dz = bz;
            dr = bufT<double>(c, "cosmo_out", n);
        }
        int *e = cosmo_err(c);
        hipchk(c, launch_comoving_distance(t, dz, dr, n, e, c->stream), "comoving distance");
        if (mem != BAOREC_MEM_DEVICE)
            hipchk(c, hipMemcpyAsync(r, dr, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream), "D2H");
        cosmo_check(c, e, "comoving_distance");
    });
}

int baorec_sky_to_cartesian(baorec_ctx *c, const baorec_cosmology *cp, int mem, const float *ra, const float *dec,
                            const float *red, int64_t n, float *x, float *y, float *z) {
    return guarded(c, [&] {
        if (n < 0) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "negative count");
        CosmoTable t = cosmo_table(c, cp);
        if (n == 0) return;
        const float *a = stage_in(c, mem, ra, n, "sky_a"), *b = stage_in(c, mem, dec, n, "sky_b");
        const float *r = stage_in(c, mem, red, n, "sky_c");
        float *ox = stage_out(c, mem, x, n, "sky_x"), *oy = stage_out(c, mem, y, n, "sky_y");
        float *oz = stage_out(c, mem, z, n, "sky_z");
        const float hh = (cp->h * 100.0f) / 100.0f;   // cosmo.H0 / 100 in Float32, as the driver uses
        int *e = cosmo_err(c);
        {
            ProfScope ps(c, "sky_to_cartesian", 24.0 * n);
            hipchk(c, launch_sky_to_cart(t, hh, a, b, r, ox, oy, oz, n, e, c->stream), "sky_to_cartesian");
        }
        copy_out(c, mem, x, ox, n);
        copy_out(c, mem, y, oy, n);
        copy_out(c, mem, z, oz, n);
        cosmo_check(c, e, "sky_to_cartesian");
    });
}

int baorec_sky_to_cartesian_noalloc(baorec_ctx *c, const baorec_cosmology *cp, const float *ra,
                                    const float *dec, const float *red, int64_t n, float *x, float *y,
                                    float *z) {
    return guarded(c, [&] {
        check_cosmology(c, cp);
        const int num = cp->z_tab_num;
        if (num > kNoallocMaxKnots)
            fail(c, BAOREC_ERR_UNSUPPORTED, "baorec_sky_to_cartesian_noalloc: z_tab_num " + std::to_string(num) +
                                                " exceeds its scratch limit of " +
                                                std::to_string(kNoallocMaxKnots) + " knots");
        if (n < 0) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "negative count");
        if (n > 0 && (!ra || !dec || !red || !x || !y || !z))
            fail(c, BAOREC_ERR_INVALID_ARGUMENT, "NULL array");
        std::lock_guard<std::mutex> lk(g_noalloc_lock);
        NoallocDevice &d = g_noalloc[c->device];
        if (!d.ptr) fail(c, BAOREC_ERR_HIP, "allocation-free scratch not set up on this device");
        double *zk = reinterpret_cast<double *>(d.ptr + kNoallocZk);
        double *rk = reinterpret_cast<double *>(d.ptr + kNoallocRk);
        double *seg = reinterpret_cast<double *>(d.ptr + kNoallocSeg);
        int *guide = reinterpret_cast<int *>(d.ptr + kNoallocGuide);
        int *err = reinterpret_cast<int *>(d.ptr + kNoallocErr);
        if (!d.table_valid || memcmp(&d.cosmo, cp, sizeof(*cp)) != 0) {
            d.table_valid = 0;
            cosmo_knots(cp, g_noalloc_knots);
            hipchk(c, hipMemcpyAsync(zk, g_noalloc_knots, sizeof(double) * num, hipMemcpyHostToDevice, c->stream),
                   "H2D knots");
            hipchk(c, launch_cosmo_table(cosmo_args(cp), zk, seg, rk, num, c->stream), "cosmology table");
            double ends[2];
            hipchk(c, hipMemcpyAsync(&ends[0], rk, sizeof(double), hipMemcpyDeviceToHost, c->stream), "D2H");
            hipchk(c, hipMemcpyAsync(&ends[1], rk + num - 1, sizeof(double), hipMemcpyDeviceToHost, c->stream),
                   "D2H");
            sync(c);
            d.r0 = ends[0];
            d.dr_inv = (double)num / (ends[1] - ends[0]);
            CosmoTable g{zk, rk, num, cp->z_tab_min, (double)(num - 1) / (cp->z_tab_max - cp->z_tab_min),
                         guide, num, d.r0, d.dr_inv};
            hipchk(c, launch_cosmo_guide(g, guide, c->stream), "cosmology guide");
            d.cosmo = *cp;
            d.table_valid = 1;
        }
        if (n == 0) {
            sync(c);
            return;
        }
        CosmoTable t{zk, rk, num, cp->z_tab_min, (double)(num - 1) / (cp->z_tab_max - cp->z_tab_min), guide,
                     num, d.r0, d.dr_inv};
        hipLaunchKernelGGL(k_zero_flag, dim3(1), dim3(64), 0, c->stream, err);
        hipchk(c, hipGetLastError(), "clear flag");
        const float hh = (cp->h * 100.0f) / 100.0f;   // cosmo.H0 / 100 in Float32, as the driver uses
        hipchk(c, launch_sky_to_cart(t, hh, ra, dec, red, x, y, z, n, err, c->stream), "sky_to_cartesian");
        int h = 0;
        hipchk(c, hipMemcpyAsync(&h, err, sizeof(int), hipMemcpyDeviceToHost, c->stream), "D2H");
        sync(c);
        if (h) fail(c, BAOREC_ERR_OUT_OF_BOX, "sky_to_cartesian: value outside the cosmology table "
                                              "(the reference raises BoundsError)");
    });
}

int baorec_cartesian_to_sky(baorec_ctx *c, const baorec_cosmology *cp, int mem, const float *x, const float *y,
                            const float *z, int64_t n, float *ra, float *dec, float *red) {
    return guarded(c, [&] {
        if (n < 0) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "negative count");
        CosmoTable t = cosmo_table(c, cp);
        if (n == 0) return;
        const float *a = stage_in(c, mem, x, n, "sky_a"), *b = stage_in(c, mem, y, n, "sky_b");
        const float *d = stage_in(c, mem, z, n, "sky_c");
        float *o1 = stage_out(c, mem, ra, n, "sky_x"), *o2 = stage_out(c, mem, dec, n, "sky_y");
        float *o3 = stage_out(c, mem, red, n, "sky_z");
        int *e = cosmo_err(c);
        {
            ProfScope ps(c, "cartesian_to_sky", 24.0 * n);
            hipchk(c, launch_cart_to_sky(t, cp->h, a, b, d, o1, o2, o3, n, e, c->stream), "cartesian_to_sky");
        }
        copy_out(c, mem, ra, o1, n);
        copy_out(c, mem, dec, o2, n);
        copy_out(c, mem, red, o3, n);
        cosmo_check(c, e, "cartesian_to_sky");
    });
}

void baorec_pk_params_default(baorec_pk_params *p) {
    if (!p) return;
    memset(p, 0, sizeof(*p));
    for (int a = 0; a < 3; a++) p->box_size[a] = 1000.f;
    p->los[2] = 1.f;
    p->mas = BAOREC_MAS_TSC;
    p->interlace = 1;
    p->kmin = 0.0;
    p->dk = 0.005;
    p->nbins = 300;
}

int baorec_power_multipoles(baorec_ctx *c, const baorec_pk_params *pk, const int32_t grid[3], int mem,
                            const float *x, const float *y, const float *z, const float *w, int64_t n,
                            double *k_mean, double *nmodes, double *p0, double *p2, double *p4) {
    return guarded(c, [&] {
        if (!pk || !k_mean || !nmodes || !p0 || !p2 || !p4) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "NULL argument");
        if (pk->nbins < 1 || pk->nbins > 1600 || !(pk->dk > 0.0) || !(pk->kmin >= 0.0))
            fail(c, BAOREC_ERR_INVALID_ARGUMENT, "need 1 <= nbins <= 1600, dk > 0, kmin >= 0");
        if (pk->mas != BAOREC_MAS_CIC && pk->mas != BAOREC_MAS_TSC)
            fail(c, BAOREC_ERR_INVALID_ARGUMENT, "unknown mass-assignment scheme");
        if (n <= 0) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "the power spectrum needs at least one particle");
        const double ln = sqrt((double)pk->los[0] * pk->los[0] + (double)pk->los[1] * pk->los[1] +
                               (double)pk->los[2] * pk->los[2]);
        if (!(ln > 0.0)) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "los is zero");
        Grid G = check_grid(c, grid, true);
        baorec_params p;
        baorec_params_default(&p, 1.f, 0.f, 0.f);
        for (int a = 0; a < 3; a++) {
            if (!(pk->box_size[a] > 0.f)) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "box_size must be positive");
            p.box_size[a] = pk->box_size[a];
            p.box_min[a] = pk->box_min[a];
        }
        p.mas = pk->mas;
        p.mutate_inputs = 0;
        // positions are only read (no wrap write-back: mutate_inputs = 0)
        Inputs in{};
        in.x = const_cast<float *>(stage_in(c, mem, x, n, "in_x"));
        in.y = const_cast<float *>(stage_in(c, mem, y, n, "in_y"));
        in.z = const_cast<float *>(stage_in(c, mem, z, n, "in_z"));
        const float *wd;
        if (w) {
            wd = stage_in(c, mem, w, n, "in_w");
        } else {
            float *ones = bufT<float>(c, "pk_ones", n);
            hipchk(c, launch_fill(ones, n, 1.f, c->stream), "fill");
            wd = ones;
        }
        const int ng = pk->interlace ? 2 : 1;
        Slot *slot = slot_buf(c) + 1;
        slot_from_params(c, slot, &p, 0.0);
        float *rho = bufT<float>(c, "pk_rho", (size_t)2 * G.R());
        scatter(c, G, &p, slot, rho, in.x, in.y, in.z, wd, n, 1, 0, "pk grid");
        if (ng == 2) {
            // the interlaced grid: box origin moved back by half a cell puts
            // every particle half a cell further along each axis
            baorec_params q = p;
            for (int a = 0; a < 3; a++) {
                const int na = a == 0 ? G.nx : (a == 1 ? G.ny : G.nz);
                q.box_min[a] = p.box_min[a] - 0.5f * (p.box_size[a] / (float)na);
            }
            Slot *slot2 = slot_buf(c) + 2;
            slot_from_params(c, slot2, &q, 0.0);
            scatter(c, G, &q, slot2, rho + G.R(), in.x, in.y, in.z, wd, n, 1, 0, "pk grid 2");
            check_err_flag(c, slot2);
        }
        check_err_flag(c, slot);
        FFTBufs f = fft_bufs(c, G);
        x_fwd(c, G, rho, f.cw, ng);
        strided_plain(c, G, f, 1, false, f.cw, ng);
        strided_plain(c, G, f, 2, false, f.cw, ng);
        // per-axis phase of half a cell and window factor
        const int pw = pk->mas == BAOREC_MAS_TSC ? 3 : 2;
        std::vector<float4> tab((size_t)G.P + G.ny + G.nz);
        const int n3[3] = {G.nx, G.ny, G.nz};
        double kf[3];
        size_t o = 0;
        for (int a = 0; a < 3; a++) {
            kf[a] = 2.0 * M_PI / (double)pk->box_size[a];
            const int cnt = a == 0 ? G.P : n3[a];
            const double hcell = (double)pk->box_size[a] / n3[a];
            for (int i = 0; i < cnt; i++) {
                const int sidx = (a == 0 || i < n3[a] / 2) ? i : i - n3[a];
                const double arg = 0.5 * sidx * kf[a] * hcell;
                const double sc = arg == 0.0 ? 1.0 : sin(arg) / arg;
                tab[o++] = make_float4((float)cos(arg), (float)sin(arg), (float)pow(sc, pw), 0.f);
            }
        }
        float4 *dtab = bufT<float4>(c, "pk_tab", tab.size());
        hipchk(c, hipMemcpyAsync(dtab, tab.data(), sizeof(float4) * tab.size(), hipMemcpyHostToDevice, c->stream),
               "H2D");
        const int nb = pk->nbins;
        PkArgs a{};
        a.F1 = f.cw;
        a.F2 = f.cw + G.C();
        a.tab = dtab;
        a.nx = G.nx;
        a.ny = G.ny;
        a.nz = G.nz;
        a.P = G.PP();
        a.PZ = G.PZ();
        a.Pv = G.P;
        for (int i = 0; i < 3; i++) {
            a.kf[i] = kf[i];
            a.los[i] = pk->los[i] / ln;
        }
        a.kmin = pk->kmin;
        a.dk = pk->dk;
        a.nbins = nb;
        a.interlace = ng == 2;
        a.part = bufT<double>(c, "pk_part", (size_t)pk_blocks() * 5 * nb);
        double *dout = bufT<double>(c, "pk_out", (size_t)5 * nb + 2);
        {
            ProfScope ps(c, "pk_modes", 8.0 * G.Cdense() * ng);
            hipchk(c, launch_pk(a, dout, c->stream), "power multipoles");
        }
        // FP64 totals sum(w), sum(w^2) for the normalisation and shot noise
        hipchk(c, launch_zero_words(dout + 5 * nb, 4, c->stream), "zero");
        hipchk(c, launch_sum<float>(wd, n, dout + 5 * nb, red_part(c), c->stream), "sum w");
        hipchk(c, launch_sumsq<float>(wd, n, dout + 5 * nb + 1, red_part(c), c->stream), "sum w2");
        std::vector<double> h((size_t)5 * nb + 2);
        hipchk(c, hipMemcpyAsync(h.data(), dout, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream), "D2H");
        sync(c);
        const double V = (double)pk->box_size[0] * pk->box_size[1] * pk->box_size[2];
        const double W = h[5 * nb], W2 = h[5 * nb + 1];
        const double norm = V / (W * W), shot = V * W2 / (W * W);
        for (int b = 0; b < nb; b++) {
            const double *q = &h[5 * (size_t)b];
            nmodes[b] = q[0];
            if (q[0] > 0.0) {
                k_mean[b] = q[1] / q[0];
                p0[b] = norm * q[2] / q[0] - shot;
                p2[b] = 5.0 * norm * q[3] / q[0];
                p4[b] = 9.0 * norm * q[4] / q[0];
            } else {
                k_mean[b] = p0[b] = p2[b] = p4[b] = NAN;
            }
        }
    });
}

// ---------------------------------------------------------------- autotune

int baorec_autotune(baorec_ctx *c, const baorec_params *p, const int32_t grid[3], int mem, const float *x,
                    const float *y, const float *z, const float *w, int64_t n) {
    return guarded(c, [&] {
        check_params(c, p);
        check_box(c, p);
        Grid G = check_grid(c, grid, true);
        if (mem != BAOREC_MEM_HOST && mem != BAOREC_MEM_DEVICE) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "mem");
        if (n <= 0) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "the catalog is empty");
        if (c->prof_on) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "profiling is on");
        // the forced choices end with this call, also when it fails
        struct Unforce {
            baorec_ctx *c;
            ~Unforce() {
                c->force_tile[0] = c->force_tile[1] = c->force_tile[2] = 0;
                c->force_pair = -1;
            }
        } unforce{c};
        baorec_params q = *p;
        q.mutate_inputs = 0;   // the timed scatters leave the inputs untouched
        float *dx = const_cast<float *>(stage_in(c, mem, x, n, "tune_x"));
        float *dy = const_cast<float *>(stage_in(c, mem, y, n, "tune_y"));
        float *dz = const_cast<float *>(stage_in(c, mem, z, n, "tune_z"));
        const float *dw = stage_in(c, mem, w, n, "tune_w");
        Slot *slot = slot_buf(c) + 2;
        float *rho = bufT<float>(c, "tune_rho", G.R());
        const int tsc = q.mas == BAOREC_MAS_TSC;
        // scatter tiles: every shape with sides in {16, 32, 64} x {8..64} x
        // {8..64} that fits the LDS tile and the sort's histogram
        std::array<int, 3> best{0, 0, 0};
        double best_ms = 1e30;
        for (int tx : {16, 32, 64})
            for (int ty : {8, 16, 32, 64})
                for (int tz : {8, 16, 32, 64}) {
                    if (tx > G.nx || ty > G.ny || tz > G.nz || tx * ty * tz > scatter_tile_max()) continue;
                    const int64_t T = (int64_t)((G.nx + tx - 1) / tx) * ((G.ny + ty - 1) / ty) * ((G.nz + tz - 1) / tz);
                    if (T > 16384) continue;
                    c->force_tile[0] = tx;
                    c->force_tile[1] = ty;
                    c->force_tile[2] = tz;
                    const double ms = time_min(c, 3, [&] {
                        slot_from_params(c, slot, &q, 0.0);
                        scatter(c, G, &q, slot, rho, dx, dy, dz, dw, n, 1, 0, "tune");
                    });
                    if (ms < best_ms) {
                        best_ms = ms;
                        best = {tx, ty, tz};
                    }
                }
        c->force_tile[0] = c->force_tile[1] = c->force_tile[2] = 0;
        check_err_flag(c, slot);   // a particle outside the grid: BoundsError, nothing stored
        if (best[0]) c->tune_tile[tile_key(G, tsc, n)] = best;
        // strided kernel of single-output passes, per axis of length >= 256
        FFTBufs f = fft_bufs(c, G);
        tables(c, G, slot, f);
        hipchk(c, launch_zero(reinterpret_cast<float *>(f.cw), 2 * G.C(), c->stream), "zero");
        // the single-output passes as the reconstruction runs them: y forward
        // and inverse; z forward * Gaussian * inverse
        for (int axis = 1; axis <= 2; axis++) {
            const int M = axis == 1 ? G.ny : G.nz;
            if (M < 256) continue;
            double t[2];
            for (int pr = 0; pr < 2; pr++) {
                c->force_pair = pr;
                t[pr] = time_min(c, 5, [&] {
                    if (axis == 1) {
                        strided_plain(c, G, f, 1, false, f.cw, 1);
                        strided_plain(c, G, f, 1, true, f.cw, 1);
                    } else {
                        StridedArgs sa = sargs(G, f, 2, 2, MUL_GAUSS);
                        sa.R2 = q.smoothing_radius * q.smoothing_radius;
                        sa.job[0].in = f.cw;
                        sa.job[0].out[0] = f.cw;
                        sa.job[0].nout = 1;
                        sa.job[0].pw[0] = 0;
                        strided(c, G, sa, 1, 1, "tune z");
                    }
                });
            }
            c->force_pair = -1;
            c->tune_pair[{axis, M}] = t[1] < t[0] ? 1 : 0;
        }
        sync(c);
    });
}

int64_t baorec_tune_report(baorec_ctx *c, char *buf, int64_t len) {
    if (!c) return BAOREC_ERR_INVALID_ARGUMENT;
    std::string js = "{\"scatter\": [";
    bool first = true;
    for (const auto &e : c->tune_tile) {
        char line[256];
        snprintf(line, sizeof line,
                 "%s\n  {\"grid\": [%lld, %lld, %lld], \"tsc\": %lld, \"log2n\": %lld, \"tile\": [%d, %d, %d]}",
                 first ? "" : ",", (long long)e.first[0], (long long)e.first[1], (long long)e.first[2],
                 (long long)e.first[3], (long long)e.first[4], e.second[0], e.second[1], e.second[2]);
        js += line;
        first = false;
    }
    js += "],\n \"strided\": [";
    first = true;
    for (const auto &e : c->tune_pair) {
        char line[128];
        snprintf(line, sizeof line, "%s\n  {\"axis\": %d, \"n\": %d, \"pair\": %d}", first ? "" : ",", e.first[0],
                 e.first[1], e.second);
        js += line;
        first = false;
    }
    js += "]}\n";
    if (buf && len > 0) {
        strncpy(buf, js.c_str(), (size_t)len - 1);
        buf[len - 1] = 0;
    }
    return (int64_t)js.size() + 1;
}

int baorec_tune_save(baorec_ctx *c, const char *path) {
    if (!c) return BAOREC_ERR_INVALID_ARGUMENT;
    if (!path) {
        c->err = "path is NULL";
        return BAOREC_ERR_INVALID_ARGUMENT;
    }
    const int64_t len = baorec_tune_report(c, nullptr, 0);
    std::string js((size_t)len, '\0');
    baorec_tune_report(c, &js[0], len);
    js.resize((size_t)len - 1);
    FILE *fp = fopen(path, "w");
    if (!fp) {
        c->err = std::string("cannot create ") + path;
        return BAOREC_ERR_INVALID_ARGUMENT;
    }
    const bool ok = fwrite(js.data(), 1, js.size(), fp) == js.size();
    if (fclose(fp) != 0 || !ok) {
        c->err = std::string("write failed: ") + path;
        return BAOREC_ERR_INVALID_ARGUMENT;
    }
    return BAOREC_OK;
}

// Reads the entries baorec_tune_save writes (one object per line); entries
// with sizes the library cannot use are rejected.
int baorec_tune_load(baorec_ctx *c, const char *path) {
    if (!c) return BAOREC_ERR_INVALID_ARGUMENT;
    FILE *fp = path ? fopen(path, "r") : nullptr;
    if (!fp) {
        c->err = std::string("cannot open ") + (path ? path : "(NULL)");
        return BAOREC_ERR_INVALID_ARGUMENT;
    }
    std::map<std::array<int64_t, 5>, std::array<int, 3>> tiles;
    std::map<std::array<int, 2>, int> pairs;
    char line[512];
    bool bad = false;
    while (fgets(line, sizeof line, fp)) {
        long long g0, g1, g2, tsc, l2;
        int t0, t1, t2, ax, nn, pr;
        const char *s = strchr(line, '{');
        if (!s) continue;
        if (sscanf(s, " {\"grid\": [%lld, %lld, %lld], \"tsc\": %lld, \"log2n\": %lld, \"tile\": [%d, %d, %d]}", &g0,
                   &g1, &g2, &tsc, &l2, &t0, &t1, &t2) == 8) {
            const int64_t T = ((g0 + t0 - 1) / (t0 > 0 ? t0 : 1)) * ((g1 + t1 - 1) / (t1 > 0 ? t1 : 1)) *
                              ((g2 + t2 - 1) / (t2 > 0 ? t2 : 1));
            if (t0 < 1 || t1 < 1 || t2 < 1 || (int64_t)t0 * t1 * t2 > scatter_tile_max() || T > 16384 || t0 > g0 ||
                t1 > g1 || t2 > g2) {
                bad = true;
                break;
            }
            tiles[{g0, g1, g2, tsc, l2}] = {t0, t1, t2};
        } else if (sscanf(s, " {\"axis\": %d, \"n\": %d, \"pair\": %d}", &ax, &nn, &pr) == 3) {
            if ((ax != 1 && ax != 2) || (pr != 0 && pr != 1)) {
                bad = true;
                break;
            }
            pairs[{ax, nn}] = pr;
        }
    }
    fclose(fp);
    if (bad) {
        c->err = std::string(path) + ": entry outside what the library supports";
        return BAOREC_ERR_INVALID_ARGUMENT;
    }
    for (const auto &e : tiles) c->tune_tile[e.first] = e.second;
    for (const auto &e : pairs) c->tune_pair[e.first] = e.second;
    return BAOREC_OK;
}

int baorec_profile(baorec_ctx *c, int enable) {
    if (!c) return BAOREC_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    prof_resolve(c);
    c->prof.clear();
    c->prof_on = enable ? 1 : 0;
    return BAOREC_OK;
}

int64_t baorec_profile_report(baorec_ctx *c, char *buf, int64_t len) {
    if (!c) return BAOREC_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    prof_resolve(c);
    std::string js = "{";
    bool first = true;
    for (auto &kv : c->prof) {
        char tmp[256];
        snprintf(tmp, sizeof tmp, "%s\"%s\": {\"count\": %lld, \"ms\": %.6f, \"bytes\": %.1f}",
                 first ? "" : ", ", kv.first.c_str(), (long long)kv.second.count, kv.second.ms,
                 kv.second.bytes);
        js += tmp;
        first = false;
    }
    js += "}";
    if (buf && len > 0) {
        strncpy(buf, js.c_str(), (size_t)len - 1);
        buf[len - 1] = 0;
    }
    return (int64_t)js.size() + 1;
}

int baorec_run_batch(baorec_ctx *c, baorec_params *p, const int32_t grid[3], int mem, int ncat,
                     const int64_t *offsets, float *x, float *y, float *z, const float *w, float *ox,
                     float *oy, float *oz) {
    return guarded(c, [&] {
        check_params(c, p);
        check_box(c, p);
        Grid G = check_grid(c, grid, true);
        if (mem != BAOREC_MEM_HOST && mem != BAOREC_MEM_DEVICE) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "mem");
        if (ncat < 0 || (ncat > 0 && !offsets)) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "ncat / offsets");
        if (ncat == 0) return;
        if (offsets[0] != 0) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "offsets[0] must be 0");
        for (int k = 0; k < ncat; k++)
            if (offsets[k + 1] < offsets[k]) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "offsets must be non-decreasing");
        const int64_t n = offsets[ncat];
        const int per = p->mas == BAOREC_MAS_TSC ? 27 : 8;
        for (int k = 0; k < ncat; k++)
            if ((offsets[k + 1] - offsets[k]) * per >= ((int64_t)1 << 32))
                fail(c, BAOREC_ERR_UNSUPPORTED, "a catalog of the batch is too large for one launch sequence");
        std::vector<int64_t> t64;
        std::vector<int> t32;
        std::vector<BatchGroup> gs = batch_groups(G, p, ncat, offsets, t64, t32);
        int maxnb = 0;
        for (const BatchGroup &g : gs) maxnb = std::max(maxnb, g.c1 - g.c0);
        int64_t *d64 = bufT<int64_t>(c, "bt_t64", t64.size());
        int *d32 = bufT<int>(c, "bt_t32", t32.size());
        hipchk(c, hipMemcpyAsync(d64, t64.data(), sizeof(int64_t) * t64.size(), hipMemcpyHostToDevice, c->stream),
               "H2D");
        hipchk(c, hipMemcpyAsync(d32, t32.data(), sizeof(int) * t32.size(), hipMemcpyHostToDevice, c->stream),
               "H2D");
        int32_t *err = bufT<int32_t>(c, "bt_err", 1);
        hipchk(c, launch_zero_words(err, 1, c->stream), "zero");
        Inputs in = stage_catalogs(c, mem, x, y, z, w, n, nullptr, nullptr, nullptr, nullptr, 0);
        float *o1 = stage_out(c, mem, ox, n, "bt_ox");
        float *o2 = stage_out(c, mem, oy, n, "bt_oy");
        float *o3 = stage_out(c, mem, oz, n, "bt_oz");
        // the host tables are read by the copies above before they go out of scope
        hipchk(c, hipStreamSynchronize(c->stream), "sync");
        for (const BatchGroup &g : gs)
            batch_group(c, G, p, offsets, g, d64, d32, in.x, in.y, in.z, in.w, o1, o2, o3, err, maxnb);
        int32_t herr = 0;
        hipchk(c, hipMemcpyAsync(&herr, err, sizeof herr, hipMemcpyDeviceToHost, c->stream), "D2H");
        sync(c);
        if (herr) fail(c, BAOREC_ERR_OUT_OF_BOX, "a particle indexes outside the grid (BoundsError)");
        mutate_back(c, mem, p, 1, x, y, z, in, n);
        copy_out(c, mem, ox, o1, n);
        copy_out(c, mem, oy, o2, n);
        copy_out(c, mem, oz, o3, n);
        sync(c);
    });
}

// ------------------------------------------------------- learned warm start

int64_t baorec_warmstart_nparams(void) { return unet_count_params(); }

int baorec_warmstart_init(baorec_ctx *c, uint64_t seed, float h_train) {
    return guarded(c, [&] {
        if (!(h_train > 0.f)) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "h_train must be positive");
        if (!c->net) c->net = unet_create();
        if (!c->net) fail(c, BAOREC_ERR_OUT_OF_MEMORY, "network");
        hipchk(c, unet_init(c->net, seed, c->stream), "network init");
        unet_set_h_train(c->net, h_train);
    });
}

namespace {
struct NetFileHeader {
    char magic[8];   // "BRUNET01"
    int64_t nparams;
    int64_t step;
    float h_train;
    int32_t reserved;
};
}  // namespace

int baorec_warmstart_save(baorec_ctx *c, const char *path) {
    return guarded(c, [&] {
        need_net(c);
        if (!path) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "path is NULL");
        const int64_t n = unet_nparams(c->net);
        std::vector<float> h((size_t)n);
        hipchk(c, unet_get_params(c->net, h.data(), 0, c->stream), "D2H parameters");
        NetFileHeader hd{};
        memcpy(hd.magic, "BRUNET01", 8);
        hd.nparams = n;
        hd.step = unet_step(c->net);
        hd.h_train = unet_h_train(c->net);
        FILE *fp = fopen(path, "wb");
        if (!fp) fail(c, BAOREC_ERR_INVALID_ARGUMENT, std::string("cannot write ") + path);
        const bool ok = fwrite(&hd, sizeof hd, 1, fp) == 1 && fwrite(h.data(), sizeof(float), (size_t)n, fp) == (size_t)n;
        if (fclose(fp) != 0 || !ok) fail(c, BAOREC_ERR_INVALID_ARGUMENT, std::string("write failed: ") + path);
    });
}

int baorec_warmstart_load(baorec_ctx *c, const char *path) {
    return guarded(c, [&] {
        if (!path) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "path is NULL");
        FILE *fp = fopen(path, "rb");
        if (!fp) fail(c, BAOREC_ERR_INVALID_ARGUMENT, std::string("cannot read ") + path);
        NetFileHeader hd{};
        const int64_t n = unet_count_params();
        std::vector<float> h((size_t)n);
        const bool ok = fread(&hd, sizeof hd, 1, fp) == 1 && memcmp(hd.magic, "BRUNET01", 8) == 0 &&
                        hd.nparams == n && fread(h.data(), sizeof(float), (size_t)n, fp) == (size_t)n;
        fclose(fp);
        if (!ok) fail(c, BAOREC_ERR_INVALID_ARGUMENT, std::string("not a warm-start network file of this build: ") + path);
        for (int64_t i = 0; i < n; i++)
            if (!std::isfinite(h[(size_t)i])) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "non-finite parameter in the file");
        if (!c->net) c->net = unet_create();
        if (!c->net) fail(c, BAOREC_ERR_OUT_OF_MEMORY, "network");
        hipchk(c, unet_set_params(c->net, h.data(), c->stream), "H2D parameters");
        unet_set_h_train(c->net, hd.h_train);
    });
}

int baorec_warmstart_params(baorec_ctx *c, int which, float *host, int64_t n) {
    return guarded(c, [&] {
        need_net(c);
        if (!host || n != unet_nparams(c->net)) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "need nparams floats");
        if (which == 1)
            hipchk(c, unet_set_params(c->net, host, c->stream), "H2D parameters");
        else if (which == 0 || which == 2)
            hipchk(c, unet_get_params(c->net, host, which == 2, c->stream), "D2H parameters");
        else
            fail(c, BAOREC_ERR_INVALID_ARGUMENT, "which must be 0 (get), 1 (set) or 2 (gradients)");
    });
}

// forward, loss against t, backward (and an Adam step when lr > 0)
static double warm_loss_backward(baorec_ctx *c, const Grid &G, int nb, const float *t, float lr) {
    const int64_t n = (int64_t)nb * G.R();
    {
        ProfScope ps(c, "unet_forward", 0.0);
        hipchk(c, unet_forward(c->net, c->stream), "network forward");
    }
    const int lb = unet_loss_blocks(n);
    double *part = bufT<double>(c, "warm_loss_part", (size_t)lb);
    double *tot = bufT<double>(c, "warm_loss", 1);
    hipchk(c, launch_unet_loss(unet_output(c->net), t, ilog2i(G.nx), ilog2i(G.ny), ilog2i(G.nz), n,
                               (float)(1.0 / (3.0 * (double)n)), unet_output_grad(c->net), part, c->stream), "loss");
    hipchk(c, launch_sum_partials(part, lb, 1, tot, 1, c->stream), "loss sum");
    {
        ProfScope ps(c, "unet_backward", 0.0);
        hipchk(c, unet_backward(c->net, c->stream), "network backward");
    }
    if (lr > 0.f) hipchk(c, unet_adam(c->net, lr, c->stream), "adam");
    double h = 0.0;
    hipchk(c, hipMemcpyAsync(&h, tot, sizeof(double), hipMemcpyDeviceToHost, c->stream), "D2H");
    sync(c);
    return h;
}

int baorec_warmstart_step(baorec_ctx *c, const baorec_params *p, const int32_t grid[3], int mem, const float *f,
                          const float *phi, int32_t nbatch, float lr, double *loss) {
    return guarded(c, [&] {
        need_net(c);
        check_params(c, p);
        check_box(c, p);
        Grid G = check_grid(c, grid, true);
        if (nbatch < 1 || !f || !phi) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "need nbatch >= 1 grids f and phi");
        if (!(lr >= 0.f)) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "lr must be >= 0");
        const int64_t N = G.R(), n = N * nbatch;
        const float *df = stage_in(c, mem, f, n, "warm_in_f");
        const float *dphi = stage_in(c, mem, phi, n, "warm_in_phi");
        Slot *slot = slot_buf(c) + 1;
        slot_from_params(c, slot, p, 0.0);
        float *phis = bufT<float>(c, "warm_phis_b", n);
        float *rs = bufT<float>(c, "warm_rs_b", n);
        float *psi = bufT<float>(c, "warm_psi_b", n);
        double *ss = bufT<double>(c, "warm_ss_b", nbatch);
        warm_inputs(c, G, p, slot, df, nbatch, phis, rs, psi, ss);
        float *t = bufT<float>(c, "warm_t", n);
        const float ih2 = cells_of(G, p).icell2[0];
        hipchk(c, launch_unet_target(dphi, phis, ss, ih2, N, n, t, c->stream), "target");
        const double l = warm_loss_backward(c, G, nbatch, t, lr);
        if (loss) *loss = l;
    });
}

int baorec_warmstart_eval(baorec_ctx *c, const int32_t grid[3], int mem, const uint16_t *features,
                          const float *target, int32_t nbatch, float *out, double *loss) {
    return guarded(c, [&] {
        need_net(c);
        Grid G = check_grid(c, grid, true);
        if (nbatch < 1 || !features || !target) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "need features and target");
        const int64_t N = G.R(), n = N * nbatch;
        uint16_t *X = nullptr;
        hipchk(c, unet_input(c->net, nbatch, ilog2i(G.nx), ilog2i(G.ny), ilog2i(G.nz), &X), "network buffers");
        hipchk(c, hipMemcpyAsync(X, features, sizeof(uint16_t) * 8 * n,
                                 mem == BAOREC_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                                 c->stream), "features");
        const float *t = stage_in(c, mem, target, n, "warm_in_t");
        const double l = warm_loss_backward(c, G, nbatch, t, 0.f);
        if (loss) *loss = l;
        if (out)
            hipchk(c, hipMemcpyAsync(out, unet_output(c->net), sizeof(float) * n,
                                     mem == BAOREC_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                                     c->stream), "output");
        sync(c);
    });
}

int baorec_warmstart_tensor(baorec_ctx *c, int kind, int index, uint16_t *host, int64_t n) {
    return guarded(c, [&] {
        need_net(c);
        const uint16_t *d = nullptr;
        int64_t cnt = 0;
        if (unet_tensor(c->net, kind, index, &d, &cnt) != hipSuccess)
            fail(c, BAOREC_ERR_INVALID_ARGUMENT, "no such tensor (run baorec_warmstart_eval or _step first)");
        if (!host || n != cnt) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "need " + std::to_string(cnt) + " entries");
        hipchk(c, hipMemcpyAsync(host, d, sizeof(uint16_t) * cnt, hipMemcpyDeviceToHost, c->stream), "D2H tensor");
        sync(c);
    });
}

int baorec_warmstart_predict(baorec_ctx *c, const baorec_params *p, const int32_t grid[3], int mem, const float *f,
                             float *seed) {
    return guarded(c, [&] {
        need_net(c);
        check_params(c, p);
        check_box(c, p);
        Grid G = check_grid(c, grid, true);
        if (!f || !seed) fail(c, BAOREC_ERR_INVALID_ARGUMENT, "f and seed are needed");
        const float *df = stage_in(c, mem, f, G.R(), "warm_in_f");
        float *out = stage_out(c, mem, seed, G.R(), "warm_out");
        Slot *slot = slot_buf(c) + 1;
        slot_from_params(c, slot, p, 0.0);
        learned_seed(c, G, p, slot, cells_of(G, p), df, out);
        copy_out(c, mem, seed, out, G.R());
        sync(c);
    });
}

}  // extern "C"
