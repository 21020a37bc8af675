// Mass assignment for gfx950: CIC (src/mas.jl:1-107) and TSC scatter with
// particles pre-binned by grid tile and accumulated in LDS, the fused
// trilinear displacement gather + RSD shift (src/mas.jl:271-327 and
// src/recon.jl:308-364), read_cic!, and the setup_box reduction
// (src/utils.jl:100-109).
//
// Scatter = four launches over a catalog batch:
//   count   per particle: apply the reference's one-sided wrap (and write it
//           back), locate the 1..8 tiles its corners touch, count them;
//   scan    exclusive prefix sum of the per-tile counts (one workgroup);
//   place   per particle again: append (u_x, u_y, u_z, w) to each touched tile;
//   accum   one workgroup per tile: 64-bit fixed-point integer LDS atomics
//           for every corner inside the tile (exact, order-independent sums),
//           then one rounding to FP32 and one coalesced write of the tile (no
//           HBM atomics).
// u is the reference's FP32 grid coordinate, so the corner indices and
// weights recomputed in accum are bit-identical to the reference's.
#include <math.h>

#include <vector>

#include "kernels.hpp"

namespace baorec {

__device__ __forceinline__ int find_cat(const int64_t *off, int ncat, int64_t i) {
    if (ncat <= 1) return 0;
    int lo = 0, hi = ncat - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Per-axis CIC corners of src/mas.jl:62-84 (0-based).  Returns false when an
// index leaves 0..n-1 (the reference throws a BoundsError).
__device__ __forceinline__ bool cic_axis(float u, int n, int wrap, int &i0, int &i1, float &w0,
                                         float &w1) {
    if (!(u >= -1.0e9f && u <= 1.0e9f)) return false;  // also rejects NaN
    const float fl = floorf(u);
    int x0 = (int)fl;
    w1 = u - fl;
    w0 = 1.0f - w1;
    if (x0 == n + 1) x0 = 1;
    const int x1 = (x0 == n && wrap) ? 1 : x0 + 1;
    if (x0 < 1 || x0 > n || x1 < 1 || x1 > n) return false;
    i0 = x0 - 1;
    i1 = x1 - 1;
    return true;
}

// TSC (extension, no reference counterpart): nodes at integer u - 1 (the same
// nodes as CIC), nearest node c, weights 0.5(0.5-d)^2, 0.75-d^2, 0.5(0.5+d)^2.
__device__ __forceinline__ bool tsc_axis(float u, int n, int wrap, int (&ii)[3], float (&ww)[3]) {
    if (!(u >= -1.0e9f && u <= 1.0e9f)) return false;
    const float g = u - 1.0f;
    const float c = floorf(g + 0.5f);
    const float d = g - c;
    ww[0] = 0.5f * (0.5f - d) * (0.5f - d);
    ww[1] = 0.75f - d * d;
    ww[2] = 0.5f * (0.5f + d) * (0.5f + d);
    const int ci = (int)c;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        int j = ci - 1 + k;
        if (wrap) {
            if (j < 0) j += n;
            else if (j >= n) j -= n;
        }
        if (j < 0 || j >= n) return false;
        ii[k] = j;
    }
    return true;
}

struct PartU {
    float u[3];
    float w;
    int cat;
    bool ok;
};

// Box of one catalog, read once into registers by a workgroup whose
// particles all belong to that catalog (the sort kernels): the wrap's
// write-back stores could alias the slot, so reading it per particle would
// reload it after every store.
struct BoxGeom {
    float box_min[3], box_size[3];
};
__device__ __forceinline__ BoxGeom box_geom(const ScatterArgs &a, int cat) {
    const Slot &sl = a.slot[cat * a.slotStride];
    BoxGeom g;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        g.box_min[k] = sl.box_min[k];
        g.box_size[k] = sl.box_size[k];
    }
    return g;
}

// Grid coordinates of particle i of catalog cat from its loaded position and
// weight (wrapped in place first when apply_wrap; see load_particle).
__device__ __forceinline__ PartU particle_at(const ScatterArgs &a, const BoxGeom &sl, int cat, int64_t i, float x,
                                             float y, float z, float w, bool apply_wrap, bool write_back) {
    PartU p;
    p.cat = cat;
    if (apply_wrap) {  // src/mas.jl:56-60: one-sided, box_min[1]/box_size[1] for all axes
        const float b0 = sl.box_min[0], L0 = sl.box_size[0];
        const bool wx = (x - b0) > L0, wy = (y - b0) > L0, wz = (z - b0) > L0;
        x = wx ? x - L0 : x;
        y = wy ? y - L0 : y;
        z = wz ? z - L0 : z;
        if (write_back) {  // only the coordinates that moved are stored (single-set sorts:
                           // the two-set sort runs without the wrap)
            if (wx) a.x[i] = x;
            if (wy) a.y[i] = y;
            if (wz) a.z[i] = z;
        }
    }
    const float pp[3] = {x, y, z};
    const int nn[3] = {a.nx, a.ny, a.nz};
#pragma unroll
    for (int k = 0; k < 3; k++)  // src/mas.jl:62-64
        p.u[k] = (pp[k] - sl.box_min[k]) * (float)nn[k] / sl.box_size[k] + 1.0f;
    p.w = w;
    p.ok = true;
    return p;
}

__device__ __forceinline__ PartU load_particle(const ScatterArgs &a, int64_t i, bool apply_wrap,
                                               bool write_back) {
    PartU p;
    p.cat = find_cat(a.offsets, a.ncat, i);
    const Slot &sl = a.slot[p.cat * a.slotStride];
    float x = a.x[i], y = a.y[i], z = a.z[i];
    if (apply_wrap) {  // src/mas.jl:56-60: one-sided, box_min[1]/box_size[1] for all axes
        const float b0 = sl.box_min[0], L0 = sl.box_size[0];
        const bool wx = (x - b0) > L0, wy = (y - b0) > L0, wz = (z - b0) > L0;
        x = wx ? x - L0 : x;
        y = wy ? y - L0 : y;
        z = wz ? z - L0 : z;
        if (write_back) {  // only the coordinates that moved are stored
            if (wx) a.x[i] = x;
            if (wy) a.y[i] = y;
            if (wz) a.z[i] = z;
        }
    }
    const float pp[3] = {x, y, z};
    const int nn[3] = {a.nx, a.ny, a.nz};
#pragma unroll
    for (int k = 0; k < 3; k++)  // src/mas.jl:62-64
        p.u[k] = (pp[k] - sl.box_min[k]) * (float)nn[k] / sl.box_size[k] + 1.0f;
    p.w = a.w[i];
    p.ok = true;
    return p;
}

// Tiles touched along one axis: up to 3 distinct tile indices.
__device__ __forceinline__ int axis_tiles(const ScatterArgs &a, float u, int n, int t, int nt,
                                          int (&tl)[3], bool &ok) {
    int cnt = 0;
    int idx[3];
    int m;
    if (a.tsc) {
        float ww[3];
        ok = tsc_axis(u, n, a.wrap, idx, ww) && ok;
        m = 3;
    } else {
        float w0, w1;
        int i0 = 0, i1 = 0;
        ok = cic_axis(u, n, a.wrap, i0, i1, w0, w1) && ok;
        idx[0] = i0;
        idx[1] = i1;
        m = 2;
    }
    if (!ok) return 0;
    for (int k = 0; k < m; k++) {
        int tt = idx[k] / t;
        bool seen = false;
        for (int q = 0; q < cnt; q++) seen |= (tl[q] == tt);
        if (!seen) tl[cnt++] = tt;
    }
    (void)nt;
    return cnt;
}

// Tile of cell index i along an axis: i >> sh for power-of-two tiles.
__device__ __forceinline__ int tile_of(int i, int t, int sh) { return sh >= 0 ? i >> sh : i / t; }

// axis_tiles for a mass assignment fixed at compile time.  CIC: the two
// corners are consecutive cells (or n-1 and 0 with the wrap), so they lie in
// one tile or two.
template <bool TSC>
__device__ __forceinline__ int axis_tiles_m(const ScatterArgs &a, float u, int n, int t, int sh, int (&tl)[3],
                                            bool &ok) {
    if constexpr (TSC) {
        int idx[3];
        float ww[3];
        ok = tsc_axis(u, n, a.wrap, idx, ww) && ok;
        if (!ok) return 0;
        int cnt = 0;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int tt = tile_of(idx[k], t, sh);
            bool seen = false;
            for (int q = 0; q < cnt; q++) seen |= (tl[q] == tt);
            if (!seen) tl[cnt++] = tt;
        }
        return cnt;
    } else {
        float w0, w1;
        int i0 = 0, i1 = 0;
        ok = cic_axis(u, n, a.wrap, i0, i1, w0, w1) && ok;
        if (!ok) return 0;
        tl[0] = tile_of(i0, t, sh);
        tl[1] = tile_of(i1, t, sh);
        return tl[1] != tl[0] ? 2 : 1;
    }
}

int tile_shift(int t) { return t > 0 && (t & (t - 1)) == 0 ? __builtin_ctz(t) : -1; }

void set_tile_shifts(ScatterArgs &a) {
    a.sx = tile_shift(a.tx);
    a.sy = tile_shift(a.ty);
    a.sz = tile_shift(a.tz);
}

__global__ void k_scatter_count(ScatterArgs a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.ntotal) return;
    PartU p = load_particle(a, i, a.wrap != 0, a.wrap && a.mutate);
    bool ok = true;
    int tx[3], ty[3], tz[3];
    const int cx = axis_tiles(a, p.u[0], a.nx, a.tx, a.ntx, tx, ok);
    const int cy = axis_tiles(a, p.u[1], a.ny, a.ty, a.nty, ty, ok);
    const int cz = axis_tiles(a, p.u[2], a.nz, a.tz, a.ntz, tz, ok);
    if (!ok) {
        atomicOr(a.err, 1);
        return;
    }
    if (p.w != 0.f) atomicMax(&a.wmax[p.cat], __float_as_uint(fabsf(p.w)));
    const int64_t tbase = (int64_t)p.cat * a.ntx * a.nty * a.ntz;
    for (int k = 0; k < cz; k++)
        for (int j = 0; j < cy; j++)
            for (int l = 0; l < cx; l++)
                atomicAdd(&a.count[tbase + tx[l] + a.ntx * (ty[j] + a.nty * tz[k])], 1u);
}

__global__ void k_scatter_place(ScatterArgs a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.ntotal) return;
    // positions were already wrapped (and written back) by the count pass
    PartU p = load_particle(a, i, a.wrap && !a.mutate, false);
    bool ok = true;
    int tx[3], ty[3], tz[3];
    const int cx = axis_tiles(a, p.u[0], a.nx, a.tx, a.ntx, tx, ok);
    const int cy = axis_tiles(a, p.u[1], a.ny, a.ty, a.nty, ty, ok);
    const int cz = axis_tiles(a, p.u[2], a.nz, a.tz, a.ntz, tz, ok);
    if (!ok) return;
    const int64_t tbase = (int64_t)p.cat * a.ntx * a.nty * a.ntz;
    const float4 rec = make_float4(p.u[0], p.u[1], p.u[2], p.w);
    for (int k = 0; k < cz; k++)
        for (int j = 0; j < cy; j++)
            for (int l = 0; l < cx; l++) {
                const uint32_t slot =
                    atomicAdd(&a.cursor[tbase + tx[l] + a.ntx * (ty[j] + a.nty * tz[k])], 1u);
                a.sorted[slot] = rec;
            }
}

// Exclusive scan of n counts into cursor (single workgroup of 1024 threads).
__global__ void __launch_bounds__(1024) k_scan(const uint32_t *__restrict__ count, uint32_t *__restrict__ cursor,
                                               int64_t n) {
    __shared__ uint32_t part[16];
    const int tid = threadIdx.x;
    const int64_t per = (n + 1023) / 1024;
    const int64_t lo = tid * per, hi = min(n, lo + per);
    // each thread's range in 16-byte loads when it is aligned (per a multiple
    // of 4: the tile counts of one catalog, 4096 or 16384 tiles)
    const bool vec = (per & 3) == 0 && (reinterpret_cast<uintptr_t>(count) & 15) == 0 &&
                     (reinterpret_cast<uintptr_t>(cursor) & 15) == 0;
    uint32_t s = 0;
    int64_t i0 = lo;
    if (vec) {
        for (; i0 + 3 < hi; i0 += 4) {
            const uint4 c = *reinterpret_cast<const uint4 *>(count + i0);
            s += ((c.x + c.y) + c.z) + c.w;
        }
    }
    for (; i0 < hi; i0++) s += count[i0];
    // exclusive prefix of s over the workgroup: scan within each wave, then
    // over the 16 wave totals
    const int lane = tid & 63, wv = tid >> 6;
    uint32_t v = s;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    if (lane == 63) part[wv] = v;
    __syncthreads();
    if (tid == 0) {
        uint32_t acc = 0;
        for (int k = 0; k < 16; k++) {
            const uint32_t t = part[k];
            part[k] = acc;
            acc += t;
        }
    }
    __syncthreads();
    uint32_t run = part[wv] + v - s;
    i0 = lo;
    if (vec) {
        for (; i0 + 3 < hi; i0 += 4) {
            const uint4 c = *reinterpret_cast<const uint4 *>(count + i0);
            uint4 o;
            o.x = run;
            o.y = o.x + c.x;
            o.z = o.y + c.y;
            o.w = o.z + c.z;
            run = o.w + c.w;
            *reinterpret_cast<uint4 *>(cursor + i0) = o;
        }
    }
    for (; i0 < hi; i0++) {
        cursor[i0] = run;
        run += count[i0];
    }
}

constexpr int kAccumThreads = 512;
constexpr int kTileMax = 32 * 16 * 16;

// Fixed-point contribution: c * 2^e rounded to an integer (exact scaling).
__device__ __forceinline__ unsigned long long fixq(float c, double S) {
    return (unsigned long long)(long long)rint((double)c * S);
}

// FP64 total of the workgroup's values into *dst (fixed order: waves, then
// their sums in order).  Every thread of the workgroup calls it.
__device__ __forceinline__ void tile_sum_store(double v, double *dst) {
    __shared__ double red[kAccumThreads / 64];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < kAccumThreads / 64; i++) s += red[i];
        *dst = s;
    }
}

// One workgroup per (catalog, tile).  Contributions (the reference's FP32
// weight products) are summed exactly in 64-bit fixed point with integer LDS
// atomics (float LDS atomics are several times slower on gfx950, and integer
// sums do not depend on the order); the scale 2^e is a per-tile power of two
// chosen from the tile's placement count and the catalog's max |w| so no cell
// can overflow.  Each cell is written once as the FP32 rounding of its sum.
// accumulate: 0 overwrite every tile, 1 add into rho, 2 rho is zero already
// (empty tiles are skipped).
template <int TSC>
__global__ void __launch_bounds__(kAccumThreads) k_scatter_accum(ScatterArgs a, int accumulate) {
    __shared__ unsigned long long tile[kTileMax];
    const int64_t tid_global = blockIdx.x;  // (catalog, tile)
    const int64_t ntiles = (int64_t)a.ntx * a.nty * a.ntz;
    const int cat = (int)(tid_global / ntiles);
    const int64_t tt = tid_global % ntiles;
    const int tix = (int)(tt % a.ntx), tiy = (int)((tt / a.ntx) % a.nty), tiz = (int)(tt / ((int64_t)a.ntx * a.nty));
    const int x0 = tix * a.tx, y0 = tiy * a.ty, z0 = tiz * a.tz;
    const int tcells = a.tx * a.ty * a.tz;
    const uint32_t end = a.cursor[tid_global];
    const uint32_t cnt = a.count[tid_global];
    const uint32_t beg = end - cnt;
    float *rho = a.rho + (int64_t)cat * a.rstride;
    // whole tile inside the grid, rows of float4: vector stores
    const bool vec = x0 + a.tx <= a.nx && y0 + a.ty <= a.ny && z0 + a.tz <= a.nz && (a.tx & 3) == 0 &&
                     (a.nx & 3) == 0 && (reinterpret_cast<uintptr_t>(rho) & 15) == 0;
    // FP64 sum of the tile's cells (overwrite mode only): the grid's total
    // without another pass over it
    const bool sums = a.part != nullptr && accumulate == 0;
    if (cnt == 0) {   // no particle touches this tile
        if (sums && threadIdx.x == 0) a.part[tid_global] = 0.0;
        if (accumulate != 0) return;
        if (vec) {
            const int tx4 = a.tx >> 2;
            for (int c4 = threadIdx.x; c4 < tcells / 4; c4 += kAccumThreads) {
                const int lx = (c4 % tx4) * 4, ly = (c4 / tx4) % a.ty, lz = c4 / (tx4 * a.ty);
                const int64_t g = (x0 + lx) + (int64_t)a.nx * ((y0 + ly) + (int64_t)a.ny * (z0 + lz));
                *reinterpret_cast<float4 *>(rho + g) = make_float4(0.f, 0.f, 0.f, 0.f);
            }
        } else {
            for (int c = threadIdx.x; c < tcells; c += kAccumThreads) {
                const int lx = c % a.tx, ly = (c / a.tx) % a.ty, lz = c / (a.tx * a.ty);
                if (x0 + lx >= a.nx || y0 + ly >= a.ny || z0 + lz >= a.nz) continue;
                rho[(x0 + lx) + (int64_t)a.nx * ((y0 + ly) + (int64_t)a.ny * (z0 + lz))] = 0.f;
            }
        }
        return;
    }
    const float bound = (float)cnt * __uint_as_float(a.wmax[cat]);
    const int e = bound > 0.f ? 61 - ilogbf(bound) : 0;
    const double S = ldexp(1.0, e);
    constexpr int U = 8;  // placements per thread and step
    auto load = [&](float4 (&rb)[U], uint32_t s0) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint32_t s = s0 + u * kAccumThreads;
            rb[u] = s < end ? a.sorted[s] : make_float4(1.f, 1.f, 1.f, 0.f);
        }
    };
    // the first step's placements are loaded while the tile is cleared, and
    // each next step's while the current one is accumulated
    float4 rbuf[U], rnext[U];
    load(rbuf, beg + threadIdx.x);
    for (int c = threadIdx.x; c < tcells; c += kAccumThreads) tile[c] = 0ull;
    __syncthreads();
    for (uint32_t s0 = beg + threadIdx.x; s0 < end; s0 += U * kAccumThreads) {
        if (s0 + U * kAccumThreads < end) load(rnext, s0 + U * kAccumThreads);
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (s0 + u * kAccumThreads >= end) break;
            const float4 r = rbuf[u];
#define TI(lx, ly, lz) ((lx) + a.tx * ((ly) + a.ty * (lz)))
            if constexpr (TSC == 0) {
                int ix0, ix1, iy0, iy1, iz0, iz1;
                float wx0, wx1, wy0, wy1, wz0, wz1;
                cic_axis(r.x, a.nx, a.wrap, ix0, ix1, wx0, wx1);
                cic_axis(r.y, a.ny, a.wrap, iy0, iy1, wy0, wy1);
                cic_axis(r.z, a.nz, a.wrap, iz0, iz1, wz0, wz1);
                wx0 *= r.w;  // src/mas.jl:86-87
                wx1 *= r.w;
                const int lx0 = ix0 - x0, lx1 = ix1 - x0, ly0 = iy0 - y0, ly1 = iy1 - y0;
                const int lz0 = iz0 - z0, lz1 = iz1 - z0;
                const bool bx0 = (unsigned)lx0 < (unsigned)a.tx, bx1 = (unsigned)lx1 < (unsigned)a.tx;
                const bool by0 = (unsigned)ly0 < (unsigned)a.ty, by1 = (unsigned)ly1 < (unsigned)a.ty;
                const bool bz0 = (unsigned)lz0 < (unsigned)a.tz, bz1 = (unsigned)lz1 < (unsigned)a.tz;
                // corner order and products of src/mas.jl:90-97
                if (bx0 && by0 && bz0) atomicAdd(&tile[TI(lx0, ly0, lz0)], fixq(wx0 * wy0 * wz0, S));
                if (bx1 && by0 && bz0) atomicAdd(&tile[TI(lx1, ly0, lz0)], fixq(wx1 * wy0 * wz0, S));
                if (bx0 && by1 && bz0) atomicAdd(&tile[TI(lx0, ly1, lz0)], fixq(wx0 * wy1 * wz0, S));
                if (bx0 && by0 && bz1) atomicAdd(&tile[TI(lx0, ly0, lz1)], fixq(wx0 * wy0 * wz1, S));
                if (bx1 && by1 && bz0) atomicAdd(&tile[TI(lx1, ly1, lz0)], fixq(wx1 * wy1 * wz0, S));
                if (bx1 && by0 && bz1) atomicAdd(&tile[TI(lx1, ly0, lz1)], fixq(wx1 * wy0 * wz1, S));
                if (bx0 && by1 && bz1) atomicAdd(&tile[TI(lx0, ly1, lz1)], fixq(wx0 * wy1 * wz1, S));
                if (bx1 && by1 && bz1) atomicAdd(&tile[TI(lx1, ly1, lz1)], fixq(wx1 * wy1 * wz1, S));
            } else {
                int ix[3], iy[3], iz[3];
                float wx[3], wy[3], wz[3];
                tsc_axis(r.x, a.nx, a.wrap, ix, wx);
                tsc_axis(r.y, a.ny, a.wrap, iy, wy);
                tsc_axis(r.z, a.nz, a.wrap, iz, wz);
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    const int lz = iz[k] - z0;
                    if ((unsigned)lz >= (unsigned)a.tz) continue;
#pragma unroll
                    for (int j = 0; j < 3; j++) {
                        const int ly = iy[j] - y0;
                        if ((unsigned)ly >= (unsigned)a.ty) continue;
#pragma unroll
                        for (int l = 0; l < 3; l++) {
                            const int lx = ix[l] - x0;
                            if ((unsigned)lx >= (unsigned)a.tx) continue;
                            atomicAdd(&tile[TI(lx, ly, lz)], fixq(((r.w * wx[l]) * wy[j]) * wz[k], S));
                        }
                    }
                }
            }
#undef TI
        }
#pragma unroll
        for (int u = 0; u < U; u++) rbuf[u] = rnext[u];
    }
    __syncthreads();
    auto cell = [&](int c) { return (float)ldexp((double)(long long)tile[c], -e); };
    double psum = 0.0;
    if (vec) {
        const int tx4 = a.tx >> 2;
        for (int c4 = threadIdx.x; c4 < tcells / 4; c4 += kAccumThreads) {
            const int lx = (c4 % tx4) * 4, ly = (c4 / tx4) % a.ty, lz = c4 / (tx4 * a.ty);
            const int c = lx + a.tx * (ly + a.ty * lz);
            const int64_t g = (x0 + lx) + (int64_t)a.nx * ((y0 + ly) + (int64_t)a.ny * (z0 + lz));
            float4 v = make_float4(cell(c), cell(c + 1), cell(c + 2), cell(c + 3));
            float4 *dst = reinterpret_cast<float4 *>(rho + g);
            if (accumulate == 1) {
                const float4 o = *dst;
                v = make_float4(o.x + v.x, o.y + v.y, o.z + v.z, o.w + v.w);
            }
            *dst = v;
            psum += ((double)v.x + (double)v.y) + ((double)v.z + (double)v.w);
        }
    } else {
        for (int c = threadIdx.x; c < tcells; c += kAccumThreads) {
            const int lx = c % a.tx, ly = (c / a.tx) % a.ty, lz = c / (a.tx * a.ty);
            if (x0 + lx >= a.nx || y0 + ly >= a.ny || z0 + lz >= a.nz) continue;  // partial edge tile
            const int64_t g = (x0 + lx) + (int64_t)a.nx * ((y0 + ly) + (int64_t)a.ny * (z0 + lz));
            const float v = cell(c);
            const float o = accumulate == 1 ? rho[g] + v : v;
            rho[g] = o;
            psum += (double)o;
        }
    }
    if (sums) tile_sum_store(psum, a.part + tid_global);
}

// ------------------------------------------------ blocked counting sort
//
// Binning without global atomics.  Particles are cut into blocks (at most
// 512 equal ranges of each catalog); block b builds an LDS histogram of the
// tiles its particles touch and stores it as row H[b][*].  A column scan
// turns H into per-block offsets inside each tile, a scan over tile totals
// gives the tile starts, and the place pass ranks each particle within its
// block with LDS atomics.

constexpr int kSortThreads = 1024;
constexpr int kSortBlockMin = 8192;        // particles per block, at least
constexpr int kSortMaxBlocks = 512;        // two 64 KB-histogram blocks per CU on 256 CUs
constexpr int kSortMaxTiles = 16384;        // LDS histogram capacity (64 KB)
constexpr int kSortUnroll = 4;             // particles in flight per thread

// Particles [beg, end) of block blk of a single catalog of n particles cut
// into nblk equal ranges.
__device__ __forceinline__ void sort_range(int64_t n, int nblk, int blk, int64_t &beg, int64_t &end) {
    beg = n * blk / nblk;
    end = n * (blk + 1) / nblk;
}

// Particle arrays of catalog cat: the second set for catalog 1 of a
// two-set sort (x2 != NULL), indexed from n1.
struct SortSrc {
    const float *x, *y, *z, *w;
    int64_t i0;
};
__device__ __forceinline__ SortSrc sort_src(const ScatterArgs &a, int cat) {
    if (a.x2 && cat == 1) return SortSrc{a.x2, a.y2, a.z2, a.w2, a.n1};
    return SortSrc{a.x, a.y, a.z, a.w, 0};
}

// f(t) for every tile t a particle's assignment touches (tiles from
// axis_tiles_m).  CIC: the base tile, then the rare second tile along each
// axis unrolled (no loops over the per-axis counts).
template <bool TSC, typename F>
__device__ __forceinline__ void each_tile(const ScatterArgs &a, const int (&tx)[3], const int (&ty)[3],
                                          const int (&tz)[3], int cx, int cy, int cz, F &&f) {
    if constexpr (TSC) {
        for (int k = 0; k < cz; k++)
            for (int j = 0; j < cy; j++)
                for (int l = 0; l < cx; l++) f(tx[l] + a.ntx * (ty[j] + a.nty * tz[k]));
    } else {
        const int b = tx[0] + a.ntx * (ty[0] + a.nty * tz[0]);
        const int dx = tx[1] - tx[0], dy = a.ntx * (ty[1] - ty[0]), dz = a.ntx * a.nty * (tz[1] - tz[0]);
        f(b);
        if (cx > 1) f(b + dx);
        if (cy > 1) {
            f(b + dy);
            if (cx > 1) f(b + dy + dx);
        }
        if (cz > 1) {
            f(b + dz);
            if (cx > 1) f(b + dz + dx);
            if (cy > 1) {
                f(b + dz + dy);
                if (cx > 1) f(b + dz + dy + dx);
            }
        }
    }
}

template <bool TSC>
__global__ void __launch_bounds__(kSortThreads) k_sort_hist(ScatterArgs a, uint32_t *H,
                                                            const int *blk_cat,
                                                            const int64_t *blk_beg) {
    __shared__ uint32_t hist[kSortMaxTiles];
    const int blk = blockIdx.x;
    const int cat = blk_cat ? blk_cat[blk] : 0;
    int64_t beg, end;
    if (blk_beg) {
        beg = blk_beg[blk];
        end = blk_beg[blk + 1];
    } else {
        sort_range(a.ntotal, gridDim.x, blk, beg, end);
    }
    const int T = a.ntx * a.nty * a.ntz;
    for (int i = threadIdx.x; i < T; i += kSortThreads) hist[i] = 0u;
    float wm = 0.f;
    const SortSrc src = sort_src(a, cat);
    const BoxGeom geo = box_geom(a, cat);
    __syncthreads();
    // kSortUnroll particles per thread per step: their loads are issued
    // together, ahead of the LDS atomics
    for (int64_t i0 = beg + threadIdx.x; i0 < end; i0 += kSortUnroll * kSortThreads) {
        float px[kSortUnroll], py[kSortUnroll], pz[kSortUnroll], pw[kSortUnroll];
#pragma unroll
        for (int u = 0; u < kSortUnroll; u++) {
            const int64_t i = i0 + (int64_t)u * kSortThreads;
            if (i < end) {
                px[u] = src.x[i - src.i0];
                py[u] = src.y[i - src.i0];
                pz[u] = src.z[i - src.i0];
                pw[u] = src.w[i - src.i0];
            }
        }
#pragma unroll
        for (int u = 0; u < kSortUnroll; u++) {
            const int64_t i = i0 + (int64_t)u * kSortThreads;
            if (i >= end) break;
            PartU p = particle_at(a, geo, cat, i, px[u], py[u], pz[u], pw[u], a.wrap != 0, a.wrap && a.mutate);
            bool ok = true;
            int tx[3], ty[3], tz[3];
            const int cx = axis_tiles_m<TSC>(a, p.u[0], a.nx, a.tx, a.sx, tx, ok);
            const int cy = axis_tiles_m<TSC>(a, p.u[1], a.ny, a.ty, a.sy, ty, ok);
            const int cz = axis_tiles_m<TSC>(a, p.u[2], a.nz, a.tz, a.sz, tz, ok);
            if (!ok) {
                atomicOr(a.err, 1);
                continue;
            }
            wm = fmaxf(wm, fabsf(p.w));
            each_tile<TSC>(a, tx, ty, tz, cx, cy, cz, [&](int t) { atomicAdd(&hist[t], 1u); });
        }
    }
    for (int o = 32; o > 0; o >>= 1) wm = fmaxf(wm, __shfl_xor(wm, o, 64));
    if ((threadIdx.x & 63) == 0 && wm > 0.f) atomicMax(&a.wmax[cat], __float_as_uint(wm));
    __syncthreads();
    uint32_t *row = H + (int64_t)blk * T;
    for (int i = threadIdx.x; i < T; i += kSortThreads) row[i] = hist[i];
}

// Column scan: for every (catalog, tile) an exclusive prefix over the blocks
// of that catalog; totals go to count[].  A workgroup takes 64 columns and
// splits the block range into 16 groups (partial sums in LDS).
constexpr int kScanCols = 64, kScanGroups = 16;

__global__ void __launch_bounds__(kScanCols * kScanGroups) k_sort_colscan(uint32_t *H, uint32_t *count,
                                                                        const int *cat_blk0, int ncat, int T,
                                                                        int nblk) {
    __shared__ uint32_t part[kScanGroups][kScanCols];
    const int col = threadIdx.x % kScanCols, grp = threadIdx.x / kScanCols;
    const int64_t g = (int64_t)blockIdx.x * kScanCols + col;
    const bool ok = g < (int64_t)ncat * T;
    const int cat = ok ? (int)(g / T) : 0, t = ok ? (int)(g % T) : 0;
    const int b0 = cat_blk0 ? cat_blk0[cat] : 0, b1 = cat_blk0 ? cat_blk0[cat + 1] : nblk;
    const int per = (b1 - b0 + kScanGroups - 1) / kScanGroups;
    const int lo = b0 + grp * per, hi = min(b1, lo + per);
    uint32_t s = 0;
    if (ok)
        for (int b = lo; b < hi; b++) s += H[(int64_t)b * T + t];
    part[grp][col] = s;
    __syncthreads();
    if (grp == 0) {
        uint32_t run = 0;
        for (int q = 0; q < kScanGroups; q++) {
            const uint32_t v = part[q][col];
            part[q][col] = run;
            run += v;
        }
        if (ok) count[g] = run;
    }
    __syncthreads();
    if (!ok) return;
    uint32_t run = part[grp][col];
    for (int b = lo; b < hi; b++) {
        uint32_t *h = H + (int64_t)b * T + t;
        const uint32_t c = *h;
        *h = run;
        run += c;
    }
}

template <bool TSC>
__global__ void __launch_bounds__(kSortThreads) k_sort_place(ScatterArgs a, const uint32_t *H,
                                                             const int *blk_cat,
                                                             const int64_t *blk_beg) {
    __shared__ uint32_t cur[kSortMaxTiles];
    const int blk = blockIdx.x;
    const int cat = blk_cat ? blk_cat[blk] : 0;
    int64_t beg, end;
    if (blk_beg) {
        beg = blk_beg[blk];
        end = blk_beg[blk + 1];
    } else {
        sort_range(a.ntotal, gridDim.x, blk, beg, end);
    }
    const int T = a.ntx * a.nty * a.ntz;
    const uint32_t *row = H + (int64_t)blk * T;
    const uint32_t *start = a.cursor + (int64_t)cat * T;
    for (int i = threadIdx.x; i < T; i += kSortThreads) cur[i] = start[i] + row[i];
    const SortSrc src = sort_src(a, cat);
    const BoxGeom geo = box_geom(a, cat);
    __syncthreads();
    for (int64_t i0 = beg + threadIdx.x; i0 < end; i0 += kSortUnroll * kSortThreads) {
        float px[kSortUnroll], py[kSortUnroll], pz[kSortUnroll], pw[kSortUnroll];
#pragma unroll
        for (int u = 0; u < kSortUnroll; u++) {
            const int64_t i = i0 + (int64_t)u * kSortThreads;
            if (i < end) {
                px[u] = src.x[i - src.i0];
                py[u] = src.y[i - src.i0];
                pz[u] = src.z[i - src.i0];
                pw[u] = src.w[i - src.i0];
            }
        }
#pragma unroll
        for (int u = 0; u < kSortUnroll; u++) {
            const int64_t i = i0 + (int64_t)u * kSortThreads;
            if (i >= end) break;
            PartU p = particle_at(a, geo, cat, i, px[u], py[u], pz[u], pw[u], a.wrap && !a.mutate, false);
            bool ok = true;
            int tx[3], ty[3], tz[3];
            const int cx = axis_tiles_m<TSC>(a, p.u[0], a.nx, a.tx, a.sx, tx, ok);
            const int cy = axis_tiles_m<TSC>(a, p.u[1], a.ny, a.ty, a.sy, ty, ok);
            const int cz = axis_tiles_m<TSC>(a, p.u[2], a.nz, a.tz, a.sz, tz, ok);
            if (!ok) continue;
            const float4 rec = make_float4(p.u[0], p.u[1], p.u[2], p.w);
            each_tile<TSC>(a, tx, ty, tz, cx, cy, cz, [&](int t) { a.sorted[atomicAdd(&cur[t], 1u)] = rec; });
        }
    }
}

// cursor[i] = start of tile i, then advanced to its end (what k_scatter_accum reads)
__global__ void k_add_count(uint32_t *cursor, const uint32_t *count, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) cursor[i] += count[i];
}

// Blocks of a single catalog: equal ranges of at least kSortBlockMin
// particles, at most kSortMaxBlocks (one wave of blocks, no ragged tail).
int sort_nblk(int64_t n) {
    const int64_t b = (n + kSortBlockMin - 1) / kSortBlockMin;
    return (int)(b < kSortMaxBlocks ? b : kSortMaxBlocks);
}

// Blocks of several catalogs: each catalog is cut into equal ranges of about
// max(kSortBlockMin, n / (kSortMaxBlocks - ncat)) particles, so that the
// per-catalog rounding up still leaves at most kSortMaxBlocks blocks (a block
// never spans two catalogs).
int sort_blocks(int64_t n, const int64_t *offsets_host, int ncat, std::vector<int> &blk_cat,
                std::vector<int64_t> &blk_beg, std::vector<int> &cat_blk0) {
    blk_cat.clear();
    blk_beg.clear();
    cat_blk0.assign(ncat + 1, 0);
    const int64_t total = offsets_host ? offsets_host[ncat] - offsets_host[0] : n;
    const int64_t slots = kSortMaxBlocks - ncat > 1 ? kSortMaxBlocks - ncat : 1;
    int64_t chunk = (total + slots - 1) / slots;
    if (chunk < kSortBlockMin) chunk = kSortBlockMin;
    for (int c = 0; c < ncat; c++) {
        const int64_t b0 = offsets_host ? offsets_host[c] : 0;
        const int64_t b1 = offsets_host ? offsets_host[c + 1] : n;
        cat_blk0[c] = (int)blk_cat.size();
        const int64_t nb = (b1 - b0 + chunk - 1) / chunk;
        for (int64_t k = 0; k < nb; k++) {
            blk_cat.push_back(c);
            blk_beg.push_back(b0 + (b1 - b0) * k / nb);
        }
    }
    cat_blk0[ncat] = (int)blk_cat.size();
    blk_beg.push_back(offsets_host ? offsets_host[ncat] : n);
    return (int)blk_cat.size();
}

bool sort_fits(const ScatterArgs &a) { return a.ntx * a.nty * a.ntz <= kSortMaxTiles; }

hipError_t launch_sort(ScatterArgs &a, int nblk, uint32_t *H, const int *blk_cat, const int64_t *blk_beg,
                       const int *cat_blk0, hipStream_t s) {
    set_tile_shifts(a);
    const int T = a.ntx * a.nty * a.ntz;
    const int64_t ntiles = (int64_t)T * a.ncat;
    hipError_t e;
    if (nblk > 0) {
        if (a.tsc)
            hipLaunchKernelGGL(k_sort_hist<true>, dim3(nblk), dim3(kSortThreads), 0, s, a, H, blk_cat, blk_beg);
        else
            hipLaunchKernelGGL(k_sort_hist<false>, dim3(nblk), dim3(kSortThreads), 0, s, a, H, blk_cat, blk_beg);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_sort_colscan, dim3((unsigned)((ntiles + kScanCols - 1) / kScanCols)),
                       dim3(kScanCols * kScanGroups), 0, s, H, a.count, cat_blk0, a.ncat, T, nblk);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, a.count, a.cursor, ntiles);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (nblk > 0) {
        if (a.tsc)
            hipLaunchKernelGGL(k_sort_place<true>, dim3(nblk), dim3(kSortThreads), 0, s, a, H, blk_cat, blk_beg);
        else
            hipLaunchKernelGGL(k_sort_place<false>, dim3(nblk), dim3(kSortThreads), 0, s, a, H, blk_cat, blk_beg);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_add_count, dim3((unsigned)((ntiles + 255) / 256)), dim3(256), 0, s, a.cursor, a.count,
                       ntiles);
    return hipGetLastError();
}

__global__ void k_zero4(float4 *a, int64_t n4) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride)
        a[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

__global__ void k_zero_words(uint32_t *a, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) a[i] = 0u;
}

// Zero fills are kernels, not hipMemsetAsync: memset nodes in captured
// graphs have given wrong results on replays.
hipError_t launch_zero_words(void *a, int64_t nwords, hipStream_t s) {
    if (nwords <= 0) return hipSuccess;
    int64_t nb = (nwords + 255) / 256;
    if (nb > 8192) nb = 8192;
    hipLaunchKernelGGL(k_zero_words, dim3((unsigned)nb), dim3(256), 0, s, static_cast<uint32_t *>(a), nwords);
    return hipGetLastError();
}

hipError_t launch_zero(float *a, int64_t n, hipStream_t s) {
    if ((reinterpret_cast<uintptr_t>(a) & 15) || (n & 3)) return launch_zero_words(a, n, s);
    const int64_t n4 = n / 4;
    int64_t nb = (n4 + 255) / 256;
    if (nb > 8192) nb = 8192;
    hipLaunchKernelGGL(k_zero4, dim3((unsigned)nb), dim3(256), 0, s, reinterpret_cast<float4 *>(a), n4);
    return hipGetLastError();
}

int scatter_tile_max() { return kTileMax; }

int64_t scatter_capacity(int64_t n, int tsc) {
    // worst case: every particle touches 8 (CIC) or 27 (TSC) tiles; in
    // practice ~1.2x for 32x16x16 tiles.  The bound used is exact-safe.
    return n * (tsc ? 27 : 8);
}

hipError_t launch_scatter(ScatterArgs &a, uint32_t * /*scan_tmp*/, hipStream_t s) {
    const int64_t ntiles = (int64_t)a.ntx * a.nty * a.ntz * a.ncat;
    hipError_t e = launch_zero_words(a.count, ntiles, s);
    if (e != hipSuccess) return e;
    if (a.ntotal > 0) {
        const unsigned nb = (unsigned)((a.ntotal + 255) / 256);
        hipLaunchKernelGGL(k_scatter_count, dim3(nb), dim3(256), 0, s, a);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, a.count, a.cursor, ntiles);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (a.ntotal > 0) {
        const unsigned nb = (unsigned)((a.ntotal + 255) / 256);
        hipLaunchKernelGGL(k_scatter_place, dim3(nb), dim3(256), 0, s, a);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_scatter_accum(const ScatterArgs &a, int accumulate, hipStream_t s) {
    const int64_t ntiles = (int64_t)a.ntx * a.nty * a.ntz * a.ncat;
    if (a.tsc)
        hipLaunchKernelGGL(k_scatter_accum<1>, dim3((unsigned)ntiles), dim3(kAccumThreads), 0, s, a,
                           accumulate);
    else
        hipLaunchKernelGGL(k_scatter_accum<0>, dim3((unsigned)ntiles), dim3(kAccumThreads), 0, s, a,
                           accumulate);
    return hipGetLastError();
}

// ---------------------------------------------------------------- gather

// 1-based index arithmetic of src/mas.jl:276-292, returned 0-based.
__device__ __forceinline__ bool gather_axis(float p, float bmin, float L, int n, int &d, int &u,
                                            float &wu, float &wd) {
    const float dist = (p - bmin) * (float)n / L;
    if (!(dist >= -1.0e9f && dist <= 1.0e9f)) return false;
    const float fl = floorf(dist);
    const int di = (int)fl;
    wu = dist - fl;
    wd = 1.0f - wu;
    int id = di + 1;
    if (id > n) id -= n;
    int iu = id + 1;
    if (iu > n) iu -= n;
    if (id < 1 || id > n || iu < 1 || iu > n) return false;
    d = id - 1;
    u = iu - 1;
    return true;
}

__device__ __forceinline__ float3 load3(const float *p) { return make_float3(p[0], p[1], p[2]); }

__global__ void k_gather(GatherArgs a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const Slot &sl = a.slot[0];
    const float px = a.x[i], py = a.y[i], pz = a.z[i];
    int dx, ux, dy, uy, dz, uz;
    float wux, wdx, wuy, wdy, wuz, wdz;
    bool ok = gather_axis(px, sl.box_min[0], sl.box_size[0], a.nx, dx, ux, wux, wdx);
    ok = gather_axis(py, sl.box_min[1], sl.box_size[1], a.ny, dy, uy, wuy, wdy) && ok;
    ok = gather_axis(pz, sl.box_min[2], sl.box_size[2], a.nz, dz, uz, wuz, wdz) && ok;
    if (!ok) {
        atomicOr(a.err, 1);
        return;
    }
    const int64_t nx = a.nx, nxy = (int64_t)a.nx * a.ny;
    // 12-byte cells: one dwordx3 load per corner
#define G(ix, iy, iz) load3(a.disp + 3 * ((ix) + nx * (iy) + nxy * (iz)))
    // term order of src/mas.jl:295-302
    const float3 f0 = G(dx, dy, dz), f1 = G(dx, dy, uz), f2 = G(dx, uy, dz), f3 = G(dx, uy, uz);
    const float3 f4 = G(ux, dy, dz), f5 = G(ux, dy, uz), f6 = G(ux, uy, dz), f7 = G(ux, uy, uz);
#undef G
    float s[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float g0 = c == 0 ? f0.x : c == 1 ? f0.y : f0.z;
        const float g1 = c == 0 ? f1.x : c == 1 ? f1.y : f1.z;
        const float g2 = c == 0 ? f2.x : c == 1 ? f2.y : f2.z;
        const float g3 = c == 0 ? f3.x : c == 1 ? f3.y : f3.z;
        const float g4 = c == 0 ? f4.x : c == 1 ? f4.y : f4.z;
        const float g5 = c == 0 ? f5.x : c == 1 ? f5.y : f5.z;
        const float g6 = c == 0 ? f6.x : c == 1 ? f6.y : f6.z;
        const float g7 = c == 0 ? f7.x : c == 1 ? f7.y : f7.z;
        s[c] = g0 * wdx * wdy * wdz + g1 * wdx * wdy * wuz + g2 * wdx * wuy * wdz +
               g3 * wdx * wuy * wuz + g4 * wux * wdy * wdz + g5 * wux * wdy * wuz +
               g6 * wux * wuy * wdz + g7 * wux * wuy * wuz;
    }
    float out[3] = {s[0], s[1], s[2]};
    if (a.field != 0) {
        float lx, ly, lz;
        if (a.has_los) {  // src/recon.jl:322-330
            lx = a.los[0];
            ly = a.los[1];
            lz = a.los[2];
        } else {          // src/recon.jl:308-320
            const float dist = sqrtf((px * px + py * py) + pz * pz);
            lx = px / dist;
            ly = py / dist;
            lz = pz / dist;
        }
        const float dot = ((s[0] * lx) + (s[1] * ly)) + (s[2] * lz);
        const float fd = a.f * dot;
        const float r[3] = {fd * lx, fd * ly, fd * lz};
        if (a.field == 1) {
            out[0] = r[0]; out[1] = r[1]; out[2] = r[2];
        } else {  // :sum, src/recon.jl:357-361
            out[0] = s[0] + r[0]; out[1] = s[1] + r[1]; out[2] = s[2] + r[2];
        }
    }
    if (a.positions) {  // src/recon.jl:376-378
        out[0] = px - out[0];
        out[1] = py - out[1];
        out[2] = pz - out[2];
    }
    a.ox[i] = out[0];
    a.oy[i] = out[1];
    a.oz[i] = out[2];
}

hipError_t launch_gather(const GatherArgs &a, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

__global__ void k_read_cic(float *out, const float *field, const float *x, const float *y,
                           const float *z, int64_t n, const Slot *slot, int nx, int ny, int nz,
                           int32_t *err) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Slot &sl = slot[0];
    int dx, ux, dy, uy, dz, uz;
    float wux, wdx, wuy, wdy, wuz, wdz;
    bool ok = gather_axis(x[i], sl.box_min[0], sl.box_size[0], nx, dx, ux, wux, wdx);
    ok = gather_axis(y[i], sl.box_min[1], sl.box_size[1], ny, dy, uy, wuy, wdy) && ok;
    ok = gather_axis(z[i], sl.box_min[2], sl.box_size[2], nz, dz, uz, wuz, wdz) && ok;
    if (!ok) {
        atomicOr(err, 1);
        return;
    }
    const int64_t sx = nx, sxy = (int64_t)nx * ny;
#define F(ix, iy, iz) field[(ix) + sx * (iy) + sxy * (iz)]
    out[i] = F(dx, dy, dz) * wdx * wdy * wdz + F(dx, dy, uz) * wdx * wdy * wuz +
             F(dx, uy, dz) * wdx * wuy * wdz + F(dx, uy, uz) * wdx * wuy * wuz +
             F(ux, dy, dz) * wux * wdy * wdz + F(ux, dy, uz) * wux * wdy * wuz +
             F(ux, uy, dz) * wux * wuy * wdz + F(ux, uy, uz) * wux * wuy * wuz;
#undef F
}

hipError_t launch_read_cic(float *out, const float *field, const float *x, const float *y,
                           const float *z, int64_t n, const Slot *slot, int nx, int ny, int nz,
                           int32_t *err, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_read_cic, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, out, field, x,
                       y, z, n, slot, nx, ny, nz, err);
    return hipGetLastError();
}

// ---------------------------------------------------------------- setup_box

__device__ __forceinline__ uint32_t fkey(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float fdekey(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

__global__ void k_box_init(Slot *slot) {
    uint32_t *mm = reinterpret_cast<uint32_t *>(slot->minmax);
    if (threadIdx.x < 3) mm[threadIdx.x] = 0xffffffffu;
    else if (threadIdx.x < 6) mm[threadIdx.x] = 0u;
}

__global__ void __launch_bounds__(256) k_box_reduce(const float *x, const float *y, const float *z,
                                                    int64_t n, Slot *slot) {
    __shared__ uint32_t red[4][6];
    uint32_t mn[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, mx[3] = {0u, 0u, 0u};
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t k[3] = {fkey(x[i]), fkey(y[i]), fkey(z[i])};
#pragma unroll
        for (int a = 0; a < 3; a++) {
            mn[a] = min(mn[a], k[a]);
            mx[a] = max(mx[a], k[a]);
        }
    }
#pragma unroll
    for (int a = 0; a < 3; a++) {
        for (int o = 32; o > 0; o >>= 1) {
            mn[a] = min(mn[a], (uint32_t)__shfl_xor((int)mn[a], o, 64));
            mx[a] = max(mx[a], (uint32_t)__shfl_xor((int)mx[a], o, 64));
        }
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            red[w][a] = mn[a];
            red[w][3 + a] = mx[a];
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int a = threadIdx.x;
        uint32_t v = red[0][a];
        for (int q = 1; q < 4; q++) v = a < 3 ? min(v, red[q][a]) : max(v, red[q][a]);
        uint32_t *mm = reinterpret_cast<uint32_t *>(slot->minmax);
        if (a < 3) atomicMin(&mm[a], v);
        else atomicMax(&mm[a], v);
    }
}

// src/utils.jl:100-109 in FP32: min - pad/2, max + pad/2, cubic side.
__global__ void k_box_final(Slot *slot, float pad) {
    if (threadIdx.x != 0) return;
    const uint32_t *mm = reinterpret_cast<const uint32_t *>(slot->minmax);
    float bmin[3], bmax[3];
    const float h = pad / 2.0f;
    for (int a = 0; a < 3; a++) {
        bmin[a] = fdekey(mm[a]) - h;
        bmax[a] = fdekey(mm[3 + a]) + h;
    }
    float L = bmax[0] - bmin[0];
    for (int a = 1; a < 3; a++) L = fmaxf(L, bmax[a] - bmin[a]);
    for (int a = 0; a < 3; a++) {
        slot->box_min[a] = bmin[a];
        slot->box_size[a] = L;
    }
}

hipError_t launch_setup_box(const float *x, const float *y, const float *z, int64_t n, Slot *slot,
                            float pad, hipStream_t s) {
    hipLaunchKernelGGL(k_box_init, dim3(1), dim3(64), 0, s, slot);
    int64_t nb = (n + 255) / 256;
    if (nb > 1024) nb = 1024;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(k_box_reduce, dim3((unsigned)nb), dim3(256), 0, s, x, y, z, n, slot);
    hipLaunchKernelGGL(k_box_final, dim3(1), dim3(64), 0, s, slot, pad);
    return hipGetLastError();
}

}  // namespace baorec
