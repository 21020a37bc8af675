// Multigrid kernels for gfx950: the anisotropic 19-point operator of
// src/multigrid.jl (damped Jacobi, residual, 27-point restriction, trilinear
// prolongation) and the multicolour Gauss-Seidel smoother of the RBGS
// schedule.  The FP32 kernels keep the reference's operation order; the
// templated ones also run FP64 coarse levels.
#include <math.h>

#include <type_traits>

#include "kernels.hpp"

namespace baorec {

template <typename T>
struct Coef {
    T px, py, pz, g, gpx2, gpy2, gpz2;
};

// Reciprocal for the Gauss-Seidel smoother's FP32 arithmetic: v_rcp_f32
// (1 ulp) instead of an IEEE division.  The smoother only approximates the
// solve inside the FP64-refined iteration, so its rounding does not reach
// the converged potential; the reference-ordered kernels (Jacobi, residual)
// keep IEEE divisions.
template <typename T, bool FAST>
__device__ __forceinline__ T mg_div(T a, T b) {
    if constexpr (FAST && std::is_same<T, float>::value) return a * __builtin_amdgcn_rcpf(b);
    return a / b;
}

// src/multigrid.jl:161-168 (per-cell coefficients) from the LOS components
// p = (px, py, pz) in cell units
template <typename T, bool FAST = false>
__device__ __forceinline__ Coef<T> coef_p(const MGLevel &L, T px, T py, T pz, T beta) {
    Coef<T> c;
    c.px = px;
    c.py = py;
    c.pz = pz;
    c.g = mg_div<T, FAST>(beta, (T)L.cell2[0] * (c.px * c.px) + (T)L.cell2[1] * (c.py * c.py) +
                                    (T)L.cell2[2] * (c.pz * c.pz));
    c.gpx2 = (T)L.icell2[0] + c.g * (c.px * c.px);
    c.gpy2 = (T)L.icell2[1] + c.g * (c.py * c.py);
    c.gpz2 = (T)L.icell2[2] + c.g * (c.pz * c.pz);
    return c;
}

// p is the fixed LOS, or the cell centre / cell size (local LOS)
template <typename T, bool FAST = false>
__device__ __forceinline__ Coef<T> coef(const MGLevel &L, int ix, int iy, int iz, T beta, int has_los) {
    if (has_los) return coef_p<T, FAST>(L, (T)L.losn[0], (T)L.losn[1], (T)L.losn[2], beta);
    if constexpr (std::is_same<T, float>::value)   // xv / cell, precomputed per level
        return coef_p<T, FAST>(L, L.pv[ix], L.pv[L.n[0] + iy], L.pv[L.n[0] + L.n[1] + iz], beta);
    return coef_p<T, FAST>(L, (T)L.xv[ix] / (T)L.cell[0], (T)L.xv[L.n[0] + iy] / (T)L.cell[1],
                           (T)L.xv[L.n[0] + L.n[1] + iz] / (T)L.cell[2], beta);
}

template <typename T>
struct Nbr {
    T A, B, C, X, G;
};

// Neighbour sums of src/multigrid.jl:170-185 in the reference's order, from
// the 18 neighbour values (p/m: +1/-1 along the two axes named).
template <typename T>
__device__ __forceinline__ Nbr<T> nbr_of(const Coef<T> &c, T vxp, T vxm, T vyp, T vym, T vzp, T vzm, T xpyp,
                                         T xmym, T xmyp, T xpym, T xpzp, T xmzm, T xmzp, T xpzm, T ypzp, T ymzm,
                                         T ymzp, T ypzm) {
    Nbr<T> r;
    r.A = c.gpx2 * (vxp + vxm);
    r.B = c.gpy2 * (vyp + vym);
    r.C = c.gpz2 * (vzp + vzm);
    r.X = (c.g / (T)2) * ((c.px * c.py) * (xpyp + xmym - xmyp - xpym) + (c.px * c.pz) * (xpzp + xmzm - xmzp - xpzm) +
                          (c.py * c.pz) * (ypzp + ymzm - ymzp - ypzm));
    r.G = c.g * (c.px * (vxp - vxm) + c.py * (vyp - vym) + c.pz * (vzp - vzm));
    return r;
}

template <typename T>
__device__ __forceinline__ Nbr<T> nbr(const T *__restrict__ v, const MGLevel &L, int ix, int iy, int iz,
                                      const Coef<T> &c) {
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    const int ixp = ix + 1 == nx ? 0 : ix + 1, ixm = ix == 0 ? nx - 1 : ix - 1;
    const int iyp = iy + 1 == ny ? 0 : iy + 1, iym = iy == 0 ? ny - 1 : iy - 1;
    const int izp = iz + 1 == nz ? 0 : iz + 1, izm = iz == 0 ? nz - 1 : iz - 1;
    const int64_t sy = nx, sz = (int64_t)nx * ny;
#define V(a, b, d) v[(a) + sy * (b) + sz * (d)]
    return nbr_of<T>(c, V(ixp, iy, iz), V(ixm, iy, iz), V(ix, iyp, iz), V(ix, iym, iz), V(ix, iy, izp),
                     V(ix, iy, izm), V(ixp, iyp, iz), V(ixm, iym, iz), V(ixm, iyp, iz), V(ixp, iym, iz),
                     V(ixp, iy, izp), V(ixm, iy, izm), V(ixm, iy, izp), V(ixp, iy, izm), V(ix, iyp, izp),
                     V(ix, iym, izm), V(ix, iym, izp), V(ix, iyp, izm));
#undef V
}

// Per-cell operations, shared by the grid-wide kernels below and the
// single-workgroup coarse-level kernel (same arithmetic in both).

// One damped Jacobi update (src/multigrid.jl:137-191 and the broadcast at
// :207): (1 - w) v + w jac(v) from the cell's f, neighbour sums and value.
__device__ __forceinline__ float jacobi_value(const Coef<float> &c, float f, const Nbr<float> &s, float v,
                                             float omega, int has_los) {
    float j = f + s.A + s.B + s.C + s.X;
    if (!has_los) j += s.G;
    j /= 2.0f * (c.gpx2 + c.gpy2 + c.gpz2);
    return (1.0f - omega) * v + omega * j;
}

// out of place: vout = (1 - w) vin + w jac(vin)
__device__ __forceinline__ void cell_jacobi(float *__restrict__ vout, const float *__restrict__ vin,
                                            const float *__restrict__ f, const MGLevel &L, int ix, int iy,
                                            int iz, float beta, float omega, int has_los) {
    const Coef<float> c = coef<float>(L, ix, iy, iz, beta, has_los);
    const Nbr<float> s = nbr<float>(vin, L, ix, iy, iz, c);
    const int64_t I = ix + (int64_t)L.n[0] * (iy + (int64_t)L.n[1] * iz);
    vout[I] = jacobi_value(c, f[I], s, vin[I], omega, has_los);
}

// r = f - A v (src/multigrid.jl:328-376 and the broadcast at :393).
template <typename T>
__device__ __forceinline__ void cell_residual(T *__restrict__ r, const T *__restrict__ v, const T *__restrict__ f,
                                              const MGLevel &L, int ix, int iy, int iz, float beta,
                                              int has_los) {
    const Coef<T> c = coef<T>(L, ix, iy, iz, (T)beta, has_los);
    const Nbr<T> s = nbr<T>(v, L, ix, iy, iz, c);
    const int64_t I = ix + (int64_t)L.n[0] * (iy + (int64_t)L.n[1] * iz);
    T t = (T)2 * (c.gpx2 + c.gpy2 + c.gpz2) * v[I] - (s.A + s.B + s.C + s.X);
    if (!has_los) t -= s.G;
    r[I] = f[I] - t;
}

// r = (f - mean f) - A v for the refined RBGS solve: v in FP64 and every
// operation in FP64, so that the residual of a converged v is not lost to
// FP32 cancellation in A v; stored as FP32 (the right-hand side of the next
// correction V-cycle).  fsum: device FP64 sum of f (NULL: no shift).
__device__ __forceinline__ void cell_residual_mixed(float *__restrict__ r, const double *__restrict__ v,
                                                    const float *__restrict__ f, double fshift,
                                                    const MGLevel &L, int ix, int iy, int iz, float beta,
                                                    int has_los) {
    const Coef<double> c = coef<double>(L, ix, iy, iz, (double)beta, has_los);
    const Nbr<double> s = nbr<double>(v, L, ix, iy, iz, c);
    const int64_t I = ix + (int64_t)L.n[0] * (iy + (int64_t)L.n[1] * iz);
    double t = 2.0 * (c.gpx2 + c.gpy2 + c.gpz2) * v[I] - (s.A + s.B + s.C + s.X);
    if (!has_los) t -= s.G;
    r[I] = (float)(((double)f[I] - fshift) - t);
}

// Multicolour Gauss-Seidel (RBGS schedule): colour = ((ix+iy)&1) + 2((iy+iz)&1)
// separates every face and edge neighbour of the 19-point stencil, so one
// colour updates in place without races.
__device__ __forceinline__ int mg_color(int ix, int iy, int iz) { return ((ix + iy) & 1) + 2 * ((iy + iz) & 1); }

// The Gauss-Seidel value of one cell (the RBGS schedule's smoother, an
// extension: the reference smooths with jacobi!, src/multigrid.jl:193-210):
// the operator of src/multigrid.jl:161-185 solved for the cell, (f +
// neighbour sums) / diagonal, evaluated with fused multiply-adds and the
// smoother's reciprocal.  The smoother only approximates the solve; the
// FP64-refined residual decides convergence.  Per axis the coefficients take
// p (the LOS component in cell units), p2 = p * p and q = cell^2 * p2.
__device__ __forceinline__ float mg_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double mg_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename T>
struct GsCoef {
    T px, py, pz, g, gx, gy, gz;
};

template <typename T>
__device__ __forceinline__ GsCoef<T> gs_coef(const MGLevel &L, T px, T py, T pz, T px2, T py2, T pz2, T qx, T qy,
                                             T qz, T beta) {
    GsCoef<T> c;
    c.px = px;
    c.py = py;
    c.pz = pz;
    c.g = mg_div<T, true>(beta, (qx + qy) + qz);
    c.gx = mg_fma(c.g, px2, (T)L.icell2[0]);
    c.gy = mg_fma(c.g, py2, (T)L.icell2[1]);
    c.gz = mg_fma(c.g, pz2, (T)L.icell2[2]);
    return c;
}

// the same from p alone (fixed LOS, or p of the cell)
template <typename T>
__device__ __forceinline__ GsCoef<T> gs_coef_p(const MGLevel &L, T px, T py, T pz, T beta) {
    const T px2 = px * px, py2 = py * py, pz2 = pz * pz;
    return gs_coef<T>(L, px, py, pz, px2, py2, pz2, (T)L.cell2[0] * px2, (T)L.cell2[1] * py2, (T)L.cell2[2] * pz2,
                      beta);
}

// the new value from f and the 18 neighbour values (argument order of nbr_of)
template <typename T>
__device__ __forceinline__ T gs_update(const GsCoef<T> &c, T f, T vxp, T vxm, T vyp, T vym, T vzp, T vzm, T xpyp,
                                       T xmym, T xmyp, T xpym, T xpzp, T xmzm, T xmzp, T xpzm, T ypzp, T ymzm, T ymzp,
                                       T ypzm, int has_los) {
    T j = mg_fma(c.gz, vzp + vzm, mg_fma(c.gy, vyp + vym, mg_fma(c.gx, vxp + vxm, f)));
    const T exy = (xpyp + xmym) - (xmyp + xpym), exz = (xpzp + xmzm) - (xmzp + xpzm),
            eyz = (ypzp + ymzm) - (ymzp + ypzm);
    const T x = mg_fma(c.py * c.pz, eyz, mg_fma(c.px * c.pz, exz, (c.px * c.py) * exy));
    j = mg_fma((T)0.5 * c.g, x, j);
    if (!has_los) j = mg_fma(c.g, mg_fma(c.pz, vzp - vzm, mg_fma(c.py, vyp - vym, c.px * (vxp - vxm))), j);
    return mg_div<T, true>(j, (T)2 * ((c.gx + c.gy) + c.gz));
}

// the coefficients of cell (ix, iy, iz)
template <typename T>
__device__ __forceinline__ GsCoef<T> gs_coef_at(const MGLevel &L, int ix, int iy, int iz, float beta, int has_los) {
    T px, py, pz;
    if (has_los) {
        px = (T)L.losn[0];
        py = (T)L.losn[1];
        pz = (T)L.losn[2];
    } else if constexpr (std::is_same<T, float>::value) {
        px = L.pv[ix];
        py = L.pv[L.n[0] + iy];
        pz = L.pv[L.n[0] + L.n[1] + iz];
    } else {
        px = (T)L.xv[ix] / (T)L.cell[0];
        py = (T)L.xv[L.n[0] + iy] / (T)L.cell[1];
        pz = (T)L.xv[L.n[0] + L.n[1] + iz] / (T)L.cell[2];
    }
    return gs_coef_p<T>(L, px, py, pz, (T)beta);
}

template <typename T>
__device__ __forceinline__ void cell_mcgs(T *v, const T *__restrict__ f, const MGLevel &L, int ix, int iy, int iz,
                                          float beta, int has_los) {
    const GsCoef<T> c = gs_coef_at<T>(L, ix, iy, iz, beta, has_los);
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    const int ixp = ix + 1 == nx ? 0 : ix + 1, ixm = ix == 0 ? nx - 1 : ix - 1;
    const int iyp = iy + 1 == ny ? 0 : iy + 1, iym = iy == 0 ? ny - 1 : iy - 1;
    const int izp = iz + 1 == nz ? 0 : iz + 1, izm = iz == 0 ? nz - 1 : iz - 1;
    const int64_t sy = nx, sz = (int64_t)nx * ny;
#define V(a, b, d) v[(a) + sy * (b) + sz * (d)]
    const T nv = gs_update<T>(c, f[ix + sy * iy + sz * iz], V(ixp, iy, iz), V(ixm, iy, iz), V(ix, iyp, iz),
                              V(ix, iym, iz), V(ix, iy, izp), V(ix, iy, izm), V(ixp, iyp, iz), V(ixm, iym, iz),
                              V(ixm, iyp, iz), V(ixp, iym, iz), V(ixp, iy, izp), V(ixm, iy, izm), V(ixm, iy, izp),
                              V(ixp, iy, izm), V(ix, iyp, izp), V(ix, iym, izm), V(ix, iym, izp), V(ix, iyp, izm),
                              has_los);
#undef V
    v[ix + sy * iy + sz * iz] = nv;
}

__global__ void k_mg_jacobi(float *__restrict__ vout, const float *__restrict__ vin,
                            const float *__restrict__ f, MGLevel L, float beta, float omega,
                            int has_los) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x, iy = blockIdx.y * blockDim.y + threadIdx.y;
    if (ix >= L.n[0] || iy >= L.n[1]) return;
    cell_jacobi(vout, vin, f, L, ix, iy, blockIdx.z, beta, omega, has_los);
}

template <typename T>
__global__ void k_mg_residual(T *__restrict__ r, const T *__restrict__ v, const T *__restrict__ f,
                              MGLevel L, float beta, int has_los) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x, iy = blockIdx.y * blockDim.y + threadIdx.y;
    if (ix >= L.n[0] || iy >= L.n[1]) return;
    cell_residual<T>(r, v, f, L, ix, iy, blockIdx.z, beta, has_los);
}

__global__ void k_mg_residual_mixed(float *__restrict__ r, const double *__restrict__ v,
                                    const float *__restrict__ f, const double *fsum, double count, MGLevel L,
                                    float beta, int has_los) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x, iy = blockIdx.y * blockDim.y + threadIdx.y;
    if (ix >= L.n[0] || iy >= L.n[1]) return;
    const double fshift = fsum ? *fsum / count : 0.0;
    cell_residual_mixed(r, v, f, fshift, L, ix, iy, blockIdx.z, beta, has_los);
}

// phi += e - mean(e): the correction of a V-cycle added to the FP64 solution,
// keeping it zero-mean (the operator annihilates constants).
template <typename T>
__global__ void k_add_correction(T *__restrict__ phi, const float *__restrict__ e, int64_t n, const double *esum) {
    const double m = *esum / (double)n;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        phi[i] = (T)((double)phi[i] + ((double)e[i] - m));
}

template <typename T>
__global__ void k_mg_mcgs(T *__restrict__ v, const T *__restrict__ f, MGLevel L, float beta,
                          int has_los, int color) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x, iy = blockIdx.y * blockDim.y + threadIdx.y;
    const int iz = blockIdx.z;
    if (ix >= L.n[0] || iy >= L.n[1] || mg_color(ix, iy, iz) != color) return;
    cell_mcgs<T>(v, f, L, ix, iy, iz, beta, has_los);
}

// One whole 4-colour Gauss-Seidel sweep in one launch, out of place: vout
// holds what the four in-place colour passes of k_mg_mcgs leave in v (the
// same cell arithmetic on the same values, so the same bits).
//
// A workgroup owns a 64 x 16 column of tz planes and marches up z with a
// ring of kGsRing planes of old values in LDS, each 72 x 24 (a halo of 4).
// The colour-c value of a cell depends on colour < c values of its 18
// neighbours (new) and colour > c values (old), so the colour-3 cells of the
// tile need colour 2 on the tile grown by 1, colour 1 grown by 2, colour 0
// grown by 3, and old values grown by 4 (every colour is computed on the halo
// cells too, redundantly with the neighbouring workgroups).  Step k loads
// plane P = z0 - 4 + k and updates colour s on plane P - 1 - 2s (s = 0..3):
// the four stages of a step read and write disjoint (plane, colour) sets, so
// one barrier per step orders everything, and the tile of plane P - 8
// (finished in the step before) is written out in coalesced rows.
//
// Latency: the next plane's float4 loads and the next step's f values are
// in flight during a step's arithmetic; the coefficients' cell positions
// come from LDS tables.  LDS rows keep even x then odd x (36 + 36), so the
// colour cells of a row, every second x, and their x neighbours are
// consecutive words.
constexpr int kGsTX = 64, kGsTY = 16, kGsHalo = 4;
constexpr int kGsLW = kGsTX + 2 * kGsHalo;     // 72 columns per plane row
// LDS rows: the even-x half (36 words + 4 spare) at 0, the odd-x half at
// kGsH = 40, a pitch of 82 words.  A half-wave of the cell list is one 32-cell
// row (32 consecutive words), and a colour's consecutive rows, 164 words
// apart, fall 4 banks further on, so the narrow columns beyond 32 spread over
// the banks too.
constexpr int kGsW = 82, kGsH = 40, kGsXWords = kGsH + kGsLW / 2;   // column words in use: 76
constexpr int kGsRows = kGsTY + 2 * kGsHalo;   // 24 rows per plane
constexpr int kGsPlane = kGsW * kGsRows;
constexpr int kGsRing = 10;                    // planes P-9 .. P
constexpr int kGsThreads = 512;   // 16 waves per CU (two workgroups)
constexpr int kGsMaxTz = 128;
constexpr int kGsChunks = (kGsLW / 4) * kGsRows;   // float4 loads per plane (432)
// colour cells per stage: (8 + e) rows x (32 + e) columns, e = 3 - s.  The
// cell list takes columns 0..31 of stages 0..3 row by row (a half-wave per
// row), then the e extra columns of stages 0..2.
constexpr int kGsA1 = 11 * 32, kGsA2 = kGsA1 + 10 * 32, kGsA3 = kGsA2 + 9 * 32, kGsB0 = kGsA3 + 8 * 32,
              kGsB1 = kGsB0 + 11 * 3, kGsB2 = kGsB1 + 10 * 2, kGsCells = kGsB2 + 9;   // 1278
constexpr int kGsIters = (kGsCells + kGsThreads - 1) / kGsThreads;
constexpr int kGsLoads = (kGsChunks + kGsThreads - 1) / kGsThreads;   // float4 plane loads per thread
constexpr int kGsCPT = kGsTX * kGsTY / kGsThreads;                     // cells written per thread (2 or 4)

__device__ __forceinline__ int gs_word(int xl, int yl) { return yl * kGsW + (xl & 1) * kGsH + (xl >> 1); }
__device__ __forceinline__ int gs_wrap(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

// The tile of this workgroup, XCD-aware: workgroup i (x fastest) runs on XCD
// i % 8, so the tiles are renumbered for each XCD to take a contiguous run of
// them (neighbouring tiles share its L2 and their halo rows).  Returns the
// tile's linear index; (tx, ty, tz) its coordinates in the launch grid.
__device__ __forceinline__ int mg_tile(int &tx, int &ty, int &tz) {
    const int nb = gridDim.x * gridDim.y * gridDim.z;
    const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    const int q = nb >> 3, r = nb & 7, x = bid & 7;
    const int t = x * q + (x < r ? x : r) + (bid >> 3);
    tx = t % gridDim.x;
    ty = (t / gridDim.x) % gridDim.y;
    tz = t / (gridDim.x * gridDim.y);
    return t;
}
__device__ __forceinline__ double wave_sum64(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Local cell (xl, yl) and local plane ql = Q - z0 + 4 of a thread's cell in
// stage s at step k (parity par of the stage planes), from its stage row r
// and column i.
__device__ __forceinline__ void gs_cell(int s, int r, int i, int par, int &xl, int &yl) {
    yl = 1 + s + 2 * r + ((((s >> 1) ^ (1 + s)) ^ par) & 1);   // rows with (y + Q) & 1 == s >> 1
    xl = 1 + s + 2 * i + (((s & 1) ^ (1 + s + yl)) & 1);       // columns with (x + y) & 1 == s & 1
}

// part (may be NULL): per workgroup the FP64 sum of the values it writes,
// in a fixed order (the total of vout without another pass over it).
// LOS: a fixed line of sight (has_los); PART: part != NULL; ZIN: vin is
// zero and not read (vin == NULL: the first sweep of a correction V-cycle,
// whose level starts from v = 0); all three template constants.
template <int LOS, int PART, int ZIN>
__global__ void __launch_bounds__(kGsThreads) k_mg_gs4(float *__restrict__ vout, const float *__restrict__ vin,
                                                       const float *__restrict__ f, MGLevel L, float beta, int tz,
                                                       double *part) {
    constexpr int has_los = LOS;
    __shared__ __align__(16) float lds[kGsRing * kGsPlane];
    __shared__ double red[kGsThreads / 64];
    double acc = 0.0;
    // (p, p^2, cell^2 p^2) of each LDS column word (76) and row; (p, p^2) of each local plane
    __shared__ float4 xytab[kGsXWords + kGsRows];
    __shared__ float2 ztab[kGsMaxTz + 8];
    // byte offsets of ring slots (j - 1, j, j + 1) mod 10 for j < 20: a cell's
    // three planes from one LDS read instead of per-lane wrap arithmetic
    __shared__ int4 slot_tab[2 * kGsRing];
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    int tile_x, tile_y, tile_z;
    const int tile = mg_tile(tile_x, tile_y, tile_z);
    const int x0 = tile_x * kGsTX, y0 = tile_y * kGsTY, z0 = tile_z * tz;
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    const int tid = threadIdx.x;
    if (!has_los)
        for (int q = tid; q < kGsXWords + kGsRows + tz + 8; q += kGsThreads) {
            if (q < kGsXWords + kGsRows) {
                int g, a;
                if (q < kGsXWords) {   // column word q: x = 2q (even half) or 2(q - kGsH) + 1
                    g = gs_wrap(x0 - kGsHalo + min(q < kGsH ? 2 * q : 2 * (q - kGsH) + 1, kGsLW - 1), nx);
                    a = 0;
                } else {
                    g = nx + gs_wrap(y0 - kGsHalo + q - kGsXWords, ny);
                    a = 1;
                }
                const float p = L.pv[g], p2 = p * p;
                xytab[q] = make_float4(p, p2, L.cell2[a] * p2, 0.f);
            } else {
                const float p = L.pv[nx + ny + gs_wrap(z0 - kGsHalo + q - kGsXWords - kGsRows, nz)];
                ztab[q - kGsXWords - kGsRows] = make_float2(p, p * p);
            }
        }
    if (tid < 2 * kGsRing) {
        constexpr int PB = kGsPlane * (int)sizeof(float);
        const int j = tid % kGsRing;
        slot_tab[tid] = make_int4((j == 0 ? kGsRing - 1 : j - 1) * PB, j * PB, (j == kGsRing - 1 ? 0 : j + 1) * PB, 0);
    }

    // plane loads: chunk q = row * 18 + j covers x0 - 4 + 4j .. +3 of LDS row
    // `row`; threads past the 432 chunks load chunk 0 again and do not store
    int64_t ld_off[kGsLoads];
    int ld_word[kGsLoads];
    bool ld_ok[kGsLoads];
#pragma unroll
    for (int u = 0; u < kGsLoads; u++) {
        const int q = tid + kGsThreads * u;
        ld_ok[u] = q < kGsChunks;
        const int qq = ld_ok[u] ? q : 0;
        const int row = qq / (kGsLW / 4), j = qq % (kGsLW / 4);
        ld_off[u] = gs_wrap(x0 - kGsHalo + 4 * j, nx) + sy * gs_wrap(y0 - kGsHalo + row, ny);
        ld_word[u] = row * kGsW + 2 * j;
    }
    float4 pre[kGsLoads];
    // unconditional (a step past the last plane reloads the last plane), so
    // that no branch makes the compiler wait for the loads before the step's
    // arithmetic
    auto fetch = [&](int P) {
        if constexpr (ZIN) {
#pragma unroll
            for (int u = 0; u < kGsLoads; u++) pre[u] = make_float4(0.f, 0.f, 0.f, 0.f);
            return;
        }
        const float *src = vin + sz * gs_wrap(min(P, z0 + tz + kGsHalo - 1), nz);
#pragma unroll
        for (int u = 0; u < kGsLoads; u++) pre[u] = *reinterpret_cast<const float4 *>(src + ld_off[u]);
    };
    auto store = [&](int slot) {
        float *pl = lds + slot * kGsPlane;
#pragma unroll
        for (int u = 0; u < kGsLoads; u++)
            if (ld_ok[u]) {
                pl[ld_word[u]] = pre[u].x;
                pl[ld_word[u] + 1] = pre[u].z;
                pl[ld_word[u] + kGsH] = pre[u].y;
                pl[ld_word[u] + kGsH + 1] = pre[u].w;
            }
    };
    // the thread's cells: iteration it handles cell jc = tid + 512 it of the
    // cell list; desc = s | r << 2 | i << 8 (stage, its row and column), or -1
    int desc[kGsIters];
#pragma unroll
    for (int it = 0; it < kGsIters; it++) {
        const int jc = tid + kGsThreads * it;
        int s, r, i;
        if (jc < kGsB0) {   // columns 0..31
            s = jc < kGsA1 ? 0 : (jc < kGsA2 ? 1 : (jc < kGsA3 ? 2 : 3));
            const int jj = jc - (s == 0 ? 0 : (s == 1 ? kGsA1 : (s == 2 ? kGsA2 : kGsA3)));
            r = jj >> 5;
            i = jj & 31;
        } else {            // columns 32 .. 34 - s of stages 0..2
            s = jc < kGsB1 ? 0 : (jc < kGsB2 ? 1 : 2);
            const int jj = jc - (s == 0 ? kGsB0 : (s == 1 ? kGsB1 : kGsB2)), e = 3 - s;
            r = jj / e;
            i = 32 + jj - r * e;
        }
        desc[it] = jc < kGsCells ? (s | (r << 2) | (i << 8)) : -1;
    }
    // stage s updates plane Q = z0 - 5 - 2s + k, needed on z0 - 3 + s .. z0 + tz + 2 - s
    auto active = [&](int s, int k) { return k >= 2 + 3 * s && k <= tz + 7 + s; };
    // Per iteration and parity par of the stage planes (z0 is even, so par =
    // (k + 1) & 1): the byte offset in a plane of the LDS word one row below
    // the cell (every neighbour then sits at a non-negative offset from it),
    // the cell's column word and row (cw | yl << 8) and its offset in a plane
    // of f (32-bit: the grid has < 2^31 cells).
    //
    // fetch_f(k, par) below runs for k = 1 - par, 3 - par, ... in turn and
    // reads plane Q = z0 - 5 - 2s + k of f, so fo[par][it], the offset in f
    // of the cell's next read, starts at k = 1 - par and advances two planes
    // per call, wrapping past the last plane.
    int cwb[2][kGsIters], cpos[2][kGsIters];
    uint32_t fo[2][kGsIters];
    const uint32_t sz32 = (uint32_t)sz, ncell = (uint32_t)(sz * nz);
#pragma unroll
    for (int par = 0; par < 2; par++)
#pragma unroll
        for (int it = 0; it < kGsIters; it++) {
            const int d = desc[it];
            int xl = kGsHalo, yl = kGsHalo;
            if (d >= 0) gs_cell(d & 3, (d >> 2) & 63, d >> 8, par, xl, yl);
            cwb[par][it] = (gs_word(xl, yl) - kGsW) * (int)sizeof(float);
            cpos[par][it] = gs_word(xl, 0) | (yl << 8);
            const int q0 = ((z0 - 5 - 2 * (d & 3) + 1 - par) % nz + nz) % nz;
            fo[par][it] = (uint32_t)gs_wrap(x0 - kGsHalo + xl, nx) + (uint32_t)nx * gs_wrap(y0 - kGsHalo + yl, ny) +
                          sz32 * (uint32_t)q0;
        }
    // f of the thread's cells at step k (a valid dummy address where inactive)
    auto fetch_f = [&](int k, auto parc, float *fr) {
        constexpr int par = decltype(parc)::value;
#pragma unroll
        for (int it = 0; it < kGsIters; it++) {
            const int d = desc[it];
            const uint32_t off = d >= 0 && active(d & 3, k) ? fo[par][it] : 0u;
            fr[it] = f[off];
            const uint32_t nxt = fo[par][it] + 2 * sz32;   // < 2^32: fo < ncell < 2^31
            fo[par][it] = min(nxt, nxt - ncell);             // nxt - ncell wraps above nxt unless nxt >= ncell
        }
    };
    // write-out: thread -> row tid / 16, x 4 (tid % 16) of the 64 x 16 tile
    constexpr int kWoTPR = kGsTX / kGsCPT;   // threads per tile row
    const int wo_row = tid / kWoTPR, wo_x = kGsCPT * (tid % kWoTPR);
    const int64_t wo_off = (x0 + wo_x) + sy * (y0 + wo_row);
    const int wo_word = (kGsHalo + wo_row) * kGsW + ((kGsHalo + wo_x) >> 1);
    const int last_load = tz + 7;

    GsCoef<float> cfix{};   // the coefficients of a fixed LOS
    if constexpr (LOS) cfix = gs_coef_p<float>(L, L.losn[0], L.losn[1], L.losn[2], beta);

    // step k; parc: the parity of its stage planes, (k + 1) & 1
    auto step = [&](int k, auto parc, const float *fcur, float *fnext) {
        constexpr int par = decltype(parc)::value;
        if (k <= last_load) store(k % kGsRing);
        fetch(z0 - kGsHalo + k + 1);
        __syncthreads();
        if (k >= 12) {   // plane z0 - 12 + k, finished in step k - 1
            const float *pl = lds + ((k - 8) % kGsRing) * kGsPlane;
            if constexpr (kGsCPT == 4) {
                const float2 ev = *reinterpret_cast<const float2 *>(pl + wo_word);
                const float2 od = *reinterpret_cast<const float2 *>(pl + wo_word + kGsH);
                *reinterpret_cast<float4 *>(vout + sz * (z0 - 12 + k) + wo_off) = make_float4(ev.x, od.x, ev.y, od.y);
                if constexpr (PART) acc += ((double)ev.x + (double)od.x) + ((double)ev.y + (double)od.y);
            } else {
                const float ev = pl[wo_word], od = pl[wo_word + kGsH];
                *reinterpret_cast<float2 *>(vout + sz * (z0 - 12 + k) + wo_off) = make_float2(ev, od);
                if constexpr (PART) acc += (double)ev + (double)od;
            }
        }
        fetch_f(k + 1, std::integral_constant<int, 1 - par>{}, fnext);
        const int k10 = k % kGsRing;
        // the update of the thread's cell of iteration it, in stage st
        auto update = [&](int it, int st) {
            const int cw = cpos[par][it] & 255, yl = cpos[par][it] >> 8;   // column word, row
            const int ql = k - 1 - 2 * st;   // local plane of Q, >= 1
            const int4 Tc = slot_tab[k10 + kGsRing - 1 - 2 * st];   // ring slots of Q - 1, Q, Q + 1
            char *lb = reinterpret_cast<char *>(lds);
            const int wb = cwb[par][it];
            // x neighbours sit in the other parity half: x - 1 at word ob, x + 1
            // at ob + 1, ob = wc + (cw >= kGsH ? -kGsH : kGsH - 1)
            const int db = (cw >= kGsH ? -kGsH : kGsH - 1) * (int)sizeof(float);
            // bases one row below the cell (a*) and below its x - 1 neighbour (b*)
            const float *am = reinterpret_cast<const float *>(lb + (Tc.x + wb));
            float *a0 = reinterpret_cast<float *>(lb + (Tc.y + wb));
            const float *ap = reinterpret_cast<const float *>(lb + (Tc.z + wb));
            const float *bm = reinterpret_cast<const float *>(lb + (Tc.x + wb + db));
            const float *b0 = reinterpret_cast<const float *>(lb + (Tc.y + wb + db));
            const float *bp = reinterpret_cast<const float *>(lb + (Tc.z + wb + db));
#define A0(o) a0[kGsW + (o)]
#define AM(o) am[kGsW + (o)]
#define AP(o) ap[kGsW + (o)]
#define B0(o) b0[kGsW + (o)]
#define BM(o) bm[kGsW + (o)]
#define BP(o) bp[kGsW + (o)]
            GsCoef<float> c;
            if constexpr (LOS) {
                c = cfix;
            } else {
                const float4 tx = xytab[cw], ty = xytab[kGsXWords + yl];
                const float2 tq = ztab[ql];
                c = gs_coef<float>(L, tx.x, ty.x, tq.x, tx.y, ty.y, tq.y, tx.z, ty.z, L.cell2[2] * tq.y, beta);
            }
            A0(0) = gs_update<float>(c, fcur[it], B0(1), B0(0), A0(kGsW), A0(-kGsW), AP(0), AM(0), B0(kGsW + 1),
                                     B0(-kGsW), B0(kGsW), B0(-kGsW + 1), BP(1), BM(0), BP(0), BM(1), AP(kGsW),
                                     AM(-kGsW), AP(-kGsW), AM(kGsW), has_los);
#undef A0
#undef AM
#undef AP
#undef B0
#undef BM
#undef BP
        };
#pragma unroll
        for (int it = 0; it < kGsIters; it++) {
            const int d = desc[it];
            if (d < 0 || !active(d & 3, k)) continue;
            update(it, d & 3);
        }
    };

    float fa[kGsIters], fb[kGsIters];
    fetch(z0 - kGsHalo);
    fetch_f(0, std::integral_constant<int, 1>{}, fa);
    for (int k = 0; k < tz + 12; k += 2) {   // tz + 12 is even: the f buffers alternate
        step(k, std::integral_constant<int, 1>{}, fa, fb);
        step(k + 1, std::integral_constant<int, 0>{}, fb, fa);
    }
    if constexpr (PART) {
        acc = wave_sum64(acc);
        if ((tid & 63) == 0) red[tid >> 6] = acc;
        __syncthreads();
        if (tid == 0) {
            double t = 0.0;
            for (int w = 0; w < kGsThreads / 64; w++) t += red[w];
            part[tile] = t;
        }
    }
}

// S damped Jacobi sweeps in one launch (temporal blocking), out of place:
// vout is what S launches of k_mg_jacobi leave (the same cell arithmetic on
// the same values, so the same bits).  A workgroup owns a 64 x 16 column of
// tz planes and marches up z; sweep level t (t = 1..S) of the tile, grown by
// S - t cells in x and y, lives in a ring of 4 LDS planes (level 0: the old
// values, loaded).  Step k loads plane P = z0 - S + k and computes level t
// on plane P - 2t + 1: level t - 1 there and on both neighbouring planes was
// finished in earlier steps, and the stages of one step touch disjoint
// planes, so one barrier per step orders them.  Level S (the tile itself)
// goes straight to vout, a row per wave.
constexpr int kJbTX = 64, kJbTY = 16, kJbRing = 4, kJbMaxTz = 128;

template <int S>
struct JbShape {
    // 512 threads at S = 3 keep its per-thread cell lists short enough for
    // two waves per SIMD without spills
    static constexpr int Threads = S == 3 ? 512 : 256;
    static constexpr int CPT = kJbTX * kJbTY / Threads;   // level-S cells per thread (one column)
    static constexpr int RS = Threads / kJbTX;             // rows between them
    static constexpr int W = kJbTX + 2 * S, H = kJbTY + 2 * S, Plane = W * H;
    // cells of the levels below S: level t covers (64 + 2e) x (16 + 2e), e = S - t
    static constexpr int cells(int t) { return (kJbTX + 2 * (S - t)) * (kJbTY + 2 * (S - t)); }
    static constexpr int Inner = S == 2 ? cells(1) : cells(1) + cells(2);
    static constexpr int Iters = (Inner + Threads - 1) / Threads;
    static constexpr int Chunks = (W + 3) / 4 * H;   // float4 columns per plane row (last may overhang)
};

// LOS: a fixed line of sight (has_los), as a template constant
template <int S, int LOS>
__global__ void __launch_bounds__(JbShape<S>::Threads, 2) k_mg_jacobi_tb(float *__restrict__ vout, const float *__restrict__ vin,
                                                             const float *__restrict__ f, MGLevel L, float beta,
                                                             float omega, int tz) {
    static_assert(S == 2 || S == 3, "2 or 3 sweeps per pass");
    constexpr int has_los = LOS;
    using J = JbShape<S>;
    __shared__ __align__(16) float lds[S * kJbRing * J::Plane];
    __shared__ float ptab[J::W + J::H + kJbMaxTz + 2 * S];
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    int tile_x, tile_y, tile_z;
    mg_tile(tile_x, tile_y, tile_z);
    const int x0 = tile_x * kJbTX, y0 = tile_y * kJbTY, z0 = tile_z * tz;
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    const int tid = threadIdx.x;
    float *pz_tab = ptab + J::W + J::H;
    if (!has_los)
        for (int q = tid; q < J::W + J::H + tz + 2 * S; q += J::Threads) {
            int g;
            if (q < J::W)
                g = gs_wrap(x0 - S + q, nx);
            else if (q < J::W + J::H)
                g = nx + gs_wrap(y0 - S + q - J::W, ny);
            else
                g = nx + ny + gs_wrap(z0 - S + q - J::W - J::H, nz);
            ptab[q] = L.pv[g];
        }
    auto plane = [&](int t, int Pl) {   // level t, local plane Pl = P - z0 + S >= 0
        return lds + (t * kJbRing + (Pl & (kJbRing - 1))) * J::Plane;
    };
    // level-0 loads: element q = row * W + col, one float per element (W is
    // not a multiple of 4 at S = 3); threads past the plane reload element 0
    constexpr int NL = (J::Plane + J::Threads - 1) / J::Threads;
    int64_t ld_off[NL];
#pragma unroll
    for (int u = 0; u < NL; u++) {
        const int q = tid + J::Threads * u, qq = q < J::Plane ? q : 0;
        ld_off[u] = gs_wrap(x0 - S + qq % J::W, nx) + sy * gs_wrap(y0 - S + qq / J::W, ny);
    }
    float pre[NL];
    auto fetch = [&](int P) {
        const float *src = vin + sz * gs_wrap(min(P, z0 + tz - 1 + S), nz);
#pragma unroll
        for (int u = 0; u < NL; u++) pre[u] = src[ld_off[u]];
    };
    auto store = [&](int Pl) {
        float *pl = plane(0, Pl);
#pragma unroll
        for (int u = 0; u < NL; u++) {
            const int q = tid + J::Threads * u;
            if (q < J::Plane) pl[q] = pre[u];
        }
    };
    // inner levels: iteration it handles cell jc = tid + 256 it of the list
    // (level 1 cells, then level 2 at S = 3); cpos = xl | yl << 8 | t << 16
    // (t = 0: no cell) and the cell's offset in a plane of f (32-bit: the
    // grid has < 2^31 cells)
    int cpos[J::Iters];
    uint32_t foff[J::Iters];
#pragma unroll
    for (int it = 0; it < J::Iters; it++) {
        const int jc = tid + J::Threads * it;
        int t = 1, jj = jc;
        if (S == 3 && jc >= J::cells(1)) {
            t = 2;
            jj = jc - J::cells(1);
        }
        const int e = S - t, w = kJbTX + 2 * e;
        const int yl = S - e + jj / w, xl = S - e + jj % w;
        cpos[it] = jc < J::Inner ? (xl | (yl << 8) | (t << 16)) : (S | (S << 8));
        foff[it] = (uint32_t)gs_wrap(x0 - S + xl, nx) + (uint32_t)nx * gs_wrap(y0 - S + yl, ny);
    }
    const uint32_t sz32 = (uint32_t)sz;
    // level t is computed on plane Q = z0 - S + k - 2t + 1, needed on
    // z0 - (S - t) .. z0 + tz - 1 + (S - t)
    auto active = [&](int t, int k) { return t > 0 && k >= 3 * t - 1 && k <= tz + 2 * S + t - 2; };
    auto fetch_f = [&](int k, float *fr) {
#pragma unroll
        for (int it = 0; it < J::Iters; it++) {
            const int t = cpos[it] >> 16;
            const uint32_t off =
                active(t, k) ? foff[it] + sz32 * (uint32_t)gs_wrap(z0 - S + k - 2 * t + 1, nz) : 0u;
            fr[it] = f[off];
        }
    };
    // level S: cells (ox, oy + RS j), j < CPT, of the tile: a wave takes a
    // row, a column per lane, so its LDS reads are consecutive words
    struct FV {
        float v[J::CPT];
    };
    const int ox = tid % kJbTX, oy = tid / kJbTX;
    const int64_t o_off = (x0 + ox) + sy * (y0 + oy);
    auto fetch_fo = [&](int k, FV &fo) {
        const int Q = z0 - S + k - 2 * S + 1;
        const float *src = f + sz * (active(S, k) ? Q : z0) + o_off;
#pragma unroll
        for (int j = 0; j < J::CPT; j++) fo.v[j] = src[sy * J::RS * j];
    };
    // one cell of level t from level t - 1 (local cell xl, yl, local plane ql)
    auto cell = [&](int t, int xl, int yl, int ql, float fv) {
        const float *pm = plane(t - 1, ql - 1), *p0 = plane(t - 1, ql), *pp = plane(t - 1, ql + 1);
        const int wc = yl * J::W + xl;
        const Coef<float> c = coef_p<float>(L, has_los ? L.losn[0] : ptab[xl], has_los ? L.losn[1] : ptab[J::W + yl],
                                            has_los ? L.losn[2] : pz_tab[ql], beta);
#define A0(o) p0[wc + (o)]
#define AM(o) pm[wc + (o)]
#define AP(o) pp[wc + (o)]
        const Nbr<float> sn = nbr_of<float>(c, A0(1), A0(-1), A0(J::W), A0(-J::W), AP(0), AM(0), A0(J::W + 1),
                                            A0(-J::W - 1), A0(J::W - 1), A0(-J::W + 1), AP(1), AM(-1), AP(-1),
                                            AM(1), AP(J::W), AM(-J::W), AP(-J::W), AM(J::W));
        const float v = jacobi_value(c, fv, sn, A0(0), omega, has_los);
#undef A0
#undef AM
#undef AP
        return v;
    };

    auto step = [&](int k, const float *fcur, float *fnext, const FV &focur, FV &fonext) {
        if (k <= tz + 2 * S - 1) store(k);
        fetch(z0 - S + k + 1);
        fetch_f(k + 1, fnext);
        fetch_fo(k + 1, fonext);
        __syncthreads();
#pragma unroll
        for (int it = 0; it < J::Iters; it++) {
            const int t = cpos[it] >> 16;
            if (!active(t, k)) continue;
            const int xl = cpos[it] & 255, yl = (cpos[it] >> 8) & 255, ql = k - 2 * t + 1;
            plane(t, ql)[yl * J::W + xl] = cell(t, xl, yl, ql, fcur[it]);
        }
        if (active(S, k)) {
            const int ql = k - 2 * S + 1;
            float *dst = vout + sz * (z0 - S + ql) + o_off;
#pragma unroll
            for (int j = 0; j < J::CPT; j++) dst[sy * J::RS * j] = cell(S, S + ox, S + oy + J::RS * j, ql, focur.v[j]);
        }
    };

    float fa[J::Iters], fb[J::Iters];
    FV foa, fob;
    fetch(z0 - S);
    fetch_f(0, fa);
    fetch_fo(0, foa);
    const int nsteps = tz + 3 * S - 1;
    int k = 0;
    for (; k + 1 < nsteps; k += 2) {   // the f buffers alternate
        step(k, fa, fb, foa, fob);
        step(k + 1, fb, fa, fob, foa);
    }
    if (k < nsteps) step(k, fa, fb, foa, fob);
}

// One FP64 refinement pass of the RBGS solve (z-marching, LDS ring):
//   phi_out = phi_in + (e - sum(e) / N)          (with e; without, phi_in is used;
//                                                 seed: phi = (double)seed; neither: phi = 0)
//   r = (float)((f - alpha) - A phi_out),  alpha = *fsum / N,
// the arithmetic of k_add_correction followed by k_mg_residual_mixed, plus
// each workgroup's FP64 sum of r in a fixed order (part[block]).  A
// workgroup owns a 64 x 16 column of tz planes; step k loads plane z0-1+k of
// phi_in (+ e) into a ring of 4 planes of 66 x 18 doubles (a halo of 1),
// writes its interior to phi_out, and computes r on the plane below.  Every
// value of phi and r is read or written once per pass.
constexpr int kRfTX = 64, kRfTY = 16;
constexpr int kRfW = kRfTX + 2, kRfRows = kRfTY + 2, kRfPlane = kRfW * kRfRows;   // 66 x 18
constexpr int kRfRing = 4, kRfThreads = 256, kRfMaxTz = 128;
constexpr int kRfLoads = (kRfPlane + kRfThreads - 1) / kRfThreads;   // 5

// The pass's form, a template constant: which of phi_in, e and seed it reads.
enum RfMode { kRfPhi = 0, kRfPhiE = 1, kRfSeed = 2, kRfZero = 3 };

template <int LOS, int MODE>
__global__ void __launch_bounds__(kRfThreads)
    k_mg_refine(double *__restrict__ phi_out, const double *__restrict__ phi_in, const float *__restrict__ seed,
                const float *__restrict__ e, const double *esum, const float *__restrict__ f, const double *fsum,
                float *__restrict__ r, double *__restrict__ part, MGLevel L, float beta, int tz) {
    constexpr int has_los = LOS;
    __shared__ double ring[kRfRing * kRfPlane];
    __shared__ double ptab[kRfW + kRfRows + kRfMaxTz + 2];   // p of LDS column | row | local plane
    __shared__ double red[kRfThreads / 64];
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    int tile_x, tile_y, tile_z;
    const int tile = mg_tile(tile_x, tile_y, tile_z);
    const int x0 = tile_x * kRfTX, y0 = tile_y * kRfTY, z0 = tile_z * tz;
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    const double count = (double)sz * nz;
    const double me = MODE == kRfPhiE ? *esum / count : 0.0;
    const double alpha = fsum ? *fsum / count : 0.0;
    const int tid = threadIdx.x;
    double *pz_tab = ptab + kRfW + kRfRows;
    if (!has_los)
        for (int q = tid; q < kRfW + kRfRows + tz + 2; q += kRfThreads) {
            int a, g;
            if (q < kRfW) {
                a = 0;
                g = gs_wrap(x0 - 1 + q, nx);
            } else if (q < kRfW + kRfRows) {
                a = 1;
                g = nx + gs_wrap(y0 - 1 + q - kRfW, ny);
            } else {
                a = 2;
                g = nx + ny + gs_wrap(z0 - 1 + q - kRfW - kRfRows, nz);
            }
            ptab[q] = (double)L.xv[g] / (double)L.cell[a];   // as coef<double>
        }
    // plane loads: element q = row * 66 + col (threads past 1188 reload 0)
    int64_t ld_off[kRfLoads];
    bool ld_in[kRfLoads];   // interior column and row: written to phi_out
#pragma unroll
    for (int u = 0; u < kRfLoads; u++) {
        const int q = tid + kRfThreads * u, qq = q < kRfPlane ? q : 0;
        const int row = qq / kRfW, col = qq % kRfW;
        ld_off[u] = gs_wrap(x0 - 1 + col, nx) + sy * gs_wrap(y0 - 1 + row, ny);
        ld_in[u] = q < kRfPlane && col >= 1 && col <= kRfTX && row >= 1 && row <= kRfTY;
    }
    double pre[kRfLoads];
    float pre_e[kRfLoads];
    auto fetch = [&](int P) {
        const int64_t zo = sz * gs_wrap(min(P, z0 + tz), nz);
#pragma unroll
        for (int u = 0; u < kRfLoads; u++) {
            if constexpr (MODE == kRfSeed)
                pre[u] = (double)seed[zo + ld_off[u]];
            else if constexpr (MODE == kRfZero)
                pre[u] = 0.0;
            else
                pre[u] = phi_in[zo + ld_off[u]];
            if constexpr (MODE == kRfPhiE) pre_e[u] = e[zo + ld_off[u]];
        }
    };
    constexpr bool write_phi = MODE != kRfPhi;   // phi_out receives the pass's phi
    auto store = [&](int P) {
        double *pl = ring + (P - z0 + 1) % kRfRing * kRfPlane;
        const int64_t zo = sz * gs_wrap(P, nz);
        const bool own = P >= z0 && P < z0 + tz;
#pragma unroll
        for (int u = 0; u < kRfLoads; u++) {
            const int q = tid + kRfThreads * u;
            if (q >= kRfPlane) continue;
            const double v = MODE == kRfPhiE ? (double)pre[u] + ((double)pre_e[u] - me) : pre[u];
            pl[q] = v;
            if (write_phi && own && ld_in[u]) phi_out[zo + ld_off[u]] = v;
        }
    };
    // f of the tile's plane Q: float4 chunk tid (row tid / 16, x 4 (tid % 16)),
    // staged through LDS (two slots); cells: x = tid % 64, rows tid / 64 + 4j
    __shared__ __align__(16) float ftile[2][kRfTX * kRfTY];
    const int64_t f_off = (x0 + 4 * (tid & 15)) + sy * (y0 + (tid >> 4));
    float4 fpre = make_float4(0.f, 0.f, 0.f, 0.f);
    auto fetch_f = [&](int Q) { fpre = *reinterpret_cast<const float4 *>(f + sz * min(Q, z0 + tz - 1) + f_off); };
    const int cx = tid & 63, cy0 = tid >> 6;
    const int64_t c_off = (x0 + cx) + sy * (y0 + cy0);
    double acc = 0.0;
    fetch(z0 - 1);
    fetch_f(z0);
    for (int k = 0; k < tz + 2; k++) {
        const int P = z0 - 1 + k;
        store(P);
        if (k >= 2) *reinterpret_cast<float4 *>(&ftile[k & 1][4 * tid]) = fpre;   // f of plane P - 1
        fetch(P + 1);
        if (k >= 2) fetch_f(P);
        __syncthreads();
        if (k < 2) continue;
        const int Q = P - 1, ql = Q - z0 + 1;
        const double *pm = ring + (ql - 1) % kRfRing * kRfPlane;
        const double *p0 = ring + ql % kRfRing * kRfPlane;
        const double *pp = ring + (ql + 1) % kRfRing * kRfPlane;
        const float *ft = ftile[k & 1];
#pragma unroll 1
        for (int j = 0; j < 4; j++) {
            const int cy = cy0 + 4 * j;
            const int xl = cx + 1, yl = cy + 1, wc = yl * kRfW + xl;
            const Coef<double> c = coef_p<double>(L, has_los ? (double)L.losn[0] : ptab[xl],
                                                  has_los ? (double)L.losn[1] : ptab[kRfW + yl],
                                                  has_los ? (double)L.losn[2] : pz_tab[ql], (double)beta);
#define A0(o) p0[wc + (o)]
#define AM(o) pm[wc + (o)]
#define AP(o) pp[wc + (o)]
            const Nbr<double> sn = nbr_of<double>(
                c, A0(1), A0(-1), A0(kRfW), A0(-kRfW), AP(0), AM(0), A0(kRfW + 1), A0(-kRfW - 1), A0(kRfW - 1),
                A0(-kRfW + 1), AP(1), AM(-1), AP(-1), AM(1), AP(kRfW), AM(-kRfW), AP(-kRfW), AM(kRfW));
            double t = 2.0 * (c.gpx2 + c.gpy2 + c.gpz2) * A0(0) - (sn.A + sn.B + sn.C + sn.X);
#undef A0
#undef AM
#undef AP
            if (!has_los) t -= sn.G;
            const float rv = (float)(((double)ft[cy * kRfTX + cx] - alpha) - t);
            r[sz * Q + c_off + sy * 4 * j] = rv;
            acc += (double)rv;
        }
    }
    acc = wave_sum64(acc);
    if ((tid & 63) == 0) red[tid >> 6] = acc;
    __syncthreads();
    if (tid == 0)
        part[tile] = (red[0] + red[1]) + (red[2] + red[3]);
}

// r = f - A v in FP32 (k_mg_residual<float>'s arithmetic, so the same bits)
// as one z-marching pass: a workgroup owns a 64 x 16 column of tz planes,
// loads each plane of v once into a ring of 4 LDS planes of 66 x 18 (a halo
// of 1) and computes the plane below.  f is staged through LDS with the
// same one-step lead; r goes out in coalesced rows.
// FAST: the coefficient g with the hardware reciprocal (mg_div<float, true>,
// 1 ulp) instead of the reference's IEEE division, for the residuals inside
// the RBGS solve
template <int FAST>
__global__ void __launch_bounds__(kRfThreads)
    k_mg_residual_zm(float *__restrict__ r, const float *__restrict__ v, const float *__restrict__ f, MGLevel L,
                     float beta, int has_los, int tz) {
    __shared__ float ring[kRfRing * kRfPlane];
    __shared__ float ptab[kRfW + kRfRows + kRfMaxTz + 2];
    __shared__ __align__(16) float ftile[2][kRfTX * kRfTY];
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    int tile_x, tile_y, tile_z;
    mg_tile(tile_x, tile_y, tile_z);
    const int x0 = tile_x * kRfTX, y0 = tile_y * kRfTY, z0 = tile_z * tz;
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    const int tid = threadIdx.x;
    float *pz_tab = ptab + kRfW + kRfRows;
    if (!has_los)
        for (int q = tid; q < kRfW + kRfRows + tz + 2; q += kRfThreads) {
            int g;
            if (q < kRfW)
                g = gs_wrap(x0 - 1 + q, nx);
            else if (q < kRfW + kRfRows)
                g = nx + gs_wrap(y0 - 1 + q - kRfW, ny);
            else
                g = nx + ny + gs_wrap(z0 - 1 + q - kRfW - kRfRows, nz);
            ptab[q] = L.pv[g];
        }
    int64_t ld_off[kRfLoads];
#pragma unroll
    for (int u = 0; u < kRfLoads; u++) {
        const int q = tid + kRfThreads * u, qq = q < kRfPlane ? q : 0;
        ld_off[u] = gs_wrap(x0 - 1 + qq % kRfW, nx) + sy * gs_wrap(y0 - 1 + qq / kRfW, ny);
    }
    float pre[kRfLoads];
    auto fetch = [&](int P) {
        const int64_t zo = sz * gs_wrap(min(P, z0 + tz), nz);
#pragma unroll
        for (int u = 0; u < kRfLoads; u++) pre[u] = v[zo + ld_off[u]];
    };
    auto store = [&](int P) {
        float *pl = ring + (P - z0 + 1) % kRfRing * kRfPlane;
#pragma unroll
        for (int u = 0; u < kRfLoads; u++) {
            const int q = tid + kRfThreads * u;
            if (q < kRfPlane) pl[q] = pre[u];
        }
    };
    const int64_t f_off = (x0 + 4 * (tid & 15)) + sy * (y0 + (tid >> 4));
    float4 fpre = make_float4(0.f, 0.f, 0.f, 0.f);
    auto fetch_f = [&](int Q) { fpre = *reinterpret_cast<const float4 *>(f + sz * min(Q, z0 + tz - 1) + f_off); };
    const int cx = tid & 63, cy0 = tid >> 6;
    const int64_t c_off = (x0 + cx) + sy * (y0 + cy0);
    fetch(z0 - 1);
    fetch_f(z0);
    for (int k = 0; k < tz + 2; k++) {
        const int P = z0 - 1 + k;
        store(P);
        if (k >= 2) *reinterpret_cast<float4 *>(&ftile[k & 1][4 * tid]) = fpre;   // f of plane P - 1
        fetch(P + 1);
        if (k >= 2) fetch_f(P);
        __syncthreads();
        if (k < 2) continue;
        const int Q = P - 1, ql = Q - z0 + 1;
        const float *pm = ring + (ql - 1) % kRfRing * kRfPlane;
        const float *p0 = ring + ql % kRfRing * kRfPlane;
        const float *pp = ring + (ql + 1) % kRfRing * kRfPlane;
        const float *ft = ftile[k & 1];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int cy = cy0 + 4 * j;
            const int xl = cx + 1, yl = cy + 1, wc = yl * kRfW + xl;
            const Coef<float> c = coef_p<float, FAST != 0>(L, has_los ? L.losn[0] : ptab[xl],
                                                           has_los ? L.losn[1] : ptab[kRfW + yl],
                                                           has_los ? L.losn[2] : pz_tab[ql], beta);
#define A0(o) p0[wc + (o)]
#define AM(o) pm[wc + (o)]
#define AP(o) pp[wc + (o)]
            const Nbr<float> sn = nbr_of<float>(
                c, A0(1), A0(-1), A0(kRfW), A0(-kRfW), AP(0), AM(0), A0(kRfW + 1), A0(-kRfW - 1), A0(kRfW - 1),
                A0(-kRfW + 1), AP(1), AM(-1), AP(-1), AM(1), AP(kRfW), AM(-kRfW), AP(-kRfW), AM(kRfW));
            float t = 2.0f * (c.gpx2 + c.gpy2 + c.gpz2) * A0(0) - (sn.A + sn.B + sn.C + sn.X);
#undef A0
#undef AM
#undef AP
            if (!has_los) t -= sn.G;
            r[sz * Q + c_off + sy * 4 * j] = ft[cy * kRfTX + cx] - t;
        }
    }
}

// r += (float)(-(*rsum) / n) when rsum (the mean removal of k_add_const),
// then per block the FP64 sums of r^2 and of (r - rprev)^2 (when rprev), in
// a fixed order: part[block] and part[nblocks + block].
__global__ void __launch_bounds__(256) k_shift_norms(float *r, const float *rprev, int64_t n, const double *rsum,
                                                     double *part) {
    __shared__ double red[2][4];
    const float cst = rsum ? (float)(-1.0 * (*rsum) / (double)n) : 0.f;
    double s2 = 0.0, d2 = 0.0;
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    auto cell = [&](float v, const float *pv, int64_t i, float &out) {
        if (rsum) v += cst;
        out = v;
        s2 += (double)v * (double)v;
        if (rprev) {
            const double d = (double)v - (double)pv[i];
            d2 += d * d;
        }
    };
    if ((n & 3) == 0 && ((uintptr_t)r & 15) == 0 && ((uintptr_t)rprev & 15) == 0) {
        float4 *r4 = reinterpret_cast<float4 *>(r);
        const float4 *p4 = reinterpret_cast<const float4 *>(rprev);
        for (int64_t i = t0; i < (n >> 2); i += stride) {
            const float4 v = r4[i];
            const float4 q = rprev ? p4[i] : v;
            const float pq[4] = {q.x, q.y, q.z, q.w};
            float4 o;
            cell(v.x, pq, 0, o.x);
            cell(v.y, pq, 1, o.y);
            cell(v.z, pq, 2, o.z);
            cell(v.w, pq, 3, o.w);
            if (rsum) r4[i] = o;
        }
    } else {
        for (int64_t i = t0; i < n; i += stride) {
            float o;
            cell(r[i], rprev, i, o);
            if (rsum) r[i] = o;
        }
    }
    s2 = wave_sum64(s2);
    d2 = wave_sum64(d2);
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = s2;
        red[1][threadIdx.x >> 6] = d2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[blockIdx.x] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
        part[gridDim.x + blockIdx.x] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
    }
}

// 27-point full weighting around fine cell 2c+1 (src/multigrid.jl:586-640).
template <typename T, typename U>
__device__ __forceinline__ void cell_reduce(T *__restrict__ coarse, const U *__restrict__ w, int fx, int fy, int fz,
                                            int kx, int ky, int kz) {
    const int cx = fx / 2, cy = fy / 2;
    const int ix = 2 * kx + 1, iy = 2 * ky + 1, iz = 2 * kz + 1;
    const int ixp = ix + 1 >= fx ? ix + 1 - fx : ix + 1, ixm = ix - 1;
    const int iyp = iy + 1 >= fy ? iy + 1 - fy : iy + 1, iym = iy - 1;
    const int izp = iz + 1 >= fz ? iz + 1 - fz : iz + 1, izm = iz - 1;
    const int64_t sy = fx, sz = (int64_t)fx * fy;
#define W(a, b, d) ((T)w[(a) + sy * (b) + sz * (d)])
    const T s = (T)8 * W(ix, iy, iz) +
                (T)4 * (W(ixp, iy, iz) + W(ixm, iy, iz) + W(ix, iyp, iz) + W(ix, iym, iz) +
                        W(ix, iy, izp) + W(ix, iy, izm)) +
                (T)2 * (W(ixp, iyp, iz) + W(ixm, iyp, iz) + W(ixp, iym, iz) + W(ixm, iym, iz) +
                        W(ixp, iy, izp) + W(ixm, iy, izp) + W(ixp, iy, izm) + W(ixm, iy, izm) +
                        W(ix, iyp, izp) + W(ix, iym, izp) + W(ix, iyp, izm) + W(ix, iym, izm)) +
                W(ixp, iyp, izp) + W(ixm, iyp, izp) + W(ixp, iym, izp) + W(ixm, iym, izp) +
                W(ixp, iyp, izm) + W(ixm, iyp, izm) + W(ixp, iym, izm) + W(ixm, iym, izm);
#undef W
    coarse[kx + (int64_t)cx * (ky + (int64_t)cy * kz)] = (T)((double)s / 64.0);
}

template <typename T, typename U>
__global__ void k_mg_reduce(T *__restrict__ coarse, const U *__restrict__ w, int fx, int fy, int fz) {
    const int kx = blockIdx.x * blockDim.x + threadIdx.x;
    if (kx >= fx / 2) return;
    cell_reduce<T, U>(coarse, w, fx, fy, fz, kx, blockIdx.y, blockIdx.z);
}

// Trilinear prolongation (src/multigrid.jl:456-509) evaluated per fine cell:
// odd fine index 2c+1 copies coarse c, even index 2c+2 averages c and c+1;
// the sums keep the reference's term order for each of the 8 cases.
template <typename U>
__device__ __forceinline__ U prolong_value(const U *__restrict__ v, int cx, int cy, int cz, int ix, int iy, int iz) {
    const int tx = (ix & 1) == 0, ty = (iy & 1) == 0, tz = (iz & 1) == 0;
    const int ax = tx ? (ix == 0 ? cx - 1 : ix / 2 - 1) : (ix - 1) / 2;
    const int ay = ty ? (iy == 0 ? cy - 1 : iy / 2 - 1) : (iy - 1) / 2;
    const int az = tz ? (iz == 0 ? cz - 1 : iz / 2 - 1) : (iz - 1) / 2;
    const int bx = ax + 1 == cx ? 0 : ax + 1, by = ay + 1 == cy ? 0 : ay + 1;
    const int bz = az + 1 == cz ? 0 : az + 1;
    const int64_t sy = cx, sz = (int64_t)cx * cy;
#define C(a, b, d) v[(a) + sy * (b) + sz * (d)]
    const U c000 = C(ax, ay, az);
    U val;
    const int code = tx + 2 * ty + 4 * tz;
    switch (code) {
        case 0: val = c000; break;
        case 1: val = (c000 + C(bx, ay, az)) / (U)2; break;
        case 2: val = (c000 + C(ax, by, az)) / (U)2; break;
        case 4: val = (c000 + C(ax, ay, bz)) / (U)2; break;
        case 3: val = (c000 + C(bx, ay, az) + C(ax, by, az) + C(bx, by, az)) / (U)4; break;
        case 6: val = (c000 + C(ax, by, az) + C(ax, ay, bz) + C(ax, by, bz)) / (U)4; break;
        case 5: val = (c000 + C(bx, ay, az) + C(ax, ay, bz) + C(bx, ay, bz)) / (U)4; break;
        default:
            val = (c000 + C(bx, ay, az) + C(ax, by, az) + C(ax, ay, bz) + C(bx, by, az) +
                   C(bx, ay, bz) + C(ax, by, bz) + C(bx, by, bz)) / (U)8;
            break;
    }
#undef C
    return val;
}

template <typename T, typename U>
__device__ __forceinline__ void cell_prolong(T *__restrict__ fine, const U *__restrict__ v, int cx, int cy, int cz,
                                             int add, int ix, int iy, int iz) {
    const U val = prolong_value<U>(v, cx, cy, cz, ix, iy, iz);
    const int64_t I = ix + (int64_t)(2 * cx) * (iy + (int64_t)(2 * cy) * iz);
    fine[I] = add ? (T)(fine[I] + (T)val) : (T)val;
}

template <typename T, typename U>
__global__ void k_mg_prolong(T *__restrict__ fine, const U *__restrict__ v, int cx, int cy, int cz,
                             int add) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    if (ix >= 2 * cx) return;
    cell_prolong<T, U>(fine, v, cx, cy, cz, add, ix, blockIdx.y, blockIdx.z);
}

// The V-cycle's residual and restriction in one z-marching pass:
// coarse = reduce(f - A v), the arithmetic of k_mg_residual<float> followed by
// k_mg_reduce<float, float> (so the same bits) without writing r.  A
// workgroup owns a 32 x 8 x tzc block of coarse cells; it streams planes of v
// (67 x 19 with the halo) through a ring of 4, computes r on the 65 x 17 fine
// cells the block's 27-point sums read (x0 .. x0 + 64: the + 1 side wraps like
// reduce's ixp) into a ring of 4 r planes, and a coarse plane every second
// step, once its three fine planes are in.  r rows keep even x (33) then odd
// x (32), so the restriction's reads, every second x, are consecutive words.
constexpr int kRrCX = 32, kRrCY = 8, kRrThreads = 256, kRrMaxTzc = 16;
constexpr int kRrVW = 2 * kRrCX + 3, kRrVH = 2 * kRrCY + 3, kRrVPlane = kRrVW * kRrVH;   // 67 x 19
constexpr int kRrW = 2 * kRrCX + 1, kRrH = 2 * kRrCY + 1, kRrPlane = kRrW * kRrH;          // 65 x 17
constexpr int kRrEven = kRrCX + 1;                                                          // 33 even x
constexpr int kRrIters = (kRrPlane + kRrThreads - 1) / kRrThreads;                          // 5
constexpr int kRrLoads = (kRrVPlane + kRrThreads - 1) / kRrThreads;                          // 5

template <int FAST>   // as k_mg_residual_zm
__global__ void __launch_bounds__(kRrThreads)
    k_mg_residual_restrict(float *__restrict__ coarse, const float *__restrict__ v, const float *__restrict__ f,
                           MGLevel L, float beta, int has_los, int tzc) {
    __shared__ float vring[4 * kRrVPlane];
    __shared__ float rring[4 * kRrPlane];
    __shared__ float ptab[kRrVW + kRrVH + 2 * kRrMaxTzc + 3];
    const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    int tile_x, tile_y, tile_z;
    mg_tile(tile_x, tile_y, tile_z);
    const int x0 = 2 * kRrCX * tile_x, y0 = 2 * kRrCY * tile_y, z0 = 2 * tzc * tile_z;   // fine origin
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    const int cx = nx / 2, cy = ny / 2;
    const int tid = threadIdx.x;
    float *pz_tab = ptab + kRrVW + kRrVH;   // local plane q = z - (z0 - 1)
    if (!has_los)
        for (int q = tid; q < kRrVW + kRrVH + 2 * tzc + 3; q += kRrThreads) {
            int g;
            if (q < kRrVW)
                g = gs_wrap(x0 - 1 + q, nx);
            else if (q < kRrVW + kRrVH)
                g = nx + gs_wrap(y0 - 1 + q - kRrVW, ny);
            else
                g = nx + ny + gs_wrap(z0 - 1 + q - kRrVW - kRrVH, nz);
            ptab[q] = L.pv[g];
        }
    // v plane loads: element q = row * 67 + col (threads past the plane reload 0)
    int64_t ld_off[kRrLoads];
#pragma unroll
    for (int u = 0; u < kRrLoads; u++) {
        const int q = tid + kRrThreads * u, qq = q < kRrVPlane ? q : 0;
        ld_off[u] = gs_wrap(x0 - 1 + qq % kRrVW, nx) + sy * gs_wrap(y0 - 1 + qq / kRrVW, ny);
    }
    float pre[kRrLoads];
    const int last_v = z0 + 2 * tzc + 1;
    auto fetch = [&](int P) {
        const int64_t zo = sz * gs_wrap(min(P, last_v), nz);
#pragma unroll
        for (int u = 0; u < kRrLoads; u++) pre[u] = v[zo + ld_off[u]];
    };
    auto store = [&](int P) {
        float *pl = vring + ((P - z0 + 1) & 3) * kRrVPlane;
#pragma unroll
        for (int u = 0; u < kRrLoads; u++) {
            const int q = tid + kRrThreads * u;
            if (q < kRrVPlane) pl[q] = pre[u];
        }
    };
    // r cells: iteration it handles cell q = tid + 256 it of the 65 x 17 plane
    int rpos[kRrIters];
    uint32_t foff[kRrIters];
#pragma unroll
    for (int it = 0; it < kRrIters; it++) {
        const int q = tid + kRrThreads * it;
        const int xr = q < kRrPlane ? q % kRrW : 0, yr = q < kRrPlane ? q / kRrW : 0;
        rpos[it] = q < kRrPlane ? (xr | (yr << 8)) : -1;
        foff[it] = (uint32_t)gs_wrap(x0 + xr, nx) + (uint32_t)nx * gs_wrap(y0 + yr, ny);
    }
    const uint32_t sz32 = (uint32_t)sz;
    float fpre[kRrIters];
    auto fetch_f = [&](int Q) {   // f of the r cells on fine plane Q
        const uint32_t zo = sz32 * (uint32_t)gs_wrap(min(Q, z0 + 2 * tzc), nz);
#pragma unroll
        for (int it = 0; it < kRrIters; it++) fpre[it] = f[foff[it] + zo];
    };
    // coarse cell of this thread: (cx0 + tid % 32, cy0 + tid / 32)
    const int kxl = tid % kRrCX, kyl = tid / kRrCX;
    const int64_t c_off = (x0 / 2 + kxl) + (int64_t)cx * (y0 / 2 + kyl);
    fetch(z0 - 1);
    for (int k = 0; k < 2 * tzc + 4; k++) {
        const int P = z0 - 1 + k;   // v plane loaded this step
        if (P <= last_v) store(P);
        float fcur[kRrIters];
#pragma unroll
        for (int it = 0; it < kRrIters; it++) fcur[it] = fpre[it];
        fetch(P + 1);
        fetch_f(P);   // f of the r plane of the next step
        __syncthreads();
        const int Q = P - 1;   // r plane computed this step
        if (Q >= z0 && Q <= z0 + 2 * tzc) {
            const int vq = Q - z0 + 1;   // local v plane of Q
            const float *pm = vring + ((vq - 1) & 3) * kRrVPlane, *p0 = vring + (vq & 3) * kRrVPlane,
                        *pp = vring + ((vq + 1) & 3) * kRrVPlane;
            float *rp = rring + ((Q - z0) & 3) * kRrPlane;
#pragma unroll
            for (int it = 0; it < kRrIters; it++) {
                if (rpos[it] < 0) continue;
                const int xr = rpos[it] & 255, yr = rpos[it] >> 8;
                const int xl = xr + 1, yl = yr + 1, wc = yl * kRrVW + xl;
                const Coef<float> c = coef_p<float, FAST != 0>(L, has_los ? L.losn[0] : ptab[xl],
                                                               has_los ? L.losn[1] : ptab[kRrVW + yl],
                                                               has_los ? L.losn[2] : pz_tab[vq], beta);
#define A0(o) p0[wc + (o)]
#define AM(o) pm[wc + (o)]
#define AP(o) pp[wc + (o)]
                const Nbr<float> sn = nbr_of<float>(
                    c, A0(1), A0(-1), A0(kRrVW), A0(-kRrVW), AP(0), AM(0), A0(kRrVW + 1), A0(-kRrVW - 1),
                    A0(kRrVW - 1), A0(-kRrVW + 1), AP(1), AM(-1), AP(-1), AM(1), AP(kRrVW), AM(-kRrVW),
                    AP(-kRrVW), AM(kRrVW));
                float t = 2.0f * (c.gpx2 + c.gpy2 + c.gpz2) * A0(0) - (sn.A + sn.B + sn.C + sn.X);
#undef A0
#undef AM
#undef AP
                if (!has_los) t -= sn.G;
                rp[yr * kRrW + (xr & 1) * kRrEven + (xr >> 1)] = fcur[it] - t;
            }
        }
        // coarse plane j once fine r planes 2j .. 2j + 2 (local) are in: Q = z0 + 2j + 3
        const int jq = Q - z0 - 3;
        if (jq >= 0 && (jq & 1) == 0 && (jq >> 1) < tzc) {
            const int j = jq >> 1;
            const float *rm = rring + ((2 * j) & 3) * kRrPlane, *r0 = rring + ((2 * j + 1) & 3) * kRrPlane,
                        *rq = rring + ((2 * j + 2) & 3) * kRrPlane;
            // the centre (local 2 kxl + 1, odd) is word kRrEven + kxl of its row, x -/+ 1
            // (even) words kxl and kxl + 1
            const int iy = 2 * kyl + 1, wc = kRrEven + kxl, we = kxl;
#define RX(a) ((a) == 0 ? wc : we + ((a) > 0))
#define W0(a, b) r0[(iy + (b)) * kRrW + RX(a)]
#define WM(a, b) rm[(iy + (b)) * kRrW + RX(a)]
#define WP(a, b) rq[(iy + (b)) * kRrW + RX(a)]
            const float sum = 8.0f * W0(0, 0) +
                              4.0f * (W0(1, 0) + W0(-1, 0) + W0(0, 1) + W0(0, -1) + WP(0, 0) + WM(0, 0)) +
                              2.0f * (W0(1, 1) + W0(-1, 1) + W0(1, -1) + W0(-1, -1) + WP(1, 0) + WP(-1, 0) +
                                      WM(1, 0) + WM(-1, 0) + WP(0, 1) + WP(0, -1) + WM(0, 1) + WM(0, -1)) +
                              WP(1, 1) + WP(-1, 1) + WP(1, -1) + WP(-1, -1) + WM(1, 1) + WM(-1, 1) + WM(1, -1) +
                              WM(-1, -1);
#undef W0
#undef WM
#undef WP
#undef RX
            coarse[c_off + (int64_t)cx * cy * (z0 / 2 + j)] = (float)((double)sum / 64.0);
        }
    }
}

// The value of prolong_value for a fine cell of parities (EX, EY, EZ) (1:
// even index) from the 2 x 2 x 2 coarse cells V[x][y][z] of its block (index
// 0: the coarse cell below the block's, with the wrap; 1: the block's own):
// an even index reads its "a" cell from 0 and its "b" cell from 1, an odd
// index its "a" cell from 1.  The same sums in the same order.
template <typename U, int EX, int EY, int EZ>
__device__ __forceinline__ U prolong_block_value(const U (&V)[2][2][2]) {
    constexpr int Ax = EX ? 0 : 1, Ay = EY ? 0 : 1, Az = EZ ? 0 : 1;
    constexpr int code = EX + 2 * EY + 4 * EZ;
    const U c000 = V[Ax][Ay][Az];
    if constexpr (code == 0) return c000;
    else if constexpr (code == 1) return (c000 + V[1][Ay][Az]) / (U)2;
    else if constexpr (code == 2) return (c000 + V[Ax][1][Az]) / (U)2;
    else if constexpr (code == 4) return (c000 + V[Ax][Ay][1]) / (U)2;
    else if constexpr (code == 3) return (c000 + V[1][Ay][Az] + V[Ax][1][Az] + V[1][1][Az]) / (U)4;
    else if constexpr (code == 6) return (c000 + V[Ax][1][Az] + V[Ax][Ay][1] + V[Ax][1][1]) / (U)4;
    else if constexpr (code == 5) return (c000 + V[1][Ay][Az] + V[Ax][Ay][1] + V[1][Ay][1]) / (U)4;
    else
        return (c000 + V[1][Ay][Az] + V[Ax][1][Az] + V[Ax][Ay][1] + V[1][1][Az] + V[1][Ay][1] + V[Ax][1][1] +
                V[1][1][1]) / (U)8;
}

// Trilinear prolongation (+add) by 2 x 2 x 2 fine blocks: a thread loads the
// 8 coarse cells its block reads, then reads, adds and writes the block as
// four aligned pairs along x; prolong_value's arithmetic, no branches.
template <typename T, typename U>
__global__ void k_mg_prolong2(T *__restrict__ fine, const U *__restrict__ v, int cx, int cy, int cz, int add) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;   // coarse x of the block
    if (c >= cx) return;
    const int cyb = blockIdx.y, czb = blockIdx.z;
    const int xs[2] = {c == 0 ? cx - 1 : c - 1, c}, ys[2] = {cyb == 0 ? cy - 1 : cyb - 1, cyb},
              zs[2] = {czb == 0 ? cz - 1 : czb - 1, czb};
    const int64_t sy = cx, sz = (int64_t)cx * cy;
    U V[2][2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int k = 0; k < 2; k++) V[i][j][k] = v[xs[i] + sy * ys[j] + sz * zs[k]];
    using T2 = typename std::conditional<sizeof(T) == 4, float2, double2>::type;
    const int64_t fsy = 2 * (int64_t)cx, fsz = fsy * 2 * cy;
    T2 *base = reinterpret_cast<T2 *>(fine + (2 * c + fsy * 2 * cyb + fsz * 2 * czb));   // 2c is even: aligned
    auto row = [&](T2 *p, U a, U b) {
        T2 o;
        if (add) {
            o = *p;
            o.x = (T)(o.x + (T)a);
            o.y = (T)(o.y + (T)b);
        } else {
            o.x = (T)a;
            o.y = (T)b;
        }
        *p = o;
    };
    const int64_t ry = fsy / 2, rz = fsz / 2;   // in pairs
    row(base, prolong_block_value<U, 1, 1, 1>(V), prolong_block_value<U, 0, 1, 1>(V));
    row(base + ry, prolong_block_value<U, 1, 0, 1>(V), prolong_block_value<U, 0, 0, 1>(V));
    row(base + rz, prolong_block_value<U, 1, 1, 0>(V), prolong_block_value<U, 0, 1, 0>(V));
    row(base + ry + rz, prolong_block_value<U, 1, 0, 0>(V), prolong_block_value<U, 0, 0, 0>(V));
}

// ------------------------------------------------------------ coarse levels
//
// The levels below a few thousand cells are latency-bound: one launch per
// sweep costs more than the sweep.  k_mg_coarse runs a whole V-cycle (or the
// whole FMG) over levels [0, nl) of an MGCoarse<T> in one 1024-thread
// workgroup, with a barrier between stages, applying the same per-cell
// operations in the same order as the per-level launches.
template <typename T>
__global__ void __launch_bounds__(1024) k_mg_coarse(MGCoarse<T> a) {
    int cur[kMGCoarseLevels];
    for (int l = 0; l < a.nl; l++) cur[l] = a.cur[l];
    auto vc = [&](int l) { return cur[l] ? a.B[l] : a.A[l]; };
    auto vo = [&](int l) { return cur[l] ? a.A[l] : a.B[l]; };
    // every cell (ix, iy, iz) of level l, then a barrier
    auto forall = [&](int l, auto &&fn) {
        const int nx = a.L[l].n[0], ny = a.L[l].n[1], n = nx * ny * a.L[l].n[2];
        for (int c = threadIdx.x; c < n; c += blockDim.x) fn(c % nx, (c / nx) % ny, c / (nx * ny));
        __syncthreads();
    };
    auto smooth = [&](int l, int n) {
        if constexpr (sizeof(T) == sizeof(float)) {
            if (!a.gs) {
                for (int i = 0; i < n; i++) {
                    T *out = vo(l);
                    const T *in = vc(l);
                    forall(l, [&](int x, int y, int z) {
                        cell_jacobi(out, in, a.F[l], a.L[l], x, y, z, a.beta, a.omega, a.has_los);
                    });
                    cur[l] ^= 1;
                }
                return;
            }
        }
        const int nx = a.L[l].n[0], ny = a.L[l].n[1], nz = a.L[l].n[2], nc = nx * ny * nz;
        if (sizeof(T) == sizeof(double) && nc <= 64) {
            // an FP64 level of at most 64 cells (the RBGS solve's coarsest,
            // with its many sweeps) in the first wave: a cell per lane, its
            // neighbours' values by cross-lane shuffles, no barrier between
            // colour passes; the same per-cell arithmetic as cell_mcgs on the
            // same values
            if (threadIdx.x < 64) {
                const int c = threadIdx.x, cc = c < nc ? c : 0;
                const int ix = cc % nx, iy = (cc / nx) % ny, iz = cc / (nx * ny);
                const int ixp = ix + 1 == nx ? 0 : ix + 1, ixm = ix == 0 ? nx - 1 : ix - 1;
                const int iyp = iy + 1 == ny ? 0 : iy + 1, iym = iy == 0 ? ny - 1 : iy - 1;
                const int izp = iz + 1 == nz ? 0 : iz + 1, izm = iz == 0 ? nz - 1 : iz - 1;
                auto id = [&](int x, int y, int z) { return x + nx * (y + ny * z); };
                T *v = vc(l);
                T val = v[cc];
                const T fv = a.F[l][cc];
                const GsCoef<T> co = gs_coef_at<T>(a.L[l], ix, iy, iz, a.beta, a.has_los);
                const int mine = c < nc ? mg_color(ix, iy, iz) : -1;
#define V(x, y, z) __shfl(val, id(x, y, z), 64)
                for (int i = 0; i < n; i++)
                    for (int col = 0; col < 4; col++) {
                        const T nv = gs_update<T>(co, fv, V(ixp, iy, iz), V(ixm, iy, iz), V(ix, iyp, iz),
                                                  V(ix, iym, iz), V(ix, iy, izp), V(ix, iy, izm), V(ixp, iyp, iz),
                                                  V(ixm, iym, iz), V(ixm, iyp, iz), V(ixp, iym, iz), V(ixp, iy, izp),
                                                  V(ixm, iy, izm), V(ixm, iy, izp), V(ixp, iy, izm), V(ix, iyp, izp),
                                                  V(ix, iym, izm), V(ix, iym, izp), V(ix, iyp, izm), a.has_los);
                        if (mine == col) val = nv;
                    }
#undef V
                if (c < nc) v[c] = val;
            }
            __syncthreads();
            return;
        }
        for (int i = 0; i < n; i++)
            for (int col = 0; col < 4; col++) {
                T *v = vc(l);
                forall(l, [&](int x, int y, int z) {
                    if (mg_color(x, y, z) == col) cell_mcgs<T>(v, a.F[l], a.L[l], x, y, z, a.beta, a.has_los);
                });
            }
    };
    auto sweeps = [&](int l) {
        if (!a.gs) return a.jacobi_n;
        return l + 1 == a.nl ? 16 * a.rb_sweeps : a.rb_sweeps;
    };
    auto vcycle = [&](int top) {
        int l = top;
        for (;; l++) {   // down: pre-smooth, residual, restrict, zero the coarse correction
            smooth(l, sweeps(l));
            if (l + 1 == a.nl) break;
            T *r = vo(l);
            const T *v = vc(l);
            forall(l, [&](int x, int y, int z) { cell_residual<T>(r, v, a.F[l], a.L[l], x, y, z, a.beta, a.has_los); });
            const int *fn = a.L[l].n;
            forall(l + 1, [&](int x, int y, int z) { cell_reduce<T, T>(a.F[l + 1], r, fn[0], fn[1], fn[2], x, y, z); });
            T *vz = vc(l + 1);
            const int nx1 = a.L[l + 1].n[0], ny1 = a.L[l + 1].n[1];
            forall(l + 1, [&](int x, int y, int z) { vz[x + nx1 * (y + ny1 * z)] = (T)0; });
        }
        smooth(l, sweeps(l));   // coarsest: post-smoothing right away
        for (l = l - 1; l >= top; l--) {   // up: prolong and add, post-smooth
            T *v = vc(l);
            const T *vcoarse = vc(l + 1);
            const int *cn = a.L[l + 1].n;
            forall(l, [&](int x, int y, int z) { cell_prolong<T, T>(v, vcoarse, cn[0], cn[1], cn[2], 1, x, y, z); });
            smooth(l, sweeps(l));
        }
    };
    if (!a.fmg) {
        vcycle(0);
        return;
    }
    // fmg: restrict f down the levels, zero the coarsest, then on the way up
    // prolong (overwrite) and run vcycles V-cycles per level
    for (int l = 0; l + 1 < a.nl; l++) {
        const int *fn = a.L[l].n;
        const T *fl = a.F[l];
        forall(l + 1, [&](int x, int y, int z) { cell_reduce<T, T>(a.F[l + 1], fl, fn[0], fn[1], fn[2], x, y, z); });
    }
    {
        T *vz = vc(a.nl - 1);
        const int nx1 = a.L[a.nl - 1].n[0], ny1 = a.L[a.nl - 1].n[1];
        forall(a.nl - 1, [&](int x, int y, int z) { vz[x + nx1 * (y + ny1 * z)] = (T)0; });
    }
    for (int l = a.nl - 1; l >= 0; l--) {
        if (l + 1 < a.nl) {
            T *v = vc(l);
            const T *vcoarse = vc(l + 1);
            const int *cn = a.L[l + 1].n;
            forall(l, [&](int x, int y, int z) { cell_prolong<T, T>(v, vcoarse, cn[0], cn[1], cn[2], 0, x, y, z); });
        }
        for (int i = 0; i < a.vcycles; i++) vcycle(l);
    }
}

// Cell centres of every level (x_vec per level, src/multigrid.jl:656).
__global__ void k_mg_xvec(float *xv, const Slot *slot, int nx, int ny, int nz) {
    const int n3[3] = {nx, ny, nz};
    const int tot = nx + ny + nz;
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < tot; idx += gridDim.x * blockDim.x) {
        int ax = idx < nx ? 0 : (idx < nx + ny ? 1 : 2);
        int i = idx - (ax == 0 ? 0 : (ax == 1 ? nx : nx + ny));
        const float h = slot->box_size[ax] / (float)n3[ax];
        const float x = (float)((double)slot->box_min[ax] + ((double)i + 0.5) * (double)h);
        xv[idx] = x;
        xv[tot + idx] = x / h;
    }
}

// FP64 reductions of a grid in a fixed order, so repeated runs agree bit for
// bit: each block sums a fixed grid-stride share (thread, then wave, then the
// four waves in order) into part[blockIdx.x]; launch_reduce then sums the
// partials with k_sum_partials (fixed order as well).
template <int OP>
__device__ __forceinline__ double red_term(double a, double b) {
    const double v = OP == 2 ? a - b : a;
    return OP == 0 ? v : v * v;
}

template <int OP, typename T>
__global__ void __launch_bounds__(256) k_reduce_part(const T *a, const T *b, int64_t n, double *part) {
    __shared__ double red[4];
    double s = 0.0;
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    bool vec = false;
    if constexpr (sizeof(T) == 4)   // float4 loads, two in flight: thread t sums chunks t, t + stride, ...
        vec = (n & 3) == 0 && ((uintptr_t)a & 15) == 0 && (OP != 2 || ((uintptr_t)b & 15) == 0);
    if (vec) {
        const float4 *a4 = reinterpret_cast<const float4 *>(a), *b4 = reinterpret_cast<const float4 *>(b);
        const int64_t n4 = n >> 2;
        double s1 = 0.0;
        int64_t i = t0;
        for (; i + stride < n4; i += 2 * stride) {
            const float4 x = a4[i], y = a4[i + stride];
            const float4 u = OP == 2 ? b4[i] : x, w = OP == 2 ? b4[i + stride] : y;
            s += red_term<OP>(x.x, u.x) + red_term<OP>(x.y, u.y) + red_term<OP>(x.z, u.z) + red_term<OP>(x.w, u.w);
            s1 += red_term<OP>(y.x, w.x) + red_term<OP>(y.y, w.y) + red_term<OP>(y.z, w.z) + red_term<OP>(y.w, w.w);
        }
        if (i < n4) {
            const float4 x = a4[i], u = OP == 2 ? b4[i] : x;
            s += red_term<OP>(x.x, u.x) + red_term<OP>(x.y, u.y) + red_term<OP>(x.z, u.z) + red_term<OP>(x.w, u.w);
        }
        s += s1;
    } else {
        for (int64_t i = t0; i < n; i += stride)
            s += red_term<OP>((double)a[i], OP == 2 ? (double)b[i] : 0.0);
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// a += sign * (*num) / den  (mean removal with a device-side FP64 total)
template <typename T>
__global__ void k_add_const(T *a, int64_t n, const double *num, double den, double sign) {
    const T c = (T)(sign * (*num) / den);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        a[i] += c;
}

template <typename T, typename U>
__global__ void k_convert(T *dst, const U *src, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = (T)src[i];
}

__global__ void k_fill(float *a, int64_t n, float v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        a[i] = v;
}

// The same two on 16-byte aligned arrays, four elements per thread (16- or
// 32-byte accesses, one pass without a loop): n4 = n / 4 groups, the tail
// (n % 4) in the first threads.
__global__ void k_convert4_f64_f32(float *__restrict__ dst, const double *__restrict__ src, int64_t n4, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n4) {
        const double2 a = reinterpret_cast<const double2 *>(src)[2 * i];
        const double2 b = reinterpret_cast<const double2 *>(src)[2 * i + 1];
        reinterpret_cast<float4 *>(dst)[i] = make_float4((float)a.x, (float)a.y, (float)b.x, (float)b.y);
    }
    if (i < n - 4 * n4) dst[4 * n4 + i] = (float)src[4 * n4 + i];
}
__global__ void k_fill4(float *a, int64_t n4, int64_t n, float v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n4) reinterpret_cast<float4 *>(a)[i] = make_float4(v, v, v, v);
    if (i < n - 4 * n4) a[4 * n4 + i] = v;
}

// ------------------------------------------------------------ launchers

// Workgroups of a kernel resident on one CU (occupancy query, once per
// kernel) and the CU count of the current device.
template <typename K>
static int blocks_per_cu(K kernel, int threads) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, threads, 0) != hipSuccess || n < 1) n = 1;
    return n;
}
static int device_cus() {
    static const int n = [] {
        int d = 0, c = 0;
        if (hipGetDevice(&d) != hipSuccess || hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, d) !=
                                                  hipSuccess)
            c = 0;
        return c > 0 ? c : 256;
    }();
    return n;
}

// z planes per workgroup of a z-marching kernel over `cols` tile columns of
// nz planes: the power of two in [min_tz, max_tz] dividing nz that minimises
// (rounds of resident workgroups) x (tz + the pipeline's extra steps), so
// that a launch neither ends on a thinly occupied last round nor pays the
// halo planes of short columns too often.
static int zm_tz(int64_t cols, int nz, int max_tz, int min_tz, int per_cu, int extra) {
    const int64_t slots = (int64_t)device_cus() * per_cu;
    int best = min_tz;
    int64_t best_cost = -1;
    for (int tz = max_tz; tz >= min_tz; tz /= 2) {
        if (nz % tz) continue;
        const int64_t rounds = (cols * (nz / tz) + slots - 1) / slots, cost = rounds * (tz + extra);
        if (best_cost < 0 || cost < best_cost) {
            best_cost = cost;
            best = tz;
        }
    }
    return best;
}

// z planes per workgroup of k_mg_refine (tz + 2 steps)
static int refine_tz(const MGLevel &L) {
    static const int per_cu = blocks_per_cu(k_mg_refine<0, kRfPhiE>, kRfThreads);
    return zm_tz((int64_t)(L.n[0] / kRfTX) * (L.n[1] / kRfTY), L.n[2], kRfMaxTz, 8, per_cu, 2);
}

// the same for k_mg_residual_zm
static int residual_zm_tz(const MGLevel &L) {
    static const int per_cu = blocks_per_cu(k_mg_residual_zm<0>, kRfThreads);
    return zm_tz((int64_t)(L.n[0] / kRfTX) * (L.n[1] / kRfTY), L.n[2], kRfMaxTz, 8, per_cu, 2);
}

bool mg_refine_ok(const MGLevel &L) { return L.n[0] % kRfTX == 0 && L.n[1] % kRfTY == 0 && L.n[2] % 8 == 0; }


static inline dim3 cell_grid(int nx, int ny, int nz, int bx) {
    return dim3((unsigned)((nx + bx - 1) / bx), (unsigned)ny, (unsigned)nz);
}
static inline int bx_for(int nx) { return nx >= 256 ? 256 : (nx >= 64 ? 64 : nx); }
// Stencil kernels: workgroups of 64 x 4 cells of one plane.  Four rows per
// workgroup share their y-neighbour rows in the CU's cache; at 512^3 this ran
// the RBGS lightcone solve in 64.0 ms against 67.0 with 256 x 1 (128 x 4,
// 256 x 4 and 32 x 8 were slower).
static inline void stencil_shape(const MGLevel &L, dim3 &grid, dim3 &block) {
    const int by = 4 < L.n[1] ? 4 : L.n[1];
    const int bx = L.n[0] >= 64 ? 64 : L.n[0];
    block = dim3((unsigned)bx, (unsigned)by);
    grid = dim3((unsigned)((L.n[0] + bx - 1) / bx), (unsigned)((L.n[1] + by - 1) / by), (unsigned)L.n[2]);
}
static inline unsigned red_blocks(int64_t n) {
    int64_t b = (n + 255) / 256;
    return (unsigned)(b > 1024 ? 1024 : (b < 1 ? 1 : b));
}

hipError_t launch_mg_jacobi(float *vout, const float *vin, const float *f, const MGLevel &L,
                            float beta, float omega, int has_los, hipStream_t s) {
    dim3 grid, block;
    stencil_shape(L, grid, block);
    hipLaunchKernelGGL(k_mg_jacobi, grid, block, 0, s, vout, vin, f, L, beta, omega, has_los);
    return hipGetLastError();
}

hipError_t launch_mg_residual(float *r, const float *v, const float *f, const MGLevel &L,
                              float beta, int has_los, hipStream_t s) {
    return launch_mg_residual_t<float>(r, v, f, L, beta, has_los, s);
}

hipError_t launch_mg_residual_fast(float *r, const float *v, const float *f, const MGLevel &L, float beta,
                                   int has_los, hipStream_t s) {
    if (mg_refine_ok(L) && (int64_t)L.n[0] * L.n[1] * L.n[2] >= ((int64_t)1 << 21)) {
        const int tz = residual_zm_tz(L);
        const dim3 grid((unsigned)(L.n[0] / kRfTX), (unsigned)(L.n[1] / kRfTY), (unsigned)(L.n[2] / tz));
        hipLaunchKernelGGL(k_mg_residual_zm<1>, grid, dim3(kRfThreads), 0, s, r, v, f, L, beta, has_los, tz);
        return hipGetLastError();
    }
    return launch_mg_residual_t<float>(r, v, f, L, beta, has_los, s);
}

template <typename T>
hipError_t launch_mg_residual_t(T *r, const T *v, const T *f, const MGLevel &L, float beta,
                                int has_los, hipStream_t s) {
    if constexpr (sizeof(T) == 4) {   // z-marching on grids of >= 128^3 cells
        if (mg_refine_ok(L) && (int64_t)L.n[0] * L.n[1] * L.n[2] >= ((int64_t)1 << 21)) {
            const int tz = residual_zm_tz(L);
            const dim3 grid((unsigned)(L.n[0] / kRfTX), (unsigned)(L.n[1] / kRfTY), (unsigned)(L.n[2] / tz));
            hipLaunchKernelGGL(k_mg_residual_zm<0>, grid, dim3(kRfThreads), 0, s, r, v, f, L, beta, has_los, tz);
            return hipGetLastError();
        }
    }
    dim3 grid, block;
    stencil_shape(L, grid, block);
    hipLaunchKernelGGL(k_mg_residual<T>, grid, block, 0, s, r, v, f, L, beta, has_los);
    return hipGetLastError();
}

// coarse planes per workgroup of k_mg_residual_restrict (2 tzc + 4 steps)
static int rr_tzc(const MGLevel &L) {
    static const int per_cu = blocks_per_cu(k_mg_residual_restrict<0>, kRrThreads);
    return zm_tz((int64_t)(L.n[0] / (2 * kRrCX)) * (L.n[1] / (2 * kRrCY)), L.n[2] / 2, kRrMaxTzc, 1, per_cu, 2);
}

bool mg_residual_restrict_ok(const MGLevel &L) {
    return L.n[0] % (2 * kRrCX) == 0 && L.n[1] % (2 * kRrCY) == 0 && L.n[2] % 2 == 0 &&
           (int64_t)L.n[0] * L.n[1] * L.n[2] < ((int64_t)1 << 31);
}

hipError_t launch_mg_residual_restrict(float *coarse, const float *v, const float *f, const MGLevel &L, float beta,
                                       int has_los, hipStream_t s, int fast) {
    if (!mg_residual_restrict_ok(L)) return hipErrorInvalidValue;
    const int tzc = rr_tzc(L);
    const dim3 grid((unsigned)(L.n[0] / (2 * kRrCX)), (unsigned)(L.n[1] / (2 * kRrCY)), (unsigned)(L.n[2] / 2 / tzc));
    hipLaunchKernelGGL(fast ? k_mg_residual_restrict<1> : k_mg_residual_restrict<0>, grid, dim3(kRrThreads), 0, s,
                       coarse, v, f, L, beta, has_los, tzc);
    return hipGetLastError();
}

// z planes per workgroup of k_mg_gs4 (at least 8; each workgroup also loads
// 8 halo planes and runs 12 pipeline steps beyond tz)
static int gs4_tz(const MGLevel &L) {
    static const int per_cu = blocks_per_cu(k_mg_gs4<0, 1, 0>, kGsThreads);
    return zm_tz((int64_t)(L.n[0] / kGsTX) * (L.n[1] / kGsTY), L.n[2], kGsMaxTz, 8, per_cu, 12);
}

bool mg_gs4_ok(const MGLevel &L) {
    return L.n[0] % kGsTX == 0 && L.n[1] % kGsTY == 0 && L.n[2] % 8 == 0 &&
           (int64_t)L.n[0] * L.n[1] * L.n[2] < ((int64_t)1 << 31);
}

int64_t mg_gs4_blocks(const MGLevel &L) {
    return (int64_t)(L.n[0] / kGsTX) * (L.n[1] / kGsTY) * (L.n[2] / gs4_tz(L));
}

hipError_t launch_mg_gs4(float *vout, const float *vin, const float *f, const MGLevel &L, float beta, int has_los,
                         hipStream_t s, double *part) {
    if (!mg_gs4_ok(L) || vout == vin || !vout) return hipErrorInvalidValue;
    const int tz = gs4_tz(L);
    const dim3 grid((unsigned)(L.n[0] / kGsTX), (unsigned)(L.n[1] / kGsTY), (unsigned)(L.n[2] / tz));
    using K = decltype(&k_mg_gs4<0, 0, 0>);
    static const K kernels[8] = {k_mg_gs4<0, 0, 0>, k_mg_gs4<0, 0, 1>, k_mg_gs4<0, 1, 0>, k_mg_gs4<0, 1, 1>,
                                 k_mg_gs4<1, 0, 0>, k_mg_gs4<1, 0, 1>, k_mg_gs4<1, 1, 0>, k_mg_gs4<1, 1, 1>};
    const K kernel = kernels[(has_los ? 4 : 0) + (part ? 2 : 0) + (vin ? 0 : 1)];
    hipLaunchKernelGGL(kernel, grid, dim3(kGsThreads), 0, s, vout, vin, f, L, beta, tz, part);
    return hipGetLastError();
}

// z planes per workgroup of k_mg_jacobi_tb<S> (tz + 3 S - 1 steps)
template <int S>
static int jacobi_tb_tz(const MGLevel &L) {
    static const int per_cu = blocks_per_cu(k_mg_jacobi_tb<S, 0>, JbShape<S>::Threads);
    return zm_tz((int64_t)(L.n[0] / kJbTX) * (L.n[1] / kJbTY), L.n[2], kJbMaxTz, 8, per_cu, 3 * S - 1);
}

bool mg_jacobi_tb_ok(const MGLevel &L) {
    return L.n[0] % kJbTX == 0 && L.n[1] % kJbTY == 0 && L.n[2] % 8 == 0 && L.n[2] >= 8 &&
           (int64_t)L.n[0] * L.n[1] * L.n[2] < ((int64_t)1 << 31);
}

hipError_t launch_mg_jacobi_tb(float *vout, const float *vin, const float *f, const MGLevel &L, float beta,
                               float omega, int has_los, int sweeps, hipStream_t s) {
    if (!mg_jacobi_tb_ok(L) || vout == vin || (sweeps != 2 && sweeps != 3)) return hipErrorInvalidValue;
    const int tz = sweeps == 2 ? jacobi_tb_tz<2>(L) : jacobi_tb_tz<3>(L);
    const dim3 grid((unsigned)(L.n[0] / kJbTX), (unsigned)(L.n[1] / kJbTY), (unsigned)(L.n[2] / tz));
    if (sweeps == 2)
        hipLaunchKernelGGL((has_los ? k_mg_jacobi_tb<2, 1> : k_mg_jacobi_tb<2, 0>), grid, dim3(JbShape<2>::Threads), 0,
                           s, vout, vin, f, L, beta, omega, tz);
    else
        hipLaunchKernelGGL((has_los ? k_mg_jacobi_tb<3, 1> : k_mg_jacobi_tb<3, 0>), grid, dim3(JbShape<3>::Threads), 0,
                           s, vout, vin, f, L, beta, omega, tz);
    return hipGetLastError();
}

int64_t mg_refine_blocks(const MGLevel &L) {
    return (int64_t)(L.n[0] / kRfTX) * (L.n[1] / kRfTY) * (L.n[2] / refine_tz(L));
}

hipError_t launch_mg_refine(double *phi_out, const double *phi_in, const float *seed, const float *e,
                            const double *esum, const float *f, const double *fsum, float *r, double *part,
                            const MGLevel &L, float beta, int has_los, hipStream_t s) {
    if (!mg_refine_ok(L) || ((e || seed || !phi_in) && phi_out == phi_in) || (seed && e) || (e && !phi_in))
        return hipErrorInvalidValue;
    const int tz = refine_tz(L);
    const dim3 grid((unsigned)(L.n[0] / kRfTX), (unsigned)(L.n[1] / kRfTY), (unsigned)(L.n[2] / tz));
    using K = decltype(&k_mg_refine<0, kRfPhi>);
    static const K kernels[2][4] = {
        {k_mg_refine<0, kRfPhi>, k_mg_refine<0, kRfPhiE>, k_mg_refine<0, kRfSeed>, k_mg_refine<0, kRfZero>},
        {k_mg_refine<1, kRfPhi>, k_mg_refine<1, kRfPhiE>, k_mg_refine<1, kRfSeed>, k_mg_refine<1, kRfZero>}};
    const int mode = seed ? kRfSeed : (e ? kRfPhiE : (phi_in ? kRfPhi : kRfZero));
    hipLaunchKernelGGL(kernels[has_los ? 1 : 0][mode], grid, dim3(kRfThreads), 0, s, phi_out, phi_in, seed, e, esum,
                       f, fsum, r, part, L, beta, tz);
    return hipGetLastError();
}

hipError_t launch_shift_norms(float *r, const float *rprev, int64_t n, const double *rsum, double *out,
                              double *part, hipStream_t s) {
    const unsigned nb = red_blocks(n);
    hipLaunchKernelGGL(k_shift_norms, dim3(nb), dim3(256), 0, s, r, rprev, n, rsum, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_sum_partials(part, nb, rprev ? 2 : 1, out, 1, s);
}

template <typename T>
hipError_t launch_mg_rbgs(T *v, const T *f, const MGLevel &L, float beta, int has_los, int color,
                          hipStream_t s) {
    dim3 grid, block;
    stencil_shape(L, grid, block);
    hipLaunchKernelGGL(k_mg_mcgs<T>, grid, block, 0, s, v, f, L, beta, has_los, color);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_mg_coarse(const MGCoarse<T> &a, hipStream_t s) {
    hipLaunchKernelGGL(k_mg_coarse<T>, dim3(1), dim3(1024), 0, s, a);
    return hipGetLastError();
}
template hipError_t launch_mg_coarse<float>(const MGCoarse<float> &, hipStream_t);
template hipError_t launch_mg_coarse<double>(const MGCoarse<double> &, hipStream_t);

hipError_t launch_mg_reduce(float *coarse, const float *fine, const int fine_n[3], hipStream_t s) {
    return launch_mg_reduce_t<float, float>(coarse, fine, fine_n, s);
}

template <typename T, typename U>
hipError_t launch_mg_reduce_t(T *coarse, const U *fine, const int fine_n[3], hipStream_t s) {
    const int cx = fine_n[0] / 2, cy = fine_n[1] / 2, cz = fine_n[2] / 2;
    const int bx = bx_for(cx);
    hipLaunchKernelGGL((k_mg_reduce<T, U>), cell_grid(cx, cy, cz, bx), dim3(bx), 0, s, coarse, fine,
                       fine_n[0], fine_n[1], fine_n[2]);
    return hipGetLastError();
}

hipError_t launch_mg_prolong(float *fine, const float *coarse, const int coarse_n[3], int add,
                             hipStream_t s) {
    return launch_mg_prolong_t<float, float>(fine, coarse, coarse_n, add, s);
}

template <typename T, typename U>
hipError_t launch_mg_prolong_t(T *fine, const U *coarse, const int coarse_n[3], int add,
                               hipStream_t s) {
    const int fx = 2 * coarse_n[0], fy = 2 * coarse_n[1], fz = 2 * coarse_n[2];
    if ((int64_t)fx * fy * fz >= ((int64_t)1 << 21) &&
        ((uintptr_t)fine % (2 * sizeof(T))) == 0) {   // 2 x 2 x 2 fine blocks, aligned pairs
        const int b2 = bx_for(coarse_n[0]);
        hipLaunchKernelGGL((k_mg_prolong2<T, U>), dim3((unsigned)((coarse_n[0] + b2 - 1) / b2),
                                                       (unsigned)coarse_n[1], (unsigned)coarse_n[2]),
                           dim3(b2), 0, s, fine, coarse, coarse_n[0], coarse_n[1], coarse_n[2], add);
        return hipGetLastError();
    }
    const int bx = bx_for(fx);
    hipLaunchKernelGGL((k_mg_prolong<T, U>), cell_grid(fx, fy, fz, bx), dim3(bx), 0, s, fine, coarse,
                       coarse_n[0], coarse_n[1], coarse_n[2], add);
    return hipGetLastError();
}

hipError_t launch_mg_xvec(float *xv, const Slot *slot, int nx, int ny, int nz, hipStream_t s) {
    hipLaunchKernelGGL(k_mg_xvec, dim3(4), dim3(256), 0, s, xv, slot, nx, ny, nz);
    return hipGetLastError();
}

template <int OP, typename T>
static hipError_t launch_reduce(const T *a, const T *b, int64_t n, double *out, double *part,
                                hipStream_t s) {
    const unsigned nb = red_blocks(n);
    hipLaunchKernelGGL((k_reduce_part<OP, T>), dim3(nb), dim3(256), 0, s, a, b, n, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_sum_partials(part, nb, 1, out, 1, s);
}
template <typename T>
hipError_t launch_sumsq(const T *a, int64_t n, double *out, double *part, hipStream_t s) {
    return launch_reduce<1, T>(a, nullptr, n, out, part, s);
}
template <typename T>
hipError_t launch_sumsq_diff(const T *a, const T *b, int64_t n, double *out, double *part, hipStream_t s) {
    return launch_reduce<2, T>(a, b, n, out, part, s);
}
template <typename T>
hipError_t launch_sum(const T *a, int64_t n, double *out, double *part, hipStream_t s) {
    return launch_reduce<0, T>(a, nullptr, n, out, part, s);
}
__global__ void k_accumulate(double *a, const double *b) {
    if (threadIdx.x == 0) *a += *b;
}
hipError_t launch_accumulate(double *a, const double *b, hipStream_t s) {
    hipLaunchKernelGGL(k_accumulate, dim3(1), dim3(64), 0, s, a, b);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_add_const(T *a, int64_t n, const double *num, double den, double sign,
                            hipStream_t s) {
    hipLaunchKernelGGL(k_add_const<T>, dim3(red_blocks(n)), dim3(256), 0, s, a, n, num, den, sign);
    return hipGetLastError();
}
template <typename T, typename U>
hipError_t launch_convert(T *dst, const U *src, int64_t n, hipStream_t s) {
    if constexpr (std::is_same<T, float>::value && std::is_same<U, double>::value) {
        const int64_t n4 = n / 4, nb = (n4 + 255) / 256;
        if (((uintptr_t)dst & 15) == 0 && ((uintptr_t)src & 15) == 0 && nb > 0 && nb < ((int64_t)1 << 31)) {
            hipLaunchKernelGGL(k_convert4_f64_f32, dim3((unsigned)nb), dim3(256), 0, s, dst, src, n4, n);
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL((k_convert<T, U>), dim3(red_blocks(n)), dim3(256), 0, s, dst, src, n);
    return hipGetLastError();
}
hipError_t launch_mg_residual_mixed(float *r, const double *v, const float *f, const double *fsum, double count,
                                    const MGLevel &L, float beta, int has_los, hipStream_t s) {
    dim3 grid, block;
    stencil_shape(L, grid, block);
    hipLaunchKernelGGL(k_mg_residual_mixed, grid, block, 0, s, r, v, f, fsum, count, L, beta, has_los);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_add_correction(T *phi, const float *e, int64_t n, const double *esum, hipStream_t s) {
    hipLaunchKernelGGL(k_add_correction<T>, dim3(red_blocks(n)), dim3(256), 0, s, phi, e, n, esum);
    return hipGetLastError();
}
template hipError_t launch_add_correction<float>(float *, const float *, int64_t, const double *, hipStream_t);
template hipError_t launch_add_correction<double>(double *, const float *, int64_t, const double *, hipStream_t);

hipError_t launch_fill(float *a, int64_t n, float v, hipStream_t s) {
    const int64_t n4 = n / 4, nb = (n4 + 255) / 256;
    if (((uintptr_t)a & 15) == 0 && nb > 0 && nb < ((int64_t)1 << 31)) {
        hipLaunchKernelGGL(k_fill4, dim3((unsigned)nb), dim3(256), 0, s, a, n4, n, v);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_fill, dim3(red_blocks(n)), dim3(256), 0, s, a, n, v);
    return hipGetLastError();
}

template hipError_t launch_mg_residual_t<float>(float *, const float *, const float *, const MGLevel &, float, int, hipStream_t);
template hipError_t launch_mg_residual_t<double>(double *, const double *, const double *, const MGLevel &, float, int, hipStream_t);
template hipError_t launch_mg_rbgs<float>(float *, const float *, const MGLevel &, float, int, int, hipStream_t);
template hipError_t launch_mg_rbgs<double>(double *, const double *, const MGLevel &, float, int, int, hipStream_t);
template hipError_t launch_mg_reduce_t<float, float>(float *, const float *, const int *, hipStream_t);
template hipError_t launch_mg_reduce_t<double, float>(double *, const float *, const int *, hipStream_t);
template hipError_t launch_mg_reduce_t<double, double>(double *, const double *, const int *, hipStream_t);
template hipError_t launch_mg_prolong_t<float, float>(float *, const float *, const int *, int, hipStream_t);
template hipError_t launch_mg_prolong_t<float, double>(float *, const double *, const int *, int, hipStream_t);
template hipError_t launch_mg_prolong_t<double, double>(double *, const double *, const int *, int, hipStream_t);
template hipError_t launch_sumsq<float>(const float *, int64_t, double *, double *, hipStream_t);
template hipError_t launch_sumsq<double>(const double *, int64_t, double *, double *, hipStream_t);
template hipError_t launch_sumsq_diff<float>(const float *, const float *, int64_t, double *, double *, hipStream_t);
template hipError_t launch_sum<float>(const float *, int64_t, double *, double *, hipStream_t);
template hipError_t launch_sum<double>(const double *, int64_t, double *, double *, hipStream_t);
template hipError_t launch_add_const<float>(float *, int64_t, const double *, double, double, hipStream_t);
template hipError_t launch_add_const<double>(double *, int64_t, const double *, double, double, hipStream_t);
template hipError_t launch_convert<double, float>(double *, const float *, int64_t, hipStream_t);
template hipError_t launch_convert<float, double>(float *, const double *, int64_t, hipStream_t);

}  // namespace baorec
