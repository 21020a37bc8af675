// Host-side launch interfaces of the HIP kernels (internal to libbaorec_hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "common.hpp"

namespace baorec {

// ------------------------------------------------------------------- FFT

// k-space multipliers applied between the forward and inverse transform of
// a fused strided pass.
enum MulKind : int {
    MUL_NONE = 0,
    MUL_GAUSS = 1,      // exp(-R^2 k^2 / 2), exponent in FP64 (src/utils.jl:52,81)
    MUL_INVK2 = 2,      // 1/k^2, 0 at k = 0 (src/iterative.jl:10-14,145-149)
    MUL_FIXEDLOS = 3,   // (sum_i k_i^2 los_i)/k^2 (src/iterative.jl:43-62)
    MUL_I_INVK2 = 4,    // i/k^2, 0 at k = 0 (src/iterative.jl:219-223)
    MUL_I = 5,          // i (src/multigrid.jl:776-779)
    MUL_MG_INV = 6,     // 1 / symbol of the fixed-LOS multigrid operator, 0 at k = 0
    MUL_GAUSS1 = 7,     // exp(-R^2 k_a^2 / 2) of the pass axis a only, exponent in FP64 (the
                        // Gaussian's factor along a: smoothing one axis at a time)
};

constexpr int kMaxJobs = 3;
constexpr int kMaxOut = 3;

struct StridedJob {
    const float2 *in;
    float2 *out[kMaxOut];
    int pw[kMaxOut];   // power of the along-axis wavenumber for each output
    int nout;
};

struct StridedArgs {
    StridedJob job[kMaxJobs];  // blockIdx.y
    int64_t cstride;           // complex elements between batch items (blockIdx.z)
    const float *ktab;         // per batch item: kx[P] | ky[ny] | kz[nz]
    int64_t ktabStride;        // floats between batch items (0: shared)
    int P;                     // complex pitch of a kx row (>= Pv, multiple of 16)
    int64_t PZ;                // complex pitch of a kz plane (P * (ny + 1))
    int Pv;                    // valid kx columns, nx/2 + 1
    int kyo;                   // offset of ky in ktab (nx/2 + 1)
    int ny, nz;
    int axis;                  // 1: y, 2: z
    int mode;                  // 0 forward, 1 inverse, 2 forward*mul*inverse
    int mul;                   // MulKind (mode 2)
    float R2;                  // smoothing radius squared (FP32)
    float los[3];
    int pair;                  // kernel of single-output jobs: -1 the library's rule,
                               // 0 one column per thread, 1 two columns per thread
    // MUL_MG_INV: symbol sum_a gp[a] (2 - 2 cos t_a) + sum_{a<b} cross[ab]
    // sin t_a sin t_b with t_a = k_a cell[a] (cross: xy, xz, yz)
    float mg_gp[3], mg_cross[3], mg_cell[3];
};

enum XEpi : int {
    XE_PLAIN = 0,     // out = scale * c2r(in0)
    XE_OD_BOX = 1,    // delta = ((rho/mean) - 1)/bias (src/recon.jl:54-55)
    XE_OD_RAND = 2,   // data - alpha*randoms normalisation (src/recon.jl:78-85)
    XE_ITER_LOCAL = 3,// delta_r = delta_s - (x^2 a + x b + c) / x^2 (inputs A, B, C)
    XE_ITER_FIXED = 4,// delta_r = delta_s - fac * D
    XE_DISP = 5,      // (Psi_x, Psi_y, Psi_z) per cell, 12 bytes
};

constexpr int kMaxXIn = 6;

struct XInvArgs {
    const float2 *in[kMaxXIn];
    // real input lines instead of spectra (XE_PLAIN, XE_OD_BOX, XE_OD_RAND):
    // input q is the R2C along x of rin[q], times exp(-R2 kx^2 / 2)
    const float *rin[2];
    float R2;
    int kxpow[kMaxXIn];
    int nin;
    int64_t cstride;      // complex elements between batch items
    int64_t rstride;      // real elements between batch items
    int P;                // complex pitch of a kx row
    int64_t PZ;           // complex pitch of a kz plane
    float *out;           // real output (3 floats per cell for XE_DISP)
    const float *aux;     // delta_s for the iterative updates
    int nx, ny, nz;
    float scale;          // 1/(nx ny nz)
    const float *ktab;    // kx at ktab[b*ktabStride + kx]
    int64_t ktabStride;
    const float *xtab;    // per batch item: x[nx] | y[ny] | z[nz] cell centres
    int64_t xtabStride;
    const Slot *slot;     // per batch item scalars (sums, counts)
    float bias;
    float fac[kMaxXIn];   // update factors (iterative)
    float ran_min;
    float2 *fwd_out;      // optional: R2C of the result line written here (pitch P),
                          // so the next transform skips reading the real grid
    int accumulate;       // XE_PLAIN: out += result instead of out = result
};

struct XFwdArgs {
    const float *in;
    float2 *out;
    int64_t rstride, cstride;
    int P;                // complex pitch of a kx row
    int64_t PZ;           // complex pitch of a kz plane
    int nx, ny, nz;
};
// out[b * stride] = sum of part[b * n .. b * n + n) in a fixed order, b < nb.
hipError_t launch_sum_partials(const double *part, int64_t n, int nb, double *out, int64_t stride, hipStream_t s);
// Half spectrum between the padded working layout (row pitch PP, plane pitch
// PZ) and the compact P x ny x nz layout: to_compact != 0 packs src (padded)
// into dst (compact), else unpacks.
hipError_t launch_spectrum_pack(float2 *dst, const float2 *src, int P, int PP, int64_t PZ, int ny, int nz,
                                int to_compact, hipStream_t s);

// Local-LOS iteration, y inverse pass: from G_m = IFFT_z(k_z^m F) (m = 0,1,2)
// it forms, per (kx, y, z), the three spectra the x pass needs
//   A = fxx IFFT_y(G0)
//   B = fxy y IFFT_y(ky G0) + fxz z IFFT_y(G1)
//   C = fyy y^2 IFFT_y(ky^2 G0) + fyz y z IFFT_y(ky G1) + fzz z^2 IFFT_y(G2)
// (y, z: cell centres), so that delta_r = delta_s - (x^2 a + x b + c)/x^2
// with a, b, c the C2R of kx^2 A, kx B and C.
struct YIterArgs {
    const float2 *g[3];
    float2 *out[3];
    int64_t cstride;
    const float *ktab;     // kx | ky | kz (per batch item)
    int64_t ktabStride;
    const float *xtab;     // x | y | z cell centres (per batch item)
    int64_t xtabStride;
    int P, Pv, nx, ny, nz;
    int64_t PZ;
    float fac[6];          // xx, xy, xz, yy, yz, zz
};
hipError_t launch_y_iter(const YIterArgs &a, int nbatch, hipStream_t s);

hipError_t launch_x_r2c(const XFwdArgs &a, int nbatch, hipStream_t s);
hipError_t launch_x_c2r(const XInvArgs &a, int epi, int nbatch, hipStream_t s);
hipError_t launch_strided(const StridedArgs &a, int njobs, int nbatch, hipStream_t s);
// True when launch_strided runs the pass with two columns per thread
// (k_strided2: lines of 256..1024, every job with one output).
bool strided_uses_pair(const StridedArgs &a, int njobs);
bool fft_side_ok(int n);

// ---------------------------------------------------------------- tables

// Builds, for each batch item, the k tables (kx[P] | ky[ny] | kz[nz]) and the
// cell-centre tables (x[nx] | y[ny] | z[nz]) from the item's box in Slot.
hipError_t launch_tables(const Slot *slot, float *ktab, float *xtab, int nx, int ny, int nz,
                         int nbatch, hipStream_t s);
// Twiddle table (4096 float2) on the host.
void fill_twiddles(float2 *tw);
void set_twiddle_table(const float2 *dev_tw);

// ------------------------------------------------------------- mass assignment

struct ScatterArgs {
    float *x, *y, *z;           // positions (mutated when wrap && mutate)
    const float *w;
    // optional second particle set (sort path, no wrap): particles
    // [n1, ntotal) are x2[i - n1], ...; catalog 1 of a two-catalog sort
    const float *x2, *y2, *z2, *w2;
    int64_t n1;
    const int64_t *offsets;     // catalog c owns [offsets[c], offsets[c+1])
    int ncat;
    int64_t ntotal;
    const Slot *slot;           // box per catalog
    int64_t slotStride;         // Slot entries between catalogs (0: shared)
    int nx, ny, nz;
    int wrap, mutate, tsc;
    int tx, ty, tz;             // tile extent (cells)
    int ntx, nty, ntz;          // tiles per axis
    int sx, sy, sz;             // log2 of the tile extents (-1: not a power of two; set by the launchers)
    uint32_t *count;            // per (catalog, tile)
    uint32_t *cursor;           // per (catalog, tile): start, advanced to end
    float4 *sorted;             // (u_x, u_y, u_z, w) per placement
    int32_t *err;               // set to 1 on an out-of-grid index
    float *rho;                 // output grids, catalog c at c*rstride
    int64_t rstride;
    uint32_t *wmax;             // per catalog: max |w| as float bits (fixed-point scale)
    double *part;               // optional: per (catalog, tile) FP64 sum of the cells written
};

hipError_t launch_scatter(ScatterArgs &a, uint32_t *scan_tmp, hipStream_t s);
// accumulate: 0 overwrite, 1 add into rho, 2 rho already zero (empty tiles skipped)
hipError_t launch_scatter_accum(const ScatterArgs &a, int accumulate, hipStream_t s);
// Zero n floats with a fill kernel (graph-capture safe).
// Zero nwords 32-bit words (a 4-byte aligned buffer) with a kernel.
hipError_t launch_zero_words(void *a, int64_t nwords, hipStream_t s);
hipError_t launch_zero(float *a, int64_t n, hipStream_t s);
// Blocked counting sort (no global atomics); valid when sort_fits(a).
bool sort_fits(const ScatterArgs &a);
// Upper bound on the blocks sort_blocks makes for a handful of catalogs.
constexpr int kSortTableBlocks = 1024;
int sort_nblk(int64_t n);  // blocks of a single catalog (tables may then be NULL)
int sort_blocks(int64_t n, const int64_t *offsets_host, int ncat, std::vector<int> &blk_cat,
                std::vector<int64_t> &blk_beg, std::vector<int> &cat_blk0);
hipError_t launch_sort(ScatterArgs &a, int nblk, uint32_t *H, const int *blk_cat, const int64_t *blk_beg,
                       const int *cat_blk0, hipStream_t s);
// Capacity (placements) the sorted buffer must hold for n particles.
int64_t scatter_capacity(int64_t n, int tsc);
// Cells of the largest tile k_scatter_accum holds in LDS.
int scatter_tile_max();

struct GatherArgs {
    const float *x, *y, *z;
    int64_t n;
    const float *disp;          // displacement grid, (Psi_x, Psi_y, Psi_z) per cell
    const Slot *slot;
    int nx, ny, nz;
    int field;                  // 0 disp, 1 rsd, 2 sum
    int positions;              // 1: write pos - shift
    int has_los;
    float los[3];
    float f;
    float *ox, *oy, *oz;
    int32_t *err;
};
hipError_t launch_gather(const GatherArgs &a, hipStream_t s);

// Scalar read-back (read_cic!) of one real grid.
hipError_t launch_read_cic(float *out, const float *field, const float *x, const float *y,
                           const float *z, int64_t n, const Slot *slot, int nx, int ny, int nz,
                           int32_t *err, hipStream_t s);

// setup_box min/max reduction into slot->minmax and the box.
hipError_t launch_setup_box(const float *x, const float *y, const float *z, int64_t n, Slot *slot,
                            float pad, hipStream_t s);

// ----------------------------------------------------------------- multigrid

struct MGLevel {
    int n[3];
    float cell[3], cell2[3], icell2[3], losn[3];
    const float *xv;   // x[n0] | y[n1] | z[n2] cell centres of this level
    const float *pv;   // the same divided by the cell size (FP32), = xv + n0 + n1 + n2
};

// ------------------------------------------------------------- cosmology

constexpr double kSpeedOfLightKms = 299792.458;   // src/cosmo.jl:12

// Derived densities of BAOrec.Cosmology (computed on the host in the Julia
// constructor's precision) and h*100 (Float32 product, as H() forms it).
struct CosmoArgs {
    double omega_nu0, omega_g0, omega_b0, omega_c0, omega_k0, omega_l0;
    double w0, wa, h100;
};
// Comoving-distance table: num knots zk (increasing, uniform) and distances
// rk (Mpc, increasing).  Interval lookup: direct index on the uniform z knots;
// for r, guide[j] = last knot with rk <= r0 + j / dr_inv (nguide entries).
struct CosmoTable {
    const double *zk, *rk;
    int num;
    double z0, dz_inv;
    const int *guide;
    int nguide;
    double r0, dr_inv;
};
hipError_t launch_cosmo_guide(const CosmoTable &t, int *guide, hipStream_t s);
hipError_t launch_cosmo_table(const CosmoArgs &c, const double *zk, double *seg, double *rk, int num,
                              hipStream_t s);
hipError_t launch_comoving_distance(const CosmoTable &t, const double *z, double *r, int64_t n, int *err,
                                    hipStream_t s);
hipError_t launch_sky_to_cart(const CosmoTable &t, float hh, const float *ra, const float *dec, const float *red,
                              float *x, float *y, float *z, int64_t n, int *err, hipStream_t s);
hipError_t launch_cart_to_sky(const CosmoTable &t, float h, const float *x, const float *y, const float *z,
                              float *ra, float *dec, float *red, int64_t n, int *err, hipStream_t s);

// ------------------------------------------------------------- power spectrum

// Multipole estimator over the half spectra F1 (and F2, interlaced grid) of
// an nx x ny x nz grid (pitch P, Pv = nx/2 + 1 valid columns).  tab holds per
// axis (x: Pv entries, y: ny, z: nz) cos and sin of k_a H_a / 2 and the
// window factor sinc(k_a H_a / 2)^p.  part: pk_blocks() * 5 * nbins doubles.
struct PkArgs {
    const float2 *F1, *F2;
    const float4 *tab;
    int nx, ny, nz, P, Pv;
    int64_t PZ;
    double kf[3], los[3];
    double kmin, dk;
    int nbins, interlace;
    double *part;
};
int pk_blocks();
hipError_t launch_pk(const PkArgs &a, double *out, hipStream_t s);

// Levels [0, nl) of a V-cycle (fmg = 0) or of the whole FMG (fmg = 1) run by
// one workgroup (the launch-latency-bound coarse levels).  cur[l] says which
// of A[l] / B[l] holds v (damped Jacobi ping-pongs; an even number of sweeps
// per visit leaves it unchanged).  gs = 0: damped Jacobi, jacobi_n sweeps
// (float only); gs = 1: four-colour Gauss-Seidel, rb_sweeps (16x on the
// coarsest level).
constexpr int kMGCoarseLevels = 4;
template <typename T>
struct MGCoarse {
    int nl;
    MGLevel L[kMGCoarseLevels];
    T *A[kMGCoarseLevels], *B[kMGCoarseLevels], *F[kMGCoarseLevels];
    int cur[kMGCoarseLevels];
    int fmg, gs;
    float beta, omega;
    int has_los, jacobi_n, rb_sweeps, vcycles;
};
template <typename T>
hipError_t launch_mg_coarse(const MGCoarse<T> &a, hipStream_t s);

hipError_t launch_mg_jacobi(float *vout, const float *vin, const float *f, const MGLevel &L,
                            float beta, float omega, int has_los, hipStream_t s);
hipError_t launch_mg_residual(float *r, const float *v, const float *f, const MGLevel &L,
                              float beta, int has_los, hipStream_t s);
// The same residual with the coefficient from the hardware reciprocal (1 ulp)
// on z-marching shapes (otherwise launch_mg_residual), for the RBGS solve.
hipError_t launch_mg_residual_fast(float *r, const float *v, const float *f, const MGLevel &L, float beta,
                                   int has_los, hipStream_t s);
hipError_t launch_mg_reduce(float *coarse, const float *fine, const int fine_n[3], hipStream_t s);
// coarse = restriction of (f - A v) in one pass, the bits of launch_mg_residual
// followed by launch_mg_reduce (sides % (64, 16, 2), < 2^31 cells)
bool mg_residual_restrict_ok(const MGLevel &L);
// fast: the coefficient from the hardware reciprocal, as launch_mg_residual_fast.
hipError_t launch_mg_residual_restrict(float *coarse, const float *v, const float *f, const MGLevel &L, float beta,
                                       int has_los, hipStream_t s, int fast = 0);
hipError_t launch_mg_prolong(float *fine, const float *coarse, const int coarse_n[3], int add,
                             hipStream_t s);
// xv[0 .. n) cell centres (x | y | z), xv[n .. 2n) centres / cell size, n = nx + ny + nz
hipError_t launch_mg_xvec(float *xv, const Slot *slot, int nx, int ny, int nz, hipStream_t s);

// FP64-capable red-black Gauss-Seidel pieces (RBGS schedule)
template <typename T>
hipError_t launch_mg_rbgs(T *v, const T *f, const MGLevel &L, float beta, int has_los, int color,
                          hipStream_t s);
// 2 or 3 damped Jacobi sweeps in one launch, out of place (vout != vin),
// bitwise that many launch_mg_jacobi calls; sides multiples of 64, 16, 8.
bool mg_jacobi_tb_ok(const MGLevel &L);
hipError_t launch_mg_jacobi_tb(float *vout, const float *vin, const float *f, const MGLevel &L, float beta,
                               float omega, int has_los, int sweeps, hipStream_t s);
// One whole 4-colour sweep out of place (vout != vin), bitwise the four
// in-place launch_mg_rbgs colour passes; for levels with mg_gs4_ok (sides
// multiples of 64, 16 and 8).
// part (may be NULL): mg_gs4_blocks(L) FP64 partial sums of vout, fixed order.
// vin == NULL: the sweep starts from v = 0 (nothing is read for it).
bool mg_gs4_ok(const MGLevel &L);
int64_t mg_gs4_blocks(const MGLevel &L);
hipError_t launch_mg_gs4(float *vout, const float *vin, const float *f, const MGLevel &L, float beta, int has_los,
                         hipStream_t s, double *part = nullptr);
template <typename T>
hipError_t launch_mg_residual_t(T *r, const T *v, const T *f, const MGLevel &L, float beta,
                                int has_los, hipStream_t s);
template <typename T, typename U>
hipError_t launch_mg_reduce_t(T *coarse, const U *fine, const int fine_n[3], hipStream_t s);
template <typename T, typename U>
hipError_t launch_mg_prolong_t(T *fine, const U *coarse, const int coarse_n[3], int add,
                               hipStream_t s);
// FP64 sum, sum of squares and sum of squared differences of a grid, written
// to *out (overwritten) in a fixed order; part holds 1024 doubles of partials
template <typename T>
hipError_t launch_sumsq(const T *a, int64_t n, double *out, double *part, hipStream_t s);
template <typename T>
hipError_t launch_sumsq_diff(const T *a, const T *b, int64_t n, double *out, double *part, hipStream_t s);
template <typename T>
hipError_t launch_sum(const T *a, int64_t n, double *out, double *part, hipStream_t s);
// *a += *b (one FP64 scalar on the device)
hipError_t launch_accumulate(double *a, const double *b, hipStream_t s);
template <typename T>
hipError_t launch_add_const(T *a, int64_t n, const double *num, double den, double sign,
                            hipStream_t s);
template <typename T, typename U>
hipError_t launch_convert(T *dst, const U *src, int64_t n, hipStream_t s);
hipError_t launch_fill(float *a, int64_t n, float v, hipStream_t s);
// r = (f - fsum/count) - A v, FP64 arithmetic, FP32 result (fsum NULL: no shift)
hipError_t launch_mg_residual_mixed(float *r, const double *v, const float *f, const double *fsum, double count,
                                    const MGLevel &L, float beta, int has_los, hipStream_t s);
// phi += e - esum/n (sum in FP64, stored in phi's type)
template <typename T>
hipError_t launch_add_correction(T *phi, const float *e, int64_t n, const double *esum, hipStream_t s);
// The FP64 refinement pass of the RBGS solve in one launch: phi = phi_in +
// (e - sum(e)/N); without e, phi = phi_in, or (double)seed when seed is
// given, or 0 when phi_in is NULL; phi goes to phi_out unless it is phi_in
// unchanged.  r = (float)((f - *fsum/N) - A phi), and the per-workgroup FP64
// sums of r in part[mg_refine_blocks(L)] (fixed order).  Sides multiples of
// 64, 16, 8.
bool mg_refine_ok(const MGLevel &L);
int64_t mg_refine_blocks(const MGLevel &L);
hipError_t launch_mg_refine(double *phi_out, const double *phi_in, const float *seed, const float *e,
                            const double *esum, const float *f, const double *fsum, float *r, double *part,
                            const MGLevel &L, float beta, int has_los, hipStream_t s);
// r += (float)(-*rsum / n) (rsum may be NULL), then out[0] = sum r^2 and,
// with rprev, out[1] = sum (r - rprev)^2, FP64 in a fixed order; part holds
// 2 x 1024 partials.
hipError_t launch_shift_norms(float *r, const float *rprev, int64_t n, const double *rsum, double *out,
                              double *part, hipStream_t s);

// ------------------------------------------------------------------ U-Net (unet.hip)
// Tensors are BF16 (stored as uint16_t), channels last [batch][z][y][x][c];
// grid sides are powers of two (log2 in lx, ly, lz).
struct ConvArgs {
    const uint16_t *in1;   // C1 channels, full resolution
    const uint16_t *in2;   // C2 channels, half resolution when up2 (nearest upsampling on read)
    const uint16_t *w;     // packed B operand (launch_unet_pack)
    const float *bias;     // CO entries or NULL
    uint16_t *out;         // CO channels
    int lx, ly, lz;
};
struct WgradArgs {
    const uint16_t *in1, *in2;   // the forward convolution's inputs
    const uint16_t *dout;        // gradient of its output (after the ReLU mask)
    float *part;                 // [nchunks][27][C1 + C2][CO] partial sums
    int lx, ly, lz, batch;
    int64_t per;                 // bricks of 16 x 4 x 4 voxels per chunk
};
struct FeatArgs {
    int lx, ly, lz, batch;
    float bmin[3], h[3];   // box_min and cell size per axis
    float nc[3];           // fixed LOS of the spectral seed (unit vector)
    float ih2;             // 1 / h^2
    float href;            // reference length of the 1/|x| feature
};
hipError_t launch_unet_conv(const ConvArgs &a, int c1, int c2, int co, bool relu, bool up2, int batch,
                            hipStream_t s);
hipError_t launch_unet_wgrad(const WgradArgs &a, int c1, int c2, int co, bool up2, int nchunks, hipStream_t s);
hipError_t launch_unet_sum_chunks(const float *part, int nch, int64_t n, float *out, hipStream_t s);
hipError_t launch_unet_pack(const float *w, int cin, int co, int transposed, uint16_t *out, hipStream_t s);
hipError_t launch_unet_relu_mask(uint16_t *dy, const uint16_t *y, int64_t n, hipStream_t s);
hipError_t launch_unet_pool(const uint16_t *in, uint16_t *out, int lx, int ly, int lz, int lc, int batch,
                            hipStream_t s);
hipError_t launch_unet_pool_bwd(const uint16_t *dc, uint16_t *df, int lx, int ly, int lz, int lc, int batch,
                                hipStream_t s);
hipError_t launch_unet_split(const uint16_t *dcat, uint16_t *dskip, uint16_t *dcoarse, int c1, int c2, int lx,
                             int ly, int lz, int batch, hipStream_t s);
hipError_t launch_unet_head(const uint16_t *x, const float *w, const float *b, float *out, int64_t n,
                            hipStream_t s);
int unet_head_bwd_blocks(int64_t n);
hipError_t launch_unet_head_bwd(const uint16_t *x, const float *w, const float *dout, uint16_t *dx, float *part,
                                int64_t n, hipStream_t s);
hipError_t launch_unet_features(const float *f, const float *phis, const float *rs, const float *psi,
                                const double *ss, const FeatArgs &a, uint16_t *X, hipStream_t s);
// a += b (FP32 grids)
hipError_t launch_add_grid(float *a, const float *b, int64_t n, hipStream_t s);
hipError_t launch_unet_target(const float *phi, const float *phis, const double *ss, float ih2, int64_t nv, int64_t n,
                              float *t, hipStream_t s);
int unet_loss_blocks(int64_t n);
hipError_t launch_unet_loss(const float *out, const float *t, int lx, int ly, int lz, int64_t n, float scale,
                            float *dout, double *part, hipStream_t s);
hipError_t launch_unet_adam(float *w, float *m, float *v, const float *g, int64_t n, float lr_t, float b1, float b2,
                            float eps, float wd, hipStream_t s);
hipError_t launch_unet_apply_seed(const float *e, const double *ss, float h2, int64_t n, int add, float *out,
                                  hipStream_t s);


// The network object (unet.hip): parameters, Adam state, activations.
struct UNet;
UNet *unet_create();
int64_t unet_count_params();
void unet_destroy(UNet *n);
int64_t unet_nparams(const UNet *n);
float unet_h_train(const UNet *n);
void unet_set_h_train(UNet *n, float h);
int64_t unet_step(const UNet *n);
bool unet_ready(const UNet *n);
hipError_t unet_init(UNet *n, uint64_t seed, hipStream_t s);
hipError_t unet_set_params(UNet *n, const float *host, hipStream_t s);
hipError_t unet_get_params(const UNet *n, float *host, int grads, hipStream_t s);
hipError_t unet_input(UNet *n, int batch, int lx, int ly, int lz, uint16_t **X);
hipError_t unet_tensor(const UNet *n, int kind, int i, const uint16_t **ptr, int64_t *count);
float *unet_output(UNet *n);
float *unet_output_grad(UNet *n);
hipError_t unet_forward(UNet *n, hipStream_t s);
hipError_t unet_backward(UNet *n, hipStream_t s);
hipError_t unet_adam(UNet *n, float lr, hipStream_t s);

}  // namespace baorec
