// Sky <-> Cartesian conversion on the GPU: the comoving-distance table of
// BAOrec.Cosmology (src/cosmo.jl:64-98) built by FP64 Gauss-Legendre
// quadrature per table interval plus an ordered prefix sum, and the per-galaxy
// conversions of the reference's lightcone driver (examples/lightcone.jl:30-80).
#include <math.h>

#include "kernels.hpp"

namespace baorec {

// 1 / H(z) in s Mpc / km, src/cosmo.jl:74-85: E(z) sums the densities in the
// reference's order; every term is FP64 (the Float32 parameters promote).
__device__ __forceinline__ double inv_hubble(const CosmoArgs &c, double z) {
    const double zp = 1.0 + z;
    const double z2 = zp * zp, z3 = z2 * zp, z4 = z2 * z2;
    const double a = 1.0 / zp;
    // f_DE(a) (src/cosmo.jl:80); at a = 1 the wa term is 0 * NaN, but
    // (1 + z)^f with z = 0 is 1 for any f, which pow() also returns
    const double fde = -3.0 * (1.0 + c.w0) + 3.0 * c.wa * ((a - 1.0) / log(a) - 1.0);
    const double ol = c.omega_l0 * pow(zp, -fde);
    const double e = sqrt(c.omega_nu0 * z4 + c.omega_g0 * z4 + c.omega_b0 * z3 + c.omega_c0 * z3 +
                          c.omega_k0 * z2 + ol);
    return 1.0 / (c.h100 * e);
}

// 4-point Gauss-Legendre rule on [a, b] of c / H (Mpc).
__device__ __forceinline__ double gl4(const CosmoArgs &c, double a, double b) {
    const double x1 = 0.33998104358485626, x2 = 0.86113631159405258;
    const double w1 = 0.65214515486254614, w2 = 0.34785484513745386;
    const double m = 0.5 * (a + b), h = 0.5 * (b - a);
    const double s = w2 * (inv_hubble(c, m - h * x2) + inv_hubble(c, m + h * x2)) +
                     w1 * (inv_hubble(c, m - h * x1) + inv_hubble(c, m + h * x1));
    return kSpeedOfLightKms * h * s;
}

// seg[i] = integral over [zk[i], zk[i+1]], i < num - 1
__global__ void k_cosmo_seg(CosmoArgs c, const double *__restrict__ zk, double *__restrict__ seg, int num) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= num - 1) return;
    seg[i] = gl4(c, zk[i], zk[i + 1]);
}

// r[k] = r0 + sum_{i<k} seg[i], with r0 the integral over [0, zk[0]]; one
// workgroup, fixed summation order (deterministic).
constexpr int kScanThreads = 1024;
__global__ void __launch_bounds__(kScanThreads) k_cosmo_scan(CosmoArgs c, const double *__restrict__ zk,
                                                            const double *__restrict__ seg,
                                                            double *__restrict__ r, int num) {
    __shared__ double part[kScanThreads];
    const int t = threadIdx.x;
    // r0: [0, zk[0]] in kScanThreads * 4 sub-intervals, tree-reduced
    double s0 = 0.0;
    if (zk[0] > 0.0) {
        const double h = zk[0] / (kScanThreads * 4);
        for (int j = 0; j < 4; j++) {
            const double a = (t * 4 + j) * h;
            s0 += gl4(c, a, a + h);
        }
    }
    part[t] = s0;
    __syncthreads();
    for (int o = kScanThreads / 2; o > 0; o >>= 1) {
        if (t < o) part[t] += part[t + o];
        __syncthreads();
    }
    const double r0 = part[0];
    __syncthreads();
    // chunked exclusive scan of the num - 1 segments
    const int nseg = num - 1;
    const int chunk = (nseg + kScanThreads - 1) / kScanThreads;
    const int b = t * chunk, e = min(nseg, b + chunk);
    double cs = 0.0;
    for (int i = b; i < e; i++) cs += seg[i];
    part[t] = cs;
    __syncthreads();
    if (t == 0) {
        double run = r0;
        for (int i = 0; i < kScanThreads; i++) {
            const double v = part[i];
            part[i] = run;
            run += v;
        }
    }
    __syncthreads();
    double run = part[t];
    if (t == 0) r[0] = r0;
    for (int i = b; i < e; i++) {
        run += seg[i];
        r[i + 1] = run;
    }
}

// Interval i of the knots xk (increasing) holding x, as
// Interpolations.Gridded(Linear()) picks it: the last i <= num - 2 with
// xk[i] <= x, starting from a guess i0 at or below it.  Then linear
// interpolation of yk.  Returns false outside [xk[0], xk[num-1]] (the
// reference's BoundsError).
__device__ __forceinline__ bool interp_from(const double *__restrict__ xk, const double *__restrict__ yk, int num,
                                            double x, int i, double &y) {
    if (!(x >= xk[0] && x <= xk[num - 1])) return false;
    i = i < 0 ? 0 : (i > num - 2 ? num - 2 : i);
    while (i > 0 && xk[i] > x) i--;
    while (i < num - 2 && xk[i + 1] <= x) i++;
    const double w = (x - xk[i]) / (xk[i + 1] - xk[i]);
    y = yk[i] + w * (yk[i + 1] - yk[i]);
    return true;
}

// r -> z: guess from the guide table (uniform in r)
__device__ __forceinline__ bool interp_r(const CosmoTable &t, double r, double &z) {
    if (!(r >= t.rk[0] && r <= t.rk[t.num - 1])) return false;
    int j = (int)((r - t.r0) * t.dr_inv);
    j = j < 0 ? 0 : (j >= t.nguide ? t.nguide - 1 : j);
    return interp_from(t.rk, t.zk, t.num, r, t.guide[j], z);
}

// z -> r: direct index on the uniform knots
__device__ __forceinline__ bool interp_z(const CosmoTable &t, double z, double &r) {
    if (!(z >= t.zk[0] && z <= t.zk[t.num - 1])) return false;
    return interp_from(t.zk, t.rk, t.num, z, (int)((z - t.z0) * t.dz_inv), r);
}

// guide[j] = last knot i <= num - 2 with rk[i] <= r0 + j / dr_inv (binary search)
__global__ void k_cosmo_guide(CosmoTable t, int *guide) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= t.nguide) return;
    const double r = t.r0 + (double)j / t.dr_inv;
    int lo = 0, hi = t.num - 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (t.rk[mid] <= r)
            lo = mid;
        else
            hi = mid;
    }
    guide[j] = lo;
}

__global__ void k_comoving_distance(CosmoTable t, const double *__restrict__ z, double *__restrict__ r,
                                    int64_t n, int *err) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = 0.0;
    if (!interp_z(t, z[i], v)) {
        atomicOr(err, 1);
        v = NAN;
    }
    r[i] = v;
}

// sky_to_cartesian, examples/lightcone.jl:30-48: angles in FP32 as the driver
// computes them (ra * pi / 180 with pi rounded to Float32), trigonometric
// functions evaluated in FP64 and rounded (Julia's Float32 cos/sin are
// correctly rounded in practice), products dist * cos(dec) * cos(ra) * h in
// FP64 (dist is the FP64 table value), stored as FP32.
__global__ void k_sky_to_cart(CosmoTable t, float hh, const float *__restrict__ ra, const float *__restrict__ dec,
                              const float *__restrict__ red, float *__restrict__ x, float *__restrict__ y,
                              float *__restrict__ z, int64_t n, int *err) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double dist = 0.0;
    if (!interp_z(t, (double)red[i], dist)) {
        atomicOr(err, 1);
        dist = NAN;
    }
    const float pif = 3.14159265358979323846f;
    const float rar = (ra[i] * pif) / 180.0f;
    const float der = (dec[i] * pif) / 180.0f;
    double sd, cd, sr, cr;
    sincos((double)der, &sd, &cd);
    sincos((double)rar, &sr, &cr);
    sd = (double)(float)sd;
    cd = (double)(float)cd;
    sr = (double)(float)sr;
    cr = (double)(float)cr;
    x[i] = (float)(((dist * cd) * cr) * (double)hh);
    y[i] = (float)(((dist * cd) * sr) * (double)hh);
    z[i] = (float)((dist * sd) * (double)hh);
}

// cartesian_to_sky, examples/lightcone.jl:50-80: r = |pos| / h in FP32, the
// redshift interpolated on the distance table, RA = rem(lon - 360, 360) as
// the driver writes it (values in (-360, 0]), Dec from atan(z, s).
__global__ void k_cart_to_sky(CosmoTable t, float h, const float *__restrict__ x, const float *__restrict__ y,
                              const float *__restrict__ z, float *__restrict__ ra, float *__restrict__ dec,
                              float *__restrict__ red, int64_t n, int *err) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float xv = x[i], yv = y[i], zv = z[i];
    const float r = sqrtf((xv * xv + yv * yv) + zv * zv) / h;
    double zz = 0.0;
    if (!interp_r(t, (double)r, zz)) {
        atomicOr(err, 1);
        zz = NAN;
    }
    const float s = sqrtf(xv * xv + yv * yv);
    const float pif = 3.14159265358979323846f;
    float lon = (float)atan2((double)yv, (double)xv);
    float lat = (float)atan2((double)zv, (double)s);
    lon = (lon * 180.0f) / pif;
    lat = (lat * 180.0f) / pif;
    lon = fmodf(lon - 360.0f, 360.0f);
    ra[i] = lon;
    dec[i] = lat;
    red[i] = (float)zz;
}

static inline unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

hipError_t launch_cosmo_table(const CosmoArgs &c, const double *zk, double *seg, double *rk, int num,
                              hipStream_t s) {
    hipLaunchKernelGGL(k_cosmo_seg, dim3(blocks_for(num)), dim3(256), 0, s, c, zk, seg, num);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_cosmo_scan, dim3(1), dim3(kScanThreads), 0, s, c, zk, seg, rk, num);
    return hipGetLastError();
}

hipError_t launch_cosmo_guide(const CosmoTable &t, int *guide, hipStream_t s) {
    hipLaunchKernelGGL(k_cosmo_guide, dim3(blocks_for(t.nguide)), dim3(256), 0, s, t, guide);
    return hipGetLastError();
}

hipError_t launch_comoving_distance(const CosmoTable &t, const double *z, double *r, int64_t n, int *err,
                                    hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_comoving_distance, dim3(blocks_for(n)), dim3(256), 0, s, t, z, r, n, err);
    return hipGetLastError();
}

hipError_t launch_sky_to_cart(const CosmoTable &t, float hh, const float *ra, const float *dec, const float *red,
                              float *x, float *y, float *z, int64_t n, int *err, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_sky_to_cart, dim3(blocks_for(n)), dim3(256), 0, s, t, hh, ra, dec, red, x, y, z, n, err);
    return hipGetLastError();
}

hipError_t launch_cart_to_sky(const CosmoTable &t, float h, const float *x, const float *y, const float *z,
                              float *ra, float *dec, float *red, int64_t n, int *err, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_cart_to_sky, dim3(blocks_for(n)), dim3(256), 0, s, t, h, x, y, z, ra, dec, red, n, err);
    return hipGetLastError();
}

}  // namespace baorec
