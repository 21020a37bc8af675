// Stockham R2C / C2R 3D FFT passes for gfx950 with fused k-space multipliers
// and real-space epilogues.
//
// A 3D real transform is three passes over HBM:
//   x pass:  R2C along the contiguous axis (n real = n/2 complex, packed trick)
//   y pass:  complex FFT along y, B consecutive kx columns per workgroup
//   z pass:  complex FFT along z, B consecutive kx columns per workgroup
// The reference's k-space kernels (gaussian_filter_kernel!, inv_laplace_kernel!,
// hessian_kernel!, displacement_kernel!, mg_gradient_kernel!) and real-space
// broadcasts are folded into these passes instead of running as extra sweeps.
#include <math.h>

#include "kernels.hpp"

namespace baorec {

__device__ float2 g_tw[4096];

void fill_twiddles(float2 *tw) {
    tw[0] = make_float2(1.f, 0.f);
    for (int S = 1; S <= 2048; S *= 2)
        for (int i = 0; i < S; i++) {
            double a = -2.0 * M_PI * (double)i / (double)S;
            tw[S + i] = make_float2((float)cos(a), (float)sin(a));
        }
}

bool fft_side_ok(int n) { return n >= 16 && n <= kMaxSide && (n & (n - 1)) == 0; }

// ------------------------------------------------------------ block sum

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// out[b] = sum of the n partials of batch item b, in a fixed order: thread t
// sums elements t + 1024 k over 8 independent accumulators (k mod 8), then
// the accumulators, the waves and the workgroup in order.
__global__ void __launch_bounds__(1024) k_sum_partials(const double *part, int64_t n, double *out,
                                                       int64_t outStride) {
    __shared__ double red[16];
    const double *p = part + (int64_t)blockIdx.x * n;
    double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int64_t i = threadIdx.x;
    for (; i + 7 * 1024 < n; i += 8 * 1024) {
#pragma unroll
        for (int u = 0; u < 8; u++) acc[u] += p[i + u * 1024];
    }
#pragma unroll
    for (int u = 0; u < 7; u++) {
        const int64_t j = i + u * 1024;
        if (j < n) acc[u] += p[j];
    }
    double s = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < 16; w++) t += red[w];
        out[blockIdx.x * outStride] = t;
    }
}

// Half spectra between the padded working layout (kx row pitch PP, kz plane
// pitch PZ) and the compact (nx/2+1) x ny x nz layout of the C ABI.
__global__ void k_spectrum_pack(float2 *__restrict__ dst, const float2 *__restrict__ src, int P, int PP, int64_t PZ,
                                int ny, int nz, int to_compact) {
    const int64_t n = (int64_t)P * ny * nz;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int kx = (int)(i % P);
        const int64_t row = i / P;
        const int ky = (int)(row % ny);
        const int64_t kz = row / ny;
        const int64_t padded = kz * PZ + (int64_t)ky * PP + kx;
        if (to_compact)
            dst[i] = src[padded];
        else
            dst[padded] = src[i];
    }
}

hipError_t launch_spectrum_pack(float2 *dst, const float2 *src, int P, int PP, int64_t PZ, int ny, int nz,
                                int to_compact, hipStream_t s) {
    const int64_t n = (int64_t)P * ny * nz;
    const int64_t nb = (n + 255) / 256;
    hipLaunchKernelGGL(k_spectrum_pack, dim3((unsigned)(nb < 16384 ? nb : 16384)), dim3(256), 0, s, dst, src, P, PP,
                       PZ, ny, nz, to_compact);
    return hipGetLastError();
}

hipError_t launch_sum_partials(const double *part, int64_t n, int nb, double *out, int64_t stride, hipStream_t s) {
    hipLaunchKernelGGL(k_sum_partials, dim3(nb), dim3(1024), 0, s, part, n, out, stride);
    return hipGetLastError();
}

// ------------------------------------------------------------ x: R2C

template <int M>
struct XShape {
    static constexpr int TPL = M / 8;                   // threads per line
    static constexpr int LPW = TPL >= 256 ? 1 : 256 / TPL;  // lines per workgroup
    static constexpr int LS = M + M / 8;                // padded LDS line stride
    static constexpr int NT = TPL * LPW;
};

template <int M>
__global__ void __launch_bounds__(XShape<M>::NT) k_x_r2c(XFwdArgs a) {
    using S = XShape<M>;
    using PF = Plan<M, false>;
    __shared__ float2 lds[S::LPW * S::LS];
    const int tid = threadIdx.x;
    const int li = tid / S::TPL, t = tid % S::TPL;
    const int64_t nlines = (int64_t)a.ny * a.nz;
    const int64_t line = (int64_t)blockIdx.x * S::LPW + li;
    const bool valid = line < nlines;
    const int b = blockIdx.y;
    const float2 *src = reinterpret_cast<const float2 *>(a.in + b * a.rstride) + line * M;
    float2 v[8];
#pragma unroll
    for (int i = 0; i < 8; i++) v[i] = valid ? src[PF::in_pos(t, i)] : make_float2(0.f, 0.f);
    LayLine lay{li * S::LS};
    fft_line<M, false>(v, lds, lay, t, g_tw);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 8; i++) lds[lay(PF::out_pos(t, i))] = v[i];
    __syncthreads();
    float2 *dst = a.out + b * a.cstride + (line % a.ny) * a.P + (line / a.ny) * a.PZ;
    if (valid) {
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int k = t + i * S::TPL;
            float2 zk = lds[lay(k)];
            float2 zc = cconj(lds[lay((M - k) & (M - 1))]);
            float2 e = cscale(cadd(zk, zc), 0.5f);
            float2 d = csub(zk, zc);
            float2 o = make_float2(0.5f * d.y, -0.5f * d.x);  // (zk - zc) / (2i)
            dst[k] = cadd(e, cmul(g_tw[2 * M + k], o));
        }
        if (t == 0) {
            float2 z0 = lds[lay(0)];
            dst[M] = make_float2(z0.x - z0.y, 0.f);
        }
    }
}

// ------------------------------------------------------------ x: C2R

template <int EPI>
struct XIn {
    static constexpr int N = EPI == XE_ITER_LOCAL ? 3 : EPI == XE_DISP ? 3 : EPI == XE_OD_RAND ? 2 : 1;
};

// Loads X[k], k = t + i*M/8, of one spectrum line (times kx^pw) and, for
// thread t == 0, the Nyquist bin X[M].
template <int M>
__device__ __forceinline__ void xline_load(const float2 *__restrict__ src, bool valid, int t,
                                           float2 (&x)[8], float2 &xn) {
#pragma unroll
    for (int i = 0; i < 8; i++) x[i] = valid ? src[t + i * (M / 8)] : make_float2(0.f, 0.f);
    xn = (valid && t == 0) ? src[M] : make_float2(0.f, 0.f);
}

// Waves per SIMD the x C2R pass is compiled for: the local-LOS iteration
// (three partial sums per cell) at 3 with n <= 512 (168 VGPRs, no spills),
// and the passes that smooth real lines (forward and inverse transform per
// input) at 3 (the randoms overdensity spills 4 VGPRs there, and still runs
// 1.09 -> 0.87 ms at 512^3); the others fit 4 or more by themselves.
// Input X'[k] = exp(-R2 kx^2 / 2) X[k] of the real line src (2M floats),
// k = t + i*M/8, and X'[M] for thread t == 0, where X is the R2C of the line:
// the packed forward transform and split of k_x_r2c (same arithmetic), then
// the Gaussian's x factor with its exponent in FP64.  Uses lds (the line's
// exchange area, barriers inside).
template <int M, class Lay>
__device__ __forceinline__ void xline_real(const float *__restrict__ src, bool valid, int t, float2 *lds,
                                           const Lay &lay, const float2 *tw, const float *kxl, float R2,
                                           float2 (&x)[8], float2 &xn) {
    using PF = Plan<M, false>;
    constexpr int TPL = M / 8;
    const float2 *s2 = reinterpret_cast<const float2 *>(src);
    float2 v[8];
#pragma unroll
    for (int i = 0; i < 8; i++) v[i] = valid ? s2[PF::in_pos(t, i)] : make_float2(0.f, 0.f);
    fft_line<M, false>(v, lds, lay, t, tw);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 8; i++) lds[lay(PF::out_pos(t, i))] = v[i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int k = t + i * TPL;
        const float2 zk = lds[lay(k)];
        const float2 zc = cconj(lds[lay((M - k) & (M - 1))]);
        const float2 e = cscale(cadd(zk, zc), 0.5f);
        const float2 d = csub(zk, zc);
        const float2 o = make_float2(0.5f * d.y, -0.5f * d.x);  // (zk - zc) / (2i)
        const float2 X = cadd(e, cmul(tw[2 * M + k], o));
        const float kk = kxl[k];
        const double g = exp(-0.5 * (double)R2 * (double)(kk * kk));
        x[i] = make_float2((float)((double)X.x * g), (float)((double)X.y * g));
    }
    xn = make_float2(0.f, 0.f);
    if (t == 0) {
        const float2 z0 = lds[lay(0)];
        const float kk = kxl[M];
        const double g = exp(-0.5 * (double)R2 * (double)(kk * kk));
        xn = make_float2((float)((double)(z0.x - z0.y) * g), 0.f);
    }
}

template <int M, int EPI, bool RIN = false>
struct XWaves {
    static constexpr int W = (EPI == XE_ITER_LOCAL || RIN) && M <= 256 ? 3 : 1;
};

// An SGPR zero the compiler cannot see through: table reads indexed with it
// are repeated per use instead of being kept live in registers.
__device__ __forceinline__ int opaque_zero() {
    int z = 0;
    asm volatile("" : "+s"(z));
    return z;
}

template <int M, int EPI, bool RIN = false>
__global__ void __launch_bounds__(XShape<M>::NT) __attribute__((amdgpu_waves_per_eu(XWaves<M, EPI, RIN>::W)))
k_x_c2r(XInvArgs a) {
    using S = XShape<M>;
    using PI = Plan<M, true>;
    constexpr int NIN = XIn<EPI>::N;
    constexpr int TPL = S::TPL;
    constexpr bool LXV = EPI == XE_ITER_LOCAL;
    __shared__ float2 lds[S::LPW * S::LS];
    // The twiddles (stage factors and exp(2 pi i k / N), 3M entries), kx and
    // (local LOS) the x coordinates in LDS, read per use: held in registers
    // they kept the local-LOS pass at 2 waves per SIMD (226 VGPRs), whose
    // arithmetic then hardly overlapped its memory traffic
    __shared__ float2 twl[3 * M];
    __shared__ float kxl[M + 1];
    __shared__ float xvl[LXV ? 2 * M : 1];
    const int tid = threadIdx.x;
    const int li = tid / TPL, t = tid % TPL;
    const int64_t nlines = (int64_t)a.ny * a.nz;
    const int64_t line = (int64_t)blockIdx.x * S::LPW + li;
    const bool valid = line < nlines;
    const int b = blockIdx.y;
    const int iy = (int)(line % a.ny), iz = (int)(line / a.ny);
    const float *kx = a.ktab + b * a.ktabStride;
    const int64_t rbase = b * a.rstride + line * (2 * M);
    LayLine lay{li * S::LS};
    // partner lane holding X[M - k] (k = t + i*M/8 -> thread TPL - t, index 7 - i)
    const int lane = threadIdx.x & 63;
    const int plane = t == 0 ? lane : lane - t + (TPL - t);

    for (int i = tid; i < 3 * M; i += S::NT) twl[i] = g_tw[i];
    for (int i = tid; i <= M; i += S::NT) kxl[i] = kx[i];
    if constexpr (LXV) {
        const float *xvg = a.xtab + b * a.xtabStride;
        for (int i = tid; i < 2 * M; i += S::NT) xvl[i] = xvg[i];
    }
    __syncthreads();

    float2 acc[8];
    const float *xv = a.xtab ? a.xtab + b * a.xtabStride : nullptr;
    float xy = 0.f, xz = 0.f;
    if constexpr (EPI == XE_ITER_LOCAL) {
        xy = xv[2 * M + iy];
        xz = xv[2 * M + a.ny + iz];
    }
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int m = PI::out_pos(t, i);
        if constexpr (EPI == XE_ITER_LOCAL || EPI == XE_ITER_FIXED) {
            acc[i] = valid ? *reinterpret_cast<const float2 *>(a.aux + rbase + 2 * m)
                           : make_float2(0.f, 0.f);
        } else {
            acc[i] = make_float2(0.f, 0.f);
        }
    }
    float3 disp[8][2];
    float rr[8][2];
    float qs[8][2];

    float2 cur[8], cn;
    const int64_t crow = (int64_t)iy * a.P + (int64_t)iz * a.PZ;
    if constexpr (!RIN) xline_load<M>(a.in[0] + b * a.cstride + crow, valid, t, cur, cn);
#pragma unroll
    for (int q = 0; q < NIN; q++) {
        float2 nxt[8], nn;
        if constexpr (RIN) {
            xline_real<M>(a.rin[q] + b * a.rstride + line * (2 * M), valid, t, lds, lay, twl, kxl, a.R2, cur, cn);
        } else {
            if (q + 1 < NIN) xline_load<M>(a.in[q + 1] + b * a.cstride + crow, valid, t, nxt, nn);
        }
        const int pw = a.kxpow[q];
        const int z0 = opaque_zero();
        float2 v[8];
        float2 xk[8];
#pragma unroll
        for (int i = 0; i < 8; i++) xk[i] = pw ? cscale(cur[i], kpow(kxl[z0 + t + i * TPL], pw)) : cur[i];
        const float2 xnyq = pw ? cscale(cn, kpow(kxl[z0 + M], pw)) : cn;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            float2 xm;
            xm.x = __shfl(xk[7 - i].x, plane, 64);
            xm.y = __shfl(xk[7 - i].y, plane, 64);
            if (t == 0) xm = i == 0 ? xnyq : xk[(8 - i) & 7];
            float2 xkk = xk[i];
            if (t == 0 && i == 0) {  // self-conjugate bins: imaginary parts ignored (FFTW/pocketfft C2R)
                xkk.y = 0.f;
                xm.y = 0.f;
            }
            const float2 xc = cconj(xm);
            const float2 s = cadd(xkk, xc);
            const float2 d = csub(xkk, xc);
            const float2 wd = cmul(cconj(twl[z0 + 2 * M + t + i * TPL]), d);  // exp(+2 pi i k / N) d
            v[i] = make_float2(s.x - wd.y, s.y + wd.x);  // s + i * w * d
        }
        fft_line<M, true>(v, lds, lay, t, twl + z0);
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const float r0 = v[i].x * a.scale, r1 = v[i].y * a.scale;
            if constexpr (EPI == XE_PLAIN) {
                acc[i] = make_float2(r0, r1);
            } else if constexpr (EPI == XE_OD_BOX || EPI == XE_OD_RAND) {
                if (q == 0) acc[i] = make_float2(r0, r1);
                else {
                    rr[i][0] = r0;  // randoms grid
                    rr[i][1] = r1;
                }
            } else if constexpr (EPI == XE_ITER_FIXED) {
                const float fac = a.fac[0];
                acc[i].x = acc[i].x - fac * r0;
                acc[i].y = acc[i].y - fac * r1;
            } else if constexpr (EPI == XE_ITER_LOCAL) {
                // q = 0: a (times x^2), 1: b (times x), 2: c; summed as
                // (x^2 a + x b) + c, then delta_s - sum/x^2 (src/iterative.jl:20-36)
                const int m = PI::out_pos(t, i);
                const float dv[2] = {r0, r1};
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const float xx = xvl[z0 + 2 * m + h];
                    if (q == 0) qs[i][h] = (xx * xx) * dv[h];
                    else if (q == 1) qs[i][h] += xx * dv[h];
                    else {
                        const float u = qs[i][h] + dv[h];
                        // 1/x^2 once per cell (the reference divides by
                        // x^2 per pair, src/iterative.jl:33-34; the product
                        // differs by at most 1 ulp)
                        const float s2 = (xx * xx + xy * xy) + xz * xz;
                        const float c = s2 > 0.f ? 1.0f / s2 : 0.f;
                        float *ac = &acc[i].x;
                        ac[h] = c > 0.f ? ac[h] - u * c : 0.f;
                    }
                }
            } else if constexpr (EPI == XE_DISP) {
                if (q == 0) { disp[i][0].x = r0; disp[i][1].x = r1; }
                else if (q == 1) { disp[i][0].y = r0; disp[i][1].y = r1; }
                else { disp[i][0].z = r0; disp[i][1].z = r1; }
            }
        }
        if (!RIN && q + 1 < NIN) {
#pragma unroll
            for (int i = 0; i < 8; i++) cur[i] = nxt[i];
            cn = nn;
        }
    }
    if constexpr (EPI == XE_DISP) {
        // 12-byte cells (Psi_x, Psi_y, Psi_z): cells 2m and 2m + 1 are three
        // float2 (rbase is even, so 3 * rbase floats is 8-byte aligned)
        if (!valid) return;
        float2 *o2 = reinterpret_cast<float2 *>(a.out + 3 * rbase);
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int m = PI::out_pos(t, i);
            const float3 d0 = disp[i][0], d1 = disp[i][1];
            o2[3 * m] = make_float2(d0.x, d0.y);
            o2[3 * m + 1] = make_float2(d0.z, d1.x);
            o2[3 * m + 2] = make_float2(d1.y, d1.z);
        }
        return;
    }
    if constexpr (EPI == XE_OD_BOX) {
        const double ncell = (double)a.nx * a.ny * a.nz;
        const float mean = (float)(a.slot[b].sum[0] / ncell);
#pragma unroll
        for (int i = 0; i < 8; i++) {
            acc[i].x = ((acc[i].x / mean) - 1.0f) / a.bias;
            acc[i].y = ((acc[i].y / mean) - 1.0f) / a.bias;
        }
    }
    if constexpr (EPI == XE_OD_RAND) {
        const Slot &sl = a.slot[b];
        const float alpha = (float)(sl.sum[0] / sl.sum[1]);
        const double thr = (double)a.ran_min * sl.sum[1] / sl.count_r;
        const float ba = a.bias * alpha;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            float *ac = &acc[i].x;
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const float d = ac[h] - alpha * rr[i][h];
                ac[h] = ((double)rr[i][h] > thr) ? d / (ba * rr[i][h]) : 0.f;
            }
        }
    }
    if (valid && a.out) {
        float2 *o2 = reinterpret_cast<float2 *>(a.out + rbase);
        if (EPI == XE_PLAIN && a.accumulate) {
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const float2 o = o2[PI::out_pos(t, i)];
                o2[PI::out_pos(t, i)] = make_float2(o.x + acc[i].x, o.y + acc[i].y);
            }
        } else {
#pragma unroll
            for (int i = 0; i < 8; i++) o2[PI::out_pos(t, i)] = acc[i];
        }
    }
    if constexpr (EPI != XE_DISP) {
        if (a.fwd_out) {
            // forward R2C of the line just produced: acc is in the forward
            // plan's input order (inverse plans end with the forward's first radix)
            using PF = Plan<M, false>;
            fft_line<M, false>(acc, lds, lay, t, twl);
            __syncthreads();
#pragma unroll
            for (int i = 0; i < 8; i++) lds[lay(PF::out_pos(t, i))] = acc[i];
            __syncthreads();
            if (valid) {
                float2 *dst = a.fwd_out + b * a.cstride + crow;
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    const int k = t + i * TPL;
                    float2 zk = lds[lay(k)];
                    float2 zc = cconj(lds[lay((M - k) & (M - 1))]);
                    float2 e = cscale(cadd(zk, zc), 0.5f);
                    float2 d = csub(zk, zc);
                    float2 o = make_float2(0.5f * d.y, -0.5f * d.x);
                    dst[k] = cadd(e, cmul(g_tw[2 * M + k], o));
                }
                if (t == 0) {
                    float2 z0 = lds[lay(0)];
                    dst[M] = make_float2(z0.x - z0.y, 0.f);
                }
            }
        }
    }
}

// ------------------------------------------------------------ strided y / z

// Columns per workgroup: 8 (64-byte row segments, 512 threads at n = 512)
// lets two workgroups share a CU (<= 128 VGPRs), so one workgroup's loads
// overlap the other's LDS exchanges; at n = 512 that measured faster than 16
// columns with one 1024-thread workgroup per CU.
template <int M>
struct SShape {
    static constexpr int B = M >= 512 ? 8 : 16;  // kx columns per workgroup
    static constexpr int NT = B * M / 8;
};

// Workgroup i is dispatched to XCD i % 8: renumber so that consecutive
// logical tiles share an XCD (and its L2).
__device__ __forceinline__ int xcd_tile(int bid, int nb) {
    const int q = nb >> 3, r = nb & 7, x = bid & 7;
    return x * q + (x < r ? x : r) + (bid >> 3);
}

// MK >= 0: the multiplier kind as a template constant (a.mul is not read)
template <int MK = -1>
__device__ __forceinline__ float2 apply_mul(float2 F, float kx, float ky, float kz, const StridedArgs &a) {
    const float R2 = a.R2, l0 = a.los[0], l1 = a.los[1], l2 = a.los[2];
    const float k2 = (kx * kx + ky * ky) + kz * kz;
    switch (MK >= 0 ? MK : a.mul) {
        case MUL_GAUSS: {
            const double g = exp(-0.5 * (double)R2 * (double)k2);
            return make_float2((float)((double)F.x * g), (float)((double)F.y * g));
        }
        case MUL_INVK2:
            return k2 > 0.f ? make_float2(F.x / k2, F.y / k2) : make_float2(0.f, 0.f);
        case MUL_FIXEDLOS: {
            if (!(k2 > 0.f)) return make_float2(0.f, 0.f);
            const float m = ((kx * kx) * l0 + (ky * ky) * l1) + (kz * kz) * l2;
            return cscale(make_float2(F.x / k2, F.y / k2), m);
        }
        case MUL_I_INVK2:
            return k2 > 0.f ? make_float2(-F.y / k2, F.x / k2) : make_float2(0.f, 0.f);
        case MUL_I:
            return make_float2(-F.y, F.x);
        case MUL_GAUSS1: {
            const float ka = a.axis == 1 ? ky : kz;
            const double g = exp(-0.5 * (double)R2 * (double)(ka * ka));
            return make_float2((float)((double)F.x * g), (float)((double)F.y * g));
        }
        case MUL_MG_INV: {
            // exact inverse of the discrete operator for a fixed LOS (the
            // multigrid stencil has constant coefficients then)
            // t_a in [-pi, pi]: the hardware sine and cosine (a seed for the
            // V-cycles, which remove its rounding)
            const float tx = kx * a.mg_cell[0], ty = ky * a.mg_cell[1], tz = kz * a.mg_cell[2];
            const float sx = __sinf(tx), sy = __sinf(ty), sz = __sinf(tz);
            const float A = (a.mg_gp[0] * (2.0f - 2.0f * __cosf(tx)) + a.mg_gp[1] * (2.0f - 2.0f * __cosf(ty))) +
                            a.mg_gp[2] * (2.0f - 2.0f * __cosf(tz)) +
                            ((a.mg_cross[0] * sx * sy + a.mg_cross[1] * sx * sz) + a.mg_cross[2] * sy * sz);
            return A > 0.f ? make_float2(F.x / A, F.y / A) : make_float2(0.f, 0.f);
        }
        default:
            return F;
    }
}

template <int M>
__global__ void __launch_bounds__(SShape<M>::NT) __attribute__((amdgpu_waves_per_eu(4)))
k_strided(StridedArgs a) {
    constexpr int B = SShape<M>::B;
    using PF = Plan<M, false>;
    using PI = Plan<M, true>;
    __shared__ float2 lds[lay_strided_size(M, B)];
    const int tid = threadIdx.x;
    const int col = tid % B, t = tid / B;
    const int ncg = (a.Pv + B - 1) / B;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int cg = tile % ncg, other = tile / ncg;
    const int kxi = cg * B + col;
    const bool valid = kxi < a.Pv;
    const StridedJob &J = a.job[blockIdx.y];
    const int64_t b = blockIdx.z;
    int64_t base, rs;
    if (a.axis == 1) {
        base = (int64_t)other * a.PZ + kxi;
        rs = a.P;
    } else {
        base = (int64_t)other * a.P + kxi;
        rs = a.PZ;
    }
    base += b * a.cstride;
    const float *kt = a.ktab + b * a.ktabStride;
    const float *kax = a.axis == 1 ? kt + a.kyo : kt + a.kyo + a.ny;
    LayStrided lay{col, B};
    float2 v[8];

    if (a.mode == 0) {
#pragma unroll
        for (int i = 0; i < 8; i++)
            v[i] = valid ? J.in[base + PF::in_pos(t, i) * rs] : make_float2(0.f, 0.f);
        fft_line<M, false>(v, lds, lay, t, g_tw);
        if (valid) {
#pragma unroll
            for (int i = 0; i < 8; i++) J.out[0][base + PF::out_pos(t, i) * rs] = v[i];
        }
        return;
    }
    if (a.mode == 2) {
#pragma unroll
        for (int i = 0; i < 8; i++)
            v[i] = valid ? J.in[base + PF::in_pos(t, i) * rs] : make_float2(0.f, 0.f);
        fft_line<M, false>(v, lds, lay, t, g_tw);
        const float kxv = valid ? kt[kxi] : 0.f;
        const float kother = a.axis == 1 ? kt[a.kyo + a.ny + other] : kt[a.kyo + other];
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const float kal = kax[PF::out_pos(t, i)];
            const float ky = a.axis == 1 ? kal : kother;
            const float kz = a.axis == 1 ? kother : kal;
            v[i] = apply_mul(v[i], kxv, ky, kz, a);
        }
    } else {
#pragma unroll
        for (int i = 0; i < 8; i++)
            v[i] = valid ? J.in[base + PI::in_pos(t, i) * rs] : make_float2(0.f, 0.f);
    }
    float2 f[8];
    float kal[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        f[i] = v[i];
        kal[i] = kax[PI::in_pos(t, i)];
    }
    TwCache<M, true> tc;
    tc.load(t, g_tw);
    for (int o = 0; o < J.nout; o++) {
        const int pw = J.pw[o];
#pragma unroll
        for (int i = 0; i < 8; i++) v[i] = pw ? cscale(f[i], kpow(kal[i], pw)) : f[i];
        fft_line_c<M, true>(v, lds, lay, t, tc);
        if (valid) {
            float2 *dst = J.out[o];
#pragma unroll
            for (int i = 0; i < 8; i++) dst[base + PI::out_pos(t, i) * rs] = v[i];
        }
    }
}

// The same pass for jobs with one output, with 16 columns per workgroup and
// two adjacent columns per thread (M threads): every row access of a wave is
// one float4 per thread, 128 contiguous bytes per row, and the two columns
// share twiddles and barriers.  Columns past Pv (inside the row pitch) are
// transformed and stored like the others; nothing reads them.
// MODE: a.mode as a template constant; MK: the multiplier kind of mode 2, or
// -1 to read a.mul
template <int M, int MODE, int MK>
__global__ void __launch_bounds__(M) __attribute__((amdgpu_waves_per_eu(4)))
k_strided2(StridedArgs a) {
    constexpr int B = 16;
    using PF = Plan<M, false>;
    using PI = Plan<M, true>;
    __shared__ float4 lds[lay_strided_size(M, B) / 2];
    const int tid = threadIdx.x;
    const int pc = tid % (B / 2), t = tid / (B / 2);
    const int ncg = (a.Pv + B - 1) / B;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int cg = tile % ncg, other = tile / ncg;
    const int kx0 = cg * B + 2 * pc;
    const bool valid = kx0 < a.Pv;
    const StridedJob &J = a.job[blockIdx.y];
    const int64_t b = blockIdx.z;
    int64_t base, rs;
    if (a.axis == 1) {
        base = (int64_t)other * a.PZ + kx0;
        rs = a.P;
    } else {
        base = (int64_t)other * a.P + kx0;
        rs = a.PZ;
    }
    base += b * a.cstride;
    const float *kt = a.ktab + b * a.ktabStride;
    const float *kax = a.axis == 1 ? kt + a.kyo : kt + a.kyo + a.ny;
    const LayPair lay{2 * pc, B};
    float2 v[2][8];
    auto load = [&](const float2 *src, int pos, int i) {
        const float4 c = valid ? *reinterpret_cast<const float4 *>(src + base + pos * rs)
                               : make_float4(0.f, 0.f, 0.f, 0.f);
        v[0][i] = make_float2(c.x, c.y);
        v[1][i] = make_float2(c.z, c.w);
    };
    auto store = [&](float2 *dst, int pos, const float2 &x0, const float2 &x1) {
        *reinterpret_cast<float4 *>(dst + base + pos * rs) = make_float4(x0.x, x0.y, x1.x, x1.y);
    };

    if constexpr (MODE == 0) {
#pragma unroll
        for (int i = 0; i < 8; i++) load(J.in, PF::in_pos(t, i), i);
        fft_line2<M, false>(v, lds, lay, t, g_tw);
        if (valid) {
#pragma unroll
            for (int i = 0; i < 8; i++) store(J.out[0], PF::out_pos(t, i), v[0][i], v[1][i]);
        }
        return;
    }
    if constexpr (MODE == 2) {
#pragma unroll
        for (int i = 0; i < 8; i++) load(J.in, PF::in_pos(t, i), i);
        fft_line2<M, false>(v, lds, lay, t, g_tw);
        const float kx0v = valid ? kt[kx0] : 0.f;
        const float kx1v = kx0 + 1 < a.Pv ? kt[kx0 + 1] : 0.f;
        const float kother = a.axis == 1 ? kt[a.kyo + a.ny + other] : kt[a.kyo + other];
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const float kal = kax[PF::out_pos(t, i)];
            const float ky = a.axis == 1 ? kal : kother;
            const float kz = a.axis == 1 ? kother : kal;
            v[0][i] = apply_mul<MK>(v[0][i], kx0v, ky, kz, a);
            v[1][i] = apply_mul<MK>(v[1][i], kx1v, ky, kz, a);
        }
    } else {
#pragma unroll
        for (int i = 0; i < 8; i++) load(J.in, PI::in_pos(t, i), i);
    }
    // one output (the launcher uses this kernel only then), k-power pw
    const int pw = J.pw[0];
    if (pw) {
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const float m = kpow(kax[PI::in_pos(t, i)], pw);
            v[0][i] = cscale(v[0][i], m);
            v[1][i] = cscale(v[1][i], m);
        }
    }
    TwCache<M, true> tc;
    tc.load(t, g_tw);
    fft_line2_c<M, true>(v, lds, lay, t, tc);
    if (valid) {
#pragma unroll
        for (int i = 0; i < 8; i++) store(J.out[0], PI::out_pos(t, i), v[0][i], v[1][i]);
    }
}

// ------------------------------------------------------------ y: local-LOS iteration

// The pointer as an opaque value: table loads through it are not hoisted
// across the transforms that use it (which would keep them all live).
template <class T>
__device__ __forceinline__ T *opaque(T *p) {
    asm volatile("" : "+s"(p));
    return p;
}

template <int M>
__global__ void __launch_bounds__(SShape<M>::NT) __attribute__((amdgpu_waves_per_eu(4)))
k_y_iter(YIterArgs a) {
    constexpr int B = SShape<M>::B;
    using PI = Plan<M, true>;
    __shared__ float2 lds[lay_strided_size(M, B)];
    const int tid = threadIdx.x;
    const int col = tid % B, t = tid / B;
    const int ncg = (a.Pv + B - 1) / B;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int cg = tile % ncg, iz = tile / ncg;
    const int kxi = cg * B + col;
    const bool valid = kxi < a.Pv;
    const int64_t b = blockIdx.z;
    const int64_t base = b * a.cstride + (int64_t)iz * a.PZ + kxi;
    const int64_t rs = a.P;
    const float *ky = a.ktab + b * a.ktabStride + a.Pv;
    const float *xv = a.xtab + b * a.xtabStride;
    const float zc = xv[a.nx + a.ny + iz];
    LayStrided lay{col, B};
    TwCache<M, true> tc;
    tc.load(t, g_tw);
    const float fxx = a.fac[0], fxy = a.fac[1], fxz = a.fac[2];
    const float fyy = a.fac[3], fyz = a.fac[4], fzz = a.fac[5];
    // in: v = G[i] * ky^p (in the inverse plan's input order)
    auto load = [&](const float2 *g0, int p, float2 (&v)[8]) {
        // opaque: each transform loads its own input (no reuse across
        // transforms that would keep a whole line live)
        const float2 *g = opaque(g0);
        const float *kyo = opaque(ky);
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int r = PI::in_pos(t, i);
            const float2 x = valid ? g[base + r * rs] : make_float2(0.f, 0.f);
            const float k = kyo[r];
            v[i] = p == 0 ? x : cscale(x, p == 1 ? k : k * k);
        }
    };
    // The six transforms in an order that keeps at most two accumulators and
    // one line live: G0 terms of A, B and C; then G1 (read twice: for B and
    // for C); then G2.
    float2 v[8], accb[8], accc[8];
    // A = fxx IFFT(G0)
    load(a.g[0], 0, v);
    fft_line_c<M, true>(v, lds, lay, t, tc);
    if (valid) {
#pragma unroll
        for (int i = 0; i < 8; i++) a.out[0][base + PI::out_pos(t, i) * rs] = cscale(v[i], fxx);
    }
    // B = fxy y IFFT(ky G0) + ...
    load(a.g[0], 1, v);
    fft_line_c<M, true>(v, lds, lay, t, tc);
#pragma unroll
    for (int i = 0; i < 8; i++) accb[i] = cscale(v[i], fxy * opaque(xv)[a.nx + PI::out_pos(t, i)]);
    // C = fyy y^2 IFFT(ky^2 G0) + ...
    load(a.g[0], 2, v);
    fft_line_c<M, true>(v, lds, lay, t, tc);
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const float yc = opaque(xv)[a.nx + PI::out_pos(t, i)];
        accc[i] = cscale(v[i], fyy * (yc * yc));
    }
    // B += fxz z IFFT(G1)
    load(a.g[1], 0, v);
    fft_line_c<M, true>(v, lds, lay, t, tc);
    if (valid) {
#pragma unroll
        for (int i = 0; i < 8; i++)
            a.out[1][base + PI::out_pos(t, i) * rs] = cadd(accb[i], cscale(v[i], fxz * zc));
    }
    // C += fyz y z IFFT(ky G1)
    load(a.g[1], 1, v);
    fft_line_c<M, true>(v, lds, lay, t, tc);
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const float yc = opaque(xv)[a.nx + PI::out_pos(t, i)];
        accc[i] = cadd(accc[i], cscale(v[i], fyz * (yc * zc)));
    }
    // C += fzz z^2 IFFT(G2)
    load(a.g[2], 0, v);
    fft_line_c<M, true>(v, lds, lay, t, tc);
    if (valid) {
#pragma unroll
        for (int i = 0; i < 8; i++)
            a.out[2][base + PI::out_pos(t, i) * rs] = cadd(accc[i], cscale(v[i], fzz * (zc * zc)));
    }
}

// ------------------------------------------------------------ launchers

template <int M>
static hipError_t y_iter_m(const YIterArgs &a, int nbatch, hipStream_t s) {
    using S = SShape<M>;
    const int ncg = (a.Pv + S::B - 1) / S::B;
    dim3 grid((unsigned)(ncg * a.nz), 1, nbatch);
    hipLaunchKernelGGL(k_y_iter<M>, grid, dim3(S::NT), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_y_iter(const YIterArgs &a, int nbatch, hipStream_t s) {
    switch (a.ny) {
        case 16: return y_iter_m<16>(a, nbatch, s);
        case 32: return y_iter_m<32>(a, nbatch, s);
        case 64: return y_iter_m<64>(a, nbatch, s);
        case 128: return y_iter_m<128>(a, nbatch, s);
        case 256: return y_iter_m<256>(a, nbatch, s);
        case 512: return y_iter_m<512>(a, nbatch, s);
        case 1024: return y_iter_m<1024>(a, nbatch, s);
        default: return hipErrorInvalidValue;
    }
}

template <int M>
static hipError_t x_r2c_m(const XFwdArgs &a, int nbatch, hipStream_t s) {
    using S = XShape<M>;
    const int64_t nlines = (int64_t)a.ny * a.nz;
    const unsigned nwg = (unsigned)((nlines + S::LPW - 1) / S::LPW);
    dim3 grid(nwg, nbatch);
    hipLaunchKernelGGL(k_x_r2c<M>, grid, dim3(S::NT), 0, s, a);
    return hipGetLastError();
}


hipError_t launch_x_r2c(const XFwdArgs &a, int nbatch, hipStream_t s) {
    switch (a.nx / 2) {
        case 8: return x_r2c_m<8>(a, nbatch, s);
        case 16: return x_r2c_m<16>(a, nbatch, s);
        case 32: return x_r2c_m<32>(a, nbatch, s);
        case 64: return x_r2c_m<64>(a, nbatch, s);
        case 128: return x_r2c_m<128>(a, nbatch, s);
        case 256: return x_r2c_m<256>(a, nbatch, s);
        case 512: return x_r2c_m<512>(a, nbatch, s);
        default: return hipErrorInvalidValue;
    }
}

template <int M, int EPI>
static hipError_t x_c2r_me(const XInvArgs &a, int nbatch, hipStream_t s) {
    using S = XShape<M>;
    const int64_t nlines = (int64_t)a.ny * a.nz;
    dim3 grid((unsigned)((nlines + S::LPW - 1) / S::LPW), nbatch);
    if constexpr (EPI == XE_PLAIN || EPI == XE_OD_BOX || EPI == XE_OD_RAND) {
        if (a.rin[0]) {
            hipLaunchKernelGGL((k_x_c2r<M, EPI, true>), grid, dim3(S::NT), 0, s, a);
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL((k_x_c2r<M, EPI>), grid, dim3(S::NT), 0, s, a);
    return hipGetLastError();
}

template <int M>
static hipError_t x_c2r_m(const XInvArgs &a, int epi, int nbatch, hipStream_t s) {
    switch (epi) {
        case XE_PLAIN: return x_c2r_me<M, XE_PLAIN>(a, nbatch, s);
        case XE_OD_BOX: return x_c2r_me<M, XE_OD_BOX>(a, nbatch, s);
        case XE_OD_RAND: return x_c2r_me<M, XE_OD_RAND>(a, nbatch, s);
        case XE_ITER_LOCAL: return x_c2r_me<M, XE_ITER_LOCAL>(a, nbatch, s);
        case XE_ITER_FIXED: return x_c2r_me<M, XE_ITER_FIXED>(a, nbatch, s);
        case XE_DISP: return x_c2r_me<M, XE_DISP>(a, nbatch, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_x_c2r(const XInvArgs &a, int epi, int nbatch, hipStream_t s) {
    switch (a.nx / 2) {
        case 8: return x_c2r_m<8>(a, epi, nbatch, s);
        case 16: return x_c2r_m<16>(a, epi, nbatch, s);
        case 32: return x_c2r_m<32>(a, epi, nbatch, s);
        case 64: return x_c2r_m<64>(a, epi, nbatch, s);
        case 128: return x_c2r_m<128>(a, epi, nbatch, s);
        case 256: return x_c2r_m<256>(a, epi, nbatch, s);
        case 512: return x_c2r_m<512>(a, epi, nbatch, s);
        default: return hipErrorInvalidValue;
    }
}

bool strided_uses_pair(const StridedArgs &a, int njobs) {
    if (a.pair == 0) return false;
    const int M = a.axis == 1 ? a.ny : a.nz;
    bool single = true;
    for (int j = 0; j < njobs; j++) single = single && a.job[j].nout == 1;
    return M >= 256 && M <= 1024 && single;
}

template <int M>
static hipError_t strided_m(const StridedArgs &a, int njobs, int nbatch, hipStream_t s) {
    using S = SShape<M>;
    const int nother = a.axis == 1 ? a.nz : a.ny;
    if constexpr (M >= 256) {
        if (strided_uses_pair(a, njobs)) {
            const int ncg = (a.Pv + 15) / 16;
            dim3 grid((unsigned)(ncg * nother), njobs, nbatch);
            using K = decltype(&k_strided2<M, 0, -1>);
            static const K fused[8] = {k_strided2<M, 2, MUL_NONE>,     k_strided2<M, 2, MUL_GAUSS>,
                                       k_strided2<M, 2, MUL_INVK2>,    k_strided2<M, 2, MUL_FIXEDLOS>,
                                       k_strided2<M, 2, MUL_I_INVK2>,  k_strided2<M, 2, MUL_I>,
                                       k_strided2<M, 2, MUL_MG_INV>,   k_strided2<M, 2, MUL_GAUSS1>};
            if (a.mode == 2 && (a.mul < 0 || a.mul > 7)) return hipErrorInvalidValue;
            const K kernel = a.mode == 0 ? k_strided2<M, 0, -1> : (a.mode == 2 ? fused[a.mul] : k_strided2<M, 1, -1>);
            hipLaunchKernelGGL(kernel, grid, dim3(M), 0, s, a);
            return hipGetLastError();
        }
    }
    const int ncg = (a.Pv + S::B - 1) / S::B;
    dim3 grid((unsigned)(ncg * nother), njobs, nbatch);
    hipLaunchKernelGGL(k_strided<M>, grid, dim3(S::NT), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_strided(const StridedArgs &a, int njobs, int nbatch, hipStream_t s) {
    const int M = a.axis == 1 ? a.ny : a.nz;
    switch (M) {
        case 16: return strided_m<16>(a, njobs, nbatch, s);
        case 32: return strided_m<32>(a, njobs, nbatch, s);
        case 64: return strided_m<64>(a, njobs, nbatch, s);
        case 128: return strided_m<128>(a, njobs, nbatch, s);
        case 256: return strided_m<256>(a, njobs, nbatch, s);
        case 512: return strided_m<512>(a, njobs, nbatch, s);
        case 1024: return strided_m<1024>(a, njobs, nbatch, s);
        default: return hipErrorInvalidValue;
    }
}

void set_twiddle_table(const float2 *host_tw) {
    (void)hipMemcpyToSymbol(HIP_SYMBOL(g_tw), host_tw, sizeof(float2) * 4096, 0,
                            hipMemcpyHostToDevice);
}

// ------------------------------------------------------------ tables

// k_vec (src/utils.jl:3-10) and x_vec (src/utils.jl:21-25) per batch item.
__global__ void k_tables(const Slot *slot, float *ktab, float *xtab, int nx, int ny, int nz) {
    const int b = blockIdx.y;
    const int P = nx / 2 + 1;
    const Slot &sl = slot[b];
    float *kt = ktab + (int64_t)b * (P + ny + nz);
    float *xt = xtab + (int64_t)b * (nx + ny + nz);
    const int n3[3] = {nx, ny, nz};
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < P + ny + nz;
         idx += gridDim.x * blockDim.x) {
        int ax, i;
        if (idx < P) { ax = 0; i = idx; }
        else if (idx < P + ny) { ax = 1; i = idx - P; }
        else { ax = 2; i = idx - P - ny; }
        const int n = n3[ax];
        const float L = sl.box_size[ax];
        const float fs = (float)((2.0 * M_PI * (double)n) / (double)L);
        const float mult = fs / (float)n;
        const int ii = (ax == 0) ? i : (i < (n + 1) / 2 ? i : i - n);
        kt[idx] = (float)ii * mult;
    }
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < nx + ny + nz;
         idx += gridDim.x * blockDim.x) {
        int ax, i;
        if (idx < nx) { ax = 0; i = idx; }
        else if (idx < nx + ny) { ax = 1; i = idx - nx; }
        else { ax = 2; i = idx - nx - ny; }
        const float h = sl.box_size[ax] / (float)n3[ax];
        xt[idx] = (float)((double)sl.box_min[ax] + ((double)i + 0.5) * (double)h);
    }
}

hipError_t launch_tables(const Slot *slot, float *ktab, float *xtab, int nx, int ny, int nz,
                         int nbatch, hipStream_t s) {
    hipLaunchKernelGGL(k_tables, dim3(4, nbatch), dim3(256), 0, s, slot, ktab, xtab, nx, ny, nz);
    return hipGetLastError();
}

}  // namespace baorec
