// Power-spectrum multipoles of a periodic box from the half spectra of one
// or two (interlaced) mass-assignment grids: the estimator the reference
// validates its reconstructions with (test_helpers/powspec_auto.conf:
// TSC, grid interlacing, multipoles 0, 2, 4, linear k bins).
//
// Per mode k of the half spectrum (Hermitian weight 2 except the kx = 0 and
// kx = Nyquist planes): delta = (F1 + F2 e^{i k.H/2}) / 2 with interlacing
// (grid 2 holds the particles shifted by half a cell), divided by the
// assignment window prod_a sinc(k_a H_a / 2)^p; then |delta|^2 times
// L_0, L_2, L_4 of mu = k.los / |k| summed per bin of |k|.
#include <math.h>

#include "kernels.hpp"

namespace baorec {

constexpr int kPkThreads = 256;

// Partial bin sums of this workgroup: part[blk][bin][q], q = (modes, sum k,
// sum P, sum P L2, sum P L4); accumulated with LDS FP64 atomics.
__global__ void __launch_bounds__(kPkThreads) k_pk_modes(PkArgs a) {
    extern __shared__ double sb[];
    const int nq = 5 * a.nbins;
    for (int i = threadIdx.x; i < nq; i += blockDim.x) sb[i] = 0.0;
    __syncthreads();
    const int64_t nmodes = (int64_t)a.Pv * a.ny * a.nz;
    const int nyq = a.nx / 2;
    for (int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; m < nmodes;
         m += (int64_t)gridDim.x * blockDim.x) {
        const int kx = (int)(m % a.Pv);
        const int64_t t = m / a.Pv;
        const int ky = (int)(t % a.ny), kz = (int)(t / a.ny);
        const float4 tx = a.tab[kx], ty = a.tab[a.Pv + ky], tz = a.tab[a.Pv + a.ny + kz];
        // tab: (cos, sin) of k_a H_a / 2 and the window factor per axis;
        // |k| in FP64 from the signed frequencies (negative Nyquist on y, z)
        const double ka = kx * a.kf[0];
        const double kb = (ky < a.ny / 2 ? ky : ky - a.ny) * a.kf[1];
        const double kc = (kz < a.nz / 2 ? kz : kz - a.nz) * a.kf[2];
        const double k2 = (ka * ka + kb * kb) + kc * kc;
        if (!(k2 > 0.0)) continue;
        const double k = sqrt(k2);
        const double bf = floor((k - a.kmin) / a.dk);
        if (!(bf >= 0.0 && bf < (double)a.nbins)) continue;
        const int bin = (int)bf;
        const int64_t I = kx + (int64_t)a.P * ky + a.PZ * kz;
        float2 d = a.F1[I];
        if (a.interlace) {
            // e^{i k.H/2} = product of the per-axis factors
            const float2 pxy = make_float2(tx.x * ty.x - tx.y * ty.y, tx.x * ty.y + tx.y * ty.x);
            const float2 ph = make_float2(pxy.x * tz.x - pxy.y * tz.y, pxy.x * tz.y + pxy.y * tz.x);
            const float2 f2 = a.F2[I];
            d = make_float2(0.5f * (d.x + (f2.x * ph.x - f2.y * ph.y)), 0.5f * (d.y + (f2.x * ph.y + f2.y * ph.x)));
        }
        const float win = (tx.z * ty.z) * tz.z;
        const double p = ((double)d.x * d.x + (double)d.y * d.y) / ((double)win * win);
        const double mu = (ka * a.los[0] + kb * a.los[1] + kc * a.los[2]) / k;
        const double mu2 = mu * mu;
        const double l2 = 1.5 * mu2 - 0.5;
        const double l4 = (35.0 * mu2 * mu2 - 30.0 * mu2 + 3.0) / 8.0;
        const double wgt = (kx == 0 || kx == nyq) ? 1.0 : 2.0;
        double *b = sb + 5 * bin;
        atomicAdd(b + 0, wgt);
        atomicAdd(b + 1, wgt * k);
        atomicAdd(b + 2, wgt * p);
        atomicAdd(b + 3, wgt * p * l2);
        atomicAdd(b + 4, wgt * p * l4);
    }
    __syncthreads();
    double *out = a.part + (int64_t)blockIdx.x * nq;
    for (int i = threadIdx.x; i < nq; i += blockDim.x) out[i] = sb[i];
}

// out[i] = sum over workgroups in order (deterministic across workgroups)
__global__ void k_pk_reduce(const double *__restrict__ part, double *__restrict__ out, int nq, int nblk) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    double s = 0.0;
    for (int b = 0; b < nblk; b++) s += part[(int64_t)b * nq + i];
    out[i] = s;
}

int pk_blocks() { return 1024; }

hipError_t launch_pk(const PkArgs &a, double *out, hipStream_t s) {
    const int nblk = pk_blocks();
    const size_t lds = sizeof(double) * 5 * a.nbins;
    hipLaunchKernelGGL(k_pk_modes, dim3(nblk), dim3(kPkThreads), lds, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int nq = 5 * a.nbins;
    hipLaunchKernelGGL(k_pk_reduce, dim3((nq + 255) / 256), dim3(256), 0, s, a.part, out, nq, nblk);
    return hipGetLastError();
}

}  // namespace baorec
