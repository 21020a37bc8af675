// Shared device helpers for libbaorec_hip (gfx950).
//
// Grid layout (all kernels): real grids are float, x fastest
// (i + nx*(j + ny*k)); the half-spectrum of an R2C transform along x is
// float2 with pitch P = nx/2+1: (kx + P*(ky + ny*kz)).  Every kernel works on
// a batch of catalogs: batch item b of a real grid starts at b*nx*ny*nz, of a
// complex grid at b*P*ny*nz.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace baorec {

constexpr int kMaxSide = 1024;

__host__ __device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    // FFT arithmetic only: fused multiply-adds are allowed here (the
    // reference's FFT library rounding is not part of its semantics)
#pragma clang fp contract(fast)
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__host__ __device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__host__ __device__ __forceinline__ float2 mul_mi(float2 a) {
    return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// Per-catalog scalars that stay on the device so that a whole
// reconstruction runs without host round trips.
struct Slot {
    float box_size[3];
    float box_min[3];
    double sum[2];     // FP64 totals of the scattered grids (data, randoms)
    double count_r;    // number of randoms (threshold uses N_r, src/recon.jl:81)
    int32_t err;       // nonzero: a particle indexed outside the grid
    int32_t pad_;
    float minmax[6];   // setup_box reduction scratch (min xyz, max xyz)
};

// ------------------------------------------------------------------ FFT core
//
// A length-M complex FFT (M = 2^m, 8 <= M <= 1024) is split over M/8 threads,
// each holding 8 values in registers.  Stages are Stockham radix-8 with one
// radix-2/4 stage when log2(M) is not a multiple of 3; between stages values
// are exchanged through LDS.  Forward plans put the odd radix first, inverse
// plans last, so the output register pattern of a forward transform is the
// input pattern of an inverse one (fused forward * multiplier * inverse).
//
// Twiddles come from one table: tw[S + i] = exp(-2 pi i * i / S) for every
// power of two S <= 2048 (4096 float2 in total).

template <int M, bool INV>
struct Plan {
    static constexpr int LOG = (M == 8) ? 3 : (M == 16) ? 4 : (M == 32) ? 5 : (M == 64) ? 6
                             : (M == 128) ? 7 : (M == 256) ? 8 : (M == 512) ? 9 : 10;
    static constexpr int NST = (LOG + 2) / 3;
    static constexpr int REM = (LOG % 3 == 0) ? 8 : ((LOG % 3 == 1) ? 2 : 4);
    static constexpr int radix(int st) {
        return INV ? (st == NST - 1 ? REM : 8) : (st == 0 ? REM : 8);
    }
    static constexpr int ns(int st) {
        int p = 1;
        for (int i = 0; i < st; i++) p *= radix(i);
        return p;
    }
    // element index held in register i by thread t before the first stage
    static __device__ __forceinline__ int in_pos(int t, int i) {
        constexpr int R0 = radix(0);
        return t + (i / R0) * (M / 8) + (i % R0) * (M / R0);
    }
    // element index held in register i by thread t after the last stage
    static __device__ __forceinline__ int out_pos(int t, int i) {
        constexpr int RL = radix(NST - 1);
        return t + (i / RL) * (M / 8) + (i % RL) * (M / RL);
    }
};

template <int R, bool INV>
__device__ __forceinline__ void dft(float2 *a) {
    if constexpr (R == 2) {
        float2 t0 = a[0], t1 = a[1];
        a[0] = cadd(t0, t1);
        a[1] = csub(t0, t1);
    } else if constexpr (R == 4) {
        float2 s02 = cadd(a[0], a[2]), d02 = csub(a[0], a[2]);
        float2 s13 = cadd(a[1], a[3]), d13 = mul_mi<INV>(csub(a[1], a[3]));
        a[0] = cadd(s02, s13);
        a[2] = csub(s02, s13);
        a[1] = cadd(d02, d13);
        a[3] = csub(d02, d13);
    } else {
        // radix 8 = 2 x radix 4 on even/odd samples, combined with W8^k
        float2 e[4] = {a[0], a[2], a[4], a[6]};
        float2 o[4] = {a[1], a[3], a[5], a[7]};
        dft<4, INV>(e);
        dft<4, INV>(o);
        const float r = 0.70710678118654752f;
        // W8^1 = (1 - i)/sqrt2 forward, (1 + i)/sqrt2 inverse
        float2 o1 = INV ? make_float2((o[1].x - o[1].y) * r, (o[1].x + o[1].y) * r)
                        : make_float2((o[1].x + o[1].y) * r, (o[1].y - o[1].x) * r);
        float2 o2 = mul_mi<INV>(o[2]);
        float2 o3 = INV ? make_float2(-(o[3].x + o[3].y) * r, (o[3].x - o[3].y) * r)
                        : make_float2((o[3].y - o[3].x) * r, -(o[3].x + o[3].y) * r);
        a[0] = cadd(e[0], o[0]);
        a[4] = csub(e[0], o[0]);
        a[1] = cadd(e[1], o1);
        a[5] = csub(e[1], o1);
        a[2] = cadd(e[2], o2);
        a[6] = csub(e[2], o2);
        a[3] = cadd(e[3], o3);
        a[7] = csub(e[3], o3);
    }
}

// Runs all stages of a length-M FFT for one line.  v: this thread's 8 values
// in Plan::in_pos order on entry, Plan::out_pos order on exit.  lay(pos) maps
// an element index of this thread's line to its LDS slot.  Every thread of the
// workgroup must call this (it contains __syncthreads()).
template <int M, bool INV, int ST, class Lay>
__device__ __forceinline__ void fft_stages(float2 (&v)[8], float2 *lds, const Lay &lay, int t,
                                           const float2 *__restrict__ tw) {
    using P = Plan<M, INV>;
    constexpr int R = P::radix(ST);
    constexpr int NS = P::ns(ST);
    constexpr int NB = 8 / R;
    if constexpr (ST > 0) {
        constexpr int RP = P::radix(ST - 1);
        constexpr int NSP = P::ns(ST - 1);
        constexpr int NBP = 8 / RP;
        __syncthreads();
#pragma unroll
        for (int s = 0; s < NBP; s++) {
            int j = t + s * (M / 8);
            int k = j & (NSP - 1);
            int base = (j - k) * RP + k;
#pragma unroll
            for (int q = 0; q < RP; q++) lds[lay(base + q * NSP)] = v[s * RP + q];
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < NB; s++) {
            int j = t + s * (M / 8);
#pragma unroll
            for (int r = 0; r < R; r++) v[s * R + r] = lds[lay(j + r * (M / R))];
        }
    }
#pragma unroll
    for (int s = 0; s < NB; s++) {
        int j = t + s * (M / 8);
        if constexpr (NS > 1) {
            int k = j & (NS - 1);
#pragma unroll
            for (int r = 1; r < R; r++) {
                float2 w = tw[NS * R + r * k];
                if (INV) w.y = -w.y;
                v[s * R + r] = cmul(v[s * R + r], w);
            }
        }
        dft<R, INV>(&v[s * R]);
    }
    if constexpr (ST + 1 < P::NST) fft_stages<M, INV, ST + 1>(v, lds, lay, t, tw);
}

template <int M, bool INV, class Lay>
__device__ __forceinline__ void fft_line(float2 (&v)[8], float2 *lds, const Lay &lay, int t,
                                         const float2 *__restrict__ tw) {
    fft_stages<M, INV, 0>(v, lds, lay, t, tw);
}

// Per-thread twiddles of every stage after the first, loaded once and reused
// by kernels that run several transforms of the same length.
template <int M, bool INV>
struct TwCache {
    using P = Plan<M, INV>;
    static constexpr int NS1 = P::NST > 1 ? P::NST - 1 : 1;
    float2 w[NS1][7];
    __device__ __forceinline__ void load(int t, const float2 *__restrict__ tw) {
        load_st<1>(t, tw);
    }
    template <int ST>
    __device__ __forceinline__ void load_st(int t, const float2 *__restrict__ tw) {
        if constexpr (ST < P::NST) {
            constexpr int R = P::radix(ST), NS = P::ns(ST), NB = 8 / R;
#pragma unroll
            for (int s = 0; s < NB; s++) {
                const int k = (t + s * (M / 8)) & (NS - 1);
#pragma unroll
                for (int r = 1; r < R; r++) {
                    float2 x = tw[NS * R + r * k];
                    if (INV) x.y = -x.y;
                    w[ST - 1][s * (R - 1) + r - 1] = x;
                }
            }
            load_st<ST + 1>(t, tw);
        }
    }
};

template <int M, bool INV, int ST, class Lay>
__device__ __forceinline__ void fft_stages_c(float2 (&v)[8], float2 *lds, const Lay &lay, int t,
                                             const TwCache<M, INV> &tc) {
    using P = Plan<M, INV>;
    constexpr int R = P::radix(ST);
    constexpr int NS = P::ns(ST);
    constexpr int NB = 8 / R;
    if constexpr (ST > 0) {
        constexpr int RP = P::radix(ST - 1);
        constexpr int NSP = P::ns(ST - 1);
        constexpr int NBP = 8 / RP;
        __syncthreads();
#pragma unroll
        for (int s = 0; s < NBP; s++) {
            int j = t + s * (M / 8);
            int k = j & (NSP - 1);
            int base = (j - k) * RP + k;
#pragma unroll
            for (int q = 0; q < RP; q++) lds[lay(base + q * NSP)] = v[s * RP + q];
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < NB; s++) {
            int j = t + s * (M / 8);
#pragma unroll
            for (int r = 0; r < R; r++) v[s * R + r] = lds[lay(j + r * (M / R))];
        }
    }
#pragma unroll
    for (int s = 0; s < NB; s++) {
        if constexpr (NS > 1) {
#pragma unroll
            for (int r = 1; r < R; r++) v[s * R + r] = cmul(v[s * R + r], tc.w[ST - 1][s * (R - 1) + r - 1]);
        }
        dft<R, INV>(&v[s * R]);
    }
    if constexpr (ST + 1 < P::NST) fft_stages_c<M, INV, ST + 1>(v, lds, lay, t, tc);
}

template <int M, bool INV, class Lay>
__device__ __forceinline__ void fft_line_c(float2 (&v)[8], float2 *lds, const Lay &lay, int t,
                                           const TwCache<M, INV> &tc) {
    fft_stages_c<M, INV, 0>(v, lds, lay, t, tc);
}

// Two lines per thread: the same stages on v[0] and v[1] (two adjacent
// columns 2c, 2c+1 of a strided layout), sharing twiddles and barriers; each
// LDS access moves both columns as one float4.
struct LayPair {
    int c2, B;   // first column (even) and columns per LDS row
    __device__ __forceinline__ int operator()(int pos) const { return ((pos + (pos >> 3)) * B + c2) >> 1; }
};

template <int M, bool INV, int ST, bool CACHED>
__device__ __forceinline__ void fft_stages2(float2 (&v)[2][8], float4 *lds, const LayPair &lay, int t,
                                            const float2 *__restrict__ tw, const TwCache<M, INV> &tc) {
    using P = Plan<M, INV>;
    constexpr int R = P::radix(ST);
    constexpr int NS = P::ns(ST);
    constexpr int NB = 8 / R;
    if constexpr (ST > 0) {
        constexpr int RP = P::radix(ST - 1);
        constexpr int NSP = P::ns(ST - 1);
        constexpr int NBP = 8 / RP;
        __syncthreads();
#pragma unroll
        for (int s = 0; s < NBP; s++) {
            int j = t + s * (M / 8);
            int k = j & (NSP - 1);
            int base = (j - k) * RP + k;
#pragma unroll
            for (int q = 0; q < RP; q++) {
                const float2 a = v[0][s * RP + q], b = v[1][s * RP + q];
                lds[lay(base + q * NSP)] = make_float4(a.x, a.y, b.x, b.y);
            }
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < NB; s++) {
            int j = t + s * (M / 8);
#pragma unroll
            for (int r = 0; r < R; r++) {
                const float4 c = lds[lay(j + r * (M / R))];
                v[0][s * R + r] = make_float2(c.x, c.y);
                v[1][s * R + r] = make_float2(c.z, c.w);
            }
        }
    }
#pragma unroll
    for (int s = 0; s < NB; s++) {
        if constexpr (NS > 1) {
            const int k = (t + s * (M / 8)) & (NS - 1);
#pragma unroll
            for (int r = 1; r < R; r++) {
                float2 w;
                if constexpr (CACHED) {
                    w = tc.w[ST - 1][s * (R - 1) + r - 1];
                } else {
                    w = tw[NS * R + r * k];
                    if (INV) w.y = -w.y;
                }
                v[0][s * R + r] = cmul(v[0][s * R + r], w);
                v[1][s * R + r] = cmul(v[1][s * R + r], w);
            }
        }
        dft<R, INV>(&v[0][s * R]);
        dft<R, INV>(&v[1][s * R]);
    }
    if constexpr (ST + 1 < P::NST) fft_stages2<M, INV, ST + 1, CACHED>(v, lds, lay, t, tw, tc);
}

// twiddles from the table
template <int M, bool INV>
__device__ __forceinline__ void fft_line2(float2 (&v)[2][8], float4 *lds, const LayPair &lay, int t,
                                          const float2 *__restrict__ tw) {
    TwCache<M, INV> none;
    fft_stages2<M, INV, 0, false>(v, lds, lay, t, tw, none);
}

// twiddles cached per thread
template <int M, bool INV>
__device__ __forceinline__ void fft_line2_c(float2 (&v)[2][8], float4 *lds, const LayPair &lay, int t,
                                            const TwCache<M, INV> &tc) {
    fft_stages2<M, INV, 0, true>(v, lds, lay, t, nullptr, tc);
}

// LDS layouts
// [pos][col], B columns per row, one pad row after every 8: the first stage
// exchange writes rows 8 apart from consecutive threads, which would
// otherwise land on the same LDS banks.
struct LayStrided {
    int col, B;
    __device__ __forceinline__ int operator()(int pos) const { return (pos + (pos >> 3)) * B + col; }
};
// LDS entries a LayStrided exchange of M positions x B columns needs
constexpr int lay_strided_size(int M, int B) { return (M + M / 8) * B; }
struct LayLine {     // contiguous lines with one pad slot every 8 entries
    int base;
    __device__ __forceinline__ int operator()(int pos) const { return base + pos + (pos >> 3); }
};

// Integer power of a real for the along-axis derivative multipliers.
__device__ __forceinline__ float kpow(float k, int m) { return m == 0 ? 1.0f : (m == 1 ? k : k * k); }

}  // namespace baorec
