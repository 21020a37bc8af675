// Learned warm start of the multigrid solve: a 3D U-Net on gfx950 matrix
// cores.  Activations and weights are BF16, channels last
// ([batch][z][y][x][channel]); every 3^3 convolution is an implicit GEMM on
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation:
//   forward      out[v][co] = b[co] + sum_{t, ci} in[v + off_t][ci] W[t][ci][co]
//                (rows = 16 voxels along x, columns = 16 output channels,
//                K = 27 taps x input channels in chunks of 8);
//   input grad   the same kernel on the output gradient with the taps mirrored
//                and W transposed;
//   weight grad  dW[t][ci][co] = sum_v in[v + off_t][ci] dout[v][co]
//                (rows = output channels, columns = input channels, K =
//                voxels of a 16x4x4 brick staged in LDS), per-chunk FP32
//                partials summed in a fixed order.
// The grid is periodic (the multigrid problem is), so every convolution wraps.
// A decoder convolution reads its skip input at full resolution and the
// coarser level's features through a nearest-neighbour upsampling on read,
// so the concatenation is never stored.  The reference has no counterpart:
// it seeds nothing (src/multigrid.jl:722-752 starts every level from zero).
#include <math.h>
#include <string.h>

#include <vector>

#include "kernels.hpp"

namespace baorec {
namespace unet {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
// round to nearest even (finite inputs; NaN stays NaN)
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ bf16x8 load8(const uint16_t *p) {
    const uint4 q = *reinterpret_cast<const uint4 *>(p);
    return __builtin_bit_cast(bf16x8, q);
}

// ------------------------------------------------------------ convolution

// Output-channel blocks of 16 per workgroup of k_conv: all of them up to 32
// channels; one from 64 on (the coarse levels' grids have few voxels, so
// splitting the channels over workgroups fills the GPU: at 16^3 a launch has
// 32 voxel workgroups).  Each output keeps its K order, so the same bits.
__host__ __device__ constexpr int conv_split(int co) { return co >= 64 ? co / 16 : 1; }

template <int C1, int C2, int CO, bool RELU, bool UP2>
__global__ void __launch_bounds__(256) k_conv(ConvArgs a) {
    constexpr int CIN = C1 + C2;
    constexpr int KB = (27 * CIN + 31) / 32;
    constexpr int NBT = CO / 16, NSPLIT = conv_split(CO), NB = NBT / NSPLIT;
    constexpr int MB = 2;
    static_assert(C1 % 8 == 0 && C2 % 8 == 0 && CO % 16 == 0, "channel blocks");
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r16 = lane & 15, g = lane >> 4;
    const int nx = 1 << a.lx, ny = 1 << a.ly, nz = 1 << a.lz;
    const int x = (blockIdx.x * 16 + r16) & (nx - 1);
    const int y0 = (blockIdx.y * 4 + wave) * MB;
    const int zb = blockIdx.z / NSPLIT, nb0 = (blockIdx.z % NSPLIT) * NB;   // first output block
    const int z = zb & (nz - 1), b = zb >> a.lz;
    f32x4 acc[MB][NB];
#pragma unroll
    for (int m = 0; m < MB; m++)
#pragma unroll
        for (int nb = 0; nb < NB; nb++) acc[m][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
    // operands of chunk kb: B (weights) and A (the input at the lane's tap and
    // channel group), 32-bit element offsets (the launcher checks the sizes)
    auto load_chunk = [&](int kb, bf16x8 *bv, bf16x8 *av) {
        const int kc = kb * 32 + g * 8;
        int t = kc / CIN;
        const int ci = kc - t * CIN;
        if (t > 26) t = 26;   // padding columns of K: their weights are zero
        const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
        const uint32_t xs = (uint32_t)((x + dx) & (nx - 1)), zs = (uint32_t)((z + dz) & (nz - 1));
#pragma unroll
        for (int nb = 0; nb < NB; nb++) bv[nb] = load8(a.w + ((uint32_t)((kb * NBT + nb0 + nb) * 4 + g) * 16 + r16) * 8);
#pragma unroll
        for (int m = 0; m < MB; m++) {
            const uint32_t ys = (uint32_t)((y0 + m + dy) & (ny - 1));
            const uint16_t *src;
            if (C2 == 0 || ci < C1) {
                src = a.in1 + ((((uint32_t)b * nz + zs) * ny + ys) * nx + xs) * C1 + ci;
            } else if (UP2) {
                const uint32_t cx = nx >> 1, cy = ny >> 1, cz = nz >> 1;
                src = a.in2 + ((((uint32_t)b * cz + (zs >> 1)) * cy + (ys >> 1)) * cx + (xs >> 1)) * C2 + (ci - C1);
            } else {
                src = a.in2 + ((((uint32_t)b * nz + zs) * ny + ys) * nx + xs) * C2 + (ci - C1);
            }
            av[m] = load8(src);
        }
    };
    // the loads of chunks kb + 1 .. kb + PF are in flight during chunk kb's
    // products (a ring of PF + 1 operand sets, indexed statically)
    constexpr int PF = 3, RING = PF + 1;
    bf16x8 bq[RING][NB], aq[RING][MB];
#pragma unroll
    for (int i = 0; i < PF; i++)
        if (i < KB) load_chunk(i, bq[i], aq[i]);
    for (int kb0 = 0; kb0 < KB; kb0 += RING) {
#pragma unroll
        for (int u = 0; u < RING; u++) {
            const int kb = kb0 + u;
            if (kb >= KB) break;
            if (kb + PF < KB) load_chunk(kb + PF, bq[(u + PF) % RING], aq[(u + PF) % RING]);
#pragma unroll
            for (int m = 0; m < MB; m++)
#pragma unroll
                for (int nb = 0; nb < NB; nb++)
                    acc[m][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aq[u][m], bq[u][nb], acc[m][nb], 0, 0, 0);
        }
    }
    // lane holds rows (voxels) 4g .. 4g+3 of column (channel) r16
#pragma unroll
    for (int m = 0; m < MB; m++) {
        const int y = y0 + m;
        if (y >= ny) continue;
#pragma unroll
        for (int nb = 0; nb < NB; nb++) {
            const int co = (nb0 + nb) * 16 + r16;
            const float bias = a.bias ? a.bias[co] : 0.f;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int xo = blockIdx.x * 16 + 4 * g + r;
                if (xo >= nx) continue;
                float v = acc[m][nb][r] + bias;
                if (RELU) v = fmaxf(v, 0.f);
                a.out[((((size_t)b * nz + z) * ny + y) * nx + xo) * CO + co] = f2bf(v);
            }
        }
    }
}

// Weight gradient.  Workgroup: 16 input channels (block blockIdx.y / NBO), 16
// output channels (block blockIdx.y % NBO), all 27 taps, bricks
// [chunk * per, (chunk + 1) * per) of 16 x 4 x 4 voxels.  Wave w owns taps
// 7w .. 7w + 6.  A = dout^T (rows: output channels, K: the brick's voxels in
// x-fastest order), B = the input at the tap's offset (columns: input
// channels); both come from LDS.
template <int C1, int C2, int CO, bool UP2>
__global__ void __launch_bounds__(256) k_conv_wgrad(WgradArgs a) {
    constexpr int CIN = C1 + C2;
    constexpr int NBO = CO / 16;
    __shared__ __align__(16) uint16_t sIn[16][6][6][18];
    __shared__ __align__(16) uint16_t sDo[16][256];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r16 = lane & 15, g = lane >> 4;
    const int cib = blockIdx.y / NBO, cob = blockIdx.y % NBO;
    const int nx = 1 << a.lx, ny = 1 << a.ly, nz = 1 << a.lz;
    const int bx = (nx + 15) >> 4, by = (ny + 3) >> 2, bz = (nz + 3) >> 2;
    const int64_t nbricks = (int64_t)bx * by * bz * a.batch;
    const int t0 = wave * 7, tn = wave == 3 ? 6 : 7;
    f32x4 acc[7];
#pragma unroll
    for (int i = 0; i < 7; i++) acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int64_t beg = (int64_t)blockIdx.x * a.per;
    const int64_t end = beg + a.per < nbricks ? beg + a.per : nbricks;
    for (int64_t br = beg; br < end; br++) {
        int64_t q = br;
        const int ix = (int)(q % bx); q /= bx;
        const int iy = (int)(q % by); q /= by;
        const int iz = (int)(q % bz);
        const int b = (int)(q / bz);
        const int x0 = ix * 16, y0 = iy * 4, z0 = iz * 4;
        __syncthreads();
        // the input block with a one-cell halo, transposed to [c][z][y][x]
        for (int idx = threadIdx.x; idx < 648 * 2; idx += 256) {
            const int v = idx >> 1, half = idx & 1;
            const int hx = v % 18, hy = (v / 18) % 6, hz = v / 108;
            const int xs = (x0 - 1 + hx) & (nx - 1), ys = (y0 - 1 + hy) & (ny - 1), zs = (z0 - 1 + hz) & (nz - 1);
            const int c = cib * 16 + half * 8;
            uint4 q4 = make_uint4(0u, 0u, 0u, 0u);
            if (c < CIN) {
                const uint16_t *src;
                if (C2 == 0 || c < C1) {
                    src = a.in1 + ((((size_t)b * nz + zs) * ny + ys) * nx + xs) * C1 + c;
                } else if (UP2) {
                    const int cx = nx >> 1, cy = ny >> 1, cz = nz >> 1;
                    src = a.in2 + ((((size_t)b * cz + (zs >> 1)) * cy + (ys >> 1)) * cx + (xs >> 1)) * C2 + (c - C1);
                } else {
                    src = a.in2 + ((((size_t)b * nz + zs) * ny + ys) * nx + xs) * C2 + (c - C1);
                }
                q4 = *reinterpret_cast<const uint4 *>(src);
            }
            const uint16_t *h = reinterpret_cast<const uint16_t *>(&q4);
#pragma unroll
            for (int j = 0; j < 8; j++) sIn[half * 8 + j][hz][hy][hx] = h[j];
        }
        // the output gradient of the brick, transposed to [c][voxel]
        for (int idx = threadIdx.x; idx < 512; idx += 256) {
            const int v = idx >> 1, half = idx & 1;
            const int x = x0 + (v & 15), y = y0 + ((v >> 4) & 3), z = z0 + (v >> 6);
            uint4 q4 = make_uint4(0u, 0u, 0u, 0u);
            if (x < nx && y < ny && z < nz)
                q4 = *reinterpret_cast<const uint4 *>(a.dout + ((((size_t)b * nz + z) * ny + y) * nx + x) * CO +
                                                     cob * 16 + half * 8);
            const uint16_t *h = reinterpret_cast<const uint16_t *>(&q4);
#pragma unroll
            for (int j = 0; j < 8; j++) sDo[half * 8 + j][v] = h[j];
        }
        __syncthreads();
        for (int kb = 0; kb < 8; kb++) {
            const bf16x8 av = *reinterpret_cast<const bf16x8 *>(&sDo[r16][kb * 32 + g * 8]);
            const int k0 = kb * 32 + g * 8;
            const int xk = k0 & 15, yk = (k0 >> 4) & 3, zk = k0 >> 6;
#pragma unroll
            for (int i = 0; i < 7; i++) {
                if (i >= tn) break;
                const int t = t0 + i;
                const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
                const uint16_t *row = &sIn[r16][zk + dz + 1][yk + dy + 1][xk + dx + 1];
                uint16_t e[8];
#pragma unroll
                for (int j = 0; j < 8; j++) e[j] = row[j];
                bf16x8 bv;
                memcpy(&bv, e, 16);
                acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[i], 0, 0, 0);
            }
        }
    }
    // D[row = output channel 4g + r][column = input channel r16]
    const int ci = cib * 16 + r16;
    if (ci >= CIN) return;
    float *part = a.part + (size_t)blockIdx.x * 27 * CIN * CO;
#pragma unroll
    for (int i = 0; i < 7; i++) {
        if (i >= tn) break;
        const int t = t0 + i;
#pragma unroll
        for (int r = 0; r < 4; r++) part[((size_t)t * CIN + ci) * CO + cob * 16 + 4 * g + r] = acc[i][r];
    }
}

// out[i] = sum over chunks c of part[c][i], c in order
__global__ void k_sum_chunks(const float *__restrict__ part, int nch, int64_t n, float *__restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        float s = 0.f;
        for (int c = 0; c < nch; c++) s += part[(size_t)c * n + i];
        out[i] = s;
    }
}

// ------------------------------------------------------------ packing

// Master FP32 weights W[t][ci][co] -> the B operand of k_conv:
// [kb][nb][g][n][8] with k = kb*32 + g*8 + j = t*CIN + ci, co = nb*16 + n.
// transposed != 0 packs the input-gradient convolution instead: its K runs
// over (mirrored tap 26 - t, fwd output channel), its columns over the fwd
// input channels.
__global__ void k_pack(const float *__restrict__ w, int cin, int co, int transposed, uint16_t *__restrict__ out) {
    const int kin = transposed ? co : cin;     // K channels of the packed conv
    const int nout = transposed ? cin : co;    // its output channels
    const int KB = (27 * kin + 31) / 32, NB = nout / 16;
    const int64_t total = (int64_t)KB * NB * 4 * 16 * 8;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(e & 7);
        int64_t q = e >> 3;
        const int n = (int)(q & 15);
        q >>= 4;
        const int gg = (int)(q & 3);
        q >>= 2;
        const int nb = (int)(q % NB);
        const int kb = (int)(q / NB);
        const int k = kb * 32 + gg * 8 + j;
        const int t = k / kin, c = k - t * kin;
        const int col = nb * 16 + n;
        float v = 0.f;
        if (t < 27) v = transposed ? w[((size_t)(26 - t) * cin + col) * co + c] : w[((size_t)t * cin + c) * co + col];
        out[e] = f2bf(v);
    }
}

// ------------------------------------------------------------ elementwise

// dy *= (y > 0)
__global__ void k_relu_mask(uint16_t *__restrict__ dy, const uint16_t *__restrict__ y, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if (!(bf2f(y[i]) > 0.f)) dy[i] = 0;
}

// 2x2x2 average pooling of a [b][z][y][x][c] tensor (c = 1 << lc)
__global__ void k_pool(const uint16_t *__restrict__ in, uint16_t *__restrict__ out, int lx, int ly, int lz, int lc,
                       int64_t nout) {
    const int cx = 1 << (lx - 1), cy = 1 << (ly - 1), cz = 1 << (lz - 1), C = 1 << lc;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nout; i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i & (C - 1));
        int64_t v = i >> lc;
        const int x = (int)(v % cx);
        v /= cx;
        const int y = (int)(v % cy);
        v /= cy;
        const int z = (int)(v % cz);
        const int64_t b = v / cz;
        const int nx = cx * 2, ny = cy * 2, nz = cz * 2;
        float s = 0.f;
        for (int d = 0; d < 8; d++) {
            const int xx = 2 * x + (d & 1), yy = 2 * y + ((d >> 1) & 1), zz = 2 * z + (d >> 2);
            s += bf2f(in[((((size_t)b * nz + zz) * ny + yy) * nx + xx) * C + c]);
        }
        out[i] = f2bf(s * 0.125f);
    }
}

// dfine += dcoarse / 8 at the parent (the adjoint of k_pool), fine side given
__global__ void k_pool_bwd(const uint16_t *__restrict__ dc, uint16_t *__restrict__ df, int lx, int ly, int lz,
                           int lc, int64_t nfine) {
    const int nx = 1 << lx, ny = 1 << ly, nz = 1 << lz, C = 1 << lc;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nfine; i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i & (C - 1));
        int64_t v = i >> lc;
        const int x = (int)(v % nx);
        v /= nx;
        const int y = (int)(v % ny);
        v /= ny;
        const int z = (int)(v % nz);
        const int64_t b = v / nz;
        const float p = bf2f(dc[((((size_t)b * (nz >> 1) + (z >> 1)) * (ny >> 1) + (y >> 1)) * (nx >> 1) + (x >> 1)) * C + c]);
        df[i] = f2bf(bf2f(df[i]) + p * 0.125f);
    }
}

// Gradient of a decoder input [skip C1 | upsampled C2] (full resolution,
// C1 + C2 channels): dskip = first C1 channels (overwrites); dcoarse = sum of
// the last C2 channels over the 8 children of each coarse cell (overwrites).
__global__ void k_split_skip(const uint16_t *__restrict__ dcat, uint16_t *__restrict__ dskip, int C1, int C2,
                             int64_t nvox) {
    const int C = C1 + C2;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nvox * C1; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t v = i / C1;
        const int c = (int)(i - v * C1);
        dskip[i] = dcat[v * C + c];
    }
}
__global__ void k_split_coarse(const uint16_t *__restrict__ dcat, uint16_t *__restrict__ dcoarse, int C1, int C2,
                               int lx, int ly, int lz, int64_t ncoarse) {
    const int C = C1 + C2;
    const int cx = 1 << (lx - 1), cy = 1 << (ly - 1), cz = 1 << (lz - 1);
    const int nx = cx * 2, ny = cy * 2, nz = cz * 2;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ncoarse * C2; i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C2);
        int64_t v = i / C2;
        const int x = (int)(v % cx);
        v /= cx;
        const int y = (int)(v % cy);
        v /= cy;
        const int z = (int)(v % cz);
        const int64_t b = v / cz;
        float s = 0.f;
        for (int d = 0; d < 8; d++) {
            const int xx = 2 * x + (d & 1), yy = 2 * y + ((d >> 1) & 1), zz = 2 * z + (d >> 2);
            s += bf2f(dcat[((((size_t)b * nz + zz) * ny + yy) * nx + xx) * C + C1 + c]);
        }
        dcoarse[i] = f2bf(s);
    }
}

// Output layer: out[v] = b + sum_c w[c] x[v][c] (16 channels, FP32 out)
__global__ void k_head(const uint16_t *__restrict__ x, const float *__restrict__ w, const float *__restrict__ b,
                       float *__restrict__ out, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint4 q0 = *reinterpret_cast<const uint4 *>(x + i * 16);
        const uint4 q1 = *reinterpret_cast<const uint4 *>(x + i * 16 + 8);
        const uint16_t *h0 = reinterpret_cast<const uint16_t *>(&q0);
        const uint16_t *h1 = reinterpret_cast<const uint16_t *>(&q1);
        float s = b[0];
        for (int c = 0; c < 8; c++) s += w[c] * bf2f(h0[c]);
        for (int c = 0; c < 8; c++) s += w[8 + c] * bf2f(h1[c]);
        out[i] = s;
    }
}
// dx[v][c] = w[c] dout[v]; per-block partials of dw[c] = sum_v x[v][c] dout[v]
// and db = sum_v dout[v] (17 values per block, summed later in block order)
__global__ void __launch_bounds__(256) k_head_bwd(const uint16_t *__restrict__ x, const float *__restrict__ w,
                                                  const float *__restrict__ dout, uint16_t *__restrict__ dx,
                                                  float *__restrict__ part, int64_t n) {
    __shared__ float red[17][4];
    float acc[17];
    for (int c = 0; c < 17; c++) acc[c] = 0.f;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float d = dout[i];
        for (int c = 0; c < 16; c++) {
            acc[c] += bf2f(x[i * 16 + c]) * d;
            dx[i * 16 + c] = f2bf(w[c] * d);
        }
        acc[16] += d;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int c = 0; c < 17; c++) {
        float v = acc[c];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if (lane == 0) red[c][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x < 17)
        part[(size_t)blockIdx.x * 17 + threadIdx.x] =
            (red[threadIdx.x][0] + red[threadIdx.x][1]) + (red[threadIdx.x][2] + red[threadIdx.x][3]);
}

// Input features of the network per voxel (8 BF16 channels):
//   0  f / s                       the solve's right-hand side
//   1  phi_s / (s h^2)             the spectral seed (fixed-LOS inverse of f)
//   2  r_s / s                     its residual under the local-LOS operator
//   3  psi / (s h^2)               the fixed-LOS inverse of that residual
//   4-6  rhat - nhat_c             local LOS minus the seed's fixed LOS
//   7  h_ref / |x|                 strength of the first-derivative term
// s = rms(f) of the item (FP64 on the device), h the cell size.
__global__ void k_features(const float *__restrict__ f, const float *__restrict__ phis, const float *__restrict__ rs,
                           const float *__restrict__ psi, const double *__restrict__ ss, FeatArgs a,
                           uint16_t *__restrict__ X) {
    const int nx = 1 << a.lx, ny = 1 << a.ly, nz = 1 << a.lz;
    const int64_t nv = (int64_t)nx * ny * nz;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv * a.batch;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t item = i / nv;
        const int64_t v = i - item * nv;
        const int x = (int)(v & (nx - 1)), y = (int)((v >> a.lx) & (ny - 1)), z = (int)(v >> (a.lx + a.ly));
        const double s = ss[item] > 0.0 ? ss[item] : 1.0;
        const float inv = (float)(1.0 / s);
        const float px = a.bmin[0] + (x + 0.5f) * a.h[0], py = a.bmin[1] + (y + 0.5f) * a.h[1],
                    pz = a.bmin[2] + (z + 0.5f) * a.h[2];
        const float r = sqrtf(px * px + py * py + pz * pz);
        const float ir = r > 0.f ? 1.f / r : 0.f;
        uint16_t o[8];
        o[0] = f2bf(f[i] * inv);
        o[1] = f2bf(phis[i] * inv * a.ih2);
        o[2] = f2bf(rs[i] * inv);
        o[3] = f2bf(psi[i] * inv * a.ih2);
        o[4] = f2bf(px * ir - a.nc[0]);
        o[5] = f2bf(py * ir - a.nc[1]);
        o[6] = f2bf(pz * ir - a.nc[2]);
        o[7] = f2bf(a.href * ir);
        uint4 q;
        memcpy(&q, o, 16);
        *reinterpret_cast<uint4 *>(X + i * 8) = q;
    }
}

// a += b
__global__ void k_add(float *__restrict__ a, const float *__restrict__ b, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        a[i] += b[i];
}

// target = (phi - phi_s) / (s h^2) and the loss
//   L = (1/N) sum_v sum_a (d(v + e_a) - d(v))^2 / 3,  d = out - target,
// the mean square error of the finite-difference gradient (the displacement
// the seed implies).  dL/dout = -(2 / 3N) lap(d).
__global__ void k_target(const float *__restrict__ phi, const float *__restrict__ phis, const double *__restrict__ ss,
                         float ih2, int64_t nv, int64_t n, float *__restrict__ t) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double s = ss[i / nv] > 0.0 ? ss[i / nv] : 1.0;
        t[i] = (float)(((double)phi[i] - (double)phis[i]) / s) * ih2;
    }
}
__global__ void __launch_bounds__(256) k_loss(const float *__restrict__ out, const float *__restrict__ t, int lx,
                                              int ly, int lz, int64_t n, float scale, float *__restrict__ dout,
                                              double *__restrict__ part) {
    __shared__ double red[4];
    const int nx = 1 << lx, ny = 1 << ly, nz = 1 << lz;
    const int64_t nv = (int64_t)nx * ny * nz;
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t base = (i / nv) * nv, v = i - base;
        const int x = (int)(v & (nx - 1)), y = (int)((v >> lx) & (ny - 1)), z = (int)(v >> (lx + ly));
        auto d = [&](int xx, int yy, int zz) {
            const int64_t j = base + (((int64_t)(zz & (nz - 1)) * ny + (yy & (ny - 1))) * nx + (xx & (nx - 1)));
            return out[j] - t[j];
        };
        const float d0 = d(x, y, z);
        const float gx = d(x + 1, y, z) - d0, gy = d(x, y + 1, z) - d0, gz = d(x, y, z + 1) - d0;
        acc += (double)gx * gx + (double)gy * gy + (double)gz * gz;
        const float lap = (d(x + 1, y, z) + d(x - 1, y, z) + d(x, y + 1, z) + d(x, y - 1, z) + d(x, y, z + 1) +
                           d(x, y, z - 1)) - 6.f * d0;
        dout[i] = -2.f * scale * lap;
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = ((red[0] + red[1]) + (red[2] + red[3])) * scale;
}

// Adam on the FP32 master copy (bias-corrected step sizes from the host)
__global__ void k_adam(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                       const float *__restrict__ gr, int64_t n, float lr_t, float b1, float b2, float eps,
                       float wd) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float g = gr[i] + wd * w[i];
        const float mi = b1 * m[i] + (1.f - b1) * g;
        const float vi = b2 * v[i] + (1.f - b2) * g * g;
        m[i] = mi;
        v[i] = vi;
        w[i] -= lr_t * mi / (sqrtf(vi) + eps);
    }
}

// the network's output in potential units: out = e s h^2 (add: out += ...)
__global__ void k_apply_seed(const float *__restrict__ e, const double *__restrict__ ss, float h2, int64_t n,
                             int add, float *__restrict__ out) {
    const double s = ss[0] > 0.0 ? ss[0] : 1.0;
    const float k = (float)s * h2;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = add ? out[i] + e[i] * k : e[i] * k;
}

static inline unsigned blocks(int64_t n) {
    int64_t b = (n + 255) / 256;
    return (unsigned)(b > 8192 ? 8192 : (b < 1 ? 1 : b));
}

}  // namespace unet

using namespace unet;

#define CONV_CASES(X)          \
    X(8, 0, 16, false)         \
    X(16, 0, 16, false)        \
    X(16, 0, 32, false)        \
    X(32, 0, 32, false)        \
    X(32, 0, 64, false)        \
    X(64, 0, 64, false)        \
    X(64, 0, 128, false)       \
    X(128, 0, 128, false)      \
    X(64, 128, 64, true)       \
    X(32, 64, 32, true)        \
    X(16, 32, 16, true)        \
    X(32, 0, 16, false)        \
    X(64, 0, 32, false)        \
    X(128, 0, 64, false)       \
    X(64, 0, 192, false)       \
    X(32, 0, 96, false)        \
    X(16, 0, 48, false)

hipError_t launch_unet_conv(const ConvArgs &a, int c1, int c2, int co, bool relu, bool up2, int batch,
                            hipStream_t s) {
    const int nx = 1 << a.lx, ny = 1 << a.ly, nz = 1 << a.lz;
    // k_conv addresses its inputs with 32-bit element offsets
    const int cmax = c1 > c2 ? c1 : c2;
    if ((int64_t)batch * nx * ny * nz * cmax >= ((int64_t)1 << 31)) return hipErrorInvalidValue;
    dim3 grid((unsigned)((nx + 15) / 16), (unsigned)((ny + 7) / 8), (unsigned)(nz * batch * conv_split(co)));
#define LAUNCH(C1_, C2_, CO_, UP_)                                                                        \
    if (c1 == C1_ && c2 == C2_ && co == CO_ && up2 == UP_) {                                              \
        if (relu)                                                                                         \
            hipLaunchKernelGGL((k_conv<C1_, C2_, CO_, true, UP_>), grid, dim3(256), 0, s, a);             \
        else                                                                                              \
            hipLaunchKernelGGL((k_conv<C1_, C2_, CO_, false, UP_>), grid, dim3(256), 0, s, a);            \
        return hipGetLastError();                                                                         \
    }
    CONV_CASES(LAUNCH)
#undef LAUNCH
    return hipErrorInvalidValue;
}

hipError_t launch_unet_wgrad(const WgradArgs &a, int c1, int c2, int co, bool up2, int nchunks, hipStream_t s) {
    const int cib = (c1 + c2 + 15) / 16, cob = co / 16;
    dim3 grid((unsigned)nchunks, (unsigned)(cib * cob));
#define LAUNCH(C1_, C2_, CO_, UP_)                                                                        \
    if (c1 == C1_ && c2 == C2_ && co == CO_ && up2 == UP_) {                                              \
        hipLaunchKernelGGL((k_conv_wgrad<C1_, C2_, CO_, UP_>), grid, dim3(256), 0, s, a);                 \
        return hipGetLastError();                                                                         \
    }
    CONV_CASES(LAUNCH)
#undef LAUNCH
    return hipErrorInvalidValue;
}

hipError_t launch_unet_sum_chunks(const float *part, int nch, int64_t n, float *out, hipStream_t s) {
    hipLaunchKernelGGL(k_sum_chunks, dim3(blocks(n)), dim3(256), 0, s, part, nch, n, out);
    return hipGetLastError();
}
hipError_t launch_unet_pack(const float *w, int cin, int co, int transposed, uint16_t *out, hipStream_t s) {
    const int kin = transposed ? co : cin, nout = transposed ? cin : co;
    const int64_t total = (int64_t)((27 * kin + 31) / 32) * (nout / 16) * 512;
    hipLaunchKernelGGL(k_pack, dim3(blocks(total)), dim3(256), 0, s, w, cin, co, transposed, out);
    return hipGetLastError();
}
hipError_t launch_unet_relu_mask(uint16_t *dy, const uint16_t *y, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_relu_mask, dim3(blocks(n)), dim3(256), 0, s, dy, y, n);
    return hipGetLastError();
}
hipError_t launch_unet_pool(const uint16_t *in, uint16_t *out, int lx, int ly, int lz, int lc, int batch,
                            hipStream_t s) {
    const int64_t n = ((int64_t)batch << (lx + ly + lz - 3)) << lc;
    hipLaunchKernelGGL(k_pool, dim3(blocks(n)), dim3(256), 0, s, in, out, lx, ly, lz, lc, n);
    return hipGetLastError();
}
hipError_t launch_unet_pool_bwd(const uint16_t *dc, uint16_t *df, int lx, int ly, int lz, int lc, int batch,
                                hipStream_t s) {
    const int64_t n = ((int64_t)batch << (lx + ly + lz)) << lc;
    hipLaunchKernelGGL(k_pool_bwd, dim3(blocks(n)), dim3(256), 0, s, dc, df, lx, ly, lz, lc, n);
    return hipGetLastError();
}
hipError_t launch_unet_split(const uint16_t *dcat, uint16_t *dskip, uint16_t *dcoarse, int c1, int c2, int lx,
                             int ly, int lz, int batch, hipStream_t s) {
    const int64_t nv = (int64_t)batch << (lx + ly + lz);
    hipLaunchKernelGGL(k_split_skip, dim3(blocks(nv * c1)), dim3(256), 0, s, dcat, dskip, c1, c2, nv);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int64_t nc = nv >> 3;
    hipLaunchKernelGGL(k_split_coarse, dim3(blocks(nc * c2)), dim3(256), 0, s, dcat, dcoarse, c1, c2, lx, ly, lz, nc);
    return hipGetLastError();
}
hipError_t launch_unet_head(const uint16_t *x, const float *w, const float *b, float *out, int64_t n,
                            hipStream_t s) {
    hipLaunchKernelGGL(k_head, dim3(blocks(n)), dim3(256), 0, s, x, w, b, out, n);
    return hipGetLastError();
}
int unet_head_bwd_blocks(int64_t n) { return (int)blocks(n); }
hipError_t launch_unet_head_bwd(const uint16_t *x, const float *w, const float *dout, uint16_t *dx, float *part,
                                int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_head_bwd, dim3(blocks(n)), dim3(256), 0, s, x, w, dout, dx, part, n);
    return hipGetLastError();
}
hipError_t launch_unet_features(const float *f, const float *phis, const float *rs, const float *psi,
                                const double *ss, const FeatArgs &a, uint16_t *X, hipStream_t s) {
    const int64_t n = ((int64_t)a.batch << (a.lx + a.ly + a.lz));
    hipLaunchKernelGGL(k_features, dim3(blocks(n)), dim3(256), 0, s, f, phis, rs, psi, ss, a, X);
    return hipGetLastError();
}
hipError_t launch_add_grid(float *a, const float *b, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_add, dim3(blocks(n)), dim3(256), 0, s, a, b, n);
    return hipGetLastError();
}
hipError_t launch_unet_target(const float *phi, const float *phis, const double *ss, float ih2, int64_t nv, int64_t n,
                              float *t, hipStream_t s) {
    hipLaunchKernelGGL(k_target, dim3(blocks(n)), dim3(256), 0, s, phi, phis, ss, ih2, nv, n, t);
    return hipGetLastError();
}
int unet_loss_blocks(int64_t n) { return (int)blocks(n); }
hipError_t launch_unet_loss(const float *out, const float *t, int lx, int ly, int lz, int64_t n, float scale,
                            float *dout, double *part, hipStream_t s) {
    hipLaunchKernelGGL(k_loss, dim3(blocks(n)), dim3(256), 0, s, out, t, lx, ly, lz, n, scale, dout, part);
    return hipGetLastError();
}
hipError_t launch_unet_adam(float *w, float *m, float *v, const float *g, int64_t n, float lr_t, float b1, float b2,
                            float eps, float wd, hipStream_t s) {
    hipLaunchKernelGGL(k_adam, dim3(blocks(n)), dim3(256), 0, s, w, m, v, g, n, lr_t, b1, b2, eps, wd);
    return hipGetLastError();
}
hipError_t launch_unet_apply_seed(const float *e, const double *ss, float h2, int64_t n, int add, float *out,
                                  hipStream_t s) {
    hipLaunchKernelGGL(k_apply_seed, dim3(blocks(n)), dim3(256), 0, s, e, ss, h2, n, add, out);
    return hipGetLastError();
}

}  // namespace baorec

// ====================================================================== host

namespace baorec {

namespace {
struct LayerDef {
    int c1, c2, co;
    bool up2, relu;
    int level;        // resolution level of the output (0 = finest)
    int in1, in2;     // input activation index (-1: network input X, -2..-4: pooled P0..P2)
};
// The network: levels 0..3 with 16, 32, 64, 128 channels; two 3^3
// convolutions per level on the way down and on the way up, 2x2x2 average
// pooling between levels, nearest upsampling into the decoder, a 1x1 head.
const LayerDef kLayers[14] = {
    {8, 0, 16, false, true, 0, -1, -1},     // 0  e0a
    {16, 0, 16, false, true, 0, 0, -1},     // 1  e0b  (skip 0)
    {16, 0, 32, false, true, 1, -2, -1},    // 2  e1a  (pooled 1)
    {32, 0, 32, false, true, 1, 2, -1},     // 3  e1b  (skip 1)
    {32, 0, 64, false, true, 2, -3, -1},    // 4  e2a
    {64, 0, 64, false, true, 2, 4, -1},     // 5  e2b  (skip 2)
    {64, 0, 128, false, true, 3, -4, -1},   // 6  e3a
    {128, 0, 128, false, true, 3, 6, -1},   // 7  e3b
    {64, 128, 64, true, true, 2, 5, 7},     // 8  d2a  [skip 2 | up e3b]
    {64, 0, 64, false, true, 2, 8, -1},     // 9  d2b
    {32, 64, 32, true, true, 1, 3, 9},      // 10 d1a  [skip 1 | up d2b]
    {32, 0, 32, false, true, 1, 10, -1},    // 11 d1b
    {16, 32, 16, true, true, 0, 1, 11},     // 12 d0a  [skip 0 | up d1b]
    {16, 0, 16, false, true, 0, 12, -1},    // 13 d0b
};
constexpr int kNL = 14;

template <typename T>
hipError_t dalloc(T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    return hipMalloc(reinterpret_cast<void **>(p), sizeof(T) * count);
}
}  // namespace

#define UCHK(x)                              \
    do {                                     \
        hipError_t e_ = (x);                 \
        if (e_ != hipSuccess) return e_;     \
    } while (0)

struct UNet {
    int64_t woff[kNL], boff[kNL], hw = 0, hb = 0, P = 0;
    float *par = nullptr, *grad = nullptr, *am = nullptr, *av = nullptr;
    uint16_t *wp[kNL] = {}, *wpt[kNL] = {};
    int64_t step = 0;
    float h_train = 0.f;
    // activations and gradients of the last forward/backward shape
    int batch = 0, lx = -1, ly = -1, lz = -1;
    uint16_t *X = nullptr, *A[kNL] = {}, *Pl[3] = {}, *dA[kNL] = {}, *dP[3] = {}, *dcat[3] = {};
    float *out = nullptr, *dout = nullptr, *wpart = nullptr, *hpart = nullptr, *cpart = nullptr;
    size_t wpart_n = 0;

    UNet() {
        int64_t o = 0;
        for (int i = 0; i < kNL; i++) {
            woff[i] = o;
            o += 27LL * (kLayers[i].c1 + kLayers[i].c2) * kLayers[i].co;
            boff[i] = o;
            o += kLayers[i].co;
        }
        hw = o;
        o += 16;
        hb = o;
        o += 1;
        P = o;
    }
    void free_acts() {
        (void)hipFree(X);
        X = nullptr;
        for (int i = 0; i < kNL; i++) {
            (void)hipFree(A[i]);
            (void)hipFree(dA[i]);
            A[i] = dA[i] = nullptr;
        }
        for (int i = 0; i < 3; i++) {
            (void)hipFree(Pl[i]);
            (void)hipFree(dP[i]);
            (void)hipFree(dcat[i]);
            Pl[i] = dP[i] = dcat[i] = nullptr;
        }
        (void)hipFree(out);
        (void)hipFree(dout);
        (void)hipFree(hpart);
        (void)hipFree(cpart);
        out = dout = hpart = cpart = nullptr;
        batch = 0;
        lx = ly = lz = -1;
    }
    ~UNet() {
        free_acts();
        (void)hipFree(par);
        (void)hipFree(grad);
        (void)hipFree(am);
        (void)hipFree(av);
        (void)hipFree(wpart);
        for (int i = 0; i < kNL; i++) {
            (void)hipFree(wp[i]);
            (void)hipFree(wpt[i]);
        }
    }
    int64_t nvox(int level) const { return (int64_t)batch << (lx + ly + lz - 3 * level); }

    hipError_t alloc_params() {
        UCHK(dalloc(&par, P));
        UCHK(dalloc(&grad, P));
        UCHK(dalloc(&am, P));
        UCHK(dalloc(&av, P));
        UCHK(hipMemset(am, 0, sizeof(float) * P));
        UCHK(hipMemset(av, 0, sizeof(float) * P));
        UCHK(hipMemset(grad, 0, sizeof(float) * P));
        for (int i = 0; i < kNL; i++) {
            const int cin = kLayers[i].c1 + kLayers[i].c2, co = kLayers[i].co;
            UCHK(dalloc(&wp[i], (size_t)((27 * cin + 31) / 32) * (co / 16) * 512));
            if (i > 0) UCHK(dalloc(&wpt[i], (size_t)((27 * co + 31) / 32) * (cin / 16) * 512));
        }
        return hipSuccess;
    }
    hipError_t pack(hipStream_t s) {
        for (int i = 0; i < kNL; i++) {
            const int cin = kLayers[i].c1 + kLayers[i].c2, co = kLayers[i].co;
            UCHK(launch_unet_pack(par + woff[i], cin, co, 0, wp[i], s));
            if (i > 0) UCHK(launch_unet_pack(par + woff[i], cin, co, 1, wpt[i], s));
        }
        return hipSuccess;
    }
    hipError_t ensure_acts(int b, int x, int y, int z) {
        if (b == batch && x == lx && y == ly && z == lz) return hipSuccess;
        free_acts();
        batch = b;
        lx = x;
        ly = y;
        lz = z;
        UCHK(dalloc(&X, (size_t)nvox(0) * 8));
        for (int i = 0; i < kNL; i++) {
            UCHK(dalloc(&A[i], (size_t)nvox(kLayers[i].level) * kLayers[i].co));
            UCHK(dalloc(&dA[i], (size_t)nvox(kLayers[i].level) * kLayers[i].co));
        }
        for (int l = 0; l < 3; l++) {
            const int c = 16 << l;
            UCHK(dalloc(&Pl[l], (size_t)nvox(l + 1) * c));
            UCHK(dalloc(&dP[l], (size_t)nvox(l + 1) * c));
            UCHK(dalloc(&dcat[l], (size_t)nvox(l) * (c + 2 * c)));
        }
        UCHK(dalloc(&out, nvox(0)));
        UCHK(dalloc(&dout, nvox(0)));
        UCHK(dalloc(&hpart, (size_t)unet_head_bwd_blocks(nvox(0)) * 17));
        UCHK(dalloc(&cpart, (size_t)4096 * 256));
        return hipSuccess;
    }
    const uint16_t *input_of(int idx) const {
        if (idx == -1) return X;
        if (idx <= -2) return Pl[-idx - 2];
        return A[idx];
    }
    int lv(int level, int which) const { return (which == 0 ? lx : which == 1 ? ly : lz) - level; }

    hipError_t forward(hipStream_t s) {
        for (int i = 0; i < kNL; i++) {
            const LayerDef &d = kLayers[i];
            ConvArgs a{};
            a.in1 = input_of(d.in1);
            a.in2 = d.in2 >= 0 ? A[d.in2] : nullptr;
            a.w = wp[i];
            a.bias = par + boff[i];
            a.out = A[i];
            a.lx = lv(d.level, 0);
            a.ly = lv(d.level, 1);
            a.lz = lv(d.level, 2);
            UCHK(launch_unet_conv(a, d.c1, d.c2, d.co, d.relu, d.up2, batch, s));
            if (i == 1 || i == 3 || i == 5) {   // pool the skip tensor of this level
                const int l = d.level, lc = 4 + l;
                UCHK(launch_unet_pool(A[i], Pl[l], lv(l, 0), lv(l, 1), lv(l, 2), lc, batch, s));
            }
        }
        return launch_unet_head(A[kNL - 1], par + hw, par + hb, out, nvox(0), s);
    }

    // per-channel sums of a [n][C] BF16 tensor into g[C] (bias gradients)
    hipError_t chan_sum(const uint16_t *x, int C, int64_t n, float *g, hipStream_t s);

    hipError_t backward(hipStream_t s) {
        const int64_t n0 = nvox(0);
        const int hbk = unet_head_bwd_blocks(n0);
        UCHK(launch_unet_head_bwd(A[kNL - 1], par + hw, dout, dA[kNL - 1], hpart, n0, s));
        UCHK(head_bias(hbk, s));
        for (int i = kNL - 1; i >= 0; i--) {
            const LayerDef &d = kLayers[i];
            const int l = d.level;
            const int64_t n = nvox(l);
            const int x = lv(l, 0), y = lv(l, 1), z = lv(l, 2);
            UCHK(launch_unet_relu_mask(dA[i], A[i], n * d.co, s));
            // weight gradient
            const int cin = d.c1 + d.c2;
            const int64_t nbricks = (int64_t)(((1 << x) + 15) >> 4) * (((1 << y) + 3) >> 2) * (((1 << z) + 3) >> 2) * batch;
            int nch = (int)(nbricks < 256 ? nbricks : 256);
            const int64_t per = (nbricks + nch - 1) / nch;
            nch = (int)((nbricks + per - 1) / per);
            const size_t need = (size_t)nch * 27 * cin * d.co;
            if (need > wpart_n) {
                (void)hipFree(wpart);
                UCHK(dalloc(&wpart, need));
                wpart_n = need;
            }
            WgradArgs g{};
            g.in1 = input_of(d.in1);
            g.in2 = d.in2 >= 0 ? A[d.in2] : nullptr;
            g.dout = dA[i];
            g.part = wpart;
            g.lx = x;
            g.ly = y;
            g.lz = z;
            g.batch = batch;
            g.per = per;
            UCHK(launch_unet_wgrad(g, d.c1, d.c2, d.co, d.up2, nch, s));
            UCHK(launch_unet_sum_chunks(wpart, nch, 27LL * cin * d.co, grad + woff[i], s));
            UCHK(chan_sum(dA[i], d.co, n, grad + boff[i], s));
            if (i == 0) break;
            // input gradient: the transposed convolution of dA[i]
            ConvArgs a{};
            a.in1 = dA[i];
            a.in2 = nullptr;
            a.w = wpt[i];
            a.bias = nullptr;
            a.lx = x;
            a.ly = y;
            a.lz = z;
            if (d.c2 > 0) {   // decoder: [skip | upsampled coarse]
                const int sl = l;   // dcat index = level
                a.out = dcat[sl];
                UCHK(launch_unet_conv(a, d.co, 0, cin, false, false, batch, s));
                UCHK(launch_unet_split(dcat[sl], dA[d.in1], dA[d.in2], d.c1, d.c2, x, y, z, batch, s));
            } else if (d.in1 <= -2) {   // encoder input from pooling
                const int pl = -d.in1 - 2;
                a.out = dP[pl];
                UCHK(launch_unet_conv(a, d.co, 0, cin, false, false, batch, s));
                const int sk = 2 * pl + 1;   // the pooled skip layer
                UCHK(launch_unet_pool_bwd(dP[pl], dA[sk], lv(pl, 0), lv(pl, 1), lv(pl, 2), 4 + pl, batch, s));
            } else {
                a.out = dA[d.in1];
                UCHK(launch_unet_conv(a, d.co, 0, cin, false, false, batch, s));
            }
        }
        return hipSuccess;
    }
    hipError_t head_bias(int hbk, hipStream_t s);
};

__global__ void k_chan_sum_part(const uint16_t *__restrict__ x, int C, int64_t n, float *__restrict__ part) {
    // block b sums rows b, b + gridDim.x, ... for channel threadIdx.x % C
    const int c = threadIdx.x % C, lanes = blockDim.x / C, r0 = threadIdx.x / C;
    __shared__ float red[256];
    float s = 0.f;
    if (r0 < lanes)
        for (int64_t i = (int64_t)blockIdx.x * lanes + r0; i < n; i += (int64_t)gridDim.x * lanes)
            s += unet::bf2f(x[i * C + c]);
    red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < C) {
        float t = 0.f;
        for (int r = 0; r < lanes; r++) t += red[r * C + threadIdx.x];
        part[(size_t)blockIdx.x * C + threadIdx.x] = t;
    }
}
// part[b][C] -> out[c] = sum_b part[b][c] in block order
__global__ void k_chan_sum_final(const float *__restrict__ part, int nb, int C, float *__restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    float t = 0.f;
    for (int b = 0; b < nb; b++) t += part[(size_t)b * C + c];
    out[c] = t;
}

hipError_t UNet::chan_sum(const uint16_t *x, int C, int64_t n, float *g, hipStream_t s) {
    const int threads = C >= 256 ? C : 256 / C * C;
    int64_t nb = n / 64;
    nb = nb < 1 ? 1 : (nb > 4096 ? 4096 : nb);
    hipLaunchKernelGGL(k_chan_sum_part, dim3((unsigned)nb), dim3((unsigned)threads), 0, s, x, C, n, cpart);
    UCHK(hipGetLastError());
    hipLaunchKernelGGL(k_chan_sum_final, dim3((unsigned)((C + 127) / 128)), dim3(128), 0, s, cpart, (int)nb, C, g);
    return hipGetLastError();
}

// the head's bias gradient: entry 16 of each block's 17 partials
__global__ void k_head_bias(const float *__restrict__ part, int nb, float *__restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float t = 0.f;
    for (int b = 0; b < nb; b++) t += part[(size_t)b * 17 + 16];
    out[0] = t;
}
// the head's weight gradient from the 17-strided partials
__global__ void k_head_w(const float *__restrict__ part, int nb, float *__restrict__ out) {
    const int c = threadIdx.x;
    if (c >= 16) return;
    float t = 0.f;
    for (int b = 0; b < nb; b++) t += part[(size_t)b * 17 + c];
    out[c] = t;
}
hipError_t UNet::head_bias(int hbk, hipStream_t s) {
    hipLaunchKernelGGL(k_head_w, dim3(1), dim3(64), 0, s, hpart, hbk, grad + hw);
    UCHK(hipGetLastError());
    hipLaunchKernelGGL(k_head_bias, dim3(1), dim3(64), 0, s, hpart, hbk, grad + hb);
    return hipGetLastError();
}

}  // namespace baorec

// ------------------------------------------------------------ C++ interface

namespace baorec {

UNet *unet_create() { return new (std::nothrow) UNet(); }
int64_t unet_count_params() {
    int64_t o = 0;
    for (int i = 0; i < kNL; i++) o += 27LL * (kLayers[i].c1 + kLayers[i].c2) * kLayers[i].co + kLayers[i].co;
    return o + 17;
}
void unet_destroy(UNet *n) { delete n; }
int64_t unet_nparams(const UNet *n) { return n->P; }
float unet_h_train(const UNet *n) { return n->h_train; }
void unet_set_h_train(UNet *n, float h) { n->h_train = h; }
int64_t unet_step(const UNet *n) { return n->step; }

// He-normal convolution weights (std sqrt(2 / fan_in)), zero biases, and a
// head of std 0.01 so that the first outputs are a small correction.
hipError_t unet_init(UNet *n, uint64_t seed, hipStream_t s) {
    if (!n->par) UCHK(n->alloc_params());
    std::vector<float> h((size_t)n->P, 0.f);
    uint64_t st = seed * 0x9E3779B97F4A7C15ull + 0x2545F4914F6CDD1Dull;
    auto next = [&]() {   // splitmix64
        uint64_t z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    auto normal = [&]() {   // Box-Muller
        const double u1 = ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        const double u2 = ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    };
    for (int i = 0; i < kNL; i++) {
        const int cin = kLayers[i].c1 + kLayers[i].c2;
        const double sd = sqrt(2.0 / (27.0 * cin));
        for (int64_t k = 0; k < 27LL * cin * kLayers[i].co; k++) h[n->woff[i] + k] = (float)(sd * normal());
    }
    for (int c = 0; c < 16; c++) h[n->hw + c] = (float)(0.01 * normal());
    n->step = 0;
    UCHK(hipMemcpyAsync(n->par, h.data(), sizeof(float) * n->P, hipMemcpyHostToDevice, s));
    UCHK(hipMemsetAsync(n->am, 0, sizeof(float) * n->P, s));
    UCHK(hipMemsetAsync(n->av, 0, sizeof(float) * n->P, s));
    UCHK(n->pack(s));
    return hipStreamSynchronize(s);
}

hipError_t unet_set_params(UNet *n, const float *host, hipStream_t s) {
    if (!n->par) UCHK(n->alloc_params());
    UCHK(hipMemcpyAsync(n->par, host, sizeof(float) * n->P, hipMemcpyHostToDevice, s));
    UCHK(hipMemsetAsync(n->am, 0, sizeof(float) * n->P, s));
    UCHK(hipMemsetAsync(n->av, 0, sizeof(float) * n->P, s));
    n->step = 0;
    UCHK(n->pack(s));
    return hipStreamSynchronize(s);
}
hipError_t unet_get_params(const UNet *n, float *host, int grads, hipStream_t s) {
    UCHK(hipMemcpyAsync(host, grads ? n->grad : n->par, sizeof(float) * n->P, hipMemcpyDeviceToHost, s));
    return hipStreamSynchronize(s);
}
bool unet_ready(const UNet *n) { return n && n->par; }

// The input tensor (8 BF16 channels per voxel) of a batch of grids with
// sides 2^lx x 2^ly x 2^lz; allocates the activations for that shape.
hipError_t unet_input(UNet *n, int batch, int lx, int ly, int lz, uint16_t **X) {
    UCHK(n->ensure_acts(batch, lx, ly, lz));
    *X = n->X;
    return hipSuccess;
}
// Device tensor of the last forward/backward (BF16 bits) and its length:
// kind 0 the output of layer i, 1 its gradient after the ReLU mask (what the
// weight and input gradients of layer i used), 2 pooled tensor i (0..2),
// 3 its gradient, 4 the input features.
hipError_t unet_tensor(const UNet *n, int kind, int i, const uint16_t **ptr, int64_t *count) {
    *ptr = nullptr;
    *count = 0;
    if (!n->X) return hipErrorNotReady;
    if ((kind == 0 || kind == 1) && i >= 0 && i < kNL) {
        *ptr = kind == 0 ? n->A[i] : n->dA[i];
        *count = n->nvox(kLayers[i].level) * kLayers[i].co;
    } else if ((kind == 2 || kind == 3) && i >= 0 && i < 3) {
        *ptr = kind == 2 ? n->Pl[i] : n->dP[i];
        *count = n->nvox(i + 1) * (16 << i);
    } else if (kind == 4) {
        *ptr = n->X;
        *count = n->nvox(0) * 8;
    } else {
        return hipErrorInvalidValue;
    }
    return hipSuccess;
}
float *unet_output(UNet *n) { return n->out; }
float *unet_output_grad(UNet *n) { return n->dout; }
hipError_t unet_forward(UNet *n, hipStream_t s) { return n->forward(s); }
hipError_t unet_backward(UNet *n, hipStream_t s) { return n->backward(s); }

hipError_t unet_adam(UNet *n, float lr, hipStream_t s) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    n->step++;
    const double c1 = 1.0 - pow((double)b1, (double)n->step), c2 = 1.0 - pow((double)b2, (double)n->step);
    const float lr_t = (float)(lr * sqrt(c2) / c1);
    UCHK(launch_unet_adam(n->par, n->am, n->av, n->grad, n->P, lr_t, b1, b2, eps, 0.f, s));
    return n->pack(s);
}

}  // namespace baorec
