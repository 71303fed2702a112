// rt_mfma.hip — BF16 matrix products on gfx950 matrix cores for the learned warm start.
//
// C[M x N] (fp32) = alpha * op(A)[M x K] * op(B)[K x N] + beta * C, A and B in bf16 with
// arbitrary element strides (so the same kernel serves A*B, A^T*B and A*B^T).  A workgroup of
// four waves computes a 64 x 64 tile; each wave owns a 32 x 32 quarter and accumulates it
// with v_mfma_f32_32x32x8_bf16 (fp32 accumulation) over K in steps of 8, the operands staged
// through LDS 32 deep in K.  Products with few output tiles and a long K (the weight
// gradients, K = pixels) split K over up to 64 slices that add their partial tiles into C
// atomically.
//
// Operand layout of v_mfma_f32_32x32x8bf16_1k (wave64): lane l holds A[l % 32][4 (l / 32) + 0..3]
// and B[4 (l / 32) + 0..3][l % 32]; accumulator register r of lane l is
// D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rt_internal.hpp"
#include "rt_mfma.hpp"

namespace rt {
namespace nn {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint16_t f2bf(float f) {
    // round to nearest even
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));  // inf / nan
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

constexpr int BM = 64, BN = 64, BK = 32;

// gridDim.z > 1: slice z covers K range [z kchunk, (z+1) kchunk) and adds into C atomically
// (C then already holds beta * C).
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g, int kchunk) {
    __shared__ uint16_t la[BM][BK + 4];  // [m][k]
    __shared__ uint16_t lb[BN][BK + 4];  // [n][k]
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int kbeg = blockIdx.z * kchunk, kend = min(g.K, kbeg + kchunk);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;  // the wave's quarter
    v16f acc;
#pragma unroll
    for (int r = 0; r < 16; r++) acc[r] = 0.f;
    for (int k0 = kbeg; k0 < kend; k0 += BK) {
        // stage A[m0.., k0..] and B[k0.., n0..] (zeros outside)
        for (int q = tid; q < BM * BK; q += 256) {
            const int m = q / BK, k = q % BK;
            const int gm = m0 + m, gk = k0 + k;
            la[m][k] = (gm < g.M && gk < kend) ? g.A[(int64_t)gm * g.sam + (int64_t)gk * g.sak] : (uint16_t)0;
        }
        for (int q = tid; q < BN * BK; q += 256) {
            const int n = q / BK, k = q % BK;
            const int gn = n0 + n, gk = k0 + k;
            lb[n][k] = (gn < g.N && gk < kend) ? g.B[(int64_t)gk * g.sbk + (int64_t)gn * g.sbn] : (uint16_t)0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; kk += 8) {
            const int kb = kk + 4 * (lane >> 5);
            v4s a, b;
#pragma unroll
            for (int x = 0; x < 4; x++) {
                a[x] = (short)la[wm + (lane & 31)][kb + x];
                b[x] = (short)lb[wn + (lane & 31)][kb + x];
            }
            acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int m = m0 + wm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const int n = n0 + wn + (lane & 31);
        if (m < g.M && n < g.N) {
            float* c = g.C + (int64_t)m * g.ldc + n;
            if (gridDim.z > 1) atomicAdd(c, g.alpha * acc[r]);
            else *c = g.alpha * acc[r] + (g.beta != 0.f ? g.beta * *c : 0.f);
        }
    }
}

__global__ void k_scale(float* __restrict__ c, int M, int N, int64_t ldc, float beta) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)M * N) return;
    float* p = c + (q / N) * ldc + q % N;
    *p = beta == 0.f ? 0.f : beta * *p;
}

__global__ void k_f32_to_bf16(const float* __restrict__ x, uint16_t* __restrict__ y, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

// ---- large tiles: 128 x 128 per workgroup, v_mfma_f32_32x32x16_bf16 -----------------------
// Both operands contiguous along K (an operand stored the other way is first transposed into
// the caller's workspace by k_transpose).  Four waves, each a 64 x 64 quarter (2 x 2 tiles of
// 32 x 32).  K advances 64 at a time through two LDS buffers holding [m][k] and [n][k] rows
// (72 bf16 = 144 B: the fragment reads of 16 consecutive rows hit 16 distinct bank quads), the
// next slab loaded into registers as 16-byte vectors while the current one is multiplied.
// Lane layout of v_mfma_f32_32x32x16_bf16 (wave64): lane l holds A[l % 32][8 (l / 32) + 0..7]
// and B[8 (l / 32) + 0..7][l % 32]; the accumulator is laid out as for the 32 x 32 x 8 product.
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

constexpr int LM = 128, LN = 128, LK = 64, LROW = LK + 8;
constexpr int LV = LM * LK / 8 / 256;  // 16-byte vectors per thread per operand and slab

__global__ __launch_bounds__(256) void k_gemm_l(GemmArgs g, int kchunk) {
    __shared__ __align__(16) uint16_t sa[2][LM * LROW];
    __shared__ __align__(16) uint16_t sb[2][LN * LROW];
    const int m0 = blockIdx.y * LM, n0 = blockIdx.x * LN;
    const int kbeg = blockIdx.z * kchunk, kend = min(g.K, kbeg + kchunk);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;

    uint4 ra[LV], rb[LV];
    auto load = [&](int k0) {
#pragma unroll
        for (int h = 0; h < LV; h++) {
            const int v = tid + h * 256;
            const int r = v >> 3, k = k0 + (v & 7) * 8;  // 8 vectors along K per row
            const bool ia = m0 + r < g.M && k < kend, ib = n0 + r < g.N && k < kend;
            ra[h] = ia ? *reinterpret_cast<const uint4*>(g.A + (int64_t)(m0 + r) * g.sam + k) : make_uint4(0, 0, 0, 0);
            rb[h] = ib ? *reinterpret_cast<const uint4*>(g.B + (int64_t)(n0 + r) * g.sbn + k) : make_uint4(0, 0, 0, 0);
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int h = 0; h < LV; h++) {
            const int v = tid + h * 256;
            const int o = (v >> 3) * LROW + (v & 7) * 8;
            *reinterpret_cast<uint4*>(&sa[buf][o]) = ra[h];
            *reinterpret_cast<uint4*>(&sb[buf][o]) = rb[h];
        }
    };

    v16f acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[i][j][r] = 0.f;

    int buf = 0;
    if (kbeg < kend) {
        load(kbeg);
        store(0);
    }
    __syncthreads();
    for (int k0 = kbeg; k0 < kend; k0 += LK) {
        const bool more = k0 + LK < kend;
        if (more) load(k0 + LK);  // in flight during the products below
#pragma unroll
        for (int kk = 0; kk < LK; kk += 16) {
            const int ko = kk + 8 * (lane >> 5);
            v8bf af[2], bf[2];
#pragma unroll
            for (int t = 0; t < 2; t++) {
                af[t] = *reinterpret_cast<const v8bf*>(&sa[buf][(wm + 32 * t + (lane & 31)) * LROW + ko]);
                bf[t] = *reinterpret_cast<const v8bf*>(&sb[buf][(wn + 32 * t + (lane & 31)) * LROW + ko]);
            }
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[i], bf[j], acc[i][j], 0, 0, 0);
        }
        if (more) store(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int m = m0 + wm + 32 * i + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                const int n = n0 + wn + 32 * j + (lane & 31);
                if (m < g.M && n < g.N) {
                    float* c = g.C + (int64_t)m * g.ldc + n;
                    if (gridDim.z > 1) atomicAdd(c, g.alpha * acc[i][j][r]);
                    else *c = g.alpha * acc[i][j][r] + (g.beta != 0.f ? g.beta * *c : 0.f);
                }
            }
}

// out[c][r] = in[r * ld + c] for an R x C bf16 block (out rows contiguous, length R), through
// 64 x 64 LDS tiles.
__global__ __launch_bounds__(256) void k_transpose(const uint16_t* __restrict__ in, int64_t ld, int R, int C,
                                                   uint16_t* __restrict__ out) {
    __shared__ uint16_t t[64][66];
    const int r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
    for (int q = threadIdx.x; q < 64 * 64; q += 256) {
        const int r = q >> 6, c = q & 63;
        t[r][c] = (r0 + r < R && c0 + c < C) ? in[(int64_t)(r0 + r) * ld + c0 + c] : (uint16_t)0;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < 64 * 64; q += 256) {
        const int c = q >> 6, r = q & 63;
        if (c0 + c < C && r0 + r < R) out[(int64_t)(c0 + c) * R + r0 + r] = t[r][c];
    }
}

// out[r][c] = in[r * ld + c]: an R x C bf16 block copied to contiguous rows (an operand stored
// along K but not 16-byte aligned)
__global__ void k_copy_rows(const uint16_t* __restrict__ in, int64_t ld, int R, int C, uint16_t* __restrict__ out) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)R * C) return;
    out[q] = in[(q / C) * ld + q % C];
}

}  // namespace nn

int64_t nn_gemm_workspace(int M, int N, int K) { return ((int64_t)M + N) * K + 16; }

hipError_t nn_gemm(const GemmArgs& g0, hipStream_t st) {
    if (g0.M <= 0 || g0.N <= 0) return hipSuccess;
    GemmArgs g = g0;
    // The large-tile kernel takes operands contiguous along K, in whole aligned 8-element
    // vectors; an operand stored along M or N is transposed into the workspace first.
    auto al = [](const void* p) { return ((uintptr_t)p & 15) == 0; };
    bool large = g.K % 8 == 0 && g.ws != nullptr && g.ws_elems >= nn_gemm_workspace(g.M, g.N, g.K) && al(g.ws);
    const bool a_k = g.sak == 1 && g.sam % 8 == 0 && al(g.A);
    const bool b_k = g.sbk == 1 && g.sbn % 8 == 0 && al(g.B);
    const bool a_t = !a_k && g.sam == 1;  // stored [K x M], row stride sak
    const bool b_t = !b_k && g.sbn == 1;  // stored [K x N], row stride sbk
    const bool a_c = !a_k && !a_t && g.sak == 1;  // along K, unaligned: copied
    const bool b_c = !b_k && !b_t && g.sbk == 1;
    large = large && (a_k || a_t || a_c) && (b_k || b_t || b_c);
    if (large) {
        uint16_t* wa = g.ws;
        uint16_t* wb = g.ws + (((int64_t)g.M * g.K + 7) / 8) * 8;
        if (a_c) {
            const int64_t n = (int64_t)g.M * g.K;
            hipLaunchKernelGGL(nn::k_copy_rows, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g.A, g.sam, g.M,
                               g.K, wa);
            g.A = wa;
            g.sam = g.K;
        }
        if (b_c) {
            const int64_t n = (int64_t)g.N * g.K;
            hipLaunchKernelGGL(nn::k_copy_rows, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g.B, g.sbn, g.N,
                               g.K, wb);
            g.B = wb;
            g.sbn = g.K;
        }
        if (a_t) {
            hipLaunchKernelGGL(nn::k_transpose, dim3((g.M + 63) / 64, (g.K + 63) / 64), dim3(256), 0, st, g.A, g.sak,
                               g.K, g.M, wa);
            g.A = wa;
            g.sam = g.K;
            g.sak = 1;
        }
        if (b_t) {
            hipLaunchKernelGGL(nn::k_transpose, dim3((g.N + 63) / 64, (g.K + 63) / 64), dim3(256), 0, st, g.B, g.sbk,
                               g.K, g.N, wb);
            g.B = wb;
            g.sbn = g.K;
            g.sbk = 1;
        }
    }
    const int bm = large ? nn::LM : nn::BM, bn = large ? nn::LN : nn::BN, bkk = large ? nn::LK : nn::BK;
    const int tiles = ((g.N + bn - 1) / bn) * ((g.M + bm - 1) / bm);
    int split = 1;
    while (split < 64 && tiles * split < 512 && g.K / (split * 2) >= 256) split *= 2;
    const int kchunk = ((g.K + split - 1) / split + bkk - 1) / bkk * bkk;
    split = (g.K + kchunk - 1) / kchunk;
    if (split > 1) {
        const int64_t n = (int64_t)g.M * g.N;
        hipLaunchKernelGGL(nn::k_scale, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g.C, g.M, g.N, g.ldc,
                           g.beta);
    }
    dim3 grid((g.N + bn - 1) / bn, (g.M + bm - 1) / bm, split);
    const int kc = split > 1 ? kchunk : g.K;
    if (large) hipLaunchKernelGGL(nn::k_gemm_l, grid, dim3(256), 0, st, g, kc);
    else hipLaunchKernelGGL(nn::k_gemm, grid, dim3(256), 0, st, g, kc);
    return hipGetLastError();
}

hipError_t nn_to_bf16(const float* x, uint16_t* y, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(nn::k_f32_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x, y, n);
    return hipGetLastError();
}

}  // namespace rt
