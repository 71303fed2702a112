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

}  // namespace nn

hipError_t nn_gemm(const GemmArgs& g, hipStream_t st) {
    if (g.M <= 0 || g.N <= 0) return hipSuccess;
    const int tiles = ((g.N + nn::BN - 1) / nn::BN) * ((g.M + nn::BM - 1) / nn::BM);
    int split = 1;
    while (split < 64 && tiles * split < 512 && g.K / (split * 2) >= 256) split *= 2;
    const int kchunk = ((g.K + split - 1) / split + nn::BK - 1) / nn::BK * nn::BK;
    split = (g.K + kchunk - 1) / kchunk;
    if (split > 1) {
        const int64_t n = (int64_t)g.M * g.N;
        hipLaunchKernelGGL(nn::k_scale, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g.C, g.M, g.N, g.ldc,
                           g.beta);
    }
    dim3 grid((g.N + nn::BN - 1) / nn::BN, (g.M + nn::BM - 1) / nn::BM, split);
    hipLaunchKernelGGL(nn::k_gemm, grid, dim3(256), 0, st, g, split > 1 ? kchunk : g.K);
    return hipGetLastError();
}

hipError_t nn_to_bf16(const float* x, uint16_t* y, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(nn::k_f32_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x, y, n);
    return hipGetLastError();
}

}  // namespace rt
