// rt_mfma.hpp — host interface of the BF16 matrix-core kernels (learned warm start).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rt {

// C[m][n] (row stride ldc) = alpha * sum_k A(m, k) B(k, n) + beta * C[m][n], with
// A(m, k) = A[m * sam + k * sak] and B(k, n) = B[k * sbk + n * sbn] in bf16.
// ws (ws_elems bf16, 16-byte aligned; may be null): room to transpose an operand that is
// stored along M or N, which lets the product use the large-tile kernel (nn_gemm_workspace
// elements suffice for any layout).
struct GemmArgs {
    const uint16_t* A;
    const uint16_t* B;
    float* C;
    int M, N, K;
    int64_t sam, sak, sbk, sbn, ldc;
    float alpha, beta;
    uint16_t* ws = nullptr;
    int64_t ws_elems = 0;
};

int64_t nn_gemm_workspace(int M, int N, int K);
hipError_t nn_gemm(const GemmArgs& g, hipStream_t st);
hipError_t nn_to_bf16(const float* x, uint16_t* y, int64_t n, hipStream_t st);

}  // namespace rt

namespace rt {

// Element-wise and data-movement kernels of the U-Net (NHWC fp32 activations).
hipError_t nn_im2col(const float* x, uint16_t* col, int B, int H, int W, int C, hipStream_t st);
hipError_t nn_col2im(const float* dcol, float* dx, int B, int H, int W, int C, hipStream_t st);
hipError_t nn_bias_relu(float* y, const float* b, int64_t P, int C, bool relu, hipStream_t st);
hipError_t nn_relu_back(float* dy, const float* y, int64_t n, hipStream_t st);
hipError_t nn_bias_grad(const float* dy, float* db, int64_t P, int C, hipStream_t st);
hipError_t nn_pool2(const float* x, float* y, int B, int H, int W, int C, hipStream_t st);
hipError_t nn_pool2_back(const float* dy, float* dx, int B, int H, int W, int C, hipStream_t st);
hipError_t nn_up2(const float* x, float* y, int B, int H, int W, int C, bool acc, hipStream_t st);
hipError_t nn_up2_back(const float* dy, float* dx, int B, int H, int W, int C, hipStream_t st);
hipError_t nn_add(float* y, const float* x, int64_t n, hipStream_t st);
hipError_t nn_mse(const float* y, const float* t, const float* m, float* dy, int64_t n, float inv_count, float* acc,
                  hipStream_t st);
hipError_t nn_adam(float* w, const float* g, float* m1, float* m2, uint16_t* wbf, int64_t n, float lr, float b1,
                   float b2, float eps, float c1, float c2, hipStream_t st);

}  // namespace rt
