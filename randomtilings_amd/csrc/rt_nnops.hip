// rt_nnops.hip — the element-wise and data-movement kernels of the warm-start U-Net
// (activations NHWC fp32: pixel-major, channel-minor; the convolutions run as bf16 matrix
// products on the matrix cores, rt_mfma.hip).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rt_internal.hpp"
#include "rt_mfma.hpp"

namespace rt {
namespace nn {

__device__ __forceinline__ uint16_t bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

#define RT_GRID(n) dim3((unsigned)(((n) + 255) / 256)), dim3(256)

// 3x3 patches with zero padding: col[p][(dy*3 + dx)*C + c] = x[pixel p shifted by (dy-1, dx-1)][c]
__global__ void k_im2col(const float* __restrict__ x, uint16_t* __restrict__ col, int B, int H, int W, int C) {
    const int64_t n = (int64_t)B * H * W * 9 * C;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int c = (int)(q % C);
    const int tap = (int)((q / C) % 9);
    const int64_t p = q / (9 * C);
    const int w = (int)(p % W), h = (int)((p / W) % H), b = (int)(p / ((int64_t)W * H));
    const int hh = h + tap / 3 - 1, ww = w + tap % 3 - 1;
    float v = 0.f;
    if (hh >= 0 && hh < H && ww >= 0 && ww < W) v = x[(((int64_t)b * H + hh) * W + ww) * C + c];
    col[q] = bf(v);
}

// Adjoint of im2col: dx[pixel][c] = sum over the 9 taps that read it.
__global__ void k_col2im(const float* __restrict__ dcol, float* __restrict__ dx, int B, int H, int W, int C) {
    const int64_t n = (int64_t)B * H * W * C;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int c = (int)(q % C);
    const int64_t p = q / C;
    const int w = (int)(p % W), h = (int)((p / W) % H), b = (int)(p / ((int64_t)W * H));
    float s = 0.f;
    for (int tap = 0; tap < 9; tap++) {
        // the output pixel whose tap `tap` reads (h, w)
        const int oh = h - (tap / 3 - 1), ow = w - (tap % 3 - 1);
        if (oh < 0 || oh >= H || ow < 0 || ow >= W) continue;
        s += dcol[((((int64_t)b * H + oh) * W + ow) * 9 + tap) * C + c];
    }
    dx[q] = s;
}

__global__ void k_bias_relu(float* __restrict__ y, const float* __restrict__ bias, int64_t P, int C, int relu) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= P * C) return;
    float v = y[q] + bias[q % C];
    if (relu) v = v > 0.f ? v : 0.f;
    y[q] = v;
}

// dy *= (y > 0) for ReLU outputs
__global__ void k_relu_back(float* __restrict__ dy, const float* __restrict__ y, int64_t n) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n && !(y[q] > 0.f)) dy[q] = 0.f;
}

// db[c] = sum_p dy[p][c] (one block per channel)
__global__ void k_bias_grad(const float* __restrict__ dy, float* __restrict__ db, int64_t P, int C) {
    const int c = blockIdx.x;
    float s = 0.f;
    for (int64_t p = threadIdx.x; p < P; p += blockDim.x) s += dy[p * C + c];
    __shared__ float red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[c] = red[0];
}

__global__ void k_pool2(const float* __restrict__ x, float* __restrict__ y, int B, int H, int W, int C) {
    const int Ho = H / 2, Wo = W / 2;
    const int64_t n = (int64_t)B * Ho * Wo * C;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int c = (int)(q % C);
    const int64_t p = q / C;
    const int w = (int)(p % Wo), h = (int)((p / Wo) % Ho), b = (int)(p / ((int64_t)Wo * Ho));
    const float* s = x + (((int64_t)b * H + 2 * h) * W + 2 * w) * C + c;
    y[q] = 0.25f * (s[0] + s[C] + s[(int64_t)W * C] + s[(int64_t)W * C + C]);
}

__global__ void k_pool2_back(const float* __restrict__ dy, float* __restrict__ dx, int B, int H, int W, int C) {
    const int64_t n = (int64_t)B * H * W * C;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int c = (int)(q % C);
    const int64_t p = q / C;
    const int w = (int)(p % W), h = (int)((p / W) % H), b = (int)(p / ((int64_t)W * H));
    dx[q] = 0.25f * dy[(((int64_t)b * (H / 2) + h / 2) * (W / 2) + w / 2) * C + c];
}

// nearest-neighbour upsampling to 2H x 2W, added into y (y += up(x)) when acc != 0
__global__ void k_up2(const float* __restrict__ x, float* __restrict__ y, int B, int H, int W, int C, int acc) {
    const int64_t n = (int64_t)B * (2 * H) * (2 * W) * C;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int c = (int)(q % C);
    const int64_t p = q / C;
    const int w = (int)(p % (2 * W)), h = (int)((p / (2 * W)) % (2 * H)), b = (int)(p / ((int64_t)4 * W * H));
    const float v = x[(((int64_t)b * H + h / 2) * W + w / 2) * C + c];
    y[q] = acc ? y[q] + v : v;
}

__global__ void k_up2_back(const float* __restrict__ dy, float* __restrict__ dx, int B, int H, int W, int C) {
    const int64_t n = (int64_t)B * H * W * C;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int c = (int)(q % C);
    const int64_t p = q / C;
    const int w = (int)(p % W), h = (int)((p / W) % H), b = (int)(p / ((int64_t)W * H));
    const int64_t W2 = 2 * W;
    const float* s = dy + (((int64_t)b * 2 * H + 2 * h) * W2 + 2 * w) * C + c;
    dx[q] = s[0] + s[C] + s[W2 * C] + s[W2 * C + C];
}

__global__ void k_add(float* __restrict__ y, const float* __restrict__ x, int64_t n) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) y[q] += x[q];
}

// masked mean squared error: per-pixel residual and d/dy; the loss sum in *acc[0], the mask
// sum in *acc[1] (atomics of wave sums)
__global__ void k_mse(const float* __restrict__ y, const float* __restrict__ t, const float* __restrict__ m,
                      float* __restrict__ dy, int64_t n, float inv_count, float* __restrict__ acc) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (q < n) {
        const float r = (y[q] - t[q]) * m[q];
        l = r * r;
        dy[q] = 2.f * r * m[q] * inv_count;
    }
    for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(acc, l);
}

// Adam (fp32 master weights) followed by the bf16 copy used by the matrix products
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, uint16_t* __restrict__ wbf, int64_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const float gq = g[q];
    const float a = b1 * m1[q] + (1.f - b1) * gq;
    const float b = b2 * m2[q] + (1.f - b2) * gq * gq;
    m1[q] = a;
    m2[q] = b;
    const float v = w[q] - lr * (a * c1) / (sqrtf(b * c2) + eps);
    w[q] = v;
    if (wbf) wbf[q] = bf(v);
}

}  // namespace nn

hipError_t nn_im2col(const float* x, uint16_t* col, int B, int H, int W, int C, hipStream_t st) {
    hipLaunchKernelGGL(nn::k_im2col, RT_GRID((int64_t)B * H * W * 9 * C), 0, st, x, col, B, H, W, C);
    return hipGetLastError();
}
hipError_t nn_col2im(const float* dcol, float* dx, int B, int H, int W, int C, hipStream_t st) {
    hipLaunchKernelGGL(nn::k_col2im, RT_GRID((int64_t)B * H * W * C), 0, st, dcol, dx, B, H, W, C);
    return hipGetLastError();
}
hipError_t nn_bias_relu(float* y, const float* b, int64_t P, int C, bool relu, hipStream_t st) {
    hipLaunchKernelGGL(nn::k_bias_relu, RT_GRID(P * C), 0, st, y, b, P, C, relu ? 1 : 0);
    return hipGetLastError();
}
hipError_t nn_relu_back(float* dy, const float* y, int64_t n, hipStream_t st) {
    hipLaunchKernelGGL(nn::k_relu_back, RT_GRID(n), 0, st, dy, y, n);
    return hipGetLastError();
}
hipError_t nn_bias_grad(const float* dy, float* db, int64_t P, int C, hipStream_t st) {
    hipLaunchKernelGGL(nn::k_bias_grad, dim3(C), dim3(256), 0, st, dy, db, P, C);
    return hipGetLastError();
}
hipError_t nn_pool2(const float* x, float* y, int B, int H, int W, int C, hipStream_t st) {
    hipLaunchKernelGGL(nn::k_pool2, RT_GRID((int64_t)B * (H / 2) * (W / 2) * C), 0, st, x, y, B, H, W, C);
    return hipGetLastError();
}
hipError_t nn_pool2_back(const float* dy, float* dx, int B, int H, int W, int C, hipStream_t st) {
    hipLaunchKernelGGL(nn::k_pool2_back, RT_GRID((int64_t)B * H * W * C), 0, st, dy, dx, B, H, W, C);
    return hipGetLastError();
}
hipError_t nn_up2(const float* x, float* y, int B, int H, int W, int C, bool acc, hipStream_t st) {
    hipLaunchKernelGGL(nn::k_up2, RT_GRID((int64_t)B * 4 * H * W * C), 0, st, x, y, B, H, W, C, acc ? 1 : 0);
    return hipGetLastError();
}
hipError_t nn_up2_back(const float* dy, float* dx, int B, int H, int W, int C, hipStream_t st) {
    hipLaunchKernelGGL(nn::k_up2_back, RT_GRID((int64_t)B * H * W * C), 0, st, dy, dx, B, H, W, C);
    return hipGetLastError();
}
hipError_t nn_add(float* y, const float* x, int64_t n, hipStream_t st) {
    hipLaunchKernelGGL(nn::k_add, RT_GRID(n), 0, st, y, x, n);
    return hipGetLastError();
}
hipError_t nn_mse(const float* y, const float* t, const float* m, float* dy, int64_t n, float inv_count, float* acc,
                  hipStream_t st) {
    hipLaunchKernelGGL(nn::k_mse, RT_GRID(n), 0, st, y, t, m, dy, n, inv_count, acc);
    return hipGetLastError();
}
hipError_t nn_adam(float* w, const float* g, float* m1, float* m2, uint16_t* wbf, int64_t n, float lr, float b1,
                   float b2, float eps, float c1, float c2, hipStream_t st) {
    hipLaunchKernelGGL(nn::k_adam, RT_GRID(n), 0, st, w, g, m1, m2, wbf, n, lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}

}  // namespace rt
