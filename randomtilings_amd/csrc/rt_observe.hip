// rt_observe.hip — height-function observables of domino chains, computed where the chains
// live (no tiling read-back).
//
// DominoTiler::TilingToHeightfunc (src/Domino/DominoTiler.cpp:423-461) walks the uncrossed
// edges from the first boundary vertex (height 0), a step from vertex (i, j) in direction nb
// changing the height by (nb > 1 ? 1 : -1) * ((i + j) even ? 1 : -1).  An edge crossed by a
// domino (its bit set in the vertex state) joins two vertices whose heights differ by -3 times
// that step (the walk around the domino takes three uncrossed steps the other way).  Boundary
// heights are the same for every tiling of the domain, so along a row each maximal run of
// vertices joined by region edges starts at a boundary vertex whose height is known (the
// domain's minimal height function holds it), and the heights of the run are a prefix sum of
// the horizontal steps: a segmented scan, one workgroup per row.  The heights of the masked
// vertices are then reduced (count, sum, sum of squares, sum of |256 h - ref|) with wave64
// reductions and one 64-bit atomic per wave.
#include <hip/hip_runtime.h>

#include "rt_internal.hpp"

namespace rt {

namespace {

constexpr int kObsThreads = 256;
constexpr int kObsWaves = kObsThreads / 64;

// (flag, value) pairs of a segmented sum: flag = a segment starts here (value absolute).
struct Seg {
    int f;
    int v;
};

__device__ __forceinline__ Seg seg_combine(Seg a, Seg b) { return Seg{a.f | b.f, b.f ? b.v : a.v + b.v}; }

// inclusive segmented scan over a wave64
__device__ __forceinline__ Seg wave_scan(Seg s, int lane) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int f = __shfl_up(s.f, d, 64);
        const int v = __shfl_up(s.v, d, 64);
        if (lane >= d) s = seg_combine(Seg{f, v}, s);
    }
    return s;
}

__device__ __forceinline__ int64_t wave_sum(int64_t x) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
    return x;
}

// The vertex state of (i, j) in the chain's class layout: class (i + j) & 1, row i, column j/2.
__device__ __forceinline__ int vstate(const uint8_t* cls, int N, int P, int i, int j) {
    return cls[(size_t)((i + j) & 1) * N * P + (size_t)i * P + (j >> 1)];
}

// the segment value of vertex (i, k): a run starts where the edge from (i, k-1) is not an edge
// of the region (neither square beside it is in the domain), or where (i, k) is outside
__device__ __forceinline__ Seg vertex_seg(const uint8_t* cls, const uint8_t* faces, const int32_t* hmin, int N, int P,
                                          int i, int k) {
    const int32_t hb = hmin[(size_t)i * N + k];
    if (hb == INT32_MAX) return Seg{1, 0};
    const int M = N - 1;
    const bool up = i >= 1 && k >= 1 && faces[(size_t)(i - 1) * M + (k - 1)];
    const bool dn = i < M && k >= 1 && faces[(size_t)i * M + (k - 1)];
    if (!(up || dn)) return Seg{1, hb};
    const int j = k - 1;
    const int step = (((i + j) & 1) == 0) ? 1 : -1;  // direction nb = 2 (towards j + 1)
    const bool crossed = (vstate(cls, N, P, i, j) & 8) != 0;
    return Seg{0, crossed ? -3 * step : step};
}

struct ObsChain {
    const uint8_t* cls;
};

__global__ __launch_bounds__(kObsThreads) void k_domino_observe(const ObsChain* __restrict__ chains, int N, int P,
                                                                 const uint8_t* __restrict__ faces,
                                                                 const int32_t* __restrict__ hmin,
                                                                 const uint8_t* __restrict__ mask,
                                                                 const int32_t* __restrict__ ref256,
                                                                 int32_t* __restrict__ heights,
                                                                 unsigned long long* __restrict__ out) {
    const int i = blockIdx.x, c = blockIdx.y;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint8_t* cls = chains[c].cls;
    const int chunk = (N + kObsThreads - 1) / kObsThreads;
    const int k0 = min(N, tid * chunk), k1 = min(N, k0 + chunk);
    // 1. this thread's run aggregate
    Seg agg{0, 0};
    for (int k = k0; k < k1; k++) agg = seg_combine(agg, vertex_seg(cls, faces, hmin, N, P, i, k));
    // 2. exclusive prefix of the aggregates over the workgroup
    __shared__ int wf[kObsWaves], wv[kObsWaves];
    Seg inc = wave_scan(agg, lane);
    if (lane == 63) {
        wf[wave] = inc.f;
        wv[wave] = inc.v;
    }
    __syncthreads();
    Seg carry{0, 0};
    for (int w = 0; w < wave; w++) carry = seg_combine(carry, Seg{wf[w], wv[w]});
    const int pf = __shfl_up(inc.f, 1, 64), pv = __shfl_up(inc.v, 1, 64);
    if (lane > 0) carry = seg_combine(carry, Seg{pf, pv});
    // 3. heights of this thread's vertices, and their statistics
    int64_t cnt = 0, s1 = 0, s2 = 0, sd = 0;
    Seg run = carry;
    for (int k = k0; k < k1; k++) {
        run = seg_combine(run, vertex_seg(cls, faces, hmin, N, P, i, k));
        const size_t x = (size_t)i * N + k;
        const bool in = hmin[x] != INT32_MAX;
        const int h = in ? run.v : INT32_MAX;
        if (heights) heights[(size_t)c * N * N + x] = h;
        if (in && mask[x]) {
            cnt += 1;
            s1 += h;
            s2 += (int64_t)h * h;
            if (ref256) {
                const int64_t dlt = (int64_t)h * 256 - ref256[x];
                sd += dlt < 0 ? -dlt : dlt;
            }
        }
    }
    cnt = wave_sum(cnt);
    s1 = wave_sum(s1);
    s2 = wave_sum(s2);
    sd = wave_sum(sd);
    if (lane == 0) {
        unsigned long long* o = out + (size_t)c * 4;
        atomicAdd(o + 0, (unsigned long long)cnt);
        atomicAdd(o + 1, (unsigned long long)s1);
        atomicAdd(o + 2, (unsigned long long)s2);
        atomicAdd(o + 3, (unsigned long long)sd);
    }
}

}  // namespace

hipError_t launch_domino_observe(const void* chains, int nchains, int N, int P, const uint8_t* faces,
                                 const int32_t* hmin, const uint8_t* mask, const int32_t* ref256, int32_t* heights,
                                 unsigned long long* out, hipStream_t st) {
    hipLaunchKernelGGL(k_domino_observe, dim3(N, nchains), dim3(kObsThreads), 0, st, (const ObsChain*)chains, N, P,
                       faces, hmin, mask, ref256, heights, out);
    return hipGetLastError();
}

}  // namespace rt
