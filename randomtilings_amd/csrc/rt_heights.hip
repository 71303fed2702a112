// rt_heights.hip — height functions on the device.
//
// Thurston's construction of the maximal and minimal tilings (DominoTiler.cpp:214-400,
// LozengeTiler.cpp:140-318 of the reference) is a shortest-path problem: the maximal height
// function is the min-plus closure of the boundary heights over the domain's directed "+1"
// edges, the minimal one the max-plus closure over the "-1" edges.  Here the boundary walk
// (O(perimeter)) runs on the host; the closure runs on the GPU as a chaotic relaxation to its
// fixpoint: every workgroup loads a 64 x 64 tile of heights with a one-vertex halo into LDS
// and relaxes it in place until the tile is stable, then writes back what dropped; launches
// repeat until one changes nothing.  Min-plus relaxation is monotone, so every fair order of
// updates reaches the same fixpoint the host's bucket-ordered search computes.  The domain
// has no tiling iff the fixpoint lowers a boundary height.  The same kernel projects an
// arbitrary height field onto the admissible ones below it (the largest admissible field
// bounded by the input), which seeds chains from predicted heights.  The tiling is then read
// off the heights per vertex (HeightfuncToTiling).
#include <algorithm>
#include <climits>
#include <cmath>
#include <string>
#include <cstring>
#include <vector>

#include "../src/common/heights.hpp"
#include "../src/common/lattice.hpp"
#include "rt_internal.hpp"

namespace rt {
namespace hf {

constexpr int kInf = INT_MAX;
constexpr int kT = 64;    // tile side
constexpr int kNT = 256;  // threads per workgroup

struct Stencil {
    int deg;
    int di[6], dj[6], bit[6];
    int lozenge;  // read-off rule: lozenge (|dh| == 2) or domino (|dh|/2)
};

// One launch of the relaxation.  g holds dir * height (kInf = unset); pull[v] bit nb means
// v takes g(v - d[nb]) + 1 as an upper bound.
__global__ __launch_bounds__(kNT) void k_relax(int32_t* __restrict__ g, const uint8_t* __restrict__ pull, int N,
                                                Stencil st, int maxit, uint32_t* __restrict__ changed) {
    constexpr int RW = kT + 2;
    __shared__ int32_t lg[RW * RW];
    __shared__ uint8_t lm[kT * kT];
    const int tiles_c = (N + kT - 1) / kT;
    const int ti = blockIdx.x / tiles_c, tj = blockIdx.x % tiles_c;
    const int i0 = ti * kT - 1, j0 = tj * kT - 1;  // global coordinates of LDS (0, 0)
    for (int q = threadIdx.x; q < RW * RW; q += kNT) {
        const int gi = i0 + q / RW, gj = j0 + q % RW;
        lg[q] = (gi >= 0 && gj >= 0 && gi < N && gj < N) ? g[(size_t)gi * N + gj] : kInf;
    }
    for (int q = threadIdx.x; q < kT * kT; q += kNT) {
        const int gi = ti * kT + q / kT, gj = tj * kT + q % kT;
        lm[q] = (gi < N && gj < N) ? pull[(size_t)gi * N + gj] : 0;
    }
    int off[6];
#pragma unroll
    for (int nb = 0; nb < 6; nb++) off[nb] = nb < st.deg ? st.di[nb] * RW + st.dj[nb] : 0;
    __syncthreads();
    for (int it = 0; it < maxit; it++) {
        bool any = false;
        for (int q = threadIdx.x; q < kT * kT; q += kNT) {
            const int m = lm[q];
            if (!m) continue;
            const int c = (q / kT + 1) * RW + q % kT + 1;
            const int cur = lg[c];
            int best = cur;
#pragma unroll
            for (int nb = 0; nb < 6; nb++) {
                if (!((m >> nb) & 1)) continue;
                const int u = lg[c - off[nb]];
                if (u != kInf && u + 1 < best) best = u + 1;
            }
            if (best < cur) {
                lg[c] = best;
                any = true;
            }
        }
        if (!__syncthreads_or(any)) break;
    }
    bool wrote = false;
    for (int q = threadIdx.x; q < kT * kT; q += kNT) {
        const int gi = ti * kT + q / kT, gj = tj * kT + q % kT;
        if (gi >= N || gj >= N) continue;
        const int v = lg[(q / kT + 1) * RW + q % kT + 1];
        const size_t x = (size_t)gi * N + gj;
        if (v < g[x]) {  // only this workgroup writes its tile
            g[x] = v;
            wrote = true;
        }
    }
    if (__syncthreads_or(wrote) && threadIdx.x == 0) atomicOr(changed, 1u);
}

// Counts fixed vertices whose value moved.
__global__ void k_fixed_moved(const int32_t* __restrict__ g, const int32_t* __restrict__ g0,
                              const uint8_t* __restrict__ pull, int64_t n, uint32_t* __restrict__ moved) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool bad = i < n && (pull[i] & 0x80) && g[i] != g0[i];
    const uint64_t b = __ballot(bad);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(moved, (uint32_t)__popcll(b));
}

// HeightfuncToTiling (DominoTiler.cpp:402-421, LozengeTiler.cpp:321-342): per vertex, the
// edges inside the domain across which the height jumps by 3 (domino) or 2 (lozenge) are
// crossed.
__global__ void k_read_tiling(const int32_t* __restrict__ g, const uint8_t* __restrict__ rmask, int N, Stencil st,
                              int dir, int32_t* __restrict__ t, int32_t* __restrict__ h) {
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= (int64_t)N * N) return;
    const int i = (int)(x / N), j = (int)(x % N);
    const int g0 = g[x];
    if (h) h[x] = (g0 == kInf) ? kInf : dir * g0;
    int s = 0;
    const int m = rmask[x];
    if (g0 != kInf) {
        for (int nb = 0; nb < st.deg; nb++) {
            if (!((m >> nb) & 1)) continue;
            const int g1 = g[(size_t)(i + st.di[nb]) * N + j + st.dj[nb]];
            if (g1 == kInf) continue;
            const int dh = dir * (g1 - g0);
            if (st.lozenge) s += (dh == 2 || dh == -2) ? st.bit[nb] : 0;
            else s += st.bit[nb] * (dh / 2 < 0 ? -(dh / 2) : dh / 2);
        }
    }
    t[x] = s;
}

// Host side: the per-vertex masks of a model's domain.
template <class Geom>
void masks(const Geom& geo, const std::vector<int>& vertices, int dir, bool lozenge, std::vector<uint8_t>& pull,
           std::vector<uint8_t>& rmask) {
    const int N = geo.N;
    pull.assign((size_t)N * N, 0);
    rmask.assign((size_t)N * N, 0);
    rtlat::for_rows(N, [&](int i) {
        for (int j = 0; j < N; j++) {
            const size_t v = (size_t)i * N + j;
            for (int nb = 0; nb < Geom::DEG; nb++) {
                // edge u -> v in direction nb, u = v - d[nb]: thurston_closure's relaxation
                const int ui = i - Geom::di[nb], uj = j - Geom::dj[nb];
                if (vertices[v] != 0 && ui >= 0 && uj >= 0 && ui < N && uj < N && vertices[(size_t)ui * N + uj] != 0 &&
                    geo.delta(ui, uj, nb) == dir) {
                    bool a, b;
                    geo.faces(ui, uj, nb, a, b);
                    if (a || b) pull[v] |= (uint8_t)(1u << nb);
                }
                // read-off edges v -> v + d[nb] (the lozenge read-off skips the outer frame)
                const int wi = i + Geom::di[nb], wj = j + Geom::dj[nb];
                if (wi < 0 || wj < 0 || wi >= N || wj >= N) continue;
                if (lozenge && (i < 1 || j < 1 || i > N - 2 || j > N - 2)) continue;
                bool a, b;
                geo.faces(i, j, nb, a, b);
                if (a || b) rmask[v] |= (uint8_t)(1u << nb);
            }
        }
    });
}

template <class Geom>
Stencil stencil(bool lozenge) {
    Stencil s{};
    s.deg = Geom::DEG;
    for (int nb = 0; nb < Geom::DEG; nb++) {
        s.di[nb] = Geom::di[nb];
        s.dj[nb] = Geom::dj[nb];
        s.bit[nb] = Geom::bit[nb];
    }
    s.lozenge = lozenge;
    return s;
}

struct DevBufs {
    int32_t *g = nullptr, *g0 = nullptr, *t = nullptr, *h = nullptr;
    uint8_t *pull = nullptr, *rmask = nullptr;
    uint32_t* flag = nullptr;
    ~DevBufs() {
        for (void* p : {(void*)g, (void*)g0, (void*)t, (void*)h, (void*)pull, (void*)rmask, (void*)flag})
            if (p) hipFree(p);
    }
};

// Relax to the fixpoint; returns the number of fixed vertices that moved, -1 on a HIP error,
// -2 if the launch bound passed before the field settled.
int64_t relax(DevBufs& b, int N, const Stencil& st, hipStream_t s, int* launches) {
    const size_t n = (size_t)N * N;
    if (hipMemcpyAsync(b.g0, b.g, n * 4, hipMemcpyDeviceToDevice, s) != hipSuccess) return -1;
    const int tiles = ((N + kT - 1) / kT) * ((N + kT - 1) / kT);
    // a launch settles every tile given its halo, so information crosses at least one tile
    // per launch; the bound covers any path through the grid
    const int max_launches = 64 * tiles + 64;
    *launches = 0;
    bool settled = false;
    for (int l = 0; l < max_launches && !settled; l++) {
        uint32_t changed = 0;
        if (launch_fill_u32(b.flag, 1, 0u, s) != hipSuccess) return -1;
        hipLaunchKernelGGL(k_relax, dim3(tiles), dim3(kNT), 0, s, b.g, b.pull, N, st, 4 * kT * kT, b.flag);
        if (hipGetLastError() != hipSuccess) return -1;
        if (hipMemcpyAsync(&changed, b.flag, 4, hipMemcpyDeviceToHost, s) != hipSuccess) return -1;
        if (hipStreamSynchronize(s) != hipSuccess) return -1;
        (*launches)++;
        settled = changed == 0;
    }
    if (!settled) return -2;
    uint32_t moved = 0;
    if (launch_fill_u32(b.flag, 1, 0u, s) != hipSuccess) return -1;
    hipLaunchKernelGGL(k_fixed_moved, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, b.g, b.g0, b.pull,
                       (int64_t)n, b.flag);
    if (hipMemcpyAsync(&moved, b.flag, 4, hipMemcpyDeviceToHost, s) != hipSuccess) return -1;
    if (hipStreamSynchronize(s) != hipSuccess) return -1;
    return moved;
}

}  // namespace hf

// Shared driver of rt_height_closure / rt_extremal_tiling.  g0h: initial heights (kInf =
// unset), fixed: vertices that may not move.  Returns RT_OK, RT_ERR_UNTILEABLE or an error.
int height_closure_device(int device, hipStream_t s, int model, const std::vector<int>& d, int dir,
                          std::vector<int32_t>& heights, const std::vector<uint8_t>& fixed, std::vector<int32_t>* tiling,
                          int* launches, std::string& err) {
    using namespace hf;
    const bool loz = model == RT_LOZENGE;
    const int N = loz ? (int)std::lround(std::sqrt(d.size() / 2.0)) + 1 : rtlat::side_of(d.size()) + 1;
    std::vector<uint8_t> pull, rmask;
    Stencil st;
    if (loz) {
        const std::vector<int> v = rtlat::lozenge_vertices(d);
        rtlat::LozGeom geo(d, N);
        masks(geo, v, dir, true, pull, rmask);
        st = stencil<rtlat::LozGeom>(true);
    } else {
        const std::vector<int> v = rtlat::domino_vertices(d);
        rtlat::DominoGeom geo(d, N);
        masks(geo, v, dir, false, pull, rmask);
        st = stencil<rtlat::DominoGeom>(false);
    }
    const size_t n = (size_t)N * N;
    std::vector<int32_t> g(n);
    rtlat::for_rows(N, [&](int i) {
        for (size_t k = (size_t)i * N; k < (size_t)(i + 1) * N; k++) {
            g[k] = heights[k] == kInf ? kInf : dir * heights[k];
            if (fixed[k]) pull[k] |= 0x80;
        }
    });
    DevBufs b;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipMalloc(&b.g, n * 4);
    if (e == hipSuccess) e = hipMalloc(&b.g0, n * 4);
    if (e == hipSuccess) e = hipMalloc(&b.pull, n);
    if (e == hipSuccess) e = hipMalloc(&b.flag, 64);
    if (tiling && e == hipSuccess) e = hipMalloc(&b.t, n * 4);
    if (tiling && e == hipSuccess) e = hipMalloc(&b.h, n * 4);
    if (tiling && e == hipSuccess) e = hipMalloc(&b.rmask, n);
    if (e == hipSuccess) e = hipMemcpyAsync(b.g, g.data(), n * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(b.pull, pull.data(), n, hipMemcpyHostToDevice, s);
    if (tiling && e == hipSuccess) e = hipMemcpyAsync(b.rmask, rmask.data(), n, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) {
        err = hipGetErrorString(e);
        return RT_ERR_HIP;
    }
    const int64_t moved = relax(b, N, st, s, launches);
    if (moved == -2) {
        err = "the relaxation did not settle within its launch bound";
        return RT_ERR_DEVICE;
    }
    if (moved < 0) {
        err = "height closure launch failed";
        return RT_ERR_HIP;
    }
    if (moved > 0) return RT_ERR_UNTILEABLE;
    if (tiling) {
        hipLaunchKernelGGL(k_read_tiling, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, b.g, b.rmask, N, st, dir,
                           b.t, b.h);
        tiling->resize(n);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(tiling->data(), b.t, n * 4, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipMemcpyAsync(heights.data(), b.h, n * 4, hipMemcpyDeviceToHost, s);
    } else {
        e = hipMemcpyAsync(g.data(), b.g, n * 4, hipMemcpyDeviceToHost, s);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        err = hipGetErrorString(e);
        return RT_ERR_HIP;
    }
    if (!tiling)
        for (size_t k = 0; k < n; k++) heights[k] = g[k] == kInf ? kInf : dir * g[k];
    return RT_OK;
}

// Boundary heights of a domain (the host walk of the max/min construction); false when the
// boundary does not close.
bool boundary_heights_host(int model, const std::vector<int>& d, std::vector<int32_t>& h, std::vector<uint8_t>& fixed) {
    std::vector<int> hf;
    bool ok;
    if (model == RT_LOZENGE) {
        const std::vector<int> v = rtlat::lozenge_vertices(d);
        rtlat::LozGeom geo(d, rtlat::side_of(v.size()));
        ok = rtapi::boundary_heights(geo, v, hf);
    } else {
        const std::vector<int> v = rtlat::domino_vertices(d);
        rtlat::DominoGeom geo(d, rtlat::side_of(v.size()));
        ok = rtapi::boundary_heights(geo, v, hf);
    }
    h.assign(hf.begin(), hf.end());
    fixed.assign(h.size(), 0);
    for (size_t k = 0; k < h.size(); k++) fixed[k] = h[k] != hf::kInf;
    return ok;
}

}  // namespace rt
