// rt_kernels.hip — gfx950 kernels of the planar flip chains (domino, lozenge,
// rectangle-triangle) and their launchers.
//
// Two execution paths share the model rules of rt_models.hpp:
//   * stepwise:   one Rotate launch and one Update launch per step, generator states in
//                 HBM (16 B AoS per generator, coalesced) — the reference's shape.
//   * persistent: one launch runs many steps.  Each workgroup owns a TR x TC tile of
//                 positions and keeps the tile plus an H-wide halo of every colour class in
//                 LDS and the TinyMT states of all those positions in VGPRs.  It advances the
//                 halo generators redundantly (their streams do not depend on the tiling, so
//                 every copy draws the same numbers), runs H steps locally, then publishes its
//                 owned positions to a double-buffered global copy, announces the block on its
//                 flag line and waits only for its (up to 8) neighbour tiles' flags before
//                 reloading its halo.  Per-step HBM traffic is zero; bit-exactness holds
//                 because each position's state after a block of H steps depends only on
//                 inputs within distance H.  Double buffering is race-free because a tile can
//                 never be more than one block ahead of a neighbour.
#include "rt_internal.hpp"
#include "rt_packed.hpp"

#include <type_traits>
#include <utility>
#include "rt_models.hpp"

#include <algorithm>

namespace rt {

// ---------------------------------------------------------------------------------
// generator initialisation (InitTinyMT, e.g. dominokernel.cl:33-40)
// ---------------------------------------------------------------------------------
__global__ void k_init_gens(int64_t G, uint32_t seed, const uint4* __restrict__ params, uint32_t m1, uint32_t m2,
                            uint32_t tm, uint4* __restrict__ gens) {
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= G) return;
    uint32_t a = m1, b = m2, c = tm;
    if (params) {
        const uint4 p = params[id];
        a = p.x;
        b = p.y;
        c = p.z;
    }
    uint32_t s0, s1, s2, s3;
    tm_init(seed + (uint32_t)id, a, b, c, s0, s1, s2, s3);
    gens[id] = make_uint4(s0, s1, s2, s3);
}

hipError_t launch_init_gens(int64_t G, uint32_t seed, const uint4* params, uint32_t m1, uint32_t m2, uint32_t tm,
                            uint4* gens, hipStream_t st) {
    if (G <= 0) return hipSuccess;
    const int nt = 256;
    const int64_t nb = (G + nt - 1) / nt;
    hipLaunchKernelGGL(k_init_gens, dim3((unsigned)nb), dim3(nt), 0, st, G, seed, params, m1, m2, tm, gens);
    return hipGetLastError();
}

__global__ void k_fill_u32(uint32_t* p, int n, uint32_t v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_add_i64(int64_t* p, int64_t v) {
    if (threadIdx.x == 0) *p += v;
}

hipError_t launch_add_i64(int64_t* p, int64_t v, hipStream_t st) {
    hipLaunchKernelGGL(k_add_i64, dim3(1), dim3(64), 0, st, p, v);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, int n, uint32_t v, hipStream_t st) {
    hipLaunchKernelGGL(k_fill_u32, dim3((n + 255) / 256), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

// Coupling test of a coupled domino pair: counts the positions whose two chains differ, one
// wave64 ballot per 64 positions and one atomic per wave.
__global__ void k_pair_differs(const uint16_t* __restrict__ a, int64_t n, uint32_t* __restrict__ count) {
    uint32_t local = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
        const int64_t i = base + threadIdx.x;
        const bool differ = i < n && ((a[i] & 0xff) != (a[i] >> 8));
        local += (uint32_t)__popcll(__ballot(differ));
    }
    if ((threadIdx.x & 63) == 0 && local) atomicAdd(count, local);
}

hipError_t launch_pair_differs(const uint16_t* a, int64_t n, uint32_t* count, hipStream_t st) {
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, count, 1, 0u);
    const int64_t nb = std::min<int64_t>((n + 255) / 256, 2048);
    hipLaunchKernelGGL(k_pair_differs, dim3((unsigned)std::max<int64_t>(nb, 1)), dim3(256), 0, st, a, n, count);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// stepwise path
// ---------------------------------------------------------------------------------
template <class M>
__global__ void k_rotate(RuleParams rp, int c, typename M::T* __restrict__ cls, int N, int W, int P,
                         uint4* __restrict__ gens, const uint4* __restrict__ params, uint32_t m1, uint32_t m2,
                         uint32_t tm) {
    const int gj = blockIdx.x * blockDim.x + threadIdx.x;
    const int gi = blockIdx.y;
    if (gj >= W || !M::has_gen(gi, gj, N, W)) return;
    const int id = M::gen_id(gi, gj, N, W);
    uint4 s = gens[id];
    uint32_t a = m1, b = m2, t = tm;
    if (params) {
        const uint4 p = params[id];
        a = p.x;
        b = p.y;
        t = p.z;
    }
    const uint32_t k = tm_draw(s.x, s.y, s.z, s.w, a, b, t);
    gens[id] = s;
    typename M::T* arr = cls + (size_t)c * N * P;
    const typename M::T st = arr[gi * P + gj];
    const typename M::T ns = M::flip(M::step_rule(rp, c), gi, st, k);
    if (ns != st) arr[gi * P + gj] = ns;
}

template <class M>
__global__ void k_update(int c, typename M::T* __restrict__ cls, int N, int W, int P) {
    const int gj = blockIdx.x * blockDim.x + threadIdx.x;
    const int gi = blockIdx.y;
    if (gj >= W || !M::updatable(gi, gj, N, W)) return;
    M::update(c, gi, cls, N * P, gi * P + gj, P);
}

template <class M>
hipError_t launch_stepwise_t(const StepwiseArgs& a, hipStream_t st) {
    using T = typename M::T;
    T* cls = (T*)a.cls;
    const int nt = 256;
    dim3 grid((a.W + nt - 1) / nt, a.N);
    for (int64_t s = 0; s < a.nsteps; s++) {
        const int c = a.sched_host[s];
        hipLaunchKernelGGL(k_rotate<M>, grid, dim3(nt), 0, st, a.rp, c, cls, a.N, a.W, a.P, a.gens, a.params, a.m1,
                           a.m2, a.tm);
        hipLaunchKernelGGL(k_update<M>, grid, dim3(nt), 0, st, c, cls, a.N, a.W, a.P);
    }
    return hipGetLastError();
}

hipError_t launch_stepwise(int model, const StepwiseArgs& a, hipStream_t st) {
    switch (model) {
        case RT_DOMINO: return launch_stepwise_t<DominoModel>(a, st);
        case RT_LOZENGE: return launch_stepwise_t<LozengeModel>(a, st);
        case RT_RECTTRI: return launch_stepwise_t<RectTriModel>(a, st);
        case kDominoPair: return launch_stepwise_t<DominoPairModel>(a, st);
    }
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------------
// persistent temporal-blocked path
// ---------------------------------------------------------------------------------
// Exchange of tile borders between workgroups goes through the chain's global copy with
// agent-coherent 32-bit accesses (xstore/xload: relaxed atomics at agent scope, coherent
// across XCDs without writing back or invalidating whole L2s), rows being 16-byte aligned.

// f(std::integral_constant<int, i>) for i = 0 .. N-1, in order
template <class F, int... I>
__device__ __forceinline__ void static_for_seq(F&& f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
    static_for_seq(f, std::make_integer_sequence<int, N>{});
}

template <class M, int TR, int TC, int H, int NT, bool UNI>
__device__ __forceinline__ void persist_body(const PersistArgs& a, const RuleParams& rules, const int tile,
                                             const int ntiles) {
    using T = typename M::T;
    constexpr int NC = M::NC;
    constexpr int ER = TR + 2 * H, EC = TC + 2 * H;
    constexpr int AREA = ER * EC;
    constexpr int GPT = (AREA + NT - 1) / NT;
    constexpr int ES = (int)sizeof(T);
    static_assert(GPT <= 32, "position masks are 32 bits");
    static_assert((TC * ES) % 4 == 0 && (H * ES) % 4 == 0, "tile columns and halo must be whole dwords");
    constexpr int EW = EC * ES / 4, TW = TC * ES / 4, HW = H * ES / 4;  // dwords per ext row, owned row, halo
    constexpr int AREAW = ER * EW;
    __shared__ __align__(16) T lds[NC * AREA];
    uint32_t* lds32 = reinterpret_cast<uint32_t*>(lds);

    const int N = a.N, W = a.W;
    const int PW = a.P * ES / 4;             // dwords per global row
    const size_t CSW = (size_t)N * PW;       // dwords per class
    const int tiles_r = ntiles / a.tiles_c;
    const int tr = tile / a.tiles_c, tc = tile - (tile / a.tiles_c) * a.tiles_c;
    const int r0 = tr * TR - H, c0 = tc * TC - H;  // global coordinates of LDS (0, 0)
    const int c0w = c0 * ES / 4;                   // ... in dwords (c0 is a whole dword)
    const int tid = threadIdx.x;
    // The rule tables are indexed per lane (rectangle-triangle states), so they are staged in
    // LDS: indexed in place they would be copied to scratch memory, and a kernel that needs
    // scratch cannot rely on all its workgroups being resident.
    __shared__ RuleParams srp;
    for (int q = tid; q < (int)(sizeof(RuleParams) / 4); q += NT)
        reinterpret_cast<uint32_t*>(&srp)[q] = reinterpret_cast<const uint32_t*>(&rules)[q];
    __syncthreads();
    const RuleParams& rp = srp;

    // Rectangle-triangle tiles whose extended row is one wavefront (EC = 64): position
    // tid + g * NT lies in row wave + g * (NT / 64), so a wave's row, its residue mod 3 and
    // with them the update's neighbour offsets are the same on every lane.  The flip reads
    // one entry of a table merged from the rule (slot h of the state hash, band b of k):
    // {state, accept below, target}, where the acceptance bound already lies at or below the
    // gate and an empty slot accepts nothing.  Bounds and bands are compared with the raw
    // tempered output t0 (k = t0 >> 9) as t << 9: every one lies at or below the gate, which
    // is below 2^23, so k < t <=> t0 < t << 9 exactly.
    constexpr bool kRowWave = std::is_same<M, RectTriModel>::value && EC == 64 && NT % 64 == 0 && ER % (NT / 64) == 0;
    constexpr int NW = NT / 64;
    __shared__ uint4 rtab[kRowWave ? 128 : 1];
    uint32_t band0 = 0, band1 = 0, band2 = 0;
    if constexpr (kRowWave) {
        const RectTriRule& r = rp.rt;
        if (tid < 128) {
            const int h = tid >> 2, b = tid & 3;
            const uint32_t key = r.hkey[h], idx = r.hval[h];
            uint4 e = make_uint4(key, 0u, 0u, 0u);
            if (key != 0xffffffffu) {
                e.y = (idx < 5 ? r.cyc_acc[idx][b] : r.tog_acc[idx - 5]) << 9;
                e.z = (uint32_t)(idx < 5 ? r.cyc_nt[idx][b] : r.tog_nt[idx - 5]);
            }
            rtab[tid] = e;
        }
        band0 = r.band[0] << 9;
        band1 = r.band[1] << 9;
        band2 = r.band[2] << 9;
    }

    // generator states of every tile + halo position this thread handles
    uint32_t s0[GPT], s1[GPT], s2[GPT], s3[GPT];
    uint32_t pm1[UNI ? 1 : GPT], pm2[UNI ? 1 : GPT], pmt[UNI ? 1 : GPT];
    uint32_t gmask = 0, umask = 0, omask = 0;
#pragma unroll
    for (int g = 0; g < GPT; g++) {
        const int p = tid + g * NT;
        s0[g] = s1[g] = s2[g] = s3[g] = 0;
        if (!UNI) pm1[g] = pm2[g] = pmt[g] = 0;
        if (p < AREA) {
            const int li = p / EC, lj = p - (p / EC) * EC;
            const int gi = r0 + li, gj = c0 + lj;
            if (M::has_gen(gi, gj, N, W)) {
                gmask |= 1u << g;
                const int id = M::gen_id(gi, gj, N, W);
                const uint4 s = gload(a.gen_in + id);
                s0[g] = s.x;
                s1[g] = s.y;
                s2[g] = s.z;
                s3[g] = s.w;
                if (!UNI) {
                    const uint4 q = gload(a.params + id);
                    pm1[g] = q.x;
                    pm2[g] = q.y;
                    pmt[g] = q.z;
                }
            }
            if (li >= 1 && li <= ER - 2 && lj >= 1 && lj <= EC - 2 && M::updatable(gi, gj, N, W)) umask |= 1u << g;
            if (li >= H && li < H + TR && lj >= H && lj < H + TC && gi < N && gj < W) omask |= 1u << g;
        }
    }

    // load tile + halo of every class (dwords outside the array read as zero; the row
    // padding past W holds zeros)
    auto load_ext = [&](int buf, bool skip_owned) {
        const uint32_t* src = reinterpret_cast<const uint32_t*>((buf ? a.cls[1] : a.cls[0]));
        for (int q = tid; q < AREAW; q += NT) {
            const int li = q / EW, lw = q - (q / EW) * EW;
            if (skip_owned && li >= H && li < H + TR && lw >= HW && lw < HW + TW) continue;
            const int gi = r0 + li, gw = c0w + lw;
            const bool in = gi >= 0 && gi < N && gw >= 0 && gw < PW;
#pragma unroll
            for (int c = 0; c < NC; c++) lds32[c * AREAW + q] = in ? xload(src + c * CSW + (size_t)gi * PW + gw) : 0u;
        }
    };
    // The colours of a block (H <= 8 steps) come as one 8-byte word fetched ahead (the first
    // before the tile loads, each next one before the block is published), so no step waits on a
    // global load; every schedule buffer is padded past its last step.
    static_assert(H <= 8, "a block's colours are fetched as one 8-byte word");
    auto fetch_block = [&](int64_t at) { return gload(reinterpret_cast<const uint2*>(a.sched + at)); };
    uint2 colours = fetch_block(0);
    load_ext(a.buf, false);
    __syncthreads();

    // neighbour tiles whose borders this tile reads
    int nbr_tile = -1;
    if (tid < 9 && tid != 4) {
        const int dr = tid / 3 - 1, dc = tid % 3 - 1;
        const int rr = tr + dr, cc = tc + dc;
        if (rr >= 0 && rr < tiles_r && cc >= 0 && cc < a.tiles_c) nbr_tile = rr * a.tiles_c + cc;
    }

    // Row-wave layout: the residue mod 3 of the wave's first row (row g lies NW rows further
    // per g, so its residue is rw_base + g * NW), and per g whether the row is updated (both the
    // same on every lane); per lane, whether its column is updated.
    int rw_base = 0;
    uint32_t rw_upd = 0;
    bool rw_col = false;
    if constexpr (kRowWave) {
        const int wave = uniform(tid >> 6);
        rw_base = (((r0 + wave) % 3) + 3) % 3;
#pragma unroll
        for (int g = 0; g < GPT; g++) {
            const int li = wave + g * NW, gi = r0 + li;
            if (li >= 1 && li <= ER - 2 && gi >= 1 && gi <= N - 2) rw_upd |= 1u << g;
        }
        const int lane = tid & 63, gj = c0 + lane;
        rw_col = lane >= 1 && lane <= EC - 2 && gj >= 1 && gj <= W - 2;
    }

    int buf = a.buf;
    uint32_t nblk = 0;
    int64_t s = 0;
    while (s < a.nsteps) {
        const int64_t blk = (a.nsteps - s < H) ? (a.nsteps - s) : H;
        // A flip reads and writes only the thread's own positions of class c, whose words the
        // thread's own update wrote in the previous step; the updates read class c' (the
        // previous step's class) of the neighbours.  So the flip must wait for the previous
        // update only when c == c', which would overwrite what a neighbour may still read; the
        // block starts after a barrier.
        int c_prev = -1;
        const uint32_t clo = (uint32_t)uniform((int)colours.x), chi = (uint32_t)uniform((int)colours.y);
        for (int64_t q = 0; q < blk; q++, s++) {
            const int c = (int)(((q < 4 ? clo : chi) >> (8 * (q & 3))) & 0xffu);
            T* lc = lds + c * AREA;
            if (c == c_prev) __syncthreads();
            c_prev = c;
            if constexpr (kRowWave) {
                // Every lane draws: a position without a generator holds state 0, which no
                // table entry matches, and its draws are discarded.
#pragma unroll
                for (int g = 0; g < GPT; g++) {
                    const int p = tid + g * NT;
                    const uint32_t t0 = tm_draw_raw(s0[g], s1[g], s2[g], s3[g], UNI ? a.m1 : pm1[g],
                                                    UNI ? a.m2 : pm2[g], UNI ? a.tm : pmt[g]);
                    const uint32_t st = (uint32_t)lc[p];
                    const uint32_t h = ((uint32_t)__umul24(st, kRtHash) >> 27) & 31u;
                    const uint32_t b = (uint32_t)(t0 >= band0) + (uint32_t)(t0 >= band1) + (uint32_t)(t0 >= band2);
                    const uint4 e = rtab[h * 4 + b];  // one read; written back unchanged if no move
                    lc[p] = (T)(((e.x == st) & (t0 < e.y)) ? e.z : st);
                }
            } else {
                const auto& sr = M::step_rule(rp, c);
#pragma unroll
                for (int g = 0; g < GPT; g++) {
                    if (gmask & (1u << g)) {
                        const int p = tid + g * NT;
                        const uint32_t k = tm_draw(s0[g], s1[g], s2[g], s3[g], UNI ? a.m1 : pm1[g],
                                                   UNI ? a.m2 : pm2[g], UNI ? a.tm : pmt[g]);
                        const int gi = r0 + p / EC;
                        const T st = lc[p];
                        const T ns = M::flip(sr, gi, st, k);
                        if (ns != st) lc[p] = ns;
                    }
                }
            }
            __syncthreads();
            if constexpr (kRowWave) {
                // recttrianglekernel.cl:120-146 with e = (row - c) mod 3 choosing the shifted
                // neighbours (the same on the whole wave); each class word is two field
                // selects and one byte permutation:
                //   c+1: bytes (u.1, u.2, x.0, 0), u = y & 0xf00f00 | z & ~0xf00f00
                //   c+2: bytes (z.2, v.0, v.1, 0), v = x & 0xf00f | y & ~0xf00f
                // One branch per step on e0 = e of row g = 0; row g then has e = (e0 + g * NW)
                // mod 3, a constant, so every row's offsets are immediates.
                const T* sc = lds + c * AREA;
                T* d1p = lds + (c == 2 ? 0 : c + 1) * AREA;
                T* d2p = lds + (c == 0 ? 2 : c - 1) * AREA;
                auto upd_rows = [&](auto e0_c) __attribute__((always_inline)) {
                    static_for<GPT>([&](auto g_c) __attribute__((always_inline)) {
                        constexpr int G = decltype(g_c)::value;
                        constexpr int E = (decltype(e0_c)::value + G * NW) % 3;
                        if (!(rw_upd & (1u << G))) return;
                        const int p = tid + G * NT;
                        const uint32_t up = (uint32_t)sc[p - EC], mid = (uint32_t)sc[p], dn = (uint32_t)sc[p + EC];
                        const uint32_t y1 = E == 0 ? (uint32_t)sc[p + 1] : mid;
                        const uint32_t z1 = E == 1 ? (uint32_t)sc[p + EC - 1] : dn;
                        const uint32_t x2 = E == 2 ? (uint32_t)sc[p - 1] : mid;
                        const uint32_t y2 = E == 1 ? (uint32_t)sc[p - EC + 1] : up;
                        const uint32_t u = (y1 & 0xf00f00u) | (z1 & ~0xf00f00u);
                        const uint32_t v = (x2 & 0xf00fu) | (y2 & ~0xf00fu);
                        if (rw_col) {
                            d1p[p] = (T)__builtin_amdgcn_perm(up, u, 0x0c040201u);
                            d2p[p] = (T)__builtin_amdgcn_perm(dn, v, 0x0c010006u);
                        }
                    });
                };
                const int e0 = (rw_base - c + 3) % 3;
                if (e0 == 0) upd_rows(std::integral_constant<int, 0>{});
                else if (e0 == 1) upd_rows(std::integral_constant<int, 1>{});
                else upd_rows(std::integral_constant<int, 2>{});
            } else {
#pragma unroll
                for (int g = 0; g < GPT; g++) {
                    if (umask & (1u << g)) {
                        const int p = tid + g * NT;
                        const int gi = r0 + p / EC;
                        M::update(c, gi, lds, AREA, p, EC);
                    }
                }
            }
        }
        if (s < a.nsteps) colours = fetch_block(s);  // in flight while the block is published
        __syncthreads();  // every update of the block is in LDS before it is published
        // publish owned positions into the other copy
        {
            uint32_t* dst = reinterpret_cast<uint32_t*>((buf ? a.cls[0] : a.cls[1]));
            for (int q = tid; q < TR * TW; q += NT) {
                const int oi = q / TW, ow = q - (q / TW) * TW;
                const int gi = tr * TR + oi, gw = tc * TW + ow;
                if (gi < N && gw < PW) {
                    const int lq = (oi + H) * EW + HW + ow;
#pragma unroll
                    for (int c = 0; c < NC; c++) xstore(dst + c * CSW + (size_t)gi * PW + gw, lds32[c * AREAW + lq]);
                }
            }
        }
        buf ^= 1;
        if (s < a.nsteps) {
            // announce this block, wait for the neighbours' announcements, reload the halo
            nblk++;
            const uint32_t epoch = a.epoch0 + nblk;
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            if (tid == 0) xstore(a.flags + (size_t)tile * kLine, epoch);
            if (nbr_tile >= 0) {
                const uint32_t* f = a.flags + (size_t)nbr_tile * kLine;
                uint32_t polls = 0;
                while (!epoch_reached(xload(f), epoch)) {
                    if (__hip_atomic_load(a.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
                    if (++polls > a.spin_limit) {
                        __hip_atomic_store(a.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            __syncthreads();
            load_ext(buf, true);
            __syncthreads();
        }
    }

    // owned generators out
#pragma unroll
    for (int g = 0; g < GPT; g++) {
        if ((gmask & omask) & (1u << g)) {
            const int p = tid + g * NT;
            const int li = p / EC, lj = p - (p / EC) * EC;
            const int id = M::gen_id(r0 + li, c0 + lj, N, W);
            gstore(a.gen_out + id, make_uint4(s0[g], s1[g], s2[g], s3[g]));
        }
    }
}

template <class M, int TR, int TC, int H, int NT, bool UNI>
__global__ __launch_bounds__(NT) void k_persist(PersistArgs a) {
    persist_body<M, TR, TC, H, NT, UNI>(a, a.rp, xcd_tile(blockIdx.x, gridDim.x), gridDim.x);
}

// batched launches (blocked path): every block runs one tile of one chain of the batch
template <class M, int TR, int TC, int H, int NT, bool UNI>
__global__ __launch_bounds__(NT) void k_persist_batch(PersistArgs a_in) {
    PersistArgs a = a_in;
    int tile, ntiles;
    const RuleParams& rules = resolve_batch(a, tile, ntiles);
    persist_body<M, TR, TC, H, NT, UNI>(a, rules, tile, ntiles);
}

// The rectangle-triangle row-wave kernels with the parameters in kernel arguments, built for 8
// waves per SIMD (at most 64 VGPRs and 96 SGPRs): with the default ~112 SGPRs only 7 fit,
// which leaves a CU one 1024-thread workgroup, idle while that workgroup loads and stores its
// tile; with two, one computes while the other streams.  (Per-generator parameters do not fit
// 64 VGPRs without scratch; those kernels keep the plain wrappers.)
template <int TR, int TC, int H, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(8))) void k_persist_rw(PersistArgs a) {
    persist_body<RectTriModel, TR, TC, H, NT, true>(a, a.rp, xcd_tile(blockIdx.x, gridDim.x), gridDim.x);
}

template <int TR, int TC, int H, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(8))) void k_persist_rw_batch(PersistArgs a_in) {
    PersistArgs a = a_in;
    int tile, ntiles;
    const RuleParams& rules = resolve_batch(a, tile, ntiles);
    persist_body<RectTriModel, TR, TC, H, NT, true>(a, rules, tile, ntiles);
}

// tile configurations per model: {TR, TC, H, NT, packed}; packed selects the specialised
// sweep: the quad kernel for lozenges and dominoes, the row-wave body (extended rows of 64
// positions) for rectangle-triangles
struct Cfg {
    int tr, tc, h, nt, packed;
};

#define RT_DOMINO_CFGS(X)    \
    X(32, 32, 8, 256, 0)     \
    X(16, 32, 4, 256, 0)     \
    X(48, 64, 8, 1024, 0)    \
    X(32, 64, 8, 512, 0)     \
    X(64, 84, 8, 1024, 1)    \
    X(56, 84, 4, 768, 1)     \
    X(32, 44, 4, 256, 1)     \
    X(16, 44, 4, 256, 1)     \
    X(16, 44, 8, 256, 1)     \
    X(32, 44, 8, 512, 1)     \
    X(8, 44, 8, 256, 1)      \
    X(16, 28, 8, 256, 1)     \
    X(16, 20, 8, 256, 1)     \
    X(8, 20, 8, 256, 1)      \
    X(16, 44, 8, 512, 1)     \
    X(24, 28, 8, 512, 1)     \
    X(32, 44, 8, 768, 1)     \
    X(40, 36, 8, 768, 1)     \
    X(48, 48, 8, 1024, 1)
#define RT_LOZENGE_CFGS(X)  \
    X(64, 84, 8, 1024, 1)   \
    X(64, 84, 4, 1024, 1)   \
    X(32, 44, 4, 256, 1)    \
    X(32, 84, 8, 512, 1)    \
    X(64, 84, 8, 1024, 0)   \
    X(32, 84, 4, 512, 1)    \
    X(64, 44, 4, 512, 1)    \
    X(48, 56, 4, 512, 1)    \
    X(68, 68, 4, 768, 1)    \
    X(56, 84, 4, 768, 1)    \
    X(80, 56, 4, 768, 1)    \
    X(68, 68, 4, 512, 1)    \
    X(16, 28, 8, 256, 1)    \
    X(8, 20, 8, 256, 1)     \
    X(32, 44, 8, 768, 1)    \
    X(40, 36, 8, 768, 1)
#define RT_RECTTRI_CFGS(X)  \
    X(32, 32, 4, 256, 0)    \
    X(48, 48, 8, 1024, 1)   \
    X(16, 32, 4, 256, 0)    \
    X(56, 56, 4, 1024, 1)   \
    X(24, 56, 4, 512, 1)    \
    X(16, 56, 4, 384, 1)    \
    X(8, 56, 4, 256, 1)

#define RT_CFG_ENTRY(a, b, c, d, e) Cfg{a, b, c, d, e},
static const Cfg kDominoCfgs[] = {RT_DOMINO_CFGS(RT_CFG_ENTRY)};
static const Cfg kLozengeCfgs[] = {RT_LOZENGE_CFGS(RT_CFG_ENTRY)};
static const Cfg kRectTriCfgs[] = {RT_RECTTRI_CFGS(RT_CFG_ENTRY)};

static const Cfg* cfg_table(int model, int* n) {
    switch (model) {
        case RT_DOMINO: *n = sizeof(kDominoCfgs) / sizeof(Cfg); return kDominoCfgs;
        case RT_LOZENGE: *n = sizeof(kLozengeCfgs) / sizeof(Cfg); return kLozengeCfgs;
        case RT_RECTTRI: *n = sizeof(kRectTriCfgs) / sizeof(Cfg); return kRectTriCfgs;
        case kDominoPair: *n = sizeof(kDominoCfgs) / sizeof(Cfg); return kDominoCfgs;
    }
    *n = 0;
    return nullptr;
}

int persist_config_count(int model) {
    int n = 0;
    cfg_table(model, &n);
    return n;
}

bool persist_config(int model, int idx, int* tr, int* tc, int* h, int* nt) {
    int n = 0;
    const Cfg* t = cfg_table(model, &n);
    if (!t || idx < 0 || idx >= n) return false;
    *tr = t[idx].tr;
    *tc = t[idx].tc;
    *h = t[idx].h;
    *nt = t[idx].nt;
    return true;
}

bool persist_config_packed(int model, int idx) {
    int n = 0;
    const Cfg* t = cfg_table(model, &n);
    return t && idx >= 0 && idx < n && t[idx].packed != 0;
}

// `one`: the launch runs one block of at most h steps (a grid that is not resident).  The packed
// domino and coupled-pair sweeps have kernels built for that, without the neighbour exchange and
// with fewer registers (the 48 x 48 pair tile on 1024 threads fits two workgroups per CU: Aztec
// diamond 2048 coupled steps 24.9 -> 17.5 us; single domino chains 2 % faster).  A lozenge chain
// keeps the full kernel: the one-block build of the same tile measured slower (hexagon 3000 on
// 32 x 44 halo-8 tiles, 4.38e11 against 4.58e11 flip attempts/s, at equal occupancy).
template <class M, int TR, int TC, int H, int NT, int PACKED>
static const void* kernel_ptr(bool uni, bool batch, bool one) {
    if constexpr (PACKED != 0 && std::is_same<M, LozengeModel>::value) {
        if (batch)
            return uni ? (const void*)k_persist_loz4_batch<TR, TC, H, NT, true>
                       : (const void*)k_persist_loz4_batch<TR, TC, H, NT, false>;
        return uni ? (const void*)k_persist_loz4<TR, TC, H, NT, true> : (const void*)k_persist_loz4<TR, TC, H, NT, false>;
    } else if constexpr (PACKED != 0 && std::is_same<M, DominoModel>::value) {
        if (batch)
            return uni ? (const void*)k_persist_dom4_batch<TR, TC, H, NT, true>
                       : (const void*)k_persist_dom4_batch<TR, TC, H, NT, false>;
        if (one)
            return uni ? (const void*)k_persist_dom4_blk<TR, TC, H, NT, true>
                       : (const void*)k_persist_dom4_blk<TR, TC, H, NT, false>;
        return uni ? (const void*)k_persist_dom4<TR, TC, H, NT, true> : (const void*)k_persist_dom4<TR, TC, H, NT, false>;
    } else if constexpr (PACKED != 0 && std::is_same<M, DominoPairModel>::value) {
        if (batch) return nullptr;  // coupled pairs run alone (coupling from the past)
        if (one)
            return uni ? (const void*)k_persist_pair2_blk<TR, TC, H, NT, true>
                       : (const void*)k_persist_pair2_blk<TR, TC, H, NT, false>;
        return uni ? (const void*)k_persist_pair2<TR, TC, H, NT, true> : (const void*)k_persist_pair2<TR, TC, H, NT, false>;
    } else if constexpr (PACKED != 0 && std::is_same<M, RectTriModel>::value) {
        // persist_body takes the row-wave path for this geometry by itself
        static_assert(TC + 2 * H == 64 && NT % 64 == 0 && (TR + 2 * H) % (NT / 64) == 0,
                      "a row-wave rectangle-triangle tile has extended rows of one wavefront");
        if (batch)
            return uni ? (const void*)k_persist_rw_batch<TR, TC, H, NT>
                       : (const void*)k_persist_batch<M, TR, TC, H, NT, false>;
        return uni ? (const void*)k_persist_rw<TR, TC, H, NT> : (const void*)k_persist<M, TR, TC, H, NT, false>;
    } else if constexpr (PACKED != 0) {
        return nullptr;  // no packed kernel for this model
    } else {
        if (batch)
            return uni ? (const void*)k_persist_batch<M, TR, TC, H, NT, true>
                       : (const void*)k_persist_batch<M, TR, TC, H, NT, false>;
        return uni ? (const void*)k_persist<M, TR, TC, H, NT, true> : (const void*)k_persist<M, TR, TC, H, NT, false>;
    }
}

static const void* persist_kernel(int model, int idx, bool uni, bool batch = false, bool one = false) {
    int i = 0;
#define RT_PICK(M_, a, b, c, d, e)                              \
    if (i++ == idx) return kernel_ptr<M_, a, b, c, d, e>(uni, batch, one);
    switch (model) {
        case RT_DOMINO: {
#define X(a, b, c, d, e) RT_PICK(DominoModel, a, b, c, d, e)
            RT_DOMINO_CFGS(X)
#undef X
        } break;
        case RT_LOZENGE: {
#define X(a, b, c, d, e) RT_PICK(LozengeModel, a, b, c, d, e)
            RT_LOZENGE_CFGS(X)
#undef X
        } break;
        case RT_RECTTRI: {
#define X(a, b, c, d, e) RT_PICK(RectTriModel, a, b, c, d, e)
            RT_RECTTRI_CFGS(X)
#undef X
        } break;
        case kDominoPair: {
#define X(a, b, c, d, e) RT_PICK(DominoPairModel, a, b, c, d, e)
            RT_DOMINO_CFGS(X)
#undef X
        } break;
    }
#undef RT_PICK
    return nullptr;
}

// Number of workgroups of this configuration a persistent launch may use: one per CU at most.
// The occupancy query bounds what a CU could hold, not what the dispatcher places: a grid of
// 256-thread workgroups at 92 % of the queried occupancy (7.4 per CU) was seen not to be
// co-resident, and a persistent grid that is not co-resident waits on tiles that never start.
// One workgroup per CU is what the persistent sweeps are built around (large tiles, 1024
// threads) and is always placed.
int persist_capacity(int model, int idx, bool uni, int device) {
    const void* k = persist_kernel(model, idx, uni);
    int tr, tc, h, nt;
    if (!k || !persist_config(model, idx, &tr, &tc, &h, &nt)) return 0;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, nt, 0) != hipSuccess) return 0;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return 0;
    return std::min(per_cu, 1) * prop.multiProcessorCount;
}

// Waves per CU that launches of the configuration's kernel (the batched one with `batch`, the
// one-block one with `one`) keep in flight, from the occupancy query: in a blocked run, where
// workgroups stream their tiles through HBM, this is what hides the load latency.
int persist_waves_per_cu(int model, int idx, bool uni, bool batch, bool one) {
    const void* k = persist_kernel(model, idx, uni, batch, one);
    int tr, tc, h, nt;
    if (!k || !persist_config(model, idx, &tr, &tc, &h, &nt)) return 0;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, nt, 0) != hipSuccess) return 0;
    return per_cu * nt / 64;
}

// Whether the configuration's kernel needs scratch (private) memory.  Waves that need scratch
// are launched only as far as the runtime's scratch allocation reaches, which the occupancy
// figure above does not bound, so such a grid is never treated as co-resident.
bool persist_uses_scratch(int model, int idx, bool uni) {
    const void* k = persist_kernel(model, idx, uni);
    if (!k) return true;
    hipFuncAttributes fa;
    if (hipFuncGetAttributes(&fa, k) != hipSuccess) return true;
    return fa.localSizeBytes > 0;
}

hipError_t launch_persist(int model, int idx, bool uni, const PersistArgs& a, int nblocks, hipStream_t st,
                          bool coop) {
    int tr, tc, h, nt;
    if (!persist_config(model, idx, &tr, &tc, &h, &nt)) return hipErrorInvalidValue;
    // a plain launch of at most one block needs no exchange: the one-block kernels
    const void* k = persist_kernel(model, idx, uni, a.batch != nullptr, !coop && a.nsteps <= h);
    if (!k) return hipErrorInvalidValue;
    void* args[] = {(void*)&a};
    if (coop) return hipLaunchCooperativeKernel(k, dim3(nblocks), dim3(nt), args, 0, st);
    return hipLaunchKernel(k, dim3(nblocks), dim3(nt), args, 0, st);
}

}  // namespace rt
