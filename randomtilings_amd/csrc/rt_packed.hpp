// rt_packed.hpp — persistent lozenge and domino sweeps on packed words (gfx950).
//
// Same scheme as the generic persistent kernel (tile + halo in LDS, generators in VGPRs,
// neighbour-flag exchange), but each thread works on whole 32-bit LDS words of one
// colour-class row: four horizontally adjacent positions ("quads": lozenge and domino, one
// byte each) or two (coupled domino pairs, 16 bits each).  All positions of a word share the
// row index, so the update's row-dependent column offsets (lozengekernel.cl:45-48,
// dominokernel.cl:22-31) are uniform over the word and become byte funnel shifts of
// neighbouring words, and the update of a word is a handful of 32-bit mask/shift operations
// instead of byte-wise passes.
//
// Flip (lozengekernel.cl:21-24 / qlozengekernel.cl:23-29, dominokernelCFTP.cl:25-29 as
// thresholds): a byte of the word toggles between the model's two flippable states (lozenge
// 25 and 38, 25 ^ 38 = 63; domino 3 and 12, 3 ^ 12 = 15) when its state and its position's
// generator's draw say so; the toggles of the word are applied with one xor.
#pragma once

#include "rt_internal.hpp"
#include "rt_models.hpp"

#include <type_traits>

namespace rt {

__device__ __forceinline__ void xstore(uint32_t* p, uint32_t v) {
    __hip_atomic_store((RT_GLOBAL uint32_t*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t xload(const uint32_t* p) {
    return __hip_atomic_load((const RT_GLOBAL uint32_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// byte-aligned funnel shift: bytes [sh, sh+4) of the little-endian pair (lo, hi)
__device__ __forceinline__ uint32_t funnel(uint32_t hi, uint32_t lo, int sh) {
    return __builtin_amdgcn_alignbyte(hi, lo, sh);
}

// The packed models.  Lozenge: three classes, the row index mod 3 selects the update's
// offsets, one rule for every class.  Domino: two classes, the row parity selects the
// offsets and (weighted rules) the rule; only rules whose two intervals split k at one
// threshold t (k < t: 3 -> 12, k >= t: 12 -> 3; dominokernelCFTP.cl's rule and the weighted
// heat-bath) run packed, with 0 < t < 2^23 (checked on the host).
struct PackedLozenge {
    using Model = LozengeModel;
    static constexpr int NC = 3, QPW = 4;  // classes, positions per 32-bit word
    static constexpr bool kLozenge = true;
    __device__ __forceinline__ static int rkey(int gi) { return ((gi % 3) + 3) % 3; }
};
struct PackedDomino {
    using Model = DominoModel;
    static constexpr int NC = 2, QPW = 4;
    static constexpr bool kLozenge = false;
    __device__ __forceinline__ static int rkey(int gi) { return gi & 1; }
};
// Coupled domino pairs (coupling from the past): a position is 16 bits, the chain from the
// maximal tiling in the low byte and the one from the minimal tiling in the high byte, so a
// word holds two positions.  One draw decides both bytes' flips and the update's bit moves
// stay inside bytes, so the domino code runs unchanged on bytes, with a position two bytes wide.
struct PackedDominoPair {
    using Model = DominoPairModel;
    static constexpr int NC = 2, QPW = 2;
    static constexpr bool kLozenge = false;
    __device__ __forceinline__ static int rkey(int gi) { return gi & 1; }
};

// ONE: the launch runs a single block of at most H steps (blocked launches, which stream every
// tile through HBM per block), so the neighbour exchange and the drawing ahead (which only hides
// exchange latency) are compiled out, and their registers with them; a launch of more steps
// raises the error flag (value 3) and runs nothing.  SKIPCHK: the check of halo words inside
// skipped neighbours is compiled in (batched launches never skip tiles and leave it out).
template <class PM, int TR, int TC, int H, int NT, bool UNI, bool ONE = false, bool SKIPCHK = true>
__device__ __forceinline__ void packed_body(const PersistArgs& a, const RuleParams& rules, const int tile,
                                            const int ntiles) {
    if constexpr (ONE) {
        if (a.nsteps > H) {
            if (threadIdx.x == 0) __hip_atomic_store(a.err, 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return;
        }
    }
    using M = typename PM::Model;
    constexpr int NC = PM::NC;
    constexpr int ER = TR + 2 * H, EC = TC + 2 * H;
    constexpr int QPW = PM::QPW, BPP = 4 / QPW;  // positions per word, bytes per position
    static_assert(TC % QPW == 0 && H % QPW == 0, "words must tile the rows");
    constexpr int EW = EC / QPW, TW = TC / QPW, HW = H / QPW;  // words per ext row, owned row, halo
    // the byte mask of position b of a word
    auto posmask = [](int b) { return (BPP == 1 ? 0xffu : 0xffffu) << (8 * BPP * b); };
    constexpr int AREAW = ER * EW;                       // words per class
    constexpr int QPT = (AREAW + NT - 1) / NT;           // words ("quads") per thread
    constexpr int GUARD = EW + 4;                        // words before and after the classes
    __shared__ uint32_t lds[GUARD + NC * AREAW + GUARD];
    uint32_t* L = lds + GUARD;

    const int N = a.N, W = a.W;
    const int PW = a.P / QPW;
    const size_t CSW = (size_t)N * PW;
    const int tiles_r = ntiles / a.tiles_c;
    const int tr = tile / a.tiles_c, tc = tile - (tile / a.tiles_c) * a.tiles_c;
    const int r0 = tr * TR - H, c0 = tc * TC - H;
    const int c0w = c0 / QPW;
    const int tid = threadIdx.x;

    // per word: a generator per position, the byte masks of positions with a generator /
    // updatable / owned, and the row key (lozenge: row mod 3; domino: row parity)
    uint32_t s0[QPT][QPW], s1[QPT][QPW], s2[QPT][QPW], s3[QPT][QPW];
    uint32_t pm1[UNI ? 1 : QPT][UNI ? 1 : QPW], pm2[UNI ? 1 : QPT][UNI ? 1 : QPW], pmt[UNI ? 1 : QPT][UNI ? 1 : QPW];
    uint32_t gbytes[QPT], ubytes[QPT];
    uint32_t omask = 0;
    int rmod[QPT];
#pragma unroll
    for (int g = 0; g < QPT; g++) {
        const int w = tid + g * NT;
        gbytes[g] = ubytes[g] = 0;
        rmod[g] = 0;
#pragma unroll
        for (int b = 0; b < QPW; b++) {
            s0[g][b] = s1[g][b] = s2[g][b] = s3[g][b] = 0;
            if (!UNI) pm1[g][b] = pm2[g][b] = pmt[g][b] = 0;
        }
        if (w < AREAW) {
            const int li = w / EW, lw = w - (w / EW) * EW;
            const int gi = r0 + li;
            rmod[g] = PM::rkey(gi);
            if (li >= H && li < H + TR && lw >= HW && lw < HW + TW) omask |= 1u << g;
#pragma unroll
            for (int b = 0; b < QPW; b++) {
                const int lj = QPW * lw + b, gj = c0 + lj;
                if (M::has_gen(gi, gj, N, W)) {
                    gbytes[g] |= posmask(b);
                    const int id = M::gen_id(gi, gj, N, W);
                    const uint4 s = gload(a.gen_in + id);
                    s0[g][b] = s.x;
                    s1[g][b] = s.y;
                    s2[g][b] = s.z;
                    s3[g][b] = s.w;
                    if (!UNI) {
                        const uint4 q = gload(a.params + id);
                        pm1[g][b] = q.x;
                        pm2[g][b] = q.y;
                        pmt[g][b] = q.z;
                    }
                }
                if (li >= 1 && li <= ER - 2 && lj >= 1 && lj <= EC - 2 && M::updatable(gi, gj, N, W))
                    ubytes[g] |= posmask(b);
            }
        }
    }

    for (int q = tid; q < GUARD; q += NT) {
        lds[q] = 0;
        lds[GUARD + NC * AREAW + q] = 0;
    }
    auto load_ext = [&](int buf, bool skip_owned) {
        const uint32_t* src = reinterpret_cast<const uint32_t*>((buf ? a.cls[1] : a.cls[0]));
        for (int q = tid; q < AREAW; q += NT) {
            const int li = q / EW, lw = q - (q / EW) * EW;
            if (skip_owned && li >= H && li < H + TR && lw >= HW && lw < HW + TW) continue;
            const int gi = r0 + li, gw = c0w + lw;
            const bool in = gi >= 0 && gi < N && gw >= 0 && gw < PW;
#pragma unroll
            for (int c = 0; c < NC; c++) L[c * AREAW + q] = in ? xload(src + c * CSW + (size_t)gi * PW + gw) : 0u;
        }
    };
    load_ext(a.buf, false);
    __syncthreads();
    // The halo reload between blocks in two halves: issue() starts the loads of the words
    // this thread reloads (into registers), land() writes them to LDS once they arrive, so
    // work placed between the two overlaps the load latency.
    constexpr int QL = (AREAW + NT - 1) / NT;
    uint32_t hv[ONE ? 1 : QL][NC];
    auto halo_issue = [&](int bufi) __attribute__((always_inline)) {
        const uint32_t* src = reinterpret_cast<const uint32_t*>((bufi ? a.cls[1] : a.cls[0]));
#pragma unroll
        for (int g = 0; g < QL; g++) {
            const int q = tid + g * NT;
            const int li = q / EW, lw = q - (q / EW) * EW;
            const int gi = r0 + li, gw = c0w + lw;
            const bool owned = li >= H && li < H + TR && lw >= HW && lw < HW + TW;
            const bool in = q < AREAW && !owned && gi >= 0 && gi < N && gw >= 0 && gw < PW;
#pragma unroll
            for (int c = 0; c < NC; c++) hv[g][c] = in ? xload(src + c * CSW + (size_t)gi * PW + gw) : 0u;
        }
    };
    auto halo_land = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int g = 0; g < QL; g++) {
            const int q = tid + g * NT;
            const int li = q / EW, lw = q - (q / EW) * EW;
            const bool owned = li >= H && li < H + TR && lw >= HW && lw < HW + TW;
            if (q < AREAW && !owned) {
#pragma unroll
                for (int c = 0; c < NC; c++) L[c * AREAW + q] = hv[g][c];
            }
        }
    };

    // a neighbour that is skipped (its cells never change) publishes nothing and is not
    // waited for; its cells keep their initial values in both copies
    auto skipped = [&](int rr, int cc) {
        return rr >= 0 && rr < tiles_r && cc >= 0 && cc < a.tiles_c && a.live_map != nullptr &&
               a.live_map[rr * a.tiles_c + cc] == 0;
    };
    int nbr_tile = -1;
    if (tid < 9 && tid != 4) {
        const int dr = tid / 3 - 1, dc = tid % 3 - 1;
        const int rr = tr + dr, cc = tc + dc;
        if (rr >= 0 && rr < tiles_r && cc >= 0 && cc < a.tiles_c && !skipped(rr, cc)) nbr_tile = rr * a.tiles_c + cc;
    }
    // Halo words inside a skipped neighbour start at zero (the host skips a tile only when
    // every cell within its halo and a margin is zero).  This tile's own computation of them
    // is exact one cell deep for a whole block, so the first change that would reach a
    // skipped tile shows here: they are checked after every block and a change raises the
    // error flag (value 2) instead of going unnoticed.
    uint32_t wmask = 0;
    if (SKIPCHK && a.live_map != nullptr) {
#pragma unroll
        for (int g = 0; g < QPT; g++) {
            const int w = tid + g * NT;
            if (w < AREAW) {
                const int li = w / EW, lw = w - (w / EW) * EW;
                const int dr = li < H ? -1 : (li >= H + TR ? 1 : 0);
                const int dc = lw < HW ? -1 : (lw >= HW + TW ? 1 : 0);
                if ((dr != 0 || dc != 0) && skipped(tr + dr, tc + dc)) wmask |= 1u << g;
            }
        }
    }

    // Transition intervals lo <= k < hi on k = t0 >> 9 (empty: lo == hi).  The lozenge rules'
    // two intervals are disjoint (checked on the host), so the state that toggles under a
    // draw is a function of the draw alone: 25 when k lies in A, 38 when it lies in B, and
    // 0xff (never a state) otherwise.  When A = [0, x) and B = [y, 2^23) are both nonempty
    // ("split": the q-rules) each test is one comparison of the raw output t0.
    const uint32_t alo = rules.fr.ab_lo[0][0], ahi = rules.fr.ab_hi[0][0];
    const uint32_t blo = rules.fr.ba_lo[0][0], bhi = rules.fr.ba_hi[0][0];
    const uint32_t asz = ahi - alo, bsz = bhi - blo;
    const bool split = !PM::kLozenge || (alo == 0u && ahi > 0u && ahi <= blo && blo < bhi && bhi == (1u << 23));
    const uint32_t aN = ahi << 9, bL = blo << 9;  // split: t0 < aN <=> k in A, t0 >= bL <=> k in B
    // per quad: byte masks of the bits each update rewrites (recomputed at each use when
    // per-generator parameters take the registers)
    uint32_t um2[UNI ? QPT : 1], um3[UNI ? QPT : 1];
#pragma unroll
    for (int g = 0; g < (UNI ? QPT : 1); g++) {
        um2[g] = ubytes[g] & 0x26262626u;
        um3[g] = ubytes[g] & 0x19191919u;
    }
    // Domino: the threshold t << 9 of the quad's row parity for class 0 and class 1 (the
    // weighted rules differ by class and parity; t0 < t << 9 <=> k < t, exact as 0 < t < 2^23)
    uint32_t dth0[PM::kLozenge ? 1 : QPT], dth1[PM::kLozenge ? 1 : QPT];
    if constexpr (!PM::kLozenge) {
#pragma unroll
        for (int g = 0; g < QPT; g++) {
            dth0[g] = rules.fr.ab_hi[0][rmod[g]] << 9;
            dth1[g] = rules.fr.ab_hi[1][rmod[g]] << 9;
        }
    }
    // dominokernel.cl's lazy rule (:11-17): below one threshold both 3 -> 12 and 12 -> 3, i.e.
    // a toggle; its two intervals both start at 0 (a split rule's 12 -> 3 interval starts at its
    // threshold t > 0).  The same for every thread of the workgroup.
    const bool toggle = !PM::kLozenge && rules.fr.ba_lo[0][0] == 0u;

    // The colours of a block (H <= 8 steps) are fetched as one 8-byte word before the
    // previous block's exchange, so no step waits on a global load.  The schedule buffer is
    // padded so the last (partial) block may read past the final step.
    static_assert(H <= 8, "a block's colours are fetched as one 8-byte word");
    auto fetch_block = [&](int64_t s_at) -> uint64_t {
        const uint2 v = gload(reinterpret_cast<const uint2*>(a.sched + s_at));
        return ((uint64_t)v.y << 32) | v.x;
    };

    // Flip decisions.  A draw depends only on its generator, and the state a draw toggles
    // only on the draw (the rule is the same for every class), so the targets of a whole
    // block can be drawn ahead of its flips.  targets() advances the generators of word g by
    // one draw each and returns the target states packed into one word (a pair position's
    // target in both its bytes): lozenge 25 when k lies in A, 38 when in B, 0xff (never a
    // state) otherwise; domino 3 below the threshold, 12 from it on.  SPLIT selects the
    // one-comparison interval tests.  Every slot draws (a slot past the class area holds zero states and its
    // draws are discarded), so no generator state flows through a divergent branch.
    auto targets = [&](int g, auto split_c, int c) __attribute__((always_inline)) -> uint32_t {
        constexpr bool SPLIT = decltype(split_c)::value;
        uint32_t x[QPW];
        const uint32_t dth = PM::kLozenge ? 0u : (c ? dth1[PM::kLozenge ? 0 : g] : dth0[PM::kLozenge ? 0 : g]);
#pragma unroll
        for (int b = 0; b < QPW; b++) {
            const uint32_t t0 = tm_draw_raw(s0[g][b], s1[g][b], s2[g][b], s3[g][b], UNI ? a.m1 : pm1[g][b],
                                            UNI ? a.m2 : pm2[g][b], UNI ? a.tm : pmt[g][b]);
            if constexpr (!PM::kLozenge) {
                // split rules: k < t: 3 -> 12; k >= t: 12 -> 3.  Toggle: 0x0f marks k < t.
                x[b] = (t0 < dth) ? (toggle ? 0x0fu : 3u) : (toggle ? 0u : 12u);
            } else if constexpr (SPLIT) {
                const uint32_t xb = (t0 >= bL) ? 38u : 0xffffffffu;
                x[b] = (t0 < aN) ? 25u : xb;
            } else {
                const uint32_t k = t0 >> 9;
                const uint32_t xb = (k - blo < bsz) ? 38u : 0xffffffffu;
                x[b] = (k - alo < asz) ? 25u : xb;
            }
        }
        if constexpr (QPW == 2) {
            return __builtin_amdgcn_perm(x[1], x[0], 0x04040000u);  // each target in both bytes
        } else {
            return __builtin_amdgcn_perm(__builtin_amdgcn_perm(x[QPW - 1], x[QPW - 2], 0x0c0c0400u),
                                         __builtin_amdgcn_perm(x[1], x[0], 0x0c0c0400u), 0x05040100u);
        }
    };
    // Toggle the bytes of quad g whose state equals its target (exact zero-byte test of
    // S ^ X; 25 ^ 38 = 63, 3 ^ 12 = 15); under the lazy toggle rule, the bytes marked 0x0f
    // whose state is 3 or 12.
    auto zero_bytes = [](uint32_t d) __attribute__((always_inline)) {
        const uint32_t t = (d & 0x7f7f7f7fu) + 0x7f7f7f7fu;
        return RT_BITOP3(t, d, 0x80808080u, ~(A | B) & C);  // 0x80 where the byte of d is 0
    };
    auto apply = [&](uint32_t* Ls, int g, uint32_t S, uint32_t X) __attribute__((always_inline)) {
        const int w = tid + g * NT;
        uint32_t z;
        if (!PM::kLozenge && toggle) z = (zero_bytes(S ^ 0x03030303u) | zero_bytes(S ^ 0x0c0c0c0cu)) & (X << 4);
        else z = zero_bytes(S ^ X);
        const uint32_t tog = ((z >> (PM::kLozenge ? 1 : 3)) - (z >> 7)) & gbytes[g];  // 63 / 15 there
        if (AREAW % NT == 0 || w < AREAW) Ls[w] = S ^ tog;
    };
    auto load_states = [&](const uint32_t* Ls, uint32_t* S) __attribute__((always_inline)) {
#pragma unroll
        for (int g = 0; g < QPT; g++) {
            const int w = tid + g * NT;
            S[g] = (AREAW % NT == 0 || w < AREAW) ? Ls[w] : 0u;
        }
    };
    // flip with draws made now
    auto flip = [&](uint32_t* Ls, auto split_c, int c) __attribute__((always_inline)) {
        uint32_t S[QPT];
        load_states(Ls, S);  // read first, so the LDS latency hides behind the draws
#pragma unroll
        for (int g = 0; g < QPT; g++) apply(Ls, g, S[g], targets(g, split_c, c));
    };
    // the targets of the first DA steps of the next full block, drawn during the exchange
    // (more would not fit the register budget; none with per-generator parameters, whose
    // registers leave no room).  Tiles of one word per thread on up to 512 threads (small domains,
    // one wave per SIMD, where every draw made inside a step lies on the step's critical path)
    // draw the whole block ahead: H more words of registers.
    constexpr int DA = (!UNI || ONE) ? 0 : (QPT == 1 && NT <= 512) ? H : H / 2;
    uint32_t XP[DA > 0 ? DA : 1][QPT];
    auto draw_ahead = [&](auto q0_c, auto q1_c, auto split_c, uint64_t cols) __attribute__((always_inline)) {
#pragma unroll
        for (int q = decltype(q0_c)::value; q < decltype(q1_c)::value; q++)
#pragma unroll
            for (int g = 0; g < QPT; g++) XP[q][g] = targets(g, split_c, (int)((cols >> (8 * q)) & 0xffu));
    };
    auto draw_ahead_any = [&](auto q0_c, auto q1_c, uint64_t cols) __attribute__((always_inline)) {
        if constexpr (DA > 0) {
            if constexpr (!PM::kLozenge) draw_ahead(q0_c, q1_c, std::true_type{}, cols);
            else if (split) draw_ahead(q0_c, q1_c, std::true_type{}, cols);
            else draw_ahead(q0_c, q1_c, std::false_type{}, cols);
        }
    };

    // Update classes c+1, c+2 from class c (lozengekernel.cl:45-52).  Every word the quad
    // needs is read first (one wait), the row-dependent offsets are funnel shifts chosen per
    // quad without branching, and the rewritten bits are merged with three-input selects
    // (a term's stray bits fall outside the bits the final select keeps).
    auto update = [&](uint32_t* Ls, uint32_t* L2, uint32_t* L3, int c) __attribute__((always_inline)) {
        if constexpr (PM::kLozenge) {
            const int t2 = (2 * c + 1) % 3, t4 = (2 * c + 2) % 3, t8 = (2 * c) % 3;
    #pragma unroll
            for (int g = 0; g < QPT; g++) {
                const int w = tid + g * NT;
                if (ubytes[g]) {
                    const int r = rmod[g];
                    const uint32_t U0 = Ls[w - EW], U1 = Ls[w - EW + 1];
                    const uint32_t Mm = Ls[w - 1], M0 = Ls[w], Mp = Ls[w + 1];
                    const uint32_t Dm = Ls[w + EW - 1], D0 = Ls[w + EW];
                    const uint32_t o2 = L2[w], o3 = L3[w];
                    const uint32_t up2 = funnel(U1, U0, (r == t2) ? 1 : 0);    // s[i-1][j+p2]
                    const uint32_t mm4 = (r == t4) ? funnel(M0, Mm, 3) : M0;   // s[i][j-p4]
                    const uint32_t mp8 = (r == t8) ? funnel(Mp, M0, 1) : M0;   // s[i][j+p8]
                    const uint32_t dm16 = (r == t2) ? funnel(D0, Dm, 3) : D0;  // s[i+1][j-p16]
                    // bits 2, 4, 32 of class c+1 and 1, 8, 16 of class c+2
                    const uint32_t n2 = RT_BITOP3(0x02020202u, up2 >> 3,
                                                  RT_BITOP3(0x04040404u, mm4 >> 1, D0 << 5, (A & B) | (~A & C)),
                                                  (A & B) | (~A & C));
                    const uint32_t n3 = RT_BITOP3(0x01010101u, U0 >> 5,
                                                  RT_BITOP3(0x08080808u, mp8 << 1, dm16 << 3, (A & B) | (~A & C)),
                                                  (A & B) | (~A & C));
                    const uint32_t m2 = UNI ? um2[UNI ? g : 0] : (ubytes[g] & 0x26262626u);
                    const uint32_t m3 = UNI ? um3[UNI ? g : 0] : (ubytes[g] & 0x19191919u);
                    L2[w] = RT_BITOP3(m2, n2, o2, (A & B) | (~A & C));
                    L3[w] = RT_BITOP3(m3, n3, o3, (A & B) | (~A & C));
                }
            }
        } else {
            // dominokernel.cl:22-31: class 1 - c (L2 here) from class c, every bit rewritten:
            // (up & 2) >> 1 | (down & 1) << 1 | (s[i][j - offL] & 8) >> 1 | (s[i][j + offR] & 4) << 1
            // with t = 1 - c, offL = (i + t + 1) & 1, offR = (i + t) & 1
#pragma unroll
            for (int g = 0; g < QPT; g++) {
                const int w = tid + g * NT;
                if (ubytes[g]) {
                    const int offR = (rmod[g] + 1 - c) & 1;
                    const uint32_t U0 = Ls[w - EW], D0 = Ls[w + EW];
                    const uint32_t Mm = Ls[w - 1], M0 = Ls[w], Mp = Ls[w + 1];
                    const uint32_t o = L2[w];
                    const uint32_t lf = offR ? M0 : funnel(M0, Mm, 4 - BPP);  // s[i][j-offL], offL = 1 - offR
                    const uint32_t rt = offR ? funnel(Mp, M0, BPP) : M0;    // s[i][j+offR]
                    const uint32_t x8 = (rt << 1) & 0x08080808u;
                    const uint32_t x4 = RT_BITOP3(0x04040404u, lf >> 1, x8, (A & B) | (~A & C));
                    const uint32_t x2 = RT_BITOP3(0x02020202u, D0 << 1, x4, (A & B) | (~A & C));
                    const uint32_t x1 = RT_BITOP3(0x01010101u, U0 >> 1, x2, (A & B) | (~A & C));
                    L2[w] = RT_BITOP3(ubytes[g], x1, o, (A & B) | (~A & C));
                }
            }
        }
    };

    // One chain step on colour class c.  A thread flips only its own quads, whose other
    // classes it updated itself in the previous step; only when the same class is rotated
    // twice in a row could a flip overwrite a neighbour that another thread's update is
    // still reading, so only then is a barrier needed before the flip.
    auto step = [&](int c, int c_prev, auto flip_c) __attribute__((always_inline)) {
        uint32_t* Ls = L + c * AREAW;
        // lozenge: classes c+1 and c+2; domino: L2 is the other class
        uint32_t* L2 = PM::kLozenge ? L + (c == 2 ? 0 : c + 1) * AREAW : L + (1 - c) * AREAW;
        uint32_t* L3 = PM::kLozenge ? L + (c == 0 ? 2 : c - 1) * AREAW : nullptr;
        if (c == c_prev) __syncthreads();
        flip_c(Ls);
        __syncthreads();
        update(Ls, L2, L3, c);
    };

    // A full block, steps unrolled (the generator states rotate through registers without
    // moves): the first DA steps use the targets drawn ahead, the others draw now.
    auto run_full = [&](uint64_t colours, int& c_prev, auto split_c) __attribute__((always_inline)) {
#pragma unroll
        for (int q = 0; q < H; q++) {
            const int c = (int)((colours >> (8 * q)) & 0xffu);
            if (q < DA) {
                step(c, c_prev, [&](uint32_t* Ls) __attribute__((always_inline)) {
                    uint32_t S[QPT];
                    load_states(Ls, S);
#pragma unroll
                    for (int g = 0; g < QPT; g++) apply(Ls, g, S[g], XP[q < DA ? q : 0][g]);
                });
            } else {
                step(c, c_prev, [&](uint32_t* Ls) __attribute__((always_inline)) { flip(Ls, split_c, c); });
            }
            c_prev = c;
        }
    };
    // a partial block (the last of a launch), drawing as it goes
    auto run_part = [&](uint64_t colours, int64_t blk, int& c_prev, auto split_c) __attribute__((always_inline)) {
        for (int64_t q = 0; q < blk; q++) {
            const int c = (int)((colours >> (8 * q)) & 0xffu);
            step(c, c_prev, [&](uint32_t* Ls) __attribute__((always_inline)) { flip(Ls, split_c, c); });
            c_prev = c;
        }
    };

    int buf = a.buf;
    uint32_t nblk = 0;
    int64_t s = 0;
    uint64_t colours = fetch_block(0);
    using Z = std::integral_constant<int, 0>;
    // the drawn-ahead steps are spread over the exchange: [0, D1) while the published words
    // drain, [D1, D3) while the neighbours' flags arrive, [D3, DA) while the halo loads
    using D1 = std::integral_constant<int, DA / 4>;
    using D3 = std::integral_constant<int, 3 * DA / 4>;
    using DF = std::integral_constant<int, DA>;
    if (a.nsteps >= H) draw_ahead_any(Z{}, DF{}, colours);
    while (s < a.nsteps) {
        const int64_t blk = (a.nsteps - s < H) ? (a.nsteps - s) : H;
        int c_prev = -1;  // the block starts after a barrier
        if (blk == H) {
            if (!PM::kLozenge || split) run_full(colours, c_prev, std::true_type{});
            else if constexpr (PM::kLozenge) run_full(colours, c_prev, std::false_type{});
        } else {
            if (!PM::kLozenge || split) run_part(colours, blk, c_prev, std::true_type{});
            else if constexpr (PM::kLozenge) run_part(colours, blk, c_prev, std::false_type{});
        }
        s += blk;
        const bool last = s >= a.nsteps;
        // the next block's colours are in flight during the exchange
        const uint64_t next_colours = last ? 0 : fetch_block(s);
        __syncthreads();  // every update of the block is in LDS before it is published
        if (wmask != 0) {
            uint32_t any = 0;
#pragma unroll
            for (int g = 0; g < QPT; g++) {
                const int w = tid + g * NT;
                if (wmask & (1u << g)) {
#pragma unroll
                    for (int c = 0; c < NC; c++) any |= L[c * AREAW + w];
                }
            }
            if (any != 0) __hip_atomic_store(a.err, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        // Publish into the other copy: after an intermediate block only the band of width H
        // along the tile's edges (the neighbours' halos read nothing else), after the last
        // block every owned word (the launch's result).
        {
            uint32_t* dst = reinterpret_cast<uint32_t*>((buf ? a.cls[0] : a.cls[1]));
            for (int q = tid; q < TR * TW; q += NT) {
                const int oi = q / TW, ow = q - (q / TW) * TW;
                const bool band = oi < H || oi >= TR - H || ow < HW || ow >= TW - HW;
                const int gi = tr * TR + oi, gw = tc * TW + ow;
                if ((last || band) && gi < N && gw < PW) {
                    const int lq = (oi + H) * EW + HW + ow;
#pragma unroll
                    for (int c = 0; c < NC; c++) xstore(dst + c * CSW + (size_t)gi * PW + gw, L[c * AREAW + lq]);
                }
            }
        }
        buf ^= 1;
        if (!ONE && !last) {
            nblk++;
            const uint32_t epoch = a.epoch0 + nblk;
            // the next full block's targets are drawn while the published words drain and
            // while the neighbours finish their blocks
            const bool full_next = a.nsteps - s >= H;
            if (full_next) draw_ahead_any(Z{}, D1{}, next_colours);
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            if (tid == 0) xstore(a.flags + (size_t)tile * kLine, epoch);
            if (full_next) draw_ahead_any(D1{}, D3{}, next_colours);
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
            if constexpr (DA > 0) {
                halo_issue(buf);
                if (full_next) draw_ahead_any(D3{}, DF{}, next_colours);
                halo_land();
            } else {
                load_ext(buf, true);
            }
            __syncthreads();
        }
        colours = next_colours;
    }

#pragma unroll
    for (int g = 0; g < QPT; g++) {
        if (!(omask & (1u << g))) continue;
        const int w = tid + g * NT;
        const int li = w / EW, lw = w - (w / EW) * EW;
#pragma unroll
        for (int b = 0; b < QPW; b++) {
            if (!(gbytes[g] & posmask(b))) continue;
            const int id = M::gen_id(r0 + li, c0 + QPW * lw + b, N, W);
            gstore(a.gen_out + id, make_uint4(s0[g][b], s1[g][b], s2[g][b], s3[g][b]));
        }
    }
}

template <int TR, int TC, int H, int NT, bool UNI>
__global__ __launch_bounds__(NT) void k_persist_loz4(PersistArgs a) {
    const int b = xcd_tile(blockIdx.x, gridDim.x);
    const bool skip = a.live != nullptr;
    packed_body<PackedLozenge, TR, TC, H, NT, UNI>(a, a.rp, skip ? a.live[b] : b,
                                                   skip ? a.ntiles : (int)gridDim.x);
}

template <int TR, int TC, int H, int NT, bool UNI>
__global__ __launch_bounds__(NT) void k_persist_loz4_batch(PersistArgs a_in) {
    PersistArgs a = a_in;
    int tile, ntiles;
    const RuleParams& rules = resolve_batch(a, tile, ntiles);
    packed_body<PackedLozenge, TR, TC, H, NT, UNI, true, false>(a, rules, tile, ntiles);
}

template <int TR, int TC, int H, int NT, bool UNI>
__global__ __launch_bounds__(NT) void k_persist_dom4(PersistArgs a) {
    packed_body<PackedDomino, TR, TC, H, NT, UNI>(a, a.rp, xcd_tile(blockIdx.x, gridDim.x), gridDim.x);
}

// One block per launch (a tile grid that is not resident): the same sweep without the exchange.
template <int TR, int TC, int H, int NT, bool UNI>
__global__ __launch_bounds__(NT) void k_persist_dom4_blk(PersistArgs a) {
    packed_body<PackedDomino, TR, TC, H, NT, UNI, true>(a, a.rp, xcd_tile(blockIdx.x, gridDim.x), gridDim.x);
}

template <int TR, int TC, int H, int NT, bool UNI>
__global__ __launch_bounds__(NT) void k_persist_dom4_batch(PersistArgs a_in) {
    PersistArgs a = a_in;
    int tile, ntiles;
    const RuleParams& rules = resolve_batch(a, tile, ntiles);
    packed_body<PackedDomino, TR, TC, H, NT, UNI, true, false>(a, rules, tile, ntiles);
}

template <int TR, int TC, int H, int NT, bool UNI>
__global__ __launch_bounds__(NT) void k_persist_pair2(PersistArgs a) {
    packed_body<PackedDominoPair, TR, TC, H, NT, UNI>(a, a.rp, xcd_tile(blockIdx.x, gridDim.x), gridDim.x);
}

template <int TR, int TC, int H, int NT, bool UNI>
__global__ __launch_bounds__(NT) void k_persist_pair2_blk(PersistArgs a) {
    packed_body<PackedDominoPair, TR, TC, H, NT, UNI, true>(a, a.rp, xcd_tile(blockIdx.x, gridDim.x), gridDim.x);
}

}  // namespace rt
