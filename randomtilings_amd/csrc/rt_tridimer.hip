// rt_tridimer.hip — the triangle-dimer chain (TriangleDimerTiler::Walk,
// TriangleDimerTiler.cpp:86-483, kernels triangledimerkernel.cl) on gfx950.
//
// State: per site (a, b) of the M x M move lattice (M = N/2) eight small integers — the
// lozenge-move states eH, eL, eR, the triangle-move states tU, tD and the butterfly-move
// states bH, bL, bR — packed into one 64-bit cell, so a site moves between HBM and LDS as one
// 8-byte access.  With 4-bit lozenge states in the input tiling (checked on the host) every
// state stays inside its field: lozenge states are sums of four distinct bits (<= 15),
// triangle states three base-4 digits of at most 6 (<= 126, 7 bits), butterfly states four
// digits of at most 6 plus 256 c (<= 766, 10 bits); the fields reproduce the reference's int
// arithmetic exactly.
//
// A reference step is one Rotate kernel plus three to nine update kernels, each reading the
// previous one's output at distance <= 1.  The Rotate kernel's random number only decides
// whether a site rotates (rd < 0.5), and a site's generator advances once per step whatever
// the tiling, so the draws are made apart from the tiling: a draw launch advances every
// generator kDrawSteps times and keeps each outcome as one bit per site and step.  A step is
// then one kernel: every workgroup loads its tile plus a halo of cells and rotation bits into
// LDS, applies the rotation, and replays the update kernels as LDS phases separated by
// workgroup barriers.  Phase p only computes sites at least p away from a halo edge that cuts
// the domain, which is exactly the region whose inputs are still valid.  The halo is as wide as
// the step's chain of phases: 4 for lozenge and triangle moves, kHalo = 6 for butterfly moves
// (the move type of a step is known when its block starts).
//
// Domains up to 64 x 64 sites run every step in one launch of one workgroup (the whole
// domain in LDS, generator states in VGPRs), with only workgroup barriers between phases.
#include "rt_internal.hpp"
#include "rt_tdsite.hpp"
#include "rt_tridimer.hpp"

namespace rt {
namespace td {

// ---- phase driver ----------------------------------------------------------------------

// x / d by a multiply-high with m = floor(2^32 / d) + 1: exact for x < 2^17 and d <= 1024
// (checked exhaustively), which covers every site index of a region (at most 88 x 88 sites).
// A hardware division would cost dozens of instructions per site and phase.  The divisors are
// region and band widths, the same for the whole workgroup and at most kRecipMax, so m comes
// from a table in constant memory (one scalar load) instead of a 64-bit division per thread.
constexpr int kRecipMax = 128;
struct Recips {
    uint32_t m[kRecipMax + 1];
};
constexpr Recips make_recips() {
    Recips r{};
    for (int d = 2; d <= kRecipMax; d++) r.m[d] = (uint32_t)(0x100000000ull / (unsigned)d + 1);
    return r;
}
__constant__ Recips kRecips = make_recips();

struct FastDiv {
    uint32_t d, m;
    // d_: the same in every lane
    __device__ __forceinline__ explicit FastDiv(uint32_t d_) : d(d_) {
        const uint32_t du = __builtin_amdgcn_readfirstlane(d_);
        m = du <= (uint32_t)kRecipMax ? kRecips.m[du] : (uint32_t)(0x100000000ull / du + 1);
    }
    __device__ __forceinline__ uint32_t div(uint32_t x) const { return d > 1 ? __umulhi(x, m) : x; }
};

// Sites of a region a phase may compute: at least `m` away from every region edge that cuts
// the domain, and inside the reference's update range 0 < gi, gj < M - 1.
struct Region {
    int gi0, gj0, rows, cols, M;
    bool cutT, cutB, cutL, cutR;
};

// Each thread takes one column of a band of NT / w rows and steps down by the band, so the
// site index is divided once per phase, not once per site; the update range is folded into
// the loop limits.  (A region wider than NT, which no tile configuration has, falls back to a
// flat index.)
template <int NT, class F>
__device__ __forceinline__ void for_sites(const Region& R, int m, F f) {
    const int i0 = max(R.cutT ? m : 0, 1 - R.gi0), i1 = min(R.rows - (R.cutB ? m : 0), R.M - 1 - R.gi0);
    const int j0 = max(R.cutL ? m : 0, 1 - R.gj0), j1 = min(R.cols - (R.cutR ? m : 0), R.M - 1 - R.gj0);
    const int w = j1 - j0;
    if (w <= 0 || i1 <= i0) return;
    const FastDiv fw((uint32_t)w);
    if (w <= NT) {
        const int band = (int)fw.div(NT);
        const int r = (int)fw.div(threadIdx.x);
        if (r >= band) return;
        const int j = j0 + (int)threadIdx.x - r * w, gj = R.gj0 + j;
        for (int i = i0 + r; i < i1; i += band) f(i, j, R.gi0 + i, gj);
        return;
    }
    const int n = (i1 - i0) * w;
    for (int idx = threadIdx.x; idx < n; idx += NT) {
        const int qi = (int)fw.div((uint32_t)idx);
        const int i = i0 + qi, j = j0 + idx - qi * w;
        f(i, j, R.gi0 + i, R.gj0 + j);
    }
}

// Like for_sites, over the sites with (a * gi + b * gj) mod q == want only (a, b in {0, 1} and
// not both 0, q in {2, 3}, 0 <= want < q): the colour classes the Rotate-adjacent phases touch.
// Rows are walked with the residue's first column and a stride of q, so threads are not spent
// on sites the phase skips; the first column of the residue moves by a fixed amount from one
// row of the thread to its next, so no remainder is taken inside the loop.
template <int NT, class F>
__device__ __forceinline__ void for_sites_mod(const Region& R, int m, int a, int b, int q, int want, F f) {
    const int i0 = max(R.cutT ? m : 0, 1 - R.gi0), i1 = min(R.rows - (R.cutB ? m : 0), R.M - 1 - R.gi0);
    const int j0 = max(R.cutL ? m : 0, 1 - R.gj0), j1 = min(R.cols - (R.cutR ? m : 0), R.M - 1 - R.gj0);
    const int w = j1 - j0;
    if (w <= 0 || i1 <= i0) return;
    if (b == 0) {  // a == 1: a condition on rows only, every column of the rows gi = want mod q
        const FastDiv fw((uint32_t)w);
        const int band = (int)fw.div(NT);
        const int r = (int)fw.div(threadIdx.x);
        if (r >= band) return;
        const int j = j0 + (int)threadIdx.x - r * w, gj = R.gj0 + j;
        const int first = i0 + modq(want - (R.gi0 + i0), q);
        for (int i = first + r * q; i < i1; i += band * q) f(i, j, R.gi0 + i, gj);
        return;
    }
    const int wc = q == 2 ? (w + 1) >> 1 : (int)((uint32_t)(w + 2) / 3u);  // columns of one residue in a row, at most
    const FastDiv fwc((uint32_t)wc);
    const int band = (int)fwc.div(NT);
    const int r = (int)fwc.div(threadIdx.x);
    if (r >= band) return;
    const int k = (int)threadIdx.x - r * wc;
    // off: first column j >= j0 with (a gi + gj) mod q == want, less j0; the next row of this
    // thread (gi + band) moves it by -a band mod q
    int off = modq(want - a * (R.gi0 + i0 + r) - (R.gj0 + j0), q);
    const int doff = modq(-a * band, q);
    for (int i = i0 + r; i < i1; i += band) {
        const int j = j0 + off + q * k;
        if (j < j1) f(i, j, R.gi0 + i, R.gj0 + j);
        off += doff;
        off -= off >= q ? q : 0;
    }
}

// Like for_sites, over pairs of sites (words of the pair view): each thread takes one pair
// column of a band of rows.  A pair may reach one site past [j0, j1): a site of the margin
// that no later phase reads (its inputs were already out of date), or a site outside the
// domain's update range 0 < gj < M - 1, which the write mask wm excludes (its value is kept).
template <int NT, class F>
__device__ __forceinline__ void for_pairs(const Region& R, int m, F f) {
    const int i0 = max(R.cutT ? m : 0, 1 - R.gi0), i1 = min(R.rows - (R.cutB ? m : 0), R.M - 1 - R.gi0);
    const int j0 = max(R.cutL ? m : 0, 1 - R.gj0), j1 = min(R.cols - (R.cutR ? m : 0), R.M - 1 - R.gj0);
    const int q0 = j0 >> 1, w = ((j1 + 1) >> 1) - q0;  // pairs covering [j0, j1)
    if (w <= 0 || i1 <= i0) return;
    const FastDiv fw((uint32_t)w);
    const int band = (int)fw.div(NT);
    const int r = (int)fw.div(threadIdx.x);
    if (r >= band) return;
    const int jp = q0 + (int)threadIdx.x - r * w, gj = R.gj0 + 2 * jp;
    const uint32_t wm = (gj >= 1 && gj <= R.M - 2 ? 0x0000ffffu : 0u) | (gj + 1 <= R.M - 2 ? 0xffff0000u : 0u);
    for (int i = i0 + r; i < i1; i += band) f(i, jp, R.gi0 + i, gj, wm);
}

// The update kernels of one step after its Rotate (TriangleDimerTiler.cpp:217-338).  m0: the
// margin earlier steps of the same launch have already used up (their phases' inputs ran out
// there), added to every phase's own.  The phases over every site (or two thirds of them) run on
// pairs of sites; the ones over a half or a third of the sites run per site on those sites only.
template <int NT>
__device__ void replay_updates(const View& V, const Region& R, const Move& mv, const int m0 = 0) {
    const PView P = V.pairs();
    if (mv.type == 0) {
        const int t = mv.r1, c = mv.r2;
        // UpdateLozengesFlipped rewrites the class ((1 - (t & 1)) gi + ((1 + t) / 2) gj) mod 2 != c
        for_sites_mod<NT>(R, m0 + 1, 1 - (t & 1), (1 + t) / 2, 2, 1 - c,
                          [&](int i, int j, int gi, int gj) { loz_flipped(V, i, j, gi, gj, t, c); });
        __syncthreads();
        for_pairs<NT>(R, m0 + 2, [&](int i, int jp, int, int, uint32_t wm) { loz_update_p(P, i, jp, t, wm); });
        __syncthreads();
        for_pairs<NT>(R, m0 + 3, [&](int i, int jp, int, int, uint32_t wm) {
            put(P, TU, i, jp, tri_u_p(P, i, jp), wm);
            put(P, BH, i, jp, bfly_from_loz_p(P, i, jp, 0), wm);
            put(P, BL, i, jp, bfly_from_loz_p(P, i, jp, 1), wm);
            put(P, BR, i, jp, bfly_from_loz_p(P, i, jp, 2), wm);
        });
        __syncthreads();
        for_pairs<NT>(R, m0 + 4, [&](int i, int jp, int, int, uint32_t wm) { put(P, TD, i, jp, tri_update_p(P, i, jp, 0), wm); });
    } else if (mv.type == 1) {
        const int c = mv.r1;
        const bool down = mv.r2 == 1;
        for_pairs<NT>(R, m0 + 1, [&](int i, int jp, int gi, int gj, uint32_t wm) { tri_flipped_p(P, i, jp, gi, gj, c, down, wm); });
        __syncthreads();
        for_pairs<NT>(R, m0 + 2, [&](int i, int jp, int, int, uint32_t wm) {
            put(P, down ? TU : TD, i, jp, tri_update_p(P, i, jp, down ? 1 : 0), wm);
        });
        __syncthreads();
        for_pairs<NT>(R, m0 + 3, [&](int i, int jp, int, int, uint32_t wm) { loz_from_tri_p(P, i, jp, wm); });
        __syncthreads();
        for_pairs<NT>(R, m0 + 4, [&](int i, int jp, int, int, uint32_t wm) {
            put(P, BH, i, jp, bfly_from_loz_p(P, i, jp, 0), wm);
            put(P, BL, i, jp, bfly_from_loz_p(P, i, jp, 1), wm);
            put(P, BR, i, jp, bfly_from_loz_p(P, i, jp, 2), wm);
        });
    } else {
        const int t = mv.r1, p1 = mv.p1, p2 = mv.p2;
        // the slice classes bslice(t) = (b1 gi + b2 gj) mod 3 these phases rewrite
        const int sa = t == 1 ? 0 : 1, sb = t == 0 ? 0 : 1;
        for_sites_mod<NT>(R, m0 + 1, sa, sb, 3, p1,
                          [&](int i, int j, int gi, int gj) { bfly_flipped1(V, i, j, gi, gj, t, p2); });
        __syncthreads();
        for_sites_mod<NT>(R, m0 + 2, sa, sb, 3, t == 0 ? (p1 + 1) % 3 : (p1 + 2) % 3,
                          [&](int i, int j, int, int) { bfly_flipped21(V, i, j, t); });
        __syncthreads();
        for_sites_mod<NT>(R, m0 + 3, sa, sb, 3, t == 0 ? (p1 + 2) % 3 : (p1 + 1) % 3,
                          [&](int i, int j, int, int) { bfly_flipped22(V, i, j, t); });
        __syncthreads();
        // UpdateButterflysFlipped23 on its slice, then every site's lozenge states from its own
        // (just final) butterfly state
        const int ts = t == 0 ? (p1 + 1) % 3 : (p1 + 2) % 3;
        for_pairs<NT>(R, m0 + 4, [&](int i, int jp, int gi, int gj, uint32_t wm) {
            const uint32_t sm = (bslice(t, gi, gj) == ts ? 0x0000ffffu : 0u) | (bslice(t, gi, gj + 1) == ts ? 0xffff0000u : 0u);
            bfly23_loz_from_bfly_p(P, i, jp, t, sm, wm);
        });
        __syncthreads();
        for_pairs<NT>(R, m0 + 5, [&](int i, int jp, int, int, uint32_t wm) {
            if (t != 0) put(P, BH, i, jp, bfly_from_loz_p(P, i, jp, 0), wm);
            if (t != 1) put(P, BL, i, jp, bfly_from_loz_p(P, i, jp, 1), wm);
            if (t != 2) put(P, BR, i, jp, bfly_from_loz_p(P, i, jp, 2), wm);
            put(P, TU, i, jp, tri_u_p(P, i, jp), wm);
        });
        __syncthreads();
        for_pairs<NT>(R, m0 + 6, [&](int i, int jp, int, int, uint32_t wm) { put(P, TD, i, jp, tri_update_p(P, i, jp, 0), wm); });
    }
}

// 64-bit cell: x = eH | eL << 4 | eR << 8 | tU << 12 | tD << 19, y = bH | bL << 10 | bR << 20.
__constant__ int kCellShift[8] = {0, 4, 8, 12, 19, 0, 10, 20};  // per plane, in x (EH..TD) or y (BH..BR)
__device__ __forceinline__ void unpack_cell(const uint2 c, int v[8]) {
    v[EH] = c.x & 15u, v[EL] = (c.x >> 4) & 15u, v[ER] = (c.x >> 8) & 15u;
    v[TU] = (c.x >> 12) & 127u, v[TD] = (c.x >> 19) & 127u;
    v[BH] = c.y & 1023u, v[BL] = (c.y >> 10) & 1023u, v[BR] = (c.y >> 20) & 1023u;
}
__device__ __forceinline__ uint2 pack_cell(const View& V, int i, int j) {
    uint2 c;
    c.x = V.g(EH, i, j) | V.g(EL, i, j) << 4 | V.g(ER, i, j) << 8 | V.g(TU, i, j) << 12 | V.g(TD, i, j) << 19;
    c.y = V.g(BH, i, j) | V.g(BL, i, j) << 10 | V.g(BR, i, j) << 20;
    return c;
}
// The cells of a pair of sites from its eight plane words (field f = 0: the low site).
__device__ __forceinline__ uint2 pack_cell_pair(const uint32_t w[8], int f) {
    const int sh = 16 * f;
    auto g = [&](int p) { return (w[p] >> sh) & 0xffffu; };
    return make_uint2(g(EH) | g(EL) << 4 | g(ER) << 8 | g(TU) << 12 | g(TD) << 19, g(BH) | g(BL) << 10 | g(BR) << 20);
}

// Write the owned tile [oi, oi1) x [oj, oj1) from the planes, a pair of sites per thread and
// iteration (the region starts at an even column, so every owned pair is one word).
template <int NT>
__device__ __forceinline__ void store_owned(const TdArgs& a, const PView& P, const Region& R, int oi, int oj, int oi1,
                                            int oj1) {
    const int ow = oj1 - oj, opc = (ow + 1) >> 1, on = (oi1 - oi) * opc;
    const int jo = (oj - R.gj0) >> 1;
    const FastDiv fo((uint32_t)opc);
    for (int idx = threadIdx.x; idx < on; idx += NT) {
        const int qi = (int)fo.div((uint32_t)idx), k = idx - qi * opc;
        const int gi = oi + qi, gj = oj + 2 * k;
        uint32_t w[8];
#pragma unroll
        for (int p = 0; p < 8; p++) w[p] = P.w(p, gi - R.gi0, jo + k);
        uint2* dst = a.cells_out + (size_t)gi * a.M + gj;
        gstore(dst, pack_cell_pair(w, 0));
        if (2 * k + 1 < ow) gstore(dst + 1, pack_cell_pair(w, 1));
    }
}

template <bool UNI>
__device__ __forceinline__ void load_params(const TdArgs& a, size_t id, uint32_t& m1, uint32_t& m2, uint32_t& tm) {
    if (UNI) {
        m1 = a.m1, m2 = a.m2, tm = a.tm;
    } else {
        const uint4 p = a.params[id];
        m1 = p.x, m2 = p.y, tm = p.z;
    }
}

// ---- rotation bits ----------------------------------------------------------------------
// One wave per 64 sites of a row: each lane advances its site's generator k times, and the
// wave's ballot of each draw's outcome (tempered output < 2^31, the Rotate kernels' rd < 0.5)
// is the row's word of that step's plane, stored by lane s for step s.
template <bool UNI>
__device__ __forceinline__ void td_draw_row(const TdArgs& a, const uint4* __restrict__ gin, uint4* __restrict__ gout,
                                            int M, int bx, int gi, int k, uint64_t* __restrict__ bits) {
    const int lane = (int)threadIdx.x & 63;
    const int gj = bx * 64 + lane;
    const bool live = gj < M;
    const size_t id = (size_t)gi * M + gj;
    uint4 g = make_uint4(0u, 0u, 0u, 0u);
    uint32_t m1 = 0, m2 = 0, tm = 0;
    if (live) {
        g = gload(gin + id);
        load_params<UNI>(a, id, m1, m2, tm);
    }
    uint64_t mine = 0;
    for (int s = 0; s < k; s++) {
        bool low = false;
        if (live) low = tm_draw_raw(g.x, g.y, g.z, g.w, m1, m2, tm) < 0x80000000u;
        const uint64_t w = __ballot(low);
        if (lane == s) mine = w;
    }
    if (live) gstore(gout + id, g);
    if (lane < k) gstore(bits + ((size_t)lane * M + gi) * td_bit_words(M) + bx, mine);
}

// Blocks of four waves: blockIdx.x = the word of the row, blockIdx.y = four rows.
template <bool UNI>
__global__ __launch_bounds__(256) void k_td_draw(TdArgs a, int k) {
    const int gi = (int)blockIdx.y * 4 + ((int)threadIdx.x >> 6);
    if (gi >= a.M) return;  // whole waves
    td_draw_row<UNI>(a, a.gen_in, a.gen_out, a.M, (int)blockIdx.x, gi, k, a.rbits);
}

// The chain of a batched block: the entry whose range [first(e), ...) holds block b.
template <class F>
__device__ __forceinline__ int td_entry(const TdBatchEntry* __restrict__ batch, int nbatch, int b, F first) {
    int lo = 0, hi = nbatch - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first(batch[mid]) <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Batched draws of the k steps from s = *step_ctr + off (s a multiple of kDrawSteps from the
// run's start): generators flip copies once per draw launch.
template <bool UNI>
__global__ __launch_bounds__(256) void k_td_draw_batch(TdArgs a, const TdBatchEntry* __restrict__ batch, int nbatch,
                                                       const int64_t* __restrict__ step_ctr, int64_t off, int k) {
    const int64_t s = *step_ctr + off;
    const int b = (int)blockIdx.x;
    const TdBatchEntry& e = batch[td_entry(batch, nbatch, b, [](const TdBatchEntry& x) { return x.dblk0; })];
    const int rw = td_bit_words(e.M), r = b - e.dblk0;
    const int gi = (r / rw) * 4 + ((int)threadIdx.x >> 6);
    if (gi >= e.M) return;  // whole waves
    const int par = (int)((s / kDrawSteps) & 1);
    td_draw_row<UNI>(a, e.gens[e.g0 ^ par], e.gens[e.g0 ^ par ^ 1], e.M, r % rw, gi, k, e.rbits);
}

// The 32-bit half of the rotation-bit word holding site (gi, gj) in plane p (bit gj & 31).
__device__ __forceinline__ const uint32_t* bit_word(const TdArgs& a, int p, int gi, int gj) {
    return reinterpret_cast<const uint32_t*>(a.rbits + ((size_t)p * a.M + gi) * td_bit_words(a.M) + (gj >> 6)) +
           ((gj >> 5) & 1);
}

// ---- one step per launch, tiles + halo --------------------------------------------------
template <int TR, int TC, int NT>
__device__ __forceinline__ void td_step_body(const TdArgs& a, const uint32_t move, const int tile) {
    constexpr int RR = TR + 2 * kHalo, RC = TC + 2 * kHalo;
    static_assert(RC % 2 == 0, "the pair view needs an even pitch");
    // 8 planes of 16-bit sites, read as pairs too (4-byte aligned), a guard of two words each side
    __shared__ uint32_t lds[8 * RR * RC / 2 + 4];
    const int M = a.M;
    const int tiles_c = (M + TC - 1) / TC;
    const int ti = tile / tiles_c, tj = tile % tiles_c;
    const int oi = ti * TR, oj = tj * TC;  // owned tile origin
    // the region: the owned tile and a halo as wide as this step's move reaches (kHalo at most,
    // what the LDS holds)
    const Move mv = decode(move);
    const int hw = move_reach(mv);
    Region R;
    R.M = M;
    R.gi0 = max(0, oi - hw);
    R.gj0 = max(0, oj - hw);
    const int gi1 = min(M, oi + TR + hw), gj1 = min(M, oj + TC + hw);
    R.rows = gi1 - R.gi0;
    R.cols = gj1 - R.gj0;
    R.cutT = R.gi0 > 0;
    R.cutL = R.gj0 > 0;
    R.cutB = gi1 < M;
    R.cutR = gj1 < M;
    const View V{reinterpret_cast<u16a*>(lds + 2), RC, RR * RC};
    const int rp = rotate_plane(mv);
    const int oi1 = min(M, oi + TR), oj1 = min(M, oj + TC);
    // load the region's cells and rotation bits a pair of sites at a time (the region starts at
    // an even column: tile origins and halo widths are even), rotate, store the pair's plane
    // words.  Every load of the thread's pairs is issued before the first is used, so the loads
    // overlap instead of waiting one by one.
    const PView P = V.pairs();
    constexpr int PPT = (RR * (RC / 2) + NT - 1) / NT;  // pairs per thread
    const int pc = (R.cols + 1) >> 1, np = R.rows * pc;
    const FastDiv fp((uint32_t)pc);
    uint2 c0[PPT], c1[PPT];
    uint32_t bw[PPT];
#pragma unroll
    for (int q = 0; q < PPT; q++) {
        const int idx = threadIdx.x + q * NT;
        if (idx < np) {
            const int qi = (int)fp.div((uint32_t)idx), jp = idx - qi * pc;
            const int gi = R.gi0 + qi, gj = R.gj0 + 2 * jp;
            const uint2* src = a.cells_in + (size_t)gi * M + gj;
            c0[q] = gload(src);
            c1[q] = make_uint2(0u, 0u);
            if (2 * jp + 1 < R.cols) c1[q] = gload(src + 1);
            bw[q] = gload(bit_word(a, a.plane, gi, gj));  // gj even: gj + 1 is in the same half
        }
    }
    // the rotated plane's field in the cell (the same in every lane): word, shift and width
    const bool rhi = rp >= BH;
    const int rsh = kCellShift[rp];
    const uint32_t rmask = rhi ? 1023u : rp >= TU ? 127u : 15u;
    auto rotate_cell = [&](uint2& c, int gi, int gj, uint32_t low) __attribute__((always_inline)) {
        uint32_t& w = rhi ? c.y : c.x;
        const int v = (int)((w >> rsh) & rmask);
        w ^= (uint32_t)rotate_xor(mv, v, gi, gj, low) << rsh;
    };
#pragma unroll
    for (int q = 0; q < PPT; q++) {
        const int idx = threadIdx.x + q * NT;
        if (idx < np) {
            const int qi = (int)fp.div((uint32_t)idx), jp = idx - qi * pc;
            const int gi = R.gi0 + qi, gj = R.gj0 + 2 * jp;
            const uint32_t b = bw[q] >> (gj & 31);
            rotate_cell(c0[q], gi, gj, b & 1u);
            rotate_cell(c1[q], gi, gj + 1, (b >> 1) & 1u);
            int v0[8], v1[8];
            unpack_cell(c0[q], v0);
            unpack_cell(c1[q], v1);
#pragma unroll
            for (int p = 0; p < 8; p++) P.set(p, qi, jp, (uint32_t)v0[p] | ((uint32_t)v1[p] << 16));
        }
    }
    __syncthreads();
    replay_updates<NT>(V, R, mv);
    __syncthreads();
    store_owned<NT>(a, P, R, oi, oj, oi1, oj1);
}

template <int TR, int TC, int NT>
__global__ __launch_bounds__(NT) void k_td_step(TdArgs a, uint32_t move) {
    td_step_body<TR, TC, NT>(a, move, xcd_tile(blockIdx.x, gridDim.x));
}

// Batched: every block runs one tile of one chain; the step index comes from a device
// counter (one captured launch replays as is), and with it the chain's moves, rotation-bit
// plane and copies: the copies flip once per launch, and the launches so far are s / spl
// (spl: steps per launch of the run's earlier launches).
__device__ __forceinline__ const TdBatchEntry& td_resolve(TdArgs& a, const TdBatchEntry* __restrict__ batch,
                                                         int nbatch, int64_t s, int spl, int& tile) {
    const int b = blockIdx.x;
    const TdBatchEntry& e = batch[td_entry(batch, nbatch, b, [](const TdBatchEntry& x) { return x.tile0; })];
    const int par = (int)((s / spl) & 1);
    a.M = uniform(e.M);
    a.cells_in = uniform(e.cells[e.buf0 ^ par]);
    a.cells_out = uniform(e.cells[e.buf0 ^ par ^ 1]);
    a.rbits = uniform(e.rbits);
    a.plane = (int)(s % kDrawSteps);
    tile = uniform(b - e.tile0);
    return e;
}

template <int TR, int TC, int NT>
__global__ __launch_bounds__(NT) void k_td_step_batch(TdArgs a0, const TdBatchEntry* __restrict__ batch, int nbatch,
                                                      const int64_t* __restrict__ step_ctr, int64_t off, int spl) {
    const int64_t s = *step_ctr + off;
    TdArgs a = a0;
    int tile;
    const TdBatchEntry& e = td_resolve(a, batch, nbatch, s, spl, tile);
    td_step_body<TR, TC, NT>(a, (uint32_t)uniform((int)e.sched[s]), tile);
}

// ---- S steps per launch, tiles + a halo of up to S * kHalo --------------------------------
// Every step's phases use up its move's reach (4, or kHalo for butterfly moves) of margin at
// each cut edge, so with a halo of the summed reaches the owned tile is exact after S steps;
// the sites in between hold values no owned site reads.  The tile's cells cross HBM once per S
// steps (at S = 2 and 64 x 64 tiles 0.95 region loads per owned site and step instead of 1.41).
template <int TR, int TC, int NT, int S>
__device__ __forceinline__ void td_multi_body(const TdArgs& a, const uint32_t moves, const int tile) {
    constexpr int HMAX = S * kHalo;
    constexpr int RR = TR + 2 * HMAX, RC = TC + 2 * HMAX;
    static_assert(RC % 2 == 0, "the pair view needs an even pitch");
    // 8 planes of 16-bit sites, read as pairs too (4-byte aligned), a guard of two words each side
    __shared__ uint32_t lds[8 * RR * RC / 2 + 4];
    const int M = a.M;
    const int tiles_c = (M + TC - 1) / TC;
    const int ti = tile / tiles_c, tj = tile % tiles_c;
    const int oi = ti * TR, oj = tj * TC;  // owned tile origin
    // the halo: the sum of the S steps' reaches (each step's phases use up its own)
    int HH = 0;
#pragma unroll
    for (int s = 0; s < S; s++) HH += move_reach(decode((moves >> (8 * s)) & 0xffu));
    Region R;
    R.M = M;
    R.gi0 = max(0, oi - HH);
    R.gj0 = max(0, oj - HH);
    const int gi1 = min(M, oi + TR + HH), gj1 = min(M, oj + TC + HH);
    R.rows = gi1 - R.gi0;
    R.cols = gj1 - R.gj0;
    R.cutT = R.gi0 > 0;
    R.cutL = R.gj0 > 0;
    R.cutB = gi1 < M;
    R.cutR = gj1 < M;
    const View V{reinterpret_cast<u16a*>(lds + 2), RC, RR * RC};
    const int oi1 = min(M, oi + TR), oj1 = min(M, oj + TC);
    // pairs of sites as in td_step_body
    const PView P = V.pairs();
    constexpr int PPT = (RR * (RC / 2) + NT - 1) / NT;  // pairs per thread
    const int pc = (R.cols + 1) >> 1, np = R.rows * pc;
    const FastDiv fp((uint32_t)pc);
    {
        uint2 c0[PPT], c1[PPT];
#pragma unroll
        for (int q = 0; q < PPT; q++) {
            const int idx = threadIdx.x + q * NT;
            if (idx < np) {
                const int qi = (int)fp.div((uint32_t)idx), jp = idx - qi * pc;
                const uint2* src = a.cells_in + (size_t)(R.gi0 + qi) * M + (R.gj0 + 2 * jp);
                c0[q] = gload(src);
                c1[q] = make_uint2(0u, 0u);
                if (2 * jp + 1 < R.cols) c1[q] = gload(src + 1);
            }
        }
#pragma unroll
        for (int q = 0; q < PPT; q++) {
            const int idx = threadIdx.x + q * NT;
            if (idx < np) {
                const int qi = (int)fp.div((uint32_t)idx), jp = idx - qi * pc;
                int v0[8], v1[8];
                unpack_cell(c0[q], v0);
                unpack_cell(c1[q], v1);
#pragma unroll
                for (int p = 0; p < 8; p++) P.set(p, qi, jp, (uint32_t)v0[p] | ((uint32_t)v1[p] << 16));
            }
        }
    }
    int used = 0;  // margin the earlier steps' phases used up
    for (int s = 0; s < S; s++) {
        const Move mv = decode((moves >> (8 * s)) & 0xffu);
        const int rp = rotate_plane(mv);
        uint32_t bw[PPT];
#pragma unroll
        for (int q = 0; q < PPT; q++) {
            const int idx = threadIdx.x + q * NT;
            if (idx < np) {
                const int qi = (int)fp.div((uint32_t)idx), jp = idx - qi * pc;
                bw[q] = gload(bit_word(a, a.plane + s, R.gi0 + qi, R.gj0 + 2 * jp));
            }
        }
        if (s > 0) __syncthreads();  // the previous step's last phase wrote neighbours' states
#pragma unroll
        for (int q = 0; q < PPT; q++) {
            const int idx = threadIdx.x + q * NT;
            if (idx < np) {
                const int qi = (int)fp.div((uint32_t)idx), jp = idx - qi * pc;
                const int gi = R.gi0 + qi, gj = R.gj0 + 2 * jp;
                const uint32_t w = P.w(rp, qi, jp), b = bw[q] >> (gj & 31);
                const uint32_t x = (uint32_t)rotate_xor(mv, (int)(w & 0xffffu), gi, gj, b & 1u) |
                                   ((uint32_t)rotate_xor(mv, (int)(w >> 16), gi, gj + 1, (b >> 1) & 1u) << 16);
                if (x) P.set(rp, qi, jp, w ^ x);
            }
        }
        __syncthreads();
        replay_updates<NT>(V, R, mv, used);
        used += move_reach(mv);
    }
    __syncthreads();
    store_owned<NT>(a, P, R, oi, oj, oi1, oj1);
}

// Two steps per launch: moves holds the two steps' move codes, the first in the low byte.
template <int TR, int TC, int NT>
__global__ __launch_bounds__(NT) void k_td_step2(TdArgs a, uint32_t moves) {
    td_multi_body<TR, TC, NT, 2>(a, moves, xcd_tile(blockIdx.x, gridDim.x));
}

template <int TR, int TC, int NT>
__global__ __launch_bounds__(NT) void k_td_step2_batch(TdArgs a0, const TdBatchEntry* __restrict__ batch, int nbatch,
                                                       const int64_t* __restrict__ step_ctr, int64_t off) {
    const int64_t s = *step_ctr + off;
    TdArgs a = a0;
    int tile;
    const TdBatchEntry& e = td_resolve(a, batch, nbatch, s, 2, tile);
    td_multi_body<TR, TC, NT, 2>(a, (uint32_t)uniform((int)e.sched[s] | ((int)e.sched[s + 1] << 8)), tile);
}

// ---- whole domain in one workgroup, every step in one launch ----------------------------
constexpr int kSmallNT = 1024;
constexpr int kSmallPer = (kSmallMax * kSmallMax + kSmallNT - 1) / kSmallNT;  // sites per thread

template <bool UNI>
__global__ __launch_bounds__(kSmallNT) void k_td_small(TdArgs a) {
    // 8 planes of M rows of an even pitch (the pair view), a guard of two words each side
    __shared__ uint32_t lds[8 * kSmallMax * kSmallMax / 2 + 4];
    const int M = a.M, n = M * M, pitch = (M + 1) & ~1;
    const View V{reinterpret_cast<u16a*>(lds + 2), pitch, M * pitch};
    Region R{0, 0, M, M, M, false, false, false, false};
    uint4 g[kSmallPer];
    uint32_t pm1[kSmallPer], pm2[kSmallPer], ptm[kSmallPer];
#pragma unroll
    for (int k = 0; k < kSmallPer; k++) {
        const int id = threadIdx.x + k * kSmallNT;
        if (id < n) {
            g[k] = a.gen_in[id];
            load_params<UNI>(a, id, pm1[k], pm2[k], ptm[k]);
            int v[8];
            unpack_cell(a.cells_in[id], v);
#pragma unroll
            for (int p = 0; p < 8; p++) V.s(p, id / M, id % M, v[p]);
        }
    }
    // the sites' coordinates, once (a division per site and step would dominate the step)
    int si[kSmallPer], sj[kSmallPer];
#pragma unroll
    for (int k = 0; k < kSmallPer; k++) {
        const int id = threadIdx.x + k * kSmallNT;
        si[k] = id / M;
        sj[k] = id - si[k] * M;
    }
    __syncthreads();
    for (int64_t s = 0; s < a.nsteps; s++) {
        const Move mv = decode(a.sched[s]);
        const int rp = rotate_plane(mv);
#pragma unroll
        for (int k = 0; k < kSmallPer; k++) {
            const int id = threadIdx.x + k * kSmallNT;
            if (id < n) {
                const uint32_t u = tm_draw_raw(g[k].x, g[k].y, g[k].z, g[k].w, pm1[k], pm2[k], ptm[k]);
                const int gi = si[k], gj = sj[k];
                const int v = V.g(rp, gi, gj);
                const int x = rotate_xor(mv, v, gi, gj, u < 0x80000000u);
                if (x) V.s(rp, gi, gj, v ^ x);
            }
        }
        __syncthreads();
        replay_updates<kSmallNT>(V, R, mv);
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < kSmallPer; k++) {
        const int id = threadIdx.x + k * kSmallNT;
        if (id < n) {
            a.gen_out[id] = g[k];
            const int i = id / M, j = id % M;
            a.cells_out[id] = pack_cell(V, i, j);
        }
    }
}

// ---- derive the move states from a tiling, and merge them back --------------------------

// TilingToTriangleFlips (TriangleDimerTiler.cpp:600-631) at (i, j) of the N/2 x N array
__device__ int tri_state(const int32_t* t, int N, int i, int j) {
    if (i < 1 || i >= N / 2 - 1 || j < 1 || j >= N - 1) return 0;
    const int a = t[(2 * i + 1) * N + j], b = t[(2 * i + 1) * N + j + 1];
    if (j % 2 == 0) {
        const int c = t[2 * i * N + j + 1];
        return ((a & 4) / 4 + (a & 8) / 4 + ((b & 8) / 8) * 3) + 4 * ((b & 4) / 4 + (a & 2) + ((b & 2) / 2) * 3) +
               16 * ((c & 1) * 2 + (a & 1) + ((c & 2) / 2) * 3);
    }
    const int d = t[(2 * i + 2) * N + j];
    return ((a & 1) + (a & 8) / 4 + ((b & 8) / 8) * 3) + 4 * ((b & 1) + (a & 2) + ((b & 2) / 2) * 3) +
           16 * ((d & 8) / 4 + ((d & 4) / 4) * 3 + (a & 4) / 4);
}

// TilingToButterflyFlips (TriangleDimerTiler.cpp:633-680) at (i, j) of the N x N array
__device__ int bfly_state(const int32_t* t, int N, int i, int j) {
    if (i < 2 || i >= N - 2 || j < 2 || j >= N - 2) return 0;
    const int x = i * N + j, s = t[x];
    int f1, f4, f16, f64, c;
    if (i % 2 == 0) {
        if (j % 2 == 0) return 0;
        const int u = t[x - N], ur = t[x - N + 1], d = t[x + N], dl = t[x + N - 1];
        f1 = (u & 8) / 8 + (u & 1) * 2 + (s & 1) * 3;
        f4 = (ur & 1) + (ur & 2) + ((s & 2) / 2) * 3;
        f16 = (d & 2) / 2 + (d & 4) / 2 + ((s & 4) / 4) * 3;
        f64 = (dl & 4) / 4 + (dl & 8) / 4 + ((s & 8) / 8) * 3;
        c = (ur & 4) / 4;
    } else if (j % 2 == 0) {
        const int ur = t[x - N + 1], r = t[x + 1], dl = t[x + N - 1], l = t[x - 1];
        f1 = (ur & 1) + (ur & 2) + (s & 1) * 3;
        f4 = (r & 2) / 2 + (r & 4) / 2 + ((s & 2) / 2) * 3;
        f16 = (dl & 4) / 4 + (dl & 8) / 4 + ((s & 4) / 4) * 3;
        f64 = (l & 8) / 8 + (l & 1) * 2 + ((s & 8) / 8) * 3;
        c = (r & 8) / 8;
    } else {
        const int u = t[x - N], r = t[x + 1], d = t[x + N], l = t[x - 1];
        f1 = (u & 1) + (u & 2) + (s & 1) * 3;
        f4 = (r & 1) + (r & 2) + ((s & 2) / 2) * 3;
        f16 = (d & 4) / 4 + (d & 8) / 4 + ((s & 4) / 4) * 3;
        f64 = (l & 4) / 4 + (l & 8) / 4 + ((s & 8) / 8) * 3;
        c = (r & 8) / 8;
    }
    return f1 + 4 * f4 + 16 * f16 + 64 * f64 + 256 * c;
}

// Walk's host split (TriangleDimerTiler.cpp:104-129): site (a, b) takes eH from t[2a][2b+1],
// eL from t[2a+1][2b], eR from t[2a+1][2b+1], tD/tU from triangle states (a, 2b)/(a, 2b+1),
// bH/bL/bR from butterfly states at the eH/eL/eR positions.
__global__ void k_td_derive(const int32_t* t, int N, uint2* cells) {
    const int M = N / 2;
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= M * M) return;
    const int a = id / M, b = id % M;
    const int r0 = 2 * a, r1 = 2 * a + 1, c0 = 2 * b, c1 = 2 * b + 1;
    const int eh = t[r0 * N + c1], el = t[r1 * N + c0], er = t[r1 * N + c1];
    const int tu = tri_state(t, N, a, c1), tdn = tri_state(t, N, a, c0);
    const int bh = bfly_state(t, N, r0, c1), bl = bfly_state(t, N, r1, c0), br = bfly_state(t, N, r1, c1);
    cells[id] = make_uint2((uint32_t)(eh | el << 4 | er << 8 | tu << 12 | tdn << 19), (uint32_t)(bh | bl << 10 | br << 20));
}

// Walk's merge (TriangleDimerTiler.cpp:437-451): only the lozenge-move states go back.
__global__ void k_td_merge(const uint2* cells, int N, int32_t* t) {
    const int M = N / 2;
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= M * M) return;
    const int a = id / M, b = id % M;
    const uint32_t c = cells[id].x;
    t[(2 * a) * N + 2 * b + 1] = c & 15u;
    t[(2 * a + 1) * N + 2 * b] = (c >> 4) & 15u;
    t[(2 * a + 1) * N + 2 * b + 1] = (c >> 8) & 15u;
}

// Tile configurations of the one-step kernel: {rows, cols, threads}.
struct TdCfg {
    int tr, tc, nt;
};
constexpr TdCfg kCfgs[] = {{32, 32, 256}, {64, 64, 1024}, {32, 64, 512}};
constexpr int kNumCfgs = sizeof(kCfgs) / sizeof(kCfgs[0]);

template <int I>
hipError_t launch_cfg(const TdArgs& a, uint32_t move, hipStream_t st) {
    constexpr TdCfg c = kCfgs[I];
    const int tiles = ((a.M + c.tr - 1) / c.tr) * ((a.M + c.tc - 1) / c.tc);
    hipLaunchKernelGGL((k_td_step<c.tr, c.tc, c.nt>), dim3(tiles), dim3(c.nt), 0, st, a, move);
    return hipGetLastError();
}

template <int I>
hipError_t launch_cfg2(const TdArgs& a, uint32_t moves, hipStream_t st) {
    constexpr TdCfg c = kCfgs[I];
    const int tiles = ((a.M + c.tr - 1) / c.tr) * ((a.M + c.tc - 1) / c.tc);
    hipLaunchKernelGGL((k_td_step2<c.tr, c.tc, c.nt>), dim3(tiles), dim3(c.nt), 0, st, a, moves);
    return hipGetLastError();
}

template <int I>
hipError_t launch_batch_cfg(const TdArgs& a, const TdBatchEntry* batch, int nbatch, int blocks,
                            const int64_t* step_ctr, int64_t off, bool two, int spl, hipStream_t st) {
    constexpr TdCfg c = kCfgs[I];
    if (two)
        hipLaunchKernelGGL((k_td_step2_batch<c.tr, c.tc, c.nt>), dim3(blocks), dim3(c.nt), 0, st, a, batch, nbatch,
                           step_ctr, off);
    else
        hipLaunchKernelGGL((k_td_step_batch<c.tr, c.tc, c.nt>), dim3(blocks), dim3(c.nt), 0, st, a, batch, nbatch,
                           step_ctr, off, spl);
    return hipGetLastError();
}

}  // namespace td

// Configurations 0 .. kNumCfgs-1 run one step per launch; kNumCfgs .. 2 kNumCfgs-1 run the same
// tiles two steps per launch (a last odd step on the one-step kernel of the same tile).
int td_config_count() { return 2 * td::kNumCfgs; }
int td_config_base(int idx) { return idx % td::kNumCfgs; }
int td_config_steps(int idx) { return idx >= td::kNumCfgs ? 2 : 1; }

bool td_config(int idx, int* tr, int* tc, int* nt) {
    if (idx < 0 || idx >= td_config_count()) return false;
    const int b = td_config_base(idx);
    *tr = td::kCfgs[b].tr;
    *tc = td::kCfgs[b].tc;
    *nt = td::kCfgs[b].nt;
    return true;
}

hipError_t td_launch_derive(const int32_t* t, int N, uint2* cells, hipStream_t st) {
    const int n = (N / 2) * (N / 2);
    hipLaunchKernelGGL(td::k_td_derive, dim3((n + 255) / 256), dim3(256), 0, st, t, N, cells);
    return hipGetLastError();
}

hipError_t td_launch_merge(const uint2* cells, int N, int32_t* t, hipStream_t st) {
    const int n = (N / 2) * (N / 2);
    hipLaunchKernelGGL(td::k_td_merge, dim3((n + 255) / 256), dim3(256), 0, st, cells, N, t);
    return hipGetLastError();
}

int td_draw_blocks(int M) { return td_bit_words(M) * ((M + 3) / 4); }

hipError_t td_launch_draw(const TdArgs& a, bool uni, int k, hipStream_t st) {
    if (k < 1 || k > kDrawSteps || !a.rbits) return hipErrorInvalidValue;
    const dim3 grid(td_bit_words(a.M), (a.M + 3) / 4);
    if (uni) hipLaunchKernelGGL(td::k_td_draw<true>, grid, dim3(256), 0, st, a, k);
    else hipLaunchKernelGGL(td::k_td_draw<false>, grid, dim3(256), 0, st, a, k);
    return hipGetLastError();
}

hipError_t td_launch_draw_batch(const TdArgs& a, bool uni, const TdBatchEntry* batch, int nbatch, int blocks,
                                const int64_t* step_ctr, int64_t off, int k, hipStream_t st) {
    if (k < 1 || k > kDrawSteps) return hipErrorInvalidValue;
    if (uni)
        hipLaunchKernelGGL(td::k_td_draw_batch<true>, dim3(blocks), dim3(256), 0, st, a, batch, nbatch, step_ctr, off, k);
    else
        hipLaunchKernelGGL(td::k_td_draw_batch<false>, dim3(blocks), dim3(256), 0, st, a, batch, nbatch, step_ctr, off, k);
    return hipGetLastError();
}

hipError_t td_launch_small(const TdArgs& a, bool uni, hipStream_t st) {
    if (a.M > kSmallMax) return hipErrorInvalidValue;
    if (uni) hipLaunchKernelGGL(td::k_td_small<true>, dim3(1), dim3(td::kSmallNT), 0, st, a);
    else hipLaunchKernelGGL(td::k_td_small<false>, dim3(1), dim3(td::kSmallNT), 0, st, a);
    return hipGetLastError();
}

// cfg: a one-step configuration (the tiles).  two: the launch runs two steps (sched[s] and
// sched[s + 1]); otherwise one.  s = *step_ctr + off.  spl: steps per launch of the run's
// earlier launches (1 or 2), which fixes the copy parity of step s.
hipError_t td_launch_step_batch(int cfg, const TdArgs& a, const TdBatchEntry* batch, int nbatch, int blocks,
                                const int64_t* step_ctr, int64_t off, bool two, int spl, hipStream_t st) {
    if (spl != 1 && spl != 2) return hipErrorInvalidValue;
    switch (cfg) {
        case 0: return td::launch_batch_cfg<0>(a, batch, nbatch, blocks, step_ctr, off, two, spl, st);
        case 1: return td::launch_batch_cfg<1>(a, batch, nbatch, blocks, step_ctr, off, two, spl, st);
        case 2: return td::launch_batch_cfg<2>(a, batch, nbatch, blocks, step_ctr, off, two, spl, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t td_launch_step2(int cfg, const TdArgs& a, uint32_t moves, hipStream_t st) {
    if (!a.rbits || a.plane < 0 || a.plane + 1 >= kDrawSteps) return hipErrorInvalidValue;
    switch (cfg) {
        case 0: return td::launch_cfg2<0>(a, moves, st);
        case 1: return td::launch_cfg2<1>(a, moves, st);
        case 2: return td::launch_cfg2<2>(a, moves, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t td_launch_step(int cfg, const TdArgs& a, uint32_t move, hipStream_t st) {
    if (!a.rbits || a.plane < 0 || a.plane >= kDrawSteps) return hipErrorInvalidValue;
    switch (cfg) {
        case 0: return td::launch_cfg<0>(a, move, st);
        case 1: return td::launch_cfg<1>(a, move, st);
        case 2: return td::launch_cfg<2>(a, move, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace rt
