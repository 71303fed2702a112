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
#include "rt_tridimer.hpp"

namespace rt {
namespace td {

enum Plane { EH = 0, EL = 1, ER = 2, TU = 3, TD = 4, BH = 5, BL = 6, BR = 7 };

struct Move {
    int type, r1, r2, p1, p2;
};
__device__ __forceinline__ Move decode(uint32_t m) {
    return {(int)(m & 3), (int)((m >> 2) & 3), (int)((m >> 4) & 1), (int)((m >> 5) & 3), (int)((m >> 7) & 1)};
}

__device__ __forceinline__ int mod3(int x) { return ((x % 3) + 3) % 3; }

// Region of the move lattice held in LDS: eight uint16 planes of `rows` x pitch.
struct View {
    uint16_t* base;
    int pitch, plane;
    __device__ __forceinline__ int g(int p, int i, int j) const { return base[p * plane + i * pitch + j]; }
    __device__ __forceinline__ void s(int p, int i, int j, int v) const { base[p * plane + i * pitch + j] = (uint16_t)v; }
};

// ---- Rotate kernels (triangledimerkernel.cl:4-22, :192-212, :330-352) ------------------
// The plane a move rotates, and the rotated value of state v at global site (gi, gj) given
// the draw's low half (rd < 0.5 <=> tempered output < 2^31).
__device__ __forceinline__ int rotate_plane(const Move& m) {
    return m.type == 0 ? EH + m.r1 : m.type == 1 ? (m.r2 ? TD : TU) : BH + m.r1;
}
__device__ __forceinline__ int rotate_value(const Move& m, int v, int gi, int gj, bool low) {
    if (!low) return v;
    if (m.type == 0) {
        const int t = m.r1;
        if (((1 - (t & 1)) * gi + ((1 + t) / 2) * gj) % 2 != m.r2) return v;
        return v == 5 ? 10 : v == 10 ? 5 : v;
    }
    if (m.type == 1) {
        if (mod3(gj - gi) != m.r1) return v;
        return v == 45 ? 54 : v == 54 ? 45 : v;
    }
    const int t = m.r1, b1 = 1 - (t & 1), b2 = (1 + t) / 2, b3 = t & 1;
    if ((b1 * gi + b2 * gj) % 3 != m.p1) return v;
    const bool par = ((b1 * gj + b3 * gi) & 1) == m.p2;
    if (v == 170) return par ? 85 : 170;
    if (v == 85) return par ? 170 : 85;
    return v;
}

// Distance a step's update chain reaches: its phases (one Rotate plus the re-derivations of
// replay_updates) each read neighbours at distance <= 1; lozenge and triangle moves take four
// phases, butterfly moves six (kHalo).
__device__ __forceinline__ int move_reach(const Move& m) { return m.type == 2 ? kHalo : 4; }

// ---- update kernels, per site (local i, j; global gi, gj) -------------------------------

// UpdateLozengesFlipped, :24-34
__device__ __forceinline__ void loz_flipped(const View& V, int i, int j, int gi, int gj, int t, int c) {
    if (((1 - (t & 1)) * gi + ((1 + t) / 2) * gj) % 2 == c) return;
    const int a = EH + t, k = (2 - t) / 2, s = t & 1;
    V.s(a, i, j,
        (V.g(a, i - 1, j + s) & 4) / 4 + (V.g(a, i - k, j + 1) & 8) / 4 + (V.g(a, i + 1, j - s) & 1) * 4 +
            (V.g(a, i + k, j - 1) & 2) * 4);
}

// UpdateLozenges0/1/2, :36-81
__device__ __forceinline__ void loz_update(const View& V, int i, int j, int t) {
    const int A = EH + t, B = EH + (t + 1) % 3, C = EH + (t + 2) % 3;
    int b = V.g(B, i, j), c = V.g(C, i, j);
    if (t == 0) {
        b = (b & ~10) | ((V.g(A, i + 1, j) & 1) * 2 + (V.g(A, i + 1, j - 1) & 1) * 8);
        c = (c & ~10) | ((V.g(A, i + 1, j) & 2) + (V.g(A, i, j) & 8));
    } else if (t == 1) {
        b = (b & ~5) | ((V.g(A, i, j) & 1) + (V.g(A, i + 1, j) & 1) * 4);
        c = (c & ~5) | ((V.g(A, i - 1, j) & 2) / 2 + (V.g(A, i, j) & 2) * 2);
    } else {
        b = (b & ~10) | ((V.g(A, i - 1, j) & 2) + (V.g(A, i, j - 1) & 2) * 4);
        c = (c & ~5) | ((V.g(A, i, j) & 1) + (V.g(A, i, j - 1) & 4));
    }
    V.s(B, i, j, b);
    V.s(C, i, j, c);
}

__device__ __forceinline__ int sel(bool c, int v) { return c ? v : 0; }

// UpdateTriangleUFromLozenges, :100-121
__device__ __forceinline__ void tri_u_from_loz(const View& V, int i, int j) {
    const int l0 = V.g(EL, i, j), l1 = V.g(EL, i, j + 1), h0 = V.g(EH, i, j), h1 = V.g(EH, i + 1, j);
    const int L = sel(l0 & 1, 1) + sel(h0 & 8, 2) + sel(h1 & 1, 3);
    const int R = sel(l1 & 1, 1) + sel(h1 & 2, 2) + sel(l1 & 2, 3);
    const int U = sel(l1 & 4, 1) + sel(h1 & 8, 2) + sel(h1 & 4, 3);
    V.s(TU, i, j, L + 4 * R + 16 * U);
}

__device__ __forceinline__ int pack(int t1, int t4, int t16, int t64, int c) {
    return t1 + 4 * t4 + 16 * t16 + 64 * t64 + 256 * c;
}

// UpdateButterflys{H,L,R}FromLozenge, :123-190
__device__ __forceinline__ void bfly_from_loz(const View& V, int i, int j, int w) {
    int c, t1, t4, t16, t64;
    if (w == 0) {
        const int lh = V.g(EH, i, j), lr1 = V.g(ER, i - 1, j), ll1 = V.g(EL, i - 1, j + 1), lr2 = V.g(ER, i, j),
                  ll2 = V.g(EL, i, j);
        c = (ll1 & 4) / 4;
        t1 = (lr1 & 8) / 8 + 2 * (lr1 & 1) + 3 * (lh & 1);
        t4 = (ll1 & 1) + (ll1 & 2) + ((lh & 2) / 2) * 3;
        t16 = (lr2 & 2) / 2 + (lr2 & 4) / 2 + ((lh & 4) / 4) * 3;
        t64 = (ll2 & 4) / 4 + (ll2 & 8) / 4 + ((lh & 8) / 8) * 3;
    } else if (w == 1) {
        const int ll = V.g(EL, i, j), lh1 = V.g(EH, i, j), lr1 = V.g(ER, i, j), lh2 = V.g(EH, i + 1, j - 1),
                  lr2 = V.g(ER, i, j - 1);
        c = (lr1 & 8) / 8;
        t1 = (lh1 & 1) + (lh1 & 2) + 3 * (ll & 1);
        t4 = (lr1 & 2) / 2 + (lr1 & 4) / 2 + ((ll & 2) / 2) * 3;
        t16 = (lh2 & 4) / 4 + (lh2 & 8) / 4 + ((ll & 4) / 4) * 3;
        t64 = (lr2 & 8) / 8 + 2 * (lr2 & 1) + ((ll & 8) / 8) * 3;
    } else {
        const int lr = V.g(ER, i, j), lh1 = V.g(EH, i, j), ll1 = V.g(EL, i, j + 1), lh2 = V.g(EH, i + 1, j),
                  ll2 = V.g(EL, i, j);
        c = (ll1 & 8) / 8;
        t1 = (lh1 & 1) + (lh1 & 2) + 3 * (lr & 1);
        t4 = (ll1 & 1) + (ll1 & 2) + ((lr & 2) / 2) * 3;
        t16 = (lh2 & 4) / 4 + (lh2 & 8) / 4 + ((lr & 4) / 4) * 3;
        t64 = (ll2 & 4) / 4 + (ll2 & 8) / 4 + ((lr & 8) / 8) * 3;
    }
    V.s(BH + w, i, j, pack(t1, t4, t16, t64, c));
}

// UpdateTrianglesFlipped0 (:214-234) / UpdateTrianglesFlipped1 (:236-256)
__device__ __forceinline__ void tri_flipped(const View& V, int i, int j, int gi, int gj, int c, bool down) {
    const int cl = mod3(gj - gi);
    if (cl == c) return;
    const int p = (cl == (c + 1) % 3) ? 1 : 0;
    // masks and the power-of-two divisors (4 - 3p) * 4, 15p + 1, 4 - 3p as shifts
    const int t1 = (4 - 3 * p) * 12, s1 = p ? 2 : 4;
    const int t4 = (15 * p + 1) * 3, s4 = p ? 4 : 0;
    const int t16 = (4 - 3 * p) * 3, s16 = p ? 0 : 2;
    const int a = down ? TD : TU;
    int x1, x4, x16;
    if (!down) {
        x1 = V.g(a, i - (1 - p), j - p);
        x4 = V.g(a, i - p, j + 1);
        x16 = V.g(a, i + 1, j - 1 + p);
    } else {
        x1 = V.g(a, i + 1 - p, j - 1);
        x4 = V.g(a, i + p, j + 1 - p);
        x16 = V.g(a, i - 1, j + p);
    }
    V.s(a, i, j, ((x1 & t1) >> s1) + (((x4 & t4) >> s4) << 2) + (((x16 & t16) >> s16) << 4));
}

// UpdateTriangles, :258-281: t = 0 down from up, t = 1 up from down
__device__ __forceinline__ int tri_digit(int v) {
    return sel((v & 48) == 16, 1) + sel((v & 3) == 3, 2) + sel((v & 12) == 8, 3);
}
__device__ __forceinline__ void tri_update(const View& V, int i, int j, int t) {
    const int src = t ? TD : TU, dst = t ? TU : TD, p = t ? 1 : -1;
    V.s(dst, i, j,
        tri_digit(V.g(src, i, j - 1 + t)) + 4 * tri_digit(V.g(src, i, j + t)) + 16 * tri_digit(V.g(src, i + p, j)));
}

// UpdateLozenge{H,L,R}FromTriangles, :283-328
__device__ __forceinline__ void loz_from_tri(const View& V, int i, int j) {
    const int d = V.g(TD, i, j), uh = V.g(TU, i - 1, j), ul = V.g(TU, i, j - 1), ur = V.g(TU, i, j);
    V.s(EH, i, j, sel((d & 48) == 32, 1) + sel((d & 48) == 48, 2) + sel((uh & 48) == 48, 4) + sel((uh & 48) == 32, 8));
    V.s(EL, i, j, sel((ul & 12) == 4, 1) + sel((ul & 12) == 12, 2) + sel((d & 3) == 1, 4) + sel((d & 3) == 2, 8));
    V.s(ER, i, j, sel((ur & 3) == 1, 1) + sel((d & 12) == 12, 2) + sel((d & 12) == 4, 4) + sel((ur & 3) == 2, 8));
}

__device__ __forceinline__ int bslice(int t, int gi, int gj) { return t == 0 ? gi % 3 : t == 1 ? gj % 3 : (gi + gj) % 3; }

// UpdateButterflysFlipped{H,L,R}1, :355-375, :454-472, :552-572
__device__ __forceinline__ void bfly_flipped1(const View& V, int i, int j, int gi, int gj, int t, int p1, int p2) {
    const int a = BH + t;
    const int par = (t == 1) ? (gi & 1) : (gj & 1);
    if (bslice(t, gi, gj) != p1 || par == p2) return;
    const int c = (V.g(a, i, j) & 256) / 256;
    int t1, t4, t16, t64;
    if (t != 2) {
        const int B1 = (t == 0) ? V.g(a, i, j - 1) : V.g(a, i - 1, j);
        const int B2 = (t == 0) ? V.g(a, i, j + 1) : V.g(a, i + 1, j);
        t1 = ((B1 & 12) / 4 == 0) ? 0 : ((B1 & 12) / 4) % 3 + 1;
        t4 = ((B2 & 3) == 0) ? 0 : ((B2 & 3) + 1) % 3 + 1;
        t16 = ((B2 & 192) / 64 == 0) ? 0 : ((B2 & 192) / 64) % 3 + 1;
        t64 = ((B1 & 48) / 16 == 0) ? 0 : ((B1 & 48) / 16 + 1) % 3 + 1;
    } else {
        const int B1 = V.g(a, i + 1, j - 1), B2 = V.g(a, i - 1, j + 1);
        t1 = ((B2 & 192) / 64 == 0) ? 0 : ((B2 & 192) / 64 + 1) % 3 + 1;
        t4 = ((B2 & 48) / 16 == 0) ? 0 : ((B2 & 48) / 16) % 3 + 1;
        t16 = ((B1 & 12) / 4 == 0) ? 0 : ((B1 & 12) / 4 + 1) % 3 + 1;
        t64 = ((B1 & 3) == 0) ? 0 : (B1 & 3) % 3 + 1;
    }
    V.s(a, i, j, pack(t1, t4, t16, t64, c));
}

// UpdateButterflysFlipped{H,L}21 (:377-398, :474-495), R21 (:574-595)
__device__ __forceinline__ void bfly_flipped21(const View& V, int i, int j, int gi, int gj, int t, int p1) {
    if (bslice(t, gi, gj) != ((t == 0) ? (p1 + 1) % 3 : (p1 + 2) % 3)) return;
    const int a = BH + t, Bf = V.g(a, i, j);
    if (t == 2) {
        const int B2 = V.g(a, i, j + 1), B3 = V.g(a, i + 1, j);
        const int c = sel((B3 & 3) == 1, 1);
        const int t4 = ((Bf & 12) / 4 == 2) ? (Bf & 12) / 4 : (B2 & 3) / 3 + ((B2 & 192) / 192) * 3;
        const int t16 = ((Bf & 48) / 16 == 1) ? (Bf & 48) / 16 : ((B3 & 3) / 3) * 3 + ((B3 & 192) / 192) * 2;
        V.s(a, i, j, pack(Bf & 3, t4, t16, (Bf & 192) / 64, c));
    } else {
        const int B1 = (t == 0) ? V.g(a, i - 1, j) : V.g(a, i - 1, j + 1);
        const int B2 = (t == 0) ? V.g(a, i - 1, j + 1) : V.g(a, i, j + 1);
        const int c = sel((B1 & 48) / 16 == 2, 1);
        const int t1 = ((Bf & 3) == 2) ? (Bf & 3) : (B1 & 192) / 192 + ((B1 & 48) / 48) * 3;
        const int t4 = ((Bf & 12) / 4 == 1) ? (Bf & 12) / 4 : ((B2 & 192) / 192) * 3 + ((B2 & 48) / 48) * 2;
        V.s(a, i, j, pack(t1, t4, (Bf & 48) / 16, (Bf & 192) / 64, c));
    }
}

// UpdateButterflysFlipped{H,L,R}22, :400-428, :497-527, :597-627
__device__ __forceinline__ void bfly_flipped22(const View& V, int i, int j, int gi, int gj, int t, int p1) {
    if (bslice(t, gi, gj) != ((t == 0) ? (p1 + 2) % 3 : (p1 + 1) % 3)) return;
    const int a = BH + t, Bf = V.g(a, i, j);
    if (t == 2) {
        const int s4 = sel((Bf & 12) / 4 != 2, 1), s16 = sel((Bf & 48) / 16 != 1, 1);
        const int B1 = V.g(a, i - 1, j), B2 = V.g(a, i, j + 1), B3 = V.g(a, i + 1, j), B4 = V.g(a, i, j - 1);
        const int c = sel((B1 & 48) / 16 == 1, 1);
        const int t1 = ((Bf & 3) == 1) ? (Bf & 3) : ((B1 & 48) / 48) * 3 + ((B1 & 12) / 12) * 2;
        const int t4 = ((B2 & 256) / 256 == 1) ? 2 : s4 * (Bf & 12) / 4;
        const int t16 = ((B3 & 256) / 256 == 1) ? 1 : s16 * (Bf & 48) / 16;
        const int t64 = ((Bf & 192) / 64 == 2) ? (Bf & 192) / 64 : (B4 & 48) / 48 + ((B4 & 12) / 12) * 3;
        V.s(a, i, j, pack(t1, t4, t16, t64, c));
    } else {
        const int s1 = sel((Bf & 3) != 2, 1), s4 = sel((Bf & 12) / 4 != 1, 1);
        const int B1 = (t == 0) ? V.g(a, i - 1, j) : V.g(a, i - 1, j + 1);
        const int B2 = (t == 0) ? V.g(a, i - 1, j + 1) : V.g(a, i, j + 1);
        const int B3 = (t == 0) ? V.g(a, i + 1, j) : V.g(a, i + 1, j - 1);
        const int B4 = (t == 0) ? V.g(a, i + 1, j - 1) : V.g(a, i, j - 1);
        const int c = sel((B3 & 3) == 2, 1);
        const int t1 = ((B1 & 256) == 256) ? 2 : s1 * (Bf & 3);
        const int t4 = ((B2 & 256) == 256) ? 1 : s4 * ((Bf & 12) / 4);
        const int t16 = ((Bf & 48) / 16 == 2) ? (Bf & 48) / 16 : ((B3 & 3) / 3) * 3 + (B3 & 12) / 12;
        const int t64 = ((Bf & 192) / 64 == 1) ? (Bf & 192) / 64 : ((B4 & 3) / 3) * 2 + ((B4 & 12) / 12) * 3;
        V.s(a, i, j, pack(t1, t4, t16, t64, c));
    }
}

// UpdateButterflysFlipped{H,L,R}23, :430-452, :529-550, :629-651
__device__ __forceinline__ void bfly_flipped23(const View& V, int i, int j, int gi, int gj, int t, int p1) {
    if (bslice(t, gi, gj) != ((t == 0) ? (p1 + 1) % 3 : (p1 + 2) % 3)) return;
    const int a = BH + t, Bf = V.g(a, i, j);
    const int c = (Bf & 256) / 256;
    if (t == 2) {
        const int s1 = sel((Bf & 3) != 1, 1), s64 = sel((Bf & 192) / 64 != 2, 1);
        const int B1 = V.g(a, i - 1, j), B4 = V.g(a, i, j - 1);
        const int t1 = ((B1 & 256) / 256 == 1) ? 1 : s1 * (Bf & 3);
        const int t64 = ((B4 & 256) / 256 == 1) ? 2 : s64 * (Bf & 192) / 64;
        V.s(a, i, j, pack(t1, (Bf & 12) / 4, (Bf & 48) / 16, t64, c));
    } else {
        const int s16 = sel((Bf & 48) / 16 != 2, 1), s64 = sel((Bf & 192) / 64 != 1, 1);
        const int B3 = (t == 0) ? V.g(a, i + 1, j) : V.g(a, i + 1, j - 1);
        const int B4 = (t == 0) ? V.g(a, i + 1, j - 1) : V.g(a, i, j - 1);
        const int t16 = ((B3 & 256) == 256) ? 2 : s16 * ((Bf & 48) / 16);
        const int t64 = ((B4 & 256) == 256) ? 1 : s64 * ((Bf & 192) / 64);
        V.s(a, i, j, pack(Bf & 3, (Bf & 12) / 4, t16, t64, c));
    }
}

// UpdateLozengeFromButterflys{H,L,R}, :654-738
__device__ __forceinline__ void loz_from_bfly(const View& V, int i, int j, int t) {
    const int F = V.g(BH + t, i, j);
    const int full = sel((F & 3) == 3, 1) + sel((F & 12) == 12, 2) + sel((F & 48) == 48, 4) + sel((F & 192) == 192, 8);
    const int c = (F & 256) == 256;
    if (t == 0) {
        V.s(EH, i, j, full);
        V.s(EL, i, j, sel(c, 1) + sel((F & 48) == 48, 2) + sel((F & 192) == 64, 4) + sel((F & 192) == 128, 8));
        V.s(ER, i, j, sel(c, 1) + sel((F & 48) == 16, 2) + sel((F & 48) == 32, 4) + sel((F & 192) == 192, 8));
    } else if (t == 1) {
        V.s(EL, i, j, full);
        V.s(ER, i, j, sel((F & 3) == 3, 1) + sel((F & 12) == 4, 2) + sel((F & 12) == 8, 4) + sel(c, 8));
        V.s(EH, i, j, sel((F & 3) == 1, 1) + sel((F & 3) == 2, 2) + sel((F & 12) == 12, 4) + sel(c, 8));
    } else {
        V.s(ER, i, j, full);
        V.s(EH, i, j, sel((F & 3) == 1, 1) + sel((F & 3) == 2, 2) + sel(c, 4) + sel((F & 192) == 192, 8));
        V.s(EL, i, j, sel((F & 3) == 3, 1) + sel(c, 2) + sel((F & 192) == 64, 4) + sel((F & 192) == 128, 8));
    }
}

// ---- phase driver ----------------------------------------------------------------------

// x / d by a multiply-high with m = floor(2^32 / d) + 1: exact for x < 2^17 and d <= 1024
// (checked exhaustively), which covers every site index of a region (at most 76 x 76 sites).
// A hardware division would cost dozens of instructions per site and phase.
struct FastDiv {
    uint32_t d, m;
    __device__ __forceinline__ explicit FastDiv(uint32_t d_)
        : d(d_), m(d_ > 1 ? (uint32_t)(0x100000000ull / d_ + 1) : 0u) {}
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
    if (w <= NT) {
        const int band = NT / w;
        const int r = (int)threadIdx.x / w;
        if (r >= band) return;
        const int j = j0 + (int)threadIdx.x - r * w, gj = R.gj0 + j;
        for (int i = i0 + r; i < i1; i += band) f(i, j, R.gi0 + i, gj);
        return;
    }
    const int n = (i1 - i0) * w;
    const FastDiv fw((uint32_t)w);
    for (int idx = threadIdx.x; idx < n; idx += NT) {
        const int qi = (int)fw.div((uint32_t)idx);
        const int i = i0 + qi, j = j0 + idx - qi * w;
        f(i, j, R.gi0 + i, R.gj0 + j);
    }
}

// Like for_sites, over the sites with (a * gi + b * gj) mod q == want only (a, b in {0, 1}, q in
// {2, 3}): the colour classes the Rotate-adjacent phases touch.  Rows are walked with the
// residue's first column and a stride of q, so threads are not spent on sites the phase skips.
template <int NT, class F>
__device__ __forceinline__ void for_sites_mod(const Region& R, int m, int a, int b, int q, int want, F f) {
    const int i0 = max(R.cutT ? m : 0, 1 - R.gi0), i1 = min(R.rows - (R.cutB ? m : 0), R.M - 1 - R.gi0);
    const int j0 = max(R.cutL ? m : 0, 1 - R.gj0), j1 = min(R.cols - (R.cutR ? m : 0), R.M - 1 - R.gj0);
    const int w = j1 - j0;
    if (w <= 0 || i1 <= i0) return;
    if (b == 0) {  // a condition on rows only: every column, rows of one residue
        const int band = NT / w;
        const int r = (int)threadIdx.x / w;
        if (r >= band) return;
        const int j = j0 + (int)threadIdx.x - r * w, gj = R.gj0 + j;
        int i = i0;
        while (i < i1 && ((a * (R.gi0 + i)) % q) != want) i++;
        for (i += r * q; i < i1; i += band * q) f(i, j, R.gi0 + i, gj);
        return;
    }
    const int wc = (w + q - 1) / q;  // columns of one residue in a row, at most
    const int band = NT / wc;
    const int r = (int)threadIdx.x / wc;
    if (r >= band) return;
    const int k = (int)threadIdx.x - r * wc;
    for (int i = i0 + r; i < i1; i += band) {
        const int gi = R.gi0 + i;
        // first column j >= j0 with (a gi + gj) mod q == want
        const int j = j0 + (((want - a * gi - (R.gj0 + j0)) % q) + q) % q + q * k;
        if (j < j1) f(i, j, gi, R.gj0 + j);
    }
}

// The update kernels of one step after its Rotate (TriangleDimerTiler.cpp:217-338).  m0: the
// margin earlier steps of the same launch have already used up (their phases' inputs ran out
// there), added to every phase's own.
template <int NT>
__device__ void replay_updates(const View& V, const Region& R, const Move& mv, const int m0 = 0) {
    if (mv.type == 0) {
        const int t = mv.r1, c = mv.r2;
        // UpdateLozengesFlipped rewrites the class ((1 - (t & 1)) gi + ((1 + t) / 2) gj) mod 2 != c
        for_sites_mod<NT>(R, m0 + 1, 1 - (t & 1), (1 + t) / 2, 2, 1 - c,
                          [&](int i, int j, int gi, int gj) { loz_flipped(V, i, j, gi, gj, t, c); });
        __syncthreads();
        for_sites<NT>(R, m0 + 2, [&](int i, int j, int, int) { loz_update(V, i, j, t); });
        __syncthreads();
        for_sites<NT>(R, m0 + 3, [&](int i, int j, int, int) {
            tri_u_from_loz(V, i, j);
            bfly_from_loz(V, i, j, 0);
            bfly_from_loz(V, i, j, 1);
            bfly_from_loz(V, i, j, 2);
        });
        __syncthreads();
        for_sites<NT>(R, m0 + 4, [&](int i, int j, int, int) { tri_update(V, i, j, 0); });
    } else if (mv.type == 1) {
        const int c = mv.r1;
        const bool down = mv.r2 == 1;
        for_sites<NT>(R, m0 + 1, [&](int i, int j, int gi, int gj) { tri_flipped(V, i, j, gi, gj, c, down); });
        __syncthreads();
        for_sites<NT>(R, m0 + 2, [&](int i, int j, int, int) { tri_update(V, i, j, down ? 1 : 0); });
        __syncthreads();
        for_sites<NT>(R, m0 + 3, [&](int i, int j, int, int) { loz_from_tri(V, i, j); });
        __syncthreads();
        for_sites<NT>(R, m0 + 4, [&](int i, int j, int, int) {
            bfly_from_loz(V, i, j, 0);
            bfly_from_loz(V, i, j, 1);
            bfly_from_loz(V, i, j, 2);
        });
    } else {
        const int t = mv.r1, p1 = mv.p1, p2 = mv.p2;
        // the slice classes bslice(t) = (b1 gi + b2 gj) mod 3 these phases rewrite
        const int sa = t == 1 ? 0 : 1, sb = t == 0 ? 0 : 1;
        for_sites_mod<NT>(R, m0 + 1, sa, sb, 3, p1,
                          [&](int i, int j, int gi, int gj) { bfly_flipped1(V, i, j, gi, gj, t, p1, p2); });
        __syncthreads();
        for_sites_mod<NT>(R, m0 + 2, sa, sb, 3, t == 0 ? (p1 + 1) % 3 : (p1 + 2) % 3,
                          [&](int i, int j, int gi, int gj) { bfly_flipped21(V, i, j, gi, gj, t, p1); });
        __syncthreads();
        for_sites_mod<NT>(R, m0 + 3, sa, sb, 3, t == 0 ? (p1 + 2) % 3 : (p1 + 1) % 3,
                          [&](int i, int j, int gi, int gj) { bfly_flipped22(V, i, j, gi, gj, t, p1); });
        __syncthreads();
        // a site's lozenge states depend only on its own (just final) butterfly state
        for_sites<NT>(R, m0 + 4, [&](int i, int j, int gi, int gj) {
            bfly_flipped23(V, i, j, gi, gj, t, p1);
            loz_from_bfly(V, i, j, t);
        });
        __syncthreads();
        for_sites<NT>(R, m0 + 5, [&](int i, int j, int, int) {
            if (t != 0) bfly_from_loz(V, i, j, 0);
            if (t != 1) bfly_from_loz(V, i, j, 1);
            if (t != 2) bfly_from_loz(V, i, j, 2);
            tri_u_from_loz(V, i, j);
        });
        __syncthreads();
        for_sites<NT>(R, m0 + 6, [&](int i, int j, int, int) { tri_update(V, i, j, 0); });
    }
}

// 64-bit cell: x = eH | eL << 4 | eR << 8 | tU << 12 | tD << 19, y = bH | bL << 10 | bR << 20.
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
        g = gin[id];
        load_params<UNI>(a, id, m1, m2, tm);
    }
    uint64_t mine = 0;
    for (int s = 0; s < k; s++) {
        bool low = false;
        if (live) low = tm_draw_raw(g.x, g.y, g.z, g.w, m1, m2, tm) < 0x80000000u;
        const uint64_t w = __ballot(low);
        if (lane == s) mine = w;
    }
    if (live) gout[id] = g;
    if (lane < k) bits[((size_t)lane * M + gi) * td_bit_words(M) + bx] = mine;
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

// Rotation bit of site (gi, gj) in plane p.
__device__ __forceinline__ const uint64_t* bit_word(const TdArgs& a, int p, int gi, int gj) {
    return a.rbits + ((size_t)p * a.M + gi) * td_bit_words(a.M) + (gj >> 6);
}

// ---- one step per launch, tiles + halo --------------------------------------------------
template <int TR, int TC, int NT>
__device__ __forceinline__ void td_step_body(const TdArgs& a, const uint32_t move, const int tile) {
    constexpr int RR = TR + 2 * kHalo, RC = TC + 2 * kHalo;
    __shared__ uint16_t lds[8 * RR * RC];
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
    const View V{lds, RC, RR * RC};
    const int rp = rotate_plane(mv);
    const int n = R.rows * R.cols;
    const int oi1 = min(M, oi + TR), oj1 = min(M, oj + TC);
    // load the region's cells and rotation bits, rotate.  Every load of the thread's sites is
    // issued before the first is used, so the loads overlap instead of waiting one by one.
    constexpr int SPT = (RR * RC + NT - 1) / NT;  // sites per thread
    const FastDiv fc((uint32_t)R.cols);
    uint2 cin[SPT];
    uint64_t bw[SPT];
#pragma unroll
    for (int q = 0; q < SPT; q++) {
        const int idx = threadIdx.x + q * NT;
        if (idx < n) {
            const int qi = (int)fc.div((uint32_t)idx);
            const int gi = R.gi0 + qi, gj = R.gj0 + idx - qi * R.cols;
            cin[q] = a.cells_in[(size_t)gi * M + gj];
            bw[q] = *bit_word(a, a.plane, gi, gj);
        }
    }
#pragma unroll
    for (int q = 0; q < SPT; q++) {
        const int idx = threadIdx.x + q * NT;
        if (idx < n) {
            const int i = (int)fc.div((uint32_t)idx), j = idx - i * R.cols;
            const int gi = R.gi0 + i, gj = R.gj0 + j;
            int v[8];
            unpack_cell(cin[q], v);
            v[rp] = rotate_value(mv, v[rp], gi, gj, (bw[q] >> (gj & 63)) & 1u);
#pragma unroll
            for (int p = 0; p < 8; p++) V.s(p, i, j, v[p]);
        }
    }
    __syncthreads();
    replay_updates<NT>(V, R, mv);
    __syncthreads();
    // write the owned tile
    const int ow = oj1 - oj, on = (oi1 - oi) * ow;
    const FastDiv fo((uint32_t)ow);
    for (int idx = threadIdx.x; idx < on; idx += NT) {
        const int qi = (int)fo.div((uint32_t)idx);
        const int gi = oi + qi, gj = oj + idx - qi * ow;
        const int i = gi - R.gi0, j = gj - R.gj0;
        a.cells_out[(size_t)gi * M + gj] = pack_cell(V, i, j);
    }
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
    a.M = e.M;
    a.cells_in = e.cells[e.buf0 ^ par];
    a.cells_out = e.cells[e.buf0 ^ par ^ 1];
    a.rbits = e.rbits;
    a.plane = (int)(s % kDrawSteps);
    tile = b - e.tile0;
    return e;
}

template <int TR, int TC, int NT>
__global__ __launch_bounds__(NT) void k_td_step_batch(TdArgs a0, const TdBatchEntry* __restrict__ batch, int nbatch,
                                                      const int64_t* __restrict__ step_ctr, int64_t off, int spl) {
    const int64_t s = *step_ctr + off;
    TdArgs a = a0;
    int tile;
    const TdBatchEntry& e = td_resolve(a, batch, nbatch, s, spl, tile);
    td_step_body<TR, TC, NT>(a, e.sched[s], tile);
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
    __shared__ uint16_t lds[8 * RR * RC];
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
    const View V{lds, RC, RR * RC};
    const int n = R.rows * R.cols;
    const int oi1 = min(M, oi + TR), oj1 = min(M, oj + TC);
    constexpr int SPT = (RR * RC + NT - 1) / NT;  // sites per thread
    const FastDiv fc((uint32_t)R.cols);
    {
        uint2 cin[SPT];
#pragma unroll
        for (int q = 0; q < SPT; q++) {
            const int idx = threadIdx.x + q * NT;
            if (idx < n) {
                const int qi = (int)fc.div((uint32_t)idx);
                cin[q] = a.cells_in[(size_t)(R.gi0 + qi) * M + (R.gj0 + idx - qi * R.cols)];
            }
        }
#pragma unroll
        for (int q = 0; q < SPT; q++) {
            const int idx = threadIdx.x + q * NT;
            if (idx < n) {
                const int i = (int)fc.div((uint32_t)idx), j = idx - i * R.cols;
                int v[8];
                unpack_cell(cin[q], v);
#pragma unroll
                for (int p = 0; p < 8; p++) V.s(p, i, j, v[p]);
            }
        }
    }
    int used = 0;  // margin the earlier steps' phases used up
    for (int s = 0; s < S; s++) {
        const Move mv = decode((moves >> (8 * s)) & 0xffu);
        const int rp = rotate_plane(mv);
        uint64_t bw[SPT];
#pragma unroll
        for (int q = 0; q < SPT; q++) {
            const int idx = threadIdx.x + q * NT;
            if (idx < n) {
                const int i = (int)fc.div((uint32_t)idx);
                bw[q] = *bit_word(a, a.plane + s, R.gi0 + i, R.gj0 + idx - i * R.cols);
            }
        }
        if (s > 0) __syncthreads();  // the previous step's last phase wrote neighbours' states
#pragma unroll
        for (int q = 0; q < SPT; q++) {
            const int idx = threadIdx.x + q * NT;
            if (idx < n) {
                const int i = (int)fc.div((uint32_t)idx), j = idx - i * R.cols;
                const int gi = R.gi0 + i, gj = R.gj0 + j;
                const int v = V.g(rp, i, j);
                const int w = rotate_value(mv, v, gi, gj, (bw[q] >> (gj & 63)) & 1u);
                if (w != v) V.s(rp, i, j, w);
            }
        }
        __syncthreads();
        replay_updates<NT>(V, R, mv, used);
        used += move_reach(mv);
    }
    __syncthreads();
    const int ow = oj1 - oj, on = (oi1 - oi) * ow;
    const FastDiv fo((uint32_t)ow);
    for (int idx = threadIdx.x; idx < on; idx += NT) {
        const int qi = (int)fo.div((uint32_t)idx);
        const int gi = oi + qi, gj = oj + idx - qi * ow;
        const int i = gi - R.gi0, j = gj - R.gj0;
        a.cells_out[(size_t)gi * M + gj] = pack_cell(V, i, j);
    }
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
    td_multi_body<TR, TC, NT, 2>(a, (uint32_t)e.sched[s] | ((uint32_t)e.sched[s + 1] << 8), tile);
}

// ---- whole domain in one workgroup, every step in one launch ----------------------------
constexpr int kSmallNT = 1024;
constexpr int kSmallPer = (kSmallMax * kSmallMax + kSmallNT - 1) / kSmallNT;  // sites per thread

template <bool UNI>
__global__ __launch_bounds__(kSmallNT) void k_td_small(TdArgs a) {
    __shared__ uint16_t lds[8 * kSmallMax * kSmallMax];
    const int M = a.M, n = M * M;
    const View V{lds, M, M * M};
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
                const int w = rotate_value(mv, v, gi, gj, u < 0x80000000u);
                if (w != v) V.s(rp, gi, gj, w);
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
