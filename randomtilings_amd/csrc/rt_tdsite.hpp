// rt_tdsite.hpp — per-site triangle-dimer update kernels (triangledimerkernel.cl) on the LDS
// planes of a tile, shared by the device kernels (rt_tridimer.hip) and host-side checks.
//
// Every function is a pure function of a few neighbouring move states, written for one site
// (int arithmetic, as the reference's kernels) and, for the phases that touch every site of a
// region, also for a pair of horizontally adjacent sites held as two 16-bit fields of one
// 32-bit word (the "pair" forms below), which halves the vector instructions of those phases.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rt_tridimer.hpp"

#define RT_TDHD __host__ __device__ __forceinline__

namespace rt {
namespace td {

enum Plane { EH = 0, EL = 1, ER = 2, TU = 3, TD = 4, BH = 5, BL = 6, BR = 7 };

struct Move {
    int type, r1, r2, p1, p2;
};
RT_TDHD Move decode(uint32_t m) {
    return {(int)(m & 3), (int)((m >> 2) & 3), (int)((m >> 4) & 1), (int)((m >> 5) & 3), (int)((m >> 7) & 1)};
}

// Residues of site coordinates.  Unsigned operands let the compiler reduce % 3 to a
// multiply-high and a subtraction (a signed remainder adds a sign fix-up on top).
RT_TDHD int umod3(int x) { return (int)((uint32_t)x % 3u); }  // x >= 0
// x mod 3 in [0, 3) for any x > -6 * 2^20 (6 * 2^20 is a multiple of 2 and of 3)
RT_TDHD int mod3(int x) { return umod3(x + 6 * (1 << 20)); }
// x mod q for q = 2 or 3, same range of x
RT_TDHD int modq(int x, int q) { return q == 2 ? (x & 1) : mod3(x); }

// Region of the move lattice held in LDS: eight uint16 planes of `rows` x pitch.
// The planes are read as 16-bit sites and, by the pair phases, as 32-bit words of two sites.
typedef uint16_t __attribute__((may_alias)) u16a;
typedef uint32_t __attribute__((may_alias)) u32a;

// Pair view of the same planes (pitch and plane size even, base 4-byte aligned): word (p, i, jp)
// holds sites (i, 2 jp) in its low 16 bits and (i, 2 jp + 1) in its high 16 bits.
struct PView {
    u32a* base;
    int pitch, plane;  // in words
    RT_TDHD uint32_t w(int p, int i, int jp) const { return base[p * plane + i * pitch + jp]; }
    RT_TDHD void set(int p, int i, int jp, uint32_t v) const { base[p * plane + i * pitch + jp] = v; }
    // sites (i, 2 jp + dj) and (i, 2 jp + dj + 1), dj in {-1, 0, 1}: a 16-bit funnel of two words
    RT_TDHD uint32_t n(int p, int i, int jp, int dj) const {
        if (dj == 0) return w(p, i, jp);
        if (dj > 0) return (w(p, i, jp) >> 16) | (w(p, i, jp + 1) << 16);
        return (w(p, i, jp - 1) >> 16) | (w(p, i, jp) << 16);
    }
};

struct View {
    u16a* base;
    int pitch, plane;
    RT_TDHD int g(int p, int i, int j) const { return base[p * plane + i * pitch + j]; }
    RT_TDHD void s(int p, int i, int j, int v) const { base[p * plane + i * pitch + j] = (uint16_t)v; }
    RT_TDHD PView pairs() const { return PView{reinterpret_cast<u32a*>(base), pitch / 2, plane / 2}; }
};

// ---- Rotate kernels (triangledimerkernel.cl:4-22, :192-212, :330-352) ------------------
// The plane a move rotates, and the rotated value of state v at global site (gi, gj) given
// the draw's low half (rd < 0.5 <=> tempered output < 2^31).
RT_TDHD int rotate_plane(const Move& m) {
    return m.type == 0 ? EH + m.r1 : m.type == 1 ? (m.r2 ? TD : TU) : BH + m.r1;
}
RT_TDHD int rotate_value(const Move& m, int v, int gi, int gj, bool low) {
    if (!low) return v;
    if (m.type == 0) {
        const int t = m.r1;
        if ((((1 - (t & 1)) * gi + ((1 + t) / 2) * gj) & 1) != m.r2) return v;
        return v == 5 ? 10 : v == 10 ? 5 : v;
    }
    if (m.type == 1) {
        if (mod3(gj - gi) != m.r1) return v;
        return v == 45 ? 54 : v == 54 ? 45 : v;
    }
    const int t = m.r1, b1 = 1 - (t & 1), b2 = (1 + t) / 2, b3 = t & 1;
    if (umod3(b1 * gi + b2 * gj) != m.p1) return v;
    const bool par = ((b1 * gj + b3 * gi) & 1) == m.p2;
    if (v == 170) return par ? 85 : 170;
    if (v == 85) return par ? 170 : 85;
    return v;
}
// The same Rotate as one xor, without branches: every move swaps two states a <-> b, which is
// v ^ (a ^ b) (5 <-> 10: 15; 45 <-> 54: 27; 85 <-> 170: 255), applied where the draw is low, the
// site is in the move's class and v is one of the two.  rotate_value(m, v, ...) ==
// v ^ rotate_xor(m, v, ...) for every move and state (tests/test_td_pairs.py).
RT_TDHD int rotate_xor(const Move& m, int v, int gi, int gj, bool low) {
    const int t = m.r1;
    bool cls;
    int a, b;
    if (m.type == 0) {
        // ((1 - (t & 1)) gi + ((1 + t) / 2) gj) mod 2 == r2
        cls = ((((t & 1) ? 0 : gi) + (t >= 1 ? gj : 0)) & 1) == m.r2;
        a = 5, b = 10;
    } else if (m.type == 1) {
        cls = mod3(gj - gi) == m.r1;
        a = 45, b = 54;
    } else {
        // b1 = 1 - (t & 1), b2 = (1 + t) / 2, b3 = t & 1: slice (b1 gi + b2 gj) mod 3 == p1 and
        // parity (b1 gj + b3 gi) mod 2 == p2
        const int b1gi = (t & 1) ? 0 : gi, b1gj = (t & 1) ? 0 : gj, b2gj = t >= 1 ? gj : 0, b3gi = (t & 1) ? gi : 0;
        cls = (umod3(b1gi + b2gj) == m.p1) & (((b1gj + b3gi) & 1) == m.p2);
        a = 85, b = 170;
    }
    return (low & cls & ((v == a) | (v == b))) ? (a ^ b) : 0;
}

// Distance a step's update chain reaches: its phases (one Rotate plus the re-derivations of
// replay_updates) each read neighbours at distance <= 1; lozenge and triangle moves take four
// phases, butterfly moves six (kHalo).
RT_TDHD int move_reach(const Move& m) { return m.type == 2 ? kHalo : 4; }

// ---- update kernels, per site (local i, j; global gi, gj) -------------------------------

// UpdateLozengesFlipped, :24-34
RT_TDHD void loz_flipped(const View& V, int i, int j, int gi, int gj, int t, int c) {
    if ((((1 - (t & 1)) * gi + ((1 + t) / 2) * gj) & 1) == c) return;
    const int a = EH + t, k = (2 - t) / 2, s = t & 1;
    V.s(a, i, j,
        (V.g(a, i - 1, j + s) & 4) / 4 + (V.g(a, i - k, j + 1) & 8) / 4 + (V.g(a, i + 1, j - s) & 1) * 4 +
            (V.g(a, i + k, j - 1) & 2) * 4);
}

// UpdateLozenges0/1/2, :36-81
RT_TDHD void loz_update(const View& V, int i, int j, int t) {
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

RT_TDHD int sel(bool c, int v) { return c ? v : 0; }

// UpdateTriangleUFromLozenges, :100-121
RT_TDHD void tri_u_from_loz(const View& V, int i, int j) {
    const int l0 = V.g(EL, i, j), l1 = V.g(EL, i, j + 1), h0 = V.g(EH, i, j), h1 = V.g(EH, i + 1, j);
    const int L = sel(l0 & 1, 1) + sel(h0 & 8, 2) + sel(h1 & 1, 3);
    const int R = sel(l1 & 1, 1) + sel(h1 & 2, 2) + sel(l1 & 2, 3);
    const int U = sel(l1 & 4, 1) + sel(h1 & 8, 2) + sel(h1 & 4, 3);
    V.s(TU, i, j, L + 4 * R + 16 * U);
}

RT_TDHD int pack(int t1, int t4, int t16, int t64, int c) {
    return t1 + 4 * t4 + 16 * t16 + 64 * t64 + 256 * c;
}

// UpdateButterflys{H,L,R}FromLozenge, :123-190
RT_TDHD void bfly_from_loz(const View& V, int i, int j, int w) {
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
RT_TDHD void tri_flipped(const View& V, int i, int j, int gi, int gj, int c, bool down) {
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
RT_TDHD int tri_digit(int v) {
    return sel((v & 48) == 16, 1) + sel((v & 3) == 3, 2) + sel((v & 12) == 8, 3);
}
RT_TDHD void tri_update(const View& V, int i, int j, int t) {
    const int src = t ? TD : TU, dst = t ? TU : TD, p = t ? 1 : -1;
    V.s(dst, i, j,
        tri_digit(V.g(src, i, j - 1 + t)) + 4 * tri_digit(V.g(src, i, j + t)) + 16 * tri_digit(V.g(src, i + p, j)));
}

// UpdateLozenge{H,L,R}FromTriangles, :283-328
RT_TDHD void loz_from_tri(const View& V, int i, int j) {
    const int d = V.g(TD, i, j), uh = V.g(TU, i - 1, j), ul = V.g(TU, i, j - 1), ur = V.g(TU, i, j);
    V.s(EH, i, j, sel((d & 48) == 32, 1) + sel((d & 48) == 48, 2) + sel((uh & 48) == 48, 4) + sel((uh & 48) == 32, 8));
    V.s(EL, i, j, sel((ul & 12) == 4, 1) + sel((ul & 12) == 12, 2) + sel((d & 3) == 1, 4) + sel((d & 3) == 2, 8));
    V.s(ER, i, j, sel((ur & 3) == 1, 1) + sel((d & 12) == 12, 2) + sel((d & 12) == 4, 4) + sel((ur & 3) == 2, 8));
}

RT_TDHD int bslice(int t, int gi, int gj) { return umod3(t == 0 ? gi : t == 1 ? gj : gi + gj); }

// UpdateButterflysFlipped{H,L,R}1, :355-375, :454-472, :552-572
RT_TDHD void bfly_flipped1(const View& V, int i, int j, int gi, int gj, int t, int p2) {
    const int a = BH + t;
    const int par = (t == 1) ? (gi & 1) : (gj & 1);
    if (par == p2) return;  // the caller walks only slice p1 (bslice(t, gi, gj) == p1)
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
// The caller walks only the slice these kernels rewrite: bslice(t, gi, gj) == (t == 0 ? p1 + 1 :
// p1 + 2) mod 3.
RT_TDHD void bfly_flipped21(const View& V, int i, int j, int t) {
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
// The caller walks only bslice(t, gi, gj) == (t == 0 ? p1 + 2 : p1 + 1) mod 3.
RT_TDHD void bfly_flipped22(const View& V, int i, int j, int t) {
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
RT_TDHD void bfly_flipped23(const View& V, int i, int j, int gi, int gj, int t, int p1) {
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
RT_TDHD void loz_from_bfly(const View& V, int i, int j, int t) {
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

// ---- pair forms: two horizontally adjacent sites per 32-bit word --------------------------
// Each function below computes, for both 16-bit fields of a word at once, exactly what the
// per-site function of the same kernel computes for one site: the reference's tests of small
// bit patterns ((v & 48) == 16 and the like) become and/and-not of two shifted bits, its sums
// of indicator terms become ors of disjoint bits or adds that stay inside a field (every value
// here is below 2^10).  P16(x) is the constant x in both fields.
RT_TDHD constexpr uint32_t P16(uint32_t x) { return x | (x << 16); }
RT_TDHD uint32_t bitp(uint32_t x, int k) { return (x >> k) & P16(1); }  // bit k of each field
RT_TDHD uint32_t x3(uint32_t x) { return x + (x << 1); }

// Store a pair: fields outside the write mask wm (sites outside the domain's update range) keep
// their value.
RT_TDHD void put(const PView& P, int p, int i, int jp, uint32_t v, uint32_t wm) {
    if (wm != 0xffffffffu) v = (v & wm) | (P.w(p, i, jp) & ~wm);
    P.set(p, i, jp, v);
}

// UpdateLozenges0/1/2 (:36-81)
RT_TDHD void loz_update_p(const PView& P, int i, int jp, int t, uint32_t wm) {
    const int A = EH + t, B = EH + (t + 1) % 3, C = EH + (t + 2) % 3;
    uint32_t b = P.w(B, i, jp), c = P.w(C, i, jp);
    if (t == 0) {
        const uint32_t d0 = P.w(A, i + 1, jp), dm = P.n(A, i + 1, jp, -1), a0 = P.w(A, i, jp);
        b = (b & ~P16(10)) | ((d0 & P16(1)) << 1) | ((dm & P16(1)) << 3);
        c = (c & ~P16(10)) | (d0 & P16(2)) | (a0 & P16(8));
    } else if (t == 1) {
        const uint32_t a0 = P.w(A, i, jp), d0 = P.w(A, i + 1, jp), u0 = P.w(A, i - 1, jp);
        b = (b & ~P16(5)) | (a0 & P16(1)) | ((d0 & P16(1)) << 2);
        c = (c & ~P16(5)) | ((u0 & P16(2)) >> 1) | ((a0 & P16(2)) << 1);
    } else {
        const uint32_t u0 = P.w(A, i - 1, jp), am = P.n(A, i, jp, -1), a0 = P.w(A, i, jp);
        b = (b & ~P16(10)) | (u0 & P16(2)) | ((am & P16(2)) << 2);
        c = (c & ~P16(5)) | (a0 & P16(1)) | (am & P16(4));
    }
    put(P, B, i, jp, b, wm);
    put(P, C, i, jp, c, wm);
}

// UpdateTriangleUFromLozenges (:100-121)
RT_TDHD uint32_t tri_u_p(const PView& P, int i, int jp) {
    const uint32_t l0 = P.w(EL, i, jp), l1 = P.n(EL, i, jp, 1), h0 = P.w(EH, i, jp), h1 = P.w(EH, i + 1, jp);
    const uint32_t L = bitp(l0, 0) + (bitp(h0, 3) << 1) + x3(bitp(h1, 0));
    const uint32_t R = bitp(l1, 0) + (bitp(h1, 1) << 1) + x3(bitp(l1, 1));
    const uint32_t U = bitp(l1, 2) + (bitp(h1, 3) << 1) + x3(bitp(h1, 2));
    return L + (R << 2) + (U << 4);
}

// 3 (b0 + 4 b1 + 16 b2 + 64 b3) for the low four bits b0..b3 of each field
RT_TDHD uint32_t spread3(uint32_t x) {
    x &= P16(15);
    x = (x | (x << 2)) & P16(0x33);
    x = (x | (x << 1)) & P16(0x55);
    return x3(x);
}

// UpdateButterflys{H,L,R}FromLozenge (:123-190).  Each base-4 digit is a sum of indicator terms
// weighted 1, 2, 3; the terms with weight 3 all come from one lozenge state (spread3), the others
// fill disjoint two-bit fields.
RT_TDHD uint32_t bfly_from_loz_p(const PView& P, int i, int jp, int w) {
    const uint32_t k3 = P16(3);
    if (w == 0) {
        const uint32_t lh = P.w(EH, i, jp), lr1 = P.w(ER, i - 1, jp), ll1 = P.n(EL, i - 1, jp, 1),
                       lr2 = P.w(ER, i, jp), ll2 = P.w(EL, i, jp);
        const uint32_t v = (bitp(lr1, 3) | (bitp(lr1, 0) << 1)) | ((ll1 & k3) << 2) | (((lr2 >> 1) & k3) << 4) |
                           (((ll2 >> 2) & k3) << 6) | (bitp(ll1, 2) << 8);
        return v + spread3(lh);
    }
    if (w == 1) {
        const uint32_t ll = P.w(EL, i, jp), lh1 = P.w(EH, i, jp), lr1 = P.w(ER, i, jp), lh2 = P.n(EH, i + 1, jp, -1),
                       lr2 = P.n(ER, i, jp, -1);
        const uint32_t v = (lh1 & k3) | (((lr1 >> 1) & k3) << 2) | (((lh2 >> 2) & k3) << 4) |
                           ((bitp(lr2, 3) | (bitp(lr2, 0) << 1)) << 6) | (bitp(lr1, 3) << 8);
        return v + spread3(ll);
    }
    const uint32_t lr = P.w(ER, i, jp), lh1 = P.w(EH, i, jp), ll1 = P.n(EL, i, jp, 1), lh2 = P.w(EH, i + 1, jp),
                   ll2 = P.w(EL, i, jp);
    const uint32_t v = (lh1 & k3) | ((ll1 & k3) << 2) | (((lh2 >> 2) & k3) << 4) | (((ll2 >> 2) & k3) << 6) |
                       (bitp(ll1, 3) << 8);
    return v + spread3(lr);
}

// tri_digit: [(v & 48) == 16] + 2 [(v & 3) == 3] + 3 [(v & 12) == 8]
RT_TDHD uint32_t tri_digit_p(uint32_t v) {
    const uint32_t e1 = bitp(v, 4) & ~bitp(v, 5), e2 = bitp(v, 0) & bitp(v, 1), e3 = bitp(v, 3) & ~bitp(v, 2);
    return e1 + (e2 << 1) + x3(e3);
}
// UpdateTriangles (:258-281): t = 0 down from up, t = 1 up from down
RT_TDHD uint32_t tri_update_p(const PView& P, int i, int jp, int t) {
    const int src = t ? TD : TU;
    const uint32_t a = P.n(src, i, jp, t - 1), b = P.n(src, i, jp, t), c = P.w(src, i + (t ? 1 : -1), jp);
    return tri_digit_p(a) + (tri_digit_p(b) << 2) + (tri_digit_p(c) << 4);
}

// UpdateLozenge{H,L,R}FromTriangles (:283-328)
RT_TDHD void loz_from_tri_p(const PView& P, int i, int jp, uint32_t wm) {
    const uint32_t d = P.w(TD, i, jp), uh = P.w(TU, i - 1, jp), ul = P.n(TU, i, jp, -1), ur = P.w(TU, i, jp);
    const uint32_t d0 = bitp(d, 0), d1 = bitp(d, 1), d2 = bitp(d, 2), d3 = bitp(d, 3), d4 = bitp(d, 4), d5 = bitp(d, 5);
    const uint32_t eh = (d5 & ~d4) | ((d5 & d4) << 1) | ((bitp(uh, 5) & bitp(uh, 4)) << 2) |
                        ((bitp(uh, 5) & ~bitp(uh, 4)) << 3);
    const uint32_t el = (bitp(ul, 2) & ~bitp(ul, 3)) | ((bitp(ul, 2) & bitp(ul, 3)) << 1) | ((d0 & ~d1) << 2) |
                        ((d1 & ~d0) << 3);
    const uint32_t er = (bitp(ur, 0) & ~bitp(ur, 1)) | ((d2 & d3) << 1) | ((d2 & ~d3) << 2) |
                        ((bitp(ur, 1) & ~bitp(ur, 0)) << 3);
    put(P, EH, i, jp, eh, wm);
    put(P, EL, i, jp, el, wm);
    put(P, ER, i, jp, er, wm);
}

// UpdateButterflysFlipped{H,L,R}23 (:430-452, :529-550, :629-651) on the fields in the slice
// mask sm, then UpdateLozengeFromButterflys{H,L,R} (:654-738) on every field of the word.
// Digit rules: "d == k ? 0 : d" is d ^ k where d == k; "flag ? k : x" replaces a two-bit digit.
RT_TDHD void bfly23_loz_from_bfly_p(const PView& P, int i, int jp, int t, uint32_t sm, uint32_t wm) {
    const uint32_t k3 = P16(3);
    const int a = BH + t;
    const uint32_t Bf = P.w(a, i, jp);
    uint32_t F = Bf;
    if (sm) {
        const uint32_t d1 = Bf & k3, d4 = (Bf >> 2) & k3, d16 = (Bf >> 4) & k3, d64 = (Bf >> 6) & k3, c = bitp(Bf, 8);
        uint32_t nb;
        if (t == 2) {
            const uint32_t b1c = bitp(P.w(a, i - 1, jp), 8), b4c = bitp(P.n(a, i, jp, -1), 8);
            const uint32_t x1 = d1 ^ (bitp(d1, 0) & ~bitp(d1, 1));           // d1 == 1 ? 0 : d1
            const uint32_t x64 = d64 ^ ((bitp(d64, 1) & ~bitp(d64, 0)) << 1);  // d64 == 2 ? 0 : d64
            const uint32_t t1 = (x1 & ~x3(b1c)) | b1c;                         // B1 flag ? 1
            const uint32_t t64 = (x64 & ~x3(b4c)) | (b4c << 1);                // B4 flag ? 2
            nb = t1 | (d4 << 2) | (d16 << 4) | (t64 << 6) | (c << 8);
        } else {
            const uint32_t b3 = t == 0 ? P.w(a, i + 1, jp) : P.n(a, i + 1, jp, -1);
            const uint32_t b4 = t == 0 ? P.n(a, i + 1, jp, -1) : P.n(a, i, jp, -1);
            const uint32_t b3c = bitp(b3, 8), b4c = bitp(b4, 8);
            const uint32_t x16 = d16 ^ ((bitp(d16, 1) & ~bitp(d16, 0)) << 1);  // d16 == 2 ? 0 : d16
            const uint32_t x64 = d64 ^ (bitp(d64, 0) & ~bitp(d64, 1));         // d64 == 1 ? 0 : d64
            const uint32_t t16 = (x16 & ~x3(b3c)) | (b3c << 1);                // B3 flag ? 2
            const uint32_t t64 = (x64 & ~x3(b4c)) | b4c;                       // B4 flag ? 1
            nb = d1 | (d4 << 2) | (t16 << 4) | (t64 << 6) | (c << 8);
        }
        F = (nb & sm) | (Bf & ~sm);
        put(P, a, i, jp, F, wm);
    }
    const uint32_t f0 = bitp(F, 0), f1 = bitp(F, 1), f2 = bitp(F, 2), f3 = bitp(F, 3), f4 = bitp(F, 4),
                   f5 = bitp(F, 5), f6 = bitp(F, 6), f7 = bitp(F, 7), c = bitp(F, 8);
    const uint32_t full = (f0 & f1) | ((f2 & f3) << 1) | ((f4 & f5) << 2) | ((f6 & f7) << 3);
    if (t == 0) {
        put(P, EH, i, jp, full, wm);
        put(P, EL, i, jp, c | ((f4 & f5) << 1) | ((f6 & ~f7) << 2) | ((f7 & ~f6) << 3), wm);
        put(P, ER, i, jp, c | ((f4 & ~f5) << 1) | ((f5 & ~f4) << 2) | ((f6 & f7) << 3), wm);
    } else if (t == 1) {
        put(P, EL, i, jp, full, wm);
        put(P, ER, i, jp, (f0 & f1) | ((f2 & ~f3) << 1) | ((f3 & ~f2) << 2) | (c << 3), wm);
        put(P, EH, i, jp, (f0 & ~f1) | ((f1 & ~f0) << 1) | ((f2 & f3) << 2) | (c << 3), wm);
    } else {
        put(P, ER, i, jp, full, wm);
        put(P, EH, i, jp, (f0 & ~f1) | ((f1 & ~f0) << 1) | (c << 2) | ((f6 & f7) << 3), wm);
        put(P, EL, i, jp, (f0 & f1) | (c << 1) | ((f6 & ~f7) << 2) | ((f7 & ~f6) << 3), wm);
    }
}

// UpdateTrianglesFlipped0/1 (:214-256) on a pair: the sites of class cl = (gj - gi) mod 3 != c
// take three two-bit fields of neighbours whose positions depend on p = [cl == c + 1]; both
// placements are computed for the whole word and each field keeps its own (class-c fields keep
// their value).  gi, gj: the low site.
RT_TDHD void tri_flipped_p(const PView& P, int i, int jp, int gi, int gj, int c, bool down, uint32_t wm) {
    const int a = down ? TD : TU;
    const int cl0 = mod3(gj - gi), cl1 = cl0 == 2 ? 0 : cl0 + 1, c1 = c == 2 ? 0 : c + 1;
    const uint32_t act = (cl0 != c ? 0x0000ffffu : 0u) | (cl1 != c ? 0xffff0000u : 0u);
    if (!act) return;
    const uint32_t one = (cl0 == c1 ? 0x0000ffffu : 0u) | (cl1 == c1 ? 0xffff0000u : 0u);  // fields with p = 1
    uint32_t y1, y4, y16, z1, z4, z16;  // p = 0 and p = 1 neighbours
    if (!down) {
        y1 = P.w(a, i - 1, jp), y4 = P.n(a, i, jp, 1), y16 = P.n(a, i + 1, jp, -1);
        z1 = P.n(a, i, jp, -1), z4 = P.n(a, i - 1, jp, 1), z16 = P.w(a, i + 1, jp);
    } else {
        y1 = P.n(a, i + 1, jp, -1), y4 = P.n(a, i, jp, 1), y16 = P.w(a, i - 1, jp);
        z1 = P.n(a, i, jp, -1), z4 = P.w(a, i + 1, jp), z16 = P.n(a, i - 1, jp, 1);
    }
    const uint32_t v0 = ((y1 >> 4) & P16(3)) | ((y4 & P16(3)) << 2) | (((y16 >> 2) & P16(3)) << 4);
    const uint32_t v1 = ((z1 >> 2) & P16(3)) | (((z4 >> 4) & P16(3)) << 2) | ((z16 & P16(3)) << 4);
    const uint32_t old = P.w(a, i, jp);
    const uint32_t nv = (((v1 & one) | (v0 & ~one)) & act) | (old & ~act);
    put(P, a, i, jp, nv, wm);
}

}  // namespace td
}  // namespace rt
