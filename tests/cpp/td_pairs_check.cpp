// Host check of the triangle-dimer pair phases (rt_tdsite.hpp): on random planes every pair form
// writes exactly what the per-site form of the same reference kernel writes, site by site, for
// inputs of any 16-bit value; built by tests/test_td_pairs.py (hipcc, host code only).
#include "rt_tdsite.hpp"
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>
using namespace rt::td;
static uint32_t td_move_code(int type, int r1, int r2, int p1, int p2) { return rt::td_move(type, r1, r2, p1, p2); }
#ifndef NTRIALS
#define NTRIALS 1500
#endif
constexpr int RR = 14, RC = 18, PL = RR * RC;
struct Buf {
    std::vector<uint32_t> mem;
    Buf() : mem(8 * PL / 2 + 8, 0) {}
    View v() { return View{reinterpret_cast<u16a*>(mem.data() + 4), RC, PL}; }
};
static int fails = 0;
static void cmp(Buf& a, Buf& b, const char* what) {
    View A = a.v(), B = b.v();
    for (int p = 0; p < 8; p++)
        for (int i = 0; i < RR; i++)
            for (int j = 0; j < RC; j++)
                if (A.g(p, i, j) != B.g(p, i, j)) {
                    if (fails++ < 20) printf("%s: plane %d (%d,%d) site %d pair %d\n", what, p, i, j, A.g(p, i, j), B.g(p, i, j));
                }
}
// every move code, state, draw half and site residue: one xor equals the branchy Rotate
static void check_rotate() {
    for (int type = 0; type < 3; type++)
        for (int r1 = 0; r1 < 3; r1++)
            for (int r2 = 0; r2 < 2; r2++)
                for (int p1 = 0; p1 < 3; p1++)
                    for (int p2 = 0; p2 < 2; p2++) {
                        const Move m = decode(td_move_code(type, r1, r2, p1, p2));
                        for (int gi = 0; gi < 12; gi++)
                            for (int gj = 0; gj < 12; gj++)
                                for (int low = 0; low < 2; low++)
                                    for (int v = 0; v < 1024; v++)
                                        if ((v ^ rotate_xor(m, v, gi, gj, low)) != rotate_value(m, v, gi, gj, low) &&
                                            fails++ < 20)
                                            printf("rotate: type %d r1 %d r2 %d p1 %d p2 %d (%d,%d) low %d v %d\n", type, r1,
                                                   r2, p1, p2, gi, gj, low, v);
                    }
}

int main() {
    check_rotate();
    std::mt19937 rng(1);
    for (int trial = 0; trial < NTRIALS; trial++) {
        Buf base;
        const int mode = trial % 3;  // 0: all 16-bit; 1: valid ranges; 2: sparse bits
        View V0 = base.v();
        for (int p = 0; p < 8; p++)
            for (int i = 0; i < RR; i++)
                for (int j = 0; j < RC; j++) {
                    uint32_t x = rng();
                    if (mode == 1) x %= (p < 3 ? 16 : p < 5 ? 127 : 767);
                    if (mode == 2) x &= rng() & rng() & 0x3ff;
                    V0.s(p, i, j, (int)(x & 0xffff));
                }
        const int gi0 = rng() % 50, gj0 = 2 * (rng() % 25);
        auto wmask = [&](int jp) {  // interior columns 1..RC-2 only
            return (2 * jp >= 1 && 2 * jp <= RC - 2 ? 0xffffu : 0u) | (2 * jp + 1 <= RC - 2 ? 0xffff0000u : 0u);
        };
        // 1) loz_update
        for (int t = 0; t < 3; t++) {
            Buf s = base, q = base;
            View S = s.v(); PView Q = q.v().pairs();
            for (int i = 1; i < RR - 1; i++) for (int j = 1; j < RC - 1; j++) loz_update(S, i, j, t);
            for (int i = 1; i < RR - 1; i++) for (int jp = 0; jp < RC / 2; jp++) loz_update_p(Q, i, jp, t, wmask(jp));
            cmp(s, q, "loz_update");
        }
        // 2) tri_u + bfly_from_loz x3
        {
            Buf s = base, q = base;
            View S = s.v(); PView Q = q.v().pairs();
            for (int i = 1; i < RR - 1; i++) for (int j = 1; j < RC - 1; j++) {
                tri_u_from_loz(S, i, j); bfly_from_loz(S, i, j, 0); bfly_from_loz(S, i, j, 1); bfly_from_loz(S, i, j, 2);
            }
            for (int i = 1; i < RR - 1; i++) for (int jp = 0; jp < RC / 2; jp++) {
                const uint32_t wm = wmask(jp);
                put(Q, TU, i, jp, tri_u_p(Q, i, jp), wm);
                put(Q, BH, i, jp, bfly_from_loz_p(Q, i, jp, 0), wm);
                put(Q, BL, i, jp, bfly_from_loz_p(Q, i, jp, 1), wm);
                put(Q, BR, i, jp, bfly_from_loz_p(Q, i, jp, 2), wm);
            }
            cmp(s, q, "tri_u+bfly");
        }
        // 3) tri_update t = 0, 1
        for (int t = 0; t < 2; t++) {
            Buf s = base, q = base;
            View S = s.v(); PView Q = q.v().pairs();
            for (int i = 1; i < RR - 1; i++) for (int j = 1; j < RC - 1; j++) tri_update(S, i, j, t);
            for (int i = 1; i < RR - 1; i++) for (int jp = 0; jp < RC / 2; jp++)
                put(Q, t ? TU : TD, i, jp, tri_update_p(Q, i, jp, t), wmask(jp));
            cmp(s, q, "tri_update");
        }
        // 4) loz_from_tri
        {
            Buf s = base, q = base;
            View S = s.v(); PView Q = q.v().pairs();
            for (int i = 1; i < RR - 1; i++) for (int j = 1; j < RC - 1; j++) loz_from_tri(S, i, j);
            for (int i = 1; i < RR - 1; i++) for (int jp = 0; jp < RC / 2; jp++) loz_from_tri_p(Q, i, jp, wmask(jp));
            cmp(s, q, "loz_from_tri");
        }
        // 6) tri_flipped, every c and direction
        for (int c = 0; c < 3; c++)
            for (int down = 0; down < 2; down++) {
                Buf s = base, q = base;
                View S = s.v(); PView Q = q.v().pairs();
                for (int i = 1; i < RR - 1; i++) for (int j = 1; j < RC - 1; j++) tri_flipped(S, i, j, gi0 + i, gj0 + j, c, down);
                for (int i = 1; i < RR - 1; i++) for (int jp = 0; jp < RC / 2; jp++)
                    tri_flipped_p(Q, i, jp, gi0 + i, gj0 + 2 * jp, c, down, wmask(jp));
                cmp(s, q, "tri_flipped");
            }
        // 5) bfly23 + loz_from_bfly, every t and p1
        for (int t = 0; t < 3; t++)
            for (int p1 = 0; p1 < 3; p1++) {
                Buf s = base, q = base;
                View S = s.v(); PView Q = q.v().pairs();
                const int ts = t == 0 ? (p1 + 1) % 3 : (p1 + 2) % 3;
                for (int i = 1; i < RR - 1; i++) for (int j = 1; j < RC - 1; j++) {
                    bfly_flipped23(S, i, j, gi0 + i, gj0 + j, t, p1);
                    loz_from_bfly(S, i, j, t);
                }
                for (int i = 1; i < RR - 1; i++) for (int jp = 0; jp < RC / 2; jp++) {
                    const int gi = gi0 + i, gj = gj0 + 2 * jp;
                    const uint32_t sm = (bslice(t, gi, gj) == ts ? 0xffffu : 0u) | (bslice(t, gi, gj + 1) == ts ? 0xffff0000u : 0u);
                    bfly23_loz_from_bfly_p(Q, i, jp, t, sm, wmask(jp));
                }
                cmp(s, q, "bfly23+loz_from_bfly");
            }
    }
    printf("fails %d\n", fails);
    return fails != 0;
}
