// rt_models.hpp — per-model flip and update rules on the colour-class layout.
//
// Each model stores its N x N vertex-state tiling as NC colour-class arrays of N rows by
// W columns ("positions"), exactly the split the reference's Walk performs on the host:
//   domino:        2 classes, W = N/2, class 0 = (i+j) even ("B"), class 1 = odd ("W")
//                  (DominoTiler.cpp:15-23), uint8 states
//   lozenge:       3 classes, W = N/3, class ((j-i) mod 3) (LozengeTiler.cpp:29-39), uint8
//   rect-triangle: 3 classes, W = N/3, class ((i-j) mod 3) (RectTriangleTiler.cpp:24-34), int32
// A step rotates one class c (every generator of the site grid draws once, whether or not
// its site can flip) and then recomputes the other class(es) from class c.  The functions
// below are the per-site bodies of the reference kernels; `idx` addresses a site in a
// row-major array of row pitch `pitch` (global memory or an LDS tile).
#pragma once

#include "rt_device.hpp"

namespace rt {

struct DominoModel {
    static constexpr int NC = 2;
    using T = uint8_t;
    // Rotate NDRange(N-2, N/2-2) at offset (1,1): dominokernel.cl:7-8, DominoTiler.cpp:37
    __device__ __forceinline__ static bool has_gen(int gi, int gj, int N, int W) {
        return gi >= 1 && gi <= N - 2 && gj >= 1 && gj <= W - 2;
    }
    __device__ __forceinline__ static int gen_id(int gi, int gj, int N, int W) { return (gi - 1) * (W - 2) + (gj - 1); }
    __device__ __forceinline__ static bool updatable(int gi, int gj, int N, int W) { return has_gen(gi, gj, N, W); }

    // Thresholds of the rotated class, resolved once per step (uniform across the grid).
    struct StepRule {
        uint32_t alo[2], ahi[2], blo[2], bhi[2];
    };
    __device__ __forceinline__ static StepRule step_rule(const RuleParams& rp, int c) {
        StepRule r;
#pragma unroll
        for (int p = 0; p < 2; p++) {
            r.alo[p] = rp.fr.ab_lo[c][p];
            r.ahi[p] = rp.fr.ab_hi[c][p];
            r.blo[p] = rp.fr.ba_lo[c][p];
            r.bhi[p] = rp.fr.ba_hi[c][p];
        }
        return r;
    }

    // dominokernel.cl:11-17 / dominokernelCFTP.cl:25-29 / weighted heat-bath, as thresholds.
    __device__ __forceinline__ static T flip(const StepRule& r, int gi, T s, uint32_t k) {
        const bool odd = gi & 1;
        const uint32_t alo = odd ? r.alo[1] : r.alo[0], ahi = odd ? r.ahi[1] : r.ahi[0];
        const uint32_t blo = odd ? r.blo[1] : r.blo[0], bhi = odd ? r.bhi[1] : r.bhi[0];
        if (s == 3 && k >= alo && k < ahi) return 12;
        if (s == 12 && k >= blo && k < bhi) return 3;
        return s;
    }

    // dominokernel.cl:22-31: the other class from the rotated class c.  The reference's
    // parity argument t is 1 when class 0 ("B") was rotated and 0 when class 1 was.
    template <class P>
    __device__ __forceinline__ static void update(int c, int gi, P* base, int cs, int idx, int pitch) {
        const int t = 1 - c;
        const P* src = base + c * cs;
        P* dst = base + (1 - c) * cs;
        const int offL = (gi + t + 1) & 1, offR = (gi + t) & 1;
        dst[idx] = (P)(((src[idx - pitch] & 2) >> 1) | ((src[idx + pitch] & 1) << 1) | ((src[idx - offL] & 8) >> 1) |
                       ((src[idx + offR] & 4) << 1));
    }
};

// Two domino chains driven by the same random numbers (the top and bottom chains of
// coupling from the past, Examples/AztecDiamondCftpEx.cpp:42-68): each position holds the
// upper chain's state in the low byte and the lower chain's in the high byte.  One TinyMT
// draw decides both flips, and the update's bit moves never cross a byte, so it runs on
// both chains at once.
struct DominoPairModel {
    static constexpr int NC = 2;
    using T = uint16_t;
    __device__ __forceinline__ static bool has_gen(int gi, int gj, int N, int W) {
        return DominoModel::has_gen(gi, gj, N, W);
    }
    __device__ __forceinline__ static int gen_id(int gi, int gj, int N, int W) {
        return DominoModel::gen_id(gi, gj, N, W);
    }
    __device__ __forceinline__ static bool updatable(int gi, int gj, int N, int W) {
        return DominoModel::updatable(gi, gj, N, W);
    }
    using StepRule = DominoModel::StepRule;
    __device__ __forceinline__ static StepRule step_rule(const RuleParams& rp, int c) {
        return DominoModel::step_rule(rp, c);
    }
    __device__ __forceinline__ static T flip(const StepRule& r, int gi, T s, uint32_t k) {
        const uint8_t lo = DominoModel::flip(r, gi, (uint8_t)(s & 0xff), k);
        const uint8_t hi = DominoModel::flip(r, gi, (uint8_t)(s >> 8), k);
        return (T)(lo | (hi << 8));
    }
    template <class P>
    __device__ __forceinline__ static void update(int c, int gi, P* base, int cs, int idx, int pitch) {
        const int t = 1 - c;
        const P* src = base + c * cs;
        P* dst = base + (1 - c) * cs;
        const int offL = (gi + t + 1) & 1, offR = (gi + t) & 1;
        dst[idx] = (P)(((src[idx - pitch] & 0x0202) >> 1) | ((src[idx + pitch] & 0x0101) << 1) |
                       ((src[idx - offL] & 0x0808) >> 1) | ((src[idx + offR] & 0x0404) << 1));
    }
};

struct LozengeModel {
    static constexpr int NC = 3;
    using T = uint8_t;
    // Rotate NDRange(N, N/3): every position has a generator (lozengekernel.cl:11-12)
    __device__ __forceinline__ static bool has_gen(int gi, int gj, int N, int W) {
        return gi >= 0 && gi < N && gj >= 0 && gj < W;
    }
    __device__ __forceinline__ static int gen_id(int gi, int gj, int N, int W) { return gi * W + gj; }
    // UpdateTiles touches the interior only (lozengekernel.cl:42)
    __device__ __forceinline__ static bool updatable(int gi, int gj, int N, int W) {
        return gi >= 1 && gi <= N - 2 && gj >= 1 && gj <= W - 2;
    }

    struct StepRule {
        uint32_t alo, ahi, blo, bhi;
    };
    __device__ __forceinline__ static StepRule step_rule(const RuleParams& rp, int c) {
        return StepRule{rp.fr.ab_lo[0][0], rp.fr.ab_hi[0][0], rp.fr.ba_lo[0][0], rp.fr.ba_hi[0][0]};
    }

    // lozengekernel.cl:21-24 / qlozengekernel.cl:23-29 / heat-bath, as thresholds.
    __device__ __forceinline__ static T flip(const StepRule& r, int gi, T s, uint32_t k) {
        if (s == 25 && k >= r.alo && k < r.ahi) return 38;
        if (s == 38 && k >= r.blo && k < r.bhi) return 25;
        return s;
    }

    // lozengekernel.cl:36-54 with tiling1 = class c, tiling2 = c+1, tiling3 = c+2, t = c.
    template <class P>
    __device__ __forceinline__ static void update(int c, int gi, P* base, int cs, int idx, int pitch) {
        const int t = c;
        const int r = gi % 3;
        const int p2 = (r == (2 * t + 1) % 3), p4 = (r == (2 * t + 2) % 3), p8 = (r == (2 * t) % 3);
        const int p16 = p2;
        const P* s = base + c * cs;
        P* d2 = base + (c == 2 ? 0 : c + 1) * cs;
        P* d3 = base + (c == 0 ? 2 : c - 1) * cs;
        d2[idx] = (P)((d2[idx] & ~38) | ((s[idx - pitch + p2] & 16) >> 3) | ((s[idx - p4] & 8) >> 1) |
                      ((s[idx + pitch] & 1) << 5));
        d3[idx] = (P)((d3[idx] & ~25) | ((s[idx - pitch] & 32) >> 5) | ((s[idx + p8] & 4) << 1) |
                      ((s[idx + pitch - p16] & 2) << 3));
    }
};

struct RectTriModel {
    static constexpr int NC = 3;
    using T = int32_t;
    __device__ __forceinline__ static bool has_gen(int gi, int gj, int N, int W) {
        return gi >= 0 && gi < N && gj >= 0 && gj < W;
    }
    __device__ __forceinline__ static int gen_id(int gi, int gj, int N, int W) { return gi * W + gj; }
    __device__ __forceinline__ static bool updatable(int gi, int gj, int N, int W) {
        return gi >= 1 && gi <= N - 2 && gj >= 1 && gj <= W - 2;
    }

    using StepRule = RectTriRule;
    __device__ __forceinline__ static const RectTriRule& step_rule(const RuleParams& rp, int c) { return rp.rt; }

    // recttrianglekernel.cl:25-118 with every float decision replaced by a threshold on k.
    // The state is looked up in the rule's hash table of the 17 states that can move (the
    // 5-cycle of three-rectangle / all-triangle hexagons, :43-64, and the 12 paired toggles,
    // :66-112); the band of k selects a 5-cycle state's target (:45-63).
    __device__ __forceinline__ static T flip(const RectTriRule& r, int gi, T s, uint32_t k) {
        if (k >= r.gate) return s;
        // __umul24 returns a signed int here: shift it as unsigned (an arithmetic shift would
        // turn products with bit 31 set into negative slots)
        const uint32_t h = ((uint32_t)__umul24((uint32_t)s, kRtHash) >> 27) & 31u;
        if ((uint32_t)s != r.hkey[h]) return s;
        const uint32_t idx = r.hval[h];
        uint32_t thr;
        int nt;
        if (idx < 5) {
            const int b = (k < r.band[0]) ? 0 : (k < r.band[1]) ? 1 : (k < r.band[2]) ? 2 : 3;
            thr = r.cyc_acc[idx][b];
            nt = r.cyc_nt[idx][b];
        } else {
            thr = r.tog_acc[idx - 5];
            nt = r.tog_nt[idx - 5];
        }
        return (k < thr) ? nt : s;
    }

    // recttrianglekernel.cl:120-146: updateTiles1 (class c -> c+1) and updateTiles2
    // (class c -> c+2), t = c.
    template <class P>
    __device__ __forceinline__ static void update(int c, int gi, P* base, int cs, int idx, int pitch) {
        const int t = c;
        const int r = gi % 3;
        const P* s = base + c * cs;
        P* d1 = base + (c == 2 ? 0 : c + 1) * cs;
        P* d2 = base + (c == 0 ? 2 : c - 1) * cs;
        {
            const int p1 = (r == t), p2 = (r == (t + 1) % 3);
            const int x = s[idx - pitch], y = s[idx + p1], z = s[idx + pitch - p2];
            d1[idx] = ((x & 0xf0) << 16) | ((x & 0xf) << 16) | ((y & 0xf00000) >> 8) | ((z & 0xf0000) >> 8) |
                      ((z & 0xf000) >> 8) | ((y & 0xf00) >> 8);
        }
        {
            const int p1 = (r == (t + 2) % 3), p2 = (r == (t + 1) % 3);
            const int x = s[idx - p1], y = s[idx - pitch + p2], z = s[idx + pitch];
            d2[idx] = ((x & 0xf000) << 8) | ((y & 0xf00) << 8) | ((y & 0xf0) << 8) | ((x & 0xf) << 8) |
                      ((z & 0xf00000) >> 16) | ((z & 0xf0000) >> 16);
        }
    }
};

}  // namespace rt
