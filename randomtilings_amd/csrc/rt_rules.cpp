// rt_rules.cpp — host derivation of the integer thresholds that replace every float
// decision of the reference kernels.
//
// The reference draws rd = as_float((u>>9)|0x3f800000) - 1.0f, which is exactly
// k * 2^-23 with k = u >> 9 in [0, 2^23).  Each decision it takes is a predicate on rd
// that is monotone in k (a comparison of rd, or of a monotone float function of rd,
// against a constant).  We evaluate the reference's own expression (float arithmetic,
// double promotion where the OpenCL source compares against a double literal) on the
// host and binary-search the k at which it changes, so the device compares integers and
// reproduces the reference's decisions bit for bit.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>

#include "rt_device.hpp"

namespace rt {

static inline float rd_of(uint32_t k) { return (float)k * (1.0f / 8388608.0f); }  // exact: k < 2^24

// smallest k in [lo, hi) where pred(k) is false, given pred true on a prefix; hi if none
static uint32_t first_false(uint32_t lo, uint32_t hi, const std::function<bool(uint32_t)>& pred) {
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (pred(mid)) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

static constexpr uint32_t K_END = 1u << 23;

// Empty intervals are stored as lo == hi, so that the device may test lo <= k < hi as the
// single unsigned comparison k - lo < hi - lo.
static void normalize(FlipRule& fr) {
    for (int c = 0; c < 3; c++)
        for (int p = 0; p < 2; p++) {
            if (fr.ab_hi[c][p] < fr.ab_lo[c][p]) fr.ab_hi[c][p] = fr.ab_lo[c][p];
            if (fr.ba_hi[c][p] < fr.ba_lo[c][p]) fr.ba_hi[c][p] = fr.ba_lo[c][p];
        }
}

static void set_all(FlipRule& fr, uint32_t alo, uint32_t ahi, uint32_t blo, uint32_t bhi) {
    for (int c = 0; c < 3; c++)
        for (int p = 0; p < 2; p++) {
            fr.ab_lo[c][p] = alo;
            fr.ab_hi[c][p] = ahi;
            fr.ba_lo[c][p] = blo;
            fr.ba_hi[c][p] = bhi;
        }
}

// Domino rules.  rule 0: dominokernel.cl:11 `rd < .8` toggles 3<->12.  rule 1:
// dominokernelCFTP.cl:25 `rd < 0.5` sends 3->12, otherwise 12->3.  rule 2: heat-bath with
// P(12) = p12[class*2 + (row&1)]: `rd < p` sends 3->12, otherwise 12->3.
bool domino_rule(int rule, const double* p12, FlipRule& fr) {
    std::memset(&fr, 0, sizeof(fr));
    if (rule == 0) {
        const uint32_t t = first_false(0, K_END, [](uint32_t k) { return (double)rd_of(k) < .8; });
        set_all(fr, 0, t, 0, t);
        normalize(fr);
        return true;
    }
    if (rule == 1 || rule == 2) {
        for (int c = 0; c < 2; c++)
            for (int p = 0; p < 2; p++) {
                const double pr = (rule == 1) ? 0.5 : p12[c * 2 + p];
                if (!(pr >= 0.0 && pr <= 1.0)) return false;
                const uint32_t t = first_false(0, K_END, [pr](uint32_t k) { return (double)rd_of(k) < pr; });
                fr.ab_lo[c][p] = 0;
                fr.ab_hi[c][p] = t;
                fr.ba_lo[c][p] = t;
                fr.ba_hi[c][p] = K_END;
            }
        normalize(fr);
        return true;
    }
    return false;
}

// Lozenge rules.  rule 0: lozengekernel.cl:21-24.  rule 1: qlozengekernel.cl:23-29 with a
// runtime q (the float variable rd is overwritten by the double expression, as there).
// rule 2: heat-bath `rd < q/(1+q)` adds a cube (25->38), otherwise removes (38->25).
static bool lozenge_rule_raw(int rule, double q, FlipRule& fr);

// The packed lozenge kernel relies on the add and remove intervals being disjoint (every
// lozenge rule's are); a rule whose intervals overlapped would be refused here.
static bool disjoint(const FlipRule& fr) {
    const uint32_t alo = fr.ab_lo[0][0], ahi = fr.ab_hi[0][0], blo = fr.ba_lo[0][0], bhi = fr.ba_hi[0][0];
    return alo == ahi || blo == bhi || ahi <= blo || bhi <= alo;
}

bool lozenge_rule(int rule, double q, FlipRule& fr) {
    if (!lozenge_rule_raw(rule, q, fr)) return false;
    return disjoint(fr);
}

static bool lozenge_rule_raw(int rule, double q, FlipRule& fr) {
    std::memset(&fr, 0, sizeof(fr));
    if (rule == 0) {
        const uint32_t t4 = first_false(0, K_END, [](uint32_t k) { return (double)rd_of(k) < .4; });
        const uint32_t t8 = first_false(t4, K_END, [](uint32_t k) { return (double)rd_of(k) < .8; });
        set_all(fr, 0, t4, t4, t8);
        normalize(fr);
        return true;
    }
    if (!(q > 0.0) || !std::isfinite(q)) return false;
    if (rule == 1) {
        const uint32_t half = first_false(0, K_END, [](uint32_t k) { return (double)rd_of(k) < .5; });
        const uint32_t ta = first_false(0, half, [q](uint32_t k) {
            const float r = (float)((double)rd_of(k) / .5);
            return q >= (double)r;
        });
        const uint32_t tr = first_false(half, K_END, [q](uint32_t k) {
            const float r = (float)(((double)rd_of(k) - .5) * .5);
            return 1 / q >= (double)r;
        });
        set_all(fr, 0, ta, half, tr);
        normalize(fr);
        return true;
    }
    if (rule == 2) {
        const double p = q / (1.0 + q);
        const uint32_t t = first_false(0, K_END, [p](uint32_t k) { return (double)rd_of(k) < p; });
        set_all(fr, 0, t, t, K_END);
        normalize(fr);
        return true;
    }
    return false;
}

// recttrianglekernel.cl:157-160
static float weight_ratio(const float* w, int a, int b) {
    float nb = w[(b & 0xf00000) / 0x100000] * w[(b & 0x0f0000) / 0x010000] * w[(b & 0x00f000) / 0x001000] *
               w[(b & 0x000f00) / 0x000100] * w[(b & 0x0000f0) / 0x000010] * w[(b & 0x00000f) / 0x000001];
    float na = w[(a & 0xf00000) / 0x100000] * w[(a & 0x0f0000) / 0x010000] * w[(a & 0x00f000) / 0x001000] *
               w[(a & 0x000f00) / 0x000100] * w[(a & 0x0000f0) / 0x000010] * w[(a & 0x00000f) / 0x000001];
    return nb / na;
}

// recttrianglekernel.cl:25-118.  The kernel gates on rd < bs (bs = 0.9 stored as float),
// divides rd /= bs in float, picks a band of a 5-cycle target (or a paired toggle) and
// accepts iff weightratio >= rd', rd' = (rd - sc)*4 for the cycle (sc = band start) and
// rd' = rd for toggles.  rd/bs is monotone in k, and so is rd' within a band.
bool recttri_rule(const float* weights16, RectTriRule& r) {
    static const int cyc[5] = {0x122133, 0x331221, 0x8ebbe8, 0xfc99cf, 0xa7dd7a};
    static const int cyc_target[5][4] = {{0xa7dd7a, 0xfc99cf, 0x8ebbe8, 0x331221},
                                         {0xa7dd7a, 0xfc99cf, 0x8ebbe8, 0x122133},
                                         {0xa7dd7a, 0xfc99cf, 0x122133, 0x331221},
                                         {0xa7dd7a, 0x122133, 0x8ebbe8, 0x331221},
                                         {0x122133, 0xfc99cf, 0x8ebbe8, 0x331221}};
    static const int tog[12][2] = {{0x47d47a, 0xd22a33}, {0xd22a33, 0x47d47a}, {0xa74d74, 0x33a22d},
                                   {0x33a22d, 0xa74d74}, {0x8eb558, 0x331eb1}, {0x331eb1, 0x8eb558},
                                   {0x855be8, 0x1be133}, {0x1be133, 0x855be8}, {0x6699cf, 0xcf1221},
                                   {0xcf1221, 0x6699cf}, {0xfc9966, 0x1221fc}, {0x1221fc, 0xfc9966}};
    float w[16] = {0.0f, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
    if (weights16) std::memcpy(w, weights16, sizeof(w));
    std::memset(&r, 0, sizeof(r));
    const float bs = 0.9;
    r.gate = first_false(0, K_END, [bs](uint32_t k) { return rd_of(k) < bs; });
    auto x_of = [bs](uint32_t k) { return rd_of(k) / bs; };
    for (int b = 0; b < 3; b++) {
        const double lim = 0.25 * (b + 1);
        r.band[b] = first_false(0, r.gate, [&](uint32_t k) { return (double)x_of(k) < lim; });
    }
    const uint32_t blo[4] = {0, r.band[0], r.band[1], r.band[2]};
    const uint32_t bhi[4] = {r.band[0], r.band[1], r.band[2], r.gate};
    for (int s = 0; s < 5; s++)
        for (int b = 0; b < 4; b++) {
            const float ratio = weight_ratio(w, cyc[s], cyc_target[s][b]);
            const float sc = 0.25f * b;
            r.cyc_acc[s][b] = first_false(blo[b], bhi[b], [&](uint32_t k) {
                const float rdp = (x_of(k) - sc) * 4;
                return ratio >= rdp;
            });
        }
    for (int t = 0; t < 12; t++) {
        const float ratio = weight_ratio(w, tog[t][0], tog[t][1]);
        r.tog_acc[t] = first_false(0, r.gate, [&](uint32_t k) { return ratio >= x_of(k); });
    }
    // the hash table of the movable states and their targets
    for (int h = 0; h < 32; h++) r.hkey[h] = 0xffffffffu;
    auto put = [&](int state, uint32_t idx) {
        const uint32_t h = (uint32_t)((uint32_t)state * kRtHash) >> 27;
        if (r.hkey[h] != 0xffffffffu) return false;  // kRtHash keeps the 17 states apart
        r.hkey[h] = (uint32_t)state;
        r.hval[h] = idx;
        return true;
    };
    for (int s = 0; s < 5; s++) {
        if (!put(cyc[s], (uint32_t)s)) return false;
        for (int b = 0; b < 4; b++) r.cyc_nt[s][b] = cyc_target[s][b];
    }
    for (int t = 0; t < 12; t++) {
        if (!put(tog[t][0], 5u + t)) return false;
        r.tog_nt[t] = tog[t][1];
    }
    return true;
}

}  // namespace rt
