// rt_oracle.cpp — CPU ORACLE (test infrastructure only; never linked into the product).
//
// A sequential restatement of the reference chain semantics, used by tests/, by
// __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker.
// Every reference OpenCL kernel is hazard-free inside one launch (it reads one colour
// class and writes another, or writes only its own site), so a sequential loop over the
// NDRange reproduces any parallel execution exactly.  Float decisions are evaluated with
// the same float/double expressions the reference kernels use (OpenCL promotes a float
// compared against a double literal to double); the product replaces them by integer
// thresholds on k = u>>9, and the parity tests pin that replacement.
//
// Pinning (what makes this restatement trustworthy, see tests/test_oracle.py):
//   * TinyMT32: known-answer outputs of tinymt32_init(seed=1) with the default parameter
//     set, and agreement with the reference's independent branchless host coding
//     (src/TinyMT/tinymt32.h:72-109) compiled into oracle/_ref.
//   * host schedule RNGs: libstdc++'s own std::mt19937 / std::minstd_rand0 /
//     std::uniform_int_distribution (the reference uses exactly these), pinned by the
//     C++-standard 10000th-output anchors.
//   * Rotate/Update rules: legal-tiling invariants after every step, exact stationary
//     laws by brute force on tiny domains, conservation laws.
#include <cstdint>
#include <cstring>
#include <cmath>
#include <random>
#include <vector>
#include <chrono>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------------------------
// TinyMT32 (src/TinyMT/tinymt32.clh)
// ---------------------------------------------------------------------------------
struct TM {
    uint32_t s0, s1, s2, s3, mat1, mat2, tmat;
};

// tinymt32.clh:39-55
inline void tm_next(TM& t) {
    uint32_t x = (t.s0 & 0x7fffffffu) ^ t.s1 ^ t.s2;
    uint32_t y = t.s3;
    x ^= x << 1;
    y ^= (y >> 1) ^ x;
    t.s0 = t.s1;
    t.s1 = t.s2;
    t.s2 = x ^ (y << 10);
    t.s3 = y;
    if (y & 1u) {
        t.s1 ^= t.mat1;
        t.s2 ^= t.mat2;
    }
}

// tinymt32.clh:62-73
inline uint32_t tm_temper(const TM& t) {
    uint32_t t0 = t.s3;
    uint32_t t1 = t.s0 + (t.s2 >> 8);
    t0 ^= t1;
    if (t1 & 1u) t0 ^= t.tmat;
    return t0;
}

// tinymt32.clh:81-93 + 125-129: as_float((t0>>9)|0x3f800000) - 1.0f
inline float tm_single01(TM& t) {
    tm_next(t);
    uint32_t bits = (tm_temper(t) >> 9) | 0x3f800000u;
    float f;
    std::memcpy(&f, &bits, 4);
    return f - 1.0f;
}

// tinymt32.clh:182-203 (with period certification :162-174)
void tm_init(TM& t, uint32_t seed) {
    uint32_t st[4] = {seed, t.mat1, t.mat2, t.tmat};
    for (int i = 1; i < 8; i++)
        st[i & 3] ^= (uint32_t)i + 1812433253u * (st[(i - 1) & 3] ^ (st[(i - 1) & 3] >> 30));
    t.s0 = st[0];
    t.s1 = st[1];
    t.s2 = st[2];
    t.s3 = st[3];
    if ((t.s0 & 0x7fffffffu) == 0 && t.s1 == 0 && t.s2 == 0 && t.s3 == 0) {
        t.s0 = 'T';
        t.s1 = 'I';
        t.s2 = 'N';
        t.s3 = 'Y';
    }
    for (int i = 0; i < 8; i++) {
        tm_next(t);
        (void)tm_temper(t);
    }
}

// A generator table as LoadTinyMT builds it (file_reader.cpp:22-38): params per id.
struct Table {
    std::vector<TM> g;
    bool ok = false;
    Table(const uint32_t* params, int64_t n) {
        if (!params || n <= 0) return;
        g.resize((size_t)n);
        for (int64_t i = 0; i < n; i++) {
            g[i].mat1 = params[3 * i + 0];
            g[i].mat2 = params[3 * i + 1];
            g[i].tmat = params[3 * i + 2];
        }
        ok = true;
    }
    // InitTinyMT kernel: tinymt32_init(&tiny, seed+id) over NDRange(count); the uint
    // kernel argument plus a size_t id wraps modulo 2^32 (e.g. dominokernel.cl:33-40).
    void init(uint32_t seed, int64_t count) {
        for (int64_t id = 0; id < count && id < (int64_t)g.size(); id++)
            tm_init(g[id], (uint32_t)(seed + (uint64_t)id));
    }
};

// The host passes a `long` seed through a `const int` kernel-functor argument into a
// `uint` kernel parameter: only the low 32 bits survive.
inline uint32_t low32(long v) { return (uint32_t)(uint64_t)v; }

// ---------------------------------------------------------------------------------
// Domino (src/Domino/DominoTiler.cpp:9-107, dominokernel.cl, dominokernelCFTP.cl)
// ---------------------------------------------------------------------------------
struct DominoRule {
    int rule;          // 0 = dominokernel.cl (lazy, 0.8), 1 = dominokernelCFTP.cl, 2 = weighted
    double p12[4];     // rule 2: probability of the 12 state, indexed by class*2 + (i&1)
};

void domino_rotate(const DominoRule& R, std::vector<TM>& gens, uint8_t* a, int N, int t) {
    const int W = N / 2;
    const int G0 = N - 2, G1 = W - 2;
    for (int g0 = 0; g0 < G0; g0++) {
        for (int g1 = 0; g1 < G1; g1++) {
            const size_t id = (size_t)g1 + (size_t)G1 * g0;  // tinymt.clh:23-30
            const int i = g0 + 1, j = g1 + 1;
            float rd = tm_single01(gens[id]);
            uint8_t& s = a[i * W + j];
            if (R.rule == 0) {  // dominokernel.cl:11-17
                if ((double)rd < .8) {
                    if (s == 3) s = 12;
                    else if (s == 12) s = 3;
                }
            } else if (R.rule == 1) {  // dominokernelCFTP.cl:25-29
                if ((double)rd < 0.5) {
                    if (s == 3) s = 12;
                } else {
                    if (s == 12) s = 3;
                }
            } else {  // weighted heat-bath (new rule; uniform weights reduce to rule 1)
                const double p = R.p12[t * 2 + (i & 1)];
                if ((double)rd < p) {
                    if (s == 3) s = 12;
                } else {
                    if (s == 12) s = 3;
                }
            }
        }
    }
}

// dominokernel.cl:22-31 (identical in dominokernelCFTP.cl:33-42)
void domino_update(uint8_t* dst, const uint8_t* ref, int N, int t) {
    const int W = N / 2;
    for (int i = 1; i <= N - 2; i++) {
        for (int j = 1; j <= W - 2; j++) {
            dst[i * W + j] = (uint8_t)((ref[(i - 1) * W + j] & 2) / 2 + 2 * (ref[(i + 1) * W + j] & 1) +
                                       (ref[i * W + j - (i + t + 1) % 2] & 8) / 2 +
                                       2 * (ref[i * W + j + (i + t) % 2] & 4));
        }
    }
}

void domino_split(const int32_t* t, int N, std::vector<uint8_t>& B, std::vector<uint8_t>& Wc) {
    const int W = N / 2;
    B.assign((size_t)W * N, 0);
    Wc.assign((size_t)W * N, 0);
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            if ((i + j) % 2 == 0) B[i * W + j / 2] = (uint8_t)t[i * N + j];
            else Wc[i * W + j / 2] = (uint8_t)t[i * N + j];
        }
}

void domino_merge(int32_t* t, int N, const std::vector<uint8_t>& B, const std::vector<uint8_t>& Wc) {
    const int W = N / 2;
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) t[i * N + j] = ((i + j) % 2 == 0) ? B[i * W + j / 2] : Wc[i * W + j / 2];
}

void domino_steps(const DominoRule& R, std::vector<TM>& gens, std::vector<uint8_t>& B, std::vector<uint8_t>& Wc,
                  int N, std::mt19937& mt, long steps) {
    for (long s = 0; s < steps; ++s) {
        int r = mt() % 2;  // DominoTiler.cpp:35
        if (r == 1) {
            domino_rotate(R, gens, B.data(), N, 0);
            domino_update(Wc.data(), B.data(), N, 1);
        } else {
            domino_rotate(R, gens, Wc.data(), N, 1);
            domino_update(B.data(), Wc.data(), N, 0);
        }
    }
}

// ---------------------------------------------------------------------------------
// Lozenge (src/Lozenge/LozengeTiler.cpp:18-88, lozengekernel.cl, qlozengekernel.cl)
// ---------------------------------------------------------------------------------
void lozenge_rotate(int rule, double q, std::vector<TM>& gens, uint8_t* a, int N) {
    const int M = N / 3;
    for (int i = 0; i < N; i++) {
        for (int j = 0; j < M; j++) {
            const size_t id = (size_t)j + (size_t)M * i;
            float rd = tm_single01(gens[id]);
            uint8_t& s = a[i * M + j];
            if (rule == 0) {  // lozengekernel.cl:21-24
                if ((double)rd < .4) {
                    if (s == 25) s = 38;
                } else if ((double)rd < .8) {
                    if (s == 38) s = 25;
                }
            } else if (rule == 1) {  // qlozengekernel.cl:23-29 with a runtime q
                if ((double)rd < .5) {
                    rd = (float)((double)rd / .5);
                    if (s == 25 && q >= (double)rd) s = 38;
                } else {
                    rd = (float)(((double)rd - .5) * .5);
                    if (s == 38 && 1 / q >= (double)rd) s = 25;
                }
            } else {  // heat-bath: add with probability q/(1+q), remove otherwise
                if ((double)rd < q / (1.0 + q)) {
                    if (s == 25) s = 38;
                } else {
                    if (s == 38) s = 25;
                }
            }
        }
    }
}

// lozengekernel.cl:36-54
void lozenge_update(const uint8_t* t1, uint8_t* t2, uint8_t* t3, int N, int t) {
    const int M = N / 3;
    for (int i = 1; i < N - 1; i++) {
        const int p2 = (i % 3 == (2 * t + 1) % 3), p4 = (i % 3 == (2 * t + 2) % 3);
        const int p8 = (i % 3 == (2 * t) % 3), p16 = (i % 3 == (2 * t + 1) % 3);
        for (int j = 1; j < M - 1; j++) {
            int v2 = t2[i * M + j];
            v2 &= ~((v2 & 2) + (v2 & 4) + (v2 & 32));
            v2 += (t1[(i - 1) * M + j + p2] & 16) / 8 + (t1[i * M + j - p4] & 8) / 2 + (t1[(i + 1) * M + j] & 1) * 32;
            t2[i * M + j] = (uint8_t)v2;
            int v3 = t3[i * M + j];
            v3 &= ~((v3 & 1) + (v3 & 8) + (v3 & 16));
            v3 += (t1[(i - 1) * M + j] & 32) / 32 + (t1[i * M + j + p8] & 4) * 2 + (t1[(i + 1) * M + j - p16] & 2) * 8;
            t3[i * M + j] = (uint8_t)v3;
        }
    }
}

inline int mod3(int x) { return ((x % 3) + 3) % 3; }

// ---------------------------------------------------------------------------------
// Rectangle-triangle (src/RectTriangle/RectTriangleTiler.cpp:17-92, recttrianglekernel.cl)
// ---------------------------------------------------------------------------------
float rt_ratio(const float* w, int a, int b) {  // recttrianglekernel.cl:157-160
    float nb = w[(b & 0xf00000) / 0x100000] * w[(b & 0x0f0000) / 0x010000] * w[(b & 0x00f000) / 0x001000] *
               w[(b & 0x000f00) / 0x000100] * w[(b & 0x0000f0) / 0x000010] * w[(b & 0x00000f) / 0x000001];
    float na = w[(a & 0xf00000) / 0x100000] * w[(a & 0x0f0000) / 0x010000] * w[(a & 0x00f000) / 0x001000] *
               w[(a & 0x000f00) / 0x000100] * w[(a & 0x0000f0) / 0x000010] * w[(a & 0x00000f) / 0x000001];
    return nb / na;
}

void recttri_flip(const float* w, std::vector<TM>& gens, int32_t* a, int N) {  // recttrianglekernel.cl:25-118
    const int M = N / 3;
    for (int i = 0; i < N; i++) {
        for (int j = 0; j < M; j++) {
            const size_t id = (size_t)j + (size_t)M * i;
            float rd = tm_single01(gens[id]);
            float bs = 0.9;
            if (!(rd < bs)) continue;
            rd /= bs;
            int hexType = a[i * M + j];
            int newType = hexType;
            int hex1, hex2;
            float sc1 = (rd < 0.75f) ? 0.5f : 0.75f;
            float sc2 = (rd < 0.5f) ? 0.25f : sc1;
            float sc = (rd < 0.25f) ? 0.0f : sc2;
            const double r = (double)rd;
            if (hexType == 0x122133) {
                hex1 = (r < 0.75) ? 0x8ebbe8 : 0x331221;
                hex2 = (r < 0.5) ? 0xfc99cf : hex1;
                newType = (r < 0.25) ? 0xa7dd7a : hex2;
            } else if (hexType == 0x331221) {
                hex1 = (r < 0.75) ? 0x8ebbe8 : 0x122133;
                hex2 = (r < 0.5) ? 0xfc99cf : hex1;
                newType = (r < 0.25) ? 0xa7dd7a : hex2;
            } else if (hexType == 0x8ebbe8) {
                hex1 = (r < 0.75) ? 0x122133 : 0x331221;
                hex2 = (r < 0.5) ? 0xfc99cf : hex1;
                newType = (r < 0.25) ? 0xa7dd7a : hex2;
            } else if (hexType == 0xfc99cf) {
                hex1 = (r < 0.75) ? 0x8ebbe8 : 0x331221;
                hex2 = (r < 0.5) ? 0x122133 : hex1;
                newType = (r < 0.25) ? 0xa7dd7a : hex2;
            } else if (hexType == 0xa7dd7a) {
                hex1 = (r < 0.75) ? 0x8ebbe8 : 0x331221;
                hex2 = (r < 0.5) ? 0xfc99cf : hex1;
                newType = (r < 0.25) ? 0x122133 : hex2;
            } else if (hexType == 0x47d47a) { newType = 0xd22a33; sc = 1.0;
            } else if (hexType == 0xd22a33) { newType = 0x47d47a; sc = 1.0;
            } else if (hexType == 0xa74d74) { newType = 0x33a22d; sc = 1.0;
            } else if (hexType == 0x33a22d) { newType = 0xa74d74; sc = 1.0;
            } else if (hexType == 0x8eb558) { newType = 0x331eb1; sc = 1.0;
            } else if (hexType == 0x331eb1) { newType = 0x8eb558; sc = 1.0;
            } else if (hexType == 0x855be8) { newType = 0x1be133; sc = 1.0;
            } else if (hexType == 0x1be133) { newType = 0x855be8; sc = 1.0;
            } else if (hexType == 0x6699cf) { newType = 0xcf1221; sc = 1.0;
            } else if (hexType == 0xcf1221) { newType = 0x6699cf; sc = 1.0;
            } else if (hexType == 0xfc9966) { newType = 0x1221fc; sc = 1.0;
            } else if (hexType == 0x1221fc) { newType = 0xfc9966; sc = 1.0;
            }
            float weightratio = rt_ratio(w, hexType, newType);
            rd = (sc == 1.0f) ? rd : (rd - sc) * 4;
            a[i * M + j] = (weightratio >= rd) ? newType : hexType;
        }
    }
}

void recttri_update1(const int32_t* t1, int32_t* t2, int N, int t) {  // recttrianglekernel.cl:120-132
    const int M = N / 3;
    for (int i = 1; i < N - 1; i++) {
        const int p1 = (i % 3 == t), p2 = (i % 3 == (t + 1) % 3);
        for (int j = 1; j < M - 1; j++) {
            t2[i * M + j] = (t1[(i - 1) * M + j] & 0x0000f0) * 0x010000 + (t1[(i - 1) * M + j] & 0x00000f) * 0x010000 +
                            (t1[i * M + j + p1] & 0xf00000) / 0x000100 + (t1[(i + 1) * M + j - p2] & 0x0f0000) / 0x000100 +
                            (t1[(i + 1) * M + j - p2] & 0x00f000) / 0x000100 + (t1[i * M + j + p1] & 0x000f00) / 0x000100;
        }
    }
}

void recttri_update2(const int32_t* t1, int32_t* t2, int N, int t) {  // recttrianglekernel.cl:134-146
    const int M = N / 3;
    for (int i = 1; i < N - 1; i++) {
        const int p1 = (i % 3 == (t + 2) % 3), p2 = (i % 3 == (t + 1) % 3);
        for (int j = 1; j < M - 1; j++) {
            t2[i * M + j] = (t1[i * M + j - p1] & 0x00f000) * 0x000100 + (t1[(i - 1) * M + j + p2] & 0x000f00) * 0x000100 +
                            (t1[(i - 1) * M + j + p2] & 0x0000f0) * 0x000100 + (t1[i * M + j - p1] & 0x00000f) * 0x000100 +
                            (t1[(i + 1) * M + j] & 0xf00000) / 0x010000 + (t1[(i + 1) * M + j] & 0x0f0000) / 0x010000;
        }
    }
}

}  // namespace

extern "C" {

// ---- TinyMT known-answer helpers ----------------------------------------------------
// init(seed) then n tinymt32_uint32 outputs
void or_tinymt_uint32(uint32_t mat1, uint32_t mat2, uint32_t tmat, uint32_t seed, int n, uint32_t* out) {
    TM t{0, 0, 0, 0, mat1, mat2, tmat};
    tm_init(t, seed);
    for (int i = 0; i < n; i++) {
        tm_next(t);
        out[i] = tm_temper(t);
    }
}

// InitTinyMT over `count` ids then `draws` single01 draws per generator; writes k=u>>9
// of each draw (row-major [draw][id]) and the final 4-word state per id.
int or_tinymt_table_draws(const uint32_t* params, int64_t nparams, uint32_t seed, int64_t count, int draws,
                          uint32_t* k_out, uint32_t* state_out) {
    if (count > nparams) return -1;
    Table tb(params, nparams);
    tb.init(seed, count);
    for (int d = 0; d < draws; d++)
        for (int64_t id = 0; id < count; id++) {
            float f = tm_single01(tb.g[id]);
            if (k_out) k_out[(size_t)d * count + id] = (uint32_t)std::lround((double)f * 8388608.0);
        }
    if (state_out)
        for (int64_t id = 0; id < count; id++) {
            state_out[4 * id + 0] = tb.g[id].s0;
            state_out[4 * id + 1] = tb.g[id].s1;
            state_out[4 * id + 2] = tb.g[id].s2;
            state_out[4 * id + 3] = tb.g[id].s3;
        }
    return 0;
}

// ---- host schedule RNG anchors -------------------------------------------------------
uint64_t or_mt19937_nth(long seed, int n) {
    std::mt19937 mt(seed);
    uint64_t v = 0;
    for (int i = 0; i < n; i++) v = mt();
    return v;
}
uint64_t or_minstd0_nth(long seed, int n) {
    std::minstd_rand0 g;
    g.seed(seed);
    uint64_t v = 0;
    for (int i = 0; i < n; i++) v = g();
    return v;
}
void or_minstd0_uniform(long seed, int lo, int hi, int n, int32_t* out) {
    std::default_random_engine g;
    std::uniform_int_distribution<int> d(lo, hi);
    g.seed(seed);
    for (int i = 0; i < n; i++) out[i] = d(g);
}

// ---- Domino --------------------------------------------------------------------------
// DominoTiler::Walk(tiling&, long steps, long seed), DominoTiler.cpp:9-55
int or_domino_walk(int rule, const double* p12, int32_t* t, int N, long steps, long seed, const uint32_t* params,
                   int64_t nparams) {
    if (N < 6 || N % 2) return -1;
    const int64_t G = (int64_t)(N - 2) * (N / 2 - 2);
    if (G > nparams) return -2;
    DominoRule R{rule, {0.5, 0.5, 0.5, 0.5}};
    if (p12) for (int k = 0; k < 4; k++) R.p12[k] = p12[k];
    std::vector<uint8_t> B, Wc;
    domino_split(t, N, B, Wc);
    Table tb(params, nparams);
    std::mt19937 mt_rand(seed);
    tb.init((uint32_t)mt_rand(), G);  // :31-32, the first mt19937 output seeds TinyMT
    domino_steps(R, tb.g, B, Wc, N, mt_rand, steps);
    domino_merge(t, N, B, Wc);
    return 0;
}

// DominoTiler::Walk(tiling&, vector<long> steps, vector<long> seeds), DominoTiler.cpp:58-107
int or_domino_walk_epochs(int rule, const double* p12, int32_t* t, int N, const long* steps, const long* seeds, int K,
                          const uint32_t* params, int64_t nparams) {
    if (N < 6 || N % 2) return -1;
    const int64_t G = (int64_t)(N - 2) * (N / 2 - 2);
    if (G > nparams) return -2;
    DominoRule R{rule, {0.5, 0.5, 0.5, 0.5}};
    if (p12) for (int k = 0; k < 4; k++) R.p12[k] = p12[k];
    std::vector<uint8_t> B, Wc;
    domino_split(t, N, B, Wc);
    Table tb(params, nparams);
    for (int k = 0; k < K; ++k) {
        std::mt19937 mt_rand(seeds[k]);
        tb.init(low32(seeds[k]), G);  // :81-82, the epoch seed itself seeds TinyMT
        domino_steps(R, tb.g, B, Wc, N, mt_rand, steps[k]);
    }
    domino_merge(t, N, B, Wc);
    return 0;
}

// ---- Lozenge -------------------------------------------------------------------------
// LozengeTiler::Walk(tiling&, int steps, long seed), LozengeTiler.cpp:18-88
int or_lozenge_walk(int rule, double q, int32_t* t, int N, long steps, long seed, const uint32_t* params,
                    int64_t nparams) {
    if (N < 6 || N % 3) return -1;
    const int M = N / 3;
    const int64_t G = (int64_t)N * M;
    if (G > nparams) return -2;
    std::vector<uint8_t> cls[3];
    for (auto& v : cls) v.assign((size_t)N * M, 0);
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) cls[mod3(j - i)][i * M + j / 3] = (uint8_t)t[i * N + j];
    Table tb(params, nparams);
    tb.init(low32(seed), G);
    std::default_random_engine generator;
    std::uniform_int_distribution<int> distribution(0, 2);
    generator.seed(seed);
    for (long s = 0; s < steps; ++s) {
        int r = distribution(generator);
        lozenge_rotate(rule, q, tb.g, cls[r].data(), N);
        lozenge_update(cls[r].data(), cls[(r + 1) % 3].data(), cls[(r + 2) % 3].data(), N, r);
    }
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) t[i * N + j] = cls[mod3(j - i)][i * M + j / 3];
    return 0;
}

// ---- Rectangle-triangle -------------------------------------------------------------
// RectTriangleTiler::Walk(tiling&, int steps, long seed), RectTriangleTiler.cpp:17-92
int or_recttri_walk(const float* weights16, int32_t* t, int N, long steps, long seed, const uint32_t* params,
                    int64_t nparams) {
    if (N < 6 || N % 3) return -1;
    const int M = N / 3;
    const int64_t G = (int64_t)N * M;
    if (G > nparams) return -2;
    float w[16] = {0.0f, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1};  // recttrianglekernel.cl:3-13
    if (weights16) for (int k = 0; k < 16; k++) w[k] = weights16[k];
    std::vector<int32_t> cls[3];
    for (auto& v : cls) v.assign((size_t)N * M, 0);
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) cls[mod3(i - j)][i * M + j / 3] = t[i * N + j];
    Table tb(params, nparams);
    std::default_random_engine generator;
    std::uniform_int_distribution<int> distribution(0, 2);
    generator.seed(seed);
    tb.init(low32(seed), G);
    for (long s = 0; s < steps; ++s) {
        int r = distribution(generator);
        recttri_flip(w, tb.g, cls[r].data(), N);
        recttri_update1(cls[r].data(), cls[(r + 1) % 3].data(), N, r);
        recttri_update2(cls[r].data(), cls[(r + 2) % 3].data(), N, r);
    }
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) t[i * N + j] = cls[mod3(i - j)][i * M + j / 3];
    return 0;
}

// Wall-clock helper for the CPU-baseline leg: seconds for `steps` lozenge steps.
double or_lozenge_time(int rule, double q, int32_t* t, int N, long steps, long seed, const uint32_t* params,
                       int64_t nparams) {
    auto t0 = std::chrono::steady_clock::now();
    int rc = or_lozenge_walk(rule, q, t, N, steps, seed, params, nparams);
    auto t1 = std::chrono::steady_clock::now();
    if (rc) return -1.0;
    return std::chrono::duration<double>(t1 - t0).count();
}

}  // extern "C"
