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
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <cmath>
#include <random>
#include <vector>
#include <chrono>

#ifdef _OPENMP
#include <omp.h>
#endif

// Every loop over one kernel's NDRange below may run its rows on several threads: within one
// launch a kernel reads one array and writes another, or writes only its own site and
// generator, so the result does not depend on the order (or the threads) that visit the
// sites.  or_set_threads chooses the thread count (1 = a plain sequential loop).
// Small loops stay sequential (a parallel region costs more than they do).
#define OR_PRAGMA(x) _Pragma(#x)
#define OR_PAR_ROWS(work) \
    OR_PRAGMA(omp parallel for schedule(static) if (g_threads > 1 && (int64_t)(work) >= 32768) num_threads(g_threads))

namespace {

int g_threads = 1;

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
        const int64_t n = std::min<int64_t>(count, (int64_t)g.size());
        OR_PAR_ROWS(n)
        for (int64_t id = 0; id < n; id++)
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
    OR_PAR_ROWS((int64_t)G0 * G1)
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
    OR_PAR_ROWS((int64_t)N * W)
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
    OR_PAR_ROWS((int64_t)N * M)
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
    OR_PAR_ROWS((int64_t)N * M)
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
    OR_PAR_ROWS((int64_t)N * M)
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
    OR_PAR_ROWS((int64_t)N * M)
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
    OR_PAR_ROWS((int64_t)N * M)
    for (int i = 1; i < N - 1; i++) {
        const int p1 = (i % 3 == (t + 2) % 3), p2 = (i % 3 == (t + 1) % 3);
        for (int j = 1; j < M - 1; j++) {
            t2[i * M + j] = (t1[i * M + j - p1] & 0x00f000) * 0x000100 + (t1[(i - 1) * M + j + p2] & 0x000f00) * 0x000100 +
                            (t1[(i - 1) * M + j + p2] & 0x0000f0) * 0x000100 + (t1[i * M + j - p1] & 0x00000f) * 0x000100 +
                            (t1[(i + 1) * M + j] & 0xf00000) / 0x010000 + (t1[(i + 1) * M + j] & 0x0f0000) / 0x010000;
        }
    }
}

// ---------------------------------------------------------------------------------
// Triangle dimers (src/TriangleDimer/TriangleDimerTiler.cpp:86-483, triangledimerkernel.cl)
//
// Eight M x M int arrays (M = N/2): lozenge-move states eH, eL, eR; triangle-move states
// tU, tD; butterfly-move states bH, bL, bR.  Each function below is one kernel of
// triangledimerkernel.cl run as a sequential loop over its NDRange(M, M); every kernel
// writes only cells it does not read in the same launch, so the loop order is immaterial.
// ---------------------------------------------------------------------------------
struct TdArrays {
    int M = 0;
    std::vector<int32_t> eH, eL, eR, tU, tD, bH, bL, bR;
};

inline bool td_inner(int M, int i, int j) { return i > 0 && j > 0 && i < M - 1 && j < M - 1; }
inline int td_sel(bool c, int v) { return c ? v : 0; }

// RotateLozenges, triangledimerkernel.cl:4-22 (every work-item draws)
void td_rotate_lozenges(std::vector<TM>& g, int32_t* a, int M, int t, int c) {
    for (int i = 0; i < M; i++)
        for (int j = 0; j < M; j++) {
            const float rd = tm_single01(g[(size_t)i * M + j]);
            if (rd < .5 && ((1 - (t & 1)) * i + ((1 + t) / 2) * j) % 2 == c) {
                int32_t& v = a[(size_t)i * M + j];
                if (v == 5) v = 10;
                else if (v == 10) v = 5;
            }
        }
}

// UpdateLozengesFlipped, :24-34
void td_update_lozenges_flipped(int32_t* a, int M, int t, int c) {
    for (int i = 0; i < M; i++)
        for (int j = 0; j < M; j++) {
            if (!td_inner(M, i, j) || ((1 - (t & 1)) * i + ((1 + t) / 2) * j) % 2 == c) continue;
            const int k = (2 - t) / 2, s = t & 1;
            a[i * M + j] = (a[(i - 1) * M + j + s] & 4) / 4 + (a[(i - k) * M + j + 1] & 8) / 4 +
                           (a[(i + 1) * M + j - s] & 1) * 4 + (a[(i + k) * M + j - 1] & 2) * 4;
        }
}

// UpdateLozenges0/1/2, :36-81: after flipping orientation `t`, rewrite the bits of the other
// two orientations that the flipped ones determine
void td_update_lozenges(int32_t* a, int32_t* b, int32_t* c, int M, int t) {
    for (int i = 1; i < M - 1; i++)
        for (int j = 1; j < M - 1; j++) {
            const int x = i * M + j;
            if (t == 0) {
                b[x] &= ~10;
                b[x] |= (a[(i + 1) * M + j] & 1) * 2 + (a[(i + 1) * M + j - 1] & 1) * 8;
                c[x] &= ~10;
                c[x] |= (a[(i + 1) * M + j] & 2) + (a[x] & 8);
            } else if (t == 1) {
                b[x] &= ~5;
                b[x] |= (a[x] & 1) + (a[(i + 1) * M + j] & 1) * 4;
                c[x] &= ~5;
                c[x] |= (a[(i - 1) * M + j] & 2) / 2 + (a[x] & 2) * 2;
            } else {
                b[x] &= ~10;
                b[x] |= (a[(i - 1) * M + j] & 2) + (a[x - 1] & 2) * 4;
                c[x] &= ~5;
                c[x] |= (a[x] & 1) + (a[x - 1] & 4);
            }
        }
}

// UpdateTriangleUFromLozenges, :100-121
void td_triangle_u_from_lozenges(const int32_t* H, const int32_t* L, int32_t* U, int M) {
    for (int i = 1; i < M - 1; i++)
        for (int j = 1; j < M - 1; j++) {
            const int32_t l0 = L[i * M + j], l1 = L[i * M + j + 1], h0 = H[i * M + j], h1 = H[(i + 1) * M + j];
            const int L1 = td_sel((l0 & 1) == 1, 1), L2 = td_sel((h0 & 8) == 8, 2), L3 = td_sel((h1 & 1) == 1, 3);
            const int R1 = td_sel((l1 & 1) == 1, 1), R2 = td_sel((h1 & 2) == 2, 2), R3 = td_sel((l1 & 2) == 2, 3);
            const int V1 = td_sel((l1 & 4) == 4, 1), V2 = td_sel((h1 & 8) == 8, 2), V3 = td_sel((h1 & 4) == 4, 3);
            U[i * M + j] = L1 + L2 + L3 + 4 * (R1 + R2 + R3) + 16 * (V1 + V2 + V3);
        }
}

// UpdateButterflys{H,L,R}FromLozenge, :123-190
void td_butterflies_from_lozenges(int which, int32_t* B, const int32_t* LH, const int32_t* LL, const int32_t* LR,
                                  int M) {
    for (int i = 1; i < M - 1; i++)
        for (int j = 1; j < M - 1; j++) {
            const int x = i * M + j;
            int c, t1, t4, t16, t64;
            if (which == 0) {
                const int lh = LH[x], lr1 = LR[x - M], ll1 = LL[x - M + 1], lr2 = LR[x], ll2 = LL[x];
                c = (ll1 & 4) / 4;
                t1 = (lr1 & 8) / 8 + 2 * (lr1 & 1) + 3 * (lh & 1);
                t4 = (ll1 & 1) + (ll1 & 2) + ((lh & 2) / 2) * 3;
                t16 = (lr2 & 2) / 2 + (lr2 & 4) / 2 + ((lh & 4) / 4) * 3;
                t64 = (ll2 & 4) / 4 + (ll2 & 8) / 4 + ((lh & 8) / 8) * 3;
            } else if (which == 1) {
                const int ll = LL[x], lh1 = LH[x], lr1 = LR[x], lh2 = LH[x + M - 1], lr2 = LR[x - 1];
                c = (lr1 & 8) / 8;
                t1 = (lh1 & 1) + (lh1 & 2) + 3 * (ll & 1);
                t4 = (lr1 & 2) / 2 + (lr1 & 4) / 2 + ((ll & 2) / 2) * 3;
                t16 = (lh2 & 4) / 4 + (lh2 & 8) / 4 + ((ll & 4) / 4) * 3;
                t64 = (lr2 & 8) / 8 + 2 * (lr2 & 1) + ((ll & 8) / 8) * 3;
            } else {
                const int lr = LR[x], lh1 = LH[x], ll1 = LL[x + 1], lh2 = LH[x + M], ll2 = LL[x];
                c = (ll1 & 8) / 8;
                t1 = (lh1 & 1) + (lh1 & 2) + 3 * (lr & 1);
                t4 = (ll1 & 1) + (ll1 & 2) + ((lr & 2) / 2) * 3;
                t16 = (lh2 & 4) / 4 + (lh2 & 8) / 4 + ((lr & 4) / 4) * 3;
                t64 = (ll2 & 4) / 4 + (ll2 & 8) / 4 + ((lr & 8) / 8) * 3;
            }
            B[x] = t1 + 4 * t4 + 16 * t16 + 64 * t64 + 256 * c;
        }
}

// RotateTriangles, :192-212
void td_rotate_triangles(std::vector<TM>& g, int32_t* a, int M, int c) {
    for (int i = 0; i < M; i++)
        for (int j = 0; j < M; j++) {
            const float rd = tm_single01(g[(size_t)i * M + j]);
            if (rd < 0.5 && mod3(j - i) == c) {
                int32_t& v = a[(size_t)i * M + j];
                if (v == 45) v = 54;
                else if (v == 54) v = 45;
            }
        }
}

// UpdateTrianglesFlipped0 (:214-234, up) and UpdateTrianglesFlipped1 (:236-256, down)
void td_update_triangles_flipped(int32_t* a, int M, int c, bool down) {
    for (int i = 1; i < M - 1; i++)
        for (int j = 1; j < M - 1; j++) {
            const int cl = mod3(j - i);
            if (cl == c) continue;
            const int p = (cl == (c + 1) % 3) ? 1 : 0;
            const int t1 = (4 - 3 * p) * 12, m1 = (4 - 3 * p) * 4;
            const int t4 = (15 * p + 1) * 3, m4 = 15 * p + 1;
            const int t16 = (4 - 3 * p) * 3, m16 = 4 - 3 * p;
            int x1, x4, x16;
            if (!down) {
                x1 = a[(i - (1 - p)) * M + j - p];
                x4 = a[(i - p) * M + j + 1];
                x16 = a[(i + 1) * M + j - 1 + p];
            } else {
                x1 = a[(i + 1 - p) * M + j - 1];
                x4 = a[(i + p) * M + j + 1 - p];
                x16 = a[(i - 1) * M + j + p];
            }
            a[i * M + j] = (x1 & t1) / m1 + ((x4 & t4) / m4) * 4 + ((x16 & t16) / m16) * 16;
        }
}

// UpdateTriangles, :258-281: t = 0 writes down triangles from up ones, t = 1 the reverse
void td_update_triangles(const int32_t* s, int32_t* d, int M, int t) {
    const int p = (t == 0) ? -1 : 1;
    for (int i = 1; i < M - 1; i++)
        for (int j = 1; j < M - 1; j++) {
            const int A = s[i * M + j - 1 + t], B = s[i * M + j + t], V = s[(i + p) * M + j];
            auto dig = [](int v) { return td_sel((v & 48) == 16, 1) + td_sel((v & 3) == 3, 2) + td_sel((v & 12) == 8, 3); };
            d[i * M + j] = dig(A) + 4 * dig(B) + 16 * dig(V);
        }
}

// UpdateLozenge{H,L,R}FromTriangles, :283-328
void td_lozenges_from_triangles(const int32_t* U, const int32_t* D, int32_t* H, int32_t* L, int32_t* R, int M) {
    for (int i = 1; i < M - 1; i++)
        for (int j = 1; j < M - 1; j++) {
            const int x = i * M + j;
            {
                const int d = D[x], u = U[x - M];
                H[x] = td_sel((d & 48) == 32, 1) + 2 * td_sel((d & 48) == 48, 1) + 4 * td_sel((u & 48) == 48, 1) +
                       8 * td_sel((u & 48) == 32, 1);
            }
            {
                const int d = D[x], u = U[x - 1];
                L[x] = td_sel((u & 12) == 4, 1) + 2 * td_sel((u & 12) == 12, 1) + 4 * td_sel((d & 3) == 1, 1) +
                       8 * td_sel((d & 3) == 2, 1);
            }
            {
                const int d = D[x], u = U[x];
                R[x] = td_sel((u & 3) == 1, 1) + 2 * td_sel((d & 12) == 12, 1) + 4 * td_sel((d & 12) == 4, 1) +
                       8 * td_sel((u & 3) == 2, 1);
            }
        }
}

// RotateButterflys, :330-352
void td_rotate_butterflies(std::vector<TM>& g, int32_t* a, int M, int t, int p1, int p2) {
    const int b1 = 1 - (t & 1), b2 = (1 + t) / 2, b3 = t & 1;
    for (int i = 0; i < M; i++)
        for (int j = 0; j < M; j++) {
            const float rd = tm_single01(g[(size_t)i * M + j]);
            if (rd < 0.5 && (b1 * i + b2 * j) % 3 == p1) {
                int32_t& v = a[(size_t)i * M + j];
                const bool par = ((b1 * j + b3 * i) & 1) == p2;
                if (v == 170) v = par ? 85 : 170;
                else if (v == 85) v = par ? 170 : 85;
            }
        }
}

inline int td_pack(int t1, int t4, int t16, int t64, int c) { return t1 + 4 * t4 + 16 * t16 + 64 * t64 + 256 * c; }

// UpdateButterflysFlipped{H,L,R}1, :355-375, :454-472, :552-572: the flipped slice's other
// parity, from its two neighbours along the slice
void td_bfly_flipped1(int32_t* a, int M, int t, int p1, int p2) {
    for (int i = 1; i < M - 1; i++)
        for (int j = 1; j < M - 1; j++) {
            const int x = i * M + j;
            const int c = (a[x] & 256) / 256;
            int t1, t4, t16, t64;
            if (t == 0 || t == 1) {
                if (t == 0 && !(i % 3 == p1 && (j & 1) != p2)) continue;
                if (t == 1 && !(j % 3 == p1 && (i & 1) != p2)) continue;
                const int B1 = (t == 0) ? a[x - 1] : a[x - M];
                const int B2 = (t == 0) ? a[x + 1] : a[x + M];
                t1 = ((B1 & 12) / 4 == 0) ? 0 : ((B1 & 12) / 4) % 3 + 1;
                t4 = ((B2 & 3) == 0) ? 0 : ((B2 & 3) + 1) % 3 + 1;
                t16 = ((B2 & 192) / 64 == 0) ? 0 : ((B2 & 192) / 64) % 3 + 1;
                t64 = ((B1 & 48) / 16 == 0) ? 0 : ((B1 & 48) / 16 + 1) % 3 + 1;
            } else {
                if (!((i + j) % 3 == p1 && (j & 1) != p2)) continue;
                const int B1 = a[x + M - 1], B2 = a[x - M + 1];
                t1 = ((B2 & 192) / 64 == 0) ? 0 : ((B2 & 192) / 64 + 1) % 3 + 1;
                t4 = ((B2 & 48) / 16 == 0) ? 0 : ((B2 & 48) / 16) % 3 + 1;
                t16 = ((B1 & 12) / 4 == 0) ? 0 : ((B1 & 12) / 4 + 1) % 3 + 1;
                t64 = ((B1 & 3) == 0) ? 0 : (B1 & 3) % 3 + 1;
            }
            a[x] = td_pack(t1, t4, t16, t64, c);
        }
}

// slice index of butterfly cell (i, j) for orientation t: rows (H), columns (L), i+j (R)
inline int td_slice(int t, int i, int j) { return t == 0 ? i % 3 : t == 1 ? j % 3 : (i + j) % 3; }

// UpdateButterflysFlipped{H,L}21 (:377-398, :474-495) and R21 (:574-595)
void td_bfly_flipped21(int32_t* a, int M, int t, int p1) {
    const int want = (t == 0) ? (p1 + 1) % 3 : (p1 + 2) % 3;
    for (int i = 1; i < M - 1; i++)
        for (int j = 1; j < M - 1; j++) {
            if (td_slice(t, i, j) != want) continue;
            const int x = i * M + j, Bf = a[x];
            if (t == 2) {
                const int B2 = a[x + 1], B3 = a[x + M];
                const int c = td_sel((B3 & 3) == 1, 1);
                const int t4 = ((Bf & 12) / 4 == 2) ? (Bf & 12) / 4 : (B2 & 3) / 3 + ((B2 & 192) / 192) * 3;
                const int t16 = ((Bf & 48) / 16 == 1) ? (Bf & 48) / 16 : ((B3 & 3) / 3) * 3 + ((B3 & 192) / 192) * 2;
                a[x] = td_pack(Bf & 3, t4, t16, (Bf & 192) / 64, c);
            } else {
                const int B1 = (t == 0) ? a[x - M] : a[x - M + 1];
                const int B2 = (t == 0) ? a[x - M + 1] : a[x + 1];
                const int c = td_sel((B1 & 48) / 16 == 2, 1);
                const int t1 = ((Bf & 3) == 2) ? (Bf & 3) : (B1 & 192) / 192 + ((B1 & 48) / 48) * 3;
                const int t4 = ((Bf & 12) / 4 == 1) ? (Bf & 12) / 4 : ((B2 & 192) / 192) * 3 + ((B2 & 48) / 48) * 2;
                a[x] = td_pack(t1, t4, (Bf & 48) / 16, (Bf & 192) / 64, c);
            }
        }
}

// UpdateButterflysFlipped{H,L,R}22, :400-428, :497-527, :597-627
void td_bfly_flipped22(int32_t* a, int M, int t, int p1) {
    const int want = (t == 0) ? (p1 + 2) % 3 : (p1 + 1) % 3;
    for (int i = 1; i < M - 1; i++)
        for (int j = 1; j < M - 1; j++) {
            if (td_slice(t, i, j) != want) continue;
            const int x = i * M + j, Bf = a[x];
            if (t == 2) {
                const int s4 = td_sel((Bf & 12) / 4 != 2, 1), s16 = td_sel((Bf & 48) / 16 != 1, 1);
                const int B1 = a[x - M], B2 = a[x + 1], B3 = a[x + M], B4 = a[x - 1];
                const int c = td_sel((B1 & 48) / 16 == 1, 1);
                const int t1 = ((Bf & 3) == 1) ? (Bf & 3) : ((B1 & 48) / 48) * 3 + ((B1 & 12) / 12) * 2;
                const int t4 = ((B2 & 256) / 256 == 1) ? 2 : s4 * (Bf & 12) / 4;
                const int t16 = ((B3 & 256) / 256 == 1) ? 1 : s16 * (Bf & 48) / 16;
                const int t64 = ((Bf & 192) / 64 == 2) ? (Bf & 192) / 64 : (B4 & 48) / 48 + ((B4 & 12) / 12) * 3;
                a[x] = td_pack(t1, t4, t16, t64, c);
            } else {
                const int s1 = td_sel((Bf & 3) != 2, 1), s4 = td_sel((Bf & 12) / 4 != 1, 1);
                const int B1 = (t == 0) ? a[x - M] : a[x - M + 1];
                const int B2 = (t == 0) ? a[x - M + 1] : a[x + 1];
                const int B3 = (t == 0) ? a[x + M] : a[x + M - 1];
                const int B4 = (t == 0) ? a[x + M - 1] : a[x - 1];
                const int c = td_sel((B3 & 3) == 2, 1);
                const int t1 = ((B1 & 256) == 256) ? 2 : s1 * (Bf & 3);
                const int t4 = ((B2 & 256) == 256) ? 1 : s4 * ((Bf & 12) / 4);
                const int t16 = ((Bf & 48) / 16 == 2) ? (Bf & 48) / 16 : ((B3 & 3) / 3) * 3 + (B3 & 12) / 12;
                const int t64 = ((Bf & 192) / 64 == 1) ? (Bf & 192) / 64 : ((B4 & 3) / 3) * 2 + ((B4 & 12) / 12) * 3;
                a[x] = td_pack(t1, t4, t16, t64, c);
            }
        }
}

// UpdateButterflysFlipped{H,L,R}23, :430-452, :529-550, :629-651
void td_bfly_flipped23(int32_t* a, int M, int t, int p1) {
    const int want = (t == 0) ? (p1 + 1) % 3 : (p1 + 2) % 3;
    for (int i = 1; i < M - 1; i++)
        for (int j = 1; j < M - 1; j++) {
            if (td_slice(t, i, j) != want) continue;
            const int x = i * M + j, Bf = a[x];
            const int c = (Bf & 256) / 256;
            if (t == 2) {
                const int s1 = td_sel((Bf & 3) != 1, 1), s64 = td_sel((Bf & 192) / 64 != 2, 1);
                const int B1 = a[x - M], B4 = a[x - 1];
                const int t1 = ((B1 & 256) / 256 == 1) ? 1 : s1 * (Bf & 3);
                const int t64 = ((B4 & 256) / 256 == 1) ? 2 : s64 * (Bf & 192) / 64;
                a[x] = td_pack(t1, (Bf & 12) / 4, (Bf & 48) / 16, t64, c);
            } else {
                const int s16 = td_sel((Bf & 48) / 16 != 2, 1), s64 = td_sel((Bf & 192) / 64 != 1, 1);
                const int B3 = (t == 0) ? a[x + M] : a[x + M - 1];
                const int B4 = (t == 0) ? a[x + M - 1] : a[x - 1];
                const int t16 = ((B3 & 256) == 256) ? 2 : s16 * ((Bf & 48) / 16);
                const int t64 = ((B4 & 256) == 256) ? 1 : s64 * ((Bf & 192) / 64);
                a[x] = td_pack(Bf & 3, (Bf & 12) / 4, t16, t64, c);
            }
        }
}

// UpdateLozengeFromButterflys{H,L,R}, :654-738
void td_lozenges_from_butterflies(int t, const int32_t* B, int32_t* LH, int32_t* LL, int32_t* LR, int M) {
    auto f = [](bool b) { return b ? 1 : 0; };
    for (int i = 1; i < M - 1; i++)
        for (int j = 1; j < M - 1; j++) {
            const int x = i * M + j, F = B[x];
            const int full = f((F & 3) == 3) + 2 * f((F & 12) == 12) + 4 * f((F & 48) == 48) + 8 * f((F & 192) == 192);
            if (t == 0) {
                LH[x] = full;
                LL[x] = f((F & 256) == 256) + 2 * f((F & 48) == 48) + 4 * f((F & 192) == 64) + 8 * f((F & 192) == 128);
                LR[x] = f((F & 256) == 256) + 2 * f((F & 48) == 16) + 4 * f((F & 48) == 32) + 8 * f((F & 192) == 192);
            } else if (t == 1) {
                LL[x] = full;
                LR[x] = f((F & 3) == 3) + 2 * f((F & 12) == 4) + 4 * f((F & 12) == 8) + 8 * f((F & 256) == 256);
                LH[x] = f((F & 3) == 1) + 2 * f((F & 3) == 2) + 4 * f((F & 12) == 12) + 8 * f((F & 256) == 256);
            } else {
                LR[x] = full;
                LH[x] = f((F & 3) == 1) + 2 * f((F & 3) == 2) + 4 * f((F & 256) == 256) + 8 * f((F & 192) == 192);
                LL[x] = f((F & 3) == 3) + 2 * f((F & 256) == 256) + 4 * f((F & 192) == 64) + 8 * f((F & 192) == 128);
            }
        }
}

// TilingToTriangleFlips, TriangleDimerTiler.cpp:600-631 (tiling side N, output N/2 x N)
std::vector<int32_t> td_triangle_flips(const int32_t* t, int N) {
    std::vector<int32_t> o((size_t)(N / 2) * N, 0);
    for (int i = 1; i < N / 2 - 1; i++)
        for (int j = 1; j < N - 1; j++) {
            const int a = t[(2 * i + 1) * N + j], b = t[(2 * i + 1) * N + j + 1];
            if (j % 2 == 0) {
                const int c = t[2 * i * N + j + 1];
                o[i * N + j] = ((a & 4) / 4 + (a & 8) / 4 + ((b & 8) / 8) * 3) + 4 * ((b & 4) / 4 + (a & 2) + ((b & 2) / 2) * 3) +
                               16 * ((c & 1) * 2 + (a & 1) + ((c & 2) / 2) * 3);
            } else {
                const int d = t[(2 * i + 2) * N + j];
                o[i * N + j] = ((a & 1) + (a & 8) / 4 + ((b & 8) / 8) * 3) + 4 * ((b & 1) + (a & 2) + ((b & 2) / 2) * 3) +
                               16 * ((d & 8) / 4 + ((d & 4) / 4) * 3 + (a & 4) / 4);
            }
        }
    return o;
}

// TilingToButterflyFlips, TriangleDimerTiler.cpp:633-680
std::vector<int32_t> td_butterfly_flips(const int32_t* t, int N) {
    std::vector<int32_t> o((size_t)N * N, 0);
    for (int i = 2; i < N - 2; i++)
        for (int j = 2; j < N - 2; j++) {
            const int x = i * N + j, s = t[x];
            int f1, f4, f16, f64, c;
            if (i % 2 == 0) {
                if (j % 2 == 0) continue;
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
            o[x] = f1 + 4 * f4 + 16 * f16 + 64 * f64 + 256 * c;
        }
    return o;
}

// One step of TriangleDimerTiler::Walk (:217-338) given the host draws.
void td_step(TdArrays& s, std::vector<TM>& g, int type, int r1, int r2, int p1, int p2) {
    const int M = s.M;
    int32_t *eH = s.eH.data(), *eL = s.eL.data(), *eR = s.eR.data(), *tU = s.tU.data(), *tD = s.tD.data();
    int32_t *bH = s.bH.data(), *bL = s.bL.data(), *bR = s.bR.data();
    if (type == 0) {
        int32_t* e[3] = {eH, eL, eR};
        td_rotate_lozenges(g, e[r1], M, r1, r2);
        td_update_lozenges_flipped(e[r1], M, r1, r2);
        td_update_lozenges(e[r1], e[(r1 + 1) % 3], e[(r1 + 2) % 3], M, r1);
        td_triangle_u_from_lozenges(eH, eL, tU, M);
        td_update_triangles(tU, tD, M, 0);
        td_butterflies_from_lozenges(0, bH, eH, eL, eR, M);
        td_butterflies_from_lozenges(1, bL, eH, eL, eR, M);
        td_butterflies_from_lozenges(2, bR, eH, eL, eR, M);
    } else if (type == 1) {
        int32_t* a = (r2 == 1) ? tD : tU;
        td_rotate_triangles(g, a, M, r1);
        td_update_triangles_flipped(a, M, r1, r2 == 1);
        if (r2 == 1) td_update_triangles(tD, tU, M, 1);
        else td_update_triangles(tU, tD, M, 0);
        td_lozenges_from_triangles(tU, tD, eH, eL, eR, M);
        td_butterflies_from_lozenges(0, bH, eH, eL, eR, M);
        td_butterflies_from_lozenges(1, bL, eH, eL, eR, M);
        td_butterflies_from_lozenges(2, bR, eH, eL, eR, M);
    } else {
        int32_t* b[3] = {bH, bL, bR};
        td_rotate_butterflies(g, b[r1], M, r1, p1, p2);
        td_bfly_flipped1(b[r1], M, r1, p1, p2);
        td_bfly_flipped21(b[r1], M, r1, p1);
        td_bfly_flipped22(b[r1], M, r1, p1);
        td_bfly_flipped23(b[r1], M, r1, p1);
        td_lozenges_from_butterflies(r1, b[r1], eH, eL, eR, M);
        for (int o = 0; o < 3; o++)
            if (o != r1) td_butterflies_from_lozenges(o, b[o], eH, eL, eR, M);
        td_triangle_u_from_lozenges(eH, eL, tU, M);
        td_update_triangles(tU, tD, M, 0);
    }
}

}  // namespace

extern "C" {

// Threads the per-kernel loops may use (1: sequential).  Returns the previous count.
int or_set_threads(int n) {
    const int old = g_threads;
    g_threads = n < 1 ? 1 : n;
    return old;
}

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

// Coupling from the past as Examples/AztecDiamondCftpEx.cpp:42-68 runs it: epoch seeds from
// std::mt19937(seed); each restart walks copies of tmax and tmin through all epochs (oldest
// first) and compares them; mode 0 prepends an epoch of twice the last epoch's length
// (:67), mode 1 one as long as all epochs together.  Returns 0 and the sample, or 1 when
// max_restarts restarts did not coalesce.
int or_domino_cftp(int rule, const double* p12, const int32_t* tmax, const int32_t* tmin, int N, long seed, int mode,
                   long first_steps, int max_restarts, const uint32_t* params, int64_t nparams, int32_t* out,
                   long* total_steps, int* restarts) {
    std::mt19937 mt_rand((std::mt19937::result_type)seed);
    std::vector<long> steps{first_steps}, seeds{(long)mt_rand()};
    const size_t NN = (size_t)N * N;
    std::vector<int32_t> top(NN), bottom(NN);
    for (int r = 0; r <= max_restarts; r++) {
        std::memcpy(top.data(), tmax, NN * 4);
        std::memcpy(bottom.data(), tmin, NN * 4);
        int rc = or_domino_walk_epochs(rule, p12, bottom.data(), N, steps.data(), seeds.data(), (int)steps.size(),
                                       params, nparams);
        if (!rc) rc = or_domino_walk_epochs(rule, p12, top.data(), N, steps.data(), seeds.data(), (int)steps.size(),
                                            params, nparams);
        if (rc) return rc;
        long total = 0;
        for (long s : steps) total += s;
        if (top == bottom) {
            std::memcpy(out, top.data(), NN * 4);
            if (total_steps) *total_steps = total;
            if (restarts) *restarts = r;
            return 0;
        }
        const long len = (mode == 0) ? 2 * steps.back() : total;
        steps.insert(steps.begin(), len);
        seeds.insert(seeds.begin(), (long)mt_rand());
    }
    return 1;
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

extern "C" {

// ---- Triangle dimers -------------------------------------------------------------------
// TriangleDimerTiler::Walk(tiling&, int steps, long seed), TriangleDimerTiler.cpp:86-483:
// split the N x N lozenge-state tiling into eH (even rows, odd columns), eL (odd, even) and
// eR (odd, odd), derive the triangle and butterfly states, and per step draw
// type = d3, r1 = d3, r2 = d2 (and p1 = d3, p2 = d2 for butterfly moves) from one
// minstd_rand0 seeded with `seed`.  Only the lozenge states are written back.
// If `moves` is non-null it receives type | r1<<2 | r2<<4 | p1<<5 | p2<<7 per step.
int or_tridimer_walk(int32_t* t, int N, long steps, long seed, const uint32_t* params, int64_t nparams,
                     uint8_t* moves) {
    if (N < 8 || N % 2) return -1;
    const int M = N / 2;
    const int64_t G = (int64_t)M * M;
    if (G > nparams) return -2;
    TdArrays s;
    s.M = M;
    for (auto* v : {&s.eH, &s.eL, &s.eR, &s.tU, &s.tD, &s.bH, &s.bL, &s.bR}) v->assign((size_t)G, 0);
    const std::vector<int32_t> tri = td_triangle_flips(t, N);
    const std::vector<int32_t> bf = td_butterfly_flips(t, N);
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            const size_t k = (size_t)(i / 2) * M + j / 2;
            if (i % 2 == 0) {
                if (j % 2 == 1) s.eH[k] = t[i * N + j], s.bH[k] = bf[i * N + j];
            } else if (j % 2 == 0) {
                s.eL[k] = t[i * N + j], s.bL[k] = bf[i * N + j];
            } else {
                s.eR[k] = t[i * N + j], s.bR[k] = bf[i * N + j];
            }
        }
    for (int i = 0; i < N / 2; i++)
        for (int j = 0; j < N; j++) (j % 2 == 0 ? s.tD : s.tU)[(size_t)i * M + j / 2] = tri[(size_t)i * N + j];
    Table tb(params, nparams);
    std::default_random_engine generator;
    std::uniform_int_distribution<int> d1(0, 2);
    std::uniform_int_distribution<int> d2(0, 1);
    generator.seed(seed);
    tb.init(low32(seed), G);
    for (long st = 0; st < steps; st++) {
        const int type = d1(generator);
        const int r1 = d1(generator);
        const int r2 = d2(generator);
        int p1 = 0, p2 = 0;
        if (type == 2) {
            p1 = d1(generator);
            p2 = d2(generator);
        }
        if (moves) moves[st] = (uint8_t)(type | r1 << 2 | r2 << 4 | p1 << 5 | p2 << 7);
        td_step(s, tb.g, type, r1, r2, p1, p2);
    }
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            const size_t k = (size_t)(i / 2) * M + j / 2;
            if (i % 2 == 0) {
                if (j % 2 == 1) t[i * N + j] = s.eH[k];
            } else {
                t[i * N + j] = (j % 2 == 0) ? s.eL[k] : s.eR[k];
            }
        }
    return 0;
}

// The derived triangle / butterfly arrays of a tiling (TilingToTriangleFlips,
// TilingToButterflyFlips), for tests of the product's statics.
void or_tridimer_flips(const int32_t* t, int N, int32_t* tri, int32_t* bfly) {
    const std::vector<int32_t> a = td_triangle_flips(t, N), b = td_butterfly_flips(t, N);
    std::memcpy(tri, a.data(), a.size() * 4);
    std::memcpy(bfly, b.data(), b.size() * 4);
}

}  // extern "C"
