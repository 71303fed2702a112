// TriangleDimerTiler.cpp — the triangle-dimer tiler over the HIP library.
//
// Walk follows TriangleDimerTiler.cpp:86-483 of the reference and runs through
// rt_tridimer_walk (the triangle and butterfly states are derived on the device).  The
// statics reproduce the reference's outputs: IceCreamCone/TestButterfly (:485-598), the
// move-state derivations (:600-680), vertices/domain/dimer indicators (:19-84, :711-770) and
// the SVG documents (:772-905).
#include "TriangleDimerTiler.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "../common/rtapi.hpp"
#include "rt_hip.h"

namespace {

inline int bit(int x, int k) { return (x >> k) & 1; }

// Dimer indicators on the edges of the triangular lattice: h[a][b] joins vertex (a,b) to
// (a,b+1), l[a][b] joins (a,b) to (a+1,b), r[a][b] joins (a,b+1) to (a+1,b).
struct Dimers {
    int E;
    std::vector<int> h, l, r;
    explicit Dimers(int e) : E(e), h((size_t)e * e, 0), l((size_t)e * e, 0), r((size_t)e * e, 0) {}
    int& H(int a, int b) { return h[(size_t)a * E + b]; }
    int& L(int a, int b) { return l[(size_t)a * E + b]; }
    int& R(int a, int b) { return r[(size_t)a * E + b]; }

    // Lozenge-move states of every rhombus (tiling side 2E): bits 1, 2, 4, 8 are the
    // rhombus' sides in the order the reference's kernels read them.
    tiling states() {
        const int T = 2 * E;
        tiling t((size_t)T * T, 0);
        for (int i = 3; i < T - 1; ++i)
            for (int j = 3; j < T - 1; ++j) {
                const int a = i / 2, b = j / 2;
                int s;
                if (i % 2 == 0) {
                    if (j % 2 == 0) continue;
                    s = R(a - 1, b) + 2 * L(a - 1, b + 1) + 4 * R(a, b) + 8 * L(a, b);  // H rhombus
                } else if (j % 2 == 0) {
                    s = H(a, b) + 2 * R(a, b) + 4 * H(a + 1, b - 1) + 8 * R(a, b - 1);  // L rhombus
                } else {
                    s = H(a, b) + 2 * L(a, b + 1) + 4 * H(a + 1, b) + 8 * L(a, b);  // R rhombus
                }
                t[(size_t)i * T + j] = s;
            }
        return t;
    }
};

// A digit of a derived move state: the sum of weight w over the listed (source, bit) pairs
// that are set.
struct Term {
    int src, bit, w;
};

// Triangle-move state digits (left, right, vertical) of a down (even column) or up (odd
// column) triangle, over the sources a = t[2i+1][j], b = t[2i+1][j+1], c = t[2i][j+1],
// d = t[2i+2][j].
constexpr Term kTriDown[3][3] = {{{0, 2, 1}, {0, 3, 2}, {1, 3, 3}},
                                 {{1, 2, 1}, {0, 1, 2}, {1, 1, 3}},
                                 {{0, 0, 1}, {2, 0, 2}, {2, 1, 3}}};
constexpr Term kTriUp[3][3] = {{{0, 0, 1}, {0, 3, 2}, {1, 3, 3}},
                               {{1, 0, 1}, {0, 1, 2}, {1, 1, 3}},
                               {{0, 2, 1}, {3, 3, 2}, {3, 2, 3}}};

// Butterfly-move states: digit k (weight 4^k) = bit lo + 2 * bit hi of a neighbouring
// rhombus state + 3 * bit k of the centre state; the centre flag (weight 256) is one bit of
// a neighbour.  Per orientation H, L, R.
struct BflyDigit {
    int di, dj, lo, hi;
};
struct BflyForm {
    BflyDigit d[4];
    int cdi, cdj, cbit;
};
constexpr BflyForm kBfly[3] = {
    {{{-1, 0, 3, 0}, {-1, 1, 0, 1}, {1, 0, 1, 2}, {1, -1, 2, 3}}, -1, 1, 2},
    {{{-1, 1, 0, 1}, {0, 1, 1, 2}, {1, -1, 2, 3}, {0, -1, 3, 0}}, 0, 1, 3},
    {{{-1, 0, 0, 1}, {0, 1, 0, 1}, {1, 0, 2, 3}, {0, -1, 2, 3}}, 0, 1, 3},
};

// The six lozenge-state bits that mark a dimer at vertex (i, j) of the (N/2+1)^2 vertex
// array: (row, column, bit) relative to the tiling.
bool vertex_covered(const tiling& t, int N, int i, int j) {
    const int probes[6][3] = {{2 * (i - 1), 2 * j + 1, 2}, {2 * (i - 1), 2 * j + 1, 3}, {2 * (i + 1), 2 * j - 1, 1},
                              {2 * (i + 1), 2 * j - 1, 0}, {2 * i - 1, 2 * j, 2},       {2 * i - 1, 2 * j + 1, 2}};
    for (const auto& p : probes)
        if (bit(t[(size_t)p[0] * N + p[1]], p[2])) return true;
    return false;
}

const char* const kSvgHead =
    "<svg xmlns='http://www.w3.org/2000/svg' xmlns:xlink=\"http://www.w3.org/1999/xlink\" height=\"433.013\" "
    "width=\"500\" viewBox=\"0 0 ";

void svg_open(std::ostream& o, int M) {
    o << kSvgHead << 3 / 2.0 * M << " " << 3 / 2.0 * M << "\" preserveAspectRatio=\"none\">\n";
}

void use(std::ostream& o, const char* id, int i, int j) {
    o << "<use xlink:href = \"#" << id << "\" x = \"" << (.5 * j + .5 * i) << "\" y = \"" << i << "\"/>\n";
}

}  // namespace

TriangleDimerTiler::TriangleDimerTiler(cl::Context context0, cl::CommandQueue queue0, std::vector<cl::Device> devices,
                                       std::string source, cl_int& err)
    : ctx(rtapi::make_ctx(queue0.device())) {
    (void)context0;
    (void)devices;
    err = rtapi::basename(source) == "triangledimerkernel.cl" ? CL_SUCCESS : -1;
}

void TriangleDimerTiler::LoadTinyMT(std::string params, int size) {
    rtapi::check(rt_load_params_file(ctx.get(), params.c_str(), size), "LoadTinyMT");
}

void TriangleDimerTiler::LoadTinyMTDefault(int size) { rtapi::load_default_params(ctx.get(), size); }

void TriangleDimerTiler::Walk(tiling& t, int steps, long seed) {
    const int N = rtapi::side(t.size());
    rtapi::check(rt_tridimer_walk(ctx.get(), t.data(), N, steps, seed), "TriangleDimerTiler::Walk");
}

// Rows i = 2 .. N+1 of the lattice hold the vertices j = 2 .. N+3-i of the triangle.  Below
// the cut rows, every row pairs its vertices horizontally from j = 2 and the odd vertex left
// at the row's end is matched up-left or up-right, alternating with period 4 in i; rows
// beside the cut-out corner pair horizontally from the cut's edge and match their last vertex
// up-right when N - M is even.
tiling TriangleDimerTiler::IceCreamCone(int N, int M) {
    if (M > N - 1) std::cout << "Cutting out too big a chunk." << std::endl;
    if ((((N + 1) * (N + 2)) / 2 + (M * (M + 1)) / 2) % 2 != 0) std::cout << "No dimer cover exists." << std::endl;
    const int E = N + 4;
    const int k = (((N + 1) * (N + 2)) / 2) % 2;
    Dimers f(E);
    for (int i = 2; i < E - 2; ++i) {
        const int last = E - i - 1;
        for (int j = 2; j <= last; ++j) {
            if (i < M + 2) {
                if (M != 0 && (i + j) % 2 == M % 2 && j > M + 3 - i) f.H(i, j) = 1;
                if ((N - M) % 2 == 0 && i % 2 == 0 && j == last) f.R(i, j) = 1;
            } else {
                if (j % 2 == 0) f.H(i, j) = 1;
                if (j == last && i % 4 == (3 + N % 2 + 2 * k) % 4) {
                    f.H(i, j) = 0;
                    f.L(i, j) = 1;
                }
                if (j == last && i % 4 == (2 + N % 2 + 2 * k) % 4) f.R(i, j) = 1;
            }
        }
    }
    return f.states();
}

// Rows r = 0 .. y + z of the hexagon, at lattice row a = 2 + r: the first y + 1 rows start one
// vertex further left each row and end at the same column, the last z rows start at the same
// column and end one vertex further left each row.  A vertex (a, b) has the lower neighbours
// (a + 1, b) (edge l[a][b]) and (a + 1, b - 1) (edge r[a][b - 1]).  Scanning the rows top down,
// each row's vertices are paired horizontally except at most one, at an even offset from the
// row's first vertex, which is matched to a lower neighbour at an even offset in the next row
// (of its two lower neighbours, which differ by one, one has it; a vertex near the middle of a
// row has both, and rows change by at most one vertex at either end).  The scan covers every
// vertex iff their number is even.
tiling TriangleDimerTiler::Hexagon(int x, int y, int z) {
    if (x < 0 || y < 0 || z < 0 || z > x + y) {
        std::cout << "Bad hexagon sides." << std::endl;
        return tiling(16 * 16, 0);
    }
    const int rows = y + z + 1;
    std::vector<int> lo(rows), hi(rows);  // lattice columns of each row's first and last vertex
    const int E = std::max(rows, x + y + 1) + 5;
    long count = 0;
    for (int r = 0; r < rows; ++r) {
        lo[r] = 2 + y - std::min(r, y);
        hi[r] = 2 + y + x - std::max(0, r - y);
        count += hi[r] - lo[r] + 1;
    }
    if (count % 2 != 0) {
        std::cout << "No dimer cover exists." << std::endl;
        return tiling((size_t)4 * E * E, 0);
    }
    Dimers f(E);
    int carry = -1;  // column of the vertex of the previous row still to be matched in this row
    for (int r = 0; r < rows; ++r) {
        const int a = 2 + r;
        int start = lo[r];  // first vertex of the run still to pair horizontally
        int skip = -1;      // the vertex the carry took
        if (carry >= 0) {
            // lower neighbours of (a - 1, carry): (a, carry) and (a, carry - 1); take the one
            // at an even offset from this row's first vertex
            const int v = ((carry - lo[r]) % 2 == 0) ? carry : carry - 1;
            if (v < lo[r] || v > hi[r]) {
                std::cout << "No dimer cover exists." << std::endl;
                return tiling((size_t)4 * E * E, 0);
            }
            if (v == carry) f.L(a - 1, carry) = 1;
            else f.R(a - 1, v) = 1;
            skip = v;
        }
        carry = -1;
        // pair [start, skip) (of even length: skip lies at an even offset) and (skip, hi]; an
        // odd run leaves one vertex, near its middle at an even offset from its first vertex,
        // for the next row
        bool ok = true;
        auto pair_run = [&](int from, int to) {
            const int n = to - from + 1;
            if (n <= 0) return;
            int keep = -1;
            if (n % 2 != 0) {
                if (r + 1 == rows) {
                    ok = false;
                    return;
                }
                keep = from + 2 * ((n - 1) / 4);
            }
            for (int b = from; b <= to;) {
                if (b == keep) {
                    b += 1;
                    continue;
                }
                f.H(a, b) = 1;
                b += 2;
            }
            if (keep >= 0) carry = keep;
        };
        if (skip >= 0) {
            pair_run(start, skip - 1);
            pair_run(skip + 1, hi[r]);
        } else {
            pair_run(start, hi[r]);
        }
        if (!ok) {
            std::cout << "No dimer cover exists." << std::endl;
            return tiling((size_t)4 * E * E, 0);
        }
    }
    return f.states();
}

tiling TriangleDimerTiler::TestButterfly(int type) {
    Dimers f(8);
    if (type == 0) {
        f.H(2, 3) = f.H(4, 3) = f.R(2, 4) = f.R(3, 2) = 1;
    } else if (type == 1) {
        f.R(2, 3) = f.R(4, 2) = f.L(3, 2) = f.L(3, 4) = 1;
    } else if (type == 2) {
        f.L(2, 4) = f.L(4, 3) = f.R(3, 2) = f.R(3, 4) = 1;
    } else {
        std::cout << "Please choose types 0 (H), 1 (L), or 2 (R)." << std::endl;
        return tiling(16 * 16, 0);
    }
    return f.states();
}

std::vector<int> TriangleDimerTiler::TilingToTriangleFlips(const tiling& t) {
    const int N = (int)std::sqrt((double)t.size());
    std::vector<int> out((size_t)(N / 2) * N, 0);
    for (int i = 1; i < N / 2 - 1; ++i)
        for (int j = 1; j < N - 1; ++j) {
            const int src[4] = {t[(size_t)(2 * i + 1) * N + j], t[(size_t)(2 * i + 1) * N + j + 1],
                                t[(size_t)(2 * i) * N + j + 1], j % 2 ? t[(size_t)(2 * i + 2) * N + j] : 0};
            const auto& form = (j % 2 == 0) ? kTriDown : kTriUp;
            int s = 0, w = 1;
            for (int dgt = 0; dgt < 3; ++dgt, w *= 4)
                for (const Term& x : form[dgt]) s += w * x.w * bit(src[x.src], x.bit);
            out[(size_t)i * N + j] = s;
        }
    return out;
}

std::vector<int> TriangleDimerTiler::TilingToButterflyFlips(const tiling& t) {
    const int N = (int)std::sqrt((double)t.size());
    std::vector<int> out((size_t)N * N, 0);
    auto at = [&](int i, int j) { return t[(size_t)i * N + j]; };
    for (int i = 2; i < N - 2; ++i)
        for (int j = 2; j < N - 2; ++j) {
            if (i % 2 == 0 && j % 2 == 0) continue;
            const BflyForm& f = kBfly[i % 2 == 0 ? 0 : (j % 2 == 0 ? 1 : 2)];
            const int s = at(i, j);
            int v = 256 * bit(at(i + f.cdi, j + f.cdj), f.cbit), w = 1;
            for (int k = 0; k < 4; ++k, w *= 4) {
                const int nb = at(i + f.d[k].di, j + f.d[k].dj);
                v += w * (bit(nb, f.d[k].lo) + 2 * bit(nb, f.d[k].hi) + 3 * bit(s, k));
            }
            out[(size_t)i * N + j] = v;
        }
    return out;
}

std::vector<int> TriangleDimerTiler::TilingToVertices(const tiling& t) {
    const int N = (int)std::sqrt((double)t.size());
    const int M = N / 2 + 1;
    std::vector<int> v((size_t)M * M, 0);
    for (int i = 2; i < M - 2; ++i)
        for (int j = 2; j < M - 2; ++j) v[(size_t)i * M + j] = vertex_covered(t, N, i, j) ? 1 : 0;
    return v;
}

domain TriangleDimerTiler::TilingToDomain(const tiling& t) {
    const int N = (int)std::sqrt((double)t.size());
    const int M = N / 2 + 1, W = 2 * (M - 1);
    const std::vector<int> v = TilingToVertices(t);
    auto V = [&](int i, int j) { return v[(size_t)i * M + j] == 1; };
    domain d((size_t)(M - 1) * W, 0);
    for (int i = 0; i < M - 1; ++i)
        for (int j = 0; j < W; ++j) {
            const int c = j / 2;
            const bool in = (j % 2 == 0) ? (V(i, c) && V(i, c + 1) && V(i + 1, c))
                                         : (V(i, c + 1) && V(i + 1, c) && V(i + 1, c + 1));
            if (in) d[(size_t)i * W + j] = 1;
        }
    return d;
}

std::vector<int> TriangleDimerTiler::DimerIndicator(const tiling& t, int c) {
    const int N = (int)std::sqrt((double)t.size());
    const int M = N / 2;
    std::vector<int> out((size_t)M * M, 0);
    if (c < 0 || c > 2) return out;
    const int row_par = (c == 0) ? 1 : 0;  // horizontal dimers sit in odd rows
    const int b = (c == 0) ? 0 : (c == 1) ? 3 : 2;
    for (int i = row_par; i < N; i += 2)
        for (int j = 1; j < N; j += 2) out[(size_t)(i / 2) * M + j / 2] = bit(t[(size_t)i * N + j], b);
    return out;
}

void TriangleDimerTiler::DomainToSVG(const domain& d, std::string filename) {
    std::ofstream o(filename.c_str());
    const int M = (int)std::sqrt((double)(d.size() / 2));
    svg_open(o, M);
    o << "<defs>\n";
    o << "<g id=\"w\">  <polygon points = \"0,1 .5,0 1,1\" fill=\"paleturquoise\"/> </g>\n";
    o << "<g id=\"b\">  <polygon points = \"0,0 1,0 .5,1\" fill=\"deepskyblue\"/> </g>\n";
    o << "</defs>\n";
    for (int i = 0; i < M; i++)
        for (int j = 0; j < 2 * M; j++)
            if (d[(size_t)i * M * 2 + j] == 1) use(o, j % 2 == 0 ? "b" : "w", i, j);
    o << "</svg>";
}

void TriangleDimerTiler::DimerToSVG(const tiling& t, std::string filename) {
    const domain d = TilingToDomain(t);
    DimerToSVG(t, d, filename);
}

void TriangleDimerTiler::TilingToSVG(const tiling& t, std::string filename) {
    std::ofstream o(filename.c_str());
    const int N = (int)std::sqrt((double)t.size());
    const int M = N / 2;
    svg_open(o, M);
    o << "<defs>\n";
    o << "<g id=\"dR\">  <polygon points = \"1,-0.6666666667 0.5,-0.3333333333 0.5,0.3333333333 0,0.6666666667 "
         "0,1.333333333 0.5,1.666666667 1,1.333333333 1,0.6666666667 1.5,0.3333333333 1.5,-0.3333333333\" "
         "fill=\"midnightblue\" stroke=\"black\" stroke-width=\".05\"/></g>\n";
    o << "<g id=\"dL\">  <polygon points = \"0,-0.6666666667 0.5,-0.3333333333 0.5,0.3333333333 1,0.6666666667 "
         "1,1.333333333 0.5,1.666666667 0,1.333333333 0,0.6666666667 -0.5,0.3333333333 -0.5,-0.3333333333\" "
         "fill=\"lightsteelblue\" stroke=\"black\" stroke-width=\".05\"/></g>\n";
    o << "<g id=\"dH\">  <polygon points = \"0,-0.6666666667 0.5,-0.3333333333 1,-0.6666666667 1.5,-0.3333333333 "
         "1.5,0.3333333333 1,0.6666666667 0.5,0.3333333333 0,0.6666666667 -0.5,0.3333333333 -0.5,-0.3333333333\" "
         "fill=\"slategrey\" stroke=\"black\" stroke-width=\".05\"/></g>\n";
    o << "</defs>\n";
    for (int i = 0; i < M - 1; ++i)
        for (int j = 1; j < 2 * M - 1; j += 2) {
            const int up = t[(size_t)2 * i * N + j], mid = t[(size_t)(2 * i + 1) * N + j];
            if (bit(up, 2)) use(o, "dR", i, j);
            if (bit(up, 3)) use(o, "dL", i, j);
            if (bit(mid, 0)) use(o, "dH", i, j);
        }
    o << "</svg>";
}

void TriangleDimerTiler::DimerToSVG(const tiling& t, const domain& d, std::string filename) {
    std::ofstream o(filename.c_str());
    const int N = (int)std::sqrt((double)t.size());
    const int M = (int)std::sqrt((double)(d.size() / 2));
    // a triangle of the domain; indices outside the array count as not in the domain
    auto D = [&](long k) { return k >= 0 && k < (long)d.size() && d[(size_t)k] == 1; };
    svg_open(o, M);
    o << "<defs>\n";
    o << "<g id=\"dR\">  <polyline points= \".5,1 1,0\" stroke=\"darkblue\" stroke-width=\".05\"/> </g>\n";
    o << "<g id=\"dL\">  <polyline points= \".5,1 0,0\" stroke=\"darkred\" stroke-width=\".05\"/> </g>\n";
    o << "<g id=\"dH\">  <polyline points= \"1,0 0,0\" stroke=\"darkgreen\" stroke-width=\".05\"/> </g>\n";
    o << "<g id=\"eR\">  <polyline points= \".5,1 1,0\" stroke=\"black\" stroke-width=\".001\"/> </g>\n";
    o << "<g id=\"eL\">  <polyline points= \".5,1 0,0\" stroke=\"black\" stroke-width=\".001\"/> </g>\n";
    o << "<g id=\"eH\">  <polyline points= \"1,0 0,0\" stroke=\"black\" stroke-width=\".001\"/> </g>\n";
    o << "<g id=\"v1\">  <ellipse cx=\".5\" cy=\"0\" rx=\".1\" ry=\".1\" stroke=\"black\" stroke-width=\".02\" "
         "fill=\"black\"/> </g> \n";
    o << "<g id=\"v2\">  <ellipse cx=\"1.5\" cy=\"0\" rx=\".1\" ry=\".1\" stroke=\"black\" stroke-width=\".02\" "
         "fill=\"black\"/> </g> \n";
    o << "<g id=\"v3\">  <ellipse cx=\"1\" cy=\"1\" rx=\".1\" ry=\".1\" stroke=\"black\" stroke-width=\".02\" "
         "fill=\"black\"/> </g> \n";
    o << "</defs>\n";
    // dimers, and the lattice edges of the domain that carry none
    for (int i = 0; i < M - 1; ++i)
        for (int j = 1; j < 2 * M - 1; j += 2) {
            const long row = (long)i * 2 * M;
            const int up = t[(size_t)2 * i * N + j], mid = t[(size_t)(2 * i + 1) * N + j];
            if (bit(up, 2)) use(o, "dR", i, j);
            else if (D(row + j - 1)) use(o, "eR", i, j);
            if (bit(up, 3)) use(o, "dL", i, j);
            else if (D(row + j - 1)) use(o, "eL", i, j);
            if (bit(mid, 0)) use(o, "dH", i, j);
            else if (D(row + j - 1) || D(row - 2 * M + j)) use(o, "eH", i, j);
            if (!bit(up, 2) && D(row + j)) use(o, "eR", i, j);
            if (!bit(up, 3) && j > 1 && D(row + j - 2)) use(o, "eL", i, j);
        }
    // vertices of the domain's triangles
    for (int i = 0; i < M; i++)
        for (int j = 2; j < 2 * M; j += 2) {
            const long row = (long)i * 2 * M;
            if (D(row + j) || D(row + j - 1)) use(o, "v1", i, j);
            if (D(row + j) || D(row + j + 1)) use(o, "v2", i, j);
            if (D(row + j) || D(row + j + 1) || D(row + j - 1)) use(o, "v3", i, j);
        }
    o << "</svg>";
}
