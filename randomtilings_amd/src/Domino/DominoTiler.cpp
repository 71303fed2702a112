// DominoTiler.cpp — the domino tiler over the HIP library (include/rt_hip.h).
//
// Walk follows DominoTiler.cpp:9-107 of the reference (colour schedule, TinyMT seeding,
// Rotate/Update semantics) and runs on the GPU through rt_domino_walk /
// rt_domino_walk_epochs.  The statics reproduce the reference's outputs: the domain
// constructors follow DominoTiler.cpp:118-193, the height conventions :202-232, and the
// SVG writers emit the same documents as :508-681.
#include "DominoTiler.h"

#include <algorithm>
#include <stdexcept>

#include "../common/heights.hpp"
#include "../common/lattice.hpp"
#include "../common/rtapi.hpp"
#include "rt_hip.h"

namespace {

using rtlat::DominoGeom;
constexpr const int* kBit = DominoGeom::bit;

tiling thurston(const domain& d, int dir, const char* name) {
    const std::vector<int> vertices = DominoTiler::DomainToVertices(d);
    const int N = rtapi::side(vertices.size());
    DominoGeom g(d, N);
    std::vector<int> hf;
    if (!rtapi::boundary_heights(g, vertices, hf)) {
        std::cout << "!!This region is NOT tileable!! " << name << ", Boundary" << std::endl;
        return tiling((size_t)N * N, 0);
    }
    if (!rtapi::thurston_closure(g, vertices, hf, dir)) {
        std::cout << "!!This region is NOT tileable!! " << name << ", Bulk" << std::endl;
        return tiling((size_t)N * N, 0);
    }
    return DominoTiler::HeightfuncToTiling(hf, d);
}

int rule_from_path(const std::string& path) {
    const std::string b = rtapi::basename(path);
    if (b == "dominokernel.cl") return RT_DOMINO_LAZY;
    if (b == "dominokernelCFTP.cl") return RT_DOMINO_CFTP;
    return -1;
}

}  // namespace

DominoTiler::DominoTiler(cl::Context context0, cl::CommandQueue queue0, std::vector<cl::Device> devices,
                         std::string source, cl_int& err)
    : ctx(rtapi::make_ctx(queue0.device())), rule(rule_from_path(source)) {
    (void)context0;
    (void)devices;
    err = (rule < 0) ? -1 : CL_SUCCESS;
    if (rule < 0) rule = RT_DOMINO_LAZY;
}

void DominoTiler::LoadTinyMT(std::string params, int size) {
    rtapi::check(rt_load_params_file(ctx.get(), params.c_str(), size), "LoadTinyMT");
}

void DominoTiler::LoadTinyMTDefault(int size) { rtapi::load_default_params(ctx.get(), size); }

void DominoTiler::Walk(tiling& t, long steps, long seed) {
    const int N = rtapi::side(t.size());
    if (N % 2 != 0) std::cout << "Careful! The tiling has odd dimension." << std::endl;
    rtapi::check(rt_domino_walk(ctx.get(), rule, nullptr, t.data(), N, steps, seed), "DominoTiler::Walk");
}

void DominoTiler::Walk(tiling& t, std::vector<long> steps, std::vector<long> seeds) {
    const int N = rtapi::side(t.size());
    if (N % 2 != 0) std::cout << "Careful! The tiling has odd dimension." << std::endl;
    if (steps.size() != seeds.size()) throw std::invalid_argument("DominoTiler::Walk: steps and seeds differ in length");
    std::vector<int64_t> st(steps.begin(), steps.end()), sd(seeds.begin(), seeds.end());
    rtapi::check(rt_domino_walk_epochs(ctx.get(), rule, nullptr, t.data(), N, st.data(), sd.data(), (int32_t)st.size()),
                 "DominoTiler::Walk");
}

void DominoTiler::WalkWeighted(tiling& t, long steps, long seed, const double p12[4]) {
    const int N = rtapi::side(t.size());
    rtapi::check(rt_domino_walk(ctx.get(), RT_DOMINO_WEIGHTED, p12, t.data(), N, steps, seed),
                 "DominoTiler::WalkWeighted");
}

tiling DominoTiler::CFTP(const tiling& tMax, const tiling& tMin, long seed, long first_steps, bool doubling,
                         int max_restarts, const double* p12, long* total_steps) {
    const int N = rtapi::side(tMax.size());
    if (tMin.size() != tMax.size()) throw std::runtime_error("DominoTiler::CFTP: tMax and tMin differ in size");
    tiling out(tMax.size(), 0);
    int64_t total = 0;
    int32_t restarts = 0;
    rtapi::check(rt_domino_cftp(ctx.get(), p12 ? RT_DOMINO_WEIGHTED : rule, p12, tMax.data(), tMin.data(), N, seed,
                                doubling ? RT_CFTP_DOUBLING : RT_CFTP_REFERENCE, first_steps, max_restarts, out.data(),
                                &total, &restarts),
                 "DominoTiler::CFTP");
    if (total_steps) *total_steps = (long)total;
    return out;
}

// ---- domains ----------------------------------------------------------------------------

domain DominoTiler::Rectangle(int M, int N) {
    int dim = std::max(M, N) + 4;
    if (dim % 2 == 0) dim++;  // odd domain side -> even tiling side
    domain d((size_t)dim * dim, 0);
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < N; ++j) d[(size_t)(i + 2) * dim + j + 2] = 1;
    return d;
}

domain DominoTiler::AztecDiamond(int N) { return AztecRectangle(N, std::vector<int>()); }

// A (2N-k) x (2N-k) diamond-shaped region whose rows widen by two squares per row in the
// top part, run at full width through the middle N rows and narrow again at the bottom;
// each entry r of `remove` deletes the square (N-k+r, r) on the lower-left staircase.  The
// region is then padded by two empty squares on every side (one more when the side would
// otherwise be even).
domain DominoTiler::AztecRectangle(int N, const std::vector<int>& remove) {
    const int k = (int)remove.size();
    const int M = 2 * N - k;
    std::vector<int> d((size_t)std::max(M, 0) * std::max(M, 0), 0);
    for (int i = 0; i < M; ++i) {
        int lo, hi;
        if (i == 0 || i == M - 1) {
            const int c = (N - k) + (i == M - 1 ? k : 0);
            lo = c - 1;
            hi = c;
        } else if (i < N - k) {
            lo = N - k - i - 1;
            hi = N - k + i;
        } else if (i > N - 1) {
            lo = i - N + k;
            hi = 3 * N - k - i - 1;
        } else {
            lo = i - (N - k);
            hi = i + (N - k);
        }
        for (int j = std::max(lo, 0); j <= std::min(hi, M - 1); ++j) d[(size_t)i * M + j] = 1;
    }
    for (int r : remove) {
        const int i0 = N - k + r, j0 = r;
        if (i0 >= 0 && i0 < M && j0 >= 0 && j0 < M) d[(size_t)i0 * M + j0] = 0;
    }
    const int M2 = M + 4 + (1 - M % 2);
    domain out((size_t)M2 * M2, 0);
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < M; ++j) out[(size_t)(i + 2) * M2 + j + 2] = d[(size_t)i * M + j];
    return out;
}

// ---- heights ----------------------------------------------------------------------------

tiling DominoTiler::MaxTiling(const domain& d) { return thurston(d, +1, "Max"); }

tiling DominoTiler::MinTiling(const domain& d) { return thurston(d, -1, "Min"); }

tiling DominoTiler::HeightfuncToTiling(const heightfunc& hf, const domain& d) {
    const int N = rtapi::side(hf.size());
    DominoGeom g(d, N);
    tiling t((size_t)N * N, 0);
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            const int h0 = hf[(size_t)i * N + j];
            if (h0 == infty) continue;
            for (int nb = 0; nb < 4; nb++) {
                const int ii = i + DominoGeom::di[nb], jj = j + DominoGeom::dj[nb];
                if (ii < 0 || jj < 0 || ii >= N || jj >= N) continue;
                const int h1 = hf[(size_t)ii * N + jj];
                if (h1 == infty) continue;
                bool a, b;
                g.faces(i, j, nb, a, b);
                if (!(a || b)) continue;
                // a domino crosses the edge iff the height jumps by 3 across it
                t[(size_t)i * N + j] += kBit[nb] * std::abs((h1 - h0) / 2);
            }
        }
    return t;
}

heightfunc DominoTiler::TilingToHeightfunc(const tiling& t, const domain& d) {
    const int N = rtapi::side(t.size());
    const std::vector<int> vertices = DomainToVertices(d);
    DominoGeom g(d, N);
    return rtapi::tiling_heights(g, vertices,
                                 [&](int i, int j, int nb) { return (t[(size_t)i * N + j] & kBit[nb]) != 0; });
}

domain DominoTiler::TilingToDomain(const tiling& t) {
    const int N = rtapi::side(t.size());
    const int M = N - 1;
    domain d((size_t)M * M, 0);
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < M; ++j) {
            // a square is covered iff its top-left corner has a crossed right/down edge or its
            // bottom-right corner a crossed up/left edge
            if ((t[(size_t)i * N + j] & (2 + 8)) != 0 || (t[(size_t)(i + 1) * N + j + 1] & (1 + 4)) != 0)
                d[(size_t)i * M + j] = 1;
        }
    return d;
}

std::vector<int> DominoTiler::DomainToVertices(const domain& d) { return rtlat::domino_vertices(d); }

// ---- SVG --------------------------------------------------------------------------------

static void svg_open(std::ofstream& f, int w, int h) {
    f << "<svg xmlns='http://www.w3.org/2000/svg' xmlns:xlink=\"http://www.w3.org/1999/xlink\" height=\"500\" "
         "width=\"500\" viewBox=\"0 0 "
      << w << " " << h << "\">\n";
}

static void use(std::ofstream& f, const char* id, double x, double y, bool spaced) {
    if (spaced) f << "<use xlink:href = \"#" << id << "\" x = \"" << x << "\" y = \"" << y << "\"/>\n";
    else f << "<use xlink:href=\"#" << id << "\" x=\"" << x << "\" y=\"" << y << "\"/>\n";
}

void DominoTiler::DomainToSVG(const domain& d, std::string filename) {
    std::ofstream f(filename.c_str());
    const int M = (int)std::sqrt((double)d.size());
    svg_open(f, M, M);
    f << "<defs>\n";
    f << "<g id=\"b\">  <polygon points = \"-.5,-.5 .5,-.5 .5,.5 -.5,.5\" fill=\"blue\"/> </g>\n";
    f << "<g id=\"w\">  <polygon points = \"-.5,-.5 .5,-.5 .5,.5 -.5,.5\" fill=\"grey\"/> </g>\n";
    f << "</defs>\n";
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < M; ++j)
            if (d[(size_t)i * M + j] == 1) use(f, ((i + j) % 2 == 0) ? "b" : "w", j + .5, i + .5, true);
    f << "</svg>";
}

void DominoTiler::TilingToSVG(const tiling& t, std::string filename) {
    std::ofstream f(filename.c_str());
    const int N = (int)std::sqrt((double)t.size());
    svg_open(f, N - 1, N - 1);
    f << "<defs>\n";
    for (const char* id : {"h1", "v1", "h2", "v2"}) {
        const bool h = id[0] == 'h';
        f << "<g id=\"" << id << "\">  <polygon points = \""
          << (h ? "-1,0 -1,-1 1,-1 1,0" : "-1,-1 0,-1 0,1 -1,1")
          << "\" fill=\"white\" stroke=\"black\" stroke-width=\".05\"/> </g>\n";
    }
    f << "</defs>\n";
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            const int s = t[(size_t)i * N + j];
            const bool odd = (i + j) % 2 == 1;
            if (s & 1) use(f, odd ? "h1" : "h2", j, i, true);
            if (s & 4) use(f, odd ? "v1" : "v2", j, i, true);
        }
    f << "</svg>";
}

void DominoTiler::MayaToSVG(const tiling& t, std::string filename) {
    std::ofstream f(filename.c_str());
    const int N = (int)std::sqrt((double)t.size());
    svg_open(f, N - 1, N - 1);
    f << "<defs>\n";
    auto circles = [&](const char* id, const char* fill, bool horizontal) {
        f << "<g id=\"" << id << "\">";
        f << "<circle cx=\"-.5\" cy=\"-.5\" r=\".4\" stroke=\"black\" stroke-width=\".02\" fill=\"" << fill << "\"/> "
          << "<circle cx=\"" << (horizontal ? ".5" : "-.5") << "\" cy=\"" << (horizontal ? "-.5" : ".5")
          << "\" r=\".4\" stroke=\"black\" stroke-width=\".02\" fill=\"" << fill << "\"/>";
        f << "</g>\n";
    };
    circles("h1", "black", true);
    circles("h2", "white", true);
    circles("v1", "black", false);
    circles("v2", "white", false);
    f << "</defs>\n";
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            const int s = t[(size_t)i * N + j];
            const bool odd = (i + j) % 2 == 1;
            if (s & 1) use(f, odd ? "h1" : "h2", j, i, true);
            if (s & 4) use(f, odd ? "v1" : "v2", j, i, true);
        }
    f << "</svg>";
}

void DominoTiler::DimerToSVG(const tiling& t, std::string filename) { DimerToSVG(t, TilingToDomain(t), filename); }

void DominoTiler::DimerToSVG(const tiling& t, const domain& d, std::string filename) {
    std::ofstream f(filename.c_str());
    const int N = (int)std::sqrt((double)t.size());
    const int M = N - 1;
    svg_open(f, M, M);
    f << "<defs>\n";
    f << "<g id=\"hd\">  <polyline points= \"-.5,0 .5,0\" stroke=\"red\" stroke-width=\".05\"/> </g>\n";
    f << "<g id=\"he\">  <polyline points= \"-.5,0 .5,0\" stroke=\"black\" stroke-width=\".01\"/> </g>\n";
    f << "<g id=\"vd\">  <polyline points= \"0,-.5 0,.5\" stroke=\"red\" stroke-width=\".05\"/> </g>\n";
    f << "<g id=\"ve\">  <polyline points= \"0,-.5 0,.5\" stroke=\"black\" stroke-width=\".01\"/> </g>\n";
    f << "<g id=\"b\">  <circle cx=\"0\" cy=\"0\" r=\".1\" stroke=\"black\" stroke-width=\".02\" fill=\"black\"/> </g> \n";
    f << "<g id=\"w\">  <circle cx=\"0\" cy=\"0\" r=\".1\" stroke=\"black\" stroke-width=\".02\" fill=\"white\" /> </g>\n";
    f << "</defs>\n";
    // dimers: a crossed up edge is a horizontal dimer above the vertex, a crossed left edge
    // a vertical one to its left
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            const int s = t[(size_t)i * N + j];
            if ((s & 1) == 1) use(f, "hd", j, i - .5, false);
            if ((s & 4) == 4) use(f, "vd", j - .5, i, false);
        }
    auto dd = [&](int i, int j) { return (i >= 0 && j >= 0 && i < M && j < M) ? d[(size_t)i * M + j] : 0; };
    for (int i = 1; i < M; ++i)
        for (int j = 1; j < M; ++j) {
            if (dd(i, j) != 1) continue;
            const double x = j + .5, y = i + .5;
            if ((i + j) % 2 == 1) {
                if (dd(i, j + 1) == 1) use(f, "he", x + .5, y, false);
                if (dd(i, j - 1) == 1) use(f, "he", x - .5, y, false);
            }
            if (dd(i - 1, j) == 1) use(f, "ve", x, y - .5, false);
            if (dd(i + 1, j) == 1) use(f, "ve", x, y + .5, false);
        }
    for (int i = 1; i < M; ++i)
        for (int j = 1; j < M; ++j)
            if (dd(i, j) == 1) use(f, ((i + j) % 2 == 0) ? "b" : "w", j + .5, i + .5, true);
    f << "</svg>";
}
