// LozengeTiler.cpp — the lozenge tiler over the HIP library.
//
// Walk follows LozengeTiler.cpp:18-88 of the reference and runs through rt_lozenge_walk.
// The statics reproduce the reference's outputs (height conventions LozengeTiler.cpp:12-16,
// domain constructor :344-389, SVG documents :480-613).
#include "LozengeTiler.h"

#include <algorithm>
#include <stdexcept>

#include "../common/heights.hpp"
#include "../common/lattice.hpp"
#include "../common/rtapi.hpp"
#include "rt_hip.h"

namespace {

using rtlat::LozGeom;
constexpr const int* kDi = LozGeom::di;
constexpr const int* kDj = LozGeom::dj;
constexpr const int* kBit = LozGeom::bit;

tiling thurston(const domain& d, int dir, const char* name) {
    const std::vector<int> vertices = LozengeTiler::DomainToVertices(d);
    const int N = rtapi::side(vertices.size());
    LozGeom g(d, N);
    std::vector<int> hf;
    if (!rtapi::boundary_heights(g, vertices, hf)) {
        std::cout << "!!This region is NOT tileable!! " << name << ", Boundary" << std::endl;
        return tiling((size_t)N * N, 0);
    }
    if (!rtapi::thurston_closure(g, vertices, hf, dir)) {
        std::cout << "!!This region is NOT tileable!! " << name << ", Bulk" << std::endl;
        return tiling((size_t)N * N, 0);
    }
    return LozengeTiler::HeightfuncToTiling(hf, d);
}

}  // namespace

LozengeTiler::LozengeTiler(cl::Context context0, cl::CommandQueue queue0, std::vector<cl::Device> devices,
                           std::string source, cl_int& err)
    : ctx(rtapi::make_ctx(queue0.device())), rule(RT_LOZENGE_UNIFORM), q(0.99) {
    (void)context0;
    (void)devices;
    const std::string b = rtapi::basename(source);
    err = CL_SUCCESS;
    if (b == "lozengekernel.cl") rule = RT_LOZENGE_UNIFORM;
    else if (b == "qlozengekernel.cl") rule = RT_LOZENGE_QREF;
    else err = -1;
}

void LozengeTiler::LoadTinyMT(std::string params, int size) {
    rtapi::check(rt_load_params_file(ctx.get(), params.c_str(), size), "LoadTinyMT");
}

void LozengeTiler::LoadTinyMTDefault(int size) { rtapi::load_default_params(ctx.get(), size); }

void LozengeTiler::SetQ(double q_, bool heatbath) {
    q = q_;
    rule = heatbath ? RT_LOZENGE_QHEAT : RT_LOZENGE_QREF;
}

void LozengeTiler::Walk(tiling& t, int steps, long seed) {
    const int N = rtapi::side(t.size());
    rtapi::check(rt_lozenge_walk(ctx.get(), rule, q, t.data(), N, steps, seed), "LozengeTiler::Walk");
}

// ---- domains ----------------------------------------------------------------------------

domain LozengeTiler::Hexagon(int N) { return AlmostHexDomain(N, 0); }

// A 2N x 4N triangle array holding a hexagon of side N (rows above the middle start one
// triangle further left each, rows below end one further left), minus an M x 2M block at
// the top right, embedded with 3 empty rows above and 6 empty triangle columns to the left,
// and widened so that the vertex array side is a multiple of 3.
domain LozengeTiler::AlmostHexDomain(int N, int M) {
    const int W2 = 4 * N + 12;
    if (M > N) {
        std::cout << "Cutting out to big a chunk." << std::endl;
        return domain((size_t)(2 * N + 6) * W2, 0);
    }
    const int side = 2 * N + 6;  // rows of the padded domain; it has 2*side columns
    const int k = (3 - (side + 1) % 3) % 3;
    const int S = side + k;
    domain d((size_t)S * 2 * S, 0);
    for (int i = 0; i < 2 * N; ++i)
        for (int j = 0; j < 4 * N; ++j) {
            bool in = (i < N) ? (2 * i + j - 1 >= 2 * N - 2) : (2 * i + j <= 6 * N - 2);
            if (i < M && j >= 4 * N - 2 * M) in = false;
            if (in) d[(size_t)(i + 3) * 2 * S + (j + 6)] = 1;
        }
    return d;
}

std::vector<int> LozengeTiler::DomainToVertices(const domain& d) { return rtlat::lozenge_vertices(d); }

domain LozengeTiler::TilingToDomain(const tiling& t) {
    const int N = rtapi::side(t.size());
    const int M = N - 1;
    domain d((size_t)M * M * 2, 0);
    auto set = [&](long k) {
        if (k >= 0 && k < (long)d.size()) d[(size_t)k] = 1;
    };
    for (int i = 0; i < N - 1; i++)
        for (int j = 0; j < N - 1; j++) {
            const int s = t[(size_t)i * N + j];
            const long up = (long)(i - 1) * 2 * M + 2 * j;
            if (s & 1) { set(up); set(up - 1); }
            if (s & 2) { set(up); set(up + 1); }
            if (s & 8) { set(up + 1); set((long)i * 2 * M + 2 * j); }
        }
    return d;
}

// ---- heights ----------------------------------------------------------------------------

tiling LozengeTiler::MaxTiling(const domain& d) { return thurston(d, +1, "Max"); }

tiling LozengeTiler::MinTiling(const domain& d) { return thurston(d, -1, "Min"); }

tiling LozengeTiler::HeightfuncToTiling(const heightfunc& hf, const domain& d) {
    const int N = rtapi::side(hf.size());
    LozGeom g(d, N);
    tiling t((size_t)N * N, 0);
    for (int i = 1; i < N - 1; ++i)
        for (int j = 1; j < N - 1; ++j) {
            const int h0 = hf[(size_t)i * N + j];
            for (int nb = 0; nb < 6; nb++) {
                bool a, b;
                g.faces(i, j, nb, a, b);
                if (!(a || b)) continue;
                const int h1 = hf[(size_t)(i + kDi[nb]) * N + j + kDj[nb]];
                if (h0 == infty || h1 == infty) continue;
                // a lozenge crosses the edge iff the height jumps by 2 across it
                if (std::abs(h1 - h0) == 2) t[(size_t)i * N + j] += kBit[nb];
            }
        }
    return t;
}

heightfunc LozengeTiler::TilingToHeightfunc(const tiling& t, const domain& d) {
    const int N = rtapi::side(t.size());
    const std::vector<int> vertices = DomainToVertices(d);
    LozGeom g(d, N);
    return rtapi::tiling_heights(g, vertices,
                                 [&](int i, int j, int nb) { return (t[(size_t)i * N + j] & kBit[nb]) != 0; });
}

// ---- text and SVG -----------------------------------------------------------------------

void LozengeTiler::PrintDomain(const domain& d, std::string filename) {
    std::ofstream f(filename.c_str());
    const int N = (int)std::sqrt(d.size() / 2);
    for (int i = 0; i < N; ++i) {
        for (int j = 0; j < 2 * N; ++j) f << d[(size_t)i * 2 * N + j] << " ";
        f << "\n";
    }
}

static void hex_svg_open(std::ofstream& f, int M) {
    f << "<svg xmlns='http://www.w3.org/2000/svg' xmlns:xlink=\"http://www.w3.org/1999/xlink\" height=\"433.013\" "
         "width=\"500\" viewBox=\"0 0 "
      << 3 / 2.0 * M << " " << 3 / 2.0 * M << "\" preserveAspectRatio=\"none\">\n";
}

static void use(std::ofstream& f, const char* id, double x, double y) {
    f << "<use xlink:href = \"#" << id << "\" x = \"" << x << "\" y = \"" << y << "\"/>\n";
}

void LozengeTiler::DimerToSVG(const tiling& t, std::string filename) { DimerToSVG(t, TilingToDomain(t), filename); }

void LozengeTiler::DimerToSVG(const tiling& t, const domain& d, std::string filename) {
    const int N = (int)std::sqrt((double)t.size());
    const int M = N - 1;
    std::ofstream f(filename.c_str());
    hex_svg_open(f, M);
    f << "<defs>\n";
    const char* seg[3] = {"-.5,-.33333333333 0,-.666666666666", "0,-.666666666 .5,-.333333333333",
                          ".5,-.333333333333 .5,.333333333333"};
    const char* nm[3] = {"1", "2", "8"};
    for (int w = 0; w < 2; w++)
        for (int k = 0; k < 3; k++)
            f << "<g id=\"" << (w ? "d" : "e") << nm[k] << "\">  <polyline points= \"" << seg[k]
              << "\" stroke=\"black\" stroke-width=\"" << (w ? ".025" : ".001") << "\"/> </g>\n";
    f << "<g id=\"b\">  <ellipse cx=\".5\" cy=\".333333333333\" rx=\".1\" ry=\".11547\" stroke=\"black\" "
         "stroke-width=\".02\" fill=\"black\"/> </g> \n";
    f << "<g id=\"w\">  <ellipse cx=\".5\" cy=\".666666666666\" rx=\".1\"  ry=\".11547\" stroke=\"black\" "
         "stroke-width=\".02\" fill=\"white\" /> </g>\n";
    f << "</defs>\n";
    for (int i = 1; i < N - 1; i++)
        for (int j = 1; j < N - 1; j++) {
            const int s = t[(size_t)i * N + j];
            if (s & 1) use(f, "d1", j + .5 * i, i);
            if (s & 2) use(f, "d2", j + .5 * i, i);
            if (s & 8) use(f, "d8", j + .5 * i, i);
        }
    // domain edges (flat indexing as the reference does; reads past the array count as empty)
    auto df = [&](long k) { return (k >= 0 && k < (long)d.size()) ? d[(size_t)k] : 0; };
    for (int i = 1; i < N - 1; i++)
        for (int j = 1; j < 2 * N - 1; j++) {
            if (df((long)i * M * 2 + j) != 1 || j % 2 != 0) continue;
            const double x = .5 + .5 * j + .5 * i;
            if (df((long)i * M * 2 + j - 1) == 1) use(f, "e1", x, i + 1);
            if (df((long)i * M * 2 + j + 1) == 1) use(f, "e2", x, i + 1);
            if (df((long)(i + 1) * M * 2 + j - 1) == 1) use(f, "e8", x, i + 1);
        }
    for (int i = 1; i < N - 1; i++)
        for (int j = 1; j < 2 * N - 1; j++)
            if (df((long)i * M * 2 + j) == 1) use(f, (j % 2 == 0) ? "b" : "w", .5 * i + .5 * j, i);
    f << "</svg>";
}

void LozengeTiler::TilingToSVG(const tiling& t, std::string filename) {
    std::ofstream f(filename.c_str());
    const int N = (int)std::sqrt((double)t.size());
    hex_svg_open(f, N - 1);
    f << "<defs>\n";
    f << "<g id=\"r\">  <polygon points = \"0,0 -1,0 -.5,-1 .5,-1\" fill=\"midnightblue\"/> </g>\n";
    f << "<g id=\"g\">  <polygon points = \"0,0 1,0 .5,-1 -.5,-1\" fill=\"lightsteelblue\"/> </g>\n";
    f << "<g id=\"b\">  <polygon points = \"0,0 .5,-1 1,0 .5,1\" fill=\"slategrey\"/> </g>\n";
    f << "</defs>\n";
    for (int i = 1; i < N - 1; i++)
        for (int j = 1; j < N - 1; j++) {
            const int s = t[(size_t)i * N + j];
            if (s & 1) use(f, "r", j + .5 * i, i);
            if (s & 2) use(f, "g", j + .5 * i, i);
            if (s & 8) use(f, "b", j + .5 * i, i);
        }
    f << "</svg>";
}

void LozengeTiler::DomainToSVG(const domain& d, std::string filename) {
    std::ofstream f(filename.c_str());
    const int M = (int)std::sqrt(d.size() / 2);
    hex_svg_open(f, M);
    f << "<defs>\n";
    f << "<g id=\"w\">  <polygon points = \"0,1 .5,0 1,1\" fill=\"lightsteelblue\"/> </g>\n";
    f << "<g id=\"b\">  <polygon points = \"0,0 1,0 .5,1\" fill=\"slategrey\"/> </g>\n";
    f << "</defs>\n";
    for (int i = 0; i < M; i++)
        for (int j = 0; j < 2 * M; j++)
            if (d[(size_t)i * M * 2 + j] == 1) use(f, (j % 2 == 0) ? "b" : "w", .5 * j + .5 * i, i);
    f << "</svg>";
}
