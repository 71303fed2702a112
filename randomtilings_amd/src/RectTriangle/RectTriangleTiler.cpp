// RectTriangleTiler.cpp — the rectangle-triangle tiler over the HIP library.
//
// Walk follows RectTriangleTiler.cpp:17-92 of the reference and runs through
// rt_recttri_walk.  The start states reproduce the reference's maxHex/minHex/slopeHex
// (RectTriangleTiler.cpp:94-505) value for value, LozengeToRectTriangle its :507-545, and the
// SVG writer its :547-697 byte for byte; the code here derives them from the hexagon's
// geometry (three rhombi for the extremal tilings, six sectors for the sloped one) instead of
// the reference's case chains.
#include "RectTriangleTiler.h"

#include <cmath>
#include <sstream>
#include <stdexcept>

#include "../common/rtapi.hpp"
#include "rt_hip.h"

namespace {

// Hexagon of side N drawn in an M x M array (M = 2N + 1) centred on (N, N), whose row i is
// then shifted right by N + 3 - i and down by 3 into the tiler's M2 x M2 frame (M2 = M + 6
// rounded up to a multiple of 3).  Only cells with |i - j| <= N are ever non-zero, and those
// land inside the frame.
struct HexCanvas {
    int N, M;
    std::vector<int> v;
    explicit HexCanvas(int n) : N(n), M(2 * n + 1), v((size_t)M * M, 0) {}
    void put(int i, int j, int x) { v[(size_t)i * M + j] = x; }
    tiling framed() const {
        const int M2 = M + 6 + (3 - M % 3) % 3;
        tiling out((size_t)M2 * M2, 0);
        for (int i = 0; i < M; ++i)
            for (int j = 0; j < M; ++j) {
                const int x = v[(size_t)i * M + j];
                if (!x) continue;
                const int j2 = j + N + 3 - i;
                if (j2 >= 0 && j2 < M2) out[(size_t)(i + 3) * M2 + j2] = x;
            }
        return out;
    }
};

// Reverse the order of the six face codes of a vertex state: the state of the same patch
// seen after a half turn of the lattice.
int half_turn(int x) {
    int r = 0;
    for (int k = 0; k < 6; ++k) r |= ((x >> (4 * k)) & 0xf) << (4 * (5 - k));
    return r;
}

// The tiling of maximal height: the hexagon split into three rhombi around the centre, each
// filled with one orientation of stacked triangles; `code(i, j)` gives a vertex state in the
// M x M picture (0 outside the hexagon).
int max_interior(int N, int i, int j) {
    if (j - i > N || i - j > N) return 0;
    if (i == N && j == N) return 0x331221;
    if (i == N && j < N) return 0x333222;  // seam between the upper-left and lower rhombi
    if (j == N && i < N) return 0x331311;  // seam between the upper-left and right rhombi
    if (i == j && i > N) return 0x211221;  // seam between the right and lower rhombi
    if (i < N && j < N) return 0x333333;
    if (i < j) return 0x111111;
    return 0x222222;
}

void max_frame(HexCanvas& c) {
    const int N = c.N, M = c.M;
    for (int k = 1; k < N; ++k) {
        c.put(k, 0, 0x003033);      // left side
        c.put(0, k, 0x000333);      // top side
        c.put(k, k + N, 0x100110);  // upper right side
        c.put(k + N, k, 0x022002);  // lower left side
    }
    for (int k = N + 1; k < M - 1; ++k) {
        c.put(k, M - 1, 0x110100);  // right side
        c.put(M - 1, k, 0x222000);  // bottom side
    }
    c.put(0, 0, 0x000033);
    c.put(N, 0, 0x003002);
    c.put(0, N, 0x000310);
    c.put(M - 1, M - 1, 0x210000);
    c.put(N, M - 1, 0x100100);
    c.put(M - 1, N, 0x022000);
}

HexCanvas extremal(int N) {
    HexCanvas c(N);
    for (int i = 1; i < c.M - 1; ++i)
        for (int j = 1; j < c.M - 1; ++j) c.put(i, j, max_interior(N, i, j));
    max_frame(c);
    return c;
}

// ---- the sloped tiling ----------------------------------------------------------------------
//
// Three rays through the centre (row i = N, column j = N, diagonal i = j) cut the hexagon
// into six sectors.  A ray cell's state depends on which side of the centre it lies, whether
// it is next to the centre, and a parity; a sector cell's state on whether it borders one,
// the other or both of the sector's rays, and a parity.

struct RayCodes {
    int far_before[2];   // index: position along the ray odd, even
    int near_before[2];  // index: N even, odd
    int near_after[2];
    int far_after[2];
};
constexpr RayCodes kRowRay = {{0x1be1fc, 0xe81c91}, {0x122133, 0xe81c91}, {0x331221, 0x29c38e}, {0xcf1eb1, 0x19c18e}};
constexpr RayCodes kColRay = {{0xd22afc, 0x7af229}, {0x122133, 0x7af229}, {0x331221, 0x913fa7}, {0xcfa22d, 0x922fa7}};
constexpr RayCodes kDiagRay = {{0x33aebd, 0xbd7833}, {0x331221, 0xbd7812}, {0x122133, 0x3387db}, {0xdbea33, 0x3387db}};

struct SectorCodes {
    int seam_a[2];  // bordering ray a only; index: parity bit set, clear
    int seam_b[2];  // bordering ray b only
    int corner[2];  // bordering both; index: N even, odd
    int bulk[2];
};
// sectors: 0 above the row ray and left of the diagonal (j < i < N), 1 below the row ray and
// right of the diagonal (N < i < j), 2/3 the two sectors between the column ray and the
// diagonal, 4/5 the two between the row and column rays
constexpr SectorCodes kSector[6] = {
    {{0x855be8, 0x58881b}, {0x8eb558, 0x58388e}, {0x583812, 0x8ebbe8}, {0x855558, 0x588885}},
    {{0x8eb558, 0xb18885}, {0x855be8, 0xe88385}, {0x218385, 0x8ebbe8}, {0x855558, 0x588885}},
    {{0x47d47a, 0x7477d2}, {0xa74d74, 0x7473a7}, {0x747312, 0xa7dd7a}, {0x474474, 0x747747}},
    {{0xa74d74, 0x2d7747}, {0x47d47a, 0x7a3747}, {0x213747, 0xa7dd7a}, {0x474474, 0x747747}},
    {{0x6699cf, 0x996f19}, {0xfc9966, 0x296c99}, {0x296319, 0xfc99cf}, {0x669966, 0x996699}},
    {{0xfc9966, 0x91f699}, {0x6699cf, 0x99c692}, {0x913692, 0xfc99cf}, {0x669966, 0x996699}},
};

int ray_code(const RayCodes& r, int N, int k) {
    const int odd = k & 1, nodd = N & 1;
    if (k < N - 1) return r.far_before[odd ? 0 : 1];
    if (k == N - 1) return r.near_before[nodd];
    if (k == N + 1) return r.near_after[nodd];
    return r.far_after[odd ? 0 : 1];
}

int slope_interior(int N, int i, int j) {
    if (i == N && j == N) return (N & 1) ? 0x122133 : 0x213312;
    if (i == N) return ray_code(kRowRay, N, j);
    if (j == N) return ray_code(kColRay, N, i);
    if (i == j) return ray_code(kDiagRay, N, i);
    int s, da, db, par;  // sector, distances to its two rays, parity bit
    if (i < N && j < i) s = 0, da = N - i, db = i - j, par = j & 1;
    else if (i > N && j > i) s = 1, da = i - N, db = j - i, par = j & 1;
    else if (j < N && i < j) s = 2, da = N - j, db = j - i, par = i & 1;
    else if (j > N && i > j) s = 3, da = j - N, db = i - j, par = i & 1;
    else if (i < N && j > N && j - i < N) s = 4, da = N - i, db = j - N, par = ((i + j) & 1) != (N & 1);
    else if (i > N && j < N && i - j < N) s = 5, da = i - N, db = N - j, par = ((i + j) & 1) != (N & 1);
    else return 0;
    const SectorCodes& c = kSector[s];
    const int p = par ? 0 : 1;
    if (da == 1 && db == 1) return c.corner[N & 1];
    if (da == 1) return c.seam_a[p];
    if (db == 1) return c.seam_b[p];
    return c.bulk[p];
}

// Boundary of the sloped tiling, written in this order (for N <= 2 some cells are written
// more than once and the last value stays).
void slope_frame(HexCanvas& c) {
    const int N = c.N, M = c.M;
    for (int k = 2; k < N - 1; ++k) {
        c.put(k, 0, 0x008085);
        c.put(0, k, 0x000747);
    }
    c.put(1, 0, 0x00308e);
    c.put(0, 1, 0x0003a7);
    c.put(N - 1, 0, 0x00801b);
    c.put(0, N - 1, 0x0007d2);
    for (int k = N + 2; k < M - 2; ++k) {
        c.put(k, M - 1, 0x580800);
        c.put(M - 1, k, 0x747000);
    }
    c.put(N + 1, M - 1, 0xb10800);
    c.put(M - 1, N + 1, 0x2d7000);
    c.put(M - 2, M - 1, 0xe80300);
    c.put(M - 1, M - 2, 0x7a3000);
    for (int k = 2; k < N - 1; ++k) {
        c.put(k, k + N, 0x900690);
        c.put(k + N, k, 0x096009);
    }
    c.put(1, N + 1, 0x200c90);
    c.put(N + 1, 1, 0x01f009);
    c.put(N - 1, M - 2, 0x900f10);
    c.put(M - 2, N - 1, 0x09c002);
    const bool unit = N == 1;
    c.put(0, 0, unit ? 0x000012 : 0x000033);
    c.put(M - 1, N, unit ? 0x013000 : 0x022000);
    c.put(N, M - 1, unit ? 0x200300 : 0x100100);
    c.put(0, N, 0x000220);
    c.put(N, 0, 0x001001);
    c.put(M - 1, M - 1, 0x330000);
}

// ---- SVG ------------------------------------------------------------------------------------

double px_x(int i, int j) { return 20 * (i * 1 / 2.0 + j); }
double px_y(int i) { return 20 * (i * sqrt3 / 2.0); }

// How a face code is drawn.  A face is the triangle P, Q, R; p, q, r are the midpoints of the
// sides opposite P, Q, R.  Codes 1-9 fill the whole face with one colour; codes 10-15 split it
// along a median into two halves.  Colours index the palette 1..3, 0 = nothing drawn.
struct FaceArt {
    int8_t whole, pPR, pPQ, rRQ, rRP, qQR, qQP;
    bool PQ, PR, QR, p, q, r;
};
constexpr FaceArt kArt[16] = {
    {0, 0, 0, 0, 0, 0, 0, false, false, false, false, false, false},
    {1, 0, 0, 0, 0, 0, 0, false, true, true, false, false, true},
    {3, 0, 0, 0, 0, 0, 0, true, false, true, false, true, false},
    {2, 0, 0, 0, 0, 0, 0, true, true, false, true, false, false},
    {2, 0, 0, 0, 0, 0, 0, false, false, false, false, false, true},
    {1, 0, 0, 0, 0, 0, 0, false, false, false, false, true, false},
    {3, 0, 0, 0, 0, 0, 0, false, false, false, true, false, false},
    {2, 0, 0, 0, 0, 0, 0, true, false, false, false, false, false},
    {1, 0, 0, 0, 0, 0, 0, false, true, false, false, false, false},
    {3, 0, 0, 0, 0, 0, 0, false, false, true, false, false, false},
    {0, 0, 0, 2, 1, 0, 0, false, true, false, false, false, true},
    {0, 0, 0, 0, 0, 3, 1, false, false, true, false, true, false},
    {0, 3, 2, 0, 0, 0, 0, true, false, false, true, false, false},
    {0, 0, 0, 1, 2, 0, 0, false, false, true, false, false, true},
    {0, 0, 0, 0, 0, 1, 3, true, false, false, false, true, false},
    {0, 2, 3, 0, 0, 0, 0, false, true, false, true, false, false},
};

const char* const kPalette[4] = {
    "",
    "\" style=\"fill:midnightblue;stroke:midnightblue;stroke-width:.4\" />\n",
    "\" style=\"fill:slategrey;stroke:slategrey;stroke-width:.4\" />\n",
    "\" style=\"fill:lightsteelblue;stroke:lightsteelblue;stroke-width:.4\" />\n",
};
const char* const kEdge = " style=\"stroke:rgb(0,0,0);stroke-width:.5\" stroke-linecap = \"round\" />\n";

struct Pt {
    double x, y;
};

void polygon(std::ostream& o, Pt a, Pt b, Pt c, int colour) {
    if (colour <= 0) return;
    o << "<polygon points=\"" << a.x << "," << a.y << " " << b.x << "," << b.y << " " << c.x << "," << c.y << " "
      << kPalette[colour];
}

void line(std::ostream& o, Pt a, Pt b) {
    o << "<line x1=\"" << a.x << "\" y1=\"" << a.y << "\" x2=\"" << b.x << "\" y2=\"" << b.y << "\"" << kEdge;
}

}  // namespace

RectTriangleTiler::RectTriangleTiler(cl::Context context0, cl::CommandQueue queue0, std::vector<cl::Device> devices,
                                     std::string source, cl_int& err)
    : ctx(rtapi::make_ctx(queue0.device())) {
    (void)context0;
    (void)devices;
    err = rtapi::basename(source) == "recttrianglekernel.cl" ? CL_SUCCESS : -1;
}

void RectTriangleTiler::LoadTinyMT(std::string params, int size) {
    rtapi::check(rt_load_params_file(ctx.get(), params.c_str(), size), "LoadTinyMT");
}

void RectTriangleTiler::LoadTinyMTDefault(int size) { rtapi::load_default_params(ctx.get(), size); }

void RectTriangleTiler::SetWeights(const std::vector<float>& w16) {
    if (!w16.empty() && w16.size() != 16) throw std::invalid_argument("SetWeights: 16 weights expected");
    weights = w16;
}

void RectTriangleTiler::Walk(tiling& t, int steps, long seed) {
    const int N = rtapi::side(t.size());
    rtapi::check(rt_recttri_walk(ctx.get(), weights.empty() ? nullptr : weights.data(), t.data(), N, steps, seed),
                 "RectTriangleTiler::Walk");
}

tiling RectTriangleTiler::maxHex(int N) {
    if (N < 0) throw std::invalid_argument("maxHex: negative side");
    return extremal(N).framed();
}

tiling RectTriangleTiler::minHex(int N) {
    if (N < 0) throw std::invalid_argument("minHex: negative side");
    // The minimal tiling is the maximal one turned by half a turn about the centre.
    const HexCanvas mx = extremal(N);
    HexCanvas mn(N);
    const int M = mn.M;
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < M; ++j) mn.put(M - 1 - i, M - 1 - j, half_turn(mx.v[(size_t)i * M + j]));
    if (N == 0) mn.put(0, 0, 0x330000);  // a single vertex: the last corner value written
    return mn.framed();
}

tiling RectTriangleTiler::slopeHex(int N) {
    if (N < 1) throw std::invalid_argument("slopeHex: side must be at least 1");
    HexCanvas c(N);
    for (int i = 1; i < c.M - 1; ++i)
        for (int j = 1; j < c.M - 1; ++j) c.put(i, j, slope_interior(N, i, j));
    slope_frame(c);
    return c.framed();
}

// Each lozenge of the lozenge tiling (states on vertices, LozengeTiler's convention) is cut
// into the two triangles of the M x 2M domain; a triangle's face code is its lozenge's
// orientation (1, 2 or 3), written into the matching hex digit of its three corner vertices.
tiling RectTriangleTiler::LozengeToRectTriangle(const tiling& tL, const domain& dL) {
    const int N = rtapi::side(tL.size());
    const int M = (int)std::sqrt((double)(dL.size() / 2));
    tiling t((size_t)N * N, 0);
    auto add = [&](int i, int j, int digit, int code) { t[(size_t)i * N + j] += code << (4 * digit); };
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < 2 * M; ++j) {
            if (dL[(size_t)i * 2 * M + j] != 1) continue;
            const int c = j / 2;
            if (j % 2 == 0) {  // corners (i,c), (i,c+1), (i+1,c)
                const int s = tL[(size_t)i * N + c];
                const int code = (s & 8) ? 1 : (s & 32) ? 3 : 2;
                add(i, c, 0, code);
                add(i, c + 1, 2, code);
                add(i + 1, c, 4, code);
            } else {  // corners (i,c+1), (i+1,c), (i+1,c+1)
                const int s = tL[(size_t)i * N + c + 1];
                const int code = (s & 16) ? 2 : (s & 32) ? 3 : 1;
                add(i, c + 1, 1, code);
                add(i + 1, c, 3, code);
                add(i + 1, c + 1, 5, code);
            }
        }
    return t;
}

std::string RectTriangleTiler::TilePicture(int TriType, int i, int j) {
    const FaceArt& a = kArt[(TriType >= 0 && TriType < 16) ? TriType : 0];
    Pt P, Q, R;
    if (j % 2 == 1) {
        P = {px_x(i + 1, j / 2), px_y(i + 1)};
        Q = {px_x(i + 1, j / 2 + 1), px_y(i + 1)};
        R = {px_x(i, j / 2 + 1), px_y(i)};
    } else {
        P = {px_x(i, j / 2 + 1), px_y(i)};
        Q = {px_x(i, j / 2), px_y(i)};
        R = {px_x(i + 1, j / 2), px_y(i + 1)};
    }
    const Pt p = {(R.x + Q.x) / 2, (R.y + Q.y) / 2};
    const Pt q = {(P.x + R.x) / 2, (P.y + R.y) / 2};
    const Pt r = {(P.x + Q.x) / 2, (P.y + Q.y) / 2};
    std::stringstream o;
    polygon(o, P, Q, R, a.whole);
    polygon(o, P, p, R, a.pPR);
    polygon(o, P, p, Q, a.pPQ);
    polygon(o, R, r, Q, a.rRQ);
    polygon(o, R, r, P, a.rRP);
    polygon(o, Q, q, R, a.qQR);
    polygon(o, Q, q, P, a.qQP);
    if (a.PQ) line(o, P, Q);
    if (a.PR) line(o, P, R);
    if (a.QR) line(o, R, Q);
    if (a.p) line(o, P, p);
    if (a.q) line(o, Q, q);
    if (a.r) line(o, R, r);
    return o.str();
}

void RectTriangleTiler::TilingToSVG(const tiling& t, std::string filename) {
    std::ofstream out(filename.c_str());
    const int N = (int)std::sqrt((double)t.size());
    const double W = sqrt3 * 2 * (N + 1), H = 3 / 2.0 * (N + 1), s = 20;
    out << "<svg xmlns='http://www.w3.org/2000/svg' height=\"" << s * H << "\" width=\"" << s * W << "\" >\n";
    for (int i = 0; i < N - 1; ++i)
        for (int j = 0; j < N - 1; ++j) {
            out << TilePicture((t[(size_t)i * N + j + 1] >> 4) & 0xf, i, 2 * j + 1);
            out << TilePicture(t[(size_t)i * N + j] & 0xf, i, 2 * j);
        }
    out << "</svg>";
}
