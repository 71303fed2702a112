#ifndef TRIANGLEDIMER_TRIANGLEDIMERTILER_H_
#define TRIANGLEDIMER_TRIANGLEDIMERTILER_H_

// TriangleDimerTiler.h — dimer covers (perfect matchings) of regions of the triangular
// lattice, on MI355X.
//
// Same class, method names and signatures as the reference's
// src/TriangleDimer/TriangleDimerTiler.h.
//
// The chain uses the three local moves of Kenyon and Remila: lozenge moves (two parallel
// dimers in a rhombus rotate), triangle moves (three dimers around a triangle of the
// lattice rotate) and butterfly moves.  A tiling is an N x N int array (N even, three
// rows/columns of zero padding) of lozenge-move states on the medial lattice: at even row /
// odd column the state of a horizontal rhombus (H), at odd row / even column an up-left one
// (L), at odd row / odd column an up-right one (R); each state holds the dimer indicators of
// the rhombus' four sides (bits 1, 2, 4, 8).  Triangle-move and butterfly-move states are
// derived from it (TilingToTriangleFlips, TilingToButterflyFlips).

#include <memory>
#include <string>

#include "../common/common.h"

struct rt_ctx;

class TriangleDimerTiler {
   private:
    std::shared_ptr<rt_ctx> ctx;

   public:
    TriangleDimerTiler(cl::Context context0, cl::CommandQueue queue0, std::vector<cl::Device> devices,
                       std::string source, cl_int &err);

    void LoadTinyMT(std::string params, int size);
    void LoadTinyMTDefault(int size);

    void Walk(tiling &t, int steps, long seed);

    // A triangle of side N edges pointing down with a triangle of side M cut from its top
    // left corner, covered by a fixed dimer configuration (tiling side 2N + 8).
    static tiling IceCreamCone(int N, int M);
    // A hexagon of the triangular lattice with sides x, y, z, x + y - z, z, y (a top row of
    // x + 1 vertices, y rows growing by one, then z rows shrinking by one) covered by dimers,
    // as a tiling of side 2 (max(y + z + 1, x + y + 1) + 5).  A hexagon with an odd number of
    // vertices has no dimer cover: the message the reference prints for a cone without one is
    // printed and an all-zero tiling returned.  Not in the reference (its only region is the
    // cone).
    static tiling Hexagon(int x, int y, int z);
    // A single butterfly of orientation 0 (H), 1 (L) or 2 (R) in a 16 x 16 tiling.
    static tiling TestButterfly(int type);

    // The triangle-move states (an N/2 x N array) and butterfly-move states (N x N) of a
    // tiling.
    static std::vector<int> TilingToTriangleFlips(const tiling &t);
    static std::vector<int> TilingToButterflyFlips(const tiling &t);
    // Vertices of the lattice covered by a dimer ((N/2+1) x (N/2+1), 1 = covered).
    static std::vector<int> TilingToVertices(const tiling &t);

    // Indicator of the dimers of one orientation: c = 0 horizontal (odd rows), c = 1 / 2 the
    // bit-8 / bit-4 sides of the even-row states.
    static std::vector<int> DimerIndicator(const tiling &t, int c);

    // Triangles whose three corners are covered vertices (an (N/2) x N array).
    static domain TilingToDomain(const tiling &t);

    static void TilingToSVG(const tiling &t, std::string filename);
    static void DimerToSVG(const tiling &t, std::string filename);
    static void DimerToSVG(const tiling &t, const domain &d, std::string filename);
    static void DomainToSVG(const domain &d, std::string filename);

    rt_ctx *Handle() const { return ctx.get(); }
};

#endif /* TRIANGLEDIMER_TRIANGLEDIMERTILER_H_ */
