#ifndef RECTTRIANGLE_RECTTRIANGLETILER_H_
#define RECTTRIANGLE_RECTTRIANGLETILER_H_

// RectTriangleTiler.h — tilings by rectangles and triangles on the triangular lattice, on
// MI355X.
//
// Same class, method names and signatures as the reference's
// src/RectTriangle/RectTriangleTiler.h.
//
// The tiling lives on the vertices of the triangular lattice (an N x N int array, N divisible
// by 3, three rows/columns of zero padding).  Each of the six faces around a vertex carries a
// code in [0, 15] describing how the tiling covers it; a vertex state packs the six codes as
// hex digits d5..d0:
//        d4
//   d5        d3
//   d2        d0
//        d1
// A Markov-chain step picks one of the three vertex colour classes ((i - j) mod 3), tries a
// local move at every vertex of that class (a rotation among five three-tile patches, or one
// of six paired toggles) and rewrites the face codes of the two other classes.

#include <memory>
#include <string>

#include "../common/common.h"

struct rt_ctx;

class RectTriangleTiler {
   private:
    std::shared_ptr<rt_ctx> ctx;
    std::vector<float> weights;  // 16 face-code weights; empty = the reference's unit weights

   public:
    RectTriangleTiler(cl::Context context0, cl::CommandQueue queue0, std::vector<cl::Device> devices,
                      std::string source, cl_int &err);

    void LoadTinyMT(std::string params, int size);
    void LoadTinyMTDefault(int size);

    // Weights of the 16 face codes (index 0 = uncovered face, normally 0); moves are accepted
    // with the ratio of the products of the weights, as recttrianglekernel.cl's
    // getWeightRatio computes it.  Without a call every weight is 1 except code 0.
    void SetWeights(const std::vector<float> &w16);

    void Walk(tiling &t, int steps, long seed);

    // Start states on a hexagon of side N (vertex array side 2N + 7, rounded up to a
    // multiple of 3): the tilings of maximal and minimal height and a sloped one.
    static tiling slopeHex(int N);
    static tiling maxHex(int N);
    static tiling minHex(int N);

    // The rectangle-triangle tiling that subdivides each lozenge of a lozenge tiling.
    static tiling LozengeToRectTriangle(const tiling &tL, const domain &dL);

    static void TilingToSVG(const tiling &t, std::string filename);
    static std::string TilePicture(int TriType, int i, int j);

    rt_ctx *Handle() const { return ctx.get(); }
};

#endif /* RECTTRIANGLE_RECTTRIANGLETILER_H_ */
