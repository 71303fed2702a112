
#ifndef LOZENGE_LOZENGETILER_H_
#define LOZENGE_LOZENGETILER_H_

// LozengeTiler.h — lozenge tilings (dimers on the hexagonal lattice) on MI355X.
//
// Same class, method names and signatures as the reference's src/Lozenge/LozengeTiler.h.
//
// The domain is a union of triangles of the triangular lattice, an M x 2M 0/1 array.  The
// tiling lives on the vertices (an N x N array, N = M + 1, N divisible by 3, with three
// rows/columns of zero padding); each vertex state sums the indicators of its crossed
// edges:
//        1   2
//     4 __\./__ 8
//         / \
//       16   32
// A step flips one of the three vertex colour classes ((j - i) mod 3) and updates the
// other two.  The kernel path's basename picks the rule: "lozengekernel.cl" (uniform) or
// "qlozengekernel.cl" (q-volume, with q settable at run time via SetQ).

#include <memory>

#include "../common/common.h"

struct rt_ctx;

class LozengeTiler {
   private:
    std::shared_ptr<rt_ctx> ctx;
    int rule;
    double q;

   public:
    LozengeTiler(cl::Context context0, cl::CommandQueue queue0, std::vector<cl::Device> devices, std::string source,
                 cl_int &err);

    void LoadTinyMT(std::string params, int size);
    void LoadTinyMTDefault(int size);

    // q of the q-volume rules (qlozengekernel.cl hard-codes 0.99); heat-bath selects the
    // add-with-probability-q/(1+q) rule instead of the reference's.
    void SetQ(double q_, bool heatbath = false);

    void Walk(tiling &t, int samples, long seed);

    // A regular hexagon of side N (the reference declares this without defining it).
    static domain Hexagon(int N);

    static tiling MaxTiling(const domain &d);
    static tiling MinTiling(const domain &d);

    static tiling HeightfuncToTiling(const heightfunc &hf, const domain &d);
    static heightfunc TilingToHeightfunc(const tiling &t, const domain &d);
    static std::vector<int> DomainToVertices(const domain &d);
    static domain TilingToDomain(const tiling &t);
    // A hexagon of side N with an M x M corner removed, padded so the tiling side is a
    // multiple of 3.
    static domain AlmostHexDomain(int N, int M);

    static void DomainToSVG(const domain &d, std::string filename);
    static void DimerToSVG(const tiling &t, std::string filename);
    static void DimerToSVG(const tiling &t, const domain &d, std::string filename);
    static void TilingToSVG(const tiling &t, std::string filename);
    static void PrintDomain(const domain &d, std::string filename);

    rt_ctx *Handle() const { return ctx.get(); }
    int Rule() const { return rule; }
};

#endif /* LOZENGE_LOZENGETILER_H_ */
