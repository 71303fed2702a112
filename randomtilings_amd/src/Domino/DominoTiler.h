#ifndef DOMINO_DOMINOTILER_H_
#define DOMINO_DOMINOTILER_H_

/*
 * DominoTiler.h — Markov chain simulation of domino tilings on MI355X, with utilities for
 * height functions, maximal/minimal tilings, standard domains and SVG drawing.
 *
 * Same class, method names and signatures as the reference's src/Domino/DominoTiler.h, so
 * callers swap this library in without source changes.  The chain runs as gfx950 HIP
 * kernels behind the C ABI in include/rt_hip.h.
 *
 * A domain is an M x M 0/1 array of the squares of the lattice.  A tiling is an N x N
 * (N = M + 1, even) array of vertex states: the sum of the indicators of the four edges at
 * the vertex that a domino crosses,
 *      1 = edge to (i-1, j)   2 = edge to (i+1, j)   4 = edge to (i, j-1)   8 = edge to (i, j+1)
 * (DominoTiler.cpp:204-206 of the reference; its header diagram is drawn mirrored).  The
 * tiling must be twice zero-padded.
 *
 * The chain rule is chosen by the basename of the kernel path the constructor receives:
 * "dominokernel.cl" (lazy: toggle with probability 0.8) or "dominokernelCFTP.cl"
 * (monotone heat-bath at 1/2, the rule used by coupling from the past).  No file is read.
 */

#include <stdint.h>

#include <memory>

#include "../common/common.h"

struct rt_ctx;

class DominoTiler {
   private:
    std::shared_ptr<rt_ctx> ctx;
    int rule;

   public:
    // The kernel path selects the chain rule; err is set to CL_SUCCESS (0) or -1 when the
    // basename names no domino rule.
    DominoTiler(cl::Context context0, cl::CommandQueue queue0, std::vector<cl::Device> devices, std::string source,
                cl_int &err);

    // Loads `size` TinyMT parameter sets from a tinymtdc-format file.
    void LoadTinyMT(std::string params, int size);
    // Uses TinyMT's default parameter set for `size` generators (seeded seed + id).
    void LoadTinyMTDefault(int size);

    // Random walk of the tiling for STEPS steps with seed SEED.
    void Walk(tiling &t, long steps, long seed);

    // Random walk with a vector of steps and a vector of seeds (one epoch per entry).
    void Walk(tiling &t, std::vector<long> steps, std::vector<long> seeds);

    // Weighted heat-bath walk: P(state 12) for each (colour class, row parity), class 0 being
    // the vertices with i+j even.  Uniform p12 = 1/2 is the dominokernelCFTP.cl rule.
    void WalkWeighted(tiling &t, long steps, long seed, const double p12[4]);

    // Coupling from the past: an exact sample of the chain's stationary law, drawn the way
    // Examples/AztecDiamondCftpEx.cpp does with Walk (epoch seeds from std::mt19937(seed),
    // epochs of first_steps and then 2*first_steps prepended per restart; `doubling` doubles
    // the total instead).  tMax/tMin are the domain's maximal and minimal tilings.  The
    // tiler's rule must be monotone (dominokernelCFTP.cl); p12 selects the weighted rule.
    // Throws std::runtime_error when max_restarts restarts do not coalesce.
    tiling CFTP(const tiling &tMax, const tiling &tMin, long seed, long first_steps = 128, bool doubling = false,
                int max_restarts = 64, const double *p12 = nullptr, long *total_steps = nullptr);

    // The rule this tiler runs (RT_DOMINO_LAZY or RT_DOMINO_CFTP of rt_hip.h).
    int Rule() const { return rule; }
    rt_ctx *Handle() const { return ctx.get(); }

    // An M x N rectangle of squares.
    static domain Rectangle(int M, int N);

    // An Aztec rectangle with the boundary points REMOVE removed (Knizel-Bufetov).
    static domain AztecRectangle(int N, const std::vector<int> &remove);

    // The Aztec diamond of order N.
    static domain AztecDiamond(int N);

    // Maximal and minimal tilings (Thurston's construction).  A domain without a tiling is
    // reported on stdout, as in the reference, and an all-zero tiling is returned.
    static tiling MaxTiling(const domain &d);
    static tiling MinTiling(const domain &d);

    // Vertex classes of a domain: 1 interior, 2 boundary, 0 outside.
    static std::vector<int> DomainToVertices(const domain &d);

    // The domain a tiling covers.
    static domain TilingToDomain(const tiling &t);

    // Between tilings and height functions.
    static tiling HeightfuncToTiling(const heightfunc &hf, const domain &d);
    static heightfunc TilingToHeightfunc(const tiling &t, const domain &d);

    // SVG drawing.
    static void DomainToSVG(const domain &d, std::string filename);
    static void TilingToSVG(const tiling &t, std::string filename);
    static void MayaToSVG(const tiling &t, std::string filename);
    static void DimerToSVG(const tiling &t, std::string filename);
    static void DimerToSVG(const tiling &t, const domain &d, std::string filename);
};

#endif /* DOMINO_DOMINOTILER_H_ */
