// A caller of the reference-compatible C++ API, written the way the reference's example
// programs use it (Examples/MinimalDominoEx.cpp, MinimalLozengeEx.cpp, RectTriangleEx.cpp,
// TriangleDimerEx.cpp, AztecDiamondCftpEx.cpp): an OpenCL-style context, queue and device
// list, a tiler constructed with a kernel path, LoadTinyMT(path, n) on a tinymtdc-format
// parameter file, a domain, an extremal start and Walk.  It writes the walked tilings as
// plain integers, one tiling per file, for the test to compare with the oracle.
//
//   tiler_walks <outdir> <tinymtdc parameter file>
#include <cstdio>
#include <fstream>
#include <random>
#include <string>
#include <vector>

#include "../../randomtilings_amd/src/Domino/DominoTiler.h"
#include "../../randomtilings_amd/src/Lozenge/LozengeTiler.h"
#include "../../randomtilings_amd/src/RectTriangle/RectTriangleTiler.h"
#include "../../randomtilings_amd/src/TriangleDimer/TriangleDimerTiler.h"

static void save(const tiling& t, const std::string& path) {
    std::ofstream f(path.c_str());
    for (size_t i = 0; i < t.size(); i++) f << t[i] << (i + 1 < t.size() ? ' ' : '\n');
}

static int side(const tiling& t) { return (int)(std::sqrt((double)t.size()) + 0.5); }

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    const std::string out = argv[1];
    const std::string params = argv[2];
    cl::Context context(CL_DEVICE_TYPE_DEFAULT);
    std::vector<cl::Device> devices;
    context.getInfo(CL_CONTEXT_DEVICES, &devices);
    std::string name;
    devices[0].getInfo(CL_DEVICE_NAME, &name);
    cl::CommandQueue queue(context);
    cl_int err = 0;

    // domino: Aztec diamond of order 40, CFTP kernel rule, from the minimal tiling
    DominoTiler dom(context, queue, devices, "./src/Domino/dominokernelCFTP.cl", err);
    domain d = DominoTiler::AztecDiamond(40);
    tiling t = DominoTiler::MinTiling(d);
    dom.LoadTinyMT(params, side(t) * side(t) / 2);
    dom.Walk(t, 3001, 20261203);
    save(t, out + "/domino.txt");

    // domino epochs: Walk(t, steps, seeds) on an Aztec diamond of order 12 from the maximal tiling
    tiling te = DominoTiler::MaxTiling(DominoTiler::AztecDiamond(12));
    std::vector<long> esteps = {256, 128, 77}, eseeds = {345903, 17, 1099511627781L};
    dom.Walk(te, esteps, eseeds);
    save(te, out + "/domino_epochs.txt");

    // coupling from the past exactly as Examples/AztecDiamondCftpEx.cpp:42-68 runs it on Walk
    // (Aztec diamond of order 5, seed 345903), and through the library call
    domain d5 = DominoTiler::AztecDiamond(5);
    tiling tMax = DominoTiler::MaxTiling(d5), tMin = DominoTiler::MinTiling(d5);
    std::mt19937 mt_rand(345903);
    std::vector<long> steps, seeds;
    steps.push_back(128);
    seeds.push_back(mt_rand());
    int restarts = 0;
    for (;; restarts++) {
        tiling top = tMax, bottom = tMin;
        dom.Walk(bottom, steps, seeds);
        dom.Walk(top, steps, seeds);
        if (top == bottom) {
            save(top, out + "/cftp_loop.txt");
            break;
        }
        steps.insert(steps.begin(), steps[steps.size() - 1] * 2);
        seeds.insert(seeds.begin(), mt_rand());
    }
    long total = 0;
    tiling s = dom.CFTP(tMax, tMin, 345903, 128, false, 64, nullptr, &total);
    save(s, out + "/cftp_call.txt");
    std::printf("cftp restarts %d total %ld\n", restarts, total);

    // lozenge: hexagon of side 30, q-volume kernel with q = 0.9, from the maximal tiling
    LozengeTiler loz(context, queue, devices, "./src/Lozenge/qlozengekernel.cl", err);
    domain h = LozengeTiler::AlmostHexDomain(30, 0);
    tiling u = LozengeTiler::MaxTiling(h);
    loz.LoadTinyMT(params, side(u) * (side(u) / 3));
    loz.SetQ(0.9);
    loz.Walk(u, 2002, 7);
    save(u, out + "/lozenge.txt");

    // rectangle-triangle: hexagon of side 12 from the minimal tiling (RectTriangleEx.cpp)
    RectTriangleTiler rtt(context, queue, devices, "./src/RectTriangle/recttrianglekernel.cl", err);
    tiling r = RectTriangleTiler::minHex(12);
    rtt.LoadTinyMT(params, (int)r.size());
    rtt.Walk(r, 1500, 31);
    save(r, out + "/recttri.txt");

    // triangle dimers: ice-cream cone of side 14 (TriangleDimerEx.cpp)
    TriangleDimerTiler tdt(context, queue, devices, "./src/TriangleDimer/triangledimerkernel.cl", err);
    tiling c = TriangleDimerTiler::IceCreamCone(14, 0);
    tdt.LoadTinyMT(params, (int)c.size());
    tdt.Walk(c, 500, 4242);
    save(c, out + "/tridimer.txt");

    std::printf("walked on %s\n", name.c_str());
    return 0;
}
