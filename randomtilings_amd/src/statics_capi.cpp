// statics_capi.cpp — rt_statics.h: C dispatch onto the tiler classes' static utilities.
#include <cstring>
#include <string>
#include <vector>

#include "Domino/DominoTiler.h"
#include "Lozenge/LozengeTiler.h"
#ifdef RT_HAVE_RECTTRI
#include "RectTriangle/RectTriangleTiler.h"
#endif
#ifdef RT_HAVE_TRIDIMER
#include "TriangleDimer/TriangleDimerTiler.h"
#endif
#include "common/common.h"
#include "rt_statics.h"

static int64_t emit(const std::vector<int>& v, int32_t* out, int64_t cap) {
    if ((int64_t)v.size() > cap) return -2;
    if (!v.empty()) std::memcpy(out, v.data(), v.size() * sizeof(int));
    return (int64_t)v.size();
}

extern "C" int64_t rt_static_call(const char* name, const int32_t* in, int64_t nin, const int32_t* in2, int64_t nin2,
                                  const int32_t* args, int32_t nargs, int32_t* out, int64_t cap) {
    const std::string f(name ? name : "");
    std::vector<int> a(in, in + (in ? nin : 0));
    std::vector<int> b(in2, in2 + (in2 ? nin2 : 0));
    auto arg = [&](int k) { return k < nargs ? args[k] : 0; };
    if (f == "domino.Rectangle") return emit(DominoTiler::Rectangle(arg(0), arg(1)), out, cap);
    if (f == "domino.AztecDiamond") return emit(DominoTiler::AztecDiamond(arg(0)), out, cap);
    if (f == "domino.AztecRectangle") return emit(DominoTiler::AztecRectangle(arg(0), a), out, cap);
    if (f == "domino.MaxTiling") return emit(DominoTiler::MaxTiling(a), out, cap);
    if (f == "domino.MinTiling") return emit(DominoTiler::MinTiling(a), out, cap);
    if (f == "domino.DomainToVertices") return emit(DominoTiler::DomainToVertices(a), out, cap);
    if (f == "domino.TilingToDomain") return emit(DominoTiler::TilingToDomain(a), out, cap);
    if (f == "domino.HeightfuncToTiling") return emit(DominoTiler::HeightfuncToTiling(a, b), out, cap);
    if (f == "domino.TilingToHeightfunc") return emit(DominoTiler::TilingToHeightfunc(a, b), out, cap);
    if (f == "lozenge.AlmostHexDomain") return emit(LozengeTiler::AlmostHexDomain(arg(0), arg(1)), out, cap);
    if (f == "lozenge.Hexagon") return emit(LozengeTiler::Hexagon(arg(0)), out, cap);
    if (f == "lozenge.MaxTiling") return emit(LozengeTiler::MaxTiling(a), out, cap);
    if (f == "lozenge.MinTiling") return emit(LozengeTiler::MinTiling(a), out, cap);
    if (f == "lozenge.DomainToVertices") return emit(LozengeTiler::DomainToVertices(a), out, cap);
    if (f == "lozenge.TilingToDomain") return emit(LozengeTiler::TilingToDomain(a), out, cap);
    if (f == "lozenge.HeightfuncToTiling") return emit(LozengeTiler::HeightfuncToTiling(a, b), out, cap);
    if (f == "lozenge.TilingToHeightfunc") return emit(LozengeTiler::TilingToHeightfunc(a, b), out, cap);
#ifdef RT_HAVE_RECTTRI
    if (f == "recttri.maxHex") return emit(RectTriangleTiler::maxHex(arg(0)), out, cap);
    if (f == "recttri.minHex") return emit(RectTriangleTiler::minHex(arg(0)), out, cap);
    if (f == "recttri.slopeHex") return emit(RectTriangleTiler::slopeHex(arg(0)), out, cap);
    if (f == "recttri.LozengeToRectTriangle") return emit(RectTriangleTiler::LozengeToRectTriangle(a, b), out, cap);
#endif
#ifdef RT_HAVE_TRIDIMER
    if (f == "tridimer.IceCreamCone") return emit(TriangleDimerTiler::IceCreamCone(arg(0), arg(1)), out, cap);
    if (f == "tridimer.Hexagon") return emit(TriangleDimerTiler::Hexagon(arg(0), arg(1), arg(2)), out, cap);
    if (f == "tridimer.TestButterfly") return emit(TriangleDimerTiler::TestButterfly(arg(0)), out, cap);
    if (f == "tridimer.TilingToTriangleFlips") return emit(TriangleDimerTiler::TilingToTriangleFlips(a), out, cap);
    if (f == "tridimer.TilingToButterflyFlips") return emit(TriangleDimerTiler::TilingToButterflyFlips(a), out, cap);
    if (f == "tridimer.TilingToVertices") return emit(TriangleDimerTiler::TilingToVertices(a), out, cap);
    if (f == "tridimer.TilingToDomain") return emit(TriangleDimerTiler::TilingToDomain(a), out, cap);
    if (f == "tridimer.DimerIndicator") return emit(TriangleDimerTiler::DimerIndicator(a, arg(0)), out, cap);
#endif
    if (f == "common.PadMatrix") {
        PadMatrix(a, arg(0), arg(1));
        return emit(a, out, cap);
    }
    if (f == "common.UnPadMatrix") {
        UnPadMatrix(a, arg(0), arg(1));
        return emit(a, out, cap);
    }
    return -1;
}

extern "C" int rt_static_write(const char* name, const int32_t* in, int64_t nin, const int32_t* in2, int64_t nin2,
                               const char* path) {
    const std::string f(name ? name : "");
    std::vector<int> a(in, in + (in ? nin : 0));
    std::vector<int> b(in2, in2 + (in2 ? nin2 : 0));
    if (f == "domino.DomainToSVG") DominoTiler::DomainToSVG(a, path);
    else if (f == "domino.TilingToSVG") DominoTiler::TilingToSVG(a, path);
    else if (f == "domino.MayaToSVG") DominoTiler::MayaToSVG(a, path);
    else if (f == "domino.DimerToSVG") DominoTiler::DimerToSVG(a, path);
    else if (f == "domino.DimerToSVG2") DominoTiler::DimerToSVG(a, b, path);
    else if (f == "lozenge.DomainToSVG") LozengeTiler::DomainToSVG(a, path);
    else if (f == "lozenge.TilingToSVG") LozengeTiler::TilingToSVG(a, path);
    else if (f == "lozenge.DimerToSVG") LozengeTiler::DimerToSVG(a, path);
    else if (f == "lozenge.DimerToSVG2") LozengeTiler::DimerToSVG(a, b, path);
    else if (f == "lozenge.PrintDomain") LozengeTiler::PrintDomain(a, path);
#ifdef RT_HAVE_RECTTRI
    else if (f == "recttri.TilingToSVG") RectTriangleTiler::TilingToSVG(a, path);
#endif
#ifdef RT_HAVE_TRIDIMER
    else if (f == "tridimer.TilingToSVG") TriangleDimerTiler::TilingToSVG(a, path);
    else if (f == "tridimer.DimerToSVG") TriangleDimerTiler::DimerToSVG(a, path);
    else if (f == "tridimer.DimerToSVG2") TriangleDimerTiler::DimerToSVG(a, b, path);
    else if (f == "tridimer.DomainToSVG") TriangleDimerTiler::DomainToSVG(a, path);
#endif
    else if (f == "common.SaveMatrix") SaveMatrix(a, path);
    else return -1;
    return 0;
}
