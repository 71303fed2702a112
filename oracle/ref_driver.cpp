// ref_driver.cpp — C entry points over the reference's own host-side statics, compiled
// from the reference sources where they lie (see oracle/Makefile) into oracle/_ref/.
// TEST INFRASTRUCTURE ONLY: it lets the tests pin this project's C++ restatement of the
// domain/tiling/height/SVG utilities against the reference itself.  The reference's
// Walk/OpenCL code is linked but never called (there is no OpenCL device).
//
// Every function here takes and returns flat int arrays; `ref_call` dispatches by name so
// the ctypes side stays small.  Returns the output length, or -1 (unknown name) / -2
// (output buffer too small).
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "Domino/DominoTiler.h"
#include "Lozenge/LozengeTiler.h"
#include "RectTriangle/RectTriangleTiler.h"
#include "TriangleDimer/TriangleDimerTiler.h"
#include "common/common.h"
#include "TinyMT/tinymt32.h"
#include "TinyMT/file_reader.h"

static int64_t put(const std::vector<int>& v, int32_t* out, int64_t cap) {
    if ((int64_t)v.size() > cap) return -2;
    if (!v.empty()) std::memcpy(out, v.data(), v.size() * sizeof(int));
    return (int64_t)v.size();
}

extern "C" {

int64_t ref_call(const char* name, const int32_t* in, int64_t nin, const int32_t* in2, int64_t nin2,
                 const int32_t* args, int nargs, int32_t* out, int64_t cap) {
    std::string f(name);
    std::vector<int> a(in, in + (in ? nin : 0));
    std::vector<int> b(in2, in2 + (in2 ? nin2 : 0));
    auto arg = [&](int k) { return k < nargs ? args[k] : 0; };
    // Domino
    if (f == "domino.Rectangle") return put(DominoTiler::Rectangle(arg(0), arg(1)), out, cap);
    if (f == "domino.AztecDiamond") return put(DominoTiler::AztecDiamond(arg(0)), out, cap);
    if (f == "domino.AztecRectangle") return put(DominoTiler::AztecRectangle(arg(0), a), out, cap);
    if (f == "domino.MaxTiling") return put(DominoTiler::MaxTiling(a), out, cap);
    if (f == "domino.MinTiling") return put(DominoTiler::MinTiling(a), out, cap);
    if (f == "domino.DomainToVertices") return put(DominoTiler::DomainToVertices(a), out, cap);
    if (f == "domino.TilingToDomain") return put(DominoTiler::TilingToDomain(a), out, cap);
    if (f == "domino.HeightfuncToTiling") return put(DominoTiler::HeightfuncToTiling(a, b), out, cap);
    if (f == "domino.TilingToHeightfunc") return put(DominoTiler::TilingToHeightfunc(a, b), out, cap);
    // Lozenge
    if (f == "lozenge.AlmostHexDomain") return put(LozengeTiler::AlmostHexDomain(arg(0), arg(1)), out, cap);
    if (f == "lozenge.MaxTiling") return put(LozengeTiler::MaxTiling(a), out, cap);
    if (f == "lozenge.MinTiling") return put(LozengeTiler::MinTiling(a), out, cap);
    if (f == "lozenge.DomainToVertices") return put(LozengeTiler::DomainToVertices(a), out, cap);
    if (f == "lozenge.TilingToDomain") return put(LozengeTiler::TilingToDomain(a), out, cap);
    if (f == "lozenge.HeightfuncToTiling") return put(LozengeTiler::HeightfuncToTiling(a, b), out, cap);
    if (f == "lozenge.TilingToHeightfunc") return put(LozengeTiler::TilingToHeightfunc(a, b), out, cap);
    // Rectangle-triangle
    if (f == "recttri.maxHex") return put(RectTriangleTiler::maxHex(arg(0)), out, cap);
    if (f == "recttri.minHex") return put(RectTriangleTiler::minHex(arg(0)), out, cap);
    if (f == "recttri.slopeHex") return put(RectTriangleTiler::slopeHex(arg(0)), out, cap);
    if (f == "recttri.LozengeToRectTriangle") return put(RectTriangleTiler::LozengeToRectTriangle(a, b), out, cap);
    // Triangle-dimer
    if (f == "tridimer.IceCreamCone") return put(TriangleDimerTiler::IceCreamCone(arg(0), arg(1)), out, cap);
    if (f == "tridimer.TestButterfly") return put(TriangleDimerTiler::TestButterfly(arg(0)), out, cap);
    if (f == "tridimer.TilingToTriangleFlips") return put(TriangleDimerTiler::TilingToTriangleFlips(a), out, cap);
    if (f == "tridimer.TilingToButterflyFlips") return put(TriangleDimerTiler::TilingToButterflyFlips(a), out, cap);
    if (f == "tridimer.TilingToVertices") return put(TriangleDimerTiler::TilingToVertices(a), out, cap);
    if (f == "tridimer.TilingToDomain") return put(TriangleDimerTiler::TilingToDomain(a), out, cap);
    if (f == "tridimer.DimerIndicator") return put(TriangleDimerTiler::DimerIndicator(a, arg(0)), out, cap);
    // common
    if (f == "common.PadMatrix") {
        PadMatrix(a, arg(0), arg(1));
        return put(a, out, cap);
    }
    if (f == "common.UnPadMatrix") {
        UnPadMatrix(a, arg(0), arg(1));
        return put(a, out, cap);
    }
    return -1;
}

// SVG / text writers: kind selects the writer; returns 0 on success, -1 if unknown.
int ref_write(const char* name, const int32_t* in, int64_t nin, const int32_t* in2, int64_t nin2, const char* path) {
    std::string f(name);
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
    else if (f == "recttri.TilingToSVG") RectTriangleTiler::TilingToSVG(a, path);
    else if (f == "tridimer.TilingToSVG") TriangleDimerTiler::TilingToSVG(a, path);
    else if (f == "tridimer.DimerToSVG") TriangleDimerTiler::DimerToSVG(a, path);
    else if (f == "tridimer.DimerToSVG2") TriangleDimerTiler::DimerToSVG(a, b, path);
    else if (f == "tridimer.DomainToSVG") TriangleDimerTiler::DomainToSVG(a, path);
    else if (f == "common.SaveMatrix") SaveMatrix(a, path);
    else return -1;
    return 0;
}

// The reference's branchless host TinyMT coding (src/TinyMT/tinymt32.h:72-109): step the
// given state n times and return n tempered outputs.  Cross-checks the oracle's coding.
void ref_tinymt_host(uint32_t s0, uint32_t s1, uint32_t s2, uint32_t s3, uint32_t mat1, uint32_t mat2, uint32_t tmat,
                     int n, uint32_t* out) {
    tinymt32_t r;
    r.status[0] = s0;
    r.status[1] = s1;
    r.status[2] = s2;
    r.status[3] = s3;
    r.mat1 = mat1;
    r.mat2 = mat2;
    r.tmat = tmat;
    for (int i = 0; i < n; i++) {
        tinymt32_next_state(&r);
        out[i] = tinymt32_temper(&r);
    }
}

// The reference's own parameter-file reader (src/TinyMT/file_reader.cpp:51-70,114-166):
// n calls of tinymt::file_reader::get into out (3n words).  Returns 0, or -1 when the reader
// threw (missing file, format error, comma not found).
int ref_tinymt_file_read(const char* path, int64_t n, uint32_t* out) {
    try {
        tinymt::file_reader fr(path);
        for (int64_t i = 0; i < n; i++) fr.get(&out[3 * i], &out[3 * i + 1], &out[3 * i + 2]);
    } catch (const std::exception&) {
        return -1;
    }
    return 0;
}

}  // extern "C"
