/*
 * rt_statics.h — C entry points of librandomtilings.so: the reference-compatible domain,
 * tiling, height-function and drawing utilities of the per-model C++ classes (DominoTiler,
 * LozengeTiler, RectTriangleTiler, TriangleDimerTiler statics), for FFI callers that cannot
 * use the C++ API.  Dispatch is by "model.Function" name, e.g. "domino.MaxTiling"
 * (DominoTiler::MaxTiling, DominoTiler.cpp:214-307 of the reference).
 */
#ifndef RT_STATICS_H_
#define RT_STATICS_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Call the static `name` with up to two int arrays (a, b) and int arguments; the result
 * array is written to out.  Returns its length, -1 for an unknown name, -2 when cap is too
 * small (nothing written). */
int64_t rt_static_call(const char* name, const int32_t* a, int64_t na, const int32_t* b, int64_t nb,
                       const int32_t* args, int32_t nargs, int32_t* out, int64_t cap);

/* Run the drawing/text writer `name` (e.g. "domino.TilingToSVG", "common.SaveMatrix") on the
 * arrays, writing `path`.  Returns 0, or -1 for an unknown name. */
int rt_static_write(const char* name, const int32_t* a, int64_t na, const int32_t* b, int64_t nb, const char* path);

#ifdef __cplusplus
}
#endif

#endif /* RT_STATICS_H_ */
