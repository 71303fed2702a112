"""ctypes access to the reference-compatible statics of librandomtilings.so
(include/rt_statics.h): domains, maximal/minimal tilings, height functions, SVG writers."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import LIB_DIR

_L = None
_I32P = C.POINTER(C.c_int32)


def lib() -> C.CDLL:
    global _L
    if _L is None:
        p = LIB_DIR / "librandomtilings.so"
        if not p.exists():
            raise RuntimeError(f"{p} not built")
        L = C.CDLL(str(p))
        L.rt_static_call.restype = C.c_int64
        L.rt_static_call.argtypes = [C.c_char_p, _I32P, C.c_int64, _I32P, C.c_int64, _I32P, C.c_int32, _I32P,
                                     C.c_int64]
        L.rt_static_write.restype = C.c_int
        L.rt_static_write.argtypes = [C.c_char_p, _I32P, C.c_int64, _I32P, C.c_int64, C.c_char_p]
        _L = L
    return _L


def _arr(a):
    if a is None:
        return None, 0
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.size


def call(name: str, a=None, b=None, args=(), cap: int | None = None) -> np.ndarray:
    """Call the static `name` ("domino.MaxTiling", ...); returns the result array."""
    a, na = _arr(a)
    b, nb = _arr(b)
    ar = np.ascontiguousarray(args if len(args) else [0], dtype=np.int32)
    if cap is None:
        cap = max(4 * (na + nb) + 4096, 1 << 16)
    while True:
        out = np.zeros(cap, dtype=np.int32)
        n = lib().rt_static_call(name.encode(), None if a is None else a.ctypes.data_as(_I32P), na,
                                 None if b is None else b.ctypes.data_as(_I32P), nb, ar.ctypes.data_as(_I32P),
                                 len(args), out.ctypes.data_as(_I32P), cap)
        if n == -2:
            cap *= 8
            continue
        if n < 0:
            raise KeyError(name)
        return out[:n].copy()


def write(name: str, a, path: str, b=None) -> None:
    a, na = _arr(a)
    b, nb = _arr(b)
    rc = lib().rt_static_write(name.encode(), a.ctypes.data_as(_I32P), na,
                               None if b is None else b.ctypes.data_as(_I32P), nb, str(path).encode())
    if rc:
        raise KeyError(name)
