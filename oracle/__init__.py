"""CPU oracle — TEST INFRASTRUCTURE ONLY.

ctypes bindings to ``_build/librt_oracle.so`` (this project's sequential restatement of the
reference chain semantics, rt_oracle.cpp) and, when it was built, ``_ref/libref.so`` (the
reference's own host-side statics compiled from its sources, see Makefile).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg may use this
module, and only as the checker or the CPU baseline; the product never imports it.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_or = None
_ref = None

_I32P = C.POINTER(C.c_int32)
_U32P = C.POINTER(C.c_uint32)
_DP = C.POINTER(C.c_double)
_FP = C.POINTER(C.c_float)
_LP = C.POINTER(C.c_long)


def lib() -> C.CDLL:
    global _or
    if _or is None:
        p = HERE / "_build" / "librt_oracle.so"
        if not p.exists():
            raise RuntimeError(f"{p} missing: run `make -C oracle`")
        L = C.CDLL(str(p))
        L.or_set_threads.argtypes = [C.c_int]
        L.or_set_threads.restype = C.c_int
        L.or_set_threads(default_threads())
        L.or_tinymt_uint32.argtypes = [C.c_uint32] * 4 + [C.c_int, _U32P]
        L.or_tinymt_table_draws.argtypes = [_U32P, C.c_int64, C.c_uint32, C.c_int64, C.c_int, _U32P, _U32P]
        L.or_tinymt_table_draws.restype = C.c_int
        L.or_mt19937_nth.argtypes = [C.c_long, C.c_int]
        L.or_mt19937_nth.restype = C.c_uint64
        L.or_minstd0_nth.argtypes = [C.c_long, C.c_int]
        L.or_minstd0_nth.restype = C.c_uint64
        L.or_minstd0_uniform.argtypes = [C.c_long, C.c_int, C.c_int, C.c_int, _I32P]
        L.or_domino_walk.argtypes = [C.c_int, _DP, _I32P, C.c_int, C.c_long, C.c_long, _U32P, C.c_int64]
        L.or_domino_walk.restype = C.c_int
        L.or_domino_walk_epochs.argtypes = [C.c_int, _DP, _I32P, C.c_int, _LP, _LP, C.c_int, _U32P, C.c_int64]
        L.or_domino_walk_epochs.restype = C.c_int
        L.or_lozenge_walk.argtypes = [C.c_int, C.c_double, _I32P, C.c_int, C.c_long, C.c_long, _U32P, C.c_int64]
        L.or_lozenge_walk.restype = C.c_int
        L.or_lozenge_time.argtypes = [C.c_int, C.c_double, _I32P, C.c_int, C.c_long, C.c_long, _U32P, C.c_int64]
        L.or_lozenge_time.restype = C.c_double
        L.or_recttri_walk.argtypes = [_FP, _I32P, C.c_int, C.c_long, C.c_long, _U32P, C.c_int64]
        L.or_recttri_walk.restype = C.c_int
        L.or_tridimer_walk.argtypes = [_I32P, C.c_int, C.c_long, C.c_long, _U32P, C.c_int64, C.POINTER(C.c_uint8)]
        L.or_tridimer_walk.restype = C.c_int
        L.or_tridimer_flips.argtypes = [_I32P, C.c_int, _I32P, _I32P]
        L.or_domino_cftp.argtypes = [C.c_int, _DP, _I32P, _I32P, C.c_int, C.c_long, C.c_int, C.c_long, C.c_int,
                                     _U32P, C.c_int64, _I32P, _LP, C.POINTER(C.c_int)]
        L.or_domino_cftp.restype = C.c_int
        _or = L
    return _or


def default_threads() -> int:
    """Threads the oracle's loops use unless set_threads says otherwise: the cores this
    process may run on, at most 16."""
    import os
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def set_threads(n: int) -> int:
    """Run the oracle's per-kernel loops on n threads (1 = sequential); returns the previous
    count.  The results do not depend on it."""
    return int(lib().or_set_threads(int(n)))


def ref_available() -> bool:
    return (HERE / "_ref" / "libref.so").exists()


def ref() -> C.CDLL:
    global _ref
    if _ref is None:
        p = HERE / "_ref" / "libref.so"
        if not p.exists():
            raise RuntimeError(f"{p} missing (reference sources were not available at build time)")
        L = C.CDLL(str(p))
        L.ref_call.argtypes = [C.c_char_p, _I32P, C.c_int64, _I32P, C.c_int64, _I32P, C.c_int, _I32P, C.c_int64]
        L.ref_call.restype = C.c_int64
        L.ref_write.argtypes = [C.c_char_p, _I32P, C.c_int64, _I32P, C.c_int64, C.c_char_p]
        L.ref_write.restype = C.c_int
        L.ref_tinymt_host.argtypes = [C.c_uint32] * 7 + [C.c_int, _U32P]
        L.ref_tinymt_file_read.argtypes = [C.c_char_p, C.c_int64, _U32P]
        L.ref_tinymt_file_read.restype = C.c_int
        _ref = L
    return _ref


def ref_tinymt_file_read(path, n: int):
    """The reference's tinymt::file_reader::get called n times: an (n, 3) uint32 array, or None
    when the reader threw."""
    out = np.zeros((n, 3), dtype=np.uint32)
    rc = ref().ref_tinymt_file_read(str(path).encode(), n, out.ctypes.data_as(_U32P))
    return None if rc else out


def _arr(a):
    if a is None:
        return None, 0
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.size


def ref_call(name: str, a=None, b=None, args=(), cap: int | None = None) -> np.ndarray:
    """Call a reference static through ref_driver.cpp's dispatcher."""
    L = ref()
    a, na = _arr(a)
    b, nb = _arr(b)
    ar = np.ascontiguousarray(args, dtype=np.int32) if len(args) else np.zeros(1, np.int32)
    if cap is None:
        cap = max(4 * (na + nb) + 1024, 1 << 20)
    out = np.zeros(cap, dtype=np.int32)
    n = L.ref_call(name.encode(), None if a is None else a.ctypes.data_as(_I32P), na,
                   None if b is None else b.ctypes.data_as(_I32P), nb, ar.ctypes.data_as(_I32P), len(args),
                   out.ctypes.data_as(_I32P), cap)
    if n == -2:
        return ref_call(name, a, b, args, cap * 8)
    if n < 0:
        raise KeyError(name)
    return out[:n].copy()


def ref_write(name: str, a, path: str, b=None) -> None:
    L = ref()
    a, na = _arr(a)
    b, nb = _arr(b)
    rc = L.ref_write(name.encode(), a.ctypes.data_as(_I32P), na, None if b is None else b.ctypes.data_as(_I32P), nb,
                     str(path).encode())
    if rc:
        raise KeyError(name)


def _params(p: np.ndarray):
    p = np.ascontiguousarray(p, dtype=np.uint32).reshape(-1, 3)
    return p, p.ctypes.data_as(_U32P), p.shape[0]


def tinymt_uint32(mat1: int, mat2: int, tmat: int, seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.uint32)
    lib().or_tinymt_uint32(mat1, mat2, tmat, seed, n, out.ctypes.data_as(_U32P))
    return out


def tinymt_table_draws(params: np.ndarray, seed: int, count: int, draws: int):
    p, pp, n = _params(params)
    k = np.zeros((draws, count), dtype=np.uint32)
    st = np.zeros((count, 4), dtype=np.uint32)
    rc = lib().or_tinymt_table_draws(pp, n, seed, count, draws, k.ctypes.data_as(_U32P), st.ctypes.data_as(_U32P))
    assert rc == 0
    return k, st


def domino_walk(t, steps: int, seed: int, params, rule: int = 0, p12=None) -> np.ndarray:
    t = np.ascontiguousarray(t, dtype=np.int32).copy()
    N = int(round(np.sqrt(t.size)))
    p, pp, n = _params(params)
    pr = None if p12 is None else np.ascontiguousarray(p12, dtype=np.float64)
    rc = lib().or_domino_walk(rule, None if pr is None else pr.ctypes.data_as(_DP), t.ctypes.data_as(_I32P), N,
                              steps, seed, pp, n)
    if rc:
        raise ValueError(f"or_domino_walk rc={rc}")
    return t


def domino_walk_epochs(t, steps, seeds, params, rule: int = 1, p12=None) -> np.ndarray:
    t = np.ascontiguousarray(t, dtype=np.int32).copy()
    N = int(round(np.sqrt(t.size)))
    p, pp, n = _params(params)
    st = np.ascontiguousarray(steps, dtype=np.int64)
    sd = np.ascontiguousarray(seeds, dtype=np.int64)
    pr = None if p12 is None else np.ascontiguousarray(p12, dtype=np.float64)
    rc = lib().or_domino_walk_epochs(rule, None if pr is None else pr.ctypes.data_as(_DP), t.ctypes.data_as(_I32P), N,
                                     st.ctypes.data_as(_LP), sd.ctypes.data_as(_LP), len(st), pp, n)
    if rc:
        raise ValueError(f"or_domino_walk_epochs rc={rc}")
    return t


def lozenge_walk(t, steps: int, seed: int, params, rule: int = 0, q: float = 0.99) -> np.ndarray:
    t = np.ascontiguousarray(t, dtype=np.int32).copy()
    N = int(round(np.sqrt(t.size)))
    p, pp, n = _params(params)
    rc = lib().or_lozenge_walk(rule, q, t.ctypes.data_as(_I32P), N, steps, seed, pp, n)
    if rc:
        raise ValueError(f"or_lozenge_walk rc={rc}")
    return t


def lozenge_time(t, steps: int, seed: int, params, rule: int = 0, q: float = 0.99) -> float:
    t = np.ascontiguousarray(t, dtype=np.int32).copy()
    N = int(round(np.sqrt(t.size)))
    p, pp, n = _params(params)
    return float(lib().or_lozenge_time(rule, q, t.ctypes.data_as(_I32P), N, steps, seed, pp, n))


def recttri_walk(t, steps: int, seed: int, params, weights16=None) -> np.ndarray:
    t = np.ascontiguousarray(t, dtype=np.int32).copy()
    N = int(round(np.sqrt(t.size)))
    p, pp, n = _params(params)
    w = None if weights16 is None else np.ascontiguousarray(weights16, dtype=np.float32)
    rc = lib().or_recttri_walk(None if w is None else w.ctypes.data_as(_FP), t.ctypes.data_as(_I32P), N, steps, seed,
                               pp, n)
    if rc:
        raise ValueError(f"or_recttri_walk rc={rc}")
    return t


def tridimer_walk(t, steps: int, seed: int, params, return_moves: bool = False):
    """TriangleDimerTiler::Walk restated; optionally also the per-step move codes
    (type | r1<<2 | r2<<4 | p1<<5 | p2<<7)."""
    t = np.ascontiguousarray(t, dtype=np.int32).copy()
    N = int(round(np.sqrt(t.size)))
    p, pp, n = _params(params)
    mv = np.zeros(max(int(steps), 1), np.uint8)
    rc = lib().or_tridimer_walk(t.ctypes.data_as(_I32P), N, steps, seed, pp, n,
                                mv.ctypes.data_as(C.POINTER(C.c_uint8)))
    if rc:
        raise ValueError(f"or_tridimer_walk rc={rc}")
    return (t, mv[:steps]) if return_moves else t


def tridimer_flips(t):
    """(TilingToTriangleFlips, TilingToButterflyFlips) of a tiling, restated."""
    t = np.ascontiguousarray(t, dtype=np.int32)
    N = int(round(np.sqrt(t.size)))
    tri = np.zeros((N // 2) * N, np.int32)
    bf = np.zeros(N * N, np.int32)
    lib().or_tridimer_flips(t.ctypes.data_as(_I32P), N, tri.ctypes.data_as(_I32P), bf.ctypes.data_as(_I32P))
    return tri, bf


def domino_cftp(tmax, tmin, seed: int, params, rule: int = 1, p12=None, mode: int = 0, first_steps: int = 128,
                max_restarts: int = 64):
    """Coupling from the past restated (AztecDiamondCftpEx.cpp:42-68); returns (sample,
    steps of the final replay, restarts)."""
    tmax = np.ascontiguousarray(tmax, dtype=np.int32)
    tmin = np.ascontiguousarray(tmin, dtype=np.int32)
    N = int(round(np.sqrt(tmax.size)))
    p, pp, n = _params(params)
    pr = None if p12 is None else np.ascontiguousarray(p12, dtype=np.float64)
    out = np.zeros_like(tmax)
    tot = C.c_long()
    rs = C.c_int()
    rc = lib().or_domino_cftp(rule, None if pr is None else pr.ctypes.data_as(_DP), tmax.ctypes.data_as(_I32P),
                              tmin.ctypes.data_as(_I32P), N, seed, mode, first_steps, max_restarts, pp, n,
                              out.ctypes.data_as(_I32P), C.byref(tot), C.byref(rs))
    if rc:
        raise ValueError(f"or_domino_cftp rc={rc}")
    return out, tot.value, rs.value
