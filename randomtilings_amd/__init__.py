"""randomtilings_amd — MI355X (gfx950) flip-MCMC for random tilings.

Python face of the C ABI in ``include/rt_hip.h`` (``lib/librt_hip.so``).  The HIP library
is the only compute path: if it is missing, importing the bindings raises, there is no
CPU fallback.  The reference-compatible C++ classes live in ``src/`` and are built into
``lib/librandomtilings.so``.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_DIR = PKG / "lib"

# status codes (rt_hip.h)
RT_OK, RT_ERR_ARG, RT_ERR_PARAMS, RT_ERR_HIP, RT_ERR_IO, RT_ERR_DEVICE, RT_ERR_UNTILEABLE = 0, -1, -2, -3, -4, -5, -6
RT_ERR_NOT_COALESCED = -7
RT_CFTP_REFERENCE, RT_CFTP_DOUBLING = 0, 1
# models
RT_DOMINO, RT_LOZENGE, RT_RECTTRI, RT_TRIDIMER, RT_DOMINO_PAIR = 0, 1, 2, 3, 4
# rules
RT_DOMINO_LAZY, RT_DOMINO_CFTP, RT_DOMINO_WEIGHTED = 0, 1, 2
RT_LOZENGE_UNIFORM, RT_LOZENGE_QREF, RT_LOZENGE_QHEAT = 0, 1, 2
RT_SEED_WALK, RT_SEED_EPOCH = 0, 1
RT_PATH_AUTO, RT_PATH_PERSISTENT, RT_PATH_STEPWISE, RT_PATH_BLOCKED = 0, 1, 2, 3

# Every symbol include/rt_hip.h declares, with (restype, argtypes).
_P = C.c_void_p
_I32P = C.POINTER(C.c_int32)
_U32P = C.POINTER(C.c_uint32)
_I64P = C.POINTER(C.c_int64)
_DP = C.POINTER(C.c_double)
_FP = C.POINTER(C.c_float)
SIGNATURES = {
    "rt_last_error": (C.c_char_p, []),
    "rt_version": (C.c_char_p, []),
    "rt_device_count": (C.c_int, []),
    "rt_device_name": (C.c_int, [C.c_int, C.c_char_p, C.c_int32]),
    "rt_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "rt_destroy": (C.c_int, [_P]),
    "rt_load_params": (C.c_int, [_P, _U32P, C.c_int64]),
    "rt_load_params_file": (C.c_int, [_P, C.c_char_p, C.c_int64]),
    "rt_parse_params_file": (C.c_int, [C.c_char_p, C.c_int64, _U32P]),
    "rt_params_count": (C.c_int64, [_P]),
    "rt_ctx_set_path": (C.c_int, [_P, C.c_int, C.c_int32]),
    "rt_ctx_mark": (C.c_int, [_P, C.c_int32]),
    "rt_ctx_elapsed_ms": (C.c_int, [_P, C.c_int32, C.c_int32, _DP]),
    "rt_ctx_sync": (C.c_int, [_P]),
    "rt_rule_thresholds": (C.c_int32, [C.c_int, C.c_int, _DP, C.c_int32, _U32P, C.c_int32]),
    "rt_domino_walk": (C.c_int, [_P, C.c_int, _DP, _I32P, C.c_int32, C.c_int64, C.c_int64]),
    "rt_domino_walk_epochs": (C.c_int, [_P, C.c_int, _DP, _I32P, C.c_int32, _I64P, _I64P, C.c_int32]),
    "rt_lozenge_walk": (C.c_int, [_P, C.c_int, C.c_double, _I32P, C.c_int32, C.c_int64, C.c_int64]),
    "rt_lozenge_walk_noalloc": (C.c_int, [_P, C.c_int, C.c_double, _I32P, C.c_int32, C.c_int64, C.c_int64]),
    "rt_recttri_walk": (C.c_int, [_P, _FP, _I32P, C.c_int32, C.c_int64, C.c_int64]),
    "rt_tridimer_walk": (C.c_int, [_P, _I32P, C.c_int32, C.c_int64, C.c_int64]),
    "rt_extremal_tiling": (C.c_int, [_P, C.c_int, _I32P, C.c_int32, C.c_int, _I32P, _I32P]),
    "rt_height_closure": (C.c_int, [_P, C.c_int, _I32P, C.c_int32, C.c_int, _I32P, C.POINTER(C.c_uint8), _I32P]),
    "rt_domino_cftp": (C.c_int, [_P, C.c_int, _DP, _I32P, _I32P, C.c_int32, C.c_int64, C.c_int, C.c_int64, C.c_int32,
                                 _I32P, _I64P, _I32P]),
    "rt_domino_cftp_ex": (C.c_int, [_P, C.c_int, _DP, _I32P, _I32P, C.c_int32, C.c_int64, C.c_int, C.c_int64,
                                    C.c_int32, _I32P, _I64P, _I32P, C.c_void_p, C.c_void_p]),
    "rt_chain_create": (C.c_int, [_P, C.c_int, C.c_int, _DP, C.c_int32, _I32P, C.c_int32, C.POINTER(_P)]),
    "rt_chain_destroy": (C.c_int, [_P]),
    "rt_chain_set_path": (C.c_int, [_P, C.c_int, C.c_int32]),
    "rt_chain_seed": (C.c_int, [_P, C.c_int, C.c_int64]),
    "rt_chain_autotune": (C.c_int, [_P, C.c_int64, _I32P]),
    "rt_chain_run": (C.c_int, [_P, C.c_int64]),
    "rt_chain_sync": (C.c_int, [_P]),
    "rt_chain_elapsed_ms": (C.c_int, [_P, _DP]),
    "rt_chain_read": (C.c_int, [_P, _I32P]),
    "rt_chain_attempts_per_step": (C.c_int64, [_P]),
    "rt_chain_state_bytes": (C.c_int64, [_P]),
    "rt_tile_config_count": (C.c_int32, [C.c_int]),
    "rt_persist_capacity": (C.c_int32, [C.c_int, C.c_int32, C.c_int, C.c_int]),
    "rt_tile_config": (C.c_int, [C.c_int, C.c_int32, _I32P, _I32P, _I32P, _I32P]),
    "rt_tile_config_packed": (C.c_int32, [C.c_int, C.c_int32]),
    "rt_chain_last_launches": (C.c_int64, [_P]),
    "rt_chain_last_config": (C.c_int32, [_P]),
    "rt_chain_coop_refusals": (C.c_int, [_P, _I64P, _I32P]),
    "rt_chain_set_skip": (C.c_int, [_P, C.c_int]),
    "rt_chain_last_grid": (C.c_int, [_P, _I32P, _I32P]),
    "rt_lozenge_skip_map": (C.c_int, [_I32P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_uint8), _I32P,
                                      _I32P]),
    "rt_chain_tile_configs": (C.c_int32, [_P]),
    "rt_chain_tile_config_info": (C.c_int, [_P, C.c_int32, _I32P, _I32P, _I32P, _I32P]),
    "rt_unet_create": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int64, C.POINTER(_P)]),
    "rt_unet_destroy": (C.c_int, [_P]),
    "rt_unet_param_count": (C.c_int64, [_P]),
    "rt_unet_get_params": (C.c_int, [_P, _FP]),
    "rt_unet_set_params": (C.c_int, [_P, _FP]),
    "rt_unet_forward": (C.c_int, [_P, _FP, C.c_int, C.c_int, C.c_int, _FP]),
    "rt_unet_loss_grad": (C.c_int, [_P, _FP, _FP, _FP, C.c_int, C.c_int, C.c_int, _FP, _FP]),
    "rt_unet_train_step": (C.c_int, [_P, _FP, _FP, _FP, C.c_int, C.c_int, C.c_int, C.c_float, _FP]),
    "rt_nn_gemm": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, _FP, _FP, _FP]),
    "rt_nn_gemm_ex": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, _FP, _FP, _FP]),
    "rt_nn_gemm_time": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "rt_observer_create": (C.c_int, [_P, _I32P, C.c_int32, _I32P, C.POINTER(C.c_uint8), _I32P, C.POINTER(_P)]),
    "rt_observer_destroy": (C.c_int, [_P]),
    "rt_observe": (C.c_int, [_P, C.POINTER(_P), C.c_int32, _I64P, _I32P]),
    "rt_batch_create": (C.c_int, [_P, C.POINTER(_P)]),
    "rt_batch_destroy": (C.c_int, [_P]),
    "rt_batch_add": (C.c_int, [_P, _P]),
    "rt_batch_run": (C.c_int, [_P, C.c_int64, C.c_int]),
    "rt_batch_sync": (C.c_int, [_P]),
    "rt_batch_autotune": (C.c_int, [_P, C.c_int64, _I32P]),
    "rt_batch_set_config": (C.c_int, [_P, C.c_int, C.c_int32]),
    "rt_batch_elapsed_ms": (C.c_int, [_P, _DP]),
    "rt_batch_last_launches": (C.c_int64, [_P]),
    "rt_batch_last_graphs": (C.c_int32, [_P]),
}

_lib = None


# rt_cftp_progress(user, restart, replay_steps, coalesced)
CFTP_PROGRESS = C.CFUNCTYPE(None, C.c_void_p, C.c_int32, C.c_int64, C.c_int32)


class RTError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def lib_path() -> Path:
    return LIB_DIR / "librt_hip.so"


def load() -> C.CDLL:
    """Load librt_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    p = lib_path()
    if not p.exists():
        raise RuntimeError(f"{p} not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc: int) -> None:
    if rc != RT_OK:
        raise RTError(rc, load().rt_last_error().decode())


def _i32(a: np.ndarray):
    return a.ctypes.data_as(_I32P)


def lozenge_skip_map(tiling: np.ndarray, tr: int, tc: int, h: int):
    """Host-only (no device): which tiles of tr x tc positions (rows x class columns, halo h)
    the packed lozenge sweep runs for this N x N tiling (rt_lozenge_skip_map).  Returns
    (map of shape (ceil(N/tr), ceil(N/3/tc)), 1 = runs; whether the tiling is a fixed point of
    the update)."""
    t = np.ascontiguousarray(tiling, dtype=np.int32).ravel()
    N = int(round(np.sqrt(t.size)))
    W = N // 3
    m = np.zeros(((N + tr - 1) // tr) * ((W + tc - 1) // tc), dtype=np.uint8)
    run, ok = C.c_int32(), C.c_int32()
    _check(load().rt_lozenge_skip_map(_i32(t), N, tr, tc, h, m.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(run),
                                      C.byref(ok)))
    return m.reshape((N + tr - 1) // tr, (W + tc - 1) // tc), bool(ok.value)


def parse_params_file(path, n: int) -> np.ndarray:
    """Host only: the first n (mat1, mat2, tmat) entries of a tinymtdc-format file, as an
    (n, 3) uint32 array (rt_parse_params_file; raises RTError RT_ERR_IO on a format error)."""
    out = np.zeros((n, 3), dtype=np.uint32)
    _check(load().rt_parse_params_file(str(path).encode(), n, out.ctypes.data_as(_U32P)))
    return out


def rule_thresholds(model: int, rule: int, rparams=None) -> np.ndarray:
    lib = load()
    rp = None if rparams is None else np.ascontiguousarray(rparams, dtype=np.float64)
    # the first call reports the size (FlipRule and RectTriRule words), the second fills it
    n = lib.rt_rule_thresholds(model, rule, None if rp is None else rp.ctypes.data_as(_DP),
                               0 if rp is None else len(rp), None, 0)
    if n < 0:
        _check(n)
    out = np.zeros(n, dtype=np.uint32)
    n = lib.rt_rule_thresholds(model, rule, None if rp is None else rp.ctypes.data_as(_DP),
                               0 if rp is None else len(rp), out.ctypes.data_as(_U32P), n)
    if n < 0:
        _check(n)
    return out


def tile_config_count(model: int) -> int:
    return int(load().rt_tile_config_count(model))


def persist_capacity(model: int, idx: int, device: int = 0, uniform: bool = True) -> int:
    """Workgroups of tile configuration idx the device holds at once (0: never persistent)."""
    return int(load().rt_persist_capacity(model, idx, device, int(uniform)))


def tile_config_packed(model: int, idx: int) -> bool:
    """Whether configuration idx runs the packed (four positions per word) sweep."""
    return bool(load().rt_tile_config_packed(model, idx))


def tile_config(model: int, idx: int):
    v = [C.c_int32() for _ in range(4)]
    _check(load().rt_tile_config(model, idx, *[C.byref(x) for x in v]))
    return tuple(x.value for x in v)


class Context:
    """One HIP device + stream + TinyMT parameter table (the reference's XTiler state)."""

    def __init__(self, device: int = 0):
        self._lib = load()
        h = _P()
        _check(self._lib.rt_create(device, C.byref(h)))
        self.h = h

    def close(self) -> None:
        if self.h:
            self._lib.rt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def load_params(self, mat3: np.ndarray) -> None:
        m = np.ascontiguousarray(mat3, dtype=np.uint32).reshape(-1, 3)
        _check(self._lib.rt_load_params(self.h, m.ctypes.data_as(_U32P), m.shape[0]))

    def load_params_file(self, path: str, n: int) -> None:
        _check(self._lib.rt_load_params_file(self.h, str(path).encode(), n))

    def params_count(self) -> int:
        return int(self._lib.rt_params_count(self.h))

    def set_path(self, path: int, tile_config: int = -1) -> None:
        _check(self._lib.rt_ctx_set_path(self.h, path, tile_config))

    def mark(self, slot: int) -> None:
        _check(self._lib.rt_ctx_mark(self.h, slot))

    def elapsed_ms(self, a: int, b: int) -> float:
        v = C.c_double()
        _check(self._lib.rt_ctx_elapsed_ms(self.h, a, b, C.byref(v)))
        return v.value

    def sync(self) -> None:
        _check(self._lib.rt_ctx_sync(self.h))

    # -- reference Walk entry points ------------------------------------------------
    def domino_walk(self, t: np.ndarray, steps: int, seed: int, rule: int = RT_DOMINO_LAZY, p12=None) -> np.ndarray:
        t = np.ascontiguousarray(t, dtype=np.int32).copy()
        N = int(round(np.sqrt(t.size)))
        rp = None if p12 is None else np.ascontiguousarray(p12, dtype=np.float64)
        _check(self._lib.rt_domino_walk(self.h, rule, None if rp is None else rp.ctypes.data_as(_DP), _i32(t), N,
                                        steps, seed))
        return t

    def domino_walk_epochs(self, t: np.ndarray, steps, seeds, rule: int = RT_DOMINO_CFTP, p12=None) -> np.ndarray:
        t = np.ascontiguousarray(t, dtype=np.int32).copy()
        N = int(round(np.sqrt(t.size)))
        st = np.ascontiguousarray(steps, dtype=np.int64)
        sd = np.ascontiguousarray(seeds, dtype=np.int64)
        rp = None if p12 is None else np.ascontiguousarray(p12, dtype=np.float64)
        _check(self._lib.rt_domino_walk_epochs(self.h, rule, None if rp is None else rp.ctypes.data_as(_DP), _i32(t),
                                               N, st.ctypes.data_as(_I64P), sd.ctypes.data_as(_I64P), len(st)))
        return t

    def lozenge_walk(self, t: np.ndarray, steps: int, seed: int, rule: int = RT_LOZENGE_UNIFORM,
                     q: float = 0.99) -> np.ndarray:
        t = np.ascontiguousarray(t, dtype=np.int32).copy()
        N = int(round(np.sqrt(t.size)))
        _check(self._lib.rt_lozenge_walk(self.h, rule, q, _i32(t), N, steps, seed))
        return t

    def lozenge_walk_noalloc(self, t: np.ndarray, steps: int, seed: int, rule: int = RT_LOZENGE_UNIFORM,
                             q: float = 0.99) -> np.ndarray:
        """lozenge_walk through rt_lozenge_walk_noalloc (no device allocation; N <= 2271)."""
        t = np.ascontiguousarray(t, dtype=np.int32).copy()
        N = int(round(np.sqrt(t.size)))
        _check(self._lib.rt_lozenge_walk_noalloc(self.h, rule, q, _i32(t), N, steps, seed))
        return t

    def recttri_walk(self, t: np.ndarray, steps: int, seed: int, weights16=None) -> np.ndarray:
        t = np.ascontiguousarray(t, dtype=np.int32).copy()
        N = int(round(np.sqrt(t.size)))
        w = None if weights16 is None else np.ascontiguousarray(weights16, dtype=np.float32)
        _check(self._lib.rt_recttri_walk(self.h, None if w is None else w.ctypes.data_as(_FP), _i32(t), N, steps,
                                         seed))
        return t

    def tridimer_walk(self, t: np.ndarray, steps: int, seed: int) -> np.ndarray:
        t = np.ascontiguousarray(t, dtype=np.int32).copy()
        N = int(round(np.sqrt(t.size)))
        _check(self._lib.rt_tridimer_walk(self.h, _i32(t), N, steps, seed))
        return t

    def domino_cftp(self, tmax: np.ndarray, tmin: np.ndarray, seed: int, rule: int = RT_DOMINO_CFTP, p12=None,
                    mode: int = RT_CFTP_REFERENCE, first_steps: int = 128, max_restarts: int = 64, progress=None):
        """Coupling from the past (rt_domino_cftp): (sample, steps of the final replay, restarts).
        progress(restart, steps, coalesced), if given, is called after every replay (coalesced
        True or False) and after every epoch but the last of a replay (coalesced None, steps so
        far)."""
        a = np.ascontiguousarray(tmax, dtype=np.int32)
        b = np.ascontiguousarray(tmin, dtype=np.int32)
        N = int(round(np.sqrt(a.size)))
        rp = None if p12 is None else np.ascontiguousarray(p12, dtype=np.float64)
        out = np.zeros_like(a)
        tot = C.c_int64()
        rs = C.c_int32()
        cb = None if progress is None else CFTP_PROGRESS(lambda u, r, n, c: progress(r, n, None if c < 0 else bool(c)))
        _check(self._lib.rt_domino_cftp_ex(self.h, rule, None if rp is None else rp.ctypes.data_as(_DP), _i32(a),
                                           _i32(b), N, seed, mode, first_steps, max_restarts, _i32(out), C.byref(tot),
                                           C.byref(rs), C.cast(cb, C.c_void_p) if cb is not None else None, None))
        return out, tot.value, rs.value

    def extremal_tiling(self, model: int, domain: np.ndarray, maximal: bool = True):
        """(tiling, heights) of the maximal / minimal tiling, built on the device."""
        d = np.ascontiguousarray(domain, dtype=np.int32)
        M = int(round(np.sqrt(d.size / (2 if model == RT_LOZENGE else 1))))
        t = np.zeros((M + 1) ** 2, np.int32)
        h = np.zeros((M + 1) ** 2, np.int32)
        _check(self._lib.rt_extremal_tiling(self.h, model, _i32(d), M, 1 if maximal else -1, _i32(t), _i32(h)))
        return t, h

    def height_closure(self, model: int, domain: np.ndarray, heights: np.ndarray, direction: int = 1, fixed=None):
        """(closed heights, relaxation launches) of rt_height_closure."""
        d = np.ascontiguousarray(domain, dtype=np.int32)
        M = int(round(np.sqrt(d.size / (2 if model == RT_LOZENGE else 1))))
        h = np.ascontiguousarray(heights, dtype=np.int32).copy()
        fx = None if fixed is None else np.ascontiguousarray(fixed, dtype=np.uint8)
        nl = C.c_int32()
        _check(self._lib.rt_height_closure(self.h, model, _i32(d), M, direction, _i32(h),
                                           None if fx is None else fx.ctypes.data_as(C.POINTER(C.c_uint8)),
                                           C.byref(nl)))
        return h, nl.value

    def chain(self, model: int, rule: int, tiling: np.ndarray, rparams=None) -> "Chain":
        return Chain(self, model, rule, tiling, rparams)


class Chain:
    """A device-resident chain (rt_chain_*)."""

    def __init__(self, ctx: Context, model: int, rule: int, tiling: np.ndarray, rparams=None):
        self.ctx = ctx
        self._lib = ctx._lib
        t = np.ascontiguousarray(tiling, dtype=np.int32)
        self.N = int(round(np.sqrt(t.size)))
        rp = None if rparams is None else np.ascontiguousarray(rparams, dtype=np.float64)
        h = _P()
        _check(self._lib.rt_chain_create(ctx.h, model, rule, None if rp is None else rp.ctypes.data_as(_DP),
                                         0 if rp is None else len(rp), _i32(t), self.N, C.byref(h)))
        self.h = h

    def close(self) -> None:
        if self.h:
            self._lib.rt_chain_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_path(self, path: int, tile_config: int = -1) -> None:
        _check(self._lib.rt_chain_set_path(self.h, path, tile_config))

    def autotune(self, probe_steps: int = 2048) -> int:
        best = C.c_int32(-1)
        _check(self._lib.rt_chain_autotune(self.h, probe_steps, C.byref(best)))
        return best.value

    def seed(self, seed: int, mode: int = RT_SEED_WALK) -> None:
        _check(self._lib.rt_chain_seed(self.h, mode, seed))

    def run(self, steps: int) -> None:
        _check(self._lib.rt_chain_run(self.h, steps))

    def sync(self) -> None:
        _check(self._lib.rt_chain_sync(self.h))

    def elapsed_ms(self) -> float:
        v = C.c_double()
        _check(self._lib.rt_chain_elapsed_ms(self.h, C.byref(v)))
        return v.value

    def read(self) -> np.ndarray:
        t = np.zeros(self.N * self.N, dtype=np.int32)
        _check(self._lib.rt_chain_read(self.h, _i32(t)))
        return t

    @property
    def attempts_per_step(self) -> int:
        return int(self._lib.rt_chain_attempts_per_step(self.h))

    @property
    def state_bytes(self) -> int:
        return int(self._lib.rt_chain_state_bytes(self.h))

    @property
    def last_launches(self) -> int:
        return int(self._lib.rt_chain_last_launches(self.h))

    @property
    def last_config(self) -> int:
        return int(self._lib.rt_chain_last_config(self.h))

    @property
    def coop_refusals(self):
        """(resident launches the runtime refused, hipError_t of the last refusal): such runs
        continued on the blocked path."""
        n, e = C.c_int64(), C.c_int32()
        _check(self._lib.rt_chain_coop_refusals(self.h, C.byref(n), C.byref(e)))
        return n.value, e.value

    def set_skip(self, enable: bool) -> None:
        """Skip (default) or run the tiles whose cells never change (lozenge, packed sweep)."""
        _check(self._lib.rt_chain_set_skip(self.h, int(bool(enable))))

    @property
    def last_grid(self):
        """(tiles of the most recent persistent run's grid, tiles that ran)."""
        t, r = C.c_int32(), C.c_int32()
        _check(self._lib.rt_chain_last_grid(self.h, C.byref(t), C.byref(r)))
        return t.value, r.value

    def tile_configs(self):
        out = []
        for i in range(self._lib.rt_chain_tile_configs(self.h)):
            v = [C.c_int32() for _ in range(4)]
            _check(self._lib.rt_chain_tile_config_info(self.h, i, *[C.byref(x) for x in v]))
            out.append(tuple(x.value for x in v))
        return out


class Observer:
    """Height-function statistics of domino chains on one domain, computed on the device
    (rt_observer_*, rt_observe)."""

    def __init__(self, ctx: Context, domain: np.ndarray, hmin: np.ndarray, mask=None, ref256=None):
        self.ctx = ctx
        self._lib = ctx._lib
        d = np.ascontiguousarray(domain, dtype=np.int32)
        M = int(round(np.sqrt(d.size)))
        self.N = M + 1
        hm = np.ascontiguousarray(hmin, dtype=np.int32)
        mk = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
        rf = None if ref256 is None else np.ascontiguousarray(ref256, dtype=np.int32)
        h = _P()
        _check(self._lib.rt_observer_create(ctx.h, _i32(d), M, _i32(hm),
                                            None if mk is None else mk.ctypes.data_as(C.POINTER(C.c_uint8)),
                                            None if rf is None else _i32(rf), C.byref(h)))
        self.h = h

    def observe(self, chains, heights: bool = False):
        """(n x 4 int64 array of count, sum h, sum h^2, sum |256 h - ref256| per chain,
        n x N^2 heights or None)."""
        n = len(chains)
        arr = (_P * n)(*[ch.h for ch in chains])
        out = np.zeros((n, 4), np.int64)
        hs = np.zeros((n, self.N * self.N), np.int32) if heights else None
        _check(self._lib.rt_observe(self.h, arr, n, out.ctypes.data_as(_I64P), None if hs is None else _i32(hs)))
        return out, hs

    def close(self) -> None:
        if self.h:
            self._lib.rt_observer_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many chains of one context run together (rt_batch_*)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self._lib = ctx._lib
        h = _P()
        _check(self._lib.rt_batch_create(ctx.h, C.byref(h)))
        self.h = h
        self.chains = []

    def add(self, ch: Chain) -> None:
        _check(self._lib.rt_batch_add(self.h, ch.h))
        self.chains.append(ch)

    def run(self, steps: int, graph: bool = True) -> None:
        _check(self._lib.rt_batch_run(self.h, steps, 1 if graph else 0))

    def sync(self) -> None:
        _check(self._lib.rt_batch_sync(self.h))

    def autotune(self, probe_steps: int = 64):
        """Time every applicable tile configuration per model on the batch's chains (put back
        as they were) and keep the fastest: {model: configuration} for the models present."""
        out = (C.c_int32 * 4)()
        _check(self._lib.rt_batch_autotune(self.h, probe_steps, out))
        return {m: out[m] for m in range(4) if out[m] >= 0}

    def set_config(self, model: int, idx: int) -> None:
        """Tile configuration of `model` for later runs (-1: the automatic choice), e.g. one an
        earlier autotune() returned for the same chains."""
        _check(self._lib.rt_batch_set_config(self.h, model, idx))

    def elapsed_ms(self) -> float:
        v = C.c_double()
        _check(self._lib.rt_batch_elapsed_ms(self.h, C.byref(v)))
        return v.value

    @property
    def last_launches(self) -> int:
        return int(self._lib.rt_batch_last_launches(self.h))

    @property
    def last_graphs(self) -> int:
        return int(self._lib.rt_batch_last_graphs(self.h))

    def close(self) -> None:
        if self.h:
            self._lib.rt_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def nn_gemm(ctx: Context, A: np.ndarray, B: np.ndarray, layout: int = 0) -> np.ndarray:
    """A @ B through the bf16 matrix-core kernel (inputs rounded to bf16, fp32 result).  With
    layout 1, A is passed as its transpose ([K x M]); with layout 2, B as its transpose."""
    A, B = _f32(A), _f32(B)
    M, K = A.shape if layout != 1 else A.shape[::-1]
    K2, N = B.shape if layout != 2 else B.shape[::-1]
    assert K == K2
    Cm = np.zeros((M, N), np.float32)
    _check(ctx._lib.rt_nn_gemm_ex(ctx.h, M, N, K, layout, A.ctypes.data_as(_FP), B.ctypes.data_as(_FP),
                                  Cm.ctypes.data_as(_FP)))
    return Cm


def nn_gemm_time(ctx: Context, M: int, N: int, K: int, layout: int = 0, reps: int = 10) -> float:
    """Milliseconds per device-resident bf16 product (rt_nn_gemm_time)."""
    ms = C.c_double()
    _check(ctx._lib.rt_nn_gemm_time(ctx.h, M, N, K, layout, reps, C.byref(ms)))
    return ms.value


class UNet:
    """The warm-start U-Net (rt_unet_*): NHWC float32 images, one output channel."""

    def __init__(self, ctx: Context, levels: int = 4, base: int = 32, in_ch: int = 4, seed: int = 0):
        self.ctx = ctx
        self._lib = ctx._lib
        self.levels, self.base, self.in_ch = levels, base, in_ch
        h = _P()
        _check(self._lib.rt_unet_create(ctx.h, levels, base, in_ch, seed, C.byref(h)))
        self.h = h

    @property
    def param_count(self) -> int:
        return int(self._lib.rt_unet_param_count(self.h))

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.param_count, np.float32)
        _check(self._lib.rt_unet_get_params(self.h, p.ctypes.data_as(_FP)))
        return p

    def set_params(self, p) -> None:
        p = _f32(p)
        _check(self._lib.rt_unet_set_params(self.h, p.ctypes.data_as(_FP)))

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = _f32(x)
        B, H, W, _ = x.shape
        y = np.zeros((B, H, W), np.float32)
        _check(self._lib.rt_unet_forward(self.h, x.ctypes.data_as(_FP), B, H, W, y.ctypes.data_as(_FP)))
        return y

    def loss_grad(self, x, target, mask, grad: bool = True):
        x, t, m = _f32(x), _f32(target), _f32(mask)
        B, H, W, _ = x.shape
        loss = C.c_float()
        g = np.zeros(self.param_count, np.float32) if grad else None
        _check(self._lib.rt_unet_loss_grad(self.h, x.ctypes.data_as(_FP), t.ctypes.data_as(_FP), m.ctypes.data_as(_FP),
                                           B, H, W, C.byref(loss), None if g is None else g.ctypes.data_as(_FP)))
        return loss.value, g

    def train_step(self, x, target, mask, lr: float = 1e-3) -> float:
        x, t, m = _f32(x), _f32(target), _f32(mask)
        B, H, W, _ = x.shape
        loss = C.c_float()
        _check(self._lib.rt_unet_train_step(self.h, x.ctypes.data_as(_FP), t.ctypes.data_as(_FP),
                                            m.ctypes.data_as(_FP), B, H, W, lr, C.byref(loss)))
        return loss.value

    def close(self) -> None:
        if self.h:
            self._lib.rt_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
