"""Legality checks and observables on the reference's tiling encodings (test helpers).

Domino (DominoTiler.cpp:204-210): vertex (i,j) carries bit 1/2/4/8 when a domino crosses
its edge to (i-1,j)/(i+1,j)/(i,j-1)/(i,j+1).  Face (fi,fj) of the M x M domain has corner
vertices (fi,fj), (fi,fj+1), (fi+1,fj), (fi+1,fj+1) (DominoTiler.cpp:479-499).

Lozenge (LozengeTiler.cpp:12-16): bit 1/2/4/8/16/32 marks the edge to
(i-1,j)/(i-1,j+1)/(i,j-1)/(i,j+1)/(i+1,j-1)/(i+1,j).  Triangle face (i,j) of the M x 2M
domain touches vertices (i,j/2),(i+1,j/2),(i,j/2+1) for even j and
(i+1,j/2+1),(i+1,j/2),(i,j/2+1) for odd j (LozengeTiler.cpp:475-478).
"""
from __future__ import annotations

import numpy as np


def _sq(a):
    a = np.asarray(a)
    if a.ndim == 2:
        return a
    n = int(round(np.sqrt(a.size)))
    return a.reshape(n, n)


def domino_consistent(t) -> bool:
    t = _sq(t)
    ok = np.array_equal((t[:-1, :] >> 1) & 1, t[1:, :] & 1)  # down bit of (i,j) == up bit of (i+1,j)
    ok &= np.array_equal((t[:, :-1] >> 3) & 1, (t[:, 1:] >> 2) & 1)  # right == left of (i,j+1)
    return bool(ok)


def domino_face_cover(t) -> np.ndarray:
    """Number of crossed boundary edges of every face (M x M, M = N-1)."""
    t = _sq(t)
    top = (t[:-1, :-1] >> 3) & 1
    bottom = (t[1:, :-1] >> 3) & 1
    left = (t[:-1, :-1] >> 1) & 1
    right = (t[:-1, 1:] >> 1) & 1
    return top + bottom + left + right


def domino_legal(t, d) -> bool:
    t = _sq(t)
    d = _sq(d)
    if not domino_consistent(t):
        return False
    cov = domino_face_cover(t)
    return bool(np.array_equal(cov, (d != 0).astype(cov.dtype)))


def domino_key(t) -> bytes:
    return np.ascontiguousarray(t, dtype=np.int32).tobytes()


def lozenge_consistent(t) -> bool:
    t = _sq(t)
    ok = np.array_equal((t[:, :-1] >> 3) & 1, (t[:, 1:] >> 2) & 1)  # 8 <-> 4
    ok &= np.array_equal((t[:-1, :] >> 5) & 1, t[1:, :] & 1)  # 32 <-> 1
    ok &= np.array_equal((t[:-1, 1:] >> 4) & 1, (t[1:, :-1] >> 1) & 1)  # 16 of (i,j) <-> 2 of (i+1,j-1)
    return bool(ok)


_LOZ_BIT = {(-1, 0): 0, (-1, 1): 1, (0, -1): 2, (0, 1): 3, (1, -1): 4, (1, 0): 5}


def lozenge_face_cover(t) -> np.ndarray:
    """Crossed edges per triangle face (M x 2M, M = N-1)."""
    t = _sq(t)
    N = t.shape[0]
    M = N - 1
    cov = np.zeros((M, 2 * M), dtype=np.int32)

    def crossed(a, b):
        (i0, j0), (i1, j1) = a, b
        bit = _LOZ_BIT[(i1 - i0, j1 - j0)]
        return (t[i0, j0] >> bit) & 1

    for i in range(M):
        for j in range(2 * M):
            if j % 2 == 0:
                v = [(i, j // 2), (i + 1, j // 2), (i, j // 2 + 1)]
            else:
                v = [(i + 1, j // 2 + 1), (i + 1, j // 2), (i, j // 2 + 1)]
            if any(x >= N or y >= N for x, y in v):
                continue
            cov[i, j] = crossed(v[0], v[1]) + crossed(v[1], v[2]) + crossed(v[0], v[2])
    return cov


def lozenge_legal(t, d=None) -> bool:
    t = _sq(t)
    if not lozenge_consistent(t):
        return False
    cov = lozenge_face_cover(t)
    if d is None:
        return bool(np.all(cov <= 1))
    d = np.asarray(d).reshape(cov.shape)
    return bool(np.array_equal(cov, (d != 0).astype(cov.dtype)))


def lozenge_cubes(t) -> int:
    """Count of vertices in the 'cube' state 38 (qlozengekernel.cl:25)."""
    return int(np.count_nonzero(_sq(t) == 38))
