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
    """Crossed edges per triangle face (M x 2M, M = N-1), whole arrays at a time: face (i, 2k)
    has corners (i, k), (i+1, k), (i, k+1), face (i, 2k+1) corners (i+1, k+1), (i+1, k),
    (i, k+1); an edge is crossed when the bit of its direction is set at its first corner
    (lozenge_face_cover_loop, the face-by-face form, is the reference for it)."""
    t = _sq(t).astype(np.int64)
    N = t.shape[0]
    M = N - 1
    cov = np.zeros((M, 2 * M), dtype=np.int32)
    a, b, c = t[:M, :M], t[1:, :M], t[1:, 1:]
    cov[:, 0::2] = ((a >> 5) & 1) + ((b >> 1) & 1) + ((a >> 3) & 1)  # (1,0), (-1,1), (0,1)
    cov[:, 1::2] = ((c >> 2) & 1) + ((b >> 1) & 1) + (c & 1)          # (0,-1), (-1,1), (-1,0)
    return cov


def lozenge_face_cover_loop(t) -> np.ndarray:
    """lozenge_face_cover one face at a time."""
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


# ---- Triangle dimers ---------------------------------------------------------------------
# TriangleDimerTiler.cpp:488-540 (IceCreamCone) encodes the dimer indicators h (horizontal),
# l (up-left) and r (up-right) of the triangular lattice into lozenge-move states:
#   t[2a][2b+1]   = r[a-1][b] + 2 l[a-1][b+1] + 4 r[a][b] + 8 l[a][b]
#   t[2a+1][2b]   = h[a][b] + 2 r[a][b] + 4 h[a+1][b-1] + 8 r[a][b-1]
#   t[2a+1][2b+1] = h[a][b] + 2 l[a][b+1] + 4 h[a+1][b] + 8 l[a][b]
# so every lattice edge appears in four states.  Vertex (a,b) meets h[a][b], h[a][b-1],
# l[a][b], l[a-1][b], r[a-1][b] and r[a][b-1] (the six edges TilingToVertices,
# TriangleDimerTiler.cpp:711-735, inspects).

def tridimer_edges(t):
    """Decode (h, l, r) from a tiling and check that all four copies of every edge agree,
    whole arrays at a time (tridimer_edges_loop, the state-by-state form, is the reference for
    it).  Returns (h, l, r, consistent)."""
    t = _sq(t).astype(np.int64)
    N = t.shape[0]
    E = N // 2
    ok = not t[0::2, 0::2].any()
    layers = {"h": [], "l": [], "r": []}

    def put(name, x, y, v):
        nonlocal ok
        inside = (x >= 0) & (x <= E) & (y >= 0) & (y <= E)
        ok &= not v[~inside].any()  # an edge outside the lattice must be absent
        lay = np.full((E + 1, E + 1), -1, np.int64)
        lay[x[inside], y[inside]] = v[inside]
        layers[name].append(lay)

    for (i0, j0), cells in (
            ((0, 1), [("r", -1, 0), ("l", -1, 1), ("r", 0, 0), ("l", 0, 0)]),
            ((1, 0), [("h", 0, 0), ("r", 0, 0), ("h", 1, -1), ("r", 0, -1)]),
            ((1, 1), [("h", 0, 0), ("l", 0, 1), ("h", 1, 0), ("l", 0, 0)])):
        s = t[i0::2, j0::2]
        a, b = np.meshgrid(np.arange(s.shape[0]), np.arange(s.shape[1]), indexing="ij")
        for k, (name, da, db) in enumerate(cells):
            put(name, (a + da).ravel(), (b + db).ravel(), ((s >> k) & 1).ravel())
    out = []
    for name in ("h", "l", "r"):
        st = np.stack(layers[name])
        one, zero = (st == 1).any(0), (st == 0).any(0)
        ok &= not (one & zero).any()
        out.append(one.astype(np.int64))
    return out[0], out[1], out[2], bool(ok)


def tridimer_edges_loop(t):
    """tridimer_edges one state at a time."""
    t = _sq(t).astype(np.int64)
    N = t.shape[0]
    E = N // 2
    h = np.zeros((E + 1, E + 1), np.int64)
    l = np.zeros_like(h)
    r = np.zeros_like(h)
    seen = {}

    def put(arr, name, a, b, v):
        if not (0 <= a <= E and 0 <= b <= E):
            return v == 0
        key = (name, a, b)
        if key in seen:
            return seen[key] == v
        seen[key] = v
        arr[a, b] = v
        return True

    ok = True
    for i in range(N):
        for j in range(N):
            s = int(t[i, j])
            a, b = i // 2, j // 2
            bits = [(s >> k) & 1 for k in range(4)]
            if i % 2 == 0 and j % 2 == 1:
                cells = [(r, "r", a - 1, b), (l, "l", a - 1, b + 1), (r, "r", a, b), (l, "l", a, b)]
            elif i % 2 == 1 and j % 2 == 0:
                cells = [(h, "h", a, b), (r, "r", a, b), (h, "h", a + 1, b - 1), (r, "r", a, b - 1)]
            elif i % 2 == 1:
                cells = [(h, "h", a, b), (l, "l", a, b + 1), (h, "h", a + 1, b), (l, "l", a, b)]
            else:
                ok &= s == 0
                continue
            for (arr, name, x, y), v in zip(cells, bits):
                ok &= put(arr, name, x, y, v)
    return h, l, r, bool(ok)


def tridimer_vertex_degree(h, l, r):
    """Number of dimers at every vertex."""
    E = h.shape[0]
    z = np.zeros((1, E), h.dtype)
    zc = np.zeros((E, 1), h.dtype)
    up = lambda a: np.vstack([z, a[:-1]])          # a[i-1][j]
    left = lambda a: np.hstack([zc, a[:, :-1]])    # a[i][j-1]
    return h + left(h) + l + up(l) + up(r) + left(r)


def tridimer_legal(t, covered=None) -> bool:
    """Consistent encoding, a matching (no vertex meets two dimers) and, when `covered` (the
    start tiling's covered vertices) is given, the same set of covered vertices."""
    h, l, r, ok = tridimer_edges(t)
    if not ok:
        return False
    deg = tridimer_vertex_degree(h, l, r)
    if deg.max() > 1:
        return False
    return covered is None or np.array_equal(deg == 1, covered)


def tridimer_covered(t):
    h, l, r, _ = tridimer_edges(t)
    return tridimer_vertex_degree(h, l, r) == 1
