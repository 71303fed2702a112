"""Exact domino placement probabilities and the statistical check of chain samples.

The weighted domino rules (RT_DOMINO_CFTP, RT_DOMINO_WEIGHTED) are heat-bath chains: at an
interior vertex v whose 2 x 2 block of squares is covered by two parallel dominoes, the block
is set to the pair of vertical dominoes (vertex state 12: the edges to the left and right of v
are crossed) with probability p_v and to the horizontal pair (state 3) otherwise, with
p_v = p12[2 ((i + j) mod 2) + (i mod 2)] (dominokernelCFTP.cl:11-31 for p = 1/2).  Its
stationary law is the dimer measure mu(T) ~ prod_{dominoes d in T} w(d) for any edge weights
w with w(vertical pair) / w(horizontal pair) = p_v / (1 - p_v) at every such block: the
heat-bath conditional of that measure is exactly p_v, and flips connect all tilings of a
simply connected region.  Such weights exist on every domain (horizontal dominoes weigh 1;
along each row of vertical-domino positions the block equations form a path, solved left to
right).

Placement probabilities then follow from Kasteleyn's theorem: with K[b, w] = w(bw) for a
horizontal domino and i * w(bw) for a vertical one (every square face gets the alternating
product -1), |det K| is the weighted number of tilings and
P(domino bw) = K[b, w] * K^-1[w, b]  (Kasteleyn 1961; Kenyon, "Local statistics of lattice
dimers", 1997).  Dense linear algebra: fine up to a few thousand squares per colour.
"""
from __future__ import annotations

import numpy as np


def _cells(domain: np.ndarray):
    d = np.asarray(domain).astype(bool)
    M = int(round(np.sqrt(d.size)))
    if M * M != d.size:
        raise ValueError("domain must be an M x M array of squares")
    return d.reshape(M, M), M


def vertical_weights(domain: np.ndarray, p12) -> np.ndarray:
    """Weights y[r, c] of the vertical domino on squares (r, c), (r + 1, c) (horizontal
    dominoes weigh 1) realising the heat-bath probabilities p12 (4 values, indexed as the
    chain indexes them) on every interior block."""
    d, M = _cells(domain)
    p = np.asarray(p12, np.float64).reshape(4)
    if np.any(p <= 0) or np.any(p >= 1):
        raise ValueError("p12 must lie strictly between 0 and 1")
    y = np.ones((M, M))
    for r in range(M - 1):
        for c in range(M - 1):
            # block of squares (r, c), (r, c+1), (r+1, c), (r+1, c+1) around vertex (r+1, c+1)
            if not (d[r, c] and d[r, c + 1] and d[r + 1, c] and d[r + 1, c + 1]):
                continue
            i, j = r + 1, c + 1
            pv = p[2 * ((i + j) % 2) + (i % 2)]
            # y[r, c] * y[r, c+1] = pv / (1 - pv), solved along the row of vertical positions
            y[r, c + 1] = pv / (1.0 - pv) / y[r, c]
    return y


def exact_placement(domain: np.ndarray, p12=(0.5, 0.5, 0.5, 0.5)):
    """(PH, PV, log_Z): PH[r, c] = P(horizontal domino on (r, c), (r, c+1)), PV[r, c] =
    P(vertical domino on (r, c), (r+1, c)) under the chain's stationary law, and the log of
    the weighted number of tilings.  Raises ValueError for a domain with no tiling."""
    d, M = _cells(domain)
    y = vertical_weights(d, p12)
    cells = np.argwhere(d)
    black = [tuple(x) for x in cells if (x[0] + x[1]) % 2 == 0]
    white = [tuple(x) for x in cells if (x[0] + x[1]) % 2 == 1]
    if len(black) != len(white):
        raise ValueError("the domain has no tiling (colour counts differ)")
    bi = {c: k for k, c in enumerate(black)}
    wi = {c: k for k, c in enumerate(white)}
    n = len(black)
    K = np.zeros((n, n), np.complex128)
    for (r, c), k in bi.items():
        for dr, dc in ((0, -1), (0, 1), (-1, 0), (1, 0)):
            q = (r + dr, c + dc)
            if q in wi:
                if dr == 0:
                    K[k, wi[q]] = 1.0
                else:
                    top = min(r, q[0])
                    K[k, wi[q]] = 1j * y[top, c]
    sign, logdet = np.linalg.slogdet(K)
    if not np.isfinite(logdet) or abs(sign) == 0:
        raise ValueError("the domain has no tiling (Kasteleyn determinant vanishes)")
    Kinv = np.linalg.inv(K)
    PH = np.zeros((M, M))
    PV = np.zeros((M, M))
    for (r, c), k in bi.items():
        for dr, dc in ((0, -1), (0, 1), (-1, 0), (1, 0)):
            q = (r + dr, c + dc)
            if q in wi:
                pr = (K[k, wi[q]] * Kinv[wi[q], k]).real
                if dr == 0:
                    PH[r, min(c, q[1])] = pr
                else:
                    PV[min(r, q[0]), c] = pr
    return PH, PV, float(logdet)


def indicators(t: np.ndarray, M: int):
    """(H, V) 0/1 arrays of a domino tiling in vertex-state form ((M+1)^2 states): H[r, c] = 1
    iff a horizontal domino covers squares (r, c), (r, c+1) (the edge between them, from vertex
    (r+1, c+1) to (r, c+1), is crossed: bit 1 of that vertex), V[r, c] = 1 iff a vertical
    domino covers (r, c), (r+1, c) (bit 8 of vertex (r+1, c))."""
    N = M + 1
    s = np.asarray(t).reshape(N, N)
    H = ((s[1:, 1:] & 1) != 0)[:, :M].astype(np.int64)
    V = ((s[1:, :M] & 8) != 0).astype(np.int64)
    return H, V


def zscores(samples, domain: np.ndarray, p12=(0.5, 0.5, 0.5, 0.5)):
    """z-scores of the observed domino frequencies in independent exact samples against the
    exact placement probabilities, over every position with 0 < P < 1 (each count is binomial).
    Returns (z array, number of positions)."""
    d, M = _cells(domain)
    PH, PV, _ = exact_placement(d, p12)
    cH = np.zeros((M, M))
    cV = np.zeros((M, M))
    S = 0
    for t in samples:
        H, V = indicators(t, M)
        cH += H
        cV += V
        S += 1
    P = np.concatenate([PH.ravel(), PV.ravel()])
    cnt = np.concatenate([cH.ravel(), cV.ravel()])
    live = (P > 1e-9) & (P < 1 - 1e-9)
    dead_ok = np.all(cnt[~live & (P <= 1e-9)] == 0) and np.all(cnt[~live & (P >= 1 - 1e-9)] == S)
    if not dead_ok:
        raise AssertionError("a domino of probability 0 or 1 was (not) observed")
    z = (cnt[live] - S * P[live]) / np.sqrt(S * P[live] * (1 - P[live]))
    return z, int(live.sum())
