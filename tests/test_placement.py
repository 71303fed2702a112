"""Exact domino placement probabilities (randomtilings_amd.placement), pinned three ways:
the Aztec diamond theorem (2^(n(n+1)/2) tilings), brute-force enumeration of every tiling of
small domains weighted by the height-function law the weighted CFTP test pins
(mu ~ prod_v (p_v / (1 - p_v))^((-1)^(i+j) h(v) / 4)), and the oracle's CFTP samples."""
import numpy as np
import pytest

import oracle as O
from randomtilings_amd import placement as PL
from randomtilings_amd import statics as S
from randomtilings_amd.params import default_table

INF = np.iinfo(np.int32).max


def enumerate_tilings(d):
    """Every domino tiling of the M x M 0/1 square array d, as vertex-state arrays
    ((M+1)^2, the bit convention of DominoTiler.cpp:204-206)."""
    M = d.shape[0]
    cells = [tuple(x) for x in np.argwhere(d)]
    out = []

    def rec(covered, dominoes):
        free = [c for c in cells if c not in covered]
        if not free:
            out.append(list(dominoes))
            return
        r, c = free[0]  # row-major first free square: its partner lies right or below
        for q in ((r, c + 1), (r + 1, c)):
            if q[0] < M and q[1] < M and d[q] and q not in covered:
                rec(covered | {(r, c), q}, dominoes + [((r, c), q)])

    rec(frozenset(), [])
    tilings = []
    for doms in out:
        t = np.zeros((M + 1, M + 1), np.int32)
        for (r, c), (r2, c2) in doms:
            if r2 == r:  # horizontal: crosses the segment (r, c+1)-(r+1, c+1)
                t[r + 1, c + 1] |= 1
                t[r, c + 1] |= 2
            else:  # vertical: crosses the segment (r+1, c)-(r+1, c+1)
                t[r + 1, c] |= 8
                t[r + 1, c + 1] |= 4
        tilings.append(t.ravel())
    return tilings


def height_law(tilings, dom, p12):
    M = int(round(np.sqrt(dom.size)))
    N = M + 1
    logw = []
    for t in tilings:
        h = S.call("domino.TilingToHeightfunc", t, dom).reshape(N, N)
        ii, jj = np.nonzero(h != INF)
        pv = np.array([p12[((i + j) % 2) * 2 + (i & 1)] for i, j in zip(ii, jj)])
        logw.append(np.sum((-1.0) ** (ii + jj) * h[ii, jj] / 4 * np.log(pv / (1 - pv))))
    w = np.exp(np.array(logw) - max(logw))
    return w / w.sum()


def notched(M):
    d = np.zeros((M, M), np.int32)
    d[2:M - 2, 2:M - 2] = 1
    d[2, 2:4] = 0
    d[M - 3, M - 4:M - 2] = 0
    return d.ravel()


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6])
def test_aztec_diamond_theorem(n):
    d = S.call("domino.AztecDiamond", args=[n])
    _, _, logz = PL.exact_placement(d)
    assert abs(logz - n * (n + 1) / 2 * np.log(2)) < 1e-8


@pytest.mark.parametrize("which", ["aztec2", "aztec3", "notched8", "square6"])
@pytest.mark.parametrize("p12", [(0.5, 0.5, 0.5, 0.5), (0.5, 0.5, 0.8, 0.2), (0.7, 0.3, 0.6, 0.15)])
def test_placement_matches_enumeration(which, p12):
    if which.startswith("aztec"):
        d = S.call("domino.AztecDiamond", args=[int(which[-1])])
    elif which == "notched8":
        d = notched(8)
    else:
        d = S.call("domino.Rectangle", args=[4, 4])
    M = int(round(np.sqrt(d.size)))
    til = enumerate_tilings(np.asarray(d).reshape(M, M))
    assert len(til) > 1
    if which == "aztec3":
        assert len(til) == 64
    w = height_law(til, d, p12)
    eH = np.zeros((M, M))
    eV = np.zeros((M, M))
    for wt, t in zip(w, til):
        H, V = PL.indicators(t, M)
        eH += wt * H
        eV += wt * V
    PH, PV, _ = PL.exact_placement(d, p12)
    assert np.allclose(PH, eH, atol=1e-9) and np.allclose(PV, eV, atol=1e-9)
    # every square covered exactly once
    cov = PH.copy()
    cov[:, 1:] += PH[:, :-1]
    cov += PV
    cov[1:, :] += PV[:-1, :]
    dm = np.asarray(d).reshape(M, M).astype(bool)
    assert np.allclose(cov[dm], 1.0) and np.allclose(cov[~dm], 0.0)


def test_indicators_round_trip(golden):
    """indicators() reads the reference's own maximal tiling of the order-16 Aztec diamond
    as a perfect cover of its squares."""
    t, d = golden["domino_aztec16_max"], golden["domino_aztec16_domain"]
    M = int(round(np.sqrt(d.size)))
    H, V = PL.indicators(t, M)
    cov = H.copy()
    cov[:, 1:] += H[:, :-1]
    cov += V
    cov[1:, :] += V[:-1, :]
    assert np.array_equal(cov, np.asarray(d).reshape(M, M))


def test_oracle_cftp_samples_match_placement(golden):
    """Independent exact samples (the oracle's coupling from the past, weighted rule) against
    the exact placement probabilities of the order-4 Aztec diamond."""
    p12 = (0.5, 0.5, 0.8, 0.2)
    d = S.call("domino.AztecDiamond", args=[4])
    tmax = S.call("domino.MaxTiling", d)
    tmin = S.call("domino.MinTiling", d)
    p = default_table(tmax.size)
    samples = [O.domino_cftp(tmax, tmin, 1000003 * s + 11, p, rule=2, p12=p12, mode=1)[0] for s in range(600)]
    z, k = PL.zscores(samples, d, p12)
    assert k > 20
    assert np.max(np.abs(z)) < 4.5
    assert abs(np.mean(z ** 2) - 1) < 0.3
