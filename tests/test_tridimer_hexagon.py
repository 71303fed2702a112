"""TriangleDimerTiler::Hexagon: dimer covers of hexagons of the triangular lattice (no
counterpart in the reference, whose only region is the cone).  Checked on the host: the
covered vertex set is exactly the hexagon (rows of the stated lengths), every vertex meets
exactly one dimer (a perfect matching), hexagons with an odd vertex count (every regular one)
are reported untileable, and the oracle's triangle-dimer chain keeps the cover legal."""
import numpy as np
import pytest

import oracle as O
from randomtilings_amd import statics as S
from randomtilings_amd.params import default_table
from tests.tilings import tridimer_covered, tridimer_edges, tridimer_legal


def region(x, y, z):
    """Lattice vertices (a, b) of the hexagon as the constructor places it."""
    pts = set()
    for r in range(y + z + 1):
        lo = 2 + y - min(r, y)
        hi = 2 + y + x - max(0, r - y)
        pts.update((2 + r, b) for b in range(lo, hi + 1))
    return pts


def sides(limit, rng, count):
    out = []
    while len(out) < count:
        x, y = (int(v) for v in rng.integers(1, limit, 2))
        z = int(rng.integers(1, x + y + 1))
        if len(region(x, y, z)) % 2 == 0:
            out.append((x, y, z))
    return out


@pytest.mark.parametrize("xyz", [(8, 3, 11), (1, 1, 2), (2, 1, 1), (6, 6, 5), (20, 20, 19), (21, 20, 20)] +
                         sides(40, np.random.default_rng(4), 12))
def test_hexagon_is_a_perfect_matching(xyz):
    x, y, z = xyz
    t = S.call("tridimer.Hexagon", args=[x, y, z])
    h, l, r, ok = tridimer_edges(t)
    assert ok
    cov = tridimer_covered(t)
    want = np.zeros_like(cov)
    for a, b in region(x, y, z):
        want[a, b] = True
    assert np.array_equal(cov, want)
    assert int(h.sum() + l.sum() + r.sum()) * 2 == len(region(x, y, z))
    assert tridimer_legal(t, want)


@pytest.mark.parametrize("xyz", [(3, 3, 3), (10, 10, 10), (4, 2, 3), (20, 20, 21)])
def test_odd_hexagon_has_no_cover(xyz):
    assert len(region(*xyz)) % 2 == 1
    assert not S.call("tridimer.Hexagon", args=list(xyz)).any()


def test_oracle_chain_keeps_hexagon_cover():
    t0 = S.call("tridimer.Hexagon", args=[12, 9, 14])
    cov = tridimer_covered(t0)
    t = O.tridimer_walk(t0, 3000, 17, default_table(t0.size))
    assert tridimer_legal(t, cov)
    assert not np.array_equal(t, t0)
