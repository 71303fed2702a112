"""Triangle-dimer statics of librandomtilings.so against the reference's own outputs: the
example program's saved start state and domain (Examples/ExampleOuts/TriangleDimer, n=6),
the committed golden fixtures, and the reference compiled into oracle/_ref when present
(every cone up to side 29 with every cut, byte-identical SVG documents).  The dimer-cover
invariant (every covered vertex meets exactly one dimer) pins the encoding itself."""
import os
import re
import tempfile
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from randomtilings_amd import statics as S
from randomtilings_amd.params import default_table
from tests.tilings import tridimer_covered, tridimer_edges, tridimer_vertex_degree

HAVE_REF = O.ref_available()
ref_only = pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref not built (reference sources absent)")
EX = Path(__file__).parent / "golden" / "exampleouts"
CONES = [(2, 0), (3, 0), (6, 0), (9, 2), (10, 0), (14, 0), (22, 0)]


def load_matrix(path):
    return np.array([int(x) for x in open(path).read().split()], dtype=np.int32)


def test_example_outputs():
    """TriangleDimerEx.cpp saves IceCreamCone(6,0) and its domain with SaveMatrix (which
    writes the first floor(sqrt(n))^2 entries of a non-square array)."""
    t = S.call("tridimer.IceCreamCone", args=[6, 0])
    assert np.array_equal(t, load_matrix(EX / "TriangleDimer_DimerStart.txt"))
    d = S.call("tridimer.TilingToDomain", t)
    saved = load_matrix(EX / "TriangleDimer_Domain.txt")
    assert np.array_equal(d[: saved.size], saved)


@pytest.mark.parametrize("n,m", CONES)
def test_cones_golden(golden, n, m):
    key = f"tridimer_cone{n}_{m}"
    t = S.call("tridimer.IceCreamCone", args=[n, m])
    assert np.array_equal(t, golden[key])
    assert np.array_equal(S.call("tridimer.TilingToTriangleFlips", t), golden[key + "_tri"])
    assert np.array_equal(S.call("tridimer.TilingToButterflyFlips", t), golden[key + "_bfly"])
    assert np.array_equal(S.call("tridimer.TilingToDomain", t), golden[key + "_domain"])
    # a perfect matching of the covered vertices, with every edge encoded consistently
    h, l, r, ok = tridimer_edges(t)
    assert ok
    deg = tridimer_vertex_degree(h, l, r)
    assert deg.max() == 1 and (deg == 1).sum() == 2 * (h.sum() + l.sum() + r.sum())
    # the cone has (n+1)(n+2)/2 - m(m+1)/2 vertices
    assert (deg == 1).sum() == (n + 1) * (n + 2) // 2 - m * (m + 1) // 2
    # TilingToVertices agrees with the decoded matching on the interior it scans
    v = S.call("tridimer.TilingToVertices", t)
    E = int(round(np.sqrt(v.size)))
    V = v.reshape(E, E)
    cov = tridimer_covered(t)
    assert np.array_equal(V[2:E - 2, 2:E - 2] == 1, cov[2:E - 2, 2:E - 2])


def test_oracle_derivations_match(golden):
    """The oracle's restated TilingToTriangleFlips / TilingToButterflyFlips agree with the
    product's and with the reference's fixtures."""
    for n, m in CONES:
        t = golden[f"tridimer_cone{n}_{m}"]
        tri, bf = O.tridimer_flips(t)
        assert np.array_equal(tri, golden[f"tridimer_cone{n}_{m}_tri"])
        assert np.array_equal(bf, golden[f"tridimer_cone{n}_{m}_bfly"])


@pytest.mark.parametrize("b", [0, 1, 2])
def test_butterflies_golden(golden, b):
    t = S.call("tridimer.TestButterfly", args=[b])
    assert np.array_equal(t, golden[f"tridimer_butterfly{b}"])
    h, l, r, ok = tridimer_edges(t)
    assert ok and tridimer_vertex_degree(h, l, r).max() == 1


def test_dimer_indicator_counts(golden):
    t = golden["tridimer_cone10_0"]
    h, l, r, _ = tridimer_edges(t)
    # the three orientations partition the dimers
    total = sum(int(S.call("tridimer.DimerIndicator", t, args=[c]).sum()) for c in range(3))
    assert total == int(h.sum() + l.sum() + r.sum())
    assert not np.any(S.call("tridimer.DimerIndicator", t, args=[5]))


@ref_only
def test_cones_match_reference_all_cuts():
    for n in range(1, 30):
        for m in range(0, n):
            t = S.call("tridimer.IceCreamCone", args=[n, m])
            assert np.array_equal(t, O.ref_call("tridimer.IceCreamCone", args=[n, m])), (n, m)
            for f in ["TilingToTriangleFlips", "TilingToButterflyFlips", "TilingToVertices", "TilingToDomain"]:
                assert np.array_equal(S.call("tridimer." + f, t), O.ref_call("tridimer." + f, t)), (f, n, m)
            for c in range(3):
                assert np.array_equal(S.call("tridimer.DimerIndicator", t, args=[c]),
                                      O.ref_call("tridimer.DimerIndicator", t, args=[c])), (n, m, c)


# DimerToSVG of the reference reads d[(i-1)*2M + j] at i = 0, before the start of the domain
# array; whatever lies there decides whether an "eH" edge is drawn on the first row.  Here
# indices outside the array count as outside the domain, so such lines are compared after
# removing them from the reference's output.
_FIRST_ROW_EH = re.compile(r'<use xlink:href = "#eH" x = "[^"]*" y = "0"/>\n')


@ref_only
def test_svg_match_reference(golden):
    t6 = golden["tridimer_cone6_0"]
    walked = O.tridimer_walk(golden["tridimer_cone10_0"], 300, 5, default_table(1 << 12))
    cases = [("tridimer.TilingToSVG", t6, None), ("tridimer.DimerToSVG", t6, None),
             ("tridimer.DimerToSVG2", t6, S.call("tridimer.TilingToDomain", t6)),
             ("tridimer.DomainToSVG", golden["tridimer_cone9_2_domain"], None),
             ("tridimer.TilingToSVG", walked, None), ("tridimer.DimerToSVG", walked, None),
             ("tridimer.DimerToSVG", golden["tridimer_cone9_2"], None)]
    with tempfile.TemporaryDirectory() as td:
        for k, (name, a, b) in enumerate(cases):
            p1, p2 = os.path.join(td, "ref.svg"), os.path.join(td, "mine.svg")
            O.ref_write(name, a, p1, b)
            S.write(name, a, p2, b)
            ref, mine = open(p1).read(), open(p2).read()
            if name.startswith("tridimer.DimerToSVG"):
                ref = _FIRST_ROW_EH.sub("", ref)
                assert not _FIRST_ROW_EH.search(mine)
            assert ref == mine, (k, name)
