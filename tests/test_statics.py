"""The reference-compatible statics of librandomtilings.so (domains, max/min tilings,
height functions, SVG/text writers) against the reference's own outputs: the committed
golden fixtures always, and the reference compiled into oracle/_ref when present (random
domains, byte-identical SVG documents).  Closed forms and invariants pin what neither
covers: Aztec diamond cell counts, brickwork extremal tilings, max >= min heights, and the
legality of every constructed tiling."""
import os
import tempfile

import numpy as np
import pytest

import oracle as O
from randomtilings_amd import statics as S
from tests.tilings import domino_legal, lozenge_legal

HAVE_REF = O.ref_available()
ref_only = pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref not built (reference sources absent)")


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 8, 16, 32, 64, 128])
def test_domino_aztec_golden(golden, n):
    d = S.call("domino.AztecDiamond", args=[n])
    assert np.array_equal(d, golden[f"domino_aztec{n}_domain"])
    assert d.sum() == 2 * n * (n + 1)  # cells of the Aztec diamond
    tmax = S.call("domino.MaxTiling", d)
    tmin = S.call("domino.MinTiling", d)
    assert np.array_equal(tmax, golden[f"domino_aztec{n}_max"])
    assert np.array_equal(tmin, golden[f"domino_aztec{n}_min"])
    assert domino_legal(tmax, d) and domino_legal(tmin, d)
    if n <= 16:
        assert np.array_equal(S.call("domino.DomainToVertices", d), golden[f"domino_aztec{n}_vertices"])
        hmax = S.call("domino.TilingToHeightfunc", tmax, d)
        hmin = S.call("domino.TilingToHeightfunc", tmin, d)
        assert np.array_equal(hmax, golden[f"domino_aztec{n}_hmax"])
        assert np.array_equal(hmin, golden[f"domino_aztec{n}_hmin"])
        fin = hmax != np.iinfo(np.int32).max
        assert np.all(hmax[fin] >= hmin[fin])  # the maximal height function dominates
        assert np.array_equal(S.call("domino.HeightfuncToTiling", hmax, d), tmax)
        assert np.array_equal(S.call("domino.TilingToDomain", tmax), golden[f"domino_aztec{n}_t2d"])


def test_domino_rectangles_and_aztec_rectangles(golden):
    for (m, n) in [(2, 2), (4, 6), (7, 4), (10, 10)]:
        d = S.call("domino.Rectangle", args=[m, n])
        assert np.array_equal(d, golden[f"domino_rect{m}x{n}_domain"])
        for kind in ["max", "min"]:
            t = S.call(f"domino.{kind.capitalize()}Tiling", d)
            assert np.array_equal(t, golden[f"domino_rect{m}x{n}_{kind}"]), (m, n, kind)
    for (n, rem) in [(4, [1]), (6, [0, 3]), (8, [1, 2, 5])]:
        d = S.call("domino.AztecRectangle", rem, args=[n])
        key = f"domino_azrect{n}_{len(rem)}"
        assert np.array_equal(d, golden[f"{key}_domain"])
        for kind in ["max", "min"]:
            assert np.array_equal(S.call(f"domino.{kind.capitalize()}Tiling", d), golden[f"{key}_{kind}"])


@pytest.mark.parametrize("nm", [(1, 0), (2, 0), (3, 0), (5, 1), (10, 2), (20, 0), (50, 0), (100, 0)])
def test_lozenge_hexagon_golden(golden, nm):
    n, m = nm
    key = f"lozenge_hex{n}_{m}"
    d = S.call("lozenge.AlmostHexDomain", args=[n, m])
    assert np.array_equal(d, golden[f"{key}_domain"])
    tmax = S.call("lozenge.MaxTiling", d)
    tmin = S.call("lozenge.MinTiling", d)
    assert np.array_equal(tmax, golden[f"{key}_max"])
    assert np.array_equal(tmin, golden[f"{key}_min"])
    if n <= 10:
        assert lozenge_legal(tmax, d) and lozenge_legal(tmin, d)
        assert np.array_equal(S.call("lozenge.DomainToVertices", d), golden[f"{key}_vertices"])
        assert np.array_equal(S.call("lozenge.TilingToHeightfunc", tmax, d), golden[f"{key}_hmax"])
        assert np.array_equal(S.call("lozenge.TilingToDomain", tmax), golden[f"{key}_t2d"])
    if m == 0:
        # empty room: the minimal tiling has exactly one addable cube and none to remove;
        # the full box has exactly one removable cube
        assert (tmin == 25).sum() == 1 and (tmin == 38).sum() == 0
        assert (tmax == 38).sum() == 1 and (tmax == 25).sum() == 0


def _staircase_rectangle(rng, m, n):
    """A rectangle with monotone staircase notches cut from two opposite corners."""
    d = S.call("domino.Rectangle", args=[m, n])
    side = int(round(np.sqrt(d.size)))
    d = d.reshape(side, side)
    cut = 0
    for i in range(m):
        cut = min(n - 1, cut + int(rng.integers(0, 2))) if i < m // 2 else max(0, cut - int(rng.integers(0, 2)))
        d[2 + i, 2:2 + cut // 2] = 0
    return d.ravel()


@ref_only
def test_domino_random_domains_match_reference():
    rng = np.random.default_rng(4)
    for trial in range(30):
        m, n = int(rng.integers(2, 12)), int(rng.integers(2, 12))
        d = _staircase_rectangle(rng, m, n)
        for f in ["MaxTiling", "MinTiling", "DomainToVertices"]:
            r = O.ref_call(f"domino.{f}", d)
            mine = S.call(f"domino.{f}", d)
            if f != "DomainToVertices" and not np.any(r):
                continue
            # the reference returns a partial height vector for domains without a tiling;
            # compare where it produced a tiling
            if f != "DomainToVertices" and not domino_legal(r, d):
                continue
            assert np.array_equal(r, mine), (trial, f)
        t = O.ref_call("domino.MaxTiling", d)
        if domino_legal(t, d):
            assert np.array_equal(O.ref_call("domino.TilingToHeightfunc", t, d), S.call("domino.TilingToHeightfunc", t, d))


@ref_only
def test_svg_documents_match_reference(golden):
    cases = [
        ("domino.DomainToSVG", golden["domino_aztec5_domain"], None),
        ("domino.TilingToSVG", golden["domino_aztec5_max"], None),
        ("domino.MayaToSVG", golden["domino_aztec5_min"], None),
        ("domino.DimerToSVG", golden["domino_aztec5_max"], None),
        ("domino.DimerToSVG2", golden["domino_aztec5_max"], golden["domino_aztec5_domain"]),
        ("lozenge.DomainToSVG", golden["lozenge_hex10_2_domain"], None),
        ("lozenge.TilingToSVG", golden["lozenge_hex10_2_max"], None),
        ("lozenge.DimerToSVG", golden["lozenge_hex10_2_max"], None),
        ("lozenge.DimerToSVG2", golden["lozenge_hex10_2_min"], golden["lozenge_hex10_2_domain"]),
        ("lozenge.PrintDomain", golden["lozenge_hex5_1_domain"], None),
        ("common.SaveMatrix", golden["domino_aztec3_max"], None),
    ]
    with tempfile.TemporaryDirectory() as td:
        for name, a, b in cases:
            p1, p2 = os.path.join(td, "ref.svg"), os.path.join(td, "mine.svg")
            if ".DimerToSVG" in name:
                # The reference's dimer drawings index the domain past its end (LozengeTiler.cpp:
                # 540-545 reads row i+1 of the last row), so what it draws there depends on
                # whatever memory follows the vector.  Ours counts those cells as empty; the
                # reference gets the same domain followed by zeros, so both read defined cells.
                d = O.ref_call(name.split(".")[0] + ".TilingToDomain", a) if b is None else b
                dpad = np.concatenate([np.asarray(d, dtype=np.int32).ravel(), np.zeros(np.size(d) + 64, np.int32)])
                O.ref_write(name.split(".")[0] + ".DimerToSVG2", a, p1, dpad)
            else:
                O.ref_write(name, a, p1, b)
            S.write(name, a, p2, b)
            assert open(p1, "rb").read() == open(p2, "rb").read(), name


def test_pad_unpad_roundtrip():
    a = np.arange(12, dtype=np.int32)
    p = S.call("common.PadMatrix", a, args=[3, 4])
    assert p.reshape(5, 6)[1:-1, 1:-1].ravel().tolist() == a.tolist()
    assert np.count_nonzero(p.reshape(5, 6)[[0, -1], :]) == 0
    assert np.array_equal(S.call("common.UnPadMatrix", p, args=[5, 6]), a)


def test_untileable_domain_reports(capfd):
    d = S.call("domino.Rectangle", args=[3, 3])  # odd number of squares
    t = S.call("domino.MaxTiling", d)
    out = capfd.readouterr().out
    assert "NOT tileable" in out
    assert not np.any(t)
