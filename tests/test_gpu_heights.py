"""Device max/min tilings and height-function closure (rt_extremal_tiling,
rt_height_closure) against the host statics, which are themselves pinned to the reference's
MaxTiling / MinTiling (DominoTiler.cpp:214-400, LozengeTiler.cpp:140-318) by the golden
fixtures: identical tilings, identical heights, identical tileability verdicts."""
import numpy as np
import pytest

import randomtilings_amd as rt
from randomtilings_amd import statics as S
from tests.tilings import domino_legal, lozenge_legal

pytestmark = pytest.mark.gpu
INF = np.iinfo(np.int32).max


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def hctx():
    c = rt.Context(0)
    yield c
    c.close()


def _side(d, model):
    return int(round(np.sqrt(d.size / (2 if model == rt.RT_LOZENGE else 1))))


def _host(model, d, maximal):
    m = "domino" if model == rt.RT_DOMINO else "lozenge"
    t = S.call(f"{m}.{'Max' if maximal else 'Min'}Tiling", d)
    return t, S.call(f"{m}.TilingToHeightfunc", t, d)


def _check(hctx, model, d, legal):
    for maximal in (True, False):
        ht, hh = _host(model, d, maximal)
        gt, gh = hctx.extremal_tiling(model, d, maximal)
        assert np.array_equal(gt, ht), maximal
        assert np.array_equal(gh, hh), maximal
        assert legal(gt, d)


@pytest.mark.parametrize("n", [1, 2, 5, 16, 64, 128])
def test_domino_aztec(hctx, golden, n):
    say("aztec", n)
    d = golden[f"domino_aztec{n}_domain"]
    _check(hctx, rt.RT_DOMINO, d, domino_legal)
    gt, _ = hctx.extremal_tiling(rt.RT_DOMINO, d, True)
    assert np.array_equal(gt, golden[f"domino_aztec{n}_max"])


@pytest.mark.parametrize("key", ["hex5_1", "hex10_2", "hex50_0", "hex100_0"])
def test_lozenge_hexagons(hctx, golden, key):
    say("hexagon", key)
    d = golden[f"lozenge_{key}_domain"]
    for maximal, kind in ((True, "max"), (False, "min")):
        gt, _ = hctx.extremal_tiling(rt.RT_LOZENGE, d, maximal)
        assert np.array_equal(gt, golden[f"lozenge_{key}_{kind}"])
    _check(hctx, rt.RT_LOZENGE, d, lambda t, dd: lozenge_legal(t, dd))


def test_large_aztec_diamond(hctx):
    """A domain spanning many relaxation tiles (tiling side 1006)."""
    d = S.call("domino.AztecDiamond", args=[500])
    say("aztec 500")
    for maximal in (True, False):
        ht, _ = _host(rt.RT_DOMINO, d, maximal)
        gt, _ = hctx.extremal_tiling(rt.RT_DOMINO, d, maximal)
        assert np.array_equal(gt, ht)


def _notched_rectangle(rng, m, n):
    d = S.call("domino.Rectangle", args=[m, n])
    side = int(round(np.sqrt(d.size)))
    d = d.reshape(side, side).copy()
    cut = 0
    for i in range(m):
        cut = min(n - 1, cut + int(rng.integers(0, 3))) if i < m // 2 else max(0, cut - int(rng.integers(0, 3)))
        d[2 + i, 2:2 + cut // 2] = 0
    return d.ravel()


def test_random_domino_domains_and_tileability(hctx, capfd):
    """Random notched rectangles, tileable or not: the device verdict matches the host's."""
    rng = np.random.default_rng(12)
    seen = {True: 0, False: 0}
    for trial in range(40):
        m, n = int(rng.integers(3, 40)), int(rng.integers(3, 40))
        d = _notched_rectangle(rng, m, n)
        for maximal in (True, False):
            ht, _ = _host(rt.RT_DOMINO, d, maximal)
            host_ok = "NOT tileable" not in capfd.readouterr().out
            try:
                gt, _ = hctx.extremal_tiling(rt.RT_DOMINO, d, maximal)
                dev_ok = True
            except rt.RTError as e:
                assert e.code == rt.RT_ERR_UNTILEABLE
                dev_ok = False
            assert dev_ok == host_ok, (trial, maximal)
            if dev_ok:
                assert np.array_equal(gt, ht), trial
            seen[host_ok] += 1
    assert seen[True] > 5 and seen[False] > 5


def test_closure_projects_onto_admissible_heights(hctx, golden):
    """Raising the interior of the maximal height function by multiples of 4 (the residue
    class of domino heights) and closing returns the maximal height function; the minimal
    one is already closed; a lowered boundary vertex is reported."""
    d = golden["domino_aztec16_domain"]
    _, hmax = hctx.extremal_tiling(rt.RT_DOMINO, d, True)
    _, hmin = hctx.extremal_tiling(rt.RT_DOMINO, d, False)
    v = S.call("domino.DomainToVertices", d)
    rng = np.random.default_rng(3)
    noisy = hmax.copy()
    inner = (v == 1) & (hmax != INF)
    noisy[inner] += 4 * rng.integers(0, 20, int(inner.sum()))
    fixed = (v == 2).astype(np.uint8)
    h, launches = hctx.height_closure(rt.RT_DOMINO, d, noisy, 1, fixed)
    assert np.array_equal(h, hmax) and launches >= 1
    h, _ = hctx.height_closure(rt.RT_DOMINO, d, hmin, 1, fixed)
    assert np.array_equal(h, hmin)
    # heights between min and max, closed from above and from below, stay between them
    mid = np.where(inner, (hmin + hmax) // 2, hmax)
    up, _ = hctx.height_closure(rt.RT_DOMINO, d, mid, 1, fixed)
    assert np.all(up[inner] >= hmin[inner]) and np.all(up[inner] <= hmax[inner])
    bad = hmax.copy()
    b = np.flatnonzero(v == 2)[5]
    bad[b] += 8
    with pytest.raises(rt.RTError) as e:
        hctx.height_closure(rt.RT_DOMINO, d, bad, -1, fixed)
    assert e.value.code == rt.RT_ERR_UNTILEABLE
