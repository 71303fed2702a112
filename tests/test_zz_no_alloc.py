"""rt_lozenge_walk_noalloc (the allocation-free lozenge walk, device memory in a file-scope
array) against rt_lozenge_walk through the C ABI: bit-identical tilings on consecutive calls,
the second on a smaller input (the array keeps the first call's contents), a walk longer than
one schedule piece, the largest hexagon the benchmark uses, and the size limit."""
import numpy as np
import pytest

import oracle as O
import randomtilings_amd as rt
from randomtilings_amd import statics as S
from randomtilings_amd.params import default_table

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def nctx():
    c = rt.Context(0)
    c.load_params(default_table(2271 * 757))
    yield c
    c.close()


def hexagon(a, maximal=False):
    d = S.call("lozenge.AlmostHexDomain", args=[a, 0])
    return S.call("lozenge.MaxTiling" if maximal else "lozenge.MinTiling", d)


def test_two_calls_second_smaller(nctx):
    big, small = hexagon(300, True), hexagon(50)
    say("noalloc: hexagon 300, 3001 q-rule steps")
    g1 = nctx.lozenge_walk_noalloc(big, 3001, 17, rule=rt.RT_LOZENGE_QREF, q=0.999)
    r1 = nctx.lozenge_walk(big, 3001, 17, rule=rt.RT_LOZENGE_QREF, q=0.999)
    assert np.array_equal(g1, r1)
    assert not np.array_equal(g1, big)
    say("noalloc: hexagon 50, 777 uniform steps")
    g2 = nctx.lozenge_walk_noalloc(small, 777, 5, rule=rt.RT_LOZENGE_UNIFORM)
    r2 = nctx.lozenge_walk(small, 777, 5, rule=rt.RT_LOZENGE_UNIFORM)
    assert np.array_equal(g2, r2)
    assert np.array_equal(g2, O.lozenge_walk(small, 777, 5, default_table(small.size), rule=rt.RT_LOZENGE_UNIFORM))


def test_walk_longer_than_one_schedule_piece(nctx):
    t0 = hexagon(5)
    steps = (1 << 20) + 1234
    say("noalloc: hexagon 5,", steps, "steps")
    g = nctx.lozenge_walk_noalloc(t0, steps, 99, rule=rt.RT_LOZENGE_QHEAT, q=0.7)
    assert np.array_equal(g, nctx.lozenge_walk(t0, steps, 99, rule=rt.RT_LOZENGE_QHEAT, q=0.7))
    assert np.array_equal(g, O.lozenge_walk(t0, steps, 99, default_table(t0.size), rule=rt.RT_LOZENGE_QHEAT, q=0.7))


def test_hexagon_1000_then_small(nctx):
    t0 = hexagon(1000)
    say("noalloc: hexagon 1000, 300 steps")
    g = nctx.lozenge_walk_noalloc(t0, 300, 20261203, rule=rt.RT_LOZENGE_QREF, q=0.999)
    assert np.array_equal(g, nctx.lozenge_walk(t0, 300, 20261203, rule=rt.RT_LOZENGE_QREF, q=0.999))
    s = hexagon(10, True)
    g = nctx.lozenge_walk_noalloc(s, 401, 3, rule=rt.RT_LOZENGE_QREF, q=0.9)
    assert np.array_equal(g, O.lozenge_walk(s, 401, 3, default_table(s.size), rule=rt.RT_LOZENGE_QREF, q=0.9))


def test_size_limit(nctx):
    t = np.zeros((2274, 2274), np.int32)
    with pytest.raises(rt.RTError) as e:
        nctx.lozenge_walk_noalloc(t, 1, 1)
    assert e.value.code == rt.RT_ERR_ARG and "scratch" in str(e.value)
