"""Device height observables (rt_observe) against the host height function.

For chains on Aztec diamonds and notched rectangles (walked some steps from the minimal
tiling, plus the extremal tilings themselves), the heights the device scan produces equal
DominoTiler::TilingToHeightfunc of the read-back tiling bit for bit (the host restatement,
itself pinned to the reference's compiled TilingToHeightfunc by the statics tests), and the
reductions equal numpy's sums over the masked vertices."""
import numpy as np
import pytest

import randomtilings_amd as rt
from randomtilings_amd import statics as S
from randomtilings_amd.params import default_table

pytestmark = pytest.mark.gpu
INF = np.iinfo(np.int32).max


@pytest.fixture(scope="module")
def octx():
    c = rt.Context(0)
    c.load_params(default_table(1 << 21))
    yield c
    c.close()


def domains():
    from randomtilings_amd import warmstart as ws
    rng = np.random.default_rng(8)
    return [("aztec5", ws.aztec(5)), ("aztec64", ws.aztec(64)), ("notched90", ws.notched_rectangle(rng, 90)),
            ("notched300", ws.notched_rectangle(rng, 300)), ("aztec700", ws.aztec(700))]


@pytest.mark.parametrize("name,d", domains(), ids=[n for n, _ in domains()])
def test_observe_matches_host_heights(octx, name, d):
    tmin, hmin = octx.extremal_tiling(rt.RT_DOMINO, d, False)
    tmax, _ = octx.extremal_tiling(rt.RT_DOMINO, d, True)
    N = int(round(np.sqrt(tmin.size)))
    rng = np.random.default_rng(3)
    mask = ((hmin != INF) & (rng.random(hmin.size) < 0.7)).astype(np.uint8)
    ref = np.where(hmin != INF, rng.integers(-50000, 50000, hmin.size), 0).astype(np.int32)
    chs = [octx.chain(rt.RT_DOMINO, rt.RT_DOMINO_CFTP, t) for t in (tmin, tmin, tmin, tmax)]
    try:
        for k, (ch, steps) in enumerate(zip(chs, (0, 37, 1001, 5))):
            ch.seed(11 + k)
            ch.run(steps)
        obs = rt.Observer(octx, d, hmin, mask, ref)
        out, hs = obs.observe(chs, heights=True)
        obs.close()
        print(name, N, out[:, :2].tolist(), flush=True)
        for k, ch in enumerate(chs):
            t = ch.read()
            h = S.call("domino.TilingToHeightfunc", t, d)
            assert np.array_equal(hs[k], h), (name, k)
            m = mask.astype(bool)
            hv = h[m].astype(np.int64)
            want = [m.sum(), hv.sum(), (hv * hv).sum(), np.abs(256 * hv - ref[m]).sum()]
            assert out[k].tolist() == [int(x) for x in want], (name, k)
    finally:
        for ch in chs:
            ch.close()


def test_observe_default_mask_no_reference(octx):
    from randomtilings_amd import warmstart as ws
    d = ws.aztec(40)
    tmin, hmin = octx.extremal_tiling(rt.RT_DOMINO, d, False)
    ch = octx.chain(rt.RT_DOMINO, rt.RT_DOMINO_CFTP, tmin)
    ch.seed(2)
    ch.run(500)
    obs = rt.Observer(octx, d, hmin)
    out, hs = obs.observe([ch])
    obs.close()
    h = S.call("domino.TilingToHeightfunc", ch.read(), d)
    ch.close()
    assert hs is None
    fin = h != INF
    assert out[0].tolist() == [int(fin.sum()), int(h[fin].astype(np.int64).sum()),
                               int((h[fin].astype(np.int64) ** 2).sum()), 0]
