"""Bit-exact parity at the benchmark workloads' full shapes (GPU vs the CPU oracle).

* Domino Aztec diamond of order 128 (N = 262), dominokernelCFTP.cl's rule, the full 10^5
  steps from the minimal tiling, on the autotuned persistent configuration.
* Lozenge hexagon a = b = c = 1000 (AlmostHexDomain(1000, 0), N = 2007), the q-rule with
  q = 0.999, from the empty room, on the autotuned packed configuration with tile skipping:
  two runs totalling 407 steps (the generator and state hand-off between runs); and the same
  workload at its stated length (10^6 sweeps), checked through the tiling's validity.
* Triangle dimers on the cone of side 1002 and a hexagon near side 2048 for long runs, checked
  through the validity of the final dimer covers.
* One run crossing the persistent kernel's launch boundary (2 * 2048 * h + 3 steps in one
  rt_chain_run) on a forced packed configuration with skipping on.
* A tile grid too large to be resident (Aztec diamond of order 700) on the AUTO path.
* A batch shaped like the batched-domains workload: one domain of each kind near tiling
  side 2048 plus small ones, one batched run, every chain against its own oracle walk.
* The coupling-from-the-past shape: Aztec diamond of order 2048 (N = 4102) under the weighted
  heat-bath p12 = (0.5, 0.5, 0.8, 0.2): Walk(t, steps[], seeds[]) and one bounded replay of
  the coupled pair (both chains in one device chain) against two oracle walks.
"""
import numpy as np
import pytest

import oracle as O
import randomtilings_amd as rt
from randomtilings_amd import statics as S
from randomtilings_amd.params import default_table
from tests.tilings import domino_legal, lozenge_legal

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def fctx():
    c = rt.Context(0)
    c.load_params(default_table(4102 * 2051))
    yield c
    c.close()


def test_c1_aztec128_full_length(fctx, golden):
    t0 = golden["domino_aztec128_min"]
    say("c1: Aztec 128, 100000 steps")
    ch = fctx.chain(rt.RT_DOMINO, rt.RT_DOMINO_CFTP, t0)
    try:
        ch.autotune(4096)
        ch.seed(20261203)
        ch.run(100000)
        g = ch.read()
    finally:
        ch.close()
    o = O.domino_walk(t0, 100000, 20261203, default_table(1 << 20), rule=rt.RT_DOMINO_CFTP)
    assert np.array_equal(g, o)
    assert domino_legal(g, golden["domino_aztec128_domain"])


def test_c2_hexagon1000_tuned(fctx):
    d = S.call("lozenge.AlmostHexDomain", args=[1000, 0])
    t0 = S.call("lozenge.MinTiling", d)
    N = int(round(np.sqrt(t0.size)))
    assert N == 2007
    say("c2: hexagon 1000, autotune")
    ch = fctx.chain(rt.RT_LOZENGE, rt.RT_LOZENGE_QREF, t0, [0.999])
    try:
        cfg = ch.autotune(16384)
        ch.seed(20261203)
        say("c2: config", cfg, ch.tile_configs()[cfg])
        ch.run(150)
        ch.run(257)
        g = ch.read()
        tiles, ran = ch.last_grid
    finally:
        ch.close()
    if rt.tile_config_packed(rt.RT_LOZENGE, cfg):
        m, ok = rt.lozenge_skip_map(t0, *rt.tile_config(rt.RT_LOZENGE, cfg)[:3])
        assert ok and ran == int(m.sum()) and ran < tiles
    o = O.lozenge_walk(t0, 407, 20261203, default_table(t0.size), rule=rt.RT_LOZENGE_QREF, q=0.999)
    assert np.array_equal(g, o)
    assert lozenge_legal(g)


def test_c2_hexagon1000_full_length(fctx):
    """The hexagon 1000 workload at its stated length, 10^6 sweeps of the three colour classes
    (3 * 10^6 chain steps, about 370 persistent launches with tile skipping), beyond what the
    oracle can follow: the run must end without the device error flag (no change reached a
    skipped tile) in a lozenge tiling of exactly the hexagon's faces, with cubes stacked."""
    from tests.tilings import lozenge_cubes
    d = S.call("lozenge.AlmostHexDomain", args=[1000, 0])
    t0 = S.call("lozenge.MinTiling", d)
    say("c2: hexagon 1000, 3e6 chain steps")
    ch = fctx.chain(rt.RT_LOZENGE, rt.RT_LOZENGE_QREF, t0, [0.999])
    try:
        ch.autotune(16384)
        ch.seed(20261204)
        ch.run(3_000_000)
        g = ch.read()
        launches = ch.last_launches
    finally:
        ch.close()
    say("c2 full length: launches", launches, "cubes", lozenge_cubes(g))
    assert launches >= 3_000_000 // (2048 * 8)
    assert lozenge_legal(g, d)
    assert lozenge_cubes(g) > lozenge_cubes(t0)


def test_tridimer_long_runs_stay_valid(fctx):
    """Triangle dimers at full size for long runs, beyond the oracle's reach: the cone of side
    1002 (the triangle-dimer workload) for 10^5 steps and a hexagon near tiling side 2048 (the
    batched-domains mix) for 2 * 10^4 steps end in dimer covers of the same vertex sets."""
    from randomtilings_amd import domains
    from tests.tilings import tridimer_covered, tridimer_legal
    cone = S.call("tridimer.IceCreamCone", args=[1002, 0])
    hexagon = S.call("tridimer.Hexagon", args=list(domains.tridimer_hexagon_sides(2048)))
    for name, t0, steps in (("cone 1002", cone, 100_000), ("hexagon", hexagon, 20_000)):
        say("tridimer long run:", name, int(round(np.sqrt(t0.size))), steps)
        g = fctx.tridimer_walk(t0, steps, 31337)
        assert tridimer_legal(g, tridimer_covered(t0)), name
        assert not np.array_equal(g, t0), name


@pytest.mark.parametrize("h", [4, 8])
def test_one_run_crosses_launch_boundaries(fctx, h):
    """A persistent run splits into launches of 2048 * h steps; one rt_chain_run of
    2 * 2048 * h + 3 steps (three launches) on a packed configuration of halo h with tile
    skipping on equals the oracle's walk."""
    cfg = next(i for i in range(rt.tile_config_count(rt.RT_LOZENGE))
               if rt.tile_config_packed(rt.RT_LOZENGE, i) and rt.tile_config(rt.RT_LOZENGE, i)[2] == h)
    t0 = S.call("lozenge.MaxTiling", S.call("lozenge.AlmostHexDomain", args=[60, 0]))
    steps = 2 * 2048 * h + 3
    say("launch boundary", h, cfg, steps)
    ch = fctx.chain(rt.RT_LOZENGE, rt.RT_LOZENGE_QREF, t0, [0.999])
    try:
        ch.set_path(rt.RT_PATH_PERSISTENT, cfg)
        ch.seed(77)
        ch.run(steps)
        g = ch.read()
        launches = ch.last_launches
    finally:
        ch.close()
    assert launches == 3
    assert np.array_equal(g, O.lozenge_walk(t0, steps, 77, default_table(t0.size), rule=1, q=0.999))


def test_auto_path_aztec700(fctx):
    """AUTO on a tile grid larger than the device holds at once (Aztec diamond of order 700,
    N = 1406): the run takes the blocked path (or a cooperative launch the runtime accepts)
    and equals the oracle."""
    t0 = S.call("domino.MinTiling", S.call("domino.AztecDiamond", args=[700]))
    say("auto aztec700")
    ch = fctx.chain(rt.RT_DOMINO, rt.RT_DOMINO_CFTP, t0)
    try:
        ch.seed(5)
        ch.run(45)
        g = ch.read()
    finally:
        ch.close()
    assert np.array_equal(g, O.domino_walk(t0, 45, 5, default_table(t0.size), rule=1))


def test_batched_domains_near_2048(fctx):
    from randomtilings_amd import domains

    rng = np.random.default_rng(2048)
    specs = []
    for side in (2048, 96):
        for kind in domains.KINDS:
            m, r, t = domains.domain(fctx, kind, side - int(rng.integers(0, 9)), rng)
            specs.append((m, r, t, kind))
    steps = 27
    chs = [fctx.chain(m, r, t) for m, r, t, _ in specs]
    try:
        for k, ch in enumerate(chs):
            ch.seed(900 + k)
        b = rt.Batch(fctx)
        for ch in chs:
            b.add(ch)
        say("batch near 2048:", [(k, int(round(np.sqrt(t.size)))) for _, _, t, k in specs])
        b.run(steps, True)
        b.sync()
        got = [ch.read() for ch in chs]
        b.close()
    finally:
        for ch in chs:
            ch.close()
    walk = {rt.RT_DOMINO: lambda t, sd, p: O.domino_walk(t, steps, sd, p, rule=1),
            rt.RT_TRIDIMER: lambda t, sd, p: O.tridimer_walk(t, steps, sd, p),
            rt.RT_RECTTRI: lambda t, sd, p: O.recttri_walk(t, steps, sd, p)}
    for k, ((m, _, t, kind), g) in enumerate(zip(specs, got)):
        say("oracle", kind, int(round(np.sqrt(t.size))))
        assert np.array_equal(g, walk[m](t, 900 + k, default_table(t.size))), (k, kind)


def test_cftp_shape_aztec2048_weighted(fctx):
    p12 = [0.5, 0.5, 0.8, 0.2]
    d = S.call("domino.AztecDiamond", args=[2048])
    tmin = fctx.extremal_tiling(rt.RT_DOMINO, d, False)[0]
    tmax = fctx.extremal_tiling(rt.RT_DOMINO, d, True)[0]
    N = int(round(np.sqrt(tmin.size)))
    assert N == 4102
    steps, seeds = [20, 13], [345903, 2**33 + 7]
    params = default_table(N * N // 2)
    say("c4 shape: walk epochs")
    g = fctx.domino_walk_epochs(tmin, steps, seeds, rule=rt.RT_DOMINO_WEIGHTED, p12=p12)
    omin = O.domino_walk_epochs(tmin, steps, seeds, params, rule=2, p12=p12)
    assert np.array_equal(g, omin)
    say("c4 shape: coupled pair replay")
    ch = fctx.chain(rt.RT_DOMINO_PAIR, rt.RT_DOMINO_WEIGHTED, tmax | (tmin << 8), p12)
    try:
        for n, sd in zip(steps, seeds):
            ch.seed(sd, mode=rt.RT_SEED_EPOCH)
            ch.run(n)
        pair = ch.read()
        cfg = ch.last_config
    finally:
        ch.close()
    say("c4 shape: pair config", cfg)
    omax = O.domino_walk_epochs(tmax, steps, seeds, params, rule=2, p12=p12)
    assert np.array_equal(pair & 0xFF, omax)
    assert np.array_equal(pair >> 8, omin)
