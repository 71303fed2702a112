"""Bit-exact parity of the HIP chains against the CPU oracle (through the C ABI).

Every case runs the reference-compatible Walk on the GPU and the oracle's sequential
restatement on the CPU with the same tiling, seed and TinyMT parameter table, and requires
identical final tilings.  Both execution paths (stepwise launches and the persistent
LDS-tiled kernel, every compiled tile configuration) are covered, with the uniform
parameter table (kernel-argument parameters) and a per-generator scrambled table.
"""
import numpy as np
import pytest

import oracle as O
import randomtilings_amd as rt
from randomtilings_amd.params import default_table, scrambled_table
from tests.tilings import domino_legal, lozenge_legal

pytestmark = pytest.mark.gpu

BIG = 1 << 20


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def gctx():
    c = rt.Context(0)
    c.load_params(default_table(BIG))
    yield c
    c.close()


@pytest.fixture(scope="module")
def sctx():
    c = rt.Context(0)
    c.load_params(scrambled_table(BIG, 11))
    yield c
    c.close()


# every compiled persistent tile configuration of each model (the C ABI enumerates them
# without a device)
N_CFGS = {m: rt.tile_config_count(m) for m in (rt.RT_DOMINO, rt.RT_LOZENGE, rt.RT_RECTTRI)}


def paths(model):
    out = [(rt.RT_PATH_STEPWISE, -1)]
    out += [(rt.RT_PATH_PERSISTENT, i) for i in range(N_CFGS[model])]
    return out


@pytest.mark.parametrize("path,cfg", paths(rt.RT_DOMINO))
@pytest.mark.parametrize("rule", [rt.RT_DOMINO_LAZY, rt.RT_DOMINO_CFTP])
def test_domino_walk_parity(gctx, golden, rule, path, cfg):
    # dominokernel.cl's lazy rule runs on the packed sweep as a toggle below one threshold
    for n, steps in [(1, 37), (3, 200), (16, 301), (64, 517)]:
        t0 = golden[f"domino_aztec{n}_min"]
        say("domino", rule, path, cfg, n, steps)
        gctx.set_path(path, cfg)
        try:
            g = gctx.domino_walk(t0, steps, 1234 + n, rule=rule)
        finally:
            gctx.set_path(rt.RT_PATH_AUTO, -1)
        o = O.domino_walk(t0, steps, 1234 + n, default_table(BIG), rule=rule)
        assert np.array_equal(g, o), (n, steps)
        assert domino_legal(g, golden[f"domino_aztec{n}_domain"])


def test_domino_aztec128_long(gctx, golden):
    t0 = golden["domino_aztec128_min"]
    say("domino aztec128 5000 steps")
    g = gctx.domino_walk(t0, 5000, 20261203, rule=rt.RT_DOMINO_CFTP)
    o = O.domino_walk(t0, 5000, 20261203, default_table(BIG), rule=1)
    assert np.array_equal(g, o)


def test_domino_scrambled_params(sctx, golden):
    t0 = golden["domino_aztec16_max"]
    for path, cfg in [(rt.RT_PATH_STEPWISE, -1), (rt.RT_PATH_PERSISTENT, 0), (rt.RT_PATH_PERSISTENT, 2)]:
        say("domino scrambled", path, cfg)
        sctx.set_path(path, cfg)
        try:
            g = sctx.domino_walk(t0, 403, 99, rule=rt.RT_DOMINO_LAZY)
        finally:
            sctx.set_path(rt.RT_PATH_AUTO, -1)
        o = O.domino_walk(t0, 403, 99, scrambled_table(BIG, 11), rule=0)
        assert np.array_equal(g, o)


def test_domino_epochs_parity(gctx, golden):
    t0 = golden["domino_aztec8_max"]
    steps = [256, 256, 128]
    seeds = [345903, 17, 2**40 + 5]
    say("domino epochs")
    g = gctx.domino_walk_epochs(t0, steps, seeds, rule=rt.RT_DOMINO_CFTP)
    o = O.domino_walk_epochs(t0, steps, seeds, default_table(BIG), rule=1)
    assert np.array_equal(g, o)


@pytest.mark.parametrize("cfg", [-1] + [i for i in range(N_CFGS[rt.RT_DOMINO])
                                        if rt.tile_config_packed(rt.RT_DOMINO, i)])
def test_domino_weighted_parity(gctx, golden, cfg):
    """Weighted heat-bath (a threshold per class and row parity), default path and every
    packed configuration, on Aztec diamonds of order 16 and 64."""
    p12 = [0.5, 0.5, 0.2, 0.8]
    for n, steps in [(16, 600), (64, 333)]:
        t0 = golden[f"domino_aztec{n}_min"]
        say("domino weighted", cfg, n)
        gctx.set_path(rt.RT_PATH_PERSISTENT if cfg >= 0 else rt.RT_PATH_AUTO, cfg)
        try:
            g = gctx.domino_walk(t0, steps, 5, rule=rt.RT_DOMINO_WEIGHTED, p12=p12)
        finally:
            gctx.set_path(rt.RT_PATH_AUTO, -1)
        o = O.domino_walk(t0, steps, 5, default_table(BIG), rule=2, p12=p12)
        assert np.array_equal(g, o), n


@pytest.mark.parametrize("path,cfg", paths(rt.RT_LOZENGE))
@pytest.mark.parametrize("rule,q", [(rt.RT_LOZENGE_UNIFORM, 0.99), (rt.RT_LOZENGE_QREF, 0.999),
                                    (rt.RT_LOZENGE_QHEAT, 0.7)])
def test_lozenge_walk_parity(gctx, golden, rule, q, path, cfg):
    for key, steps in [("hex3_0_min", 50), ("hex10_2_max", 333), ("hex50_0_min", 407)]:
        t0 = golden[f"lozenge_{key}"]
        say("lozenge", rule, q, path, cfg, key, steps)
        gctx.set_path(path, cfg)
        try:
            g = gctx.lozenge_walk(t0, steps, 77, rule=rule, q=q)
        finally:
            gctx.set_path(rt.RT_PATH_AUTO, -1)
        o = O.lozenge_walk(t0, steps, 77, default_table(BIG), rule=rule, q=q)
        assert np.array_equal(g, o), key
        assert lozenge_legal(g)


@pytest.mark.parametrize("cfg", range(N_CFGS[rt.RT_LOZENGE]))
@pytest.mark.parametrize("rule,q", [(rt.RT_LOZENGE_UNIFORM, 0.99), (rt.RT_LOZENGE_QREF, 0.999),
                                    (rt.RT_LOZENGE_QHEAT, 5.0)])
def test_lozenge_multitile_parity(gctx, rule, q, cfg):
    """A hexagon of side 150 (N = 307: a grid of several tiles in both directions for every
    configuration), a step count that is no multiple of the block length, persistent and
    blocked paths: neighbours exchange only their edge bands between blocks."""
    from randomtilings_amd import statics as St
    t0 = St.call("lozenge.MinTiling", St.call("lozenge.AlmostHexDomain", args=[150, 0]))
    o = O.lozenge_walk(t0, 203, 19, default_table(BIG), rule=rule, q=q)
    for path in (rt.RT_PATH_PERSISTENT, rt.RT_PATH_BLOCKED):
        say("lozenge multitile", rule, q, path, cfg)
        gctx.set_path(path, cfg)
        try:
            g = gctx.lozenge_walk(t0, 203, 19, rule=rule, q=q)
        finally:
            gctx.set_path(rt.RT_PATH_AUTO, -1)
        assert np.array_equal(g, o), path


@pytest.mark.parametrize("rule,q", [(rt.RT_LOZENGE_QREF, 0.999), (rt.RT_LOZENGE_QHEAT, 5.0)])
def test_lozenge_skipped_tiles_parity(gctx, rule, q):
    """A hexagon of side 300 (N = 609): the packed sweep leaves out the corner tiles that lie
    outside the hexagon (fewer workgroups ran than the grid has tiles, exactly the tiles the
    host skip map keeps), with skipping on and off, over every tile configuration: 203 steps
    equal to the oracle's."""
    from randomtilings_amd import statics as St
    t0 = St.call("lozenge.MaxTiling", St.call("lozenge.AlmostHexDomain", args=[300, 0]))
    o = O.lozenge_walk(t0, 203, 23, default_table(BIG), rule=rule, q=q)
    skipped_cfgs = 0
    for cfg in range(N_CFGS[rt.RT_LOZENGE]):
        tr, tc, h, _ = rt.tile_config(rt.RT_LOZENGE, cfg)
        m, ok = rt.lozenge_skip_map(t0, tr, tc, h)
        assert ok
        for skip in (True, False):
            say("lozenge skipped tiles", rule, q, cfg, skip)
            ch = gctx.chain(rt.RT_LOZENGE, rule, t0, [q])
            try:
                ch.set_skip(skip)
                ch.set_path(rt.RT_PATH_PERSISTENT, cfg)
                ch.seed(23)
                ch.run(203)
                g = ch.read()
                tiles, ran = ch.last_grid
            finally:
                ch.close()
            assert tiles == m.size
            if skip and ran < tiles:
                assert ran == int(m.sum())
                skipped_cfgs += 1
            else:
                assert ran == tiles
            assert np.array_equal(g, o), (cfg, skip)
    assert skipped_cfgs >= 8


def test_lozenge_inconsistent_tiling_runs_every_tile(gctx):
    """A hexagon-300 tiling with one edge bit that no longer agrees with its mirror is not a
    fixed point of the update, so nothing is skipped (every tile runs) and the walk still
    equals the oracle's, which repairs the bit in its first steps exactly as the sweep does."""
    from randomtilings_amd import statics as St
    t0 = St.call("lozenge.MaxTiling", St.call("lozenge.AlmostHexDomain", args=[300, 0])).copy()
    N = int(round(np.sqrt(t0.size)))
    t0.reshape(N, N)[N // 2, N // 2] ^= 1
    m, ok = rt.lozenge_skip_map(t0, 56, 84, 4)
    assert not ok and m.min() == 1
    o = O.lozenge_walk(t0, 150, 29, default_table(BIG), rule=rt.RT_LOZENGE_QREF, q=0.999)
    ch = gctx.chain(rt.RT_LOZENGE, rt.RT_LOZENGE_QREF, t0, [0.999])
    try:
        ch.set_path(rt.RT_PATH_PERSISTENT, 0)
        ch.seed(29)
        ch.run(150)
        g = ch.read()
        tiles, ran = ch.last_grid
    finally:
        ch.close()
    assert ran == tiles
    assert np.array_equal(g, o)


def test_lozenge_scrambled_params(sctx, golden):
    t0 = golden["lozenge_hex20_0_max"]
    for path, cfg in [(rt.RT_PATH_STEPWISE, -1), (rt.RT_PATH_PERSISTENT, 0), (rt.RT_PATH_PERSISTENT, 2)]:
        say("lozenge scrambled", path, cfg)
        sctx.set_path(path, cfg)
        try:
            g = sctx.lozenge_walk(t0, 250, 3, rule=rt.RT_LOZENGE_QREF, q=0.9)
        finally:
            sctx.set_path(rt.RT_PATH_AUTO, -1)
        o = O.lozenge_walk(t0, 250, 3, scrambled_table(BIG, 11), rule=1, q=0.9)
        assert np.array_equal(g, o)


@pytest.mark.parametrize("path,cfg", paths(rt.RT_RECTTRI))
def test_recttri_walk_parity(gctx, golden, path, cfg):
    for key, steps in [("minHex5", 100), ("maxHex10", 300), ("slopeHex20", 250)]:
        t0 = golden[f"recttri_{key}"]
        say("recttri", path, cfg, key, steps)
        gctx.set_path(path, cfg)
        try:
            g = gctx.recttri_walk(t0, steps, 31)
        finally:
            gctx.set_path(rt.RT_PATH_AUTO, -1)
        o = O.recttri_walk(t0, steps, 31, default_table(BIG))
        assert np.array_equal(g, o), key


def test_recttri_weighted_parity(gctx, golden):
    w = np.random.default_rng(3).uniform(0.3, 1.7, 16).astype(np.float32)
    w[0] = 0
    t0 = golden["recttri_minHex10"]
    say("recttri weighted")
    g = gctx.recttri_walk(t0, 400, 8, weights16=w)
    o = O.recttri_walk(t0, 400, 8, default_table(BIG), weights16=w)
    assert np.array_equal(g, o)


@pytest.mark.parametrize("cfg", [i for i in range(N_CFGS[rt.RT_RECTTRI]) if rt.tile_config_packed(rt.RT_RECTTRI, i)])
def test_recttri_rowwave_multitile(gctx, sctx, cfg):
    """The row-wave sweep (extended tile rows of 64 positions) on maxHex(60) (129 x 129: several
    tiles in both directions), a step count that is no multiple of the block length, persistent
    and blocked paths, with the default rule, random weights, and a per-generator parameter
    table."""
    from randomtilings_amd import statics as St
    t0 = St.call("recttri.maxHex", args=[60])
    w = np.random.default_rng(5).uniform(0.3, 1.7, 16).astype(np.float32)
    w[0] = 0
    cases = [(gctx, None, default_table(BIG)), (gctx, w, default_table(BIG)), (sctx, w, scrambled_table(BIG, 11))]
    for ctx, wts, table in cases:
        o = O.recttri_walk(t0, 203, 17, table, weights16=wts)
        for path in (rt.RT_PATH_PERSISTENT, rt.RT_PATH_BLOCKED):
            say("recttri row-wave", cfg, path, wts is not None, ctx is sctx)
            ctx.set_path(path, cfg)
            try:
                g = ctx.recttri_walk(t0, 203, 17, weights16=wts)
            finally:
                ctx.set_path(rt.RT_PATH_AUTO, -1)
            assert np.array_equal(g, o), (path, wts is not None, ctx is sctx)


def test_chain_runs_compose(gctx, golden):
    """run(a) then run(b) on a device chain equals one Walk of a+b steps."""
    t0 = golden["lozenge_hex20_0_min"]
    ch = gctx.chain(rt.RT_LOZENGE, rt.RT_LOZENGE_QREF, t0, [0.999])
    ch.seed(42)
    for s in [1, 7, 100, 13]:
        ch.run(s)
    g = ch.read()
    ch.close()
    o = O.lozenge_walk(t0, 121, 42, default_table(BIG), rule=1, q=0.999)
    assert np.array_equal(g, o)


def test_zero_steps_is_identity(gctx, golden):
    t0 = golden["domino_aztec5_max"]
    assert np.array_equal(gctx.domino_walk(t0, 0, 1), t0)


def test_short_param_table_is_an_error(golden):
    c = rt.Context(0)
    c.load_params(default_table(10))
    with pytest.raises(rt.RTError) as e:
        c.domino_walk(golden["domino_aztec5_max"], 10, 1)
    assert e.value.code == rt.RT_ERR_PARAMS
    c.close()


def test_autotune_leaves_chain_untouched(gctx, golden):
    """Autotuning probes a scratch clone: the tuned chain's trajectory equals the oracle's."""
    t0 = golden["lozenge_hex50_0_min"]
    ch = gctx.chain(rt.RT_LOZENGE, rt.RT_LOZENGE_QREF, t0, [0.999])
    ch.seed(9)
    ch.run(50)
    best = ch.autotune(64)
    assert 0 <= best < N_CFGS[rt.RT_LOZENGE]
    ch.run(70)
    g = ch.read()
    ch.close()
    o = O.lozenge_walk(t0, 120, 9, default_table(BIG), rule=1, q=0.999)
    assert np.array_equal(g, o)


# ---- triangle dimers ------------------------------------------------------------------------
N_TD = rt.tile_config_count(rt.RT_TRIDIMER)
TD_PATHS = [(rt.RT_PATH_AUTO, -1)] + [(rt.RT_PATH_STEPWISE, i) for i in range(N_TD)]


@pytest.mark.parametrize("path,cfg", TD_PATHS)
def test_tridimer_walk_parity(gctx, golden, path, cfg):
    """Small cones (one workgroup holds the whole move lattice on the AUTO path; one tiled
    launch per step on the STEPWISE path) against the oracle."""
    from tests.tilings import tridimer_covered, tridimer_legal
    for key, steps in [("cone2_0", 40), ("cone6_0", 300), ("cone9_2", 333), ("cone22_0", 517)]:
        t0 = golden[f"tridimer_{key}"]
        say("tridimer", path, cfg, key, steps)
        gctx.set_path(path, cfg)
        try:
            g = gctx.tridimer_walk(t0, steps, 4242)
        finally:
            gctx.set_path(rt.RT_PATH_AUTO, -1)
        o = O.tridimer_walk(t0, steps, 4242, default_table(BIG))
        assert np.array_equal(g, o), key
        assert tridimer_legal(g, tridimer_covered(t0))


@pytest.mark.parametrize("path,cfg", TD_PATHS)
def test_tridimer_hexagon_parity(gctx, path, cfg):
    """Triangle-dimer hexagons (TriangleDimerTiler::Hexagon covers) on every path: a small one
    (one workgroup on the AUTO path) and one spanning several tiles."""
    from randomtilings_amd import statics as S
    from tests.tilings import tridimer_covered, tridimer_legal
    for xyz, steps in [((8, 3, 11), 301), ((41, 40, 40), 150)]:
        t0 = S.call("tridimer.Hexagon", args=list(xyz))
        assert t0.any()
        say("tridimer hexagon", path, cfg, xyz)
        gctx.set_path(path, cfg)
        try:
            g = gctx.tridimer_walk(t0, steps, 515)
        finally:
            gctx.set_path(rt.RT_PATH_AUTO, -1)
        assert np.array_equal(g, O.tridimer_walk(t0, steps, 515, default_table(BIG))), xyz
        assert tridimer_legal(g, tridimer_covered(t0))


@pytest.mark.parametrize("cfg", range(N_TD))
def test_tridimer_large_cone_tiled(gctx, cfg):
    """A cone whose move lattice spans several tiles of every configuration (M = 154): the
    halo replay must reproduce the reference's kernel sequence across tile borders."""
    from randomtilings_amd import statics as S
    t0 = S.call("tridimer.IceCreamCone", args=[150, 0])
    say("tridimer large", cfg)
    gctx.set_path(rt.RT_PATH_STEPWISE, cfg)
    try:
        g = gctx.tridimer_walk(t0, 120, 99)
    finally:
        gctx.set_path(rt.RT_PATH_AUTO, -1)
    o = O.tridimer_walk(t0, 120, 99, default_table(BIG))
    assert np.array_equal(g, o)


def test_tridimer_scrambled_params(sctx, golden):
    t0 = golden["tridimer_cone14_0"]
    # per-generator parameters on the one-step and two-step tiled kernels (3, 4: two steps)
    for path, cfg in [(rt.RT_PATH_AUTO, -1), (rt.RT_PATH_STEPWISE, 0), (rt.RT_PATH_STEPWISE, 3),
                      (rt.RT_PATH_STEPWISE, 4)]:
        say("tridimer scrambled", path, cfg)
        sctx.set_path(path, cfg)
        try:
            g = sctx.tridimer_walk(t0, 250, 8)
        finally:
            sctx.set_path(rt.RT_PATH_AUTO, -1)
        assert np.array_equal(g, O.tridimer_walk(t0, 250, 8, scrambled_table(BIG, 11)))


def test_tridimer_chain_runs_compose(gctx, golden):
    """Runs of 1, 9, 150 and 40 steps end where one 200-step walk ends: the tiled path's draw
    blocks (32 steps of rotation bits) restart with every run, on one- and two-step tiles."""
    from randomtilings_amd import statics as S
    t0 = golden["tridimer_cone10_0"]
    t_big = S.call("tridimer.IceCreamCone", args=[150, 0])
    for t, path, cfg in [(t0, rt.RT_PATH_AUTO, -1), (t0, rt.RT_PATH_STEPWISE, -1), (t_big, rt.RT_PATH_STEPWISE, 0),
                         (t_big, rt.RT_PATH_STEPWISE, 3)]:
        say("tridimer compose", path, cfg)
        ch = gctx.chain(rt.RT_TRIDIMER, 0, t)
        ch.set_path(path, cfg)
        ch.seed(31)
        for s in [1, 9, 150, 40]:
            ch.run(s)
        g = ch.read()
        ch.close()
        assert np.array_equal(g, O.tridimer_walk(t, 200, 31, default_table(BIG))), (path, cfg)


@pytest.mark.parametrize("path,cfg", [(rt.RT_PATH_AUTO, -1), (rt.RT_PATH_STEPWISE, 0), (rt.RT_PATH_STEPWISE, 4)])
def test_tridimer_arbitrary_states_parity(gctx, path, cfg):
    """Uniformly random 4-bit lozenge states (no tiling at all): the derived triangle and
    butterfly states reach their largest values (up to 126 and 766), which the 64-bit cells
    must hold without loss; the walk still equals the oracle's int arithmetic."""
    rng = np.random.default_rng(2024)
    for N in (60, 180):
        t0 = rng.integers(0, 16, N * N).astype(np.int32)
        say("tridimer arbitrary states", path, cfg, N)
        gctx.set_path(path, cfg)
        try:
            g = gctx.tridimer_walk(t0, 97, 123)
        finally:
            gctx.set_path(rt.RT_PATH_AUTO, -1)
        assert np.array_equal(g, O.tridimer_walk(t0, 97, 123, default_table(BIG))), N


# ---- coupling from the past -------------------------------------------------------------------
@pytest.mark.parametrize("mode", [rt.RT_CFTP_REFERENCE, rt.RT_CFTP_DOUBLING])
@pytest.mark.parametrize("rule,p12", [(rt.RT_DOMINO_CFTP, None), (rt.RT_DOMINO_WEIGHTED, [0.6, 0.4, 0.45, 0.55])])
def test_domino_cftp_parity(gctx, golden, mode, rule, p12):
    """rt_domino_cftp (coupled device chains, device coupling test) returns the oracle's
    sample, final replay length and restart count (the cases coalesce within 17 restarts of
    the reference schedule and 8 doublings)."""
    cases = [(2, 345903), (5, 345903), (16, 11)]
    if mode == rt.RT_CFTP_DOUBLING:  # the reference schedule grows only linearly per restart
        cases.append((32, 7))
    bound = 24 if mode == rt.RT_CFTP_REFERENCE else 9
    for n, seed in cases:
        tmax, tmin = golden[f"domino_aztec{n}_max"], golden[f"domino_aztec{n}_min"]
        say("cftp", mode, rule, n, seed)
        g = gctx.domino_cftp(tmax, tmin, seed, rule=rule, p12=p12, mode=mode, max_restarts=bound)
        o = O.domino_cftp(tmax, tmin, seed, default_table(BIG), rule=rule, p12=p12, mode=mode, max_restarts=bound)
        assert np.array_equal(g[0], o[0]) and g[1:] == o[1:], (n, g[1:], o[1:])


@pytest.mark.parametrize("cfg", [i for i in range(N_CFGS[rt.RT_DOMINO]) if rt.tile_config_packed(rt.RT_DOMINO, i)])
@pytest.mark.parametrize("rule,p12", [(rt.RT_DOMINO_CFTP, None), (rt.RT_DOMINO_WEIGHTED, [0.6, 0.4, 0.45, 0.55])])
def test_domino_cftp_packed_configs(gctx, golden, cfg, rule, p12):
    """Coupled pairs on the packed sweep (two 16-bit positions per word) in every packed
    configuration, persistent and blocked (the one-block pair kernel): the oracle's sample,
    replay length and restart count."""
    for n, seed in [(16, 11), (32, 7)]:
        tmax, tmin = golden[f"domino_aztec{n}_max"], golden[f"domino_aztec{n}_min"]
        o = O.domino_cftp(tmax, tmin, seed, default_table(BIG), rule=rule, p12=p12, mode=rt.RT_CFTP_DOUBLING,
                          max_restarts=9)
        for path in (rt.RT_PATH_PERSISTENT, rt.RT_PATH_BLOCKED):
            say("cftp packed", cfg, rule, n, path)
            gctx.set_path(path, cfg)
            try:
                g = gctx.domino_cftp(tmax, tmin, seed, rule=rule, p12=p12, mode=rt.RT_CFTP_DOUBLING, max_restarts=9)
            finally:
                gctx.set_path(rt.RT_PATH_AUTO, -1)
            assert np.array_equal(g[0], o[0]) and g[1:] == o[1:], (n, path, g[1:], o[1:])


def test_domino_cftp_progress(gctx, golden):
    """rt_domino_cftp_ex reports every replay (index, length, coalesced) and returns the same
    sample as rt_domino_cftp: doubling from 8 steps, replays of 8, 16, 32, ... steps."""
    tmax, tmin = golden["domino_aztec16_max"], golden["domino_aztec16_min"]
    seen = []
    inside = []
    g = gctx.domino_cftp(tmax, tmin, 11, mode=rt.RT_CFTP_DOUBLING, first_steps=8, max_restarts=20,
                         progress=lambda r, n, c: (inside if c is None else seen).append((r, n, c)))
    h = gctx.domino_cftp(tmax, tmin, 11, mode=rt.RT_CFTP_DOUBLING, first_steps=8, max_restarts=20)
    assert np.array_equal(g[0], h[0]) and g[1:] == h[1:]
    assert [r for r, _, _ in seen] == list(range(g[2] + 1))
    assert [n for _, n, _ in seen] == [8 << r for r in range(g[2] + 1)]
    assert [c for _, _, c in seen] == [False] * g[2] + [True]
    assert seen[-1][1] == g[1]
    # within replay r (epochs 8 << (r - 1), ..., 8 newest first): the steps after each epoch
    assert inside == [(r, sum(8 << e for e in range(r - 1, r - 1 - k, -1)), None)
                      for r in range(1, g[2] + 1) for k in range(1, r + 1)]


def test_domino_cftp_bound(gctx, golden):
    tmax, tmin = golden["domino_aztec16_max"], golden["domino_aztec16_min"]
    with pytest.raises(rt.RTError) as e:
        gctx.domino_cftp(tmax, tmin, 11, max_restarts=0, first_steps=4)
    assert e.value.code == rt.RT_ERR_NOT_COALESCED


@pytest.mark.parametrize("cfg", [0, 1] + [i for i in range(N_CFGS[rt.RT_DOMINO]) if rt.tile_config_packed(rt.RT_DOMINO, i)])
def test_domino_blocked_path(gctx, cfg):
    """The blocked path (one launch per block of h steps, the launch boundary being the
    exchange; what a tile grid too large to be resident runs) on an Aztec diamond of order
    700 (~2000 tiles of 16 x 32)."""
    from randomtilings_amd import statics as S
    d = S.call("domino.AztecDiamond", args=[700])
    t0 = S.call("domino.MinTiling", d)
    say("domino blocked", cfg)
    ch = gctx.chain(rt.RT_DOMINO, rt.RT_DOMINO_CFTP, t0)
    try:
        ch.set_path(rt.RT_PATH_BLOCKED, cfg)
        ch.seed(5)
        ch.run(37)
        launches = ch.last_launches
        g = ch.read()
    finally:
        ch.close()
    h = rt.tile_config(rt.RT_DOMINO, cfg)[2]
    assert launches == (37 + h - 1) // h
    assert np.array_equal(g, O.domino_walk(t0, 37, 5, default_table(BIG), rule=1))


def test_context_outlives_its_chains(golden):
    """Destroying a context while a chain is live defers the release to the chain's
    destruction; the chain keeps working meanwhile."""
    t0 = golden["lozenge_hex10_2_min"]
    c = rt.Context(0)
    c.load_params(default_table(BIG))
    ch = c.chain(rt.RT_LOZENGE, rt.RT_LOZENGE_QREF, t0, [0.9])
    c.close()
    ch.seed(3)
    ch.run(77)
    g = ch.read()
    ch.close()
    assert np.array_equal(g, O.lozenge_walk(t0, 77, 3, default_table(BIG), rule=1, q=0.9))
