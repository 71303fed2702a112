"""Batched runs (rt_batch_*): many chains of different models, sizes and rules advanced
together — one blocked launch per block of h steps per model, optionally replayed from a
captured HIP graph — end exactly where individual rt_chain_run calls end, run after run."""
import numpy as np
import pytest

import oracle as O
import randomtilings_amd as rt
from randomtilings_amd import statics as S
from randomtilings_amd.params import default_table

pytestmark = pytest.mark.gpu
BIG = 1 << 20


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def bctx():
    c = rt.Context(0)
    c.load_params(default_table(BIG))
    yield c
    c.close()


def _specs(golden):
    specs = []
    for n in (3, 16, 64, 130):
        d = S.call("domino.AztecDiamond", args=[n])
        specs.append((rt.RT_DOMINO, rt.RT_DOMINO_CFTP, S.call("domino.MinTiling", d), None))
    specs.append((rt.RT_DOMINO, rt.RT_DOMINO_WEIGHTED, golden["domino_aztec32_max"], [0.6, 0.4, 0.45, 0.55]))
    specs.append((rt.RT_DOMINO, rt.RT_DOMINO_LAZY, golden["domino_rect10x10_min"], None))
    for key, q in (("hex10_2_min", 0.9), ("hex50_0_min", 0.999), ("hex100_0_max", 1.2)):
        specs.append((rt.RT_LOZENGE, rt.RT_LOZENGE_QREF, golden[f"lozenge_{key}"], [q]))
    specs.append((rt.RT_LOZENGE, rt.RT_LOZENGE_UNIFORM, golden["lozenge_hex20_0_max"], [0.99]))
    for key in ("minHex10", "slopeHex20"):
        specs.append((rt.RT_RECTTRI, 0, golden[f"recttri_{key}"], None))
    specs.append((rt.RT_TRIDIMER, 0, golden["tridimer_cone22_0"], None))
    return specs


@pytest.mark.parametrize("graph", [True, False])
def test_batch_equals_individual_runs(bctx, golden, graph):
    specs = _specs(golden)
    solo, batched = [], []
    b = rt.Batch(bctx)
    try:
        for k, (m, rule, t0, rp) in enumerate(specs):
            a = bctx.chain(m, rule, t0, rp)
            c = bctx.chain(m, rule, t0, rp)
            a.seed(1000 + k)
            c.seed(1000 + k)
            solo.append(a)
            batched.append(c)
            b.add(c)
        built = []
        for steps in (37, 200, 1, 64, 33):
            say("batch", graph, steps)
            b.run(steps, graph)
            built.append(b.last_graphs)
            for a in solo:
                a.run(steps)
            b.sync()
            for k, (a, c) in enumerate(zip(solo, batched)):
                assert np.array_equal(a.read(), c.read()), (k, steps)
        assert b.last_launches > 0
        # one graph (a 32-step round of every model) captured by the first run long enough to
        # use it and replayed by the later ones
        assert built == ([1, 0, 0, 0, 0] if graph else [0] * 5)
        if graph:  # a chain added later changes the batch: the next run captures a new graph
            m, rule, t0, rp = specs[0]
            a, c = bctx.chain(m, rule, t0, rp), bctx.chain(m, rule, t0, rp)
            a.seed(5)
            c.seed(5)
            solo.append(a)
            batched.append(c)
            b.add(c)
            b.run(70, True)
            assert b.last_graphs == 1
            for x in solo:
                x.run(70)
            b.sync()
            for k, (a, c) in enumerate(zip(solo, batched)):
                assert np.array_equal(a.read(), c.read()), (k, 70)
    finally:
        b.close()
        for ch in solo + batched:
            ch.close()


def test_batch_matches_oracle(bctx, golden):
    t0 = golden["lozenge_hex50_0_min"]
    c = bctx.chain(rt.RT_LOZENGE, rt.RT_LOZENGE_QREF, t0, [0.999])
    d = bctx.chain(rt.RT_DOMINO, rt.RT_DOMINO_CFTP, golden["domino_aztec16_min"])
    b = rt.Batch(bctx)
    try:
        c.seed(5)
        d.seed(6)
        b.add(c)
        b.add(d)
        b.run(150)
        b.run(91)
        assert np.array_equal(c.read(), O.lozenge_walk(t0, 241, 5, default_table(BIG), rule=1, q=0.999))
        assert np.array_equal(d.read(), O.domino_walk(golden["domino_aztec16_min"], 241, 6, default_table(BIG), rule=1))
    finally:
        b.close()
        c.close()
        d.close()


@pytest.mark.parametrize("cfg", range(rt.tile_config_count(rt.RT_TRIDIMER)))
def test_batch_tridimer_configs(bctx, golden, cfg):
    """Triangle-dimer chains batched on every tile configuration (3 and up: two steps per
    launch, an odd last step on the one-step kernel), over odd and even step counts: each chain
    ends where its own run ends.  The largest chain's configuration serves the batch."""
    t_big = S.call("tridimer.IceCreamCone", args=[150, 0])  # move lattice 154: several tiles
    ts = [golden["tridimer_cone22_0"], t_big, golden["tridimer_cone9_2"]]
    solo, batched = [], []
    b = rt.Batch(bctx)
    try:
        for k, t0 in enumerate(ts):
            a = bctx.chain(rt.RT_TRIDIMER, 0, t0)
            c = bctx.chain(rt.RT_TRIDIMER, 0, t0)
            a.seed(77 + k)
            c.seed(77 + k)
            c.set_path(rt.RT_PATH_STEPWISE, cfg)
            solo.append(a)
            batched.append(c)
            b.add(c)
        for steps in (37, 1, 64, 9):
            say("batch tridimer", cfg, steps)
            b.run(steps, True)
            for a in solo:
                a.run(steps)
            b.sync()
            for k, (a, c) in enumerate(zip(solo, batched)):
                assert np.array_equal(a.read(), c.read()), (k, steps)
            assert batched[1].last_config == cfg
    finally:
        b.close()
        for ch in solo + batched:
            ch.close()


def test_batch_autotune_leaves_chains_untouched(bctx, golden):
    """rt_batch_autotune times every applicable configuration of each model on the batch's own
    chains and puts them back as they were: after it, the batch continues exactly where
    individual runs of the same chains end, on the configurations it chose."""
    specs = _specs(golden)
    solo, batched = [], []
    b = rt.Batch(bctx)
    try:
        for k, (m, rule, t0, rp) in enumerate(specs):
            a, c = bctx.chain(m, rule, t0, rp), bctx.chain(m, rule, t0, rp)
            a.seed(300 + k)
            c.seed(300 + k)
            solo.append(a)
            batched.append(c)
            b.add(c)
        b.run(45)
        chosen = b.autotune(40)
        say("batch autotune", chosen)
        assert set(chosen) == {rt.RT_DOMINO, rt.RT_LOZENGE, rt.RT_RECTTRI, rt.RT_TRIDIMER}
        b.run(77)
        b.sync()
        for a in solo:
            a.run(45)
            a.run(77)
        for k, (a, c) in enumerate(zip(solo, batched)):
            assert np.array_equal(a.read(), c.read()), k
        for m, cfg in chosen.items():
            assert all(c.last_config == cfg for c, (mm, _, _, _) in zip(batched, specs) if mm == m)
    finally:
        b.close()
        for ch in solo + batched:
            ch.close()


def test_batch_set_config_every_configuration(bctx, golden):
    """rt_batch_set_config: the mixed batch runs on configuration i of every model at once (i
    cycling through each model's compiled configurations), run after run, and still ends exactly
    where individual runs of the same chains end; out-of-range indices and a packed
    configuration for a rule that cannot use it are refused."""
    specs = _specs(golden)
    ncfg = {m: rt.tile_config_count(m) for m in (rt.RT_DOMINO, rt.RT_LOZENGE, rt.RT_RECTTRI, rt.RT_TRIDIMER)}
    solo, batched = [], []
    b = rt.Batch(bctx)
    try:
        for k, (m, rule, t0, rp) in enumerate(specs):
            a, c = bctx.chain(m, rule, t0, rp), bctx.chain(m, rule, t0, rp)
            a.seed(500 + k)
            c.seed(500 + k)
            solo.append(a)
            batched.append(c)
            b.add(c)
        total = 0
        for i in range(max(ncfg.values())):
            for m, n in ncfg.items():
                b.set_config(m, i % n)
            steps = 19 + 3 * i
            say("batch set_config", i, steps)
            b.run(steps, i % 2 == 0)
            total += steps
            for m, n in ncfg.items():
                assert all(c.last_config == i % n for c, s in zip(batched, specs) if s[0] == m), (i, m)
        b.sync()
        for a in solo:
            a.run(total)
        for k, (a, c) in enumerate(zip(solo, batched)):
            assert np.array_equal(a.read(), c.read()), k
        for m, n in ncfg.items():
            with pytest.raises(rt.RTError):
                b.set_config(m, n)
        with pytest.raises(rt.RTError):
            b.set_config(rt.RT_DOMINO_PAIR, 0)
    finally:
        b.close()
        for ch in solo + batched:
            ch.close()
    # a weighted domino rule with P(12) = 1 in one class does not split k inside (0, 2^23): the
    # packed sweep cannot run it, so forcing a packed configuration fails the run
    packed = [i for i in range(ncfg[rt.RT_DOMINO]) if rt.tile_config_packed(rt.RT_DOMINO, i)]
    ch = bctx.chain(rt.RT_DOMINO, rt.RT_DOMINO_WEIGHTED, golden["domino_aztec16_min"], [1.0, 0.5, 0.5, 0.5])
    b = rt.Batch(bctx)
    try:
        ch.seed(3)
        b.add(ch)
        b.set_config(rt.RT_DOMINO, packed[0])
        with pytest.raises(rt.RTError):
            b.run(8)
        b.set_config(rt.RT_DOMINO, -1)
        b.run(8)
        b.sync()
    finally:
        b.close()
        ch.close()
