"""Learned warm start end to end at a small scale: the U-Net trained on the GPU from chain
samples, its prediction projected exactly onto an admissible height function (a legal
tiling of the domain between the minimal and maximal ones), and the chain started there
reaching the observable's equilibrium in fewer steps than the plain chain from the minimal
tiling."""
import numpy as np
import pytest

import randomtilings_amd as rt
from randomtilings_amd import warmstart as ws
from randomtilings_amd.params import default_table
from tests.tilings import domino_legal

pytestmark = pytest.mark.gpu


def test_warm_start_beats_the_minimal_start():
    rng = np.random.default_rng(5)
    ctx = rt.Context(0)
    ctx.load_params(default_table(32 * 32))
    canvas = 32
    doms = [ws.Domain(ctx, d) for _, d in ws.domain_bank(rng, 32, canvas)]
    samples = ws.equilibrium_samples(ctx, doms, 6000, 8, 11)
    X, T, M = ws.build_dataset(ctx, doms, samples, canvas, average=True)
    assert len(X) == 32
    net = rt.UNet(ctx, 3, 16, ws.IN_CH, seed=2)
    losses = ws.train(net, X, T, M, 150, 8, 3e-3, rng)
    print("loss", losses[0], losses[-1], flush=True)
    assert np.mean(losses[-5:]) < 0.5 * losses[0]
    gains = []
    for k, (_, d) in enumerate(ws.domain_bank(np.random.default_rng(123), 2, canvas)):
        dom = ws.Domain(ctx, d)
        h = ws.warm_heights(ctx, net, dom, canvas)
        fin = dom.hmax != ws.INF
        assert np.all(h[fin] >= dom.hmin[fin]) and np.all(h[fin] <= dom.hmax[fin])
        tw = ws.warm_tiling(ctx, net, dom, canvas)
        assert domino_legal(tw, dom.d)
        plain = ws.trajectory(ctx, dom, dom.tmin, 8000, 50, 7 + k, chains=64)
        warm = ws.trajectory(ctx, dom, tw, 8000, 50, 7 + k, chains=64)
        eq = 0.5 * (plain[len(plain) // 2:].mean() + warm[len(warm) // 2:].mean())
        sp, sw = ws.steps_to_equilibrium(plain, eq, 50), ws.steps_to_equilibrium(warm, eq, 50)
        print("held-out", dom.N, sp, sw, plain[0], warm[0], eq, flush=True)
        assert abs(warm[0] - eq) < abs(plain[0] - eq)
        gains.append((sp, sw))
    assert sum(s for s, _ in gains) > 1.5 * sum(w for _, w in gains)
    net.close()
    ctx.close()


def test_learned_start_beats_midpoint_on_shape():
    """The ablation: on held-out Aztec diamonds under 2-periodic weights, the projected U-Net
    prediction starts nearer the long-run mean height field than the non-learned midpoint
    start (projected and mean-matched the same way) and reaches equilibrium of the mean
    |h - h_bar| observable, which mean matching does not target, in fewer chain steps; both
    beat the minimal tiling."""
    rng = np.random.default_rng(21)
    ctx = rt.Context(0)
    ctx.load_params(default_table(128 * 128))
    canvas = 32
    doms = [ws.Domain(ctx, d, ws.periodic_weights(rng, 0.6, 1.0))
            for _, d in ws.domain_bank(rng, 48, 2 * canvas)]
    samples = ws.equilibrium_samples(ctx, doms, 30000, 8, 3)
    X, T, M = ws.build_dataset(ctx, doms, samples, canvas, average=True)
    net = rt.UNet(ctx, 3, 16, ws.IN_CH, seed=4)
    ws.train(net, X, T, M, 200, 8, 3e-3, rng)
    hrng = np.random.default_rng(77)
    tot = {"plain": 0, "midpoint": 0, "learned": 0}
    start = {"midpoint": 0.0, "learned": 0.0}
    for n in (36, 44):
        dom = ws.Domain(ctx, ws.aztec(n), ws.periodic_weights(hrng, 0.6, 1.0))
        starts = {"plain": dom.tmin, "midpoint": ws.midpoint_tiling(ctx, dom),
                  "learned": ws.warm_tiling(ctx, net, dom, canvas)}
        every = 100
        res = ws.compare_starts(ctx, dom, starts, 16 * dom.N * dom.N // every * every, every, 9 + n, chains=16)
        print("held-out", n, {k: res[k] for k in ("plain", "midpoint", "learned")}, res["settled"], flush=True)
        assert res["settled"]
        for k in tot:
            tot[k] += res[k]["steps_deviation"]
        for k in start:
            start[k] += res[k]["start_deviation"]
    print("totals", tot, start, flush=True)
    assert start["learned"] < start["midpoint"]
    assert tot["learned"] < tot["midpoint"] < tot["plain"]
    net.close()
    ctx.close()
