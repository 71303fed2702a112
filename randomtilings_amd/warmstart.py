"""Learned warm start for domino chains.

A U-Net (rt_unet, trained on the GPU by this library's kernels) maps a domain — its
minimal and maximal height functions, its vertex mask and its weight field — to the
limit-shape height, expressed as the fraction r = (h - h_min) / (h_max - h_min).  The
prediction is turned into a start state exactly: rounded to the residue class of domino
heights (h = h_min mod 4), clamped between h_min and h_max, and closed on the device
(rt_height_closure: the largest admissible height function below it, which keeps the
boundary since it stays above h_min).  The chain then runs from that tiling; the plain
chain from the minimal tiling remains the fallback and the yardstick.

Chains run the weighted heat-bath (RT_DOMINO_WEIGHTED) with a 2-periodic weight pattern drawn
per domain: squares of one colour class keep probability 1/2, those of the other alternate
row by row between a^2/(a^2+b^2) and b^2/(a^2+b^2), a and b uniform in [0.3, 1] (a = b is the
uniform rule of dominokernelCFTP.cl).  The network sees the pattern's contrast as an input
channel.  A domain larger than the training canvas is predicted on its features sampled down
to the canvas (heights are given as h/N, which the limit shape keeps under scaling) and the
predicted fraction is interpolated back.

Observables are computed on the device (rt_observe): the mean interior height as a fraction of
the way from h_min to h_max, which the projection is matched to, and the mean |h - h_bar| over
the interior vertices as a fraction of the mean h_max - h_min, h_bar being the long-run mean
height field, which it is not.

Everything here drives the C ABI; the domain / height utilities are the reference-compatible
statics of librandomtilings.so.
"""
from __future__ import annotations

import sys

import numpy as np

import randomtilings_amd as rt
from randomtilings_amd import statics as S

INF = np.iinfo(np.int32).max
IN_CH = 4


# ---- domains ---------------------------------------------------------------------------

def aztec(n: int) -> np.ndarray:
    return S.call("domino.AztecDiamond", args=[n])


def notched_rectangle(rng, m: int) -> np.ndarray:
    """An m x m square (m even) with staircase notches of even depth cut from its left side:
    every row keeps an even length, so it is tileable."""
    d = S.call("domino.Rectangle", args=[m, m])
    side = int(round(np.sqrt(d.size)))
    d = d.reshape(side, side).copy()
    cut = 0
    for i in range(m):
        step = int(rng.integers(0, 3))
        cut = min(m - 2, cut + step) if i < m // 2 else max(0, cut - step)
        d[2 + i, 2:2 + 2 * (cut // 2)] = 0
    return d.ravel()


def domain_bank(rng, count: int, max_side: int, min_side: int = 0):
    """`count` random domains whose vertex arrays fit a max_side canvas (and, with min_side,
    are at least that large): Aztec diamonds and notched squares, alternately."""
    out = []
    for k in range(count):
        if k % 2 == 0:
            lo = max(4, max_side // 4, (min_side - 6) // 2)
            n = int(rng.integers(lo, max(lo, (max_side - 6) // 2) + 1))
            out.append(("aztec", aztec(n)))
        else:
            lo = max(4, max_side // 8, (min_side - 6) // 2)
            m = 2 * int(rng.integers(lo, max(lo + 1, (max_side - 6) // 2)))
            out.append(("notched", notched_rectangle(rng, m)))
    return out


def periodic_weights(rng, lo: float = 0.3, hi: float = 1.0):
    """p12 of the weighted rule for edge weights a, b drawn uniformly from [lo, hi]: class 0
    squares 1/2, class 1 squares a^2/(a^2+b^2) on even rows and b^2/(a^2+b^2) on odd rows."""
    a, b = rng.uniform(lo, hi, 2)
    p = a * a / (a * a + b * b)
    return [0.5, 0.5, float(p), float(1.0 - p)]


# ---- heights and features ----------------------------------------------------------------

class Domain:
    """A domain with its extremal tilings and height functions (built on the device) and the
    chain rule its chains run (p12 None: dominokernelCFTP.cl's uniform rule)."""

    def __init__(self, ctx: rt.Context, d: np.ndarray, p12=None):
        self.p12 = None if p12 is None else [float(x) for x in p12]
        self.rule = rt.RT_DOMINO_CFTP if p12 is None else rt.RT_DOMINO_WEIGHTED
        self.d = np.ascontiguousarray(d, np.int32)
        self.tmin, self.hmin = ctx.extremal_tiling(rt.RT_DOMINO, self.d, False)
        self.tmax, self.hmax = ctx.extremal_tiling(rt.RT_DOMINO, self.d, True)
        self.N = int(round(np.sqrt(self.tmin.size)))
        self.v = S.call("domino.DomainToVertices", self.d)
        self.inner = (self.v == 1) & (self.hmax != INF) & (self.hmax != self.hmin)
        self.fixed = (self.v == 2).astype(np.uint8)

    def heights(self, t: np.ndarray) -> np.ndarray:
        return S.call("domino.TilingToHeightfunc", t, self.d)

    def chain(self, ctx: rt.Context, t: np.ndarray) -> "rt.Chain":
        return ctx.chain(rt.RT_DOMINO, self.rule, t, self.p12)

    @property
    def contrast(self) -> float:
        """The weight pattern's contrast 2 p - 1 (0 for the uniform rule)."""
        return 0.0 if self.p12 is None else 2.0 * self.p12[2] - 1.0

    def observable(self, h: np.ndarray) -> float:
        """Mean interior height as a fraction of the way from h_min to h_max."""
        m = self.inner
        lo, hi = self.hmin[m].mean(), self.hmax[m].mean()
        return float((h[m].mean() - lo) / (hi - lo))

    def observer(self, ctx: rt.Context, ref=None) -> "rt.Observer":
        """Device observer over the interior vertices; ref: a real height field (the second
        observable measures |h - ref|)."""
        r256 = None if ref is None else np.where(self.hmin != INF, np.round(256.0 * ref), 0).astype(np.int32)
        return rt.Observer(ctx, self.d, self.hmin, self.inner.astype(np.uint8), r256)

    def observables(self, out: np.ndarray):
        """(mean-height fraction, mean |h - ref| / mean range) per row of rt_observe's output."""
        m = self.inner
        lo, hi = float(self.hmin[m].mean()), float(self.hmax[m].mean())
        cnt = out[:, 0].astype(np.float64)
        mean = out[:, 1] / cnt
        dev = out[:, 3] / (256.0 * cnt)
        return (mean - lo) / (hi - lo), dev / (hi - lo)


def features(dom: Domain, canvas: int) -> np.ndarray:
    """canvas x canvas x 4: h_min and h_max (scaled by 1/N), the interior mask and the weight
    pattern's contrast, sampled down by the smallest integer stride that fits the canvas."""
    N = dom.N
    s = max(1, -(-N // canvas))
    n = -(-N // s)
    x = np.zeros((canvas, canvas, IN_CH), np.float32)
    fin = (dom.hmax != INF).reshape(N, N)[::s, ::s]
    hmin = np.where(dom.hmax != INF, dom.hmin, 0).reshape(N, N)[::s, ::s]
    hmax = np.where(dom.hmax != INF, dom.hmax, 0).reshape(N, N)[::s, ::s]
    x[:n, :n, 0] = hmin / N
    x[:n, :n, 1] = hmax / N
    x[:n, :n, 2] = dom.inner.reshape(N, N)[::s, ::s]
    x[:n, :n, 3] = dom.contrast * fin
    return x


def target_fraction(dom: Domain, h: np.ndarray, canvas: int):
    """The fraction field of heights h and the interior mask, sampled down as features() samples
    the inputs."""
    N = dom.N
    s = max(1, -(-N // canvas))
    n = -(-N // s)
    t = np.zeros((canvas, canvas), np.float32)
    m = np.zeros((canvas, canvas), np.float32)
    r = np.zeros(N * N, np.float32)
    inn = dom.inner
    r[inn] = (h[inn] - dom.hmin[inn]) / (dom.hmax[inn] - dom.hmin[inn])
    t[:n, :n] = r.reshape(N, N)[::s, ::s]
    m[:n, :n] = inn.reshape(N, N)[::s, ::s]
    return t, m


# ---- data from the chain itself --------------------------------------------------------------

def equilibrium_samples(ctx: rt.Context, doms, steps: int, per_domain: int, seed: int):
    """Final tilings of per_domain chains per domain, each run `steps` steps under the domain's
    rule, alternately from the minimal and the maximal tiling (a chain not yet at equilibrium
    errs towards its start, so the two halves' errors cancel in an average), all advanced
    together in one batch."""
    chains, owner = [], []
    for i, dom in enumerate(doms):
        for k in range(per_domain):
            ch = dom.chain(ctx, dom.tmin if k % 2 == 0 else dom.tmax)
            ch.seed(seed + 1000 * i + k)
            chains.append(ch)
            owner.append(i)
    b = rt.Batch(ctx)
    for ch in chains:
        b.add(ch)
    chunk = 20000  # runs continue one another; a line per chunk shows progress on long runs
    for s0 in range(0, steps, chunk):
        b.run(min(chunk, steps - s0), True)
        b.sync()
        print(f"# equilibrium samples: {min(steps, s0 + chunk)}/{steps} steps, {len(chains)} chains",
              file=sys.stderr, flush=True)
    out = [(owner[k], ch.read()) for k, ch in enumerate(chains)]
    b.close()
    for ch in chains:
        ch.close()
    return out


def build_dataset(ctx, doms, samples, canvas, average: bool = False):
    """Inputs, targets and masks.  With average, one example per domain whose target is the
    mean fraction over that domain's samples: an estimate of the expected height with far less
    noise than any single sample, which is what the network is asked to predict."""
    X, T, M = [], [], []
    acc = {}
    for i, t in samples:
        dom = doms[i]
        tt, mm = target_fraction(dom, dom.heights(t), canvas)
        if average:
            s, n, _ = acc.get(i, (0.0, 0, mm))
            acc[i] = (s + tt, n + 1, mm)
        else:
            X.append(features(dom, canvas))
            T.append(tt)
            M.append(mm)
    for i, (s, n, mm) in sorted(acc.items()):
        X.append(features(doms[i], canvas))
        T.append((s / n).astype(np.float32))
        M.append(mm)
    return np.stack(X), np.stack(T), np.stack(M)


def train(net: rt.UNet, X, T, M, epochs: int, batch: int, lr: float, rng) -> list:
    losses = []
    n = len(X)
    for e in range(epochs):
        order = rng.permutation(n)
        for s in range(0, n - batch + 1, batch):
            idx = order[s:s + batch]
            losses.append(net.train_step(X[idx], T[idx], M[idx], lr))
        if (e + 1) % 25 == 0:
            print(f"# training: epoch {e + 1}/{epochs}, loss {losses[-1]:.4g}", file=sys.stderr, flush=True)
    return losses


# ---- warm start ------------------------------------------------------------------------------

def predicted_fraction(net: rt.UNet, dom: Domain, canvas: int) -> np.ndarray:
    """The network's fraction field on the domain's vertices (clipped to [0, 1]); a domain
    larger than the canvas is predicted on its sampled-down features and interpolated back
    (bilinearly between the sampled vertices)."""
    N = dom.N
    s = max(1, -(-N // canvas))
    n = -(-N // s)
    r = net.forward(features(dom, canvas)[None])[0, :n, :n].astype(np.float64)
    if s > 1:
        x = np.minimum(np.arange(N) / s, n - 1)
        i0 = np.minimum(np.floor(x).astype(int), n - 2)
        f = x - i0
        r = r[i0] * (1 - f)[:, None] + r[i0 + 1] * f[:, None]
        r = r[:, i0] * (1 - f)[None, :] + r[:, i0 + 1] * f[None, :]
    return np.clip(r.ravel(), 0.0, 1.0)


def project(ctx: rt.Context, dom: Domain, hp: np.ndarray) -> np.ndarray:
    """An admissible height function near the real-valued field hp: rounded to the residue
    class of domino heights (h = h_min mod 4), clamped between h_min and h_max, and closed on
    the device (the largest admissible function below it, which keeps the boundary)."""
    h = dom.hmin.astype(np.float64).copy()
    inn = dom.inner
    h[inn] = dom.hmin[inn] + 4 * np.round((hp[inn] - dom.hmin[inn]) / 4.0)
    h = np.clip(h, dom.hmin, dom.hmax).astype(np.int64)
    h = np.where(dom.hmax == INF, INF, h).astype(np.int32)
    closed, _ = ctx.height_closure(rt.RT_DOMINO, dom.d, h, 1, dom.fixed)
    return closed


def warm_heights(ctx: rt.Context, net, dom: Domain, canvas: int, rounds: int = 3, fraction=None) -> np.ndarray:
    """The network's prediction (or a given fraction field, e.g. the constant 1/2 of the
    midpoint start) made an admissible height function whose mean-height observable matches
    the field's: rounding and closure move the mean height, so the real-valued field is
    shifted by the difference and projected again (a few rounds)."""
    r = predicted_fraction(net, dom, canvas) if fraction is None else np.broadcast_to(
        np.asarray(fraction, np.float64), dom.hmin.shape)
    inn = dom.inner
    hp = dom.hmin.astype(np.float64).copy()
    hp[inn] = dom.hmin[inn] + r[inn] * (dom.hmax[inn] - dom.hmin[inn])
    lo, hi = dom.hmin[inn].mean(), dom.hmax[inn].mean()
    want = (hp[inn].mean() - lo) / (hi - lo)
    shift = 0.0
    best, best_err = None, np.inf
    for _ in range(max(1, rounds)):
        h = project(ctx, dom, hp + shift)
        err = dom.observable(h) - want
        if abs(err) < best_err:
            best, best_err = h, abs(err)
        shift -= err * (hi - lo)
    return best


def warm_tiling(ctx, net, dom: Domain, canvas: int, fraction=None) -> np.ndarray:
    return S.call("domino.HeightfuncToTiling", warm_heights(ctx, net, dom, canvas, fraction=fraction), dom.d)


def midpoint_tiling(ctx, dom: Domain) -> np.ndarray:
    """The non-learned start: (h_min + h_max) / 2, projected and mean-matched as the learned
    prediction is."""
    return warm_tiling(ctx, None, dom, 0, fraction=0.5)


def trajectories(ctx: rt.Context, dom: Domain, starts, steps: int, every: int, seed: int, chains: int = 16,
                 ref=None, mean_from: int = -1, settle=None, progress=None):
    """Observables of `chains` independent chains from each start tiling, all in one batch and
    read on the device every `every` chain steps (index 0: the starts).  Returns (per start:
    (mean-height fractions, mean |h - ref| fractions or None) averaged over its chains; the mean
    height field over every chain and reading from step mean_from on, or None; the steps run).
    Chain k of every start has the same seed, so the starts share their random numbers, and a
    repeated call repeats every trajectory bit for bit.

    settle = (i, j, tol): run until the mean-height fractions of starts i and j (5-reading
    running means) first agree within tol, at step T, then as many steps again, at most `steps`
    in all; the mean field is then taken from T on.  Under a monotone rule with shared random
    numbers, chains from the minimal and the maximal tiling bracket the chains from any other
    start, so once they agree every start has forgotten where it began.

    progress(step, total): called after a reading when 20 s have passed since the last call."""
    import time
    chs, owner = [], []
    for i, t0 in enumerate(starts):
        for k in range(chains):
            ch = dom.chain(ctx, t0)
            ch.seed(seed + 7919 * k)
            chs.append(ch)
            owner.append(i)
    owner = np.array(owner)
    b = rt.Batch(ctx)
    for ch in chs:
        b.add(ch)
    obs = dom.observer(ctx, ref)
    o1 = [[] for _ in starts]
    o2 = [[] for _ in starts]
    acc, nacc = None, 0
    total = (steps // every) * every
    # the mean field starts at mean_from; when settling, at the settling step, or at half the
    # cap if the brackets never agree
    mean_start = mean_from if settle is None else total // 2
    settled = False

    def read(step):
        nonlocal acc, nacc
        want = 0 <= mean_start <= step
        out, hs = obs.observe(chs, heights=want)
        a, d = dom.observables(out)
        for i in range(len(starts)):
            o1[i].append(float(a[owner == i].mean()))
            o2[i].append(float(d[owner == i].mean()))
        if want:
            s = hs.astype(np.float64).sum(axis=0)
            acc = s if acc is None else acc + s
            nacc += len(chs)

    read(0)
    step = 0
    said = time.monotonic()
    while step < total:
        b.run(every, True)
        step += every
        read(step)
        if progress is not None and time.monotonic() - said >= 20.0:
            progress(step, total)
            said = time.monotonic()
        if settle is not None and not settled:
            i, j, tol = settle
            k = min(5, len(o1[i]))
            if abs(np.mean(o1[i][-k:]) - np.mean(o1[j][-k:])) <= tol:
                settled = True
                if step < mean_start:
                    mean_start = step
                total = min(total, 2 * step)
    b.close()
    obs.close()
    for ch in chs:
        ch.close()
    mean = None if acc is None else np.where(dom.hmin != INF, acc / max(nacc, 1), 0.0)
    res = [(np.array(o1[i]), np.array(o2[i]) if ref is not None else None) for i in range(len(starts))]
    return res, mean, step


def trajectory(ctx: rt.Context, dom: Domain, t0: np.ndarray, steps: int, every: int, seed: int,
               chains: int = 16) -> np.ndarray:
    """The mean-height observable averaged over `chains` chains from t0, every `every` steps."""
    return trajectories(ctx, dom, [t0], steps, every, seed, chains)[0][0][0]


def compare_starts(ctx: rt.Context, dom: Domain, starts: dict, steps: int, every: int, seed: int,
                   chains: int = 16, tol: float = 0.01, progress=None) -> dict:
    """Chain steps each start needs to reach equilibrium on both observables.  The chains from
    the minimal and the maximal tiling run beside the starts as brackets; a first pass runs
    until the brackets' mean heights agree within tol / 4 (at most `steps`), then as long
    again, and takes the long-run mean height field h_bar over that second half; a second pass
    repeats the same trajectories (same seeds and length, bit for bit) measuring |h - h_bar|.
    Each observable's equilibrium value is its average over the second half of every
    trajectory, brackets included.  progress(pass, step, total) reports long passes."""
    names = list(starts)
    ts = [starts[n] for n in names] + [dom.tmin, dom.tmax]
    lo_i, hi_i = len(names), len(names) + 1
    def pr(k):
        return None if progress is None else (lambda step, total: progress(k, step, total))

    res1, hbar, run = trajectories(ctx, dom, ts, steps, every, seed, chains, settle=(lo_i, hi_i, tol / 4),
                                   progress=pr(1))
    res2, _, _ = trajectories(ctx, dom, ts, run, every, seed, chains, ref=hbar, progress=pr(2))
    o1 = [r[0] for r in res1]
    o2 = [r[1] for r in res2]
    half = len(o1[0]) // 2
    eq1 = float(np.mean([x[half:].mean() for x in o1]))
    eq2 = float(np.mean([x[half:].mean() for x in o2]))
    gap = float(abs(o1[lo_i][-1] - o1[hi_i][-1]))
    out = {"eq_height_fraction": eq1, "eq_deviation": eq2, "steps_run": run, "bracket_gap_at_end": gap,
           "settled": bool(run < (steps // every) * every or gap <= tol / 4)}
    for i, n in enumerate(names):
        out[n] = {"steps_height": steps_to_equilibrium(o1[i], eq1, every, tol),
                  "steps_deviation": steps_to_equilibrium(o2[i], eq2, every, tol),
                  "start_height_fraction": float(o1[i][0]), "start_deviation": float(o2[i][0])}
    return out


def steps_to_equilibrium(obs: np.ndarray, eq: float, every: int, tol: float = 0.01) -> int:
    """Steps until the observable first comes within `tol` of its equilibrium value (the
    observable is already a fraction of the max - min range, so `tol` is relative to that
    range), read from a running mean over 5 samples; the full length when it never does."""
    k = min(5, len(obs))
    run = np.convolve(obs, np.ones(k) / k, mode="valid")
    ok = np.flatnonzero(np.abs(run - eq) <= tol)
    first = len(run) if ok.size == 0 else int(ok[0])
    return first * every
