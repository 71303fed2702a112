"""Learned warm start for domino chains.

A U-Net (rt_unet, trained on the GPU by this library's kernels) maps a domain — its
minimal and maximal height functions, its vertex mask and its weight field — to the
limit-shape height, expressed as the fraction r = (h - h_min) / (h_max - h_min).  The
prediction is turned into a start state exactly: rounded to the residue class of domino
heights (h = h_min mod 4), clamped between h_min and h_max, and closed on the device
(rt_height_closure: the largest admissible height function below it, which keeps the
boundary since it stays above h_min).  The chain then runs from that tiling; the plain
chain from the minimal tiling remains the fallback and the yardstick.

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


def domain_bank(rng, count: int, max_side: int):
    """`count` random domains whose vertex arrays fit a max_side canvas: Aztec diamonds and
    notched squares, alternately."""
    out = []
    for k in range(count):
        if k % 2 == 0:
            n = int(rng.integers(max(4, max_side // 4), (max_side - 6) // 2 + 1))
            out.append(("aztec", aztec(n)))
        else:
            m = 2 * int(rng.integers(max(4, max_side // 8), (max_side - 6) // 2))
            out.append(("notched", notched_rectangle(rng, m)))
    return out


# ---- heights and features ----------------------------------------------------------------

class Domain:
    """A domain with its extremal tilings and height functions (built on the device)."""

    def __init__(self, ctx: rt.Context, d: np.ndarray):
        self.d = np.ascontiguousarray(d, np.int32)
        self.tmin, self.hmin = ctx.extremal_tiling(rt.RT_DOMINO, self.d, False)
        self.tmax, self.hmax = ctx.extremal_tiling(rt.RT_DOMINO, self.d, True)
        self.N = int(round(np.sqrt(self.tmin.size)))
        self.v = S.call("domino.DomainToVertices", self.d)
        self.inner = (self.v == 1) & (self.hmax != INF) & (self.hmax != self.hmin)
        self.fixed = (self.v == 2).astype(np.uint8)

    def heights(self, t: np.ndarray) -> np.ndarray:
        return S.call("domino.TilingToHeightfunc", t, self.d)

    def observable(self, h: np.ndarray) -> float:
        """Mean interior height as a fraction of the way from h_min to h_max."""
        m = self.inner
        lo, hi = self.hmin[m].mean(), self.hmax[m].mean()
        return float((h[m].mean() - lo) / (hi - lo))


def features(dom: Domain, canvas: int, weight: float = 0.5) -> np.ndarray:
    """canvas x canvas x 4: h_min and h_max (scaled by 1/N), the interior mask, the weight."""
    N = dom.N
    x = np.zeros((canvas, canvas, IN_CH), np.float32)
    fin = dom.hmax != INF
    hmin = np.where(fin, dom.hmin, 0).reshape(N, N)
    hmax = np.where(fin, dom.hmax, 0).reshape(N, N)
    x[:N, :N, 0] = hmin / N
    x[:N, :N, 1] = hmax / N
    x[:N, :N, 2] = dom.inner.reshape(N, N)
    x[:N, :N, 3] = weight * fin.reshape(N, N)
    return x


def target_fraction(dom: Domain, h: np.ndarray, canvas: int):
    N = dom.N
    t = np.zeros((canvas, canvas), np.float32)
    m = np.zeros((canvas, canvas), np.float32)
    r = np.zeros(N * N, np.float32)
    inn = dom.inner
    r[inn] = (h[inn] - dom.hmin[inn]) / (dom.hmax[inn] - dom.hmin[inn])
    t[:N, :N] = r.reshape(N, N)
    m[:N, :N] = inn.reshape(N, N)
    return t, m


# ---- data from the chain itself --------------------------------------------------------------

def equilibrium_samples(ctx: rt.Context, doms, steps: int, per_domain: int, seed: int):
    """Final tilings of per_domain chains per domain, each run `steps` steps from the minimal
    tiling (dominokernelCFTP.cl's rule), all advanced together in one batch."""
    chains, owner = [], []
    for i, dom in enumerate(doms):
        for k in range(per_domain):
            ch = ctx.chain(rt.RT_DOMINO, rt.RT_DOMINO_CFTP, dom.tmin)
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
    """The network's fraction field on the domain's vertices (clipped to [0, 1])."""
    N = dom.N
    return np.clip(net.forward(features(dom, canvas)[None])[0, :N, :N].ravel(), 0.0, 1.0)


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


def warm_heights(ctx: rt.Context, net: rt.UNet, dom: Domain, canvas: int, rounds: int = 3) -> np.ndarray:
    """The network's prediction made an admissible height function whose observable matches
    the prediction's: rounding and closure move the mean height, so the real-valued field is
    shifted by the difference and projected again (a few rounds)."""
    r = predicted_fraction(net, dom, canvas)
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


def warm_tiling(ctx, net, dom: Domain, canvas: int) -> np.ndarray:
    return S.call("domino.HeightfuncToTiling", warm_heights(ctx, net, dom, canvas), dom.d)


def trajectory(ctx: rt.Context, dom: Domain, t0: np.ndarray, steps: int, every: int, seed: int,
               chains: int = 16) -> np.ndarray:
    """The observable averaged over `chains` independent chains from t0 (one batch), every
    `every` chain steps (index 0: the start)."""
    chs = [ctx.chain(rt.RT_DOMINO, rt.RT_DOMINO_CFTP, t0) for _ in range(chains)]
    b = rt.Batch(ctx)
    for k, ch in enumerate(chs):
        ch.seed(seed + 7919 * k)
        b.add(ch)
    obs = [dom.observable(dom.heights(t0))]
    for _ in range(steps // every):
        b.run(every, True)
        b.sync()
        obs.append(float(np.mean([dom.observable(dom.heights(ch.read())) for ch in chs])))
    b.close()
    for ch in chs:
        ch.close()
    return np.array(obs)


def steps_to_equilibrium(obs: np.ndarray, eq: float, every: int, tol: float = 0.01) -> int:
    """Steps until the observable first comes within `tol` of its equilibrium value (the
    observable is already a fraction of the max - min range, so `tol` is relative to that
    range), read from a running mean over 5 samples; the full length when it never does."""
    k = min(5, len(obs))
    run = np.convolve(obs, np.ones(k) / k, mode="valid")
    ok = np.flatnonzero(np.abs(run - eq) <= tol)
    first = len(run) if ok.size == 0 else int(ok[0])
    return first * every
