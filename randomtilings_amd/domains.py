"""Domain mixes for batched runs: Aztec diamonds, domino rectangles with staircase notches
(tileability decided by the device maximal/minimal construction), triangle-dimer hexagons
(covered by TriangleDimerTiler::Hexagon, sides chosen so a dimer cover exists) and
rectangle-triangle hexagons, drawn from a seed (lozenge hexagons on request).

The reference builds one domain per tiler and loops over them in the caller
(src/Domino/DominoTiler.cpp:118-193, src/TriangleDimer/TriangleDimerTiler.cpp:485-581,
src/RectTriangle/RectTriangleTiler.cpp:94-200); these helpers produce the start tilings of
such a mix so that many chains can run as one batch.
"""
from __future__ import annotations

import numpy as np

from . import statics

KINDS = ("aztec", "notched", "tridimer", "recttri")


def notched_rectangle(ctx, side: int, rng: np.random.Generator, lo: int = 8):
    """A square domino domain of about `side` tiling positions whose left edge is cut by a
    random staircase (cuts of 0-2 squares per row, growing then shrinking, in pairs of
    columns); returns its minimal tiling, built on the device.  A staircase whose region has
    no tiling (the device construction says RT_ERR_UNTILEABLE) is redrawn one size smaller."""
    import randomtilings_amd as rt

    while True:
        m = side - 6
        d = statics.call("domino.Rectangle", args=[m, m])
        sd = int(round(np.sqrt(d.size)))
        d = d.reshape(sd, sd).copy()
        cut = 0
        for i in range(m):
            step = int(rng.integers(0, 3))
            cut = min(m - 1, cut + step) if i < m // 2 else max(0, cut - step)
            d[2 + i, 2:2 + 2 * (cut // 2)] = 0
        try:
            return ctx.extremal_tiling(rt.RT_DOMINO, d.ravel(), False)[0]
        except rt.RTError as e:
            if e.code != rt.RT_ERR_UNTILEABLE:
                raise
            side = max(lo, side - 1)


def hexagon_vertices(x: int, y: int, z: int) -> int:
    """Vertices of the triangle-dimer hexagon with sides x, y, z (TriangleDimerTiler::Hexagon):
    a top row of x + 1, y rows growing by one, z rows shrinking by one."""
    return sum(x + 1 + min(r, y) - max(0, r - y) for r in range(y + z + 1))


def tridimer_hexagon_sides(side: int):
    """Sides (x, y, z) of a nearly regular triangle-dimer hexagon whose tiling side is about
    `side` (4 m + 12 for sides near m) and whose vertex count is even (a regular hexagon's,
    3 m^2 + 3 m + 1, never is)."""
    m = max(2, (side - 12) // 4)
    for xyz in ((m + 1, m, m), (m, m, m - 1), (m, m, m + 1), (m + 1, m, m + 1), (m, m + 1, m)):
        if hexagon_vertices(*xyz) % 2 == 0:
            return xyz
    raise ValueError(side)


def domain(ctx, kind: str, side: int, rng: np.random.Generator):
    """(model, rule, start tiling) of one domain of the mix with a tiling side near `side`."""
    import randomtilings_amd as rt

    if kind == "aztec":
        d = statics.call("domino.AztecDiamond", args=[max(1, side // 2 - 3)])
        return rt.RT_DOMINO, rt.RT_DOMINO_CFTP, ctx.extremal_tiling(rt.RT_DOMINO, d, False)[0]
    if kind == "notched":
        return rt.RT_DOMINO, rt.RT_DOMINO_CFTP, notched_rectangle(ctx, side, rng)
    if kind == "tridimer":
        return rt.RT_TRIDIMER, 0, statics.call("tridimer.Hexagon", args=list(tridimer_hexagon_sides(side)))
    if kind == "recttri":
        return rt.RT_RECTTRI, 0, statics.call("recttri.maxHex", args=[max(1, (side - 7) // 2)])
    if kind == "lozenge":  # not in the default mix: lozenge batches on their own
        d = statics.call("lozenge.AlmostHexDomain", args=[max(1, (side - 7) // 2), 0])
        return rt.RT_LOZENGE, rt.RT_LOZENGE_UNIFORM, statics.call("lozenge.MinTiling", d)
    raise ValueError(kind)


def mix(ctx, count: int, sides, seed: int, kinds=KINDS):
    """`count` domains, kinds cycling through `kinds`, tiling sides uniform in [lo, hi] drawn
    from `seed`: a list of (model, rule, tiling, kind)."""
    rng = np.random.default_rng(seed)
    lo, hi = sides
    out = []
    for k in range(count):
        kind = kinds[k % len(kinds)]
        side = int(rng.integers(lo, hi + 1))
        m, r, t = domain(ctx, kind, side, rng)
        out.append((m, r, t, kind))
    return out
