"""Write the golden fixtures under tests/golden/ from the reference's own host statics.

Run in a container that has the reference sources (``make -C oracle`` builds
oracle/_ref/libref.so from them).  The fixtures are data: domains and tilings the
reference's constructors produce (``DominoTiler::AztecDiamond``, ``MaxTiling`` ...), used
as start states for the chains and to pin this project's C++ restatement of the statics.
The reference's example outputs (Examples/ExampleOuts/*.txt) are copied verbatim as data.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import shutil
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
import oracle as O  # noqa: E402

OUT = Path(__file__).resolve().parent
REF = Path("/root/reference")


def main() -> None:
    arrays: dict[str, np.ndarray] = {}

    def put(key, a):
        arrays[key] = np.asarray(a, dtype=np.int32)

    # domino: Aztec diamonds, rectangles, Aztec rectangles
    for n in [1, 2, 3, 4, 5, 8, 16, 32, 64, 128]:
        d = O.ref_call("domino.AztecDiamond", args=[n])
        put(f"domino_aztec{n}_domain", d)
        put(f"domino_aztec{n}_max", O.ref_call("domino.MaxTiling", d))
        put(f"domino_aztec{n}_min", O.ref_call("domino.MinTiling", d))
        if n <= 16:
            t = arrays[f"domino_aztec{n}_max"]
            put(f"domino_aztec{n}_vertices", O.ref_call("domino.DomainToVertices", d))
            put(f"domino_aztec{n}_hmax", O.ref_call("domino.TilingToHeightfunc", t, d))
            put(f"domino_aztec{n}_hmin", O.ref_call("domino.TilingToHeightfunc", arrays[f"domino_aztec{n}_min"], d))
            put(f"domino_aztec{n}_t2d", O.ref_call("domino.TilingToDomain", t))
    for (m, n) in [(2, 2), (4, 6), (7, 4), (10, 10)]:
        d = O.ref_call("domino.Rectangle", args=[m, n])
        put(f"domino_rect{m}x{n}_domain", d)
        put(f"domino_rect{m}x{n}_max", O.ref_call("domino.MaxTiling", d))
        put(f"domino_rect{m}x{n}_min", O.ref_call("domino.MinTiling", d))
    for (n, rem) in [(4, [1]), (6, [0, 3]), (8, [1, 2, 5])]:
        d = O.ref_call("domino.AztecRectangle", rem, args=[n])
        put(f"domino_azrect{n}_{len(rem)}_domain", d)
        put(f"domino_azrect{n}_{len(rem)}_remove", rem)
        put(f"domino_azrect{n}_{len(rem)}_max", O.ref_call("domino.MaxTiling", d))
        put(f"domino_azrect{n}_{len(rem)}_min", O.ref_call("domino.MinTiling", d))

    # lozenge: almost-hexagons
    for (n, m) in [(1, 0), (2, 0), (3, 0), (5, 1), (10, 2), (20, 0), (50, 0), (100, 0)]:
        d = O.ref_call("lozenge.AlmostHexDomain", args=[n, m])
        put(f"lozenge_hex{n}_{m}_domain", d)
        put(f"lozenge_hex{n}_{m}_max", O.ref_call("lozenge.MaxTiling", d))
        put(f"lozenge_hex{n}_{m}_min", O.ref_call("lozenge.MinTiling", d))
        if n <= 10:
            t = arrays[f"lozenge_hex{n}_{m}_max"]
            put(f"lozenge_hex{n}_{m}_vertices", O.ref_call("lozenge.DomainToVertices", d))
            put(f"lozenge_hex{n}_{m}_hmax", O.ref_call("lozenge.TilingToHeightfunc", t, d))
            put(f"lozenge_hex{n}_{m}_t2d", O.ref_call("lozenge.TilingToDomain", t))

    # rectangle-triangle hexagons
    for n in [1, 2, 3, 5, 10, 20]:
        for kind in ["maxHex", "minHex", "slopeHex"]:
            put(f"recttri_{kind}{n}", O.ref_call(f"recttri.{kind}", args=[n]))

    # triangle-dimer
    for (n, m) in [(2, 0), (3, 0), (6, 0), (9, 2), (10, 0), (14, 0), (22, 0)]:
        t = O.ref_call("tridimer.IceCreamCone", args=[n, m])
        put(f"tridimer_cone{n}_{m}", t)
        put(f"tridimer_cone{n}_{m}_tri", O.ref_call("tridimer.TilingToTriangleFlips", t))
        put(f"tridimer_cone{n}_{m}_bfly", O.ref_call("tridimer.TilingToButterflyFlips", t))
        put(f"tridimer_cone{n}_{m}_domain", O.ref_call("tridimer.TilingToDomain", t))
    for k in [0, 1, 2]:
        put(f"tridimer_butterfly{k}", O.ref_call("tridimer.TestButterfly", args=[k]))

    np.savez_compressed(OUT / "statics.npz", **arrays)
    print(f"wrote {len(arrays)} arrays to {OUT / 'statics.npz'}")

    # the reference's deterministic example outputs, as data
    ex = REF / "Examples" / "ExampleOuts"
    (OUT / "exampleouts").mkdir(exist_ok=True)
    for rel in ["RectTriangle/MaxTiling.txt", "RectTriangle/MinTiling.txt", "RectTriangle/SlopeTiling.txt",
                "TriangleDimer/DimerStart.txt", "TriangleDimer/Domain.txt"]:
        src = ex / rel
        if src.exists():
            shutil.copy(src, OUT / "exampleouts" / rel.replace("/", "_"))


if __name__ == "__main__":
    main()
