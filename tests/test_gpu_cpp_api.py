"""The reference-compatible C++ API end to end: tests/cpp/tiler_walks (built by build(),
linked against librandomtilings.so) constructs DominoTiler and LozengeTiler the way the
reference's example programs do, walks an Aztec diamond and a hexagon on the GPU, and writes
the tilings; they must equal the oracle's walks bit for bit."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from randomtilings_amd import statics as S
from randomtilings_amd.params import default_table

pytestmark = pytest.mark.gpu

EXE = Path(__file__).parent / "cpp" / "tiler_walks"


def test_cpp_tiler_walks_match_oracle(tmp_path):
    assert EXE.exists(), f"{EXE} not built: run build()"
    print("cpp api: tiler_walks", flush=True)
    r = subprocess.run([str(EXE), str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("walked on ")

    t0 = S.call("domino.MinTiling", S.call("domino.AztecDiamond", args=[40]))
    N = int(round(np.sqrt(t0.size)))
    o = O.domino_walk(t0, 3001, 20261203, default_table(N * N // 2), rule=1)
    g = np.loadtxt(tmp_path / "domino.txt", dtype=np.int64)
    assert np.array_equal(g, o)

    u0 = S.call("lozenge.MaxTiling", S.call("lozenge.AlmostHexDomain", args=[30, 0]))
    N = int(round(np.sqrt(u0.size)))
    o = O.lozenge_walk(u0, 2002, 7, default_table(N * (N // 3)), rule=1, q=0.9)
    g = np.loadtxt(tmp_path / "lozenge.txt", dtype=np.int64)
    assert np.array_equal(g, o)
