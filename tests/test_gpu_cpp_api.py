"""The reference-compatible C++ API end to end: tests/cpp/tiler_walks (built by build(),
linked against librandomtilings.so) constructs DominoTiler, LozengeTiler, RectTriangleTiler
and TriangleDimerTiler the way the reference's example programs do, loads a tinymtdc-format
parameter file with LoadTinyMT(path, n), and runs Walk (single seed and epochs), the
CFTP driver loop of Examples/AztecDiamondCftpEx.cpp and DominoTiler::CFTP on the GPU.  Every
tiling it writes must equal the oracle's bit for bit under the same parameter table.

Before the program starts, this process runs a persistent walk, so it holds the device's
cooperative queue: the program's resident grids may then be refused a cooperative launch and
must continue, with the same result, on the blocked path."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle as O
import randomtilings_amd as rt
from randomtilings_amd import statics as S
from randomtilings_amd.params import scrambled_table, write_tinymtdc

pytestmark = pytest.mark.gpu

EXE = Path(__file__).parent / "cpp" / "tiler_walks"


def test_cpp_tiler_walks_match_oracle(tmp_path, golden):
    assert EXE.exists(), f"{EXE} not built: run build()"
    table = scrambled_table(8192, 5)
    pfile = tmp_path / "params.txt"
    write_tinymtdc(pfile, table)
    assert np.array_equal(rt.parse_params_file(pfile, 8192), table)

    with rt.Context(0) as c:  # hold a cooperative queue in this process
        c.load_params(table)
        ch = c.chain(rt.RT_LOZENGE, rt.RT_LOZENGE_QREF, golden["lozenge_hex20_0_min"], [0.9])
        ch.set_path(rt.RT_PATH_PERSISTENT, 0)
        ch.seed(1)
        ch.run(100)
        ch.sync()
        print("cpp api: parent cooperative refusals", ch.coop_refusals, flush=True)
        print("cpp api: tiler_walks", flush=True)
        r = subprocess.run([str(EXE), str(tmp_path), str(pfile)], capture_output=True, text=True, timeout=300)
        ch.close()
    assert r.returncode == 0, r.stdout + r.stderr
    assert "walked on " in r.stdout
    print(r.stdout, flush=True)

    def load(name):
        return np.loadtxt(tmp_path / name, dtype=np.int64)

    t0 = S.call("domino.MinTiling", S.call("domino.AztecDiamond", args=[40]))
    assert np.array_equal(load("domino.txt"), O.domino_walk(t0, 3001, 20261203, table, rule=1))

    te = S.call("domino.MaxTiling", S.call("domino.AztecDiamond", args=[12]))
    o = O.domino_walk_epochs(te, [256, 128, 77], [345903, 17, 1099511627781], table, rule=1)
    assert np.array_equal(load("domino_epochs.txt"), o)

    d5 = S.call("domino.AztecDiamond", args=[5])
    sample, total, restarts = O.domino_cftp(S.call("domino.MaxTiling", d5), S.call("domino.MinTiling", d5), 345903,
                                            table, rule=1, mode=rt.RT_CFTP_REFERENCE, first_steps=128)
    assert np.array_equal(load("cftp_loop.txt"), sample)
    assert np.array_equal(load("cftp_call.txt"), sample)
    assert f"cftp restarts {restarts} total {total}" in r.stdout

    u0 = S.call("lozenge.MaxTiling", S.call("lozenge.AlmostHexDomain", args=[30, 0]))
    assert np.array_equal(load("lozenge.txt"), O.lozenge_walk(u0, 2002, 7, table, rule=1, q=0.9))

    r0 = S.call("recttri.minHex", args=[12])
    assert np.array_equal(load("recttri.txt"), O.recttri_walk(r0, 1500, 31, table))

    c0 = S.call("tridimer.IceCreamCone", args=[14, 0])
    assert np.array_equal(load("tridimer.txt"), O.tridimer_walk(c0, 500, 4242, table))
