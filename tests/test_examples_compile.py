"""The reference's example programs compile and link, unchanged, against this project's
drop-in headers (randomtilings_amd/src, at the reference's header paths) and libraries.

Each example includes its headers as "../src/<Model>/<Model>Tiler.h" relative to its own
directory, so the test places a byte-for-byte copy of the example in a temporary
Examples/ folder next to a src/ link to randomtilings_amd/src; nothing else changes.  The
programs are only built (they read a tinymtdc parameter file the reference does not ship,
and would need a GPU): the C++ GPU test runs the same API calls.  Skipped where the
reference sources are absent; nothing from them is kept.
"""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
REF = Path("/root/reference")
LIB = ROOT / "randomtilings_amd" / "lib"
EXAMPLES = ["MinimalDominoEx.cpp", "MinimalLozengeEx.cpp", "RectTriangleEx.cpp", "TriangleDimerEx.cpp",
            "AztecDiamondCftpEx.cpp"]

pytestmark = pytest.mark.skipif(not (REF / "Examples").is_dir(), reason="reference sources absent")


@pytest.mark.parametrize("name", EXAMPLES)
def test_reference_example_builds_unchanged(tmp_path, name):
    if not (LIB / "librandomtilings.so").exists():
        pytest.skip("librandomtilings.so not built")
    ex = tmp_path / "Examples"
    ex.mkdir()
    shutil.copyfile(REF / "Examples" / name, ex / name)
    os.symlink(ROOT / "randomtilings_amd" / "src", tmp_path / "src")
    exe = tmp_path / name.replace(".cpp", "")
    r = subprocess.run(["g++", "-std=gnu++11", "-O0", "-I", str(ROOT / "include"), str(ex / name), "-L", str(LIB),
                        "-lrandomtilings", "-lrt_hip", f"-Wl,-rpath,{LIB}", "-o", str(exe)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    assert (REF / "Examples" / name).read_bytes() == (ex / name).read_bytes()
    assert exe.exists()
