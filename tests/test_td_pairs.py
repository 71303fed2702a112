"""The triangle-dimer kernels run their all-site update phases on pairs of sites packed into
32-bit words (randomtilings_amd/csrc/rt_tdsite.hpp).  This CPU test builds tests/cpp/
td_pairs_check.cpp as a host program and checks that every pair form writes, site for site,
what the per-site form of the same reference kernel (triangledimerkernel.cl, cited per
function) writes, on random planes of arbitrary 16-bit states, every move variant and slice.
The per-site forms are the ones the GPU parity tests pin to the oracle."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


def test_pair_phases_equal_site_phases(tmp_path):
    if not Path(HIPCC).exists():
        pytest.skip("hipcc not available")
    exe = tmp_path / "td_pairs_check"
    r = subprocess.run([HIPCC, "-O2", "-std=c++17", "-x", "hip", "--offload-arch=gfx950",
                        "-I", str(ROOT / "include"), "-I", str(ROOT / "randomtilings_amd" / "csrc"),
                        str(ROOT / "tests" / "cpp" / "td_pairs_check.cpp"), "-o", str(exe)],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-4000:]
    assert r.stdout.strip().endswith("fails 0")


def test_pair_check_catches_a_wrong_digit(tmp_path):
    """The same check on a copy of the header with one bit test of tri_digit_p inverted fails."""
    if not Path(HIPCC).exists():
        pytest.skip("hipcc not available")
    src = (ROOT / "randomtilings_amd" / "csrc" / "rt_tdsite.hpp").read_text()
    bad = src.replace("e3 = bitp(v, 3) & ~bitp(v, 2)", "e3 = bitp(v, 3) & bitp(v, 2)")
    assert bad != src
    (tmp_path / "rt_tdsite.hpp").write_text(bad)
    shutil.copyfile(ROOT / "randomtilings_amd" / "csrc" / "rt_tridimer.hpp", tmp_path / "rt_tridimer.hpp")
    exe = tmp_path / "td_pairs_bad"
    r = subprocess.run([HIPCC, "-O2", "-std=c++17", "-x", "hip", "--offload-arch=gfx950", "-DNTRIALS=50",
                        "-I", str(tmp_path), "-I", str(ROOT / "include"),
                        str(ROOT / "tests" / "cpp" / "td_pairs_check.cpp"), "-o", str(exe)],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    assert r.returncode != 0
