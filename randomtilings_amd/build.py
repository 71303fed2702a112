"""Build the native parts of randomtilings_amd in-tree.

* ``lib/librt_hip.so`` — the gfx950 HIP kernels and the C ABI (include/rt_hip.h).
* ``lib/librandomtilings.so`` — the reference-compatible C++ classes (src/), linked
  against librt_hip.so.

Every call recompiles (the build check must exercise the compiler).  Each translation
unit compiles in parallel.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
SRC = PKG / "src"
LIB = PKG / "lib"
OBJ = PKG / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # no packed FP32 instructions: some of their forms have given wrong results on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", str(ROOT / "include"),
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-I", str(ROOT / "include"), "-I", str(SRC)]

HIP_UNITS = ["rt_kernels.hip", "rt_tridimer.hip", "rt_heights.hip", "rt_mfma.hip", "rt_nnops.hip", "rt_capi.cpp",
             "rt_rules.cpp", "rt_unet.cpp", "rt_skip.cpp", "rt_noalloc.hip", "rt_observe.hip"]
CXX_UNITS = [
    "common/common.cpp", "TinyMT/file_reader.cpp", "Domino/DominoTiler.cpp",
    "Lozenge/LozengeTiler.cpp", "RectTriangle/RectTriangleTiler.cpp",
    "TriangleDimer/TriangleDimerTiler.cpp", "statics_capi.cpp",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"compile failed: {cmd[-1] if cmd else ''}")


def _hip_obj(unit: str) -> Path:
    src = CSRC / unit
    out = OBJ / (Path(unit).stem + ".o")
    lang = ["-x", "hip"] if unit.endswith(".hip") else []
    _run([HIPCC, *HIP_FLAGS, *lang, "-c", str(src), "-o", str(out)])
    return out


def _cxx_obj(unit: str) -> Path:
    src = SRC / unit
    out = OBJ / ("api_" + Path(unit).stem + ".o")
    have = []
    if (SRC / "RectTriangle" / "RectTriangleTiler.cpp").exists():
        have.append("-DRT_HAVE_RECTTRI")
    if (SRC / "TriangleDimer" / "TriangleDimerTiler.cpp").exists():
        have.append("-DRT_HAVE_TRIDIMER")
    _run(["g++", *HOST_FLAGS, *have, "-c", str(src), "-o", str(out)])
    return out


def build_native(jobs: int = 8) -> Path:
    OBJ.mkdir(exist_ok=True)
    LIB.mkdir(exist_ok=True)
    units = [u for u in HIP_UNITS if (CSRC / u).exists()]
    cxx_units = [u for u in CXX_UNITS if (SRC / u).exists()]
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        hip_objs = list(ex.map(_hip_obj, units))
        cxx_objs = list(ex.map(_cxx_obj, cxx_units))
    so = LIB / "librt_hip.so"
    tmp = LIB / "librt_hip.so.tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, hip_objs)])
    os.replace(tmp, so)
    if cxx_objs:
        api = LIB / "librandomtilings.so"
        tmp = LIB / "librandomtilings.so.tmp"
        _run(["g++", "-shared", "-o", str(tmp), *map(str, cxx_objs), "-L", str(LIB), "-lrt_hip",
              f"-Wl,-rpath,$ORIGIN"])
        os.replace(tmp, api)
        build_cpp_caller()
    return so


def build_cpp_caller() -> None:
    """The C++ caller program the GPU tests run against librandomtilings.so (tests/cpp): it
    uses the reference-compatible classes the way the reference's example programs do."""
    src = ROOT / "tests" / "cpp" / "tiler_walks.cpp"
    if not src.exists():
        return
    exe = src.with_suffix("")
    _run(["g++", "-O2", "-std=gnu++17", str(src), "-L", str(LIB), "-lrandomtilings", "-lrt_hip",
          "-Wl,-rpath,$ORIGIN/../../randomtilings_amd/lib", "-o", str(exe)])


def build_oracle() -> None:
    _run(["make", "-B", "-C", str(ROOT / "oracle")])


if __name__ == "__main__":
    build_native()
    build_oracle()
    print("built", LIB)
