"""TinyMT32 parameter tables in the tinymtdc format.

The reference loads per-generator (mat1, mat2, tmat) triples from
``tinymt32dc.0.1048576.txt`` (a ``jj1bdx/tinymtdc-longbatch`` output; not shipped with
the reference and holding fewer sets than large chains need).  These helpers build tables
of any size:

* ``default_table(n)`` — every generator uses TinyMT32's published default parameter set
  (src/TinyMT/tinymt32def.clh:48-50) and differs only by its seed ``seed + id``.
* ``scrambled_table(n, seed)`` — pseudo-random triples.  They are NOT certified TinyMT
  parameter sets (no period guarantee); they exist to exercise the per-generator
  parameter path bit-exactly in tests.
* ``write_tinymtdc`` / ``read_tinymtdc`` — the text format file_reader.cpp parses: lines
  starting with ``#`` are comments; mat1, mat2, tmat are the hex fields after the 3rd,
  4th and 5th comma.
"""
from __future__ import annotations

import numpy as np

DEFAULT_MAT1 = 0x8F7011EE
DEFAULT_MAT2 = 0xFC78FF1F
DEFAULT_TMAT = 0x3793FDFF


def default_table(n: int) -> np.ndarray:
    t = np.empty((n, 3), dtype=np.uint32)
    t[:, 0] = DEFAULT_MAT1
    t[:, 1] = DEFAULT_MAT2
    t[:, 2] = DEFAULT_TMAT
    return t


def scrambled_table(n: int, seed: int = 12345) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return rng.integers(0, 2**32, size=(n, 3), dtype=np.uint64).astype(np.uint32)


def write_tinymtdc(path: str, table: np.ndarray) -> None:
    with open(path, "w") as f:
        f.write("# characteristic, type, id, mat1, mat2, tmat, weight, delta\n")
        for i, (a, b, c) in enumerate(np.asarray(table, dtype=np.uint32)):
            f.write(f"0,32,{i},{int(a):08x},{int(b):08x},{int(c):08x},0,0\n")


def read_tinymtdc(path: str, n: int) -> np.ndarray:
    """The first n (mat1, mat2, tmat) entries of a tinymtdc file, read by the library's own
    reader (rt_parse_params_file, the one LoadTinyMT uses)."""
    from . import parse_params_file

    return parse_params_file(path, n)
