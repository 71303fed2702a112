"""The tinymtdc parameter-file reader (rt_parse_params_file, behind rt_load_params_file and
every XTiler::LoadTinyMT(path, n)) against the reference's own reader,
tinymt::file_reader::get (src/TinyMT/file_reader.cpp:51-70,114-166), compiled from its
sources into oracle/_ref.  Host only: no device is touched.

Each file is synthetic, in the tinymtdc-longbatch line shape
``<charpoly>,<type>,<id>,<mat1>,<mat2>,<tmat>,<weight>,<delta>``: the test knows the values
it wrote, and where the reference library is present its reader must return the same table
or fail on the same files.
"""
import numpy as np
import pytest

import oracle as O
import randomtilings_amd as rt

ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference sources absent)")


def _line(rng, i, fmt="{:08x}", m=None):
    m = m if m is not None else [int(x) for x in rng.integers(0, 2**32, 3, dtype=np.uint64)]
    poly = "".join(f"{int(x):08x}" for x in rng.integers(0, 2**32, 4, dtype=np.uint64))
    return f"{poly},3,{i}," + ",".join(fmt.format(v) for v in m) + ",63,0", m


def _write(tmp_path, lines, name="p.txt", newline="\n"):
    p = tmp_path / name
    p.write_bytes(newline.join(lines).encode() + newline.encode())
    return p


def _both(path, n):
    """(ours or the RTError code, the reference's or None when it threw / is absent)."""
    try:
        mine = rt.parse_params_file(path, n)
    except rt.RTError as e:
        mine = e.code
    theirs = O.ref_tinymt_file_read(path, n) if O.ref_available() else "absent"
    return mine, theirs


def test_reads_the_fields_after_commas_3_4_5(tmp_path):
    rng = np.random.default_rng(1)
    lines, want = ["# tinymt32dc parameters", "# charpoly,type,id,mat1,mat2,tmat,weight,delta"], []
    fmts = ["{:08x}", "{:08X}", "{:x}", "0x{:x}", " {:08x}"]  # strtoul: case, no padding, 0x, leading blank
    for i in range(200):
        ln, m = _line(rng, i, fmts[i % len(fmts)])
        lines.append(ln)
        want.append(m)
        if i % 37 == 5:
            lines.append("# a comment between entries")
    p = _write(tmp_path, lines)
    mine, theirs = _both(p, 200)
    assert np.array_equal(mine, np.array(want, np.uint32))
    if isinstance(theirs, np.ndarray):
        assert np.array_equal(mine, theirs)
    # the first n entries of a longer file
    assert np.array_equal(rt.parse_params_file(p, 17), np.array(want[:17], np.uint32))


def test_default_parameter_line():
    """TinyMT's published default parameters (src/TinyMT/tinymt32def.clh:48-50)."""
    import tempfile
    from pathlib import Path
    with tempfile.TemporaryDirectory() as d:
        p = Path(d) / "one.txt"
        p.write_text("# x\n0,3,0,8f7011ee,fc78ff1f,3793fdff,63,0\n")
        assert rt.parse_params_file(p, 1).tolist() == [[0x8F7011EE, 0xFC78FF1F, 0x3793FDFF]]


@ref
def test_crlf_lines_and_long_fields_match_reference(tmp_path):
    rng = np.random.default_rng(2)
    lines = [_line(rng, i)[0] for i in range(10)]
    # a 9-digit field fits an unsigned long and is truncated to 32 bits by both readers
    lines.append("abcd,3,10,123456789,fc78ff1f,3793fdff,63,0")
    p = _write(tmp_path, lines, newline="\r\n")
    mine, theirs = _both(p, 11)
    assert isinstance(mine, np.ndarray) and np.array_equal(mine, theirs)
    assert mine[10, 0] == 0x23456789


@pytest.mark.parametrize("case", ["short", "missing_comma", "empty_line", "overflow", "no_file"])
def test_format_errors(tmp_path, case):
    rng = np.random.default_rng(3)
    lines = [_line(rng, i)[0] for i in range(8)]
    n = 8
    if case == "short":
        n = 9
    elif case == "missing_comma":
        lines[4] = "abcd,3,4,8f7011ee,fc78ff1f"
    elif case == "empty_line":
        lines[5] = ""
    elif case == "overflow":  # beyond unsigned long: strtoul sets ERANGE
        lines[6] = "abcd,3,6,8f7011ee,1" + "0" * 17 + ",3793fdff,63,0"
    p = tmp_path / "absent.txt" if case == "no_file" else _write(tmp_path, lines)
    mine, theirs = _both(p, n)
    assert mine == rt.RT_ERR_IO
    if theirs != "absent":
        assert theirs is None
    if case != "no_file":  # the entries before the bad one still read
        good = min(4, n - 1)
        assert rt.parse_params_file(p, good).shape == (good, 3)
