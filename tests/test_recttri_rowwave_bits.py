"""The rect-triangle row-wave update builds each class word with two field selects and one
byte permutation (csrc/rt_kernels.hip, kRowWave): class c+1 from x = up, y, z and class c+2
from x, y, z = dn.  This pins those constants, read from the kernel source, against the
reference's mask-and-shift update (recttrianglekernel.cl:120-146, as oracle/rt_oracle.cpp's
recttri_update1/2 restate it) on random 24-bit states, with v_perm_b32 and v_bfi_b32 emulated
bit for bit; a changed selector or mask fails."""
import re
from pathlib import Path

import numpy as np
import pytest

SRC = Path(__file__).resolve().parents[1] / "randomtilings_amd" / "csrc" / "rt_kernels.hip"


def _constants():
    s = SRC.read_text()
    u = re.search(r"const uint32_t u = \(y1 & (0x[0-9a-f]+)u\) \| \(z1 & ~(0x[0-9a-f]+)u\);", s)
    v = re.search(r"const uint32_t v = \(x2 & (0x[0-9a-f]+)u\) \| \(y2 & ~(0x[0-9a-f]+)u\);", s)
    p1 = re.search(r"__builtin_amdgcn_perm\(up, u, (0x[0-9a-f]+)u\)", s)
    p2 = re.search(r"__builtin_amdgcn_perm\(dn, v, (0x[0-9a-f]+)u\)", s)
    assert u and v and p1 and p2, "row-wave update lines not found"
    assert u.group(1) == u.group(2) and v.group(1) == v.group(2)
    return int(u.group(1), 16), int(v.group(1), 16), int(p1.group(1), 16), int(p2.group(1), 16)


def perm(s0, s1, sel):
    """v_perm_b32 for selectors 0-7 (bytes of {s0:s1}, s1 the low word) and 0x0c (zero)."""
    out = np.zeros_like(s0)
    for i in range(4):
        x = (sel >> (8 * i)) & 0xFF
        if x < 4:
            byte = (s1 >> np.uint32(8 * x)) & np.uint32(0xFF)
        elif x < 8:
            byte = (s0 >> np.uint32(8 * (x - 4))) & np.uint32(0xFF)
        elif x == 0x0C:
            byte = np.zeros_like(s0)
        else:
            raise ValueError(f"selector byte {x:#x} not emulated")
        out |= byte << np.uint32(8 * i)
    return out


def bfi(mask, a, b):
    m = np.uint32(mask)
    return (a & m) | (b & ~m)


def ref_update1(x, y, z):  # recttri_update1: x = s[i-1][j], y = s[i][j+p1], z = s[i+1][j-p2]
    return ((x & 0x0000F0) * 0x010000 + (x & 0x00000F) * 0x010000 + (y & 0xF00000) // 0x000100 +
            (z & 0x0F0000) // 0x000100 + (z & 0x00F000) // 0x000100 + (y & 0x000F00) // 0x000100)


def ref_update2(x, y, z):  # recttri_update2: x = s[i][j-p1], y = s[i-1][j+p2], z = s[i+1][j]
    return ((x & 0x00F000) * 0x000100 + (y & 0x000F00) * 0x000100 + (y & 0x0000F0) * 0x000100 +
            (x & 0x00000F) * 0x000100 + (z & 0xF00000) // 0x010000 + (z & 0x0F0000) // 0x010000)


def kernel_update(consts, x, y, z, which):
    mu, mv, s1, s2 = consts
    if which == 1:
        return perm(x, bfi(mu, y, z), s1)  # d1 = perm(up, u(y1, z1))
    return perm(z, bfi(mv, x, y), s2)      # d2 = perm(dn, v(x2, y2))


@pytest.mark.parametrize("which", [1, 2])
def test_rowwave_update_matches_reference(which):
    consts = _constants()
    rng = np.random.default_rng(7 + which)
    x, y, z = (rng.integers(0, 1 << 24, 200000, dtype=np.uint64).astype(np.uint32) for _ in range(3))
    ref = (ref_update1 if which == 1 else ref_update2)(x.astype(np.int64), y.astype(np.int64), z.astype(np.int64))
    got = kernel_update(consts, x, y, z, which).astype(np.int64)
    assert np.array_equal(got, ref)


def test_a_changed_selector_fails():
    mu, mv, s1, s2 = _constants()
    rng = np.random.default_rng(3)
    x, y, z = (rng.integers(0, 1 << 24, 4096, dtype=np.uint64).astype(np.uint32) for _ in range(3))
    ref = ref_update1(x.astype(np.int64), y.astype(np.int64), z.astype(np.int64))
    for bad in (s1 ^ 0x01, s1 ^ 0x0100, (s1 & ~0xFF0000) | 0x050000):
        assert not np.array_equal(kernel_update((mu, mv, bad, s2), x, y, z, 1).astype(np.int64), ref)
    assert not np.array_equal(kernel_update((mu ^ 0x100, mv, s1, s2), x, y, z, 1).astype(np.int64), ref)


@pytest.mark.parametrize("seed", [None, 1, 2, 3])
def test_raw_output_thresholds_exact(seed):
    """The row-wave flip compares the raw tempered output t0 with t << 9 (k = t0 >> 9): exact
    when every band limit and acceptance bound is below 2^23 (k < t <=> t0 < t << 9, and t << 9
    does not wrap).  They lie at or below the gate, which is below 2^23, for unit and random
    weights (including weights far from 1, where acceptance saturates)."""
    import randomtilings_amd as rt

    if seed is None:
        th = rt.rule_thresholds(rt.RT_RECTTRI, 0)
    else:
        w = np.random.default_rng(seed).uniform(0.05, 20.0, 16).astype(np.float32)
        w[0] = 0
        th = rt.rule_thresholds(rt.RT_RECTTRI, 0, w.astype(np.float64))
    r = th[24:]
    gate, band, cyc_acc, tog_acc = int(r[0]), r[1:4], r[4:24], r[24:36]
    assert 0 < gate < (1 << 23)
    for t in list(band) + list(cyc_acc) + list(tog_acc):
        assert int(t) <= gate
        k = np.array([0, max(int(t) - 1, 0), int(t), (1 << 23) - 1], dtype=np.uint64)
        for base in k << np.uint64(9):  # the smallest and largest raw outputs of each k
            for t0 in (int(base), int(base) | 511):
                assert (t0 < (int(t) << 9)) == ((t0 >> 9) < int(t))
