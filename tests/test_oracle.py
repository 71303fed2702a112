"""The CPU oracle, pinned before it is trusted.

* TinyMT32: TinyMT's published check output for tinymt32_init(1) with the default
  parameter set, and agreement with the reference's own branchless host coding
  (src/TinyMT/tinymt32.h) compiled into oracle/_ref.
* host schedule RNGs: the C++ standard's 10000th-output anchors, and the libstdc++
  uniform_int_distribution sequence the reference's Walk draws its colours from.
* chain rules: legality invariants after every walk, and the exact stationary laws of tiny
  domains by brute force (all tilings enumerated, frequencies chi-square tested).
"""
import numpy as np
import pytest

import oracle as O
from randomtilings_amd.params import default_table, scrambled_table
from tests.conftest import square
from tests.tilings import domino_legal, lozenge_legal, domino_key, lozenge_cubes


def test_tinymt_kat_default_params():
    out = O.tinymt_uint32(0x8F7011EE, 0xFC78FF1F, 0x3793FDFF, 1, 5)
    assert list(out) == [2545341989, 981918433, 3715302833, 2387538352, 3591001365]


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference sources absent)")
def test_tinymt_matches_reference_host_coding():
    import ctypes as C
    rng = np.random.default_rng(7)
    for _ in range(20):
        mat = [int(x) for x in rng.integers(0, 2**32, 3)]
        seed = int(rng.integers(0, 2**32))
        k, st = O.tinymt_table_draws(np.array([mat], np.uint32), seed, 1, 0)
        # oracle: init then 64 outputs; reference host coding from the same post-init state
        ours = O.tinymt_uint32(mat[0], mat[1], mat[2], seed, 64)
        out = np.zeros(64, np.uint32)
        O.ref().ref_tinymt_host(*[int(x) for x in st[0]], *mat, 64, out.ctypes.data_as(C.POINTER(C.c_uint32)))
        assert np.array_equal(out, ours)


def test_single01_is_k_times_2pow23():
    k, _ = O.tinymt_table_draws(default_table(4), 99, 4, 16)
    u = np.concatenate([O.tinymt_uint32(0x8F7011EE, 0xFC78FF1F, 0x3793FDFF, 99 + i, 16)[None] for i in range(4)])
    assert np.array_equal(k, (u >> 9).T)


def test_schedule_rng_anchors():
    L = O.lib()
    assert L.or_mt19937_nth(5489, 10000) == 4123659995
    assert L.or_minstd0_nth(1, 10000) == 1043618065
    out = np.zeros(12, np.int32)
    L.or_minstd0_uniform(12345, 0, 2, 12, out.ctypes.data_as(O._I32P))
    assert list(out) == [0, 2, 2, 0, 0, 0, 2, 1, 2, 2, 0, 0]


@pytest.mark.parametrize("rule", [0, 1])
def test_domino_walk_stays_legal(golden, rule):
    for n in [1, 2, 5, 16]:
        t0 = golden[f"domino_aztec{n}_min"]
        d = golden[f"domino_aztec{n}_domain"]
        for seed in [1, 2]:
            t = O.domino_walk(t0, 300, seed, default_table(t0.size), rule=rule)
            assert domino_legal(square(t), square(d))


def test_domino_aztec2_uniform_law(golden):
    """Aztec diamond of order 2 has 8 tilings; the lazy chain's law is uniform."""
    t0 = golden["domino_aztec2_min"]
    params = default_table(t0.size)
    counts = {}
    S = 1600
    for seed in range(S):
        t = O.domino_walk(t0, 60, 1000 + seed, params, rule=0)
        k = domino_key(t)
        counts[k] = counts.get(k, 0) + 1
    assert len(counts) == 8
    exp = S / 8
    chi2 = sum((c - exp) ** 2 / exp for c in counts.values())
    assert chi2 < 24.3  # chi-square, 7 dof, p = 0.001


def test_lozenge_hexagon_q_law(golden):
    """The 1x1x1 hexagon has 2 tilings (with or without the cube); the q-rule's
    stationary ratio P(cube)/P(no cube) is q."""
    t0 = golden["lozenge_hex1_0_min"]
    params = default_table(t0.size)
    q = 0.6
    S = 2000
    n_cube = 0
    base = lozenge_cubes(square(t0))
    for seed in range(S):
        t = O.lozenge_walk(t0, 40, 5000 + seed, params, rule=1, q=q)
        assert lozenge_legal(square(t))
        n_cube += lozenge_cubes(square(t)) != base
    p = n_cube / S
    expect = q / (1 + q)
    assert abs(p - expect) < 4 * np.sqrt(expect * (1 - expect) / S)


def test_oracle_per_generator_params_change_the_walk(golden):
    t0 = golden["domino_aztec8_min"]
    a = O.domino_walk(t0, 500, 3, default_table(t0.size), rule=1)
    b = O.domino_walk(t0, 500, 3, scrambled_table(t0.size, 1), rule=1)
    assert not np.array_equal(a, b)
