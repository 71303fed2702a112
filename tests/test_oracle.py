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


def _matchings(cov):
    """Number of perfect matchings of the covered vertices (brute force, tiny regions)."""
    V = {tuple(v) for v in np.argwhere(cov)}
    nb = {v: [w for w in [(v[0], v[1] + 1), (v[0] + 1, v[1]), (v[0] + 1, v[1] - 1),
                          (v[0], v[1] - 1), (v[0] - 1, v[1]), (v[0] - 1, v[1] + 1)] if w in V] for v in V}

    def count(left):
        if not left:
            return 1
        v = min(left)
        return sum(count(left - {v, w}) for w in nb[v] if w in left)

    return count(frozenset(V))


def test_tridimer_walk_keeps_a_dimer_cover(golden):
    """Every step keeps a perfect matching of the same vertices (TriangleDimerTiler.cpp:86-483
    with all three move types), and all three move types occur."""
    from tests.tilings import tridimer_covered, tridimer_legal
    for key in ["cone6_0", "cone9_2", "cone10_0"]:
        t0 = golden[f"tridimer_{key}"]
        cov = tridimer_covered(t0)
        t = t0
        for seed in range(3):
            t, mv = O.tridimer_walk(t, 150, 77 + seed, default_table(t0.size), return_moves=True)
            assert tridimer_legal(t, cov), (key, seed)
            assert set(np.unique(mv & 3)) == {0, 1, 2}
        assert not np.array_equal(t, t0)


def test_tridimer_move_schedule_draw_order():
    """Walk draws type, r1, r2 (+ p1, p2 for butterflies) from one minstd_rand0 through
    uniform_int_distribution(0,2)/(0,1) (TriangleDimerTiler.cpp:211-222, :291-292)."""
    t0 = np.zeros(16 * 16, np.int32)  # an empty region: the schedule alone is observable
    _, mv = O.tridimer_walk(t0, 200, 12345, default_table(64), return_moves=True)
    # minstd_rand0 (x <- 16807 x mod 2^31-1) restated independently of libstdc++
    x = 12345 % 2147483647

    def nxt():
        nonlocal x
        x = (x * 16807) % 2147483647
        return x

    def uni(hi):  # libstdc++ downscaling for a range of hi+1 values over minstd (max - min = 2^31 - 3)
        urange = hi + 1
        scaling = (2147483646 - 1) // urange
        past = urange * scaling
        while True:
            r = nxt() - 1
            if r < past:
                return r // scaling
    exp = []
    for _ in range(200):
        ty, r1, r2 = uni(2), uni(2), uni(1)
        p1 = p2 = 0
        if ty == 2:
            p1, p2 = uni(2), uni(1)
        exp.append(ty | r1 << 2 | r2 << 4 | p1 << 5 | p2 << 7)
    assert mv.tolist() == exp


def test_tridimer_uniform_law(golden):
    """The cone of side 3 has exactly 6 dimer covers (brute force); the chain's moves are
    symmetric involutions accepted with probability 1/2, so its law is uniform over them."""
    t0 = golden["tridimer_cone3_0"]
    from tests.tilings import tridimer_covered
    assert _matchings(tridimer_covered(t0)) == 6
    params = default_table(t0.size)
    counts = {}
    S = 1800
    for s in range(S):
        t = O.tridimer_walk(t0, 800, 7919 * s + 1, params)
        counts[t.tobytes()] = counts.get(t.tobytes(), 0) + 1
    assert len(counts) == 6
    exp = S / 6
    chi2 = sum((c - exp) ** 2 / exp for c in counts.values())
    assert chi2 < 20.5  # chi-square, 5 dof, p = 0.001


def test_cftp_schedules(golden):
    """Final replay lengths: 128 + 256 r for the reference schedule
    (AztecDiamondCftpEx.cpp:47-67), 128 * 2^r when doubling."""
    tmax, tmin = golden["domino_aztec16_max"], golden["domino_aztec16_min"]
    p = default_table(tmax.size)
    _, tot, r = O.domino_cftp(tmax, tmin, 345903, p, mode=0)
    assert r > 0 and tot == 128 + 256 * r
    _, tot, r = O.domino_cftp(tmax, tmin, 345903, p, mode=1)
    assert r > 0 and tot == 128 * 2 ** r


def test_cftp_uniform_law(golden):
    """Coupling from the past with dominokernelCFTP.cl's rule samples the 8 tilings of the
    order-2 Aztec diamond uniformly."""
    tmax, tmin = golden["domino_aztec2_max"], golden["domino_aztec2_min"]
    p = default_table(tmax.size)
    counts = {}
    S = 2400
    for s in range(S):
        t, _, _ = O.domino_cftp(tmax, tmin, 1000003 * s + 7, p)
        counts[t.tobytes()] = counts.get(t.tobytes(), 0) + 1
    assert len(counts) == 8
    exp = S / 8
    assert sum((c - exp) ** 2 / exp for c in counts.values()) < 24.3  # 7 dof, p = 0.001


def test_cftp_weighted_law(golden):
    """The weighted heat-bath (P(12) = p_v per vertex class and row parity) has stationary
    law mu(T) ~ prod_v (p_v / (1 - p_v))^((-1)^(i+j) h_T(v) / 4), h the domino height
    function: one flip at v changes only h(v), by 4 in the direction set by the
    checkerboard colour of v.  CFTP samples on the order-2 Aztec diamond follow it."""
    from randomtilings_amd import statics as St
    n = 2
    tmax, tmin, d = (golden[f"domino_aztec{n}_{k}"] for k in ("max", "min", "domain"))
    p12 = [0.8, 0.2, 0.3, 0.6]
    p = default_table(tmax.size)
    counts, tiles = {}, {}
    S = 4000
    for s in range(S):
        t, _, _ = O.domino_cftp(tmax, tmin, 1000003 * s + 7, p, rule=2, p12=p12)
        counts[t.tobytes()] = counts.get(t.tobytes(), 0) + 1
        tiles[t.tobytes()] = t
    assert len(counts) == 8
    N = int(round(np.sqrt(tmax.size)))
    inf = np.iinfo(np.int32).max
    keys = list(counts)
    logw = []
    for k in keys:
        h = St.call("domino.TilingToHeightfunc", tiles[k], d).reshape(N, N)
        ii, jj = np.nonzero(h != inf)
        pv = np.array([p12[((i + j) % 2) * 2 + (i & 1)] for i, j in zip(ii, jj)])
        logw.append(np.sum((-1.0) ** (ii + jj) * h[ii, jj] / 4 * np.log(pv / (1 - pv))))
    w = np.exp(np.array(logw) - max(logw))
    w /= w.sum()
    obs = np.array([counts[k] for k in keys])
    chi2 = float(((obs - S * w) ** 2 / (S * w)).sum())
    assert chi2 < 24.3, chi2  # 7 dof, p = 0.001


def _recttri_law_chi2(t0, w16, samples, steps):
    """Chi-square of the oracle's final states against mu(T) ~ prod_faces w(face).  The flip
    (recttrianglekernel.cl:25-118) proposes symmetrically (each of the five all-triangle /
    single-rectangle states of a hexagon goes to each other one with probability 0.9/4; the
    six rectangle pairs toggle with probability 0.9) and accepts iff the six-digit weight
    ratio >= a uniform draw, i.e. Metropolis for the product of the weights of the six faces
    around the vertex.  Every face's digit is copied unchanged into the two other vertices
    that see it (updateTiles1/2), so prod_v prod_digits w = prod_faces w^3."""
    p = default_table(t0.size)
    counts, tiles = {}, {}
    for s in range(samples):
        t = O.recttri_walk(t0, steps, 7919 * s + 1, p, weights16=w16)
        counts[t.tobytes()] = counts.get(t.tobytes(), 0) + 1
        tiles[t.tobytes()] = t
    keys = list(counts)
    lw = np.log(np.where(np.arange(16) == 0, 1.0, np.asarray(w16, np.float64)))
    logw = np.array([lw[np.stack([(tiles[k] >> (4 * q)) & 15 for q in range(6)])].sum() / 3 for k in keys])
    w = np.exp(logw - logw.max())
    w /= w.sum()
    obs = np.array([counts[k] for k in keys])
    return len(keys), float(((obs - samples * w) ** 2 / (samples * w)).sum())


@pytest.mark.parametrize("weighted", [False, True])
def test_recttri_law(golden, weighted):
    """The side-2 hexagon: 368 tilings reached; uniform for unit weights, the face-weight
    product law for the weights below (a wrong exponent of the vertex product, 1/2 or 1
    instead of 1/3, gives chi-square above 1200)."""
    from scipy.stats import chi2
    t0 = golden["recttri_minHex2"]
    w16 = np.ones(16, np.float32)
    if weighted:
        w16 = np.array([0, 1.0, 0.6, 1.4, 0.7, 1.3, 0.9, 0.7, 1.3, 0.9, 1.2, 0.8, 1.1, 1.2, 0.8, 1.1], np.float32)
    k, c = _recttri_law_chi2(t0, w16, 12000, 200)
    assert k == 368
    assert c < chi2.ppf(0.999, k - 1)
