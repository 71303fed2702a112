"""librt_hip.so loads and exports every symbol include/rt_hip.h declares; the host-only
threshold derivation reproduces the reference's float decisions for every k (brute force
over all 2^23 mantissas, evaluated with IEEE float32/float64 as the reference kernels do).
No GPU is needed."""
import re
from pathlib import Path

import numpy as np
import pytest

import randomtilings_amd as rt

ROOT = Path(__file__).resolve().parents[1]
K = np.arange(1 << 23, dtype=np.uint32)
RD = K.astype(np.float32) * np.float32(2.0**-23)  # the reference's rd, exactly k * 2^-23
RD64 = RD.astype(np.float64)


def header_symbols():
    txt = (ROOT / "include" / "rt_hip.h").read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(rt_[a-z0-9_]+)\s*\(", txt)))


def test_all_header_symbols_exported():
    lib = rt.load()
    syms = header_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) <= set(rt.SIGNATURES), set(syms) - set(rt.SIGNATURES)


def test_version():
    assert b"gfx950" in rt.load().rt_version()


def _interval(mask):
    """[lo, hi) of a boolean mask over k that must be one contiguous run."""
    idx = np.flatnonzero(mask)
    if idx.size == 0:
        return None
    lo, hi = int(idx[0]), int(idx[-1]) + 1
    assert idx.size == hi - lo, "decision is not an interval of k"
    return lo, hi


def _fr(th):
    fr = th[:24].reshape(4, 3, 2)  # ab_lo, ab_hi, ba_lo, ba_hi x [class][parity]
    return fr


def check_transition(lo, hi, mask):
    iv = _interval(mask)
    if iv is None:
        assert lo >= hi
    else:
        assert (lo, hi) == iv


def test_domino_lazy_thresholds():
    fr = _fr(rt.rule_thresholds(rt.RT_DOMINO, rt.RT_DOMINO_LAZY))
    mask = RD64 < 0.8  # dominokernel.cl:11
    for c in range(2):
        for p in range(2):
            check_transition(fr[0, c, p], fr[1, c, p], mask)
            check_transition(fr[2, c, p], fr[3, c, p], mask)
    assert fr[1, 0, 0] == 6710887


def test_domino_cftp_thresholds():
    fr = _fr(rt.rule_thresholds(rt.RT_DOMINO, rt.RT_DOMINO_CFTP))
    for c in range(2):
        for p in range(2):
            check_transition(fr[0, c, p], fr[1, c, p], RD64 < 0.5)  # dominokernelCFTP.cl:25
            check_transition(fr[2, c, p], fr[3, c, p], ~(RD64 < 0.5))
    assert fr[1, 0, 0] == 1 << 22


def test_domino_weighted_thresholds_and_uniform_reduction():
    p12 = [0.5, 0.5, 0.2, 0.8]
    fr = _fr(rt.rule_thresholds(rt.RT_DOMINO, rt.RT_DOMINO_WEIGHTED, p12))
    for c in range(2):
        for par in range(2):
            pr = p12[c * 2 + par]
            check_transition(fr[0, c, par], fr[1, c, par], RD64 < pr)
            check_transition(fr[2, c, par], fr[3, c, par], ~(RD64 < pr))
    a = rt.rule_thresholds(rt.RT_DOMINO, rt.RT_DOMINO_WEIGHTED, [0.5] * 4)
    b = rt.rule_thresholds(rt.RT_DOMINO, rt.RT_DOMINO_CFTP)
    assert np.array_equal(a[:24], b[:24])


def test_lozenge_uniform_thresholds():
    fr = _fr(rt.rule_thresholds(rt.RT_LOZENGE, rt.RT_LOZENGE_UNIFORM))
    add = RD64 < 0.4  # lozengekernel.cl:21-24
    rem = ~add & (RD64 < 0.8)
    check_transition(fr[0, 0, 0], fr[1, 0, 0], add)
    check_transition(fr[2, 0, 0], fr[3, 0, 0], rem)
    assert (fr[1, 0, 0], fr[3, 0, 0]) == (3355444, 6710887)


@pytest.mark.parametrize("q", [0.99, 0.999, 0.5, 1.0, 2.0, 3.9, 7.5])
def test_lozenge_qref_thresholds(q):
    fr = _fr(rt.rule_thresholds(rt.RT_LOZENGE, rt.RT_LOZENGE_QREF, [q]))
    low = RD64 < 0.5  # qlozengekernel.cl:23-29, rd re-stored into the float variable
    r_add = (RD64 / 0.5).astype(np.float32).astype(np.float64)
    r_rem = ((RD64 - 0.5) * 0.5).astype(np.float32).astype(np.float64)
    check_transition(fr[0, 0, 0], fr[1, 0, 0], low & (q >= r_add))
    check_transition(fr[2, 0, 0], fr[3, 0, 0], ~low & (1 / q >= r_rem))
    if q == 0.999:
        assert fr[1, 0, 0] == 4190110  # k <= 4,190,109 adds


@pytest.mark.parametrize("q", [0.999, 0.3, 1.7])
def test_lozenge_heatbath_thresholds(q):
    fr = _fr(rt.rule_thresholds(rt.RT_LOZENGE, rt.RT_LOZENGE_QHEAT, [q]))
    add = RD64 < q / (1 + q)
    check_transition(fr[0, 0, 0], fr[1, 0, 0], add)
    check_transition(fr[2, 0, 0], fr[3, 0, 0], ~add)


CYC = [0x122133, 0x331221, 0x8EBBE8, 0xFC99CF, 0xA7DD7A]
CYC_T = [[0xA7DD7A, 0xFC99CF, 0x8EBBE8, 0x331221], [0xA7DD7A, 0xFC99CF, 0x8EBBE8, 0x122133],
         [0xA7DD7A, 0xFC99CF, 0x122133, 0x331221], [0xA7DD7A, 0x122133, 0x8EBBE8, 0x331221],
         [0x122133, 0xFC99CF, 0x8EBBE8, 0x331221]]
TOG = [(0x47D47A, 0xD22A33), (0xD22A33, 0x47D47A), (0xA74D74, 0x33A22D), (0x33A22D, 0xA74D74),
       (0x8EB558, 0x331EB1), (0x331EB1, 0x8EB558), (0x855BE8, 0x1BE133), (0x1BE133, 0x855BE8),
       (0x6699CF, 0xCF1221), (0xCF1221, 0x6699CF), (0xFC9966, 0x1221FC), (0x1221FC, 0xFC9966)]


def _ratio(w, a, b):
    def prod(x):
        p = np.float32(1.0)
        first = True
        for sh in (20, 16, 12, 8, 4, 0):
            v = w[(x >> sh) & 0xF]
            p = v if first else np.float32(p * v)
            first = False
        return p
    return np.float32(prod(b) / prod(a))


@pytest.mark.parametrize("weights", [None, "random"])
def test_recttri_thresholds(weights):
    if weights is None:
        w = np.array([0.0] + [1.0] * 15, np.float32)
        th = rt.rule_thresholds(rt.RT_RECTTRI, 0)
    else:
        w = np.random.default_rng(3).uniform(0.3, 1.7, 16).astype(np.float32)
        w[0] = 0
        th = rt.rule_thresholds(rt.RT_RECTTRI, 0, w.astype(np.float64))
    r = th[24:]
    gate, band = r[0], r[1:4]
    cyc_acc = r[4:24].reshape(5, 4)
    tog_acc = r[24:36]
    bs = np.float32(0.9)
    g = RD < bs  # recttrianglekernel.cl:36-37
    assert _interval(g) == (0, gate)
    x = (RD / bs).astype(np.float32)
    for b in range(3):
        assert _interval(g & (x.astype(np.float64) < 0.25 * (b + 1))) == (0, band[b])
    bands = [(0, band[0]), (band[0], band[1]), (band[1], band[2]), (band[2], gate)]
    for s in range(5):
        for b in range(4):
            lo, hi = bands[b]
            ratio = _ratio(w, CYC[s], CYC_T[s][b])
            rdp = ((x[lo:hi] - np.float32(0.25 * b)) * np.float32(4)).astype(np.float32)
            acc = ratio >= rdp
            m = np.zeros(1 << 23, bool)
            m[lo:hi] = acc
            check_transition(lo, cyc_acc[s, b], m)
    for t, (a, bb) in enumerate(TOG):
        ratio = _ratio(w, a, bb)
        m = g & (ratio >= x)
        check_transition(0, tog_acc[t], m)


def test_bad_arguments_raise():
    with pytest.raises(rt.RTError):
        rt.rule_thresholds(rt.RT_LOZENGE, rt.RT_LOZENGE_QREF, [-1.0])
    with pytest.raises(rt.RTError):
        rt.rule_thresholds(rt.RT_DOMINO, rt.RT_DOMINO_WEIGHTED, [1.5, 0.5, 0.5, 0.5])


def test_recttri_state_table():
    """The rect-triangle rule's hash table (rt_rule_thresholds words 60..155): the 17 states
    that can move sit at their slot h = (s * 0xc9c201 mod 2^32) >> 27 with their index, the
    other slots are empty, and the toggle targets pair the 12 toggle states with each other."""
    w = rt.rule_thresholds(rt.RT_RECTTRI, 0)
    hkey, hval = w[60:92], w[92:124]
    cyc_nt, tog_nt = w[124:144].reshape(5, 4), w[144:156]
    cyc = [0x122133, 0x331221, 0x8ebbe8, 0xfc99cf, 0xa7dd7a]
    tog = [0x47d47a, 0xd22a33, 0xa74d74, 0x33a22d, 0x8eb558, 0x331eb1, 0x855be8, 0x1be133, 0x6699cf,
           0xcf1221, 0xfc9966, 0x1221fc]
    slots = set()
    for idx, s in enumerate(cyc + tog):
        h = ((s * 0xc9c201) & 0xffffffff) >> 27
        assert hkey[h] == s and hval[h] == idx
        slots.add(h)
    assert len(slots) == 17
    assert all(hkey[h] == 0xffffffff for h in range(32) if h not in slots)
    assert set(cyc_nt.ravel()) <= set(cyc)  # a 5-cycle state moves to another 5-cycle state
    assert all(cyc_nt[i, b] != cyc[i] for i in range(5) for b in range(4))
    for t in range(12):
        partner = tog.index(int(tog_nt[t]))
        assert tog_nt[partner] == tog[t]
