"""Pins the oracle's double-double arithmetic and canonical activity sum.

- Known-answer vectors of the reference's own unit test (tests/src/misc/dbldblarith.c),
  committed as tests/golden/dbldblarith_kat.json.
- Bit-exact agreement with the reference's src/scip/dbldblarith.h macros compiled from the
  SCIP tree (oracle/_ref/libdbldbl_ref.so) on random inputs.
- Mathematical properties of the canonical pairwise sum (exactness on integers, zero padding,
  tiling invariance, accuracy against math.fsum).
"""
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as orc

GOLD = os.path.join(os.path.dirname(__file__), "golden", "dbldblarith_kat.json")


def _kat():
    with open(GOLD) as f:
        return json.load(f)


def test_kat_sum_matches_reference_vector():
    k = _kat()
    a, b, c = (float.fromhex(k["inputs"][n]) for n in "abc")
    L = orc.lib()
    hi, lo = orc.dd2(L.orc_dd_sum, c, b)
    hi, lo = orc.dd2(L.orc_dd_sum21, hi, lo, a)
    hi, lo = orc.dd2(L.orc_dd_sum21, hi, lo, a)
    exp = float.fromhex(k["sum"]["expected"])
    assert hi + lo == exp
    assert ((c + b) + a) + a != exp  # the vector is one where plain double arithmetic fails


def test_kat_product_matches_reference_vector():
    k = _kat()
    a, b, c = (float.fromhex(k["inputs"][n]) for n in "abc")
    L = orc.lib()
    hi, lo = orc.dd2(L.orc_dd_prod, a, b)
    hi, lo = orc.dd2(L.orc_dd_prod21, hi, lo, c)
    exp = float.fromhex(k["product"]["expected"])
    assert hi + lo == exp
    assert (a * b) * c != exp


def test_twosum_error_is_exact():
    rng = np.random.default_rng(7)
    L = orc.lib()
    for _ in range(2000):
        a = float(rng.standard_normal() * 10.0 ** rng.integers(-8, 9))
        b = float(rng.standard_normal() * 10.0 ** rng.integers(-8, 9))
        hi, lo = orc.dd2(L.orc_dd_sum, a, b)
        assert Fraction(hi) + Fraction(lo) == Fraction(a) + Fraction(b)
        assert hi == a + b


def test_twoprod_error_is_exact():
    rng = np.random.default_rng(8)
    L = orc.lib()
    for _ in range(2000):
        a = float(rng.standard_normal() * 10.0 ** rng.integers(-6, 7))
        b = float(rng.standard_normal() * 10.0 ** rng.integers(-6, 7))
        hi, lo = orc.dd2(L.orc_dd_prod, a, b)
        assert Fraction(hi) + Fraction(lo) == Fraction(a) * Fraction(b)


@pytest.mark.skipif(orc.reflib() is None, reason="reference tree not present (oracle/_ref not built)")
def test_restatement_bit_exact_against_reference_header():
    rng = np.random.default_rng(11)
    L, R = orc.lib(), orc.reflib()
    for _ in range(5000):
        xs = [float(rng.standard_normal() * 10.0 ** rng.integers(-12, 13)) for _ in range(3)]
        assert orc.dd2(L.orc_dd_sum, xs[0], xs[1]) == orc.dd2(R.ref_dd_sum, xs[0], xs[1])
        assert orc.dd2(L.orc_dd_sum21, *xs) == orc.dd2(R.ref_dd_sum21, *xs)
        assert orc.dd2(L.orc_dd_prod, xs[0], xs[1]) == orc.dd2(R.ref_dd_prod, xs[0], xs[1])
        assert orc.dd2(L.orc_dd_prod21, *xs) == orc.dd2(R.ref_dd_prod21, *xs)


def _seq_activity(x):
    """SCIP's sequential Sum21 accumulation of nonzero contributions (cons_linear.c:2283 order)."""
    L = orc.lib()
    hi, lo = 0.0, 0.0
    for v in x:
        if v != 0.0:
            hi, lo = orc.dd2(L.orc_dd_sum21, hi, lo, float(v))
    return hi, lo


@pytest.mark.skipif(orc.reflib() is None, reason="reference tree not present (oracle/_ref not built)")
def test_sequential_activity_matches_reference_macro_loop():
    import ctypes as C
    rng = np.random.default_rng(12)
    R = orc.reflib()
    for n in (1, 2, 7, 64, 1000):
        x = rng.standard_normal(n) * 10.0 ** rng.integers(-3, 4, size=n)
        x[rng.random(n) < 0.1] = 0.0
        hi, lo = C.c_double(), C.c_double()
        R.ref_activity_sum(x.ctypes.data_as(C.POINTER(C.c_double)), n, C.byref(hi), C.byref(lo))
        assert _seq_activity(x) == (hi.value, lo.value)


def test_canonical_sum_exact_on_integers():
    rng = np.random.default_rng(3)
    for n in (0, 1, 2, 3, 5, 64, 65, 1000, 10000):
        x = rng.integers(-10**6, 10**6, size=n).astype(float)
        hi, lo = orc.canonical_sum(x)
        assert hi + lo == float(sum(int(v) for v in x))
        assert lo == 0.0


def test_canonical_sum_zero_padding_and_tiling():
    rng = np.random.default_rng(4)
    L = orc.lib()
    for n in (3, 64, 100, 777):
        x = rng.standard_normal(n) * 10.0 ** rng.integers(-3, 4, size=n)
        base = orc.canonical_sum(x)
        for extra in (1, 5, 64):
            assert orc.canonical_sum(np.concatenate([x, np.zeros(extra)])) == base
        # the root is the pair of the two aligned halves of the padded tree
        P = 1
        while P < n:
            P *= 2
        if P >= 2:
            xp = np.concatenate([x, np.zeros(P - n)])
            left = orc.canonical_sum(xp[: P // 2])
            right = orc.canonical_sum(xp[P // 2:])
            assert orc.dd2(L.orc_dd_pairadd, *left, *right) == base


def test_canonical_sum_accuracy_vs_fsum():
    rng = np.random.default_rng(5)
    for n in (10, 100, 1000, 5000):
        for _ in range(5):
            x = rng.standard_normal(n) * 10.0 ** rng.integers(-3, 8, size=n)
            hi, lo = orc.canonical_sum(x)
            exact = math.fsum(x)
            tol = 4 * np.finfo(float).eps * max(abs(exact), 1e-300) + 1e-26 * np.abs(x).sum()
            assert abs((hi + lo) - exact) <= tol
            shi, slo = _seq_activity(x)
            assert abs((shi + slo) - exact) <= tol


def test_pairadd_commutative_and_zero_identity():
    rng = np.random.default_rng(6)
    L = orc.lib()
    for _ in range(2000):
        a = (float(rng.standard_normal() * 1e6), float(rng.standard_normal() * 1e-11))
        b = (float(rng.standard_normal() * 1e3), float(rng.standard_normal() * 1e-14))
        assert orc.dd2(L.orc_dd_pairadd, *a, *b) == orc.dd2(L.orc_dd_pairadd, *b, *a)
        assert orc.dd2(L.orc_dd_pairadd, *a, 0.0, 0.0) == a
