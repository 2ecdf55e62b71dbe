"""Ranged-row propagation of the oracle's Jacobi twin (params rangedrows=1, after
rangedRowPropagation, cons_linear.c:5715-6697): the two examples of the reference's own
documentation of that function (cons_linear.c:5702-5711), and soundness against brute-force
enumeration of every integer point of small ranged systems with gcd structure."""
import itertools

import numpy as np
import pytest

from oracle import oracle as orc
from tests.util import INF, make_instance


def prop(inst, ranged=True):
    return orc.propagate_jacobi(orc.OMatrix.from_instance(inst), inst.lb, inst.ub, orc.params(rangedrows=int(ranged)))


def test_reference_example_infeasible():
    # 12 x1 + 9 x2 - x3 = 0, x1, x2 free integers, 1 <= x3 <= 2: x3 must be a multiple of 3
    inst = make_instance([([(0, 12.0), (1, 9.0), (2, -1.0)], 0.0, 0.0)], [-INF, -INF, 1.0], [INF, INF, 2.0], [1, 1, 1])
    assert prop(inst)[2]["cutoff"] == 1
    assert prop(inst, ranged=False)[2]["cutoff"] == 0   # activity-based propagation alone sees nothing


def test_reference_example_fixing():
    # 12 x1 + 9 x2 - x3 = 1: the only possible value of x3 is 2
    inst = make_instance([([(0, 12.0), (1, 9.0), (2, -1.0)], 1.0, 1.0)], [-INF, -INF, 1.0], [INF, INF, 2.0], [1, 1, 1])
    lb, ub, r = prop(inst)
    assert r["cutoff"] == 0 and lb[2] == 2.0 and ub[2] == 2.0
    lb0, ub0, _ = prop(inst, ranged=False)
    assert (lb0[2], ub0[2]) == (1.0, 2.0)


def test_single_column_outside_the_check_group_is_bounded():
    # 3 y + x1 + x2 in [4, 5], x1, x2 binary, y integer in [0, 10]: x1 + x2 in [0, 2] leaves
    # 3 y in [2, 5], so y = 1 (one check-group column count 2: w = y)
    inst = make_instance([([(0, 3.0), (1, 1.0), (2, 1.0)], 4.0, 5.0)], [0.0, 0.0, 0.0], [10.0, 1.0, 1.0], [1, 1, 1])
    lb, ub, r = prop(inst)
    assert r["cutoff"] == 0 and (lb[0], ub[0]) == (1.0, 1.0)


def ranged_system(rng, nrows, ncols, shift=0, slack=True):
    """Ranged rows and equations of 3-5 columns through a random integer point (the last row's
    sides moved by shift; slack=False: the point is feasible), in the two shapes that give deductions: one unit coefficient among
    multiples of a gcd >= 2 (the unit column is bounded), or one multiple of the gcd among unit
    coefficients (that column is bounded)."""
    lb = rng.integers(-2, 1, size=ncols).astype(float)
    ub = lb + rng.integers(1, 5, size=ncols)
    xs = np.array([rng.integers(int(l), int(u) + 1) for l, u in zip(lb, ub)], dtype=float)
    rows = []
    for _ in range(nrows):
        L = int(rng.integers(3, 6))
        cols = rng.choice(ncols, size=L, replace=False)
        g = int(rng.integers(2, 7))
        sign = lambda: int(rng.choice([-1, 1]))
        if rng.integers(0, 2):
            coefs = [g * int(rng.integers(1, 3)) * sign() for _ in range(L - 1)] + [sign()]
        else:
            coefs = [g * sign()] + [sign() for _ in range(L - 1)]
        act = float(np.dot(coefs, xs[cols])) - (float(rng.integers(0, 2)) if slack else 0.0)
        width = float(rng.integers(0, 3))
        rows.append(([(int(c), float(a)) for c, a in zip(cols, coefs)], act, act + width))
    entries, lo, hi = rows[-1]
    rows[-1] = (entries, lo + shift, hi + shift)
    return make_instance(rows, lb, ub, [1] * ncols)


def feasible_points(inst):
    doms = [range(int(l), int(u) + 1) for l, u in zip(inst.lb, inst.ub)]
    pts = np.array(list(itertools.product(*doms)), dtype=float)
    ok = np.ones(len(pts), dtype=bool)
    for i in range(len(inst.lhs)):
        a, e = inst.rowptr[i], inst.rowptr[i + 1]
        act = pts[:, inst.colidx[a:e]] @ inst.vals[a:e]
        ok &= (act >= inst.lhs[i] - 1e-9) & (act <= inst.rhs[i] + 1e-9)
    return pts[ok]


@pytest.mark.parametrize("seed", range(40))
def test_ranged_sound_against_enumeration(seed):
    rng = np.random.default_rng(1200 + seed)
    inst = ranged_system(rng, 4, 7, shift=(seed % 4 == 3) * float(seed % 3 + 1))
    F = feasible_points(inst)
    lb, ub, r = prop(inst)
    if r["cutoff"]:
        assert len(F) == 0, seed
    else:
        assert np.all((F >= lb - 1e-9) & (F <= ub + 1e-9)), seed


def test_ranged_rows_add_deductions():
    # over the random systems, ranged propagation finds more than activity propagation alone
    extra = 0
    for seed in range(40):
        inst = ranged_system(np.random.default_rng(1200 + seed), 4, 7, shift=(seed % 4 == 3) * float(seed % 3 + 1))
        a = prop(inst)
        b = prop(inst, ranged=False)
        if a[2]["cutoff"] and not b[2]["cutoff"]:
            extra += 1
        elif not a[2]["cutoff"] and not b[2]["cutoff"]:
            extra += int(np.sum(a[0] > b[0]) + np.sum(a[1] < b[1]))
    assert extra > 5, extra


def test_large_system_gains_from_ranged_rows():
    inst = ranged_system(np.random.default_rng(1300), 1500, 3000, slack=False)
    a = prop(inst)
    b = prop(inst, ranged=False)
    assert not a[2]["cutoff"] and not b[2]["cutoff"]
    assert np.all(a[0] >= b[0]) and np.all(a[1] <= b[1])
    assert int(np.sum(a[0] > b[0]) + np.sum(a[1] < b[1])) > 20
