"""Where the propagation fixpoint is unique, both oracle schedules must reach it.

With integer columns, integer coefficients, integral sides and numerics/boundstreps = 0, every
strict improvement of an integer bound is accepted (SCIPsetIsUbBetter, set.c:7735-7753, reduces
to newub < oldub), all activities are exact integers, and SCIP's tightening of a column's bound
is the largest (smallest) integer value that the row still supports.  The propagation operator
is then monotone, so every fair order of row visits -- SCIP's Gauss-Seidel order
(orc_propagate_scip) as well as the device's Jacobi rounds (orc_propagate_jacobi) -- ends in the
same greatest fixpoint.

That fixpoint is computed here without any of the oracle's code: as the join of all boxes in
which every bound value of every column is supported by each of its rows (bounds consistency:
with column j at the bound value v, the other columns of the row can still reach an activity
inside [lhs, rhs] within the box).  The join of such boxes has the property again, so it is the
greatest one; if no such box is non-empty, propagation must end in a cutoff.  This extends the
pure-binary uniqueness pin (test_binary_fixpoint_unique_jacobi_equals_scip_order) to general
integers.
"""
import itertools

import numpy as np
import pytest

from oracle import oracle as orc
from tests.util import INF, make_instance

BS0 = dict(boundstreps=0.0)


def integer_instance(rng, nrows, ncols, maxlen, coefrange=5, boundrange=3, ubmax=None, planted=False):
    """Integer columns, integer coefficients, integral sides (ranged, <= or >= rows); with
    planted, every row's sides hold a random integer point of the box, so the system is
    feasible."""
    lb = rng.integers(-boundrange, 1, size=ncols).astype(float)
    ub = lb + rng.integers(0, (ubmax if ubmax is not None else boundrange) + 1, size=ncols)
    x = np.array([rng.integers(int(l), int(u) + 1) for l, u in zip(lb, ub)])
    rows = []
    for _ in range(nrows):
        L = int(rng.integers(1, maxlen + 1))
        cols = rng.choice(ncols, size=min(L, ncols), replace=False)
        coefs = rng.integers(-coefrange, coefrange + 1, size=len(cols))
        coefs[coefs == 0] = 1
        mn = int(sum(a * (lb[c] if a > 0 else ub[c]) for c, a in zip(cols, coefs)))
        mx = int(sum(a * (ub[c] if a > 0 else lb[c]) for c, a in zip(cols, coefs)))
        kind = int(rng.integers(0, 3))
        if planted:
            act = int(sum(a * x[c] for c, a in zip(cols, coefs)))
            lo_side = act - int(rng.integers(0, (mx - mn) // 8 + 1))
            hi_side = act + int(rng.integers(0, (mx - mn) // 8 + 1))
        else:
            lo_side = int(rng.integers(mn, mx + 1))
            hi_side = int(rng.integers(lo_side, mx + 1))
        lo = float(lo_side) if kind != 1 else -INF
        hi = float(hi_side) if kind != 0 else INF
        rows.append(([(int(c), float(a)) for c, a in zip(cols, coefs)], lo, hi))
    return make_instance(rows, lb, ub, [1] * ncols), rows


def supported(box, rows):
    """Every bound value of every column is supported by every row (exact integers)."""
    for entries, lhs, rhs in rows:
        lo = [a * (box[c][0] if a > 0 else box[c][1]) for c, a in entries]
        hi = [a * (box[c][1] if a > 0 else box[c][0]) for c, a in entries]
        smin, smax = sum(lo), sum(hi)
        if not entries:
            if smin > rhs or smax < lhs:
                return False
            continue
        for q, (c, a) in enumerate(entries):
            minres, maxres = smin - lo[q], smax - hi[q]
            for v in box[c]:
                if a * v + minres > rhs or a * v + maxres < lhs:
                    return False
    return True


def greatest_supported_box(lb, ub, rows):
    """Join of all supported boxes inside [lb, ub]; None if there is none."""
    n = len(lb)
    ivals = [[(l, u) for l in range(int(lb[j]), int(ub[j]) + 1) for u in range(l, int(ub[j]) + 1)]
             for j in range(n)]
    best = None
    for box in itertools.product(*ivals):
        if supported(box, rows):
            if best is None:
                best = [list(b) for b in box]
            else:
                for j in range(n):
                    best[j][0] = min(best[j][0], box[j][0])
                    best[j][1] = max(best[j][1], box[j][1])
    return best


@pytest.mark.parametrize("seed", range(60))
def test_greatest_fixpoint_by_enumeration(seed):
    rng = np.random.default_rng(5000 + seed)
    inst, rows = integer_instance(rng, nrows=int(rng.integers(1, 5)), ncols=int(rng.integers(2, 5)), maxlen=4)
    g = greatest_supported_box(inst.lb, inst.ub, rows)
    om = orc.OMatrix.from_instance(inst)
    for fn in (orc.propagate_jacobi, orc.propagate_scip):
        lb, ub, r = fn(om, inst.lb, inst.ub, orc.params(**BS0))
        if g is None:
            assert r["cutoff"], (fn.__name__, seed)
        else:
            assert not r["cutoff"], (fn.__name__, seed)
            assert np.array_equal(lb, [b[0] for b in g]) and np.array_equal(ub, [b[1] for b in g]), \
                (fn.__name__, seed, lb, ub, g)


def test_enumeration_sees_an_unpropagated_box():
    # the checker rejects the start box of a system that propagation tightens (x0 + x1 <= 1 on
    # [0, 2] x [1, 2] forces x0 <= 0 and x1 <= 1), so agreement above is not vacuous
    rows = [([(0, 1.0), (1, 1.0)], -INF, 1.0)]
    assert not supported(((0, 2), (1, 2)), rows)
    assert greatest_supported_box([0, 1], [2, 2], rows) == [[0, 0], [1, 1]]


@pytest.mark.parametrize("seed", range(16))
def test_general_integer_schedules_agree_with_boundstreps_zero(seed):
    # larger systems with bounds up to 1000, as the integer columns of the mixed generator:
    # no enumeration, but the fixpoint is unique, so both schedules must agree exactly
    rng = np.random.default_rng(6000 + seed)
    inst, _ = integer_instance(rng, nrows=300, ncols=200, maxlen=8, coefrange=20, boundrange=0, ubmax=1000,
                               planted=True)
    om = orc.OMatrix.from_instance(inst)
    jl, ju, jr = orc.propagate_jacobi(om, inst.lb, inst.ub, orc.params(**BS0))
    gl, gu, gr = orc.propagate_scip(om, inst.lb, inst.ub, orc.params(**BS0))
    assert not jr["cutoff"] and not gr["cutoff"]
    assert np.array_equal(jl, gl) and np.array_equal(ju, gu)
    assert jr["nchgbds"] > 0 and jr["rounds"] > 2
