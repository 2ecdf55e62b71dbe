"""Exact checks of propagation reasons, independent of the propagation code: with rational
arithmetic (fractions.Fraction of the stored doubles), a reason row and the bounds in force
when a change was made must imply the change; an infeasibility certificate must prove
infeasibility.  Used by the oracle's and the engine's reason tests."""
import math
from fractions import Fraction

import numpy as np

INF = 1e20
FEASTOL = 1e-6


def row_entries(inst, i):
    a, e = inst.rowptr[i], inst.rowptr[i + 1]
    return list(zip(inst.colidx[a:e].tolist(), inst.vals[a:e].tolist()))


def residual(entries, lb, ub, skip, maximal):
    """Exact minimal (maximal) activity of the row without column skip; None if infinite."""
    tot = Fraction(0)
    for k, a in entries:
        if k == skip:
            continue
        use = (ub[k] if a > 0 else lb[k]) if maximal else (lb[k] if a > 0 else ub[k])
        if abs(use) >= INF:
            return None
        tot += Fraction(a) * Fraction(use)
    return tot


def implied(inst, i, j, bound, lb, ub):
    """The bound of column j (0 lower, 1 upper) that row i implies from the other columns'
    bounds lb/ub, as a Fraction (None: the row implies nothing on that side)."""
    ent = row_entries(inst, i)
    a = dict(ent)[j]
    lhs, rhs = inst.lhs[i], inst.rhs[i]
    # lower bound: a > 0 from lhs with the maximal residual, a < 0 from rhs with the minimal
    use_lhs = (bound == 0) == (a > 0)
    side = lhs if use_lhs else rhs
    if abs(side) >= INF:
        return None
    res = residual(ent, lb, ub, j, maximal=use_lhs)
    if res is None:
        return None
    return (Fraction(side) - res) / Fraction(a)


def check_implied(inst, i, j, bound, val, lb, ub, tol=1e-9):
    """The accepted bound val is no stronger than what row i implies (integral columns:
    after SCIP's feasibility rounding)."""
    t = implied(inst, i, j, bound, lb, ub)
    assert t is not None, (i, j, bound)
    tf = float(t)
    slack = tol * max(1.0, abs(tf))
    if inst.integral[j]:
        tf = math.ceil(tf - FEASTOL - slack) if bound == 0 else math.floor(tf + FEASTOL + slack)
        slack = 0.0
    if bound == 0:
        assert val <= tf + slack, (i, j, val, tf)
    else:
        assert val >= tf - slack, (i, j, val, tf)


def replay_history(inst, lb0, ub0, hist):
    """Walks the history in order; each change must be implied by its row and the bounds of
    the rounds before it (the bounds its row sweep read).  Returns the bounds at the start
    of every round, and the final bounds."""
    lb, ub = np.array(lb0, dtype=float), np.array(ub0, dtype=float)
    starts = {}
    last = 0
    for r, j, b, v, i in zip(hist["round"], hist["col"], hist["bound"], hist["val"], hist["row"]):
        r = int(r)
        assert r >= last, "history out of round order"
        if r not in starts:
            starts[r] = (lb.copy(), ub.copy())
        last = r
        start_lb, start_ub = starts[r]
        check_implied(inst, int(i), int(j), int(b), float(v), start_lb, start_ub)
        if b == 0:
            assert v > lb[j]
            lb[j] = v
        else:
            assert v < ub[j]
            ub[j] = v
    return starts, lb, ub


def check_conflict(inst, conflict, lb, ub):
    """lb/ub: the bounds at the start of the conflict's round.  The certificate must prove
    infeasibility exactly."""
    if conflict["kind"] == 1:
        i = conflict["row"]
        ent = row_entries(inst, i)
        if conflict["side"] == 0:
            act = residual(ent, lb, ub, -1, maximal=False)
            assert act is not None and act > Fraction(inst.rhs[i]), i
        else:
            act = residual(ent, lb, ub, -1, maximal=True)
            assert act is not None and act < Fraction(inst.lhs[i]), i
    elif conflict["kind"] == 2:
        j = conflict["col"]
        lo = Fraction(lb[j]) if conflict["rowlb"] < 0 else implied(inst, conflict["rowlb"], j, 0, lb, ub)
        hi = Fraction(ub[j]) if conflict["rowub"] < 0 else implied(inst, conflict["rowub"], j, 1, lb, ub)
        assert lo is not None and hi is not None
        lo, hi = float(lo), float(hi)
        if inst.integral[j]:
            lo, hi = math.ceil(lo - FEASTOL), math.floor(hi + FEASTOL)
        assert lo > hi, (j, lo, hi)
    else:
        raise AssertionError("no certificate")
