"""Exact checks of propagation reasons, independent of the propagation code: with rational
arithmetic (fractions.Fraction of the stored doubles), a reason row and the bounds in force
when a change was made must imply the change; an infeasibility certificate must prove
infeasibility.  Used by the oracle's and the engine's reason tests."""
import itertools
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


def row_points(inst, i, lb, ub, maxpoints=200000):
    """All integer assignments of row i's columns within lb/ub that satisfy the row (rows of
    integral columns with finite, small domains; None when that does not hold)."""
    ent = row_entries(inst, i)
    doms = []
    size = 1
    for k, _ in ent:
        if not inst.integral[k] or abs(lb[k]) >= INF or abs(ub[k]) >= INF:
            return None
        lo, hi = math.ceil(lb[k] - FEASTOL), math.floor(ub[k] + FEASTOL)
        size *= max(hi - lo + 1, 0)
        if size > maxpoints:
            return None
        doms.append(range(lo, hi + 1))
    cols = [k for k, _ in ent]
    a = np.array([v for _, v in ent])
    pts = np.array(list(itertools.product(*doms)), dtype=float).reshape(-1, len(ent))
    act = pts @ a
    ok = (act >= inst.lhs[i] - 1e-9) & (act <= inst.rhs[i] + 1e-9)
    return cols, pts[ok]


def check_ranged_implied(inst, i, j, bound, val, lb, ub):
    """A ranged-row deduction: every integer point of row i within the bounds (column j's own
    bounds before the change included) satisfies the new bound.  Returns False when the row
    cannot be enumerated."""
    r = row_points(inst, i, lb, ub)
    if r is None:
        return False
    cols, pts = r
    x = pts[:, cols.index(j)]
    assert np.all(x >= val - 1e-9) if bound == 0 else np.all(x <= val + 1e-9), (i, j, bound, val)
    return True


def replay_history(inst, lb0, ub0, hist):
    """Walks the history in order; each change must be implied by its row and the bounds of
    the rounds before it (the bounds its row sweep read): rule 0 by the residual activity in
    exact arithmetic, rule 1 (ranged-row test) by enumeration of the row's integer points when
    the row is small enough.  Returns the bounds at the start of every round, the final
    bounds, and the number of ranged entries that were enumerated."""
    lb, ub = np.array(lb0, dtype=float), np.array(ub0, dtype=float)
    rules = hist.get("rule")
    starts = {}
    last = 0
    enumerated = 0
    for q, (r, j, b, v, i) in enumerate(zip(hist["round"], hist["col"], hist["bound"], hist["val"], hist["row"])):
        r = int(r)
        assert r >= last, "history out of round order"
        if r not in starts:
            starts[r] = (lb.copy(), ub.copy())
        last = r
        start_lb, start_ub = starts[r]
        if rules is not None and rules[q]:
            enumerated += int(check_ranged_implied(inst, int(i), int(j), int(b), float(v), start_lb, start_ub))
        else:
            check_implied(inst, int(i), int(j), int(b), float(v), start_lb, start_ub)
        if b == 0:
            assert v > lb[j]
            lb[j] = v
        else:
            assert v < ub[j]
            ub[j] = v
    replay_history.enumerated = enumerated
    return starts, lb, ub


def check_conflict(inst, conflict, lb, ub):
    """lb/ub: the bounds at the start of the conflict's round.  The certificate must prove
    infeasibility: exactly for activity arguments, by enumeration for the ranged-row test
    (returns False when that row cannot be enumerated)."""
    if conflict["kind"] == 1:
        i = conflict["row"]
        ent = row_entries(inst, i)
        if conflict["side"] == 2:
            r = row_points(inst, i, lb, ub)
            if r is None:
                return False
            assert len(r[1]) == 0, i
        elif conflict["side"] == 0:
            act = residual(ent, lb, ub, -1, maximal=False)
            assert act is not None and act > Fraction(inst.rhs[i]), i
        else:
            act = residual(ent, lb, ub, -1, maximal=True)
            assert act is not None and act < Fraction(inst.lhs[i]), i
    elif conflict["kind"] == 2:
        j = conflict["col"]
        rules = conflict.get("rules", 0)

        def side(row, bound, rule):
            if row < 0:
                return float(lb[j] if bound == 0 else ub[j])
            if rule:
                r = row_points(inst, row, lb, ub)
                if r is None:
                    return None
                x = r[1][:, r[0].index(j)]
                if len(x) == 0:
                    return INF if bound == 0 else -INF
                return float(x.min() if bound == 0 else x.max())
            t = implied(inst, row, j, bound, lb, ub)
            assert t is not None
            return float(t)

        lo = side(conflict["rowlb"], 0, rules & 1)
        hi = side(conflict["rowub"], 1, (rules >> 1) & 1)
        if lo is None or hi is None:
            return False
        if inst.integral[j]:
            lo, hi = math.ceil(lo - FEASTOL), math.floor(hi + FEASTOL)
        assert lo > hi, (j, lo, hi)
    else:
        raise AssertionError("no certificate")
    return True
