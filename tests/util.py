"""Helpers shared by the tests: hand-built instances and random small instances."""
import numpy as np

import scip_amd as sa

INF = 1e20


def make_instance(rows, lb, ub, integral):
    """rows: list of (entries, lhs, rhs) with entries a list of (col, coef)."""
    rowptr = [0]
    colidx, vals, lhs, rhs = [], [], [], []
    for entries, lo, hi in rows:
        for c, a in sorted(entries):
            colidx.append(c)
            vals.append(a)
        rowptr.append(len(colidx))
        lhs.append(lo)
        rhs.append(hi)
    return sa.Instance(rowptr=np.array(rowptr, dtype=np.int64), colidx=np.array(colidx, dtype=np.int32),
                       vals=np.array(vals, dtype=np.float64), lhs=np.array(lhs, dtype=np.float64),
                       rhs=np.array(rhs, dtype=np.float64), integral=np.array(integral, dtype=np.uint8),
                       lb=np.array(lb, dtype=np.float64), ub=np.array(ub, dtype=np.float64))


def random_integer_instance(rng, nrows, ncols, maxlen, coefrange=5, boundrange=3, sidejitter=0.5):
    """Small pure-integer instance with random ranged/one-sided rows (for brute force)."""
    rows = []
    lb = rng.integers(-boundrange, 1, size=ncols).astype(float)
    ub = lb + rng.integers(0, boundrange + 1, size=ncols)
    for _ in range(nrows):
        L = int(rng.integers(1, maxlen + 1))
        cols = rng.choice(ncols, size=min(L, ncols), replace=False)
        coefs = rng.integers(-coefrange, coefrange + 1, size=len(cols))
        coefs[coefs == 0] = 1
        mn = sum(a * (lb[c] if a > 0 else ub[c]) for c, a in zip(cols, coefs))
        mx = sum(a * (ub[c] if a > 0 else lb[c]) for c, a in zip(cols, coefs))
        kind = rng.integers(0, 3)
        mid = mn + (mx - mn) * rng.uniform(0.2, 0.8)
        span = (mx - mn) * sidejitter * rng.uniform(0.1, 1.0)
        lo = mid - span if kind != 1 else -INF
        hi = mid + span if kind != 0 else INF
        rows.append(([(int(c), float(a)) for c, a in zip(cols, coefs)], lo, hi))
    return make_instance(rows, lb, ub, [1] * ncols)


def random_binary_instance(rng, nrows, ncols, maxlen, coefmax=20, hidden=True, point=False):
    """Pure-binary rows with random sides tight enough to propagate; with hidden=True every
    row is satisfied by a planted 0/1 point, so the system is feasible.  point=True returns
    (instance, planted point)."""
    rows = []
    xs = rng.integers(0, 2, size=ncols).astype(float)
    for _ in range(nrows):
        L = int(rng.integers(2, maxlen + 1))
        cols = rng.choice(ncols, size=min(L, ncols), replace=False)
        coefs = rng.integers(1, coefmax + 1, size=len(cols)).astype(float)
        coefs *= rng.choice([-1.0, 1.0], size=len(cols))
        pos = coefs[coefs > 0].sum()
        neg = coefs[coefs < 0].sum()
        side = neg + (pos - neg) * rng.uniform(0.05, 0.95)
        packing = bool(rng.integers(0, 2))
        if hidden:
            act = float(coefs @ xs[cols])
            slack = float(rng.integers(0, max(2, int(0.3 * (pos - neg)))))
            side = act + (slack if packing else -slack)
        if packing:
            rows.append(([(int(c), float(a)) for c, a in zip(cols, coefs)], -INF, side))
        else:
            rows.append(([(int(c), float(a)) for c, a in zip(cols, coefs)], side, INF))
    inst = make_instance(rows, [0.0] * ncols, [1.0] * ncols, [1] * ncols)
    return (inst, xs) if point else inst


def random_fixings(rng, ncols, nvec, maxfix):
    ptr = [0]
    cols, vals = [], []
    for _ in range(nvec):
        f = int(rng.integers(1, maxfix + 1))
        cs = rng.choice(ncols, size=f, replace=False)
        cols.extend(int(c) for c in cs)
        vals.extend(float(rng.integers(0, 2)) for _ in cs)
        ptr.append(len(cols))
    return np.array(ptr, dtype=np.int64), np.array(cols, dtype=np.int32), np.array(vals, dtype=np.float64)


def apply_fixings(lb, ub, ptr, cols, vals, v):
    lb = lb.copy()
    ub = ub.copy()
    for q in range(ptr[v], ptr[v + 1]):
        lb[cols[q]] = vals[q]
        ub[cols[q]] = vals[q]
    return lb, ub


def same_bounds(a, b):
    """Numeric equality of bound arrays (-0.0 == +0.0)."""
    return np.array_equal(np.asarray(a), np.asarray(b))


def planted_integer_instance(rng, nrows, ncols, maxlen, coefrange=6, boundrange=6):
    """Small general-integer instance with sides tight around a planted integer point (feasible,
    with several propagation rounds)."""
    lb = -rng.integers(0, boundrange, size=ncols).astype(float)
    ub = lb + rng.integers(1, 2 * boundrange, size=ncols)
    xs = np.array([rng.integers(int(l), int(u) + 1) for l, u in zip(lb, ub)], dtype=float)
    rows = []
    for _ in range(nrows):
        L = int(rng.integers(2, maxlen + 1))
        cols = rng.choice(ncols, size=min(L, ncols), replace=False)
        coefs = rng.integers(1, coefrange + 1, size=len(cols)) * rng.choice([-1, 1], size=len(cols))
        act = float(coefs @ xs[cols])
        kind = rng.integers(0, 3)
        lo = act - float(rng.integers(0, 3)) if kind != 1 else -INF
        hi = act + float(rng.integers(0, 3)) if kind != 0 else INF
        rows.append(([(int(c), float(a)) for c, a in zip(cols, coefs)], lo, hi))
    return make_instance(rows, lb, ub, [1] * ncols)


def late_cut_instance(inst, rounds, cols, bounds, vals):
    """inst plus one row x_k + x_m <= 1 over two binary columns that its propagation lifts to 1
    in the same round r >= 2 (from a reason history: rounds, cols, bounds, vals), so the system
    is found infeasible by that row in round r + 1, after r rounds of history; None if no such
    pair exists."""
    byround = {}
    for r, j, b, v in zip(rounds, cols, bounds, vals):
        if r >= 2 and b == 0 and v == 1.0 and inst.lb[j] == 0.0 and inst.ub[j] == 1.0:
            byround.setdefault(int(r), []).append(int(j))
    for r in sorted(byround):
        js = byround[r]
        if len(js) >= 2:
            rows = []
            for i in range(inst.nrows):
                a, e = inst.rowptr[i], inst.rowptr[i + 1]
                rows.append((list(zip(inst.colidx[a:e].tolist(), inst.vals[a:e].tolist())), inst.lhs[i], inst.rhs[i]))
            rows.append(([(js[0], 1.0), (js[1], 1.0)], -INF, 1.0))
            return make_instance(rows, inst.lb, inst.ub, inst.integral)
    return None
