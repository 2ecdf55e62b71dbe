"""Pins the oracle's propagation semantics on the CPU.

Closed-form single-row cases (textbook bound tightening with SCIP's acceptance rules),
brute-force soundness on tiny integer instances, agreement of the Jacobi twin with the
SCIP-order restatement where the fixpoint is unique (pure binary), the mutual-fixpoint
property where it is not, and worklist == full sweep.
"""
import itertools

import numpy as np
import pytest

import scip_amd as sa
from oracle import oracle as orc
from tests.util import (INF, apply_fixings, make_instance, random_binary_instance, random_fixings,
                        random_integer_instance, same_bounds)


def both(inst, lb=None, ub=None, **kw):
    om = orc.OMatrix.from_instance(inst)
    lb = inst.lb if lb is None else lb
    ub = inst.ub if ub is None else ub
    prm = orc.params(**kw)
    return orc.propagate_jacobi(om, lb, ub, prm), orc.propagate_scip(om, lb, ub, prm)


# --- closed forms ------------------------------------------------------------------------------
def test_binary_packing_fixes_complement():
    # x + y <= 1 with x fixed to 1 forces y = 0
    inst = make_instance([([(0, 1.0), (1, 1.0)], -INF, 1.0)], [1.0, 0.0], [1.0, 1.0], [1, 1])
    for lb, ub, r in both(inst):
        assert r["cutoff"] == 0 and ub[1] == 0.0 and lb[0] == 1.0


def test_continuous_upper_bound_from_rhs():
    # x + y <= 10, x in [0,100], y in [5,8]: x <= 5 (95% of the domain removed, accepted)
    inst = make_instance([([(0, 1.0), (1, 1.0)], -INF, 10.0)], [0.0, 5.0], [100.0, 8.0], [0, 0])
    for lb, ub, r in both(inst):
        assert ub[0] == 5.0 and r["cutoff"] == 0


def test_boundstreps_rejects_small_improvement():
    # x + y <= 10, x in [0,10.2], y in [0,10]: candidate ub 10 improves by < 5% of 10.2 -> rejected
    inst = make_instance([([(0, 1.0), (1, 1.0)], -INF, 10.0)], [0.0, 0.0], [10.2, 10.0], [0, 0])
    for lb, ub, r in both(inst):
        assert ub[0] == 10.2 and r["nchgbds"] == 0
    # with boundstreps 0 the same change is accepted
    for lb, ub, r in both(inst, boundstreps=0.0):
        assert ub[0] == 10.0


def test_integer_rounding_with_feastol():
    # 2x <= 5 -> x <= 2 ;  3y <= 5.9999999 -> y <= 2 (floor(1.99999997 + 1e-6)), not 1
    inst = make_instance([([(0, 2.0)], -INF, 5.0), ([(1, 3.0)], -INF, 5.9999999)], [0.0, 0.0], [10.0, 10.0], [1, 1])
    for lb, ub, r in both(inst):
        assert ub[0] == 2.0 and ub[1] == 2.0


def test_lower_bound_from_lhs_negative_coefficient():
    # -x + y >= 3, y in [0,4], x in [0,10] (integer): x <= y - 3 <= 1
    inst = make_instance([([(0, -1.0), (1, 1.0)], 3.0, INF)], [0.0, 0.0], [10.0, 4.0], [1, 1])
    for lb, ub, r in both(inst):
        assert ub[0] == 1.0 and lb[1] == 3.0


def test_singleton_row_forces_small_change():
    # x <= 9.9 as a one-variable row: force=TRUE bypasses the 5% rule (cons_linear.c:7029)
    inst = make_instance([([(0, 1.0)], -INF, 9.9)], [0.0], [10.0], [0])
    for lb, ub, r in both(inst):
        assert ub[0] == 9.9
    # the same bound from a two-variable row is too small a change
    inst2 = make_instance([([(0, 1.0), (1, 1.0)], -INF, 9.9)], [0.0, 0.0], [10.0, 0.0], [0, 0])
    for lb, ub, r in both(inst2):
        assert ub[0] == 10.0


def test_infeasible_row_is_cutoff():
    inst = make_instance([([(0, 1.0), (1, 1.0)], 3.0, INF)], [0.0, 0.0], [1.0, 1.0], [1, 1])
    for lb, ub, r in both(inst):
        assert r["cutoff"] == 1


def test_infinite_bound_blocks_tightening():
    # x + y <= 10, y in [-inf, 5]: the min activity has one -inf contribution, so no bound on x;
    # y's residual is x's min 0 -> candidate ub 10, not better than 5
    inst = make_instance([([(0, 1.0), (1, 1.0)], -INF, 10.0)], [0.0, -INF], [100.0, 5.0], [0, 0])
    for lb, ub, r in both(inst):
        assert ub[0] == 100.0 and ub[1] == 5.0 and r["nchgbds"] == 0


def test_huge_contribution_relaxes_activity():
    # 1e10 * z with z in [0, 1e6] gives a huge (>= 1e15) max contribution; x + 1e10 z <= 1e16
    # cannot bound x (max side) but z's bound comes from x's min residual 0: z <= 1e6 already.
    inst = make_instance([([(0, 1.0), (1, 1e10)], -INF, 1e16)], [0.0, 0.0], [5.0, 1e6], [0, 0])
    om = orc.OMatrix.from_instance(inst)
    mn, mx, cnt, mad = orc.row_activities(om, inst.lb, inst.ub)
    assert cnt[0].tolist() == [0, 0, 0, 0, 0, 0, 1, 0]
    assert mx[0] == 5.0 and mn[0] == 0.0


def test_maxactdelta_binary_rule_and_general_rule():
    bin_inst = make_instance([([(0, 3.0), (1, -7.0), (2, 2.0)], -INF, 100.0)], [0.0, 0.0, 1.0], [1.0, 1.0, 1.0],
                             [1, 1, 1])
    om = orc.OMatrix.from_instance(bin_inst)
    assert orc.row_activities(om, bin_inst.lb, bin_inst.ub)[3][0] == 7.0   # fixed var 2 ignored
    gen = make_instance([([(0, 3.0), (1, -7.0)], -INF, 100.0)], [0.0, -2.0], [4.0, 1.0], [1, 0])
    om = orc.OMatrix.from_instance(gen)
    assert orc.row_activities(om, gen.lb, gen.ub)[3][0] == 21.0
    inf = make_instance([([(0, 3.0), (1, -7.0)], -INF, 100.0)], [0.0, -INF], [4.0, 1.0], [1, 0])
    om = orc.OMatrix.from_instance(inf)
    assert orc.row_activities(om, inf.lb, inf.ub)[3][0] == 1e20


def test_empty_row_feasibility():
    ok = make_instance([([], -1.0, 1.0)], [0.0], [1.0], [1])
    bad = make_instance([([], 1.0, 2.0)], [0.0], [1.0], [1])
    for lb, ub, r in both(ok):
        assert r["cutoff"] == 0
    for lb, ub, r in both(bad):
        assert r["cutoff"] == 1


# --- brute force -------------------------------------------------------------------------------
def _feasible_points(inst, lb, ub):
    cols = range(inst.ncols)
    ranges = [range(int(lb[j]), int(ub[j]) + 1) for j in cols]
    pts = []
    for x in itertools.product(*ranges):
        ok = True
        for i in range(inst.nrows):
            s, e = inst.rowptr[i], inst.rowptr[i + 1]
            act = sum(inst.vals[k] * x[inst.colidx[k]] for k in range(s, e))
            if act < inst.lhs[i] - 1e-9 or act > inst.rhs[i] + 1e-9:
                ok = False
                break
        if ok:
            pts.append(x)
    return pts


@pytest.mark.parametrize("seed", range(40))
def test_bruteforce_soundness(seed):
    rng = np.random.default_rng(seed)
    inst = random_integer_instance(rng, nrows=int(rng.integers(1, 5)), ncols=int(rng.integers(2, 5)), maxlen=4)
    pts = _feasible_points(inst, inst.lb, inst.ub)
    for lb, ub, r in both(inst):
        if r["cutoff"]:
            assert pts == [], "cutoff declared on a feasible system"
        else:
            for x in pts:
                assert all(lb[j] <= x[j] <= ub[j] for j in range(inst.ncols)), "propagation removed a solution"


@pytest.mark.parametrize("seed", range(20))
def test_bruteforce_binary_complete_when_no_point(seed):
    # on pure-binary single-row systems propagation detects infeasibility exactly
    rng = np.random.default_rng(100 + seed)
    inst = random_binary_instance(rng, nrows=1, ncols=5, maxlen=5)
    pts = _feasible_points(inst, inst.lb, inst.ub)
    for lb, ub, r in both(inst):
        if not pts:
            assert r["cutoff"] == 1


# --- schedules ---------------------------------------------------------------------------------
@pytest.mark.parametrize("seed", range(12))
def test_binary_fixpoint_unique_jacobi_equals_scip_order(seed):
    rng = np.random.default_rng(200 + seed)
    inst = random_binary_instance(rng, nrows=300, ncols=200, maxlen=8)
    om = orc.OMatrix.from_instance(inst)
    rl, ru, rr = orc.propagate_jacobi(om, inst.lb, inst.ub)
    assert rr["cutoff"] == 0
    free = np.nonzero(rl < ru)[0]
    ptr, cols, vals = random_fixings(rng, len(free), 12, 4)
    cols = free[cols].astype(np.int32)
    nontrivial = 0
    for v in range(12):
        lb, ub = apply_fixings(rl, ru, ptr, cols, vals, v)
        jl, ju, jr = orc.propagate_jacobi(om, lb, ub)
        gl, gu, gr = orc.propagate_scip(om, lb, ub)
        assert jr["cutoff"] == gr["cutoff"]
        if not jr["cutoff"]:
            assert same_bounds(jl, gl) and same_bounds(ju, gu)
        nontrivial += jr["cutoff"] or jr["nchgbds"] > 0
    assert nontrivial > 0


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_covering_packing_generator_schedules_agree(seed):
    inst = sa.gen_covering_packing(400, 1000, 10, seed=seed)
    om = orc.OMatrix.from_instance(inst)
    ptr, cols, vals = sa.gen_node_fixings(inst, 16, 1, 50, seed=seed)
    for v in range(16):
        lb, ub = apply_fixings(inst.lb, inst.ub, ptr, cols, vals, v)
        jl, ju, jr = orc.propagate_jacobi(om, lb, ub)
        gl, gu, gr = orc.propagate_scip(om, lb, ub)
        assert jr["cutoff"] == gr["cutoff"]
        if not jr["cutoff"]:
            assert same_bounds(jl, gl) and same_bounds(ju, gu)


@pytest.mark.parametrize("seed", range(4))
def test_worklist_equals_full_sweep(seed):
    inst = sa.gen_mixed(2000, 2000, seed=seed)
    om = orc.OMatrix.from_instance(inst)
    a = orc.propagate_jacobi(om, inst.lb, inst.ub)
    b = orc.propagate_jacobi(om, inst.lb, inst.ub, full_sweep=True)
    assert same_bounds(a[0], b[0]) and same_bounds(a[1], b[1])
    assert a[2]["cutoff"] == b[2]["cutoff"] and a[2]["nchgbds"] == b[2]["nchgbds"]


@pytest.mark.parametrize("seed", range(4))
def test_mutual_fixpoint_mixed(seed):
    """Where the 5% filter makes the fixpoint order dependent, each schedule's fixpoint is a
    fixpoint of the other, up to changes that SCIP itself finds when restarted from its own
    fixpoint (stale incremental activities / maxactdelta)."""
    inst = sa.gen_mixed(2000, 2000, seed=10 + seed)
    om = orc.OMatrix.from_instance(inst)
    jl, ju, jr = orc.propagate_jacobi(om, inst.lb, inst.ub)
    gl, gu, gr = orc.propagate_scip(om, inst.lb, inst.ub)
    a, b, r1 = orc.propagate_scip(om, jl, ju)
    assert r1["nchgbds"] == 0 and r1["cutoff"] == 0
    c, d, r2 = orc.propagate_jacobi(om, gl, gu)
    e, f, r3 = orc.propagate_scip(om, gl, gu)
    assert r2["cutoff"] == 0
    changed_j = set(np.nonzero((c != gl) | (d != gu))[0])
    changed_s = set(np.nonzero((e != gl) | (f != gu))[0])
    assert changed_j <= changed_s


def test_integer_bounds_mostly_agree_across_schedules():
    inst = sa.gen_mixed(3000, 3000, seed=5)
    om = orc.OMatrix.from_instance(inst)
    jl, ju, _ = orc.propagate_jacobi(om, inst.lb, inst.ub)
    gl, gu, _ = orc.propagate_scip(om, inst.lb, inst.ub)
    it = inst.integral.astype(bool)
    agree = np.mean((jl[it] == gl[it]) & (ju[it] == gu[it]))
    assert agree > 0.9


@pytest.mark.parametrize("seed", range(4))
def test_notighten_only_tests_rows(seed):
    # consPropLinear with tightenbounds FALSE (cons_linear.c:16147-16161): propagateCons only
    # runs its final test (:7717-7742), so no bound moves and the verdict is "some row's
    # activity range misses its sides", with exact integer activities on this instance
    rng = np.random.default_rng(900 + seed)
    inst = random_integer_instance(rng, nrows=60, ncols=30, maxlen=6)
    if seed % 2:
        # a row whose rhs lies below its minimal activity
        s, e = inst.rowptr[7], inst.rowptr[8]
        a = inst.vals[s:e]
        inst.rhs[7] = float(a @ np.where(a > 0, inst.lb[inst.colidx[s:e]], inst.ub[inst.colidx[s:e]])) - 1.0
        inst.lhs[7] = -INF
    om = orc.OMatrix.from_instance(inst)
    expect_cut = False
    for i in range(inst.nrows):
        s, e = inst.rowptr[i], inst.rowptr[i + 1]
        a = inst.vals[s:e]
        lo = np.where(a > 0, inst.lb[inst.colidx[s:e]], inst.ub[inst.colidx[s:e]])
        hi = np.where(a > 0, inst.ub[inst.colidx[s:e]], inst.lb[inst.colidx[s:e]])
        mn, mx = float(a @ lo), float(a @ hi)
        expect_cut |= mn > inst.rhs[i] + 1e-6 * max(1.0, abs(mn), abs(inst.rhs[i])) or \
            mx < inst.lhs[i] - 1e-6 * max(1.0, abs(mx), abs(inst.lhs[i]))
    for fn in (orc.propagate_jacobi, orc.propagate_scip):
        for ranged in (0, 1):
            lb, ub, r = fn(om, inst.lb, inst.ub, orc.params(notighten=1, rangedrows=ranged))
            assert np.array_equal(lb, inst.lb) and np.array_equal(ub, inst.ub)
            assert r["nchgbds"] == 0 and r["rounds"] <= 1
            assert bool(r["cutoff"]) == expect_cut == bool(seed % 2)
    # with tightening the same instances do move bounds or cut off
    lb, ub, r = orc.propagate_jacobi(om, inst.lb, inst.ub)
    assert r["cutoff"] or r["nchgbds"] > 0 or not expect_cut


# consdataGetReliableResidualActivity (cons_linear.c:2580-2657) forced by cancellation: the
# row x0 + x1 - x2 + a*y with x0 = x2 = 1e14 and x1 = 0.3 fixed, y in [1, 2] continuous, |a| =
# 1e7 (maxactdelta 1e7: the hard path).  y's residual activity is 0.3 in double-double, a
# 1/3.3e7 fraction of the activity, so SCIPisUpdateUnreliable (set.c:8032, recomputefac 1e6)
# holds and the residual is recomputed as a plain double sum in row order:
# fl(fl(1e14 + 0.3) - 1e14) = 0.296875 (the spacing of doubles near 1e14 is 1/64).  The bound
# from the recomputed residual differs from the double-double one ((1.5e7 - 0.3) / 1e7 =
# 1.49999997), so the expected values below pin that the recompute ran and what it summed.
RELIABLE = {
    # name: (coefficient of y, lhs, rhs, bound of y, expected value)
    "rhs_ub": (1e7, -INF, 1.5e7, "ub", (1.5e7 - 0.296875) / 1e7),      # = 1.4999999703125
    "lhs_lb": (1e7, 1.5e7, INF, "lb", (1.5e7 - 0.296875) / 1e7),
    "lhs_ub_neg": (-1e7, -1.5e7, INF, "ub", (-1.5e7 - 0.296875) / -1e7),  # = 1.5000000296875
    "rhs_lb_neg": (-1e7, -INF, -1.5e7, "lb", (-1.5e7 - 0.296875) / -1e7),
}


def reliable_instance(name):
    a, lhs, rhs, _, _ = RELIABLE[name]
    return make_instance([([(0, 1.0), (1, 1.0), (2, -1.0), (3, a)], lhs, rhs)],
                         [1e14, 0.3, 1e14, 1.0], [1e14, 0.3, 1e14, 2.0], [0, 0, 0, 0])


def test_reliable_residual_constants():
    assert np.float64(1e14) + np.float64(0.3) - np.float64(1e14) == 0.296875
    assert (1.5e7 - 0.296875) / 1e7 == 1.4999999703125 != (1.5e7 - 0.3) / 1e7
    assert (-1.5e7 - 0.296875) / -1e7 == 1.5000000296875 != (-1.5e7 - 0.3) / -1e7


@pytest.mark.parametrize("name", sorted(RELIABLE))
def test_reliable_residual_recompute(name):
    _, _, _, side, expect = RELIABLE[name]
    inst = reliable_instance(name)
    om = orc.OMatrix.from_instance(inst)
    for fn in (orc.propagate_jacobi, orc.propagate_scip):
        lb, ub, r = fn(om, inst.lb, inst.ub)
        assert not r["cutoff"] and r["nchgbds"] == 1 and r["nrecompute"] == 1, (fn.__name__, r)
        got = lb[3] if side == "lb" else ub[3]
        assert got == expect, (fn.__name__, got, expect)
        other = ub[3] if side == "lb" else lb[3]
        assert other == (2.0 if side == "lb" else 1.0)
