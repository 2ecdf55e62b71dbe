"""Reasons of the oracle's Jacobi twin (oracle.propagate_reasons): the history reproduces the
propagation's bounds, every entry is implied by its row and the bounds of the earlier rounds,
and every cutoff certificate proves infeasibility -- all checked in exact rational arithmetic
(tests/reasons_check.py), so a wrong row, side or round fails."""
import numpy as np
import pytest

import scip_amd as sa
from oracle import oracle as orc
from tests.reasons_check import check_conflict, replay_history
from tests.util import (INF, late_cut_instance, make_instance, planted_integer_instance, random_binary_instance,
                        random_integer_instance)


def run_and_check(inst, prm=None):
    om = orc.OMatrix.from_instance(inst)
    r = orc.propagate_reasons(om, inst.lb, inst.ub, prm)
    jl, ju, jr = orc.propagate_jacobi(om, inst.lb, inst.ub, prm)
    assert r["result"]["cutoff"] == jr["cutoff"] and r["result"]["rounds"] == jr["rounds"]
    starts, lb, ub = replay_history(inst, inst.lb, inst.ub, r)
    if jr["cutoff"]:
        # the last round's start: every change of the earlier rounds applied
        r["certified"] = check_conflict(inst, r["conflict"], lb, ub)
        assert r["conflict"]["round"] == jr["rounds"]
    else:
        assert r["conflict"]["kind"] == 0
        assert np.array_equal(lb, jl) and np.array_equal(ub, ju)
        assert r["nhist"] == jr["nchgbds"]
    return r


@pytest.mark.parametrize("seed", range(8))
def test_binary_reasons(seed):
    inst = random_binary_instance(np.random.default_rng(900 + seed), 400, 300, 9, hidden=(seed % 2 == 0))
    run_and_check(inst)


@pytest.mark.parametrize("seed", range(8))
def test_integer_reasons(seed):
    inst = random_integer_instance(np.random.default_rng(950 + seed), 60, 40, 6, sidejitter=0.3)
    run_and_check(inst)


@pytest.mark.parametrize("seed", range(4))
def test_planted_integer_reasons(seed):
    r = run_and_check(planted_integer_instance(np.random.default_rng(950 + seed), 400, 300, 6))
    assert r["nhist"] > 100 and r["result"]["rounds"] > 5


@pytest.mark.parametrize("seed", range(3))
def test_row_certificate_after_rounds_of_history(seed):
    inst = random_binary_instance(np.random.default_rng(900 + seed), 1500, 1000, 10, hidden=True)
    o = orc.propagate_reasons(orc.OMatrix.from_instance(inst), inst.lb, inst.ub)
    t = late_cut_instance(inst, o["round"], o["col"], o["bound"], o["val"])
    r = run_and_check(t)
    assert r["conflict"]["kind"] == 1 and r["conflict"]["round"] >= 3 and r["nhist"] > 0


@pytest.mark.parametrize("seed", range(4))
def test_mixed_reasons(seed):
    inst = sa.gen_mixed(600, 600, seed=40 + seed, planted=bool(seed % 2))
    run_and_check(inst)


def test_known_history_with_a_forced_row():
    # x0 integer [0,5], x1, x2 binary, x3 continuous [0,10]
    # r0: 2 x0 <= 3 (one variable: forced);  r1: x1 + x2 >= 2;  r2: x3 + 4 x0 <= 6;  r3: x3 - 5 x1 >= 0
    inst = make_instance([([(0, 2.0)], -INF, 3.0),
                          ([(1, 1.0), (2, 1.0)], 2.0, INF),
                          ([(3, 1.0), (0, 4.0)], -INF, 6.0),
                          ([(3, 1.0), (1, -5.0)], 0.0, INF)],
                         [0.0, 0.0, 0.0, 0.0], [5.0, 1.0, 1.0, 10.0], [1, 1, 1, 0])
    r = run_and_check(inst)
    got = list(zip(r["round"].tolist(), r["col"].tolist(), r["bound"].tolist(), r["val"].tolist(), r["row"].tolist()))
    # round 1: x0 <= 1 (r0 and r2 both give 1; the forced candidate of r0 wins), x1, x2 >= 1
    # (r1), x3 <= 6 (r2); round 2, after x1 = 1: x3 >= 5 (r3); round 3, after x3 >= 5:
    # 4 x0 <= 1, so x0 <= 0 (r2)
    assert got == [(1, 0, 1, 1.0, 0), (1, 1, 0, 1.0, 1), (1, 2, 0, 1.0, 1), (1, 3, 1, 6.0, 2), (2, 3, 0, 5.0, 3),
                   (3, 0, 1, 0.0, 2)]


def test_known_certificates():
    # a row whose maximal activity misses its lhs in round 1
    inst = make_instance([([(0, 1.0), (1, 1.0)], 3.0, INF), ([(1, 1.0), (2, 1.0)], -INF, 1.0)],
                         [0.0] * 3, [1.0] * 3, [1] * 3)
    r = run_and_check(inst)
    assert r["conflict"] == dict(kind=1, round=1, row=0, side=1, col=-1, rowlb=-1, rowub=-1, rules=0)
    # x + y >= 2 and x + z <= 0.5: row 0 lifts x to 1 while row 1 pushes it to 0 in the same round
    inst = make_instance([([(0, 1.0), (1, 1.0)], 2.0, INF), ([(0, 1.0), (2, 1.0)], -INF, 0.5)],
                         [0.0] * 3, [1.0] * 3, [1] * 3)
    r = run_and_check(inst)
    assert r["conflict"] == dict(kind=2, round=1, row=-1, side=0, col=0, rowlb=0, rowub=1, rules=0)
    assert r["nhist"] == 0


def test_certificate_kinds_occur_in_random_cases():
    kinds = set()
    for seed in range(8):
        for inst in (random_binary_instance(np.random.default_rng(900 + seed), 400, 300, 9, hidden=(seed % 2 == 0)),
                     random_integer_instance(np.random.default_rng(950 + seed), 60, 40, 6, sidejitter=0.3)):
            kinds.add(orc.propagate_reasons(orc.OMatrix.from_instance(inst), inst.lb, inst.ub)["conflict"]["kind"])
    assert {0, 2} <= kinds, kinds


def test_checks_reject_wrong_reasons():
    # the exact checks are not vacuous: the other side of a row certificate, a swapped column
    # certificate and a history entry credited to another row all fail
    inst = make_instance([([(0, 1.0), (1, 1.0)], 3.0, INF), ([(1, 1.0), (2, 1.0)], -INF, 1.0)],
                         [0.0] * 3, [1.0] * 3, [1] * 3)
    with pytest.raises(AssertionError):
        check_conflict(inst, dict(kind=1, row=0, side=0), inst.lb, inst.ub)
    with pytest.raises(AssertionError):
        check_conflict(inst, dict(kind=1, row=1, side=1), inst.lb, inst.ub)
    inst = make_instance([([(0, 1.0), (1, 1.0)], 2.0, INF), ([(0, 1.0), (2, 1.0)], -INF, 0.5)],
                         [0.0] * 3, [1.0] * 3, [1] * 3)
    with pytest.raises(AssertionError):
        check_conflict(inst, dict(kind=2, col=0, rowlb=-1, rowub=1), inst.lb, inst.ub)
    inst = make_instance([([(0, 2.0)], -INF, 3.0), ([(1, 1.0), (2, 1.0)], 2.0, INF),
                          ([(3, 1.0), (0, 4.0)], -INF, 6.0), ([(3, 1.0), (1, -5.0)], 0.0, INF)],
                         [0.0, 0.0, 0.0, 0.0], [5.0, 1.0, 1.0, 10.0], [1, 1, 1, 0])
    r = orc.propagate_reasons(orc.OMatrix.from_instance(inst), inst.lb, inst.ub)
    bad = dict(r, row=r["row"].copy())
    bad["row"][4] = 2           # x3 >= 5 credited to r2
    with pytest.raises(AssertionError):
        replay_history(inst, inst.lb, inst.ub, bad)
    bad = dict(r, round=r["round"].copy())
    bad["round"][4] = 1         # x3 >= 5 read in round 1, before x1 = 1
    with pytest.raises(AssertionError):
        replay_history(inst, inst.lb, inst.ub, bad)


RANGED = dict(rangedrows=1)


def test_ranged_reference_examples_with_reasons():
    inst = make_instance([([(0, 12.0), (1, 9.0), (2, -1.0)], 0.0, 0.0)], [-INF, -INF, 1.0], [INF, INF, 2.0], [1, 1, 1])
    r = run_and_check(inst, orc.params(**RANGED))
    assert r["conflict"] == dict(kind=1, round=1, row=0, side=2, col=-1, rowlb=-1, rowub=-1, rules=0)
    inst = make_instance([([(0, 12.0), (1, 9.0), (2, -1.0)], 1.0, 1.0)], [-INF, -INF, 1.0], [INF, INF, 2.0], [1, 1, 1])
    r = run_and_check(inst, orc.params(**RANGED))
    got = list(zip(r["round"].tolist(), r["col"].tolist(), r["bound"].tolist(), r["val"].tolist(), r["row"].tolist(),
                   r["rule"].tolist()))
    assert got == [(1, 2, 0, 2.0, 0, 1)]


@pytest.mark.parametrize("seed", range(40))
def test_ranged_reasons_on_small_systems(seed):
    from tests.test_oracle_ranged import ranged_system
    inst = ranged_system(np.random.default_rng(1200 + seed), 4, 7, shift=(seed % 4 == 3) * float(seed % 3 + 1))
    r = run_and_check(inst, orc.params(**RANGED))
    if r["result"]["cutoff"]:
        assert r["certified"]


def test_ranged_reasons_are_exercised():
    from tests.test_oracle_ranged import ranged_system
    rules = sides = 0
    for seed in range(40):
        inst = ranged_system(np.random.default_rng(1200 + seed), 4, 7, shift=(seed % 4 == 3) * float(seed % 3 + 1))
        r = orc.propagate_reasons(orc.OMatrix.from_instance(inst), inst.lb, inst.ub, orc.params(**RANGED))
        rules += int(np.sum(r["rule"]))
        sides += int(r["conflict"]["kind"] == 1 and r["conflict"]["side"] == 2)
    assert rules > 5 and sides > 0, (rules, sides)
