"""Probing restatement of the oracle (oracle.probe, after prop_probing.c SCIPapplyProbingVar
and SCIPanalyzeDeductionsProbing): hand-computed known answers (an equivalence, a fixing
common to both branches, an implication of one branch only, an infeasible branch), and
soundness against brute-force enumeration of every 0/1 point of small random systems."""
import itertools

import numpy as np
import pytest

from oracle import oracle as orc
from tests.util import INF, make_instance, random_binary_instance


def test_known_answers_equivalence_and_common_fixing():
    # x=0, y=1, z=2, w=3 binary: y = x; z >= x and z >= 1 - x; w <= 1 - x
    inst = make_instance([([(1, 1.0), (0, -1.0)], 0.0, 0.0),
                          ([(2, 1.0), (0, -1.0)], 0.0, INF),
                          ([(2, 1.0), (0, 1.0)], 1.0, INF),
                          ([(3, 1.0), (0, 1.0)], -INF, 1.0)],
                         [0.0] * 4, [1.0] * 4, [1] * 4)
    om = orc.OMatrix.from_instance(inst)
    (r,) = orc.probe(om, inst.lb, inst.ub, [0])
    assert r["verdict"] == 0
    # y follows x in both branches (an aggregation y = x); z is 1 in both (a fixing); w changes
    # only in the x = 1 branch (an implication: not listed)
    assert list(r["cols"]) == [1, 2]
    assert r["b0"].tolist() == [[0.0, 0.0], [1.0, 1.0]]
    assert r["b1"].tolist() == [[1.0, 1.0], [1.0, 1.0]]


def test_known_answers_infeasible_branch():
    # a + b >= 1, a + c >= 1, b + c <= 1: a = 0 forces b = c = 1, which violates the third row
    inst = make_instance([([(0, 1.0), (1, 1.0)], 1.0, INF),
                          ([(0, 1.0), (2, 1.0)], 1.0, INF),
                          ([(1, 1.0), (2, 1.0)], -INF, 1.0)],
                         [0.0] * 3, [1.0] * 3, [1] * 3)
    om = orc.OMatrix.from_instance(inst)
    ra, rb, rc = orc.probe(om, inst.lb, inst.ub, [0, 1, 2])
    assert ra["verdict"] == 1 and len(ra["cols"]) == 0
    # b = 0 gives a = 1; b = 1 gives c = 0 and then a = 1: a = 1 in both branches
    assert rb["verdict"] == 0 and list(rb["cols"]) == [0]
    assert rb["b0"].tolist() == [[1.0, 1.0]] and rb["b1"].tolist() == [[1.0, 1.0]]
    # c = 1 gives b = 0 and a = 1; c = 0 gives a = 1: again a = 1 in both, b only in one branch
    assert rc["verdict"] == 0 and list(rc["cols"]) == [0]


def feasible_points(inst):
    n = len(inst.lb)
    pts = np.array(list(itertools.product([0.0, 1.0], repeat=n)))
    ok = np.ones(len(pts), dtype=bool)
    for i in range(len(inst.lhs)):
        a, e = inst.rowptr[i], inst.rowptr[i + 1]
        act = pts[:, inst.colidx[a:e]] @ inst.vals[a:e]
        ok &= (act >= inst.lhs[i] - 1e-9) & (act <= inst.rhs[i] + 1e-9)
    return pts[ok]


def small_case(seed, n=11):
    rng = np.random.default_rng(700 + seed)
    inst = random_binary_instance(rng, nrows=9, ncols=n, maxlen=5, coefmax=6, hidden=bool(seed % 3))
    om = orc.OMatrix.from_instance(inst)
    return inst, orc.probe(om, inst.lb, inst.ub, list(range(n)))


@pytest.mark.parametrize("seed", range(12))
def test_probing_sound_against_enumeration(seed):
    inst, res = small_case(seed)
    F = feasible_points(inst)
    for x, r in enumerate(res):
        for val, bit in ((0.0, 1), (1.0, 2)):
            if r["verdict"] & bit:
                assert not np.any(F[:, x] == val), (seed, x, val)   # an infeasible branch holds no point
        for k, j in enumerate(r["cols"]):
            assert j != x
            for val, b in ((0.0, r["b0"][k]), (1.0, r["b1"][k])):
                Fx = F[F[:, x] == val]
                assert np.all((Fx[:, j] >= b[0]) & (Fx[:, j] <= b[1])), (seed, x, j, val)


def test_enumeration_cases_are_not_vacuous():
    pairs = verdicts = 0
    for seed in range(12):
        _, res = small_case(seed)
        pairs += sum(len(r["cols"]) for r in res)
        verdicts += sum(r["verdict"] != 0 for r in res)
    assert pairs > 20 and verdicts > 5, (pairs, verdicts)


def test_split_known_answer():
    # x + y = 4 with x, y integers in [0, 4]; branches x <= 1 and x >= 2
    inst = make_instance([([(0, 1.0), (1, 1.0)], 4.0, 4.0)], [0.0, 0.0], [4.0, 4.0], [1, 1])
    om = orc.OMatrix.from_instance(inst)
    (r,) = orc.probe(om, inst.lb, inst.ub, [0], downub=[1.0], uplb=[2.0])
    assert r["verdict"] == 0
    # the probe column itself and y, changed in both branches
    assert list(r["cols"]) == [0, 1]
    assert r["b0"].tolist() == [[0.0, 1.0], [3.0, 4.0]]
    assert r["b1"].tolist() == [[2.0, 4.0], [0.0, 2.0]]


def integer_points(inst):
    doms = [range(int(l), int(u) + 1) for l, u in zip(inst.lb, inst.ub)]
    pts = np.array(list(itertools.product(*doms)), dtype=float)
    ok = np.ones(len(pts), dtype=bool)
    for i in range(len(inst.lhs)):
        a, e = inst.rowptr[i], inst.rowptr[i + 1]
        act = pts[:, inst.colidx[a:e]] @ inst.vals[a:e]
        ok &= (act >= inst.lhs[i] - 1e-9) & (act <= inst.rhs[i] + 1e-9)
    return pts[ok]


@pytest.mark.parametrize("seed", range(10))
def test_split_probing_sound_against_enumeration(seed):
    from tests.util import planted_integer_instance
    inst = planted_integer_instance(np.random.default_rng(1500 + seed), 6, 6, 4, coefrange=4, boundrange=3)
    om = orc.OMatrix.from_instance(inst)
    F = integer_points(inst)
    cols = [j for j in range(6) if inst.ub[j] > inst.lb[j]]
    d = [float(np.floor((inst.lb[j] + inst.ub[j]) / 2)) for j in cols]
    res = orc.probe(om, inst.lb, inst.ub, cols, downub=d, uplb=[v + 1.0 for v in d])
    for q, (x, r) in enumerate(zip(cols, res)):
        down = F[F[:, x] <= d[q]]
        up = F[F[:, x] >= d[q] + 1]
        if r["verdict"] & 1:
            assert len(down) == 0
        if r["verdict"] & 2:
            assert len(up) == 0
        if r["verdict"] == 0:
            assert x in r["cols"]
        for k, j in enumerate(r["cols"]):
            for pts, b in ((down, r["b0"][k]), (up, r["b1"][k])):
                assert np.all((pts[:, j] >= b[0] - 1e-9) & (pts[:, j] <= b[1] + 1e-9)), (seed, x, j)
