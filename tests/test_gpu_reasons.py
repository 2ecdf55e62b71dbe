"""GPU parity of propagation with reasons (sap_propagate_reasons): the bounds, the history of
accepted changes with their reason rows and the cutoff certificate equal the oracle's
(oracle.propagate_reasons) exactly, and the engine's own history passes the exact
rational checks of tests/reasons_check.py."""
import os

import numpy as np
import pytest

import scip_amd as sa
from oracle import oracle as orc
from tests.reasons_check import check_conflict, replay_history
from tests.util import (INF, late_cut_instance, make_instance, planted_integer_instance, random_binary_instance,
                        random_integer_instance)

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def say(*a):
    print(*a, flush=True)


def compare(gpu_ctx, inst, name, exact=True, ranged=False):
    say(f"[case] reasons {name} rows={inst.nrows} cols={inst.ncols} ranged={ranged}")
    mat = sa.Matrix(gpu_ctx, inst)
    prm = sa.default_params(flags=sa.SAP_FLAG_RANGEDROWS if ranged else 0)
    g = mat.propagate_reasons(params=prm)
    o = orc.propagate_reasons(orc.OMatrix.from_instance(inst), inst.lb, inst.ub, orc.params(rangedrows=int(ranged)))
    cut = g["stats"]["result"] == sa.SAP_CUTOFF
    assert cut == bool(o["result"]["cutoff"]), name
    assert g["stats"]["rounds"] == o["result"]["rounds"], name
    assert g["conflict"] == o["conflict"], (name, g["conflict"], o["conflict"])
    assert g["nhist"] == o["nhist"], name
    for k in ("round", "col", "bound", "row", "rule"):
        assert np.array_equal(g[k], o[k]), (name, k)
    assert np.array_equal(g["val"].view(np.int64), o["val"].view(np.int64)), name
    if not cut:
        assert np.array_equal(g["lb"], o["lb"]) and np.array_equal(g["ub"], o["ub"]), name
        # the same propagation as without reasons
        pl, pu, ps = mat.propagate(params=prm)
        assert np.array_equal(pl, g["lb"]) and np.array_equal(pu, g["ub"]) and ps["rounds"] == g["stats"]["rounds"]
    if exact:
        _, lb, ub = replay_history(inst, inst.lb, inst.ub, g)
        if cut:
            check_conflict(inst, g["conflict"], lb, ub)
    mat.close()
    return g


def test_known_cases(gpu_ctx):
    inst = make_instance([([(0, 2.0)], -INF, 3.0), ([(1, 1.0), (2, 1.0)], 2.0, INF),
                          ([(3, 1.0), (0, 4.0)], -INF, 6.0), ([(3, 1.0), (1, -5.0)], 0.0, INF)],
                         [0.0, 0.0, 0.0, 0.0], [5.0, 1.0, 1.0, 10.0], [1, 1, 1, 0])
    g = compare(gpu_ctx, inst, "forced")
    assert g["row"].tolist() == [0, 1, 1, 2, 3, 2]
    inst = make_instance([([(0, 1.0), (1, 1.0)], 3.0, INF), ([(1, 1.0), (2, 1.0)], -INF, 1.0)],
                         [0.0] * 3, [1.0] * 3, [1] * 3)
    assert compare(gpu_ctx, inst, "rowcut")["conflict"]["kind"] == sa.SAP_CONFLICT_ROW
    inst = make_instance([([(0, 1.0), (1, 1.0)], 2.0, INF), ([(0, 1.0), (2, 1.0)], -INF, 0.5)],
                         [0.0] * 3, [1.0] * 3, [1] * 3)
    g = compare(gpu_ctx, inst, "colcut")
    assert g["conflict"]["kind"] == sa.SAP_CONFLICT_COLUMN and (g["conflict"]["rowlb"], g["conflict"]["rowub"]) == (0, 1)


@pytest.mark.parametrize("seed", range(6))
def test_random_binary(gpu_ctx, seed):
    inst = random_binary_instance(np.random.default_rng(900 + seed), 3000, 2000, 10, hidden=(seed % 2 == 0))
    compare(gpu_ctx, inst, f"binary{seed}")


@pytest.mark.parametrize("seed", range(6))
def test_random_integer(gpu_ctx, seed):
    inst = random_integer_instance(np.random.default_rng(950 + seed), 400, 300, 8, sidejitter=0.3)
    compare(gpu_ctx, inst, f"integer{seed}")


@pytest.mark.parametrize("seed", range(4))
def test_planted_integer(gpu_ctx, seed):
    compare(gpu_ctx, planted_integer_instance(np.random.default_rng(950 + seed), 400, 300, 6), f"pint{seed}")


@pytest.mark.parametrize("seed", range(3))
def test_row_certificate_after_history(gpu_ctx, seed):
    inst = random_binary_instance(np.random.default_rng(900 + seed), 1500, 1000, 10, hidden=True)
    o = orc.propagate_reasons(orc.OMatrix.from_instance(inst), inst.lb, inst.ub)
    g = compare(gpu_ctx, late_cut_instance(inst, o["round"], o["col"], o["bound"], o["val"]), f"late{seed}")
    assert g["conflict"]["kind"] == sa.SAP_CONFLICT_ROW and g["nhist"] > 0


@pytest.mark.parametrize("seed", range(4))
def test_mixed(gpu_ctx, seed):
    # long rows (wave and workgroup paths), hard-path rows, infinite and huge bounds
    inst = sa.gen_mixed(20000, 20000, seed=60 + seed, planted=bool(seed % 2))
    compare(gpu_ctx, inst, f"mixed{seed}", exact=(seed < 2))


@pytest.mark.parametrize("name", ["bell5", "p0548", "misc03", "egout", "lseu"])
def test_real_instances(gpu_ctx, name):
    inst, _ = sa.read_mps(os.path.join(HERE, "golden", "mps", name + ".mps"))
    compare(gpu_ctx, inst, name)


def test_history_capacity(gpu_ctx):
    inst = random_binary_instance(np.random.default_rng(901), 3000, 2000, 10, hidden=True)
    mat = sa.Matrix(gpu_ctx, inst)
    full = mat.propagate_reasons()
    n = full["nhist"]
    assert n > 2
    with pytest.raises(sa.SapError) as e:
        mat.propagate_reasons(histcap=n - 1)
    assert e.value.rc == sa.SAP_NOMEMORY
    again = mat.propagate_reasons(histcap=n)
    for k in ("round", "col", "bound", "val", "row"):
        assert np.array_equal(full[k], again[k]), k
    mat.close()


def test_ranged_small_systems(gpu_ctx):
    from tests.test_oracle_ranged import ranged_system
    rules = 0
    for seed in range(40):
        inst = ranged_system(np.random.default_rng(1200 + seed), 4, 7, shift=(seed % 4 == 3) * float(seed % 3 + 1))
        rules += int(np.sum(compare(gpu_ctx, inst, f"rsmall{seed}", ranged=True)["rule"]))
    assert rules > 5


@pytest.mark.parametrize("seed", range(2))
def test_ranged_large(gpu_ctx, seed):
    from tests.test_oracle_ranged import ranged_system
    g = compare(gpu_ctx, ranged_system(np.random.default_rng(1300 + seed), 1500, 3000, slack=False), f"rlarge{seed}",
                ranged=True)
    assert int(np.sum(g["rule"])) > 0


@pytest.mark.parametrize("name", ["bell5", "p0548", "misc03", "rgn", "gt2"])
def test_ranged_real_instances(gpu_ctx, name):
    inst, _ = sa.read_mps(os.path.join(HERE, "golden", "mps", name + ".mps"))
    compare(gpu_ctx, inst, name, ranged=True)


def test_ranged_mixed(gpu_ctx):
    compare(gpu_ctx, sa.gen_mixed(20000, 20000, seed=71, planted=True), "rmixed", exact=False, ranged=True)
