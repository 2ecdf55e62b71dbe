"""GPU parity: the HIP engine against the oracle's Jacobi twin, through the C ABI.

Integer/index/verdict work is compared exactly; bounds are compared with numeric equality
(identical doubles, -0.0 == +0.0), i.e. bit-exact up to the sign of zero.
"""
import numpy as np
import pytest

import scip_amd as sa
from oracle import oracle as orc
from tests.util import (INF, apply_fixings, make_instance, random_binary_instance, random_fixings,
                        random_integer_instance, same_bounds)

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def check_single(ctx, inst, lb=None, ub=None, name="", **kw):
    say(f"[case] {name} rows={inst.nrows} cols={inst.ncols} nnz={inst.nnz} {kw}")
    lb = inst.lb if lb is None else lb
    ub = inst.ub if ub is None else ub
    om = orc.OMatrix.from_instance(inst)
    ol, ou, orr = orc.propagate_jacobi(om, lb, ub, orc.params(**kw))
    mat = sa.Matrix(ctx, inst)
    gl, gu, gs = mat.propagate(lb, ub, sa.default_params(**kw))
    assert (gs["result"] == sa.SAP_CUTOFF) == bool(orr["cutoff"]), (gs, orr)
    assert gs["rounds"] == orr["rounds"], (gs, orr)
    if not orr["cutoff"]:
        assert same_bounds(gl, ol) and same_bounds(gu, ou)
        assert gs["nchgbds"] == orr["nchgbds"]
        assert gs["rows_swept"] == orr["rows_swept"] and gs["nnz_swept"] == orr["nnz_swept"]
        assert gs["result"] == (sa.SAP_REDUCEDDOM if orr["nchgbds"] > 0 else sa.SAP_DIDNOTFIND)
    mat.close()
    return gl, gu, gs


# --- activities --------------------------------------------------------------------------------
@pytest.mark.parametrize("kind", ["cover", "mixed", "long"])
def test_row_activities_bit_exact(gpu_ctx, kind):
    if kind == "cover":
        inst = sa.gen_covering_packing(2000, 5000, 10, seed=1)
    elif kind == "mixed":
        inst = sa.gen_mixed(20000, 20000, seed=2)
    else:
        inst = sa.gen_mixed(300, 20000, seed=9, lenmode=1, minlen=60, maxlen=9000)
    say(f"[case] activities {kind} nnz={inst.nnz}")
    om = orc.OMatrix.from_instance(inst)
    mn, mx, cnt, mad = orc.row_activities(om, inst.lb, inst.ub)
    mat = sa.Matrix(gpu_ctx, inst)
    gmn, gmx, gcnt, gmad = mat.row_activities()
    assert np.array_equal(gcnt, cnt)
    assert same_bounds(gmn, mn) and same_bounds(gmx, mx) and same_bounds(gmad, mad)
    mat.close()


# --- closed-form and edge cases ------------------------------------------------------------------
EDGE = {
    "packing": ([([(0, 1.0), (1, 1.0)], -INF, 1.0)], [1.0, 0.0], [1.0, 1.0], [1, 1]),
    "cont_rhs": ([([(0, 1.0), (1, 1.0)], -INF, 10.0)], [0.0, 5.0], [100.0, 8.0], [0, 0]),
    "boundstreps": ([([(0, 1.0), (1, 1.0)], -INF, 10.0)], [0.0, 0.0], [10.2, 10.0], [0, 0]),
    "int_round": ([([(0, 2.0)], -INF, 5.0), ([(1, 3.0)], -INF, 5.9999999)], [0.0, 0.0], [10.0, 10.0], [1, 1]),
    "neg_coef": ([([(0, -1.0), (1, 1.0)], 3.0, INF)], [0.0, 0.0], [10.0, 4.0], [1, 1]),
    "singleton": ([([(0, 1.0)], -INF, 9.9)], [0.0], [10.0], [0]),
    "infeasible": ([([(0, 1.0), (1, 1.0)], 3.0, INF)], [0.0, 0.0], [1.0, 1.0], [1, 1]),
    "inf_bound": ([([(0, 1.0), (1, 1.0)], -INF, 10.0)], [0.0, -INF], [100.0, 5.0], [0, 0]),
    "huge": ([([(0, 1.0), (1, 1e10)], -INF, 1e16)], [0.0, 0.0], [5.0, 1e6], [0, 0]),
    "empty_ok": ([([], -1.0, 1.0), ([(0, 1.0)], -INF, 0.5)], [0.0], [1.0], [1]),
    "empty_bad": ([([], 1.0, 2.0)], [0.0], [1.0], [1]),
}


@pytest.mark.parametrize("name", sorted(EDGE))
def test_edge_cases(gpu_ctx, name):
    rows, lb, ub, it = EDGE[name]
    check_single(gpu_ctx, make_instance(rows, lb, ub, it), name=name)


def test_boundstreps_zero(gpu_ctx):
    rows, lb, ub, it = EDGE["boundstreps"]
    check_single(gpu_ctx, make_instance(rows, lb, ub, it), name="boundstreps0", boundstreps=0.0)


@pytest.mark.parametrize("seed", range(25))
def test_random_small_integer(gpu_ctx, seed):
    rng = np.random.default_rng(seed)
    inst = random_integer_instance(rng, nrows=int(rng.integers(1, 8)), ncols=int(rng.integers(2, 8)), maxlen=6)
    check_single(gpu_ctx, inst, name=f"int{seed}")


@pytest.mark.parametrize("seed", range(6))
def test_random_binary_with_fixings(gpu_ctx, seed):
    rng = np.random.default_rng(300 + seed)
    inst = random_binary_instance(rng, nrows=400, ncols=300, maxlen=12)
    check_single(gpu_ctx, inst, name=f"bin{seed}")


# --- generated workloads -------------------------------------------------------------------------
def test_config1_covering_packing(gpu_ctx):
    inst = sa.gen_covering_packing(2000, 5000, 10, seed=1)
    check_single(gpu_ctx, inst, name="cfg1")


@pytest.mark.parametrize("seed", range(4))
def test_mixed_lognormal(gpu_ctx, seed):
    inst = sa.gen_mixed(20000, 20000, seed=seed)
    check_single(gpu_ctx, inst, name=f"mixed{seed}")


@pytest.mark.parametrize("seed", range(4))
def test_mixed_planted_feasible(gpu_ctx, seed):
    inst = sa.gen_mixed(60000, 60000, seed=seed, planted=True)
    gl, gu, gs = check_single(gpu_ctx, inst, name=f"planted{seed}")
    assert gs["result"] == sa.SAP_REDUCEDDOM and gs["rounds"] >= 3


def test_mixed_long_rows_all_bins(gpu_ctx):
    # rows 60..9000 long exercise the wavefront and workgroup bins
    inst = sa.gen_mixed(400, 30000, seed=7, lenmode=1, minlen=60, maxlen=9000)
    check_single(gpu_ctx, inst, name="longrows")


def test_mixed_uneven_lengths(gpu_ctx):
    inst = sa.gen_mixed(5000, 7500, seed=4004, lenmode=1, minlen=4, maxlen=64)
    check_single(gpu_ctx, inst, name="uniform4-64")


@pytest.mark.parametrize("maxrounds", [0, 1, 2])
def test_maxrounds(gpu_ctx, maxrounds):
    inst = sa.gen_mixed(5000, 5000, seed=3)
    check_single(gpu_ctx, inst, name="maxrounds", maxrounds=maxrounds)


def test_determinism_and_resident_path(gpu_ctx):
    inst = sa.gen_mixed(50000, 50000, seed=5)
    say("[case] determinism")
    mat = sa.Matrix(gpu_ctx, inst)
    a = mat.propagate()
    b = mat.propagate()
    assert a[0].tobytes() == b[0].tobytes() and a[1].tobytes() == b[1].tobytes()
    bnd = sa.Bounds(mat, 1)
    bnd.upload(inst.lb, inst.ub)
    st = bnd.propagate()
    lb, ub = bnd.download()
    assert lb.tobytes() == a[0].tobytes() and ub.tobytes() == a[1].tobytes()
    assert st["rounds"] == a[2]["rounds"]
    # a fixpoint stays a fixpoint
    c = mat.propagate(a[0], a[1])
    assert c[2]["nchgbds"] == 0 and c[2]["rounds"] == 1
    bnd.close()
    mat.close()


# --- batches -------------------------------------------------------------------------------------
def test_batch_nodes_match_single_vector_oracle(gpu_ctx):
    inst = sa.gen_covering_packing(2000, 3000, 10, seed=3)
    rng = np.random.default_rng(3)
    # tighten sides so fixings cascade: use the planted-point binary generator instead
    inst = random_binary_instance(rng, nrows=3000, ncols=2000, maxlen=12)
    om = orc.OMatrix.from_instance(inst)
    rl, ru, rr = orc.propagate_jacobi(om, inst.lb, inst.ub)
    free = np.nonzero(rl < ru)[0]
    nvec = 48
    ptr, cols, vals = random_fixings(rng, len(free), nvec, 6)
    cols = free[cols].astype(np.int32)
    say(f"[case] batch nvec={nvec} free={len(free)}")
    mat = sa.Matrix(gpu_ctx, inst)
    for mode, flags in [(m, f) for m in (sa.SAP_MODE_DENSE, sa.SAP_MODE_SPARSE)
                        for f in (0, sa.SAP_FLAG_BASE_IS_FIXPOINT)]:
        say(f"[case] batch mode={mode} flags={flags}")
        bnd = sa.Bounds(mat, nvec, mode)
        out = bnd.propagate_batch(rl, ru, ptr, cols, vals, vals, sa.default_params(flags=flags))
        for v in range(nvec):
            lb, ub = apply_fixings(rl, ru, ptr, cols, vals, v)
            ol, ou, orr = orc.propagate_jacobi(om, lb, ub)
            assert (out["result"][v] == sa.SAP_CUTOFF) == bool(orr["cutoff"]), (v, flags)
            if not orr["cutoff"]:
                assert same_bounds(out["lb"][v], ol) and same_bounds(out["ub"][v], ou), (v, flags)
                if flags == 0:
                    assert out["rounds"][v] == orr["rounds"]
        bnd.close()
    mat.close()


@pytest.mark.parametrize("mode", ["dense", "sparse"])
def test_single_vector_both_modes(gpu_ctx, mode):
    inst = sa.gen_mixed(30000, 30000, seed=21, planted=True)
    say(f"[case] single vector mode={mode}")
    om = orc.OMatrix.from_instance(inst)
    ol, ou, orr = orc.propagate_jacobi(om, inst.lb, inst.ub)
    mat = sa.Matrix(gpu_ctx, inst)
    bnd = sa.Bounds(mat, 1, sa.SAP_MODE_DENSE if mode == "dense" else sa.SAP_MODE_SPARSE)
    bnd.upload(inst.lb, inst.ub)
    st = bnd.propagate()
    lb, ub = bnd.download()
    assert st["rounds"] == orr["rounds"] and st["nchgbds"] == orr["nchgbds"]
    assert st["rows_swept"] == orr["rows_swept"] and st["nnz_swept"] == orr["nnz_swept"]
    assert same_bounds(lb, ol) and same_bounds(ub, ou)
    bnd.close()
    mat.close()


def test_batch_mixed_instance(gpu_ctx):
    inst = sa.gen_mixed(8000, 8000, seed=11)
    om = orc.OMatrix.from_instance(inst)
    rng = np.random.default_rng(11)
    nvec = 8
    it = np.nonzero(inst.integral)[0]
    ptr = np.arange(0, 3 * nvec + 1, 3, dtype=np.int64)
    cols = rng.choice(it, size=3 * nvec, replace=False).astype(np.int32)
    vals = np.zeros(3 * nvec)
    say("[case] batch mixed")
    mat = sa.Matrix(gpu_ctx, inst)
    for mode in (sa.SAP_MODE_DENSE, sa.SAP_MODE_SPARSE):
        bnd = sa.Bounds(mat, nvec, mode)
        out = bnd.propagate_batch(inst.lb, inst.ub, ptr, cols, vals, vals)
        for v in range(nvec):
            lb, ub = apply_fixings(inst.lb, inst.ub, ptr, cols, vals, v)
            ol, ou, orr = orc.propagate_jacobi(om, lb, ub)
            assert (out["result"][v] == sa.SAP_CUTOFF) == bool(orr["cutoff"])
            if not orr["cutoff"]:
                assert same_bounds(out["lb"][v], ol) and same_bounds(out["ub"][v], ou)
                assert out["rounds"][v] == orr["rounds"]
        # a smaller batch on the same (larger) bounds object
        out2 = bnd.propagate_batch(inst.lb, inst.ub, ptr[:4], cols, vals, vals)
        assert np.array_equal(out2["lb"], out["lb"][:3]) and np.array_equal(out2["ub"], out["ub"][:3])
        bnd.close()
    mat.close()


# --- input validation ------------------------------------------------------------------------------
def test_invalid_inputs_rejected(gpu_ctx):
    say("[case] invalid inputs")
    good = make_instance([([(0, 1.0), (1, 2.0)], -INF, 3.0)], [0.0, 0.0], [1.0, 1.0], [1, 1])
    dup = make_instance([([(0, 1.0), (1, 2.0)], -INF, 3.0)], [0.0, 0.0], [1.0, 1.0], [1, 1])
    dup.colidx[1] = 0
    with pytest.raises(sa.SapError) as e:
        sa.Matrix(gpu_ctx, dup)
    assert e.value.rc == sa.SAP_INVALIDDATA
    bad = make_instance([([(0, 1.0), (1, 2.0)], -INF, 3.0)], [0.0, 0.0], [1.0, 1.0], [1, 1])
    bad.colidx[1] = 7
    with pytest.raises(sa.SapError):
        sa.Matrix(gpu_ctx, bad)
    nanv = make_instance([([(0, 1.0), (1, 2.0)], -INF, 3.0)], [0.0, 0.0], [1.0, 1.0], [1, 1])
    nanv.vals[0] = np.nan
    with pytest.raises(sa.SapError):
        sa.Matrix(gpu_ctx, nanv)
    mat = sa.Matrix(gpu_ctx, good)
    with pytest.raises(sa.SapError):
        mat.propagate(np.array([np.nan, 0.0]), np.array([1.0, 1.0]))
    with pytest.raises(sa.SapError):
        mat.propagate(params=sa.default_params(maxrounds=-5))
    mat.close()


@pytest.mark.parametrize("planted", [False, True])
def test_graph_rounds_equal_host_loop(gpu_ctx, planted):
    inst = sa.gen_mixed(40000, 40000, seed=31, planted=planted)
    say(f"[case] graph vs host loop planted={planted}")
    mat = sa.Matrix(gpu_ctx, inst)
    a = mat.propagate(params=sa.default_params())
    b = mat.propagate(params=sa.default_params(flags=sa.SAP_FLAG_HOSTLOOP))
    assert a[0].tobytes() == b[0].tobytes() and a[1].tobytes() == b[1].tobytes()
    for k in ("result", "rounds", "nchgbds", "rows_swept", "nnz_swept"):
        assert a[2][k] == b[2][k], k
    assert a[2]["sweep_launches"] == b[2]["sweep_launches"] == a[2]["rounds"]
    # different parameters on the same matrix re-capture the round graphs
    c = mat.propagate(params=sa.default_params(maxrounds=2))
    d = mat.propagate(params=sa.default_params(maxrounds=2, flags=sa.SAP_FLAG_HOSTLOOP))
    assert c[0].tobytes() == d[0].tobytes() and c[2]["rounds"] == d[2]["rounds"] <= 2
    mat.close()


@pytest.mark.parametrize("badrow", [False, True])
def test_notighten_flag(gpu_ctx, badrow):
    # SAP_FLAG_NOTIGHTEN (consPropLinear's tightenbounds FALSE): rows are only tested; with
    # badrow one row's rhs lies below its finite minimal activity, so the final test cuts off
    inst = sa.gen_mixed(20000, 20000, seed=6, planted=True)
    if badrow:
        mn, _, _, _ = orc.row_activities(orc.OMatrix.from_instance(inst), inst.lb, inst.ub)
        i = int(np.nonzero(np.abs(mn) < 1e15)[0][3])
        inst.rhs[i] = mn[i] - 1.0
        inst.lhs[i] = -INF
    say(f"[case] notighten badrow={badrow}")
    ol, ou, orr = orc.propagate_jacobi(orc.OMatrix.from_instance(inst), inst.lb, inst.ub,
                                       orc.params(notighten=1, rangedrows=1))
    assert bool(orr["cutoff"]) == badrow and orr["rounds"] == 1 and orr["nchgbds"] == 0
    mat = sa.Matrix(gpu_ctx, inst)
    for flags in (sa.SAP_FLAG_NOTIGHTEN, sa.SAP_FLAG_NOTIGHTEN | sa.SAP_FLAG_RANGEDROWS):
        gl, gu, gs = mat.propagate(inst.lb, inst.ub, sa.default_params(flags=flags))
        assert (gs["result"] == sa.SAP_CUTOFF) == badrow
        assert gs["rounds"] == 1 and gs["nchgbds"] == 0
        assert np.array_equal(gl, inst.lb) and np.array_equal(gu, inst.ub)
    mat.close()


@pytest.mark.parametrize("name", ["rhs_ub", "lhs_lb", "lhs_ub_neg", "rhs_lb_neg"])
def test_reliable_residual_recompute(gpu_ctx, name):
    # the cancellation cases whose bound only the plain-double recompute of
    # consdataGetReliableResidualActivity gives; the device counts the recomputes it made
    from tests.test_oracle_prop import RELIABLE, reliable_instance
    _, _, _, side, expect = RELIABLE[name]
    inst = reliable_instance(name)
    gl, gu, gs = check_single(gpu_ctx, inst, name=f"reliable_{name}")
    assert (gl[3] if side == "lb" else gu[3]) == expect
    assert gs["nrecompute"] == 1 and gs["nchgbds"] == 1


def test_recompute_count_matches_twin(gpu_ctx):
    # on the mixed generator the reliability recompute is rare; the device counts the same
    # recomputes as the twin, round by round and row by row
    inst = sa.gen_mixed(60000, 60000, seed=1, planted=True)
    ol, ou, orr = orc.propagate_jacobi(orc.OMatrix.from_instance(inst), inst.lb, inst.ub)
    gl, gu, gs = check_single(gpu_ctx, inst, name="recompute_count")
    assert gs["nrecompute"] == orr["nrecompute"]


@pytest.mark.parametrize("seed", range(6))
def test_unique_integer_fixpoint_equals_scip_order(gpu_ctx, seed):
    # integer data with boundstreps 0: the fixpoint is unique, so the device must reach SCIP's
    # own (Gauss-Seidel) fixpoint exactly, not only the Jacobi twin's
    from tests.test_oracle_fixpoint import integer_instance
    rng = np.random.default_rng(6000 + seed)
    inst, _ = integer_instance(rng, nrows=300, ncols=200, maxlen=8, coefrange=20, boundrange=0, ubmax=1000,
                               planted=True)
    gl, gu, gs = check_single(gpu_ctx, inst, name=f"unique_int{seed}", boundstreps=0.0)
    sl, su, sr = orc.propagate_scip(orc.OMatrix.from_instance(inst), inst.lb, inst.ub, orc.params(boundstreps=0.0))
    assert not sr["cutoff"] and gs["result"] == sa.SAP_REDUCEDDOM
    assert np.array_equal(gl, sl) and np.array_equal(gu, su)


@pytest.mark.parametrize("blocks", [0, 3])
def test_gather_floor_entry_point(gpu_ctx, blocks):
    # the memory-floor measurement (sap_bench_gather) on a matrix whose nonzero count is not a
    # multiple of four (the kernel's vector loads and its one-by-one tail), with one lane per
    # four nonzeros and with a grid-stride loop over three workgroups; it leaves the bounds alone
    inst = sa.gen_mixed(3001, 2500, seed=8)
    assert inst.nnz % 4 != 0
    mat = sa.Matrix(gpu_ctx, inst)
    b = sa.Bounds(mat, 1)
    b.upload(inst.lb, inst.ub)
    ms = b.bench_gather(3, blocks)
    assert 0.0 < ms < 1000.0
    gl, gu = b.download(0)
    assert np.array_equal(gl, inst.lb) and np.array_equal(gu, inst.ub)
    b.close()
    mat.close()
