"""GPU parity of ranged-row propagation (SAP_FLAG_RANGEDROWS) with the oracle's Jacobi twin
(rangedrows=1): bit-exact bounds, verdicts and round counts on small gcd-structured systems,
a large one, the mixed generator, MIPLIB instances, and node batches from a resident base."""
import os

import numpy as np
import pytest

import scip_amd as sa
from oracle import oracle as orc
from tests.test_oracle_ranged import ranged_system
from tests.util import apply_fixings, random_fixings

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
RANGED = sa.SAP_FLAG_RANGEDROWS


def say(*a):
    print(*a, flush=True)


def check(gpu_ctx, inst, name, modes=(sa.SAP_MODE_DENSE, sa.SAP_MODE_SPARSE)):
    say(f"[case] ranged {name} rows={inst.nrows}")
    ol, ou, orr = orc.propagate_jacobi(orc.OMatrix.from_instance(inst), inst.lb, inst.ub, orc.params(rangedrows=1))
    mat = sa.Matrix(gpu_ctx, inst)
    for mode in modes:
        b = sa.Bounds(mat, 1, mode)
        b.upload(inst.lb, inst.ub)
        st = b.propagate(sa.default_params(flags=RANGED))
        assert (st["result"] == sa.SAP_CUTOFF) == bool(orr["cutoff"]), (name, mode)
        assert st["rounds"] == orr["rounds"], (name, mode, st["rounds"], orr["rounds"])
        if not orr["cutoff"]:
            gl, gu = b.download(0)
            assert np.array_equal(gl, ol) and np.array_equal(gu, ou), (name, mode)
            assert st["nchgbds"] == orr["nchgbds"]
        b.close()
    mat.close()
    return orr


def test_small_systems(gpu_ctx):
    cut = 0
    for seed in range(40):
        inst = ranged_system(np.random.default_rng(1200 + seed), 4, 7, shift=(seed % 4 == 3) * float(seed % 3 + 1))
        cut += check(gpu_ctx, inst, f"small{seed}")["cutoff"]
    assert cut > 0


@pytest.mark.parametrize("seed", range(3))
def test_large_system(gpu_ctx, seed):
    # feasible, with 40-100 bounds that only ranged-row propagation finds
    inst = ranged_system(np.random.default_rng(1300 + seed), 1500, 3000, slack=False)
    check(gpu_ctx, inst, f"large{seed}")


@pytest.mark.parametrize("seed", range(2))
def test_mixed(gpu_ctx, seed):
    check(gpu_ctx, sa.gen_mixed(20000, 20000, seed=70 + seed, planted=True), f"mixed{seed}")


@pytest.mark.parametrize("name", ["bell5", "blend2", "dcmulti", "egout", "enigma", "flugpl", "gt2", "lseu", "misc03",
                                  "p0548", "rgn"])
def test_real_instances(gpu_ctx, name):
    path = os.path.join(HERE, "golden", "mps", name + ".mps")
    if not os.path.exists(path):
        pytest.skip("instance not among the fixtures")
    inst, _ = sa.read_mps(path)
    check(gpu_ctx, inst, name)


def test_nodes(gpu_ctx):
    inst = ranged_system(np.random.default_rng(1310), 1500, 3000, slack=False)
    om = orc.OMatrix.from_instance(inst)
    prm = orc.params(rangedrows=1)
    rl, ru, rr = orc.propagate_jacobi(om, inst.lb, inst.ub, prm)
    assert not rr["cutoff"]
    free = np.nonzero(rl < ru)[0]
    rng = np.random.default_rng(5)
    nvec = 48
    ptr, cols, vals = random_fixings(rng, len(free), nvec, 4)
    cols = free[cols].astype(np.int32)
    vals = np.array([rl[c] if v == 0 else ru[c] for c, v in zip(cols, vals)])
    say(f"[case] ranged nodes nvec={nvec}")
    mat = sa.Matrix(gpu_ctx, inst)
    base = sa.Bounds(mat, 1)
    base.upload(rl, ru)
    for mode in (sa.SAP_MODE_DENSE, sa.SAP_MODE_SPARSE):
        nodes = sa.Bounds(mat, nvec, mode)
        out = nodes.propagate_nodes(base, 0, ptr, cols, vals, vals, sa.default_params(flags=RANGED))
        for v in range(nvec):
            lb0, ub0 = apply_fixings(rl, ru, ptr, cols, vals, v)
            ol, ou, orr = orc.propagate_jacobi(om, lb0, ub0, prm)
            assert (out["result"][v] == sa.SAP_CUTOFF) == bool(orr["cutoff"]), (mode, v)
            if orr["cutoff"]:
                continue
            a, e = out["chgptr"][v], out["chgptr"][v + 1]
            gl, gu = lb0.copy(), ub0.copy()
            gl[out["chgcol"][a:e]] = out["chglb"][a:e]
            gu[out["chgcol"][a:e]] = out["chgub"][a:e]
            assert np.array_equal(gl, ol) and np.array_equal(gu, ou), (mode, v)
        nodes.close()
    base.close()
    mat.close()


@pytest.mark.parametrize("mode", [sa.SAP_MODE_DENSE, sa.SAP_MODE_SPARSE])
def test_slots(gpu_ctx, mode):
    # continuous batching with ranged rows: nodes admitted over several steps, each equal to the
    # oracle's result for that node alone
    inst = ranged_system(np.random.default_rng(1311), 1500, 3000, slack=False)
    om = orc.OMatrix.from_instance(inst)
    prm = orc.params(rangedrows=1)
    rl, ru, rr = orc.propagate_jacobi(om, inst.lb, inst.ub, prm)
    free = np.nonzero(rl < ru)[0]
    rng = np.random.default_rng(9)
    nnode = 30
    ptr, cols, vals = random_fixings(rng, len(free), nnode, 4)
    cols = free[cols].astype(np.int32)
    vals = np.array([rl[c] if v == 0 else ru[c] for c, v in zip(cols, vals)])
    say(f"[case] ranged slots mode={mode}")
    mat = sa.Matrix(gpu_ctx, inst)
    base = sa.Bounds(mat, 1)
    base.upload(rl, ru)
    sl = sa.Bounds(mat, 8, mode)
    p = sa.default_params(flags=RANGED)
    pending = list(range(nnode))
    free_slots = list(range(8))
    owner = {}
    results = {}
    for _ in range(2000):
        while pending and free_slots:
            v = pending.pop(0)
            s = free_slots.pop(0)
            owner[s] = v
            sl.admit([s], base, 0, [0, ptr[v + 1] - ptr[v]], cols[ptr[v]:ptr[v + 1]], vals[ptr[v]:ptr[v + 1]],
                     vals[ptr[v]:ptr[v + 1]])
        fin, _ = sl.step(p, rounds=sa.SAP_STEP_SHORT)
        if len(fin):
            c = sl.collect(fin)
            for i, s in enumerate(fin.tolist()):
                v = owner.pop(s)
                a, e = c["chgptr"][i], c["chgptr"][i + 1]
                results[v] = (int(c["result"][i]), c["chgcol"][a:e].copy(), c["chglb"][a:e].copy(),
                              c["chgub"][a:e].copy())
                free_slots.append(s)
        if len(results) == nnode:
            break
    assert len(results) == nnode
    for v in range(nnode):
        lb0, ub0 = apply_fixings(rl, ru, ptr, cols, vals, v)
        ol, ou, orr = orc.propagate_jacobi(om, lb0, ub0, prm)
        res, cc, cl, cu = results[v]
        assert (res == sa.SAP_CUTOFF) == bool(orr["cutoff"]), v
        if not orr["cutoff"]:
            gl, gu = lb0.copy(), ub0.copy()
            gl[cc] = cl
            gu[cc] = cu
            assert np.array_equal(gl, ol) and np.array_equal(gu, ou), v
    sl.close()
    base.close()
    mat.close()
