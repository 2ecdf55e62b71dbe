"""GPU parity on real MIPLIB instances (the MPS files SCIP's own tests use): root propagation,
batched node fixings, all instances packed into one launch, and the SCIP plugin path --
each bit-exact against the oracle's Jacobi twin."""
import glob
import os

import numpy as np
import pytest

import scip_amd as sa
from oracle import oracle as orc
from tests import plugin_host as ph
from tests.util import same_bounds

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
MPS = sorted(glob.glob(os.path.join(HERE, "golden", "mps", "*.mps")))


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("path", MPS, ids=[os.path.basename(p) for p in MPS])
def test_root_propagation(gpu_ctx, path):
    inst, _ = sa.read_mps(path)
    say(f"[case] {os.path.basename(path)} rows={inst.nrows} nnz={inst.nnz}")
    ol, ou, orr = orc.propagate_jacobi(orc.OMatrix.from_instance(inst), inst.lb, inst.ub)
    for mode in (sa.SAP_MODE_DENSE, sa.SAP_MODE_SPARSE):
        mat = sa.Matrix(gpu_ctx, inst)
        b = sa.Bounds(mat, 1, mode)
        b.upload(inst.lb, inst.ub)
        st = b.propagate()
        gl, gu = b.download(0)
        assert (st["result"] == sa.SAP_CUTOFF) == bool(orr["cutoff"])
        assert st["rounds"] == orr["rounds"] and st["nchgbds"] == orr["nchgbds"]
        assert st["rows_swept"] == orr["rows_swept"] and st["nnz_swept"] == orr["nnz_swept"]
        assert same_bounds(gl, ol) and same_bounds(gu, ou)
        b.close()
        mat.close()


@pytest.mark.parametrize("name", ["bell5", "misc03", "p0548", "lseu", "dcmulti"])
def test_nodes_on_real_instances(gpu_ctx, name):
    inst, _ = sa.read_mps(os.path.join(HERE, "golden", "mps", name + ".mps"))
    om = orc.OMatrix.from_instance(inst)
    rl, ru, rr = orc.propagate_jacobi(om, inst.lb, inst.ub)
    rng = np.random.default_rng(len(name))
    free = np.nonzero((inst.integral != 0) & (rl < ru) & (ru - rl < 1e6))[0]
    nvec = 32
    ptr = [0]
    cols, vals = [], []
    for _ in range(nvec):
        k = int(rng.integers(1, min(8, len(free)) + 1))
        cs = rng.choice(free, size=k, replace=False)
        for c in cs:
            cols.append(int(c))
            vals.append(float(rl[c] if rng.integers(0, 2) == 0 else ru[c]))
        ptr.append(len(cols))
    ptr = np.array(ptr, dtype=np.int64)
    cols = np.array(cols, dtype=np.int32)
    vals = np.array(vals)
    say(f"[case] nodes {name} nvec={nvec}")
    mat = sa.Matrix(gpu_ctx, inst)
    base = sa.Bounds(mat, 1)
    base.upload(rl, ru)
    nodes = sa.Bounds(mat, nvec)
    out = nodes.propagate_nodes(base, 0, ptr, cols, vals, vals, sa.default_params())
    for v in range(nvec):
        lb0, ub0 = rl.copy(), ru.copy()
        lb0[cols[ptr[v]:ptr[v + 1]]] = vals[ptr[v]:ptr[v + 1]]
        ub0[cols[ptr[v]:ptr[v + 1]]] = vals[ptr[v]:ptr[v + 1]]
        ol, ou, orr = orc.propagate_jacobi(om, lb0, ub0)
        assert (out["result"][v] == sa.SAP_CUTOFF) == bool(orr["cutoff"]), v
        assert out["rounds"][v] == orr["rounds"], v
        if not orr["cutoff"]:
            a, e = out["chgptr"][v], out["chgptr"][v + 1]
            gl, gu = lb0.copy(), ub0.copy()
            gl[out["chgcol"][a:e]] = out["chglb"][a:e]
            gu[out["chgcol"][a:e]] = out["chgub"][a:e]
            assert same_bounds(gl, ol) and same_bounds(gu, ou), v
    nodes.close()
    base.close()
    mat.close()


def test_all_instances_packed(gpu_ctx):
    parts = [sa.read_mps(p)[0] for p in MPS]
    pm = sa.PackedMatrix(gpu_ctx, parts)
    lb, ub, res, rnd, st = pm.propagate_packed()
    for i, t in enumerate(parts):
        ol, ou, orr = orc.propagate_jacobi(orc.OMatrix.from_instance(t), t.lb, t.ub)
        c0, c1 = pm.coloff[i], pm.coloff[i + 1]
        assert (res[i] == sa.SAP_CUTOFF) == bool(orr["cutoff"]) and rnd[i] == orr["rounds"], i
        assert same_bounds(lb[c0:c1], ol) and same_bounds(ub[c0:c1], ou), i
    pm.close()


def test_plugin_on_bell5(gpu_ctx):
    inst, _ = sa.read_mps(os.path.join(HERE, "golden", "mps", "bell5.mps"))
    ol, ou, orr = orc.propagate_jacobi(orc.OMatrix.from_instance(inst), inst.lb, inst.ub)
    h = ph.Host(inst)
    assert h.call("th_initpre") == ph.SCIP_OKAY
    rc, res, nchg = h.presol()
    assert rc == ph.SCIP_OKAY
    glb, gub = h.bounds(glob=True)
    assert same_bounds(glb, ol) and same_bounds(gub, ou)
    assert nchg == int((ol != inst.lb).sum() + (ou != inst.ub).sum())
    assert h.call("th_exitpre") == ph.SCIP_OKAY
    assert h.close() == ph.SCIP_OKAY
