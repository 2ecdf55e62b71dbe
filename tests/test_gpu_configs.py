"""The benchmark workloads at their stated sizes, on the GPU, against the oracle.

  mixed 1M x 1M (seed 2): the large single instance, literal sides (cut off in round 1) and
      sides around a planted feasible point (7 rounds) -- bit-exact with the Jacobi twin;
  covering/packing 200k x 300k (seed 3) with 1,024 node vectors fixing 1..50 binaries each,
      propagated in one call from the resident root fixpoint -- every node against the twin;
  512 uneven instances (seed 4), packed into one matrix and propagated in one launch per round
      -- a sample spanning the size law (the largest included) against the twin, and every
      instance's verdict and round count against the engine on that instance alone.

On the mixed shapes the device fixpoint is also compared with SCIP's own order
(orc_propagate_scip): integer bounds exactly, continuous bounds within 1e-9 relative, and the
fractions that agree must stay at the level measured for these seeds.  Where they differ, the
difference is one of order (the 5 % acceptance filter of numerics/boundstreps makes the
reachable fixpoint order dependent): the device's fixpoint is a fixpoint of SCIP's own
propagation -- SCIP order restarted from the device bounds changes nothing.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import scip_amd as sa
from oracle import oracle as orc
from tests.util import apply_fixings

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def rel_close(a, b, tol=1e-9):
    """|a - b| <= tol * max(1, |a|, |b|), elementwise; equal infinities count as close."""
    with np.errstate(invalid="ignore"):
        d = np.abs(a - b) <= tol * np.maximum(1.0, np.maximum(np.abs(a), np.abs(b)))
    return d | (a == b)


def scip_order_agreement(inst, gl, gu):
    """Fractions of integer columns whose bounds equal SCIP order's exactly, and of continuous
    columns within 1e-9; asserts that SCIP order restarted from the device bounds changes
    nothing (the device fixpoint is a fixpoint of SCIP's propagation)."""
    om = orc.OMatrix.from_instance(inst)
    sl, su, sr = orc.propagate_scip(om, inst.lb, inst.ub)
    assert not sr["cutoff"]
    _, _, again = orc.propagate_scip(om, gl, gu)
    assert again["nchgbds"] == 0 and not again["cutoff"], again
    it = inst.integral.astype(bool)
    fi = float(np.mean((gl[it] == sl[it]) & (gu[it] == su[it]))) if it.any() else 1.0
    fc = float(np.mean(rel_close(gl[~it], sl[~it]) & rel_close(gu[~it], su[~it]))) if (~it).any() else 1.0
    return fi, fc


def propagate_and_check(ctx, inst, name):
    say(f"[case] {name} rows={inst.nrows} cols={inst.ncols} nnz={inst.nnz}")
    om = orc.OMatrix.from_instance(inst)
    ol, ou, orr = orc.propagate_jacobi(om, inst.lb, inst.ub)
    mat = sa.Matrix(ctx, inst)
    gl, gu, gs = mat.propagate()
    mat.close()
    assert (gs["result"] == sa.SAP_CUTOFF) == bool(orr["cutoff"])
    assert gs["rounds"] == orr["rounds"]
    if not orr["cutoff"]:
        assert np.array_equal(gl, ol) and np.array_equal(gu, ou)
        assert gs["nchgbds"] == orr["nchgbds"] and gs["nrecompute"] == orr["nrecompute"]
        assert gs["rows_swept"] == orr["rows_swept"] and gs["nnz_swept"] == orr["nnz_swept"]
    return gl, gu, gs, orr


def test_mixed_1m_literal(gpu_ctx):
    inst = sa.gen_mixed(1_000_000, 1_000_000, seed=2)
    _, _, gs, orr = propagate_and_check(gpu_ctx, inst, "mixed1M literal")
    assert orr["cutoff"] and orr["rounds"] == 1 and gs["result"] == sa.SAP_CUTOFF


def test_mixed_1m_planted_and_scip_order(gpu_ctx):
    inst = sa.gen_mixed(1_000_000, 1_000_000, seed=2, planted=True)
    gl, gu, gs, orr = propagate_and_check(gpu_ctx, inst, "mixed1M planted")
    assert gs["result"] == sa.SAP_REDUCEDDOM and gs["rounds"] == 7
    fi, fc = scip_order_agreement(inst, gl, gu)
    say(f"[scip order] mixed1M planted: integer bounds exact {fi:.4f}, continuous within 1e-9 {fc:.4f}")
    assert fi >= 0.96 and fc >= 0.69   # 0.9669 and 0.6975 for this seed


@pytest.mark.parametrize("mode", [sa.SAP_MODE_SPARSE, sa.SAP_MODE_DENSE])
def test_covering_packing_1024_nodes(gpu_ctx, mode):
    inst = sa.gen_covering_packing(200_000, 300_000, 10, seed=3)
    om = orc.OMatrix.from_instance(inst)
    rl, ru, rr = orc.propagate_jacobi(om, inst.lb, inst.ub)
    assert not rr["cutoff"]
    ptr, cols, vals = sa.gen_node_fixings(inst, 1024, 1, 50, seed=3)
    say(f"[case] cover200k nodes mode={mode} fixings={len(cols)}")
    mat = sa.Matrix(gpu_ctx, inst)
    base = sa.Bounds(mat, 1)
    base.upload(rl, ru)
    nodes = sa.Bounds(mat, 1024, mode)
    out = nodes.propagate_nodes(base, 0, ptr, cols, vals, vals, sa.default_params(flags=sa.SAP_FLAG_BASE_IS_FIXPOINT))
    nodes.close()
    base.close()
    mat.close()

    def one(v):
        lb0, ub0 = apply_fixings(rl, ru, ptr, cols, vals, v)
        ol, ou, orr = orc.propagate_jacobi_marked(om, lb0, ub0, orc.rows_of_columns(om, cols[ptr[v]:ptr[v + 1]]))
        bad = []
        if (out["result"][v] == sa.SAP_CUTOFF) != bool(orr["cutoff"]):
            bad.append("verdict")
        elif not orr["cutoff"]:
            a, e = out["chgptr"][v], out["chgptr"][v + 1]
            gl, gu = lb0.copy(), ub0.copy()
            gl[out["chgcol"][a:e]] = out["chglb"][a:e]
            gu[out["chgcol"][a:e]] = out["chgub"][a:e]
            changed = np.nonzero((ol != lb0) | (ou != ub0))[0]
            if not (np.array_equal(gl, ol) and np.array_equal(gu, ou)):
                bad.append("bounds")
            if not np.array_equal(out["chgcol"][a:e], changed):
                bad.append("change list")
            if out["rounds"][v] != orr["rounds"]:
                bad.append(f"rounds {out['rounds'][v]} vs {orr['rounds']}")
        return v, bad, bool(orr["cutoff"])

    with ThreadPoolExecutor(16) as ex:
        res = list(ex.map(one, range(1024)))
    bad = [(v, b) for v, b, _ in res if b]
    assert not bad, bad[:10]
    ncut = sum(c for _, _, c in res)
    assert 0 < ncut < 1024


def _sample_indices(sizes, k=10):
    order = np.argsort(sizes, kind="stable")
    pick = set(int(order[int(q)]) for q in np.linspace(0, len(sizes) - 1, k))
    return sorted(pick | {int(np.argmax(sizes)), int(np.argmin(sizes))})


@pytest.fixture(scope="module")
def uneven512():
    parts = sa.gen_uneven_instances(512, 4, threads=16)
    yield parts
    del parts


def test_uneven_512_packed(gpu_ctx, uneven512):
    sizes = sa.uneven_sizes(512, 4)
    parts = uneven512
    say(f"[case] uneven512 rows={int(sizes.sum())} nnz={sum(p.nnz for p in parts)}")
    pm = sa.PackedMatrix(gpu_ctx, parts)
    plb, pub, pres, prnd, pst = pm.propagate_packed()
    coloff = pm.coloff
    pm.close()
    # every instance alone on the engine: same verdict and rounds (and bounds)
    for i, p in enumerate(parts):
        mat = sa.Matrix(gpu_ctx, p)
        gl, gu, gs = mat.propagate()
        mat.close()
        assert (gs["result"] == sa.SAP_CUTOFF) == (pres[i] == sa.SAP_CUTOFF), i
        assert gs["rounds"] == prnd[i], (i, gs["rounds"], prnd[i])
        if gs["result"] != sa.SAP_CUTOFF:
            a, e = coloff[i], coloff[i + 1]
            assert np.array_equal(gl, plb[a:e]) and np.array_equal(gu, pub[a:e]), i
    # a sample spanning the size law against the twin, and against SCIP order
    pick = _sample_indices(sizes)
    say(f"[case] uneven512 oracle sample {pick}")

    def one(i):
        p = parts[i]
        ol, ou, orr = orc.propagate_jacobi(orc.OMatrix.from_instance(p), p.lb, p.ub)
        a, e = coloff[i], coloff[i + 1]
        ok = (pres[i] == sa.SAP_CUTOFF) == bool(orr["cutoff"]) and prnd[i] == orr["rounds"]
        if ok and not orr["cutoff"]:
            ok = np.array_equal(plb[a:e], ol) and np.array_equal(pub[a:e], ou)
            fi, fc = scip_order_agreement(p, plb[a:e], pub[a:e])
        else:
            fi = fc = None
        return i, ok, fi, fc

    with ThreadPoolExecutor(8) as ex:
        res = list(ex.map(one, pick))
    for i, ok, fi, fc in res:
        say(f"[scip order] uneven instance {i} rows={parts[i].nrows}: integer exact {fi}, continuous {fc}")
        assert ok, i
        if fi is not None:
            assert fi >= 0.99 and fc >= 0.87, (i, fi, fc)   # 0.9966..0.9995 and 0.8765..0.8886


def _same_request(a, b, i):
    assert a["result"][i] == b["result"][i], i
    assert a["rounds"][i] == b["rounds"][i], i
    if a["result"][i] == sa.SAP_CUTOFF:
        return
    x0, x1 = a["chgptr"][i], a["chgptr"][i + 1]
    y0, y1 = b["chgptr"][i], b["chgptr"][i + 1]
    assert np.array_equal(a["chgcol"][x0:x1], b["chgcol"][y0:y1]), i
    assert np.array_equal(a["chglb"][x0:x1], b["chglb"][y0:y1]) and np.array_equal(a["chgub"][x0:x1], b["chgub"][y0:y1])


def test_serving_mix_slice(gpu_ctx, uneven512):
    # two seconds of the 2,000 req/s serving mix (covering/packing nodes and Zipf-popular
    # uneven-instance nodes) under a budget of 30 % of the total footprint, from a cold start:
    # problems are uploaded on demand and evicted while others run.  Every request gets the
    # same verdict, rounds and changed bounds in the mix as alone, and a sample equals the
    # Jacobi twin on that node alone.
    cp = sa.gen_covering_packing(200_000, 300_000, 10, seed=3)
    probs = [cp] + list(uneven512)
    arrival, prob, ptr, col, val, _ = sa.gen_serving_trace(probs, 2000.0, 2.0, seed=5)
    n = len(arrival)
    say(f"[case] serving slice {n} requests over {len(probs)} problems")

    def replay(mode):
        srv = sa.Server(0, 4, 1 << 50, 8)
        for p in probs:
            srv.add_problem(p)
        srv.set_slots(0, 256)
        foot = sum(srv.footprint(i) for i in range(len(probs)))
        srv.set_budget(int(0.3 * foot))
        out = srv.replay(arrival, prob, ptr, col, val, val, mode=mode, want_changes=True)
        srv.close()
        return out

    mix = replay(sa.SAP_SERVE_CONTINUOUS)
    alone = replay(sa.SAP_SERVE_PLAIN)
    say(f"[info] slice mix: {mix['stats']}")
    assert mix["stats"]["evictions"] > 0 and mix["stats"]["max_running"] > 1
    for i in range(n):
        _same_request(mix, alone, i)
    # the twin on a sample of nodes, from each problem's root fixpoint
    roots = {}

    def one(i):
        q = int(prob[i])
        p = probs[q]
        om = orc.OMatrix.from_instance(p)
        if q not in roots:
            roots[q] = orc.propagate_jacobi(om, p.lb, p.ub)
        rl, ru, rr = roots[q]
        if rr["cutoff"]:
            return i, mix["result"][i] == sa.SAP_CUTOFF
        fc = col[ptr[i]:ptr[i + 1]]
        lb0, ub0 = rl.copy(), ru.copy()
        lb0[fc] = val[ptr[i]:ptr[i + 1]]
        ub0[fc] = val[ptr[i]:ptr[i + 1]]
        ol, ou, orr = orc.propagate_jacobi_marked(om, lb0, ub0, orc.rows_of_columns(om, fc))
        if (mix["result"][i] == sa.SAP_CUTOFF) != bool(orr["cutoff"]) or mix["rounds"][i] != orr["rounds"]:
            return i, False
        if orr["cutoff"]:
            return i, True
        a, e = mix["chgptr"][i], mix["chgptr"][i + 1]
        gl, gu = lb0.copy(), ub0.copy()
        gl[mix["chgcol"][a:e]] = mix["chglb"][a:e]
        gu[mix["chgcol"][a:e]] = mix["chgub"][a:e]
        return i, bool(np.array_equal(gl, ol) and np.array_equal(gu, ou))

    sample = list(range(0, n, 40))
    res = [one(i) for i in sample]
    bad = [i for i, ok in res if not ok]
    assert not bad, bad[:10]
