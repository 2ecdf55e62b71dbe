"""GPU parity of the batched forms: node vectors from a resident base with sparse results
(sap_propagate_nodes), independent instances packed into one launch (sap_propagate_packed)
and continuous batching (sap_slots_*).  Every result is compared with the oracle's Jacobi
twin run on that node / instance alone, and with the engine's own one-at-a-time result:
bit-exact bounds, identical verdicts and round counts.
"""
import numpy as np
import pytest

import scip_amd as sa
from oracle import oracle as orc
from tests.util import apply_fixings, make_instance, random_binary_instance, random_fixings, same_bounds

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def node_problem(seed, nrows=2500, ncols=1800, nvec=40, maxfix=8):
    rng = np.random.default_rng(seed)
    inst = random_binary_instance(rng, nrows=nrows, ncols=ncols, maxlen=12)
    om = orc.OMatrix.from_instance(inst)
    rl, ru, rr = orc.propagate_jacobi(om, inst.lb, inst.ub)
    assert not rr["cutoff"]
    free = np.nonzero(rl < ru)[0]
    ptr, cols, vals = random_fixings(rng, len(free), nvec, maxfix)
    cols = free[cols].astype(np.int32)
    return inst, om, rl, ru, ptr, cols, vals


def apply_changes(lb, ub, out, v):
    lb = lb.copy()
    ub = ub.copy()
    a, e = out["chgptr"][v], out["chgptr"][v + 1]
    lb[out["chgcol"][a:e]] = out["chglb"][a:e]
    ub[out["chgcol"][a:e]] = out["chgub"][a:e]
    return lb, ub


def check_node_out(om, rl, ru, ptr, cols, vals, out, v, check_rounds, oparams=None):
    lb0, ub0 = apply_fixings(rl, ru, ptr, cols, vals, v)
    if oparams is None:
        ol, ou, orr = orc.propagate_jacobi(om, lb0, ub0)
    else:
        ol, ou, orr = orc.propagate_jacobi(om, lb0, ub0, oparams)
    assert (out["result"][v] == sa.SAP_CUTOFF) == bool(orr["cutoff"]), v
    if orr["cutoff"]:
        assert out["chgptr"][v + 1] == out["chgptr"][v]
        return
    gl, gu = apply_changes(lb0, ub0, out, v)
    assert same_bounds(gl, ol) and same_bounds(gu, ou), v
    # the list holds exactly the columns whose bounds changed, ascending
    a, e = out["chgptr"][v], out["chgptr"][v + 1]
    changed = np.nonzero((ol != lb0) | (ou != ub0))[0]
    assert np.array_equal(out["chgcol"][a:e], changed), v
    assert out["result"][v] == (sa.SAP_REDUCEDDOM if len(changed) else sa.SAP_DIDNOTFIND)
    if check_rounds:
        assert out["rounds"][v] == orr["rounds"], (v, out["rounds"][v], orr["rounds"])


@pytest.mark.parametrize("mode", [sa.SAP_MODE_DENSE, sa.SAP_MODE_SPARSE])
@pytest.mark.parametrize("flags", [0, sa.SAP_FLAG_BASE_IS_FIXPOINT])
def test_nodes_from_resident_base(gpu_ctx, mode, flags):
    inst, om, rl, ru, ptr, cols, vals = node_problem(5)
    nvec = len(ptr) - 1
    say(f"[case] nodes mode={mode} flags={flags} nvec={nvec}")
    mat = sa.Matrix(gpu_ctx, inst)
    base = sa.Bounds(mat, 1)
    base.upload(rl, ru)
    nodes = sa.Bounds(mat, nvec, mode)
    out = nodes.propagate_nodes(base, 0, ptr, cols, vals, vals, sa.default_params(flags=flags))
    for v in range(nvec):
        check_node_out(om, rl, ru, ptr, cols, vals, out, v, check_rounds=(flags == 0))
    # a second call on the same vectors gives the same answer (logs are cleared per call)
    out2 = nodes.propagate_nodes(base, 0, ptr, cols, vals, vals, sa.default_params(flags=flags))
    for k in ("result", "rounds", "chgptr", "chgcol", "chglb", "chgub"):
        assert np.array_equal(out[k], out2[k]), k
    # other nodes on the same vectors (each vector is restored from its previous node), then
    # a changed base (full copy)
    perm = np.random.default_rng(2).permutation(nvec)
    p2 = np.zeros(nvec + 1, dtype=np.int64)
    c2, v2 = [], []
    for i, v in enumerate(perm):
        c2.extend(cols[ptr[v]:ptr[v + 1]])
        v2.extend(1.0 - vals[ptr[v]:ptr[v + 1]])
        p2[i + 1] = len(c2)
    c2 = np.array(c2, dtype=np.int32)
    v2 = np.array(v2)
    out3 = nodes.propagate_nodes(base, 0, p2, c2, v2, v2, sa.default_params(flags=flags))
    for v in range(nvec):
        check_node_out(om, rl, ru, p2, c2, v2, out3, v, check_rounds=(flags == 0))
    base.upload(inst.lb, inst.ub)   # the root bounds, not the fixpoint: every vector is copied
    out4 = nodes.propagate_nodes(base, 0, p2, c2, v2, v2, sa.default_params(flags=0))
    for v in range(nvec):
        check_node_out(om, inst.lb, inst.ub, p2, c2, v2, out4, v, check_rounds=True)
    nodes.close()
    base.close()
    mat.close()


@pytest.mark.parametrize("mode", [sa.SAP_MODE_DENSE, sa.SAP_MODE_SPARSE])
@pytest.mark.parametrize("maxrounds", [1, 2, 3])
def test_nodes_maxrounds(gpu_ctx, mode, maxrounds):
    # a batched call launches its first round and a two-round group before it first waits:
    # the round cap still ends every vector after maxrounds rounds
    inst, om, rl, ru, ptr, cols, vals = node_problem(15)
    nvec = len(ptr) - 1
    say(f"[case] nodes maxrounds={maxrounds} mode={mode} nvec={nvec}")
    mat = sa.Matrix(gpu_ctx, inst)
    base = sa.Bounds(mat, 1)
    base.upload(rl, ru)
    nodes = sa.Bounds(mat, nvec, mode)
    out = nodes.propagate_nodes(base, 0, ptr, cols, vals, vals, sa.default_params(maxrounds=maxrounds))
    assert out["rounds"].max() <= maxrounds
    for v in range(nvec):
        check_node_out(om, rl, ru, ptr, cols, vals, out, v, check_rounds=True,
                       oparams=orc.params(maxrounds=maxrounds))
    full = nodes.propagate_nodes(base, 0, ptr, cols, vals, vals, sa.default_params())
    assert full["rounds"].max() > maxrounds   # the cap binds for some node
    nodes.close()
    base.close()
    mat.close()


@pytest.mark.parametrize("mode", [sa.SAP_MODE_DENSE, sa.SAP_MODE_SPARSE])
def test_nodes_host_loop_and_probe_equal_graph_rounds(gpu_ctx, mode):
    # the host-driven round loop reads the outcomes and change lists after its rounds, the
    # graph-driven one with its last wait: same answers
    inst, om, rl, ru, ptr, cols, vals = node_problem(16)
    nvec = len(ptr) - 1
    say(f"[case] nodes host loop mode={mode} nvec={nvec}")
    mat = sa.Matrix(gpu_ctx, inst)
    base = sa.Bounds(mat, 1)
    base.upload(rl, ru)
    nodes = sa.Bounds(mat, nvec, mode)
    for flags in (0, sa.SAP_FLAG_BASE_IS_FIXPOINT):
        a = nodes.propagate_nodes(base, 0, ptr, cols, vals, vals, sa.default_params(flags=flags))
        b = nodes.propagate_nodes(base, 0, ptr, cols, vals, vals,
                                  sa.default_params(flags=flags | sa.SAP_FLAG_HOSTLOOP))
        for k in ("result", "chgptr", "chgcol", "chglb", "chgub"):
            assert np.array_equal(a[k], b[k]), (flags, k)
        if flags == 0:
            assert np.array_equal(a["rounds"], b["rounds"])
            for v in range(nvec):
                check_node_out(om, rl, ru, ptr, cols, vals, b, v, check_rounds=True)
    free = np.nonzero(rl < ru)[0].astype(np.int32)[:24]
    pa = nodes.probe(base, 0, free, sa.default_params())
    pb = nodes.probe(base, 0, free, sa.default_params(flags=sa.SAP_FLAG_HOSTLOOP))
    for k in ("verdict", "rounds", "dedptr", "dedcol", "dlb0", "dub0", "dlb1", "dub1"):
        assert np.array_equal(pa[k], pb[k]), k
    nodes.close()
    base.close()
    mat.close()


def tight_long_rows(rng, nrows, ncols):
    """Binary rows, three in four of 2..399 nonzeros, with sides at most 3 from the activity
    of a hidden 0/1 point: heavy fixings cut many nodes off, in different rounds."""
    xs = rng.integers(0, 2, size=ncols).astype(float)
    rows = []
    for i in range(nrows):
        n = int(rng.integers(2, 400)) if i % 4 else int(rng.integers(2, 40))
        cs = rng.choice(ncols, size=n, replace=False)
        a = rng.integers(1, 6, size=n).astype(float) * rng.choice([-1.0, 1.0], size=n)
        act = float(a @ xs[cs])
        slack = float(rng.integers(0, 4))
        entries = [(int(c), float(x)) for c, x in zip(cs, a)]
        rows.append((entries, -1e20, act + slack) if i % 2 else (entries, act - slack, 1e20))
    return make_instance(rows, [0.0] * ncols, [1.0] * ncols, [1] * ncols)


@pytest.mark.parametrize("mode", [sa.SAP_MODE_DENSE, sa.SAP_MODE_SPARSE])
def test_nodes_cut_off_while_long_rows_are_swept(gpu_ctx, mode):
    # rows longer than longrow = 64 are swept by a whole workgroup; vectors of the same batch
    # are cut off in rounds 1..4 while these sweeps run, and the other vectors' results must
    # not depend on it
    rng = np.random.default_rng(41)
    inst = tight_long_rows(rng, 800, 600)
    om = orc.OMatrix.from_instance(inst)
    rl, ru, rr = orc.propagate_jacobi(om, inst.lb, inst.ub)
    assert not rr["cutoff"]
    free = np.nonzero(rl < ru)[0]
    ptr, cols, vals = random_fixings(rng, len(free), 64, 150)
    cols = free[cols].astype(np.int32)
    say(f"[case] long-row nodes mode={mode}")
    gpu_ctx.set_tuning(3, 64, 50, 0, 4)
    try:
        mat = sa.Matrix(gpu_ctx, inst)
        base = sa.Bounds(mat, 1)
        base.upload(rl, ru)
        nodes = sa.Bounds(mat, 64, mode)
        for _ in range(2):
            out = nodes.propagate_nodes(base, 0, ptr, cols, vals, vals, sa.default_params())
            ncut = int((out["result"] == sa.SAP_CUTOFF).sum())
            assert 10 <= ncut <= 54, ncut
            for v in range(64):
                check_node_out(om, rl, ru, ptr, cols, vals, out, v, check_rounds=True)
        nodes.close()
        base.close()
        mat.close()
    finally:
        gpu_ctx.set_tuning()


def test_redundant_row_flags_over_many_calls(gpu_ctx):
    # the locally-redundant row flags carry a generation per call and are cleared only every
    # 255 calls: 300 calls on one bounds object, alternating two node sets and two batch sizes
    # (so vectors sit out calls), must each give the answer of the first call with those inputs
    inst, om, rl, ru, ptr, cols, vals = node_problem(7, nvec=40, maxfix=12)
    vals2 = 1.0 - vals
    mat = sa.Matrix(gpu_ctx, inst)
    base = sa.Bounds(mat, 1)
    base.upload(rl, ru)
    nodes = sa.Bounds(mat, 40)
    p = sa.default_params(flags=sa.SAP_FLAG_BASE_IS_FIXPOINT)
    cases = [(ptr, vals), (ptr, vals2), (ptr[:21], vals), (ptr[:21], vals2)]
    first = [None] * len(cases)
    say("[case] 300 node calls on one bounds object")
    for call in range(300):
        k = call % len(cases)
        pp, vv = cases[k]
        nv = len(pp) - 1
        out = nodes.propagate_nodes(base, 0, pp, cols[:pp[-1]], vv[:pp[-1]], vv[:pp[-1]], p)
        if first[k] is None:
            first[k] = out
            for v in range(nv):
                check_node_out(om, rl, ru, pp, cols, vv, out, v, check_rounds=False)
            continue
        for key in ("result", "rounds", "chgptr", "chgcol", "chglb", "chgub"):
            assert np.array_equal(out[key], first[k][key]), (call, key)
    nodes.close()
    base.close()
    mat.close()


def test_nodes_change_capacity(gpu_ctx):
    inst, om, rl, ru, ptr, cols, vals = node_problem(6, nvec=16)
    mat = sa.Matrix(gpu_ctx, inst)
    base = sa.Bounds(mat, 1)
    base.upload(rl, ru)
    nodes = sa.Bounds(mat, 16)
    full = nodes.propagate_nodes(base, 0, ptr, cols, vals, vals, sa.default_params(flags=sa.SAP_FLAG_BASE_IS_FIXPOINT))
    need = int(full["chgptr"][-1])
    say(f"[case] nodes capacity need={need}")
    if need > 0:
        with pytest.raises(sa.SapError) as e:
            nodes.propagate_nodes(base, 0, ptr, cols, vals, vals,
                                  sa.default_params(flags=sa.SAP_FLAG_BASE_IS_FIXPOINT), chgcap=need - 1)
        assert e.value.rc == sa.SAP_NOMEMORY
    exact = nodes.propagate_nodes(base, 0, ptr, cols, vals, vals,
                                  sa.default_params(flags=sa.SAP_FLAG_BASE_IS_FIXPOINT), chgcap=need)
    assert np.array_equal(exact["chgcol"], full["chgcol"])
    verd = nodes.propagate_nodes(base, 0, ptr, cols, vals, vals,
                                 sa.default_params(flags=sa.SAP_FLAG_BASE_IS_FIXPOINT), want_changes=False)
    assert np.array_equal(verd["result"], full["result"])
    nodes.close()
    base.close()
    mat.close()


def test_batch_verdicts_without_bound_download(gpu_ctx):
    inst, om, rl, ru, ptr, cols, vals = node_problem(7, nvec=24)
    mat = sa.Matrix(gpu_ctx, inst)
    bnd = sa.Bounds(mat, 24)
    p = sa.default_params(flags=sa.SAP_FLAG_BASE_IS_FIXPOINT)
    a = bnd.propagate_batch(rl, ru, ptr, cols, vals, vals, p, want_bounds=False)
    b = bnd.propagate_batch(rl, ru, ptr, cols, vals, vals, p, want_bounds=True)
    assert np.array_equal(a["result"], b["result"])
    for v in range(24):
        lb0, ub0 = apply_fixings(rl, ru, ptr, cols, vals, v)
        if b["result"][v] != sa.SAP_CUTOFF:
            ch = (b["lb"][v] != lb0).any() or (b["ub"][v] != ub0).any()
            assert b["result"][v] == (sa.SAP_REDUCEDDOM if ch else sa.SAP_DIDNOTFIND)
    bnd.close()
    mat.close()


# --- packed independent instances --------------------------------------------------------------
def packed_parts():
    parts = []
    sizes = [(300, 450), (1200, 1800), (50, 75), (5000, 7500), (800, 1200), (20000, 30000), (2, 3)]
    for i, (m, n) in enumerate(sizes):
        planted = (i % 3) != 1
        parts.append(sa.gen_mixed(m, n, seed=4000 + i, lenmode=1, minlen=4, maxlen=64, planted=planted))
    rng = np.random.default_rng(9)
    parts.append(random_binary_instance(rng, nrows=600, ncols=500, maxlen=10))
    parts.append(sa.gen_covering_packing(400, 1000, 10, seed=11))
    # literal sides at this size: infeasible, found in round 1
    parts.insert(3, sa.gen_mixed(30000, 30000, seed=2, planted=False))
    return parts


def test_packed_instances_match_alone(gpu_ctx):
    parts = packed_parts()
    say(f"[case] packed {len(parts)} instances, rows {[p.nrows for p in parts]}")
    pm = sa.PackedMatrix(gpu_ctx, parts)
    lb, ub, res, rnd, st = pm.propagate_packed()
    ncut = 0
    for i, t in enumerate(parts):
        c0, c1 = pm.coloff[i], pm.coloff[i + 1]
        ol, ou, orr = orc.propagate_jacobi(orc.OMatrix.from_instance(t), t.lb, t.ub)
        assert (res[i] == sa.SAP_CUTOFF) == bool(orr["cutoff"]), i
        assert rnd[i] == orr["rounds"], (i, rnd[i], orr["rounds"])
        if orr["cutoff"]:
            ncut += 1
            continue
        assert same_bounds(lb[c0:c1], ol) and same_bounds(ub[c0:c1], ou), i
        assert res[i] == (sa.SAP_REDUCEDDOM if orr["nchgbds"] > 0 else sa.SAP_DIDNOTFIND), i
        # the engine alone on the same instance
        m1 = sa.Matrix(gpu_ctx, t)
        gl, gu, gs = m1.propagate()
        assert same_bounds(gl, ol) and same_bounds(gu, ou) and gs["rounds"] == rnd[i]
        m1.close()
    assert 0 < ncut < len(parts)
    pm.close()


def test_packed_maxrounds_per_instance(gpu_ctx):
    parts = packed_parts()[:5]
    pm = sa.PackedMatrix(gpu_ctx, parts)
    lb, ub, res, rnd, st = pm.propagate_packed(params=sa.default_params(maxrounds=2))
    for i, t in enumerate(parts):
        c0, c1 = pm.coloff[i], pm.coloff[i + 1]
        ol, ou, orr = orc.propagate_jacobi(orc.OMatrix.from_instance(t), t.lb, t.ub, orc.params(maxrounds=2))
        assert (res[i] == sa.SAP_CUTOFF) == bool(orr["cutoff"]) and rnd[i] == orr["rounds"], i
        if not orr["cutoff"]:
            assert same_bounds(lb[c0:c1], ol) and same_bounds(ub[c0:c1], ou), i
    pm.close()


def test_packed_rejects_shared_columns(gpu_ctx):
    parts = packed_parts()[:2]
    inst, rowoff, coloff = sa.pack_instances(parts)
    bad = inst.colidx.copy()
    bad[0] = coloff[1]   # row 0 of instance 0 points into instance 1
    L = sa.lib()
    import ctypes as C
    h = C.c_void_p()
    rc = L.sap_matrix_create_packed(gpu_ctx.h, 2, sa._i64p(rowoff), sa._i64p(coloff), inst.nrows, inst.ncols,
                                    sa._i64p(inst.rowptr), sa._i32p(bad), sa._dp(inst.vals), sa._dp(inst.lhs),
                                    sa._dp(inst.rhs), sa._u8p(inst.integral), sa._dp(inst.lb), sa._dp(inst.ub),
                                    C.byref(h))
    assert rc == sa.SAP_INVALIDDATA


# --- continuous batching -----------------------------------------------------------------------
def run_slots(ctx, inst, rl, ru, ptr, cols, vals, nslots, mode, order, admit_per_step, rounds):
    """Admit the nodes in `order`, at most admit_per_step per step, into nslots slots;
    returns per node (result, rounds, lb, ub) reconstructed from the collected changes."""
    mat = sa.Matrix(ctx, inst)
    base = sa.Bounds(mat, 1)
    base.upload(rl, ru)
    sl = sa.Bounds(mat, nslots, mode)
    free = list(range(nslots))
    slot_node = {}
    queue = list(order)
    out = {}
    steps = 0
    while queue or slot_node:
        adm_nodes, adm_slots = [], []
        while queue and free and len(adm_nodes) < admit_per_step:
            adm_nodes.append(queue.pop(0))
            adm_slots.append(free.pop(0))
        if adm_nodes:
            optr = [0]
            ocol, oval = [], []
            for v in adm_nodes:
                ocol.extend(cols[ptr[v]:ptr[v + 1]])
                oval.extend(vals[ptr[v]:ptr[v + 1]])
                optr.append(len(ocol))
            sl.admit(adm_slots, base, 0, optr, ocol, oval, oval)
            for s_, v in zip(adm_slots, adm_nodes):
                slot_node[s_] = v
        fin, st = sl.step(rounds=rounds)
        steps += 1
        assert steps < 10000
        if len(fin):
            c = sl.collect(fin)
            for i, s_ in enumerate(fin):
                v = slot_node.pop(int(s_))
                lb0, ub0 = apply_fixings(rl, ru, ptr, cols, vals, v)
                a, e = c["chgptr"][i], c["chgptr"][i + 1]
                lb1, ub1 = lb0.copy(), ub0.copy()
                lb1[c["chgcol"][a:e]] = c["chglb"][a:e]
                ub1[c["chgcol"][a:e]] = c["chgub"][a:e]
                out[v] = (int(c["result"][i]), int(c["rounds"][i]), lb1, ub1)
                free.append(int(s_))
    sl.close()
    base.close()
    mat.close()
    return out


@pytest.mark.parametrize("mode", [sa.SAP_MODE_SPARSE, sa.SAP_MODE_DENSE])
def test_slots_match_oracle_and_alone(gpu_ctx, mode):
    inst, om, rl, ru, ptr, cols, vals = node_problem(8, nvec=48)
    nvec = len(ptr) - 1
    say(f"[case] slots mode={mode} nodes={nvec}")
    mixed = run_slots(gpu_ctx, inst, rl, ru, ptr, cols, vals, nslots=12, mode=mode,
                      order=list(np.random.default_rng(1).permutation(nvec)), admit_per_step=5,
                      rounds=sa.SAP_STEP_SHORT)
    assert sorted(mixed) == list(range(nvec))
    for v in range(nvec):
        lb0, ub0 = apply_fixings(rl, ru, ptr, cols, vals, v)
        ol, ou, orr = orc.propagate_jacobi(om, lb0, ub0)
        res, rnd, gl, gu = mixed[v]
        assert (res == sa.SAP_CUTOFF) == bool(orr["cutoff"]), v
        if not orr["cutoff"]:
            assert same_bounds(gl, ol) and same_bounds(gu, ou), v
    # alone: one slot, one node at a time, long steps
    for v in range(0, nvec, 7):
        alone = run_slots(gpu_ctx, inst, rl, ru, ptr, cols, vals, nslots=1, mode=mode, order=[v],
                          admit_per_step=1, rounds=sa.SAP_STEP_LONG)
        a = alone[v]
        m = mixed[v]
        assert a[0] == m[0] and a[1] == m[1], v
        if a[0] != sa.SAP_CUTOFF:
            assert same_bounds(a[2], m[2]) and same_bounds(a[3], m[3]), v


def test_slots_full_sweep_admission(gpu_ctx):
    # without the fixpoint flag the admitted slot sweeps every row first: root propagation
    inst = sa.gen_mixed(3000, 3000, seed=12, planted=True)
    ol, ou, orr = orc.propagate_jacobi(orc.OMatrix.from_instance(inst), inst.lb, inst.ub)
    mat = sa.Matrix(gpu_ctx, inst)
    base = sa.Bounds(mat, 1)
    base.upload(inst.lb, inst.ub)
    sl = sa.Bounds(mat, 4)
    sl.admit([2], base, 0, [0, 0], [], [], [], flags=0)
    fin = []
    for _ in range(1000):
        f, _ = sl.step(rounds=sa.SAP_STEP_LONG)
        fin.extend(int(x) for x in f)
        if fin:
            break
    assert fin == [2]
    c = sl.collect([2])
    assert (c["result"][0] == sa.SAP_CUTOFF) == bool(orr["cutoff"]) and c["rounds"][0] == orr["rounds"]
    gl, gu = sl.download(2)
    assert same_bounds(gl, ol) and same_bounds(gu, ou)
    sl.close()
    base.close()
    mat.close()


def test_slots_api_errors(gpu_ctx):
    inst, om, rl, ru, ptr, cols, vals = node_problem(9, nvec=4)
    mat = sa.Matrix(gpu_ctx, inst)
    base = sa.Bounds(mat, 1)
    base.upload(rl, ru)
    sl = sa.Bounds(mat, 4)
    with pytest.raises(sa.SapError) as e:
        sl.launch()
    assert e.value.rc == sa.SAP_INVALIDCALL   # not yet used as slots
    sl.admit([0, 1], base, 0, ptr[:3], cols[:ptr[2]], vals[:ptr[2]], vals[:ptr[2]])
    with pytest.raises(sa.SapError):
        sl.admit([1], base, 0, [0, 0], [], [], [])   # slot 1 is running
    with pytest.raises(sa.SapError):
        sl.collect([0])   # not finished yet
    with pytest.raises(sa.SapError) as e:
        sl.propagate_batch(rl, ru, ptr, cols, vals, vals)
    assert e.value.rc == sa.SAP_INVALIDCALL   # slots and one-shot calls do not mix
    done = set()
    for _ in range(1000):
        f, _ = sl.step()
        done.update(int(x) for x in f)
        if done == {0, 1}:
            break
    assert done == {0, 1}
    sl.collect([0, 1])
    sl.close()
    base.close()
    mat.close()


def test_parents_outlive_children(gpu_ctx):
    """A context with live matrices, or a matrix with live bound sets, is not freed."""
    import ctypes as C
    ctx = sa.Context(0)
    inst = sa.gen_covering_packing(100, 200, 5, seed=1)
    mat = sa.Matrix(ctx, inst)
    b = sa.Bounds(mat, 2)
    L = sa.lib()
    assert L.sap_matrix_free(mat.h) == sa.SAP_INVALIDCALL
    assert L.sap_context_free(ctx.h) == sa.SAP_INVALIDCALL
    lb, ub, st = mat.propagate()       # still usable
    assert st["rounds"] >= 1
    b.close()
    mat.close()
    ctx.close()


def test_log_clearing_after_uncollected_calls(gpu_ctx):
    # the host keeps a bound on the vectors whose change-log bits may be uncollected: a call
    # that stops for lack of room (SAP_NOMEMORY) leaves them, a smaller call must then still
    # start from clean logs for every vector, and later calls of any size give the same
    # answers as a fresh bound set
    inst, om, rl, ru, ptr, cols, vals = node_problem(11, nvec=16)
    mat = sa.Matrix(gpu_ctx, inst)
    base = sa.Bounds(mat, 1)
    base.upload(rl, ru)
    p = sa.default_params(flags=sa.SAP_FLAG_BASE_IS_FIXPOINT)

    def sub(n):
        return ptr[:n + 1], cols[:ptr[n]], vals[:ptr[n]]

    fresh = {}
    for n in (16, 8, 3):
        f = sa.Bounds(mat, 16)
        sp, sc, sv = sub(n)
        fresh[n] = f.propagate_nodes(base, 0, sp, sc, sv, sv, p)
        f.close()
    need = int(fresh[16]["chgptr"][-1])
    assert need > 1
    nodes = sa.Bounds(mat, 16)
    for n in (16, 8, 16, 3, 8):
        if n == 16:
            sp, sc, sv = sub(16)
            with pytest.raises(sa.SapError) as e:
                nodes.propagate_nodes(base, 0, sp, sc, sv, sv, p, chgcap=1)
            assert e.value.rc == sa.SAP_NOMEMORY
        sp, sc, sv = sub(n)
        out = nodes.propagate_nodes(base, 0, sp, sc, sv, sv, p)
        for key in ("result", "rounds", "chgptr", "chgcol", "chglb", "chgub"):
            assert np.array_equal(out[key], fresh[n][key]), (n, key)
        for v in range(n):
            check_node_out(om, rl, ru, sp, sc, sv, out, v, check_rounds=True)
    nodes.close()
    base.close()
    mat.close()
