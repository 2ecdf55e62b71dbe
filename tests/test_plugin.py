"""The SCIP propagator plugin (scip_amd/plugin/prop_gpulinear.c) driven through the test
double of the SCIP API: registration and error handling on the CPU, and on the GPU the
bounds it hands to SCIP (SCIPinferVarLbProp/UbProp in PROPEXEC, SCIPtightenVarLb/Ub in
PRESOL) equal the engine's and the oracle's fixpoint."""
import os

import numpy as np
import pytest

import scip_amd as sa
from oracle import oracle as orc
from tests import plugin_host as ph
from tests.util import INF, apply_fixings, make_instance, random_binary_instance, random_fixings


def test_testhost_builds_and_links_the_plugin():
    assert os.path.exists(ph.LIBPATH)
    L = ph.lib()
    for sym in ("th_create", "th_include_gpulinear", "th_propexec", "th_presol", "SCIPincludePropGpulinear"):
        assert hasattr(L, sym)


def test_plugin_registers_parameters():
    inst = sa.gen_covering_packing(50, 40, 4, seed=3)
    h = ph.Host(inst)
    assert h.call("th_set_int_param", b"propagating/gpulinear/maxrounds", 5) == ph.SCIP_OKAY
    assert h.call("th_set_int_param", b"propagating/gpulinear/device", 0) == ph.SCIP_OKAY
    assert h.call("th_set_int_param", b"propagating/gpulinear/probing", 1) == ph.SCIP_OKAY
    assert h.call("th_set_int_param", b"propagating/gpulinear/probingvectors", 64) == ph.SCIP_OKAY
    assert h.call("th_set_int_param", b"propagating/gpulinear/nosuch", 0) != ph.SCIP_OKAY
    assert h.close() == ph.SCIP_OKAY


def test_conflict_resolution_without_a_matrix_reports_unexplained_change():
    inst = sa.gen_covering_packing(50, 40, 4, seed=3)
    h = ph.Host(inst)
    for upper in (False, True):
        rc, res = h.resprop(0, upper)
        assert rc == ph.SCIP_OKAY and res == ph.SCIP_DIDNOTFIND
    assert h.close() == ph.SCIP_OKAY


@pytest.mark.skipif(sa.device_count() > 0, reason="checks the no-device error path")
def test_plugin_without_device_returns_a_scip_error():
    inst = sa.gen_covering_packing(50, 40, 4, seed=3)
    h = ph.Host(inst)
    assert h.call("th_initsol") == sa.SAP_INVALIDCALL  # SCIP_INVALIDCALL, no abort
    rc, res = h.propexec()
    assert rc == ph.SCIP_OKAY and res == ph.SCIP_DIDNOTRUN  # no matrix: the propagator does not run
    assert h.close() == ph.SCIP_OKAY


def _oracle(inst, lb, ub, ranged=True):
    """The oracle's fixpoint; PROPEXEC propagates ranged rows (SCIP's default
    constraints/linear/rangedrowpropagation), presolving does not."""
    return orc.propagate_jacobi(orc.OMatrix.from_instance(inst), lb, ub, orc.params(rangedrows=int(ranged)))


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_propexec_root_matches_engine_and_oracle(gpu_ctx, seed):
    inst = sa.gen_mixed(3000, 3000, seed=seed, planted=True)
    ol, ou, orr = _oracle(inst, inst.lb, inst.ub)
    h = ph.Host(inst)
    assert h.call("th_initsol") == ph.SCIP_OKAY
    rc, res = h.propexec()
    assert rc == ph.SCIP_OKAY
    assert not orr["cutoff"]
    changed = (ol != inst.lb).any() or (ou != inst.ub).any()
    assert res == (ph.SCIP_REDUCEDDOM if changed else ph.SCIP_DIDNOTFIND)
    lb, ub = h.bounds()
    assert np.array_equal(lb, ol) and np.array_equal(ub, ou)
    # global bounds are untouched by PROPEXEC
    glb, gub = h.bounds(glob=True)
    assert np.array_equal(glb, inst.lb) and np.array_equal(gub, inst.ub)
    # a second call at the fixpoint finds nothing
    rc, res = h.propexec()
    assert rc == ph.SCIP_OKAY and res == ph.SCIP_DIDNOTFIND
    assert h.call("th_exitsol") == ph.SCIP_OKAY
    assert h.close() == ph.SCIP_OKAY


@pytest.mark.gpu
def test_propexec_nodes_match_oracle(gpu_ctx):
    rng = np.random.default_rng(11)
    inst = random_binary_instance(rng, 400, 300, 12)
    rl, ru, _ = _oracle(inst, inst.lb, inst.ub)
    h = ph.Host(inst)
    assert h.call("th_initsol") == ph.SCIP_OKAY
    free = np.nonzero(rl < ru)[0]
    nvec = 12
    ptr, cols, vals = random_fixings(rng, len(free), nvec, 20)
    cols = free[cols].astype(np.int32)
    ncut = 0
    for k in range(nvec):
        nl, nu = apply_fixings(rl, ru, ptr, cols, vals, k)
        h.set_local(nl, nu)
        rc, res = h.propexec()
        assert rc == ph.SCIP_OKAY
        ol, ou, orr = _oracle(inst, nl, nu)
        if orr["cutoff"]:
            assert res == ph.SCIP_CUTOFF
            ncut += 1
            continue
        lb, ub = h.bounds()
        assert np.array_equal(lb, ol) and np.array_equal(ub, ou), f"node {k}"
    assert ncut < nvec
    assert h.close() == ph.SCIP_OKAY


@pytest.mark.gpu
def test_presol_tightens_global_bounds(gpu_ctx):
    inst = sa.gen_mixed(2000, 2000, seed=4, planted=True)
    ol, ou, _ = _oracle(inst, inst.lb, inst.ub, ranged=False)
    h = ph.Host(inst)
    assert h.call("th_initpre") == ph.SCIP_OKAY
    rc, res, nchg = h.presol()
    assert rc == ph.SCIP_OKAY
    expect = int((ol != inst.lb).sum() + (ou != inst.ub).sum())
    assert nchg == expect
    assert res == (ph.SCIP_SUCCESS if expect else ph.SCIP_DIDNOTFIND)
    glb, gub = h.bounds(glob=True)
    assert np.array_equal(glb, ol) and np.array_equal(gub, ou)
    assert h.call("th_exitpre") == ph.SCIP_OKAY
    assert h.close() == ph.SCIP_OKAY


@pytest.mark.gpu
def test_cutoff_reported(gpu_ctx):
    inst = sa.gen_mixed(30000, 30000, seed=2, planted=False)  # literal sides: infeasible at this size
    _, _, orr = _oracle(inst, inst.lb, inst.ub)
    h = ph.Host(inst)
    assert h.call("th_initsol") == ph.SCIP_OKAY
    rc, res = h.propexec()
    assert rc == ph.SCIP_OKAY
    assert orr["cutoff"] and res == ph.SCIP_CUTOFF
    assert h.close() == ph.SCIP_OKAY


def expected_probing(om, rl, ru, integral):
    """What the plugin's presolve probing hands to SCIP, in order: probes in column order over
    the unfixed binaries of the fixpoint, skipping a column fixed by an earlier probe; a cut-off
    branch fixes the column (prop_probing.c:477-556); otherwise the columns changed in both
    branches go to SCIPanalyzeDeductionsProbing, whose union fixings and tightenings are
    applied as the test host applies them (binary columns here: a tightening is a fixing)."""
    probecols = np.nonzero((integral != 0) & (rl == 0.0) & (ru == 1.0))[0]
    ref = dict(zip(probecols.tolist(), orc.probe(om, rl, ru, probecols)))
    glb, gub = rl.copy(), ru.copy()
    fix, ded = [], []
    for x in probecols.tolist():
        if glb[x] > 0.5 or gub[x] < 0.5:
            continue
        r = ref[x]
        assert r["verdict"] != 3
        if r["verdict"]:
            val = 0.0 if r["verdict"] & 2 else 1.0
            fix.append((x, val))
            glb[x] = gub[x] = val
            continue
        for k, j in enumerate(r["cols"].tolist()):
            (l0, u0), (l1, u1) = r["b0"][k], r["b1"][k]
            ded.append((x, j, l0, u0, l1, u1))
            lo, hi = min(l0, l1), max(u0, u1)
            if lo == hi:
                fix.append((j, lo))
                glb[j] = gub[j] = lo
            else:
                if lo > glb[j] + 0.5:
                    glb[j] = lo
                if hi < gub[j] - 0.5:
                    gub[j] = hi
    return fix, ded, glb, gub


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [21, 25, 27])
def test_presolve_probing_hands_the_device_branches_to_scip(gpu_ctx, seed):
    rng = np.random.default_rng(seed)
    inst = random_binary_instance(rng, nrows=900, ncols=600, maxlen=10)
    om = orc.OMatrix.from_instance(inst)
    rl, ru, rr = orc.propagate_jacobi(om, inst.lb, inst.ub)
    assert not rr["cutoff"]
    print(f"[case] plugin probing seed={seed}", flush=True)
    h = ph.Host(inst)
    assert h.call("th_set_int_param", b"propagating/gpulinear/probing", 1) == ph.SCIP_OKAY
    assert h.call("th_set_int_param", b"propagating/gpulinear/probingvectors", 128) == ph.SCIP_OKAY
    assert h.call("th_initpre") == ph.SCIP_OKAY
    rc, res, nchg = h.presol()
    assert rc == ph.SCIP_OKAY and res == ph.SCIP_SUCCESS
    fix, ded, naggr, nimpl = h.probe_records()
    efix, eded, glb, gub = expected_probing(om, rl, ru, inst.integral)
    assert fix == efix
    assert len(ded) == len(eded) and all(np.array_equal(np.array(a), np.array(b)) for a, b in zip(ded, eded))
    gl, gu = h.bounds(glob=True)
    assert np.array_equal(gl, glb) and np.array_equal(gu, gub)
    # the case finds fixings from cut-off branches and pairs for SCIP's analysis
    assert len(efix) > 0 and len(eded) > 0
    assert h.call("th_exitpre") == ph.SCIP_OKAY
    assert h.close() == ph.SCIP_OKAY


def conflict_arrays(n, cset):
    """Bounds of a conflict set as arrays; columns not in the set get infinite bounds, so a
    check that needs them fails."""
    lb = np.full(n, -1e20)
    ub = np.full(n, 1e20)
    for j, b, v in cset:
        if b == 0:
            lb[j] = v
        else:
            ub[j] = v
    return lb, ub


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["binary", "mixed", "integer"])
def test_propexec_changes_carry_reasons_that_explain_them(gpu_ctx, case):
    from tests.reasons_check import check_implied, check_ranged_implied
    from tests.util import planted_integer_instance
    if case == "binary":
        inst = random_binary_instance(np.random.default_rng(904), 1500, 1000, 10, hidden=True)
    elif case == "mixed":
        inst = sa.gen_mixed(3000, 3000, seed=61, planted=True)
    else:
        inst = planted_integer_instance(np.random.default_rng(951), 400, 300, 6)
    mat = sa.Matrix(gpu_ctx, inst)
    ref = mat.propagate_reasons(params=sa.default_params(flags=sa.SAP_FLAG_RANGEDROWS))
    mat.close()
    assert ref["stats"]["result"] != sa.SAP_CUTOFF and ref["nhist"] > 0
    print(f"[case] plugin reasons {case} nhist={ref['nhist']}", flush=True)
    h = ph.Host(inst)
    assert h.call("th_initsol") == ph.SCIP_OKAY
    rc, res = h.propexec()
    assert rc == ph.SCIP_OKAY and res == ph.SCIP_REDUCEDDOM
    var, bound, info, old, new = h.changes()
    # the device history, in order, with reason row * 2 + rule as inferinfo
    assert np.array_equal(var, ref["col"]) and np.array_equal(bound, ref["bound"])
    assert np.array_equal(info, 2 * ref["row"] + ref["rule"]) and np.array_equal(new, ref["val"])
    # every change is explained by its row and the bounds in force at its time
    for pos in range(len(var)):
        rc, r = h.resprop_change(pos)
        assert rc == ph.SCIP_OKAY and r == ph.SCIP_SUCCESS
        cset, _ = h.conflict()
        lb, ub = conflict_arrays(inst.ncols, cset)
        if info[pos] & 1:
            check_ranged_implied(inst, int(info[pos]) >> 1, int(var[pos]), int(bound[pos]), float(new[pos]), lb, ub)
        else:
            check_implied(inst, int(info[pos]) >> 1, int(var[pos]), int(bound[pos]), float(new[pos]), lb, ub)
    assert h.call("th_exitsol") == ph.SCIP_OKAY
    assert h.close() == ph.SCIP_OKAY


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(4))
def test_propexec_cutoff_starts_conflict_analysis(gpu_ctx, seed):
    from tests.reasons_check import check_conflict
    from tests.util import late_cut_instance, random_integer_instance
    if seed < 2:
        # a row certificate after rounds of history
        base = random_binary_instance(np.random.default_rng(900 + seed), 1500, 1000, 10, hidden=True)
        o = orc.propagate_reasons(orc.OMatrix.from_instance(base), base.lb, base.ub)
        inst = late_cut_instance(base, o["round"], o["col"], o["bound"], o["val"])
    elif seed == 2:
        inst = random_binary_instance(np.random.default_rng(901), 1500, 1000, 10, hidden=False)
    else:
        inst = random_integer_instance(np.random.default_rng(953), 400, 300, 8, sidejitter=0.3)
    mat = sa.Matrix(gpu_ctx, inst)
    ref = mat.propagate_reasons(params=sa.default_params(flags=sa.SAP_FLAG_RANGEDROWS))
    mat.close()
    if ref["stats"]["result"] != sa.SAP_CUTOFF:
        pytest.skip("instance not infeasible")
    print(f"[case] plugin conflict seed={seed} kind={ref['conflict']['kind']}", flush=True)
    h = ph.Host(inst)
    assert h.call("th_initsol") == ph.SCIP_OKAY
    rc, res = h.propexec()
    assert rc == ph.SCIP_OKAY and res == ph.SCIP_CUTOFF
    var, bound, info, old, new = h.changes()
    assert np.array_equal(var, ref["col"]) and np.array_equal(info, 2 * ref["row"] + ref["rule"])
    cset, nanalyzed = h.conflict()
    assert nanalyzed == 1 and len(cset) > 0
    # the conflict set alone proves the node infeasible
    lb, ub = conflict_arrays(inst.ncols, cset)
    check_conflict(inst, ref["conflict"], lb, ub)
    assert h.call("th_exitsol") == ph.SCIP_OKAY
    assert h.close() == ph.SCIP_OKAY


@pytest.mark.gpu
def test_propexec_without_ranged_rows_follows_scip_parameter(gpu_ctx):
    # constraints/linear/rangedrowpropagation = FALSE: the device skips the ranged-row test
    from tests.test_oracle_ranged import ranged_system
    inst = ranged_system(np.random.default_rng(1300), 1500, 3000, slack=False)
    with_r = _oracle(inst, inst.lb, inst.ub, ranged=True)
    without = _oracle(inst, inst.lb, inst.ub, ranged=False)
    assert not np.array_equal(with_r[0], without[0]) or not np.array_equal(with_r[1], without[1])
    for ranged, (ol, ou, _) in ((True, with_r), (False, without)):
        for reasons in (0, 1):
            h = ph.Host(inst)
            ph.lib().th_set_rangedrowprop(h.h, int(ranged))
            assert h.call("th_set_int_param", b"propagating/gpulinear/reasons", reasons) == ph.SCIP_OKAY
            assert h.call("th_initsol") == ph.SCIP_OKAY
            rc, res = h.propexec()
            assert rc == ph.SCIP_OKAY and res == ph.SCIP_REDUCEDDOM
            lb, ub = h.bounds()
            assert np.array_equal(lb, ol) and np.array_equal(ub, ou), (ranged, reasons)
            assert h.close() == ph.SCIP_OKAY


def _ranged_instance():
    from tests.test_oracle_ranged import ranged_system
    inst = ranged_system(np.random.default_rng(1300), 1500, 3000, slack=False)
    with_r = _oracle(inst, inst.lb, inst.ub, ranged=True)
    without = _oracle(inst, inst.lb, inst.ub, ranged=False)
    assert not np.array_equal(with_r[0], without[0]) or not np.array_equal(with_r[1], without[1])
    return inst, with_r, without


# consPropLinear's gating (cons_linear.c:16147-16171) at the node PROPEXEC runs at:
# (stage, in probing, depth, in repropagation, linear int parameters) -> ranged rows expected
GATING = {
    "root": ((ph.SCIP_STAGE_SOLVING, False, 0, False, {}), True),
    "probing": ((ph.SCIP_STAGE_SOLVING, True, 0, False, {}), False),
    "probing_presolving": ((ph.SCIP_STAGE_PRESOLVING, True, 0, False, {}), False),
    "presolving_stage": ((ph.SCIP_STAGE_PRESOLVING, False, 0, False, {}), False),
    "repropagation": ((ph.SCIP_STAGE_SOLVING, False, 0, True, {}), False),
    "beyond_maxdepth": ((ph.SCIP_STAGE_SOLVING, False, 3, False, {"rangedrowmaxdepth": 2}), False),
    "within_maxdepth": ((ph.SCIP_STAGE_SOLVING, False, 2, False, {"rangedrowmaxdepth": 2}), True),
    "freq_skips_depth": ((ph.SCIP_STAGE_SOLVING, False, 3, False, {"rangedrowfreq": 2}), False),
    "freq_hits_depth": ((ph.SCIP_STAGE_SOLVING, False, 4, False, {"rangedrowfreq": 2}), True),
}


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(GATING))
def test_propexec_ranged_rows_follow_conspropLinear_gating(gpu_ctx, case):
    (stage, probing, depth, reprop, ints), ranged = GATING[case]
    inst, with_r, without = _ranged_instance()
    ol, ou, _ = with_r if ranged else without
    for reasons in (0, 1):
        h = ph.Host(inst)
        assert h.call("th_set_int_param", b"propagating/gpulinear/reasons", reasons) == ph.SCIP_OKAY
        for k, v in ints.items():
            h.set_linear_int(k, v)
        assert h.call("th_initsol") == ph.SCIP_OKAY
        h.set_node(stage, probing, depth, reprop)
        rc, res = h.propexec()
        assert rc == ph.SCIP_OKAY and res == ph.SCIP_REDUCEDDOM
        lb, ub = h.bounds()
        assert np.array_equal(lb, ol) and np.array_equal(ub, ou), (case, reasons)
        assert h.close() == ph.SCIP_OKAY


@pytest.mark.gpu
@pytest.mark.parametrize("depth,tightens", [(1, False), (2, True), (0, True)])
def test_propexec_tightenboundsfreq(gpu_ctx, depth, tightens):
    # tightenboundsfreq 2 with the propagator's frequency 1: bounds are tightened at depths 0, 2,
    # 4, ...; elsewhere rows are only tested (cons_linear.c:16155-16157)
    inst = sa.gen_mixed(3000, 3000, seed=0, planted=True)
    ol, ou, orr = orc.propagate_jacobi(orc.OMatrix.from_instance(inst), inst.lb, inst.ub,
                                       orc.params(rangedrows=1, notighten=int(not tightens)))
    assert not orr["cutoff"]
    for reasons in (0, 1):
        h = ph.Host(inst)
        assert h.call("th_set_int_param", b"propagating/gpulinear/reasons", reasons) == ph.SCIP_OKAY
        h.set_linear_int("tightenboundsfreq", 2)
        assert h.call("th_initsol") == ph.SCIP_OKAY
        h.set_node(ph.SCIP_STAGE_SOLVING, False, depth, False)
        rc, res = h.propexec()
        assert rc == ph.SCIP_OKAY
        assert res == (ph.SCIP_REDUCEDDOM if tightens else ph.SCIP_DIDNOTFIND)
        lb, ub = h.bounds()
        assert np.array_equal(lb, ol) and np.array_equal(ub, ou)
        if not tightens:
            assert np.array_equal(lb, inst.lb) and np.array_equal(ub, inst.ub)
        assert h.close() == ph.SCIP_OKAY


@pytest.mark.gpu
def test_propexec_without_tightening_still_finds_infeasible_rows(gpu_ctx):
    # tightenboundsfreq -1: no tightening at any depth, but the final test of propagateCons
    # (cons_linear.c:7717-7742) still cuts off a row whose minimal activity exceeds its rhs
    inst = make_instance([([(0, 1.0), (1, 1.0)], 3.0, INF), ([(1, 1.0), (2, 1.0)], -INF, 1.0)],
                         [0.0, 0.0, 0.0], [1.0, 1.0, 1.0], [1, 1, 1])
    h = ph.Host(inst)
    h.set_linear_int("tightenboundsfreq", -1)
    assert h.call("th_initsol") == ph.SCIP_OKAY
    rc, res = h.propexec()
    assert rc == ph.SCIP_OKAY and res == ph.SCIP_CUTOFF
    assert h.close() == ph.SCIP_OKAY


def _strongbranch_expectation(inst, lb, ub, cands, solvals):
    """Per candidate, both children propagated alone by the oracle's Jacobi twin; then the
    deductions of a strong-branching rule (branch_relpscost.c:1992-2070) against the node."""
    import math
    om = orc.OMatrix.from_instance(inst)
    feastol = 1e-6
    nl, nu = lb.copy(), ub.copy()
    di, ui = [], []
    cut = False
    for c, x in zip(cands, solvals):
        dub = math.ceil(x - 1.0 - feastol)
        ulb = math.floor(x + 1.0 + feastol)
        kids = []
        for side, b in (("down", dub), ("up", ulb)):
            l, u = lb.copy(), ub.copy()
            if side == "down":
                u[c] = b
            else:
                l[c] = b
            ol, ou, r = orc.propagate_jacobi(om, l, u)
            kids.append((bool(r["cutoff"]), ol, ou))
        (dcut, dl, du), (ucut, ul, uu) = kids
        di.append(dcut)
        ui.append(ucut)
        if dcut and ucut:
            cut = True
        elif dcut:
            nl[c] = max(nl[c], ulb)
        elif ucut:
            nu[c] = min(nu[c], dub)
        else:
            nl = np.maximum(nl, np.minimum(dl, ul))
            nu = np.minimum(nu, np.maximum(du, uu))
    cut = cut or bool(np.any(nl > nu + feastol))
    return np.array(di), np.array(ui), nl, nu, cut


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_strongbranch_children_match_the_oracle(gpu_ctx, seed):
    # the strong-branching consumer of the batched split probing: every candidate's two
    # children against the oracle's propagation of that child alone
    from tests.test_oracle_fixpoint import integer_instance
    rng = np.random.default_rng(7000 + seed)
    inst, _ = integer_instance(rng, nrows=300, ncols=200, maxlen=8, coefrange=20, boundrange=0, ubmax=10,
                               planted=True)
    rl, ru, rr = _oracle(inst, inst.lb, inst.ub, ranged=False)
    assert not rr["cutoff"]
    free = np.nonzero(rl < ru)[0]
    cands = rng.choice(free, size=min(24, len(free)), replace=False).astype(np.int32)
    solvals = rl[cands] + (ru[cands] - rl[cands]) * rng.uniform(0.05, 0.95, size=len(cands))
    solvals = np.where(np.abs(solvals - np.round(solvals)) < 0.01, solvals + 0.3, solvals)
    edi, eui, enl, enu, ecut = _strongbranch_expectation(inst, rl, ru, cands, solvals)
    h = ph.Host(inst)
    assert h.call("th_initsol") == ph.SCIP_OKAY
    h.set_local(rl, ru)
    di, ui, nl, nu, cut, rc = h.strongbranch(cands, solvals)
    assert rc == ph.SCIP_OKAY
    assert np.array_equal(di, edi) and np.array_equal(ui, eui)
    assert cut == ecut
    if not ecut:
        assert np.array_equal(nl, enl) and np.array_equal(nu, enu)
        # the deductions are real: some candidate has an infeasible child or a tighter union
        assert edi.any() or eui.any() or (enl > rl).any() or (enu < ru).any()
    assert h.close() == ph.SCIP_OKAY


def _dive_expectation(inst, lb, ub, cands, order, fixvals, maxdepth):
    """Each dive emulated with the oracle's Jacobi twin: the node propagated, then every step
    starts from it with all of the dive's fixings so far (only their rows marked first)."""
    om = orc.OMatrix.from_instance(inst)
    rl, ru, rr = orc.propagate_jacobi(om, lb, ub)
    nd = order.shape[0]
    if rr["cutoff"]:
        return np.zeros(nd, dtype=int), np.ones(nd, dtype=bool), None, None
    depth = np.zeros(nd, dtype=int)
    cut = np.zeros(nd, dtype=bool)
    out_l = np.tile(rl, (nd, 1))
    out_u = np.tile(ru, (nd, 1))
    for k in range(nd):
        cl, cu = rl.copy(), ru.copy()
        fix = []
        pos = 0
        while maxdepth < 0 or len(fix) < maxdepth:
            c = None
            while pos < len(cands) and c is None:
                cand = order[k, pos]
                pos += 1
                if cl[cands[cand]] < cu[cands[cand]]:
                    c = cand
            if c is None:
                break
            col = cands[c]
            v = fixvals[k, c]
            if inst.integral[col]:
                v = np.floor(v + 0.5)
            v = min(max(v, cl[col]), cu[col])
            fix.append((col, v))
            sl, su = rl.copy(), ru.copy()
            for fc, fv in fix:
                sl[fc] = su[fc] = fv
            nl, nu, r = orc.propagate_jacobi_marked(om, sl, su, orc.rows_of_columns(om, [f[0] for f in fix]))
            if r["cutoff"]:
                cut[k] = True
                break
            cl, cu = nl, nu
        depth[k] = len(fix)
        out_l[k], out_u[k] = cl, cu
    return depth, cut, out_l, out_u


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["planted_binary", "covering", "integer"])
def test_batched_dives_match_the_oracle(gpu_ctx, kind):
    # fix-and-propagate dives (heur_fixandinfer.c) advanced together on the device, every dive
    # against the oracle's propagation of its own fixings: on a system with a planted 0/1 point
    # the dives that fix towards that point fix every candidate and the others mostly end in
    # cutoffs, on covering/packing rows with random values every dive ends in a cutoff, on
    # general integers they fix a few and propagation does the rest
    rng = np.random.default_rng({"planted_binary": 7100, "covering": 7300, "integer": 7200}[kind])
    if kind == "planted_binary":
        inst, xs = random_binary_instance(rng, 500, 300, 12, point=True)
    elif kind == "covering":
        inst = sa.gen_covering_packing(300, 200, 10, seed=11)
    else:
        from tests.test_oracle_fixpoint import integer_instance
        inst, _ = integer_instance(rng, nrows=300, ncols=200, maxlen=8, coefrange=20, boundrange=0, ubmax=6,
                                   planted=True)
    cands = np.nonzero(inst.integral)[0][:80].astype(np.int32)
    nd = 24
    order = np.array([rng.permutation(len(cands)) for _ in range(nd)], dtype=np.int32)
    fixvals = inst.lb[cands] + (inst.ub[cands] - inst.lb[cands]) * rng.uniform(0, 1, size=(nd, len(cands)))
    if kind == "planted_binary":
        fixvals[::2] = xs[cands]
    edepth, ecut, el, eu = _dive_expectation(inst, inst.lb, inst.ub, cands, order, fixvals, -1)
    if kind == "covering":
        assert ecut.all()
    elif kind == "integer":
        assert not ecut.any()
    else:
        assert not ecut[::2].any() and ecut[1::2].any()
    h = ph.Host(inst)
    assert h.call("th_initsol") == ph.SCIP_OKAY
    depth, cut, dl, du, rc = h.dive(cands, order, fixvals)
    assert rc == ph.SCIP_OKAY
    assert np.array_equal(depth, edepth) and np.array_equal(cut, ecut)
    for k in range(nd):
        if not ecut[k]:
            assert np.array_equal(dl[k], el[k]) and np.array_equal(du[k], eu[k]), k
            assert np.all(dl[k][cands] == du[k][cands])   # every candidate ended fixed
    # a depth limit stops every dive there
    depth2, cut2, _, _, rc = h.dive(cands, order, fixvals, maxdepth=3)
    assert rc == ph.SCIP_OKAY and depth2.max() <= 3
    assert np.array_equal(depth2, np.minimum(edepth, 3)) or ecut.any()
    assert h.close() == ph.SCIP_OKAY
