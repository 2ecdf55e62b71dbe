#!/usr/bin/env python3
"""Benchmark: linear-constraint domain propagation to fixpoint on one MI355X.

Workload (one step = one propagation of the root bound vector to its fixpoint, bounds and
matrix resident in HBM): plain GPU, large single instance -- 1,000,000 rows x 1,000,000
columns, log-normal row lengths (median 8, sigma 1.0, max 10,000), 40% integer columns
[0, U{1..1000}], 60% continuous [-U[0,1e4], U[0,1e4]] with 5% one-sided infinite,
coefficients +-10^U[-3,3], ranged rows at +-0.3 of the finite activity span, seed 2.

Prints one JSON line (rank 0).  `value` is swept nonzeros per second (rows actually swept
by the worklist, summed over rounds) over the K timed steps.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "nnz-sweeps/s to fixpoint, batched vs plain GPU kernel; p50/p99 request latency"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes of the row sweep: 8 (value) + 4 (column index) + 16 (lb, ub gather) per
# swept nonzero; 4 (row pointer) + 16 (lhs, rhs) per swept row
SWEEP_BYTES_PER_NNZ = 28
SWEEP_BYTES_PER_ROW = 20


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(budget_s: float = 12.0, planted: bool = False):
    """SCIP-order (Gauss-Seidel) CPU restatement, single thread, on a bounded sample of the
    same workload: the same generator at 100,000 x 100,000, repeated until the budget."""
    import scip_amd as sa
    from oracle import oracle as orc
    inst = sa.gen_mixed(100_000, 100_000, seed=2, planted=planted)
    om = orc.OMatrix.from_instance(inst)
    nnz = 0
    runs = 0
    t0 = time.perf_counter()
    ttf = []
    while True:
        t = time.perf_counter()
        _, _, r = orc.propagate_scip(om, inst.lb, inst.ub)
        ttf.append(time.perf_counter() - t)
        nnz += r["nnz_swept"]
        runs += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    el = time.perf_counter() - t0
    return {
        "value": nnz / el,
        "unit": "nnz-sweeps/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"SCIP-order restatement (oracle/oracle.c orc_propagate_scip), 1 thread, same generator at "
                   f"100000x100000 (nnz {inst.nnz}), {runs} runs to fixpoint in {el:.1f}s; "
                   f"time-to-fixpoint {1000 * min(ttf):.1f} ms"),
        "time_to_fixpoint_ms": 1000 * min(ttf),
    }


def traffic_from_profiles():
    """HBM bytes per sweep launch from the committed rocprofv3 PMC summary, if present."""
    path = os.path.join(ROOT, "profiles", "k_sweep_mixed1M_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    except Exception:
        return None


def copy_probe(nbytes: int = 1 << 30, reps: int = 5):
    """A plain device copy of known size, for scaling rocprof FETCH_SIZE / WRITE_SIZE."""
    import torch
    a = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes, "reps": reps}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--rows", type=int, default=1_000_000)
    ap.add_argument("--cols", type=int, default=1_000_000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--copy-probe", action="store_true", help="only run the copy kernel used to scale PMC bytes")
    ap.add_argument("--variant", choices=("both", "literal", "planted", "none"), default="both",
                    help="literal: only the headline instance (for profiling); planted: only the feasible variant; "
                         "none: only the batched legs")
    ap.add_argument("--legs", default="nodes,probe,packed,serve",
                    help="batched legs to run after the headline: nodes, probe, packed, serve (comma list, or none)")
    ap.add_argument("--nodes", type=int, default=1024, help="node vectors of the batched-nodes leg")
    ap.add_argument("--probe-vectors", type=int, default=2048, help="branch vectors per launch of the probing leg")
    ap.add_argument("--packed-instances", type=int, default=512, help="instances of the packed leg")
    ap.add_argument("--serve-seconds", type=float, default=10.0, help="length of the served arrival trace")
    ap.add_argument("--serve-rate", type=float, default=2000.0, help="Poisson arrival rate (requests/s)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU (one process per GPU with its own share of "
                                   "the rows or of the node vectors is the sharding it would use, not built); "
                                   "run with --gpus 1"}), flush=True)
        sys.exit(2)
    if args.copy_probe:
        copy_probe()
        return

    import scip_amd as sa

    ctx = sa.Context(0)
    if args.variant == "none":
        print(json.dumps(run_legs(sa, ctx, args, {})), flush=True)
        ctx.close()
        return
    if args.variant != "both":
        out = run_workload(sa, ctx, args, planted=(args.variant == "planted"))
        print(json.dumps(out), flush=True)
        ctx.close()
        return
    main_out = run_workload(sa, ctx, args, planted=False)
    feas = run_workload(sa, ctx, args, planted=True)
    main_out["config"]["feasible_variant"] = {
        "workload": "same generator with sides centred on a planted feasible point (sidemode 1)",
        "value": feas["value"], "unit": "nnz-sweeps/s", "ms_per_step": feas["ms_per_step"],
        "rounds_to_fixpoint": feas["config"]["rounds_to_fixpoint"],
        "nnz_swept_per_step": feas["config"]["nnz_swept_per_step"],
        "result": feas["config"]["result"], "roofline_frac": feas["roofline"]["frac"],
    }
    if not args.no_cpu:
        log("[bench] CPU baseline leg")
        main_out["cpu_baseline"] = cpu_baseline()
        fb = cpu_baseline(budget_s=6.0, planted=True)
        main_out["config"]["feasible_variant"]["cpu_baseline_value"] = fb["value"]
        main_out["config"]["feasible_variant"]["cpu_time_to_fixpoint_ms_100k_sample"] = fb["time_to_fixpoint_ms"]
    print(json.dumps(run_legs(sa, ctx, args, main_out)), flush=True)
    ctx.close()


def run_legs(sa, ctx, args, main_out):
    legs = [x for x in args.legs.split(",") if x and x != "none"]
    parts = None
    if "nodes" in legs:
        main_out["batched_nodes"] = run_batched_nodes(sa, ctx, args)
    if "probe" in legs:
        main_out["probing"] = run_probing(sa, ctx, args)
    if "packed" in legs or "serve" in legs:
        parts = gen_uneven_instances(sa, args.packed_instances)
    if "packed" in legs:
        main_out["packed_instances"] = run_packed(sa, ctx, args, parts)
    if "serve" in legs:
        main_out["serving"] = run_serving(sa, args, parts)
    return main_out


def run_workload(sa, ctx, args, planted: bool):
    log(f"[bench] generating {args.rows}x{args.cols} instance (planted={planted})")
    t = time.perf_counter()
    inst = sa.gen_mixed(args.rows, args.cols, seed=2, planted=planted)
    log(f"[bench] nnz={inst.nnz} generated in {time.perf_counter() - t:.1f}s")
    mat = sa.Matrix(ctx, inst)
    init = sa.Bounds(mat, 1)
    init.upload(inst.lb, inst.ub)
    work = sa.Bounds(mat, 1)
    params = sa.default_params()

    def step():
        work.copy_from(init)
        return work.propagate(params)

    for w in range(args.warmup):
        st = step()
        log(f"[bench] warmup {w}: rounds={st['rounds']} nnz_swept={st['nnz_swept']} dev_ms={st['device_ms']:.3f}")

    tot_nnz = tot_rows = tot_launch = 0
    tot_sweep_ms = tot_dev_ms = 0.0
    rounds = None
    t0 = time.perf_counter()
    for k in range(args.steps):
        st = step()
        tot_nnz += st["nnz_swept"]
        tot_rows += st["rows_swept"]
        tot_launch += st["sweep_launches"]
        tot_sweep_ms += st["sweep_ms"]
        tot_dev_ms += st["device_ms"]
        rounds = st["rounds"]
    el = time.perf_counter() - t0
    log(f"[bench] {args.steps} steps in {el:.3f}s; rounds={rounds} result={st['result']} nchg={st['nchgbds']}")

    value = tot_nnz / el
    sweep_bytes = SWEEP_BYTES_PER_NNZ * tot_nnz + SWEEP_BYTES_PER_ROW * tot_rows
    achieved = sweep_bytes / (tot_sweep_ms * 1e-3) / 1e9 if tot_sweep_ms > 0 else 0.0
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic_from_profiles(),
        "kernel": "k_sweep",
        "launches": tot_launch,
        "avg_launch_us": round(1000 * tot_sweep_ms / max(tot_launch, 1), 2),
        "algorithmic_bytes_per_launch": int(sweep_bytes / max(tot_launch, 1)),
    }
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "nnz-sweeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000 * el / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "plain GPU, large single instance (1M rows x 1M cols, log-normal rows median 8 sigma 1 "
                        "max 10k, 40% int / 60% cont, ranged rows +-0.3 span, seed 2), propagate to fixpoint "
                        "from root bounds" + (" [planted feasible sides]" if planted else ""),
            "nrows": inst.nrows,
            "ncols": inst.ncols,
            "nnz": inst.nnz,
            "rounds_to_fixpoint": rounds,
            "nnz_swept_per_step": tot_nnz // args.steps,
            "nominal_nnz_sweeps_per_step": rounds * inst.nnz,
            "device_ms_per_step": tot_dev_ms / args.steps,
            "result": {sa.SAP_CUTOFF: "cutoff", sa.SAP_REDUCEDDOM: "reduceddom"}.get(st["result"], "didnotfind"),
            "parallelism": "single GPU",
        },
        "roofline": roofline,
    }
    work.close()
    init.close()
    mat.close()
    return out



# ------------------------------------------------------------------------------------------
# batched legs: node vectors on one matrix, packed uneven instances, served requests
# ------------------------------------------------------------------------------------------
def _rate(nnz, seconds):
    return nnz / seconds if seconds > 0 else 0.0


def run_batched_nodes(sa, ctx, args):
    """Covering/packing rows 200,000 x 300,000 (10 nonzeros per row, seed 3) and node vectors
    fixing 1..50 random binaries to 0/1 (seed 3), propagated from the root fixpoint held on the
    device.  Batched: all vectors in one call (one launch per round for all of them).  Plain:
    one node per launch sequence, from the same resident root (C loop of the server's plain
    mode), so both sweep exactly the same rows."""
    nn = args.nodes
    log(f"[bench] batched nodes: 200000x300000 covering/packing, {nn} nodes")
    inst = sa.gen_covering_packing(200_000, 300_000, 10, seed=3)
    mat = sa.Matrix(ctx, inst)
    base = sa.Bounds(mat, 1)
    base.upload(inst.lb, inst.ub)
    rst = base.propagate()
    ptr, col, val = sa.gen_node_fixings(inst, nn, 1, 50, seed=3)
    p = sa.default_params(flags=sa.SAP_FLAG_BASE_IS_FIXPOINT)
    nodes = sa.Bounds(mat, nn)
    for _ in range(max(args.warmup, 1)):
        out = nodes.propagate_nodes(base, 0, ptr, col, val, val, p)
    steps = max(args.steps, 1)
    nnz = rows = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        out = nodes.propagate_nodes(base, 0, ptr, col, val, val, p)
        nnz += out["stats"]["nnz_swept"]
        rows += out["stats"]["rows_swept"]
    el = time.perf_counter() - t0
    res = out["result"]
    batched = {"value": _rate(nnz, el), "unit": "nnz-sweeps/s", "ms_per_batch": 1000 * el / steps,
               "nodes_per_s": nn * steps / el, "nnz_swept_per_batch": nnz // steps,
               "rows_swept_per_batch": rows // steps,
               "cutoff_nodes": int((res == sa.SAP_CUTOFF).sum()),
               "reduceddom_nodes": int((res == sa.SAP_REDUCEDDOM).sum()),
               "max_rounds": int(out["rounds"].max()), "changed_bounds": int(out["chgptr"][-1])}
    nodes.close()
    # plain: the same nodes one at a time through the server's C loop (PLAIN mode, all
    # arriving at once) -- admission, launch sequence and collection per node
    srv = sa.Server(0, 1, 64 << 30, 1)
    pid = srv.add_problem(inst)
    srv.warm(pid)
    arr = np.zeros(nn)
    prob = np.full(nn, pid, dtype=np.int32)
    srv.replay(arr[:8], prob[:8], ptr[:9], col[:ptr[8]], val[:ptr[8]], val[:ptr[8]], mode=sa.SAP_SERVE_PLAIN,
               steprounds=sa.SAP_STEP_LONG)
    t0 = time.perf_counter()
    pl = srv.replay(arr, prob, ptr, col, val, val, mode=sa.SAP_SERVE_PLAIN, steprounds=sa.SAP_STEP_LONG)
    elp = time.perf_counter() - t0
    same = bool(np.array_equal(pl["result"], res))
    srv.close()
    plain = {"value": _rate(pl["stats"]["nnz_swept"], elp), "unit": "nnz-sweeps/s",
             "ms_per_node": 1000 * elp / nn, "nodes_per_s": nn / elp, "nnz_swept": pl["stats"]["nnz_swept"],
             "verdicts_equal_batched": same}
    base.close()
    mat.close()
    return {"workload": "covering/packing 200000x300000, 10 nnz/row, seed 3; nodes fix 1..50 binaries, seed 3; "
                        "from the root fixpoint (resident)",
            "nodes": nn, "root_rounds": rst["rounds"], "batched": batched, "plain": plain,
            "speedup_nnz_sweeps": batched["value"] / plain["value"] if plain["value"] > 0 else None}


def run_probing(sa, ctx, args):
    """Probing of every unfixed binary of the covering/packing 200,000 x 300,000 instance (seed 3)
    at its root fixpoint: both branches x = 0 / x = 1 of each, propagated from the resident
    fixpoint.  Batched: sap_probe with --probe-vectors branch vectors per launch.  Plain: one
    probe (two branches) per call on a sample of the same columns, as a loop of
    SCIPapplyProbingVar would run them; the sample's verdicts and pair lists are checked equal."""
    log("[bench] probing: 200000x300000 covering/packing, all unfixed binaries")
    inst = sa.gen_covering_packing(200_000, 300_000, 10, seed=3)
    mat = sa.Matrix(ctx, inst)
    base = sa.Bounds(mat, 1)
    base.upload(inst.lb, inst.ub)
    base.propagate()
    rl, ru = base.download(0)
    cols = np.nonzero((inst.integral != 0) & (rl == 0.0) & (ru == 1.0))[0].astype(np.int32)
    p = sa.default_params(flags=sa.SAP_FLAG_BASE_IS_FIXPOINT)
    pv = sa.Bounds(mat, args.probe_vectors)
    pv.probe(base, 0, cols[:4 * args.probe_vectors], p)   # warm-up (graphs, buffers)
    t0 = time.perf_counter()
    out = pv.probe(base, 0, cols, p)
    el = time.perf_counter() - t0
    st = out["stats"]
    ver = out["verdict"]
    batched = {"probes_per_s": len(cols) / el, "branches_per_s": 2 * len(cols) / el, "seconds": el,
               "value": _rate(st["nnz_swept"], el), "unit": "nnz-sweeps/s", "nnz_swept": st["nnz_swept"],
               "down_infeasible": int((ver == 1).sum()), "up_infeasible": int((ver == 2).sum()),
               "both_infeasible": int((ver == 3).sum()), "pairs": int(out["dedptr"][-1]),
               "max_rounds": int(out["rounds"].max()) if len(cols) else 0}
    pv.close()
    # plain: one probe per call, on an evenly spaced sample
    ns = min(len(cols), 2000)
    sample = cols[np.linspace(0, len(cols) - 1, ns).astype(np.int64)] if ns else cols
    one = sa.Bounds(mat, 2)
    one.probe(base, 0, sample[:16], p)
    nnz = 0
    same = True
    pos = np.searchsorted(cols, sample)
    t0 = time.perf_counter()
    for k, x in enumerate(sample):
        o = one.probe(base, 0, sample[k:k + 1], p)
        nnz += o["stats"]["nnz_swept"]
        q = int(pos[k])
        a, e = out["dedptr"][q], out["dedptr"][q + 1]
        same &= bool(o["verdict"][0] == ver[q] and np.array_equal(o["dedcol"], out["dedcol"][a:e]))
    elp = time.perf_counter() - t0
    one.close()
    plain = {"probes_per_s": ns / elp if elp > 0 else None, "sample": int(ns), "seconds": elp,
             "value": _rate(nnz, elp), "unit": "nnz-sweeps/s", "equal_to_batched": same}
    base.close()
    mat.close()
    return {"workload": "covering/packing 200000x300000, 10 nnz/row, seed 3; probe every unfixed binary at the "
                        "root fixpoint (resident)",
            "probes": int(len(cols)), "batched": batched, "plain": plain,
            "speedup_probes_per_s": (batched["probes_per_s"] / plain["probes_per_s"]
                                     if plain["probes_per_s"] else None)}


def gen_uneven_instances(sa, count, seed=4):
    """Instances with rows log-uniform in [500, 500,000], columns 1.5 x rows, 4..64 nonzeros per
    row and the mixed integer/continuous distributions of the large instance; instance i uses
    seed 4000+i.  Sides are centred on a planted feasible point (sidemode 1)."""
    rng = np.random.default_rng(seed)
    m = np.rint(np.exp(rng.uniform(np.log(500), np.log(500_000), size=count))).astype(np.int64)
    log(f"[bench] generating {count} uneven instances (rows {m.sum()})")
    t = time.perf_counter()
    with ThreadPoolExecutor(16) as ex:
        parts = list(ex.map(lambda i: sa.gen_mixed(int(m[i]), int(round(1.5 * m[i])), seed=4000 + i, lenmode=1,
                                                   minlen=4, maxlen=64, planted=True), range(count)))
    log(f"[bench] generated nnz={sum(p.nnz for p in parts)} in {time.perf_counter() - t:.1f}s")
    return parts


def run_packed(sa, ctx, args, parts):
    """All instances packed block-diagonally into one resident matrix, propagated to their
    fixpoints from the root bounds in one launch per round (per-instance verdicts); against one
    resident matrix per instance propagated one after another."""
    # row binning and tile shape from the autotuner, on a row sample of the packed system; the
    # plain form below runs under the same tuning
    t = time.perf_counter()
    packed_inst, _, _ = sa.pack_instances(parts)
    tuning, best_ms, default_ms = ctx.autotune(packed_inst, samplerows=100_000, apply=True)
    del packed_inst
    log(f"[bench] packed autotune {time.perf_counter() - t:.1f}s: {tuning} ({best_ms:.3f} vs default {default_ms:.3f} ms)")
    t = time.perf_counter()
    pm = sa.PackedMatrix(ctx, parts)
    log(f"[bench] packed upload {time.perf_counter() - t:.1f}s, device bytes {pm.info()['device_bytes']}")
    plb, pub, pres, prnd, pst = pm.propagate_packed()
    init = sa.Bounds(pm, 1)
    init.upload(pm.inst.lb, pm.inst.ub)
    work = sa.Bounds(pm, 1)
    for _ in range(max(args.warmup, 1)):
        work.copy_from(init)
        work.propagate()
    steps = max(args.steps, 1)
    nnz = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        work.copy_from(init)
        st = work.propagate()
        nnz += st["nnz_swept"]
    el = time.perf_counter() - t0
    packed = {"value": _rate(nnz, el), "unit": "nnz-sweeps/s", "ms_per_pass": 1000 * el / steps,
              "nnz_swept_per_pass": nnz // steps, "rounds": st["rounds"], "sweep_launches_per_pass":
              st["sweep_launches"] // 1, "cutoff_instances": int((pres == sa.SAP_CUTOFF).sum())}
    work.close()
    init.close()
    pm.close()
    # plain: one matrix per instance, one propagation after another
    t = time.perf_counter()
    mats = [sa.Matrix(ctx, p) for p in parts]
    log(f"[bench] per-instance uploads {time.perf_counter() - t:.1f}s")
    inits, works = [], []
    for mt, p in zip(mats, parts):
        a = sa.Bounds(mt, 1)
        a.upload(p.lb, p.ub)
        inits.append(a)
        works.append(sa.Bounds(mt, 1))
    for a, w in zip(inits, works):
        w.copy_from(a)
        w.propagate()
    nnzp = 0
    same = True
    t0 = time.perf_counter()
    for k, (a, w) in enumerate(zip(inits, works)):
        w.copy_from(a)
        st = w.propagate()
        nnzp += st["nnz_swept"]
        same &= (st["result"] == sa.SAP_CUTOFF) == (pres[k] == sa.SAP_CUTOFF)
    elp = time.perf_counter() - t0
    for x in inits + works + mats:
        x.close()
    plain = {"value": _rate(nnzp, elp), "unit": "nnz-sweeps/s", "ms_all_instances": 1000 * elp,
             "nnz_swept": nnzp, "verdicts_equal_packed": bool(same)}
    ctx.set_tuning()
    return {"workload": f"{len(parts)} instances, rows log-uniform [500, 500000] (seed 4), cols 1.5 x rows, "
                        "4..64 nnz/row, mixed int/cont, instance i seed 4000+i, planted-feasible sides; "
                        "root propagation to fixpoint",
            "instances": len(parts), "rows": int(sum(p.nrows for p in parts)),
            "nnz": int(sum(p.nnz for p in parts)), "packed": packed, "plain": plain,
            "tuning": {"autotuned": tuning, "sample_rows": 100_000, "best_ms": best_ms, "default_ms": default_ms},
            "speedup_nnz_sweeps": packed["value"] / plain["value"] if plain["value"] > 0 else None}


def run_serving(sa, args, parts):
    """Poisson arrivals (rate, seconds from the arguments); half the requests are nodes of the
    covering/packing 200,000 x 300,000 problem (1..50 binaries fixed), half are nodes of a
    uniformly chosen uneven instance (1..50 integer columns fixed to 0/1).  All problems are
    resident (warmed) before the trace; continuous batching against the plain
    one-request-at-a-time form on the same trace."""
    cp = sa.gen_covering_packing(200_000, 300_000, 10, seed=3)
    srv = sa.Server(0, 4, 200 << 30, 16)
    pids = [srv.add_problem(cp)] + [srv.add_problem(p) for p in parts]
    # half of all requests are nodes of the covering/packing problem: its pool gets more slots
    # (256 x ~26 MB of per-slot state); the 512 instances keep 16 each
    srv.set_slots(pids[0], 256)
    t = time.perf_counter()
    for pid in pids:
        srv.warm(pid)
    log(f"[bench] serving: {len(pids)} problems resident in {time.perf_counter() - t:.1f}s")
    rng = np.random.default_rng(5)
    n = int(args.serve_rate * args.serve_seconds)
    arrival = np.cumsum(rng.exponential(1.0 / args.serve_rate, size=n))
    arrival -= arrival[0]
    kind = rng.integers(0, 2, size=n)
    prob = np.where(kind == 0, pids[0], rng.integers(1, len(pids), size=n)).astype(np.int32)
    probs = [cp] + list(parts)
    intcols = [np.nonzero(p.integral)[0] for p in probs]
    ptr = np.zeros(n + 1, dtype=np.int64)
    cols, vals = [], []
    for i in range(n):
        ic = intcols[prob[i]]
        k = int(min(len(ic), rng.integers(1, 51)))
        c = ic[np.unique(rng.integers(0, len(ic), size=k))] if k else np.zeros(0, dtype=np.int64)
        cols.append(c.astype(np.int32))
        vals.append(rng.integers(0, 2, size=len(c)).astype(np.float64))
        ptr[i + 1] = ptr[i] + len(c)
    col = np.concatenate(cols)
    val = np.concatenate(vals)
    log(f"[bench] serving trace: {n} requests over {args.serve_seconds}s")
    cont = srv.replay(arrival, prob, ptr, col, val, val, mode=sa.SAP_SERVE_CONTINUOUS, steprounds=sa.SAP_STEP_SHORT)
    cst = cont["stats"]
    # saturated capacity: the same requests all arriving at once
    sat = srv.replay(np.zeros(n), prob, ptr, col, val, val, mode=sa.SAP_SERVE_CONTINUOUS,
                     steprounds=sa.SAP_STEP_LONG)
    sst = sat["stats"]
    # plain on a prefix of the same trace (it falls behind the arrivals)
    npl = min(n, max(200, n // 10))
    pl = srv.replay(arrival[:npl], prob[:npl], ptr[:npl + 1], col[:ptr[npl]], val[:ptr[npl]], val[:ptr[npl]],
                    mode=sa.SAP_SERVE_PLAIN, steprounds=sa.SAP_STEP_LONG)
    pst = pl["stats"]
    same = bool(np.array_equal(pl["result"], cont["result"][:npl]) and np.array_equal(pl["nchg"], cont["nchg"][:npl]))
    srv.close()
    return {"workload": f"Poisson {args.serve_rate:.0f} req/s for {args.serve_seconds:.0f}s; 50% nodes of the "
                        f"200000x300000 covering/packing problem, 50% nodes of {len(parts)} uneven instances; "
                        "1..50 fixings per request",
            "requests": n,
            "continuous": {"value": _rate(cst["nnz_swept"], cst["wall_s"]), "unit": "nnz-sweeps/s",
                           "requests_per_s": n / cst["wall_s"], "p50_ms": cst["p50_ms"], "p99_ms": cst["p99_ms"],
                           "max_ms": cst["max_ms"], "max_running": cst["max_running"], "steps": cst["steps"],
                           "nnz_swept": cst["nnz_swept"]},
            "saturated": {"value": _rate(sst["nnz_swept"], sst["wall_s"]), "unit": "nnz-sweeps/s",
                          "requests_per_s": n / sst["wall_s"], "p50_ms": sst["p50_ms"], "p99_ms": sst["p99_ms"],
                          "max_running": sst["max_running"]},
            "plain": {"requests": npl, "value": _rate(pst["nnz_swept"], pst["wall_s"]), "unit": "nnz-sweeps/s",
                      "requests_per_s": npl / pst["wall_s"], "p50_ms": pst["p50_ms"], "p99_ms": pst["p99_ms"],
                      "results_equal_continuous": same},
            "speedup_nnz_sweeps_saturated_vs_plain": (_rate(sst["nnz_swept"], sst["wall_s"])
                                                      / max(_rate(pst["nnz_swept"], pst["wall_s"]), 1e-9))}


if __name__ == "__main__":
    main()
