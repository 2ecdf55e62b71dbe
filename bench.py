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


def host_cpu():
    """The host the CPU legs ran on: CPU model, logical CPUs, CPUs this process may use."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return {"model": model, "nproc": os.cpu_count(), "usable": usable}


def cpu_baseline(rows: int, cols: int, budget_s: float = 20.0):
    """SCIP-order (Gauss-Seidel) CPU restatement (oracle/oracle.c orc_propagate_scip), single
    thread, on the timed workload itself: the large instance with literal sides and with
    planted-feasible sides, each propagated to its fixpoint from the root bounds, repeated
    until half the budget is spent (at least once)."""
    import scip_amd as sa
    from oracle import oracle as orc
    legs = {}
    for name, planted in (("literal", False), ("planted", True)):
        inst = sa.gen_mixed(rows, cols, seed=2, planted=planted)
        om = orc.OMatrix.from_instance(inst)
        ttf = []
        t0 = time.perf_counter()
        while True:
            t = time.perf_counter()
            _, _, r = orc.propagate_scip(om, inst.lb, inst.ub)
            ttf.append(time.perf_counter() - t)
            if time.perf_counter() - t0 >= budget_s / 2:
                break
        legs[name] = {"time_to_fixpoint_ms": 1000 * min(ttf), "runs": len(ttf), "rounds": r["rounds"],
                      "nnz_swept": r["nnz_swept"], "result": "cutoff" if r["cutoff"] else "reduceddom",
                      "nnz_sweeps_per_s": r["nnz_swept"] / min(ttf)}
        del om, inst
    pl = legs["planted"]
    return {
        "value": pl["nnz_sweeps_per_s"],
        "unit": "nnz-sweeps/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"SCIP-order restatement (oracle/oracle.c orc_propagate_scip), 1 thread, on the timed "
                   f"{rows}x{cols} instance itself: literal sides (cut off after "
                   f"{legs['literal']['nnz_swept']} swept nonzeros) and planted-feasible sides "
                   f"({pl['rounds']} rounds, {pl['nnz_swept']} swept nonzeros); value is the planted "
                   "variant's nnz-sweeps/s, the one that runs the fixpoint loop; time to fixpoint of both "
                   "beside it"),
        "time_to_fixpoint_ms": {k: v["time_to_fixpoint_ms"] for k, v in legs.items()},
        "legs": legs,
        "host": host_cpu(),
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
    ap.add_argument("--gather-probe", action="store_true",
                    help="only run the memory-floor kernel of the row sweep on the large instance (for profiling)")
    ap.add_argument("--variant", choices=("both", "literal", "planted", "none"), default="both",
                    help="literal: only the headline instance (for profiling); planted: only the feasible variant; "
                         "none: only the batched legs")
    ap.add_argument("--legs", default="cfg1,nodes,probe,packed,serve",
                    help="batched legs to run after the headline: nodes, probe, packed, serve (comma list, or none)")
    ap.add_argument("--nodes", type=int, default=1024, help="node vectors of the batched-nodes leg")
    ap.add_argument("--probe-vectors", type=int, default=2048, help="branch vectors per launch of the probing leg")
    ap.add_argument("--packed-instances", type=int, default=512, help="instances of the packed leg")
    ap.add_argument("--serve-seconds", type=float, default=60.0, help="length of the served arrival trace")
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

    if args.gather_probe:
        ctx = sa.Context(0)
        inst = sa.gen_mixed(args.rows, args.cols, seed=2)
        mat = sa.Matrix(ctx, inst)
        b = sa.Bounds(mat, 1)
        b.upload(inst.lb, inst.ub)
        by = SWEEP_BYTES_PER_NNZ * inst.nnz + SWEEP_BYTES_PER_ROW * inst.nrows
        for blocks in (0, 1024, 4096):
            ms = b.bench_gather(max(args.steps, 1), blocks)
            print(json.dumps({"blocks": blocks, "gather_floor_ms": ms, "algorithmic_bytes": by,
                              "achieved_GBs": by / (ms * 1e-3) / 1e9}), flush=True)
        b.close()
        mat.close()
        ctx.close()
        return

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
        cb = cpu_baseline(args.rows, args.cols)
        main_out["cpu_baseline"] = cb
        main_out["config"]["gpu_vs_cpu_time_to_fixpoint"] = {
            "literal": cb["time_to_fixpoint_ms"]["literal"] / main_out["ms_per_step"],
            "planted": cb["time_to_fixpoint_ms"]["planted"] / feas["ms_per_step"]}
    print(json.dumps(run_legs(sa, ctx, args, main_out)), flush=True)
    ctx.close()


def run_config1(sa, ctx, args):
    """The reference's CPU-runnable case: covering/packing 2,000 x 5,000, 10 nonzeros per row,
    seed 1, root bounds to fixpoint.  Time to fixpoint of the plain GPU path (bounds resident,
    one copy of the root bounds and one propagation per step) against the SCIP-order CPU
    restatement, single thread.  Pure binary: the fixpoint is unique, so both must agree."""
    from oracle import oracle as orc
    inst = sa.gen_covering_packing(2000, 5000, 10, seed=1)
    mat = sa.Matrix(ctx, inst)
    init = sa.Bounds(mat, 1)
    init.upload(inst.lb, inst.ub)
    work = sa.Bounds(mat, 1)
    for _ in range(20):
        work.copy_from(init)
        st = work.propagate()
    reps = 500
    t0 = time.perf_counter()
    dev = 0.0
    for _ in range(reps):
        work.copy_from(init)
        st = work.propagate()
        dev += st["device_ms"]
    gpu_ms = 1000 * (time.perf_counter() - t0) / reps
    gl, gu = work.download(0)
    work.close()
    init.close()
    mat.close()
    om = orc.OMatrix.from_instance(inst)
    ttf = []
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < 2.0:
        t = time.perf_counter()
        sl, su, sr = orc.propagate_scip(om, inst.lb, inst.ub)
        ttf.append(time.perf_counter() - t)
    same = bool((st["result"] == sa.SAP_CUTOFF) == bool(sr["cutoff"])
                and (sr["cutoff"] or (np.array_equal(gl, sl) and np.array_equal(gu, su))))
    return {"workload": "covering/packing 2000x5000, 10 nnz/row, coefficients U{1..100}, sides 0.4 x row sum, "
                        "seed 1; root bounds to fixpoint",
            "gpu_ms_to_fixpoint": gpu_ms, "gpu_device_ms": dev / reps, "gpu_rounds": st["rounds"],
            "gpu_nchgbds": st["nchgbds"], "gpu_nnz_swept": st["nnz_swept"],
            "cpu_ms_to_fixpoint": 1000 * min(ttf), "cpu_runs": len(ttf), "cpu_rounds": sr["rounds"], "cpu_cores": 1,
            "cpu_kind": "SCIP-order restatement (oracle/oracle.c orc_propagate_scip)",
            "speedup_time_to_fixpoint": (1000 * min(ttf)) / gpu_ms, "same_fixpoint": same}


def run_legs(sa, ctx, args, main_out):
    legs = [x for x in args.legs.split(",") if x and x != "none"]
    parts = None
    if "cfg1" in legs:
        main_out["config1_cpu_vs_gpu"] = run_config1(sa, ctx, args)
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
    # the memory floor: the sweep's streams and bound gathers over every nonzero, no arithmetic
    floor_ms = min(init.bench_gather(20, blocks) for blocks in (0, 1024, 4096))
    floor_bytes = SWEEP_BYTES_PER_NNZ * inst.nnz + SWEEP_BYTES_PER_ROW * inst.nrows
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
        "gather_floor": {
            "what": "k_gather_floor: the sweep's value/column streams and (lb, ub) gather of every nonzero, "
                    "no arithmetic (sap_bench_gather, HIP events, 20 launches; the fastest of one lane per "
                    "4 nonzeros, 1,024 and 4,096 workgroups)",
            "ms": round(floor_ms, 4),
            "achieved": round(floor_bytes / (floor_ms * 1e-3) / 1e9, 1) if floor_ms > 0 else None,
            "frac": round(floor_bytes / (floor_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if floor_ms > 0 else None,
        },
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
            "note": ("one-round cutoff: the literal sides are jointly infeasible at this size (the oracle agrees), "
                     "so a step is one sweep plus the column step that finds it; the planted variant "
                     "(feasible_variant) runs the fixpoint loop" if st["result"] == sa.SAP_CUTOFF and rounds == 1
                     else "propagated to its fixpoint"),
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
    devms = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        out = nodes.propagate_nodes(base, 0, ptr, col, val, val, p)
        nnz += out["stats"]["nnz_swept"]
        rows += out["stats"]["rows_swept"]
        devms += out["stats"]["device_ms"]
    el = time.perf_counter() - t0
    res = out["result"]
    # batched algorithmic bytes per call: the matrix once for the union of rows any vector
    # swept (12 B per nonzero: value and column index), per vector and swept nonzero the bound
    # gather (16 B), per swept row its pointer and sides (20 B), per changed column its bounds
    # read and written and its candidate keys (32 B).  A row is swept only when it holds an
    # overlay or a changed column of its vector, so the rows holding any of those bound the union.
    hit = np.zeros(inst.ncols, dtype=bool)
    hit[col] = True
    hit[out["chgcol"]] = True
    nzrow = np.repeat(np.arange(inst.nrows), np.diff(inst.rowptr))
    rowhit = np.zeros(inst.nrows, dtype=bool)
    rowhit[nzrow[hit[inst.colidx]]] = True
    nnz_union = int(np.diff(inst.rowptr)[rowhit].sum())
    bbytes = 12 * nnz_union + 16 * (nnz // steps) + 20 * (rows // steps) + 32 * int(out["chgptr"][-1])
    bach = bbytes / (devms / steps * 1e-3) / 1e9 if devms > 0 else 0.0
    batched = {"value": _rate(nnz, el), "unit": "nnz-sweeps/s", "ms_per_batch": 1000 * el / steps,
               "nodes_per_s": nn * steps / el, "nnz_swept_per_batch": nnz // steps,
               "rows_swept_per_batch": rows // steps,
               "cutoff_nodes": int((res == sa.SAP_CUTOFF).sum()),
               "reduceddom_nodes": int((res == sa.SAP_REDUCEDDOM).sum()),
               "max_rounds": int(out["rounds"].max()), "changed_bounds": int(out["chgptr"][-1]),
               "roofline": {"bound": "hbm", "algorithmic_bytes_per_call": bbytes, "nnz_union": nnz_union,
                            "device_ms_per_call": devms / steps, "achieved": round(bach, 1), "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": round(bach / HBM_PEAK_GBS, 4),
                            "formula": "12*nnz_union + 16*sum nnz_swept + 20*sum rows_swept + 32*changed columns, "
                                       "over the device time of the whole call"}}
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
    log(f"[bench] generating {count} uneven instances (rows {int(sa.uneven_sizes(count, seed).sum())})")
    t = time.perf_counter()
    parts = sa.gen_uneven_instances(count, seed, threads=16)
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


def _pct(x, q):
    return float(np.percentile(x, q)) if len(x) else None


def _serve_summary(r, n, sel=None):
    """Throughput and latency of one replay (over the requests sel, default all)."""
    st = r["stats"]
    lat = r["latency_ms"] if sel is None else r["latency_ms"][sel]
    return {"value": _rate(st["nnz_swept"], st["wall_s"]), "unit": "nnz-sweeps/s", "requests": n,
            "requests_per_s": n / st["wall_s"] if st["wall_s"] > 0 else 0.0, "wall_s": st["wall_s"],
            "p50_ms": _pct(lat, 50), "p99_ms": _pct(lat, 99), "max_ms": float(lat.max()) if len(lat) else None,
            "steps": st["steps"], "max_running": st["max_running"], "uploads": st["uploads"],
            "evictions": st["evictions"], "peak_bytes": st["peak_bytes"], "nnz_swept": st["nnz_swept"]}


def run_serving(sa, args, parts):
    """Poisson arrivals (rate, seconds from the arguments).  Half the requests are nodes of the
    covering/packing 200,000 x 300,000 problem (1..50 binaries fixed); half are nodes (1..50
    integer columns fixed to 0/1) of one of the 512 uneven instances, chosen with Zipf(1.2)
    popularity over a random order, so request sizes follow the instances' size law; slot pools
    are sized by each problem's share of the requests.  The HBM budget holds the
    covering/packing problem and 85 % of the instances' footprint: instances
    are uploaded on demand and evicted least recently used during the trace.  Every replay
    starts from the same resident set (the most popular instances that fit).  Continuous
    batching against the plain one-request-at-a-time form: each at the offered rate (latency)
    and saturated, all requests at t = 0 (capacity)."""
    cp = sa.gen_covering_packing(200_000, 300_000, 10, seed=3)
    probs = [cp] + list(parts)
    arrival, prob, ptr, col, val, order = sa.gen_serving_trace(probs, args.serve_rate, args.serve_seconds, seed=5)
    n = len(arrival)
    srv = sa.Server(0, 4, 1 << 50, 8)
    pids = [srv.add_problem(p) for p in probs]
    assert pids == list(range(len(probs)))
    # slot pools sized by demand, as a deployment would size them from its request mix: the
    # covering/packing problem 256 slots, an instance 8 to 128 (a slot per ~64 of its requests
    # in the trace, rounded up to a power of two)
    srv.set_slots(0, 256)
    demand = np.bincount(prob, minlength=len(probs))
    for pid in range(1, len(probs)):
        want = int(2 ** np.ceil(np.log2(max(demand[pid] / 64.0, 1.0))))
        srv.set_slots(pid, int(min(128, max(8, want))))
    foot = np.array([srv.footprint(pid) for pid in pids], dtype=np.int64)
    budget = int(foot[0] + 0.85 * foot[1:].sum())
    srv.set_budget(budget)
    # the steady resident set: the covering/packing problem and the most popular instances
    # that fit the budget, warmed least popular first so LRU order follows popularity
    hot, used = [], foot[0]
    for pid in order:
        if used + foot[pid] > budget:
            break
        hot.append(int(pid))
        used += foot[pid]

    def warm_state():
        srv.warm(0)
        for pid in reversed(hot):
            srv.warm(pid)

    t = time.perf_counter()
    warm_state()
    log(f"[bench] serving: budget {budget / 2**30:.1f} GiB of {foot.sum() / 2**30:.1f} GiB; "
        f"{len(hot)} instances warm in {time.perf_counter() - t:.1f}s")
    is_cp = prob == 0
    log(f"[bench] serving trace: {n} requests over {args.serve_seconds}s")

    def prefix(m, at_zero=False):
        return (np.zeros(m) if at_zero else arrival[:m], prob[:m], ptr[:m + 1], col[:ptr[m]], val[:ptr[m]],
                val[:ptr[m]])

    # continuous batching at the offered rate
    warm_state()
    cont = srv.replay(arrival, prob, ptr, col, val, val, mode=sa.SAP_SERVE_CONTINUOUS, steprounds=sa.SAP_STEP_SHORT)
    log(f"[bench] serving continuous: {cont['stats']}")
    # saturated capacity: the same requests all at t = 0
    warm_state()
    sat = srv.replay(np.zeros(n), prob, ptr, col, val, val, mode=sa.SAP_SERVE_CONTINUOUS,
                     steprounds=sa.SAP_STEP_LONG)
    # plain: saturated on a prefix (capacity), and at the offered rate on a 10 s prefix (latency)
    npl = min(n, max(2000, n // 12))
    warm_state()
    plsat = srv.replay(*prefix(npl, at_zero=True), mode=sa.SAP_SERVE_PLAIN, steprounds=sa.SAP_STEP_LONG)
    nof = min(n, int(10 * args.serve_rate))
    warm_state()
    plof = srv.replay(*prefix(nof), mode=sa.SAP_SERVE_PLAIN, steprounds=sa.SAP_STEP_LONG)
    same = bool(np.array_equal(plof["result"], cont["result"][:nof]) and np.array_equal(plof["nchg"], cont["nchg"][:nof])
                and np.array_equal(sat["result"], cont["result"]) and np.array_equal(sat["nchg"], cont["nchg"]))
    srv.close()
    c_all = _serve_summary(cont, n)
    c_all.update({
        "covering_packing_nodes": {"requests": int(is_cp.sum()), "p50_ms": _pct(cont["latency_ms"][is_cp], 50),
                                   "p99_ms": _pct(cont["latency_ms"][is_cp], 99)},
        "instance_nodes_resident": {"requests": int((~is_cp & ~cont["cold"]).sum()),
                                    "p50_ms": _pct(cont["latency_ms"][~is_cp & ~cont["cold"]], 50),
                                    "p99_ms": _pct(cont["latency_ms"][~is_cp & ~cont["cold"]], 99)},
        "instance_nodes_cold": {"requests": int(cont["cold"].sum()),
                                "p50_ms": _pct(cont["latency_ms"][cont["cold"]], 50),
                                "p99_ms": _pct(cont["latency_ms"][cont["cold"]], 99)},
    })
    p_of = _serve_summary(plof, nof)
    p_of["p50_ms_resident"] = _pct(plof["latency_ms"][~plof["cold"]], 50)
    p_of["p99_ms_resident"] = _pct(plof["latency_ms"][~plof["cold"]], 99)
    c_sat = _serve_summary(sat, n)
    p_sat = _serve_summary(plsat, npl)
    return {"workload": f"Poisson {args.serve_rate:.0f} req/s for {args.serve_seconds:.0f}s; 50% nodes of the "
                        f"200000x300000 covering/packing problem (1..50 binaries fixed), 50% nodes of the "
                        f"{len(parts)} uneven instances (Zipf 1.2 popularity, 1..50 integer columns fixed); "
                        f"HBM budget {budget / 2**30:.1f} GiB of {foot.sum() / 2**30:.1f} GiB total footprint "
                        "(LRU eviction and on-demand upload); same warm set before each replay",
            "requests": n,
            "continuous": c_all,
            "continuous_saturated": c_sat,
            "plain_offered_rate": p_of,
            "plain_saturated": p_sat,
            "results_equal_across_forms": same,
            "speedup_nnz_sweeps_saturated": c_sat["value"] / max(p_sat["value"], 1e-9),
            "speedup_requests_saturated": c_sat["requests_per_s"] / max(p_sat["requests_per_s"], 1e-9)}


if __name__ == "__main__":
    main()
