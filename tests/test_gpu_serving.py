"""Serving (include/scip_amd_serve.h): continuous batching of node requests over several
problems and streams.  Each request's result -- verdict, rounds and every changed bound -- is
the same alone (plain mode: one request at a time) and in the mix (continuous batching), the
same under a tight HBM budget that forces evictions, and equal to the oracle's Jacobi twin on
that node alone.
"""
import numpy as np
import pytest

import scip_amd as sa
from oracle import oracle as orc
from tests.util import make_instance, random_binary_instance, same_bounds

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def problems():
    rng = np.random.default_rng(21)
    ps = [random_binary_instance(rng, nrows=3000, ncols=2000, maxlen=12)]
    for i, n in enumerate((800, 2500, 6000)):
        ps.append(sa.gen_mixed(n, int(1.5 * n), seed=5000 + i, lenmode=1, minlen=4, maxlen=64, planted=True))
    # an infeasible root: x0 + x1 >= 3 with binaries
    ps.append(make_instance([([(0, 1.0), (1, 1.0)], 3.0, 1e20)], [0.0, 0.0], [1.0, 1.0], [1, 1]))
    return ps


def roots(ps):
    out = []
    for p in ps:
        l, u, r = orc.propagate_jacobi(orc.OMatrix.from_instance(p), p.lb, p.ub)
        out.append((l, u, bool(r["cutoff"])))
    return out


def trace(ps, rts, n, rate, seed):
    rng = np.random.default_rng(seed)
    arrival = np.cumsum(rng.exponential(1.0 / rate, size=n))
    arrival -= arrival[0]
    prob = rng.integers(0, len(ps), size=n).astype(np.int32)
    ptr = [0]
    col, val = [], []
    for i in range(n):
        p = ps[prob[i]]
        l, u, _ = rts[prob[i]]
        cand = np.nonzero((p.integral != 0) & (l < u))[0]
        k = min(len(cand), int(rng.integers(1, 12)))
        if k:
            cs = rng.choice(cand, size=k, replace=False)
            for c in cs:
                col.append(int(c))
                val.append(float(l[c] if rng.integers(0, 2) == 0 else u[c]))
        ptr.append(len(col))
    return (arrival, prob, np.array(ptr, dtype=np.int64), np.array(col, dtype=np.int32),
            np.array(val, dtype=np.float64))


def run(ps, tr, mode, budget=8 << 30, slots=16, nstreams=3, pool=None):
    srv = sa.Server(0, nstreams, budget, slots)
    for p in ps:
        srv.add_problem(p)
    for pid, n in (pool or {}).items():
        srv.set_slots(pid, n)
    arrival, prob, ptr, col, val = tr
    out = srv.replay(arrival, prob, ptr, col, val, val, mode=mode, want_changes=True)
    srv.close()
    return out


def same_request(a, b, i):
    assert a["result"][i] == b["result"][i], i
    assert a["rounds"][i] == b["rounds"][i], i
    if a["result"][i] == sa.SAP_CUTOFF:
        return
    x0, x1 = a["chgptr"][i], a["chgptr"][i + 1]
    y0, y1 = b["chgptr"][i], b["chgptr"][i + 1]
    assert np.array_equal(a["chgcol"][x0:x1], b["chgcol"][y0:y1]), i
    assert same_bounds(a["chglb"][x0:x1], b["chglb"][y0:y1]) and same_bounds(a["chgub"][x0:x1], b["chgub"][y0:y1])


def test_same_alone_and_in_the_mix():
    ps = problems()
    rts = roots(ps)
    tr = trace(ps, rts, n=240, rate=20000.0, seed=3)
    say(f"[case] serving mix {len(tr[0])} requests over {len(ps)} problems")
    mix = run(ps, tr, sa.SAP_SERVE_CONTINUOUS)
    alone = run(ps, tr, sa.SAP_SERVE_PLAIN)
    st = mix["stats"]
    say(f"[info] mix: {st}")
    assert st["max_running"] > 1          # requests really overlapped
    for i in range(len(tr[0])):
        same_request(mix, alone, i)
    assert mix["stats"]["nnz_swept"] == alone["stats"]["nnz_swept"]
    # against the oracle on each node alone
    arrival, prob, ptr, col, val = tr
    for i in range(0, len(arrival), 3):
        p = ps[prob[i]]
        l, u, rootcut = rts[prob[i]]
        if rootcut:
            assert mix["result"][i] == sa.SAP_CUTOFF
            continue
        lb0, ub0 = l.copy(), u.copy()
        lb0[col[ptr[i]:ptr[i + 1]]] = val[ptr[i]:ptr[i + 1]]
        ub0[col[ptr[i]:ptr[i + 1]]] = val[ptr[i]:ptr[i + 1]]
        ol, ou, orr = orc.propagate_jacobi(orc.OMatrix.from_instance(p), lb0, ub0)
        assert (mix["result"][i] == sa.SAP_CUTOFF) == bool(orr["cutoff"]), i
        if orr["cutoff"]:
            continue
        a, e = mix["chgptr"][i], mix["chgptr"][i + 1]
        gl, gu = lb0.copy(), ub0.copy()
        gl[mix["chgcol"][a:e]] = mix["chglb"][a:e]
        gu[mix["chgcol"][a:e]] = mix["chgub"][a:e]
        assert same_bounds(gl, ol) and same_bounds(gu, ou), i
        assert np.array_equal(mix["chgcol"][a:e], np.nonzero((ol != lb0) | (ou != ub0))[0]), i


def test_admission_under_a_tight_budget():
    ps = problems()
    rts = roots(ps)
    tr = trace(ps, rts, n=120, rate=5000.0, seed=4)
    roomy = run(ps, tr, sa.SAP_SERVE_CONTINUOUS)
    # room for about two of the problems at a time: requests wait and problems are evicted
    tight_budget = 20 << 20
    tight = run(ps, tr, sa.SAP_SERVE_CONTINUOUS, budget=tight_budget, slots=16)
    say(f"[info] tight: {tight['stats']}")
    assert tight["stats"]["evictions"] > 0 and tight["stats"]["uploads"] > len(ps)
    assert tight["stats"]["peak_bytes"] <= tight_budget
    for i in range(len(tr[0])):
        same_request(roomy, tight, i)


def test_problem_larger_than_budget_is_an_error():
    ps = problems()[:2]
    rts = roots(ps)
    tr = trace(ps, rts, n=4, rate=1000.0, seed=5)
    srv = sa.Server(0, 1, 1 << 20, 4)
    for p in ps:
        srv.add_problem(p)
    with pytest.raises(sa.SapError) as e:
        srv.replay(*tr[:4], tr[4], tr[4])
    assert e.value.rc == sa.SAP_NOMEMORY
    srv.close()


def test_pool_sizes_do_not_change_results():
    # per-problem pools: a large one for problem 0, a single slot for problem 2, and a resize
    # of a resident problem between two replays
    ps = problems()
    rts = roots(ps)
    tr = trace(ps, rts, n=160, rate=20000.0, seed=6)
    say(f"[case] serving pool sizes {len(tr[0])} requests")
    base = run(ps, tr, sa.SAP_SERVE_CONTINUOUS)
    sized = run(ps, tr, sa.SAP_SERVE_CONTINUOUS, pool={0: 64, 2: 1})
    for i in range(len(tr[0])):
        same_request(base, sized, i)
    srv = sa.Server(0, 2, 8 << 30, 8)
    for p in ps:
        srv.add_problem(p)
    srv.warm(0)
    arrival, prob, ptr, col, val = tr
    a = srv.replay(arrival, prob, ptr, col, val, val, mode=sa.SAP_SERVE_CONTINUOUS, want_changes=True)
    srv.set_slots(0, 40)
    b = srv.replay(arrival, prob, ptr, col, val, val, mode=sa.SAP_SERVE_CONTINUOUS, want_changes=True)
    assert b["stats"]["uploads"] >= 1
    with pytest.raises(sa.SapError):
        srv.set_slots(0, 0)
    srv.close()
    for i in range(len(tr[0])):
        same_request(a, b, i)
        same_request(a, base, i)
