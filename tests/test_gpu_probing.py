"""GPU parity of probing (sap_probe): both branches of every probe column propagated in
batched launches, compared with the oracle's probing restatement (oracle.probe) -- identical
verdicts, identical lists of columns changed in both branches, bit-exact branch bounds, and
identical round counts when every row is swept first."""
import os

import numpy as np
import pytest

import scip_amd as sa
from oracle import oracle as orc
from tests.util import INF, make_instance, random_binary_instance

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def say(*a):
    print(*a, flush=True)


def check_against_oracle(om, lb, ub, probecols, out, check_rounds):
    ref = orc.probe(om, lb, ub, probecols)
    for p, r in enumerate(ref):
        assert out["verdict"][p] == r["verdict"], (p, out["verdict"][p], r["verdict"])
        if check_rounds:
            assert tuple(out["rounds"][p]) == r["rounds"], p
        a, e = out["dedptr"][p], out["dedptr"][p + 1]
        assert np.array_equal(out["dedcol"][a:e], r["cols"]), p
        for k, (lo, hi) in (("0", ("dlb0", "dub0")), ("1", ("dlb1", "dub1"))):
            got = np.stack([out[lo][a:e], out[hi][a:e]], 1)
            assert np.array_equal(got.view(np.int64), r["b" + k].view(np.int64)), (p, k)
    return ref


def binary_problem(seed, nrows=2500, ncols=1800):
    rng = np.random.default_rng(seed)
    inst = random_binary_instance(rng, nrows=nrows, ncols=ncols, maxlen=12)
    om = orc.OMatrix.from_instance(inst)
    rl, ru, rr = orc.propagate_jacobi(om, inst.lb, inst.ub)
    assert not rr["cutoff"]
    return inst, om, rl, ru


@pytest.mark.parametrize("mode", [sa.SAP_MODE_DENSE, sa.SAP_MODE_SPARSE])
@pytest.mark.parametrize("flags", [0, sa.SAP_FLAG_BASE_IS_FIXPOINT])
def test_probe_matches_oracle(gpu_ctx, mode, flags):
    inst, om, rl, ru = binary_problem(11)
    free = np.nonzero(rl < ru)[0].astype(np.int32)
    probecols = free[np.random.default_rng(3).permutation(len(free))[:300]]
    say(f"[case] probe mode={mode} flags={flags} nprobe={len(probecols)}")
    mat = sa.Matrix(gpu_ctx, inst)
    base = sa.Bounds(mat, 1)
    base.upload(rl, ru)
    # 64 vectors: 32 probes per launch, so the 300 probes take 10 chunks
    b = sa.Bounds(mat, 64, mode)
    out = b.probe(base, 0, probecols, sa.default_params(flags=flags))
    ref = check_against_oracle(om, rl, ru, probecols, out, check_rounds=(flags == 0))
    # the case exercises every outcome
    assert any(r["verdict"] == 1 for r in ref) and any(r["verdict"] == 2 for r in ref)
    assert sum(len(r["cols"]) for r in ref) > 50
    # again on the same vectors (restored from the change logs): same answer
    out2 = b.probe(base, 0, probecols, sa.default_params(flags=flags))
    for k in ("verdict", "dedptr", "dedcol", "dlb0", "dub0", "dlb1", "dub1"):
        assert np.array_equal(out[k], out2[k]), k
    b.close()
    base.close()
    mat.close()


def test_probe_base_inside_the_probe_set_and_capacity(gpu_ctx):
    inst, om, rl, ru = binary_problem(12, nrows=1200, ncols=900)
    free = np.nonzero(rl < ru)[0].astype(np.int32)
    mat = sa.Matrix(gpu_ctx, inst)
    b = sa.Bounds(mat, 16)
    b.upload(rl, ru)   # every vector holds the root fixpoint; vector 3 is the base
    say(f"[case] probe base inside, nprobe={len(free)}")
    out = b.probe(b, 3, free, sa.default_params(flags=sa.SAP_FLAG_BASE_IS_FIXPOINT))
    check_against_oracle(om, rl, ru, free, out, check_rounds=False)
    # too little room: SAP_NOMEMORY with the full counts, then the same answer with room
    need = int(out["dedptr"][-1])
    assert need > 1
    b.upload(rl, ru)
    with pytest.raises(sa.SapError) as e:
        b.probe(b, 3, free, sa.default_params(flags=sa.SAP_FLAG_BASE_IS_FIXPOINT), dedcap=need - 1)
    assert e.value.rc == sa.SAP_NOMEMORY
    b.upload(rl, ru)
    out2 = b.probe(b, 3, free, sa.default_params(flags=sa.SAP_FLAG_BASE_IS_FIXPOINT), dedcap=need)
    for k in ("verdict", "dedptr", "dedcol", "dlb0", "dub0", "dlb1", "dub1"):
        assert np.array_equal(out[k], out2[k]), k
    b.close()
    mat.close()


def test_probe_long_pair_lists(gpu_ctx):
    # x + y = 1 written as two rows for every column y of a probe's fan: each branch of probe 0
    # fixes its 5,000 fan columns, of probe 1 its ten, so the chunk's pair lists (5,010 entries)
    # are longer than those read back together with their counts
    k0, k1 = 5000, 10
    rows = []
    for k in range(k0 + k1):
        x = 0 if k < k0 else 1
        rows.append(([(x, 1.0), (2 + k, 1.0)], 1.0, INF))
        rows.append(([(x, 1.0), (2 + k, 1.0)], -INF, 1.0))
    n = 2 + k0 + k1
    inst = make_instance(rows, [0.0] * n, [1.0] * n, [1] * n)
    om = orc.OMatrix.from_instance(inst)
    probecols = np.array([0, 1], dtype=np.int32)
    say(f"[case] probe long pair lists, fans {k0} and {k1}")
    mat = sa.Matrix(gpu_ctx, inst)
    base = sa.Bounds(mat, 1)
    base.upload(inst.lb, inst.ub)
    b = sa.Bounds(mat, 4)
    out = b.probe(base, 0, probecols, sa.default_params())
    ref = check_against_oracle(om, inst.lb, inst.ub, probecols, out, check_rounds=True)
    assert [len(r["cols"]) for r in ref] == [k0, k1]
    assert list(out["verdict"]) == [0, 0]
    b.close()
    base.close()
    mat.close()


def test_probe_rejects_columns_that_are_not_unfixed_binaries(gpu_ctx):
    inst, om, rl, ru = binary_problem(13, nrows=300, ncols=200)
    fixed = np.nonzero(rl == ru)[0]
    free = np.nonzero(rl < ru)[0]
    mat = sa.Matrix(gpu_ctx, inst)
    base = sa.Bounds(mat, 1)
    base.upload(rl, ru)
    b = sa.Bounds(mat, 4)
    say("[case] probe invalid columns")
    for cols in ([int(fixed[0])] if len(fixed) else [], [inst.ncols], [-1]):
        if not cols:
            continue
        with pytest.raises(sa.SapError) as e:
            b.probe(base, 0, np.array([int(free[0])] + cols, dtype=np.int32))
        assert e.value.rc == sa.SAP_INVALIDDATA
    # a general-integer column is not a binary either
    lb2, ub2 = rl.copy(), ru.copy()
    ub2[free[1]] = 2.0
    base.upload(lb2, ub2)
    with pytest.raises(sa.SapError) as e:
        b.probe(base, 0, np.array([int(free[1])], dtype=np.int32))
    assert e.value.rc == sa.SAP_INVALIDDATA
    # an empty probe list is a no-op
    out = b.probe(base, 0, np.zeros(0, dtype=np.int32))
    assert len(out["verdict"]) == 0 and out["dedptr"].tolist() == [0]
    b.close()
    base.close()
    mat.close()


@pytest.mark.parametrize("name", ["p0548", "bell5", "misc03", "lseu"])
def test_probe_real_instances(gpu_ctx, name):
    inst, _ = sa.read_mps(os.path.join(HERE, "golden", "mps", name + ".mps"))
    om = orc.OMatrix.from_instance(inst)
    rl, ru, rr = orc.propagate_jacobi(om, inst.lb, inst.ub)
    assert not rr["cutoff"]
    probecols = np.nonzero((inst.integral != 0) & (rl == 0.0) & (ru == 1.0))[0].astype(np.int32)
    say(f"[case] probe {name} nprobe={len(probecols)}")
    mat = sa.Matrix(gpu_ctx, inst)
    base = sa.Bounds(mat, 1)
    base.upload(rl, ru)
    b = sa.Bounds(mat, 256)
    out = b.probe(base, 0, probecols, sa.default_params(flags=0))
    check_against_oracle(om, rl, ru, probecols, out, check_rounds=True)
    b.close()
    base.close()
    mat.close()


@pytest.mark.parametrize("seed", range(2))
def test_split_probe_matches_oracle(gpu_ctx, seed):
    # strong-branching children x <= floor(mid) / x >= floor(mid) + 1 of general-integer columns
    from tests.util import planted_integer_instance
    inst = planted_integer_instance(np.random.default_rng(960 + seed), 400, 300, 6)
    om = orc.OMatrix.from_instance(inst)
    # from the unpropagated bounds (every row swept first): both infeasible children and
    # long pair lists occur
    rl, ru = inst.lb, inst.ub
    cols = np.nonzero(ru - rl >= 1)[0].astype(np.int32)
    down = np.floor((rl[cols] + ru[cols]) / 2)
    up = down + 1
    say(f"[case] split probe seed={seed} nprobe={len(cols)}")
    mat = sa.Matrix(gpu_ctx, inst)
    base = sa.Bounds(mat, 1)
    base.upload(rl, ru)
    b = sa.Bounds(mat, 32)
    out = b.probe(base, 0, cols, sa.default_params(flags=0), downub=down, uplb=up)
    ref = orc.probe(om, rl, ru, cols, downub=down, uplb=up)
    for p, r in enumerate(ref):
        assert out["verdict"][p] == r["verdict"], p
        assert tuple(out["rounds"][p]) == r["rounds"], p
        a, e = out["dedptr"][p], out["dedptr"][p + 1]
        assert np.array_equal(out["dedcol"][a:e], r["cols"]), p
        for k, (lo, hi) in (("0", ("dlb0", "dub0")), ("1", ("dlb1", "dub1"))):
            got = np.stack([out[lo][a:e], out[hi][a:e]], 1)
            assert np.array_equal(got.view(np.int64), r["b" + k].view(np.int64)), (p, k)
    assert any(r["verdict"] for r in ref) and sum(len(r["cols"]) for r in ref) > len(ref)
    # a split outside the base bounds, or down >= up, is rejected
    for dn, u in ((down[0] + 1, up[0]), (rl[cols[0]] - 1, up[0])):
        with pytest.raises(sa.SapError) as e:
            b.probe(base, 0, cols[:1], downub=np.array([dn]), uplb=np.array([u]))
        assert e.value.rc == sa.SAP_INVALIDDATA
    b.close()
    base.close()
    mat.close()
