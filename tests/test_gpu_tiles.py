"""Tile layouts of the dense sweep (one, two or four nonzeros per lane, rows packed first-fit
decreasing, rows of up to 128 nonzeros in tiles) against the oracle's Jacobi twin: row-length
patterns that exercise every padding (0..3 empty positions at a row's end), full tiles of
single-nonzero rows, a row that fills a whole tile, and rows just above the tiled lengths."""
import numpy as np
import pytest

import scip_amd as sa
from oracle import oracle as orc
from tests.util import INF, make_instance, same_bounds

pytestmark = pytest.mark.gpu

LAYOUTS = [(5, 1), (5, 2), (6, 2), (5, 4), (6, 4), (3, 4)]   # (tilebins, tilewidth)


def say(*a):
    print(*a, flush=True)


def planted(rng, lengths, ncols, intfrac=0.5, inffrac=0.03, hugefrac=0.0):
    """Rows of the given lengths over ncols columns, mixed integral/continuous, sides around the
    activity of a hidden point so that rows propagate."""
    integral = (rng.uniform(size=ncols) < intfrac).astype(np.uint8)
    lb = np.where(integral == 1, 0.0, -rng.uniform(0, 50, ncols))
    ub = np.where(integral == 1, rng.integers(1, 20, ncols).astype(float), rng.uniform(0, 50, ncols))
    # the hidden point: an integer in [lb, ub], or a value in [max(lb, -10), min(ub, 10)]
    x = np.where(integral == 1, np.floor(lb + rng.uniform(size=ncols) * (ub - lb + 1)),
                 rng.uniform(np.maximum(lb, -10.0), np.minimum(ub, 10.0)))
    x = np.minimum(x, ub)
    # infinite bounds on continuous columns only, beyond the hidden point
    u = rng.uniform(size=ncols)
    cont = integral == 0
    lb[cont & (u < inffrac)] = -INF
    ub[cont & (u >= inffrac) & (u < 2 * inffrac)] = INF
    if hugefrac > 0:
        h = cont & (rng.uniform(size=ncols) < hugefrac)
        lb[h] = 1e16
        ub[h] = 1e16
        x[h] = 1e16
    rows = []
    for L in lengths:
        cols = rng.choice(ncols, size=min(int(L), ncols), replace=False)
        a = rng.uniform(0.5, 9.5, len(cols)) * rng.choice([-1.0, 1.0], len(cols))
        if rng.uniform() < 0.5:
            a = np.round(a)
            a[a == 0] = 1.0
        act = float(a @ x[cols])
        slack = abs(act) * 0.02 + rng.uniform(0, 3)
        kind = rng.integers(0, 3)
        lo = act - slack if kind != 1 else -INF
        hi = act + slack if kind != 0 else INF
        rows.append(([(int(c), float(v)) for c, v in zip(cols, a)], lo, hi))
    return make_instance(rows, lb, ub, integral)


def check(ctx, inst, name):
    om = orc.OMatrix.from_instance(inst)
    ol, ou, orr = orc.propagate_jacobi(om, inst.lb, inst.ub)
    for tb, tw in LAYOUTS:
        say(f"[case] {name} tilebins={tb} tilewidth={tw} rows={inst.nrows} nnz={inst.nnz}")
        ctx.set_tuning(tilebins=tb, tilewidth=tw)
        try:
            mat = sa.Matrix(ctx, inst)
            b = sa.Bounds(mat, 1, sa.SAP_MODE_DENSE)
            b.upload(inst.lb, inst.ub)
            st = b.propagate()
            gl, gu = b.download(0)
            assert (st["result"] == sa.SAP_CUTOFF) == bool(orr["cutoff"]), (name, tb, tw)
            assert st["rounds"] == orr["rounds"], (name, tb, tw)
            if not orr["cutoff"]:
                assert same_bounds(gl, ol) and same_bounds(gu, ou), (name, tb, tw)
                assert st["nchgbds"] == orr["nchgbds"]
                assert st["rows_swept"] == orr["rows_swept"] and st["nnz_swept"] == orr["nnz_swept"]
            b.close()
            mat.close()
        finally:
            ctx.set_tuning()
    return orr


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_every_length_up_to_130(gpu_ctx, seed):
    rng = np.random.default_rng(seed)
    lengths = np.tile(np.arange(1, 131), 6)
    rng.shuffle(lengths)
    orr = check(gpu_ctx, planted(rng, lengths, 3000), f"lengths 1..130 seed {seed}")
    assert orr["nchgbds"] > 0 or orr["cutoff"]


@pytest.mark.parametrize("pattern", ["ones", "ones_twos", "pad3", "pad1", "full128", "odd127_129", "huge"])
def test_length_patterns(gpu_ctx, pattern):
    rng = np.random.default_rng(7)
    hugefrac = 0.0
    if pattern == "ones":
        lengths = np.ones(700, dtype=int)
    elif pattern == "ones_twos":
        lengths = rng.integers(1, 3, 1500)
    elif pattern == "pad3":
        lengths = 4 * rng.integers(0, 30, 600) + 1
    elif pattern == "pad1":
        lengths = 4 * rng.integers(0, 30, 600) + 3
    elif pattern == "full128":
        lengths = np.full(200, 128)
    elif pattern == "odd127_129":
        lengths = rng.choice([127, 128, 129, 63, 65, 1], 400)
    else:
        lengths = rng.integers(1, 100, 800)
        hugefrac = 0.01
    check(gpu_ctx, planted(rng, lengths, 2500, hugefrac=hugefrac), pattern)


def test_generated_instance_all_layouts(gpu_ctx):
    inst = sa.gen_mixed(30000, 30000, seed=77, planted=True, maxlen=400)
    orr = check(gpu_ctx, inst, "mixed 30k planted")
    assert orr["rounds"] >= 2
