"""Row binning / launch shape (sap_tuning) changes only how rows are swept, never the result:
every tuning gives the oracle's bounds bit for bit; the autotuner returns a valid tuning."""
import numpy as np
import pytest

import scip_amd as sa
from oracle import oracle as orc
from tests.util import same_bounds

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("tilebins,longrow,capscale,rowbins,tilewidth", [
    (6, 4096, 100, 0, 2), (6, 64, 100, 0, 2), (6, 100, 100, 0, 2), (6, 4096, 100, 3, 2), (5, 4096, 100, 0, 2),
    (6, 4096, 100, 0, 4), (6, 100, 50, 0, 4), (3, 4096, 100, 0, 4), (6, 4096, 100, 2, 4),
    (5, 4096, 100, 0, 1), (0, 4096, 100, 0, 2), (3, 64, 10, 0, 2), (3, 64, 10, 0, 1),
    (4, 1024, 200, 0, 2), (1, 128, 1000, 0, 2), (1, 128, 1000, 0, 1), (5, 65536, 50, 0, 2), (5, 4096, 100, 1, 2),
    (5, 4096, 100, 2, 2), (5, 4096, 100, 3, 2), (5, 4096, 100, 3, 1), (2, 4096, 100, 3, 2), (0, 256, 100, 3, 2)])
def test_tunings_give_identical_results(gpu_ctx, tilebins, longrow, capscale, rowbins, tilewidth):
    inst = sa.gen_mixed(20000, 20000, seed=31, planted=True, maxlen=3000)
    ol, ou, orr = orc.propagate_jacobi(orc.OMatrix.from_instance(inst), inst.lb, inst.ub)
    say(f"[case] tuning {tilebins} {longrow} {capscale} {rowbins} {tilewidth}")
    gpu_ctx.set_tuning(tilebins, longrow, capscale, rowbins, tilewidth)
    try:
        mat = sa.Matrix(gpu_ctx, inst)
        for mode in (sa.SAP_MODE_DENSE, sa.SAP_MODE_SPARSE):
            b = sa.Bounds(mat, 1, mode)
            b.upload(inst.lb, inst.ub)
            st = b.propagate()
            gl, gu = b.download(0)
            assert (st["result"] == sa.SAP_CUTOFF) == bool(orr["cutoff"])
            assert st["rounds"] == orr["rounds"] and st["nnz_swept"] == orr["nnz_swept"]
            if not orr["cutoff"]:
                assert same_bounds(gl, ol) and same_bounds(gu, ou), mode
            b.close()
        mat.close()
    finally:
        gpu_ctx.set_tuning()


def test_autotune_returns_a_valid_tuning(gpu_ctx):
    inst = sa.gen_mixed(60000, 60000, seed=32, planted=True)
    best, bms, dms = gpu_ctx.autotune(inst, samplerows=20000, apply=False)
    say(f"[info] autotune best={best} best_ms={bms:.3f} default_ms={dms:.3f}")
    assert best["tilebins"] in (3, 4, 5, 6) and best["longrow"] in (1024, 4096, 16384) and best["rowbins"] in (0, 2, 3)
    assert best["capscale"] in (25, 50, 100) and best["tilewidth"] in (2, 4)
    assert 0 < bms
    assert gpu_ctx.tuning() == {"tilebins": 6, "longrow": 1024, "capscale": 50, "rowbins": 0, "tilewidth": 4}


def test_width_one_tiles_hold_rows_of_at_most_64(gpu_ctx):
    with pytest.raises(sa.SapError):
        gpu_ctx.set_tuning(6, 1024, 50, 0, 1)
    assert gpu_ctx.tuning()["tilebins"] == 6 and gpu_ctx.tuning()["tilewidth"] == 4
