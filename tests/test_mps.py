"""MPS reader (scip_amd/csrc/mps.c): the MIPLIB header counts of the instances SCIP's own tests
use, the semantics of SCIP's reader_mps.c on a hand-written file, error codes, and the oracle's
propagation on the real instances."""
import glob
import os
import re

import numpy as np
import pytest

import scip_amd as sa
from oracle import oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))
MPS = sorted(glob.glob(os.path.join(HERE, "golden", "mps", "*.mps")))
INF = 1e20


def header(path):
    return {k: int(v) for k, v in re.findall(r"^\*(ROWS|COLUMNS|INTEGER|NONZERO):\s+(\d+)", open(path).read(), re.M)}


@pytest.mark.parametrize("path", MPS, ids=[os.path.basename(p) for p in MPS])
def test_counts_match_miplib_header(path):
    inst, nsc = sa.read_mps(path)
    h = header(path)
    if not h:
        pytest.skip("no MIPLIB header")
    assert (inst.nrows, inst.ncols, int(inst.integral.sum()), inst.nnz) == \
        (h["ROWS"], h["COLUMNS"], h["INTEGER"], h["NONZERO"])
    # CSR well formed: ascending distinct columns, nonzero values
    for i in range(inst.nrows):
        c = inst.colidx[inst.rowptr[i]:inst.rowptr[i + 1]]
        assert np.all(np.diff(c) > 0)
    assert np.all(inst.vals != 0)


def binaries(inst):
    return int(((inst.integral != 0) & (inst.lb == 0) & (inst.ub == 1)).sum())


def test_binary_counts_stated_in_the_files():
    d = {os.path.basename(p)[:-4]: sa.read_mps(p)[0] for p in MPS}
    assert binaries(d["bell5"]) == 30                      # "30 of the integer variables are binary"
    assert binaries(d["misc03"]) == 159                    # "all integer variables are binary"
    assert binaries(d["egout"]) == 55
    assert binaries(d["p0548"]) == 548 == d["p0548"].ncols  # "pure 0/1 IP"
    inst, nsc = sa.read_mps([p for p in MPS if p.endswith("semicon1.mps")][0])
    assert nsc == 1 and inst.nrows == 4 and inst.ncols == 5


SEMANTICS = """NAME          TOY
ROWS
 N  obj
 N  free
 L  r1
 G  r2
 E  r3
 E  r4
 L  r5
COLUMNS
    x         obj       1.0        r1        1.0
    x         r2        2.0        r3        1.0
    x         r1        3.0        free      9.0
    MARKER    'MARKER'  'INTORG'
    y         r1        1.0        r2        1e-12
    y         r4        1.0
    z         r3        1.0        r5        1.0
    w         r5        2.0
    v         r5        1.0
    u         r4        1.0
    MARKER    'MARKER'  'INTEND'
    c         r4        1.0
RHS
    rhs       obj       5.0        r1        10.0
    rhs       r2        -3.0       r3        4.0
    rhs       r4        2.0        r5        7.0
    other     r1        99.0
RANGES
    rng       r1        4.0        r2        -2.0
    rng       r3        -1.5       r4        3.0
BOUNDS
 UP BND       y         5.5
 LO BND       z         2.3
 BV BND       w
 UP BND       v         1.0
 MI BND       x
 FR BND       c
 FX BND       u         1.0
 UP OTHER     v         0.0
ENDATA
"""


def test_reader_semantics(tmp_path):
    f = tmp_path / "toy.mps"
    f.write_text(SEMANTICS)
    inst, nsc = sa.read_mps(str(f))
    # rows r1..r5 (objective and the free row are not part of the system)
    assert inst.nrows == 5 and inst.ncols == 7
    col = {n: i for i, n in enumerate(["x", "y", "z", "w", "v", "u", "c"])}
    # r1: x appears twice (1 + 3 = 4), y 1; the 1e-12 entry of y in r2 is dropped
    r = {i: dict(zip(inst.colidx[inst.rowptr[i]:inst.rowptr[i + 1]], inst.vals[inst.rowptr[i]:inst.rowptr[i + 1]]))
         for i in range(5)}
    assert r[0] == {col["x"]: 4.0, col["y"]: 1.0}
    assert r[1] == {col["x"]: 2.0}
    # sides: L r1 rhs 10, range 4 -> [6, 10]; G r2 lhs -3, range -2 -> [-3, -1];
    # E r3 = 4, range -1.5 -> [2.5, 4]; E r4 = 2, range 3 -> [2, 5]; L r5 <= 7; 2nd RHS vector ignored
    assert list(zip(inst.lhs, inst.rhs)) == [(6.0, 10.0), (-3.0, -1.0), (2.5, 4.0), (2.0, 5.0), (-INF, 7.0)]
    # bounds: x continuous, MI -> (-inf, inf); y integer marker + UP 5.5 -> general integer [0, 5]
    # (rounded inward); z integer + LO 2.3 -> general integer [3, inf); w BV -> binary;
    # v UP 1 keeps it binary (the 2nd bound vector is ignored); u FX 1 -> [1, 1];
    # c continuous FR -> free
    exp = {"x": (-INF, INF, 0), "y": (0.0, 5.0, 1), "z": (3.0, INF, 1), "w": (0.0, 1.0, 1), "v": (0.0, 1.0, 1),
           "u": (1.0, 1.0, 1), "c": (-INF, INF, 0)}
    for n, (lo, hi, it) in exp.items():
        j = col[n]
        assert (inst.lb[j], inst.ub[j], int(inst.integral[j])) == (lo, hi, it), n
    assert nsc == 0


def test_reader_errors(tmp_path):
    with pytest.raises(sa.SapError) as e:
        sa.read_mps(str(tmp_path / "missing.mps"))
    assert e.value.rc == sa.SAP_NOFILE
    f = tmp_path / "bad.mps"
    f.write_text("NAME X\nROWS\n Q  r1\nENDATA\n")
    with pytest.raises(sa.SapError) as e:
        sa.read_mps(str(f))
    assert e.value.rc == sa.SAP_READERROR
    f.write_text("NAME X\nROWS\n N  obj\n L  r1\nCOLUMNS\n    x  r1  1.0\n")   # no ENDATA
    with pytest.raises(sa.SapError) as e:
        sa.read_mps(str(f))
    assert e.value.rc == sa.SAP_READERROR


def pure_binary(inst):
    return bool(np.all(inst.integral != 0) and np.all(inst.lb == 0) and np.all(inst.ub == 1))


@pytest.mark.parametrize("path", MPS, ids=[os.path.basename(p) for p in MPS])
def test_oracle_propagates_real_instances(path):
    inst, _ = sa.read_mps(path)
    om = orc.OMatrix.from_instance(inst)
    jl, ju, jr = orc.propagate_jacobi(om, inst.lb, inst.ub)
    sl, su, sr = orc.propagate_scip(om, inst.lb, inst.ub)
    # these instances are feasible: neither schedule may cut the root off
    assert not jr["cutoff"] and not sr["cutoff"]
    # propagation only tightens
    assert np.all(jl >= inst.lb) and np.all(ju <= inst.ub) and np.all(sl >= inst.lb) and np.all(su <= inst.ub)
    if pure_binary(inst):
        # unique binary fixpoint: both schedules agree exactly
        assert np.array_equal(jl, sl) and np.array_equal(ju, su)
