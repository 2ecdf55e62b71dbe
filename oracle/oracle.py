"""ctypes binding of the oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REFLIB = os.path.join(HERE, "_ref", "libdbldbl_ref.so")


class orc_params(C.Structure):
    _fields_ = [
        ("infinity", C.c_double), ("epsilon", C.c_double), ("sumepsilon", C.c_double), ("feastol", C.c_double),
        ("boundstreps", C.c_double), ("hugeval", C.c_double), ("recompfac", C.c_double),
        ("maxeasyactivitydelta", C.c_double), ("maxrounds", C.c_int32), ("rangedrows", C.c_int32),
        ("notighten", C.c_int32), ("pad", C.c_int32),
    ]


class orc_matrix(C.Structure):
    _fields_ = [
        ("nrows", C.c_int64), ("ncols", C.c_int64), ("rowptr", C.POINTER(C.c_int64)),
        ("colidx", C.POINTER(C.c_int32)), ("vals", C.POINTER(C.c_double)), ("lhs", C.POINTER(C.c_double)),
        ("rhs", C.POINTER(C.c_double)), ("integral", C.POINTER(C.c_uint8)), ("binary", C.POINTER(C.c_uint8)),
    ]


class orc_conflict(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("kind", "round", "row", "side", "col", "rowlb", "rowub", "rules")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class orc_result(C.Structure):
    _fields_ = [("cutoff", C.c_int32), ("rounds", C.c_int32), ("nchgbds", C.c_int64), ("rows_swept", C.c_int64),
                ("nnz_swept", C.c_int64), ("nrecompute", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class orc_rowact(C.Structure):
    _fields_ = [("minact", C.c_double), ("maxact", C.c_double), ("minact_hi", C.c_double), ("minact_lo", C.c_double),
                ("maxact_hi", C.c_double), ("maxact_lo", C.c_double)] + \
               [(n, C.c_int32) for n in ("minposinf", "minneginf", "minposhuge", "minneghuge", "maxposinf",
                                         "maxneginf", "maxposhuge", "maxneghuge")] + [("maxactdelta", C.c_double)]


_DP = C.POINTER(C.c_double)
_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(LIB)
        L.orc_params_default.argtypes = [C.POINTER(orc_params)]
        L.orc_propagate_jacobi.argtypes = [C.POINTER(orc_matrix), C.POINTER(orc_params), _DP, _DP, C.c_int,
                                           C.POINTER(orc_result)]
        L.orc_propagate_scip.argtypes = [C.POINTER(orc_matrix), C.POINTER(orc_params), _DP, _DP, C.POINTER(orc_result)]
        L.orc_propagate_jacobi_marked.argtypes = [C.POINTER(orc_matrix), C.POINTER(orc_params), _DP, _DP,
                                                  C.POINTER(C.c_uint8), C.POINTER(orc_result)]
        L.orc_row_activities.argtypes = [C.POINTER(orc_matrix), C.POINTER(orc_params), _DP, _DP, C.c_int,
                                         C.POINTER(orc_rowact)]
        L.orc_compute_binary.argtypes = [C.c_int64, C.POINTER(C.c_uint8), _DP, _DP, C.POINTER(C.c_uint8)]
        for n in ("orc_dd_sum", "orc_dd_prod"):
            getattr(L, n).argtypes = [C.c_double, C.c_double, _DP, _DP]
        for n in ("orc_dd_sum21", "orc_dd_prod21"):
            getattr(L, n).argtypes = [C.c_double, C.c_double, C.c_double, _DP, _DP]
        L.orc_dd_pairadd.argtypes = [C.c_double] * 4 + [_DP, _DP]
        L.orc_canonical_sum.argtypes = [_DP, C.c_int64, _DP, _DP]
        _I32P = C.POINTER(C.c_int32)
        L.orc_propagate_jacobi_reasons.argtypes = [C.POINTER(orc_matrix), C.POINTER(orc_params), _DP, _DP, C.c_int64,
                                                   C.POINTER(C.c_int64), _I32P, _I32P, _I32P, _DP, _I32P, _I32P,
                                                   C.POINTER(orc_conflict), C.POINTER(orc_result)]
        _lib = L
    return _lib


def reflib():
    """The reference's own dbldblarith.h macros (None when the reference is absent)."""
    global _ref
    if _ref is None and os.path.exists(REFLIB):
        R = C.CDLL(REFLIB)
        for n in ("ref_dd_sum", "ref_dd_prod"):
            getattr(R, n).argtypes = [C.c_double, C.c_double, _DP, _DP]
        for n in ("ref_dd_sum21", "ref_dd_prod21"):
            getattr(R, n).argtypes = [C.c_double, C.c_double, C.c_double, _DP, _DP]
        R.ref_dd_sum22.argtypes = [C.c_double] * 4 + [_DP, _DP]
        R.ref_activity_sum.argtypes = [_DP, C.c_long, _DP, _DP]
        _ref = R
    return _ref


def params(**kw) -> orc_params:
    p = orc_params()
    lib().orc_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class OMatrix:
    """Keeps the numpy arrays alive for the orc_matrix view."""

    def __init__(self, rowptr, colidx, vals, lhs, rhs, integral, glb, gub):
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
        self.colidx = np.ascontiguousarray(colidx, dtype=np.int32)
        self.vals = np.ascontiguousarray(vals, dtype=np.float64)
        self.lhs = np.ascontiguousarray(lhs, dtype=np.float64)
        self.rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        self.integral = np.ascontiguousarray(integral, dtype=np.uint8)
        self.glb = np.ascontiguousarray(glb, dtype=np.float64)
        self.gub = np.ascontiguousarray(gub, dtype=np.float64)
        self.binary = np.zeros(len(self.glb), dtype=np.uint8)
        lib().orc_compute_binary(len(self.glb), self.integral.ctypes.data_as(C.POINTER(C.c_uint8)),
                                 self.glb.ctypes.data_as(_DP), self.gub.ctypes.data_as(_DP),
                                 self.binary.ctypes.data_as(C.POINTER(C.c_uint8)))
        m = orc_matrix()
        m.nrows = len(self.lhs)
        m.ncols = len(self.glb)
        m.rowptr = self.rowptr.ctypes.data_as(C.POINTER(C.c_int64))
        m.colidx = self.colidx.ctypes.data_as(C.POINTER(C.c_int32))
        m.vals = self.vals.ctypes.data_as(_DP)
        m.lhs = self.lhs.ctypes.data_as(_DP)
        m.rhs = self.rhs.ctypes.data_as(_DP)
        m.integral = self.integral.ctypes.data_as(C.POINTER(C.c_uint8))
        m.binary = self.binary.ctypes.data_as(C.POINTER(C.c_uint8))
        self.m = m

    @classmethod
    def from_instance(cls, inst):
        return cls(inst.rowptr, inst.colidx, inst.vals, inst.lhs, inst.rhs, inst.integral, inst.lb, inst.ub)


def propagate_jacobi(om: OMatrix, lb, ub, prm: orc_params | None = None, full_sweep: bool = False):
    lb = np.array(lb, dtype=np.float64, copy=True)
    ub = np.array(ub, dtype=np.float64, copy=True)
    r = orc_result()
    lib().orc_propagate_jacobi(C.byref(om.m), C.byref(prm or params()), lb.ctypes.data_as(_DP),
                               ub.ctypes.data_as(_DP), 1 if full_sweep else 0, C.byref(r))
    return lb, ub, r.as_dict()


def propagate_jacobi_marked(om: OMatrix, lb, ub, mark0, prm: orc_params | None = None):
    """The Jacobi twin started with only the rows flagged in mark0 marked (the device's
    SAP_FLAG_BASE_IS_FIXPOINT start)."""
    lb = np.array(lb, dtype=np.float64, copy=True)
    ub = np.array(ub, dtype=np.float64, copy=True)
    mk = np.ascontiguousarray(mark0, dtype=np.uint8)
    assert len(mk) == om.m.nrows
    r = orc_result()
    lib().orc_propagate_jacobi_marked(C.byref(om.m), C.byref(prm or params()), lb.ctypes.data_as(_DP),
                                      ub.ctypes.data_as(_DP), mk.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(r))
    return lb, ub, r.as_dict()


def rows_of_columns(om: OMatrix, cols):
    """nrows flags: the rows that contain one of the given columns."""
    hit = np.zeros(int(om.m.ncols), dtype=bool)
    hit[np.asarray(cols, dtype=np.int64)] = True
    nz = hit[om.colidx]
    rows = getattr(om, "_nzrow", None)
    if rows is None:
        rows = om._nzrow = np.repeat(np.arange(int(om.m.nrows), dtype=np.int32), np.diff(om.rowptr))
    out = np.zeros(int(om.m.nrows), dtype=np.uint8)
    out[rows[nz]] = 1
    return out


def propagate_reasons(om: OMatrix, lb, ub, prm: orc_params | None = None):
    """The Jacobi twin with reasons: bounds, result, history (round, col, bound, val, row) and
    the conflict of a cutoff -- the same record the engine's sap_propagate_reasons returns."""
    lb = np.array(lb, dtype=np.float64, copy=True)
    ub = np.array(ub, dtype=np.float64, copy=True)
    I32 = C.POINTER(C.c_int32)
    cap = max(1024, 2 * int(om.m.ncols))
    while True:
        l, u = lb.copy(), ub.copy()
        hr, hc, hb, hw, hu = (np.zeros(cap, dtype=np.int32) for _ in range(5))
        hv = np.zeros(cap)
        n = C.c_int64(0)
        cf = orc_conflict()
        r = orc_result()
        lib().orc_propagate_jacobi_reasons(C.byref(om.m), C.byref(prm or params()), l.ctypes.data_as(_DP),
                                           u.ctypes.data_as(_DP), cap, C.byref(n), hr.ctypes.data_as(I32),
                                           hc.ctypes.data_as(I32), hb.ctypes.data_as(I32), hv.ctypes.data_as(_DP),
                                           hw.ctypes.data_as(I32), hu.ctypes.data_as(I32), C.byref(cf), C.byref(r))
        if n.value > cap:
            cap = n.value
            continue
        k = n.value
        return dict(lb=l, ub=u, result=r.as_dict(), conflict=cf.as_dict(), nhist=k,
                    round=hr[:k], col=hc[:k], bound=hb[:k], val=hv[:k], row=hw[:k], rule=hu[:k])


def propagate_scip(om: OMatrix, lb, ub, prm: orc_params | None = None):
    lb = np.array(lb, dtype=np.float64, copy=True)
    ub = np.array(ub, dtype=np.float64, copy=True)
    r = orc_result()
    lib().orc_propagate_scip(C.byref(om.m), C.byref(prm or params()), lb.ctypes.data_as(_DP),
                             ub.ctypes.data_as(_DP), C.byref(r))
    return lb, ub, r.as_dict()


def row_activities(om: OMatrix, lb, ub, canonical: bool = True, prm: orc_params | None = None):
    lb = np.ascontiguousarray(lb, dtype=np.float64)
    ub = np.ascontiguousarray(ub, dtype=np.float64)
    nr = om.m.nrows
    out = (orc_rowact * max(nr, 1))()
    lib().orc_row_activities(C.byref(om.m), C.byref(prm or params()), lb.ctypes.data_as(_DP),
                             ub.ctypes.data_as(_DP), 1 if canonical else 0, out)
    mn = np.array([out[i].minact for i in range(nr)])
    mx = np.array([out[i].maxact for i in range(nr)])
    cnt = np.array([[out[i].minposinf, out[i].minneginf, out[i].minposhuge, out[i].minneghuge, out[i].maxposinf,
                     out[i].maxneginf, out[i].maxposhuge, out[i].maxneghuge] for i in range(nr)], dtype=np.int32)
    mad = np.array([out[i].maxactdelta for i in range(nr)])
    return mn, mx, cnt.reshape(nr, 8), mad


def dd2(fn, *args):
    hi, lo = C.c_double(), C.c_double()
    fn(*args, C.byref(hi), C.byref(lo))
    return hi.value, lo.value


def canonical_sum(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    hi, lo = C.c_double(), C.c_double()
    lib().orc_canonical_sum(x.ctypes.data_as(_DP), len(x), C.byref(hi), C.byref(lo))
    return hi.value, lo.value


def probe(om: OMatrix, lb, ub, probecols, prm: orc_params | None = None, downub=None, uplb=None):
    """Probing on binary columns, restated from prop_probing.c: each probe column x is fixed
    to 0 and to 1 (SCIPapplyProbingVar, prop_probing.c:1203-1311, here with linear
    propagation as the only propagator: the Jacobi twin), and the two branches are compared
    as SCIPanalyzeDeductionsProbing (prop_probing.c:1314-1500) compares them.  With
    downub/uplb the branches are x <= downub[q] and x >= uplb[q] (the children of strong
    branching with propagation, scip_var.c:4139-4440).

    Returns per probe: verdict bits (1: x = 0 infeasible, 2: x = 1 infeasible), the rounds of
    both branches, and -- when neither branch is infeasible -- the columns other than x whose
    bounds differ from (lb, ub) in both branches, ascending, with their bounds in each branch.
    """
    lb = np.asarray(lb, dtype=np.float64)
    ub = np.asarray(ub, dtype=np.float64)
    split = downub is not None
    out = []
    for q, x in enumerate(probecols):
        branch = []
        for side in (0, 1):
            l0 = lb.copy()
            u0 = ub.copy()
            if split and side == 0:
                u0[x] = downub[q]
            elif split:
                l0[x] = uplb[q]
            else:
                l0[x] = u0[x] = float(side)
            pl, pu, r = propagate_jacobi(om, l0, u0, prm)
            branch.append((pl, pu, r))
        verdict = (1 if branch[0][2]["cutoff"] else 0) | (2 if branch[1][2]["cutoff"] else 0)
        rec = dict(verdict=verdict, rounds=(branch[0][2]["rounds"], branch[1][2]["rounds"]),
                   cols=np.zeros(0, dtype=np.int32), b0=np.zeros((0, 2)), b1=np.zeros((0, 2)))
        if verdict == 0:
            (l0, u0, _), (l1, u1, _) = branch
            ch0 = (l0 != lb) | (u0 != ub)
            ch1 = (l1 != lb) | (u1 != ub)
            # binary probing: the probe column is fixed in both branches and not listed; split
            # probing (strong branching): it is always listed, as SCIP's analysis of a
            # non-binary probing variable includes it
            ch0[x] = ch1[x] = split
            cols = np.nonzero(ch0 & ch1)[0].astype(np.int32)
            rec.update(cols=cols, b0=np.stack([l0[cols], u0[cols]], 1), b1=np.stack([l1[cols], u1[cols]], 1))
        out.append(rec)
    return out
