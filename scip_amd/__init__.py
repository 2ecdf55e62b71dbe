"""scip_amd -- MI355X linear-constraint domain propagation for SCIP.

Thin ctypes binding of the C ABI declared in include/scip_amd_prop.h and
include/scip_amd_gen.h (library: scip_amd/libscipamd.so, built by
scip_amd/build.py).  This module is a convenience for tests and benchmarks; the
product boundary is the C ABI, which the SCIP propagator plugin
(scip_amd/plugin/prop_gpulinear.c) calls directly.

There is no fallback: if the HIP library is missing, importing the binding
raises, and every propagation call goes through the device engine.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libscipamd.so")

SAP_OKAY = 1
SAP_ERROR = 0
SAP_NOMEMORY = -1
SAP_INVALIDCALL = -8
SAP_INVALIDDATA = -9
SAP_NOTIMPLEMENTED = -18
SAP_READERROR = -2
SAP_NOFILE = -4

SAP_DIDNOTFIND = 3
SAP_CUTOFF = 7
SAP_REDUCEDDOM = 10

SAP_FLAG_BASE_IS_FIXPOINT = 0x1
SAP_FLAG_HOSTLOOP = 0x2
SAP_FLAG_RANGEDROWS = 0x4
SAP_FLAG_NOTIGHTEN = 0x8

SAP_STEP_SHORT = 2
SAP_STEP_LONG = 8

SAP_CONFLICT_NONE = 0
SAP_CONFLICT_ROW = 1
SAP_CONFLICT_COLUMN = 2

SAP_PROBE_DOWN_INFEASIBLE = 1
SAP_PROBE_UP_INFEASIBLE = 2

SAP_MODE_AUTO = 0
SAP_MODE_DENSE = 1
SAP_MODE_SPARSE = 2

INFINITY = 1e20


class SapError(RuntimeError):
    def __init__(self, rc: int, what: str):
        self.rc = rc
        super().__init__(f"{what} failed with code {rc}")


class sap_params(C.Structure):
    _fields_ = [
        ("infinity", C.c_double),
        ("epsilon", C.c_double),
        ("sumepsilon", C.c_double),
        ("feastol", C.c_double),
        ("boundstreps", C.c_double),
        ("hugeval", C.c_double),
        ("recompfac", C.c_double),
        ("maxeasyactivitydelta", C.c_double),
        ("maxrounds", C.c_int32),
        ("flags", C.c_int32),
    ]


class sap_tuning(C.Structure):
    _fields_ = [("tilebins", C.c_int32), ("longrow", C.c_int32), ("capscale", C.c_int32), ("rowbins", C.c_int32),
                ("tilewidth", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class sap_conflict(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("kind", "round", "row", "side", "col", "rowlb", "rowub", "rules")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class sap_stats(C.Structure):
    _fields_ = [
        ("result", C.c_int32),
        ("rounds", C.c_int32),
        ("nchgbds", C.c_int64),
        ("rows_swept", C.c_int64),
        ("nnz_swept", C.c_int64),
        ("device_ms", C.c_double),
        ("sweep_ms", C.c_double),
        ("sweep_launches", C.c_int64),
        ("nrecompute", C.c_int64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class sap_instance(C.Structure):
    _fields_ = [
        ("nrows", C.c_int64),
        ("ncols", C.c_int64),
        ("nnz", C.c_int64),
        ("rowptr", C.POINTER(C.c_int64)),
        ("colidx", C.POINTER(C.c_int32)),
        ("vals", C.POINTER(C.c_double)),
        ("lhs", C.POINTER(C.c_double)),
        ("rhs", C.POINTER(C.c_double)),
        ("integral", C.POINTER(C.c_uint8)),
        ("lb", C.POINTER(C.c_double)),
        ("ub", C.POINTER(C.c_double)),
    ]


class sap_serve_stats(C.Structure):
    _fields_ = [
        ("wall_s", C.c_double),
        ("requests", C.c_int64),
        ("nnz_swept", C.c_int64),
        ("rows_swept", C.c_int64),
        ("steps", C.c_int64),
        ("admissions", C.c_int64),
        ("uploads", C.c_int64),
        ("evictions", C.c_int64),
        ("max_running", C.c_int64),
        ("peak_bytes", C.c_int64),
        ("p50_ms", C.c_double),
        ("p99_ms", C.c_double),
        ("max_ms", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


SAP_SERVE_CONTINUOUS = 0
SAP_SERVE_PLAIN = 1

_P = C.c_void_p
_I64P = C.POINTER(C.c_int64)
_I32P = C.POINTER(C.c_int32)
_U8P = C.POINTER(C.c_uint8)
_DP = C.POINTER(C.c_double)

# name -> (restype, argtypes); every symbol of include/scip_amd_prop.h and scip_amd_gen.h
EXPORTS = {
    "sap_params_default": (None, [C.POINTER(sap_params)]),
    "sap_error_string": (C.c_char_p, [C.c_int]),
    "sap_device_count": (C.c_int, [_I32P]),
    "sap_context_create": (C.c_int, [C.c_int32, C.POINTER(_P)]),
    "sap_context_free": (C.c_int, [_P]),
    "sap_matrix_create": (C.c_int, [_P, C.c_int64, C.c_int64, _I64P, _I32P, _DP, _DP, _DP, _U8P, _DP, _DP,
                                    C.POINTER(_P)]),
    "sap_matrix_free": (C.c_int, [_P]),
    "sap_tuning_default": (None, [C.POINTER(sap_tuning)]),
    "sap_context_set_tuning": (C.c_int, [_P, C.POINTER(sap_tuning)]),
    "sap_context_get_tuning": (C.c_int, [_P, C.POINTER(sap_tuning)]),
    "sap_autotune": (C.c_int, [_P, C.c_int64, C.c_int64, _I64P, _I32P, _DP, _DP, _DP, _U8P, _DP, _DP,
                               C.POINTER(sap_params), C.c_int64, C.c_uint64, C.POINTER(sap_tuning), _DP, _DP]),
    "sap_matrix_info": (C.c_int, [_P, _I64P, _I64P, _I64P, _I64P]),
    "sap_propagate": (C.c_int, [_P, C.POINTER(sap_params), _DP, _DP, C.POINTER(sap_stats)]),
    "sap_bounds_create": (C.c_int, [_P, C.c_int32, C.POINTER(_P)]),
    "sap_bounds_create_mode": (C.c_int, [_P, C.c_int32, C.c_int32, C.POINTER(_P)]),
    "sap_bounds_free": (C.c_int, [_P]),
    "sap_bounds_upload": (C.c_int, [_P, C.c_int32, _DP, _DP]),
    "sap_bounds_download": (C.c_int, [_P, C.c_int32, _DP, _DP]),
    "sap_bounds_copy": (C.c_int, [_P, C.c_int32, _P, C.c_int32]),
    "sap_propagate_resident": (C.c_int, [_P, C.POINTER(sap_params), C.POINTER(sap_stats)]),
    "sap_propagate_batch": (C.c_int, [_P, C.POINTER(sap_params), C.c_int32, _DP, _DP, _I64P, _I32P, _DP, _DP,
                                      _I32P, _I32P, _DP, _DP, C.POINTER(sap_stats)]),
    "sap_propagate_nodes": (C.c_int, [_P, C.POINTER(sap_params), C.c_int32, _P, C.c_int32, _I64P, _I32P, _DP, _DP,
                                      _I32P, _I32P, C.c_int64, _I64P, _I32P, _DP, _DP, C.POINTER(sap_stats)]),
    "sap_propagate_reasons": (C.c_int, [_P, C.POINTER(sap_params), _DP, _DP, C.c_int64, _I64P, _I32P, _I32P, _I32P,
                                        _DP, _I32P, _I32P, C.POINTER(sap_conflict), C.POINTER(sap_stats)]),
    "sap_probe": (C.c_int, [_P, C.POINTER(sap_params), C.c_int32, _I32P, _P, C.c_int32, _I32P, _I32P, C.c_int64,
                            _I64P, _I32P, _DP, _DP, _DP, _DP, C.POINTER(sap_stats)]),
    "sap_probe_split": (C.c_int, [_P, C.POINTER(sap_params), C.c_int32, _I32P, _DP, _DP, _P, C.c_int32, _I32P, _I32P,
                                  C.c_int64, _I64P, _I32P, _DP, _DP, _DP, _DP, C.POINTER(sap_stats)]),
    "sap_matrix_create_packed": (C.c_int, [_P, C.c_int32, _I64P, _I64P, C.c_int64, C.c_int64, _I64P, _I32P, _DP,
                                           _DP, _DP, _U8P, _DP, _DP, C.POINTER(_P)]),
    "sap_propagate_packed": (C.c_int, [_P, C.POINTER(sap_params), _DP, _DP, _I32P, _I32P, C.POINTER(sap_stats)]),
    "sap_slots_admit": (C.c_int, [_P, C.c_int32, _I32P, _P, C.c_int32, _I64P, _I32P, _DP, _DP, C.c_int32]),
    "sap_slots_launch": (C.c_int, [_P, C.POINTER(sap_params), C.c_int32]),
    "sap_slots_query": (C.c_int, [_P, _I32P]),
    "sap_slots_wait": (C.c_int, [_P, _I32P, _I32P, C.POINTER(sap_stats)]),
    "sap_slots_step": (C.c_int, [_P, C.POINTER(sap_params), C.c_int32, _I32P, _I32P, C.POINTER(sap_stats)]),
    "sap_slots_collect": (C.c_int, [_P, C.c_int32, _I32P, _I32P, _I32P, C.c_int64, _I64P, _I32P, _DP, _DP]),
    "sap_row_activities": (C.c_int, [_P, C.POINTER(sap_params), _DP, _DP, _DP, _DP, _I32P, _DP]),
    "sap_server_create": (C.c_int, [C.c_int32, C.c_int32, C.c_int64, C.c_int32, C.POINTER(_P)]),
    "sap_server_free": (C.c_int, [_P]),
    "sap_server_add_problem": (C.c_int, [_P, C.c_int64, C.c_int64, _I64P, _I32P, _DP, _DP, _DP, _U8P, _DP, _DP,
                                         _I32P]),
    "sap_server_warm": (C.c_int, [_P, C.c_int32]),
    "sap_server_footprint": (C.c_int, [_P, C.c_int32, _I64P]),
    "sap_bench_gather": (C.c_int, [_P, C.c_int32, C.c_int32, _DP]),
    "sap_matrix_set_context": (C.c_int, [_P, _P]),
    "sap_propagate_device_noalloc": (C.c_int, [_P, C.POINTER(sap_params), C.c_void_p, C.c_void_p,
                                               C.POINTER(sap_stats)]),
    "sap_server_set_budget": (C.c_int, [_P, C.c_int64]),
    "sap_server_cold": (C.c_int, [_P, C.c_int32, _U8P]),
    "sap_server_set_slots": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "sap_server_replay": (C.c_int, [_P, C.POINTER(sap_params), C.c_int32, C.c_int32, C.c_int32, _DP, _I32P, _I64P,
                                    _I32P, _DP, _DP, _I32P, _I32P, _I32P, _DP, C.c_int64, _I64P, _I32P, _DP, _DP,
                                    C.POINTER(sap_serve_stats)]),
    "sap_gen_covering_packing": (C.c_int, [C.c_int64, C.c_int64, C.c_int32, C.c_uint64, C.POINTER(sap_instance)]),
    "sap_gen_mixed": (C.c_int, [C.c_int64, C.c_int64, C.c_int32, C.c_double, C.c_double, C.c_int32, C.c_int32,
                                C.c_int32, C.c_uint64, C.POINTER(sap_instance)]),
    "sap_gen_node_fixings": (C.c_int, [C.POINTER(sap_instance), C.c_int32, C.c_int32, C.c_int32, C.c_uint64,
                                       _I64P, _I32P, _DP]),
    "sap_instance_free": (None, [C.POINTER(sap_instance)]),
    "sap_read_mps": (C.c_int, [C.c_char_p, C.POINTER(sap_instance), _I32P]),
}

_lib = None

# live handles, closed children-first when the interpreter exits (a context outlives its
# matrices, a matrix its bound sets; finalizers alone run in no particular order)
_LIVE = {k: weakref.WeakSet() for k in ("server", "bounds", "matrix", "context")}


@atexit.register
def _close_live_handles():
    for kind in ("server", "bounds", "matrix", "context"):
        for obj in list(_LIVE[kind]):
            try:
                obj.close()
            except Exception:
                pass


def lib() -> C.CDLL:
    """Load libscipamd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise ImportError(f"{LIBPATH} is missing: run scip_amd/build.py (or __graft_entry__.build())")
        L = C.CDLL(LIBPATH)
        for name, (res, args) in EXPORTS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _chk(rc: int, what: str):
    if rc != SAP_OKAY:
        raise SapError(rc, what)


def _dp(a):
    return a.ctypes.data_as(_DP) if a is not None else None


def _i64p(a):
    return a.ctypes.data_as(_I64P) if a is not None else None


def _i32p(a):
    return a.ctypes.data_as(_I32P) if a is not None else None


def _u8p(a):
    return a.ctypes.data_as(_U8P) if a is not None else None


def default_params(**kw) -> sap_params:
    p = sap_params()
    lib().sap_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def device_count() -> int:
    n = C.c_int32(0)
    lib().sap_device_count(C.byref(n))
    return n.value


# ------------------------------------------------------------------------------------------
# instances
# ------------------------------------------------------------------------------------------
@dataclass
class Instance:
    """Linear system lhs <= A x <= rhs (CSR) with column integrality and global bounds."""
    rowptr: np.ndarray
    colidx: np.ndarray
    vals: np.ndarray
    lhs: np.ndarray
    rhs: np.ndarray
    integral: np.ndarray
    lb: np.ndarray
    ub: np.ndarray

    @property
    def nrows(self):
        return len(self.lhs)

    @property
    def ncols(self):
        return len(self.lb)

    @property
    def nnz(self):
        return int(self.rowptr[-1])


def _from_c(inst: sap_instance) -> Instance:
    nr, nc, nnz = inst.nrows, inst.ncols, inst.nnz

    def arr(ptr, n, dt):
        if n == 0:
            return np.zeros(0, dtype=dt)
        return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt, copy=True)

    out = Instance(
        rowptr=arr(inst.rowptr, nr + 1, np.int64),
        colidx=arr(inst.colidx, nnz, np.int32),
        vals=arr(inst.vals, nnz, np.float64),
        lhs=arr(inst.lhs, nr, np.float64),
        rhs=arr(inst.rhs, nr, np.float64),
        integral=arr(inst.integral, nc, np.uint8),
        lb=arr(inst.lb, nc, np.float64),
        ub=arr(inst.ub, nc, np.float64),
    )
    lib().sap_instance_free(C.byref(inst))
    return out


def gen_covering_packing(nrows: int, ncols: int, nnz_per_row: int = 10, seed: int = 1) -> Instance:
    inst = sap_instance()
    _chk(lib().sap_gen_covering_packing(nrows, ncols, nnz_per_row, seed, C.byref(inst)), "sap_gen_covering_packing")
    return _from_c(inst)


def gen_mixed(nrows: int, ncols: int, seed: int = 2, lenmode: int = 0, median: float = 8.0, sigma: float = 1.0,
              minlen: int = 1, maxlen: int = 10000, planted: bool = False) -> Instance:
    """planted=False: sides centred on the activity span (the stated workload); planted=True:
    sides centred on the activity of a hidden feasible point."""
    inst = sap_instance()
    _chk(lib().sap_gen_mixed(nrows, ncols, lenmode, median, sigma, minlen, maxlen, 1 if planted else 0, seed,
                             C.byref(inst)), "sap_gen_mixed")
    return _from_c(inst)


def read_mps(path: str):
    """Read an MPS file (include/scip_amd_mps.h); returns (Instance, number of semi-continuous columns)."""
    inst = sap_instance()
    nsc = C.c_int32(0)
    _chk(lib().sap_read_mps(os.fsencode(path), C.byref(inst), C.byref(nsc)), f"sap_read_mps({path})")
    return _from_c(inst), nsc.value


def gen_node_fixings(inst: Instance, nnodes: int, minfix: int = 1, maxfix: int = 50, seed: int = 3):
    """Returns (ovptr, ovcol, ovval) with lb = ub = ovval for each fixed column."""
    ci = sap_instance()
    ci.nrows, ci.ncols, ci.nnz = inst.nrows, inst.ncols, inst.nnz
    ci.integral = _u8p(inst.integral)
    ptr = np.zeros(nnodes + 1, dtype=np.int64)
    col = np.zeros(max(nnodes * maxfix, 1), dtype=np.int32)
    val = np.zeros(max(nnodes * maxfix, 1), dtype=np.float64)
    _chk(lib().sap_gen_node_fixings(C.byref(ci), nnodes, minfix, maxfix, seed, _i64p(ptr), _i32p(col), _dp(val)),
         "sap_gen_node_fixings")
    n = int(ptr[-1])
    return ptr, col[:n].copy(), val[:n].copy()


# ------------------------------------------------------------------------------------------
# engine
# ------------------------------------------------------------------------------------------
class Context:
    def __init__(self, device: int = 0):
        self.h = _P()
        _chk(lib().sap_context_create(device, C.byref(self.h)), "sap_context_create")
        _LIVE["context"].add(self)

    def set_tuning(self, tilebins: int = 6, longrow: int = 1024, capscale: int = 50, rowbins: int = 0,
                   tilewidth: int = 4):
        t = sap_tuning(tilebins, longrow, capscale, rowbins, tilewidth)
        _chk(lib().sap_context_set_tuning(self.h, C.byref(t)), "sap_context_set_tuning")

    def tuning(self):
        t = sap_tuning()
        _chk(lib().sap_context_get_tuning(self.h, C.byref(t)), "sap_context_get_tuning")
        return t.as_dict()

    def autotune(self, inst: "Instance", lb=None, ub=None, params: sap_params | None = None,
                 samplerows: int = 200_000, seed: int = 1, apply: bool = True):
        """Measure the candidate tunings on a row sample of inst; returns (best, best_ms, default_ms)."""
        rp, ci, va, lh, rh, it = [np.ascontiguousarray(a) for a in (inst.rowptr, inst.colidx, inst.vals, inst.lhs,
                                                                     inst.rhs, inst.integral)]
        lb = np.ascontiguousarray(inst.lb if lb is None else lb, dtype=np.float64)
        ub = np.ascontiguousarray(inst.ub if ub is None else ub, dtype=np.float64)
        best = sap_tuning()
        bms, dms = C.c_double(0), C.c_double(0)
        p = params or default_params()
        _chk(lib().sap_autotune(self.h, inst.nrows, inst.ncols, _i64p(rp), _i32p(ci), _dp(va), _dp(lh), _dp(rh),
                                _u8p(it), _dp(lb), _dp(ub), C.byref(p), samplerows, seed, C.byref(best),
                                C.byref(bms), C.byref(dms)), "sap_autotune")
        if apply:
            _chk(lib().sap_context_set_tuning(self.h, C.byref(best)), "sap_context_set_tuning")
        return best.as_dict(), bms.value, dms.value

    def close(self):
        if self.h:
            lib().sap_context_free(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Matrix:
    """Device-resident CSR matrix (sap_matrix)."""

    def __init__(self, ctx: Context, inst: Instance):
        self.ctx = ctx
        self.inst = inst
        self.h = _P()
        self._keep = [np.ascontiguousarray(a) for a in (inst.rowptr, inst.colidx, inst.vals, inst.lhs, inst.rhs,
                                                         inst.integral, inst.lb, inst.ub)]
        rp, ci, va, lh, rh, it, lb, ub = self._keep
        _chk(lib().sap_matrix_create(ctx.h, inst.nrows, inst.ncols, _i64p(rp), _i32p(ci), _dp(va), _dp(lh), _dp(rh),
                                     _u8p(it), _dp(lb), _dp(ub), C.byref(self.h)), "sap_matrix_create")
        _LIVE["matrix"].add(self)

    def info(self):
        a, b, c, d = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        _chk(lib().sap_matrix_info(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)), "sap_matrix_info")
        return dict(nrows=a.value, ncols=b.value, nnz=c.value, device_bytes=d.value)

    def propagate(self, lb=None, ub=None, params: sap_params | None = None):
        """Propagate one bound vector; returns (lb, ub, stats dict)."""
        lb = np.array(self.inst.lb if lb is None else lb, dtype=np.float64, copy=True)
        ub = np.array(self.inst.ub if ub is None else ub, dtype=np.float64, copy=True)
        st = sap_stats()
        p = params or default_params()
        _chk(lib().sap_propagate(self.h, C.byref(p), _dp(lb), _dp(ub), C.byref(st)), "sap_propagate")
        return lb, ub, st.as_dict()

    def propagate_device_noalloc(self, dlb: int, dub: int, params: sap_params | None = None):
        """sap_propagate_device_noalloc on caller-owned device arrays (addresses dlb, dub of
        ncols doubles each, overwritten with the fixpoint unless cut off); returns the stats."""
        st = sap_stats()
        p = params or default_params()
        _chk(lib().sap_propagate_device_noalloc(self.h, C.byref(p), C.c_void_p(dlb), C.c_void_p(dub), C.byref(st)),
             "sap_propagate_device_noalloc")
        return st.as_dict()

    def propagate_reasons(self, lb=None, ub=None, params: sap_params | None = None, histcap: int | None = None):
        """Propagate one bound vector with reasons: bounds, stats, the history of accepted
        changes (round, col, bound, val, row arrays) and the conflict of a cutoff."""
        lb0 = np.array(self.inst.lb if lb is None else lb, dtype=np.float64, copy=True)
        ub0 = np.array(self.inst.ub if ub is None else ub, dtype=np.float64, copy=True)
        cap = int(histcap if histcap is not None else max(1024, 2 * self.inst.ncols))
        p = params or default_params()
        while True:
            lb, ub = lb0.copy(), ub0.copy()
            n = C.c_int64(0)
            hr, hc, hb, hw, hu = (np.zeros(max(cap, 1), dtype=np.int32) for _ in range(5))
            hv = np.zeros(max(cap, 1))
            cf = sap_conflict()
            st = sap_stats()
            rc = lib().sap_propagate_reasons(self.h, C.byref(p), _dp(lb), _dp(ub), cap, C.byref(n), _i32p(hr),
                                             _i32p(hc), _i32p(hb), _dp(hv), _i32p(hw), _i32p(hu), C.byref(cf),
                                             C.byref(st))
            if rc == SAP_NOMEMORY and histcap is None and n.value > cap:
                cap = n.value   # again with room for the whole history (same results)
                continue
            _chk(rc, "sap_propagate_reasons")
            k = n.value
            return dict(lb=lb, ub=ub, stats=st.as_dict(), conflict=cf.as_dict(), nhist=k,
                        round=hr[:k], col=hc[:k], bound=hb[:k], val=hv[:k], row=hw[:k], rule=hu[:k])

    def row_activities(self, lb=None, ub=None, params: sap_params | None = None):
        nr = self.inst.nrows
        lb = np.ascontiguousarray(self.inst.lb if lb is None else lb, dtype=np.float64)
        ub = np.ascontiguousarray(self.inst.ub if ub is None else ub, dtype=np.float64)
        mn, mx, mad = np.zeros(nr), np.zeros(nr), np.zeros(nr)
        cnt = np.zeros(8 * nr, dtype=np.int32)
        p = params or default_params()
        _chk(lib().sap_row_activities(self.h, C.byref(p), _dp(lb), _dp(ub), _dp(mn), _dp(mx), _i32p(cnt), _dp(mad)),
             "sap_row_activities")
        return mn, mx, cnt.reshape(nr, 8), mad

    def close(self):
        if self.h:
            lib().sap_matrix_free(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def uneven_sizes(count: int, seed: int = 4):
    """Row counts of the uneven-instances workload: log-uniform in [500, 500,000]."""
    rng = np.random.default_rng(seed)
    return np.rint(np.exp(rng.uniform(np.log(500), np.log(500_000), size=count))).astype(np.int64)


def gen_uneven_instances(count: int = 512, seed: int = 4, threads: int = 16, which=None):
    """The uneven-instances workload: instance i has rows from uneven_sizes, 1.5 x as many
    columns, 4..64 nonzeros per row, the integer/continuous value, bound and coefficient
    distributions of gen_mixed, sides centred on a planted feasible point, seed 4000 + i
    (seed 4 draws the sizes).  which: the instance indices to build (default all)."""
    from concurrent.futures import ThreadPoolExecutor
    m = uneven_sizes(count, seed)
    idx = range(count) if which is None else list(which)
    one = lambda i: gen_mixed(int(m[i]), int(round(1.5 * m[i])), seed=1000 * seed + i, lenmode=1, minlen=4,
                              maxlen=64, planted=True)
    with ThreadPoolExecutor(max(1, threads)) as ex:
        return list(ex.map(one, idx))


def gen_serving_trace(probs, rate: float, seconds: float, seed: int = 5, zipf: float = 1.2):
    """The serving workload's request trace over probs = [covering/packing problem, instance
    1, ..., instance k]: Poisson arrivals at rate for seconds; half the requests are nodes of
    probs[0], half nodes of an instance chosen with Zipf(zipf) popularity over a random order of
    the instances; a node fixes 1..50 distinct integer columns to 0 or 1.  Returns (arrival_s,
    prob, ovptr, ovcol, ovval, popularity order of the instances (indices into probs))."""
    rng = np.random.default_rng(seed)
    k = len(probs) - 1
    order = 1 + rng.permutation(k)
    w = 1.0 / np.arange(1, k + 1) ** zipf
    w /= w.sum()
    n = int(rate * seconds)
    arrival = np.cumsum(rng.exponential(1.0 / rate, size=n))
    arrival -= arrival[0]
    kind = rng.integers(0, 2, size=n)
    prob = np.where(kind == 0, 0, order[rng.choice(k, size=n, p=w)]).astype(np.int32)
    intcols = [np.nonzero(p.integral)[0] for p in probs]
    ptr = np.zeros(n + 1, dtype=np.int64)
    cols, vals = [], []
    for i in range(n):
        ic = intcols[prob[i]]
        m = int(min(len(ic), rng.integers(1, 51)))
        c = ic[np.unique(rng.integers(0, len(ic), size=m))] if m else np.zeros(0, dtype=np.int64)
        cols.append(c.astype(np.int32))
        vals.append(rng.integers(0, 2, size=len(c)).astype(np.float64))
        ptr[i + 1] = ptr[i] + len(c)
    col = np.concatenate(cols) if cols else np.zeros(0, dtype=np.int32)
    val = np.concatenate(vals) if vals else np.zeros(0)
    return arrival, prob, ptr, col, val, order


def pack_instances(insts):
    """Concatenate instances block-diagonally: returns (Instance, rowoff, coloff)."""
    rowoff = np.zeros(len(insts) + 1, dtype=np.int64)
    coloff = np.zeros(len(insts) + 1, dtype=np.int64)
    for i, t in enumerate(insts):
        rowoff[i + 1] = rowoff[i] + t.nrows
        coloff[i + 1] = coloff[i] + t.ncols
    nnzoff = np.cumsum([0] + [t.nnz for t in insts])
    rowptr = np.concatenate([[0]] + [t.rowptr[1:] + nnzoff[i] for i, t in enumerate(insts)]).astype(np.int64)
    colidx = np.concatenate([t.colidx.astype(np.int64) + coloff[i] for i, t in enumerate(insts)]).astype(np.int32)
    cat = lambda f, dt: np.concatenate([getattr(t, f) for t in insts]).astype(dt)
    inst = Instance(rowptr=rowptr, colidx=colidx, vals=cat("vals", np.float64), lhs=cat("lhs", np.float64),
                    rhs=cat("rhs", np.float64), integral=cat("integral", np.uint8), lb=cat("lb", np.float64),
                    ub=cat("ub", np.float64))
    return inst, rowoff, coloff


class PackedMatrix(Matrix):
    """Several independent instances in one resident matrix (sap_matrix_create_packed)."""

    def __init__(self, ctx: Context, insts):
        self.ctx = ctx
        self.parts = list(insts)
        inst, self.rowoff, self.coloff = pack_instances(self.parts)
        self.inst = inst
        self.h = _P()
        self._keep = [np.ascontiguousarray(a) for a in (inst.rowptr, inst.colidx, inst.vals, inst.lhs, inst.rhs,
                                                         inst.integral, inst.lb, inst.ub)]
        rp, ci, va, lh, rh, it, lb, ub = self._keep
        _chk(lib().sap_matrix_create_packed(ctx.h, len(self.parts), _i64p(self.rowoff), _i64p(self.coloff),
                                            inst.nrows, inst.ncols, _i64p(rp), _i32p(ci), _dp(va), _dp(lh), _dp(rh),
                                            _u8p(it), _dp(lb), _dp(ub), C.byref(self.h)), "sap_matrix_create_packed")
        _LIVE["matrix"].add(self)

    def propagate_packed(self, lb=None, ub=None, params: sap_params | None = None):
        """Returns (lb, ub, result[ninst], rounds[ninst], stats)."""
        lb = np.array(self.inst.lb if lb is None else lb, dtype=np.float64, copy=True)
        ub = np.array(self.inst.ub if ub is None else ub, dtype=np.float64, copy=True)
        k = len(self.parts)
        res = np.zeros(k, dtype=np.int32)
        rnd = np.zeros(k, dtype=np.int32)
        st = sap_stats()
        p = params or default_params()
        _chk(lib().sap_propagate_packed(self.h, C.byref(p), _dp(lb), _dp(ub), _i32p(res), _i32p(rnd), C.byref(st)),
             "sap_propagate_packed")
        return lb, ub, res, rnd, st.as_dict()


def _overlay_arrays(ovptr, ovcol, ovlb, ovub):
    return (np.ascontiguousarray(ovptr, dtype=np.int64), np.ascontiguousarray(ovcol, dtype=np.int32),
            np.ascontiguousarray(ovlb, dtype=np.float64), np.ascontiguousarray(ovub, dtype=np.float64))


class Bounds:
    """Device-resident bound vectors (sap_bounds)."""

    def __init__(self, mat: Matrix, nvec: int = 1, mode: int = 0):
        """mode: SAP_MODE_AUTO (0), SAP_MODE_DENSE (1) or SAP_MODE_SPARSE (2)."""
        self.mat = mat
        self.nvec = nvec
        self.h = _P()
        _chk(lib().sap_bounds_create_mode(mat.h, nvec, mode, C.byref(self.h)), "sap_bounds_create_mode")
        _LIVE["bounds"].add(self)

    def upload(self, lb, ub, v: int = -1):
        lb = np.ascontiguousarray(lb, dtype=np.float64)
        ub = np.ascontiguousarray(ub, dtype=np.float64)
        _chk(lib().sap_bounds_upload(self.h, v, _dp(lb), _dp(ub)), "sap_bounds_upload")

    def download(self, v: int = 0):
        n = self.mat.inst.ncols
        lb, ub = np.zeros(n), np.zeros(n)
        _chk(lib().sap_bounds_download(self.h, v, _dp(lb), _dp(ub)), "sap_bounds_download")
        return lb, ub

    def bench_gather(self, reps: int = 20, blocks: int = 0) -> float:
        """Mean time (ms) of the row sweep's bare memory traffic over this matrix (sap_bench_gather)."""
        ms = C.c_double(0.0)
        _chk(lib().sap_bench_gather(self.h, reps, blocks, C.byref(ms)), "sap_bench_gather")
        return ms.value

    def copy_from(self, src: "Bounds", dv: int = 0, sv: int = 0):
        _chk(lib().sap_bounds_copy(self.h, dv, src.h, sv), "sap_bounds_copy")

    def propagate(self, params: sap_params | None = None):
        st = sap_stats()
        p = params or default_params()
        _chk(lib().sap_propagate_resident(self.h, C.byref(p), C.byref(st)), "sap_propagate_resident")
        return st.as_dict()

    def propagate_batch(self, base_lb, base_ub, ovptr, ovcol, ovlb, ovub, params: sap_params | None = None,
                        want_bounds: bool = True, want_result: bool = True):
        nvec = len(ovptr) - 1
        n = self.mat.inst.ncols
        base_lb = np.ascontiguousarray(base_lb, dtype=np.float64)
        base_ub = np.ascontiguousarray(base_ub, dtype=np.float64)
        ovptr = np.ascontiguousarray(ovptr, dtype=np.int64)
        ovcol = np.ascontiguousarray(ovcol, dtype=np.int32)
        ovlb = np.ascontiguousarray(ovlb, dtype=np.float64)
        ovub = np.ascontiguousarray(ovub, dtype=np.float64)
        res = np.zeros(nvec, dtype=np.int32) if want_result else None
        rnd = np.zeros(nvec, dtype=np.int32)
        olb = np.zeros(nvec * n) if want_bounds else None
        oub = np.zeros(nvec * n) if want_bounds else None
        st = sap_stats()
        p = params or default_params()
        _chk(lib().sap_propagate_batch(self.h, C.byref(p), nvec, _dp(base_lb), _dp(base_ub), _i64p(ovptr),
                                       _i32p(ovcol), _dp(ovlb), _dp(ovub), _i32p(res), _i32p(rnd), _dp(olb), _dp(oub),
                                       C.byref(st)), "sap_propagate_batch")
        out = dict(result=res, rounds=rnd, stats=st.as_dict())
        if want_bounds:
            out["lb"] = olb.reshape(nvec, n)
            out["ub"] = oub.reshape(nvec, n)
        return out

    def propagate_nodes(self, base: "Bounds", basevec, ovptr, ovcol, ovlb, ovub, params: sap_params | None = None,
                        chgcap: int | None = None, want_changes: bool = True):
        """Node vectors from a resident base vector; sparse changed-bound lists per node."""
        nvec = len(ovptr) - 1
        ovptr, ovcol, ovlb, ovub = _overlay_arrays(ovptr, ovcol, ovlb, ovub)
        res = np.zeros(nvec, dtype=np.int32)
        rnd = np.zeros(nvec, dtype=np.int32)
        cap = int(chgcap if chgcap is not None else max(1024, 64 * nvec))
        while True:
            cptr = np.zeros(nvec + 1, dtype=np.int64)
            ccol = np.zeros(max(cap, 1), dtype=np.int32) if want_changes else None
            clb = np.zeros(max(cap, 1)) if want_changes else None
            cub = np.zeros(max(cap, 1)) if want_changes else None
            st = sap_stats()
            p = params or default_params()
            rc = lib().sap_propagate_nodes(self.h, C.byref(p), nvec, base.h, basevec, _i64p(ovptr), _i32p(ovcol),
                                           _dp(ovlb), _dp(ovub), _i32p(res), _i32p(rnd), cap, _i64p(cptr),
                                           _i32p(ccol), _dp(clb), _dp(cub), C.byref(st))
            if rc == SAP_NOMEMORY and chgcap is None and int(cptr[-1]) > cap:
                cap = int(cptr[-1])   # run again with room for every change (same results)
                continue
            _chk(rc, "sap_propagate_nodes")
            n = int(cptr[-1])
            out = dict(result=res, rounds=rnd, chgptr=cptr, stats=st.as_dict())
            if want_changes:
                out.update(chgcol=ccol[:n], chglb=clb[:n], chgub=cub[:n])
            return out

    def probe(self, base: "Bounds", basevec, probecol, params: sap_params | None = None,
              dedcap: int | None = None, want_pairs: bool = True, downub=None, uplb=None):
        """Both branches x = 0 / x = 1 of every probe column (with downub/uplb: x <= downub,
        x >= uplb) from a resident base vector: per-probe verdict bits, per-branch rounds and
        the columns changed in both branches."""
        probecol = np.ascontiguousarray(probecol, dtype=np.int32)
        n = len(probecol)
        split = downub is not None
        if split:
            downub = np.ascontiguousarray(downub, dtype=np.float64)
            uplb = np.ascontiguousarray(uplb, dtype=np.float64)
        ver = np.zeros(max(n, 1), dtype=np.int32)
        rnd = np.zeros(max(2 * n, 1), dtype=np.int32)
        cap = int(dedcap if dedcap is not None else max(1024, 16 * n))
        while True:
            dptr = np.zeros(n + 1, dtype=np.int64)
            arrs = [np.zeros(max(cap, 1), dtype=np.int32)] + [np.zeros(max(cap, 1)) for _ in range(4)] \
                if want_pairs else [None] * 5
            st = sap_stats()
            p = params or default_params()
            if split:
                rc = lib().sap_probe_split(self.h, C.byref(p), n, _i32p(probecol), _dp(downub), _dp(uplb), base.h,
                                           basevec, _i32p(ver), _i32p(rnd), cap, _i64p(dptr), _i32p(arrs[0]),
                                           _dp(arrs[1]), _dp(arrs[2]), _dp(arrs[3]), _dp(arrs[4]), C.byref(st))
            else:
                rc = lib().sap_probe(self.h, C.byref(p), n, _i32p(probecol), base.h, basevec, _i32p(ver),
                                     _i32p(rnd), cap, _i64p(dptr), _i32p(arrs[0]), _dp(arrs[1]), _dp(arrs[2]),
                                     _dp(arrs[3]), _dp(arrs[4]), C.byref(st))
            if rc == SAP_NOMEMORY and dedcap is None and int(dptr[-1]) > cap:
                cap = int(dptr[-1])   # run again with room for every pair (same results)
                continue
            _chk(rc, "sap_probe_split" if split else "sap_probe")
            k = int(dptr[-1])
            out = dict(verdict=ver[:n], rounds=rnd[:2 * n].reshape(n, 2), dedptr=dptr, stats=st.as_dict())
            if want_pairs:
                out.update(dedcol=arrs[0][:k], dlb0=arrs[1][:k], dub0=arrs[2][:k], dlb1=arrs[3][:k], dub1=arrs[4][:k])
            return out

    # ---- continuous batching ----
    def admit(self, slots, base: "Bounds", basevec, ovptr, ovcol, ovlb, ovub, flags: int = SAP_FLAG_BASE_IS_FIXPOINT):
        slots = np.ascontiguousarray(slots, dtype=np.int32)
        ovptr, ovcol, ovlb, ovub = _overlay_arrays(ovptr, ovcol, ovlb, ovub)
        _chk(lib().sap_slots_admit(self.h, len(slots), _i32p(slots), base.h, basevec, _i64p(ovptr), _i32p(ovcol),
                                   _dp(ovlb), _dp(ovub), flags), "sap_slots_admit")

    def launch(self, params: sap_params | None = None, rounds: int = SAP_STEP_LONG):
        p = params or default_params()
        _chk(lib().sap_slots_launch(self.h, C.byref(p), rounds), "sap_slots_launch")

    def query(self) -> bool:
        r = C.c_int32(0)
        _chk(lib().sap_slots_query(self.h, C.byref(r)), "sap_slots_query")
        return bool(r.value)

    def wait(self):
        nf = C.c_int32(0)
        fin = np.zeros(self.nvec, dtype=np.int32)
        st = sap_stats()
        _chk(lib().sap_slots_wait(self.h, C.byref(nf), _i32p(fin), C.byref(st)), "sap_slots_wait")
        return fin[:nf.value].copy(), st.as_dict()

    def step(self, params: sap_params | None = None, rounds: int = SAP_STEP_LONG):
        self.launch(params, rounds)
        return self.wait()

    def collect(self, slots, chgcap: int | None = None):
        slots = np.ascontiguousarray(slots, dtype=np.int32)
        n = len(slots)
        res = np.zeros(n, dtype=np.int32)
        rnd = np.zeros(n, dtype=np.int32)
        cptr = np.zeros(n + 1, dtype=np.int64)
        cap = int(chgcap if chgcap is not None else max(1024, 64 * n))
        for _ in range(2):
            ccol = np.zeros(max(cap, 1), dtype=np.int32)
            clb = np.zeros(max(cap, 1))
            cub = np.zeros(max(cap, 1))
            rc = lib().sap_slots_collect(self.h, n, _i32p(slots), _i32p(res), _i32p(rnd), cap, _i64p(cptr),
                                         _i32p(ccol), _dp(clb), _dp(cub))
            if rc == SAP_NOMEMORY and chgcap is None:
                cap = int(cptr[-1])
                continue
            _chk(rc, "sap_slots_collect")
            m = int(cptr[-1])
            return dict(result=res, rounds=rnd, chgptr=cptr.copy(), chgcol=ccol[:m], chglb=clb[:m], chgub=cub[:m])
        raise SapError(SAP_NOMEMORY, "sap_slots_collect")

    def close(self):
        if self.h:
            lib().sap_bounds_free(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------------------------------
# serving
# ------------------------------------------------------------------------------------------
class Server:
    """Request server (include/scip_amd_serve.h): resident problems, admission by HBM
    footprint, continuous batching of node requests."""

    def __init__(self, device: int = 0, nstreams: int = 4, hbm_budget: int = 64 << 30, slots: int = 64):
        self.h = _P()
        self.slots = slots
        _chk(lib().sap_server_create(device, nstreams, hbm_budget, slots, C.byref(self.h)), "sap_server_create")
        _LIVE["server"].add(self)

    def add_problem(self, inst: Instance) -> int:
        rp, ci, va, lh, rh, it, lb, ub = [np.ascontiguousarray(a) for a in (
            inst.rowptr, inst.colidx, inst.vals, inst.lhs, inst.rhs, inst.integral, inst.lb, inst.ub)]
        pid = C.c_int32(0)
        _chk(lib().sap_server_add_problem(self.h, inst.nrows, inst.ncols, _i64p(rp), _i32p(ci), _dp(va), _dp(lh),
                                          _dp(rh), _u8p(it), _dp(lb), _dp(ub), C.byref(pid)), "sap_server_add_problem")
        return pid.value

    def set_budget(self, hbm_budget: int):
        _chk(lib().sap_server_set_budget(self.h, hbm_budget), "sap_server_set_budget")

    def footprint(self, pid: int) -> int:
        """Device bytes the problem occupies while resident (what admission counts)."""
        b = C.c_int64(0)
        _chk(lib().sap_server_footprint(self.h, pid, C.byref(b)), "sap_server_footprint")
        return b.value

    def warm(self, pid: int):
        _chk(lib().sap_server_warm(self.h, pid), "sap_server_warm")

    def set_slots(self, pid: int, slots: int):
        _chk(lib().sap_server_set_slots(self.h, pid, slots), "sap_server_set_slots")

    def replay(self, arrival_s, prob, ovptr, ovcol, ovlb, ovub, mode: int = SAP_SERVE_CONTINUOUS,
               steprounds: int = SAP_STEP_SHORT, params: sap_params | None = None, want_changes: bool = False):
        arrival_s = np.ascontiguousarray(arrival_s, dtype=np.float64)
        prob = np.ascontiguousarray(prob, dtype=np.int32)
        ovptr, ovcol, ovlb, ovub = _overlay_arrays(ovptr, ovcol, ovlb, ovub)
        n = len(arrival_s)
        res = np.zeros(n, dtype=np.int32)
        rnd = np.zeros(n, dtype=np.int32)
        nch = np.zeros(n, dtype=np.int32)
        lat = np.zeros(n)
        st = sap_serve_stats()
        p = params or default_params()
        if want_changes:
            cap = max(1024, 256 * n)
            while True:
                cptr = np.zeros(n + 1, dtype=np.int64)
                ccol = np.zeros(cap, dtype=np.int32)
                clb = np.zeros(cap)
                cub = np.zeros(cap)
                rc = lib().sap_server_replay(self.h, C.byref(p), mode, steprounds, n, _dp(arrival_s), _i32p(prob),
                                             _i64p(ovptr), _i32p(ovcol), _dp(ovlb), _dp(ovub), _i32p(res), _i32p(rnd),
                                             _i32p(nch), _dp(lat), cap, _i64p(cptr), _i32p(ccol), _dp(clb), _dp(cub),
                                             C.byref(st))
                if rc == SAP_NOMEMORY and int(cptr[-1]) > cap:
                    cap = int(cptr[-1])
                    continue
                _chk(rc, "sap_server_replay")
                m = int(cptr[-1])
                return dict(result=res, rounds=rnd, nchg=nch, latency_ms=lat, stats=st.as_dict(), chgptr=cptr,
                            chgcol=ccol[:m], chglb=clb[:m], chgub=cub[:m])
        _chk(lib().sap_server_replay(self.h, C.byref(p), mode, steprounds, n, _dp(arrival_s), _i32p(prob),
                                     _i64p(ovptr), _i32p(ovcol), _dp(ovlb), _dp(ovub), _i32p(res), _i32p(rnd),
                                     _i32p(nch), _dp(lat), 0, None, None, None, None, C.byref(st)),
             "sap_server_replay")
        cold = np.zeros(n, dtype=np.uint8)
        _chk(lib().sap_server_cold(self.h, n, _u8p(cold)), "sap_server_cold")
        return dict(result=res, rounds=rnd, nchg=nch, latency_ms=lat, stats=st.as_dict(), cold=cold.astype(bool))

    def close(self):
        if self.h:
            lib().sap_server_free(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
