"""ctypes driver for plugin/libgpulinear_testhost.so: the SCIP propagator plugin
(scip_amd/plugin/prop_gpulinear.c) running inside the test double of the SCIP API subset
it calls (scip_amd/plugin/testhost/testhost.c)."""
import ctypes as C
import os

import numpy as np

import scip_amd as sa

LIBPATH = os.path.join(os.path.dirname(sa.__file__), "plugin", "libgpulinear_testhost.so")

# SCIP_RETCODE / SCIP_RESULT values (src/scip/type_retcode.h, type_result.h)
SCIP_OKAY = 1
SCIP_DIDNOTRUN = 1
SCIP_DIDNOTFIND = 3
SCIP_CUTOFF = 7
SCIP_REDUCEDDOM = 10
SCIP_SUCCESS = 17

# SCIP_STAGE values (src/scip/type_set.h:41-56)
SCIP_STAGE_PRESOLVING = 5
SCIP_STAGE_SOLVING = 9

_lib = None
_P = np.ctypeslib.ndpointer


def lib():
    global _lib
    if _lib is None:
        sa.lib()  # the engine library first, so the plugin's dependency resolves
        L = C.CDLL(LIBPATH)
        L.th_create.restype = C.c_void_p
        L.th_create.argtypes = []
        L.th_add_vars.argtypes = [C.c_void_p, C.c_int, _P(np.float64), _P(np.float64), _P(np.uint8)]
        L.th_add_rows.argtypes = [C.c_void_p, C.c_int, _P(np.int64), _P(np.int32), _P(np.float64),
                                  _P(np.float64), _P(np.float64)]
        for f in ("th_include_gpulinear", "th_initpre", "th_exitpre", "th_initsol", "th_exitsol", "th_free"):
            getattr(L, f).argtypes = [C.c_void_p]
        L.th_set_int_param.argtypes = [C.c_void_p, C.c_char_p, C.c_int]
        L.th_propexec.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.th_presol.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.th_resprop.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_int)]
        L.th_set_local_bounds.argtypes = [C.c_void_p, _P(np.float64), _P(np.float64)]
        L.th_set_local_bounds.restype = None
        L.th_get_bounds.argtypes = [C.c_void_p, C.c_int, _P(np.float64), _P(np.float64)]
        L.th_get_bounds.restype = None
        IP = C.POINTER(C.c_int)
        L.th_get_probe_records.argtypes = [C.c_void_p, IP, C.c_void_p, C.c_void_p, IP, C.c_void_p, C.c_void_p,
                                           C.c_void_p, IP, IP]
        L.th_set_rangedrowprop.argtypes = [C.c_void_p, C.c_int]
        L.th_set_rangedrowprop.restype = None
        L.th_set_linear_int.argtypes = [C.c_void_p, C.c_char_p, C.c_int]
        L.th_set_node.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
        L.th_set_node.restype = None
        IPn = _P(np.int32)
        L.th_strongbranch.argtypes = [C.c_void_p, C.c_int, IPn, _P(np.float64), IPn, IPn, _P(np.float64),
                                      _P(np.float64), C.POINTER(C.c_int)]
        L.th_dive.argtypes = [C.c_void_p, C.c_int, C.c_int, IPn, IPn, _P(np.float64), C.c_int, IPn, IPn,
                              _P(np.float64), _P(np.float64)]
        L.th_resprop_change.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int)]
        L.th_get_changes.argtypes = [C.c_void_p] + [C.c_void_p] * 5
        L.th_get_conflict.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_int)]
        _lib = L
    return _lib


class Host:
    """One SCIP problem (variables + linear rows) with the gpulinear propagator included."""

    def __init__(self, inst):
        L = lib()
        self.n = int(len(inst.lb))
        self.h = L.th_create()
        assert self.h
        assert L.th_add_vars(self.h, self.n, np.ascontiguousarray(inst.lb, np.float64),
                             np.ascontiguousarray(inst.ub, np.float64),
                             np.ascontiguousarray(inst.integral, np.uint8)) == SCIP_OKAY
        nrows = len(inst.rowptr) - 1
        assert L.th_add_rows(self.h, nrows, np.ascontiguousarray(inst.rowptr, np.int64),
                             np.ascontiguousarray(inst.colidx, np.int32), np.ascontiguousarray(inst.vals, np.float64),
                             np.ascontiguousarray(inst.lhs, np.float64),
                             np.ascontiguousarray(inst.rhs, np.float64)) == SCIP_OKAY
        assert L.th_include_gpulinear(self.h) == SCIP_OKAY

    def call(self, name, *args):
        return getattr(lib(), name)(self.h, *args)

    def set_node(self, stage=SCIP_STAGE_SOLVING, inprobing=False, depth=0, inreprop=False):
        lib().th_set_node(self.h, stage, 1 if inprobing else 0, depth, 1 if inreprop else 0)

    def set_linear_int(self, name, value):
        assert lib().th_set_linear_int(self.h, name.encode(), value) == SCIP_OKAY

    def strongbranch(self, cands, solvals):
        """SCIPgpulinearStrongbranch at the current local bounds: (downinf, upinf, newlbs, newubs,
        cutoff, rc)."""
        n = len(cands)
        di = np.zeros(max(n, 1), dtype=np.int32)
        ui = np.zeros(max(n, 1), dtype=np.int32)
        nl = np.zeros(self.n)
        nu = np.zeros(self.n)
        cut = C.c_int(0)
        rc = lib().th_strongbranch(self.h, n, np.ascontiguousarray(cands, dtype=np.int32),
                                   np.ascontiguousarray(solvals, dtype=np.float64), di, ui, nl, nu, C.byref(cut))
        return di[:n].astype(bool), ui[:n].astype(bool), nl, nu, bool(cut.value), rc

    def dive(self, cands, order, fixvals, maxdepth=-1):
        """SCIPgpulinearDive at the current local bounds: (depth, cutoff, divelbs, diveubs, rc)."""
        order = np.ascontiguousarray(order, dtype=np.int32)
        nd = order.shape[0]
        depth = np.zeros(max(nd, 1), dtype=np.int32)
        cut = np.zeros(max(nd, 1), dtype=np.int32)
        dl = np.zeros(max(nd, 1) * self.n)
        du = np.zeros(max(nd, 1) * self.n)
        rc = lib().th_dive(self.h, nd, len(cands), np.ascontiguousarray(cands, dtype=np.int32), order.ravel(),
                           np.ascontiguousarray(fixvals, dtype=np.float64).ravel(), maxdepth, depth, cut, dl, du)
        return depth[:nd], cut[:nd].astype(bool), dl.reshape(-1, self.n)[:nd], du.reshape(-1, self.n)[:nd], rc

    def propexec(self):
        r = C.c_int(0)
        rc = lib().th_propexec(self.h, C.byref(r))
        return rc, r.value

    def presol(self):
        r = C.c_int(0)
        n = C.c_int(0)
        rc = lib().th_presol(self.h, C.byref(r), C.byref(n))
        return rc, r.value, n.value

    def resprop(self, j, upper):
        r = C.c_int(0)
        rc = lib().th_resprop(self.h, j, 1 if upper else 0, C.byref(r))
        return rc, r.value

    def set_local(self, lb, ub):
        lib().th_set_local_bounds(self.h, np.ascontiguousarray(lb, np.float64), np.ascontiguousarray(ub, np.float64))

    def bounds(self, glob=False):
        lb = np.empty(self.n)
        ub = np.empty(self.n)
        lib().th_get_bounds(self.h, 1 if glob else 0, lb, ub)
        return lb, ub

    def probe_records(self):
        """(fixings [(var, value)], deductions [(probing var, var, lb0, ub0, lb1, ub1)], naggr, nimpl)"""
        L = lib()
        nf, nd, na, ni = C.c_int(0), C.c_int(0), C.c_int(0), C.c_int(0)
        L.th_get_probe_records(self.h, C.byref(nf), None, None, C.byref(nd), None, None, None, C.byref(na),
                               C.byref(ni))
        fv = np.zeros(max(nf.value, 1), dtype=np.int32)
        fx = np.zeros(max(nf.value, 1))
        dp = np.zeros(max(nd.value, 1), dtype=np.int32)
        dv = np.zeros(max(nd.value, 1), dtype=np.int32)
        db = np.zeros((max(nd.value, 1), 4))
        L.th_get_probe_records(self.h, C.byref(nf), fv.ctypes.data, fx.ctypes.data, C.byref(nd), dp.ctypes.data,
                               dv.ctypes.data, db.ctypes.data, C.byref(na), C.byref(ni))
        fix = [(int(fv[i]), float(fx[i])) for i in range(nf.value)]
        ded = [(int(dp[i]), int(dv[i])) + tuple(float(x) for x in db[i]) for i in range(nd.value)]
        return fix, ded, na.value, ni.value

    def changes(self):
        """Local bound changes handed to SCIP: (var, bound 0/1, inferinfo, old, new) arrays."""
        L = lib()
        n = L.th_get_changes(self.h, None, None, None, None, None)
        v, b, i = (np.zeros(max(n, 1), dtype=np.int32) for _ in range(3))
        o, w = np.zeros(max(n, 1)), np.zeros(max(n, 1))
        L.th_get_changes(self.h, v.ctypes.data, b.ctypes.data, i.ctypes.data, o.ctypes.data, w.ctypes.data)
        return v[:n], b[:n], i[:n], o[:n], w[:n]

    def resprop_change(self, pos):
        r = C.c_int(0)
        rc = lib().th_resprop_change(self.h, pos, C.byref(r))
        return rc, r.value

    def conflict(self):
        """Conflict candidate set [(var, bound 0/1, value)] and the number of analyses."""
        L = lib()
        na = C.c_int(0)
        n = L.th_get_conflict(self.h, None, None, None, C.byref(na))
        v, b = np.zeros(max(n, 1), dtype=np.int32), np.zeros(max(n, 1), dtype=np.int32)
        x = np.zeros(max(n, 1))
        L.th_get_conflict(self.h, v.ctypes.data, b.ctypes.data, x.ctypes.data, C.byref(na))
        return [(int(v[k]), int(b[k]), float(x[k])) for k in range(n)], na.value

    def close(self):
        if self.h:
            rc = lib().th_free(self.h)
            self.h = None
            return rc
        return SCIP_OKAY
