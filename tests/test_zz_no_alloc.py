"""The no-allocation entry point (sap_propagate_device_noalloc) against sap_propagate, through
the C ABI: two calls in a row on the same device scratch, the second on a smaller matrix, and
the size and ranged-row refusals.  The caller's device buffers are allocated here with the HIP
runtime the engine itself links (libamdhip64), as a caller that owns device memory would."""
import ctypes as C

import numpy as np
import pytest

import scip_amd as sa
from tests.util import INF, make_instance

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


_hip = None


def hip():
    global _hip
    if _hip is None:
        sa.lib()
        _hip = C.CDLL("libamdhip64.so.7")   # the runtime libscipamd.so is linked against
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


class DeviceArray:
    """ncols doubles of caller-owned device memory."""

    def __init__(self, host):
        host = np.ascontiguousarray(host, dtype=np.float64)
        self.n = len(host)
        self.ptr = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.ptr), max(8 * self.n, 8)) == 0
        assert hip().hipMemcpy(self.ptr, host.ctypes.data, 8 * self.n, 1) == 0   # host to device

    def data_ptr(self):
        return self.ptr.value

    def numpy(self):
        out = np.empty(self.n)
        assert hip().hipMemcpy(out.ctypes.data, self.ptr, 8 * self.n, 2) == 0   # device to host
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()


def device_pair(lb, ub):
    dl, du = DeviceArray(lb), DeviceArray(ub)
    assert hip().hipDeviceSynchronize() == 0
    return dl, du


def check_against_plain(ctx, inst, name, flags=0):
    say(f"[case] noalloc {name} rows={inst.nrows} cols={inst.ncols}")
    mat = sa.Matrix(ctx, inst)
    p = sa.default_params(flags=flags)
    gl, gu, gs = mat.propagate(inst.lb, inst.ub, p)
    dl, du = device_pair(inst.lb, inst.ub)
    ns = mat.propagate_device_noalloc(dl.data_ptr(), du.data_ptr(), p)
    nl, nu = dl.numpy(), du.numpy()
    dl.free()
    du.free()
    mat.close()
    assert ns["result"] == gs["result"], (ns, gs)
    for k in ("rounds", "nchgbds", "rows_swept", "nnz_swept", "nrecompute"):
        if gs["result"] != sa.SAP_CUTOFF or k == "rounds":
            assert ns[k] == gs[k], (k, ns[k], gs[k])
    if gs["result"] == sa.SAP_CUTOFF:
        assert np.array_equal(nl, inst.lb) and np.array_equal(nu, inst.ub)   # untouched
    else:
        assert np.array_equal(nl, gl) and np.array_equal(nu, gu)
    return gs


def test_noalloc_equals_sap_propagate_twice_then_smaller(gpu_ctx):
    big = sa.gen_mixed(60000, 60000, seed=1, planted=True)
    small = sa.gen_mixed(5000, 7000, seed=3, planted=True)
    a = check_against_plain(gpu_ctx, big, "big")
    assert a["result"] == sa.SAP_REDUCEDDOM and a["rounds"] >= 3
    check_against_plain(gpu_ctx, small, "smaller")
    check_against_plain(gpu_ctx, big, "big again")


def test_noalloc_cutoff_and_no_tightening(gpu_ctx):
    check_against_plain(gpu_ctx, sa.gen_mixed(30000, 30000, seed=2), "literal cutoff")
    check_against_plain(gpu_ctx, sa.gen_mixed(20000, 20000, seed=6, planted=True), "notighten",
                        flags=sa.SAP_FLAG_NOTIGHTEN)
    check_against_plain(gpu_ctx, sa.gen_covering_packing(2000, 5000, 10, seed=1), "cover 2000x5000")


def test_noalloc_refuses_what_does_not_fit(gpu_ctx):
    # more columns than the scratch holds
    n = (1 << 20) + 1
    inst = sa.Instance(rowptr=np.array([0, 2], dtype=np.int64), colidx=np.array([0, n - 1], dtype=np.int32),
                       vals=np.array([1.0, 1.0]), lhs=np.array([-INF]), rhs=np.array([1.0]),
                       integral=np.zeros(n, dtype=np.uint8), lb=np.zeros(n), ub=np.ones(n))
    mat = sa.Matrix(gpu_ctx, inst)
    dl, du = device_pair(inst.lb, inst.ub)
    with pytest.raises(sa.SapError) as e:
        mat.propagate_device_noalloc(dl.data_ptr(), du.data_ptr())
    assert e.value.rc == sa.SAP_INVALIDDATA
    dl.free()
    du.free()
    mat.close()
    # ranged-row propagation needs per-nonzero flags
    rows = [([(0, 2.0), (1, 2.0), (2, 3.0)], 3.0, 3.0)]
    inst = make_instance(rows, [0.0, 0.0, 0.0], [1.0, 1.0, 1.0], [1, 1, 1])
    mat = sa.Matrix(gpu_ctx, inst)
    dl, du = device_pair(inst.lb, inst.ub)
    with pytest.raises(sa.SapError) as e:
        mat.propagate_device_noalloc(dl.data_ptr(), du.data_ptr(), sa.default_params(flags=sa.SAP_FLAG_RANGEDROWS))
    assert e.value.rc == sa.SAP_INVALIDDATA
    dl.free()
    du.free()
    mat.close()
