"""The C-ABI library loads and exports every symbol include/*.h declares; host-side
behaviour that needs no device (defaults, error strings, generators)."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

import scip_amd as sa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DECL = re.compile(r"^\s*(?:const\s+)?\w+\s*\*?\s*(sap_\w+)\s*\(", re.M)


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        with open(h) as f:
            names |= set(DECL.findall(f.read()))
    return sorted(names)


def test_headers_declare_the_abi():
    names = declared_symbols()
    assert "sap_propagate" in names and "sap_matrix_create" in names and "sap_gen_mixed" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    L = sa.lib()
    missing = [n for n in declared_symbols() if not hasattr(L, n)]
    assert missing == []


def test_binding_covers_every_declared_symbol():
    assert sorted(sa.EXPORTS) == declared_symbols()


def test_param_defaults_are_scips():
    p = sa.default_params()
    assert (p.infinity, p.epsilon, p.sumepsilon, p.feastol) == (1e20, 1e-9, 1e-6, 1e-6)
    assert (p.boundstreps, p.hugeval, p.recompfac, p.maxeasyactivitydelta) == (0.05, 1e15, 1e6, 1e6)
    assert p.maxrounds == -1 and p.flags == 0


def test_error_strings_and_codes():
    L = sa.lib()
    for rc in (sa.SAP_OKAY, sa.SAP_ERROR, sa.SAP_NOMEMORY, sa.SAP_INVALIDCALL, sa.SAP_INVALIDDATA,
               sa.SAP_NOTIMPLEMENTED):
        assert L.sap_error_string(rc)
    assert L.sap_error_string(12345) == b"unknown return code"
    # SCIP_RETCODE values (type_retcode.h:42-60)
    assert (sa.SAP_OKAY, sa.SAP_ERROR, sa.SAP_NOMEMORY, sa.SAP_INVALIDCALL, sa.SAP_INVALIDDATA) == (1, 0, -1, -8, -9)


def test_null_arguments_are_rejected_without_a_device():
    L = sa.lib()
    assert L.sap_context_create(0, None) == sa.SAP_INVALIDDATA
    assert L.sap_propagate(None, None, None, None, None) == sa.SAP_INVALIDDATA
    assert L.sap_matrix_info(None, None, None, None, None) == sa.SAP_INVALIDDATA
    assert L.sap_context_free(None) == sa.SAP_OKAY


def test_context_create_without_device_fails_cleanly():
    if sa.device_count() > 0:
        pytest.skip("a device is visible")
    h = C.c_void_p()
    assert sa.lib().sap_context_create(0, C.byref(h)) == sa.SAP_INVALIDCALL


def test_generators_deterministic_and_well_formed():
    a = sa.gen_covering_packing(500, 1200, 10, seed=1)
    b = sa.gen_covering_packing(500, 1200, 10, seed=1)
    c = sa.gen_covering_packing(500, 1200, 10, seed=2)
    assert np.array_equal(a.colidx, b.colidx) and np.array_equal(a.vals, b.vals)
    assert not np.array_equal(a.colidx, c.colidx)
    for i in range(a.nrows):
        s, e = a.rowptr[i], a.rowptr[i + 1]
        cols = a.colidx[s:e]
        assert e - s == 10 and np.all(np.diff(cols) > 0)
        side = 0.4 * a.vals[s:e].sum()
        assert (a.lhs[i] == -1e20 and a.rhs[i] == side) or (a.lhs[i] == side and a.rhs[i] == 1e20)
    assert np.all((a.vals >= 1) & (a.vals <= 100) & (a.vals == np.round(a.vals)))

    m = sa.gen_mixed(3000, 3000, seed=2)
    assert np.all(np.abs(m.vals) >= 1e-3 - 1e-15) and np.all(np.abs(m.vals) <= 1e3 + 1e-9)
    lens = np.diff(m.rowptr)
    assert lens.min() >= 1 and 6 <= np.median(lens) <= 10
    it = m.integral.astype(bool)
    assert 0.35 < it.mean() < 0.45
    assert np.all(m.lb[it] == 0) and np.all((m.ub[it] >= 1) & (m.ub[it] <= 1000))
    cont = ~it
    inf_share = np.mean((m.lb[cont] <= -1e20) | (m.ub[cont] >= 1e20))
    assert 0.02 < inf_share < 0.09
    assert np.all(m.lhs <= m.rhs)


def test_node_fixings():
    inst = sa.gen_covering_packing(200, 500, 10, seed=3)
    ptr, cols, vals = sa.gen_node_fixings(inst, 32, 1, 50, seed=3)
    assert len(ptr) == 33 and ptr[0] == 0
    sizes = np.diff(ptr)
    assert sizes.min() >= 1 and sizes.max() <= 50
    for v in range(32):
        cs = cols[ptr[v]:ptr[v + 1]]
        assert len(set(cs.tolist())) == len(cs)
    assert set(np.unique(vals).tolist()) <= {0.0, 1.0}
