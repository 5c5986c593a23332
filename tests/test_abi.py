"""The C ABI library without a GPU: it loads, exports every function that
include/cable_gpu.h declares, and its field and table metadata agree with
the Python mirror.  Nothing here launches device work."""
import ctypes
import os
import re

import numpy as np
import pytest

from cable_amd import gpu
from cable_amd.fields import FIELDS, NFIELDS

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cable_gpu.h")


@pytest.fixture(scope="module")
def L():
    if not os.path.exists(gpu.LIBPATH):
        from cable_amd import build
        build.build_gpu()
    return gpu.lib()


def declared(header=HEADER, prefix="cable_gpu_"):
    txt = open(header).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(" + prefix + r"\w+)\s*\(", txt)))


def test_every_declared_function_is_exported(L):
    names = declared()
    assert len(names) >= 19
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(gpu.EXPORTS) == names


def test_surrogate_functions_are_exported(L):
    from cable_amd import surrogate
    names = declared(os.path.join(ROOT, "include", "cable_mlp.h"), "cable_mlp_")
    assert len(names) >= 13
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert surrogate.lib().cable_mlp_nparam() == surrogate.NPARAM == 37256


def test_field_table_matches_python_mirror(L):
    assert L.cable_gpu_nfields() == NFIELDS
    nm = ctypes.c_char_p()
    pr = ctypes.create_string_buffer(1)
    n = ctypes.c_int()
    role = ctypes.c_int()
    roles = ["PARAM", "FORC", "STATE", "DIAG"]
    for f in FIELDS:
        assert L.cable_gpu_field_info(f.fid, ctypes.byref(nm), pr, ctypes.byref(n), ctypes.byref(role)) == 0
        assert nm.value.decode() == f.name
        assert pr.value.decode() == f.prec
        assert n.value == f.n and roles[role.value] == f.role
        assert L.cable_gpu_field_id(f.name.encode()) == f.fid
    assert L.cable_gpu_field_id(b"canopy%nonexistent") == -1
    assert L.cable_gpu_field_info(NFIELDS, ctypes.byref(nm), pr, ctypes.byref(n), ctypes.byref(role)) != 0
    assert b"field" in L.cable_gpu_last_error()


def test_class_table_layout(L):
    from cable_amd.workload import Workload
    wl = Workload(34, 1)
    stab = dict(wl.stab)
    stab["zse"] = wl.params["soil%zse"]
    V, S = gpu.class_tables(wl.vtab, stab)
    assert V.shape == (17, L.cable_gpu_veg_cols()) and S.shape == (9, L.cable_gpu_soil_cols())
    # every tabulated per-tile parameter of the workload equals its class row
    for f in FIELDS:
        for table, T, cls in ((0, V, wl.iveg), (1, S, wl.isoilm)):
            c = L.cable_gpu_table_column(table, f.fid)
            if c < 0 or f.name not in wl.params:
                continue
            per_tile = np.asarray(wl.params[f.name]).reshape(f.n, -1)
            np.testing.assert_array_equal(T[cls - 1, c:c + f.n].T, per_tile, err_msg=f.name)


def test_errors_without_a_context(L):
    assert L.cable_gpu_step(None, 1, 1800.0) == 1
    assert b"null" in L.cable_gpu_last_error()
    assert L.cable_gpu_run(None, 0, 1, 1800.0) == 1


def test_new_entry_points_reject_bad_arguments(L):
    """Argument checks of the forcing generator, the consistency check and
    the surrogate evaluation, before any device work."""
    import ctypes
    d = ctypes.c_double
    assert L.cable_gpu_synth_forcing(None, 0, 1, 1, d(1800.0), d(172.0), d(0.0), None, None) == 1
    assert L.cable_gpu_download_forcing(None, 0, 0, None, 0) == 1
    assert L.cable_gpu_consistency_check(None, 1) == 1
    assert L.cable_gpu_consistency_values(None, None) == 1
    assert L.cable_gpu_surrogate_eval(None, 0, None, None) == 1
    assert L.cable_gpu_run_hybrid(None, 0, 1, d(1800.0), 4, d(0.5), 1) == 1
    assert L.cable_gpu_create(ctypes.byref(ctypes.c_void_p()), 0, 0) == 1
    assert b"bad arguments" in L.cable_gpu_last_error()
    assert L.cable_gpu_create(ctypes.byref(ctypes.c_void_p()), 16, 7) == 1
    assert b"precision" in L.cable_gpu_last_error()


def test_class_order_is_a_permutation_grouped_by_type():
    from cable_amd.workload import Workload
    wl = Workload(3000, 1, forcing=False)
    cost = {c: float(c % 4) for c in range(1, 18)}
    for o in (gpu.class_order(wl.iveg, wl.isoilm, wl.vlai), gpu.class_order(wl.iveg, wl.isoilm, wl.vlai, cost)):
        assert sorted(o.tolist()) == list(range(wl.n))
        v = wl.iveg[o]
        # each vegetation type is one contiguous run
        runs = v[np.r_[True, v[1:] != v[:-1]]]
        assert len(runs) == len(set(runs.tolist()))
    # with costs, the types come most expensive first (ties by type)
    v = wl.iveg[gpu.class_order(wl.iveg, wl.isoilm, wl.vlai, cost)]
    runs = v[np.r_[True, v[1:] != v[:-1]]].tolist()
    assert runs == sorted(runs, key=lambda c: (-cost[c], c))
    # with a characteristic air temperature, a type's tiles go in 4 K bands
    # of it, then soil type, then leaf area
    o = gpu.class_order(wl.iveg, wl.isoilm, wl.vlai, cost, tair=wl.t0)
    assert sorted(o.tolist()) == list(range(wl.n))
    v, band, soil, lai = wl.iveg[o], np.floor(wl.t0[o] / 4.0), wl.isoilm[o], wl.vlai[o]
    for c in np.unique(v):
        m = v == c
        key = list(zip(band[m], soil[m], lai[m]))
        assert key == sorted(key)
