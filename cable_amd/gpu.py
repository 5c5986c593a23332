"""Python host side of the MI355X CABLE timestep (ctypes over libcable_gpu.so).

The product path: every call goes to the HIP library; when the library is
missing the import fails loudly (there is no CPU fallback).

``CableGPU`` mirrors the reference's step interface: the caller owns the
derived-type fields (here {Fortran name: numpy array}), attaches them, and
calls ``cbm(ktau, dels)`` once per step exactly as the offline driver calls
``cbm`` (src/offline/cable_serial.F90:594); ``run`` is the device-resident
multi-step path used for production runs and the benchmark.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .fields import BY_NAME, FIELDS, NFIELDS

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libcable_gpu.so")

PREC_FP64, PREC_MIXED = 0, 1
HOST_F64, HOST_F32, HOST_I32 = 0, 1, 2

EXPORTS = [
    "cable_gpu_nfields", "cable_gpu_field_info", "cable_gpu_field_id",
    "cable_gpu_create", "cable_gpu_destroy", "cable_gpu_upload",
    "cable_gpu_download", "cable_gpu_upload_fields", "cable_gpu_download_fields", "cable_gpu_set_globals",
    "cable_gpu_step", "cable_gpu_cbm",
    "cable_gpu_veg_cols", "cable_gpu_soil_cols", "cable_gpu_table_column",
    "cable_gpu_set_class_tables", "cable_gpu_set_forcing", "cable_gpu_run",
    "cable_gpu_sync", "cable_gpu_last_run_ms", "cable_gpu_last_kernel_ms", "cable_gpu_set_launch_timing",
    "cable_gpu_counters",
    "cable_gpu_set_tile_order",
    "cable_gpu_run_capture", "cable_gpu_samples", "cable_gpu_copy_samples", "cable_gpu_set_surrogate",
    "cable_gpu_run_hybrid", "cable_gpu_run_hybrid_noalloc", "cable_gpu_hybrid_counters", "cable_gpu_synth_forcing",
    "cable_gpu_download_forcing", "cable_gpu_surrogate_eval", "cable_gpu_consistency_check",
    "cable_gpu_consistency_values", "cable_gpu_state_bytes", "cable_gpu_save_state", "cable_gpu_load_state",
    "cable_gpu_set_met", "cable_gpu_met_clock",
    "cable_gpu_last_error",
]

_lib = None


class CableError(RuntimeError):
    pass


class MetVars(ctypes.Structure):
    """cable_gpu_met (include/cable_gpu.h)."""
    _fields_ = [(n, ctypes.c_void_p) for n in
                ("swdown", "lwdown", "tair", "qair", "wind", "wind_e", "psurf", "rainf")] + [
        ("tair_offset", ctypes.c_double), ("psurf_scale", ctypes.c_double), ("rainf_scale", ctypes.c_double),
        ("qair_is_rh", ctypes.c_int), ("elevation", ctypes.c_void_p), ("latitude", ctypes.c_void_p),
        ("shod", ctypes.c_double), ("sdoy", ctypes.c_int), ("smoy", ctypes.c_int), ("syear", ctypes.c_int),
        ("fixed_co2", ctypes.c_double)]


def lib():
    """Load libcable_gpu.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise CableError(f"{LIBPATH} is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(LIBPATH)
    vp, i64, c_int, c_double = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double
    sig = {
        "cable_gpu_nfields": ([], c_int),
        "cable_gpu_field_info": ([c_int, ctypes.POINTER(ctypes.c_char_p), ctypes.c_char_p,
                                  ctypes.POINTER(c_int), ctypes.POINTER(c_int)], c_int),
        "cable_gpu_field_id": ([ctypes.c_char_p], c_int),
        "cable_gpu_create": ([ctypes.POINTER(vp), i64, c_int], c_int),
        "cable_gpu_destroy": ([vp], c_int),
        "cable_gpu_upload": ([vp, c_int, vp, c_int], c_int),
        "cable_gpu_download": ([vp, c_int, vp, c_int], c_int),
        "cable_gpu_upload_fields": ([vp, c_int, vp, vp, vp], c_int),
        "cable_gpu_download_fields": ([vp, c_int, vp, vp, vp], c_int),
        "cable_gpu_set_globals": ([vp, vp, vp, vp, vp], c_int),
        "cable_gpu_step": ([vp, c_int, c_double], c_int),
        "cable_gpu_cbm": ([vp, c_int, c_double], c_int),
        "cable_gpu_veg_cols": ([], c_int),
        "cable_gpu_soil_cols": ([], c_int),
        "cable_gpu_table_column": ([c_int, c_int], c_int),
        "cable_gpu_set_class_tables": ([vp, vp, vp], c_int),
        "cable_gpu_set_forcing": ([vp, c_int] + [vp] * 8 + [c_int, c_double, c_double], c_int),
        "cable_gpu_run": ([vp, c_int, c_int, c_double], c_int),
        "cable_gpu_sync": ([vp], c_int),
        "cable_gpu_last_run_ms": ([vp, ctypes.POINTER(c_double)], c_int),
        "cable_gpu_last_kernel_ms": ([vp, ctypes.POINTER(c_double), ctypes.POINTER(c_double)], c_int),
        "cable_gpu_set_launch_timing": ([vp, c_int], c_int),
        "cable_gpu_counters": ([vp, ctypes.POINTER(i64)], c_int),
        "cable_gpu_last_error": ([], ctypes.c_char_p),
        "cable_gpu_set_tile_order": ([vp, vp], c_int),
        "cable_gpu_run_capture": ([vp, c_int, c_int, c_double, i64], c_int),
        "cable_gpu_samples": ([vp, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(i64)], c_int),
        "cable_gpu_copy_samples": ([vp, i64, i64, vp, vp], c_int),
        "cable_gpu_set_surrogate": ([vp, vp, vp], c_int),
        "cable_gpu_run_hybrid": ([vp, c_int, c_int, c_double, c_int, c_double, c_int], c_int),
        "cable_gpu_run_hybrid_noalloc": ([vp, c_int, c_int, c_double, c_int, c_double, c_int], c_int),
        "cable_gpu_hybrid_counters": ([vp, ctypes.POINTER(i64), ctypes.POINTER(i64)], c_int),
        "cable_gpu_synth_forcing": ([vp, c_int, c_int, ctypes.c_uint64, c_double, c_double, c_double, vp, vp],
                                    c_int),
        "cable_gpu_download_forcing": ([vp, c_int, c_int, vp, c_int], c_int),
        "cable_gpu_surrogate_eval": ([vp, c_int, vp, vp], c_int),
        "cable_gpu_consistency_check": ([vp, c_int], c_int),
        "cable_gpu_state_bytes": ([vp], i64),
        "cable_gpu_set_met": ([vp, c_int, c_int, c_double, ctypes.POINTER(MetVars)], c_int),
        "cable_gpu_met_clock": ([c_int, c_double, c_int, c_int, c_int, c_double, c_int, c_int, vp, vp], c_int),
        "cable_gpu_save_state": ([vp, vp, i64], c_int),
        "cable_gpu_load_state": ([vp, vp, i64, c_int], c_int),
        "cable_gpu_consistency_values": ([vp, vp], c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    if L.cable_gpu_nfields() != NFIELDS:
        raise CableError("libcable_gpu.so was built from a different field table")
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise CableError(lib().cable_gpu_last_error().decode())


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def class_tables(vtab: dict, stab: dict):
    """Pack the workload's per-class tables into the library's column layout."""
    L = lib()
    nv, ns = L.cable_gpu_veg_cols(), L.cable_gpu_soil_cols()
    V = np.zeros((17, nv))
    S = np.zeros((9, ns))
    vmap = {"veg%canst1": "canst1", "veg%dleaf": "dleaf", "veg%ejmax": "ejmax", "veg%frac4": "frac4",
            "veg%hc": "hc", "veg%rp20": "rp20", "veg%shelrb": "shelrb", "veg%vegcf": "vegcf",
            "veg%vbeta": "vbeta", "veg%vcmax": "vcmax", "veg%xfang": "xfang", "veg%extkn": "extkn",
            "veg%a1gs": "a1gs", "veg%d0gs": "d0gs", "veg%alpha": "alpha", "veg%convex": "convex",
            "veg%cfrd": "cfrd", "veg%gswmin": "gswmin", "veg%conkc0": "conkc0", "veg%conko0": "conko0",
            "veg%ekc": "ekc", "veg%eko": "eko"}
    for fname, key in vmap.items():
        V[:, L.cable_gpu_table_column(0, BY_NAME[fname].fid)] = vtab[key]
    V[:, L.cable_gpu_table_column(0, BY_NAME["veg%meth"].fid)] = 1.0
    c = L.cable_gpu_table_column(0, BY_NAME["veg%refl"].fid)
    V[:, c:c + 2] = vtab["refl"][:2].T
    c = L.cable_gpu_table_column(0, BY_NAME["veg%taul"].fid)
    V[:, c:c + 2] = vtab["taul"][:2].T
    c = L.cable_gpu_table_column(0, BY_NAME["veg%froot"].fid)
    V[:, c:c + 6] = vtab["froot"].T
    V[:, L.cable_gpu_table_column(0, BY_NAME["rough%za_uv"].fid)] = vtab.get("za_uv", 40.0)
    V[:, L.cable_gpu_table_column(0, BY_NAME["rough%za_tq"].fid)] = vtab.get("za_tq", 40.0)
    for key in ("css", "hsbh", "hyds", "i2bp3", "ibp2", "rhosoil", "sfc", "ssat", "swilt", "cnsd", "pwb_min"):
        S[:, L.cable_gpu_table_column(1, BY_NAME[f"soil%{key}"].fid)] = stab[key]
    for key in ("sfc", "swilt"):
        c = L.cable_gpu_table_column(1, BY_NAME[f"soil%{key}_vec"].fid)
        S[:, c:c + 6] = np.tile(stab[key][:, None], (1, 6))
    c = L.cable_gpu_table_column(1, BY_NAME["soil%zse_vec"].fid)
    S[:, c:c + 6] = np.array(stab["zse"])[None, :]
    return np.ascontiguousarray(V), np.ascontiguousarray(S)


class CableGPU:
    """Device-resident state of ``mp`` tiles."""

    def __init__(self, mp: int, precision: int = PREC_FP64):
        self.mp = int(mp)
        self.precision = precision
        h = ctypes.c_void_p()
        _check(lib().cable_gpu_create(ctypes.byref(h), self.mp, int(precision)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().cable_gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- fields ------------------------------------------------------------
    def upload(self, name: str, arr):
        f = BY_NAME[name]
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
        if a.size != f.n * self.mp:
            raise CableError(f"{name}: {a.size} elements, expected {f.n * self.mp}")
        _check(lib().cable_gpu_upload(self.h, f.fid, _ptr(a), HOST_F64))

    def download(self, name: str):
        f = BY_NAME[name]
        out = np.empty(f.shape(self.mp), dtype=np.float64)
        _check(lib().cable_gpu_download(self.h, f.fid, _ptr(out), HOST_F64))
        return out

    def upload_fields(self, fields: dict):
        """Several fields in one batched transfer (cable_gpu_upload_fields);
        fields: {name: array} (float64)."""
        names = [k for k in fields if k in BY_NAME]
        arrs = []
        for k in names:
            a = np.ascontiguousarray(fields[k], dtype=np.float64).reshape(-1)
            if a.size != BY_NAME[k].n * self.mp:
                raise CableError(f"{k}: {a.size} elements, expected {BY_NAME[k].n * self.mp}")
            arrs.append(a)
        self._batched(lib().cable_gpu_upload_fields, names, arrs)

    def download_fields(self, names, out=None):
        """Several fields in one batched transfer (cable_gpu_download_fields):
        {name: float64 array}; out (optional) supplies the arrays to fill."""
        out = {} if out is None else out
        for k in names:
            if k not in out:
                out[k] = np.empty(BY_NAME[k].shape(self.mp), dtype=np.float64)
        self._batched(lib().cable_gpu_download_fields, list(names), [out[k] for k in names])
        return out

    def _batched(self, fn, names, arrs):
        n = len(names)
        ids = np.array([BY_NAME[k].fid for k in names], dtype=np.int32)
        ptrs = (ctypes.c_void_p * max(n, 1))(*[_ptr(a) for a in arrs])
        types = np.full(n, HOST_F64, dtype=np.int32)
        _check(fn(self.h, n, _ptr(ids), ptrs, _ptr(types)))

    def upload_all(self, fields: dict):
        for k, v in fields.items():
            if k in BY_NAME:
                self.upload(k, v)

    def set_tile_order(self, order):
        """Device position j holds tile order[j] (see class_order)."""
        o = np.ascontiguousarray(order, dtype=np.int64)
        if o.shape != (self.mp,):
            raise CableError(f"tile order of {o.shape} for {self.mp} tiles")
        _check(lib().cable_gpu_set_tile_order(self.h, _ptr(o)))

    def set_globals(self, zse, zshh, ratecp, ratecs):
        arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (zse, zshh, ratecp, ratecs)]
        self._globals = arrs
        _check(lib().cable_gpu_set_globals(self.h, *[_ptr(a) for a in arrs]))

    # -- one step, every diagnostic ------------------------------------------
    def step(self, ktau: int, dels: float):
        """One offline-driver step: cbm with the driver's contract around it
        and the balance checks (cable_serial.F90:572-605, 725-729)."""
        _check(lib().cable_gpu_step(self.h, int(ktau), float(dels)))

    def cbm(self, ktau: int, dels: float):
        """cbm alone, as the Fortran drop-in calls it."""
        _check(lib().cable_gpu_cbm(self.h, int(ktau), float(dels)))

    def dryleaf_passes(self):
        v = ctypes.c_int64()
        _check(lib().cable_gpu_counters(self.h, ctypes.byref(v)))
        return v.value

    # -- device-resident run -----------------------------------------------
    def set_class_tables(self, V, S):
        self._tabs = (np.ascontiguousarray(V, dtype=np.float64), np.ascontiguousarray(S, dtype=np.float64))
        _check(lib().cable_gpu_set_class_tables(self.h, _ptr(self._tabs[0]), _ptr(self._tabs[1])))

    def set_forcing(self, forcing: dict, ca=400.0e-6, doy=172.0):
        """forcing: {met%fsd (ns,2,mp) or fsd1 (ns,mp), met%fld, met%tk, met%qv, met%ua,
        met%pmb, met%precip, met%coszen}, each (nsteps, mp)."""
        fsd = forcing["met%fsd"]
        fsd1 = fsd[:, 0, :] if fsd.ndim == 3 else fsd
        arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in
                (fsd1, forcing["met%fld"], forcing["met%tk"], forcing["met%qv"], forcing["met%ua"],
                 forcing["met%pmb"], forcing["met%precip"], forcing["met%coszen"])]
        ns = arrs[0].shape[0]
        _check(lib().cable_gpu_set_forcing(self.h, ns, *[_ptr(a) for a in arrs], HOST_F64,
                                           float(ca), float(doy)))
        self.nsteps = ns

    def run(self, k0: int, nsteps: int, dels: float):
        _check(lib().cable_gpu_run(self.h, int(k0), int(nsteps), float(dels)))

    def synth_forcing(self, k_first: int, nsteps: int, seed: int, t0, lat, dels=1800.0, doy=172.0, hod0=0.0):
        """Generate the synthetic forcing of steps k_first.. on the device
        (workload.synth_forcing is the same generator on the host)."""
        t0 = np.ascontiguousarray(t0, dtype=np.float64)
        lat = np.ascontiguousarray(lat, dtype=np.float64)
        if t0.shape != (self.mp,) or lat.shape != (self.mp,):
            raise CableError("t0 and lat need one value per tile")
        _check(lib().cable_gpu_synth_forcing(self.h, int(k_first), int(nsteps), int(seed) & (2 ** 64 - 1),
                                             float(dels), float(doy), float(hod0), _ptr(t0), _ptr(lat)))
        self.nsteps = int(nsteps)

    def set_met(self, k_first: int, nsteps: int, dels: float, met: dict, latitude, elevation=None,
                start=(0.0, 1, 1, 2000), fixed_co2=400.0):
        """Resident forcing from met variables in the file's units
        (cable_gpu_set_met).  met: swdown, tair, qair, wind, rainf and
        optionally lwdown, wind_e, psurf, each (nsteps, mp), plus the
        conversions tair_offset, psurf_scale, rainf_scale, qair_is_rh;
        start = (shod, sdoy, smoy, syear) of the run's first step."""
        keep = []

        def arr(x, dt):
            if x is None:
                return None
            a = np.ascontiguousarray(x, dtype=dt)
            keep.append(a)
            return a.ctypes.data
        m = MetVars()
        for k in ("swdown", "lwdown", "tair", "qair", "wind", "wind_e", "psurf", "rainf"):
            v = met.get(k)
            if v is not None and np.shape(v) != (nsteps, self.mp):
                raise CableError(f"{k}: shape {np.shape(v)}, expected {(nsteps, self.mp)}")
            setattr(m, k, arr(v, np.float32))
        m.tair_offset = float(met.get("tair_offset", 0.0))
        m.psurf_scale = float(met.get("psurf_scale", 0.01))
        m.rainf_scale = float(met.get("rainf_scale", dels))
        m.qair_is_rh = 1 if met.get("qair_is_rh") else 0
        m.elevation = arr(elevation, np.float64)
        m.latitude = arr(latitude, np.float64)
        m.shod, m.sdoy, m.smoy, m.syear = float(start[0]), int(start[1]), int(start[2]), int(start[3])
        m.fixed_co2 = float(fixed_co2)
        _check(lib().cable_gpu_set_met(self.h, int(k_first), int(nsteps), float(dels), ctypes.byref(m)))
        self.nsteps = int(nsteps)

    FORCING = ["met%fsd", "met%fld", "met%tk", "met%qv", "met%ua", "met%pmb", "met%precip", "met%coszen"]

    def download_forcing(self, k: int):
        """The resident forcing of step k as {name: array(mp)} (met%fsd: band 1)."""
        out = {}
        for i, name in enumerate(self.FORCING):
            a = np.empty(self.mp)
            _check(lib().cable_gpu_download_forcing(self.h, int(k), i, _ptr(a), HOST_F64))
            out[name] = a
        return out

    # -- surrogate (mixed contexts) --------------------------------------------
    def run_capture(self, k0: int, nsteps: int, dels: float, capacity: int):
        """Resident steps that also record one training sample per vegetated
        tile-step; returns the number of samples held."""
        _check(lib().cable_gpu_run_capture(self.h, int(k0), int(nsteps), float(dels), int(capacity)))
        return self.samples()[2]

    def samples(self):
        """(x device pointer, y device pointer, n) of the captured samples."""
        x, y, n = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_int64()
        _check(lib().cable_gpu_samples(self.h, ctypes.byref(x), ctypes.byref(y), ctypes.byref(n)))
        return x.value, y.value, n.value

    def copy_samples(self, first: int, n: int):
        x = np.empty((n, 24), dtype=np.float32)
        y = np.empty((n, 8), dtype=np.float32)
        _check(lib().cable_gpu_copy_samples(self.h, int(first), int(n), _ptr(x), _ptr(y)))
        return x, y

    def set_surrogate(self, params, norm):
        self._sur = (np.ascontiguousarray(params, dtype=np.float32), np.ascontiguousarray(norm, dtype=np.float32))
        _check(lib().cable_gpu_set_surrogate(self.h, _ptr(self._sur[0]), _ptr(self._sur[1])))

    def run_hybrid(self, k0: int, nsteps: int, dels: float, max_passes=2, resid_max=0.5, use_surrogate=True):
        _check(lib().cable_gpu_run_hybrid(self.h, int(k0), int(nsteps), float(dels), int(max_passes),
                                          float(resid_max), 1 if use_surrogate else 0))

    def run_hybrid_noalloc(self, k0: int, nsteps: int, dels: float, max_passes=2, resid_max=0.5,
                           use_surrogate=True):
        """run_hybrid through the entry point that allocates nothing (at most
        1,048,576 tiles; the surrogate must be set first)."""
        _check(lib().cable_gpu_run_hybrid_noalloc(self.h, k0, nsteps, dels, max_passes, resid_max,
                                                  1 if use_surrogate else 0))

    def surrogate_eval(self, k: int):
        """(raw features (mp, 24), predictions (mp, 8)) of step k from the
        current state, as the hybrid step's inference computes them."""
        x = np.empty((self.mp, 24), dtype=np.float32)
        p = np.empty((self.mp, 8), dtype=np.float32)
        _check(lib().cable_gpu_surrogate_eval(self.h, int(k), _ptr(x), _ptr(p)))
        return x, p

    def hybrid_counters(self):
        """(fallback tile-steps, tile-steps) of the last hybrid run."""
        fb, n = ctypes.c_int64(), ctypes.c_int64()
        _check(lib().cable_gpu_hybrid_counters(self.h, ctypes.byref(fb), ctypes.byref(n)))
        return fb.value, n.value

    def last_run_ms(self):
        v = ctypes.c_double()
        _check(lib().cable_gpu_last_run_ms(self.h, ctypes.byref(v)))
        return v.value

    def last_kernel_ms(self):
        """(step kernels, inference kernels) device ms of the last run."""
        a, b = ctypes.c_double(), ctypes.c_double()
        _check(lib().cable_gpu_last_kernel_ms(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def set_launch_timing(self, on):
        """Per-launch events in hybrid runs (on by default; they cost the run
        time, see include/cable_gpu.h)."""
        _check(lib().cable_gpu_set_launch_timing(self.h, 1 if on else 0))

    def sync(self):
        _check(lib().cable_gpu_sync(self.h))

    def save_state(self) -> bytes:
        """The device state (every STATE and DIAG field) as one buffer."""
        n = lib().cable_gpu_state_bytes(self.h)
        if n < 0:
            _check(-n)
        buf = ctypes.create_string_buffer(n)
        _check(lib().cable_gpu_save_state(self.h, buf, n))
        return buf.raw

    def load_state(self, state: bytes, continue_run=True):
        """Load a saved state; continue_run: the next step continues the
        saved run exactly (else it is the first step of a new run)."""
        buf = ctypes.create_string_buffer(state, len(state))
        _check(lib().cable_gpu_load_state(self.h, buf, len(state), 1 if continue_run else 0))

    def consistency_check(self, enable=True):
        """Start (zeroing the sums) or stop the driver's consistency sums."""
        _check(lib().cable_gpu_consistency_check(self.h, 1 if enable else 0))

    def consistency_values(self):
        """{count_bal, new_sumbal, sum_fe, sum_fpn, ebal_tot, wbal_tot} as the
        offline driver forms them (cable_serial.F90:748-771)."""
        out = np.zeros(6)
        _check(lib().cable_gpu_consistency_values(self.h, _ptr(out)))
        return dict(zip(["count_bal", "new_sumbal", "sum_fe", "sum_fpn", "ebal_tot", "wbal_tot"], out.tolist()))


def met_clock(precision, start, dels, k_first, nsteps):
    """(hod, doy) of steps k_first.. on the offline reader's clock, as the
    library computes it (no device work)."""
    hod, doy = np.empty(nsteps), np.empty(nsteps)
    _check(lib().cable_gpu_met_clock(int(precision), float(start[0]), int(start[1]), int(start[2]),
                                     int(start[3]), float(dels), int(k_first), int(nsteps), _ptr(hod), _ptr(doy)))
    return hod, doy


def class_order(iveg, isoilm=None, vlai=None, cost=None, tair=None):
    """A tile order that groups tiles by vegetation type (then soil type, then
    leaf area): the wavefronts then run one type's canopy iteration counts and
    branches, and tiles of similar leaf area side by side.  cost: optional
    {vegetation type: time per tile-step} (class_costs); the types then come
    most expensive first, so that the last workgroups the dispatcher starts
    are the short ones (longest-first scheduling).  tair: optional per-tile
    characteristic air temperature (K); within a type, tiles then go in 4 K
    bands of it before soil type, so that a wavefront's tiles share a snow
    regime (snowfall at or below 273.16 K, snow layers, frozen soil) --
    measured 2.7% faster in mixed precision and 5% in FP64 at 1,048,576
    tiles."""
    iveg = np.asarray(iveg)
    if isoilm is None:
        return np.argsort(iveg, kind="stable")
    if vlai is None:
        return np.lexsort((np.asarray(isoilm), iveg))
    keys = [np.asarray(vlai), np.asarray(isoilm)]
    if tair is not None:
        keys.append(np.floor(np.asarray(tair, dtype=np.float64) / 4.0).astype(np.int64))
    keys.append(iveg)
    if cost is not None:
        rank = {c: r for r, c in enumerate(sorted(cost, key=lambda c: (-cost[c], c)))}
        keys.append(np.array([rank.get(int(c), len(rank)) for c in iveg]))
    return np.lexsort(tuple(keys))


def class_costs(wl, precision=PREC_FP64, sample=4096, steps=8):
    """Measured device time per tile-step of each vegetation type of workload
    wl: the resident run on up to `sample` tiles of the type for `steps`
    steps (a calibration for class_order)."""
    out = {}
    for c in np.unique(wl.iveg):
        idx = np.nonzero(wl.iveg == c)[0][:sample]
        sub = wl.subset(idx)
        nst = min(steps, sub.nsteps)
        g = from_workload(sub, precision=precision, full=False)
        g.run(0, nst, sub.dels)
        out[int(c)] = g.last_run_ms() / (nst * sub.n)
        g.close()
    return out


def from_workload(wl, precision=PREC_FP64, full=True, order=None):
    """A CableGPU holding workload ``wl``'s parameters and initial state (every
    other field zero, as the reference's derived types start); ``order``: the
    device tile order (class_order), or None for the caller's order."""
    g = CableGPU(wl.n, precision)
    if order is not None:
        g.set_tile_order(order)
    g.upload_all(wl.params)
    g.upload_all(wl.state0)
    p = wl.params
    g.set_globals(p["soil%zse"], p["soil%zshh"], p["bgc%ratecp"], p["bgc%ratecs"])
    if not full:
        stab = dict(wl.stab)
        stab["zse"] = p["soil%zse"]
        g.set_class_tables(*class_tables(wl.vtab, stab))
        if getattr(wl, "forcing", None) is not None:   # else: synth_forcing
            g.set_forcing(wl.forcing)
    return g
