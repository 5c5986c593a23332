"""Synthetic CABLE workload: per-tile parameters, initial state and forcing.

The generator follows the offline driver's recipes so that every implementation
(the compiled reference, the CPU restatement and the GPU kernels) is fed the
same FP64 numbers:

* parameters from the default tables (src/offline/pft_params.nml,
  src/offline/cable_soilparm.nml) through the formulas of
  ``init_veg_from_vegin`` (src/offline/cable_parameters.F90:3277-3345),
  ``write_default_params`` (:1129-1693), ``consistency_ice_veg_soil``
  (:2467-2650) and ``derived_parameters`` (:1759-2343);
* forcing as ``get_met_data`` forms it (src/offline/cable_input.F90:1880-1883
  SW split, :2666-2672 snow partition, :2677 coszen via ``sinbet``
  src/science/radiation/cbl_sinbet.F90:8-28, :2679-2680 tvair = tvrad = tk)
  with specific humidity from relative humidity by ``rh_sh``/``svp``
  (src/offline/cable_checks.F90:633-696).

Random draws come from counter-based Philox streams, one per quantity, drawn
tile-major, so the first N tiles of a larger workload are the N-tile workload.

Array convention: a Fortran field ``x(mp, k1, k2)`` is a C-ordered numpy array
of shape ``(k2, k1, mp)`` (same bytes, tile index fastest).
"""
from __future__ import annotations

import copy
import math

import numpy as np

from . import tables as T

MS = 6        # soil layers
MSN = 3       # snow layers
NRB = 3       # radiation bands
MF = 2        # leaves
NITER = 4     # stability iterations
SWB = 2       # shortwave bands
TFRZ = 273.16
PI_F = 3.1415927                  # src/params/cable_maths_constants_mod.F90:32
PI180_F = PI_F / 180.0

DEFAULT_SEED = 0xCAB1E2026
DOY = 172.0
DELS = 1800.0

# stream ids for the Philox generator
_S_VEG, _S_SOIL, _S_LAI, _S_LAT, _S_T0, _S_ALB, _S_WB = range(1, 8)
_F_SW, _F_LW, _F_RH, _F_UA, _F_PB, _F_PA, _F_PS = range(20, 27)


def _rng(seed: int, stream: int) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(key=(int(seed) << 8) | stream))


def _uniform(seed, stream, shape):
    """U(0,1) drawn tile-major: shape[-1] is the tile axis."""
    n = shape[-1]
    rest = int(np.prod(shape[:-1])) if len(shape) > 1 else 1
    x = _rng(seed, stream).random((n, rest))
    return np.ascontiguousarray(x.T.reshape(shape))


def svp(tk):
    """Saturation vapour pressure (hPa), src/offline/cable_checks.F90:656-696."""
    tk = np.asarray(tk, dtype=np.float64)
    temp = tk - 273.15
    out = np.empty_like(tk)
    lo = temp < -20.0
    log10 = math.log(10.0)
    toot = 273.15 / tk[lo]
    toto = 1.0 / toot
    eilog = (-9.09718 * (toot - 1) - 3.56654 * (np.log(toot) / log10)
             + 0.876793 * (1 - toto) + (math.log(6.1071) / log10))
    out[lo] = np.power(10.0, eilog)
    hi = ~lo
    tsot = 373.15 / tk[hi]
    ewlog = -7.90298 * (tsot - 1) + 5.02808 * (np.log(tsot) / log10)
    ewlog2 = ewlog - 1.3816e-07 * (np.power(10.0, 11.344 * (1 - (1 / tsot))) - 1)
    ewlog3 = ewlog2 + 0.0081328 * (np.power(10.0, -3.49149 * (tsot - 1)) - 1)
    ewlog4 = ewlog3 + (math.log(1013.246) / log10)
    out[hi] = np.power(10.0, ewlog4)
    return out


def rh_sh(rel_hum_pct, tk, psurf_hpa):
    """Specific humidity from relative humidity, cable_checks.F90:633-651."""
    es = svp(tk)
    ws = 0.622 * es / (psurf_hpa - es)
    return (rel_hum_pct / 100.0) * ws


def sinbet(doy, xslat, hod):
    """cos(solar zenith), src/science/radiation/cbl_sinbet.F90:8-28."""
    sindec = -math.sin(23.45 * PI180_F) * math.cos(2.0 * PI_F * (doy + 10.0) / 365.0)
    z = (np.sin(PI180_F * xslat) * sindec
         + np.cos(PI180_F * xslat) * math.sqrt(1.0 - sindec * sindec)
         * np.cos(PI_F * (hod - 12.0) / 12.0))
    return np.maximum(z, 1e-8)


def tile_classes(n: int, seed: int = DEFAULT_SEED):
    """PFT and soil index per tile: stratified over the 17 x 9 table entries,
    then made consistent for ice (iveg 17 <=> isoilm 9)."""
    t = np.arange(n)
    iveg = 1 + (t % T.NVEG)
    isoilm = 1 + ((t // T.NVEG) % T.NSOIL)
    # consistency_ice_veg_soil, src/offline/cable_parameters.F90:2527-2549
    isoilm = np.where(iveg == T.ICE_VEG, T.ICE_SOIL, isoilm)
    iveg = np.where(isoilm == T.ICE_SOIL, T.ICE_VEG, iveg)
    return iveg.astype(np.int32), isoilm.astype(np.int32)


def fortran_nint(x):
    """Fortran NINT: the nearest integer, halves rounded away from zero."""
    x = np.asarray(x, dtype=np.float64)
    return np.floor(np.abs(x) + 0.5) * np.sign(x)


def froot_from_rootbeta(rootbeta, zse):
    """Root fraction per soil layer from the PFT's rootbeta (Jackson et al.
    1996): the cumulative fraction 1 - rootbeta**d at the bottom of layers
    1..5 (d in cm), the last layer the rest, then differenced
    (init_veg_from_vegin, src/offline/cable_parameters.F90:3333-3342).
    Returns (6, n)."""
    rootbeta = np.atleast_1d(np.asarray(rootbeta, dtype=np.float64))
    froot = np.zeros((MS, rootbeta.size))
    totdepth = 0.0
    for k in range(MS - 1):
        totdepth = totdepth + zse[k] * 100.0
        froot[k] = np.minimum(1.0, 1.0 - np.power(rootbeta, totdepth))
    froot[MS - 1] = 1.0 - froot[MS - 2]
    for k in range(MS - 2, 0, -1):
        froot[k] = froot[k] - froot[k - 1]
    return froot


def initial_owetfac(wb1, wbice1, swilt, sfc):
    """The surface wetness factor's starting value: the top layer's water
    between wilting point and field capacity, clamped to [0, 1], reduced by
    (1 - ice fraction)**2 where the layer holds ice (derived_parameters,
    src/offline/cable_parameters.F90:2245-2254)."""
    wb1, wbice1 = np.asarray(wb1, dtype=np.float64), np.asarray(wbice1, dtype=np.float64)
    o = np.maximum(0.0, np.minimum(1.0, (wb1 - swilt) / (sfc - swilt)))
    ice = wbice1 > 0.0
    frac = np.where(ice, wbice1 / np.where(ice, wb1, 1.0), 0.0)
    return np.where(ice, o * (1.0 - frac) ** 2, o)


def veg_class_table():
    """Per-PFT vegetation parameters as cbm sees them (after
    init_veg_from_vegin and write_default_params).  Returns {name: array(17)}
    plus froot (6,17), refl/taul (3,17)."""
    V = {k: np.array(v, dtype=np.float64) for k, v in T.VEG.items()}
    out = {}
    for k in ("canst1", "frac4", "vcmax", "hc", "xfang", "vbeta", "xalbnir",
              "rp20", "rpcoef", "rs20", "shelrb", "wai", "a1gs", "d0gs",
              "vegcf", "extkn", "tminvj", "tmaxvj", "g0", "g1", "alpha",
              "convex", "cfrd", "gswmin", "conkc0", "conko0", "ekc", "eko",
              "rootbeta", "zr", "clitt"):
        out[k] = V[k].copy()
    # cable_pft_params.F90:110
    out["dleaf"] = np.sqrt(V["width"] * V["length"])
    # cable_parameters.F90:1543 (ejmax = 2 vcmax after init_veg_from_vegin)
    out["ejmax"] = 2.0 * V["vcmax"]
    out["refl"] = np.stack([V["refl1"], V["refl2"], V["refl3"]])
    out["taul"] = np.stack([V["taul1"], V["taul2"], V["taul3"]])
    out["froot"] = froot_from_rootbeta(V["rootbeta"], T.ZSE)
    out["cplant"] = np.stack([V["cplant1"], V["cplant2"], V["cplant3"]])
    out["csoil"] = np.stack([V["csoil1"], V["csoil2"]])
    out["ratecp"] = np.array([V["ratecp1"][0], V["ratecp2"][0], V["ratecp3"][0]])
    out["ratecs"] = np.array([V["ratecs1"][0], V["ratecs2"][0]])
    return out


def soil_class_table():
    """Per-soil parameters as cbm sees them (consistency_ice_veg_soil and
    derived_parameters applied).  {name: array(9)}."""
    S = {k: np.array(v, dtype=np.float64) for k, v in T.SOIL.items()}
    out = {k: S[k].copy() for k in ("bch", "clay", "css", "hyds", "rhosoil",
                                     "sand", "sfc", "silt", "ssat", "sucs",
                                     "swilt")}
    ice = T.ICE_SOIL - 1
    # cable_parameters.F90:2504-2507
    out["rhosoil"][ice] = 921.0     # density_ice
    out["css"][ice] = 2100.0        # csice
    # derived_parameters, cable_parameters.F90:2236-2241
    out["hsbh"] = out["hyds"] * np.abs(out["sucs"]) * out["bch"]
    nint = fortran_nint(out["bch"])
    out["ibp2"] = nint + 2.0
    out["pwb_min"] = np.power(out["swilt"] / out["ssat"], out["ibp2"])
    out["i2bp3"] = 2.0 * nint + 3.0
    # dry-soil thermal conductivity: the gridinfo file supplies it in the
    # reference; the synthetic workload uses the reference's own soil_thermal_fix
    # expression (cable_parameters.F90:2185-2187) on the table ssat and rhosoil.
    ssb = np.minimum(0.8, np.maximum(0.1, out["ssat"]))
    out["cnsd"] = ((0.135 * (1.0 - ssb)) + (64.7 / out["rhosoil"])) / (1.0 - 0.947 * (1.0 - ssb))
    return out


def zshh():
    z = np.zeros(MS + 1)
    zse = T.ZSE
    z[0] = 0.5 * zse[0]
    z[MS] = 0.5 * zse[MS - 1]
    for k in range(1, MS):
        z[k] = 0.5 * (zse[k - 1] + zse[k])
    return z


class Workload:
    """All inputs of an N-tile, nsteps-step run.

    ``params``  : {fortran name: array} per-tile parameters (reference layout)
    ``state0``  : {fortran name: array} initial state
    ``forcing`` : {fortran name: array(nsteps, ...)} per-step met fields
    ``vtab``/``stab``: class tables; ``iveg``/``isoilm`` per-tile indices.
    """

    def __init__(self, n: int, nsteps: int = 48, seed: int = DEFAULT_SEED,
                 dels: float = DELS, forcing: bool = True):
        self.n, self.nsteps, self.seed, self.dels = n, nsteps, seed, dels
        self.iveg, self.isoilm = tile_classes(n, seed)
        self.vtab = veg_class_table()
        self.stab = soil_class_table()
        self.vlai = 0.1 + 5.9 * _uniform(seed, _S_LAI, (n,))
        self.lat = -55.0 + 125.0 * _uniform(seed, _S_LAT, (n,))
        self.t0 = 273.0 + 24.0 * _uniform(seed, _S_T0, (n,))
        a = _uniform(seed, _S_ALB, (n,))
        self.albsoil = np.zeros((NRB, n))
        self.albsoil[0] = 0.06 + 0.10 * a
        self.albsoil[1] = 2.0 * self.albsoil[0]
        self._params()
        self._state0()
        self.forcing = self.make_forcing(0, nsteps) if forcing else None

    # ------------------------------------------------------------------
    def _params(self):
        n = self.n
        iv = self.iveg - 1
        isl = self.isoilm - 1
        V, S = self.vtab, self.stab
        p = {}
        p["veg%iveg"] = self.iveg.astype(np.float64)
        p["soil%isoilm"] = self.isoilm.astype(np.float64)
        for k in ("canst1", "dleaf", "ejmax", "frac4", "hc", "xalbnir", "rp20",
                  "rpcoef", "rs20", "shelrb", "vegcf", "tminvj", "tmaxvj",
                  "vbeta", "vcmax", "xfang", "extkn", "wai", "a1gs", "d0gs",
                  "alpha", "convex", "cfrd", "gswmin", "conkc0", "conko0",
                  "ekc", "eko", "g0", "g1", "rootbeta", "zr", "clitt"):
            p[f"veg%{k}"] = V[k][iv]
        p["veg%meth"] = np.ones(n)
        p["veg%vlai"] = self.vlai.copy()
        p["veg%refl"] = V["refl"][:2, iv]
        p["veg%taul"] = V["taul"][:2, iv]
        p["veg%froot"] = V["froot"][:, iv]
        for k in ("bch", "clay", "css", "hsbh", "hyds", "i2bp3", "ibp2",
                  "rhosoil", "sand", "sfc", "silt", "ssat", "sucs", "swilt",
                  "cnsd", "pwb_min"):
            p[f"soil%{k}"] = S[k][isl]
        p["soil%zse"] = np.array(T.ZSE)
        p["soil%zshh"] = zshh()
        p["soil%albsoil"] = self.albsoil.copy()
        for k in ("swilt", "sfc", "ssat", "sucs", "bch", "hyds"):
            p[f"soil%{k}_vec"] = np.tile(S[k][isl], (MS, 1))
        p["soil%zse_vec"] = np.tile(np.array(T.ZSE)[:, None], (1, n))
        p["rough%za_uv"] = np.full(n, 40.0)
        p["rough%za_tq"] = np.full(n, 40.0)
        p["bgc%ratecp"] = V["ratecp"].copy()
        p["bgc%ratecs"] = V["ratecs"].copy()
        p["rad%latitude"] = self.lat.copy()
        self.params = p

    def _state0(self):
        n = self.n
        p = self.params
        u = _uniform(self.seed, _S_WB, (MS, n))
        swilt, sfc, ssat = p["soil%swilt"], p["soil%sfc"], p["soil%ssat"]
        wb = swilt[None, :] + (sfc - swilt)[None, :] * u
        tair1 = self.t0 + 5.0 * np.sin(2.0 * math.pi * (0.0 - 9.0) / 24.0)
        s = {}
        s["ssnow%wb"] = wb
        s["ssnow%wbice"] = np.zeros((MS, n))
        s["ssnow%tgg"] = np.tile(tair1, (MS, 1))
        s["ssnow%tggsn"] = np.full((MSN, n), TFRZ)
        s["ssnow%ssdn"] = np.full((MSN, n), 120.0)
        s["ssnow%ssdnn"] = np.full(n, 120.0)
        s["ssnow%isflag"] = np.zeros(n)
        s["ssnow%snowd"] = np.zeros(n)
        s["ssnow%rtsoil"] = np.full(n, 100.0)
        s["ssnow%sconds"] = np.full((MSN, n), 0.06)
        s["ssnow%albsoilsn"] = self.albsoil.copy()
        s["canopy%us"] = np.full(n, 0.1)
        s["canopy%fwsoil"] = np.ones(n)
        s["bgc%cplant"] = self.vtab["cplant"][:, self.iveg - 1]
        s["bgc%csoil"] = self.vtab["csoil"][:, self.iveg - 1]
        # derived_parameters, cable_parameters.F90:2245-2254, 2296-2302
        s["ssnow%owetfac"] = initial_owetfac(wb[0], s["ssnow%wbice"][0], swilt, sfc)
        s["ssnow%wbliq"] = wb - s["ssnow%wbice"]
        s["ssnow%wblf"] = np.maximum(0.01, s["ssnow%wbliq"] / ssat[None, :])
        s["ssnow%wbfice"] = s["ssnow%wbice"] / ssat[None, :]
        # before the time loop, src/offline/cable_serial.F90:437-442
        s["ssnow%otss_0"] = s["ssnow%tgg"][0].copy()
        s["ssnow%otss"] = s["ssnow%tgg"][0].copy()
        s["ssnow%tss"] = s["ssnow%tgg"][0].copy()
        s["met%ofsd"] = np.full(n, 0.1)
        self.state0 = s

    # ------------------------------------------------------------------
    def make_forcing(self, k0: int, k1: int):
        """Forcing of steps k0..k1-1 (0-based) as {name: array(nk, ...)}."""
        n, seed, dels = self.n, self.seed, self.dels
        ns = self.nsteps
        ks = np.arange(k0, k1)
        hod = 0.5 * ks                                   # half-hourly, from 0h
        # all streams are drawn for the full run, tile-major, then sliced
        u_sw = _uniform(seed, _F_SW, (ns, n))[k0:k1]
        u_lw = _uniform(seed, _F_LW, (ns, n))[k0:k1]
        u_rh = _uniform(seed, _F_RH, (ns, n))[k0:k1]
        u_ua = _rng(seed, _F_UA).standard_normal((n, ns)).T[k0:k1]
        u_pb = _uniform(seed, _F_PB, (ns, n))[k0:k1]
        u_pa = _rng(seed, _F_PA).standard_exponential((n, ns)).T[k0:k1]
        u_ps = _uniform(seed, _F_PS, (ns, n))[k0:k1]
        diurnal = np.maximum(0.0, 1000.0 * np.sin(2.0 * math.pi * (hod - 6.0) / 24.0))
        swdown = diurnal[:, None] * (0.5 + 0.5 * u_sw)
        tk = self.t0[None, :] + 5.0 * np.sin(2.0 * math.pi * (hod[:, None] - 9.0) / 24.0)
        pmb = 950.0 + 70.0 * u_ps
        qv = rh_sh(100.0 * (0.3 + 0.6 * u_rh), tk, pmb)
        sigma = 0.5
        ua = np.exp((math.log(3.0) - 0.5 * sigma * sigma) + sigma * u_ua)
        # Bernoulli(0.1) * Exp(mean 2 mm/h), converted to mm per step
        precip = np.where(u_pb < 0.1, 2.0 * u_pa, 0.0) * (dels / 3600.0)
        precip_sn = np.where(tk <= TFRZ, precip, 0.0)
        f = {}
        f["met%fsd"] = np.stack([0.5 * swdown, 0.5 * swdown], axis=1)   # (nk, 2, n)
        f["met%fld"] = 250.0 + 170.0 * u_lw
        f["met%tk"] = tk
        f["met%tvair"] = tk.copy()
        f["met%tvrad"] = tk.copy()
        f["met%qv"] = qv
        f["met%ua"] = ua
        f["met%pmb"] = pmb
        f["met%precip"] = precip
        f["met%precip_sn"] = precip_sn
        f["met%ca"] = np.full((k1 - k0, n), 400.0e-6)
        f["met%coszen"] = np.stack([sinbet(DOY, self.lat, h) for h in hod])
        f["met%doy"] = np.full((k1 - k0, n), DOY)
        f["met%hod"] = np.tile(hod[:, None], (1, n))
        return f

    # ------------------------------------------------------------------
    def subset(self, idx):
        """The same run restricted to tiles ``idx`` (grid-wide parameters such
        as soil%zse and bgc%ratecp are kept whole)."""
        idx = np.asarray(idx)
        n = self.n
        sub = copy.copy(self)
        sub.n = int(idx.size)

        def cut(d):
            return {k: (v[..., idx] if k not in GRID_WIDE and v.shape[-1] == n else v.copy())
                    for k, v in d.items()}
        sub.params = cut(self.params)
        sub.state0 = cut(self.state0)
        sub.forcing = cut(self.forcing) if self.forcing is not None else None
        for a in ("iveg", "isoilm", "vlai", "lat", "t0"):
            setattr(sub, a, getattr(self, a)[idx])
        sub.albsoil = self.albsoil[:, idx]
        sub.make_forcing = None
        return sub


# parameters that are one value per grid, not per tile
GRID_WIDE = ("soil%zse", "soil%zshh", "bgc%ratecp", "bgc%ratecs")


def add_snow(wl, every=3, deep=150.0, thin=5.0):
    """Start a workload's tiles with a mix of deep snow (three layers once
    snowcheck has run), thin single-layer snow and none, in turn."""
    sn = np.zeros(wl.n)
    sn[::every] = deep
    sn[1::every] = thin
    wl.state0["ssnow%snowd"] = sn
    return wl


# ----------------------------------------------------------------------
# synthetic forcing keyed by (seed, draw, step, tile): the host form of the
# device generator behind cable_gpu_synth_forcing (include/cable_gpu.h)

_M64 = (1 << 64) - 1


def mix64(z):
    """SplitMix64's increment and finaliser (Steele, Lea and Flood, OOPSLA
    2014) on Python ints or numpy uint64 arrays (wrapping arithmetic)."""
    if isinstance(z, np.ndarray):
        with np.errstate(over="ignore"):
            z = z + np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            return z ^ (z >> np.uint64(31))
    z = (int(z) + 0x9E3779B97F4A7C15) & _M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return z ^ (z >> 31)


def synth_keys(seed):
    """Key of each of the 8 draws (sw, lw, rh, ua Box-Muller pair, rain
    occurrence, rain amount, psurf)."""
    return [mix64(mix64(int(seed) & _M64) ^ (s + 1)) for s in range(8)]


def synth_u(key, tiles, steps):
    """U(0,1), open at both ends: (steps x tiles) array."""
    t = np.asarray(tiles, dtype=np.uint64)[None, :]
    k = np.asarray(steps, dtype=np.uint64)[:, None]
    h = mix64(np.uint64(key) ^ (k << np.uint64(32)) ^ t)
    return ((h >> np.uint64(11)).astype(np.float64) + 0.5) * 2.0 ** -53


def synth_forcing(t0, lat, k_first, nsteps, seed=DEFAULT_SEED, dels=DELS, doy=DOY, hod0=0.0, tiles=None):
    """Forcing of steps k_first.. for tiles with mean air temperature t0 (K)
    and latitude lat (degrees), {name: array(nsteps, n)}, in the device
    generator's formulas; ``tiles``: the caller's tile indices (default
    0..n-1), which key the draws."""
    t0 = np.asarray(t0, dtype=np.float64)
    lat = np.asarray(lat, dtype=np.float64)
    n = t0.size
    tiles = np.arange(n) if tiles is None else np.asarray(tiles)
    ks = np.arange(k_first, k_first + nsteps)
    K = synth_keys(seed)
    u = [synth_u(K[s], tiles, ks) for s in range(8)]
    hod = (hod0 + ks * (dels / 3600.0))[:, None]
    two_pi = 2.0 * math.pi
    diurnal = np.maximum(0.0, 1000.0 * np.sin(two_pi * (hod - 6.0) / 24.0))
    sw = diurnal * (0.5 + 0.5 * u[0])
    tair = t0[None, :] + 5.0 * np.sin(two_pi * (hod - 9.0) / 24.0)
    ps = 950.0 + 70.0 * u[7]
    qv = rh_sh(100.0 * (0.3 + 0.6 * u[2]), tair, ps)
    z = np.sqrt(-2.0 * np.log(u[3])) * np.cos(two_pi * u[4])
    sigma = 0.5
    ua = np.exp((math.log(3.0) - 0.5 * sigma * sigma) + sigma * z)
    rain = np.where(u[5] < 0.1, 2.0 * -np.log(u[6]) * (dels / 3600.0), 0.0)
    cz = np.stack([sinbet(doy, lat, h) for h in hod[:, 0]])
    return {"met%fsd": 0.5 * sw, "met%fld": 250.0 + 170.0 * u[1], "met%tk": tair, "met%qv": qv,
            "met%ua": ua, "met%pmb": ps, "met%precip": rain, "met%coszen": cz}
