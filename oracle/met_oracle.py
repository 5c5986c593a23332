"""CPU oracle of the met-variable forcing path (cable_gpu_set_met).

Test infrastructure only: tests/ use it to check the library.  It restates
what the offline reader get_met_data (src/offline/cable_input.F90) does with
a met file's variables for a run on local time:

* the clock: at the run's first step hod = shod, doy = sdoy (:1688-1693),
  then hod = hod + dels/3600 per step (:1707); at hod >= 24 the day rolls
  over and the month and year follow the reader's day-of-year tables
  (:1780-1846), at hod < 0 it rolls back (:1710-1779), with IS_LEAPYEAR
  (src/util/cable_common.F90:226-235);
* fsd(:,1) = fsd(:,2) = SWdown/2 (:1880-1883); tk = Tair + convert%Tair
  (:1923); pmb = PSurf * convert%PSurf (:1959) or, without PSurf,
  1013.25 (tk/(tk + 0.0065 elev))**(9.80665/287.04/0.0065) (:1979-1981);
  qv = Qair, or rh_sh(Qair, tk, pmb) for relative humidity (:2000-2003,
  src/offline/cable_checks.F90:633-696); ua = Wind, or
  sqrt(Wind_N**2 + Wind_E**2) (:2094); precip = Rainf * convert%Rainf
  (:2202-2204); fld = LWdown or 0.0000094 * 0.0000000567 * tk**6 (:2257);
  ca = fixedCO2/1e6 (:2295); coszen = sinbet(doy, lat, hod) (:2677,
  src/science/radiation/cbl_sinbet.F90:8-28).

`dtype` is the reader's REAL: float64 for the promoted build, float32 for
the shipped one.
"""
from __future__ import annotations

import numpy as np


def is_leapyear(y):
    return (y % 4 == 0 and y % 100 != 0) or y % 400 == 0


_FWD = {True: [32, 61, 92, 122, 153, 183, 214, 245, 275, 306, 336],
        False: [32, 60, 91, 121, 152, 182, 213, 244, 274, 305, 335]}
_BWD = {True: [31, 60, 91, 121, 152, 182, 213, 244, 274, 305, 335],
        False: [31, 59, 90, 120, 151, 181, 212, 243, 273, 304, 334]}


def _roll(hod, doy, moy, year, T):
    if hod < T(0.0):
        doy = T(doy - T(1))
        hod = T(hod + T(24.0))
        lp = is_leapyear(year)
        d = int(doy)
        if d == 0:
            moy, year = 12, year - 1
            doy = T(365) if lp else (T(366) if is_leapyear(year) else T(365))
        elif d in _BWD[lp]:
            moy = _BWD[lp].index(d) + 1
    elif hod >= T(24.0):
        doy = T(doy + T(1))
        hod = T(hod - T(24.0))
        lp = is_leapyear(year)
        d = int(doy)
        if d in _FWD[lp]:
            moy = _FWD[lp].index(d) + 2
        if d == (367 if lp else 366):
            year, moy, doy = year + 1, 1, T(1)
    return hod, doy, moy, year


def clock(shod, sdoy, smoy, syear, dels, nsteps, dtype=np.float64):
    """(hod, doy, moy, year) of steps 0..nsteps-1, each a list."""
    T = np.dtype(dtype).type
    hod, doy, moy, year = T(shod), T(sdoy), smoy, syear
    hod, doy, moy, year = _roll(hod, doy, moy, year, T)
    out = []
    for k in range(nsteps):
        if k > 0:
            hod = T(hod + T(dels) / T(3600.0))
            hod, doy, moy, year = _roll(hod, doy, moy, year, T)
        out.append((float(hod), float(doy), moy, year))
    return [list(x) for x in zip(*out)]


def svp(tk, T):
    tk = np.asarray(tk, dtype=T)
    temp = tk - T(273.15)
    l10 = np.log(T(10.0))
    with np.errstate(all="ignore"):
        toot = T(273.15) / tk
        toto = T(1.) / toot
        ei = T(-9.09718) * (toot - T(1)) - T(3.56654) * (np.log(toot) / l10) + T(0.876793) * (T(1) - toto) + \
            (np.log(T(6.1071)) / l10)
        tsot = T(373.15) / tk
        ew = T(-7.90298) * (tsot - T(1)) + T(5.02808) * (np.log(tsot) / l10)
        ew = ew - T(1.3816e-07) * (np.power(T(10), T(11.344) * (T(1) - (T(1) / tsot))) - T(1))
        ew = ew + T(0.0081328) * (np.power(T(10), T(-3.49149) * (tsot - T(1))) - T(1))
        ew = ew + (np.log(T(1013.246)) / l10)
    return np.where(temp < T(-20.0), np.power(T(10.0), ei), np.power(T(10.0), ew)).astype(T)


def forcing(met, lat, hod, doy, elevation=None, dtype=np.float64):
    """{name: array(nsteps, mp)} of the resident forcing from met variables
    (dict of (nsteps, mp) arrays and the conversion constants)."""
    T = np.dtype(dtype).type
    f = lambda k: np.asarray(met[k], dtype=np.float32).astype(T)
    tk = f("tair") + T(met["tair_offset"])
    if met.get("psurf") is not None:
        pmb = f("psurf") * T(met["psurf_scale"])
    else:
        e = np.asarray(elevation, dtype=T)[None, :]
        pmb = T(1013.25) * np.power(tk / (tk + T(0.0065) * e), T(9.80665) / T(287.04) / T(0.0065))
    qv = f("qair")
    if met.get("qair_is_rh"):
        es = svp(tk, T)
        qv = (qv / T(100.0)) * (T(0.622) * es / (pmb - es))
    ua = f("wind")
    if met.get("wind_e") is not None:
        we = f("wind_e")
        ua = np.sqrt(ua * ua + we * we)
    fld = f("lwdown") if met.get("lwdown") is not None else T(0.0000094) * T(0.0000000567) * np.power(tk, T(6.0))
    pi = T(3.1415927)
    pi180 = pi / T(180.0)
    d = np.asarray(doy, dtype=T)[:, None]
    h = np.asarray(hod, dtype=T)[:, None]
    x = np.asarray(lat, dtype=T)[None, :]
    sindec = -np.sin(T(23.45) * pi180) * np.cos(T(2.) * pi * (d + T(10.0)) / T(365.0))
    z = np.sin(pi180 * x) * sindec + np.cos(pi180 * x) * np.sqrt(T(1.) - sindec * sindec) * \
        np.cos(pi * (h - T(12.0)) / T(12.0))
    return {"met%fsd": T(0.5) * f("swdown"), "met%fld": fld, "met%tk": tk, "met%qv": qv, "met%ua": ua,
            "met%pmb": pmb, "met%precip": f("rainf") * T(met["rainf_scale"]),
            "met%coszen": np.maximum(z, T(1e-8))}
