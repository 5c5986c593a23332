"""Forcing from a met file's variables (cable_gpu_set_met): the reader's
clock and unit conversions.  The oracle (oracle/met_oracle.py) is pinned to
the calendar (Python's datetime) and to closed forms; the library's clock
(host code) and the device conversions are checked against the oracle."""
import datetime as dt

import numpy as np
import pytest

from oracle import met_oracle as M


@pytest.mark.parametrize("start", [(22.0, 365, 12, 2019), (23.5, 59, 2, 2020), (0.0, 1, 1, 2100),
                                   (12.0, 366, 12, 2000)])
def test_oracle_clock_follows_the_calendar(start):
    """Half-hourly steps across month, year and leap-day boundaries: hod,
    day of year and month are those of the calendar."""
    shod, sdoy, smoy, syear = start
    hod, doy, moy, year = M.clock(shod, sdoy, smoy, syear, 1800.0, 400)
    t0 = dt.datetime(syear, 1, 1) + dt.timedelta(days=sdoy - 1, hours=shod)
    for k in range(400):
        t = t0 + dt.timedelta(minutes=30 * k)
        assert (hod[k], doy[k], year[k]) == (t.hour + t.minute / 60.0, t.timetuple().tm_yday, t.year), k
        assert moy[k] == t.month, k


def test_oracle_clock_rolls_back_before_midnight():
    # a negative start hour (a local-time offset) moves to the previous day;
    # from 1 January the reader assumes the year before a leap year is not one
    hod, doy, moy, year = M.clock(-1.0, 1, 1, 2021, 1800.0, 1)
    assert (hod[0], doy[0], moy[0], year[0]) == (23.0, 366.0, 12, 2020)
    hod, doy, moy, year = M.clock(-1.0, 1, 1, 2020, 1800.0, 1)
    assert (hod[0], doy[0], moy[0], year[0]) == (23.0, 365.0, 12, 2019)
    hod, doy, moy, year = M.clock(-0.5, 61, 3, 2020, 1800.0, 1)
    assert (doy[0], moy[0]) == (60.0, 2)


def test_oracle_conversions_closed_forms():
    rng = np.random.default_rng(0)
    n, ns = 50, 3
    met = {"swdown": rng.uniform(0, 1000, (ns, n)), "tair": rng.uniform(-30, 35, (ns, n)),
           "qair": rng.uniform(20, 95, (ns, n)), "wind": rng.uniform(-5, 5, (ns, n)),
           "wind_e": rng.uniform(-5, 5, (ns, n)), "rainf": rng.exponential(1e-4, (ns, n)),
           "tair_offset": 273.16, "rainf_scale": 1800.0, "qair_is_rh": True}
    elev = rng.uniform(0, 3000, n)
    lat = rng.uniform(-60, 70, n)
    f = M.forcing(met, lat, [6.0, 12.0, 18.0], [172.0] * 3, elevation=elev)
    f32 = lambda k: np.asarray(met[k], dtype=np.float32).astype(np.float64)
    np.testing.assert_array_equal(f["met%tk"], f32("tair") + 273.16)
    np.testing.assert_array_equal(f["met%fsd"], 0.5 * f32("swdown"))
    np.testing.assert_allclose(f["met%ua"], np.hypot(f32("wind"), f32("wind_e")), rtol=1e-15)
    np.testing.assert_allclose(f["met%precip"], f32("rainf") * 1800.0, rtol=1e-15)
    # sea-level pressure at zero elevation, and the barometric decrease
    g = M.forcing(dict(met, qair_is_rh=False), lat, [12.0] * 3, [172.0] * 3, elevation=np.zeros(n))
    np.testing.assert_allclose(g["met%pmb"], 1013.25, rtol=1e-15)
    # the standard atmosphere (288.15 K at sea level, 6.5 K/km): 898.75 hPa and
    # 281.65 K at 1000 m, the reader's formula taking the station temperature
    h = M.forcing(dict(met, qair_is_rh=False, tair=np.full((ns, n), 8.5), tair_offset=273.15), lat, [12.0] * 3, [172.0] * 3,
                  elevation=np.full(n, 1000.0))
    np.testing.assert_allclose(h["met%pmb"], 898.75, atol=0.05)
    # relative humidity comes back from qv through the saturation pressure
    es = M.svp(f["met%tk"], np.float64)
    rh = 100.0 * f["met%qv"] / (0.622 * es / (f["met%pmb"] - es))
    np.testing.assert_allclose(rh, f32("qair"), rtol=1e-12)
    # Swinbank's longwave, and cos(zenith) at the equinox noon on the equator
    np.testing.assert_allclose(f["met%fld"], 0.0000094 * 0.0000000567 * f["met%tk"] ** 6, rtol=1e-14)
    z = M.forcing(met, np.zeros(n), [12.0] * 3, [80.0] * 3, elevation=elev)["met%coszen"]
    assert np.abs(z - 1.0).max() < 2e-3


@pytest.mark.parametrize("precision", [0, 1])
def test_library_clock_matches_oracle(precision):
    from cable_amd import gpu
    T = np.float64 if precision == 0 else np.float32
    for start in [(22.0, 365, 12, 2019), (0.25, 59, 2, 2024), (23.0, 366, 12, 2000)]:
        hod, doy, _, _ = M.clock(*start, 1800.0, 300, dtype=T)
        h, d = gpu.met_clock(precision, start, 1800.0, 100, 200)
        np.testing.assert_array_equal(h, hod[100:])
        np.testing.assert_array_equal(d, doy[100:])


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("variant", ["full", "derived"])
def test_device_met_matches_oracle(precision, variant):
    """Every resident forcing variable from met variables: all given (K,
    kg/kg, Pa, LWdown, vector wind), or the reader's derived forms (deg C,
    relative humidity, PSurf from elevation, Swinbank's longwave)."""
    from cable_amd import gpu
    from cable_amd.workload import Workload
    print(f"[met] device conversions precision={precision} {variant}", flush=True)
    rng = np.random.default_rng(7)
    wl = Workload(1500, 1, forcing=False)
    n, ns, k0 = wl.n, 6, 46
    met = {"swdown": rng.uniform(0, 1000, (ns, n)), "wind": rng.lognormal(1.0, 0.5, (ns, n)),
           "rainf": np.where(rng.random((ns, n)) < 0.1, rng.exponential(5e-4, (ns, n)), 0.0)}
    if variant == "full":
        met.update(tair=rng.uniform(250, 310, (ns, n)), qair=rng.uniform(1e-3, 2e-2, (ns, n)),
                   psurf=rng.uniform(9.5e4, 1.02e5, (ns, n)), lwdown=rng.uniform(250, 420, (ns, n)),
                   wind_e=rng.normal(0, 2, (ns, n)), tair_offset=0.0, psurf_scale=0.01, rainf_scale=1800.0)
    else:
        met.update(tair=rng.uniform(-25, 35, (ns, n)), qair=rng.uniform(30, 90, (ns, n)), qair_is_rh=True,
                   tair_offset=273.16, rainf_scale=0.5)
    elev = rng.uniform(0, 2500, n)
    start = (23.0, 59, 2, 2024)
    g = gpu.from_workload(wl, precision=precision, full=False, order=gpu.class_order(wl.iveg, wl.isoilm, wl.vlai))
    g.set_met(k0, ns, 1800.0, met, wl.lat, elevation=elev, start=start)
    T = np.float64 if precision == 0 else np.float32
    hod, doy, _, _ = M.clock(*start, 1800.0, k0 + ns, dtype=T)
    ref = M.forcing(met, wl.lat, hod[k0:], doy[k0:], elevation=elev, dtype=T)
    rtol = 1e-13 if precision == 0 else 4e-6
    for i in range(ns):
        got = g.download_forcing(k0 + i)
        for k, a in got.items():
            np.testing.assert_allclose(a, ref[k][i], rtol=rtol, atol=1e-30, err_msg=f"{k} step {k0 + i}")
    # the window's steps run, and use the window's day of year
    g.run(k0, ns, 1800.0)
    assert np.isfinite(g.download("canopy%fe")).all()
    g.close()
