"""Synthetic forcing keyed by (seed, draw, step, tile): the host generator
(cable_amd/workload.py) against known answers and its distributions, and the
device generator (cable_gpu_synth_forcing) against the host one."""
import math

import numpy as np
import pytest
from scipy import stats

from cable_amd import workload as W


def test_mix64_is_splitmix64():
    # SplitMix64 seeded with 0: the first output is mix64(0)
    assert W.mix64(0) == 0xE220A8397B1DCDAF
    z = np.array([0, 1, 2 ** 63, 2 ** 64 - 1, 0x9E3779B97F4A7C15], dtype=np.uint64)
    assert [int(v) for v in W.mix64(z)] == [W.mix64(int(v)) for v in z]


def test_uniform_draws_are_uniform_and_independent():
    keys = W.synth_keys(W.DEFAULT_SEED)
    assert len(set(keys)) == 8
    tiles = np.arange(200000)
    u = W.synth_u(keys[0], tiles, [5])[0]
    v = W.synth_u(keys[1], tiles, [5])[0]
    w = W.synth_u(keys[0], tiles, [6])[0]
    assert 0.0 < u.min() and u.max() < 1.0
    for x in (u, v, w):
        assert stats.kstest(x, "uniform").pvalue > 1e-4
    for a, b in ((u, v), (u, w)):
        assert abs(np.corrcoef(a, b)[0, 1]) < 0.01
    # a tile's draws do not depend on which other tiles are generated
    np.testing.assert_array_equal(W.synth_u(keys[3], tiles[1000:1010], [7, 9]),
                                  W.synth_u(keys[3], tiles, [7, 9])[:, 1000:1010])


def test_forcing_distributions():
    n, ns = 20000, 48
    rng = np.random.default_rng(0)
    t0 = rng.uniform(273.0, 297.0, n)
    lat = rng.uniform(-55.0, 70.0, n)
    f = W.synth_forcing(t0, lat, 0, ns)
    hod = 0.5 * np.arange(ns)
    # SWdown = max(0, 1000 sin(2 pi (hod - 6)/24)) x U(0.5, 1), half per band
    peak = np.maximum(0.0, 1000.0 * np.sin(2 * math.pi * (hod - 6.0) / 24.0))[:, None]
    ratio = np.divide(2.0 * f["met%fsd"], peak, out=np.full_like(peak * np.ones(n), 0.75), where=peak > 0)
    assert ratio.min() >= 0.5 and ratio.max() <= 1.0
    assert abs(ratio[peak[:, 0] > 0].mean() - 0.75) < 0.005
    np.testing.assert_allclose(f["met%tk"], t0[None, :] + 5.0 * np.sin(2 * math.pi * (hod[:, None] - 9.0) / 24.0))
    assert 250.0 < f["met%fld"].min() and f["met%fld"].max() < 420.0
    assert 950.0 < f["met%pmb"].min() and f["met%pmb"].max() < 1020.0
    # lognormal wind with mean 3 m/s, sigma 0.5
    lw = np.log(f["met%ua"]).ravel()
    assert abs(lw.std() - 0.5) < 0.005 and abs(f["met%ua"].mean() - 3.0) < 0.02
    # rain on 10% of tile-steps, Exp(mean 2 mm/h) over half an hour
    wet = f["met%precip"] > 0
    assert abs(wet.mean() - 0.1) < 0.003
    assert abs(f["met%precip"][wet].mean() - 1.0) < 0.02
    # relative humidity recovered from qv lies in [30, 90]%
    es = W.svp(f["met%tk"])
    rh = 100.0 * f["met%qv"] / (0.622 * es / (f["met%pmb"] - es))
    assert rh.min() >= 30.0 - 1e-9 and rh.max() <= 90.0 + 1e-9
    np.testing.assert_array_equal(f["met%coszen"][3], W.sinbet(W.DOY, lat, 1.5))


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_device_generator_matches_host(precision):
    from cable_amd import gpu
    print(f"[synth] device generator, precision {precision}", flush=True)
    wl = W.Workload(4099, 1, forcing=False)     # not a multiple of 256
    g = gpu.from_workload(wl, precision=precision, full=False, order=gpu.class_order(wl.iveg, wl.isoilm))
    g.synth_forcing(48, 8, W.DEFAULT_SEED, wl.t0, wl.lat)
    ref = W.synth_forcing(wl.t0, wl.lat, 48, 8)
    for i, k in enumerate((48, 51, 55)):
        got = g.download_forcing(k)
        for name, a in got.items():
            r = ref[name][k - 48]
            if precision == 0:
                if name in ("met%fld", "met%pmb", "met%fsd"):
                    # affine in the draw (and a sine of the step): the same operations
                    np.testing.assert_allclose(a, r, rtol=1e-15, atol=0, err_msg=name)
                else:
                    np.testing.assert_allclose(a, r, rtol=1e-13, atol=1e-300, err_msg=name)
            else:
                np.testing.assert_allclose(a, r.astype(np.float32), rtol=2.0 ** -22, atol=1e-30, err_msg=name)
    # steps run inside the resident window only
    g.run(48, 8, wl.dels)
    with pytest.raises(gpu.CableError):
        g.run(40, 8, wl.dels)
    out = g.download("canopy%fe")
    assert np.isfinite(out).all()
    g.close()
