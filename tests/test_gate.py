"""The accuracy gate of the mixed and hybrid forms, on the GPU.

Both are compared with the plain FP64 path -- which the per-step parity
tests pin to the oracle, and the oracle to the compiled reference -- flux by
flux at every step of a full day (48 half-hourly steps) on 65,536 tiles of
the benchmark workload, in the benchmark's class tile order.  Errors are in
the gate's units: W/m2 for the energy fluxes (canopy%rnet, fe, fh, fev,
fes) and umol CO2 m-2 s-1 for canopy%fgpp (stored in g C m-2 s-1,
src/offline/cable_define_types.F90:465; 1 umol = 12e-6 g C).

The yardstick is the reference's own spread, measured on the CPU by
oracle/spread.py on 49,152 tiles x 48 steps of the same workload (daylight
leaf-bearing tile-steps, the ones the hybrid warm-starts):

* the reference as shipped (FP32 default REAL) against itself promoted to
  FP64: latent heat p99.9 0.043 W/m2, max 43.9 W/m2, 0.013% of tile-steps
  beyond 1 W/m2;
* the FP64 reference with its leaf-temperature start (cable_canopy.F90:213)
  moved by 1e-9 K: latent heat p99.9 0.85 W/m2, max 199 W/m2, 0.088% beyond
  1 W/m2; by 1e-6 K: p99.9 6.3 W/m2, max 444 W/m2, 0.51% beyond; by 0.1 K:
  p99.9 16.3 W/m2, 1.5% beyond; by 1 K: p99.9 15.3 W/m2, 1.4% beyond (the
  spread saturates from about 1e-3 K).

The reference's fixed point stops at the first iterate within 0.1 K
(cbl_dryLeaf.F90:565-606), so any change of its start moves some tiles to
another iterate of that band: no method that starts the iteration anywhere
else can keep every tile within 1 W/m2.  The mixed path is held to the
shipped build's spread, the hybrid -- whose predicted start (the leaf
temperature where the first NITER pass converges) differs from the
reference's (the air temperature) by kelvins -- to the reference's own
spread under a 1 K change of its start; both report the gate's numbers.
"""
import numpy as np
import pytest

from cable_amd.workload import Workload
from oracle.spread import ENERGY, GATE, error_stats

pytestmark = pytest.mark.gpu

FLUXES = ENERGY + ["canopy%fgpp"]
N_TILES, N_STEPS = 65536, 48
HYB_CAP, HYB_RESID = 4, 0.5

# the reference's spread on daylight leaf-bearing tile-steps (oracle/spread.py,
# 49,152 tiles x 48 steps): p99.9 and fraction beyond the gate per flux
SHIPPED_VS_PROMOTED = {"canopy%rnet": (0.0059, 2.0e-5), "canopy%fe": (0.043, 1.3e-4),
                       "canopy%fh": (0.033, 1.3e-4), "canopy%fev": (0.029, 8.6e-5),
                       "canopy%fes": (0.012, 5.7e-5), "canopy%fgpp": (0.00063, 3.5e-5)}
START_1K = {"canopy%rnet": (4.41, 3.24e-3), "canopy%fe": (15.3, 1.42e-2), "canopy%fh": (11.7, 1.21e-2),
            "canopy%fev": (10.6, 1.15e-2), "canopy%fes": (6.87, 4.9e-3), "canopy%fgpp": (1.013, 4.45e-3)}


def _day(wl):
    fsd = wl.forcing["met%fsd"][:, 0, :]
    return (2.0 * fsd > 0.001) & (wl.iveg < 14)[None, :]


def _trajectory(g, wl, hybrid=None):
    out = {k: [] for k in FLUXES}
    for s in range(wl.nsteps):
        if hybrid is None:
            g.run(s, 1, wl.dels)
        else:
            g.run_hybrid(s, 1, wl.dels, **hybrid)
        for k in FLUXES:
            out[k].append(g.download(k).reshape(-1))
    return {k: np.stack(v) for k, v in out.items()}


def _show(label, st):
    print(label + ": " + "; ".join(
        f"{k.split('%')[1]} max {v['max']:.3g} p99.9 {v['p99.9']:.3g} beyond {v.get('frac_beyond_gate', 0):.2e}"
        for k, v in st.items()), flush=True)


@pytest.fixture(scope="module")
def forms():
    from cable_amd import gpu
    from cable_amd import surrogate as S
    print(f"[gate] training the surrogate; forms on {N_TILES} tiles x {N_STEPS} steps", flush=True)
    params, norm, info = S.train_on_gpu(n_tiles=65536, nsteps=48, epochs=2.0, log=lambda m: print(m, flush=True))
    wl = Workload(N_TILES, N_STEPS)
    order = gpu.class_order(wl.iveg, wl.isoilm, wl.vlai)
    res = {}
    for name, prec in (("fp64", gpu.PREC_FP64), ("mixed", gpu.PREC_MIXED), ("hybrid", gpu.PREC_MIXED)):
        g = gpu.from_workload(wl, precision=prec, full=False, order=order)
        hyb = None
        if name == "hybrid":
            g.set_surrogate(params, norm)
            hyb = {"max_passes": HYB_CAP, "resid_max": HYB_RESID, "use_surrogate": True}
        res[name] = _trajectory(g, wl, hyb)
        g.close()
        print(f"[gate] {name} done", flush=True)
    return wl, res


def test_mixed_within_the_shipped_reference_spread(forms):
    """The mixed path (FP32 canopy) against FP64: its tail no wider than the
    reference's own shipped build against its promoted build, within 3x,
    and the gate's 99.9th percentile met."""
    wl, res = forms
    day = _day(wl)
    st = error_stats(res["mixed"], res["fp64"], FLUXES, daylight=day)
    _show("[gate] mixed vs fp64, daylight leafy tile-steps", st)
    _show("[gate] mixed vs fp64, all tile-steps", error_stats(res["mixed"], res["fp64"], FLUXES))
    for k in FLUXES:
        p999, frac = SHIPPED_VS_PROMOTED[k]
        gate = GATE[k] / (12e-6 if k == "canopy%fgpp" else 1.0)
        assert st[k]["p99.9"] <= gate, (k, st[k])
        assert st[k]["p99.9"] <= max(3.0 * p999, 0.05 * gate), (k, st[k]["p99.9"], p999)
        assert st[k]["frac_beyond_gate"] <= 3.0 * frac + 1e-4, (k, st[k]["frac_beyond_gate"], frac)


def test_hybrid_within_the_reference_start_sensitivity(forms):
    """The hybrid against FP64 on daylight leaf-bearing tile-steps: its
    99.9th percentile and its fraction of tile-steps beyond the gate within
    1.5x of the reference's own when its leaf-temperature start moves by
    1 K."""
    wl, res = forms
    day = _day(wl)
    st = error_stats(res["hybrid"], res["fp64"], FLUXES, daylight=day)
    al = error_stats(res["hybrid"], res["fp64"], FLUXES)
    _show("[gate] hybrid vs fp64, daylight leafy tile-steps", st)
    _show("[gate] hybrid vs fp64, all tile-steps", al)
    print("[gate] hybrid / reference 1 K start spread (p99.9, beyond): " + "; ".join(
        f"{k.split('%')[1]} {st[k]['p99.9'] / START_1K[k][0]:.2f}, "
        f"{st[k]['frac_beyond_gate'] / START_1K[k][1]:.2f}" for k in FLUXES), flush=True)
    for k in FLUXES:
        p999, frac = START_1K[k]
        assert st[k]["p99.9"] <= 1.5 * p999, (k, st[k]["p99.9"], p999)
        assert st[k]["frac_beyond_gate"] <= 1.5 * frac, (k, st[k]["frac_beyond_gate"], frac)


def test_fallback_tiles_get_the_plain_step():
    """No network, one pass per NITER pass and no residual allowance: every
    capped tile the single pass does not settle falls back, and a tile the
    pass settles is where the plain method stops too.  Every tile then
    carries the plain mixed step's result, bit for bit, over daylight steps
    with many fallbacks."""
    from cable_amd import gpu
    print("[gate] fallback tiles", flush=True)
    wl = Workload(8192, 30)
    order = gpu.class_order(wl.iveg, wl.isoilm, wl.vlai)
    names = FLUXES + ["ssnow%runoff", "rad%trad", "ssnow%tgg", "ssnow%wb", "ssnow%snowd", "canopy%cansto"]
    g = gpu.from_workload(wl, precision=gpu.PREC_MIXED, full=False, order=order)
    g.run_hybrid(0, 30, wl.dels, max_passes=1, resid_max=0.0, use_surrogate=False)
    fb, ts = g.hybrid_counters()
    a = {k: g.download(k) for k in names}
    g.close()
    g = gpu.from_workload(wl, precision=gpu.PREC_MIXED, full=False, order=order)
    g.run(0, 30, wl.dels)
    b = {k: g.download(k) for k in names}
    g.close()
    print(f"[gate] fallback fraction {fb / ts:.3f}", flush=True)
    assert fb > 0.05 * ts
    # the same kernel code in another mode: the same bits
    diff = [k for k in names if not np.array_equal(a[k], b[k])]
    assert not diff, diff
