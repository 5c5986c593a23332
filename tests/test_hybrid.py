"""The surrogate's use in the step (mixed-precision contexts, on the GPU):
sample capture, the hybrid step's fallback machinery, and what the trained
network gains over the same capped iteration without it."""
import numpy as np
import pytest

from cable_amd.workload import Workload

pytestmark = pytest.mark.gpu

HEADLINE = ["canopy%rnet", "canopy%fe", "canopy%fh", "canopy%fev", "canopy%fes", "canopy%fgpp",
            "ssnow%runoff", "rad%trad"]


def _ctx(wl, precision=1):
    from cable_amd import gpu
    return gpu.from_workload(wl, precision=precision, full=False, order=gpu.class_order(wl.iveg, wl.isoilm))


def _outputs(g, names=HEADLINE + ["ssnow%tgg", "ssnow%wb"]):
    return {k: g.download(k) for k in names}


def test_capture_keeps_the_plain_trajectory_and_writes_samples():
    """The capture step writes samples and leaves the plain step's bits."""
    print("[hybrid] capture", flush=True)
    wl = Workload(8192, 16)
    g = _ctx(wl)
    n = g.run_capture(0, 16, wl.dels, 8192 * 16)
    a = _outputs(g)
    x, y = g.copy_samples(0, n)
    g.close()
    g = _ctx(wl)
    g.run(0, 16, wl.dels)
    b = _outputs(g)
    g.close()
    # at most one sample per tile-step of a type that can carry leaves
    # (iveg < 14), and only where the canopy above the snow has leaves
    veg = (wl.iveg < 14).sum()
    assert 0.5 * veg * 16 <= n <= veg * 16
    assert np.isfinite(x).all() and np.isfinite(y).all()
    # positive leaf-surface CO2; leaf temperature within 20 K of the air
    assert (y[:, 1] > 0).all() and (y[:, 2] > 0).all()
    assert np.abs(y[:, 0]).max() < 20.0
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_uncapped_hybrid_without_network_is_the_plain_step():
    """max_passes = 20 and no network: the hybrid step is the plain method,
    bit for bit, and nothing falls back."""
    print("[hybrid] uncapped without network", flush=True)
    wl = Workload(8192, 12)
    g = _ctx(wl)
    g.run_hybrid(0, 12, wl.dels, max_passes=20, resid_max=0.5, use_surrogate=False)
    fb, ts = g.hybrid_counters()
    a = _outputs(g)
    g.close()
    g = _ctx(wl)
    g.run(0, 12, wl.dels)
    b = _outputs(g)
    g.close()
    assert fb == 0 and ts == 8192 * 12
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_night_steps_run_the_reference_iteration():
    """Without daylight no tile is warm-started or capped: with a network and
    a cap of one pass, the night steps (00:00 to 05:30) send nothing to the
    fallback and give the plain step's bits."""
    from cable_amd.surrogate import NIN, NOUT
    from oracle import mlp_oracle as O
    print("[hybrid] night steps", flush=True)
    wl = Workload(8192, 12)
    p = O.init_params(5).astype(np.float32)
    unit = np.concatenate([np.zeros(NIN), np.ones(NIN), np.zeros(NOUT), np.ones(NOUT)]).astype(np.float32)
    g = _ctx(wl)
    g.set_surrogate(p, unit)
    g.run_hybrid(0, 12, wl.dels, max_passes=1, resid_max=0.5, use_surrogate=True)
    fb, ts = g.hybrid_counters()
    a = _outputs(g)
    g.close()
    g = _ctx(wl)
    g.run(0, 12, wl.dels)
    b = _outputs(g)
    g.close()
    assert fb == 0 and ts == 8192 * 12
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_unlit_steps_keep_the_leaf_state_for_the_morning():
    """A hybrid step with no leaf-bearing tile in daylight runs no network and
    the plain step, which must still keep the leaf state the next step's
    features read.  A hybrid run from midnight into the morning (the first
    daylight step is 06:30) against the capture step, which keeps that state
    too, through the night and the hybrid from 06:30: with a network and a
    cap of one pass the morning's results depend on the features, and they
    give the same bits and the same fallbacks."""
    from cable_amd.surrogate import NIN, NOUT
    from oracle import mlp_oracle as O
    print("[hybrid] unlit steps into the morning", flush=True)
    wl = Workload(8192, 16)
    p = O.init_params(7).astype(np.float32)
    unit = np.concatenate([np.zeros(NIN), np.ones(NIN), np.zeros(NOUT), np.ones(NOUT)]).astype(np.float32)
    g = _ctx(wl)
    g.set_surrogate(p, unit)
    g.run_hybrid(0, 16, wl.dels, max_passes=1, resid_max=0.5, use_surrogate=True)
    ca = g.hybrid_counters()
    a = _outputs(g)
    g.close()
    g = _ctx(wl)
    g.set_surrogate(p, unit)
    g.run_capture(0, 13, wl.dels, 8192 * 13)
    g.run_hybrid(13, 3, wl.dels, max_passes=1, resid_max=0.5, use_surrogate=True)
    cb = g.hybrid_counters()
    b = _outputs(g)
    g.close()
    assert ca[0] == cb[0] and ca[0] > 0, (ca, cb)
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_daylight_dispatch_across_days_and_chunks():
    """A hybrid run of 70 steps in one call (from 22:30: night, a day, night
    and the next morning), whose daylight scan covers 64 steps at a time,
    against the same steps in 70 calls of one step each (every call's step
    keeps the leaf state): the same bits and the same fallbacks, so no step
    that a later daylight step reads drops the state.  The scan's 64th step
    is the last night step before a dawn (06:00), so the case where the next
    step lies beyond the scan is covered too."""
    from cable_amd.surrogate import NIN, NOUT
    from oracle import mlp_oracle as O
    print("[hybrid] daylight dispatch over 70 steps", flush=True)
    k0 = 45                                   # 22:30; step k0 + 63 is 06:00
    wl = Workload(3000, k0 + 70)
    p = O.init_params(13).astype(np.float32)
    unit = np.concatenate([np.zeros(NIN), np.ones(NIN), np.zeros(NOUT), np.ones(NOUT)]).astype(np.float32)
    g = _ctx(wl)
    g.set_surrogate(p, unit)
    g.run_hybrid(k0, 70, wl.dels, max_passes=1, resid_max=0.5, use_surrogate=True)
    ca = g.hybrid_counters()
    a = _outputs(g)
    g.close()
    g = _ctx(wl)
    g.set_surrogate(p, unit)
    fb = 0
    for k in range(k0, k0 + 70):
        g.run_hybrid(k, 1, wl.dels, max_passes=1, resid_max=0.5, use_surrogate=True)
        fb += g.hybrid_counters()[0]
    b = _outputs(g)
    g.close()
    assert ca[0] == fb and fb > 0, (ca, fb)
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_launch_timing_changes_only_the_timing():
    """Without the per-launch events a hybrid run gives the same bits and
    counters, and reports the whole run as step time."""
    from cable_amd.surrogate import NIN, NOUT
    from oracle import mlp_oracle as O
    print("[hybrid] launch timing", flush=True)
    wl = Workload(8192, 16)
    p = O.init_params(3).astype(np.float32)
    unit = np.concatenate([np.zeros(NIN), np.ones(NIN), np.zeros(NOUT), np.ones(NOUT)]).astype(np.float32)
    runs = []
    for timing in (True, False):
        g = _ctx(wl)
        g.set_surrogate(p, unit)
        g.set_launch_timing(timing)
        g.run_hybrid(12, 4, wl.dels, max_passes=2, resid_max=0.5, use_surrogate=True)
        runs.append((g.hybrid_counters(), _outputs(g), g.last_kernel_ms(), g.last_run_ms()))
        g.close()
    (c0, a, (s0, i0), _), (c1, b, (s1, i1), r1) = runs
    assert c0 == c1
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    assert s0 > 0 and i0 > 0
    assert i1 == 0 and s1 == pytest.approx(r1)


def test_hybrid_is_deterministic():
    """The fallback count is summed by atomics in whatever order the waves
    reach it; no tile's result depends on that order.  Two runs of the same
    daylight steps, with a cap that sends tiles to the fallback, give the
    same bits and the same counters."""
    from cable_amd.surrogate import NIN, NOUT
    from oracle import mlp_oracle as O
    print("[hybrid] determinism", flush=True)
    wl = Workload(8192, 24)
    p = O.init_params(9).astype(np.float32)
    unit = np.concatenate([np.zeros(NIN), np.ones(NIN), np.zeros(NOUT), np.ones(NOUT)]).astype(np.float32)
    runs = []
    for _ in range(2):
        g = _ctx(wl)
        g.set_surrogate(p, unit)
        g.run_hybrid(14, 10, wl.dels, max_passes=2, resid_max=0.5, use_surrogate=True)
        runs.append((g.hybrid_counters(), _outputs(g)))
        g.close()
    (c0, a), (c1, b) = runs
    assert c0 == c1 and c0[0] > 0
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_network_warm_start_cuts_fallbacks_and_passes():
    """Train on samples of one workload, run another: at the same pass cap the
    network's starting point leaves fewer tiles to the fallback, and with no
    cap it needs fewer leaf passes, than the reference's own start.  The
    network warm-starts sunlit leaf-bearing tiles only, so the comparison
    runs the daylight steps 12..35 (06:00 to 17:30)."""
    from cable_amd import surrogate as S
    print("[hybrid] network gain", flush=True)
    params, norm, info = S.train_on_gpu(n_tiles=65536, nsteps=48, epochs=2.0,
                                        log=lambda m: print(m, flush=True))
    assert info["loss_last"] < 0.2 * info["loss_first"]
    wl = Workload(65536, 36)
    res = {}
    for use in (True, False):
        for cap in (3, 20):
            g = _ctx(wl)
            g.set_surrogate(params, norm)
            g.run_hybrid(12, 24, wl.dels, max_passes=cap, resid_max=0.5, use_surrogate=use)
            fb, ts = g.hybrid_counters()
            res[(use, cap)] = (fb / ts, g.dryleaf_passes())
            g.close()
    print({f"surrogate={u} cap={c}": (round(v[0], 4), v[1]) for (u, c), v in res.items()}, flush=True)
    assert res[(True, 3)][0] < 0.5 * res[(False, 3)][0]
    assert res[(True, 20)][1] < res[(False, 20)][1]


def test_inference_kernel_matches_oracle():
    """The hybrid step's inference (features from the device state, then the
    LDS-staged matrix-core network) against the CPU oracle of the network on
    the same features, for every tile of a step."""
    from cable_amd.surrogate import NIN, NOUT, norm_stats
    from oracle import mlp_oracle as O
    print("[hybrid] inference vs oracle", flush=True)
    wl = Workload(3000, 2)                     # not a multiple of the 256-tile block
    g = _ctx(wl)
    p = O.init_params(11).astype(np.float32)
    unit = np.concatenate([np.zeros(NIN), np.ones(NIN), np.zeros(NOUT), np.ones(NOUT)]).astype(np.float32)
    g.set_surrogate(p, unit)
    x0, _ = g.surrogate_eval(0)
    assert np.isfinite(x0).all()
    rng = np.random.default_rng(12)
    norm = norm_stats(x0, rng.normal(3.0, 2.0, size=(x0.shape[0], NOUT)))
    g.set_surrogate(p, norm)
    x, pred = g.surrogate_eval(0)
    g.close()
    np.testing.assert_array_equal(x, x0)
    mu_x, sd_x = norm[:NIN], norm[NIN:2 * NIN]
    mu_y, sd_y = norm[2 * NIN:2 * NIN + NOUT], norm[2 * NIN + NOUT:]
    ref, _ = O.forward(p.astype(np.float64), (x.astype(np.float64) - mu_x) / sd_x, bf16_ops=True)
    got = (pred.astype(np.float64) - mu_y) / sd_y
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"inference max error / max |y| = {err:.2e}", flush=True)
    assert err < 2e-3
