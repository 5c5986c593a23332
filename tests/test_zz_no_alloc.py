"""The hybrid run that allocates nothing (cable_gpu_run_hybrid_noalloc)
against cable_gpu_run_hybrid on the same inputs, through the C ABI: the same
kernels on the same inputs, so the same bits and counters, on two calls of
the allocation-free entry point in a row, the second with a smaller context
(whose scratch then still holds the first call's predictions)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NAMES = ["canopy%rnet", "canopy%fe", "canopy%fh", "canopy%fev", "canopy%fes", "canopy%fgpp",
         "ssnow%runoff", "rad%trad", "ssnow%tgg", "ssnow%wb", "canopy%cansto"]


def _pair(n, k0, nsteps, params, norm):
    from cable_amd import gpu
    from cable_amd.workload import Workload
    wl = Workload(n, k0 + nsteps)
    order = gpu.class_order(wl.iveg, wl.isoilm, wl.vlai)
    out = []
    for noalloc in (False, True):
        g = gpu.from_workload(wl, precision=gpu.PREC_MIXED, full=False, order=order)
        g.set_surrogate(params, norm)
        g.run(0, k0, wl.dels)                    # the night, plain
        if noalloc:
            g.run_hybrid_noalloc(k0, nsteps, wl.dels, 4, 0.5)
        else:
            g.run_hybrid(k0, nsteps, wl.dels, 4, 0.5)
        out.append(({k: g.download(k) for k in NAMES}, g.hybrid_counters(), g.dryleaf_passes()))
        g.close()
    return out


def test_hybrid_without_allocation_matches_the_hybrid_run():
    from cable_amd import gpu
    from cable_amd import surrogate as S
    from cable_amd.workload import Workload
    print("[noalloc] training a small network", flush=True)
    params, norm, _ = S.train_on_gpu(n_tiles=65536, nsteps=48, epochs=1.0)
    for n, k0, nsteps in ((65536, 12, 8), (3000, 14, 4)):
        print(f"[noalloc] {n} tiles, steps {k0}..{k0 + nsteps - 1}", flush=True)
        (a, ca, pa), (b, cb, pb) = _pair(n, k0, nsteps, params, norm)
        assert ca == cb and pa == pb, (ca, cb, pa, pb)
        assert ca[0] > 0                          # some tiles fell back
        diff = [k for k in NAMES if not np.array_equal(a[k], b[k])]
        assert not diff, diff
    # the network's state must exist: without cable_gpu_set_surrogate the
    # call is refused before any work
    wl = Workload(256, 2)
    g = gpu.from_workload(wl, precision=gpu.PREC_MIXED, full=False)
    with pytest.raises(gpu.CableError):
        g.run_hybrid_noalloc(0, 1, wl.dels, 4, 0.5, use_surrogate=True)
    g.close()
