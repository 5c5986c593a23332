"""The reference's own spread: how far the reference's results move under
changes that leave its method unchanged.  Test infrastructure (the yardstick
the GPU forms' flux errors are read against); nothing that ships uses it.

Two measurements on the same synthetic workload (cable_amd/workload.py):

* precision: the reference as shipped (default REAL = FP32, REAL(r_2) = FP64;
  oracle/_ref/r4) against the reference promoted to FP64 (oracle/_ref/r8);
* start: the FP64 oracle (bit-for-bit the promoted reference) with the leaf
  temperature the canopy's fixed point starts from (tlfx = tlfy = met%tk,
  src/science/canopy/cable_canopy.F90:213-214) raised by a tiny offset,
  against the same oracle without it.  The fixed point
  (src/science/canopy/cbl_dryLeaf.F90:239-606) stops at the first iterate
  within 0.1 K of the previous one (:565-606), so its result depends on where
  it starts.

Both compare the eight headline fluxes at every step.  Runs are split over
tile blocks in parallel processes (tiles are independent).

  python -m oracle.spread [--tiles 49152] [--steps 48] [--procs 8] [--offset 1e-9]

prints one JSON object.
"""
from __future__ import annotations

import argparse
import json
import multiprocessing as mp_
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HEADLINE = ["canopy%rnet", "canopy%fe", "canopy%fh", "canopy%fev", "canopy%fes", "canopy%fgpp",
            "ssnow%runoff", "rad%trad"]
ENERGY = ["canopy%rnet", "canopy%fe", "canopy%fh", "canopy%fev", "canopy%fes"]
# canopy%fgpp is in g C m-2 s-1; 1 umol CO2 m-2 s-1 = 12e-6 g C m-2 s-1
GC_PER_UMOL = 12.0e-6
# the gate: 1 W/m2 for the energy fluxes, 0.1 umol/m2/s for carbon
GATE = dict({k: 1.0 for k in ENERGY}, **{"canopy%fgpp": 0.1 * GC_PER_UMOL})


def in_gate_units(name, d):
    """|difference| in the gate's units: W/m2, umol/m2/s for fgpp, else as is."""
    return d / GC_PER_UMOL if name == "canopy%fgpp" else d


def error_stats(a, b, names=HEADLINE, daylight=None):
    """Per field: max, p99.9, p99, RMS of |a - b| and the fraction of values
    beyond the gate, over every tile-step (arrays (steps, tiles)); fgpp in
    umol/m2/s.  ``daylight`` (steps, tiles) bool restricts the statistics."""
    out = {}
    for k in names:
        d = np.abs(np.asarray(a[k], dtype=np.float64) - np.asarray(b[k], dtype=np.float64))
        if daylight is not None:
            d = d[daylight]
        d = in_gate_units(k, d.ravel())
        r = {"max": float(d.max()), "p99.9": float(np.quantile(d, 0.999)), "p99": float(np.quantile(d, 0.99)),
             "rms": float(np.sqrt(np.mean(d * d)))}
        if k in GATE:
            r["frac_beyond_gate"] = float((d > in_gate_units(k, GATE[k])).mean())
        out[k] = r
    return out


def _ref_block(args):
    from cable_amd.workload import Workload
    from oracle import refrun
    n, steps, idx, build = args
    wl = Workload(n, steps).subset(idx)
    out, _ = refrun.run(wl, steps, HEADLINE, every_step=True, build=build)
    return out


def _oracle_block(args):
    import ctypes

    from cable_amd.workload import Workload
    from oracle import oracle as O
    n, steps, idx, offset = args
    L = O.lib()
    L.oracle_set_leaf_start_offset.argtypes = [ctypes.c_double]
    L.oracle_set_leaf_start_offset(offset)
    wl = Workload(n, steps).subset(idx)
    out, _ = O.OracleRun(wl).run(nsteps=steps, outputs=HEADLINE, every_step=True)
    L.oracle_set_leaf_start_offset(0.0)
    return out


def _gather(parts):
    return {k: np.concatenate([p[k].reshape(p[k].shape[0], -1) for p in parts], axis=1) for k in HEADLINE}


def _run_blocks(fn, n, steps, procs, extra):
    blocks = np.array_split(np.arange(n), procs)
    with mp_.get_context("fork").Pool(procs) as pool:
        parts = pool.map(fn, [(n, steps, b, extra) for b in blocks])
    return _gather(parts)


def daylight_mask(n, steps):
    """Tile-steps with sunlight on a leaf-bearing tile (shortwave above the
    step's radiation threshold and iveg < 14)."""
    from cable_amd.workload import Workload
    wl = Workload(n, steps)
    fsd = wl.forcing["met%fsd"][:, 0, :]
    return (2.0 * fsd > 0.001) & (wl.iveg < 14)[None, :]


def measure(n=49152, steps=48, procs=8, offset=1e-9, log=print):
    res = {"tiles": n, "steps": steps, "processes": procs}
    t = time.perf_counter()
    r8 = _run_blocks(_ref_block, n, steps, procs, "r8")
    r4 = _run_blocks(_ref_block, n, steps, procs, "r4")
    log(f"[spread] reference r8 and r4 in {time.perf_counter() - t:.1f}s")
    t = time.perf_counter()
    o0 = _run_blocks(_oracle_block, n, steps, procs, 0.0)
    o1 = _run_blocks(_oracle_block, n, steps, procs, offset)
    log(f"[spread] oracle with and without the start offset in {time.perf_counter() - t:.1f}s")
    day = daylight_mask(n, steps)
    res["oracle_equals_r8"] = bool(all(np.array_equal(o0[k], r8[k]) for k in HEADLINE))
    res["shipped_vs_promoted"] = {"all": error_stats(r4, r8), "daylight_leafy": error_stats(r4, r8, daylight=day),
                                  "last_step": error_stats({k: v[-1:] for k, v in r4.items()},
                                                           {k: v[-1:] for k, v in r8.items()})}
    res["leaf_start_offset_K"] = offset
    res["start_offset_vs_none"] = {"all": error_stats(o1, o0), "daylight_leafy": error_stats(o1, o0, daylight=day),
                                   "last_step": error_stats({k: v[-1:] for k, v in o1.items()},
                                                            {k: v[-1:] for k, v in o0.items()})}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--tiles", type=int, default=49152)
    ap.add_argument("--steps", type=int, default=48)
    ap.add_argument("--procs", type=int, default=8)
    ap.add_argument("--offset", type=float, default=1e-9)
    a = ap.parse_args()
    print(json.dumps(measure(a.tiles, a.steps, a.procs, a.offset,
                             log=lambda m: print(m, file=sys.stderr, flush=True))), flush=True)


if __name__ == "__main__":
    main()
