"""Throughput of the CABLE timestep (cbm) on one MI355X, in tile-steps/s.

A step is one cbm pass over every tile of the synthetic workload: the
device-resident path (cable_gpu_run) with state, parameters, class tables and
the run's forcing already in HBM when the timed region starts.  The headline
line is the plain FP64 path on 262,144 tiles x 48 half-hourly steps; the other
forms -- FP64, mixed precision and the surrogate-warm-started hybrid at
1,048,576 tiles, with the mixed and hybrid fluxes' distance from FP64 after
the run and the surrogate's training on the GPU -- are reported beside it
under "forms".

  python bench.py [--gpus 1] [--steps 48] [--warmup 4] [--no-forms] [--no-cpu]

Prints one JSON line.  "roofline" rates the kernel against HBM bandwidth
using the algorithmic bytes per tile-step of the path (state read and
written once, forcing, per-tile parameters and the eight outputs once);
"cpu_baseline" times the reference's own cbm (built from its sources) on a
prefix of the same workload on one host core.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "tile-steps/sec on 1 MI355X: hybrid vs plain FP64, at ≤1 W/m² flux error"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes per tile-step: prognostic state read and written once,
# 8 forcing fields and the per-tile parameters read once, 8 outputs written
ALGO_BYTES = {"fp64": 1016, "mixed": 756, "hybrid": 764}
CPU_SAMPLE_TILES = 49152


HEADLINE = ["canopy%rnet", "canopy%fe", "canopy%fh", "canopy%fev", "canopy%fes", "canopy%fgpp",
            "ssnow%runoff", "rad%trad"]


# compared between forms: the headline fluxes, soil temperature (K) and soil
# moisture (m3/m3) of all six layers
COMPARED = HEADLINE + ["ssnow%tgg", "ssnow%wb"]


def flux_errors(got, ref):
    """max, 99.9th and 99th percentile and RMS of |got - ref| per flux
    (and per soil state over all layers), and the fraction of values more
    than 1 (W/m² for the energy fluxes) away."""
    out = {}
    for k in COMPARED:
        d = np.abs(np.asarray(got[k], dtype=np.float64) - np.asarray(ref[k], dtype=np.float64)).ravel()
        out[k] = {"max": float(d.max()), "p99.9": float(np.quantile(d, 0.999)), "p99": float(np.quantile(d, 0.99)),
                  "rms": float(np.sqrt(np.mean(d * d))), "frac_above_1": float((d > 1.0).mean())}
    return out


_COSTS = {}


def class_costs(wl, prec, log):
    """Per-vegetation-type step cost for the longest-first tile order,
    measured once per precision (outside every timed region)."""
    from cable_amd import gpu
    if prec not in _COSTS:
        t = time.perf_counter()
        _COSTS[prec] = gpu.class_costs(wl, prec)
        log(f"[bench] class costs ({'fp64' if prec == gpu.PREC_FP64 else 'mixed'}) measured in "
            f"{time.perf_counter() - t:.1f}s")
    return _COSTS[prec]


ORDER_LABEL = "grouped by vegetation type (measured most expensive first), then soil type and leaf area"


def run_form(n, steps, warmup, precision, log, class_order=True, hybrid=None, keep=False):
    from cable_amd import gpu
    from cable_amd.workload import Workload
    nsteps = max(steps, warmup, 1)
    t = time.perf_counter()
    wl = Workload(n, nsteps)
    log(f"[bench] workload {n} tiles x {nsteps} steps generated in {time.perf_counter() - t:.1f}s")
    prec = gpu.PREC_FP64 if precision == "fp64" else gpu.PREC_MIXED
    order = gpu.class_order(wl.iveg, wl.isoilm, wl.vlai, class_costs(wl, prec, log)) if class_order else None

    def go(g, k):
        if hybrid is None:
            g.run(0, k, wl.dels)
        else:
            g.set_surrogate(hybrid["params"], hybrid["norm"])
            g.run_hybrid(0, k, wl.dels, hybrid["max_passes"], hybrid["resid_max"])
    if warmup:
        g = gpu.from_workload(wl, precision=prec, full=False, order=order)
        go(g, warmup)
        g.sync()
        g.close()
    g = gpu.from_workload(wl, precision=prec, full=False, order=order)
    g.sync()
    t0 = time.perf_counter()
    go(g, steps)
    g.sync()
    wall = time.perf_counter() - t0
    ev_ms = g.last_run_ms()
    passes = g.dryleaf_passes()
    extra = {}
    if hybrid is not None:
        fb, ts = g.hybrid_counters()
        extra["fallback_fraction"] = fb / ts
    if keep:
        extra["_out"] = {k: g.download(k) for k in COMPARED}
    g.close()
    per_launch_s = ev_ms * 1e-3 / steps
    achieved = ALGO_BYTES[precision] * n / per_launch_s / 1e9
    log(f"[bench] {precision} n={n}: wall {wall * 1e3:.2f} ms for {steps} steps, "
        f"events {ev_ms:.2f} ms, {n * steps / wall:.4g} tile-steps/s")
    return dict({"tiles": n, "steps": steps, "value": n * steps / wall, "ms_per_step": wall * 1e3 / steps,
                 "kernel_ms": ev_ms / steps, "achieved_GBs": achieved, "dryleaf_passes": passes}, **extra)


def dropin_form(n, steps, log):
    """The drop-in step as the Fortran cbm replacement drives it: every field
    uploaded from host arrays, one cable_gpu_step, every state and diagnostic
    field downloaded -- the PCIe-inclusive rate of the host-buffer boundary."""
    from cable_amd import gpu
    from cable_amd.fields import FIELDS
    from cable_amd.workload import Workload
    wl = Workload(n, steps)
    host = {f.name: np.zeros(f.shape(n)) for f in FIELDS}
    for d in (wl.params, wl.state0):
        for k, v in d.items():
            if k in host:
                host[k] = np.asarray(v, dtype=np.float64).reshape(host[k].shape)
    g = gpu.CableGPU(n, gpu.PREC_FP64)
    p = wl.params
    g.set_globals(p["soil%zse"], p["soil%zshh"], p["bgc%ratecp"], p["bgc%ratecs"])
    up = dn = dev = 0.0
    for k in range(steps):
        for name, v in wl.forcing.items():
            if name in host:
                host[name] = np.asarray(v[k], dtype=np.float64).reshape(host[name].shape)
        t0 = time.perf_counter()
        for f in FIELDS:
            g.upload(f.name, host[f.name])
        t1 = time.perf_counter()
        g.step(k + 1, wl.dels)
        g.sync()
        t2 = time.perf_counter()
        for f in FIELDS:
            if f.role in ("STATE", "DIAG"):
                host[f.name] = g.download(f.name)
        t3 = time.perf_counter()
        up, dev, dn = up + t1 - t0, dev + t2 - t1, dn + t3 - t2
    g.close()
    wall = up + dev + dn
    log(f"[bench] drop-in fp64 n={n}: {wall / steps * 1e3:.1f} ms per step "
        f"(upload {up / steps * 1e3:.1f}, step {dev / steps * 1e3:.1f}, download {dn / steps * 1e3:.1f})")
    return {"tiles": n, "steps": steps, "value": n * steps / wall, "ms_per_step": wall * 1e3 / steps,
            "upload_ms": up * 1e3 / steps, "step_ms": dev * 1e3 / steps, "download_ms": dn * 1e3 / steps}


def profiled_traffic(tiles, precision):
    """HBM bytes per launch measured by rocprofv3 (FETCH_SIZE + WRITE_SIZE,
    scaled by a plain copy) for this workload, from the committed profile
    summary of the same kernel, or None when there is none."""
    path = os.path.join(ROOT, "profiles", f"cbm_kernel_{precision}_{tiles}x48.md")
    if not os.path.exists(path):
        return None, None
    with open(path) as fh:
        for line in fh:
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1]), os.path.relpath(path, ROOT)
    return None, None


def cpu_baseline(steps, log):
    """The reference's cbm (oracle/_ref, compiled from its sources) on the
    first CPU_SAMPLE_TILES tiles of the workload, one process on one core;
    the CPU restatement (oracle/) when the reference build is absent."""
    from cable_amd.workload import Workload
    wl = Workload(CPU_SAMPLE_TILES, steps)
    try:
        from oracle import refrun
        if refrun.available("r8"):
            _, sec = refrun.run(wl, steps, [], every_step=False, build="r8")
            kind, what = "reference", "reference cbm promoted to FP64 (amdflang -O2), driver step loop"
        else:
            raise FileNotFoundError
    except (FileNotFoundError, OSError):
        from oracle.oracle import OracleRun
        _, sec = OracleRun(wl).run(nsteps=steps, outputs=[], every_step=False)
        kind, what = "port", "FP64 C restatement of the reference cbm (gcc -O2)"
    v = CPU_SAMPLE_TILES * steps / sec
    log(f"[bench] cpu baseline ({kind}): {sec:.2f}s, {v:.4g} tile-steps/s")
    out = {"value": v, "unit": "tile-steps/s", "cores": 1, "kind": kind,
           "sample": f"first {CPU_SAMPLE_TILES} tiles of the workload x {steps} steps; {what}; "
                     f"one process; {sec:.2f} s"}
    if kind == "reference":
        out.update(_cpu_reference_more(wl, steps, log))
    return out


def _cpu_reference_more(wl, steps, log):
    """The shipped-precision reference on one core, and the promoted one on
    all host cores (one process per contiguous block of tiles, as the MPI
    workers split land points; cbm keeps module state, so no threads)."""
    import concurrent.futures as cf
    from oracle import refrun
    more = {}
    if refrun.available("r4"):
        _, sec = refrun.run(wl, steps, [], every_step=False, build="r4")
        more["shipped_precision_one_core"] = {"value": CPU_SAMPLE_TILES * steps / sec, "seconds": sec}
        log(f"[bench] cpu baseline shipped precision: {sec:.2f}s")
    p = max(1, min(16, os.cpu_count() or 1))
    blocks = np.array_split(np.arange(CPU_SAMPLE_TILES), p)
    with cf.ThreadPoolExecutor(p) as ex:
        secs = list(ex.map(lambda b: refrun.run(wl.subset(b), steps, [], every_step=False, build="r8")[1], blocks))
    more["all_cores"] = {"value": CPU_SAMPLE_TILES * steps / max(secs), "processes": p,
                         "seconds": max(secs)}
    log(f"[bench] cpu baseline on {p} processes: {max(secs):.2f}s")
    return more


TRAIN = {"n_tiles": 262144, "nsteps": 96, "epochs": 2.0, "max_samples": 16777216}
# the hybrid's polish passes per NITER pass and its leaf energy residual limit
# (W/m2).  Over the week-long run a cap of 3 was more accurate but slower
# than 4 (5.5% of tile-steps fall back instead of 0.4%)
HYB_CAP, HYB_RESID = 4, 0.5


def production(args, log):
    """The hybrid production run: 4,194,304 tiles x 336 half-hourly steps
    (one week) with forcing generated on the device in 48-step windows, run
    side by side in plain FP64, mixed precision and the surrogate-warm-started
    hybrid; returns the JSON line (value: the hybrid's tile-steps/s)."""
    from cable_amd import gpu, surrogate
    from cable_amd.workload import DEFAULT_SEED, Workload
    n, total, win = args.tiles, args.steps, 48
    t = time.perf_counter()
    wl = Workload(n, 1, forcing=False)
    log(f"[bench] production workload {n} tiles generated in {time.perf_counter() - t:.1f}s")
    params, norm, info = surrogate.train_on_gpu(log=log, **TRAIN)
    cal = Workload(65536, 8)        # calibration of the longest-first order
    o64 = gpu.class_order(wl.iveg, wl.isoilm, wl.vlai, class_costs(cal, gpu.PREC_FP64, log))
    o32 = gpu.class_order(wl.iveg, wl.isoilm, wl.vlai, class_costs(cal, gpu.PREC_MIXED, log))
    ctx = {"fp64": gpu.from_workload(wl, precision=gpu.PREC_FP64, full=False, order=o64),
           "mixed": gpu.from_workload(wl, precision=gpu.PREC_MIXED, full=False, order=o32),
           "hybrid": gpu.from_workload(wl, precision=gpu.PREC_MIXED, full=False, order=o32)}
    ctx["hybrid"].set_surrogate(params, norm)
    wall = {k: 0.0 for k in ctx}
    dev_ms = {k: 0.0 for k in ctx}
    passes = {k: 0 for k in ctx}
    fb = ts = 0
    synth_s = 0.0
    # the run ends at midnight; the last day is run in two halves so that the
    # forms are also compared at noon (the downloads are outside the timing)
    noon = {}
    for k0 in range(0, total, win):
        nk = min(win, total - k0)
        parts = [(k0, nk // 2), (k0 + nk // 2, nk - nk // 2)] if k0 + win >= total and nk >= 2 else [(k0, nk)]
        for name, g in ctx.items():
            t = time.perf_counter()
            g.synth_forcing(k0, nk, DEFAULT_SEED, wl.t0, wl.lat, dels=wl.dels)
            synth_s += time.perf_counter() - t
            for j, (p0, pn) in enumerate(parts):
                g.sync()
                t = time.perf_counter()
                if name == "hybrid":
                    g.run_hybrid(p0, pn, wl.dels, HYB_CAP, HYB_RESID)
                else:
                    g.run(p0, pn, wl.dels)
                g.sync()
                wall[name] += time.perf_counter() - t
                dev_ms[name] += g.last_run_ms()
                passes[name] += g.dryleaf_passes()
                if name == "hybrid":
                    f, s_ = g.hybrid_counters()
                    fb, ts = fb + f, ts + s_
                if len(parts) > 1 and j == 0:
                    noon[name] = {h: g.download(h) for h in COMPARED}
        log(f"[bench] production steps {k0}..{k0 + nk - 1} done")
    out = {k: {h: g.download(h) for h in COMPARED} for k, g in ctx.items()}
    for g in ctx.values():
        g.close()
    forms = {}
    for k in ctx:
        forms[f"{k}_{n}"] = {"tiles": n, "steps": total, "value": n * total / wall[k],
                             "ms_per_step": wall[k] * 1e3 / total, "kernel_ms": dev_ms[k] / total,
                             "dryleaf_passes": passes[k]}
    forms[f"mixed_{n}"]["flux_error_vs_fp64_after_run"] = flux_errors(out["mixed"], out["fp64"])
    forms[f"hybrid_{n}"]["flux_error_vs_fp64_after_run"] = flux_errors(out["hybrid"], out["fp64"])
    if noon:
        forms[f"mixed_{n}"]["flux_error_vs_fp64_noon_last_day"] = flux_errors(noon["mixed"], noon["fp64"])
        forms[f"hybrid_{n}"]["flux_error_vs_fp64_noon_last_day"] = flux_errors(noon["hybrid"], noon["fp64"])
    forms[f"hybrid_{n}"]["fallback_fraction"] = fb / ts
    forms[f"hybrid_{n}"]["max_passes"], forms[f"hybrid_{n}"]["resid_max_Wm2"] = HYB_CAP, HYB_RESID
    forms[f"hybrid_{n}"]["surrogate_training"] = info
    forms["hybrid_vs_fp64"] = wall["fp64"] / wall["hybrid"]
    forms["mixed_vs_fp64"] = wall["fp64"] / wall["mixed"]
    forms["forcing_synthesis_s"] = synth_s
    hyb = forms[f"hybrid_{n}"]
    achieved = ALGO_BYTES["hybrid"] * n / (hyb["kernel_ms"] * 1e-3) / 1e9
    return {
        "metric": METRIC, "value": hyb["value"], "unit": "tile-steps/s", "n_gpus": 1, "steps": total,
        "warmup": 0, "ms_per_step": hyb["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16+f32+f64",
        "data": "synthetic (forcing generated on the device, parameters from the reference's PFT and soil tables)",
        "config": {"workload": f"hybrid production run, {n} tiles x {total} half-hourly steps",
                   "tiles": n, "precision": "mixed + BF16 surrogate warm start", "path": "device-resident run",
                   "tile_order": ORDER_LABEL, "forcing_window_steps": win},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None},
        "cpu_baseline": None,
        "forms": forms,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=48)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--tiles", type=int, default=262144)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64",
                    help="path of the headline line (fp64: the plain FP64 step)")
    ap.add_argument("--no-forms", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--caller-order", action="store_true",
                    help="keep the tiles in the caller's order on the device (no class grouping)")
    ap.add_argument("--production", action="store_true",
                    help="the hybrid production run instead: 4,194,304 tiles x 336 steps by default "
                         "(--tiles, --steps), FP64, mixed and hybrid side by side")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a plain device-to-device copy of BYTES bytes (torch), "
                         "the reference kernel for scaling HBM counters, and exit")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        n = args.copy_probe // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": 4 * n, "copy_probe_bytes_written": 4 * n,
                          "launches": 3}), flush=True)
        return
    if args.gpus != 1:
        sys.stderr.write("bench.py: the library steps the tiles of one GPU; sharding across GPUs is not "
                         "implemented (tiles are independent). Run with --gpus 1.\n")
        sys.exit(2)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    if args.production:
        if args.tiles == 262144:
            args.tiles = 4194304
        if args.steps == 48:
            args.steps = 336
        print(json.dumps(production(args, log)), flush=True)
        return
    # the CPU leg first, while this process has not touched the GPU
    cpu = None if args.no_cpu else cpu_baseline(args.steps, log)
    co = not args.caller_order
    prec = args.precision
    head = run_form(args.tiles, args.steps, args.warmup, prec, log, co)
    forms = {f"{prec}_{args.tiles}": head}
    if not args.no_forms:
        from cable_amd import surrogate
        f64 = run_form(1048576, args.steps, args.warmup, "fp64", log, co, keep=True)
        mix = run_form(1048576, args.steps, args.warmup, "mixed", log, co, keep=True)
        # the surrogate: samples from the plain mixed kernel on a fresh workload
        # of the same distribution, trained by the GPU kernels
        params, norm, info = surrogate.train_on_gpu(log=log, **TRAIN)
        flops = 3 * 2 * (24 * 128 + 2 * 128 * 128 + 128 * 8) * info["batches"] * 2048
        info["training_TFLOPs"] = flops / info["train_s"] / 1e12
        hyb = run_form(1048576, args.steps, args.warmup, "mixed", log, co, keep=True,
                       hybrid={"params": params, "norm": norm, "max_passes": HYB_CAP, "resid_max": HYB_RESID})
        ref = f64.pop("_out")
        mix["flux_error_vs_fp64_after_run"] = flux_errors(mix.pop("_out"), ref)
        hyb["flux_error_vs_fp64_after_run"] = flux_errors(hyb.pop("_out"), ref)
        hyb["surrogate_training"] = info
        hyb["max_passes"], hyb["resid_max_Wm2"] = HYB_CAP, HYB_RESID
        forms["dropin_fp64_262144"] = dropin_form(262144, 3, log)
        forms["fp64_1048576"] = f64
        forms["mixed_1048576"] = mix
        forms["hybrid_1048576"] = hyb
        forms["mixed_vs_fp64_at_1048576"] = mix["value"] / f64["value"]
        forms["hybrid_vs_fp64_at_1048576"] = hyb["value"] / f64["value"]
    traffic, traffic_src = profiled_traffic(args.tiles, prec) if not args.caller_order else (None, None)
    what = {"fp64": "cbm plain FP64", "mixed": "cbm mixed precision (FP32 canopy, FP64 solves and balances)"}[prec]
    out = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "tile-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == "fp64" else "f32+f64",
        "data": "synthetic (seeded forcing, parameters from the reference's PFT and soil tables)",
        "config": {"workload": f"{what}, {args.tiles} tiles x {args.steps} half-hourly steps",
                   "tiles": args.tiles, "precision": prec, "path": "device-resident run",
                   "tile_order": "caller" if args.caller_order else ORDER_LABEL},
        "roofline": {"bound": "hbm", "achieved": head["achieved_GBs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": head["achieved_GBs"] / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_src},
        "cpu_baseline": cpu,
        "forms": forms,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
