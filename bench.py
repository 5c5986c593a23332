"""Throughput of the CABLE timestep (cbm) on one MI355X, in tile-steps/s.

A step is one cbm pass over every tile of the synthetic workload: the
device-resident path (cable_gpu_run, cable_gpu_run_hybrid) with state,
parameters, class tables and the run's forcing already in HBM when the timed
region starts.  The headline line is the metric's own form: the
surrogate-warm-started hybrid step (mixed precision, BF16 network trained on
the GPU beforehand) on 1,048,576 tiles, with the plain FP64 step on the same
tiles beside it ("hybrid_vs_fp64") and the flux error of the hybrid and mixed
forms against FP64 over every tile-step of the run ("gate").  The other forms
-- plain FP64 on 262,144 tiles, mixed precision, the surrogate's training and
the drop-in step through host buffers -- are under "forms".

  python bench.py [--gpus 1] [--steps 48] [--warmup 4] [--form hybrid|fp64|mixed]
                  [--tiles N] [--no-forms] [--no-cpu]

Prints one JSON line.  "roofline" rates the dominant kernel (the step kernel
of the headline form) against HBM bandwidth with the algorithmic bytes per
tile-step of the path (state read and written once, forcing, per-tile
parameters and the eight outputs once); the kernel is bound by vector-ALU
issue, not by HBM ("bound": "valu", figures from profiles/).
"cpu_baseline" times the reference's own cbm (built from its sources) on the
GPU box's host cores.
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

METRIC = "tile-steps/sec on 1 MI355X: hybrid vs plain FP64, at \u22641 W/m\u00b2 flux error"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes per tile-step: prognostic state read and written once,
# 8 forcing fields and the per-tile parameters read once, 8 outputs written
# (the hybrid also reads and writes the previous step's leaf state)
ALGO_BYTES = {"fp64": 1016, "mixed": 756, "hybrid": 764}
# the reference sample timed on the host: the first tiles of the headline
# workload, split over the host cores
CPU_SAMPLE_TILES = 65536
# the box's CPU share for one GPU
CPU_SHARE = 16

HEADLINE = ["canopy%rnet", "canopy%fe", "canopy%fh", "canopy%fev", "canopy%fes", "canopy%fgpp",
            "ssnow%runoff", "rad%trad"]
# the gate: 1 W/m2 for the energy fluxes, 0.1 umol CO2 m-2 s-1 for carbon
# (canopy%fgpp is in g C m-2 s-1: 1 umol = 12e-6 g C)
GATE_FLUXES = ["canopy%rnet", "canopy%fe", "canopy%fh", "canopy%fev", "canopy%fes", "canopy%fgpp"]
GC_PER_UMOL = 12.0e-6
GATE = dict({k: 1.0 for k in GATE_FLUXES[:5]}, **{"canopy%fgpp": 0.1})
# the reference's own spread over daylight leaf-bearing tile-steps, measured
# on 49,152 tiles x 48 steps of this workload (reference built from its
# sources, oracle/spread.py): p99.9 and fraction beyond the gate, for its
# shipped build against the promoted one and for a 0.1 K and a 1 K change of
# its leaf temperature start
REFERENCE_SPREAD = {
    "shipped_vs_promoted": {"canopy%rnet": (0.0059, 2.0e-5), "canopy%fe": (0.043, 1.3e-4),
                            "canopy%fh": (0.033, 1.3e-4), "canopy%fev": (0.029, 8.6e-5),
                            "canopy%fes": (0.012, 5.7e-5), "canopy%fgpp": (0.00063, 3.5e-5)},
    "leaf_start_0.1K": {"canopy%rnet": (4.83, 3.7e-3), "canopy%fe": (16.3, 1.48e-2),
                        "canopy%fh": (11.9, 1.26e-2), "canopy%fev": (11.1, 1.22e-2),
                        "canopy%fes": (7.62, 5.3e-3), "canopy%fgpp": (0.857, 4.3e-3)},
    "leaf_start_1K": {"canopy%rnet": (4.41, 3.24e-3), "canopy%fe": (15.3, 1.42e-2),
                      "canopy%fh": (11.7, 1.21e-2), "canopy%fev": (10.6, 1.15e-2),
                      "canopy%fes": (6.87, 4.9e-3), "canopy%fgpp": (1.013, 4.45e-3)},
}

# compared between forms: the headline fluxes, soil temperature (K) and soil
# moisture (m3/m3) of all six layers
COMPARED = HEADLINE + ["ssnow%tgg", "ssnow%wb"]

TRAIN = {"n_tiles": 262144, "nsteps": 96, "epochs": 2.0, "max_samples": 16777216}
# the hybrid's polish passes per NITER pass and its leaf energy residual limit
# (W/m2): after HYB_CAP passes a warm-started tile whose residual exceeds
# HYB_RESID falls back to the reference's iteration
HYB_CAP, HYB_RESID = 4, 0.5


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


ORDER_LABEL = ("grouped by vegetation type (measured most expensive first), then 4 K band of the tile's mean air "
               "temperature, soil type and leaf area")


def run_form(n, steps, warmup, precision, log, class_order=True, hybrid=None):
    """Time ``steps`` resident steps of ``precision`` ("fp64", "mixed"; with
    ``hybrid``, the hybrid method) on the n-tile workload, after ``warmup``
    untimed steps in another context of the same workload."""
    from cable_amd import gpu
    from cable_amd.workload import Workload
    nsteps = max(steps, warmup, 1)
    t = time.perf_counter()
    wl = Workload(n, nsteps)
    log(f"[bench] workload {n} tiles x {nsteps} steps generated in {time.perf_counter() - t:.1f}s")
    prec = gpu.PREC_FP64 if precision == "fp64" else gpu.PREC_MIXED
    order = (gpu.class_order(wl.iveg, wl.isoilm, wl.vlai, class_costs(wl, prec, log), tair=wl.t0) if class_order
             else None)

    def ctx():
        g = gpu.from_workload(wl, precision=prec, full=False, order=order)
        if hybrid is not None:
            g.set_surrogate(hybrid["params"], hybrid["norm"])
        g.sync()
        return g

    def go(g, k):
        if hybrid is None:
            g.run(0, k, wl.dels)
        else:
            g.run_hybrid(0, k, wl.dels, hybrid["max_passes"], hybrid["resid_max"])
    if warmup:
        g = ctx()
        go(g, warmup)
        g.sync()
        g.close()
    g = ctx()
    if hybrid is not None:
        g.set_launch_timing(False)
    t0 = time.perf_counter()
    go(g, steps)
    g.sync()
    wall = time.perf_counter() - t0
    ev_ms = g.last_run_ms()
    step_ms, infer_ms = g.last_kernel_ms()
    passes = g.dryleaf_passes()
    extra = {}
    if hybrid is not None:
        fb, ts = g.hybrid_counters()
        extra["fallback_fraction"] = fb / ts
        g.close()
        # the kernels' own device times: the same steps again from the same
        # start, with an event pair around each launch (the events cost the
        # run about 4%, so the timed run records none)
        g = ctx()
        go(g, steps)
        g.sync()
        step_ms, infer_ms = g.last_kernel_ms()
        extra["inference_ms"] = infer_ms / steps
        extra["kernel_times"] = "per-launch HIP events in a second, untimed run of the same steps"
    g.close()
    form = "hybrid" if hybrid is not None else precision
    achieved = ALGO_BYTES[form] * n / (step_ms * 1e-3 / steps) / 1e9
    log(f"[bench] {form} n={n}: wall {wall * 1e3:.2f} ms for {steps} steps, "
        f"events {ev_ms:.2f} ms (step kernels {step_ms:.2f}), {n * steps / wall:.4g} tile-steps/s")
    return dict({"tiles": n, "steps": steps, "value": n * steps / wall, "ms_per_step": wall * 1e3 / steps,
                 "device_ms_per_step": ev_ms / steps, "step_kernel_ms": step_ms / steps,
                 "achieved_GBs": achieved, "dryleaf_passes": passes}, **extra)


def _quantile(d, q):
    k = min(d.size - 1, int(np.ceil(q * d.size)) - 1)
    return float(np.partition(d, k)[k]) if d.size else 0.0


def gate_stats(n, steps, params, norm, log):
    """Flux error of the mixed and hybrid forms against plain FP64 at every
    tile-step of an n-tile, ``steps``-step run (untimed; the three forms are
    stepped side by side and downloaded after each step).  Energy fluxes in
    W/m2, canopy%fgpp in umol CO2 m-2 s-1; over every tile-step and over the
    daylight leaf-bearing ones (those the hybrid warm-starts)."""
    from cable_amd import gpu
    from cable_amd.workload import Workload
    t = time.perf_counter()
    wl = Workload(n, steps)
    order = gpu.class_order(wl.iveg, wl.isoilm, wl.vlai, tair=wl.t0)
    ctx = {"fp64": gpu.from_workload(wl, precision=gpu.PREC_FP64, full=False, order=order),
           "mixed": gpu.from_workload(wl, precision=gpu.PREC_MIXED, full=False, order=order),
           "hybrid": gpu.from_workload(wl, precision=gpu.PREC_MIXED, full=False, order=order)}
    ctx["hybrid"].set_surrogate(params, norm)
    diffs = {f: {k: [] for k in GATE_FLUXES} for f in ("mixed", "hybrid")}
    fsd = wl.forcing["met%fsd"][:, 0, :]
    day = (2.0 * fsd > 0.001) & (wl.iveg < 14)[None, :]
    for s in range(steps):
        ctx["fp64"].run(s, 1, wl.dels)
        ctx["mixed"].run(s, 1, wl.dels)
        ctx["hybrid"].run_hybrid(s, 1, wl.dels, HYB_CAP, HYB_RESID)
        ref = {k: ctx["fp64"].download(k).reshape(-1) for k in GATE_FLUXES}
        for f in ("mixed", "hybrid"):
            for k in GATE_FLUXES:
                d = np.abs(ctx[f].download(k).reshape(-1).astype(np.float64) - ref[k])
                if k == "canopy%fgpp":
                    d = d / GC_PER_UMOL
                diffs[f][k].append(d.astype(np.float32))
    for g in ctx.values():
        g.close()
    out = {"tiles": n, "steps": steps, "units": "W/m2; canopy%fgpp in umol CO2 m-2 s-1",
           "thresholds": GATE, "reference_spread_daylight_leafy": REFERENCE_SPREAD}
    for f in ("mixed", "hybrid"):
        out[f] = {}
        for sub, mask in (("all", None), ("daylight_leafy", day)):
            r = {}
            for k in GATE_FLUXES:
                d = np.stack(diffs[f][k])
                d = (d if mask is None else d[mask]).ravel()
                r[k] = {"max": float(d.max()), "p99.9": _quantile(d, 0.999), "p99": _quantile(d, 0.99),
                        "frac_beyond": float((d > GATE[k]).mean())}
            out[f][sub] = r
        # the gate as stated (every flux on every tile-step) and at the 99.9th
        # percentile of the daylight leaf-bearing tile-steps
        out[f]["every_value_within_gate"] = all(out[f]["all"][k]["max"] <= GATE[k] for k in GATE_FLUXES)
        out[f]["p99.9_within_gate"] = all(out[f]["daylight_leafy"][k]["p99.9"] <= GATE[k] for k in GATE_FLUXES)
    log(f"[bench] gate statistics over {n} tiles x {steps} steps in {time.perf_counter() - t:.1f}s")
    return out


def dropin_form(n, steps, log):
    """The drop-in boundary at n FP64 tiles: host arrays in, one
    cable_gpu_step, host arrays out -- the PCIe-inclusive rate of the
    host-buffer boundary -- in three transfer configurations, each direction
    one batched transfer through the context's pinned staging buffer:
      all        every field up, every state and diagnostic field down (the
                 host's arrays are the state, as in the reference);
      resident   after the first step only the forcing up (the device keeps
                 the state), every state and diagnostic field down;
      headline   the forcing up, the eight headline fluxes down.
    The first step of each (every field up) is set-up and not timed."""
    from cable_amd import gpu
    from cable_amd.fields import FIELDS
    from cable_amd.workload import Workload
    wl = Workload(n, steps + 1)
    host0 = {f.name: np.zeros(f.shape(n)) for f in FIELDS}
    for d in (wl.params, wl.state0):
        for k, v in d.items():
            if k in host0:
                host0[k] = np.asarray(v, dtype=np.float64).reshape(host0[k].shape)
    written = [f.name for f in FIELDS if f.role in ("STATE", "DIAG")]
    headline = ["canopy%rnet", "canopy%fe", "canopy%fh", "canopy%fev", "canopy%fes", "canopy%fgpp",
                "ssnow%runoff", "rad%trad"]
    p = wl.params
    out = {"tiles": n, "steps": steps}
    for mode in ("all", "resident", "headline"):
        host = {k: v.copy() for k, v in host0.items()}
        g = gpu.CableGPU(n, gpu.PREC_FP64)
        g.set_globals(p["soil%zse"], p["soil%zshh"], p["bgc%ratecp"], p["bgc%ratecs"])
        down = headline if mode == "headline" else written
        up = dev = dn = 0.0
        for k in range(steps + 1):
            for name, v in wl.forcing.items():
                if name in host:
                    host[name] = np.asarray(v[k], dtype=np.float64).reshape(host[name].shape)
            t0 = time.perf_counter()
            if mode == "all" or k == 0:
                g.upload_fields(host)
            else:
                g.upload_fields({name: host[name] for name in wl.forcing if name in host})
            t1 = time.perf_counter()
            g.step(k + 1, wl.dels)
            g.sync()
            t2 = time.perf_counter()
            g.download_fields(down, out=host)
            t3 = time.perf_counter()
            if k > 0:
                up, dev, dn = up + t1 - t0, dev + t2 - t1, dn + t3 - t2
        g.close()
        wall = up + dev + dn
        log(f"[bench] drop-in fp64 n={n} transfers={mode}: {wall / steps * 1e3:.1f} ms per step "
            f"(upload {up / steps * 1e3:.1f}, step {dev / steps * 1e3:.1f}, download {dn / steps * 1e3:.1f})")
        out[mode] = {"value": n * steps / wall, "ms_per_step": wall * 1e3 / steps, "upload_ms": up * 1e3 / steps,
                     "step_ms": dev * 1e3 / steps, "download_ms": dn * 1e3 / steps,
                     "fields_down": len(down)}
    return out


def profiled(name):
    """Figures of the committed profile summary profiles/<name>.md (lines
    'key: value'), or {} when there is none."""
    path = os.path.join(ROOT, "profiles", name + ".md")
    out = {}
    if os.path.exists(path):
        with open(path) as fh:
            for line in fh:
                k, sep, v = line.partition(":")
                if sep and k.strip() in ("traffic_bytes_per_launch", "valu_issue_utilisation", "fp32_valu_frac",
                                         "fp64_valu_frac", "mfma_busy_frac", "mfma_bf16_frac"):
                    try:
                        out[k.strip()] = float(v)
                    except ValueError:
                        pass
        out["source"] = os.path.relpath(path, ROOT)
    return out


def cpu_baseline(steps, log):
    """The reference's cbm (oracle/_ref/r8, promoted to FP64, compiled from
    its sources) on the host: the first CPU_SAMPLE_TILES tiles of the
    workload, one process per host core of the box's share, each pinned to
    its core and owning a contiguous block of tiles (as the MPI workers
    split land points, src/offline/cable_mpimaster.F90:1399-1487; cbm keeps
    module state, so no threads); the reference's own configuration 0
    (4,096 tiles x 48 steps) the same way; one core alone, promoted and
    shipped precision.  The CPU restatement (oracle/) stands in when the
    reference build is absent."""
    import concurrent.futures as cf

    from cable_amd.workload import Workload
    from oracle import refrun
    cores = sorted(os.sched_getaffinity(0))[:CPU_SHARE]
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            model = next((l.split(":", 1)[1].strip() for l in fh if l.startswith("model name")), model)
    except OSError:
        pass
    host = {"cpu_model": model, "host_cores": os.cpu_count(), "cores_used": len(cores)}
    if not refrun.available("r8"):
        from oracle.oracle import OracleRun
        wl = Workload(4096, steps)
        _, sec = OracleRun(wl).run(nsteps=steps, outputs=[], every_step=False)
        return dict(host, value=4096 * steps / sec, unit="tile-steps/s", cores=1, kind="port",
                    sample=f"first 4096 tiles x {steps} steps, FP64 C restatement of the reference cbm, one core")

    def spread(wl, build="r8"):
        blocks = np.array_split(np.arange(wl.n), len(cores))
        with cf.ThreadPoolExecutor(len(cores)) as ex:
            secs = list(ex.map(lambda bc: refrun.run(wl.subset(bc[0]), wl.nsteps, [], every_step=False, build=build,
                                                     cpu=bc[1])[1], zip(blocks, cores)))
        return max(secs)

    wl = Workload(CPU_SAMPLE_TILES, steps)
    sec = spread(wl)
    v = CPU_SAMPLE_TILES * steps / sec
    log(f"[bench] cpu baseline: reference on {len(cores)} pinned processes, {sec:.2f}s, {v:.4g} tile-steps/s")
    out = dict(host, value=v, unit="tile-steps/s", cores=len(cores), kind="reference",
               sample=f"first {CPU_SAMPLE_TILES} tiles of the workload x {steps} steps; reference cbm promoted to "
                      f"FP64 (amdflang -O2), driver step loop; {len(cores)} processes of contiguous tile blocks, "
                      f"each pinned to one core ({model}); slowest process {sec:.2f} s")
    cfg0 = Workload(4096, 48)
    s0 = spread(cfg0)
    out["cfg0_4096x48"] = {"value": 4096 * 48 / s0, "seconds": s0, "processes": len(cores)}
    log(f"[bench] cpu baseline cfg 0 (4096 x 48): {s0:.3f}s on {len(cores)} processes")
    one = Workload(16384, steps)
    _, s1 = refrun.run(one, steps, [], every_step=False, build="r8", cpu=cores[0])
    out["one_core"] = {"value": 16384 * steps / s1, "seconds": s1, "sample": f"16384 tiles x {steps} steps"}
    if refrun.available("r4"):
        _, s4 = refrun.run(one, steps, [], every_step=False, build="r4", cpu=cores[0])
        out["one_core_shipped_precision"] = {"value": 16384 * steps / s4, "seconds": s4}
    log(f"[bench] cpu baseline one core: promoted {s1:.2f}s" + (f", shipped {s4:.2f}s" if refrun.available("r4") else ""))
    return out


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
    o64 = gpu.class_order(wl.iveg, wl.isoilm, wl.vlai, class_costs(cal, gpu.PREC_FP64, log), tair=wl.t0)
    o32 = gpu.class_order(wl.iveg, wl.isoilm, wl.vlai, class_costs(cal, gpu.PREC_MIXED, log), tair=wl.t0)
    ctx = {"fp64": gpu.from_workload(wl, precision=gpu.PREC_FP64, full=False, order=o64),
           "mixed": gpu.from_workload(wl, precision=gpu.PREC_MIXED, full=False, order=o32),
           "hybrid": gpu.from_workload(wl, precision=gpu.PREC_MIXED, full=False, order=o32)}
    ctx["hybrid"].set_surrogate(params, norm)
    # no per-launch events in the timed runs (they cost the hybrid about 4%):
    # its step time is then the whole step, inference included
    ctx["hybrid"].set_launch_timing(False)
    wall = {k: 0.0 for k in ctx}
    dev_ms = {k: 0.0 for k in ctx}
    step_ms = {k: 0.0 for k in ctx}
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
                step_ms[name] += g.last_kernel_ms()[0]
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
                             "ms_per_step": wall[k] * 1e3 / total, "device_ms_per_step": dev_ms[k] / total,
                             "step_kernel_ms": step_ms[k] / total,
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
    achieved = ALGO_BYTES["hybrid"] * n / (hyb["step_kernel_ms"] * 1e-3) / 1e9
    return {
        "metric": METRIC, "value": hyb["value"], "unit": "tile-steps/s", "n_gpus": 1, "steps": total,
        "warmup": 0, "ms_per_step": hyb["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16+f32+f64",
        "data": "synthetic (forcing generated on the device, parameters from the reference's PFT and soil tables)",
        "config": {"workload": f"hybrid production run, {n} tiles x {total} half-hourly steps",
                   "tiles": n, "precision": "mixed + BF16 surrogate warm start", "path": "device-resident run",
                   "tile_order": ORDER_LABEL, "forcing_window_steps": win},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "the hybrid's whole step, step and inference kernels (events around each run)"},
        "cpu_baseline": None,
        "forms": forms,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=48)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--form", choices=["hybrid", "fp64", "mixed"], default="hybrid",
                    help="form of the headline line (hybrid: the metric's own form)")
    ap.add_argument("--tiles", type=int, default=None,
                    help="tiles of the headline form (default 1,048,576; the production run 4,194,304)")
    ap.add_argument("--no-forms", action="store_true", help="only the headline form (for profiling)")
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
        args.tiles = args.tiles or 4194304
        if args.steps == 48:
            args.steps = 336
        print(json.dumps(production(args, log)), flush=True)
        return
    n = args.tiles or 1048576
    co = not args.caller_order
    # the CPU leg first, while this process has not touched the GPU
    cpu = None if args.no_cpu else cpu_baseline(args.steps, log)
    forms = {}
    if args.form == "hybrid":
        from cable_amd import surrogate
        # the surrogate: samples from the plain mixed kernel on a fresh
        # workload of the same distribution, trained by the GPU kernels
        params, norm, info = surrogate.train_on_gpu(log=log, **TRAIN)
        flops = 3 * 2 * (24 * 128 + 2 * 128 * 128 + 128 * 8) * info["batches"] * 2048
        info["training_TFLOPs"] = flops / info["train_s"] / 1e12
        hyb = {"params": params, "norm": norm, "max_passes": HYB_CAP, "resid_max": HYB_RESID}
        head = run_form(n, args.steps, args.warmup, "mixed", log, co, hybrid=hyb)
        head.update(max_passes=HYB_CAP, resid_max_Wm2=HYB_RESID, surrogate_training=info)
        forms[f"hybrid_{n}"] = head
        f64 = run_form(n, args.steps, args.warmup, "fp64", log, co)
        forms[f"fp64_{n}"] = f64
    else:
        head = run_form(n, args.steps, args.warmup, args.form, log, co)
        forms[f"{args.form}_{n}"] = head
    if not args.no_forms:
        if args.form == "hybrid":
            mix = run_form(n, args.steps, args.warmup, "mixed", log, co)
            forms[f"mixed_{n}"] = mix
            forms[f"mixed_vs_fp64_at_{n}"] = mix["value"] / f64["value"]
        forms["fp64_262144"] = run_form(262144, args.steps, args.warmup, "fp64", log, co)
        forms["dropin_fp64_262144"] = dropin_form(262144, 3, log)
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
    }
    if args.form == "hybrid":
        out["hybrid_vs_fp64"] = head["value"] / f64["value"]
        out["gate"] = gate_stats(n, args.steps, params, norm, log)
        what = (f"hybrid cbm step (mixed precision, BF16 surrogate warm start of the leaf fixed point, "
                f"{HYB_CAP} polish passes, fallback above {HYB_RESID} W/m2), {n} tiles x {args.steps} "
                f"half-hourly steps from midnight, plain FP64 on the same tiles beside it")
        dtype, kern = "bf16+f32+f64", "cbm_kernel_hybrid"
    else:
        what = {"fp64": "cbm plain FP64", "mixed": "cbm mixed precision (FP32 canopy, FP64 solves and balances)"}[
            args.form] + f", {n} tiles x {args.steps} half-hourly steps"
        dtype, kern = ("f64" if args.form == "fp64" else "f32+f64"), f"cbm_kernel_{args.form}"
    prof = profiled(f"{kern}_{n}x{args.steps}") if co else {}
    out.update({
        "dtype": dtype,
        "data": "synthetic (seeded forcing, parameters from the reference's PFT and soil tables)",
        "config": {"workload": what, "tiles": n, "form": args.form, "path": "device-resident run",
                   "tile_order": "caller" if args.caller_order else ORDER_LABEL},
        "roofline": {"bound": "valu", "achieved": head["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": head["achieved_GBs"] / HBM_PEAK_GBS, "traffic": prof.get("traffic_bytes_per_launch"),
                     "kernel": ("step kernels of the headline form, every step's launch (per-launch HIP events on "
                                "the context stream); traffic and valu: the hybrid step kernel of the daylight steps"
                                if args.form == "hybrid" else
                                "step kernel of the headline form (per-launch HIP events on the context stream)"),
                     "valu": {k: prof[k] for k in ("valu_issue_utilisation", "fp32_valu_frac", "fp64_valu_frac")
                              if k in prof},
                     "mfma": ({"kernel": "the surrogate's inference (mlp_infer_kernel), daylight launches",
                               "busy_frac": prof["mfma_busy_frac"], "bf16_frac": prof["mfma_bf16_frac"]}
                              if "mfma_bf16_frac" in prof else None),
                     "profile": prof.get("source")},
        "cpu_baseline": cpu,
        "forms": forms,
    })
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
