"""Run the compiled reference cbm() (oracle/_ref) on a synthetic workload.

Test infrastructure only: tests/, the golden-vector script and bench.py's
cpu_baseline leg use it.  Each call runs the driver in a fresh process
(the reference keeps first-call state in SAVE variables, e.g.
src/science/soilsnow/cbl_soilsnow_main.F90:60-62, so one process is one run).
"""
from __future__ import annotations

import os
import resource
import struct
import subprocess
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")


def _big_stack():
    _, hard = resource.getrlimit(resource.RLIMIT_STACK)
    resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))


def driver_path(build: str = "r8") -> str:
    return os.path.join(REF_DIR, build, "cable_ref_driver")


def available(build: str = "r8") -> bool:
    return os.access(driver_path(build), os.X_OK)


def _name(s: str) -> bytes:
    b = s.encode()
    assert len(b) <= 40
    return b + b" " * (40 - len(b))


def _set(f, name, arr):
    a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
    f.write(struct.pack("<i", 1) + _name(name) + struct.pack("<q", a.size))
    f.write(a.tobytes())


def run(wl, nsteps=None, outputs=(), every_step=True, build="r8",
        extra_sets=None, timeout=3600, poison=None, cpu=None, switches=None, env=None):
    """Run ``nsteps`` steps of the reference on workload ``wl``.

    outputs: field names to read back; with every_step they are read after
    each step (result[name] has a leading step axis), else after the last.
    poison: optional (step, name, value) -- overwrite a field before that
    step (0-based), used to find which fields carry state between steps.
    cpu: optional host core to pin the driver process to (taskset).
    switches: optional {switch: value} away from the shipped configuration
    (the driver's op 7), set before the first step.
    env: optional extra environment variables of the driver process.
    Returns (dict, seconds the driver spent in its step operations -- cbm
    and the driver contract around it -- measured inside the driver process,
    without its file input and output).
    """
    nsteps = wl.nsteps if nsteps is None else nsteps
    with tempfile.TemporaryDirectory() as td:
        pin, pout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(pin, "wb") as f:
            f.write(struct.pack("<q", wl.n))
            for k, v in wl.params.items():
                _set(f, k, v)
            for k, v in wl.state0.items():
                _set(f, k, v)
            for k, v in (extra_sets or {}).items():
                _set(f, k, v)
            f.write(struct.pack("<i", 2))
            for k, v in (switches or {}).items():
                f.write(struct.pack("<i", 7) + _name(k) + _name(str(v)))
            fr = wl.forcing if wl.forcing is not None else wl.make_forcing(0, nsteps)
            for s in range(nsteps):
                for k, v in fr.items():
                    _set(f, k, v[s])
                if poison is not None and poison[0] == s:
                    _set(f, poison[1], poison[2])
                f.write(struct.pack("<iid", 3, s + 1, float(wl.dels)))
                if every_step or s == nsteps - 1:
                    for nm in outputs:
                        f.write(struct.pack("<i", 4) + _name(nm))
            f.write(struct.pack("<i", 6))
            f.write(struct.pack("<i", 0))
        t0 = time.perf_counter()
        # the reference keeps its per-step temporaries on the stack (automatic
        # arrays of mp elements), so a large mp needs the stack limit raised,
        # as the offline model is run with `ulimit -s unlimited`
        pin_cmd = ["taskset", "-c", str(cpu)] if cpu is not None else []
        subprocess.run(pin_cmd + [driver_path(build), pin, pout], check=True, timeout=timeout,
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, preexec_fn=_big_stack,
                       env=None if env is None else {**os.environ, **env})
        dt = time.perf_counter() - t0
        res = {}
        with open(pout, "rb") as f:
            data = f.read()
        pos = 0
        while pos < len(data):
            nm = data[pos:pos + 40].decode().strip()
            n = struct.unpack_from("<q", data, pos + 40)[0]
            arr = np.frombuffer(data, dtype=np.float64, count=n, offset=pos + 48).copy()
            pos += 48 + 8 * n
            res.setdefault(nm, []).append(arr)
    step_seconds = float(res.pop("drv%step_seconds")[0][0])
    out = {}
    for nm, lst in res.items():
        out[nm] = np.stack(lst) if every_step else lst[-1]
    return out, step_seconds
