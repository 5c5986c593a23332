import os
import sys
from types import SimpleNamespace

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libcable_gpu.so)")


from cable_amd.workload import add_snow as snowy  # noqa: E402

GOLDEN = os.path.join(ROOT, "tests", "golden", "cbm_ref_77x28.npz")


def golden():
    """The reference's golden vectors (tests/golden/make_golden.py) as
    (case, expected): ``case`` has the attributes of a Workload that the
    oracle and the GPU wrapper read; ``expected[prefix][field]`` holds the
    reference outputs (prefix r8: all fields after the steps in
    ``case.full_steps``; r8h, r4h: headline fluxes after every step; r4:
    state after the last step)."""
    z = np.load(GOLDEN)
    inputs = {k[3:]: z[k] for k in z.files if k.startswith("in:")}
    forcing = {k[2:]: z[k] for k in z.files if k.startswith("f:")}
    case = SimpleNamespace(n=int(inputs["veg%iveg"].size), nsteps=int(z["meta:nsteps"]),
                           dels=float(z["meta:dels"]), params=inputs, state0={},
                           forcing=forcing, full_steps=[int(x) for x in z["meta:full_steps"]])
    expected = {}
    for k in z.files:
        pre, _, name = k.partition(":")
        if pre in ("r8", "r8h", "r4h", "r4"):
            expected.setdefault(pre, {})[name] = z[k]
    return case, expected


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = np.abs(a - b)
    s = np.maximum(np.abs(a), np.abs(b))
    with np.errstate(invalid="ignore", divide="ignore"):
        r = np.where(d == 0, 0.0, d / s)
    return r
