import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libcable_gpu.so)")


def snowy(wl, every=3, deep=150.0, thin=5.0):
    """Give a workload's tiles a mix of deep (3-layer) snow, thin snow and none."""
    n = wl.n
    sn = np.zeros(n)
    sn[::every] = deep
    sn[1::every] = thin
    wl.state0["ssnow%snowd"] = sn
    return wl


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = np.abs(a - b)
    s = np.maximum(np.abs(a), np.abs(b))
    with np.errstate(invalid="ignore", divide="ignore"):
        r = np.where(d == 0, 0.0, d / s)
    return r
