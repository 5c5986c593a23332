"""The CPU oracle (oracle/cbm_oracle.c) pinned against the reference.

* bit-for-bit against the golden vectors written by the compiled reference
  (tests/golden/cbm_ref_77x28.npz, tests/golden/make_golden.py);
* bit-for-bit against the compiled reference itself on a 612-tile, 48-step
  run with deep, thin and no snow, when oracle/_ref is built (here only);
* the Thomas solve against a dense solve, and the balance closures the
  reference's own energy/water checks (src/offline/cable_checks.F90:472-618)
  compute, which must vanish to rounding in FP64.
"""
import ctypes
import os

import numpy as np
import pytest

from conftest import golden, snowy
from cable_amd.fields import FIELDS
from cable_amd.workload import Workload
from oracle import refrun
from oracle.oracle import OracleRun, lib

REF_NAMES = [f.name for f in FIELDS if f.role in ("STATE", "DIAG") and not f.name.startswith("drv%")]


def _mismatch(a, b):
    a = np.asarray(a).reshape(-1)
    b = np.asarray(b).reshape(-1)
    same = (a == b) | (np.isnan(a) & np.isnan(b))
    return int((~same).sum())


def test_oracle_reproduces_golden_vectors():
    case, exp = golden()
    out, _ = OracleRun(case).run(nsteps=case.nsteps, outputs=REF_NAMES)
    bad = []
    for k in REF_NAMES:
        got = out[k][[i - 1 for i in case.full_steps]]
        if _mismatch(got, exp["r8"][k]):
            bad.append(k)
    for k, v in exp["r8h"].items():
        if _mismatch(out[k], v):
            bad.append(k + " (every step)")
    assert not bad, f"oracle differs from the reference's golden vectors in {bad[:10]}"


@pytest.mark.skipif(not refrun.available("r8"), reason="oracle/_ref/r8 not built (no reference here)")
def test_oracle_matches_compiled_reference_48_steps():
    wl = snowy(Workload(612, 48))
    ref, _ = refrun.run(wl, 48, REF_NAMES, every_step=True, build="r8")
    ora, _ = OracleRun(wl).run(nsteps=48, outputs=REF_NAMES)
    bad = [k for k in REF_NAMES if _mismatch(ora[k], ref[k])]
    assert not bad, f"fields differing: {bad[:10]}"


def _trimb(a, b, c, r):
    L = lib()
    L.oracle_trimb.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int]
    x = np.ascontiguousarray(r, dtype=np.float64).copy()
    arrs = [np.ascontiguousarray(v, dtype=np.float64) for v in (a, b, c)]
    L.oracle_trimb(*[v.ctypes.data for v in arrs], x.ctypes.data, len(x))
    return x


@pytest.mark.parametrize("kmax", [6, 7, 9])
def test_trimb_is_a_tridiagonal_solve(kmax):
    rng = np.random.default_rng(kmax)
    for _ in range(50):
        a = rng.uniform(-1, 1, kmax)
        c = rng.uniform(-1, 1, kmax)
        b = np.abs(a) + np.abs(c) + rng.uniform(0.5, 2, kmax)
        r = rng.uniform(-10, 10, kmax)
        A = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
        x = _trimb(a, b, c, r)
        np.testing.assert_allclose(x, np.linalg.solve(A, r), rtol=1e-12, atol=1e-12)


def test_balances_close_in_fp64():
    """Energy and radiation budgets close to rounding at FP64; the water
    budget closes to 1e-4 mm on snow-free land tiles once the driver's
    previous-step water store exists (step 2 on).  Lakes (iveg 16) are topped
    up by cbm itself and snow tiles carry snow water outside wbtot, so the
    reference's own check does not close there."""
    wl = snowy(Workload(306, 12))
    names = ["bal%ebal", "bal%radbal", "bal%drybal", "bal%wetbal", "bal%wbal",
             "canopy%rnet", "canopy%fe", "canopy%fh", "canopy%ga"]
    out, _ = OracleRun(wl).run(nsteps=12, outputs=names)
    scale = max(np.abs(out[k]).max() for k in ("canopy%rnet", "canopy%fe", "canopy%fh", "canopy%ga"))
    assert scale > 100.0
    for k in ("bal%ebal", "bal%radbal", "bal%drybal", "bal%wetbal"):
        assert np.abs(out[k]).max() < 1e-12 * scale * 100, k
    land = (wl.state0["ssnow%snowd"] == 0) & (wl.iveg != 16)
    assert np.abs(out["bal%wbal"][1:, land]).max() < 1e-4


def test_oracle_catches_a_planted_mistake():
    """The golden comparison is sensitive: a 1e-12 relative change to one
    parameter (vcmax) shows up in the outputs of a daylight step."""
    case, exp = golden()
    case.params = {k: v.copy() for k, v in case.params.items()}
    case.params["veg%vcmax"] *= 1.0 + 1e-12
    step = case.full_steps[1]        # a daylight step
    out, _ = OracleRun(case).run(nsteps=step, outputs=["canopy%fpn"])
    assert _mismatch(out["canopy%fpn"][step - 1], exp["r8"]["canopy%fpn"][1]) >= 1


def test_workload_is_a_prefix_and_consistent():
    a, b = Workload(40, 4), Workload(90, 4)
    for d1, d2 in ((a.params, b.params), (a.state0, b.state0), (a.forcing, b.forcing)):
        for k, v in d1.items():
            w = d2[k]
            if v.shape == w.shape:
                np.testing.assert_array_equal(v, w, err_msg=k)
            else:
                np.testing.assert_array_equal(v, w[..., :40], err_msg=k)
    # ice: iveg 17 <=> isoilm 9 (cable_parameters.F90:2467-2650)
    w = Workload(2000, 1)
    assert np.array_equal(w.iveg == 17, w.isoilm == 9)
    f = w.forcing
    assert f["met%fld"].min() >= 250 and f["met%fld"].max() <= 420
    assert f["met%pmb"].min() >= 950 and f["met%pmb"].max() <= 1020
    assert np.all(f["met%fsd"][:, 0] == f["met%fsd"][:, 1])


REF_NML = "/root/reference/src/offline"


def _parse_nml(path, prefix):
    import re
    txt = open(path).read()
    out = {}
    for m in re.finditer(prefix + r"%(\w+)\s*=\s*([^%/]*?)(?=\n\s*\w+%|\n/|\Z)", txt, re.S):
        k = m.group(1).lower()
        if k == "desc":
            continue
        vals = []
        for tok in m.group(2).replace("\n", " ").split(","):
            tok = tok.strip()
            if not tok:
                continue
            if "*" in tok:
                n, x = tok.split("*")
                vals += [float(x)] * int(n)
            else:
                vals.append(float(tok))
        out[k] = vals
    return out


@pytest.mark.skipif(not os.path.isdir(REF_NML), reason="reference namelists not present")
def test_class_tables_match_reference_namelists():
    from cable_amd import tables
    v = _parse_nml(os.path.join(REF_NML, "pft_params.nml"), "vegin")
    s = _parse_nml(os.path.join(REF_NML, "cable_soilparm.nml"), "soilin")
    assert len(v) > 40 and len(s) > 10
    bad = [k for k, x in v.items() if tables.VEG.get(k) != x]
    bad += [k for k, x in s.items() if tables.SOIL.get(k) != x]
    assert not bad, bad
