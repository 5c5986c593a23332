"""The GPU's step code on the CPU: tests/host_step.hip compiles step_tile
(cable_amd/csrc/cbm_kernel.h, the per-tile body of the timestep kernel) for
the host, and these tests run it on host arrays

* against the oracle, step by step from identical state, in the drop-in
  (every field) and resident (compact set, class tables) layouts;
* under AddressSanitizer and UndefinedBehaviorSanitizer (host code only);
* under MemorySanitizer, which reports any branch or output that depends
  on a Tile member the gather did not set.

No GPU: these are CPU tests.
"""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from conftest import snowy
from cable_amd.fields import FIELDS
from cable_amd.workload import Workload
from oracle import parity

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "host_step.hip")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CLANG = os.path.join(os.path.dirname(os.path.realpath(HIPCC)), "..", "llvm", "bin", "clang++")
BASE = ["-x", "hip", "--cuda-host-only", "-std=c++17", "-g", "-O1", "-fno-omit-frame-pointer",
        "-Wno-duplicate-decl-specifier"]
# compile flags (the sanitizers instrument the host code only: each -fsanitize=
# follows -Xarch_host) and link flags.  The program needs no HIP runtime, so it
# is linked without it: the runtime's own (uninstrumented) start-up would
# otherwise trip MemorySanitizer before main
SANITIZERS = {
    "plain": ([], []),
    "asan_ubsan": (["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
                    "-Xarch_host", "-fno-sanitize-recover=all"], ["-fsanitize=address,undefined"]),
    "msan": (["-Xarch_host", "-fsanitize=memory", "-Xarch_host", "-fsanitize-memory-track-origins"],
             ["-fsanitize=memory"]),
}

pytestmark = pytest.mark.skipif(not (os.path.exists(HIPCC) and os.path.exists(CLANG)), reason="hipcc not found")


@pytest.fixture(scope="module")
def builds(tmp_path_factory):
    d = tmp_path_factory.mktemp("host_step")
    out = {}
    for name, (cflags, lflags) in SANITIZERS.items():
        obj, exe = str(d / f"host_step_{name}.o"), str(d / f"host_step_{name}")
        r = subprocess.run([HIPCC] + BASE + cflags + ["-c", SRC, "-o", obj], capture_output=True, text=True)
        if r.returncode == 0:
            r = subprocess.run([CLANG] + lflags + [obj, "-o", exe], capture_output=True, text=True)
        out[name] = exe if r.returncode == 0 else r.stderr[-2000:]
    return out


def _tables():
    from cable_amd import gpu
    wl = Workload(1, 1)
    stab = dict(wl.stab)
    stab["zse"] = wl.params["soil%zse"]
    return gpu.class_tables(wl.vtab, stab)


def _write_input(path, wl, nsteps, full, precision, states):
    V, S = _tables()
    p = wl.params
    glob = np.concatenate([p["soil%zse"], p["soil%zshh"], p["bgc%ratecp"], p["bgc%ratecs"]]).astype(np.float64)
    with open(path, "wb") as f:
        f.write(struct.pack("<qiiid", wl.n, nsteps, int(full), precision, float(wl.dels)))
        f.write(glob.tobytes())
        f.write(np.ascontiguousarray(V, dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(S, dtype=np.float64).tobytes())
        for s, st in enumerate(states):
            f.write(struct.pack("<ii", s + 1, 1 if s == 0 else 0))
            for fl in FIELDS:
                f.write(np.ascontiguousarray(st[fl.name], dtype=np.float64).tobytes())


def _read_output(path, n, nsteps):
    data = np.fromfile(path, dtype=np.float64)
    out, pos = [], 0
    for _ in range(nsteps):
        d = {}
        for fl in FIELDS:
            m = fl.n * n
            d[fl.name] = data[pos:pos + m]
            pos += m
        out.append(d)
    assert pos == data.size
    return out


def _oracle_steps(wl, nsteps):
    """The oracle's state before and after each step."""
    from oracle.oracle import OracleRun
    orun = OracleRun(wl)
    pre, post = [], []
    for s in range(nsteps):
        fk = {k: v[s] for k, v in wl.forcing.items() if k in orun.f}
        for k, v in fk.items():
            orun.f[k][...] = v
        pre.append({k: v.copy() for k, v in orun.f.items()})
        orun.step(fk)
        post.append({k: v.reshape(-1).copy() for k, v in orun.f.items()})
    return pre, post


def _run(exe, tmp_path, wl, nsteps, full, precision, pre, env=None):
    """full: True/1 the drop-in layout, False/0 the resident one, 2 the
    hybrid method (mixed precision)."""
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _write_input(fin, wl, nsteps, full, precision, pre)
    r = subprocess.run([exe, fin, fout], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    return _read_output(fout, wl.n, nsteps), r.stderr


# the resident run writes the prognostic state and these fluxes
RESIDENT_OUT = [f.name for f in FIELDS if f.role == "STATE" and f.name != "ssnow%wbliq"] + [
    "canopy%rnet", "canopy%fe", "canopy%fh", "canopy%fev", "canopy%fes", "canopy%fgpp", "ssnow%runoff",
    "rad%trad"]
COMPARE = [f.name for f in FIELDS if f.role in ("STATE", "DIAG") and f.name != "drv%owb"]


@pytest.fixture(scope="module")
def case():
    wl = snowy(Workload(306, 48))
    pre, post = _oracle_steps(wl, 48)
    return wl, pre, post


def _worst(got, want, names):
    w = []
    for s in range(len(want)):
        for k in names:
            e = parity.worst(got[s][k], want[s][k], k)
            w.append((e[0], k, s + 1))
    return max(w)


def test_host_build_of_the_device_step_matches_the_oracle(builds, case, tmp_path):
    """The device physics, built for the host, against the oracle: every
    value of every state and diagnostic field within 1e-10 relative plus
    its floor, at each of 48 steps of 306 tiles with deep, thin and no snow;
    the resident layout gives the same bits as the drop-in layout."""
    exe = builds["plain"]
    assert os.path.exists(exe), exe
    wl, pre, post = case
    full, _ = _run(exe, tmp_path, wl, 48, True, 0, pre)
    worst = _worst(full, post, COMPARE)
    print(f"host build vs oracle: largest |host - oracle| / (1e-10 |oracle| + floor) {worst[0]:.3f} "
          f"({worst[1]}, step {worst[2]})")
    assert worst[0] <= 1.0, worst
    res, _ = _run(exe, tmp_path, wl, 48, False, 0, pre)
    diff = [(s + 1, k) for s in range(48) for k in RESIDENT_OUT
            if not np.array_equal(res[s][k][:wl.n] if k == "ssnow%gammzz" else res[s][k],
                                  full[s][k][:wl.n] if k == "ssnow%gammzz" else full[s][k])]
    assert not diff, diff[:10]


@pytest.mark.parametrize("precision,full", [(0, 1), (0, 0), (1, 1), (1, 0), (1, 2)],
                         ids=["fp64-dropin", "fp64-resident", "mixed-dropin", "mixed-resident", "hybrid"])
def test_device_step_under_address_and_undefined_behaviour_sanitizers(builds, case, tmp_path, precision, full):
    exe = builds["asan_ubsan"]
    if not os.path.exists(exe):
        pytest.skip("sanitizer build unavailable: " + exe[-300:])
    wl, pre, _ = case
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    steps = slice(12, 24) if full == 2 else slice(0, 12)
    _, err = _run(exe, tmp_path, wl, 12, full, precision, pre[steps], env=env)
    assert "runtime error" not in err and "ERROR: AddressSanitizer" not in err, err[-3000:]


@pytest.mark.parametrize("precision,full", [(0, 1), (0, 0), (1, 1), (1, 0), (1, 2)],
                         ids=["fp64-dropin", "fp64-resident", "mixed-dropin", "mixed-resident", "hybrid"])
def test_device_step_reads_no_unset_tile_member(builds, case, tmp_path, precision, full):
    """MemorySanitizer: no branch, address or stored output of the step
    depends on a value the gather left unset (the resident layout loads only
    the compact set)."""
    exe = builds["msan"]
    if not os.path.exists(exe):
        pytest.skip("MemorySanitizer build unavailable: " + exe[-300:])
    wl, pre, _ = case
    env = dict(os.environ, MSAN_OPTIONS="halt_on_error=1")
    # steps 12-23: daylight from 06:00, where the hybrid warm-starts and falls back
    steps = slice(12, 24) if full == 2 else slice(0, 12)
    _, err = _run(exe, tmp_path, wl, 12, full, precision, pre[steps], env=env)
    assert "MemorySanitizer" not in err, err[-3000:]
    if full == 2:
        assert int(err.split("fallbacks")[1].split()[0]) > 0, err[-500:]
