"""The Fortran drop-in for cbm (cable_amd/fortran/cbm_gpu.F90).

oracle/_ref/gpu8 and gpu4 are the reference's own cbm closure and driver
built with cbm_gpu.F90 in place of src/offline/cbl_model_driver_offline.F90,
so there the reference's code calls the HIP library through the
ISO_C_BINDING layer.  The GPU tests run them on the golden case and compare
with the reference's own outputs stored in tests/golden.
"""
import os

import numpy as np
import pytest

from conftest import golden
from cable_amd.fields import FIELDS
from oracle import parity, refrun

REF_NAMES = [f.name for f in FIELDS if f.role in ("STATE", "DIAG") and not f.name.startswith("drv%")]


def test_generated_bindings_are_current():
    from cable_amd.fortran import gen_fields
    with open(gen_fields.OUT) as fh:
        assert fh.read() == gen_fields.generate(), "run python cable_amd/fortran/gen_fields.py"


# balance diagnostics the offline driver's cable_checks computes after cbm,
# and met%ofsd, which it sets before cbm; no file under the reference's
# src/science touches them
CHECKS = {"bal%wbal", "bal%wbal_tot", "bal%precip_tot", "bal%rnoff_tot", "bal%evap_tot", "bal%radbal",
          "bal%ebal", "bal%ebalsoil", "bal%ebalveg", "bal%ebal_tot", "bal%radbalsum", "met%ofsd"}


def test_bindings_cover_every_field():
    """The drop-in moves every field cbm reads or writes, and none of the
    driver's own: its state (drv%, sum_flux%) and the balance diagnostics
    it computes itself (moving those would overwrite the host's running
    totals with the device's stale copies)."""
    from cable_amd.fortran import gen_fields
    txt = gen_fields.generate()
    for f in FIELDS:
        if f.name.startswith(("drv%", "sum_flux%")) or f.name in CHECKS:
            assert f"C_LOC({f.name})" not in txt, f.name
            continue
        assert f"C_LOC({f.name})" in txt, f.name
    assert all(n in {f.name for f in FIELDS} for n in CHECKS)


def _flang():
    import shutil
    for c in (shutil.which("amdflang"), "/opt/rocm/llvm/bin/amdflang"):
        if c and os.path.exists(c):
            return c
    return None


@pytest.mark.skipif(_flang() is None, reason="no Fortran compiler")
def test_fortran_met_type_has_the_c_layout(tmp_path):
    """TYPE(cable_gpu_met) of the ISO_C_BINDING module and struct
    cable_gpu_met of include/cable_gpu.h: same size and field offsets
    (the ctypes mirror follows the C rules)."""
    import ctypes
    import subprocess
    from cable_amd import gpu
    src = tmp_path / "layout.F90"
    src.write_text("""PROGRAM layout
USE, INTRINSIC :: ISO_C_BINDING
USE cable_gpu_c, ONLY: cable_gpu_met
IMPLICIT NONE
TYPE(cable_gpu_met), TARGET :: m
INTEGER(C_INTPTR_T) :: b
b = TRANSFER(C_LOC(m), b)
WRITE(*, '(I0)') C_SIZEOF(m)
WRITE(*, '(I0)') TRANSFER(C_LOC(m%tair_offset), b) - b
WRITE(*, '(I0)') TRANSFER(C_LOC(m%qair_is_rh), b) - b
WRITE(*, '(I0)') TRANSFER(C_LOC(m%latitude), b) - b
WRITE(*, '(I0)') TRANSFER(C_LOC(m%sdoy), b) - b
WRITE(*, '(I0)') TRANSFER(C_LOC(m%syear), b) - b
WRITE(*, '(I0)') TRANSFER(C_LOC(m%fixed_co2), b) - b
END PROGRAM layout
""")
    mod = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cable_amd", "fortran",
                       "cable_gpu_c.F90")
    exe = tmp_path / "layout"
    libdir = os.path.dirname(gpu.LIBPATH)
    gpu.lib()  # built
    subprocess.run([_flang(), "-J", str(tmp_path), "-o", str(exe), mod, str(src), f"-L{libdir}", "-lcable_gpu",
                    f"-Wl,-rpath,{libdir}"], check=True, cwd=tmp_path)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    M = gpu.MetVars
    want = [ctypes.sizeof(M)] + [getattr(M, f).offset for f in
                                 ("tair_offset", "qair_is_rh", "latitude", "sdoy", "syear", "fixed_co2")]
    assert got == want


def _scaled_err(got, exp, name):
    """largest |got - exp| / (1e-10 |exp| + floor) of a field
    (oracle/parity.py): at most 1 where every value passes"""
    return parity.worst(got, exp, name)[0]


@pytest.mark.gpu
@pytest.mark.skipif(not refrun.available("gpu8"), reason="oracle/_ref/gpu8 not built")
def test_reference_driver_through_fortran_dropin_fp64():
    """Step 1 of the promoted reference driver with the GPU cbm matches the
    CPU reference's golden outputs within 1e-10; the later stored steps
    within the drift bounds of the plain path."""
    print("[fortran] gpu8 driver on the golden case", flush=True)
    case, exp = golden()
    out, _ = refrun.run(case, case.nsteps, REF_NAMES, every_step=True, build="gpu8", timeout=300)
    worst = sorted(((_scaled_err(out[k][0], exp["r8"][k][0], k), k) for k in REF_NAMES), reverse=True)
    print("worst step-1 fields:", worst[:5], flush=True)
    assert worst[0][0] <= 1.0, worst[:5]
    for j, step in enumerate(case.full_steps[1:], start=1):
        for k, bound in (("canopy%fe", 1e-4), ("canopy%fh", 1e-4), ("canopy%rnet", 1e-4),
                         ("canopy%fgpp", 1e-8), ("ssnow%tgg", 1e-5), ("ssnow%wb", 1e-7)):
            d = np.abs(out[k][step - 1] - exp["r8"][k][j]).max()
            assert d <= bound, (step, k, d)


@pytest.mark.gpu
@pytest.mark.skipif(not refrun.available("gpu8"), reason="oracle/_ref/gpu8 not built")
def test_resident_transfers_give_the_same_run():
    """CABLE_GPU_SYNC=resident: after the first call the drop-in uploads only
    what the offline driver writes between calls (forcing, oldcansto, vlai)
    and the device keeps the rest -- the same bits as uploading every field
    each call.  With a download list as well, the listed fields still come
    back bit for bit (the host's copies of the others go stale, which only
    the host's own balance diagnostics would see)."""
    print("[fortran] gpu8 driver, resident transfers", flush=True)
    case, _ = golden()
    ref, _ = refrun.run(case, case.nsteps, REF_NAMES, every_step=True, build="gpu8", timeout=300)
    res, _ = refrun.run(case, case.nsteps, REF_NAMES, every_step=True, build="gpu8", timeout=300,
                        env={"CABLE_GPU_SYNC": "resident"})
    for k in REF_NAMES:
        np.testing.assert_array_equal(res[k], ref[k], err_msg=k)
    few = ["canopy%fe", "canopy%fh", "canopy%rnet", "canopy%fgpp", "ssnow%tgg", "ssnow%wb", "rad%trad"]
    lst, _ = refrun.run(case, case.nsteps, few, every_step=True, build="gpu8", timeout=300,
                        env={"CABLE_GPU_SYNC": "resident", "CABLE_GPU_DOWNLOAD": ",".join(few)})
    for k in few:
        np.testing.assert_array_equal(lst[k], ref[k], err_msg=k)


@pytest.mark.gpu
@pytest.mark.skipif(not refrun.available("gpu4"), reason="oracle/_ref/gpu4 not built")
def test_shipped_precision_driver_through_fortran_dropin():
    """The shipped-precision build (host REAL = FP32, REAL(r_2) = FP64) with
    the GPU cbm: the drop-in selects the mixed device path, which keeps the
    same precision split, so its fluxes sit as close to the FP32 reference
    as the FP32 reference sits to the FP64 one, and within the 1 W/m2
    target of the FP64 reference."""
    print("[fortran] gpu4 driver on the golden case", flush=True)
    case, exp = golden()
    names = list(exp["r8h"])
    out, _ = refrun.run(case, case.nsteps, names, every_step=True, build="gpu4", timeout=300)
    for k in ("canopy%rnet", "canopy%fe", "canopy%fh", "canopy%fev", "canopy%fes"):
        d_gpu = np.abs(out[k] - exp["r8h"][k]).max()
        d_r4 = np.abs(exp["r4h"][k] - exp["r8h"][k]).max()
        d_gr4 = np.abs(out[k] - exp["r4h"][k]).max()
        print(f"{k}: |gpu4 - r8| {d_gpu:.3e}  |gpu4 - r4| {d_gr4:.3e}  |r4 - r8| {d_r4:.3e}", flush=True)
        assert d_gpu <= 1.0, (k, d_gpu)
        assert d_gr4 <= max(2.0 * d_r4, 0.05), (k, d_gr4, d_r4)


@pytest.mark.skipif(not refrun.available("gpu8"), reason="oracle/_ref/gpu8 not built")
@pytest.mark.parametrize("switch,value", [("gs_switch", "medlyn"), ("icycle", "1"), ("or_evap", "T"),
                                          ("soil_struc", "sli")])
def test_dropin_refuses_a_configuration_it_does_not_implement(switch, value):
    """The reference driver built with the GPU drop-in, run with one switch
    away from the shipped configuration: the first cbm call stops (ERROR STOP
    999) naming the switch, before any device call -- so this runs without a
    GPU."""
    import subprocess
    case, _ = golden()
    with pytest.raises(subprocess.CalledProcessError) as e:
        refrun.run(case, 1, [], every_step=False, build="gpu8", switches={switch: value}, timeout=120)
    out = (e.value.stdout or b"").decode() + (e.value.stderr or b"").decode()
    assert "does not implement" in out and switch.split("_")[0] in out, out[-500:]
