"""Restart files in the reference's variables (cable_amd/restart.py): a
state round-trips through the reference's restart names, dimensions and
kinds.  CPU only."""
import numpy as np
import pytest

from cable_amd import restart as R
from cable_amd.fields import BY_NAME, zero_fields
from cable_amd.workload import Workload, add_snow


def _state(n):
    wl = add_snow(Workload(n, 1))
    f = zero_fields(n)
    for src in (wl.params, wl.state0):
        for k, v in src.items():
            if k in f:
                f[k][...] = v
    rng = np.random.default_rng(3)
    # diagnostics a restart also carries: give them distinct values
    for name, field in R.RESTART:
        if field in f and not np.any(f[field]):
            integer = BY_NAME[field].prec == "I"
            f[field][...] = (rng.integers(0, 2, size=f[field].shape) if integer
                             else rng.uniform(0.1, 2.0, size=f[field].shape))
    f["soil%zse"] = np.asarray(wl.params["soil%zse"])
    return f


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_state_round_trips_through_the_reference_restart_names(tmp_path, precision):
    from scipy.io import netcdf_file
    n = 97
    f = _state(n)
    path = str(tmp_path / "restart.nc")
    R.write_restart(path, R.restart_variables(f, n, precision), n)
    with netcdf_file(path, "r", mmap=False) as nc:
        # the reference's restart grid: patch dimension "mp", the others by name
        assert nc.dimensions["mp"] == n and nc.dimensions["soil"] == 6 and nc.dimensions["snow"] == 3
        assert nc.variables["wb"].dimensions == ("soil", "mp")
        assert nc.variables["tggsn"].dimensions == ("snow", "mp")
        assert nc.variables["albsoilsn"].dimensions == ("rad", "mp")
        assert nc.variables["cplant"].dimensions == ("plant_carbon_pools", "mp")
        assert nc.variables["csoil"].dimensions == ("soil_carbon_pools", "mp")
        assert nc.variables["trad"].dimensions == ("mp",)
        # native kinds: integers, REAL(r_2) in FP64, default REAL per build
        assert nc.variables["isflag"].data.dtype.kind == "i"
        assert nc.variables["wb"].data.dtype.itemsize == 8
        assert nc.variables["tgg"].data.dtype.itemsize == (8 if precision == 0 else 4)
        names = set(nc.variables)
    assert names == {name for name, _ in R.RESTART}
    assert not names & set(R.NOT_CARRIED)
    back = R.fields_from_restart(R.read_restart(path))
    for name, field in R.RESTART:
        want = np.asarray(f[field], dtype=np.float64).reshape(back[field].shape)
        if precision == 1 and field != "soil%zse" and BY_NAME[field].prec == "F":
            want = want.astype(np.float32).astype(np.float64)
        np.testing.assert_array_equal(back[field], want, err_msg=name)
