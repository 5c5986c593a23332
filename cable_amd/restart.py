"""Restart files in the reference's variables.

The reference writes its restart file from the output variables registered
with restart=.true. (src/offline/cable_diagnostics.F90; the stream is written
by cable_output_write_restart, src/util/output/cable_output.F90:373-376) and
reads it back in get_restart_data (src/offline/cable_initialise.F90:163-).
Each variable is named by its field_name and laid out on the restart grid:
the patch dimension is "mp" (src/util/output/cable_output_common.F90:176-180),
the others keep their names ("soil", "snow", "rad", "plant_carbon_pools",
"soil_carbon_pools"), and each is stored in its native kind.

This module maps the library's fields to those variables and writes and
reads them as a NetCDF (classic format) file.  The variables the reference's
restart holds that this path does not carry are left out: bal%wbtot0 and
bal%osnowd0 (initial totals of the balance diagnostics, outside cbm),
ssnow%GWwb and the SLI variables (groundwater and SLI are not built), and the
grid description (patchfrac, nap, mvtype, mstype).  A run loaded from such a
file starts as a reference run restarted from its restart file does: the
next step is a run's first step (cbl_soilsnow_main.F90:60-62); the library's
own saved state (cable_gpu_save_state) is the exact continuation.
"""
from __future__ import annotations

import numpy as np

from .fields import BY_NAME

# (restart variable, field), in the order get_restart_data reads them
RESTART = [
    ("tgg", "ssnow%tgg"), ("wb", "ssnow%wb"), ("wbice", "ssnow%wbice"), ("gammzz", "ssnow%gammzz"),
    ("tss", "ssnow%tss"), ("ssdnn", "ssnow%ssdnn"), ("ssdn", "ssnow%ssdn"), ("osnowd", "ssnow%osnowd"),
    ("smass", "ssnow%smass"), ("sdepth", "ssnow%sdepth"), ("tggsn", "ssnow%tggsn"), ("snage", "ssnow%snage"),
    ("snowd", "ssnow%snowd"), ("rtsoil", "ssnow%rtsoil"), ("isflag", "ssnow%isflag"),
    ("albsoilsn", "ssnow%albsoilsn"), ("rnof1", "ssnow%rnof1"), ("rnof2", "ssnow%rnof2"),
    ("runoff", "ssnow%runoff"), ("cansto", "canopy%cansto"), ("sghflux", "canopy%sghflux"),
    ("ghflux", "canopy%ghflux"), ("ga", "canopy%ga"), ("dgdtg", "canopy%dgdtg"), ("fev", "canopy%fev"),
    ("fes", "canopy%fes"), ("fhs", "canopy%fhs"), ("cplant", "bgc%cplant"), ("csoil", "bgc%csoil"),
    ("albedo", "rad%albedo"), ("trad", "rad%trad"), ("iveg", "veg%iveg"), ("isoil", "soil%isoilm"),
    ("albsoil", "soil%albsoil"), ("zse", "soil%zse"),
]
# the reference's restart variables this path does not carry
NOT_CARRIED = ["wbtot0", "osnowd0", "GWwb", "patchfrac", "nap", "mvtype", "mstype"]

# the second (slow) dimension of a field, by its length
_DIM = {("ssnow%albsoilsn", 3): "rad", ("ssnow%", 6): "soil", ("ssnow%", 3): "snow", ("rad%", 3): "rad",
        ("soil%", 3): "rad",
        ("bgc%cplant", 3): "plant_carbon_pools", ("bgc%csoil", 2): "soil_carbon_pools"}


def _dims(field):
    if field == "soil%zse":
        return ("soil",)
    f = BY_NAME[field]
    if f.n == 1:
        return ("mp",)
    for (prefix, n), d in _DIM.items():
        if field.startswith(prefix) and f.n == n:
            return (d, "mp")
    raise ValueError(f"no restart dimension for {field} ({f.n} values per tile)")


def _kind(field, precision):
    """The native kind of the variable: int for integer fields, FP64 for
    REAL(r_2) fields (and the grid's layer depths) and for default REAL in
    FP64 (the promoted build), FP32 for default REAL in the mixed precision
    (the shipped build)."""
    if field == "soil%zse":
        return np.float64
    p = BY_NAME[field].prec
    if p == "I":
        return np.int32
    if p in ("D", "B") or precision == 0:
        return np.float64
    return np.float32


def restart_variables(fields, mp, precision=0):
    """{restart variable: array} from {field: array} (the fields' arrays as
    the library holds them, (n, mp) with the tile fastest; soil%zse (6,))."""
    out = {}
    for name, field in RESTART:
        a = np.asarray(fields[field])
        shape = (6,) if field == "soil%zse" else ((mp,) if BY_NAME[field].n == 1 else (BY_NAME[field].n, mp))
        out[name] = np.ascontiguousarray(a.reshape(shape), dtype=_kind(field, precision))
    return out


def fields_from_restart(variables):
    """{field: float64 array} from {restart variable: array}."""
    return {field: np.asarray(variables[name], dtype=np.float64) for name, field in RESTART if name in variables}


def write_restart(path, variables, mp):
    """Write restart variables to a NetCDF (classic) file, on the reference's
    restart grid and dimension names."""
    from scipy.io import netcdf_file
    with netcdf_file(path, "w", version=2) as nc:
        dims = {"mp": mp, "soil": 6, "snow": 3, "rad": 3, "plant_carbon_pools": 3, "soil_carbon_pools": 2}
        for d, n in dims.items():
            nc.createDimension(d, n)
        for name, field in RESTART:
            a = variables[name]
            v = nc.createVariable(name, a.dtype, _dims(field))
            v[:] = a


def read_restart(path):
    """{restart variable: array} from a restart file written by write_restart
    (or any file with the reference's restart variables on that grid)."""
    from scipy.io import netcdf_file
    out = {}
    with netcdf_file(path, "r", mmap=False) as nc:
        for name, _ in RESTART:
            if name in nc.variables:
                out[name] = np.array(nc.variables[name][:])
    return out


def save_restart(g, path, zse):
    """The state of a context (cable_amd.gpu.CableGPU) as a restart file in
    the reference's variables; zse: the grid's soil layer depths (6)."""
    fields = {field: g.download(field) for _, field in RESTART if field != "soil%zse"}
    fields["soil%zse"] = np.asarray(zse, dtype=np.float64)
    write_restart(path, restart_variables(fields, g.mp, g.precision), g.mp)


def load_restart(g, path):
    """Upload a restart file's variables into a context's fields (all but the
    grid's layer depths, which set_globals takes)."""
    for field, a in fields_from_restart(read_restart(path)).items():
        if field != "soil%zse":
            g.upload(field, a)
