"""The per-tile field table shared by every implementation.

Parsed from ``include/cable_fields.h`` (the single source of truth), so the
C oracle, the HIP kernels and the Python host side agree on field order,
element counts, precision and role.
"""
from __future__ import annotations

import os
import re
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(os.path.dirname(_HERE), "include", "cable_fields.h")


@dataclass(frozen=True)
class Field:
    fid: int
    ident: str
    name: str      # Fortran component name, e.g. "ssnow%wb"
    prec: str      # "F" default REAL, "D" REAL(r_2), "I" INTEGER, "B" FP64 accumulator
    n: int         # elements per tile
    role: str      # PARAM, FORC, STATE, DIAG

    def shape(self, mp: int):
        """numpy shape of the Fortran (mp, ...) array (tile index last)."""
        if self.name == "rad%qcan":
            return (3, 2, mp)
        return (mp,) if self.n == 1 else (self.n, mp)


def _parse():
    txt = open(HEADER).read()
    body = txt[txt.index("#define CABLE_FIELDS(X)"):txt.index("enum cable_field_id")]
    out = []
    for m in re.finditer(r'X\((\w+),\s*"([^"]+)",\s*([FDIB]),\s*(\d+),\s*(\w+)\)', body):
        out.append(Field(len(out), m.group(1), m.group(2), m.group(3), int(m.group(4)), m.group(5)))
    return out


FIELDS = _parse()
BY_NAME = {f.name: f for f in FIELDS}
NFIELDS = len(FIELDS)


def zero_fields(mp: int):
    """{name: zero array} for every field (the reference harness zero-fills
    every component before the run)."""
    return {f.name: np.zeros(f.shape(mp)) for f in FIELDS}
