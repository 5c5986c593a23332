"""The FP64 parity criterion: per value, relative to the oracle, with an
absolute floor per field.  Test infrastructure (tests/ and
__graft_entry__.smoke()); nothing that ships uses it.

A GPU value g passes against the oracle's o when

    |g - o| <= TOL * |o| + ATOL(field) + 2 * SENS

so every value is held to TOL relative on its own (not to a fraction of the
field's largest value), down to an absolute floor below which a difference
is physically nothing.  SENS, where a test supplies it, is the oracle's own
change of that value when every exp, log, pow and atan result of the step
moves by one ulp (OracleRun.step_sensitivity): the device's math library
rounds these functions differently from the host's by about an ulp, and in
the leaf fixed point an unconverged iterate amplifies such a difference by
many orders of magnitude (cbl_dryLeaf.F90:239-606 stops on a 0.1 K test,
not at the fixed point).  There the oracle itself cannot fix the value to
1e-10, and the GPU is held to twice the oracle's own spread instead.  The
floors, by the field's unit:

* energy fluxes and radiation (W/m2), temperatures (K), water amounts and
  rates (mm, kg/m2): 1e-9;
* carbon fluxes (canopy%fgpp and relatives, g C m-2 s-1): 1e-11, about
  1e-6 umol CO2 m-2 s-1;
* carbon amounts (bgc pools, sum_flux running sums, g C m-2): 1e-11 per
  second of a step (2e-8 for 1,800 s steps);
* every other field (dimensionless fractions, soil moisture m3/m3,
  conductances, lengths): 1e-12;
* the balance residuals: bal%ebal and its relatives are differences of
  terms of O(1000) W/m2 (cable_checks.F90:565-618), bal%wbal/wbal_tot of
  terms of O(10) mm (:515-547); a residual's rounding is set by those terms,
  so their floor is TOL times the closing terms (1e-7 W/m2, 1e-9 mm).

The floors matter only where the computation is ill conditioned: net
photosynthesis near the light compensation point (canopy%fpn, the difference
of gross assimilation and respiration, and the carbon sums it feeds), and
differences of nearly equal fluxes such as canopy%ga.  There a one-ulp
difference in exp, log or pow (the device's math library against glibc) is
amplified beyond 1e-10 of the small result.
"""
from __future__ import annotations

import re

import numpy as np

TOL = 1e-10
DELS = 1800.0

_ENERGY = re.compile(
    r"^(canopy%(fe|fh|fev|fes|fess|fesp|fevc|fevw|fhvw|fhv|fnv|fns|fhs|rnet|rniso|epot|fevw_pot|ghflux|"
    r"sghflux|segg|ga|fhs_cor|fns_cor|ga_cor)|rad%(rniso|lwabv|swnet|lwnet|rnet|flws|qcan|qssabs)|"
    r"ssnow%(potev|dfn_dtg|dfh_dtg|dfe_ddq|dfe_dtg))$")
_TEMP = re.compile(r"^(ssnow%(tss|tgg|tggsn|otss|otss_0|tggav|t_snwlr|deltss)|canopy%(tv|tscrn)|rad%trad)$")
_WATER = re.compile(
    r"^(ssnow%(pudsto|wbtot1|wbtot2|wb_lake|snowd|smass|osnowd|runoff|rnof1|rnof2|smelt|evapsn|sinfil|"
    r"fwtop1|fwtop2|fwtop3|wbtot|evapfbl|rex|sdepth|totsdepth)|canopy%(cansto|oldcansto|wcint|through|spill|"
    r"precis|delwc|dewmm)|bal%(precip_tot|rnoff_tot|evap_tot))$")
_CFLUX = re.compile(r"^canopy%(frday|fpn|frp|frpw|frpr|frs|fnpp|fgpp|fnee|fra)$")
_CAMOUNT = re.compile(r"^(bgc%(cplant|csoil)|sum_flux%.*)$")
_RESID = {
    "bal%ebal": 1e3, "bal%ebal_tot": 1e3, "bal%ebalveg": 1e3, "bal%ebalsoil": 1e3,
    "bal%drybal": 1e3, "bal%wetbal": 1e3, "bal%radbal": 1e3, "bal%radbalsum": 1e3,
    "bal%wbal": 1e1, "bal%wbal_tot": 1e1,
}


def atol(name):
    """The absolute floor of a field, in its own unit."""
    if name in _RESID:
        return TOL * _RESID[name]
    if _ENERGY.match(name) or _TEMP.match(name) or _WATER.match(name):
        return 1e-9
    if _CFLUX.match(name):
        return 1e-11
    if _CAMOUNT.match(name):
        return 1e-11 * DELS
    return 1e-12


def excess(g, o, name, sens=None):
    """Per value |g - o| / (TOL |o| + atol + 2 sens): at most 1 where the
    value passes."""
    g = np.asarray(g, dtype=np.float64).reshape(-1)
    o = np.asarray(o, dtype=np.float64).reshape(-1)
    bound = TOL * np.abs(o) + atol(name)
    if sens is not None:
        bound = bound + 2.0 * np.asarray(sens, dtype=np.float64).reshape(-1)
    return np.abs(g - o) / bound


def beyond_tolerance(g, o, name):
    """Values the TOL-relative-plus-floor bound alone does not cover (the
    ones that need the oracle's one-ulp spread)."""
    return excess(g, o, name) > 1.0


def rel_above_floor(g, o, name):
    """Per value |g - o| / |o| for the values whose relative term is the
    larger part of the bound (|o| >= atol/TOL); 0 elsewhere."""
    g = np.asarray(g, dtype=np.float64).reshape(-1)
    o = np.asarray(o, dtype=np.float64).reshape(-1)
    big = np.abs(o) >= atol(name) / TOL
    return np.where(big, np.abs(g - o) / np.where(big, np.abs(o), 1.0), 0.0)


def worst(g, o, name, sens=None):
    """(largest excess, index, gpu value, oracle value, largest relative
    error above the floor) of a field."""
    e = excess(g, o, name, sens)
    if e.size == 0:
        return 0.0, -1, 0.0, 0.0, 0.0
    i = int(np.argmax(e))
    return (float(e[i]), i, float(np.asarray(g).reshape(-1)[i]), float(np.asarray(o).reshape(-1)[i]),
            float(rel_above_floor(g, o, name).max()))
