/*
 * cbm_oracle.c -- CPU restatement of CABLE's per-tile timestep (cbm) in FP64.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle the GPU kernels are
 * checked against; it is loaded by tests/, by __graft_entry__.smoke() and by
 * bench.py's cpu_baseline leg, and by nothing that ships.
 *
 * It restates, one tile at a time, what the reference computes for the
 * shipped configuration (src/offline/cable.nml: soil_struc 'default',
 * FWSOIL_SWITCH 'standard', GS_SWITCH 'leuning', SSNOW_POTEV 'HDM',
 * DIAG_SOIL_RESP 'ON', icycle 0, every other switch at its default), at the
 * precision of the promoted build (-fdefault-real-8: every REAL is FP64).
 * Each function cites the reference routine it follows.  Expressions keep
 * the reference's left-to-right evaluation order, and powers are evaluated
 * the way the reference compiler lowers them: integer exponents (x**2,
 * x**3.0, x**4) as multiplication chains x*(x*(x*x)), x**0.5 as sqrt, 2.0**y
 * as exp2, every other real exponent through pow.  Fortran MAX(a,b) is
 * (a > b ? a : b) and MIN(a,b) is (a < b ? a : b), as the compiler emits
 * them.  Compile with -ffp-contract=off (no FMA).
 *
 * Pinned against the compiled reference (oracle/_ref, tests/golden).
 *
 * Per-tile data: every field of include/cable_fields.h, as double.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "../include/cable_fields.h"

#define MS 6
#define MSN 3

/* Test hook for measuring the reference's own sensitivity: an offset (K)
   added to the leaf temperature the canopy's fixed point starts from
   (tlfx = tlfy = met%tk, src/science/canopy/cable_canopy.F90:213-214).  It
   is 0 unless oracle_set_leaf_start_offset sets it; x + 0.0 == x, so the
   default leaves every result bit-identical. */
static double leaf_start_offset = 0.0;
void oracle_set_leaf_start_offset(double dk) { leaf_start_offset = dk; }

/* Test hook for measuring how far one-ulp differences of the elementary
   functions carry: with a shift of +1 (-1) every exp, exp2, log, pow and
   atan result is moved one ulp up (down).  A device math library rounds
   these functions differently from the host's by about an ulp, so the
   oracle's own response to the shift bounds the part of a device-host
   difference that the computation, not the device code, produces.  0 (the
   default) returns the library's value unchanged. */
static int ulp_shift = 0;
void oracle_set_ulp_shift(int s) { ulp_shift = s; }
static inline double shifted(double y) {
  return ulp_shift == 0 ? y : nextafter(y, ulp_shift > 0 ? INFINITY : -INFINITY);
}
static inline double o_exp(double x) { return shifted(exp(x)); }
static inline double o_exp2(double x) { return shifted(exp2(x)); }
static inline double o_log(double x) { return shifted(log(x)); }
static inline double o_pow(double x, double y) { return shifted(pow(x, y)); }
static inline double o_atan(double x) { return shifted(atan(x)); }
#define exp(x) o_exp(x)
#define exp2(x) o_exp2(x)
#define log(x) o_log(x)
#define pow(x, y) o_pow(x, y)
#define atan(x) o_atan(x)

/* ---- constants: src/params/ (phys, photo, other, maths) and src/util/cable_common.F90 ---- */
static const double TFRZ = 273.16, SBOLTZ = 5.67e-8, EMSOIL = 1.0,
                    EMLEAF = 1.0, CAPP = 1004.64, HL = 2.5014e6,
                    HLF = 0.334e6, DHEAT = 21.5e-6, GRAV = 9.8086,
                    RGAS = 8.3143, RMAIR = 0.02897, RMH2O = 0.018016,
                    CGSNOW = 2090.0, CSICE = 2.100e3, CSWAT = 4.218e3,
                    DENSITY_LIQ = 1000.0, DENSITY_ICE = 921.0,
                    TETENA = 6.106, TETENB = 17.27, TETENC = 237.3,
                    VONK = 0.40, A33 = 1.25, CSW = 0.50, CTL = 0.40,
                    APOL = 0.70, PRANDT = 0.71, CRD = 0.3, CSD = 0.003,
                    CCD = 15.0, CCW_C = 2.0, USUHM = 0.3, ZETA0 = 0.0,
                    ZETNEG = -15.0, ZETPOS = 1.0, ZDLIN = 1.0, UMIN = 0.1;
static const double PI_C = 3.1415927;           /* cable_maths_constants_mod */
static const double GAUSS_W[3] = {0.308, 0.514, 0.178};
static const double LAI_THRESH = 0.001, RAD_THRESH = 0.001,
                    COSZEN_TOLS = 1.0e-4;
/* photosynthesis constants, cable_photo_constants_mod.F90:29-41 */
static const int MAXITER = 20;
static const double GAM0 = 28.0e-6, GAM1 = 0.0509, GAM2 = 0.0010,
                    RGBWC = 1.32, RGSWC = 1.57, TREFK = 298.2;
/* snow constants, cable_common.F90:217-222; snmin from cable.nml */
static const double MAX_GLACIER_SNOWD = 1100.0, SNOW_CCNSW = 2.0,
                    MAX_SSDN = 750.0, MAX_SCONDS = 2.51,
                    FROZEN_LIMIT = 0.85, SNMIN = 1.0;
enum { LAKES = 16, ICE_VEG = 17, ICE_SOIL = 9 };

typedef struct {
#define X(id, name, prec, n, role) double id[n];
  CABLE_FIELDS(X)
#undef X
} Tile;

typedef struct {
  double zse[MS], zshh[MS + 1], ratecp[3], ratecs[2];
} Globals;

static inline double MAX(double a, double b) { return a > b ? a : b; }
static inline double MIN(double a, double b) { return a < b ? a : b; }
static inline double SIGN(double a, double b) { return copysign(fabs(a), b); }
static inline double p2(double x) { return x * x; }
static inline double p3(double x) { return x * (x * x); }
static inline double p4(double x) { return x * (x * (x * x)); }

/* ===================================================================== */
/* src/science/roughness/roughnessHGT_effLAI_cbl.F90:42-149 and
   src/science/roughness/cable_roughness.F90:64-317 (soil_struc default,
   l_new_roughness_soil and or_evap false) */
static void ruff_resist(Tile *t) {
  const double z0soilsn_min = 1.e-7, z0soilsn_min_PF = 1.e-4;
  /* HgtAboveSnow */
  double sde = MAX(100.0, t->ssnow_ssdnn[0]);
  double hmin = 10.0 * z0soilsn_min;
  double hcomp = t->veg_hc[0] - (1.2 * t->ssnow_snowd[0] / sde);
  double hruff = MAX(hmin, hcomp);
  t->rough_hruff[0] = hruff;
  /* LAI_eff (cable_runtime%offline) */
  double frac = hruff / MAX(0.01, t->veg_hc[0]);
  double vlaiw = t->veg_vlai[0] * frac;
  t->canopy_vlaiw[0] = vlaiw;
  t->canopy_rghlai[0] = vlaiw;
  double z0soil = 0.0009 * MIN(1.0, vlaiw) + 1.e-4;
  double z0soilsn = z0soil;
  double snowd = t->ssnow_snowd[0];
  if (snowd > 0.01)
    z0soilsn = MAX(z0soilsn_min, z0soil - z0soil * MIN(snowd, 10.) / 10.);
  if (snowd > 0.01 && (int)t->veg_iveg[0] == ICE_VEG)
    z0soilsn = MAX(z0soilsn, z0soilsn_min_PF);
  t->rough_z0soil[0] = z0soil;
  t->rough_z0soilsn[0] = z0soilsn;
  double rghlai = vlaiw;
  if (vlaiw <= LAI_THRESH || hruff < z0soilsn) { /* bare soil, :222-254 */
    double z0m = z0soilsn;
    t->rough_z0m[0] = z0m;
    t->rough_rt0us[0] = 0.0;
    t->rough_disp[0] = 0.0;
    double zuv = MAX(3.5 + z0m, t->rough_za_uv[0]);
    double ztq = MAX(3.5 + z0m, t->rough_za_tq[0]);
    zuv = MAX(zuv, hruff - 0.0);
    ztq = MAX(ztq, hruff - 0.0);
    t->rough_zref_uv[0] = zuv;
    t->rough_zref_tq[0] = ztq;
    t->rough_zruffs[0] = 0.0;
    t->rough_rt1usa[0] = 0.0;
    t->rough_rt1usb[0] = 0.0;
    double usuh = MIN(sqrt(CSD + CRD * (vlaiw * 0.5)), USUHM);
    double xx = sqrt(CCD * MAX((vlaiw * 0.5), 0.0005));
    double dh = 1.0 - (1.0 - exp(-xx)) / xx;
    t->rough_usuh[0] = usuh;
    t->rough_coexp[0] = usuh / (VONK * CCW_C * (1.0 - dh));
  } else { /* vegetated, :256-316 */
    double usuh = MIN(sqrt(CSD + CRD * (rghlai * 0.5)), USUHM);
    double xx = sqrt(CCD * MAX((rghlai * 0.5), 0.0005));
    double dh = 1.0 - (1.0 - exp(-xx)) / xx;
    double disp = dh * hruff;
    double z0m =
        ((1.0 - dh) * exp(log(CCW_C) - 1. + 1. / CCW_C - VONK / usuh)) * hruff;
    double zuv = MAX(3.5 + z0m, t->rough_za_uv[0]);
    double ztq = MAX(3.5 + z0m, t->rough_za_tq[0]);
    zuv = MAX(zuv, hruff - disp);
    ztq = MAX(ztq, hruff - disp);
    double coexp = usuh / (VONK * CCW_C * (1.0 - dh));
    double term2 = exp(2 * CSW * rghlai * (1 - disp / hruff));
    double term3 = A33 * A33 * CTL * 2 * CSW * rghlai;
    double term5 = MAX((2. / 3.) * hruff / disp, 1.0);
    double term6 = exp(3. * coexp * (disp / hruff - 1.));
    double term6a = exp(coexp * (0.1 * hruff / hruff - 1.));
    double rt0us = term5 * (ZDLIN * log(ZDLIN * disp / z0soilsn) + (1 - ZDLIN)) *
                   (exp(2 * CSW * rghlai) - term2) / term3;
    double zruffs = disp + hruff * (A33 * A33) * CTL / VONK / term5;
    double rt1usa = term5 * (term2 - 1.0) / term3;
    double rt1usb = term5 * (MIN(ztq + disp, zruffs) - hruff) /
                    (A33 * A33 * CTL * hruff);
    rt1usb = MAX(rt1usb, 0.0);
    t->rough_usuh[0] = usuh;
    t->rough_disp[0] = disp;
    t->rough_z0m[0] = z0m;
    t->rough_zref_uv[0] = zuv;
    t->rough_zref_tq[0] = ztq;
    t->rough_coexp[0] = coexp;
    t->rough_term2[0] = term2;
    t->rough_term3[0] = term3;
    t->rough_term5[0] = term5;
    t->rough_term6[0] = term6;
    t->rough_term6a[0] = term6a;
    t->rough_rt0us[0] = rt0us;
    t->rough_zruffs[0] = zruffs;
    t->rough_rt1usa[0] = rt1usa;
    t->rough_rt1usb[0] = rt1usb;
  }
}

/* src/science/misc/cable_air.F90:51-97 */
static void define_air(Tile *t) {
  double tv = t->met_tvair[0], pmb = t->met_pmb[0];
  double es = TETENA * exp(TETENB * (tv - TFRZ) / (TETENC + (tv - TFRZ)));
  double cmolar = pmb * 100.0 / (RGAS * (tv));
  t->air_cmolar[0] = cmolar;
  t->air_rho[0] = MIN(1.3, RMAIR * cmolar);
  t->air_volm[0] = RGAS * (tv) / (100.0 * pmb);
  double rlam = HL;
  t->air_rlam[0] = rlam;
  t->air_qsat[0] = (RMH2O / RMAIR) * es / pmb;
  t->air_epsi[0] = (rlam / CAPP) * (RMH2O / RMAIR) * es * TETENB * TETENC /
                   p2(TETENC + (tv - TFRZ)) / pmb;
  t->air_visc[0] = 1e-5 * MAX(1.0, 1.35 + 0.0092 * (tv - TFRZ));
  t->air_psyc[0] = pmb * 100.0 * CAPP * RMAIR / rlam / RMH2O;
  t->air_dsatdk[0] = 100.0 * (TETENA * TETENB * TETENC) /
                     p2((tv - TFRZ) + TETENC) *
                     exp(TETENB * (tv - TFRZ) / ((tv - TFRZ) + TETENC));
}

/* src/science/canopy/cbl_qsat.F90:16-53 */
static inline double qsatfjh(double tair, double pmb) {
  return (RMH2O / RMAIR) * (TETENA * exp(TETENB * tair / (TETENC + tair))) / pmb;
}

/* src/science/radiation/cbl_spitter.F90:36-76 */
static double spitter(int doy, double coszen, double fsd) {
  const double solcon = 1370.0;
  double fbeam = 0.0, tmprat;
  double tmpr = 0.847 + coszen * (1.04 * coszen - 1.61);
  double tmpk = (1.47 - tmpr) / 1.66;
  if (coszen > 1.0e-10 && fsd > 10.0)
    tmprat = fsd / (solcon * (1.0 + 0.033 * cos(2.0 * PI_C * ((double)doy - 10.0) /
                                                365.0)) * coszen);
  else
    tmprat = 0.0;
  if (tmprat > 0.22) fbeam = 6.4 * p2(tmprat - 0.22);
  if (tmprat > 0.35) fbeam = MIN(1.66 * tmprat - 0.4728, 1.0);
  if (tmprat > tmpk) fbeam = MAX(1.0 - tmpr, 0.0);
  return fbeam;
}

/* Radiation set-up and albedo, with the local coefficients the two share:
   src/science/radiation/cbl_init_radiation.F90:30-392,
   src/science/radiation/cbl_rhoch.F90:38-62,
   src/science/albedo/cbl_albedo.F90:56-386,
   src/science/albedo/cbl_snow_albedo.F90:36-161 */
static void init_radiation_albedo(Tile *t, int veg_mask) {
  const double pi180 = PI_C / 180.0;
  double vlaiw = t->canopy_vlaiw[0], coszen = t->met_coszen[0];
  double xk[3], c1[3], rhoch[3], xvlai2[3];
  double cos3[3] = {cos(pi180 * 15.0), cos(pi180 * 45.0), cos(pi180 * 75.0)};
  double xphi1 = 0.0, xphi2 = 0.0;
  if (veg_mask) {
    double xf = t->veg_xfang[0];
    xphi1 = 0.5 - xf * (0.633 + 0.33 * xf);
    xphi2 = 0.877 * (1.0 - 2.0 * xphi1);
  }
  for (int b = 0; b < 3; b++) {
    xvlai2[b] = vlaiw;
    xk[b] = (xvlai2[b] > LAI_THRESH) ? xphi1 / cos3[b] + xphi2 : 0.0;
  }
  c1[0] = sqrt(1.0 - t->veg_taul[0] - t->veg_refl[0]);
  c1[1] = sqrt(1.0 - t->veg_taul[1] - t->veg_refl[1]);
  c1[2] = 1.0;
  for (int b = 0; b < 3; b++) rhoch[b] = (1.0 - c1[b]) / (1.0 + c1[b]);
  /* ExtinctionCoeff */
  double tiny = COSZEN_TOLS * 1e-2, huge = COSZEN_TOLS * 1e2;
  double extkb = 0.5, extkd = 0.7;
  if (veg_mask) {
    double s = GAUSS_W[0] * exp(-xk[0] * xvlai2[0]);
    s = s + GAUSS_W[1] * exp(-xk[1] * xvlai2[1]);
    s = s + GAUSS_W[2] * exp(-xk[2] * xvlai2[2]);
    extkd = -log(s) / vlaiw;
  }
  if (veg_mask && coszen > tiny) extkb = xphi1 / coszen + xphi2;
  if (coszen < tiny) extkb = 1.0e5;
  if (fabs(extkb - extkd) < 0.001) extkb = extkd + 0.001;
  t->rad_extkb[0] = extkb;
  t->rad_extkd[0] = extkd;
  /* EffectiveExtinctCoeffs */
  for (int b = 0; b < 3; b++) t->rad_extkbm[b] = 0.0, t->rad_extkdm[b] = 0.0;
  for (int b = 0; b < 2; b++) {
    if (veg_mask) t->rad_extkbm[b] = extkb * c1[b];
    t->rad_extkdm[b] = extkd * c1[b];
  }
  /* BeamFraction */
  double fb = spitter((int)t->met_doy[0], coszen, t->met_fsd[0] + t->met_fsd[1]);
  t->rad_fbeam[0] = fb;
  t->rad_fbeam[1] = fb;
  if (coszen < huge) t->rad_fbeam[0] = 0.0, t->rad_fbeam[1] = 0.0;

  /* ---- Albedo ---- */
  double alb[3] = {0.0, 0.0, 0.0};
  for (int b = 0; b < 3; b++) t->rad_rhocbm[b] = 0.0, t->rad_rhocdf[b] = 0.0;
  { /* surface_albedosn */
    const double alvo = 0.95, aliro = 0.70;
    int iveg = (int)t->veg_iveg[0];
    double snowd = t->ssnow_snowd[0];
    double sa = t->soil_albsoil[0];
    if (iveg == LAKES)
      sa = -0.022 * (MIN(275.0, MAX(260.0, t->ssnow_tgg[0])) - 260.0) + 0.45;
    if (snowd > 1.0 && iveg == LAKES) sa = 0.85;
    double sfact = 0.68;
    if (sa <= 0.14)
      sfact = 0.5;
    else if (sa > 0.14 && sa <= 0.20)
      sfact = 0.62;
    alb[1] = 2.0 * sa / (1.0 + sfact);
    alb[0] = sfact * alb[1];
    double snrat = 0.0, alir = 0.0, alv = 0.0;
    if (snowd > 1.0) {
      double tmp = snowd / MAX(t->ssnow_ssdnn[0], 200.0);
      snrat = MIN(1.0, tmp / (tmp + 0.1));
      double fage = 1.0 - 1.0 / (1.0 + t->ssnow_snage[0]);
      tmp = MAX(0.17365, coszen);
      double fz = (tmp > 0.5) ? 0.0 : 1.5 / (1.0 + 4.0 * tmp) - 0.5;
      double fzenm = MAX(0.0, fz);
      tmp = alvo * (1.0 - 0.2 * fage);
      alv = 0.4 * fzenm * (1.0 - tmp) + tmp;
      tmp = aliro * (1.0 - 0.5 * fage);
      alir = 0.4 * fzenm * (1.0 - tmp) + tmp;
    }
    alb[1] = MIN(aliro, (1.0 - snrat) * alb[1] + snrat * alir);
    alb[0] = MIN(alvo, (1.0 - snrat) * alb[0] + snrat * alv);
    if ((int)t->soil_isoilm[0] == ICE_SOIL) {
      alb[0] = alvo - 0.05;
      alb[1] = aliro - 0.05;
    }
  }
  /* CanopyReflectance */
  for (int b = 0; b < 2; b++) {
    if (veg_mask)
      t->rad_rhocbm[b] = 2.0 * extkb / (extkb + extkd) * rhoch[b];
    t->rad_rhocdf[b] =
        rhoch[b] * 2.0 *
        (GAUSS_W[0] * xk[0] / (xk[0] + extkd) + GAUSS_W[1] * xk[1] / (xk[1] + extkd) +
         GAUSS_W[2] * xk[2] / (xk[2] + extkd));
  }
  /* CanopyTransmitance */
  for (int b = 0; b < 2; b++) {
    if (veg_mask) {
      double d = MIN(t->rad_extkbm[b] * vlaiw, 20.0);
      t->rad_cexpkbm[b] = exp(-1.0 * d);
    }
    double d = t->rad_extkdm[b] * vlaiw;
    t->rad_cexpkdm[b] = exp(-1.0 * d);
  }
  /* EffectiveSurfaceReflectance */
  for (int b = 0; b < 3; b++) t->rad_reffbm[b] = alb[b], t->rad_reffdf[b] = alb[b];
  for (int b = 0; b < 2; b++) {
    if (veg_mask) {
      t->rad_reffdf[b] = t->rad_rhocdf[b] +
                         (alb[b] - t->rad_rhocdf[b]) * p2(t->rad_cexpkdm[b]);
      t->rad_reffbm[b] = t->rad_rhocbm[b] +
                         (alb[b] - t->rad_rhocbm[b]) * p2(t->rad_cexpkbm[b]);
    }
  }
  /* FbeamRadAlbedo */
  for (int b = 0; b < 3; b++) t->rad_albedo[b] = alb[b];
  for (int b = 0; b < 2; b++)
    if (veg_mask)
      t->rad_albedo[b] = (1.0 - t->rad_fbeam[b]) * t->rad_reffdf[b] +
                         t->rad_fbeam[b] * t->rad_reffbm[b];
  for (int b = 0; b < 3; b++) t->ssnow_albsoilsn[b] = alb[b];
}

/* src/science/radiation/cbl_radiation.F90:30-218 */
static void radiation(Tile *t, int sunlit_veg) {
  double vlaiw = t->canopy_vlaiw[0];
  double extkb = t->rad_extkb[0], extkd = t->rad_extkd[0];
  double cf2n = exp(-t->veg_extkn[0] * vlaiw);
  double transd = 1.0;
  if (vlaiw > LAI_THRESH) transd = exp(-extkd * vlaiw);
  double dummy2 = MIN(extkb * vlaiw, 30.);
  double transb = exp(-dummy2);
  t->rad_transd[0] = transd;
  t->rad_transb[0] = transb;
  double tvrad = t->met_tvrad[0];
  double flpwb = SBOLTZ * p4(tvrad);
  double flwv = EMLEAF * flpwb;
  double flws = SBOLTZ * EMSOIL * p4(t->ssnow_tss[0]);
  t->rad_flws[0] = flws;
  double emair = t->met_fld[0] / flpwb;
  double *gradis = t->rad_gradis, *qcan = t->rad_qcan; /* qcan[l + 2*b] */
  gradis[0] = gradis[1] = 0.0;
  for (int k = 0; k < 6; k++) qcan[k] = 0.0;
  double rho = t->air_rho[0];
  if (vlaiw > LAI_THRESH) {
    gradis[0] = (4.0 * EMLEAF / (CAPP * rho)) * flpwb / (tvrad)*extkd *
                ((1.0 - transb * transd) / (extkb + extkd) +
                 (transd - transb) / (extkb - extkd));
    gradis[1] = (8.0 * EMLEAF / (CAPP * rho)) * flpwb / tvrad * extkd *
                    (1.0 - transd) / extkd - gradis[0];
    qcan[0 + 2 * 2] = (flws - flwv) * extkd * (transd - transb) / (extkb - extkd) +
                      (emair - EMLEAF) * extkd * flpwb * (1.0 - transd * transb) /
                          (extkb + extkd);
    qcan[1 + 2 * 2] =
        (1.0 - transd) * (flws + t->met_fld[0] - 2.0 * flwv) - qcan[0 + 2 * 2];
  }
  gradis[0] = t->air_cmolar[0] * gradis[0];
  gradis[1] = t->air_cmolar[0] * gradis[1];
  gradis[0] = MAX(1.0e-3, gradis[0]);
  gradis[1] = MAX(1.0e-3, gradis[1]);
  double cf1 = 0.0, cf3 = 0.0;
  for (int b = 0; b < 2; b++) {
    if (sunlit_veg) {
      double kdm = t->rad_extkdm[b], kbm = t->rad_extkbm[b], fb = t->rad_fbeam[b];
      double rdf = t->rad_reffdf[b], rbm = t->rad_reffbm[b];
      double ckd = t->rad_cexpkdm[b], ckb = t->rad_cexpkbm[b];
      double tl = t->veg_taul[b], rl = t->veg_refl[b];
      cf1 = (1.0 - transb * ckd) / (extkb + kdm);
      cf3 = (1.0 - transb * ckb) / (extkb + kbm);
      double fsd = t->met_fsd[b];
      qcan[0 + 2 * b] =
          fsd * ((1.0 - fb) * (1.0 - rdf) * kdm * cf1 + fb * (1.0 - rbm) * kbm * cf3 +
                 fb * (1.0 - tl - rl) * extkb *
                     ((1 - transb) / extkb - (1 - p2(transb)) / (extkb + extkb)));
      qcan[1 + 2 * b] =
          fsd * ((1.0 - fb) * (1.0 - rdf) * kdm * ((1.0 - ckd) / kdm - cf1) +
                 fb * (1. - rbm) * kbm * ((1.0 - ckb) / kbm - cf3) -
                 fb * (1.0 - tl - rl) * extkb *
                     ((1 - transb) / extkb - (1 - p2(transb)) / (extkb + extkb)));
    }
  }
  double *fsd = t->met_fsd;
  if (sunlit_veg) {
    t->rad_qssabs[0] =
        fsd[0] * (t->rad_fbeam[0] * (1. - t->rad_reffbm[0]) *
                      exp(-MIN(t->rad_extkbm[0] * vlaiw, 20.)) +
                  (1. - t->rad_fbeam[0]) * (1. - t->rad_reffdf[0]) *
                      exp(-MIN(t->rad_extkdm[0] * vlaiw, 20.))) +
        fsd[1] * (t->rad_fbeam[1] * (1. - t->rad_reffbm[1]) * t->rad_cexpkbm[1] +
                  (1. - t->rad_fbeam[1]) * (1. - t->rad_reffdf[1]) * t->rad_cexpkdm[1]);
    t->rad_scalex[0] = (1.0 - transb * cf2n) / (extkb + t->veg_extkn[0]);
    t->rad_fvlai[0] = (1.0 - transb) / extkb;
    t->rad_fvlai[1] = vlaiw - t->rad_fvlai[0];
  } else {
    t->rad_qssabs[0] = (1.0 - t->ssnow_albsoilsn[0]) * fsd[0] +
                       (1.0 - t->ssnow_albsoilsn[1]) * fsd[1];
    t->rad_scalex[0] = 0.0;
    t->rad_fvlai[0] = 0.0;
    t->rad_fvlai[1] = vlaiw;
  }
  t->rad_scalex[1] = (1.0 - cf2n) / t->veg_extkn[0] - t->rad_scalex[0];
  for (int l = 0; l < 2; l++)
    t->rad_rniso[l] = qcan[l + 0] + qcan[l + 2] + qcan[l + 4];
}

/* src/science/canopy/cbl_friction_vel.F90:112-166 */
static double psim(double zeta) {
  const double gu = 16.0, a = 1.0, b = 0.667, xc = 5.0, d = 0.35;
  double z = 0.5 + SIGN(0.5, zeta);
  double stable = -a * zeta - b * (zeta - xc / d) * exp(-d * zeta) - b * xc / d;
  double x = pow(1.0 + gu * fabs(zeta), 0.25);
  double unstable = log((1.0 + x * x) * p2(1.0 + x) / 8) - 2.0 * atan(x) + PI_C * 0.5;
  return z * stable + (1.0 - z) * unstable;
}

/* src/science/canopy/cbl_friction_vel.F90:168-221 */
static double psis(double zeta) {
  const double gu = 16.0, a = 1.0, b = 0.667, c = 5.0, d = 0.35;
  double z = 0.5 + SIGN(0.5, zeta);
  double stzeta = MAX(0., zeta);
  double stable = -pow(1. + 2. / 3. * a * stzeta, 3. / 2.) -
                  b * (stzeta - c / d) * exp(-d * stzeta) - b * c / d + 1.;
  double y = sqrt(1.0 + gu * fabs(zeta));
  double unstable = 2.0 * log((1 + y) * 0.5);
  return z * stable + (1.0 - z) * unstable;
}

/* src/science/canopy/cbl_dryLeaf.F90:779-877 */
static inline double ej3x(double parx, double alpha, double convex, double x) {
  return MAX(0.0, 0.25 * ((alpha * parx + x -
                           sqrt(p2(alpha * parx + x) - 4.0 * convex * alpha * parx * x)) /
                          (2.0 * convex)));
}
static inline double ej4x(double parx, double alpha, double convex, double x) {
  return MAX(0.0, (alpha * parx + x -
                   sqrt(p2(alpha * parx + x) - 4.0 * convex * alpha * parx * x)) /
                      (2.0 * convex));
}
static inline double xvcmxt4(double x) {
  /* q10c4 ** (0.1*x - 2.5) with q10c4 = 2.0 */
  return exp2(0.1 * x - 2.5) /
         ((1.0 + exp(0.3 * (13.0 - x))) * (1.0 + exp(0.3 * (x - 36.0))));
}
static inline double xvcmxt3(double x) {
  const double EHaVc = 73637.0, EHdVc = 149252.0, EntropVc = 486.0,
               xVccoef = 1.17461;
  double num = xVccoef * exp((EHaVc / (RGAS * TREFK)) * (1. - TREFK / x));
  double den = 1.0 + exp((EntropVc * x - EHdVc) / (RGAS * x));
  return MAX(0.0, num / den);
}
static inline double xejmxt3(double x) {
  const double EHaJx = 50300.0, EHdJx = 152044.0, EntropJx = 495.0,
               xjxcoef = 1.16715;
  double num = xjxcoef * exp((EHaJx / (RGAS * TREFK)) * (1. - TREFK / x));
  double den = 1.0 + exp((EntropJx * x - EHdJx) / (RGAS * x));
  return MAX(0.0, num / den);
}

/* src/science/canopy/cbl_photosynthesis.F90:10-226, one leaf */
static double photosynthesis_leaf(double csxz, double cx1z, double cx2z,
                                  double gswminz, double rdxz, double vcmxt3z,
                                  double vcmxt4z, double vx3z, double vx4z,
                                  double gs_coeffz, double fwsoilz) {
  const double effc4 = 4000.0;
  double coef0z, coef1z, coef2z, ciz, delcxz;
  double anrubiscoz = 0.0, anrubpz = 0.0, ansinkz = 0.0;
  coef2z = gswminz * fwsoilz / RGSWC + gs_coeffz * (vcmxt3z - (rdxz - vcmxt4z));
  coef1z = (1.0 - csxz * gs_coeffz) * (vcmxt3z + vcmxt4z - rdxz) +
           (gswminz * fwsoilz / RGSWC) * (cx1z - csxz) -
           gs_coeffz * (vcmxt3z * cx2z / 2.0 + cx1z * (rdxz - vcmxt4z));
  coef0z = -(1.0 - csxz * gs_coeffz) * (vcmxt3z * cx2z / 2.0 + cx1z * (rdxz - vcmxt4z)) -
           (gswminz * fwsoilz / RGSWC) * cx1z * csxz;
  if (fabs(coef2z) > 1.0e-9 && fabs(coef1z) < 1.0e-9) {
    ciz = 99999.0;
    anrubiscoz = 99999.0;
  }
  if (fabs(coef2z) < 1.e-9 && fabs(coef1z) >= 1e-9) {
    ciz = -1.0 * coef0z / coef1z;
    ciz = MAX(0.0, ciz);
    anrubiscoz = vcmxt3z * (ciz - cx2z / 2.0) / (ciz + cx1z) + vcmxt4z - rdxz;
  }
  if (fabs(coef2z) >= 1.e-9) {
    delcxz = p2(coef1z) - 4.0 * coef0z * coef2z;
    ciz = (-coef1z + sqrt(MAX(0.0, delcxz))) / (2.0 * coef2z);
    ciz = MAX(0.0, ciz);
    anrubiscoz = vcmxt3z * (ciz - cx2z / 2.0) / (ciz + cx1z) + vcmxt4z - rdxz;
  }
  coef2z = gswminz * fwsoilz / RGSWC + gs_coeffz * (vx3z - (rdxz - vx4z));
  coef1z = (1.0 - csxz * gs_coeffz) * (vx3z + vx4z - rdxz) +
           (gswminz * fwsoilz / RGSWC) * (cx2z - csxz) -
           gs_coeffz * (vx3z * cx2z / 2.0 + cx2z * (rdxz - vx4z));
  coef0z = -(1.0 - csxz * gs_coeffz) * (vx3z * cx2z / 2.0 + cx2z * (rdxz - vx4z)) -
           (gswminz * fwsoilz / RGSWC) * cx2z * csxz;
  ciz = 99999.0;
  anrubpz = 99999.0;
  if (fabs(coef2z) < 1.e-9 && fabs(coef1z) >= 1.e-9) {
    ciz = -1.0 * coef0z / coef1z;
    ciz = MAX(0.0, ciz);
    anrubpz = vx3z * (ciz - cx2z / 2.0) / (ciz + cx2z) + vx4z - rdxz;
  }
  if (fabs(coef2z) >= 1.e-9) {
    delcxz = p2(coef1z) - 4.0 * coef0z * coef2z;
    ciz = (-coef1z + sqrt(MAX(0.0, delcxz))) / (2.0 * coef2z);
    ciz = MAX(0.0, ciz);
    anrubpz = vx3z * (ciz - cx2z / 2.0) / (ciz + cx2z) + vx4z - rdxz;
  }
  coef2z = gs_coeffz;
  coef1z = gswminz * fwsoilz / RGSWC + gs_coeffz * (rdxz - 0.5 * vcmxt3z) +
           effc4 * vcmxt4z - gs_coeffz * csxz * effc4 * vcmxt4z;
  coef0z = -(gswminz * fwsoilz / RGSWC) * csxz * effc4 * vcmxt4z +
           (rdxz - 0.5 * vcmxt3z) * gswminz * fwsoilz / RGSWC;
  if (fabs(coef2z) < 1.0e-9 && fabs(coef1z) < 1.0e-9) {
    ciz = 99999.0;
    ansinkz = 99999.0;
  }
  if (fabs(coef2z) < 1.e-9 && fabs(coef1z) >= 1.e-9) {
    ciz = -1.0 * coef0z / coef1z;
    ansinkz = ciz;
  }
  if (fabs(coef2z) >= 1.e-9) {
    delcxz = p2(coef1z) - 4.0 * coef0z * coef2z;
    ciz = (-coef1z + sqrt(MAX(0.0, delcxz))) / (2.0 * coef2z);
    ansinkz = ciz;
  }
  (void)ciz;
  return MIN(MIN(anrubiscoz, anrubpz), ansinkz);
}

/* src/science/soilsnow/cbl_remove_trans.F90:43-93 */
static void transp_soil_water(double dels, const double *swilt,
                              const double *froot, const double *zse,
                              double fevc, const double *wbliq, double *evapfbl) {
  double diff[MS + 1];
  double xx = 0., xxd = 0.;
  for (int k = 0; k <= MS; k++) diff[k] = 0.;
  if (fevc > 0.0) {
    for (int k = 1; k <= MS; k++) {
      xx = fevc * dels / HL * froot[k - 1] + diff[k - 1];
      diff[k] = MAX(0.0, wbliq[k - 1] - 1.1 * swilt[k - 1]) * zse[k - 1] * DENSITY_LIQ;
      xxd = xx - diff[k];
      if (xxd > 0.0) {
        evapfbl[k - 1] = diff[k];
        diff[k] = xxd;
      } else {
        evapfbl[k - 1] = xx;
        diff[k] = 0.0;
      }
    }
  }
}

/* Working variables define_canopy shares with dryLeaf and wetLeaf. */
typedef struct {
  double cansat, dsx, fwsoil, tlfx, tlfy, ecy, hcy, rny;
  double gbhu[2], gbhf[2], csx[2], ghwet;
  double sum_rad_rniso, sum_rad_gradis;
} Canopy;

/* src/science/canopy/cbl_dryLeaf.F90:10-666 (GS_SWITCH leuning,
   FWSOIL_SWITCH standard, CALL_climate false) */
static void dryLeaf(Tile *t, Canopy *c, int iter, double dels,
                    int *npass_out) {
  const double co2cp3 = 0.0, jtomol = 4.6e-6;
  double vlaiw = t->canopy_vlaiw[0];
  double rlam = t->air_rlam[0];
  double gs_coeff[2] = {0.0, 0.0};
  if (iter == 1) {
    /* fwsoil_calc_std, src/science/canopy/cbl_fwsoil.F90:13-38 */
    double s = 0.0;
    for (int k = 0; k < MS; k++) {
      double r = (t->ssnow_wbliq[k] - t->soil_swilt_vec[k]) /
                 (t->soil_sfc_vec[k] - t->soil_swilt_vec[k]);
      double term = t->veg_froot[k] * MAX(1.0e-9, MIN(1.0, r));
      s = (k == 0) ? term : s + term;
    }
    double rwater = MAX(1.0e-9, s);
    c->fwsoil = MAX(1.0e-9, MIN(1.0, t->veg_vbeta[0] * rwater));
    t->canopy_fwsoil[0] = c->fwsoil;
  }
  double fwsoil = c->fwsoil;
  double gswmin[2];
  for (int l = 0; l < 2; l++) gswmin[l] = MAX(1.e-6, t->rad_scalex[l] * t->veg_gswmin[0]);
  double gw[2] = {1.0e-3, 1.0e-3}, gh[2] = {1.0e-3, 1.0e-3}, ghr[2] = {1.0e-3, 1.0e-3};
  double rdx[2] = {0.0, 0.0}, anx[2] = {0.0, 0.0}, an_y[2] = {0.0, 0.0}, rdy[2] = {0.0, 0.0};
  double rnx = c->sum_rad_rniso, abs_deltlf = 999.0, gras = 1.0e-6;
  double hcx = 0.0, ecx = c->sum_rad_rniso;
  c->hcy = 0.0;
  double tlfxx = c->tlfx;
  double psycst[2] = {t->air_psyc[0], t->air_psyc[0]};
  double ejmxt3[2], vcmxt3[2], vcmxt4[2], vx3[2], vx4[2], temp2[2];
  double cx1 = 0.0, cx2 = 0.0, deltlf = 0.0;
  double oldevapfbl[MS];
  t->canopy_fevc[0] = 0.0;
  for (int k = 0; k < MS; k++) t->ssnow_evapfbl[k] = 0.0;
  c->ghwet = 1.0e-3;
  t->canopy_fevw[0] = 0.0;
  t->canopy_fhvw[0] = 0.0;
  double sum_gbh = (c->gbhu[0] + c->gbhf[0]) + (c->gbhu[1] + c->gbhf[1]);
  if (vlaiw <= LAI_THRESH) {
    rnx = 0.0;
    ecx = 0.0;
    c->ecy = ecx;
    abs_deltlf = 0.0;
    c->rny = rnx;
  }
  double deltlfy = abs_deltlf;
  int k = 0, npass = 0;
  while (k < MAXITER) {
    k = k + 1;
    int active = (vlaiw > LAI_THRESH && abs_deltlf > 0.1);
    if (active) {
      npass++;
      double tlfx = c->tlfx;
      c->ghwet = 2.0 * sum_gbh;
      gras = MAX(1.0e-6, 1.595E8 * fabs(tlfx - t->met_tvair[0]) * p3(t->veg_dleaf[0]));
      c->gbhf[0] = t->rad_fvlai[0] * t->air_cmolar[0] * 0.5 * DHEAT * (pow(gras, 0.25)) /
                   t->veg_dleaf[0];
      c->gbhf[1] = t->rad_fvlai[1] * t->air_cmolar[0] * 0.5 * DHEAT * (pow(gras, 0.25)) /
                   t->veg_dleaf[0];
      c->gbhf[0] = MAX(1.e-6, c->gbhf[0]);
      c->gbhf[1] = MAX(1.e-6, c->gbhf[1]);
      for (int l = 0; l < 2; l++) {
        gh[l] = 2.0 * (c->gbhu[l] + c->gbhf[l]);
        ghr[l] = t->rad_gradis[l] + gh[l];
      }
      double frac4 = t->veg_frac4[0];
      double temp = xvcmxt3(tlfx) * t->veg_vcmax[0] * (1.0 - frac4);
      vcmxt3[0] = t->rad_scalex[0] * temp;
      vcmxt3[1] = t->rad_scalex[1] * temp;
      temp = xvcmxt4(tlfx - TFRZ) * t->veg_vcmax[0] * frac4;
      vcmxt4[0] = t->rad_scalex[0] * temp;
      vcmxt4[1] = t->rad_scalex[1] * temp;
      temp = xejmxt3(tlfx) * t->veg_ejmax[0] * (1.0 - frac4);
      ejmxt3[0] = t->rad_scalex[0] * temp;
      ejmxt3[1] = t->rad_scalex[1] * temp;
      double tdiff = tlfx - TREFK;
      double conkct = t->veg_conkc0[0] * exp((t->veg_ekc[0] / (RGAS * TREFK)) *
                                             (1.0 - TREFK / tlfx));
      double conkot = t->veg_conko0[0] * exp((t->veg_eko[0] / (RGAS * TREFK)) *
                                             (1.0 - TREFK / tlfx));
      tlfxx = tlfx;
      cx1 = conkct * (1.0 + 0.21 / conkot);
      cx2 = 2.0 * GAM0 * (1.0 + GAM1 * tdiff + GAM2 * tdiff * tdiff);
      double alpha = t->veg_alpha[0], convex = t->veg_convex[0];
      temp2[0] = t->rad_qcan[0] * jtomol * (1.0 - frac4);
      temp2[1] = t->rad_qcan[1] * jtomol * (1.0 - frac4);
      vx3[0] = ej3x(temp2[0], alpha, convex, ejmxt3[0]);
      vx3[1] = ej3x(temp2[1], alpha, convex, ejmxt3[1]);
      temp2[0] = t->rad_qcan[0] * jtomol * frac4;
      temp2[1] = t->rad_qcan[1] * jtomol * frac4;
      vx4[0] = ej4x(temp2[0], alpha, convex, vcmxt4[0]);
      vx4[1] = ej4x(temp2[1], alpha, convex, vcmxt4[1]);
      double cfrd = t->veg_cfrd[0];
      rdx[0] = (cfrd * vcmxt3[0] + cfrd * vcmxt4[0]);
      rdx[1] = (cfrd * vcmxt3[1] + cfrd * vcmxt4[1]);
      gs_coeff[0] = (fwsoil / (c->csx[0] - co2cp3)) *
                    (t->veg_a1gs[0] / (1.0 + c->dsx / t->veg_d0gs[0]));
      gs_coeff[1] = (fwsoil / (c->csx[1] - co2cp3)) *
                    (t->veg_a1gs[0] / (1.0 + c->dsx / t->veg_d0gs[0]));
    }
    /* photosynthesis, for every tile: anx zeroed, computed where active */
    anx[0] = anx[1] = 0.0;
    if (t->rad_fvlai[0] + t->rad_fvlai[1] > LAI_THRESH) {
      for (int l = 0; l < 2; l++)
        if (t->rad_fvlai[l] > LAI_THRESH && abs_deltlf > 0.1)
          anx[l] = photosynthesis_leaf(c->csx[l], cx1, cx2, gswmin[l], rdx[l], vcmxt3[l],
                                       vcmxt4[l], vx3[l], vx4[l], gs_coeff[l], fwsoil);
    }
    if (active) {
      for (int l = 0; l < 2; l++) {
        if (t->rad_fvlai[l] > LAI_THRESH) {
          c->csx[l] = t->met_ca[0] - RGBWC * anx[l] / (c->gbhu[l] + c->gbhf[l]);
          c->csx[l] = MAX(1.0e-4, c->csx[l]);
          t->canopy_gswx[l] =
              MAX(1.e-3, gswmin[l] * fwsoil + MAX(0.0, RGSWC * gs_coeff[l] * anx[l]));
          gw[l] = 1.0 / (1.0 / t->canopy_gswx[l] + 1.0 / (1.075 * (c->gbhu[l] + c->gbhf[l])));
          gw[l] = MAX(gw[l], 0.00001);
          psycst[l] = t->air_psyc[0] * (ghr[l] / gw[l]);
        }
      }
      double dsatdk = t->air_dsatdk[0], dtv = t->met_tvair[0] - t->met_tk[0];
      ecx = (dsatdk * (t->rad_rniso[0] - CAPP * RMAIR * dtv * t->rad_gradis[0]) +
             CAPP * RMAIR * t->met_dva[0] * ghr[0]) /
                (dsatdk + psycst[0]) +
            (dsatdk * (t->rad_rniso[1] - CAPP * RMAIR * dtv * t->rad_gradis[1]) +
             CAPP * RMAIR * t->met_dva[0] * ghr[1]) /
                (dsatdk + psycst[1]);
      double fwet = t->canopy_fwet[0];
      double local_fevc = (1.0 - fwet) * ecx;
      if (local_fevc > 0.0) {
        transp_soil_water(dels, t->soil_swilt_vec, t->veg_froot, t->soil_zse_vec,
                          local_fevc, t->ssnow_wbliq, t->ssnow_evapfbl);
        double s = t->ssnow_evapfbl[0];
        for (int q = 1; q < MS; q++) s = s + t->ssnow_evapfbl[q];
        t->canopy_fevc[0] = s * rlam / dels;
        ecx = t->canopy_fevc[0] / (1.0 - fwet);
      }
      hcx = (c->sum_rad_rniso - ecx - CAPP * RMAIR * dtv * c->sum_rad_gradis) *
            (gh[0] + gh[1]) / (ghr[0] + ghr[1]);
      c->tlfx = t->met_tvair[0] + hcx / (CAPP * RMAIR * (gh[0] + gh[1]));
      rnx = c->sum_rad_rniso - CAPP * RMAIR * (c->tlfx - t->met_tk[0]) * c->sum_rad_gradis;
      c->dsx = t->met_dva[0] + dsatdk * (c->tlfx - t->met_tvair[0]);
      c->dsx = MAX(c->dsx, 0.0);
      deltlf = tlfxx - c->tlfx;
      abs_deltlf = fabs(deltlf);
    }
    if (abs_deltlf < fabs(deltlfy)) {
      deltlfy = deltlf;
      c->tlfy = c->tlfx;
      c->rny = rnx;
      c->hcy = hcx;
      c->ecy = ecx;
      rdy[0] = rdx[0];
      rdy[1] = rdx[1];
      an_y[0] = anx[0];
      an_y[1] = anx[1];
      for (int q = 0; q < MS; q++) oldevapfbl[q] = t->ssnow_evapfbl[q];
    }
    if (abs_deltlf > 0.1) {
      double w = 0.5 * ((double)(k - 5 > 0 ? k - 5 : 0) / ((double)k - 4.9999));
      c->tlfx = w * tlfxx + (1.0 - w) * c->tlfx;
    }
    if (k == 1) {
      c->tlfy = c->tlfx;
      c->rny = rnx;
      c->hcy = hcx;
      c->ecy = ecx;
      rdy[0] = rdx[0];
      rdy[1] = rdx[1];
      an_y[0] = anx[0];
      an_y[1] = anx[1];
      for (int q = 0; q < MS; q++) oldevapfbl[q] = t->ssnow_evapfbl[q];
    }
  }
  double fwet = t->canopy_fwet[0];
  t->canopy_fevc[0] = (1.0 - fwet) * c->ecy;
  if (c->ecy > 0.0 && fwet < 1.0) {
    if (fabs(c->ecy - ecx) > 1.0e-6) {
      double s = oldevapfbl[0];
      for (int q = 1; q < MS; q++) s = s + oldevapfbl[q];
      if (fabs(t->canopy_fevc[0] - (s * rlam / dels)) > 1.0e-4) {
        /* the reference prints 'Error! oldevapfbl not right.' here */
      } else {
        for (int q = 0; q < MS; q++) t->ssnow_evapfbl[q] = oldevapfbl[q];
      }
    }
  }
  t->canopy_frday[0] = 12.0 * (rdy[0] + rdy[1]);
  t->canopy_fpn[0] = MIN(-12.0 * (an_y[0] + an_y[1]), t->canopy_frday[0]);
  if (npass_out) *npass_out += npass;
}

/* src/science/canopy/cbl_wetleaf.F90:9-111 */
static void wetLeaf(Tile *t, Canopy *c, double dels) {
  c->ghwet = 1.0e-3;
  double gwwet = 1.0e-3, ghrwet = 1.0e-3;
  t->canopy_fevw[0] = 0.0;
  t->canopy_fhvw[0] = 0.0;
  double sum_gbh = (c->gbhu[0] + c->gbhf[0]) + (c->gbhu[1] + c->gbhf[1]);
  if (t->canopy_vlaiw[0] > LAI_THRESH) {
    double rlam = t->air_rlam[0], dsatdk = t->air_dsatdk[0], psyc = t->air_psyc[0];
    double srn = c->sum_rad_rniso, sg = c->sum_rad_gradis, fwet = t->canopy_fwet[0];
    c->ghwet = 2.0 * sum_gbh;
    gwwet = 1.075 * sum_gbh;
    ghrwet = sg + c->ghwet;
    double ccfevw = MIN(t->canopy_cansto[0] * rlam / dels,
                        2.0 / (1440.0 / (dels / 60.0)) * rlam);
    double dtv = t->met_tvair[0] - t->met_tk[0];
    t->canopy_fevw[0] = MIN(fwet * (dsatdk * (srn - CAPP * RMAIR * dtv * sg) +
                                    CAPP * RMAIR * t->met_dva[0] * ghrwet) /
                                (dsatdk + psyc * ghrwet / gwwet),
                            ccfevw);
    t->canopy_fevw_pot[0] = (dsatdk * (srn - CAPP * RMAIR * dtv * sg) +
                             CAPP * RMAIR * t->met_dva[0] * ghrwet) /
                            (dsatdk + psyc * ghrwet / gwwet);
    t->canopy_fhvw[0] =
        fwet * (srn - CAPP * RMAIR * (c->tlfy - t->met_tk[0]) * sg) - t->canopy_fevw[0];
  }
}

/* src/science/canopy/cbl_pot_evap_snow.F90:79-167 (HDM, no litter/or_evap) */
static double humidity_deficit_method(Tile *t, double dq, double dqu) {
  if (t->ssnow_snowd[0] > 1.0 || t->ssnow_tgg[0] == TFRZ) {
    dq = MAX(-0.1e-3, dq);
    dqu = MAX(-0.1e-3, dqu);
  }
  if (dq <= 0.0 && dqu < dq) dqu = dq;
  if (dq >= 0.0 && dqu < 0.0) dqu = 0.0;
  (void)dqu;
  return t->air_rho[0] * t->air_rlam[0] * dq / t->ssnow_rtsoil[0];
}

/* src/science/canopy/cbl_latent_heat.F90:15-285 */
static void latent_heat_flux(Tile *t, const Globals *g, double dels) {
  double potev = t->ssnow_potev[0], rlam = t->air_rlam[0], snowd = t->ssnow_snowd[0];
  if (potev < 0.) t->ssnow_wetfac[0] = 1.0;
  double wetfac = t->ssnow_wetfac[0];
  double fess = wetfac * potev;
  double pwet = MAX(0., MIN(0.2, t->ssnow_pudsto[0] / MAX(1., t->ssnow_pudsmx[0])));
  fess = fess * (1. - pwet);
  double frescale = g->zse[0] * DENSITY_LIQ * rlam / dels;
  if (snowd < 0.1 && fess > 0.) {
    double flower = t->ssnow_wb[0] - t->soil_swilt[0] / 2.0;
    double fupper = MAX(0., flower * frescale - t->ssnow_evapfbl[0] * rlam / dels);
    fess = MIN(fess, fupper);
    fupper = (t->ssnow_wb[0] - t->ssnow_wbice[0] / FROZEN_LIMIT) * frescale;
    fupper = MAX(fupper, 0.);
    fess = MIN(fess, fupper);
  }
  double cls = 1.;
  if (snowd >= 0.1) {
    cls = 1.1335;
    fess = cls * potev;
  }
  if (snowd < 0.1 && potev < 0. && t->ssnow_tss[0] < TFRZ) {
    cls = 1.1335;
    fess = cls * potev;
  }
  if (snowd >= 0.1 && potev > 0.) {
    cls = 1.1335;
    fess = MIN((wetfac * potev) * cls, snowd / dels * rlam * cls);
  }
  t->ssnow_cls[0] = cls;
  double fesp = MIN(t->ssnow_pudsto[0] / dels * rlam, MAX(pwet * potev, 0.));
  t->canopy_fess[0] = fess;
  t->canopy_fesp[0] = fesp;
  t->canopy_fes[0] = fess + fesp;
}

/* src/science/canopy/cbl_within_canopy.F90:10-159 (or_evap false) */
static void within_canopy(Tile *t, Canopy *c, double rt0, double *qstvair) {
  const double rhlitt = 0.0, relitt = 0.0;
  double cmolar = t->air_cmolar[0];
  double rrbw = ((c->gbhu[0] + c->gbhf[0]) + (c->gbhu[1] + c->gbhf[1])) / cmolar;
  double rrsw = (t->canopy_gswx[0] + t->canopy_gswx[1]) / cmolar;
  double fix_eqn = t->ssnow_cls[0] * rt0 / (rt0 + relitt);
  if (t->ssnow_potev[0] > 0.) fix_eqn = fix_eqn * t->ssnow_wetfac[0];
  double fix_eqn2 = rt0 / (rt0 + rhlitt);
  if (t->veg_meth[0] > 0 && t->canopy_vlaiw[0] > LAI_THRESH &&
      t->rough_hruff[0] > t->rough_z0soilsn[0]) {
    double epsi = t->air_epsi[0], rt1 = t->rough_rt1[0], rlam = t->air_rlam[0];
    double rho = t->air_rho[0];
    double dmah = (rt0 + fix_eqn2 * rt1) * ((1. + epsi) * rrsw + rrbw) +
                  epsi * (rt0 * rt1) * (rrbw * rrsw);
    double dmbh = (-rlam / CAPP) * (rt0 * rt1) * (rrbw * rrsw);
    double dmch = ((1. + epsi) * rrsw + rrbw) * rt0 * rt1 *
                  (t->canopy_fhv[0] + t->canopy_fhs[0]) / (rho * CAPP);
    double dmae = (-epsi * CAPP / rlam) * (rt0 * rt1) * (rrbw * rrsw);
    double dmbe = (rt0 + fix_eqn * rt1) * ((1. + epsi) * rrsw + rrbw) +
                  (rt0 * rt1) * (rrbw * rrsw);
    double dmce = ((1. + epsi) * rrsw + rrbw) * rt0 * rt1 *
                  (t->canopy_fev[0] + t->canopy_fes[0] / t->ssnow_cls[0]) / (rho * rlam);
    double tk = t->met_tk[0], tss = t->ssnow_tss[0];
    double tvair = tk + (dmbe * dmch - dmbh * dmce) /
                            (dmah * dmbe - dmae * dmbh + 1.0e-12);
    double lo = MIN(tss, tk) - 5.0, hi = MAX(tss, tk) + 5.0;
    tvair = MAX(tvair, lo);
    tvair = MIN(tvair, hi);
    t->met_tvair[0] = tvair;
    double qv = t->met_qv[0];
    double qvair = qv + (dmah * dmce - dmae * dmch) / (dmah * dmbe - dmae * dmbh + 1.0e-12);
    qvair = MAX(0.0, qvair);
    lo = MIN(t->ssnow_qstss[0], qv);
    hi = MAX(t->ssnow_qstss[0], qv);
    qvair = MAX(qvair, lo);
    qvair = MIN(qvair, hi);
    t->met_qvair[0] = qvair;
    *qstvair = qsatfjh(tvair - TFRZ, t->met_pmb[0]);
    t->met_dva[0] = (*qstvair - qvair) * RMAIR / RMH2O * t->met_pmb[0] * 100.;
  }
}

/* src/science/canopy/cbl_zetar.F90:13-159 (NITER = 4, soil_struc default) */
static void update_zetar(Tile *t, int iter, int *iterplus) {
  if (iter < 4) {
    *iterplus = (iter + 1 > 2) ? iter + 1 : 2;
    double us = t->canopy_us[0];
    double z = -(VONK * GRAV * t->rough_zref_tq[0] *
                 (t->canopy_fh[0] + 0.07 * t->canopy_fe[0])) /
               (t->air_rho[0] * CAPP * t->met_tk[0] * p3(us));
    z = MIN(ZETPOS, z);
    z = MAX(ZETNEG, z);
    t->canopy_zetar[*iterplus - 1] = z;
  }
}

/* src/science/canopy/cable_canopy.F90:10-1048 */
static void define_canopy(Tile *t, const Globals *g, double dels, int sunlit_veg,
                          int *npass) {
  Canopy C, *c = &C;
  memset(c, 0, sizeof(C));
  double vlaiw = t->canopy_vlaiw[0];
  int veg = vlaiw > LAI_THRESH;
  t->canopy_cansto[0] = t->canopy_oldcansto[0];
  c->cansat = t->veg_canst1[0] * vlaiw;
  { /* Surf_wetness_fact, src/science/canopy/cbl_SurfaceWetness.F90:10-79 */
    double precip = t->met_precip[0], psn = t->met_precip_sn[0];
    double upper = 4.0 * MIN(dels, 1800.0) / (60.0 * 1440.0);
    double ftemp = MIN(precip - psn, upper);
    double lower = c->cansat - t->canopy_cansto[0];
    upper = MAX(lower, 0.0);
    double wcint = (ftemp > 0.0 && t->met_tk[0] > TFRZ) ? MIN(upper, ftemp) : 0.0;
    t->canopy_wcint[0] = wcint;
    t->canopy_through[0] = psn + MIN(precip - psn, MAX(0.0, precip - psn - wcint));
    t->canopy_cansto[0] = t->canopy_cansto[0] + wcint;
    t->canopy_fwet[0] =
        MAX(0.0, MIN(0.9, 0.8 * t->canopy_cansto[0] / MAX(c->cansat, 0.01)));
    t->ssnow_satfrac[0] = 1.0e-8;
    t->ssnow_rh_srf[0] = 1.0;
    /* initialize_wetfac, src/science/canopy/cbl_init_wetfac_mod.F90:9-116 */
    double wilting_pt = t->soil_swilt[0] / 2.0;
    double num = t->ssnow_wb[0] - wilting_pt;
    double den = t->soil_sfc[0] - wilting_pt;
    den = MAX(0.0830, den);
    double wf = num / den;
    wf = MIN(1.0, wf);
    wf = MAX(0.0, wf);
    if (t->ssnow_wbice[0] > 0.0) {
      double ice_ratio = p2(t->ssnow_wbice[0] / t->ssnow_wb[0]);
      double ice_factor = 1.0 - MIN(0.2, ice_ratio);
      ice_factor = MAX(0.5, ice_factor);
      wf = wf * ice_factor;
    }
    if (t->ssnow_snowd[0] > 0.1) wf = 0.9;
    if ((int)t->veg_iveg[0] == LAKES) {
      if (t->met_tk[0] >= TFRZ + 5.) wf = 1.0;
      if (t->met_tk[0] < TFRZ + 5.) wf = 0.7;
    }
    t->ssnow_wetfac[0] = 0.5 * (wf + t->ssnow_owetfac[0]);
  }
  t->canopy_fevw_pot[0] = 0.0;
  t->canopy_gswx[0] = t->canopy_gswx[1] = 1e-3;
  c->gbhf[0] = c->gbhf[1] = 1e-3;
  c->gbhu[0] = c->gbhu[1] = 1e-3;
  for (int k = 0; k < MS; k++) t->ssnow_evapfbl[k] = 0.0, t->ssnow_rex[k] = 0.0;
  c->csx[0] = c->csx[1] = t->met_ca[0];
  t->met_tvair[0] = t->met_tk[0];
  t->met_qvair[0] = t->met_qv[0];
  t->canopy_tv[0] = t->met_tvair[0];
  t->canopy_fwsoil[0] = 1.0;
  t->canopy_kthlitt[0] = 0.3;
  t->canopy_dvlitt[0] = 3.1415841138194147e-05;
  define_air(t);
  double qstvair = qsatfjh(t->met_tvair[0] - TFRZ, t->met_pmb[0]);
  t->met_dva[0] = (qstvair - t->met_qvair[0]) * RMAIR / RMH2O * t->met_pmb[0] * 100.0;
  c->dsx = t->met_dva[0];
  c->dsx = MAX(c->dsx, 0.0);
  c->tlfx = t->met_tk[0] + leaf_start_offset;
  c->tlfy = t->met_tk[0] + leaf_start_offset;
  double ortsoil = t->ssnow_rtsoil[0];
  int isflag = (int)t->ssnow_isflag[0];
  t->ssnow_tss[0] = (double)(1 - isflag) * t->ssnow_tgg[0] + (double)isflag * t->ssnow_tggsn[0];
  double tss = t->ssnow_tss[0];
  double tss4 = p4(tss);
  t->canopy_fes[0] = 0.;
  t->canopy_fess[0] = 0.;
  t->canopy_fesp[0] = 0.;
  t->ssnow_potev[0] = 0.;
  t->canopy_fevw_pot[0] = 0.;
  t->canopy_fhs_cor[0] = 0.0;
  t->canopy_fns_cor[0] = 0.0;
  t->canopy_ga_cor[0] = 0.0;
  radiation(t, sunlit_veg);
  t->canopy_zetar[0] = ZETA0;
  t->canopy_zetar[1] = ZETPOS + 1;
  t->canopy_zetash[0] = ZETA0;
  t->canopy_zetash[1] = ZETPOS + 1;
  c->sum_rad_rniso = t->rad_rniso[0] + t->rad_rniso[1];
  c->sum_rad_gradis = t->rad_gradis[0] + t->rad_gradis[1];
  double rt0 = 0.0, rt1usc = 0.0;
  int iterplus = 0;
  for (int iter = 1; iter <= 4; iter++) {
    /* comp_friction_vel, src/science/canopy/cbl_friction_vel.F90:19-108 */
    double zr = t->canopy_zetar[iter - 1];
    double zuv = t->rough_zref_uv[0], ztq = t->rough_zref_tq[0], z0m = t->rough_z0m[0];
    double psim_1 = psim(zr * zuv / ztq);
    double rescale = VONK * MAX(t->met_ua[0], UMIN);
    double z_eff = zuv / z0m;
    double psim_2 = psim(zr * z0m / ztq);
    double ll = rescale / (log(z_eff) - psim_1 + psim_2);
    t->canopy_us[0] = MIN(MAX(1.e-6, ll), 10.0);
    double us = t->canopy_us[0];
    double disp = t->rough_disp[0], zruffs = t->rough_zruffs[0],
           z0soilsn = t->rough_z0soilsn[0];
    double xx = 0.5 + SIGN(0.5, ztq + disp - zruffs);
    rt1usc = xx * (log(ztq / MAX(zruffs - disp, z0soilsn)) - psis(zr) +
                   psis(zr * (MAX(zruffs - disp, z0soilsn)) / ztq)) /
             VONK;
    const double rt_min = 5.;
    rt0 = MAX(rt_min, t->rough_rt0us[0] / us);
    t->rough_rt1[0] = MAX(5., (t->rough_rt1usa[0] + t->rough_rt1usb[0] + rt1usc) / us);
    double rtsoil = veg ? rt0 : rt0 + t->rough_rt1[0];
    rtsoil = MAX(rt_min, rtsoil);
    if (rtsoil > 2. * ortsoil || rtsoil < 0.5 * ortsoil)
      rtsoil = MAX(rt_min, 0.5 * (rtsoil + ortsoil));
    t->ssnow_rtsoil[0] = rtsoil;
    if (veg) {
      double gbvtop = t->air_cmolar[0] * APOL * t->air_visc[0] / PRANDT / t->veg_dleaf[0] *
                      sqrt(us / MAX(t->rough_usuh[0], 1.e-6) * t->veg_dleaf[0] /
                           t->air_visc[0]) *
                      pow(PRANDT, 1.0 / 3.0) / t->veg_shelrb[0];
      gbvtop = MAX(0.05, gbvtop);
      double coexp = t->rough_coexp[0], extkb = t->rad_extkb[0];
      c->gbhu[0] = gbvtop * (1.0 - exp(-MIN(vlaiw * (0.5 * coexp + extkb), 20.0))) /
                   (extkb + 0.5 * coexp);
      c->gbhu[1] = (2.0 / coexp) * gbvtop *
                       (1.0 - exp(-MIN(0.5 * coexp * vlaiw, 20.0))) -
                   c->gbhu[0];
    }
    c->rny = c->sum_rad_rniso;
    c->hcy = 0.0;
    c->ecy = c->rny - c->hcy;
    dryLeaf(t, c, iter, dels, npass);
    wetLeaf(t, c, dels);
    t->canopy_fev[0] = t->canopy_fevc[0] + t->canopy_fevw[0];
    double fwet = t->canopy_fwet[0];
    t->canopy_fhv[0] = (1.0 - fwet) * c->hcy + t->canopy_fhvw[0];
    t->canopy_fnv[0] = (1.0 - fwet) * c->rny + t->canopy_fevw[0] + t->canopy_fhvw[0];
    double tvrad = t->met_tvrad[0], transd = t->rad_transd[0];
    if (veg && t->rough_hruff[0] > z0soilsn) {
      t->rad_lwabv[0] = CAPP * RMAIR * (c->tlfy - t->met_tk[0]) * c->sum_rad_gradis;
      double q = t->rad_lwabv[0] / (2.0 * (1.0 - transd) * SBOLTZ * EMLEAF) + p4(tvrad);
      if (q > 0.0)
        t->canopy_tv[0] = pow(q, 0.25);
      else
        t->canopy_tv[0] = tvrad;
    } else {
      t->canopy_tv[0] = tvrad;
    }
    t->canopy_fns[0] = t->rad_qssabs[0] + transd * t->met_fld[0] +
                       (1.0 - transd) * EMLEAF * SBOLTZ * p4(t->canopy_tv[0]) -
                       EMSOIL * SBOLTZ * tss4;
    t->ssnow_qstss[0] = qsatfjh(tss - TFRZ, t->met_pmb[0]);
    {
      double dq = t->ssnow_qstss[0] - t->met_qv[0];
      double dqu = t->ssnow_rh_srf[0] * t->ssnow_qstss[0] - t->met_qv[0];
      t->ssnow_potev[0] = humidity_deficit_method(t, dq, dqu);
    }
    latent_heat_flux(t, g, dels);
    t->canopy_fhs[0] =
        t->air_rho[0] * CAPP * (tss - t->met_tvair[0]) / t->ssnow_rtsoil[0];
    within_canopy(t, c, rt0, &qstvair);
    t->ssnow_qstss[0] = qsatfjh(tss - TFRZ, t->met_pmb[0]);
    {
      double dq = t->ssnow_qstss[0] - t->met_qvair[0];
      double dqu = t->ssnow_rh_srf[0] * t->ssnow_qstss[0] - t->met_qvair[0];
      t->ssnow_potev[0] = humidity_deficit_method(t, dq, dqu);
    }
    latent_heat_flux(t, g, dels);
    t->canopy_fhs[0] =
        t->air_rho[0] * CAPP * (tss - t->met_tvair[0]) / t->ssnow_rtsoil[0];
    t->canopy_ga[0] = t->canopy_fns[0] - t->canopy_fhs[0] - t->canopy_fes[0];
    t->canopy_fe[0] = t->canopy_fev[0] + t->canopy_fes[0];
    t->canopy_fh[0] = t->canopy_fhv[0] + t->canopy_fhs[0];
    if (t->ssnow_potev[0] >= 0.)
      t->ssnow_potev[0] = MAX(0.00001, t->ssnow_potev[0]);
    else
      t->ssnow_potev[0] = MIN(-0.0002, t->ssnow_potev[0]);
    if (t->canopy_fevw_pot[0] >= 0.)
      t->canopy_fevw_pot[0] = MAX(0.000001, t->canopy_fevw_pot[0]);
    else
      t->canopy_fevw_pot[0] = MIN(-0.002, t->canopy_fevw_pot[0]);
    t->canopy_rnet[0] = t->canopy_fnv[0] + t->canopy_fns[0];
    t->canopy_rniso[0] = c->sum_rad_rniso + t->rad_qssabs[0] + transd * t->met_fld[0] +
                         (1.0 - transd) * EMLEAF * SBOLTZ * p4(tvrad) -
                         EMSOIL * SBOLTZ * p4(tvrad);
    t->canopy_epot[0] =
        (t->canopy_fevw_pot[0] + t->ssnow_potev[0] / t->ssnow_cls[0]) * dels / t->air_rlam[0];
    double rll = t->canopy_epot[0] * t->air_rlam[0] / dels;
    if (rll == 0) rll = 1.e-7;
    t->canopy_wetfac_cs[0] = MAX(0., MIN(1.0, t->canopy_fe[0] / rll));
    if (t->canopy_wetfac_cs[0] <= 0.)
      t->canopy_wetfac_cs[0] =
          MAX(0., MIN(1., MAX(t->canopy_fev[0] / t->canopy_fevw_pot[0],
                              t->canopy_fes[0] / t->ssnow_potev[0])));
    update_zetar(t, iter, &iterplus);
  }
  double us = t->canopy_us[0], ua = t->met_ua[0];
  t->canopy_cduv[0] = us * us / p2(MAX(ua, UMIN));
  {
    double lai_min = MAX(LAI_THRESH, vlaiw);
    double rsun = t->rad_fvlai[0] / lai_min, rshd = t->rad_fvlai[1] / lai_min;
    double cc = rsun * t->canopy_gswx[0] + rshd * t->canopy_gswx[1];
    double mincc = MAX(1.e-06, cc);
    double transd = t->rad_transd[0];
    cc = (1. - transd) * mincc;
    double relm = t->ssnow_wb[0] / t->soil_sfc[0];
    double sc = transd * p2(0.01 * relm);
    double surf = cc + sc;
    if ((int)t->soil_isoilm[0] == ICE_SOIL) surf = 1.e6;
    t->canopy_gswx_t[0] = surf;
  }
  double zuv = t->rough_zref_uv[0], ztq = t->rough_zref_tq[0], z0m = t->rough_z0m[0];
  double z4 = t->canopy_zetar[3];
  t->canopy_cdtq[0] = t->canopy_cduv[0] *
                      (log(zuv / z0m) - psim(z4 * zuv / ztq) + psim(z4 * z0m / ztq)) /
                      (log(ztq / (0.1 * z0m)) - psis(z4) + psis(z4 * 0.1 * z0m / ztq));
  double rho = t->air_rho[0];
  double tstar = -t->canopy_fh[0] / (rho * CAPP * us);
  double qstar = -t->canopy_fe[0] / (rho * t->air_rlam[0] * us * t->ssnow_cls[0]);
  double disp = t->rough_disp[0], hruff = t->rough_hruff[0], z0soilsn = t->rough_z0soilsn[0];
  double zscrn = MAX(z0m, 2.0 - disp);
  double zp = t->canopy_zetar[iterplus - 1];
  double ftemp = (log(ztq / zscrn) - psis(zp) + psis(zp * zscrn / ztq)) / VONK;
  t->canopy_tscrn[0] = t->met_tk[0] - TFRZ - tstar * ftemp;
  double term1 = 0., term2 = 0., term5 = 0., term3 = 0., r_sc = 0.;
  double zscl = MAX(z0soilsn, 2.0);
  double rghlai = t->canopy_rghlai[0];
  double rt0us = t->rough_rt0us[0], rt1usa = t->rough_rt1usa[0],
         rt1usb = t->rough_rt1usb[0], zruffs = t->rough_zruffs[0];
  int scr = (vlaiw > LAI_THRESH && hruff > 0.01);
  if (scr) {
    if (disp > 0.0) {
      term1 = exp(2 * CSW * rghlai * (1 - zscl / hruff));
      term2 = exp(2 * CSW * rghlai * (1 - disp / hruff));
      term5 = MAX(2. / 3. * hruff / disp, 1.);
    }
    term3 = A33 * A33 * CTL * 2 * CSW * rghlai;
    if (zscl < disp) {
      r_sc = term5 * log(zscl / z0soilsn) * (exp(2 * CSW * rghlai) - term2) / term3;
      r_sc = r_sc + term5 * log(disp / zscl) * (exp(2 * CSW * rghlai) - term1) / term3;
    } else if (disp <= zscl && zscl < hruff) {
      r_sc = rt0us + term5 * (term2 - term1) / term3;
    } else if (hruff <= zscl && zscl < zruffs) {
      r_sc = rt0us + rt1usa + term5 * (zscl - hruff) / (A33 * A33 * CTL * hruff);
    } else if (zscl >= zruffs) {
      r_sc = rt0us + rt1usa + rt1usb +
             (log((zscl - disp) / MAX(zruffs - disp, z0soilsn)) -
              psis((zscl - disp) * zp / ztq) + psis((zruffs - disp) * zp / ztq)) /
                 VONK;
    }
    double tss = t->ssnow_tss[0];
    t->canopy_tscrn[0] =
        tss + (t->met_tk[0] - tss) *
                  MIN(1., (r_sc / MAX(1., rt0us + rt1usa + rt1usb + rt1usc))) -
        TFRZ;
  }
  double rsts = qsatfjh(t->canopy_tscrn[0], t->met_pmb[0]);
  double wetfac = t->ssnow_wetfac[0], qv = t->met_qv[0];
  double qtgnet = rsts * wetfac - qv;
  double qsurf = (qtgnet > 0.) ? rsts * wetfac : 0.1 * rsts * wetfac + 0.9 * qv;
  t->canopy_qmom[0] = rho * (us * us);
  t->canopy_qscrn[0] = qv - qstar * ftemp;
  if (scr)
    t->canopy_qscrn[0] =
        qsurf + (qv - qsurf) * MIN(1., (r_sc / MAX(1., rt0us + rt1usa + rt1usb + rt1usc)));
  double rlam = t->air_rlam[0];
  t->canopy_dewmm[0] =
      -(MIN(0.0, t->canopy_fevw[0]) + MIN(0.0, t->canopy_fevc[0])) * dels / rlam;
  t->canopy_cansto[0] = t->canopy_cansto[0] + t->canopy_dewmm[0];
  t->canopy_cansto[0] =
      MAX(t->canopy_cansto[0] - MAX(0.0, t->canopy_fevw[0]) * dels / rlam, 0.0);
  t->canopy_spill[0] = MAX(0.0, t->canopy_cansto[0] - c->cansat);
  t->canopy_through[0] = t->canopy_through[0] + t->canopy_spill[0];
  t->canopy_precis[0] = MAX(0., t->canopy_through[0]);
  t->canopy_cansto[0] = t->canopy_cansto[0] - t->canopy_spill[0];
  t->canopy_delwc[0] = t->canopy_cansto[0] - t->canopy_oldcansto[0];
  double tssn = t->ssnow_tss[0];
  t->ssnow_dfn_dtg[0] = (-1.) * 4. * EMSOIL * SBOLTZ * tss4 / tssn;
  double rttsoil = t->ssnow_rtsoil[0];
  t->ssnow_dfh_dtg[0] = rho * CAPP / rttsoil;
  t->ssnow_dfe_ddq[0] = t->ssnow_wetfac[0] * rho * rlam * t->ssnow_cls[0] / rttsoil;
  t->ssnow_ddq_dtg[0] = (RMH2O / RMAIR) / t->met_pmb[0] * TETENA * TETENB * TETENC /
                        (p2(TETENC + tssn - TFRZ)) *
                        exp(TETENB * (tssn - TFRZ) / (TETENC + tssn - TFRZ));
  t->ssnow_dfe_dtg[0] = t->ssnow_dfe_ddq[0] * t->ssnow_ddq_dtg[0];
  t->canopy_dgdtg[0] = t->ssnow_dfn_dtg[0] - t->ssnow_dfh_dtg[0] - t->ssnow_dfe_dtg[0];
  double dtl = c->tlfy - t->met_tk[0];
  t->bal_drybal[0] = (c->ecy + c->hcy) - c->sum_rad_rniso + CAPP * RMAIR * dtl * c->sum_rad_gradis;
  t->bal_wetbal[0] = t->canopy_fevw[0] + t->canopy_fhvw[0] - c->sum_rad_rniso * t->canopy_fwet[0] +
                     CAPP * RMAIR * dtl * c->sum_rad_gradis * t->canopy_fwet[0];
  double *qc = t->rad_qcan;
  t->rad_swnet[0] = (qc[0] + qc[1]) + (qc[2] + qc[3]) + t->rad_qssabs[0];
  double transd = t->rad_transd[0];
  t->rad_lwnet[0] = t->met_fld[0] - SBOLTZ * EMLEAF * p4(t->canopy_tv[0]) * (1 - transd) -
                    t->rad_flws[0] * transd;
  t->rad_rnet[0] = t->rad_swnet[0] + t->rad_lwnet[0];
}

/* ===================================================================== */
/* soil and snow */

/* src/science/soilsnow/cbl_trimb.F90:17-53 */
static void trimb(const double *a, const double *b, const double *c, double *rhs,
                  int kmax) {
  double e[9], temp[9], g[9];
  e[0] = c[0] / b[0];
  for (int k = 1; k < kmax - 1; k++) {
    temp[k] = 1. / (b[k] - a[k] * e[k - 1]);
    e[k] = c[k] * temp[k];
  }
  g[0] = rhs[0] / b[0];
  for (int k = 1; k < kmax - 1; k++) g[k] = (rhs[k] - a[k] * g[k - 1]) * temp[k];
  rhs[kmax - 1] = (rhs[kmax - 1] - a[kmax - 1] * g[kmax - 2]) /
                  (b[kmax - 1] - a[kmax - 1] * e[kmax - 2]);
  for (int k = kmax - 2; k >= 0; k--) rhs[k] = g[k] - e[k] * rhs[k + 1];
}

/* src/science/soilsnow/cbl_snowCheck.F90:9-100 */
static void snowcheck(Tile *t) {
  double snowd = t->ssnow_snowd[0];
  double *ssdn = t->ssnow_ssdn, *tggsn = t->ssnow_tggsn, *sdepth = t->ssnow_sdepth,
         *smass = t->ssnow_smass;
  if (snowd <= 0.0) {
    t->ssnow_isflag[0] = 0;
    ssdn[0] = ssdn[1] = ssdn[2] = 120.0;
    t->ssnow_ssdnn[0] = 120.0;
    tggsn[0] = tggsn[1] = tggsn[2] = TFRZ;
    sdepth[0] = snowd / ssdn[0];
    sdepth[1] = 0.;
    sdepth[2] = 0.;
    smass[0] = snowd;
    smass[1] = 0.0;
    smass[2] = 0.0;
  } else if (snowd < SNMIN * t->ssnow_ssdnn[0]) {
    if ((int)t->ssnow_isflag[0] == 1) {
      ssdn[0] = t->ssnow_ssdnn[0];
      t->ssnow_tgg[0] = tggsn[0];
    }
    t->ssnow_isflag[0] = 0;
    t->ssnow_ssdnn[0] = MIN(400.0, MAX(120.0, ssdn[0]));
    double tm = MIN(TFRZ, t->ssnow_tgg[0]);
    tggsn[0] = tggsn[1] = tggsn[2] = tm;
    sdepth[0] = snowd / ssdn[0];
    sdepth[1] = 0.0;
    sdepth[2] = 0.0;
    smass[0] = snowd;
    smass[1] = 0.0;
    smass[2] = 0.0;
    ssdn[0] = ssdn[1] = ssdn[2] = t->ssnow_ssdnn[0];
  } else {
    if ((int)t->ssnow_isflag[0] == 0) {
      double tm = MIN(TFRZ, t->ssnow_tgg[0]);
      tggsn[0] = tggsn[1] = tggsn[2] = tm;
      ssdn[1] = ssdn[0];
      ssdn[2] = ssdn[0];
      sdepth[0] = t->ssnow_t_snwlr[0];
      smass[0] = t->ssnow_t_snwlr[0] * ssdn[0];
      smass[1] = (snowd - smass[0]) * 0.4;
      smass[2] = (snowd - smass[0]) * 0.6;
      sdepth[1] = smass[1] / ssdn[1];
      sdepth[2] = smass[2] / ssdn[2];
      t->ssnow_ssdnn[0] =
          (ssdn[0] * smass[0] + ssdn[1] * smass[1] + ssdn[2] * smass[2]) / snowd;
    }
    t->ssnow_isflag[0] = 1;
  }
}

/* src/science/soilsnow/cbl_snowDensity.F90:9-102 */
static void snowdensity(Tile *t, double dels) {
  double *ssdn = t->ssnow_ssdn, *tggsn = t->ssnow_tggsn, *smass = t->ssnow_smass;
  double snowd = t->ssnow_snowd[0];
  int isflag = (int)t->ssnow_isflag[0];
  double tmin1 = MIN(TFRZ, t->ssnow_tgg[0]);
  if (snowd > 0.1 && isflag == 0) {
    ssdn[0] = MIN(MAX_SSDN,
                  MAX(120.0, ssdn[0] + dels * ssdn[0] * 3.1e-6 *
                                           exp(-0.03 * (273.15 - tmin1) -
                                               (ssdn[0] >= 150.0 ? 0.046 : 0.0) *
                                                   (ssdn[0] - 150.0))));
    ssdn[0] = MIN(MAX_SSDN, ssdn[0] + dels * 9.806 * ssdn[0] * 0.75 * snowd /
                                          (3.0e7 * exp(0.021 * ssdn[0] +
                                                       0.081 * (273.15 - MIN(TFRZ, t->ssnow_tgg[0])))));
    if ((int)t->soil_isoilm[0] != 9) ssdn[0] = MIN(450.0, ssdn[0]);
    t->ssnow_sconds[0] = MAX(0.2, MIN(2.876e-6 * p2(ssdn[0]) + 0.074, MAX_SCONDS));
    t->ssnow_sconds[1] = t->ssnow_sconds[0];
    t->ssnow_sconds[2] = t->ssnow_sconds[0];
    t->ssnow_ssdnn[0] = ssdn[0];
    ssdn[1] = ssdn[0];
    ssdn[2] = ssdn[0];
  }
  if (isflag == 1) {
    for (int k = 0; k < 3; k++)
      ssdn[k] = ssdn[k] + dels * ssdn[k] * 3.1e-6 *
                              exp(-0.03 * (273.15 - MIN(TFRZ, tggsn[k])) -
                                  (ssdn[k] >= 150.0 ? 0.046 : 0.0) * (ssdn[k] - 150.0));
    double tw = t->ssnow_t_snwlr[0];
    ssdn[0] = ssdn[0] + dels * 9.806 * ssdn[0] * tw * ssdn[0] /
                            (3.0e7 * exp(.021 * ssdn[0] + 0.081 * (273.15 - MIN(TFRZ, tggsn[0]))));
    ssdn[1] = ssdn[1] + dels * 9.806 * ssdn[1] * (tw * ssdn[0] + 0.5 * smass[1]) /
                            (3.0e7 * exp(.021 * ssdn[1] + 0.081 * (273.15 - MIN(TFRZ, tggsn[1]))));
    ssdn[2] = ssdn[2] + dels * 9.806 * ssdn[2] * (tw * ssdn[0] + smass[1] + 0.5 * smass[2]) /
                            (3.0e7 * exp(.021 * ssdn[2] + 0.081 * (273.15 - MIN(TFRZ, tggsn[2]))));
    for (int k = 0; k < 3; k++) t->ssnow_sdepth[k] = smass[k] / ssdn[k];
    t->ssnow_ssdnn[0] = (ssdn[0] * smass[0] + ssdn[1] * smass[1] + ssdn[2] * smass[2]) / snowd;
    for (int k = 0; k < 3; k++)
      t->ssnow_sconds[k] = MAX(0.2, MIN(2.876e-6 * p2(ssdn[k]) + 0.074, MAX_SCONDS));
  }
}

/* src/science/soilsnow/cbl_snowAccum.F90:10-188 */
static void snow_accum(Tile *t, double dels) {
  double *ssdn = t->ssnow_ssdn, *smass = t->ssnow_smass, *sdepth = t->ssnow_sdepth,
         *tggsn = t->ssnow_tggsn;
  double psn = t->met_precip_sn[0];
  int isflag = (int)t->ssnow_isflag[0];
  int notice = (int)t->soil_isoilm[0] != 9;
  double osnowd = t->ssnow_osnowd[0];
  if (t->canopy_precis[0] > 0.0 && isflag == 0) {
    t->ssnow_snowd[0] = MAX(t->ssnow_snowd[0] + psn, 0.0);
    t->canopy_precis[0] = t->canopy_precis[0] - psn;
    ssdn[0] = MAX(120.0, ssdn[0] * osnowd / MAX(0.01, t->ssnow_snowd[0]) +
                             120.0 * psn / MAX(0.01, t->ssnow_snowd[0]));
    t->ssnow_ssdnn[0] = ssdn[0];
    if (t->canopy_precis[0] > 0.0 && t->ssnow_tgg[0] < TFRZ) {
      double pr = t->canopy_precis[0];
      t->ssnow_snowd[0] = MAX(t->ssnow_snowd[0] + pr, 0.0);
      t->ssnow_tgg[0] = t->ssnow_tgg[0] + pr * HLF / ((t->ssnow_gammzz[0]) + CSWAT * pr);
      t->ssnow_dtmlt[0] = t->ssnow_dtmlt[0] + pr * HLF / ((t->ssnow_gammzz[0]) + CSWAT * pr);
      ssdn[0] = MIN(MAX_SSDN, MAX(120.0, ssdn[0] * osnowd / MAX(0.01, t->ssnow_snowd[0]) +
                                             DENSITY_LIQ * pr / MAX(0.01, t->ssnow_snowd[0])));
      if (notice) ssdn[0] = MIN(450.0, ssdn[0]);
      t->canopy_precis[0] = 0.0;
      t->ssnow_ssdnn[0] = ssdn[0];
    }
  }
  if (t->canopy_precis[0] > 0.0 && isflag > 0) {
    t->ssnow_snowd[0] = MAX(t->ssnow_snowd[0] + psn, 0.0);
    t->canopy_precis[0] = t->canopy_precis[0] - psn;
    double osm = smass[0];
    smass[0] = smass[0] + psn;
    ssdn[0] = MAX(120.0, ssdn[0] * osm / smass[0] + 120.0 * psn / smass[0]);
    sdepth[0] = MAX(0.02, smass[0] / ssdn[0]);
    if (t->canopy_precis[0] > 0.0) {
      double pr = t->canopy_precis[0];
      t->ssnow_snowd[0] = MAX(t->ssnow_snowd[0] + pr, 0.0);
      double sgamm = ssdn[0] * CGSNOW * sdepth[0];
      osm = smass[0];
      tggsn[0] = tggsn[0] + pr * HLF * osm / (sgamm * osnowd);
      t->ssnow_dtmlt[0] = t->ssnow_dtmlt[0] + pr * HLF * osm / (sgamm * osnowd);
      smass[0] = smass[0] + pr * osm / osnowd;
      ssdn[0] = MAX(120.0, MIN(ssdn[0] * osm / smass[0] + DENSITY_LIQ * (1.0 - osm / smass[0]),
                               MAX_SSDN));
      if (notice) ssdn[0] = MIN(450.0, ssdn[0]);
      sdepth[0] = smass[0] / ssdn[0];
      for (int k = 1; k < 3; k++) {
        sgamm = ssdn[k] * CGSNOW * sdepth[k];
        osm = smass[k];
        tggsn[k] = tggsn[k] + pr * HLF * osm / (sgamm * osnowd);
        smass[k] = smass[k] + pr * osm / osnowd;
        ssdn[k] = MAX(120.0, MIN(ssdn[k] * osm / smass[k] + DENSITY_LIQ * (1.0 - osm / smass[k]),
                                 MAX_SSDN));
        if (notice) ssdn[k] = MIN(450.0, ssdn[k]);
        sdepth[k] = smass[k] / ssdn[k];
      }
      t->canopy_precis[0] = 0.0;
    }
  }
  double fsc = t->canopy_fess[0] + t->canopy_fes_cor[0];
  t->canopy_segg[0] = (t->canopy_fess[0] + t->canopy_fes_cor[0]) / HL;
  t->ssnow_evapsn[0] = 0;
  if (t->ssnow_cls[0] == 1.1335) {
    t->ssnow_evapsn[0] = dels * (t->canopy_fess[0] + t->canopy_fes_cor[0]) / (HL + HLF);
    double xxx = t->ssnow_evapsn[0];
    isflag = (int)t->ssnow_isflag[0];
    if (isflag == 0 && fsc > 0.0) t->ssnow_evapsn[0] = MIN(t->ssnow_snowd[0], xxx);
    if (isflag > 0 && fsc > 0.0) t->ssnow_evapsn[0] = MIN(0.9 * smass[0], xxx);
    t->ssnow_snowd[0] = t->ssnow_snowd[0] - t->ssnow_evapsn[0];
    if (isflag > 0) {
      smass[0] = smass[0] - t->ssnow_evapsn[0];
      sdepth[0] = MAX(0.02, smass[0] / ssdn[0]);
    }
    t->canopy_segg[0] = (HL + HLF) * (xxx - t->ssnow_evapsn[0]) / HL / dels;
  }
}

/* src/science/soilsnow/cbl_snowMelt.F90:9-120 */
static double snow_melting(Tile *t, double dels) {
  double snowmlt = 0.0, smelt1[4] = {0.0, 0.0, 0.0, 0.0};
  int isflag = (int)t->ssnow_isflag[0];
  if (t->ssnow_snowd[0] > 0.0 && isflag == 0 && t->ssnow_tgg[0] >= TFRZ) {
    double gz = t->ssnow_gammzz[0];
    double snowflx = (t->ssnow_tgg[0] - TFRZ) * gz;
    snowmlt = MIN(snowflx / HLF, t->ssnow_snowd[0]);
    t->ssnow_dtmlt[0] = t->ssnow_dtmlt[0] + snowmlt * HLF / gz;
    t->ssnow_snowd[0] = t->ssnow_snowd[0] - snowmlt;
    t->ssnow_tgg[0] = t->ssnow_tgg[0] - snowmlt * HLF / gz;
  }
  int three = (t->ssnow_snowd[0] > 0.0 && isflag > 0);
  if (three) {
    double *ssdn = t->ssnow_ssdn, *smass = t->ssnow_smass, *sdepth = t->ssnow_sdepth,
           *tggsn = t->ssnow_tggsn;
    int notice = (int)t->soil_isoilm[0] != 9;
    for (int k = 1; k <= 3; k++) {
      int q = k - 1;
      double sgamm = ssdn[q] * CGSNOW * sdepth[q];
      double snowflx = smelt1[k - 1] * HLF / dels;
      tggsn[q] = tggsn[q] + (snowflx * dels + smelt1[k - 1] * CSWAT * (TFRZ - tggsn[q])) /
                                (sgamm + CSWAT * smelt1[k - 1]);
      double osm = smass[q];
      smass[q] = smass[q] + smelt1[k - 1];
      ssdn[q] = MAX(120.0, MIN(ssdn[q] * osm / smass[q] + DENSITY_LIQ * (1.0 - osm / smass[q]),
                               MAX_SSDN));
      if (notice) ssdn[q] = MIN(450.0, ssdn[q]);
      sdepth[q] = smass[q] / ssdn[q];
      sgamm = smass[q] * CGSNOW;
      smelt1[k - 1] = 0.0;
      smelt1[k] = 0.0;
      if (tggsn[q] > TFRZ) {
        snowflx = (tggsn[q] - TFRZ) * sgamm;
        smelt1[k] = MIN(snowflx / HLF, 0.6 * smass[q]);
        t->ssnow_dtmlt[q] = t->ssnow_dtmlt[q] + smelt1[k] * HLF / sgamm;
        osm = smass[q];
        smass[q] = smass[q] - smelt1[k];
        tggsn[q] = tggsn[q] - smelt1[k] * HLF / sgamm;
        sdepth[q] = smass[q] / ssdn[q];
      }
    }
    snowmlt = smelt1[1] + smelt1[2] + smelt1[3];
    t->ssnow_snowd[0] = t->ssnow_snowd[0] - snowmlt;
  }
  return snowmlt;
}

/* src/science/soilsnow/cbl_snowl_adjust.F90:9-156 */
static void snowl_adjust(Tile *t) {
  if ((int)t->ssnow_isflag[0] <= 0) return;
  double *ssdn = t->ssnow_ssdn, *smass = t->ssnow_smass, *sdepth = t->ssnow_sdepth,
         *tggsn = t->ssnow_tggsn;
  double tw = t->ssnow_t_snwlr[0], snowd = t->ssnow_snowd[0];
  double excd, excm, osm;
  if (sdepth[0] > tw) {
    excd = sdepth[0] - tw;
    excm = excd * ssdn[0];
    sdepth[0] = sdepth[0] - excd;
    osm = smass[0];
    smass[0] = smass[0] - excm;
    osm = smass[1];
    smass[1] = MAX(0.01, smass[1] + excm);
    ssdn[1] = MAX(120.0, MIN(MAX_SSDN, ssdn[1] * osm / smass[1] + ssdn[0] * excm / smass[1]));
    sdepth[1] = smass[1] / ssdn[1];
    tggsn[1] = tggsn[1] * osm / smass[1] + tggsn[0] * excm / smass[1];
    smass[2] = MAX(0.01, snowd - smass[0] - smass[1]);
  } else {
    excd = tw - sdepth[0];
    excm = excd * ssdn[1];
    osm = smass[0];
    smass[0] = smass[0] + excm;
    sdepth[0] = tw;
    ssdn[0] = MAX(120.0, MIN(MAX_SSDN, ssdn[0] * osm / smass[0] + ssdn[1] * excm / smass[0]));
    tggsn[0] = tggsn[0] * osm / smass[0] + tggsn[1] * excm / smass[0];
    smass[1] = MAX(0.01, smass[1] - excm);
    sdepth[1] = smass[1] / ssdn[1];
    smass[2] = MAX(0.01, snowd - smass[0] - smass[1]);
  }
  double frac = smass[1] / MAX(0.02, smass[2]);
  double xfrac = 2.0 / 3.0 / frac;
  if (xfrac > 1.0) {
    excm = (xfrac - 1.0) * smass[1];
    osm = smass[1];
    smass[1] = MAX(0.01, smass[1] + excm);
    tggsn[1] = tggsn[1] * osm / smass[1] + tggsn[2] * excm / smass[1];
    ssdn[1] = MAX(120.0, MIN(MAX_SSDN, ssdn[1] * osm / smass[1] + ssdn[2] * excm / smass[1]));
    smass[2] = MAX(0.01, snowd - smass[0] - smass[1]);
    sdepth[2] = MAX(0.02, smass[2] / ssdn[2]);
  } else {
    excm = (1 - xfrac) * smass[1];
    smass[1] = MAX(0.01, smass[1] - excm);
    sdepth[1] = MAX(0.02, smass[1] / ssdn[1]);
    osm = smass[2];
    smass[2] = MAX(0.01, snowd - smass[0] - smass[1]);
    tggsn[2] = tggsn[2] * osm / smass[2] + tggsn[1] * excm / smass[2];
    ssdn[2] = MAX(120.0, MIN(MAX_SSDN, ssdn[2] * osm / smass[2] + ssdn[1] * excm / smass[2]));
    sdepth[2] = smass[2] / ssdn[2];
  }
  t->ssnow_isflag[0] = 1;
  t->ssnow_ssdnn[0] = (ssdn[0] * sdepth[0] + ssdn[1] * sdepth[1] + ssdn[2] * sdepth[2]) /
                      (sdepth[0] + sdepth[1] + sdepth[2]);
}

/* src/science/soilsnow/cbl_stempv.F90:13-221 with old_soil_conductivity,
   src/science/soilsnow/cbl_Oldconductivity.F90:7-59 */
static void stempv(Tile *t, const Globals *g, double dels) {
  const double *zse = g->zse;
  double at[9], bt[9], ct[9], coeff[10], ccnsw[MS];
  /* row r of the 9-row system is layer r-2 of the reference (-2..6);
     coeff[r] is reference coeff(r-2), r = 0..9 (-2..7) */
  for (int r = 0; r < 9; r++) at[r] = 0.0, bt[r] = 1.0, ct[r] = 0.0;
  for (int r = 0; r < 10; r++) coeff[r] = 0.0;
  double ssat = t->soil_ssat[0];
  for (int k = 0; k < MS; k++) {
    if ((int)t->soil_isoilm[0] == 9) {
      ccnsw[k] = SNOW_CCNSW;
    } else {
      double ew = t->ssnow_wblf[k] * ssat;
      double exp_arg = (ew * log(60.0)) + (t->ssnow_wbfice[k] * ssat * log(250.0));
      double f = MAX(1.0, sqrt(MIN(2.0, 0.5 * ssat / MIN(ew, 0.5 * ssat))));
      if (exp_arg > 30)
        ccnsw[k] = 1.5 * f;
      else
        ccnsw[k] = MIN(t->soil_cnsd[0] * exp(exp_arg), 1.5) * f;
    }
  }
  double xx = 0.;
  int isflag = (int)t->ssnow_isflag[0];
  double coefa = 0.0, coefb = 0.0;
  double *gammzz = t->ssnow_gammzz, *hcll = t->soil_heat_cap_lower_limit;
  double css = t->soil_css[0], rhosoil = t->soil_rhosoil[0];
#define CO(k) coeff[(k) + 2]
#define AT(k) at[(k) + 2]
#define BT(k) bt[(k) + 2]
#define CT(k) ct[(k) + 2]
  if (isflag == 0) {
    xx = MAX(0., t->ssnow_snowd[0] / t->ssnow_ssdnn[0]);
    ccnsw[0] = (ccnsw[0] - 0.2) * (zse[0] / (zse[0] + xx)) + 0.2;
    for (int k = 3; k <= MS; k++)
      CO(k) = 2.0 / (zse[k - 2] / ccnsw[k - 2] + zse[k - 1] / ccnsw[k - 1]);
    CO(2) = 2.0 / ((zse[0] + xx) / ccnsw[0] + zse[1] / ccnsw[1]);
    coefa = 0.0;
    coefb = CO(2);
    for (int k = 1; k <= MS; k++) {
      double wblfsp = t->ssnow_wblf[k - 1];
      gammzz[k - 1] = MAX(hcll[k - 1], (1.0 - ssat) * css * rhosoil +
                                           ssat * (wblfsp * CSWAT * DENSITY_LIQ +
                                                   t->ssnow_wbfice[k - 1] * CSICE * DENSITY_ICE)) *
                      zse[k - 1];
      if (k == 1) gammzz[0] = gammzz[0] + CGSNOW * t->ssnow_snowd[0];
      double dtg = dels / gammzz[k - 1];
      AT(k) = -dtg * CO(k);
      CT(k) = -dtg * CO(k + 1);
      BT(k) = 1.0 - AT(k) - CT(k);
    }
    BT(1) = BT(1) - t->canopy_dgdtg[0] * dels / gammzz[0];
    t->ssnow_tgg[0] = t->ssnow_tgg[0] +
                      (t->canopy_ga[0] - t->ssnow_tgg[0] * t->canopy_dgdtg[0]) * dels / gammzz[0];
  } else {
    double *ssdn = t->ssnow_ssdn, *sd = t->ssnow_sdepth, *sc = t->ssnow_sconds;
    for (int k = 0; k < 3; k++) sc[k] = MAX(0.2, MIN(2.876e-6 * p2(ssdn[k]) + 0.074, MAX_SCONDS));
    CO(-1) = 2.0 / (sd[0] / sc[0] + sd[1] / sc[1]);
    CO(0) = 2.0 / (sd[1] / sc[1] + sd[2] / sc[2]);
    CO(1) = 2.0 / (sd[2] / sc[2] + zse[0] / ccnsw[0]);
    for (int k = 2; k <= MS; k++)
      CO(k) = 2.0 / (zse[k - 2] / ccnsw[k - 2] + zse[k - 1] / ccnsw[k - 1]);
    coefa = CO(-1);
    coefb = CO(1);
    for (int k = 1; k <= 3; k++) {
      double sgamm = ssdn[k - 1] * CGSNOW * sd[k - 1];
      double dtg = dels / sgamm;
      AT(k - 3) = -dtg * CO(k - 3);
      CT(k - 3) = -dtg * CO(k - 2);
      BT(k - 3) = 1.0 - AT(k - 3) - CT(k - 3);
    }
    for (int k = 1; k <= MS; k++) {
      double wblfsp = t->ssnow_wblf[k - 1];
      gammzz[k - 1] = MAX((1.0 - ssat) * css * rhosoil +
                              ssat * (wblfsp * CSWAT * DENSITY_LIQ +
                                      t->ssnow_wbfice[k - 1] * CSICE * DENSITY_ICE),
                          hcll[k - 1]) *
                      zse[k - 1];
      double dtg = dels / gammzz[k - 1];
      AT(k) = -dtg * CO(k);
      CT(k) = -dtg * CO(k + 1);
      BT(k) = 1.0 - AT(k) - CT(k);
    }
    double sgamm = ssdn[0] * CGSNOW * sd[0];
    BT(-2) = BT(-2) - t->canopy_dgdtg[0] * dels / sgamm;
    t->ssnow_tggsn[0] = t->ssnow_tggsn[0] +
                        (t->canopy_ga[0] - t->ssnow_tggsn[0] * t->canopy_dgdtg[0]) * dels / sgamm;
  }
#undef CO
#undef AT
#undef BT
#undef CT
  double rhs[9];
  for (int k = 0; k < 3; k++) rhs[k] = t->ssnow_tggsn[k];
  for (int k = 0; k < MS; k++) rhs[3 + k] = t->ssnow_tgg[k];
  trimb(at, bt, ct, rhs, MS + 3);
  for (int k = 0; k < 3; k++) t->ssnow_tggsn[k] = rhs[k];
  for (int k = 0; k < MS; k++) t->ssnow_tgg[k] = rhs[3 + k];
  t->canopy_sghflux[0] = coefa * (t->ssnow_tggsn[0] - t->ssnow_tggsn[1]);
  t->canopy_ghflux[0] = coefb * (t->ssnow_tgg[0] - t->ssnow_tgg[1]);
}

/* src/science/soilsnow/cbl_remove_trans.F90:9-40 */
static void remove_trans(Tile *t) {
  if (t->canopy_fevc[0] < 0.0) {
    t->canopy_fevw[0] = t->canopy_fevw[0] + t->canopy_fevc[0];
    t->canopy_fevc[0] = 0.0;
  }
  for (int k = 0; k < MS; k++) {
    t->ssnow_wbliq[k] =
        t->ssnow_wbliq[k] - t->ssnow_evapfbl[k] / (t->soil_zse_vec[k] * DENSITY_LIQ);
    t->ssnow_wb[k] = t->ssnow_wbliq[k] + t->ssnow_wbice[k];
  }
}

/* src/science/soilsnow/cbl_soilfreeze.F90:9-81 */
static void soilfreeze(Tile *t, const Globals *g) {
  const double *zse = g->zse;
  double ssat = t->soil_ssat[0], css = t->soil_css[0], rhosoil = t->soil_rhosoil[0];
  for (int k = 0; k < MS; k++) {
    double *wb = &t->ssnow_wb[k], *wbice = &t->ssnow_wbice[k], *tgg = &t->ssnow_tgg[k],
           *gz = &t->ssnow_gammzz[k];
    if (*tgg < TFRZ && FROZEN_LIMIT * *wb - *wbice > .001) {
      double sf = MIN(FROZEN_LIMIT * *wb - *wbice,
                      (ssat - *wb) / MAX((1.0 - DENSITY_ICE / DENSITY_LIQ), 1.0E-3));
      sf = MIN(MAX(0.0, sf) * zse[k] * DENSITY_ICE, (TFRZ - *tgg) * *gz / HLF);
      *wbice = MIN(*wbice + sf / (zse[k] * DENSITY_ICE), FROZEN_LIMIT * *wb);
      *wb = *wb + sf / (zse[k] * DENSITY_ICE) - sf / (zse[k] * DENSITY_LIQ);
      double a1 = t->soil_heat_cap_lower_limit[k];
      double a2 = (1.0 - ssat) * css * rhosoil + (*wb - *wbice) * (CSWAT * DENSITY_LIQ) +
                  *wbice * (CSICE * DENSITY_ICE);
      *gz = MAX(a1, a2) * zse[k];
      if (k == 0 && (int)t->ssnow_isflag[0] == 0) *gz = *gz + CGSNOW * t->ssnow_snowd[0];
      *tgg = *tgg + sf * HLF / *gz;
    } else if (*tgg > TFRZ && *wbice > 0.) {
      double sm = MIN(*wbice * zse[k] * DENSITY_ICE, (*tgg - TFRZ) * *gz / HLF);
      *wbice = MAX(0.0, *wbice - sm / (zse[k] * DENSITY_ICE));
      *wb = *wb - sm / (zse[k] * DENSITY_ICE) + sm / (zse[k] * DENSITY_LIQ);
      double a1 = t->soil_heat_cap_lower_limit[k];
      double a2 = (1.0 - ssat) * css * rhosoil + (*wb - *wbice) * (CSWAT * DENSITY_LIQ) +
                  *wbice * (CSICE * DENSITY_ICE);
      *gz = MAX(a1, a2) * zse[k];
      if (k == 0 && (int)t->ssnow_isflag[0] == 0) *gz = *gz + CGSNOW * t->ssnow_snowd[0];
      *tgg = *tgg - sm * HLF / *gz;
    }
  }
}

/* src/science/soilsnow/cbl_smoisturev.F90:11-444 (nmeth = -1,
   l_new_runoff_speed false) */
static void smoisturev(Tile *t, const Globals *g, double dels) {
  const double *zse = g->zse;
  double at[MS], bt[MS], ct[MS], fluxh[MS + 1], delt[MS + 1], dtt[MS + 1];
  double *wb = t->ssnow_wb, *wbice = t->ssnow_wbice, *wblf = t->ssnow_wblf;
  double ssat = t->soil_ssat[0], hyds = t->soil_hyds[0];
  for (int k = 0; k < MS; k++) at[k] = 0.0, bt[k] = 1.0, ct[k] = 0.0;
  delt[0] = 0.0;
  fluxh[0] = 0.0;
  fluxh[MS] = 0.0;
  for (int k = 1; k <= MS - 1; k++) {
    double wbl_k = MAX(0.01, wb[k - 1] - wbice[k - 1]);
    double wbl_kp = MAX(0.01, wb[k] - wbice[k]);
    delt[k] = wbl_kp - wbl_k;
    double wh = MIN(wbl_k, wbl_kp);
    if (wbice[k - 1] > 0.05 || wbice[k] > 0.01) wh = 0.9 * wbl_k + 0.1 * wbl_kp;
    double speed_k = hyds * pow(wh / ssat, t->soil_i2bp3[0] - 1);
    double rat = delt[k - 1] / (delt[k] + SIGN(1.0e-20, delt[k]));
    double phi = MAX(MAX(0.0, MIN(1.0, 2.0 * rat)), MIN(2.0, rat));
    double fluxhi = wh, fluxlo = wbl_k;
    speed_k = MIN(speed_k, 0.5 * zse[k - 1] / dels);
    fluxh[k] = speed_k * (fluxlo + phi * (fluxhi - fluxlo));
  }
  if (wb[MS - 1] > t->soil_sfc[0]) {
    double wbl_k = MAX(0.001, wb[MS - 1] - wbice[MS - 1]);
    double wbl_kp = MAX(0.001, ssat - wbice[MS - 1]);
    double wh = MIN(wbl_k, wbl_kp);
    if (wbice[MS - 1] > 0.05) wh = 0.9 * wbl_k + 0.1 * wbl_kp;
    double speed_k = hyds * pow(wh / ssat, t->soil_i2bp3[0] - 1);
    speed_k = 0.5 * speed_k / (1. - MIN(0.5, 10. * wbice[MS - 1]));
    double fluxlo = wbl_k;
    speed_k = MIN(0.5 * speed_k, 0.5 * zse[MS - 1] / dels);
    fluxh[MS] = MAX(0.0, speed_k * fluxlo);
  }
  for (int k = MS; k >= 1; k--) {
    fluxh[k - 1] = MIN(fluxh[k - 1], (ssat - wb[k - 1]) * zse[k - 1] / dels + fluxh[k]);
    wb[k - 1] = wb[k - 1] + dels * (fluxh[k - 1] - fluxh[k]) / zse[k - 1];
    double ssatcurr_k = ssat - wbice[k - 1];
    dtt[k] = dels / (zse[k - 1] * ssatcurr_k);
    wblf[k - 1] = (wb[k - 1] - wbice[k - 1]) / ssatcurr_k;
  }
  t->ssnow_rnof2[0] = dels * fluxh[MS] * DENSITY_LIQ;
  for (int k = 2; k <= MS; k++) {
    double wbh_k = (zse[k - 1] * wblf[k - 2] + zse[k - 2] * wblf[k - 1]) / (zse[k - 1] + zse[k - 2]);
    double fact = pow(wbh_k, t->soil_ibp2[0] - 1);
    double pwb_wbh =
        (t->soil_hsbh[0] *
         (1. - MIN(2. * MIN(0.1, MAX(wbice[k - 2] / MAX(0.01, wb[k - 2]),
                                     wbice[k - 1] / MAX(0.01, wb[k - 1]))),
                   0.1))) *
        MAX(t->soil_pwb_min[0], wbh_k * fact);
    double z3_k = pwb_wbh / g->zshh[k - 1];
    at[k - 1] = -dtt[k] * z3_k;
    ct[k - 2] = -dtt[k - 1] * z3_k;
  }
  for (int k = 0; k < MS; k++) bt[k] = 1. - at[k] - ct[k];
  wblf[0] = wblf[0] + dtt[1] * t->ssnow_fwtop1[0] / DENSITY_LIQ;
  wblf[1] = wblf[1] + dtt[2] * t->ssnow_fwtop2[0] / DENSITY_LIQ;
  wblf[2] = wblf[2] + dtt[3] * t->ssnow_fwtop3[0] / DENSITY_LIQ;
  trimb(at, bt, ct, wblf, MS);
  for (int k = 0; k < MS; k++) {
    double ssatcurr = ssat - wbice[k];
    wb[k] = wblf[k] * ssatcurr + wbice[k];
  }
  double dfactor = 1.0 - DENSITY_ICE / DENSITY_LIQ;
  for (int k = 0; k < MS; k++) {
    if (wbice[k] > FROZEN_LIMIT * wb[k]) {
      double sicemelt = (wbice[k] - FROZEN_LIMIT * wb[k]) / (1.0 - FROZEN_LIMIT * dfactor);
      wbice[k] = wbice[k] - sicemelt;
      wb[k] = wb[k] - dfactor * sicemelt;
      t->ssnow_tgg[k] = t->ssnow_tgg[k] -
                        sicemelt * zse[k] * DENSITY_ICE * HLF / (t->ssnow_gammzz[k]);
    }
  }
}

/* src/science/soilsnow/cbl_surfbv.F90:9-146 (nglacier = 2, offline) */
static void surfbv(Tile *t, const Globals *g, double dels) {
  const double *zse = g->zse;
  smoisturev(t, g, dels);
  double ssat = t->soil_ssat[0];
  for (int k = 0; k < MS; k++) {
    double xxx = ssat;
    t->ssnow_rnof1[0] = t->ssnow_rnof1[0] + MAX(t->ssnow_wb[k] - xxx, 0.0) * DENSITY_LIQ * zse[k];
    t->ssnow_wb[k] = MAX(t->soil_swilt[0] / (2. * 2.0), MIN(t->ssnow_wb[k], xxx));
  }
  double rnof5 = 0.;
  double smelt1[4] = {0., 0., 0., 0.};
  double smasstot = 0.0;
  int isflag = (int)t->ssnow_isflag[0];
  double *smass = t->ssnow_smass;
  if (t->ssnow_snowd[0] > MAX_GLACIER_SNOWD) {
    rnof5 = MIN(0.1, t->ssnow_snowd[0] - MAX_GLACIER_SNOWD);
    if (isflag == 0) {
      smasstot = 0.0;
      t->ssnow_tgg[0] = t->ssnow_tgg[0] - rnof5 * HLF / (t->ssnow_gammzz[0]);
      t->ssnow_snowd[0] = t->ssnow_snowd[0] - rnof5;
    } else {
      smasstot = smass[0] + smass[1] + smass[2];
    }
  }
  for (int k = 1; k <= 3; k++) {
    if (t->ssnow_snowd[0] > MAX_GLACIER_SNOWD && isflag > 0) {
      smelt1[k] = MIN(rnof5 * smass[k - 1] / smasstot, 0.2 * smass[k - 1]);
      smass[k - 1] = smass[k - 1] - smelt1[k];
      t->ssnow_snowd[0] = t->ssnow_snowd[0] - smelt1[k];
    }
  }
  if (isflag > 0) rnof5 = smelt1[1] + smelt1[2] + smelt1[3];
  t->ssnow_sinfil[0] = 0.0;
  if ((int)t->veg_iveg[0] == LAKES) {
    double *wbl = &t->ssnow_wb_lake[0], *si = &t->ssnow_sinfil[0];
    *si = MIN(t->ssnow_rnof1[0], *wbl);
    t->ssnow_rnof1[0] = MAX(0.0, t->ssnow_rnof1[0] - *si);
    *wbl = MAX(0.0, *wbl - *si);
    *si = MIN(t->ssnow_rnof2[0], *wbl);
    t->ssnow_rnof2[0] = MAX(0.0, t->ssnow_rnof2[0] - *si);
    *wbl = MAX(0.0, *wbl - *si);
    double xxx = MAX(0.0, (t->ssnow_wb[MS - 1] - t->soil_sfc[0]) * zse[MS - 1] * DENSITY_LIQ);
    *si = MIN(xxx, *wbl);
    t->ssnow_wb[MS - 1] = t->ssnow_wb[MS - 1] - *si / (zse[MS - 1] * DENSITY_LIQ);
    *wbl = MAX(0.0, *wbl - *si);
    xxx = MAX(0.0, (t->ssnow_wb[MS - 1] - 0.5 * (t->soil_sfc[0] + t->soil_swilt[0])) *
                       zse[MS - 1] * DENSITY_LIQ);
    *si = MIN(xxx, *wbl);
    t->ssnow_wb[MS - 1] = t->ssnow_wb[MS - 1] - *si / (zse[MS - 1] * DENSITY_LIQ);
    *wbl = MAX(0.0, *wbl - *si);
  }
  t->ssnow_rnof1[0] = t->ssnow_rnof1[0] / dels + rnof5 / dels;
  t->ssnow_rnof2[0] = t->ssnow_rnof2[0] / dels;
  t->ssnow_runoff[0] = t->ssnow_rnof1[0] + t->ssnow_rnof2[0];
}

/* src/science/soilsnow/cbl_soilsnow_main.F90:28-207 */
static void soil_snow(Tile *t, const Globals *g, double dels, int first_call) {
  const double *zse = g->zse;
  double zsetot = zse[0];
  for (int k = 1; k < MS; k++) zsetot = zsetot + zse[k];
  t->ssnow_tggav[0] = 0.;
  for (int k = 0; k < MS; k++) {
    t->ssnow_tggav[0] = t->ssnow_tggav[0] + ((zse[k] / zsetot) * t->ssnow_tgg[k]);
    t->soil_heat_cap_lower_limit[k] = MAX(0.01, t->soil_css[0] * t->soil_rhosoil[0]);
  }
  t->ssnow_t_snwlr[0] = 0.05;
  t->ssnow_fwtop1[0] = t->ssnow_fwtop2[0] = t->ssnow_fwtop3[0] = 0.0;
  t->ssnow_runoff[0] = 0.0;
  t->ssnow_rnof1[0] = 0.0;
  t->ssnow_rnof2[0] = 0.0;
  t->ssnow_smelt[0] = 0.0;
  for (int k = 0; k < 3; k++) t->ssnow_dtmlt[k] = 0.0;
  t->ssnow_osnowd[0] = t->ssnow_snowd[0];
  for (int k = 0; k < MS; k++) t->ssnow_wbliq[k] = t->ssnow_wb[k] - t->ssnow_wbice[k];
  double ssat = t->soil_ssat[0], css = t->soil_css[0], rhosoil = t->soil_rhosoil[0];
  double xx = css * rhosoil;
  if (first_call)
    t->ssnow_gammzz[0] =
        MAX((1.0 - ssat) * css * rhosoil +
                (t->ssnow_wb[0] - t->ssnow_wbice[0]) * CSWAT * DENSITY_LIQ +
                t->ssnow_wbice[0] * CSICE * DENSITY_ICE,
            xx) * zse[0] +
        (1. - t->ssnow_isflag[0]) * CGSNOW * t->ssnow_snowd[0];
  for (int k = 0; k < MS; k++) {
    t->ssnow_wblf[k] = MAX(0.01, (t->ssnow_wb[k] - t->ssnow_wbice[k])) / ssat;
    t->ssnow_wbfice[k] = t->ssnow_wbice[k] / ssat;
  }
  snowcheck(t);
  snowdensity(t, dels);
  snow_accum(t, dels);
  double snowmlt = snow_melting(t, dels);
  t->ssnow_smelt[0] = snowmlt;
  snowl_adjust(t);
  stempv(t, g, dels);
  int isflag = (int)t->ssnow_isflag[0];
  t->ssnow_tss[0] = (double)(1 - isflag) * t->ssnow_tgg[0] + (double)isflag * t->ssnow_tggsn[0];
  snowmlt = snow_melting(t, dels);
  t->ssnow_smelt[0] = t->ssnow_smelt[0] + snowmlt;
  remove_trans(t);
  soilfreeze(t, g);
  double totwet = t->canopy_precis[0] + t->ssnow_smelt[0];
  double weting = totwet + MAX(0.0, t->ssnow_pudsto[0] - t->canopy_fesp[0] / HL * dels);
  double xxx = ssat - t->ssnow_wb[0];
  double sinfil1 = MIN(0.95 * xxx * zse[0] * DENSITY_LIQ, weting);
  xxx = ssat - t->ssnow_wb[1];
  double sinfil2 = MIN(0.95 * xxx * zse[1] * DENSITY_LIQ, weting - sinfil1);
  xxx = ssat - t->ssnow_wb[2];
  double sinfil3 = MIN(0.95 * xxx * zse[2] * DENSITY_LIQ, weting - sinfil1 - sinfil2);
  t->ssnow_fwtop1[0] = sinfil1 / dels - t->canopy_segg[0];
  t->ssnow_fwtop2[0] = sinfil2 / dels;
  t->ssnow_fwtop3[0] = sinfil3 / dels;
  t->ssnow_pudsto[0] = MAX(0.0, weting - sinfil1 - sinfil2 - sinfil3);
  t->ssnow_rnof1[0] = MAX(0., t->ssnow_pudsto[0] - t->ssnow_pudsmx[0]);
  t->ssnow_pudsto[0] = t->ssnow_pudsto[0] - t->ssnow_rnof1[0];
  surfbv(t, g, dels);
  t->ssnow_smelt[0] = t->ssnow_smelt[0] / dels;
  isflag = (int)t->ssnow_isflag[0];
  t->ssnow_tss[0] = (double)(1 - isflag) * t->ssnow_tgg[0] + (double)isflag * t->ssnow_tggsn[0];
  for (int k = 0; k < MS; k++) t->ssnow_wbliq[k] = t->ssnow_wb[k] - t->ssnow_wbice[k];
  t->ssnow_totsdepth[0] = t->ssnow_sdepth[0] + t->ssnow_sdepth[1] + t->ssnow_sdepth[2];
  t->ssnow_wbtot[0] = 0.0;
  for (int k = 0; k < MS; k++)
    t->ssnow_wbtot[0] = t->ssnow_wbtot[0] + (t->ssnow_wbliq[k] * DENSITY_LIQ +
                                             t->ssnow_wbice[k] * DENSITY_ICE) * zse[k];
}

/* src/science/soilsnow/cbl_snow_aging.F90:11-83 */
static void snow_aging(Tile *t, double dels) {
  double snowd = t->ssnow_snowd[0];
  if (snowd > 1.0) {
    double dnsnow = MIN(1.0, 0.1 * MAX(0.0, snowd - t->ssnow_osnowd[0]));
    int fl = (int)t->ssnow_isflag[0];
    double tmp = fl * t->ssnow_tggsn[0] + (1 - fl) * t->ssnow_tgg[0];
    tmp = MIN(tmp, TFRZ);
    double ar1 = 5000.0 * (1.0 / (TFRZ - 0.01) - 1.0 / tmp);
    double ar2 = 10.0 * ar1;
    double ar3;
    if ((int)t->soil_isoilm[0] == 9) {
      ar3 = 0.0000001;
      dnsnow = 1.0;
    } else {
      ar3 = 0.1;
    }
    double dtau = 1.0e-6 * (exp(ar1) + exp(ar2) + ar3) * dels;
    t->ssnow_snage[0] = MAX(0.0, (t->ssnow_snage[0] + dtau) * (1.0 - dnsnow));
  }
}

/* src/science/misc/cable_carbon.F90:319-360, 220-314, 38-216 (mvtype 17,
   DIAG_SOIL_RESP 'ON') */
static const double RW17[17] = {16., 16., 18., 8.7, 12.5, 15.1, 10.4, 7.5, 6.1,
                                6.1, 0.001, 5.8, 0.001, 5.8, 0.001, 9.0, 0.001};
static const double TFCL17[17] = {0.42, 0.248, 0.38, 0.345, 0.31, 0.73, 2.4, 0.35, 0.997,
                                  0.997, 0.9500, 0.55, 0.9500, 0.55, 0.9500, 2.4, 0.9500};
static const double TVCLST17[17] = {235., 283., 268., 278., 278., 235., 278.0, 278.0,
                                    278.0, 278.0, 278.0, 278., 278., 278., 268., 278., 278.};

static void carbon(Tile *t, const Globals *g, double dels) {
  const double *rc = g->ratecp, *rs = g->ratecs;
  double *cp = t->bgc_cplant, *cs = t->bgc_csoil;
  /* plantcarb */
  double S = rc[0] * cp[0] + rc[1] * cp[1] + rc[2] * cp[2];
  double pc1 = (S - rc[0] * cp[0]);
  double pc1w = (S - rc[0] * cp[0] - rc[2] * cp[2]);
  double pc1r = (S - rc[0] * cp[0] - rc[1] * cp[1]);
  double tk = t->met_tk[0];
  double tmp1 = MAX(3.22 - 0.046 * (tk - TFRZ), 1e-6);
  double tmp2 = 0.1 * (tk - TFRZ - 20.0);
  double tmp3 = pow(tmp1, tmp2);
  const double spy = 365.0 * 24.0 * 3600.0;
  t->canopy_frp[0] = t->veg_rp20[0] * tmp3 * pc1 / spy;
  t->canopy_frpw[0] = t->veg_rp20[0] * tmp3 * pc1w / spy;
  t->canopy_frpr[0] = t->veg_rp20[0] * tmp3 * pc1r / spy;
  (void)rs;
  /* soilcarb, DIAG_SOIL_RESP on */
  {
    const double t0 = -46.0, rswch = 0.16, soilcf = 1.0;
    double swilt = t->soil_swilt[0];
    double den = MAX(0.07, t->soil_sfc[0] - swilt);
    double rswc = MAX(0.0001, t->veg_froot[0] * (t->ssnow_wb[1] - swilt)) / den;
    double tsoil = t->veg_froot[0] * t->ssnow_tgg[1] - TFRZ;
    double tref = MAX(0., t->ssnow_tgg[MS - 1] - (TFRZ - .05));
    for (int k = 1; k < MS; k++) {
      rswc = rswc + MAX(0.0001, t->veg_froot[k] * (t->ssnow_wb[k] - swilt)) / den;
      tsoil = tsoil + t->veg_froot[k] * t->ssnow_tgg[k];
    }
    rswc = MIN(1., rswc);
    tsoil = MAX(t0 + 2., tsoil);
    double e0rswc = 52.4 + 285. * rswc;
    double ftsoil = MIN(0.0015, 1. / (tref - t0) - 1. / (tsoil - t0));
    double sss = MAX(-15., MIN(1., e0rswc * ftsoil));
    double ftsrs = exp(sss);
    t->canopy_frs[0] = t->veg_vegcf[0] * (144.0 / 44.0e6) * soilcf *
                       MIN(1., 1.4 * MAX(.3, .0278 * tsoil + .5)) * ftsrs * rswc /
                       (rswch + rswc);
  }
  /* carbon_pl */
  {
    const double beta = 0.9, trnl = 3.17e-8, trnr = 4.53e-9, trnsf = 1.057e-10,
                 trnw = 6.342e-10;
    int iv = (int)t->veg_iveg[0] - 1;
    double coef_cold = exp(MIN(1., -(t->canopy_tv[0] - TVCLST17[iv])));
    double wbav = t->veg_froot[0] * t->ssnow_wb[0];
    for (int k = 1; k < MS; k++) wbav = wbav + t->veg_froot[k] * t->ssnow_wb[k];
    wbav = MAX(0.01, wbav);
    double ce = 2.0 - t->soil_ibp2[0];
    double esw = pow(wbav, ce) - 1.0;
    esw = MAX(1.0, esw);
    double eswilt = pow(t->soil_swilt[0], ce) - 1.0;
    double rel = esw / eswilt - 1.0;
    rel = MIN(1.0, rel);
    double coef_drght = exp(5.0 * rel);
    double coef_cd = (coef_cold + coef_drght) * 2.0e-7;
    double fcl = exp(-TFCL17[iv] * t->veg_vlai[0]);
    double clitt = (coef_cd + trnl) * cp[0];
    double fpn = t->canopy_fpn[0];
    cp[0] = cp[0] - dels * (fpn * fcl + clitt);
    double fr = MIN(1., exp(-RW17[iv] * beta * 0.0001 * cp[2] / MAX(cp[1], 0.01)) / beta);
    double cfwd = trnw * cp[1];
    cp[1] = cp[1] - dels * (fpn * (1. - fcl) * (1. - fr) + t->canopy_frpw[0] + cfwd);
    double cfrts = trnr * cp[2];
    cp[2] = cp[2] - dels * (fpn * (1. - fcl) * fr + cfrts + t->canopy_frpr[0]);
    double cfsf = trnsf * cs[0];
    double frs = t->canopy_frs[0];
    cs[0] = cs[0] + dels * (0.98 * clitt + 0.9 * cfrts + cfwd - cfsf - 0.98 * frs);
    cs[1] = cs[1] + dels * (0.02 * clitt + 0.1 * cfrts + cfsf - 0.02 * frs);
    for (int k = 0; k < 3; k++) cp[k] = MAX(0.00, cp[k]);
    for (int k = 0; k < 2; k++) cs[k] = MAX(0.00, cs[k]);
  }
}

/* ===================================================================== */
/* One timestep of one tile: src/offline/cbl_model_driver_offline.F90:38-231
   plus the offline driver's contract around it (src/offline/cable_serial.F90
   :572-575 before, :602-605 and :725-729 after). */
static void cbm_tile(Tile *t, const Globals *g, int ktau, double dels, int first_call,
                     int *npass) {
  /* driver, before cbm */
  t->met_ofsd[0] = t->met_fsd[0] + t->met_fsd[1];
  t->canopy_oldcansto[0] = t->canopy_cansto[0];
  if ((int)t->veg_iveg[0] >= 14) t->veg_vlai[0] = 0.;
  /* lake water fix-up, :116-121 */
  if ((int)t->veg_iveg[0] == LAKES && t->ssnow_wb[0] < t->soil_sfc[0]) {
    t->ssnow_wbtot1[0] = (t->ssnow_wb[0]) * DENSITY_LIQ * g->zse[0];
    t->ssnow_wb[0] = t->soil_sfc[0];
    t->ssnow_wbtot2[0] = (t->ssnow_wb[0]) * DENSITY_LIQ * g->zse[0];
  }
  t->ssnow_wb_lake[0] = t->ssnow_wb_lake[0] + MAX(t->ssnow_wbtot2[0] - t->ssnow_wbtot1[0], 0.);
  ruff_resist(t);
  define_air(t);
  int veg_mask = t->canopy_vlaiw[0] > LAI_THRESH;
  int sunlit = (t->met_fsd[0] + t->met_fsd[1]) > RAD_THRESH;
  int sunlit_veg = veg_mask && sunlit;
  init_radiation_albedo(t, veg_mask);
  t->rad_albedo_t[0] = (t->rad_albedo[0] + t->rad_albedo[1]) * 0.5;
  t->ssnow_otss_0[0] = t->ssnow_otss[0];
  t->ssnow_otss[0] = t->ssnow_tss[0];
  define_canopy(t, g, dels, sunlit_veg, npass);
  t->ssnow_owetfac[0] = t->ssnow_wetfac[0];
  soil_snow(t, g, dels, first_call);
  snow_aging(t, dels);
  t->ssnow_deltss[0] = t->ssnow_tss[0] - t->ssnow_otss[0];
  t->canopy_fev[0] = t->canopy_fevc[0] + t->canopy_fevw[0];
  t->canopy_fe[0] = t->canopy_fev[0] + t->canopy_fes[0];
  t->canopy_rnet[0] = t->canopy_fns[0] + t->canopy_fnv[0];
  double transd = t->rad_transd[0];
  t->rad_trad[0] = pow((1. - transd) * p4(t->canopy_tv[0]) + transd * p4(t->ssnow_tss[0]), 0.25);
  carbon(t, g, dels);
  double fpn = t->canopy_fpn[0];
  t->canopy_fnpp[0] = -1.0 * fpn - t->canopy_frp[0];
  t->canopy_fgpp[0] = -1.0 * fpn + t->canopy_frday[0];
  t->canopy_fnee[0] = fpn + t->canopy_frs[0] + t->canopy_frp[0];
  t->canopy_fra[0] = t->canopy_frp[0] + t->canopy_frday[0];
  /* driver, after cbm */
  t->ssnow_smelt[0] = t->ssnow_smelt[0] * dels;
  t->ssnow_rnof1[0] = t->ssnow_rnof1[0] * dels;
  t->ssnow_rnof2[0] = t->ssnow_rnof2[0] * dels;
  t->ssnow_runoff[0] = t->ssnow_runoff[0] * dels;
  /* sumcflux, src/offline/casa_cable.F90:482-560 (icycle = 0, kstart = 1),
     called at src/offline/cable_serial.F90:713-715: running sums of the
     carbon fluxes times the step length; it sets fnee to cbm's value again */
  {
    double pn = t->canopy_fpn[0] * dels, rd = t->canopy_frday[0] * dels;
    double rpw = t->canopy_frpw[0] * dels, rpr = t->canopy_frpr[0] * dels;
    double rp = t->canopy_frp[0] * dels, rs = t->canopy_frs[0] * dels;
    if (ktau == 1) {
      t->sum_flux_sumpn[0] = pn;
      t->sum_flux_sumrd[0] = rd;
      t->sum_flux_dsumpn[0] = pn;
      t->sum_flux_dsumrd[0] = rd;
      t->sum_flux_sumrpw[0] = rpw;
      t->sum_flux_sumrpr[0] = rpr;
      t->sum_flux_sumrp[0] = rp;
      t->sum_flux_dsumrp[0] = rp;
      t->sum_flux_sumrs[0] = rs;
    } else {
      t->sum_flux_sumpn[0] = t->sum_flux_sumpn[0] + pn;
      t->sum_flux_sumrd[0] = t->sum_flux_sumrd[0] + rd;
      t->sum_flux_dsumpn[0] = t->sum_flux_dsumpn[0] + pn;
      t->sum_flux_dsumrd[0] = t->sum_flux_dsumrd[0] + rd;
      t->sum_flux_sumrpw[0] = t->sum_flux_sumrpw[0] + rpw;
      t->sum_flux_sumrpr[0] = t->sum_flux_sumrpr[0] + rpr;
      t->sum_flux_sumrp[0] = t->sum_flux_sumrp[0] + rp;
      t->sum_flux_dsumrp[0] = t->sum_flux_dsumrp[0] + rp;
      t->sum_flux_sumrs[0] = t->sum_flux_sumrs[0] + rs;
    }
  }
  /* mass_balance, src/offline/cable_checks.F90:472-551 */
  {
    if (ktau == 1) t->drv_owb[0] = t->ssnow_wbtot[0];
    double delwb = t->ssnow_wbtot[0] - t->drv_owb[0];
    t->drv_owb[0] = t->ssnow_wbtot[0];
    double rlam = t->air_rlam[0], cls = t->ssnow_cls[0];
    t->bal_wbal[0] = t->met_precip[0] - t->canopy_delwc[0] - t->ssnow_snowd[0] +
                     t->ssnow_osnowd[0] - t->ssnow_runoff[0] -
                     (t->canopy_fevw[0] + t->canopy_fevc[0] + t->canopy_fes[0] / cls) * dels / rlam -
                     delwb - t->ssnow_qrecharge[0];
    if (ktau == 1) {
      t->bal_wbal_tot[0] = 0.;
      t->bal_precip_tot[0] = 0.;
      t->bal_rnoff_tot[0] = 0.;
      t->bal_evap_tot[0] = 0.;
    }
    if (ktau > 10) {
      t->bal_wbal_tot[0] = t->bal_wbal_tot[0] + t->bal_wbal[0];
      t->bal_precip_tot[0] = t->bal_precip_tot[0] + t->met_precip[0];
      t->bal_rnoff_tot[0] = t->bal_rnoff_tot[0] + t->ssnow_rnof1[0] + t->ssnow_rnof2[0];
      t->bal_evap_tot[0] =
          t->bal_evap_tot[0] + (t->canopy_fev[0] + t->canopy_fes[0] / cls) * dels / rlam;
    }
  }
  /* energy_balance, src/offline/cable_checks.F90:565-618 */
  {
    double *fsd = t->met_fsd, *alb = t->rad_albedo, *qc = t->rad_qcan;
    double tv4 = p4(t->canopy_tv[0]);
    t->bal_radbal[0] = fsd[0] + fsd[1] + t->met_fld[0] - alb[0] * fsd[0] - alb[1] * fsd[1] -
                       (EMSOIL * SBOLTZ * transd * p4(t->ssnow_otss[0])) -
                       (EMLEAF * SBOLTZ * (1 - transd) * tv4) - t->canopy_fnv[0] -
                       t->canopy_fns[0];
    t->bal_ebalsoil[0] = t->canopy_fns[0] - t->canopy_fes[0] - t->canopy_fhs[0] - t->canopy_ga[0];
    t->bal_ebalveg[0] = t->canopy_fnv[0] - t->canopy_fev[0] - t->canopy_fhv[0];
    t->bal_ebal[0] = (qc[0] + qc[1]) + (qc[2] + qc[3]) + t->rad_qssabs[0] + t->met_fld[0] -
                     SBOLTZ * EMLEAF * tv4 * (1 - transd) - t->rad_flws[0] * transd -
                     t->canopy_fev[0] - t->canopy_fes[0] - t->canopy_fh[0] - t->canopy_ga[0];
    t->bal_ebal_tot[0] = t->bal_ebal_tot[0] + t->bal_ebal[0];
    t->bal_radbalsum[0] = t->bal_radbalsum[0] + t->bal_radbal[0];
  }
}

/* ===================================================================== */
/* C interface for the tests: fields[id] points at the (n, mp) array of
   field id (double), globals at zse(6), zshh(7), ratecp(3), ratecs(2).
   Returns the number of dryLeaf passes performed, summed over tiles. */
int64_t oracle_cbm_step(int64_t mp, int ktau, double dels, int first_call,
                        double *const *fields, const double *globals) {
  Globals g;
  memcpy(g.zse, globals, sizeof(g.zse));
  memcpy(g.zshh, globals + 6, sizeof(g.zshh));
  memcpy(g.ratecp, globals + 13, sizeof(g.ratecp));
  memcpy(g.ratecs, globals + 16, sizeof(g.ratecs));
  static const int nel[] = {
#define X(id, name, prec, n, role) n,
      CABLE_FIELDS(X)
#undef X
  };
  int64_t passes = 0;
  Tile t;
  for (int64_t i = 0; i < mp; i++) {
    double *dst = (double *)&t;
    for (int f = 0; f < CABLE_NFIELDS; f++) {
      const double *src = fields[f];
      for (int k = 0; k < nel[f]; k++) dst[k] = src[(int64_t)k * mp + i];
      dst += nel[f];
    }
    int np = 0;
    cbm_tile(&t, &g, ktau, dels, first_call, &np);
    passes += np;
    const double *s = (const double *)&t;
    for (int f = 0; f < CABLE_NFIELDS; f++) {
      double *d = fields[f];
      for (int k = 0; k < nel[f]; k++) d[(int64_t)k * mp + i] = s[k];
      s += nel[f];
    }
  }
  return passes;
}

int oracle_nfields(void) { return CABLE_NFIELDS; }

/* test hook: the Thomas solve of a single kmax-row system (kmax <= 9), in
   place on rhs */
void oracle_trimb(const double *a, const double *b, const double *c, double *rhs, int kmax) {
  trimb(a, b, c, rhs, kmax);
}
