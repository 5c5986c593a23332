// cbm_physics.h -- one CABLE timestep (cbm) for one tile, as device code.
//
// One GPU thread owns one tile for the whole step: every intermediate of the
// radiation, roughness, canopy energy-balance iteration and soil/snow update
// lives in that thread's registers, and HBM is touched only to read the tile's
// state, parameters and forcing and to write its state and outputs.
//
// The code is a template on
//   R     the working precision: double for the plain FP64 path, float for
//         the mixed path.  Default-REAL quantities, the whole canopy iteration
//         (its REAL(r_2) quantities included) and the soil's elementary
//         functions are computed in R.  REAL(r_2) state (soil moisture, ice,
//         heat capacity) and its updates, the two tridiagonal (Thomas) systems
//         and the balance accumulators are double in both paths.
//   FULL  write every diagnostic of the reference's derived types (the drop-in
//         mode used by the Fortran cbm() replacement and the parity tests), or
//         only the prognostic state and the eight headline fluxes.
//   TABLES  read the vegetation and soil parameters from per-class tables
//         staged in LDS, indexed by the tile's vegetation and soil type, or
//         from per-tile arrays.
//
// The physics follows the reference routine by routine (file:line cited at
// each function); where the reference's order of floating-point operations
// matters for FP64 parity it is kept, and powers are evaluated as the
// reference compiler lowers them (integer exponents as multiplication chains,
// x**0.5 as sqrt, 2.0**y as exp2).  Build with -ffp-contract=off.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

#include "../../include/cable_fields.h"

// device code; the same functions also compile for the host so the test
// suite can run the kernel's physics on the CPU.  Everything is inlined into
// the step kernel: the tile's working set only stays in registers when no
// routine is left as a call (with several kernel instantiations in one
// translation unit the compiler would otherwise outline shared routines and
// pass the working set through scratch memory)
#ifndef CBM_HD
#define CBM_HD __host__ __device__ __forceinline__
#endif

namespace cbm {

// float arguments take the float overloads, on the host and on the device
using std::atan;
using std::copysign;
using std::cos;
using std::exp;
using std::exp2;
using std::fabs;
using std::log;
using std::pow;
using std::sqrt;

constexpr int MS = 6;

// ---- class tables (LDS) ---------------------------------------------------
// Column of each vegetation / soil parameter in the per-class tables.  A
// parameter held per tile (iveg, isoilm, vlai, albsoil) has no column.
enum VegCol {
  VC_canst1, VC_dleaf, VC_ejmax, VC_frac4, VC_hc, VC_rp20, VC_shelrb, VC_vegcf,
  VC_vbeta, VC_vcmax, VC_xfang, VC_extkn, VC_a1gs, VC_d0gs, VC_alpha, VC_convex,
  VC_cfrd, VC_gswmin, VC_conkc0, VC_conko0, VC_ekc, VC_eko, VC_meth,
  VC_refl, VC_taul = VC_refl + 2, VC_froot = VC_taul + 2,
  VC_za_uv = VC_froot + 6, VC_za_tq, VC_fes_cor, VC_qrecharge, VC_pudsmx,
  VC_N
};
enum SoilCol {
  SC_css, SC_hsbh, SC_hyds, SC_i2bp3, SC_ibp2, SC_rhosoil, SC_sfc, SC_ssat,
  SC_swilt, SC_cnsd, SC_pwb_min, SC_sfc_vec, SC_swilt_vec = SC_sfc_vec + 6,
  SC_zse_vec = SC_swilt_vec + 6, SC_N = SC_zse_vec + 6
};
constexpr int NVEGCLASS = 17, NSOILCLASS = 9;

constexpr int veg_col(int id) {
  switch (id) {
#define C_(f) case CABLE_F_veg_##f: return VC_##f;
    C_(canst1) C_(dleaf) C_(ejmax) C_(frac4) C_(hc) C_(rp20) C_(shelrb) C_(vegcf)
    C_(vbeta) C_(vcmax) C_(xfang) C_(extkn) C_(a1gs) C_(d0gs) C_(alpha) C_(convex)
    C_(cfrd) C_(gswmin) C_(conkc0) C_(conko0) C_(ekc) C_(eko) C_(meth) C_(refl)
    C_(taul) C_(froot)
#undef C_
    case CABLE_F_rough_za_uv: return VC_za_uv;
    case CABLE_F_rough_za_tq: return VC_za_tq;
    case CABLE_F_canopy_fes_cor: return VC_fes_cor;
    case CABLE_F_ssnow_qrecharge: return VC_qrecharge;
    case CABLE_F_ssnow_pudsmx: return VC_pudsmx;
    default: return -1;
  }
}
constexpr int soil_col(int id) {
  switch (id) {
#define C_(f) case CABLE_F_soil_##f: return SC_##f;
    C_(css) C_(hsbh) C_(hyds) C_(i2bp3) C_(ibp2) C_(rhosoil) C_(sfc) C_(ssat)
    C_(swilt) C_(cnsd) C_(pwb_min) C_(sfc_vec) C_(swilt_vec) C_(zse_vec)
#undef C_
    default: return -1;
  }
}

// ---- per-tile working set -------------------------------------------------
#define CBM_TY_F R
#define CBM_TY_D double
#define CBM_TY_I int
#define CBM_TY_B double
template <class R>
struct Tile {
#define X(id, name, prec, n, role) CBM_TY_##prec id[n];
  CABLE_FIELDS(X)
#undef X
};

// what a step does besides the physics
enum StepMode {
  MODE_PLAIN = 0,    // the reference's method
  MODE_CAPTURE = 1,  // the reference's method, and write surrogate training samples
  MODE_HYBRID = 2,   // surrogate warm start, capped polish passes checked by
                     // the leaf energy residual, the reference's iteration
                     // where the check rejects them
  MODE_UNLIT = 3     // a hybrid step in which no leaf-bearing tile is in
                     // daylight: the reference's method, and the leaf state
                     // the next step's surrogate features read
};

// globals of one step
struct StepArgs {
  void *const *F;        // device pointer table: F[field id] -> (n, mp) array
  int64_t mp;
  int64_t foff;          // element offset of this step in the forcing arrays
  int ktau;
  int first_call;        // first soil_snow call of the run (gammzz(:,1) init)
  int driver;            // 1: the offline driver's step (its contract around
                         // cbm and the balance checks); 0: cbm alone
  double dels;
  double zse[MS], zshh[MS + 1], ratecp[3], ratecs[2];
  const double *vtab;    // NVEGCLASS x VC_N   (TABLES)
  const double *stab;    // NSOILCLASS x SC_N  (TABLES)
  int compact_forcing;   // FAST mode: fsd2 = fsd1, tvrad = tvair = tk,
                         // precip_sn from tk, ca and doy scalars
  double ca, doy;
  unsigned long long *npass;  // optional: sum of dryLeaf passes (work counter)
  double *cons_part;     // optional: per-workgroup sums of bal%wbal, bal%ebal,
                         // canopy%fpn, canopy%fe ([gridDim.x][4]) for the
                         // driver's consistency check
  // ---- surrogate (mixed path only) ----
  float *hyb;                // [3][mp]: previous step's leaf and canopy-air
                             // temperature, canopy-air humidity (or null)
  int hyb_valid;             // 0: hyb not yet written (first step)
  const float *norm;         // feature/target normalisation (64 values)
  const float *pred;         // [mp][8] surrogate predictions (MODE_HYBRID)
  float *sx, *sy;            // [cap][24], [cap][8] training samples (MODE_CAPTURE)
  unsigned long long *scount;  // samples written so far
  unsigned long long scap;     // capacity of sx/sy
  int max_passes;            // MODE_HYBRID: dryLeaf passes per NITER pass
  float resid_max;           // MODE_HYBRID: leaf energy residual limit, W/m2
  unsigned long long *nfallback;  // MODE_HYBRID: tiles that fell back (counter)
  const unsigned long long *lit_blocks;  // optional: per block of 256 tiles, bit
                             // `lit_bit` set where a leaf-bearing tile is in
                             // daylight at this step (the inference's vote)
  int lit_bit;
};

// literal of the working precision
#define K(x) R(x)

template <class R>
struct Consts {
  static constexpr R TFRZ = R(273.16), SBOLTZ = R(5.67e-8), EMSOIL = R(1.0),
                     EMLEAF = R(1.0), CAPP = R(1004.64), HL = R(2.5014e6),
                     HLF = R(0.334e6), DHEAT = R(21.5e-6), GRAV = R(9.8086),
                     RGAS = R(8.3143), RMAIR = R(0.02897), RMH2O = R(0.018016),
                     CGSNOW = R(2090.0), CSICE = R(2.100e3), CSWAT = R(4.218e3),
                     DENSITY_LIQ = R(1000.0), DENSITY_ICE = R(921.0),
                     TETENA = R(6.106), TETENB = R(17.27), TETENC = R(237.3),
                     VONK = R(0.40), A33 = R(1.25), CSW = R(0.50), CTL = R(0.40),
                     APOL = R(0.70), PRANDT = R(0.71), CRD = R(0.3), CSD = R(0.003),
                     CCD = R(15.0), CCW_C = R(2.0), USUHM = R(0.3), ZETA0 = R(0.0),
                     ZETNEG = R(-15.0), ZETPOS = R(1.0), ZDLIN = R(1.0), UMIN = R(0.1),
                     PI_C = R(3.1415927), LAI_THRESH = R(0.001), RAD_THRESH = R(0.001),
                     COSZEN_TOLS = R(1.0e-4), GAM0 = R(28.0e-6), GAM1 = R(0.0509),
                     GAM2 = R(0.0010), RGBWC = R(1.32), RGSWC = R(1.57),
                     TREFK = R(298.2), MAX_GLACIER_SNOWD = R(1100.0),
                     SNOW_CCNSW = R(2.0), MAX_SSDN = R(750.0), MAX_SCONDS = R(2.51),
                     FROZEN_LIMIT = R(0.85), SNMIN = R(1.0);
};

template <class T> CBM_HD inline T MAX(T a, T b) { return a > b ? a : b; }
template <class T> CBM_HD inline T MIN(T a, T b) { return a < b ? a : b; }
CBM_HD inline double MAX(double a, float b) { return a > (double)b ? a : (double)b; }
CBM_HD inline double MAX(float a, double b) { return (double)a > b ? (double)a : b; }
CBM_HD inline double MIN(double a, float b) { return a < (double)b ? a : (double)b; }
CBM_HD inline double MIN(float a, double b) { return (double)a < b ? (double)a : b; }
template <class T> CBM_HD inline T SIGN(T a, T b) { return copysign(fabs(a), b); }
template <class T> CBM_HD inline T p2(T x) { return x * x; }
template <class T> CBM_HD inline T p3(T x) { return x * (x * x); }
template <class T> CBM_HD inline T p4(T x) { return x * (x * (x * x)); }
// x**0.25 and x**1.5 for x >= 0 through the square root, which the device
// evaluates in hardware: within about an ulp of pow (two correctly rounded
// steps), at a small part of the cost of the FP64 pow sequence
template <class T> CBM_HD inline T root4(T x) { return sqrt(sqrt(x)); }
template <class T> CBM_HD inline T pow1p5(T x) { return x * sqrt(x); }

// parameter access: class table row (TABLES) or the tile's own copy; the
// table holds FP64 values, converted to the field's working type on use
#define PV(id, k)                                                              \
  ((TABLES && veg_col(CABLE_F_##id) >= 0)                                      \
       ? (decltype(+t.id[0]))(vrow[veg_col(CABLE_F_##id) + (k)])              \
   : (TABLES && soil_col(CABLE_F_##id) >= 0)                                   \
       ? (decltype(+t.id[0]))(srow[soil_col(CABLE_F_##id) + (k)])             \
       : t.id[k])

}  // namespace cbm
