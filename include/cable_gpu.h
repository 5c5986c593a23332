/*
 * cable_gpu.h -- C ABI of the MI355X CABLE timestep library (libcable_gpu.so).
 *
 * The library replaces one call of the reference's per-tile physics,
 *   SUBROUTINE cbm(ktau, dels, air, bgc, canopy, met, bal, rad, rough, soil,
 *                  ssnow, sum_flux, veg, climate, xk, c1, rhoch)
 *   (src/offline/cbl_model_driver_offline.F90:38-231),
 * together with the offline driver's contract around it
 * (src/offline/cable_serial.F90:572-575 before the call, :602-605 after it)
 * and the per-tile balance accumulators (src/offline/cable_checks.F90:472-618,
 * called at cable_serial.F90:725-729).
 *
 * Data model.  Every per-tile field of include/cable_fields.h lives in device
 * memory as the Fortran (mp, ...) array, tile index fastest.  A host array is
 * attached by field id and copied in or out; the Fortran wrapper passes
 * C_LOC of the derived-type component (ssnow%wb, canopy%fe, ...).  The
 * device copy is authoritative between a step and the next download.
 *
 * Errors.  Every function returns 0 on success and a nonzero CABLE_E_* code
 * otherwise; cable_gpu_last_error() gives the message of the calling
 * thread's last failure.  The reference aborts (cable_abort -> error stop,
 * src/util/error_handler/cable_error_handler.F90:48-61); the Fortran wrapper
 * turns a nonzero code into that abort.
 *
 * Threading.  A context is used by one host thread at a time; its work is
 * ordered on the context's own HIP stream.
 */
#ifndef CABLE_GPU_H
#define CABLE_GPU_H

#include <stdint.h>

#include "cable_fields.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cable_gpu_ctx cable_gpu_ctx;

/* precision of the reference's default-REAL quantities on the device */
enum {
  CABLE_PREC_FP64 = 0,  /* plain: everything in FP64 */
  CABLE_PREC_MIXED = 1  /* FP32 radiation, roughness, canopy iteration; FP64
                           soil-moisture and soil-temperature solves and the
                           balance accumulators */
};

/* host element types for upload/download */
enum { CABLE_HOST_F64 = 0, CABLE_HOST_F32 = 1, CABLE_HOST_I32 = 2 };

enum {
  CABLE_OK = 0,
  CABLE_E_ARG = 1,      /* invalid argument (null pointer, bad id, size) */
  CABLE_E_HIP = 2,      /* HIP runtime error (allocation, copy, launch) */
  CABLE_E_STATE = 3,    /* call out of order (e.g. run before forcing) */
  CABLE_E_NOSYM = 4
};

/* Field table (include/cable_fields.h).  prec: 'F','D','I','B'; role:
   0 PARAM, 1 FORC, 2 STATE, 3 DIAG.  Returns CABLE_E_ARG for a bad id. */
int cable_gpu_nfields(void);
int cable_gpu_field_info(int id, const char **name, char *prec, int *n, int *role);
/* id of a Fortran component name such as "ssnow%wb", or -1 */
int cable_gpu_field_id(const char *name);

/* Create a context for mp tiles.  Allocates the device copy of every field
   (zero-filled, as the reference's harness zero-fills its derived types) on
   the current HIP device.  Replaces the allocation the reference performs in
   alloc_cbm_var (src/offline/cable_define_types.F90:725-741). */
int cable_gpu_create(cable_gpu_ctx **ctx, int64_t mp, int precision);
int cable_gpu_destroy(cable_gpu_ctx *ctx);

/* Copy one field between a host array of mp*n elements and the device. */
int cable_gpu_upload(cable_gpu_ctx *ctx, int field_id, const void *host, int host_type);
int cable_gpu_download(cable_gpu_ctx *ctx, int field_id, void *host, int host_type);

/* Several fields in one transfer each way (what a drop-in moves around each
   call): hosts[k] is the caller's array of field ids[k] in host type
   host_types[k].  The library packs the fields into one pinned staging
   buffer of the context (type conversion and tile order applied on up to
   eight host threads), moves it with one copy and scatters it on the device
   with one kernel; a download is the same in reverse.  Returns when the
   host arrays (download) or the device fields (upload) hold the values.
   Replaces the per-field transfers of the drop-in's host boundary
   (cable_gpu_upload/download above); no counterpart in the reference, whose
   cbm works on host memory (src/offline/cbl_model_driver_offline.F90:38). */
int cable_gpu_upload_fields(cable_gpu_ctx *ctx, int n, const int *field_ids, const void *const *hosts,
                            const int *host_types);
int cable_gpu_download_fields(cable_gpu_ctx *ctx, int n, const int *field_ids, void *const *hosts,
                              const int *host_types);

/* Order of the tiles in device memory: device position j holds the caller's
   tile order[j] (a permutation of 0..mp-1).  Uploads, downloads and the
   resident forcing are permuted on the way in and out, so the caller keeps
   its own tile order; the physics of each tile is unchanged.  Grouping tiles
   of one vegetation type lets a wavefront run one type's iteration counts
   and branches instead of all of them.  May be called any time before
   cable_gpu_set_forcing; device contents are carried over.  (No counterpart
   in the reference, whose tiles are independent in any order.) */
int cable_gpu_set_tile_order(cable_gpu_ctx *ctx, const int64_t *order);

/* Saved state (restart).  cable_gpu_save_state packs every STATE and DIAG
   field of include/cable_fields.h -- the prognostic state, the hidden state
   the reference carries between steps (wbtot1/wbtot2, gammzz(:,1), owetfac,
   the driver's owb) and the diagnostics that keep values on tiles a step
   does not touch -- in the device's element types and the caller's tile
   order, behind a header that records mp, the precision and the field
   table.  cable_gpu_state_bytes gives the size (negative: error).
   cable_gpu_load_state checks the header and uploads the fields; with
   continue_run = 1 the next step continues the saved run exactly, with 0 it
   is the first step of a new run, as the reference started from a restart
   (first-call gammzz(:,1), src/science/soilsnow/cbl_soilsnow_main.F90:60-62).
   Parameters and forcing are not part of the state.  Replaces, for this
   path, the reference's restart output and input
   (src/util/output/cable_output.F90:373-376, load_parameters in
   src/offline/cable_parameters.F90). */
int64_t cable_gpu_state_bytes(cable_gpu_ctx *ctx);
int cable_gpu_save_state(cable_gpu_ctx *ctx, void *buf, int64_t bytes);
int cable_gpu_load_state(cable_gpu_ctx *ctx, const void *buf, int64_t bytes, int continue_run);

/* Grid-wide parameters: soil%zse(6), soil%zshh(7), bgc%ratecp(3),
   bgc%ratecs(2) (src/offline/cable_define_types.F90:119-213, 706-717). */
int cable_gpu_set_globals(cable_gpu_ctx *ctx, const double *zse, const double *zshh,
                          const double *ratecp, const double *ratecs);

/* One call of the reference's cbm (src/offline/cbl_model_driver_offline.F90:
   38-231) with every diagnostic of the derived types written; ktau counts
   from 1, dels in seconds.  The first call of a context performs the
   reference's first-call initialisation of gammzz(:,1)
   (src/science/soilsnow/cbl_soilsnow_main.F90:60-62, 92-96).  This is what
   the Fortran drop-in (cable_amd/fortran/cbm_gpu.F90) calls. */
int cable_gpu_cbm(cable_gpu_ctx *ctx, int ktau, double dels);

/* One step of the offline driver: its contract before cbm (ofsd, oldcansto,
   vlai = 0 where iveg >= 14; src/offline/cable_serial.F90:572-575), cbm, the
   unit conversions after it (:602-605), the running carbon sums of sumcflux
   (src/offline/casa_cable.F90:482-560, called at :713-715; the sum_flux%*
   fields, restarted at ktau = 1) and the per-tile water and energy balance
   checks (src/offline/cable_checks.F90:472-618, called at
   cable_serial.F90:725-729; the driver's saved previous water store is the
   field drv%owb).  cable_gpu_run performs the same driver step. */
int cable_gpu_step(cable_gpu_ctx *ctx, int ktau, double dels);

/* ---- device-resident run (production / benchmark path) ---------------- */
/* Per-class parameter tables: veg (17 x cable_gpu_veg_cols()) and soil
   (9 x cable_gpu_soil_cols()), row-major FP64, in the column order given by
   cable_gpu_table_column(); the per-tile iveg, isoilm, vlai and albsoil stay
   per-tile fields. */
int cable_gpu_veg_cols(void);
int cable_gpu_soil_cols(void);
/* column of field `field_id` in the veg (table=0) or soil (table=1) table,
   -1 if the field is not tabulated */
int cable_gpu_table_column(int table, int field_id);
int cable_gpu_set_class_tables(cable_gpu_ctx *ctx, const double *veg, const double *soil);

/* Forcing for nsteps steps kept in device memory: each array is
   (nsteps, mp), tile fastest.  The remaining met fields follow the offline
   forcing reader (src/offline/cable_input.F90:1880-1883, 2666-2680):
   fsd(:,2) = fsd(:,1), tvair = tvrad = tk, precip_sn = precip where
   tk <= 273.16, ca and doy uniform. */
int cable_gpu_set_forcing(cable_gpu_ctx *ctx, int nsteps, const void *fsd1, const void *fld,
                          const void *tk, const void *qv, const void *ua, const void *pmb,
                          const void *precip, const void *coszen, int host_type, double ca,
                          double doy);

/* Synthetic forcing generated on the device for steps k_first ..
   k_first+nsteps-1, replacing the resident forcing (runs of more tiles and
   steps than a host-side forcing file would hold; no counterpart in the
   reference, which reads met files in get_met_data,
   src/offline/cable_input.F90:1650-2692).  The met variables follow the
   forcing reader: SWdown split evenly over the two bands (:1880-1883),
   coszen by sinbet (:2677) and, in the step, tvair = tvrad = tk and
   precip_sn = precip where tk <= 273.16 (:2666-2680), ca = 400 ppm.
   Per (caller tile t, step k) the draws are U(0,1) values keyed by
   SplitMix64 on (seed, draw, k, t): SWdown = max(0, 1000 sin(2 pi (hod-6)/24))
   x U(0.5, 1); LWdown ~ U(250, 420); Tair = t0[t] + 5 sin(2 pi (hod-9)/24);
   RH ~ U(30, 90)% -> qv by rh_sh (src/offline/cable_checks.F90:633-696);
   wind lognormal with mean 3 m/s and sigma 0.5; precip Bernoulli(0.1) x
   Exp(mean 2 mm/h) in mm per step; PSurf ~ U(950, 1020) hPa; hod = hod0 +
   k dels/3600; lat[t] in degrees.  t0 and lat are the caller's per-tile
   arrays (mp values).  cable_gpu_run then accepts steps in the window. */
int cable_gpu_synth_forcing(cable_gpu_ctx *ctx, int k_first, int nsteps, uint64_t seed, double dels,
                            double doy, double hod0, const double *t0, const double *lat);
/* Met forcing in the met file's own variables and units, turned into the
   resident forcing as get_met_data forms met% (src/offline/cable_input.F90:
   1650-2692) for the offline driver's local-time clock.  Arrays are
   (nsteps, mp) FP32 in the caller's tile order, as a met file holds them;
   null where the file lacks the variable.  No counterpart as a function in
   the reference, which reads the file itself. */
typedef struct {
  const float *swdown;   /* W/m2; fsd(:,1) = fsd(:,2) = SWdown/2 (:1880-1883) */
  const float *lwdown;   /* W/m2; null: Swinbank's 0.0000094*0.0000000567*tk**6 (:2257) */
  const float *tair;     /* Tair + tair_offset = tk (:1923) */
  const float *qair;     /* kg/kg, or relative humidity in % (qair_is_rh) through rh_sh (:2000-2003) */
  const float *wind;     /* m/s: scalar wind, or the northward part when wind_e is given */
  const float *wind_e;   /* m/s eastward part: ua = sqrt(wind**2 + wind_e**2) (:2094); or null */
  const float *psurf;    /* PSurf * psurf_scale = pmb in hPa (:1959); null: from elevation and tk (:1979-1981) */
  const float *rainf;    /* Rainf * rainf_scale = precip in mm per step (:2202-2204) */
  double tair_offset;    /* convert%Tair: 273.16 for deg C, 0 for K (:1051-1058) */
  double psurf_scale;    /* convert%PSurf: 0.01 for Pa, 10 for kPa, 1 for hPa (:1197-1210) */
  double rainf_scale;    /* convert%Rainf: dels for per-second rates, dels/3600 for per-hour (:1108-1115) */
  int qair_is_rh;        /* convert%Qair == -999 (:1075-1083) */
  const double *elevation; /* m per tile (needed when psurf is null) */
  const double *latitude;  /* degrees per tile, for coszen = sinbet(doy, lat, hod) (:2677) */
  double shod;           /* local hour of day of the run's first step (ktau = kstart) */
  int sdoy, smoy, syear; /* its day of year, month and year */
  double fixed_co2;      /* ppmv: ca = fixedCO2/1e6 (:2295) */
} cable_gpu_met;
/* Steps k_first .. k_first+nsteps-1 of a run whose first step is k = 0: the
   clock (hod, doy; :1685-1850) advances dels/3600 per step from shod and
   rolls days, months and years over as the reader does, and the step uses
   the window's day of year.  Snowf is not taken: precip_sn is partitioned
   by temperature in the step (:2666-2672), the reader's rule when the file
   has no snowfall. */
int cable_gpu_set_met(cable_gpu_ctx *ctx, int k_first, int nsteps, double dels, const cable_gpu_met *met);
/* The reader's clock alone, in the precision's REAL (no device work): hod
   and doy of steps k_first .. k_first+nsteps-1. */
int cable_gpu_met_clock(int precision, double shod, int sdoy, int smoy, int syear, double dels, int k_first,
                        int nsteps, double *hod, double *doy);

/* One resident forcing variable of step k in the caller's tile order;
   which: 0 fsd(:,1), 1 fld, 2 tk, 3 qv, 4 ua, 5 pmb, 6 precip, 7 coszen. */
int cable_gpu_download_forcing(cable_gpu_ctx *ctx, int k, int which, void *host, int host_type);

/* Run steps k0..k0+nsteps-1 of the resident forcing (ktau = k0+1 ...),
   writing only the prognostic state and the eight headline fluxes
   (canopy%rnet, fe, fh, fev, fes, fgpp, ssnow%runoff, rad%trad).  The
   launches are queued on the context stream and the call returns when they
   have completed (it reads the device time and the work counter). */
int cable_gpu_run(cable_gpu_ctx *ctx, int k0, int nsteps, double dels);
int cable_gpu_sync(cable_gpu_ctx *ctx);

/* The offline driver's consistency check (src/offline/cable_serial.F90:
   748-799, run when cable_user%consistency_check is set).  enable = 1 zeroes
   the running sums; each later step of cable_gpu_run then adds, over all
   tiles, new_sumbal += SUM(bal%wbal)/mp + SUM(bal%ebal)/mp,
   new_sumfpn += SUM(canopy%fpn)/mp, new_sumfe += SUM(canopy%fe)/mp and
   count_bal += 1, in FP64 and in a fixed order on the device.  enable = 0
   stops adding. */
int cable_gpu_consistency_check(cable_gpu_ctx *ctx, int enable);
/* out[6]: count_bal, new_sumbal (the value compare_consistency_check_values
   checks), new_sumfe/count_bal, new_sumfpn/count_bal,
   SUM(bal%ebal_tot)/mp/count_bal and SUM(bal%wbal_tot)/mp/count_bal (what the
   driver prints at the last step, cable_serial.F90:756-771). */
int cable_gpu_consistency_values(cable_gpu_ctx *ctx, double *out);

/* Device time of the last cable_gpu_run, in milliseconds, measured with HIP
   events on the context stream around its kernels. */
int cable_gpu_last_run_ms(cable_gpu_ctx *ctx, double *ms);

/* Device time of the last run's kernels, in milliseconds, summed over its
   steps: the step kernels, and the surrogate inference kernels of a hybrid
   run (0 otherwise); HIP events on the context stream around each launch of
   a hybrid run, around the run otherwise.  (No counterpart in the
   reference; a measurement.) */
int cable_gpu_last_kernel_ms(cable_gpu_ctx *ctx, double *step_ms, double *infer_ms);
/* Whether a hybrid run records those per-launch events (on = 1, the
   default).  Each event is a timestamp packet the stream waits on: four per
   step measured 4% of the hybrid's time per step at 1,048,576 tiles.  Off,
   cable_gpu_last_kernel_ms gives the whole run as step time and no inference
   time.  (No counterpart in the reference; a measurement.) */
int cable_gpu_set_launch_timing(cable_gpu_ctx *ctx, int on);

/* Total dryLeaf fixed-point passes of the last step or run (summed over tiles
   and the four stability iterations), a work counter. */
int cable_gpu_counters(cable_gpu_ctx *ctx, int64_t *dryleaf_passes);

/* ---- canopy surrogate (mixed-precision contexts; no counterpart in the
   reference) -------------------------------------------------------------
   The surrogate (include/cable_mlp.h) maps a tile's forcing, state and
   parameters at the start of a step to the converged state of the leaf
   fixed point (src/science/canopy/cbl_dryLeaf.F90:10-666).  Its 24 inputs:
   fsd(1), fld, tk - 273.16, 1000 qv, ua, pmb, coszen, vlai, 1e6 vcmax,
   frac4, hc, a1gs, d0gs, dleaf, xfang, extkn, fwsoil (standard scheme),
   the previous step's leaf temperature, canopy-air temperature (both minus
   tk) and 1000 x (canopy-air humidity - qv), tss - tk, snowd, cansto,
   rtsoil.  Its 8 outputs: leaf temperature - tk after the first and the last
   NITER pass, leaf-surface CO2 / ca of the sunlit and shaded leaves after
   the first pass (which carries the net assimilation, csx = ca - 1.32 An /
   gb), the final stomatal conductances of both leaves, and the canopy's
   latent and sensible heat. */

/* Run steps k0..k0+nsteps-1 as cable_gpu_run, and append one training
   sample (24 inputs, 8 converged outputs) per vegetated tile-step to a
   device buffer of `capacity` samples (reallocated if larger). */
int cable_gpu_run_capture(cable_gpu_ctx *ctx, int k0, int nsteps, double dels, int64_t capacity);
/* device pointers to the captured samples (row-major [n][24] and [n][8]),
   valid until the next capture or destroy */
int cable_gpu_samples(cable_gpu_ctx *ctx, const float **x, const float **y, int64_t *n);
/* copy captured samples first..first+n-1 to host arrays [n][24], [n][8] */
int cable_gpu_copy_samples(cable_gpu_ctx *ctx, int64_t first, int64_t n, float *x, float *y);
/* network parameters and normalisation for the hybrid step (layouts as in
   include/cable_mlp.h) */
int cable_gpu_set_surrogate(cable_gpu_ctx *ctx, const float *params, const float *norm);
/* Run steps k0..k0+nsteps-1 with the hybrid method: per step the network
   predicts the leaf state of each sunlit leaf-bearing tile and the leaf
   fixed point of those tiles starts there.  After max_passes polish passes
   (per NITER pass) such a tile is checked by its leaf energy-balance
   residual beyond the reference's 0.1 K tolerance, cr (ghr1 + ghr2)
   (|dT| - 0.1 K): within resid_max (W/m2) it stops; otherwise it falls back
   to the reference's iteration, which goes on with the reference's damping
   up to its 20 passes (src/science/canopy/cbl_dryLeaf.F90:239-606).  Every
   other tile runs the reference's iteration.  use_surrogate = 0 runs the
   same capped method from the reference's own starting point (the
   comparison without the network).  Which steps have a leaf-bearing tile
   in daylight is found by one scan of the forcing window's shortwave, and
   one wait, per 64 steps: such a step is two launches, the inference and
   the step; any other step is one launch of the plain step, which before a
   daylight step (and at the end of the run or of the 64) also keeps the
   leaf state the network's inputs read. */
int cable_gpu_run_hybrid(cable_gpu_ctx *ctx, int k0, int nsteps, double dels, int max_passes,
                         double resid_max, int use_surrogate);
/* cable_gpu_run_hybrid that allocates nothing.  No counterpart in the
   reference: for callers that cannot allocate device memory while they run
   (code under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc).  The step's scratch -- the network's predictions and the work
   and fallback counters -- lives in a device array of the library's code
   object, sized for at most 1,048,576 tiles (32 MiB, held on each GPU the
   library runs on until the process ends); a larger context is refused with
   CABLE_E_ARG.  Everything else must exist before the call: the context and
   its fields (cable_gpu_create), class tables and forcing, and, with
   use_surrogate = 1, the network (cable_gpu_set_surrogate, which also sets up
   the surrogate's state; without it the call returns CABLE_E_STATE).  Every
   call on a GPU shares that GPU's scratch: calls that overlap on one GPU
   race.  Results, counters (cable_gpu_hybrid_counters, cable_gpu_counters)
   and cable_gpu_last_run_ms as cable_gpu_run_hybrid; no per-kernel split
   (cable_gpu_last_kernel_ms gives the whole run as step time).  It makes no
   host wait during the run, so it does not look for the steps without
   daylight: every step runs the inference and the hybrid step, which give
   the plain step's bits there, at the cost of the night steps' extra time. */
int cable_gpu_run_hybrid_noalloc(cable_gpu_ctx *ctx, int k0, int nsteps, double dels, int max_passes,
                                 double resid_max, int use_surrogate);
/* tile-steps that fell back, and all tile-steps, of the last hybrid run */
int cable_gpu_hybrid_counters(cable_gpu_ctx *ctx, int64_t *fallback_tile_steps, int64_t *tile_steps);
/* The surrogate's view of step k from the current state, as the hybrid step
   computes it: x (mp x 24) the raw features, pred (mp x 8) the denormalised
   predictions, host arrays in the caller's tile order.  Predictions are
   computed for every tile, including those the hybrid step does not
   warm-start (no leaves, or no daylight). */
int cable_gpu_surrogate_eval(cable_gpu_ctx *ctx, int k, float *x, float *pred);

const char *cable_gpu_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* CABLE_GPU_H */
