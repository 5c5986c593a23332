// cbm_kernel.h -- the timestep kernel: one thread per tile.
//
// cbm_kernel<R, FULL, TABLES> gathers one tile's fields from the device
// arrays, runs Stepper<R, FULL, TABLES> (cbm_step.h) and scatters the
// results.  FULL: every field of include/cable_fields.h in and every state
// and diagnostic field out (the drop-in step).  !FULL: the resident run's
// compact set -- prognostic state, eight forcing fields and the per-tile
// parameters in, state and eight headline fluxes out.
//
// The kernels of each working precision are instantiated in their own
// translation unit (cbm_kernels_fp64.hip, cbm_kernels_mixed.hip), so each can
// be compiled with the floating-point options its precision contract allows.
#pragma once
#include <hip/hip_runtime.h>

#include "cbm_step.h"
#include "mlp_kernels.h"

namespace cbm {

namespace mlp = cbm_mlp;
using cbm_mlp::f32_to_bf16;
using cbm_mlp::hswish;
using cbm_mlp::lds_barrier;
using cbm_mlp::v16f;
using cbm_mlp::v4s;

enum { ROLE_PARAM = 0, ROLE_FORC = 1, ROLE_STATE = 2, ROLE_DIAG = 3 };


// field arrays are device (global) memory: global loads and stores rather
// than flat ones, which also count against the LDS/scalar wait counter
template <class T>
using gptr = __attribute__((address_space(1))) T *;
template <class T>
CBM_HD T ld(const void *p, int64_t i) {
  return ((const gptr<const T>)p)[i];
}
template <class T, class V>
CBM_HD void st(void *p, int64_t i, V v) {
  ((gptr<T>)p)[i] = static_cast<T>(v);
}

// device type of a field of precision code P for working precision R
template <class R, char P> struct DevT { using type = R; };
template <class R> struct DevT<R, 'D'> { using type = double; };
template <class R> struct DevT<R, 'B'> { using type = double; };
template <class R> struct DevT<R, 'I'> { using type = int; };

constexpr int role_id(int r) { return r; }

// FP32 working precision: the register budget of two waves per SIMD (the
// FP32 live set fits it and the extra wave hides the latency of the LDS
// table reads and transcendental sequences); FP64 keeps one wave per SIMD,
// which measured faster than spilling its larger live set
// ---------------------------------------------------------------------------
// surrogate inputs: 24 features of a tile at the start of a step, read from
// the device arrays (the resident run's layout: forcing at a.foff) and the
// class-table rows.  The capture step and the inference kernel both call
// this, so training and use see the same features.  The loads and the
// arithmetic are separate steps, so that the inference can issue one block's
// loads before the previous block's network and finish them after it.
template <class R>
struct FeatureRaw {
  using DT = typename DevT<R, 'F'>::type;
  DT tk, qv, fsd, fld, ua, pmb, coszen, vlai, tss, snowd, cansto, rtsoil;
  double wb[MS], wbice[MS];
  float hyb[3];
  int iveg, isoilm;
};

// the device arrays the features read, looked up once: a kernel that stores
// between feature loads would otherwise reload every pointer from a.F
struct FeatureSrc {
  const void *tk, *qv, *fsd, *fld, *ua, *pmb, *coszen, *iveg, *isoilm, *vlai, *wb, *wbice, *tss, *snowd, *cansto,
      *rtsoil;
};
CBM_HD inline FeatureSrc feature_src(const StepArgs &a) {
  void *const *F = a.F;
  return {F[CABLE_F_met_tk],       F[CABLE_F_met_qv],      F[CABLE_F_met_fsd],   F[CABLE_F_met_fld],
          F[CABLE_F_met_ua],       F[CABLE_F_met_pmb],     F[CABLE_F_met_coszen], F[CABLE_F_veg_iveg],
          F[CABLE_F_soil_isoilm],  F[CABLE_F_veg_vlai],    F[CABLE_F_ssnow_wb],  F[CABLE_F_ssnow_wbice],
          F[CABLE_F_ssnow_tss],    F[CABLE_F_ssnow_snowd], F[CABLE_F_canopy_cansto], F[CABLE_F_ssnow_rtsoil]};
}

template <class R>
CBM_HD void load_feature_raw(const StepArgs &a, const FeatureSrc &p, int64_t t, FeatureRaw<R> &r) {
  using DT = typename DevT<R, 'F'>::type;
  const int64_t mp = a.mp, f = t + a.foff;
  r.tk = ld<DT>(p.tk, f);
  r.qv = ld<DT>(p.qv, f);
  r.fsd = ld<DT>(p.fsd, f);
  r.fld = ld<DT>(p.fld, f);
  r.ua = ld<DT>(p.ua, f);
  r.pmb = ld<DT>(p.pmb, f);
  r.coszen = ld<DT>(p.coszen, f);
  r.iveg = ld<int>(p.iveg, t);
  r.isoilm = ld<int>(p.isoilm, t);
  r.vlai = ld<DT>(p.vlai, t);
#pragma unroll
  for (int k = 0; k < MS; k++) {
    r.wb[k] = ld<double>(p.wb, k * mp + t);
    r.wbice[k] = ld<double>(p.wbice, k * mp + t);
  }
  const bool h = a.hyb_valid && a.hyb;
#pragma unroll
  for (int c = 0; c < 3; c++) r.hyb[c] = h ? a.hyb[c * mp + t] : 0.0f;
  r.tss = ld<DT>(p.tss, t);
  r.snowd = ld<DT>(p.snowd, t);
  r.cansto = ld<DT>(p.cansto, t);
  r.rtsoil = ld<DT>(p.rtsoil, t);
}

// vrow, srow: the rows of the tile's vegetation type and soil type
template <class R>
CBM_HD void finish_features(const StepArgs &a, const FeatureRaw<R> &r, const double *vrow, const double *srow,
                            float *x) {
  const float tk = (float)r.tk, qv = (float)r.qv;
  const float vlai = r.iveg >= 14 ? 0.0f : (float)r.vlai;
  x[0] = (float)r.fsd;
  x[1] = (float)r.fld;
  x[2] = tk - 273.16f;
  x[3] = qv * 1.0e3f;
  x[4] = (float)r.ua;
  x[5] = (float)r.pmb;
  x[6] = (float)r.coszen;
  x[7] = vlai;
  x[8] = (float)vrow[VC_vcmax] * 1.0e6f;
  x[9] = (float)vrow[VC_frac4];
  x[10] = (float)vrow[VC_hc];
  x[11] = (float)vrow[VC_a1gs];
  x[12] = (float)vrow[VC_d0gs];
  x[13] = (float)vrow[VC_dleaf];
  x[14] = (float)vrow[VC_xfang];
  x[15] = (float)vrow[VC_extkn];
  // root-zone water stress of the standard scheme (cbl_fwsoil.F90:13-38)
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < MS; k++) {
    const double wl = r.wb[k] - r.wbice[k];
    const double sw = srow[SC_swilt_vec + k], fc = srow[SC_sfc_vec + k];
    float q = (float)((wl - sw) / (fc - sw));
    q = q < 1.0e-9f ? 1.0e-9f : (q > 1.0f ? 1.0f : q);
    s += (float)vrow[VC_froot + k] * q;
  }
  float fw = (float)vrow[VC_vbeta] * (s > 1.0e-9f ? s : 1.0e-9f);
  x[16] = fw < 1.0e-9f ? 1.0e-9f : (fw > 1.0f ? 1.0f : fw);
  const bool h = a.hyb_valid && a.hyb;
  x[17] = (h ? r.hyb[0] : tk) - tk;
  x[18] = (h ? r.hyb[1] : tk) - tk;
  x[19] = ((h ? r.hyb[2] : qv) - qv) * 1.0e3f;
  x[20] = (float)r.tss - tk;
  x[21] = (float)r.snowd;
  x[22] = (float)r.cansto;
  x[23] = (float)r.rtsoil;
}

template <class R>
CBM_HD void surrogate_features(const StepArgs &a, int64_t t, const double *vrow, const double *srow, float *x) {
  FeatureRaw<R> r;
  load_feature_raw<R>(a, feature_src(a), t, r);
  finish_features<R>(a, r, vrow, srow, x);
}

// wave-wide append: lanes with `take` get consecutive slots from *count
__device__ inline unsigned long long wave_append(bool take, unsigned long long *count) {
  const unsigned long long mask = __ballot(take);
  const int lane = threadIdx.x & 63;
  unsigned long long base = 0;
  if (lane == 0 && mask) base = atomicAdd(count, (unsigned long long)__popcll(mask));
  base = __shfl(base, 0, 64);
  return base + (unsigned long long)__popcll(mask & ((1ull << lane) - 1ull));
}

// what a tile's step reports to the kernel around it
template <int MODE, bool CONS>
struct TileOut {
  int passes = 0;                      // dryLeaf passes (work counter)
  bool fallback = false;               // MODE_HYBRID: the tile fell back to the
                                       // reference's leaf iteration
  bool emit = false;                   // MODE_CAPTURE: the tile yields a sample
  double cons[CONS ? 4 : 1] = {0.0};   // CONS: the tile's wbal, ebal, fpn, fe
  float feat[MODE == MODE_CAPTURE ? mlp::NIN : 1];
  float tgt[MODE == MODE_CAPTURE ? mlp::NOUT : 1];
};

// One tile's step: gather its fields from the arrays of a.F (FULL: every
// field; otherwise the resident run's compact set), run the physics, scatter
// the results.  vtab/stab are the class tables (TABLES), staged in LDS by the
// kernel.  Host and device code: the test suite also runs it on the CPU, on
// host arrays, under the sanitizers.
template <class R, bool FULL, bool TABLES, int MODE, bool CONS>
CBM_HD void step_tile(const StepArgs &a, int64_t t, const double *vtab, const double *stab,
                      TileOut<MODE, CONS> &o) {
  const int64_t mp = a.mp;
  void *const *F = a.F;
  Tile<R> T;
  // ---- gather ----
  if constexpr (FULL) {
#define PARAM ROLE_PARAM
#define FORC ROLE_FORC
#define STATE ROLE_STATE
#define DIAG ROLE_DIAG
#define X(id, name, prec, n, role)                                                    \
  {                                                                                   \
    using DT = typename DevT<R, #prec[0]>::type;                                      \
    const int64_t off = (role == ROLE_FORC) ? a.foff : 0;                             \
    _Pragma("unroll") for (int k = 0; k < n; k++) T.id[k] =                           \
        (decltype(T.id[0] + 0))ld<DT>(F[CABLE_F_##id], (int64_t)k * mp + t + off);     \
  }
    CABLE_FIELDS(X)
#undef X
#undef PARAM
#undef FORC
#undef STATE
#undef DIAG
  } else {
    // prognostic state
#define LDS_(id, n)                                                                   \
  {                                                                                   \
    using DT = typename DevT<R, 'F'>::type;                                           \
    _Pragma("unroll") for (int k = 0; k < n; k++) T.id[k] = ld<DT>(F[CABLE_F_##id], (int64_t)k * mp + t); \
  }
#define LDD_(id, n)                                                                   \
  { _Pragma("unroll") for (int k = 0; k < n; k++) T.id[k] = ld<double>(F[CABLE_F_##id], (int64_t)k * mp + t); }
    T.ssnow_isflag[0] = ld<int>(F[CABLE_F_ssnow_isflag], t);
    LDS_(ssnow_pudsto, 1) LDS_(ssnow_rtsoil, 1) LDS_(ssnow_wbtot1, 1) LDS_(ssnow_wbtot2, 1)
    LDS_(ssnow_wb_lake, 1) LDS_(ssnow_owetfac, 1) LDS_(ssnow_snage, 1) LDS_(ssnow_snowd, 1)
    LDS_(ssnow_ssdnn, 1) LDS_(ssnow_tss, 1)
    LDS_(ssnow_tgg, 1) LDS_(ssnow_tggsn, 1) LDS_(canopy_cansto, 1)
    LDD_(ssnow_wb, 1) LDD_(ssnow_wbice, 1)
    // the canopy reads only the top soil and snow layer; the deeper layers,
    // snow layers, carbon pools, gammzz(:,1) and the balance accumulators
    // are loaded after the canopy (Stepper::run's `late`).  Soil liquid
    // water as the previous step left it (wb - wbice), before this step's
    // lake fix-up:
#pragma unroll
    for (int k = 0; k < MS; k++)
      T.ssnow_wbliq[k] = ld<double>(F[CABLE_F_ssnow_wb], (int64_t)k * mp + t) -
                         ld<double>(F[CABLE_F_ssnow_wbice], (int64_t)k * mp + t);
    // per-tile parameters
    T.veg_iveg[0] = ld<int>(F[CABLE_F_veg_iveg], t);
    T.soil_isoilm[0] = ld<int>(F[CABLE_F_soil_isoilm], t);
    LDS_(veg_vlai, 1)
    T.soil_albsoil[0] = ld<typename DevT<R, 'F'>::type>(F[CABLE_F_soil_albsoil], t);
    // compact forcing (offline reader's derivations)
    {
      using DT = typename DevT<R, 'F'>::type;
      const int64_t i = t + a.foff;
      R fsd = ld<DT>(F[CABLE_F_met_fsd], i);
      T.met_fsd[0] = fsd;
      T.met_fsd[1] = fsd;
      R tk = ld<DT>(F[CABLE_F_met_tk], i);
      T.met_tk[0] = tk;
      T.met_tvair[0] = tk;
      T.met_tvrad[0] = tk;
      T.met_fld[0] = ld<DT>(F[CABLE_F_met_fld], i);
      T.met_qv[0] = ld<DT>(F[CABLE_F_met_qv], i);
      T.met_ua[0] = ld<DT>(F[CABLE_F_met_ua], i);
      T.met_pmb[0] = ld<DT>(F[CABLE_F_met_pmb], i);
      R pr = ld<DT>(F[CABLE_F_met_precip], i);
      T.met_precip[0] = pr;
      T.met_precip_sn[0] = (tk <= R(273.16)) ? pr : R(0.0);
      T.met_coszen[0] = ld<DT>(F[CABLE_F_met_coszen], i);
      T.met_ca[0] = (R)a.ca;
      T.met_doy[0] = (R)a.doy;
    }
#undef LDS_
#undef LDD_
  }
  const int iv = T.veg_iveg[0] - 1, is = T.soil_isoilm[0] - 1;
  const double *vrow = TABLES ? vtab + (iv < 0 ? 0 : iv) * VC_N : nullptr;
  const double *srow = TABLES ? stab + (is < 0 ? 0 : is) * SC_N : nullptr;
  Stepper<R, FULL, TABLES, MODE> S(T, a, vrow, srow);
  auto late = [&]() {
    if constexpr (!FULL) {
      void *const *F = a.F;
      using DT = typename DevT<R, 'F'>::type;
#define LDS_(id, n) \
  { _Pragma("unroll") for (int k = 0; k < n; k++) T.id[k] = ld<DT>(F[CABLE_F_##id], (int64_t)k * mp + t); }
#define LDD_(id, n) \
  { _Pragma("unroll") for (int k = 0; k < n; k++) T.id[k] = ld<double>(F[CABLE_F_##id], (int64_t)k * mp + t); }
      // a fresh read: nothing has written these since the gather, and the
      // compiler must not keep the gather's values live to reuse them
      __asm__ __volatile__("" ::: "memory");
#define LDL_(T_, id, k0, n) \
  { _Pragma("unroll") for (int k = k0; k < n; k++) T.id[k] = ld<T_>(F[CABLE_F_##id], (int64_t)k * mp + t); }
      LDL_(DT, ssnow_tgg, 1, 6) LDL_(DT, ssnow_tggsn, 1, 3) LDL_(double, ssnow_wb, 1, 6)
      LDL_(double, ssnow_wbice, 1, 6)
#undef LDL_
      // values the canopy reads only before its stability iterations and
      // soil, snow and the balances read again: the shortwave forcing, the
      // precipitation and its snow part, and the step's starting surface
      // temperature (otss, which the step sets from the stored tss).  FP64
      // only: the FP32 kernel's register allocation spills more with it
      if constexpr (sizeof(R) == 8) {
        const int64_t i = t + a.foff;
        const R fsd = ld<DT>(F[CABLE_F_met_fsd], i), pr = ld<DT>(F[CABLE_F_met_precip], i);
        T.met_fsd[0] = fsd;
        T.met_fsd[1] = fsd;
        T.met_precip[0] = pr;
        T.met_precip_sn[0] = (T.met_tk[0] <= R(273.16)) ? pr : R(0.0);
        T.ssnow_otss[0] = ld<DT>(F[CABLE_F_ssnow_tss], t);
      }
      LDS_(ssnow_smass, 3) LDS_(ssnow_ssdn, 3) LDS_(bgc_cplant, 3) LDS_(bgc_csoil, 2)
      LDD_(drv_owb, 1) LDD_(bal_ebal_tot, 1) LDD_(bal_wbal_tot, 1) LDD_(bal_precip_tot, 1)
      LDD_(bal_rnoff_tot, 1) LDD_(bal_evap_tot, 1) LDD_(bal_radbalsum, 1)
      T.ssnow_gammzz[0] = ld<double>(F[CABLE_F_ssnow_gammzz], t);
#undef LDS_
#undef LDD_
    }
  };
  if constexpr (MODE == MODE_CAPTURE) surrogate_features<R>(a, t, vrow, srow, o.feat);
  if constexpr (MODE == MODE_HYBRID) {
    if (a.pred) {
#pragma unroll
      for (int c = 0; c < mlp::NOUT; c++) S.pred[c] = a.pred[t * mlp::NOUT + c];
    }
  }
  // the offline driver's running carbon sums (driver mode): the resident
  // run reads the previous step's sums only after soil and snow
  auto tail = [&]() {
    if constexpr (!FULL) {
      if (a.driver) {
        void *const *F = a.F;
        using DT = typename DevT<R, 'F'>::type;
        __asm__ __volatile__("" ::: "memory");
#define LDF_(id) T.id[0] = ld<DT>(F[CABLE_F_##id], t);
        LDF_(sum_flux_sumpn) LDF_(sum_flux_sumrp) LDF_(sum_flux_sumrpw) LDF_(sum_flux_sumrpr)
        LDF_(sum_flux_sumrs) LDF_(sum_flux_sumrd) LDF_(sum_flux_dsumpn) LDF_(sum_flux_dsumrp)
        LDF_(sum_flux_dsumrd)
#undef LDF_
      }
    }
  };
  // the previous-step canopy state the surrogate's features read (written
  // by the step at the end of the canopy)
  if constexpr (MODE != MODE_PLAIN) S.hyb_out = a.hyb ? a.hyb + t : nullptr;
  S.run(late, tail);
  if (a.driver) S.sumcflux();
  o.passes = S.npass;
  if constexpr (CONS) {
    o.cons[0] = (double)T.bal_wbal[0];
    o.cons[1] = (double)T.bal_ebal[0];
    o.cons[2] = (double)T.canopy_fpn[0];
    o.cons[3] = (double)T.canopy_fe[0];
  }
  if constexpr (MODE == MODE_HYBRID) o.fallback = S.fell_back;
  if constexpr (MODE == MODE_CAPTURE) {
    o.emit = S.vegetated;
#pragma unroll
    for (int c = 0; c < mlp::NOUT; c++) o.tgt[c] = S.tgt[c];
  }
  // ---- scatter ----
  if constexpr (FULL) {
#define PARAM ROLE_PARAM
#define FORC ROLE_FORC
#define STATE ROLE_STATE
#define DIAG ROLE_DIAG
#define X(id, name, prec, n, role)                                                      \
  if (role == ROLE_STATE || role == ROLE_DIAG || CABLE_F_##id == CABLE_F_met_tvair ||   \
      CABLE_F_##id == CABLE_F_veg_vlai) {                                               \
    using DT = typename DevT<R, #prec[0]>::type;                                        \
    const int64_t off = (role == ROLE_FORC) ? a.foff : 0;                               \
    _Pragma("unroll") for (int k = 0; k < n; k++)                                       \
        st<DT>(F[CABLE_F_##id], (int64_t)k * mp + t + off, T.id[k]);                    \
  }
    CABLE_FIELDS(X)
#undef X
#undef PARAM
#undef FORC
#undef STATE
#undef DIAG
  } else {
    using DT = typename DevT<R, 'F'>::type;
#define STS_(id, n) \
  { _Pragma("unroll") for (int k = 0; k < n; k++) st<DT>(F[CABLE_F_##id], (int64_t)k * mp + t, T.id[k]); }
#define STD_(id, n) \
  { _Pragma("unroll") for (int k = 0; k < n; k++) st<double>(F[CABLE_F_##id], (int64_t)k * mp + t, T.id[k]); }
    st<int>(F[CABLE_F_ssnow_isflag], t, T.ssnow_isflag[0]);
    STS_(ssnow_pudsto, 1) STS_(ssnow_rtsoil, 1) STS_(ssnow_wbtot1, 1) STS_(ssnow_wbtot2, 1)
    STS_(ssnow_wb_lake, 1) STS_(ssnow_owetfac, 1) STS_(ssnow_snage, 1) STS_(ssnow_snowd, 1)
    STS_(ssnow_ssdnn, 1) STS_(ssnow_tss, 1) STS_(ssnow_smass, 3) STS_(ssnow_ssdn, 3)
    STS_(ssnow_tgg, 6) STS_(ssnow_tggsn, 3) STS_(canopy_cansto, 1) STS_(bgc_cplant, 3)
    STS_(bgc_csoil, 2)
    STD_(ssnow_wb, 6) STD_(ssnow_wbice, 6) STD_(drv_owb, 1)
    STD_(bal_ebal_tot, 1) STD_(bal_wbal_tot, 1) STD_(bal_precip_tot, 1) STD_(bal_rnoff_tot, 1)
    STD_(bal_evap_tot, 1) STD_(bal_radbalsum, 1)
    st<double>(F[CABLE_F_ssnow_gammzz], t, T.ssnow_gammzz[0]);
    if (a.driver) {
      STS_(sum_flux_sumpn, 1) STS_(sum_flux_sumrp, 1) STS_(sum_flux_sumrpw, 1) STS_(sum_flux_sumrpr, 1)
      STS_(sum_flux_sumrs, 1) STS_(sum_flux_sumrd, 1) STS_(sum_flux_dsumpn, 1) STS_(sum_flux_dsumrp, 1)
      STS_(sum_flux_dsumrd, 1)
    }
    // the eight headline fluxes
    STS_(canopy_rnet, 1) STS_(canopy_fe, 1) STS_(canopy_fh, 1) STS_(canopy_fev, 1)
    STD_(canopy_fes, 1) STS_(canopy_fgpp, 1) STS_(ssnow_runoff, 1) STS_(rad_trad, 1)
#undef STS_
#undef STD_
  }
}

// CONS: also form the workgroup's sums for the driver's consistency check
// (a separate instantiation: the check's code costs the plain step 2-3% even
// when it is switched off at run time)
template <class R, bool FULL, bool TABLES, int MODE = MODE_PLAIN, bool CONS = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 2 : 1)))
cbm_kernel(StepArgs a) {
  __shared__ double lds_v[TABLES ? NVEGCLASS * VC_N : 1];
  __shared__ double lds_s[TABLES ? NSOILCLASS * SC_N : 1];
  const int64_t mp = a.mp;
  if constexpr (TABLES) {
    for (int i = threadIdx.x; i < NVEGCLASS * VC_N; i += blockDim.x) lds_v[i] = a.vtab[i];
    for (int i = threadIdx.x; i < NSOILCLASS * SC_N; i += blockDim.x) lds_s[i] = a.stab[i];
    __syncthreads();
  }
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = t < mp;
  TileOut<MODE, CONS> o;
  if (live) step_tile<R, FULL, TABLES, MODE, CONS>(a, t, lds_v, lds_s, o);
  if constexpr (CONS) {
    // workgroup sums in a fixed order: a butterfly within each wave, then
    // the waves in order
    __shared__ double wsum[4][4];
#pragma unroll
    for (int q = 0; q < 4; q++)
      for (int k = 32; k > 0; k >>= 1) o.cons[q] += __shfl_xor(o.cons[q], k, 64);
    const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
      for (int q = 0; q < 4; q++) wsum[w][q] = o.cons[q];
    __syncthreads();
    if (threadIdx.x < 4) {
      double s = 0.0;
      for (int i = 0; i < nw; i++) s += wsum[i][threadIdx.x];
      a.cons_part[(int64_t)blockIdx.x * 4 + threadIdx.x] = s;
    }
  }
  if (a.npass) {
    // wave-level sum, one atomic per wave
    int passes = o.passes;
    for (int k = 32; k > 0; k >>= 1) passes += __shfl_xor(passes, k, 64);
    if ((threadIdx.x & 63) == 0 && passes) atomicAdd(a.npass, (unsigned long long)passes);
  }
  if constexpr (MODE == MODE_HYBRID) {
    // tiles that fell back: a wave-level count, one atomic per wave
    const unsigned long long fb = __ballot(o.fallback);
    if ((threadIdx.x & 63) == 0 && fb) atomicAdd(a.nfallback, (unsigned long long)__popcll(fb));
  }
  if constexpr (MODE == MODE_CAPTURE) {
    const unsigned long long slot = wave_append(o.emit, a.scount);
    if (o.emit && slot < a.scap) {
#pragma unroll
      for (int c = 0; c < mlp::NIN; c++) a.sx[slot * mlp::NIN + c] = o.feat[c];
#pragma unroll
      for (int c = 0; c < mlp::NOUT; c++) a.sy[slot * mlp::NOUT + c] = o.tgt[c];
    }
  }
}

// ---------------------------------------------------------------------------
// surrogate inference for every tile of a step: 24 features -> network ->
// 8 denormalised predictions in pred[t*8 ..] (and, when xraw is given, the
// raw features in xraw[t*24 ..]).
//
// One workgroup per CU of eight waves steps blocks of 256 tiles, 32 per
// wave, two waves per SIMD so that one wave's matrix-core products overlap
// the other's activations.  The weights (BF16, transposed [out][in], rows padded by
// 8) and biases are staged once per workgroup into LDS; each wave keeps its 32 activation rows
// in LDS between layers.  A layer is computed transposed, D = W^T x A^T on
// v_mfma_f32_32x32x16_bf16: a lane then holds four consecutive outputs of
// one tile per register group, which it writes back as one 8-byte row
// segment.  Layer 1's K (24) is padded with zeros to 32.
namespace infer {
using namespace mlp;
constexpr int WAVES = 8, TILES = 32 * WAVES;
static_assert(TILES == 256, "the hybrid run's daylight scan votes on blocks of 256 tiles");
constexpr int VCHUNK = 64;                       // blocks voted on at once
constexpr int K1 = 32;                           // layer-1 K padded to a multiple of 16
constexpr int P1 = K1 + 8, PH = HID + 8;         // weight row pitches (BF16)
constexpr int AP = HID + 8;                      // activation row pitch
constexpr int OFF1 = 0, OFF2 = OFF1 + HID * P1, OFF3 = OFF2 + HID * PH, OFF4 = OFF3 + HID * PH;
constexpr int WLDS = OFF4 + NOUT * PH;           // 41,024 BF16
constexpr int ALDS = WAVES * 32 * AP;            // 34,816 BF16
constexpr int BLDS = 3 * HID + NOUT;             // biases, FP32
static_assert(2 * (WLDS + ALDS) + 4 * BLDS + 8 * (NVEGCLASS * VC_N + NSOILCLASS * SC_N) + 4 * (VCHUNK + NNORM + 3 * HID) <=
                  160 * 1024,
              "fits the 160 KB of LDS");
// the transposed global layout: per layer [out][in], layers back to back
constexpr int G1 = 0, G2 = G1 + HID * NIN, G3 = G2 + HID * HID, G4 = G3 + HID * HID, GTOT = G4 + NOUT * HID;
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf4 __attribute__((ext_vector_type(4)));
typedef uint32_t v4u __attribute__((ext_vector_type(4)));
}  // namespace infer

// One hidden layer of the inference for a wave's 32 tiles, in place in
// their activation rows A (pitch AP): Z = W^T A^T + b on
// v_mfma_f32_32x32x16_bf16, then hard-swish, as BF16, back into A.  Lane:
// tile `col`; register group g of output block nb: outputs nb*32 + 8g + kq +
// 0..3.  Software-pipelined over the four blocks of 32 outputs: the inputs
// are read into registers first (the rows are then free to be overwritten),
// block nb+1's weights are read while block nb's products run, and block
// nb's activations (vector work) are formed while block nb+1's products run
// on the matrix cores.  (A second weight buffer spilled registers.)  A block's accumulators start at its bias, formed by
// one more product: each output's BF16 pair (hi, lo) in inputs 0 and 1 of
// its row, against inputs fixed at one (a lane of the upper half holds
// inputs 8..15, all zero).
template <int KS>
__device__ __forceinline__ void infer_hidden(const __bf16 *W, int pw, __bf16 *A, const uint32_t *bpk, int col,
                                             int kh, int kq) {
  using namespace infer;
  bf8 xv[KS], w[KS];
#pragma unroll
  for (int s = 0; s < KS; s++) xv[s] = *(const bf8 *)&A[col * AP + 16 * s + kh];
  auto weights = [&](int nb, bf8 *w) {
#pragma unroll
    for (int s = 0; s < KS; s++) w[s] = *(const bf8 *)&W[(nb * 32 + col) * pw + 16 * s + kh];
  };
  const v4u ones = {kh == 0 ? 0x3F803F80u : 0u, 0u, 0u, 0u};
  const v16f zero = {};
  auto products = [&](int nb, const bf8 *w) {
    const v4u bw = {kh == 0 ? bpk[nb * 32 + col] : 0u, 0u, 0u, 0u};
    v16f acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf8, bw), __builtin_bit_cast(bf8, ones),
                                                       zero, 0, 0, 0);
#pragma unroll
    for (int s = 0; s < KS; s++) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w[s], xv[s], acc, 0, 0, 0);
    return acc;
  };
  auto activations = [&](int nb, const v16f &acc) {
#pragma unroll
    for (int g = 0; g < 4; g++) {
      bf4 v;
#pragma unroll
      for (int q = 0; q < 4; q++) v[q] = (__bf16)cbm_mlp::hswish(acc[4 * g + q]);
      *(bf4 *)&A[col * AP + nb * 32 + 8 * g + kq] = v;
    }
  };
  weights(0, w);
  v16f a0 = products(0, w);
  weights(1, w);
  v16f a1 = products(1, w);
  activations(0, a0);
  weights(2, w);
  a0 = products(2, w);
  activations(1, a1);
  weights(3, w);
  a1 = products(3, w);
  activations(2, a0);
  activations(3, a1);
}

// (a template so that the header's kernel has one definition across the
// translation units that include it)
template <int NWAVES = infer::WAVES>
__global__ void __launch_bounds__(64 * NWAVES) mlp_infer_kernel(StepArgs a, const uint16_t *__restrict__ wt,
                                                               const float *__restrict__ params,
                                                               float *__restrict__ pred, float *__restrict__ xraw) {
  using namespace mlp;
  using namespace infer;
  static_assert(NWAVES == WAVES, "LDS sized for WAVES");
  __shared__ __attribute__((aligned(16))) __bf16 wl[WLDS];
  __shared__ __attribute__((aligned(16))) __bf16 act[ALDS];
  __shared__ __attribute__((aligned(16))) float bl[BLDS];
  __shared__ uint32_t bpk[3 * HID];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, col = lane & 31;
  const int kh = 8 * (lane >> 5), kq = 4 * (lane >> 5);
  __bf16 *A = &act[wave * 32 * AP];
  bool staged = false;
  const int64_t nblk = (a.mp + TILES - 1) / TILES;
  // A block with no leaf-bearing tile in daylight has nothing to predict:
  // the warm start applies to sunlit vegetated tiles only, and the test is
  // the mixed step's own (both shortwave bands, fsd(:,2) = fsd(:,1), above
  // RAD_THRESH in FP32).  The workgroup takes the votes of VCHUNK of its
  // blocks at once (loads issued together), then steps the blocks with work.
  __shared__ int vote[VCHUNK];
  // the class tables (the features' table columns) and the outputs'
  // normalisation, so that no global load but the features' own is waited
  // for between a block's feature loads and its network
  __shared__ double tv[NVEGCLASS * VC_N], ts[NSOILCLASS * SC_N];
  __shared__ float nrm[NNORM];
  for (int i = threadIdx.x; i < NVEGCLASS * VC_N; i += blockDim.x) tv[i] = a.vtab[i];
  for (int i = threadIdx.x; i < NSOILCLASS * SC_N; i += blockDim.x) ts[i] = a.stab[i];
  // (the inputs' standard deviations as reciprocals)
  if (threadIdx.x < NNORM)
    nrm[threadIdx.x] = threadIdx.x >= NIN && threadIdx.x < 2 * NIN ? 1.0f / a.norm[threadIdx.x] : a.norm[threadIdx.x];
  const FeatureSrc src = feature_src(a);
  const int64_t G = gridDim.x;
  const bool feat_lane = threadIdx.x < TILES;
  for (int64_t c0 = blockIdx.x; c0 < nblk; c0 += G * VCHUNK) {
    __syncthreads();  // the previous chunk's blocks are done with vote[]
    if (threadIdx.x < VCHUNK) {
      // the votes the caller's daylight scan already took, or none yet
      const int64_t b = c0 + (int64_t)threadIdx.x * G;
      vote[threadIdx.x] = a.lit_blocks && b < nblk ? (int)((a.lit_blocks[b] >> a.lit_bit) & 1ull) : 0;
    }
    __syncthreads();
    for (int i = threadIdx.x; !a.lit_blocks && i < VCHUNK * TILES; i += blockDim.x) {
      const int64_t b = c0 + (int64_t)(i / TILES) * G, t = b * TILES + i % TILES;
      if (b < nblk && t < a.mp && ld<int>(src.iveg, t) < 14) {
        const float fsd = ld<float>(src.fsd, t + a.foff);
        if (fsd + fsd > 0.001f) vote[i / TILES] = 1;
      }
    }
    __syncthreads();
    // the chunk's first block at or after qb with work (VCHUNK: none); the
    // same for every thread
    auto next_block = [&](int qb) {
      for (; qb < VCHUNK; qb++) {
        if (c0 + (int64_t)qb * G >= nblk) return VCHUNK;
        if (vote[qb] || xraw) return qb;
      }
      return VCHUNK;
    };
    // A block's features are loaded before the previous block's network and
    // finished (the arithmetic, the normalisation, BF16) in its middle, by
    // the threads of waves 0-3 while the waves sharing their SIMDs compute;
    // they wait in registers for the rows to be free.
    FeatureRaw<float> raw;
    bf8 fr[3];
    auto finish = [&](int64_t t) {
      float x[NIN];
      if (t < a.mp) {
        const int iv = raw.iveg - 1, is = raw.isoilm - 1;
        finish_features<float>(a, raw, &tv[(iv < 0 ? 0 : iv) * VC_N], &ts[(is < 0 ? 0 : is) * SC_N], x);
        if (xraw) {
#pragma unroll
          for (int c = 0; c < NIN; c++) xraw[t * NIN + c] = x[c];
        }
#pragma unroll
        for (int c = 0; c < NIN; c++) x[c] = (x[c] - nrm[c]) * nrm[NIN + c];
      } else {
#pragma unroll
        for (int c = 0; c < NIN; c++) x[c] = 0.0f;
      }
#pragma unroll
      for (int j = 0; j < 3; j++)
#pragma unroll
        for (int e = 0; e < 8; e++) fr[j][e] = (__bf16)x[8 * j + e];
    };
    int qb = next_block(0);
    if (qb < VCHUNK && feat_lane) {
      const int64_t t = (c0 + (int64_t)qb * G) * TILES + threadIdx.x;
      if (t < a.mp) load_feature_raw<float>(a, src, t, raw);
      finish(t);
    }
    while (qb < VCHUNK) {
      const int64_t t0 = (c0 + (int64_t)qb * G) * TILES;
      // the previous block's layers are done with the activation rows
      lds_barrier();
      // the weights are staged in LDS once per workgroup (the grid is one
      // workgroup per CU, each stepping every gridDim.x-th block of 256
      // tiles), at its first block with work, so a night step reads none; the
      // barrier after the feature rows below also publishes them
      if (!staged) {
        // 8-byte units of four consecutive inputs, loads issued together
        // before the stores
        constexpr int U = GTOT / 4, PER = (U + 64 * WAVES - 1) / (64 * WAVES);
        uint2 v[PER];
#pragma unroll
        for (int q = 0; q < PER; q++) {
          const int u = threadIdx.x + q * 64 * WAVES;
          if (u < U) v[q] = *(const uint2 *)&wt[4 * u];
        }
#pragma unroll
        for (int q = 0; q < PER; q++) {
          const int u = threadIdx.x + q * 64 * WAVES, e = 4 * u;
          if (u < U) {
            int dst;
            if (e < G2) dst = OFF1 + (e / NIN) * P1 + e % NIN;
            else if (e < G3) dst = OFF2 + ((e - G2) / HID) * PH + (e - G2) % HID;
            else if (e < G4) dst = OFF3 + ((e - G3) / HID) * PH + (e - G3) % HID;
            else dst = OFF4 + ((e - G4) / HID) * PH + (e - G4) % HID;
            *(uint2 *)&wl[dst] = v[q];
          }
        }
        for (int u = threadIdx.x; u < HID * 2; u += blockDim.x)
          *(uint2 *)&wl[OFF1 + (u >> 1) * P1 + NIN + 4 * (u & 1)] = make_uint2(0u, 0u);
        for (int u = threadIdx.x; u < BLDS; u += blockDim.x)
          bl[u] = params[u < 3 * HID ? LAYER_B[u / HID] + u % HID : LAYER_B[3] + u - 3 * HID];
        // the hidden layers' biases as BF16 pairs: hi = b rounded, lo = the
        // rest rounded (b - hi is exact in FP32), so hi + lo keeps 16 bits
        for (int u = threadIdx.x; u < 3 * HID; u += blockDim.x) {
          const float b = params[LAYER_B[u / HID] + u % HID];
          const uint16_t hi = f32_to_bf16(b), lo = f32_to_bf16(b - mlp::bf16_to_f32(hi));
          bpk[u] = (uint32_t)hi | ((uint32_t)lo << 16);
        }
        staged = true;
      }
      // features of tile threadIdx.x as BF16 row threadIdx.x (24 inputs and
      // zeros to the padded K of 32)
      if (feat_lane) {
        __bf16 *row = &act[threadIdx.x * AP];
#pragma unroll
        for (int j = 0; j < 3; j++) *(bf8 *)&row[8 * j] = fr[j];
        *(uint4 *)&row[NIN] = make_uint4(0u, 0u, 0u, 0u);
      }
      // (LDS only: the staged weights and the feature rows)
      lds_barrier();
      const int qn = next_block(qb + 1);
      const int64_t tn = (c0 + (int64_t)qn * G) * TILES + threadIdx.x;
      if (qn < VCHUNK && feat_lane && tn < a.mp) load_feature_raw<float>(a, src, tn, raw);
      // from here on a wave touches only its own 32 activation rows: its LDS
      // operations complete in order, so the layers need no workgroup barrier
      infer_hidden<K1 / 16>(&wl[OFF1], P1, A, &bpk[0], col, kh, kq);
      infer_hidden<HID / 16>(&wl[OFF2], PH, A, &bpk[HID], col, kh, kq);
      if (qn < VCHUNK && feat_lane) finish(tn);
      infer_hidden<HID / 16>(&wl[OFF3], PH, A, &bpk[2 * HID], col, kh, kq);
      // output layer: 8 outputs, rows 8..31 of the weight operand are zero
      {
        // every operand read first, then the eight products on two
        // accumulators (two chains of dependent products instead of one)
        constexpr int KS = HID / 16;
        bf8 xo[KS], wo[KS];
#pragma unroll
        for (int s = 0; s < KS; s++) {
          xo[s] = *(const bf8 *)&A[col * AP + 16 * s + kh];
          if (col < NOUT) wo[s] = *(const bf8 *)&wl[OFF4 + col * PH + 16 * s + kh];
          else
#pragma unroll
            for (int e = 0; e < 8; e++) wo[s][e] = (__bf16)0.0f;
        }
        v16f acc = {}, acc1 = {};
#pragma unroll
        for (int s = 0; s < KS; s += 2) {
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wo[s], xo[s], acc, 0, 0, 0);
          acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wo[s + 1], xo[s + 1], acc1, 0, 0, 0);
        }
        __builtin_amdgcn_sched_group_barrier(0x0100, 2 * KS, 0);  // LDS reads
        __builtin_amdgcn_sched_group_barrier(0x0008, KS, 0);      // matrix products
#pragma unroll
        for (int q = 0; q < 4; q++) acc[q] += acc1[q];
        // lane: tile `col`, outputs kq + 0..3 in registers 0..3
        const int64_t t = t0 + wave * 32 + col;
        if (t < a.mp) {
          float4 o;
          o.x = (acc[0] + bl[3 * HID + kq + 0]) * nrm[2 * NIN + NOUT + kq + 0] + nrm[2 * NIN + kq + 0];
          o.y = (acc[1] + bl[3 * HID + kq + 1]) * nrm[2 * NIN + NOUT + kq + 1] + nrm[2 * NIN + kq + 1];
          o.z = (acc[2] + bl[3 * HID + kq + 2]) * nrm[2 * NIN + NOUT + kq + 2] + nrm[2 * NIN + kq + 2];
          o.w = (acc[3] + bl[3 * HID + kq + 3]) * nrm[2 * NIN + NOUT + kq + 3] + nrm[2 * NIN + kq + 3];
          *(float4 *)&pred[t * NOUT + kq] = o;
        }
      }
      qb = qn;
    }
  }
}


// launch one step over mp tiles on `stream` (256 threads per workgroup)
template <class R, bool FULL, bool TABLES, int MODE = MODE_PLAIN, bool CONS = false>
hipError_t launch_cbm(const StepArgs &a, int64_t mp, hipStream_t stream) {
  // FP64 runs one wave per SIMD: one-wave workgroups let the dispatcher fill
  // SIMDs wave by wave (measured 2-4% faster than four-wave workgroups); the
  // FP32 kernels keep four waves per workgroup, two resident per SIMD
  const int block = sizeof(R) == 8 ? 64 : 256;
  const int64_t grid = (mp + block - 1) / block;
  hipLaunchKernelGGL((cbm_kernel<R, FULL, TABLES, MODE, CONS>), dim3((unsigned)grid), dim3(block), 0, stream, a);
  return hipGetLastError();
}

// defined in cbm_kernels_fp64.hip and cbm_kernels_mixed.hip
hipError_t launch_fp64(bool full, const StepArgs &a, int64_t mp, hipStream_t stream);
hipError_t launch_mixed(bool full, const StepArgs &a, int64_t mp, hipStream_t stream);
// mixed path, resident layout: a step in MODE_CAPTURE or MODE_HYBRID, and the
// surrogate inference over mp tiles (cbm_kernels_mixed.hip)
hipError_t launch_mixed_mode(int mode, const StepArgs &a, int64_t mp, hipStream_t stream);
hipError_t launch_mlp_infer(const StepArgs &a, const uint16_t *wt, const float *params, float *pred,
                            float *xraw, int64_t mp, hipStream_t stream);

}  // namespace cbm
