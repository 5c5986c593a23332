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

namespace cbm {

enum { ROLE_PARAM = 0, ROLE_FORC = 1, ROLE_STATE = 2, ROLE_DIAG = 3 };


template <class T>
__device__ __forceinline__ T ld(const void *p, int64_t i) {
  return static_cast<const T *>(p)[i];
}
template <class T, class V>
__device__ __forceinline__ void st(void *p, int64_t i, V v) {
  static_cast<T *>(p)[i] = static_cast<T>(v);
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
template <class R, bool FULL, bool TABLES>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 2 : 1)))
cbm_kernel(StepArgs a) {
  __shared__ double lds_v[TABLES ? NVEGCLASS * VC_N : 1];
  __shared__ double lds_s[TABLES ? NSOILCLASS * SC_N : 1];
  if constexpr (TABLES) {
    for (int i = threadIdx.x; i < NVEGCLASS * VC_N; i += blockDim.x) lds_v[i] = a.vtab[i];
    for (int i = threadIdx.x; i < NSOILCLASS * SC_N; i += blockDim.x) lds_s[i] = a.stab[i];
    __syncthreads();
  }
  const int64_t mp = a.mp;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int passes = 0;
  if (t < mp) {
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
      LDS_(ssnow_ssdnn, 1) LDS_(ssnow_tss, 1) LDS_(ssnow_smass, 3) LDS_(ssnow_ssdn, 3)
      LDS_(ssnow_tgg, 6) LDS_(ssnow_tggsn, 3) LDS_(canopy_cansto, 1) LDS_(bgc_cplant, 3)
      LDS_(bgc_csoil, 2)
      LDD_(ssnow_wb, 6) LDD_(ssnow_wbice, 6) LDD_(drv_owb, 1)
      LDD_(bal_ebal_tot, 1) LDD_(bal_wbal_tot, 1) LDD_(bal_precip_tot, 1) LDD_(bal_rnoff_tot, 1)
      LDD_(bal_evap_tot, 1) LDD_(bal_radbalsum, 1)
      T.ssnow_gammzz[0] = ld<double>(F[CABLE_F_ssnow_gammzz], t);
      // soil liquid water as the previous step left it (wb - wbice), before
      // this step's lake fix-up
#pragma unroll
      for (int k = 0; k < MS; k++) T.ssnow_wbliq[k] = T.ssnow_wb[k] - T.ssnow_wbice[k];
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
    const double *vrow = TABLES ? lds_v + (iv < 0 ? 0 : iv) * VC_N : nullptr;
    const double *srow = TABLES ? lds_s + (is < 0 ? 0 : is) * SC_N : nullptr;
    Stepper<R, FULL, TABLES> S(T, a, vrow, srow);
    S.run();
    passes = S.npass;
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
      // the eight headline fluxes
      STS_(canopy_rnet, 1) STS_(canopy_fe, 1) STS_(canopy_fh, 1) STS_(canopy_fev, 1)
      STD_(canopy_fes, 1) STS_(canopy_fgpp, 1) STS_(ssnow_runoff, 1) STS_(rad_trad, 1)
#undef STS_
#undef STD_
    }
  }
  if (a.npass) {
    // wave-level sum, one atomic per wave
    for (int o = 32; o > 0; o >>= 1) passes += __shfl_xor(passes, o, 64);
    if ((threadIdx.x & 63) == 0 && passes) atomicAdd(a.npass, (unsigned long long)passes);
  }
}


// launch one step over mp tiles on `stream` (256 threads per workgroup)
template <class R, bool FULL, bool TABLES>
hipError_t launch_cbm(const StepArgs &a, int64_t mp, hipStream_t stream) {
  const int block = 256;
  const int64_t grid = (mp + block - 1) / block;
  hipLaunchKernelGGL((cbm_kernel<R, FULL, TABLES>), dim3((unsigned)grid), dim3(block), 0, stream, a);
  return hipGetLastError();
}

// defined in cbm_kernels_fp64.hip and cbm_kernels_mixed.hip
hipError_t launch_fp64(bool full, const StepArgs &a, int64_t mp, hipStream_t stream);
hipError_t launch_mixed(bool full, const StepArgs &a, int64_t mp, hipStream_t stream);

}  // namespace cbm
