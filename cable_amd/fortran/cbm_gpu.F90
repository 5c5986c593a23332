! cbm_gpu.F90 -- the reference's cbm() with its physics on the GPU.
!
! Drop-in for MODULE cable_cbm_module of src/offline/cbl_model_driver_offline.F90
! (:29-231): the same module name, the same procedure and the same dummy
! arguments, so a reference build links this file in place of that one and
! its drivers (cable_serial.F90:594-595, cable_mpiworker.F90:503-504) call
! the GPU without change.
!
! Each call hands every per-tile component the step reads to the device
! (parameters, forcing, state and the diagnostics that keep values on tiles a
! step does not touch), runs one timestep of the reference's cbm -- ruff_resist,
! define_air, init_radiation/albedo, define_canopy, soil_snow, snow_aging,
! carbon and the flux totals -- and copies every component it writes back
! into the derived types.  The caller's c1, rhoch and xk scratch arrays are
! not used: the kernel keeps those coefficients in registers.
!
! Device precision follows the host build unless cable_gpu_precision is set
! before the first call: a build with 8-byte default REAL (-fdefault-real-8)
! gets the FP64 path; the shipped build (default REAL FP32, REAL(r_2) FP64)
! gets the mixed path, which stores the fields with that split on the device
! and computes the canopy in FP32, the soil state and solves in FP64.  FP64
! arithmetic on FP32-stored fields is not offered: the reference compares
! some carried values exactly (e.g. ssnow%cls against 1.1335 in
! src/science/soilsnow/cbl_snowAccum.F90), and those tests flip when an FP64
! value is rounded to FP32 between steps.
MODULE cable_cbm_module
USE, INTRINSIC :: ISO_C_BINDING
USE cable_gpu_c
IMPLICIT NONE
PRIVATE
PUBLIC :: cbm, cbm_gpu_finalise
INTEGER(C_INT), PUBLIC, SAVE :: cable_gpu_precision = -1   ! -1: from KIND(1.0)

TYPE(C_PTR), SAVE :: ctx = C_NULL_PTR

CONTAINS

SUBROUTINE cbm(ktau, dels, air, bgc, canopy, met, bal, rad, rough, soil, &
               ssnow, sum_flux, veg, climate, xk, c1, rhoch)
  USE cable_def_types_mod
  USE cable_climate_type_mod, ONLY: climate_type
  USE cable_gpu_fields, ONLY: cable_gpu_put_fields, cable_gpu_get_fields

  TYPE (air_type),       INTENT(INOUT) :: air
  TYPE (bgc_pool_type),  INTENT(INOUT) :: bgc
  TYPE (canopy_type),    INTENT(INOUT) :: canopy
  TYPE (met_type),       INTENT(INOUT) :: met
  TYPE (balances_type),  INTENT(INOUT) :: bal
  TYPE (radiation_type), INTENT(INOUT) :: rad
  TYPE (roughness_type), INTENT(INOUT) :: rough
  TYPE (soil_snow_type), INTENT(INOUT) :: ssnow
  TYPE (sum_flux_type),  INTENT(INOUT) :: sum_flux
  TYPE (climate_type),   INTENT(IN)    :: climate
  TYPE (soil_parameter_type), INTENT(INOUT) :: soil
  TYPE (veg_parameter_type),  INTENT(INOUT) :: veg
  REAL, INTENT(IN) :: dels
  INTEGER, INTENT(IN) :: ktau
  REAL :: c1(mp,nrb)
  REAL :: rhoch(mp,nrb)
  REAL :: xk(mp,nrb)

  REAL(C_DOUBLE) :: zse(ms), zshh(ms+1), ratecp(ncp), ratecs(ncs)

  IF (.NOT. C_ASSOCIATED(ctx)) THEN
    IF (cable_gpu_precision < 0) &
      cable_gpu_precision = MERGE(CABLE_PREC_FP64, CABLE_PREC_MIXED, KIND(1.0) == 8)
    CALL cable_gpu_check(cable_gpu_create(ctx, INT(mp, C_INT64_T), cable_gpu_precision))
    zse = soil%zse
    zshh = soil%zshh
    ratecp = bgc%ratecp
    ratecs = bgc%ratecs
    CALL cable_gpu_check(cable_gpu_set_globals(ctx, zse, zshh, ratecp, ratecs))
  END IF
  CALL cable_gpu_put_fields(ctx, air, bgc, canopy, met, bal, rad, rough, soil, ssnow, veg)
  CALL cable_gpu_check(cable_gpu_cbm(ctx, INT(ktau, C_INT), REAL(dels, C_DOUBLE)))
  CALL cable_gpu_get_fields(ctx, air, bgc, canopy, met, bal, rad, rough, soil, ssnow, veg)
END SUBROUTINE cbm

! release the device copy (optional; the process end releases it too)
SUBROUTINE cbm_gpu_finalise()
  IF (C_ASSOCIATED(ctx)) THEN
    CALL cable_gpu_check(cable_gpu_destroy(ctx))
    ctx = C_NULL_PTR
  END IF
END SUBROUTINE cbm_gpu_finalise

END MODULE cable_cbm_module
