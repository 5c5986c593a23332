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
!
! Transfers.  By default every call uploads every field the step reads and
! downloads every field it writes, so the host's derived types are the
! state, as the reference's are.  With cable_gpu_transfers = 'resident' (or the
! environment variable CABLE_GPU_SYNC=resident) the device holds the state:
! after the first call only the components the offline driver writes between
! calls go up (the forcing, canopy%oldcansto, veg%vlai;
! src/offline/cable_serial.F90:572-575), and a host that reads back only some
! fields lists their ids in cable_gpu_download_only (or their names, comma
! separated, in CABLE_GPU_DOWNLOAD); canopy%cansto always comes back, since
! the driver copies it into oldcansto.  Each direction is one batched
! transfer through a pinned staging buffer (cable_gpu_upload_fields).
!
! The device implements the shipped configuration (src/offline/cable.nml):
! icycle = 0 and the cable_user switches the path's science branches on at
! their shipped values (soil_struc 'default', fwsoil_switch 'standard',
! gs_switch 'leuning', ssnow_potev 'HDM', diag_soil_resp 'ON', every logical
! option off, snmin = 1, no hydraulic redistribution).  The first call checks
! them before it touches the device and stops with a message (ERROR STOP 999,
! as cable_abort does) if the run asks for anything else.
MODULE cable_cbm_module
USE, INTRINSIC :: ISO_C_BINDING
USE cable_gpu_c
IMPLICIT NONE
PRIVATE
PUBLIC :: cbm, cbm_gpu_finalise
INTEGER(C_INT), PUBLIC, SAVE :: cable_gpu_precision = -1   ! -1: from KIND(1.0)
CHARACTER(LEN=16), PUBLIC, SAVE :: cable_gpu_transfers = ''      ! 'all' or 'resident'; blank: CABLE_GPU_SYNC or 'all'
INTEGER(C_INT), ALLOCATABLE, PUBLIC, SAVE :: cable_gpu_download_only(:)   ! field ids (resident only)
LOGICAL, SAVE :: resident = .FALSE.

TYPE(C_PTR), SAVE :: ctx = C_NULL_PTR

CONTAINS

SUBROUTINE cbm(ktau, dels, air, bgc, canopy, met, bal, rad, rough, soil, &
               ssnow, sum_flux, veg, climate, xk, c1, rhoch)
  USE cable_def_types_mod
  USE cable_climate_type_mod, ONLY: climate_type
  USE cable_gpu_fields, ONLY: cable_gpu_put_fields, cable_gpu_get_fields, CABLE_F_canopy_cansto

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
  LOGICAL :: first

  first = .NOT. C_ASSOCIATED(ctx)
  IF (first) THEN
    CALL check_configuration()
    CALL transfer_options()
    IF (cable_gpu_precision < 0) &
      cable_gpu_precision = MERGE(CABLE_PREC_FP64, CABLE_PREC_MIXED, KIND(1.0) == 8)
    CALL cable_gpu_check(cable_gpu_create(ctx, INT(mp, C_INT64_T), cable_gpu_precision))
    zse = soil%zse
    zshh = soil%zshh
    ratecp = bgc%ratecp
    ratecs = bgc%ratecs
    CALL cable_gpu_check(cable_gpu_set_globals(ctx, zse, zshh, ratecp, ratecs))
  END IF
  CALL cable_gpu_put_fields(ctx, air, bgc, canopy, met, bal, rad, rough, soil, ssnow, veg, &
                            MERGE(1, 0, resident .AND. .NOT. first))
  CALL cable_gpu_check(cable_gpu_cbm(ctx, INT(ktau, C_INT), REAL(dels, C_DOUBLE)))
  IF (ALLOCATED(cable_gpu_download_only)) THEN
    CALL cable_gpu_get_fields(ctx, air, bgc, canopy, met, bal, rad, rough, soil, ssnow, veg, &
                              [cable_gpu_download_only, CABLE_F_canopy_cansto])
  ELSE
    CALL cable_gpu_get_fields(ctx, air, bgc, canopy, met, bal, rad, rough, soil, ssnow, veg)
  END IF
END SUBROUTINE cbm

! cable_gpu_transfers and the download list, from the environment when the host
! has not set them
SUBROUTINE transfer_options()
  CHARACTER(LEN=4096) :: names
  INTEGER :: st, i, j, k
  IF (LEN_TRIM(cable_gpu_transfers) == 0) THEN
    CALL GET_ENVIRONMENT_VARIABLE('CABLE_GPU_SYNC', cable_gpu_transfers, STATUS=st)
    IF (st /= 0) cable_gpu_transfers = 'all'
  END IF
  CALL need(TRIM(cable_gpu_transfers) == 'all' .OR. TRIM(cable_gpu_transfers) == 'resident', &
            "cable_gpu_transfers = 'all' or 'resident'")
  resident = TRIM(cable_gpu_transfers) == 'resident'
  IF (.NOT. ALLOCATED(cable_gpu_download_only)) THEN
    CALL GET_ENVIRONMENT_VARIABLE('CABLE_GPU_DOWNLOAD', names, STATUS=st)
    IF (st == 0 .AND. LEN_TRIM(names) > 0) THEN
      ALLOCATE(cable_gpu_download_only(0))
      i = 1
      DO WHILE (i <= LEN_TRIM(names))
        j = INDEX(names(i:), ',')
        j = MERGE(LEN_TRIM(names), i + j - 2, j == 0)
        k = cable_gpu_field_id(TRIM(ADJUSTL(names(i:j))) // C_NULL_CHAR)
        CALL need(k >= 0, 'a field name in CABLE_GPU_DOWNLOAD: ' // TRIM(ADJUSTL(names(i:j))))
        cable_gpu_download_only = [cable_gpu_download_only, INT(k, C_INT)]
        i = j + 2
      END DO
    END IF
  END IF
  CALL need(resident .OR. .NOT. ALLOCATED(cable_gpu_download_only), &
            "cable_gpu_transfers = 'resident' for a download list (each call would upload the rest from the host)")
END SUBROUTINE transfer_options

! The run's configuration against the one the device implements; stops on
! the first difference (before any device call).
SUBROUTINE check_configuration()
  USE cable_runtime_opts_mod, ONLY: cable_user, snmin
  USE cable_common_module, ONLY: redistrb
  USE casadimension, ONLY: icycle
  CALL need(icycle == 0, 'icycle = 0 (no CASA-CNP)')
  CALL need(TRIM(cable_user%soil_struc) == 'default', "cable_user%soil_struc = 'default'")
  CALL need(TRIM(cable_user%fwsoil_switch) == 'standard', "cable_user%fwsoil_switch = 'standard'")
  CALL need(TRIM(cable_user%gs_switch) == 'leuning', "cable_user%gs_switch = 'leuning'")
  CALL need(TRIM(cable_user%ssnow_potev) == 'HDM', "cable_user%ssnow_potev = 'HDM'")
  CALL need(TRIM(cable_user%diag_soil_resp) == 'ON', "cable_user%diag_soil_resp = 'ON'")
  CALL need(.NOT. cable_user%l_rev_corr, 'cable_user%l_rev_corr off')
  CALL need(.NOT. cable_user%l_new_roughness_soil, 'cable_user%l_new_roughness_soil off')
  CALL need(.NOT. cable_user%l_new_runoff_speed, 'cable_user%l_new_runoff_speed off')
  CALL need(.NOT. cable_user%l_new_reduce_soilevp, 'cable_user%l_new_reduce_soilevp off')
  CALL need(.NOT. cable_user%soil_thermal_fix, 'cable_user%soil_thermal_fix off')
  CALL need(.NOT. cable_user%or_evap, 'cable_user%or_evap off')
  CALL need(.NOT. cable_user%gw_model, 'cable_user%gw_model off')
  CALL need(.NOT. cable_user%litter, 'cable_user%litter off')
  CALL need(.NOT. cable_user%call_climate, 'cable_user%call_climate off')
  CALL need(.NOT. cable_user%soilsnow_init_spec, 'cable_user%soilsnow_init_spec off')
  CALL need(.NOT. redistrb, 'redistrb off (no hydraulic redistribution)')
  CALL need(snmin == 1.0, 'snmin = 1')
END SUBROUTINE check_configuration

SUBROUTINE need(ok, what)
  LOGICAL, INTENT(IN) :: ok
  CHARACTER(LEN=*), INTENT(IN) :: what
  IF (.NOT. ok) THEN
    WRITE(*, '(A)') 'cbm (GPU): the run asks for a configuration the device does not implement; it needs ' // what
    ERROR STOP 999
  END IF
END SUBROUTINE need

! release the device copy (optional; the process end releases it too)
SUBROUTINE cbm_gpu_finalise()
  IF (C_ASSOCIATED(ctx)) THEN
    CALL cable_gpu_check(cable_gpu_destroy(ctx))
    ctx = C_NULL_PTR
  END IF
END SUBROUTINE cbm_gpu_finalise

END MODULE cable_cbm_module
