! cable_gpu_c.F90 -- ISO_C_BINDING interfaces to libcable_gpu.so
! (include/cable_gpu.h), for Fortran hosts of the reference model.
!
! Every function returns 0 on success; cable_gpu_check turns a nonzero code
! into the reference's abort semantics (message, then ERROR STOP 999, as
! cable_abort does in src/util/error_handler/cable_error_handler.F90:48-61).
MODULE cable_gpu_c
USE, INTRINSIC :: ISO_C_BINDING
IMPLICIT NONE
PRIVATE
PUBLIC :: cable_gpu_create, cable_gpu_destroy, cable_gpu_upload, cable_gpu_download, &
          cable_gpu_upload_fields, cable_gpu_download_fields, cable_gpu_field_id, &
          cable_gpu_set_globals, cable_gpu_cbm, cable_gpu_step, cable_gpu_sync, &
          cable_gpu_check, cable_gpu_error_message
! the device-resident run (a driver that owns its time loop)
PUBLIC :: cable_gpu_set_tile_order, cable_gpu_veg_cols, cable_gpu_soil_cols, cable_gpu_table_column, &
          cable_gpu_set_class_tables, cable_gpu_set_forcing, cable_gpu_set_met, cable_gpu_run, &
          cable_gpu_last_run_ms, cable_gpu_consistency_check, cable_gpu_consistency_values, &
          cable_gpu_state_bytes, cable_gpu_save_state, cable_gpu_load_state, cable_gpu_met

! the met variables of a window in the file's units (struct cable_gpu_met)
TYPE, BIND(C) :: cable_gpu_met
  TYPE(C_PTR) :: swdown = C_NULL_PTR, lwdown = C_NULL_PTR, tair = C_NULL_PTR, qair = C_NULL_PTR, &
                 wind = C_NULL_PTR, wind_e = C_NULL_PTR, psurf = C_NULL_PTR, rainf = C_NULL_PTR
  REAL(C_DOUBLE) :: tair_offset = 0.0_C_DOUBLE, psurf_scale = 0.01_C_DOUBLE, rainf_scale = 1.0_C_DOUBLE
  INTEGER(C_INT) :: qair_is_rh = 0
  TYPE(C_PTR) :: elevation = C_NULL_PTR, latitude = C_NULL_PTR
  REAL(C_DOUBLE) :: shod = 0.0_C_DOUBLE
  INTEGER(C_INT) :: sdoy = 1, smoy = 1, syear = 2000
  REAL(C_DOUBLE) :: fixed_co2 = 400.0_C_DOUBLE
END TYPE cable_gpu_met

INTEGER(C_INT), PARAMETER, PUBLIC :: CABLE_PREC_FP64 = 0, CABLE_PREC_MIXED = 1
INTEGER(C_INT), PARAMETER, PUBLIC :: CABLE_HOST_F64 = 0, CABLE_HOST_F32 = 1, CABLE_HOST_I32 = 2

INTERFACE
  INTEGER(C_INT) FUNCTION cable_gpu_create(ctx, mp, precision) BIND(C, NAME='cable_gpu_create')
    IMPORT :: C_PTR, C_INT64_T, C_INT
    TYPE(C_PTR), INTENT(OUT) :: ctx
    INTEGER(C_INT64_T), VALUE :: mp
    INTEGER(C_INT), VALUE :: precision
  END FUNCTION cable_gpu_create

  INTEGER(C_INT) FUNCTION cable_gpu_destroy(ctx) BIND(C, NAME='cable_gpu_destroy')
    IMPORT :: C_PTR, C_INT
    TYPE(C_PTR), VALUE :: ctx
  END FUNCTION cable_gpu_destroy

  INTEGER(C_INT) FUNCTION cable_gpu_upload(ctx, field_id, host, host_type) &
      BIND(C, NAME='cable_gpu_upload')
    IMPORT :: C_PTR, C_INT
    TYPE(C_PTR), VALUE :: ctx, host
    INTEGER(C_INT), VALUE :: field_id, host_type
  END FUNCTION cable_gpu_upload

  INTEGER(C_INT) FUNCTION cable_gpu_download(ctx, field_id, host, host_type) &
      BIND(C, NAME='cable_gpu_download')
    IMPORT :: C_PTR, C_INT
    TYPE(C_PTR), VALUE :: ctx, host
    INTEGER(C_INT), VALUE :: field_id, host_type
  END FUNCTION cable_gpu_download

  INTEGER(C_INT) FUNCTION cable_gpu_field_id(name) BIND(C, NAME='cable_gpu_field_id')
    IMPORT :: C_INT, C_CHAR
    CHARACTER(KIND=C_CHAR), INTENT(IN) :: name(*)
  END FUNCTION cable_gpu_field_id

  INTEGER(C_INT) FUNCTION cable_gpu_upload_fields(ctx, n, field_ids, hosts, host_types) &
      BIND(C, NAME='cable_gpu_upload_fields')
    IMPORT :: C_PTR, C_INT
    TYPE(C_PTR), VALUE :: ctx
    INTEGER(C_INT), VALUE :: n
    INTEGER(C_INT), INTENT(IN) :: field_ids(*), host_types(*)
    TYPE(C_PTR), INTENT(IN) :: hosts(*)
  END FUNCTION cable_gpu_upload_fields

  INTEGER(C_INT) FUNCTION cable_gpu_download_fields(ctx, n, field_ids, hosts, host_types) &
      BIND(C, NAME='cable_gpu_download_fields')
    IMPORT :: C_PTR, C_INT
    TYPE(C_PTR), VALUE :: ctx
    INTEGER(C_INT), VALUE :: n
    INTEGER(C_INT), INTENT(IN) :: field_ids(*), host_types(*)
    TYPE(C_PTR), INTENT(IN) :: hosts(*)
  END FUNCTION cable_gpu_download_fields

  INTEGER(C_INT) FUNCTION cable_gpu_set_globals(ctx, zse, zshh, ratecp, ratecs) &
      BIND(C, NAME='cable_gpu_set_globals')
    IMPORT :: C_PTR, C_INT, C_DOUBLE
    TYPE(C_PTR), VALUE :: ctx
    REAL(C_DOUBLE), INTENT(IN) :: zse(*), zshh(*), ratecp(*), ratecs(*)
  END FUNCTION cable_gpu_set_globals

  INTEGER(C_INT) FUNCTION cable_gpu_cbm(ctx, ktau, dels) BIND(C, NAME='cable_gpu_cbm')
    IMPORT :: C_PTR, C_INT, C_DOUBLE
    TYPE(C_PTR), VALUE :: ctx
    INTEGER(C_INT), VALUE :: ktau
    REAL(C_DOUBLE), VALUE :: dels
  END FUNCTION cable_gpu_cbm

  INTEGER(C_INT) FUNCTION cable_gpu_step(ctx, ktau, dels) BIND(C, NAME='cable_gpu_step')
    IMPORT :: C_PTR, C_INT, C_DOUBLE
    TYPE(C_PTR), VALUE :: ctx
    INTEGER(C_INT), VALUE :: ktau
    REAL(C_DOUBLE), VALUE :: dels
  END FUNCTION cable_gpu_step

  INTEGER(C_INT) FUNCTION cable_gpu_sync(ctx) BIND(C, NAME='cable_gpu_sync')
    IMPORT :: C_PTR, C_INT
    TYPE(C_PTR), VALUE :: ctx
  END FUNCTION cable_gpu_sync

  INTEGER(C_INT) FUNCTION cable_gpu_set_tile_order(ctx, order) BIND(C, NAME='cable_gpu_set_tile_order')
    IMPORT :: C_PTR, C_INT
    TYPE(C_PTR), VALUE :: ctx, order
  END FUNCTION cable_gpu_set_tile_order

  INTEGER(C_INT) FUNCTION cable_gpu_veg_cols() BIND(C, NAME='cable_gpu_veg_cols')
    IMPORT :: C_INT
  END FUNCTION cable_gpu_veg_cols

  INTEGER(C_INT) FUNCTION cable_gpu_soil_cols() BIND(C, NAME='cable_gpu_soil_cols')
    IMPORT :: C_INT
  END FUNCTION cable_gpu_soil_cols

  INTEGER(C_INT) FUNCTION cable_gpu_table_column(table, field_id) BIND(C, NAME='cable_gpu_table_column')
    IMPORT :: C_INT
    INTEGER(C_INT), VALUE :: table, field_id
  END FUNCTION cable_gpu_table_column

  INTEGER(C_INT) FUNCTION cable_gpu_set_class_tables(ctx, veg, soil) BIND(C, NAME='cable_gpu_set_class_tables')
    IMPORT :: C_PTR, C_INT
    TYPE(C_PTR), VALUE :: ctx, veg, soil
  END FUNCTION cable_gpu_set_class_tables

  INTEGER(C_INT) FUNCTION cable_gpu_set_forcing(ctx, nsteps, fsd1, fld, tk, qv, ua, pmb, precip, coszen, &
                                                host_type, ca, doy) BIND(C, NAME='cable_gpu_set_forcing')
    IMPORT :: C_PTR, C_INT, C_DOUBLE
    TYPE(C_PTR), VALUE :: ctx, fsd1, fld, tk, qv, ua, pmb, precip, coszen
    INTEGER(C_INT), VALUE :: nsteps, host_type
    REAL(C_DOUBLE), VALUE :: ca, doy
  END FUNCTION cable_gpu_set_forcing

  INTEGER(C_INT) FUNCTION cable_gpu_set_met(ctx, k_first, nsteps, dels, met) BIND(C, NAME='cable_gpu_set_met')
    IMPORT :: C_PTR, C_INT, C_DOUBLE, cable_gpu_met
    TYPE(C_PTR), VALUE :: ctx
    INTEGER(C_INT), VALUE :: k_first, nsteps
    REAL(C_DOUBLE), VALUE :: dels
    TYPE(cable_gpu_met), INTENT(IN) :: met
  END FUNCTION cable_gpu_set_met

  INTEGER(C_INT) FUNCTION cable_gpu_run(ctx, k0, nsteps, dels) BIND(C, NAME='cable_gpu_run')
    IMPORT :: C_PTR, C_INT, C_DOUBLE
    TYPE(C_PTR), VALUE :: ctx
    INTEGER(C_INT), VALUE :: k0, nsteps
    REAL(C_DOUBLE), VALUE :: dels
  END FUNCTION cable_gpu_run

  INTEGER(C_INT) FUNCTION cable_gpu_last_run_ms(ctx, ms) BIND(C, NAME='cable_gpu_last_run_ms')
    IMPORT :: C_PTR, C_INT, C_DOUBLE
    TYPE(C_PTR), VALUE :: ctx
    REAL(C_DOUBLE), INTENT(OUT) :: ms
  END FUNCTION cable_gpu_last_run_ms

  INTEGER(C_INT) FUNCTION cable_gpu_consistency_check(ctx, enable) BIND(C, NAME='cable_gpu_consistency_check')
    IMPORT :: C_PTR, C_INT
    TYPE(C_PTR), VALUE :: ctx
    INTEGER(C_INT), VALUE :: enable
  END FUNCTION cable_gpu_consistency_check

  INTEGER(C_INT) FUNCTION cable_gpu_consistency_values(ctx, out) BIND(C, NAME='cable_gpu_consistency_values')
    IMPORT :: C_PTR, C_INT, C_DOUBLE
    TYPE(C_PTR), VALUE :: ctx
    REAL(C_DOUBLE), INTENT(OUT) :: out(6)
  END FUNCTION cable_gpu_consistency_values

  INTEGER(C_INT64_T) FUNCTION cable_gpu_state_bytes(ctx) BIND(C, NAME='cable_gpu_state_bytes')
    IMPORT :: C_PTR, C_INT64_T
    TYPE(C_PTR), VALUE :: ctx
  END FUNCTION cable_gpu_state_bytes

  INTEGER(C_INT) FUNCTION cable_gpu_save_state(ctx, buf, bytes) BIND(C, NAME='cable_gpu_save_state')
    IMPORT :: C_PTR, C_INT, C_INT64_T
    TYPE(C_PTR), VALUE :: ctx, buf
    INTEGER(C_INT64_T), VALUE :: bytes
  END FUNCTION cable_gpu_save_state

  INTEGER(C_INT) FUNCTION cable_gpu_load_state(ctx, buf, bytes, continue_run) BIND(C, NAME='cable_gpu_load_state')
    IMPORT :: C_PTR, C_INT, C_INT64_T
    TYPE(C_PTR), VALUE :: ctx, buf
    INTEGER(C_INT64_T), VALUE :: bytes
    INTEGER(C_INT), VALUE :: continue_run
  END FUNCTION cable_gpu_load_state

  TYPE(C_PTR) FUNCTION c_last_error() BIND(C, NAME='cable_gpu_last_error')
    IMPORT :: C_PTR
  END FUNCTION c_last_error

  INTEGER(C_SIZE_T) FUNCTION c_strlen(s) BIND(C, NAME='strlen')
    IMPORT :: C_PTR, C_SIZE_T
    TYPE(C_PTR), VALUE :: s
  END FUNCTION c_strlen
END INTERFACE

CONTAINS

FUNCTION cable_gpu_error_message() RESULT(msg)
  CHARACTER(LEN=:), ALLOCATABLE :: msg
  TYPE(C_PTR) :: p
  CHARACTER(KIND=C_CHAR), POINTER :: s(:)
  INTEGER :: n, i
  p = c_last_error()
  IF (.NOT. C_ASSOCIATED(p)) THEN
    msg = ''
    RETURN
  END IF
  n = INT(c_strlen(p))
  CALL C_F_POINTER(p, s, [n])
  ALLOCATE(CHARACTER(LEN=n) :: msg)
  DO i = 1, n
    msg(i:i) = s(i)
  END DO
END FUNCTION cable_gpu_error_message

SUBROUTINE cable_gpu_check(rc)
  INTEGER(C_INT), INTENT(IN) :: rc
  IF (rc /= 0) THEN
    WRITE(*, '(A)') 'libcable_gpu: ' // cable_gpu_error_message()
    ERROR STOP 999
  END IF
END SUBROUTINE cable_gpu_check

END MODULE cable_gpu_c
