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
          cable_gpu_set_globals, cable_gpu_cbm, cable_gpu_step, cable_gpu_sync, &
          cable_gpu_check, cable_gpu_error_message

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
