! Driver that runs the reference cbm() on inputs written by the test suite.
!
! Test infrastructure only: this program is linked against objects compiled
! from the reference sources by oracle/Makefile and is used to produce the
! golden vectors the CPU restatement (oracle/cbm_oracle.c) and the GPU kernels
! are checked against.  It is never part of the shipped library.
!
! It reproduces the offline serial driver's contract around the cbm call
! (src/offline/cable_serial.F90:437-442 before the time loop, :572-575 before
! each call, :594-605 around it, :713-715 and :725-729 after it) and the
! cable.nml switches of the shipped configuration (src/offline/cable.nml).
!
! Protocol (stream files, little endian):
!   input : int64 mp, then records  int32 op, ...
!           op 1 SET   : char(40) name, int64 n, real(8) data(n)
!           op 2 BEGIN : (the pre-loop initialisation of cable_serial.F90:437-442)
!           op 3 STEP  : int32 ktau, real(8) dels
!           op 4 GET   : char(40) name   -> output record: char(40) name, int64 n, data
!           op 5 NAMES : write the component list to stdout
!           op 7 SWITCH: char(40) switch, char(40) value -- a configuration
!                        switch away from the shipped one (a cable_user
!                        character switch, a cable_user logical 'T'/'F',
!                        or icycle)
!           op 0 END
PROGRAM cable_ref_driver
USE cable_def_types_mod
USE cable_common_module, ONLY: cable_user, cable_runtime, snmin, wiltParam, satuParam, ktau_gl
USE cable_cbm_module, ONLY: cbm
USE cable_climate_type_mod, ONLY: climate_type
USE cable_phys_constants_mod, ONLY: CSBOLTZ => SBOLTZ, CEMLEAF => EMLEAF, CEMSOIL => EMSOIL
USE casadimension, ONLY: icycle
USE casavariable, ONLY: casa_flux
USE sumcflux_mod, ONLY: sumcflux
USE ref_fields
IMPLICIT NONE
CHARACTER(LEN=512) :: inpath, outpath
INTEGER :: uin, uout, op, ios
INTEGER(8) :: n, mp8
INTEGER(4) :: ktau4
REAL(8) :: dels8
REAL :: dels
CHARACTER(LEN=40) :: name, value
REAL(8), ALLOCATABLE :: buf(:)
TYPE (climate_type) :: climate
TYPE (casa_flux) :: casaflux   ! not referenced with icycle = 0
REAL, ALLOCATABLE :: c1(:,:), rhoch(:,:), xk(:,:)
INTEGER(8) :: clk0, clk1, clkrate
REAL(8) :: step_seconds = 0.0_8

CALL GET_COMMAND_ARGUMENT(1, inpath)
CALL GET_COMMAND_ARGUMENT(2, outpath)
OPEN(NEWUNIT=uin, FILE=TRIM(inpath), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
OPEN(NEWUNIT=uout, FILE=TRIM(outpath), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
READ(uin) mp8
mp = INT(mp8)
mvtype = 17
mstype = 9

! configuration of the shipped cable.nml (src/offline/cable.nml:53-72)
cable_runtime%offline = .TRUE.
cable_user%fwsoil_switch = 'standard'
cable_user%gs_switch = 'leuning'
cable_user%diag_soil_resp = 'ON '
cable_user%leaf_respiration = 'ON '
cable_user%ssnow_potev = 'HDM'
cable_user%soil_struc = 'default'
cable_user%consistency_check = .TRUE.
snmin = 1.0
wiltParam = 0.5
satuParam = 0.8
icycle = 0

CALL alloc_cbm_var(air, mp)
CALL alloc_cbm_var(bgc, mp)
CALL alloc_cbm_var(canopy, mp)
CALL alloc_cbm_var(met, mp)
CALL alloc_cbm_var(bal, mp)
CALL alloc_cbm_var(rad, mp)
CALL alloc_cbm_var(rough, mp)
CALL alloc_cbm_var(soil, mp)
CALL alloc_cbm_var(ssnow, mp)
CALL alloc_cbm_var(sum_flux, mp)
CALL alloc_cbm_var(veg, mp)
CALL alloc_cbm_var(climate, mp)
CALL zero_all()
ALLOCATE(c1(mp,nrb), rhoch(mp,nrb), xk(mp,nrb))
c1 = 0.0; rhoch = 0.0; xk = 0.0

DO
  READ(uin, IOSTAT=ios) op
  IF (ios /= 0) EXIT
  SELECT CASE (op)
  CASE (0)
    EXIT
  CASE (1)
    READ(uin) name, n
    IF (ALLOCATED(buf)) DEALLOCATE(buf)
    ALLOCATE(buf(n))
    READ(uin) buf
    IF (field_size(name) /= n) THEN
      WRITE(0,*) 'size mismatch for ', TRIM(name), field_size(name), n
      ERROR STOP 2
    END IF
    CALL field_set(name, buf, n)
  CASE (2)
    ! src/offline/cable_serial.F90:437-442
    ssnow%otss_0 = ssnow%tgg(:,1)
    ssnow%otss = ssnow%tgg(:,1)
    ssnow%tss = ssnow%tgg(:,1)
    canopy%fes_cor = 0.
    canopy%fhs_cor = 0.
    met%ofsd = 0.1
  CASE (3)
    READ(uin) ktau4, dels8
    CALL SYSTEM_CLOCK(clk0, clkrate)
    dels = REAL(dels8)
    ktau_gl = ktau4
    ! src/offline/cable_serial.F90:572-575
    met%ofsd = met%fsd(:,1) + met%fsd(:,2)
    canopy%oldcansto = canopy%cansto
    WHERE (veg%iveg(:) .GE. 14) veg%vlai = 0.
    CALL cbm(ktau4, dels, air, bgc, canopy, met, bal, rad, rough, soil, ssnow, &
             sum_flux, veg, climate, xk, c1, rhoch)
    ! src/offline/cable_serial.F90:602-605
    ssnow%smelt = ssnow%smelt*dels
    ssnow%rnof1 = ssnow%rnof1*dels
    ssnow%rnof2 = ssnow%rnof2*dels
    ssnow%runoff = ssnow%runoff*dels
    ! src/offline/cable_serial.F90:713-715 (kstart = 1)
    CALL sumcflux(ktau4, 1, ktau4, dels, bgc, canopy, soil, ssnow, sum_flux, veg, &
                  met, casaflux, .FALSE.)
    ! src/offline/cable_serial.F90:725-729 (check%mass_bal, check%energy_bal)
    CALL drv_mass_balance(dels, ktau4)
    CALL drv_energy_balance()
    CALL SYSTEM_CLOCK(clk1)
    step_seconds = step_seconds + REAL(clk1 - clk0, 8) / REAL(clkrate, 8)
  CASE (4)
    READ(uin) name
    n = field_size(name)
    IF (n < 0) THEN
      WRITE(0,*) 'unknown field ', TRIM(name)
      ERROR STOP 3
    END IF
    IF (ALLOCATED(buf)) DEALLOCATE(buf)
    ALLOCATE(buf(n))
    CALL field_get(name, buf, n)
    WRITE(uout) name, n, buf
  CASE (5)
    CALL field_names(6)
  CASE (7)
    READ(uin) name, value
    CALL set_switch(name, value)
  CASE (6)
    ! wall time spent in the step operations (cbm and the driver contract
    ! around it), excluding file input and output
    name = 'drv%step_seconds'
    WRITE(uout) name, 1_8, step_seconds
  CASE DEFAULT
    WRITE(0,*) 'bad op ', op
    ERROR STOP 4
  END SELECT
END DO
CLOSE(uin)
CLOSE(uout)

CONTAINS

SUBROUTINE set_switch(sw, val)
  CHARACTER(LEN=*), INTENT(IN) :: sw, val
  LOGICAL :: on
  on = TRIM(val) == 'T'
  SELECT CASE (TRIM(sw))
  CASE ('gs_switch'); cable_user%gs_switch = TRIM(val)
  CASE ('fwsoil_switch'); cable_user%fwsoil_switch = TRIM(val)
  CASE ('ssnow_potev'); cable_user%ssnow_potev = TRIM(val)
  CASE ('soil_struc'); cable_user%soil_struc = TRIM(val)
  CASE ('diag_soil_resp'); cable_user%diag_soil_resp = TRIM(val)
  CASE ('or_evap'); cable_user%or_evap = on
  CASE ('litter'); cable_user%litter = on
  CASE ('gw_model'); cable_user%gw_model = on
  CASE ('icycle'); READ(val, *) icycle
  CASE DEFAULT
    WRITE(0,*) 'unknown switch ', TRIM(sw)
    ERROR STOP 5
  END SELECT
END SUBROUTINE set_switch

! The balance accumulators of src/offline/cable_checks.F90 cannot be compiled
! here (the module USEs cable_abort_module, which needs netCDF), so the driver
! states them itself, statement for statement.
SUBROUTINE drv_mass_balance(dels, ktau)
  ! src/offline/cable_checks.F90:472-551 (soil_struc = 'default' branch)
  REAL, INTENT(IN) :: dels
  INTEGER(4), INTENT(IN) :: ktau
  REAL(r_2), DIMENSION(:), POINTER, SAVE :: owb
  REAL(r_2), DIMENSION(mp) :: delwb
  IF (ktau == 1) THEN
    ALLOCATE(owb(mp))
    owb = ssnow%wbtot
  END IF
  delwb = ssnow%wbtot - owb
  owb = ssnow%wbtot
  bal%wbal = REAL(met%precip - canopy%delwc - ssnow%snowd + ssnow%osnowd        &
       - ssnow%runoff - (canopy%fevw + canopy%fevc                              &
       + canopy%fes/ssnow%cls)*dels/air%rlam - delwb - ssnow%qrecharge)
  IF (ktau == 1) THEN
    bal%wbal_tot = 0.; bal%precip_tot = 0.
    bal%rnoff_tot = 0.; bal%evap_tot = 0.
  END IF
  IF (ktau > 10) THEN
    bal%wbal_tot = bal%wbal_tot + bal%wbal
    bal%precip_tot = bal%precip_tot + met%precip
    bal%rnoff_tot = bal%rnoff_tot + ssnow%rnof1 + ssnow%rnof2
    bal%evap_tot = bal%evap_tot + (canopy%fev + canopy%fes/ssnow%cls)*dels/air%rlam
  END IF
END SUBROUTINE drv_mass_balance

SUBROUTINE drv_energy_balance()
  ! src/offline/cable_checks.F90:565-618 (L_REV_CORR = .FALSE.)
  bal%Radbal = met%fsd(:,1) + met%fsd(:,2) + met%fld - rad%albedo(:,1)*met%fsd(:,1) &
       - rad%albedo(:,2)*met%fsd(:,2)                                                &
       - (CEMSOIL*CSBOLTZ*rad%transd*ssnow%otss**4)                                  &
       - (CEMLEAF*CSBOLTZ*(1-rad%transd)*canopy%tv**4)                               &
       - canopy%fnv - canopy%fns
  bal%EbalSoil = canopy%fns - canopy%fes - canopy%fhs - canopy%ga
  bal%Ebalveg = canopy%fnv - canopy%fev - canopy%fhv
  bal%Ebal = SUM(rad%qcan(:,:,1),2) + SUM(rad%qcan(:,:,2),2) + rad%qssabs          &
       + met%fld - CSBOLTZ*CEMLEAF*canopy%tv**4*(1-rad%transd)                       &
       - rad%flws*rad%transd                                                         &
       - canopy%fev - canopy%fes                                                     &
       - canopy%fh - canopy%ga
  bal%ebal_tot = bal%ebal_tot + bal%ebal
  bal%RadbalSum = bal%RadbalSum + bal%Radbal
END SUBROUTINE drv_energy_balance

END PROGRAM cable_ref_driver
