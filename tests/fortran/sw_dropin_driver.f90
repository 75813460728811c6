! sw_dropin_driver.f90 -- test driver for the reference-named shallow-water
! drop-in (self_amd/fortran/sw_dropin.f90): the run sequence of the
! reference's shallow-water drivers with its module, type and procedure
! names -- Build from shallowwater.params and the pickup at iterInit, one
! GlobalTimeDerivative and a Tecplot dump, nTimeSteps RK3 steps with a dump
! every dumpFreq steps, the final pickup, the mesh file, Trash.
PROGRAM sw_dropin_driver
  USE ModelPrecision
  USE QuadMeshClass
  USE SWParamsClass
  USE ConservativeShallowWaterClass
  IMPLICIT NONE
  TYPE(ShallowWater) :: sw
  INTEGER :: first, nSteps, every, it
  REAL(prec) :: t, dt
  CHARACTER(LEN=10) :: tag

  CALL sw % Build()
  first = sw % params % iterInit
  nSteps = sw % params % nTimeSteps
  every = sw % params % dumpFreq
  dt = sw % params % dt
  WRITE(tag, '(I10.10)') first
  CALL sw % GlobalTimeDerivative(ZERO)
  CALL sw % WriteTecplot('ShallowWater.'//tag)
  DO it = first, first + nSteps - 1
    t = REAL(it, prec) * dt
    CALL sw % ForwardStepRK3(t)
    IF (MOD(it + 1, every) == 0) THEN
      WRITE(tag, '(I10.10)') it + 1
      CALL sw % WriteTecplot('ShallowWater.'//tag)
    ENDIF
  ENDDO
  CALL sw % WritePickup(first + nSteps)
  CALL sw % mesh % WriteTecplot()
  CALL sw % Trash()
END PROGRAM sw_dropin_driver
