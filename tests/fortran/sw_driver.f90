! sw_driver.f90 -- a reference-style driver program on the drop-in type:
! Build, per-element SetSolution/SetBathymetry, CalculateBathymetryGradient,
! ForwardStepRK3(tn) in a loop, per-element GetSolution (the pattern of
! src/highend/shallowwater/ShallowWater_Driver.f90:60-66).
! Usage: sw_driver OUTFILE PRECISION(0|1) FORM(0|1) NSTEPS [MESHFILE]
! With MESHFILE the model reads that ISM-V2 mesh (SpecMeshFile), scaled by
! the same xScale, yScale, instead of the structured 6 x 5 mesh.
! Writes (stream, native endian): initial solution, bathymetry (after the
! gradient), final solution, each (0:N,0:N,1:3,1:nElems) FP64; then
! WritePickup(NSTEPS) (ShallowWater.<NSTEPS>.pickup in the working
! directory), ReadPickup of it into a second model, one more step of both,
! and both solutions (two more blocks).
PROGRAM sw_driver
  USE ISO_C_BINDING
  USE SELF_AMD_CABI
  USE ShallowWaterAMD_Class
  IMPLICIT NONE
  TYPE(ShallowWaterAMD) :: sw, sw2
  TYPE(SWParamsAMD)     :: p
  REAL(C_DOUBLE), ALLOCATABLE :: sol(:,:,:), bathy(:,:,:), allSol(:,:,:,:), allBathy(:,:,:,:)
  REAL(C_DOUBLE) :: x, y, tn, kx, ky
  REAL(C_DOUBLE), PARAMETER :: pi = 3.14159265358979323846_C_DOUBLE
  INTEGER :: iEl, i, j, iStep, nSteps, u, N
  CHARACTER(256) :: outFile, arg

  CALL GET_COMMAND_ARGUMENT(1, outFile)
  CALL GET_COMMAND_ARGUMENT(2, arg)
  READ(arg, *) p % precision
  CALL GET_COMMAND_ARGUMENT(3, arg)
  READ(arg, *) p % form
  CALL GET_COMMAND_ARGUMENT(4, arg)
  READ(arg, *) nSteps
  IF (COMMAND_ARGUMENT_COUNT() >= 5) CALL GET_COMMAND_ARGUMENT(5, p % SpecMeshFile)
  p % polyDeg = 5
  p % nXElem = 6
  p % nYElem = 5
  p % xScale = 2.0_C_DOUBLE
  p % yScale = 3.0_C_DOUBLE
  p % dt = 1.0e-3_C_DOUBLE
  p % g = 9.81_C_DOUBLE
  p % f0 = 0.1_C_DOUBLE
  p % warp = 0.03_C_DOUBLE
  N = p % polyDeg
  kx = 2.0_C_DOUBLE * pi / p % xScale
  ky = 2.0_C_DOUBLE * pi / p % yScale

  CALL self_check(self_init(0_C_INT), 'self_init')
  CALL sw % Build(p)
  ALLOCATE(sol(0:N,0:N,3), bathy(0:N,0:N,3))
  ALLOCATE(allSol(0:N,0:N,3,sw % nElems), allBathy(0:N,0:N,3,sw % nElems))
  DO iEl = 1, sw % nElems
    DO j = 0, N
      DO i = 0, N
        CALL sw % GetPositionAtNode(iEl, x, y, i, j)
        sol(i,j,3) = 10.0_C_DOUBLE + 0.7_C_DOUBLE * SIN(kx * x + 0.3_C_DOUBLE) * COS(ky * y)
        sol(i,j,1) = 0.6_C_DOUBLE * COS(kx * x) * SIN(ky * y + 0.2_C_DOUBLE) * sol(i,j,3)
        sol(i,j,2) = 0.8_C_DOUBLE * SIN(kx * x - 0.1_C_DOUBLE) * COS(2.0_C_DOUBLE * ky * y) * sol(i,j,3)
        bathy(i,j,1) = 10.0_C_DOUBLE + 0.5_C_DOUBLE * COS(kx * x + ky * y)
        bathy(i,j,2:3) = 0.0_C_DOUBLE
      ENDDO
    ENDDO
    CALL sw % SetSolution(iEl, sol)
    CALL sw % SetBathymetry(iEl, bathy)
    allSol(:,:,:,iEl) = sol
  ENDDO
  CALL sw % CalculateBathymetryGradient()
  DO iEl = 1, sw % nElems
    CALL sw % GetBathymetry(iEl, bathy)
    allBathy(:,:,:,iEl) = bathy
  ENDDO

  tn = 0.0_C_DOUBLE
  DO iStep = 1, nSteps
    CALL sw % ForwardStepRK3(tn)
    tn = tn + p % dt
  ENDDO

  OPEN(NEWUNIT=u, FILE=TRIM(outFile), ACCESS='stream', FORM='unformatted', STATUS='replace')
  WRITE(u) allSol
  WRITE(u) allBathy
  DO iEl = 1, sw % nElems
    CALL sw % GetSolution(iEl, sol)
    allSol(:,:,:,iEl) = sol
  ENDDO
  WRITE(u) allSol

  CALL sw % WritePickup(nSteps)
  CALL sw2 % Build(p)
  CALL sw2 % ReadPickup(nSteps)
  CALL sw % ForwardStepRK3(tn)
  CALL sw2 % ForwardStepRK3(tn)
  DO iEl = 1, sw % nElems
    CALL sw % GetSolution(iEl, sol)
    allSol(:,:,:,iEl) = sol
  ENDDO
  WRITE(u) allSol
  DO iEl = 1, sw2 % nElems
    CALL sw2 % GetSolution(iEl, sol)
    allSol(:,:,:,iEl) = sol
  ENDDO
  WRITE(u) allSol
  CLOSE(u)
  CALL sw2 % Trash()
  CALL sw % Trash()
  PRINT '(A,I0,A,I0,A)', 'sw_driver: ', nSteps, ' steps on ', sw % nEdges, ' edges done'
END PROGRAM sw_driver
