! adv_dropin.f90 -- the 2-D advection-diffusion model under the reference's
! own module, type and procedure names, so that a reference driver program
! compiles unchanged against it and runs on the GPU.
!
! Modules (each mirrors the reference module of the same name, as far as the
! model's driver uses it; ModelPrecision, ConstantsDictionary and
! QuadMeshClass come from sw_dropin.f90):
!   AdvectionParamsClass  src/highend/advection-diffusion2d/AdvectionParamsClass.f90
!                         (Build_AdvectionParams: the advection.params
!                         namelists TracerSetup, TimeManagement,
!                         SpaceManagement and their defaults)
!   AdvectionClass        src/highend/advection-diffusion2d/AdvectionClass.f90:
!                         TYPE Advection with Build(forceNoPickup) (:176-252,
!                         BuildQuadMesh :292-334), Trash, the per-element
!                         accessors, GlobalTimeDerivative (:1316-1378),
!                         ForwardStepRK3 (:1268-1312), CoarseToFine
!                         (:1894-1977), WriteTecplot (:1981-2036), WritePickup
!                         (:2040-2103) and ReadPickup (:2107-2171).
! The model itself is the C ABI's (self_adv2d_*), with the nodal storage as
! the reference builds it: its p-direction derivative matrix transposed
! (NodalStorage_2D, the reference's quirk), so that a driver reproduces the
! reference's results.  The element arrays a driver reads and writes
! (solution, velocity, relaxation field and factor) are mirrored on the host: a Set marks the host copy newer, a device step
! marks it older, and each side is refreshed only when the other is used.
MODULE AdvectionParamsClass
  USE ModelPrecision
  USE ConstantsDictionary
  IMPLICIT NONE
  CHARACTER(1), PARAMETER, PRIVATE :: blank = ' '

  TYPE AdvectionParams
    ! TracerSetup
    INTEGER    :: nTracers
    REAL(prec) :: kappa
    ! TimeManagement
    REAL(prec) :: dt
    INTEGER    :: iterInit, nTimeSteps, dumpFreq
    ! SpaceManagement
    CHARACTER(50) :: SpecMeshFile
    INTEGER       :: polyDeg, nXElem, nYElem, nPlot
    REAL(prec)    :: dxPlot, xScale, yScale
  CONTAINS
    PROCEDURE :: Build => Build_AdvectionParams
  END TYPE AdvectionParams

CONTAINS

  SUBROUTINE Build_AdvectionParams(thisParam)
    CLASS(AdvectionParams), INTENT(out) :: thisParam
    INTEGER       :: nUnit
    INTEGER       :: nTracers, iterInit, nTimeSteps, dumpFreq, polyDeg, nXElem, nYElem, nPlot
    REAL(prec)    :: kappa, dt, xScale, yScale
    CHARACTER(50) :: SpecMeshFile
    NAMELIST / TracerSetup / nTracers, kappa
    NAMELIST / TimeManagement / dt, iterInit, nTimeSteps, dumpFreq
    NAMELIST / SpaceManagement / SpecMeshFile, polyDeg, nXElem, nYElem, nPlot, xScale, yScale

    nTracers = 2
    kappa = ZERO
    dt = ZERO
    iterInit = 0
    nTimeSteps = 0
    dumpFreq = 1
    SpecMeshFile = blank
    polyDeg = 5
    nXElem = 5
    nYElem = 5
    nPlot = 10
    xScale = ONE
    yScale = ONE

    OPEN(NEWUNIT=nUnit, FILE='advection.params', STATUS='old', ACTION='read')
    READ(UNIT=nUnit, NML=TracerSetup)
    READ(UNIT=nUnit, NML=TimeManagement)
    READ(UNIT=nUnit, NML=SpaceManagement)
    CLOSE(UNIT=nUnit)
    WRITE(UNIT=*, NML=TracerSetup)
    WRITE(UNIT=*, NML=TimeManagement)
    WRITE(UNIT=*, NML=SpaceManagement)

    thisParam % nTracers = nTracers
    thisParam % kappa = kappa
    thisParam % dt = dt
    thisParam % iterInit = iterInit
    thisParam % nTimeSteps = nTimeSteps
    thisParam % dumpFreq = dumpFreq
    thisParam % SpecMeshFile = SpecMeshFile
    thisParam % polyDeg = polyDeg
    thisParam % nXElem = nXElem
    thisParam % nYElem = nYElem
    thisParam % nPlot = nPlot
    thisParam % dxPlot = 2.0_prec / REAL(nPlot, prec)
    thisParam % xScale = xScale
    thisParam % yScale = yScale
  END SUBROUTINE Build_AdvectionParams

END MODULE AdvectionParamsClass

MODULE AdvectionClass
  USE ISO_C_BINDING
  USE ModelPrecision
  USE ConstantsDictionary
  USE QuadMeshClass
  USE AdvectionParamsClass
  USE SELF_AMD_CABI
  IMPLICIT NONE

  ! fields of self_adv2d_set_field / self_adv2d_get_field
  INTEGER(C_INT), PARAMETER, PRIVATE :: F_SOL = 0, F_VEL = 1, F_RELAX = 2, F_RFAC = 3
  INTEGER(C_INT), PARAMETER, PRIVATE :: G_SOL = 0, G_TEND = 1, G_DCDX = 2, G_DCDY = 3, G_VEL = 4, G_RELAX = 5, &
                                        G_RFAC = 6
  ! the host mirrors: solution, velocity, relaxation field, relaxation factor
  INTEGER, PARAMETER, PRIVATE :: M_SOL = 1, M_VEL = 2, M_RELAX = 3, M_RFAC = 4

  TYPE Advection
    TYPE(AdvectionParams) :: params
    TYPE(QuadMesh)        :: mesh
    INTEGER               :: nS, nP, nEq, nPlot
    REAL(prec), ALLOCATABLE :: plMatS(:,:), plMatP(:,:)
    TYPE(C_PTR) :: handle = C_NULL_PTR
    ! host mirrors (0:N,0:N,1:nEq or 2,nElems) and their state: 0 equal to
    ! the device, 1 host newer (upload before the device uses it), -1
    ! device newer (download before the host reads it)
    REAL(prec), ALLOCATABLE :: sol(:,:,:,:), vel(:,:,:,:), relax(:,:,:,:), rFac(:,:,:,:)
    INTEGER :: state(4) = 0
  CONTAINS
    PROCEDURE :: Build => Build_Advection
    PROCEDURE :: Trash => Trash_Advection
    PROCEDURE :: GetSolution => GetSolution_Advection
    PROCEDURE :: SetSolution => SetSolution_Advection
    PROCEDURE :: GetVelocity => GetVelocity_Advection
    PROCEDURE :: SetVelocity => SetVelocity_Advection
    PROCEDURE :: GetRelaxationField => GetRelaxationField_Advection
    PROCEDURE :: SetRelaxationField => SetRelaxationField_Advection
    PROCEDURE :: GetRelaxationFactor => GetRelaxationFactor_Advection
    PROCEDURE :: SetRelaxationFactor => SetRelaxationFactor_Advection
    PROCEDURE :: GetGradient => GetGradient_Advection
    PROCEDURE :: GlobalTimeDerivative => GlobalTimeDerivative_Advection
    PROCEDURE :: ForwardStepRK3 => ForwardStepRK3_Advection
    PROCEDURE :: CoarseToFine => CoarseToFine_Advection
    PROCEDURE :: WriteTecplot => WriteTecplot_Advection
    PROCEDURE :: WritePickup => WritePickup_Advection
    PROCEDURE :: ReadPickup => ReadPickup_Advection
  END TYPE Advection

CONTAINS

  ! Build (:176-252): parameters from advection.params; the default mesh
  ! (LoadDefaultMesh, walled, every boundary edge whose primary side is west
  ! or north flagged InflowOne, east or south InflowTwo, :303-320) or the
  ! ISM-V2 file SpecMeshFile, scaled by xScale, yScale (:328-330); the model
  ! on the device with zero fields; unless forceNoPickup is present, the
  ! pickup at iterInit; the plotting matrices to nPlot + 1 uniform points.
  SUBROUTINE Build_Advection(myDGSEM, forceNoPickup)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in), OPTIONAL   :: forceNoPickup
    INTEGER(C_INT) :: N, nE, nEd, nNd, gDeg, k
    INTEGER(C_INT), ALLOCATABLE :: edges(:,:), elemNodes(:), nodeType(:)
    REAL(C_DOUBLE), ALLOCATABLE :: nodes(:), qWeight(:), dMatS(:,:), dMatP(:,:), lW(:), lE(:), lS(:), lN(:)
    REAL(C_DOUBLE), ALLOCATABLE :: nHat(:), nLen(:), nodeX(:), nodeY(:), bw(:), sNew(:)
    CHARACTER(KIND=C_CHAR, LEN=256) :: path
    INTEGER :: i

    CALL myDGSEM % params % Build()
    N = myDGSEM % params % polyDeg
    myDGSEM % nS = N
    myDGSEM % nP = N
    myDGSEM % nEq = myDGSEM % params % nTracers
    myDGSEM % nPlot = myDGSEM % params % nPlot
    ALLOCATE(nodes(0:N), qWeight(0:N), dMatS(0:N, 0:N), dMatP(0:N, 0:N), lW(0:N), lE(0:N), lS(0:N), lN(0:N))
    CALL self_check(self_nodal2d_build(N, SELF_GAUSS, SELF_DG, 1_C_INT, nodes, qWeight, dMatS, dMatP, lW, lE, lS, &
                                       lN), 'Build')
    PRINT *, 'Module AdvectionClass.f90 : S/R BuildQuadMesh :'
    IF (LEN_TRIM(myDGSEM % params % SpecMeshFile) == 0) THEN
      PRINT *, ' Loading default mesh.'
      CALL self_check(self_quadmesh_counts(myDGSEM % params % nXElem, myDGSEM % params % nYElem, 0_C_INT, 0_C_INT, &
                                           nE, nEd), 'Build')
      CALL alloc_mesh()
      CALL self_check(self_quadmesh_structured(myDGSEM % params % nXElem, myDGSEM % params % nYElem, N, 0_C_INT, &
                                               0_C_INT, myDGSEM % params % xScale, myDGSEM % params % yScale, ZERO, &
                                               nodes, edges, myDGSEM % mesh % x, myDGSEM % mesh % y, &
                                               myDGSEM % mesh % dxds, myDGSEM % mesh % dxdp, myDGSEM % mesh % dyds, &
                                               myDGSEM % mesh % dydp, myDGSEM % mesh % J, nHat, nLen), 'Build')
      ! sides: south 1, east 2, north 3, west 4
      DO k = 1, nEd
        IF (edges(3, k) == SELF_NO_NORMAL_FLOW) THEN
          IF (edges(2, k) == 4 .OR. edges(2, k) == 3) THEN
            edges(3, k) = SELF_INFLOW_ONE
          ELSEIF (edges(2, k) == 2 .OR. edges(2, k) == 1) THEN
            edges(3, k) = SELF_INFLOW_TWO
          ENDIF
        ENDIF
      ENDDO
    ELSE
      PRINT *, 'Reading mesh from '//TRIM(myDGSEM % params % SpecMeshFile)//'.'
      path = TRIM(myDGSEM % params % SpecMeshFile)//C_NULL_CHAR
      CALL self_check(self_specmesh_counts(path, nNd, nEd, nE, gDeg), 'Build')
      CALL alloc_mesh()
      ALLOCATE(elemNodes(4 * nE), nodeType(nNd), nodeX(nNd), nodeY(nNd))
      CALL self_check(self_specmesh_read(path, N, nodes, edges, elemNodes, nodeType, nodeX, nodeY, &
                                         myDGSEM % mesh % x, myDGSEM % mesh % y, myDGSEM % mesh % dxds, &
                                         myDGSEM % mesh % dxdp, myDGSEM % mesh % dyds, myDGSEM % mesh % dydp, &
                                         myDGSEM % mesh % J, nHat, nLen), 'Build')
      CALL self_check(self_quadmesh_scale(N, nodes, nE, nNd, myDGSEM % params % xScale, myDGSEM % params % yScale, &
                                          nodeX, nodeY, myDGSEM % mesh % x, myDGSEM % mesh % y, myDGSEM % mesh % dxds, &
                                          myDGSEM % mesh % dxdp, myDGSEM % mesh % dyds, myDGSEM % mesh % dydp, &
                                          myDGSEM % mesh % J, nHat, nLen), 'Build')
    ENDIF
    myDGSEM % mesh % nElems = nE
    myDGSEM % mesh % nS = N
    myDGSEM % mesh % nP = N
    CALL self_check(self_adv2d_create(N, INT(myDGSEM % nEq, C_INT), nE, nEd, edges, myDGSEM % mesh % dxds, &
                                      myDGSEM % mesh % dxdp, myDGSEM % mesh % dyds, myDGSEM % mesh % dydp, &
                                      myDGSEM % mesh % J, nHat, nLen, dMatS, dMatP, qWeight, lW, lE, lS, lN, &
                                      myDGSEM % params % kappa, myDGSEM % params % dt, myDGSEM % handle), 'Build')
    ALLOCATE(myDGSEM % sol(0:N, 0:N, myDGSEM % nEq, nE), myDGSEM % vel(0:N, 0:N, 2, nE), &
             myDGSEM % relax(0:N, 0:N, myDGSEM % nEq, nE), myDGSEM % rFac(0:N, 0:N, myDGSEM % nEq, nE))
    myDGSEM % sol = ZERO
    myDGSEM % vel = ZERO
    myDGSEM % relax = ZERO
    myDGSEM % rFac = ZERO
    myDGSEM % state = 0

    IF (.NOT. PRESENT(forceNoPickup)) CALL myDGSEM % ReadPickup(myDGSEM % params % iterInit)

    ALLOCATE(bw(0:N), sNew(0:myDGSEM % nPlot), myDGSEM % plMatS(0:myDGSEM % nPlot, 0:N), &
             myDGSEM % plMatP(0:myDGSEM % nPlot, 0:N))
    DO i = 0, myDGSEM % nPlot  ! UniformPoints(-1, 1, nPlot), src/common/CommonRoutines.f90:460-482
      sNew(i) = -ONE + (TWO / REAL(myDGSEM % nPlot, prec)) * REAL(i, prec)
    ENDDO
    CALL self_check(self_barycentric_weights(N, nodes, bw), 'Build')
    CALL self_check(self_interpolation_matrix(N, nodes, bw, INT(myDGSEM % nPlot, C_INT), sNew, myDGSEM % plMatS), &
                    'Build')
    myDGSEM % plMatP = myDGSEM % plMatS
  CONTAINS
    SUBROUTINE alloc_mesh()
      ALLOCATE(edges(6, nEd), nHat(nE * 4 * (N + 1) * 2), nLen(nE * 4 * (N + 1)))
      ALLOCATE(myDGSEM % mesh % x(0:N, 0:N, nE), myDGSEM % mesh % y(0:N, 0:N, nE), &
               myDGSEM % mesh % dxds(0:N, 0:N, nE), myDGSEM % mesh % dxdp(0:N, 0:N, nE), &
               myDGSEM % mesh % dyds(0:N, 0:N, nE), myDGSEM % mesh % dydp(0:N, 0:N, nE), &
               myDGSEM % mesh % J(0:N, 0:N, nE))
    END SUBROUTINE alloc_mesh
  END SUBROUTINE Build_Advection

  SUBROUTINE Trash_Advection(myDGSEM)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    IF (C_ASSOCIATED(myDGSEM % handle)) CALL self_check(self_adv2d_destroy(myDGSEM % handle), 'Trash')
    myDGSEM % handle = C_NULL_PTR
    IF (ALLOCATED(myDGSEM % sol)) DEALLOCATE(myDGSEM % sol, myDGSEM % vel, myDGSEM % relax, myDGSEM % rFac)
    IF (ALLOCATED(myDGSEM % plMatS)) DEALLOCATE(myDGSEM % plMatS, myDGSEM % plMatP)
    IF (ALLOCATED(myDGSEM % mesh % x)) DEALLOCATE(myDGSEM % mesh % x, myDGSEM % mesh % y, myDGSEM % mesh % dxds, &
                                                  myDGSEM % mesh % dxdp, myDGSEM % mesh % dyds, &
                                                  myDGSEM % mesh % dydp, myDGSEM % mesh % J)
  END SUBROUTINE Trash_Advection

  ! ------------------------------------------------ host mirrors
  ! the device copy of every field the host changed
  SUBROUTINE push(myDGSEM)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    IF (myDGSEM % state(M_SOL) == 1) CALL self_check(self_adv2d_set_field(myDGSEM % handle, F_SOL, myDGSEM % sol), &
                                                     'SetSolution')
    IF (myDGSEM % state(M_VEL) == 1) CALL self_check(self_adv2d_set_field(myDGSEM % handle, F_VEL, myDGSEM % vel), &
                                                     'SetVelocity')
    IF (myDGSEM % state(M_RELAX) == 1) CALL self_check(self_adv2d_set_field(myDGSEM % handle, F_RELAX, &
                                                                            myDGSEM % relax), 'SetRelaxationField')
    IF (myDGSEM % state(M_RFAC) == 1) CALL self_check(self_adv2d_set_field(myDGSEM % handle, F_RFAC, &
                                                                           myDGSEM % rFac), 'SetRelaxationFactor')
    WHERE (myDGSEM % state == 1) myDGSEM % state = 0
  END SUBROUTINE push

  ! the host copy of field m, if the device's is newer
  SUBROUTINE pull(myDGSEM, m)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in) :: m
    IF (myDGSEM % state(m) /= -1) RETURN
    SELECT CASE (m)
    CASE (M_SOL)
      CALL self_check(self_adv2d_get_field(myDGSEM % handle, G_SOL, myDGSEM % sol), 'GetSolution')
    CASE (M_VEL)
      CALL self_check(self_adv2d_get_field(myDGSEM % handle, G_VEL, myDGSEM % vel), 'GetVelocity')
    CASE (M_RELAX)
      CALL self_check(self_adv2d_get_field(myDGSEM % handle, G_RELAX, myDGSEM % relax), 'GetRelaxationField')
    CASE (M_RFAC)
      CALL self_check(self_adv2d_get_field(myDGSEM % handle, G_RFAC, myDGSEM % rFac), 'GetRelaxationFactor')
    END SELECT
    myDGSEM % state(m) = 0
  END SUBROUTINE pull

  ! before the host changes one element of field m: the rest of it current
  SUBROUTINE touch(myDGSEM, m)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in) :: m
    CALL pull(myDGSEM, m)
    myDGSEM % state(m) = 1
  END SUBROUTINE touch

  SUBROUTINE GetSolution_Advection(myDGSEM, iEl, sol)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in) :: iEl
    REAL(prec), INTENT(out) :: sol(0:myDGSEM % nS, 0:myDGSEM % nP, 1:myDGSEM % nEq)
    CALL pull(myDGSEM, M_SOL)
    sol = myDGSEM % sol(:, :, :, iEl)
  END SUBROUTINE GetSolution_Advection

  SUBROUTINE SetSolution_Advection(myDGSEM, iEl, sol)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in) :: iEl
    REAL(prec), INTENT(in) :: sol(0:myDGSEM % nS, 0:myDGSEM % nP, 1:myDGSEM % nEq)
    CALL touch(myDGSEM, M_SOL)
    myDGSEM % sol(:, :, :, iEl) = sol
  END SUBROUTINE SetSolution_Advection

  SUBROUTINE GetVelocity_Advection(myDGSEM, iEl, u, v)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in) :: iEl
    REAL(prec), INTENT(out) :: u(0:myDGSEM % nS, 0:myDGSEM % nP), v(0:myDGSEM % nS, 0:myDGSEM % nP)
    CALL pull(myDGSEM, M_VEL)
    u = myDGSEM % vel(:, :, 1, iEl)
    v = myDGSEM % vel(:, :, 2, iEl)
  END SUBROUTINE GetVelocity_Advection

  ! SetVelocity (:434-449); the velocity's boundary values follow on the device
  SUBROUTINE SetVelocity_Advection(myDGSEM, iEl, u, v)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in) :: iEl
    REAL(prec), INTENT(in) :: u(0:myDGSEM % nS, 0:myDGSEM % nP), v(0:myDGSEM % nS, 0:myDGSEM % nP)
    CALL touch(myDGSEM, M_VEL)
    myDGSEM % vel(:, :, 1, iEl) = u
    myDGSEM % vel(:, :, 2, iEl) = v
  END SUBROUTINE SetVelocity_Advection

  SUBROUTINE GetRelaxationField_Advection(myDGSEM, iEl, f)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in) :: iEl
    REAL(prec), INTENT(out) :: f(0:myDGSEM % nS, 0:myDGSEM % nP, 1:myDGSEM % nEq)
    CALL pull(myDGSEM, M_RELAX)
    f = myDGSEM % relax(:, :, :, iEl)
  END SUBROUTINE GetRelaxationField_Advection

  SUBROUTINE SetRelaxationField_Advection(myDGSEM, iEl, f)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in) :: iEl
    REAL(prec), INTENT(in) :: f(0:myDGSEM % nS, 0:myDGSEM % nP, 1:myDGSEM % nEq)
    CALL touch(myDGSEM, M_RELAX)
    myDGSEM % relax(:, :, :, iEl) = f
  END SUBROUTINE SetRelaxationField_Advection

  SUBROUTINE GetRelaxationFactor_Advection(myDGSEM, iEl, f)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in) :: iEl
    REAL(prec), INTENT(out) :: f(0:myDGSEM % nS, 0:myDGSEM % nP, 1:myDGSEM % nEq)
    CALL pull(myDGSEM, M_RFAC)
    f = myDGSEM % rFac(:, :, :, iEl)
  END SUBROUTINE GetRelaxationFactor_Advection

  SUBROUTINE SetRelaxationFactor_Advection(myDGSEM, iEl, f)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in) :: iEl
    REAL(prec), INTENT(in) :: f(0:myDGSEM % nS, 0:myDGSEM % nP, 1:myDGSEM % nEq)
    CALL touch(myDGSEM, M_RFAC)
    myDGSEM % rFac(:, :, :, iEl) = f
  END SUBROUTINE SetRelaxationFactor_Advection

  ! GetGradient: the concentration gradient of the last time derivative
  ! (CalculateConcentrationGradient :1123-1264, run inside each one)
  SUBROUTINE GetGradient_Advection(myDGSEM, iEl, dcdx, dcdy)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in) :: iEl
    REAL(prec), INTENT(out), DIMENSION(0:myDGSEM % nS, 0:myDGSEM % nP, 1:myDGSEM % nEq) :: dcdx, dcdy
    REAL(prec), ALLOCATABLE :: gx(:,:,:,:), gy(:,:,:,:)
    ALLOCATE(gx(0:myDGSEM % nS, 0:myDGSEM % nP, myDGSEM % nEq, myDGSEM % mesh % nElems), &
             gy(0:myDGSEM % nS, 0:myDGSEM % nP, myDGSEM % nEq, myDGSEM % mesh % nElems))
    CALL self_check(self_adv2d_get_field(myDGSEM % handle, G_DCDX, gx), 'GetGradient')
    CALL self_check(self_adv2d_get_field(myDGSEM % handle, G_DCDY, gy), 'GetGradient')
    dcdx = gx(:, :, :, iEl)
    dcdy = gy(:, :, :, iEl)
  END SUBROUTINE GetGradient_Advection

  ! ------------------------------------------------ time stepping on the device
  SUBROUTINE GlobalTimeDerivative_Advection(myDGSEM, tn)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    REAL(prec), INTENT(in) :: tn
    CALL push(myDGSEM)
    CALL self_check(self_adv2d_global_time_derivative(myDGSEM % handle, tn), 'GlobalTimeDerivative')
  END SUBROUTINE GlobalTimeDerivative_Advection

  SUBROUTINE ForwardStepRK3_Advection(myDGSEM, tn)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    REAL(prec), INTENT(in) :: tn
    CALL push(myDGSEM)
    CALL self_check(self_adv2d_forward_step_rk3(myDGSEM % handle, tn, 1_C_INT, C_NULL_PTR), 'ForwardStepRK3')
    myDGSEM % state(M_SOL) = -1
  END SUBROUTINE ForwardStepRK3_Advection

  ! ------------------------------------------------ output
  ! CoarseToFine (:1894-1977): positions, velocity, tracers and their
  ! gradients of element iEl at the plotting points
  SUBROUTINE CoarseToFine_Advection(myDGSEM, iEl, x, y, u, v, c, dcdx, dcdy)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in) :: iEl
    REAL(prec), INTENT(out), DIMENSION(0:myDGSEM % nPlot, 0:myDGSEM % nPlot) :: x, y, u, v
    REAL(prec), INTENT(out), DIMENSION(0:myDGSEM % nPlot, 0:myDGSEM % nPlot, 1:myDGSEM % nEq) :: c, dcdx, dcdy
    REAL(prec), DIMENSION(0:myDGSEM % nS, 0:myDGSEM % nP, 1:myDGSEM % nEq) :: s, gx, gy
    REAL(prec), DIMENSION(0:myDGSEM % nS, 0:myDGSEM % nP) :: lu, lv
    INTEGER :: k
    CALL myDGSEM % GetSolution(iEl, s)
    CALL myDGSEM % GetGradient(iEl, gx, gy)
    CALL myDGSEM % GetVelocity(iEl, lu, lv)
    CALL c2f(myDGSEM % mesh % x(:, :, iEl), x)
    CALL c2f(myDGSEM % mesh % y(:, :, iEl), y)
    CALL c2f(lu, u)
    CALL c2f(lv, v)
    DO k = 1, myDGSEM % nEq
      CALL c2f(s(:, :, k), c(:, :, k))
      CALL c2f(gx(:, :, k), dcdx(:, :, k))
      CALL c2f(gy(:, :, k), dcdy(:, :, k))
    ENDDO
  CONTAINS
    SUBROUTINE c2f(fOld, fNew)
      REAL(prec), INTENT(in)  :: fOld(0:myDGSEM % nS, 0:myDGSEM % nP)
      REAL(prec), INTENT(out) :: fNew(0:myDGSEM % nPlot, 0:myDGSEM % nPlot)
      REAL(prec) :: a(0:myDGSEM % nS, 0:myDGSEM % nP), b(0:myDGSEM % nPlot, 0:myDGSEM % nPlot)
      a = fOld
      CALL self_check(self_coarse_to_fine_2d(INT(myDGSEM % nS, C_INT), INT(myDGSEM % nP, C_INT), 1_C_SIZE_T, &
                                             myDGSEM % plMatS, myDGSEM % plMatP, INT(myDGSEM % nPlot, C_INT), &
                                             INT(myDGSEM % nPlot, C_INT), a, b), 'CoarseToFine')
      fNew = b
    END SUBROUTINE c2f
  END SUBROUTINE CoarseToFine_Advection

  ! WriteTecplot (:1981-2036): <filename>.tec (or Advection.tec), per element
  ! a zone of X, Y, U, V, C1, C2 and the gradients of C1 and C2 at the
  ! plotting points; the whole mesh goes through one batched CoarseToFine
  ! per field.
  SUBROUTINE WriteTecplot_Advection(myDGSEM, filename)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    CHARACTER(*), INTENT(in), OPTIONAL :: filename
    INTEGER :: iS, iP, iEl, fUnit, nPlot, N, nE, nq
    CHARACTER(len=5) :: zoneID
    REAL(prec), ALLOCATABLE :: gx(:,:,:,:), gy(:,:,:,:), tmp(:,:,:), xf(:,:,:), yf(:,:,:), uf(:,:,:), vf(:,:,:)
    REAL(prec), ALLOCATABLE :: cf(:,:,:,:), gxf(:,:,:,:), gyf(:,:,:,:)
    nPlot = myDGSEM % nPlot
    N = myDGSEM % nS
    nE = myDGSEM % mesh % nElems
    nq = myDGSEM % nEq
    CALL pull(myDGSEM, M_SOL)
    CALL pull(myDGSEM, M_VEL)
    ALLOCATE(gx(0:N, 0:N, nq, nE), gy(0:N, 0:N, nq, nE), tmp(0:N, 0:N, nE))
    ALLOCATE(xf(0:nPlot, 0:nPlot, nE), yf(0:nPlot, 0:nPlot, nE), uf(0:nPlot, 0:nPlot, nE), vf(0:nPlot, 0:nPlot, nE))
    ALLOCATE(cf(0:nPlot, 0:nPlot, nE, 2), gxf(0:nPlot, 0:nPlot, nE, 2), gyf(0:nPlot, 0:nPlot, nE, 2))
    CALL self_check(self_adv2d_get_field(myDGSEM % handle, G_DCDX, gx), 'WriteTecplot')
    CALL self_check(self_adv2d_get_field(myDGSEM % handle, G_DCDY, gy), 'WriteTecplot')
    CALL c2f_all(myDGSEM % mesh % x, xf)
    CALL c2f_all(myDGSEM % mesh % y, yf)
    tmp = myDGSEM % vel(:, :, 1, :)
    CALL c2f_all(tmp, uf)
    tmp = myDGSEM % vel(:, :, 2, :)
    CALL c2f_all(tmp, vf)
    DO iS = 1, 2
      tmp = myDGSEM % sol(:, :, MIN(iS, nq), :)
      CALL c2f_all(tmp, cf(:, :, :, iS))
      tmp = gx(:, :, MIN(iS, nq), :)
      CALL c2f_all(tmp, gxf(:, :, :, iS))
      tmp = gy(:, :, MIN(iS, nq), :)
      CALL c2f_all(tmp, gyf(:, :, :, iS))
    ENDDO
    IF (PRESENT(filename)) THEN
      OPEN(NEWUNIT=fUnit, FILE=TRIM(filename)//'.tec', FORM='formatted', STATUS='replace')
    ELSE
      OPEN(NEWUNIT=fUnit, FILE='Advection.tec', FORM='formatted', STATUS='replace')
    ENDIF
    WRITE(fUnit, *) 'VARIABLES = "X", "Y", "U", "V", "C1", "C2", "dC1dx", "dC1dy", "dC2dx", "dC2dy" '
    DO iEl = 1, nE
      WRITE(zoneID, '(I5.5)') iEl
      WRITE(fUnit, *) 'ZONE T="el'//TRIM(zoneID)//'", I=', nPlot + 1, ', J=', nPlot + 1, ',F=POINT'
      DO iP = 0, nPlot
        DO iS = 0, nPlot
          WRITE(fUnit, *) xf(iS, iP, iEl), yf(iS, iP, iEl), uf(iS, iP, iEl), vf(iS, iP, iEl), &
                          cf(iS, iP, iEl, 1), cf(iS, iP, iEl, 2), gxf(iS, iP, iEl, 1), gyf(iS, iP, iEl, 1), &
                          gxf(iS, iP, iEl, 2), gyf(iS, iP, iEl, 2)
        ENDDO
      ENDDO
    ENDDO
    CLOSE(UNIT=fUnit)
  CONTAINS
    SUBROUTINE c2f_all(fOld, fNew)
      REAL(prec), INTENT(in)  :: fOld(0:N, 0:N, nE)
      REAL(prec), INTENT(out) :: fNew(0:nPlot, 0:nPlot, nE)
      REAL(prec), ALLOCATABLE :: a(:,:,:)
      ALLOCATE(a(0:N, 0:N, nE))
      a = fOld
      CALL self_check(self_coarse_to_fine_2d(INT(N, C_INT), INT(N, C_INT), INT(nE, C_SIZE_T), myDGSEM % plMatS, &
                                             myDGSEM % plMatP, INT(nPlot, C_INT), INT(nPlot, C_INT), a, fNew), &
                      'WriteTecplot')
    END SUBROUTINE c2f_all
  END SUBROUTINE WriteTecplot_Advection

  ! WritePickup (:2040-2103) / ReadPickup (:2107-2171):
  ! Advection.<iter, 10 digits>.pickup in the reference's record layout
  SUBROUTINE WritePickup_Advection(myDGSEM, iter)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in) :: iter
    CHARACTER(len=10) :: iterChar
    CALL push(myDGSEM)
    WRITE(iterChar, '(I10.10)') iter
    CALL self_check(self_adv2d_write_pickup(myDGSEM % handle, 'Advection.'//iterChar//'.pickup'//C_NULL_CHAR), &
                    'WritePickup')
  END SUBROUTINE WritePickup_Advection

  SUBROUTINE ReadPickup_Advection(myDGSEM, iter)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in) :: iter
    CHARACTER(len=10) :: iterChar
    WRITE(iterChar, '(I10.10)') iter
    CALL self_check(self_adv2d_read_pickup(myDGSEM % handle, 'Advection.'//iterChar//'.pickup'//C_NULL_CHAR), &
                    'ReadPickup')
    myDGSEM % state = -1
  END SUBROUTINE ReadPickup_Advection

END MODULE AdvectionClass
