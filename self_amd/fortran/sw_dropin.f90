! sw_dropin.f90 -- the shallow-water model under the reference's own module,
! type and procedure names, so that a reference driver program compiles
! unchanged against it and runs on the GPU.
!
! Modules (each mirrors the reference module of the same name, as far as a
! shallow-water driver uses it):
!   ModelPrecision         src/common/ModelPrecision.f90 (kinds)
!   ConstantsDictionary    src/common/ConstantsDictionary.f90:35-40 (constants)
!   SWParamsClass          examples/shallowwater/*/localmods/SWParamsClass.f90
!                          (the shallowwater.params namelist file and defaults)
!   QuadMeshClass          src/geom/QuadMeshClass.f90 (QuadMesh % nElems and
!                          WriteTecplot, :2476-2520, the output a driver calls)
!   ConservativeShallowWaterClass
!                          ConservativeShallowWaterClass.f90 (:60-151): TYPE
!                          ShallowWater with Build(forceNoPickup) (:163-245),
!                          Trash, the per-element accessors,
!                          CalculateBathymetryGradient, GlobalTimeDerivative,
!                          ForwardStepRK3, CoarseToFine, WriteTecplot
!                          (:1918-1970), WritePickup, ReadPickup.
! The device model is ShallowWaterAMD (shallowwater_amd.f90) over the C ABI,
! with the nodal storage as the reference builds it (its p-direction
! derivative matrix transposed, NodalStorage_2D);
! this layer reads the parameters, builds it, and writes the reference's
! output formats.  The reference's mesh is walled (no periodic boundaries).
MODULE ModelPrecision
  IMPLICIT NONE
  INTEGER, PARAMETER :: sp = SELECTED_REAL_KIND(6, 37)
  INTEGER, PARAMETER :: dp = SELECTED_REAL_KIND(15, 307)
  INTEGER, PARAMETER :: prec = dp
END MODULE ModelPrecision

MODULE ConstantsDictionary
  USE ModelPrecision
  IMPLICIT NONE
  REAL(prec), PARAMETER :: pi = 4.0_prec * ATAN(1.0_prec)
  REAL(prec), PARAMETER :: ZERO = 0.0_prec, ONE = 1.0_prec, TWO = 2.0_prec, HALF = 0.5_prec
  REAL(prec), PARAMETER :: TOL = EPSILON(1.0_prec)
END MODULE ConstantsDictionary

MODULE SWParamsClass
  USE ModelPrecision
  USE ConstantsDictionary
  IMPLICIT NONE

  CHARACTER(1), PARAMETER :: nada = ' '

  TYPE SWParams
    INTEGER       :: ModelFormulation
    REAL(prec)    :: dt
    INTEGER       :: iterInit, nTimeSteps, dumpFreq
    CHARACTER(50) :: SpecMeshFile
    INTEGER       :: polyDeg, nXElem, nYElem, geomPolyDeg, nPlot
    REAL(prec)    :: dxPlot, xScale, yScale
    REAL(prec)    :: g, f0, betaX, betaY, linearDrag
    REAL(prec)    :: eta0, x0, y0, x1, y1, R0, R1, delta, u0, u1, h0
  CONTAINS
    PROCEDURE :: Build => Build_SWParams
  END TYPE SWParams

  ! ModelFormulation keywords (the reference maps them to its model flags)
  INTEGER, PARAMETER :: LINEAR_FORM = 1, NONLINEAR_FORM = 2, CONSERVATIVE_FORM = 3

CONTAINS

  ! The reference's Build: defaults, then the five namelist groups of
  ! ./shallowwater.params in order, echoed to standard output.
  SUBROUTINE Build_SWParams(thisParam)
    CLASS(SWParams), INTENT(out) :: thisParam
    CHARACTER(50) :: ModelFormulation, SpecMeshFile
    REAL(prec)    :: dt, xScale, yScale, g, f0, betaX, betaY, linearDrag
    REAL(prec)    :: eta0, x0, y0, x1, y1, R0, R1, delta, u0, u1, h0
    INTEGER       :: iterInit, nTimeSteps, dumpFreq, polyDeg, nXElem, nYElem, nPlot, nUnit
    CHARACTER(50) :: up
    INTEGER :: k
    NAMELIST / MODEL_FORM / ModelFormulation
    NAMELIST / TIME_MANAGEMENT / dt, iterInit, nTimeSteps, dumpFreq
    NAMELIST / SPACE_MANAGEMENT / SpecMeshFile, polyDeg, nXElem, nYElem, nPlot, xScale, yScale
    NAMELIST / PHYSICAL_PARAMETERS / g, f0, betaX, betaY, linearDrag
    NAMELIST / DIPOLE_PARAMETERS / eta0, x0, y0, x1, y1, R0, R1, delta, u0, u1, h0

    ModelFormulation = 'LINEAR'
    dt = ZERO
    iterInit = 0
    nTimeSteps = 0
    dumpFreq = 1
    SpecMeshFile = nada
    polyDeg = 5
    nXElem = 5
    nYElem = 5
    nPlot = 10
    xScale = ONE
    yScale = ONE
    g = ONE
    f0 = 10.0_prec
    betaX = ZERO
    betaY = ZERO
    linearDrag = ZERO
    eta0 = ZERO
    x0 = 5.0_prec
    y0 = 5.5_prec
    x1 = 5.0_prec
    y1 = 4.5_prec
    R0 = 0.25_prec
    R1 = R0
    delta = 0.1_prec
    u0 = 0.5_prec
    u1 = 0.5_prec
    h0 = ONE

    OPEN(NEWUNIT=nUnit, FILE='shallowwater.params', STATUS='old', ACTION='read')
    READ(nUnit, NML=MODEL_FORM)
    READ(nUnit, NML=TIME_MANAGEMENT)
    READ(nUnit, NML=SPACE_MANAGEMENT)
    READ(nUnit, NML=PHYSICAL_PARAMETERS)
    READ(nUnit, NML=DIPOLE_PARAMETERS)
    CLOSE(nUnit)
    WRITE(*, NML=MODEL_FORM)
    WRITE(*, NML=TIME_MANAGEMENT)
    WRITE(*, NML=SPACE_MANAGEMENT)
    WRITE(*, NML=PHYSICAL_PARAMETERS)
    WRITE(*, NML=DIPOLE_PARAMETERS)

    up = ModelFormulation
    DO k = 1, LEN_TRIM(up)
      IF (up(k:k) >= 'a' .AND. up(k:k) <= 'z') up(k:k) = ACHAR(IACHAR(up(k:k)) - 32)
    ENDDO
    SELECT CASE (TRIM(up))
    CASE ('NONLINEAR')
      thisParam % ModelFormulation = NONLINEAR_FORM
    CASE ('CONSERVATIVE')
      thisParam % ModelFormulation = CONSERVATIVE_FORM
    CASE DEFAULT
      thisParam % ModelFormulation = LINEAR_FORM
    END SELECT
    thisParam % dt = dt
    thisParam % iterInit = iterInit
    thisParam % nTimeSteps = nTimeSteps
    thisParam % dumpFreq = dumpFreq
    thisParam % SpecMeshFile = SpecMeshFile
    thisParam % polyDeg = polyDeg
    thisParam % nXElem = nXElem
    thisParam % nYElem = nYElem
    thisParam % geomPolyDeg = 1
    thisParam % nPlot = nPlot
    thisParam % dxPlot = 2.0_prec / REAL(nPlot, prec)
    thisParam % xScale = xScale
    thisParam % yScale = yScale
    thisParam % g = g
    thisParam % f0 = f0
    thisParam % betaX = betaX
    thisParam % betaY = betaY
    thisParam % linearDrag = linearDrag
    thisParam % eta0 = eta0
    thisParam % x0 = x0
    thisParam % y0 = y0
    thisParam % x1 = x1
    thisParam % y1 = y1
    thisParam % R0 = R0
    thisParam % R1 = R1
    thisParam % delta = delta
    thisParam % u0 = u0
    thisParam % u1 = u1
    thisParam % h0 = h0
  END SUBROUTINE Build_SWParams

END MODULE SWParamsClass

MODULE QuadMeshClass
  USE ModelPrecision
  USE ConstantsDictionary
  IMPLICIT NONE

  ! The mesh as a driver sees it: element count, degree and the nodal
  ! geometry (positions, covariant metric terms, Jacobian).
  TYPE QuadMesh
    INTEGER :: nElems = 0, nS = 0, nP = 0
    REAL(prec), ALLOCATABLE :: x(:,:,:), y(:,:,:), dxds(:,:,:), dxdp(:,:,:), dyds(:,:,:), dydp(:,:,:), J(:,:,:)
  CONTAINS
    PROCEDURE :: GetPositionAtNode => GetPositionAtNode_QuadMesh
    PROCEDURE :: WriteTecplot => WriteTecplot_QuadMesh
  END TYPE QuadMesh

CONTAINS

  SUBROUTINE GetPositionAtNode_QuadMesh(myQuadMesh, iEl, x, y, i, j)
    CLASS(QuadMesh), INTENT(in) :: myQuadMesh
    INTEGER, INTENT(in)         :: iEl, i, j
    REAL(prec), INTENT(out)     :: x, y
    x = myQuadMesh % x(i, j, iEl)
    y = myQuadMesh % y(i, j, iEl)
  END SUBROUTINE GetPositionAtNode_QuadMesh

  ! mesh.tec (or <filename>.tec): per element a zone of the node values of
  ! X, Y, Jacobian and the four metric terms, the reference's layout
  SUBROUTINE WriteTecplot_QuadMesh(myQuadMesh, filename)
    CLASS(QuadMesh), INTENT(inout)     :: myQuadMesh
    CHARACTER(*), INTENT(in), OPTIONAL :: filename
    INTEGER :: iS, iP, iEl, fUnit
    CHARACTER(7) :: zoneID
    IF (PRESENT(filename)) THEN
      OPEN(NEWUNIT=fUnit, FILE=TRIM(filename)//'.tec', FORM='formatted')
    ELSE
      OPEN(NEWUNIT=fUnit, FILE='mesh.tec', FORM='formatted')
    ENDIF
    WRITE(fUnit, *) 'VARIABLES = "X", "Y", "Jacobian", "dxds", "dxdp", "dyds", "dydp" '
    DO iEl = 1, myQuadMesh % nElems
      WRITE(zoneID, '(I7.7)') iEl
      WRITE(fUnit, *) 'ZONE T="el'//TRIM(zoneID)//'", I=', myQuadMesh % nS + 1, ', J=', myQuadMesh % nP + 1, &
                      ',F=POINT'
      DO iP = 0, myQuadMesh % nP
        DO iS = 0, myQuadMesh % nS
          WRITE(fUnit, *) myQuadMesh % x(iS, iP, iEl), myQuadMesh % y(iS, iP, iEl), myQuadMesh % J(iS, iP, iEl), &
                          myQuadMesh % dxds(iS, iP, iEl), myQuadMesh % dxdp(iS, iP, iEl), &
                          myQuadMesh % dyds(iS, iP, iEl), myQuadMesh % dydp(iS, iP, iEl)
        ENDDO
      ENDDO
    ENDDO
    CLOSE(fUnit)
  END SUBROUTINE WriteTecplot_QuadMesh

END MODULE QuadMeshClass

MODULE ConservativeShallowWaterClass
  USE ISO_C_BINDING
  USE ModelPrecision
  USE ConstantsDictionary
  USE SWParamsClass
  USE QuadMeshClass
  USE SELF_AMD_CABI
  USE ShallowWaterAMD_Class
  IMPLICIT NONE

  INTEGER, PARAMETER :: nSWEq = 3
  REAL(prec), PARAMETER :: hDefault = ONE

  TYPE ShallowWater
    INTEGER :: nEq = nSWEq, nS = 0, nP = 0, nPlot = 0
    TYPE(SWParams)        :: params
    TYPE(QuadMesh)        :: mesh
    TYPE(ShallowWaterAMD) :: model
    REAL(prec), ALLOCATABLE :: plMatS(:,:), plMatP(:,:)
    LOGICAL :: gradientStale = .TRUE.
  CONTAINS
    PROCEDURE :: Build => Build_ShallowWater
    PROCEDURE :: Trash => Trash_ShallowWater
    PROCEDURE :: GetSolution => GetSolution_ShallowWater
    PROCEDURE :: SetSolution => SetSolution_ShallowWater
    PROCEDURE :: GetTendency => GetTendency_ShallowWater
    PROCEDURE :: GetBathymetry => GetBathymetry_ShallowWater
    PROCEDURE :: SetBathymetry => SetBathymetry_ShallowWater
    PROCEDURE :: SetPlanetaryVorticity => SetPlanetaryVorticity_ShallowWater
    PROCEDURE :: CalculateBathymetryGradient => CalculateBathymetryGradient_ShallowWater
    PROCEDURE :: GlobalTimeDerivative => GlobalTimeDerivative_ShallowWater
    PROCEDURE :: ForwardStepRK3 => ForwardStepRK3_ShallowWater
    PROCEDURE :: CoarseToFine => CoarseToFine_ShallowWater
    PROCEDURE :: WriteTecplot => WriteTecplot_ShallowWater
    PROCEDURE :: WritePickup => WritePickup_ShallowWater
    PROCEDURE :: ReadPickup => ReadPickup_ShallowWater
  END TYPE ShallowWater

CONTAINS

  ! Build_ShallowWater (:163-245): parameters from shallowwater.params, the
  ! model on the device with depth hDefault and zero gradient, then (unless
  ! forceNoPickup is present) the pickup at iterInit and the bathymetry
  ! gradient; the plotting matrices to nPlot + 1 uniform points.
  SUBROUTINE Build_ShallowWater(myDGSEM, forceNoPickup)
    CLASS(ShallowWater), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in), OPTIONAL      :: forceNoPickup
    TYPE(SWParamsAMD) :: p
    INTEGER :: N, nPlot, i
    REAL(C_DOUBLE), ALLOCATABLE :: bw(:), sNew(:)

    CALL myDGSEM % params % Build()
    p % polyDeg = myDGSEM % params % polyDeg
    p % nXElem = myDGSEM % params % nXElem
    p % nYElem = myDGSEM % params % nYElem
    p % xScale = myDGSEM % params % xScale
    p % yScale = myDGSEM % params % yScale
    p % dt = myDGSEM % params % dt
    p % g = myDGSEM % params % g
    p % f0 = myDGSEM % params % f0
    p % betaX = myDGSEM % params % betaX
    p % betaY = myDGSEM % params % betaY
    p % linearDrag = myDGSEM % params % linearDrag
    p % periodicX = .FALSE.
    p % periodicY = .FALSE.
    p % precision = SELF_FP64
    p % form = SELF_FUSED
    p % SpecMeshFile = TRIM(myDGSEM % params % SpecMeshFile)
    p % quirkDMatP = 1  ! the reference's nodal storage, so that a driver reproduces its results
    CALL myDGSEM % model % Build(p)

    N = myDGSEM % params % polyDeg
    myDGSEM % nS = N
    myDGSEM % nP = N
    myDGSEM % mesh % nElems = myDGSEM % model % nElems
    myDGSEM % mesh % nS = N
    myDGSEM % mesh % nP = N
    myDGSEM % mesh % x = myDGSEM % model % x
    myDGSEM % mesh % y = myDGSEM % model % y
    myDGSEM % mesh % dxds = myDGSEM % model % dxds
    myDGSEM % mesh % dxdp = myDGSEM % model % dxdp
    myDGSEM % mesh % dyds = myDGSEM % model % dyds
    myDGSEM % mesh % dydp = myDGSEM % model % dydp
    myDGSEM % mesh % J = myDGSEM % model % J
    myDGSEM % gradientStale = .TRUE.

    IF (.NOT. PRESENT(forceNoPickup)) THEN
      CALL myDGSEM % ReadPickup(myDGSEM % params % iterInit)
      CALL myDGSEM % CalculateBathymetryGradient()
    ENDIF

    nPlot = myDGSEM % params % nPlot
    myDGSEM % nPlot = nPlot
    ALLOCATE(bw(0:N), sNew(0:nPlot), myDGSEM % plMatS(0:nPlot, 0:N), myDGSEM % plMatP(0:nPlot, 0:N))
    DO i = 0, nPlot
      sNew(i) = -ONE + TWO * REAL(i, prec) / REAL(nPlot, prec)
    ENDDO
    CALL self_check(self_barycentric_weights(INT(N, C_INT), myDGSEM % model % nodes, bw), 'Build')
    CALL self_check(self_interpolation_matrix(INT(N, C_INT), myDGSEM % model % nodes, bw, INT(nPlot, C_INT), &
                                              sNew, myDGSEM % plMatS), 'Build')
    myDGSEM % plMatP = myDGSEM % plMatS
  END SUBROUTINE Build_ShallowWater

  SUBROUTINE Trash_ShallowWater(myDGSEM)
    CLASS(ShallowWater), INTENT(inout) :: myDGSEM
    CALL myDGSEM % model % Trash()
    IF (ALLOCATED(myDGSEM % plMatS)) DEALLOCATE(myDGSEM % plMatS, myDGSEM % plMatP)
    IF (ALLOCATED(myDGSEM % mesh % x)) DEALLOCATE(myDGSEM % mesh % x, myDGSEM % mesh % y, myDGSEM % mesh % dxds, &
                                                  myDGSEM % mesh % dxdp, myDGSEM % mesh % dyds, &
                                                  myDGSEM % mesh % dydp, myDGSEM % mesh % J)
  END SUBROUTINE Trash_ShallowWater

  SUBROUTINE GetSolution_ShallowWater(myDGSEM, iEl, theSolution)
    CLASS(ShallowWater), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in)                :: iEl
    REAL(prec), INTENT(out)            :: theSolution(0:myDGSEM % nS, 0:myDGSEM % nP, 1:myDGSEM % nEq)
    CALL myDGSEM % model % GetSolution(iEl, theSolution)
  END SUBROUTINE GetSolution_ShallowWater

  SUBROUTINE SetSolution_ShallowWater(myDGSEM, iEl, theSolution)
    CLASS(ShallowWater), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in)                :: iEl
    REAL(prec), INTENT(in)             :: theSolution(0:myDGSEM % nS, 0:myDGSEM % nP, 1:myDGSEM % nEq)
    CALL myDGSEM % model % SetSolution(iEl, theSolution)
  END SUBROUTINE SetSolution_ShallowWater

  SUBROUTINE GetTendency_ShallowWater(myDGSEM, iEl, theTend)
    CLASS(ShallowWater), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in)                :: iEl
    REAL(prec), INTENT(out)            :: theTend(0:myDGSEM % nS, 0:myDGSEM % nP, 1:myDGSEM % nEq)
    CALL myDGSEM % model % GetTendency(iEl, theTend)
  END SUBROUTINE GetTendency_ShallowWater

  SUBROUTINE GetBathymetry_ShallowWater(myDGSEM, iEl, theBathymetry)
    CLASS(ShallowWater), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in)                :: iEl
    REAL(prec), INTENT(out)            :: theBathymetry(0:myDGSEM % nS, 0:myDGSEM % nP, 1:3)
    CALL myDGSEM % model % GetBathymetry(iEl, theBathymetry)
  END SUBROUTINE GetBathymetry_ShallowWater

  SUBROUTINE SetBathymetry_ShallowWater(myDGSEM, iEl, theBathymetry)
    CLASS(ShallowWater), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in)                :: iEl
    REAL(prec), INTENT(in)             :: theBathymetry(0:myDGSEM % nS, 0:myDGSEM % nP, 1:3)
    CALL myDGSEM % model % SetBathymetry(iEl, theBathymetry)
    myDGSEM % gradientStale = .TRUE.
  END SUBROUTINE SetBathymetry_ShallowWater

  SUBROUTINE SetPlanetaryVorticity_ShallowWater(myDGSEM, iEl, theVorticity)
    CLASS(ShallowWater), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in)                :: iEl
    REAL(prec), INTENT(in)             :: theVorticity(0:myDGSEM % nS, 0:myDGSEM % nP)
    CALL myDGSEM % model % SetPlanetaryVorticity(iEl, theVorticity)
  END SUBROUTINE SetPlanetaryVorticity_ShallowWater

  ! CalculateBathymetryGradient(iEl) (:1733-1868): the device computes every
  ! element at once, so a driver's loop over elements costs one pass; iEl is
  ! accepted for the reference's call form.
  SUBROUTINE CalculateBathymetryGradient_ShallowWater(myDGSEM, iEl)
    CLASS(ShallowWater), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in), OPTIONAL      :: iEl
    IF (myDGSEM % gradientStale) THEN
      CALL myDGSEM % model % CalculateBathymetryGradient()
      myDGSEM % gradientStale = .FALSE.
    ENDIF
  END SUBROUTINE CalculateBathymetryGradient_ShallowWater

  ! GlobalTimeDerivative(tn[, m]): some of the reference's drivers pass the
  ! RK stage m as well; the time derivative does not depend on it.
  SUBROUTINE GlobalTimeDerivative_ShallowWater(myDGSEM, tn, m)
    CLASS(ShallowWater), INTENT(inout) :: myDGSEM
    REAL(prec), INTENT(in)             :: tn
    INTEGER, INTENT(in), OPTIONAL      :: m
    CALL myDGSEM % model % GlobalTimeDerivative(tn)
  END SUBROUTINE GlobalTimeDerivative_ShallowWater

  SUBROUTINE ForwardStepRK3_ShallowWater(myDGSEM, tn)
    CLASS(ShallowWater), INTENT(inout) :: myDGSEM
    REAL(prec), INTENT(in)             :: tn
    CALL myDGSEM % model % ForwardStepRK3(tn)
  END SUBROUTINE ForwardStepRK3_ShallowWater

  ! CoarseToFine (the reference's per-element interpolation to the plotting
  ! points): positions, depth, the three solution fields and the planetary
  ! vorticity of element iEl on (0:nPlot, 0:nPlot), through the batched
  ! CoarseToFine_2D of the C ABI (which returns the reference's orientation).
  SUBROUTINE CoarseToFine_ShallowWater(myDGSEM, iEl, x, y, depth, u, v, eta, vorticity)
    CLASS(ShallowWater), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in)                :: iEl
    REAL(prec), INTENT(out), DIMENSION(0:myDGSEM % nPlot, 0:myDGSEM % nPlot) :: x, y, depth, u, v, eta, vorticity
    REAL(prec) :: sol(0:myDGSEM % nS, 0:myDGSEM % nP, 1:3), h(0:myDGSEM % nS, 0:myDGSEM % nP, 1:3)
    CALL myDGSEM % model % GetSolution(iEl, sol)
    CALL myDGSEM % model % GetBathymetry(iEl, h)
    CALL c2f(myDGSEM % mesh % x(:, :, iEl), x)
    CALL c2f(myDGSEM % mesh % y(:, :, iEl), y)
    CALL c2f(h(:, :, 1), depth)
    CALL c2f(sol(:, :, 1), u)
    CALL c2f(sol(:, :, 2), v)
    CALL c2f(sol(:, :, 3), eta)
    CALL c2f(myDGSEM % model % vorticity(:, :, iEl), vorticity)
  CONTAINS
    SUBROUTINE c2f(fOld, fNew)
      REAL(prec), INTENT(in)  :: fOld(0:myDGSEM % nS, 0:myDGSEM % nP)
      REAL(prec), INTENT(out) :: fNew(0:myDGSEM % nPlot, 0:myDGSEM % nPlot)
      CALL self_check(self_coarse_to_fine_2d(INT(myDGSEM % nS, C_INT), INT(myDGSEM % nP, C_INT), 1_C_SIZE_T, &
                                             myDGSEM % plMatS, myDGSEM % plMatP, INT(myDGSEM % nPlot, C_INT), &
                                             INT(myDGSEM % nPlot, C_INT), fOld, fNew), 'CoarseToFine')
    END SUBROUTINE c2f
  END SUBROUTINE CoarseToFine_ShallowWater

  ! WriteTecplot (:1918-1970): <filename>.tec (or ShallowWater.tec), per
  ! element a zone of X, Y, depth, U/H, V/H and H - depth at the plotting
  ! points.  The whole mesh goes through one batched CoarseToFine per field.
  SUBROUTINE WriteTecplot_ShallowWater(myDGSEM, filename)
    CLASS(ShallowWater), INTENT(inout) :: myDGSEM
    CHARACTER(*), INTENT(in), OPTIONAL :: filename
    INTEGER :: iS, iP, iEl, fUnit, nPlot, N, nE
    CHARACTER(len=5) :: zoneID
    REAL(prec), ALLOCATABLE :: sol(:,:,:,:), h(:,:,:,:), xf(:,:,:), yf(:,:,:), df(:,:,:), uf(:,:,:), vf(:,:,:), &
                               ef(:,:,:), tmp(:,:,:)
    nPlot = myDGSEM % nPlot
    N = myDGSEM % nS
    nE = myDGSEM % mesh % nElems
    ALLOCATE(sol(0:N, 0:N, 3, nE), h(0:N, 0:N, 3, nE), tmp(0:N, 0:N, nE))
    ALLOCATE(xf(0:nPlot, 0:nPlot, nE), yf(0:nPlot, 0:nPlot, nE), df(0:nPlot, 0:nPlot, nE))
    ALLOCATE(uf(0:nPlot, 0:nPlot, nE), vf(0:nPlot, 0:nPlot, nE), ef(0:nPlot, 0:nPlot, nE))
    DO iEl = 1, nE
      CALL myDGSEM % model % GetSolution(iEl, sol(:, :, :, iEl))
      CALL myDGSEM % model % GetBathymetry(iEl, h(:, :, :, iEl))
    ENDDO
    CALL c2f_all(myDGSEM % mesh % x, xf)
    CALL c2f_all(myDGSEM % mesh % y, yf)
    tmp = h(:, :, 1, :)
    CALL c2f_all(tmp, df)
    tmp = sol(:, :, 1, :)
    CALL c2f_all(tmp, uf)
    tmp = sol(:, :, 2, :)
    CALL c2f_all(tmp, vf)
    tmp = sol(:, :, 3, :)
    CALL c2f_all(tmp, ef)
    IF (PRESENT(filename)) THEN
      OPEN(NEWUNIT=fUnit, FILE=TRIM(filename)//'.tec', FORM='formatted', STATUS='replace')
    ELSE
      OPEN(NEWUNIT=fUnit, FILE='ShallowWater.tec', FORM='formatted', STATUS='replace')
    ENDIF
    WRITE(fUnit, *) 'VARIABLES = "X", "Y", "Depth" "U", "V", "Eta" '
    DO iEl = 1, nE
      WRITE(zoneID, '(I5.5)') iEl
      WRITE(fUnit, *) 'ZONE T="el'//TRIM(zoneID)//'", I=', nPlot + 1, ', J=', nPlot + 1, ',F=POINT'
      DO iP = 0, nPlot
        DO iS = 0, nPlot
          WRITE(fUnit, *) xf(iS, iP, iEl), yf(iS, iP, iEl), df(iS, iP, iEl), uf(iS, iP, iEl) / ef(iS, iP, iEl), &
                          vf(iS, iP, iEl) / ef(iS, iP, iEl), ef(iS, iP, iEl) - df(iS, iP, iEl)
        ENDDO
      ENDDO
    ENDDO
    CLOSE(fUnit)
  CONTAINS
    SUBROUTINE c2f_all(fOld, fNew)
      REAL(prec), INTENT(in)  :: fOld(0:N, 0:N, nE)
      REAL(prec), INTENT(out) :: fNew(0:nPlot, 0:nPlot, nE)
      CALL self_check(self_coarse_to_fine_2d(INT(N, C_INT), INT(N, C_INT), INT(nE, C_SIZE_T), myDGSEM % plMatS, &
                                             myDGSEM % plMatP, INT(nPlot, C_INT), INT(nPlot, C_INT), fOld, fNew), &
                      'WriteTecplot')
    END SUBROUTINE c2f_all
  END SUBROUTINE WriteTecplot_ShallowWater

  SUBROUTINE WritePickup_ShallowWater(myDGSEM, iter)
    CLASS(ShallowWater), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in)                :: iter
    CALL myDGSEM % model % WritePickup(iter)
  END SUBROUTINE WritePickup_ShallowWater

  SUBROUTINE ReadPickup_ShallowWater(myDGSEM, iter)
    CLASS(ShallowWater), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in)                :: iter
    CALL myDGSEM % model % ReadPickup(iter)
    myDGSEM % gradientStale = .TRUE.
  END SUBROUTINE ReadPickup_ShallowWater

END MODULE ConservativeShallowWaterClass
