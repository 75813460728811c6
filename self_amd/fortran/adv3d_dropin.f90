! adv3d_dropin.f90 -- the 3-D advection model under the reference's own
! module, type and procedure names, so that the reference's driver and
! initial-condition programs compile unchanged against it and run on the GPU.
!
! Modules (each mirrors the reference module of the same name, as far as the
! model's programs use it; ModelPrecision and ConstantsDictionary come from
! sw_dropin.f90):
!   Timing                src/common/Timing.f90: TYPE MultiTimers with Build,
!                         Trash, AddTimer, StartThisTimer, StopThisTimer
!                         (which also accumulates, :408-421),
!                         AccumulateTimings and Write_MultiTimers
!                         (Timing.stats); timers are wall-clock, since the
!                         model's work runs on the device
!   NodalStorage_3D_Class src/nodal/NodalStorage_3D_Class.f90: the nodes,
!                         weights, DG derivative matrix and boundary
!                         interpolants of degree N (GAUSS, DG), with
!                         GetNumberOfNodes and GetQuadratureWeights
!   HexMeshClass          src/geom/HexMeshClass.f90: the default structured
!                         mesh (LoadDefaultMesh + ScaleTheMesh, walled, every
!                         boundary face NO_NORMAL_FLOW) with GetPositionAtNode,
!                         GetPositions, GetJacobian, GetJacobianAtNode,
!                         GetCovariantMetricsAtNode and WriteTecplot
!                         (:2239-2299)
!   AdvectionParamsClass  src/highend/advection3d/Advection3DParamsClass.f90
!                         (the advection.params namelists TracerSetup,
!                         TimeManagement, SpaceManagement and their defaults)
!   Advection3DClass      src/highend/advection3d/Advection3DClass.f90: TYPE
!                         Advection with Build(forceNoPickup) (:167-272),
!                         Trash, the per-element accessors,
!                         GlobalTimeDerivative(tn, m) (:1049-1121),
!                         ForwardStepRK3 (:1000-1045), CoarseToFine
!                         (:1747-1842), WriteTecplot (:1846-1901), WritePickup
!                         (:1905-1972) and ReadPickup (:1976-2047).
! The model itself is the C ABI's (self_adv3d_*).  The reference's
! SetRelaxationFactor and GetRelaxationFactor(AtNode) use the relaxation
! field's storage (:898-962), so the relaxation source is (r - c) r: one host
! array stands for both and every upload writes it to both device fields.
! The element arrays a program reads and writes (solution, velocity,
! relaxation) are mirrored on the host: a Set marks the host copy newer, a
! device step marks it older, and each side is refreshed only when the other
! is used.  The time derivative filters the solution in place (the adaptive
! roll-off filter, :1125-1233), so it too marks the host solution older.
MODULE Timing
  USE ModelPrecision
  USE ConstantsDictionary
  IMPLICIT NONE

  INTEGER, PARAMETER :: timerNameLength = 30
  INTEGER, PARAMETER, PRIVATE :: maxTimers = 64

  TYPE MultiTimers
    INTEGER :: nTimers = 0, current = 0
    INTEGER :: timerID(maxTimers) = 0
    CHARACTER(timerNameLength) :: timerName(maxTimers)
    LOGICAL :: stopped(maxTimers) = .FALSE.
    REAL(prec) :: startTime(maxTimers) = ZERO, stopTime(maxTimers) = ZERO
    REAL(prec) :: accumulatedTime(maxTimers) = ZERO, nObs(maxTimers) = ZERO
  CONTAINS
    PROCEDURE :: Build => Build_MultiTimers
    PROCEDURE :: Trash => Trash_MultiTimers
    PROCEDURE :: SetName => SetThisTimerName
    PROCEDURE :: GetName => GetThisTimerName
    PROCEDURE :: ThereAreNoTimers
    PROCEDURE :: AddTimer
    PROCEDURE :: PointToTimer
    PROCEDURE :: StartThisTimer
    PROCEDURE :: StopThisTimer
    PROCEDURE :: AccumulateTimings
    PROCEDURE :: Write_MultiTimers
  END TYPE MultiTimers

CONTAINS

  REAL(prec) FUNCTION wall_seconds()
    INTEGER(8) :: count, rate
    CALL SYSTEM_CLOCK(count, rate)
    wall_seconds = REAL(count, prec) / REAL(rate, prec)
  END FUNCTION wall_seconds

  SUBROUTINE Build_MultiTimers(theTimers)
    CLASS(MultiTimers) :: theTimers
    theTimers % nTimers = 0
    theTimers % current = 0
  END SUBROUTINE Build_MultiTimers

  SUBROUTINE Trash_MultiTimers(theTimers)
    CLASS(MultiTimers) :: theTimers
    theTimers % nTimers = 0
    theTimers % current = 0
  END SUBROUTINE Trash_MultiTimers

  LOGICAL FUNCTION ThereAreNoTimers(theTimers)
    CLASS(MultiTimers) :: theTimers
    ThereAreNoTimers = theTimers % nTimers == 0
  END FUNCTION ThereAreNoTimers

  ! AddTimer (:286-344): a new timer, named and keyed, becomes current
  SUBROUTINE AddTimer(theTimers, timername, inKey)
    CLASS(MultiTimers), INTENT(inout) :: theTimers
    CHARACTER(*) :: timername
    INTEGER :: inKey
    IF (theTimers % nTimers == maxTimers) THEN
      PRINT *, 'MODULE Timing.f90 : S/R AddTimer : no room for another timer.'
      RETURN
    ENDIF
    theTimers % nTimers = theTimers % nTimers + 1
    theTimers % current = theTimers % nTimers
    theTimers % timerID(theTimers % current) = inKey
    CALL theTimers % SetName(timername)
  END SUBROUTINE AddTimer

  ! the first timer keyed tID becomes current (none: no timer is current)
  SUBROUTINE PointToTimer(theTimers, tID)
    CLASS(MultiTimers) :: theTimers
    INTEGER, INTENT(in) :: tID
    INTEGER :: k
    theTimers % current = 0
    DO k = 1, theTimers % nTimers
      IF (theTimers % timerID(k) == tID) THEN
        theTimers % current = k
        RETURN
      ENDIF
    ENDDO
  END SUBROUTINE PointToTimer

  SUBROUTINE SetThisTimerName(theTimers, nameInput, tID)
    CLASS(MultiTimers) :: theTimers
    CHARACTER(*), INTENT(in) :: nameInput
    INTEGER, OPTIONAL, INTENT(in) :: tID
    IF (PRESENT(tID)) CALL theTimers % PointToTimer(tID)
    IF (theTimers % current == 0) RETURN
    theTimers % timerName(theTimers % current) = nameInput
    theTimers % nObs(theTimers % current) = ZERO
  END SUBROUTINE SetThisTimerName

  SUBROUTINE GetThisTimerName(theTimers, nameOutput, tID)
    CLASS(MultiTimers) :: theTimers
    CHARACTER(*), INTENT(out) :: nameOutput
    INTEGER, OPTIONAL, INTENT(in) :: tID
    IF (PRESENT(tID)) CALL theTimers % PointToTimer(tID)
    nameOutput = ' '
    IF (theTimers % current > 0) nameOutput = theTimers % timerName(theTimers % current)
  END SUBROUTINE GetThisTimerName

  SUBROUTINE StartThisTimer(theTimers, tID)
    CLASS(MultiTimers), INTENT(inout) :: theTimers
    INTEGER, INTENT(in) :: tID
    CALL theTimers % PointToTimer(tID)
    IF (theTimers % current == 0) RETURN
    theTimers % stopped(theTimers % current) = .FALSE.
    theTimers % startTime(theTimers % current) = wall_seconds()
  END SUBROUTINE StartThisTimer

  ! StopThisTimer (:408-421) accumulates the interval itself, so a program
  ! that calls AccumulateTimings after it (as the reference's drivers do)
  ! counts the interval twice, as with the reference's module.
  SUBROUTINE StopThisTimer(theTimers, tID)
    CLASS(MultiTimers), INTENT(inout) :: theTimers
    INTEGER, INTENT(in) :: tID
    CALL theTimers % PointToTimer(tID)
    IF (theTimers % current == 0) RETURN
    theTimers % stopTime(theTimers % current) = wall_seconds()
    theTimers % stopped(theTimers % current) = .TRUE.
    CALL theTimers % AccumulateTimings()
  END SUBROUTINE StopThisTimer

  SUBROUTINE AccumulateTimings(theTimers, tID)
    CLASS(MultiTimers) :: theTimers
    INTEGER, INTENT(in), OPTIONAL :: tID
    INTEGER :: k
    IF (PRESENT(tID)) CALL theTimers % PointToTimer(tID)
    k = theTimers % current
    IF (k == 0) RETURN
    IF (theTimers % stopped(k)) THEN
      theTimers % accumulatedTime(k) = theTimers % accumulatedTime(k) + theTimers % stopTime(k) - &
                                       theTimers % startTime(k)
    ELSE
      PRINT *, 'Module Timing.f90 : S/R ElapsedTime : Warning! Timer "', TRIM(theTimers % timerName(k)), &
               '" is not stopped'
    ENDIF
    theTimers % nObs(k) = theTimers % nObs(k) + ONE
  END SUBROUTINE AccumulateTimings

  ! Write_MultiTimers (:509-567): Timing.stats and the same table on
  ! standard output.  Timers that never ran report an average of zero.
  SUBROUTINE Write_MultiTimers(theTimers)
    CLASS(MultiTimers), INTENT(inout) :: theTimers
    INTEGER :: k, fUnit
    OPEN(NEWUNIT=fUnit, FILE='Timing.stats')
    CALL table(fUnit)
    CLOSE(fUnit)
    CALL table(6)
    PRINT *, '------------------------------------------------------------'
    PRINT *, ' Timing Results saved to  "Timing.Stats" '
  CONTAINS
    SUBROUTINE table(u)
      INTEGER, INTENT(in) :: u
      REAL(prec) :: avg
      WRITE(u, *) '====================== Timing Results ======================'
      WRITE(u, *) ' '
      DO k = 1, theTimers % nTimers
        avg = ZERO
        IF (theTimers % nObs(k) > ZERO) avg = theTimers % accumulatedTime(k) / theTimers % nObs(k)
        WRITE(u, *) theTimers % timerName(k)
        WRITE(u, *) 'Number of Measurements : ', theTimers % nObs(k)
        WRITE(u, *) 'Accumulated Time       : ', theTimers % accumulatedTime(k)
        WRITE(u, *) 'Average Time           : ', avg
        WRITE(u, *) '------------------------------------------------------------'
      ENDDO
    END SUBROUTINE table
  END SUBROUTINE Write_MultiTimers

END MODULE Timing

MODULE NodalStorage_3D_Class
  USE ISO_C_BINDING
  USE ModelPrecision
  USE SELF_AMD_CABI
  IMPLICIT NONE

  TYPE NodalStorage_3D
    INTEGER :: N = -1
    REAL(prec), ALLOCATABLE :: nodes(:), qWeight(:), dMat(:,:), lagLeft(:), lagRight(:)
  CONTAINS
    PROCEDURE :: Build => Build_NodalStorage_3D
    PROCEDURE :: Trash => Trash_NodalStorage_3D
    PROCEDURE :: GetNumberOfNodes => GetNumberOfNodes_NodalStorage_3D
    PROCEDURE :: GetQuadratureWeights => GetQuadratureWeights_NodalStorage_3D
  END TYPE NodalStorage_3D

CONTAINS

  ! Build (:107-264) with equal degree in s, p and q
  SUBROUTINE Build_NodalStorage_3D(myNodal, nS, nP, nQ, quadrature, approxForm)
    CLASS(NodalStorage_3D), INTENT(inout) :: myNodal
    INTEGER, INTENT(in) :: nS, nP, nQ, quadrature, approxForm
    INTEGER(C_INT) :: N
    IF (nP /= nS .OR. nQ /= nS) CALL self_check(-1_C_INT, 'NodalStorage_3D % Build: nS = nP = nQ only')
    N = nS
    CALL myNodal % Trash()
    myNodal % N = N
    ALLOCATE(myNodal % nodes(0:N), myNodal % qWeight(0:N), myNodal % dMat(0:N, 0:N), myNodal % lagLeft(0:N), &
             myNodal % lagRight(0:N))
    CALL self_check(self_nodal3d_build(N, INT(quadrature, C_INT), INT(approxForm, C_INT), myNodal % nodes, &
                                       myNodal % qWeight, myNodal % dMat, myNodal % lagLeft, myNodal % lagRight), &
                    'NodalStorage_3D % Build')
  END SUBROUTINE Build_NodalStorage_3D

  SUBROUTINE Trash_NodalStorage_3D(myNodal)
    CLASS(NodalStorage_3D), INTENT(inout) :: myNodal
    IF (ALLOCATED(myNodal % nodes)) DEALLOCATE(myNodal % nodes, myNodal % qWeight, myNodal % dMat, &
                                               myNodal % lagLeft, myNodal % lagRight)
    myNodal % N = -1
  END SUBROUTINE Trash_NodalStorage_3D

  SUBROUTINE GetNumberOfNodes_NodalStorage_3D(myNodal, nS, nP, nQ)
    CLASS(NodalStorage_3D), INTENT(in) :: myNodal
    INTEGER, INTENT(out) :: nS, nP, nQ
    nS = myNodal % N
    nP = myNodal % N
    nQ = myNodal % N
  END SUBROUTINE GetNumberOfNodes_NodalStorage_3D

  SUBROUTINE GetQuadratureWeights_NodalStorage_3D(myNodal, wS, wP, wQ)
    CLASS(NodalStorage_3D), INTENT(in) :: myNodal
    REAL(prec), INTENT(out) :: wS(0:myNodal % N), wP(0:myNodal % N), wQ(0:myNodal % N)
    wS = myNodal % qWeight
    wP = myNodal % qWeight
    wQ = myNodal % qWeight
  END SUBROUTINE GetQuadratureWeights_NodalStorage_3D

END MODULE NodalStorage_3D_Class

MODULE HexMeshClass
  USE ISO_C_BINDING
  USE ModelPrecision
  USE ConstantsDictionary
  USE SELF_AMD_CABI
  IMPLICIT NONE

  ! Per element: positions and Jacobian (0:N,0:N,0:N); covariant metric
  ! terms (9, 0:N,0:N,0:N) in the order dxds dxdp dxdq dyds dydp dydq dzds
  ! dzdp dzdq.  The face table (9, nFaces) and the boundary geometry the
  ! device model is built from: nHat and bpos (3, 0:N,0:N, 6, nElems), nLen
  ! (0:N,0:N, 6, nElems).
  TYPE HexMesh
    INTEGER :: nElems = 0, nFaces = 0, N = -1
    REAL(prec), ALLOCATABLE :: x(:,:,:,:), y(:,:,:,:), z(:,:,:,:), J(:,:,:,:), cov(:,:,:,:,:)
    INTEGER(C_INT), ALLOCATABLE :: faces(:,:)
    REAL(prec), ALLOCATABLE :: nHat(:,:,:,:,:), nLen(:,:,:,:), bpos(:,:,:,:,:)
  CONTAINS
    PROCEDURE :: LoadDefaultMesh => LoadDefaultMesh_HexMesh
    PROCEDURE :: Trash => Trash_HexMesh
    PROCEDURE :: GetNumberOfElements => GetNumberOfElements_HexMesh
    PROCEDURE :: GetNumberOfFaces => GetNumberOfFaces_HexMesh
    PROCEDURE :: GetPositions => GetPositions_HexMesh
    PROCEDURE :: GetPositionAtNode => GetPositionAtNode_HexMesh
    PROCEDURE :: GetJacobian => GetJacobian_HexMesh
    PROCEDURE :: GetJacobianAtNode => GetJacobianAtNode_HexMesh
    PROCEDURE :: GetCovariantMetricsAtNode => GetCovariantMetricsAtNode_HexMesh
    PROCEDURE :: WriteTecplot => WriteTecplot_HexMesh
  END TYPE HexMesh

CONTAINS

  ! LoadDefaultMesh followed by ScaleTheMesh: nX x nY x nZ elements on the
  ! unit cube, scaled by (xScale, yScale, zScale); nodes are the degree-N
  ! interpolation nodes.
  SUBROUTINE LoadDefaultMesh_HexMesh(myMesh, N, nodes, nX, nY, nZ, xScale, yScale, zScale)
    CLASS(HexMesh), INTENT(inout) :: myMesh
    INTEGER, INTENT(in) :: N, nX, nY, nZ
    REAL(prec), INTENT(in) :: nodes(0:N), xScale, yScale, zScale
    INTEGER(C_INT) :: nE, nF, warpSign(3)
    CALL myMesh % Trash()
    CALL self_check(self_hexmesh_counts(INT(nX, C_INT), INT(nY, C_INT), INT(nZ, C_INT), 0_C_INT, 0_C_INT, 0_C_INT, &
                                        nE, nF), 'LoadDefaultMesh')
    myMesh % nElems = nE
    myMesh % nFaces = nF
    myMesh % N = N
    ALLOCATE(myMesh % x(0:N, 0:N, 0:N, nE), myMesh % y(0:N, 0:N, 0:N, nE), myMesh % z(0:N, 0:N, 0:N, nE), &
             myMesh % J(0:N, 0:N, 0:N, nE), myMesh % cov(9, 0:N, 0:N, 0:N, nE), myMesh % faces(9, nF), &
             myMesh % nHat(3, 0:N, 0:N, 6, nE), myMesh % nLen(0:N, 0:N, 6, nE), myMesh % bpos(3, 0:N, 0:N, 6, nE))
    warpSign = 1
    CALL self_check(self_hexmesh_structured(INT(nX, C_INT), INT(nY, C_INT), INT(nZ, C_INT), INT(N, C_INT), 0_C_INT, &
                                            0_C_INT, 0_C_INT, xScale, yScale, zScale, ZERO, warpSign, nodes, &
                                            myMesh % faces, myMesh % x, myMesh % y, myMesh % z, myMesh % cov, &
                                            myMesh % J, myMesh % nHat, myMesh % nLen, myMesh % bpos), 'LoadDefaultMesh')
  END SUBROUTINE LoadDefaultMesh_HexMesh

  SUBROUTINE Trash_HexMesh(myMesh)
    CLASS(HexMesh), INTENT(inout) :: myMesh
    IF (ALLOCATED(myMesh % x)) DEALLOCATE(myMesh % x, myMesh % y, myMesh % z, myMesh % J, myMesh % cov, &
                                          myMesh % faces, myMesh % nHat, myMesh % nLen, myMesh % bpos)
    myMesh % nElems = 0
    myMesh % nFaces = 0
  END SUBROUTINE Trash_HexMesh

  SUBROUTINE GetNumberOfElements_HexMesh(myMesh, nElems)
    CLASS(HexMesh), INTENT(in) :: myMesh
    INTEGER, INTENT(out) :: nElems
    nElems = myMesh % nElems
  END SUBROUTINE GetNumberOfElements_HexMesh

  SUBROUTINE GetNumberOfFaces_HexMesh(myMesh, nFaces)
    CLASS(HexMesh), INTENT(in) :: myMesh
    INTEGER, INTENT(out) :: nFaces
    nFaces = myMesh % nFaces
  END SUBROUTINE GetNumberOfFaces_HexMesh

  SUBROUTINE GetPositions_HexMesh(myMesh, iEl, x, y, z)
    CLASS(HexMesh), INTENT(in) :: myMesh
    INTEGER, INTENT(in) :: iEl
    REAL(prec), INTENT(out), DIMENSION(0:myMesh % N, 0:myMesh % N, 0:myMesh % N) :: x, y, z
    x = myMesh % x(:, :, :, iEl)
    y = myMesh % y(:, :, :, iEl)
    z = myMesh % z(:, :, :, iEl)
  END SUBROUTINE GetPositions_HexMesh

  SUBROUTINE GetPositionAtNode_HexMesh(myMesh, iEl, x, y, z, i, j, k)
    CLASS(HexMesh), INTENT(in) :: myMesh
    INTEGER, INTENT(in) :: iEl, i, j, k
    REAL(prec), INTENT(out) :: x, y, z
    x = myMesh % x(i, j, k, iEl)
    y = myMesh % y(i, j, k, iEl)
    z = myMesh % z(i, j, k, iEl)
  END SUBROUTINE GetPositionAtNode_HexMesh

  SUBROUTINE GetJacobian_HexMesh(myMesh, iEl, J)
    CLASS(HexMesh), INTENT(in) :: myMesh
    INTEGER, INTENT(in) :: iEl
    REAL(prec), INTENT(out) :: J(0:myMesh % N, 0:myMesh % N, 0:myMesh % N)
    J = myMesh % J(:, :, :, iEl)
  END SUBROUTINE GetJacobian_HexMesh

  SUBROUTINE GetJacobianAtNode_HexMesh(myMesh, iEl, J, iS, iP, iQ)
    CLASS(HexMesh), INTENT(in) :: myMesh
    INTEGER, INTENT(in) :: iEl, iS, iP, iQ
    REAL(prec), INTENT(out) :: J
    J = myMesh % J(iS, iP, iQ, iEl)
  END SUBROUTINE GetJacobianAtNode_HexMesh

  SUBROUTINE GetCovariantMetricsAtNode_HexMesh(myMesh, iEl, dxds, dxdp, dxdq, dyds, dydp, dydq, dzds, dzdp, dzdq, &
                                               i, j, k)
    CLASS(HexMesh), INTENT(in) :: myMesh
    INTEGER, INTENT(in) :: iEl, i, j, k
    REAL(prec), INTENT(out) :: dxds, dxdp, dxdq, dyds, dydp, dydq, dzds, dzdp, dzdq
    dxds = myMesh % cov(1, i, j, k, iEl)
    dxdp = myMesh % cov(2, i, j, k, iEl)
    dxdq = myMesh % cov(3, i, j, k, iEl)
    dyds = myMesh % cov(4, i, j, k, iEl)
    dydp = myMesh % cov(5, i, j, k, iEl)
    dydq = myMesh % cov(6, i, j, k, iEl)
    dzds = myMesh % cov(7, i, j, k, iEl)
    dzdp = myMesh % cov(8, i, j, k, iEl)
    dzdq = myMesh % cov(9, i, j, k, iEl)
  END SUBROUTINE GetCovariantMetricsAtNode_HexMesh

  ! WriteTecplot (:2239-2299): <filename>.tec (or mesh.tec), per element a
  ! zone of X, Y, Z, the Jacobian and the nine metric terms at the nodes
  SUBROUTINE WriteTecplot_HexMesh(myMesh, filename)
    CLASS(HexMesh), INTENT(inout) :: myMesh
    CHARACTER(*), INTENT(in), OPTIONAL :: filename
    INTEGER :: iS, iP, iQ, iEl, fUnit, N
    CHARACTER(7) :: zoneID
    N = myMesh % N
    IF (PRESENT(filename)) THEN
      OPEN(NEWUNIT=fUnit, FILE=TRIM(filename)//'.tec', FORM='formatted')
    ELSE
      OPEN(NEWUNIT=fUnit, FILE='mesh.tec', FORM='formatted')
    ENDIF
    WRITE(fUnit, *) 'VARIABLES = "X", "Y", "Z", "Jacobian", "dxds", "dxdp", "dxdq", "dyds", "dydp", "dydq", "dzds", '// &
                    '"dzdp", "dzdq" '
    DO iEl = 1, myMesh % nElems
      WRITE(zoneID, '(I7.7)') iEl
      WRITE(fUnit, *) 'ZONE T="el'//TRIM(zoneID)//'", I=', N + 1, ', J=', N + 1, ', K=', N + 1, ',F=POINT'
      DO iQ = 0, N
        DO iP = 0, N
          DO iS = 0, N
            WRITE(fUnit, *) myMesh % x(iS, iP, iQ, iEl), myMesh % y(iS, iP, iQ, iEl), myMesh % z(iS, iP, iQ, iEl), &
                            myMesh % J(iS, iP, iQ, iEl), myMesh % cov(:, iS, iP, iQ, iEl)
          ENDDO
        ENDDO
      ENDDO
    ENDDO
    CLOSE(UNIT=fUnit)
  END SUBROUTINE WriteTecplot_HexMesh

END MODULE HexMeshClass

MODULE AdvectionParamsClass
  USE ModelPrecision
  USE ConstantsDictionary
  IMPLICIT NONE
  CHARACTER(1), PARAMETER, PRIVATE :: blank = ' '

  TYPE AdvectionParams
    ! TracerSetup
    INTEGER :: nTracers
    ! TimeManagement
    REAL(prec) :: dt
    INTEGER    :: iterInit, nTimeSteps, dumpFreq
    ! SpaceManagement
    CHARACTER(50) :: SpecMeshFile
    INTEGER       :: polyDeg, nCutoff, nXElem, nYElem, nZElem, geomPolyDeg, nPlot
    REAL(prec)    :: dxPlot, xScale, yScale, zScale
  CONTAINS
    PROCEDURE :: Build => Build_AdvectionParams
  END TYPE AdvectionParams

CONTAINS

  SUBROUTINE Build_AdvectionParams(thisParam)
    CLASS(AdvectionParams), INTENT(out) :: thisParam
    INTEGER       :: nUnit
    INTEGER       :: nTracers, iterInit, nTimeSteps, dumpFreq, polyDeg, nCutoff, nXElem, nYElem, nZElem, nPlot
    REAL(prec)    :: dt, xScale, yScale, zScale
    CHARACTER(50) :: SpecMeshFile
    NAMELIST / TracerSetup / nTracers
    NAMELIST / TimeManagement / dt, iterInit, nTimeSteps, dumpFreq
    NAMELIST / SpaceManagement / SpecMeshFile, polyDeg, nCutoff, nXElem, nYElem, nZElem, nPlot, xScale, yScale, &
                                 zScale

    nTracers = 2
    dt = ZERO
    iterInit = 0
    nTimeSteps = 0
    dumpFreq = 1
    SpecMeshFile = blank
    polyDeg = 5
    nCutoff = 3
    nXElem = 2
    nYElem = 2
    nZElem = 2
    nPlot = 10
    xScale = ONE
    yScale = ONE
    zScale = ONE

    OPEN(NEWUNIT=nUnit, FILE='advection.params', STATUS='old', ACTION='read')
    READ(UNIT=nUnit, NML=TracerSetup)
    READ(UNIT=nUnit, NML=TimeManagement)
    READ(UNIT=nUnit, NML=SpaceManagement)
    CLOSE(UNIT=nUnit)
    WRITE(UNIT=*, NML=TracerSetup)
    WRITE(UNIT=*, NML=TimeManagement)
    WRITE(UNIT=*, NML=SpaceManagement)

    thisParam % nTracers = nTracers
    thisParam % dt = dt
    thisParam % iterInit = iterInit
    thisParam % nTimeSteps = nTimeSteps
    thisParam % dumpFreq = dumpFreq
    thisParam % SpecMeshFile = SpecMeshFile
    thisParam % polyDeg = polyDeg
    thisParam % nCutoff = nCutoff
    thisParam % nXElem = nXElem
    thisParam % nYElem = nYElem
    thisParam % nZElem = nZElem
    thisParam % geomPolyDeg = polyDeg
    thisParam % nPlot = nPlot
    thisParam % dxPlot = 2.0_prec / REAL(nPlot, prec)
    thisParam % xScale = xScale
    thisParam % yScale = yScale
    thisParam % zScale = zScale
  END SUBROUTINE Build_AdvectionParams

END MODULE AdvectionParamsClass

MODULE Advection3DClass
  USE ISO_C_BINDING
  USE ModelPrecision
  USE ConstantsDictionary
  USE Timing
  USE NodalStorage_3D_Class
  USE HexMeshClass
  USE AdvectionParamsClass
  USE SELF_AMD_CABI
  IMPLICIT NONE

  ! fields of self_adv3d_set_field / self_adv3d_get_field
  INTEGER(C_INT), PARAMETER, PRIVATE :: F_SOL = 0, F_VEL = 1, F_RELAX = 2, F_RFAC = 3
  INTEGER(C_INT), PARAMETER, PRIVATE :: G_SOL = 0, G_TEND = 1
  ! the host mirrors: solution, velocity, relaxation (field and factor)
  INTEGER, PARAMETER, PRIVATE :: M_SOL = 1, M_VEL = 2, M_RELAX = 3

  TYPE Advection
    INTEGER                :: nEq, nPlot, nS, nP, nQ
    TYPE(HexMesh)          :: mesh
    TYPE(NodalStorage_3D)  :: dgStorage
    TYPE(AdvectionParams)  :: params
    TYPE(MultiTimers)      :: clocks
    REAL(prec), ALLOCATABLE :: plMatS(:,:), plMatP(:,:), plMatQ(:,:)
    TYPE(C_PTR) :: handle = C_NULL_PTR
    ! host mirrors (0:N,0:N,0:N, 1:nEq or 3, nElems) and their state: 0
    ! equal to the device, 1 host newer (upload before the device uses it),
    ! -1 device newer (download before the host reads it)
    REAL(prec), ALLOCATABLE :: sol(:,:,:,:,:), vel(:,:,:,:,:), relax(:,:,:,:,:)
    INTEGER :: state(3) = 0
  CONTAINS
    PROCEDURE :: Build => Build_Advection
    PROCEDURE :: Trash => Trash_Advection
    PROCEDURE :: GetSolution => GetSolution_Advection
    PROCEDURE :: SetSolution => SetSolution_Advection
    PROCEDURE :: GetSolutionWithVarID => GetSolutionWithVarID_Advection
    PROCEDURE :: SetSolutionWithVarID => SetSolutionWithVarID_Advection
    PROCEDURE :: GetSolutionAtNode => GetSolutionAtNode_Advection
    PROCEDURE :: SetSolutionAtNode => SetSolutionAtNode_Advection
    PROCEDURE :: GetTendency => GetTendency_Advection
    PROCEDURE :: GetVelocity => GetVelocity_Advection
    PROCEDURE :: SetVelocity => SetVelocity_Advection
    PROCEDURE :: GetRelaxationField => GetRelaxationField_Advection
    PROCEDURE :: SetRelaxationField => SetRelaxationField_Advection
    PROCEDURE :: GetRelaxationFactor => GetRelaxationFactor_Advection
    PROCEDURE :: SetRelaxationFactor => SetRelaxationFactor_Advection
    PROCEDURE :: GlobalTimeDerivative => GlobalTimeDerivative_Advection
    PROCEDURE :: ForwardStepRK3 => ForwardStepRK3_Advection
    PROCEDURE :: CoarseToFine => CoarseToFine_Advection
    PROCEDURE :: WriteTecplot => WriteTecplot_Advection
    PROCEDURE :: WritePickup => WritePickup_Advection
    PROCEDURE :: ReadPickup => ReadPickup_Advection
  END TYPE Advection

CONTAINS

  ! Build (:167-272): parameters from advection.params; the clocks the
  ! reference registers (File-I/O is timer 0; the four phases of the time
  ! derivative, timers 1-4, run on the device and stay unmeasured here);
  ! nodal storage (GAUSS, DG) and the roll-off filter of cutoff nCutoff; the
  ! default mesh scaled by xScale, yScale, zScale (BuildHexMesh :316-341);
  ! the model on the device with zero fields; unless forceNoPickup is
  ! present, the pickup at iterInit; the plotting matrix to nPlot + 1
  ! uniform points.
  SUBROUTINE Build_Advection(myDGSEM, forceNoPickup)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in), OPTIONAL   :: forceNoPickup
    INTEGER(C_INT) :: N, nE
    REAL(C_DOUBLE), ALLOCATABLE :: F1(:,:), bw(:), sNew(:)
    INTEGER :: i, nPlot

    CALL myDGSEM % params % Build()
    CALL myDGSEM % clocks % Build()
    CALL myDGSEM % clocks % AddTimer('DoTheAdaptiveFilter', 1)
    CALL myDGSEM % clocks % AddTimer('CalculateSolutionAtBoundaries', 2)
    CALL myDGSEM % clocks % AddTimer('EdgeFlux', 3)
    CALL myDGSEM % clocks % AddTimer('MappedTimeDerivative', 4)
    CALL myDGSEM % clocks % AddTimer('File-I/O', 0)

    N = myDGSEM % params % polyDeg
    myDGSEM % nS = N
    myDGSEM % nP = N
    myDGSEM % nQ = N
    myDGSEM % nEq = myDGSEM % params % nTracers
    myDGSEM % nPlot = myDGSEM % params % nPlot
    nPlot = myDGSEM % nPlot

    CALL myDGSEM % dgStorage % Build(N, N, N, INT(SELF_GAUSS), INT(SELF_DG))
    ALLOCATE(F1(0:N, 0:N))
    CALL self_check(self_rolloff_filter1d(N, INT(myDGSEM % params % nCutoff, C_INT), myDGSEM % dgStorage % nodes, &
                                          myDGSEM % dgStorage % qWeight, F1), 'Build')

    PRINT *, 'Module AdvectionClass.f90 : S/R BuildQuadMesh :'
    PRINT *, ' Loading default mesh.'
    CALL myDGSEM % mesh % LoadDefaultMesh(INT(N), myDGSEM % dgStorage % nodes, myDGSEM % params % nXElem, &
                                          myDGSEM % params % nYElem, myDGSEM % params % nZElem, &
                                          myDGSEM % params % xScale, myDGSEM % params % yScale, &
                                          myDGSEM % params % zScale)
    nE = myDGSEM % mesh % nElems
    CALL self_check(self_adv3d_create(N, INT(myDGSEM % nEq, C_INT), nE, INT(myDGSEM % mesh % nFaces, C_INT), &
                                      myDGSEM % mesh % faces, myDGSEM % mesh % cov, myDGSEM % mesh % J, &
                                      myDGSEM % mesh % nHat, myDGSEM % mesh % nLen, myDGSEM % mesh % bpos, &
                                      myDGSEM % dgStorage % dMat, myDGSEM % dgStorage % qWeight, &
                                      myDGSEM % dgStorage % lagLeft, myDGSEM % dgStorage % lagRight, F1, 1_C_INT, &
                                      myDGSEM % params % dt, myDGSEM % handle), 'Build')
    ALLOCATE(myDGSEM % sol(0:N, 0:N, 0:N, myDGSEM % nEq, nE), myDGSEM % vel(0:N, 0:N, 0:N, 3, nE), &
             myDGSEM % relax(0:N, 0:N, 0:N, myDGSEM % nEq, nE))
    myDGSEM % sol = ZERO
    myDGSEM % vel = ZERO
    myDGSEM % relax = ZERO
    myDGSEM % state = 0

    IF (.NOT. PRESENT(forceNoPickup)) CALL myDGSEM % ReadPickup(myDGSEM % params % iterInit)

    ALLOCATE(bw(0:N), sNew(0:nPlot), myDGSEM % plMatS(0:nPlot, 0:N), myDGSEM % plMatP(0:nPlot, 0:N), &
             myDGSEM % plMatQ(0:nPlot, 0:N))
    DO i = 0, nPlot  ! UniformPoints(-1, 1, nPlot), src/common/CommonRoutines.f90:460-482
      sNew(i) = -ONE + (TWO / REAL(nPlot, prec)) * REAL(i, prec)
    ENDDO
    CALL self_check(self_barycentric_weights(N, myDGSEM % dgStorage % nodes, bw), 'Build')
    CALL self_check(self_interpolation_matrix(N, myDGSEM % dgStorage % nodes, bw, INT(nPlot, C_INT), sNew, &
                                              myDGSEM % plMatS), 'Build')
    myDGSEM % plMatP = myDGSEM % plMatS
    myDGSEM % plMatQ = myDGSEM % plMatS
  END SUBROUTINE Build_Advection

  ! Trash (:276-312): the timing table (Write_MultiTimers), then the model
  SUBROUTINE Trash_Advection(myDGSEM)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    CALL myDGSEM % clocks % Write_MultiTimers()
    IF (C_ASSOCIATED(myDGSEM % handle)) CALL self_check(self_adv3d_destroy(myDGSEM % handle), 'Trash')
    myDGSEM % handle = C_NULL_PTR
    IF (ALLOCATED(myDGSEM % sol)) DEALLOCATE(myDGSEM % sol, myDGSEM % vel, myDGSEM % relax)
    IF (ALLOCATED(myDGSEM % plMatS)) DEALLOCATE(myDGSEM % plMatS, myDGSEM % plMatP, myDGSEM % plMatQ)
    CALL myDGSEM % dgStorage % Trash()
    CALL myDGSEM % mesh % Trash()
    CALL myDGSEM % clocks % Trash()
  END SUBROUTINE Trash_Advection

  ! ------------------------------------------------ host mirrors
  ! the device copy of every field the host changed; the relaxation array
  ! goes to both the field and the factor
  SUBROUTINE push(myDGSEM)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    IF (myDGSEM % state(M_SOL) == 1) CALL self_check(self_adv3d_set_field(myDGSEM % handle, F_SOL, myDGSEM % sol), &
                                                     'SetSolution')
    IF (myDGSEM % state(M_VEL) == 1) CALL self_check(self_adv3d_set_field(myDGSEM % handle, F_VEL, myDGSEM % vel), &
                                                     'SetVelocity')
    IF (myDGSEM % state(M_RELAX) == 1) THEN
      CALL self_check(self_adv3d_set_field(myDGSEM % handle, F_RELAX, myDGSEM % relax), 'SetRelaxationField')
      CALL self_check(self_adv3d_set_field(myDGSEM % handle, F_RFAC, myDGSEM % relax), 'SetRelaxationFactor')
    ENDIF
    WHERE (myDGSEM % state == 1) myDGSEM % state = 0
  END SUBROUTINE push

  ! the host copy of the solution, if the device's is newer (velocity and
  ! relaxation change only on the host)
  SUBROUTINE pull(myDGSEM, m)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in) :: m
    IF (myDGSEM % state(m) /= -1) RETURN
    IF (m == M_SOL) CALL self_check(self_adv3d_get_field(myDGSEM % handle, G_SOL, myDGSEM % sol), 'GetSolution')
    myDGSEM % state(m) = 0
  END SUBROUTINE pull

  ! before the host changes part of field m: the rest of it current
  SUBROUTINE touch(myDGSEM, m)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in) :: m
    CALL pull(myDGSEM, m)
    myDGSEM % state(m) = 1
  END SUBROUTINE touch

  SUBROUTINE GetSolution_Advection(myDGSEM, iEl, theSolution)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in) :: iEl
    REAL(prec), INTENT(out) :: theSolution(0:myDGSEM % nS, 0:myDGSEM % nP, 0:myDGSEM % nQ, 1:myDGSEM % nEq)
    CALL pull(myDGSEM, M_SOL)
    theSolution = myDGSEM % sol(:, :, :, :, iEl)
  END SUBROUTINE GetSolution_Advection

  SUBROUTINE SetSolution_Advection(myDGSEM, iEl, theSolution)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in) :: iEl
    REAL(prec), INTENT(in) :: theSolution(0:myDGSEM % nS, 0:myDGSEM % nP, 0:myDGSEM % nQ, 1:myDGSEM % nEq)
    CALL touch(myDGSEM, M_SOL)
    myDGSEM % sol(:, :, :, :, iEl) = theSolution
  END SUBROUTINE SetSolution_Advection

  SUBROUTINE GetSolutionWithVarID_Advection(myDGSEM, iEl, varID, theSolution)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in) :: iEl, varID
    REAL(prec), INTENT(out) :: theSolution(0:myDGSEM % nS, 0:myDGSEM % nP, 0:myDGSEM % nQ)
    CALL pull(myDGSEM, M_SOL)
    theSolution = myDGSEM % sol(:, :, :, varID, iEl)
  END SUBROUTINE GetSolutionWithVarID_Advection

  SUBROUTINE SetSolutionWithVarID_Advection(myDGSEM, iEl, varID, theSolution)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in) :: iEl, varID
    REAL(prec), INTENT(in) :: theSolution(0:myDGSEM % nS, 0:myDGSEM % nP, 0:myDGSEM % nQ)
    CALL touch(myDGSEM, M_SOL)
    myDGSEM % sol(:, :, :, varID, iEl) = theSolution
  END SUBROUTINE SetSolutionWithVarID_Advection

  SUBROUTINE GetSolutionAtNode_Advection(myDGSEM, iEl, i, j, k, theSolution)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in) :: iEl, i, j, k
    REAL(prec), INTENT(out) :: theSolution(1:myDGSEM % nEq)
    CALL pull(myDGSEM, M_SOL)
    theSolution = myDGSEM % sol(i, j, k, :, iEl)
  END SUBROUTINE GetSolutionAtNode_Advection

  SUBROUTINE SetSolutionAtNode_Advection(myDGSEM, iEl, i, j, k, theSolution)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in) :: iEl, i, j, k
    REAL(prec), INTENT(in) :: theSolution(1:myDGSEM % nEq)
    CALL touch(myDGSEM, M_SOL)
    myDGSEM % sol(i, j, k, :, iEl) = theSolution
  END SUBROUTINE SetSolutionAtNode_Advection

  ! GetTendency (:580-593): the tendency of the last time derivative
  SUBROUTINE GetTendency_Advection(myDGSEM, iEl, theTend)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in) :: iEl
    REAL(prec), INTENT(out) :: theTend(0:myDGSEM % nS, 0:myDGSEM % nP, 0:myDGSEM % nQ, 1:myDGSEM % nEq)
    REAL(prec), ALLOCATABLE :: t(:,:,:,:,:)
    ALLOCATE(t(0:myDGSEM % nS, 0:myDGSEM % nP, 0:myDGSEM % nQ, myDGSEM % nEq, myDGSEM % mesh % nElems))
    CALL self_check(self_adv3d_get_field(myDGSEM % handle, G_TEND, t), 'GetTendency')
    theTend = t(:, :, :, :, iEl)
  END SUBROUTINE GetTendency_Advection

  SUBROUTINE GetVelocity_Advection(myDGSEM, iEl, u, v, w)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in) :: iEl
    REAL(prec), INTENT(out), DIMENSION(0:myDGSEM % nS, 0:myDGSEM % nP, 0:myDGSEM % nQ) :: u, v, w
    u = myDGSEM % vel(:, :, :, 1, iEl)
    v = myDGSEM % vel(:, :, :, 2, iEl)
    w = myDGSEM % vel(:, :, :, 3, iEl)
  END SUBROUTINE GetVelocity_Advection

  ! SetVelocity (:477-494); the face velocities follow on the device
  SUBROUTINE SetVelocity_Advection(myDGSEM, iEl, u, v, w)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in) :: iEl
    REAL(prec), INTENT(in), DIMENSION(0:myDGSEM % nS, 0:myDGSEM % nP, 0:myDGSEM % nQ) :: u, v, w
    CALL touch(myDGSEM, M_VEL)
    myDGSEM % vel(:, :, :, 1, iEl) = u
    myDGSEM % vel(:, :, :, 2, iEl) = v
    myDGSEM % vel(:, :, :, 3, iEl) = w
  END SUBROUTINE SetVelocity_Advection

  SUBROUTINE GetRelaxationField_Advection(myDGSEM, iEl, relaxField)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in) :: iEl
    REAL(prec), INTENT(out) :: relaxField(0:myDGSEM % nS, 0:myDGSEM % nP, 0:myDGSEM % nQ, 1:myDGSEM % nEq)
    relaxField = myDGSEM % relax(:, :, :, :, iEl)
  END SUBROUTINE GetRelaxationField_Advection

  SUBROUTINE SetRelaxationField_Advection(myDGSEM, iEl, relaxField)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in) :: iEl
    REAL(prec), INTENT(in) :: relaxField(0:myDGSEM % nS, 0:myDGSEM % nP, 0:myDGSEM % nQ, 1:myDGSEM % nEq)
    CALL touch(myDGSEM, M_RELAX)
    myDGSEM % relax(:, :, :, :, iEl) = relaxField
  END SUBROUTINE SetRelaxationField_Advection

  ! GetRelaxationFactor / SetRelaxationFactor (:898-928): the relaxation
  ! field's storage, as the reference's
  SUBROUTINE GetRelaxationFactor_Advection(myDGSEM, iEl, rFac)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in) :: iEl
    REAL(prec), INTENT(out) :: rFac(0:myDGSEM % nS, 0:myDGSEM % nP, 0:myDGSEM % nQ, 1:myDGSEM % nEq)
    rFac = myDGSEM % relax(:, :, :, :, iEl)
  END SUBROUTINE GetRelaxationFactor_Advection

  SUBROUTINE SetRelaxationFactor_Advection(myDGSEM, iEl, rFac)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in) :: iEl
    REAL(prec), INTENT(in) :: rFac(0:myDGSEM % nS, 0:myDGSEM % nP, 0:myDGSEM % nQ, 1:myDGSEM % nEq)
    CALL touch(myDGSEM, M_RELAX)
    myDGSEM % relax(:, :, :, :, iEl) = rFac
  END SUBROUTINE SetRelaxationFactor_Advection

  ! ------------------------------------------------ time stepping on the device
  ! GlobalTimeDerivative(tn, m): the adaptive filter of RK stage m applied
  ! to the solution in place, then the tendency
  SUBROUTINE GlobalTimeDerivative_Advection(myDGSEM, tn, m)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    REAL(prec), INTENT(in) :: tn
    INTEGER, INTENT(in) :: m
    CALL push(myDGSEM)
    CALL self_check(self_adv3d_global_time_derivative(myDGSEM % handle, tn, INT(m, C_INT)), 'GlobalTimeDerivative')
    myDGSEM % state(M_SOL) = -1
  END SUBROUTINE GlobalTimeDerivative_Advection

  SUBROUTINE ForwardStepRK3_Advection(myDGSEM, tn)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    REAL(prec), INTENT(in) :: tn
    CALL push(myDGSEM)
    CALL self_check(self_adv3d_forward_step_rk3(myDGSEM % handle, tn, 1_C_INT, C_NULL_PTR), 'ForwardStepRK3')
    myDGSEM % state(M_SOL) = -1
  END SUBROUTINE ForwardStepRK3_Advection

  ! ------------------------------------------------ output
  ! CoarseToFine (:1747-1842): positions, velocity and tracers of element
  ! iEl at the plotting points
  SUBROUTINE CoarseToFine_Advection(myDGSEM, iEl, x, y, z, u, v, w, c)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in) :: iEl
    REAL(prec), INTENT(out), DIMENSION(0:myDGSEM % nPlot, 0:myDGSEM % nPlot, 0:myDGSEM % nPlot) :: x, y, z, u, v, w
    REAL(prec), INTENT(out) :: c(0:myDGSEM % nPlot, 0:myDGSEM % nPlot, 0:myDGSEM % nPlot, 1:myDGSEM % nEq)
    INTEGER :: k
    CALL pull(myDGSEM, M_SOL)
    CALL c2f(myDGSEM % mesh % x(:, :, :, iEl), x)
    CALL c2f(myDGSEM % mesh % y(:, :, :, iEl), y)
    CALL c2f(myDGSEM % mesh % z(:, :, :, iEl), z)
    CALL c2f(myDGSEM % vel(:, :, :, 1, iEl), u)
    CALL c2f(myDGSEM % vel(:, :, :, 2, iEl), v)
    CALL c2f(myDGSEM % vel(:, :, :, 3, iEl), w)
    DO k = 1, myDGSEM % nEq
      CALL c2f(myDGSEM % sol(:, :, :, k, iEl), c(:, :, :, k))
    ENDDO
  CONTAINS
    SUBROUTINE c2f(fOld, fNew)
      REAL(prec), INTENT(in)  :: fOld(0:myDGSEM % nS, 0:myDGSEM % nP, 0:myDGSEM % nQ)
      REAL(prec), INTENT(out) :: fNew(0:myDGSEM % nPlot, 0:myDGSEM % nPlot, 0:myDGSEM % nPlot)
      REAL(prec) :: a(0:myDGSEM % nS, 0:myDGSEM % nP, 0:myDGSEM % nQ)
      REAL(prec) :: b(0:myDGSEM % nPlot, 0:myDGSEM % nPlot, 0:myDGSEM % nPlot)
      a = fOld
      CALL self_check(self_coarse_to_fine_3d(INT(myDGSEM % nS, C_INT), 1_C_SIZE_T, myDGSEM % plMatS, &
                                             INT(myDGSEM % nPlot, C_INT), a, b), 'CoarseToFine')
      fNew = b
    END SUBROUTINE c2f
  END SUBROUTINE CoarseToFine_Advection

  ! WriteTecplot (:1846-1901): <filename>.tec (or Advection.tec), per element
  ! a zone of X, Y, Z, U, V, W, C1, C2 at the plotting points; the whole mesh
  ! goes through one batched CoarseToFine per field.
  SUBROUTINE WriteTecplot_Advection(myDGSEM, filename)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    CHARACTER(*), INTENT(in), OPTIONAL :: filename
    INTEGER :: iS, iP, iQ, iEl, fUnit, nPlot, N, nE, nq, k
    CHARACTER(len=5) :: zoneID
    REAL(prec), ALLOCATABLE :: tmp(:,:,:,:), f(:,:,:,:,:)
    nPlot = myDGSEM % nPlot
    N = myDGSEM % nS
    nE = myDGSEM % mesh % nElems
    nq = myDGSEM % nEq
    CALL pull(myDGSEM, M_SOL)
    ! f(:,:,:,:,k): X, Y, Z, U, V, W, C1, C2 at the plotting points
    ALLOCATE(tmp(0:N, 0:N, 0:N, nE), f(0:nPlot, 0:nPlot, 0:nPlot, nE, 8))
    CALL c2f_all(myDGSEM % mesh % x, f(:, :, :, :, 1))
    CALL c2f_all(myDGSEM % mesh % y, f(:, :, :, :, 2))
    CALL c2f_all(myDGSEM % mesh % z, f(:, :, :, :, 3))
    DO k = 1, 3
      tmp = myDGSEM % vel(:, :, :, k, :)
      CALL c2f_all(tmp, f(:, :, :, :, 3 + k))
    ENDDO
    DO k = 1, 2
      tmp = myDGSEM % sol(:, :, :, MIN(k, nq), :)
      CALL c2f_all(tmp, f(:, :, :, :, 6 + k))
    ENDDO
    IF (PRESENT(filename)) THEN
      OPEN(NEWUNIT=fUnit, FILE=TRIM(filename)//'.tec', FORM='formatted', STATUS='replace')
    ELSE
      OPEN(NEWUNIT=fUnit, FILE='Advection.tec', FORM='formatted', STATUS='replace')
    ENDIF
    WRITE(fUnit, *) 'VARIABLES = "X", "Y","Z", "U", "V", "W", "C1", "C2" '
    DO iEl = 1, nE
      WRITE(zoneID, '(I5.5)') iEl
      WRITE(fUnit, *) 'ZONE T="el'//TRIM(zoneID)//'", I=', nPlot + 1, ', J=', nPlot + 1, ', K=', nPlot + 1, ',F=POINT'
      DO iQ = 0, nPlot
        DO iP = 0, nPlot
          DO iS = 0, nPlot
            WRITE(fUnit, *) f(iS, iP, iQ, iEl, :)
          ENDDO
        ENDDO
      ENDDO
    ENDDO
    CLOSE(UNIT=fUnit)
  CONTAINS
    SUBROUTINE c2f_all(fOld, fNew)
      REAL(prec), INTENT(in)  :: fOld(0:N, 0:N, 0:N, nE)
      REAL(prec), INTENT(out) :: fNew(0:nPlot, 0:nPlot, 0:nPlot, nE)
      REAL(prec), ALLOCATABLE :: a(:,:,:,:), b(:,:,:,:)
      ALLOCATE(a(0:N, 0:N, 0:N, nE), b(0:nPlot, 0:nPlot, 0:nPlot, nE))
      a = fOld
      CALL self_check(self_coarse_to_fine_3d(INT(N, C_INT), INT(nE, C_SIZE_T), myDGSEM % plMatS, INT(nPlot, C_INT), &
                                             a, b), 'WriteTecplot')
      fNew = b
    END SUBROUTINE c2f_all
  END SUBROUTINE WriteTecplot_Advection

  ! WritePickup (:1905-1972) / ReadPickup (:1976-2047):
  ! Advection.<iter, 10 digits>.pickup, unformatted direct access, big
  ! endian, one record per (0:N)^3 array: per element the nEq tracers, the
  ! nEq relaxation fields, the nEq relaxation factors (the same storage),
  ! then u, v, w.
  SUBROUTINE WritePickup_Advection(myDGSEM, iter)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in) :: iter
    CHARACTER(len=10) :: iterChar
    INTEGER :: fUnit, thisRec, iEl, iEq, k, n1
    CALL pull(myDGSEM, M_SOL)
    n1 = myDGSEM % nS + 1
    WRITE(iterChar, '(I10.10)') iter
    OPEN(NEWUNIT=fUnit, FILE='Advection.'//iterChar//'.pickup', FORM='unformatted', ACCESS='direct', &
         STATUS='replace', ACTION='write', CONVERT='big_endian', RECL=prec * n1 * n1 * n1)
    thisRec = 1
    DO iEl = 1, myDGSEM % mesh % nElems
      DO iEq = 1, myDGSEM % nEq
        WRITE(fUnit, REC=thisRec) myDGSEM % sol(:, :, :, iEq, iEl)
        thisRec = thisRec + 1
      ENDDO
      DO k = 1, 2
        DO iEq = 1, myDGSEM % nEq
          WRITE(fUnit, REC=thisRec) myDGSEM % relax(:, :, :, iEq, iEl)
          thisRec = thisRec + 1
        ENDDO
      ENDDO
      DO k = 1, 3
        WRITE(fUnit, REC=thisRec) myDGSEM % vel(:, :, :, k, iEl)
        thisRec = thisRec + 1
      ENDDO
    ENDDO
    CLOSE(UNIT=fUnit)
  END SUBROUTINE WritePickup_Advection

  ! The relaxation factor records are read into the relaxation storage
  ! after the field records, so the factor is what remains, as with the
  ! reference's SetRelaxationFactor.
  SUBROUTINE ReadPickup_Advection(myDGSEM, iter)
    CLASS(Advection), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in) :: iter
    CHARACTER(len=10) :: iterChar
    INTEGER :: fUnit, thisRec, iEl, iEq, k, n1
    n1 = myDGSEM % nS + 1
    WRITE(iterChar, '(I10.10)') iter
    OPEN(NEWUNIT=fUnit, FILE='Advection.'//iterChar//'.pickup', FORM='unformatted', ACCESS='direct', &
         STATUS='old', ACTION='read', CONVERT='big_endian', RECL=prec * n1 * n1 * n1)
    thisRec = 1
    DO iEl = 1, myDGSEM % mesh % nElems
      DO iEq = 1, myDGSEM % nEq
        READ(fUnit, REC=thisRec) myDGSEM % sol(:, :, :, iEq, iEl)
        thisRec = thisRec + 1
      ENDDO
      DO k = 1, 2
        DO iEq = 1, myDGSEM % nEq
          READ(fUnit, REC=thisRec) myDGSEM % relax(:, :, :, iEq, iEl)
          thisRec = thisRec + 1
        ENDDO
      ENDDO
      DO k = 1, 3
        READ(fUnit, REC=thisRec) myDGSEM % vel(:, :, :, k, iEl)
        thisRec = thisRec + 1
      ENDDO
    ENDDO
    CLOSE(UNIT=fUnit)
    myDGSEM % state = 1
  END SUBROUTINE ReadPickup_Advection

END MODULE Advection3DClass
