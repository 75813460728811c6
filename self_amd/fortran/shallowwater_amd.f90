! shallowwater_amd.f90 -- drop-in ShallowWater type over the C ABI.
!
! Mirrors the ShallowWater type of
! src/highend/shallowwater/ConservativeShallowWaterClass.f90 (:60-151) with
! the same type-bound procedure names and per-element accessor signatures:
!   Build (:163-245), Trash (:249-268), GetSolution/SetSolution (:316-346),
!   GetTendency (:516-529), SetBathymetry (:929-942),
!   SetPlanetaryVorticity (:1035-1048), CalculateBathymetryGradient
!   (:1733-1868, here for all elements at once), ForwardStepRK3 (:1173-1218),
!   GlobalTimeDerivative (:1222-1262), WritePickup/ReadPickup (:2013-2125).
! The type keeps reference-shaped host arrays as a mirror; the device holds
! the state.  Set* write the mirror and mark it newer; the next device call
! uploads it once.  Device calls mark the device newer; the next Get*
! downloads once.  Parameters come in an explicit SWParamsAMD argument
! instead of the reference's runtime.params namelist.
MODULE ShallowWaterAMD_Class
  USE ISO_C_BINDING
  USE SELF_AMD_CABI
  IMPLICIT NONE

  TYPE SWParamsAMD
    INTEGER        :: polyDeg = 7, nXElem = 4, nYElem = 4
    REAL(C_DOUBLE) :: xScale = 1.0_C_DOUBLE, yScale = 1.0_C_DOUBLE, dt = 1.0_C_DOUBLE
    REAL(C_DOUBLE) :: g = 9.81_C_DOUBLE, f0 = 0.0_C_DOUBLE, betaX = 0.0_C_DOUBLE, betaY = 0.0_C_DOUBLE
    REAL(C_DOUBLE) :: linearDrag = 0.0_C_DOUBLE, warp = 0.0_C_DOUBLE
    LOGICAL        :: periodicX = .TRUE., periodicY = .TRUE.
    INTEGER        :: precision = SELF_FP64, form = SELF_FUSED
    ! an ISM-V2 mesh file instead of the structured nXElem x nYElem mesh
    ! (SWParams % SpecMeshFile; blank = the structured mesh)
    CHARACTER(LEN=256) :: SpecMeshFile = ' '
    ! 1: the reference's nodal storage, whose p-direction derivative matrix
    ! is transposed (NodalStorage_2D); 0: the derivative matrix itself
    INTEGER        :: quirkDMatP = 0
  END TYPE SWParamsAMD

  TYPE ShallowWaterAMD
    INTEGER           :: nEq = 3, nS, nP, nElems, nEdges
    TYPE(SWParamsAMD) :: params
    TYPE(C_PTR)       :: handle = C_NULL_PTR
    REAL(C_DOUBLE), ALLOCATABLE :: nodes(:), qWeight(:), x(:,:,:), y(:,:,:)
    ! covariant metric terms and Jacobian at the nodes (the mesh's, kept for output)
    REAL(C_DOUBLE), ALLOCATABLE :: dxds(:,:,:), dxdp(:,:,:), dyds(:,:,:), dydp(:,:,:), J(:,:,:)
    REAL(C_DOUBLE), ALLOCATABLE :: solution(:,:,:,:), tendency(:,:,:,:), bathymetry(:,:,:,:), vorticity(:,:,:)
    LOGICAL :: hostSolutionNewer = .FALSE., deviceSolutionNewer = .FALSE.
    LOGICAL :: hostBathymetryNewer = .FALSE., hostVorticityNewer = .FALSE.
    LOGICAL :: deviceBathymetryNewer = .FALSE.
  CONTAINS
    PROCEDURE :: Build => Build_ShallowWaterAMD
    PROCEDURE :: Trash => Trash_ShallowWaterAMD
    PROCEDURE :: GetSolution => GetSolution_ShallowWaterAMD
    PROCEDURE :: SetSolution => SetSolution_ShallowWaterAMD
    PROCEDURE :: GetTendency => GetTendency_ShallowWaterAMD
    PROCEDURE :: GetBathymetry => GetBathymetry_ShallowWaterAMD
    PROCEDURE :: SetBathymetry => SetBathymetry_ShallowWaterAMD
    PROCEDURE :: SetPlanetaryVorticity => SetPlanetaryVorticity_ShallowWaterAMD
    PROCEDURE :: GetPositionAtNode => GetPositionAtNode_ShallowWaterAMD
    PROCEDURE :: CalculateBathymetryGradient => CalculateBathymetryGradient_ShallowWaterAMD
    PROCEDURE :: GlobalTimeDerivative => GlobalTimeDerivative_ShallowWaterAMD
    PROCEDURE :: ForwardStepRK3 => ForwardStepRK3_ShallowWaterAMD
    PROCEDURE :: WritePickup => WritePickup_ShallowWaterAMD
    PROCEDURE :: ReadPickup => ReadPickup_ShallowWaterAMD
    PROCEDURE, PRIVATE :: Upload => Upload_ShallowWaterAMD
  END TYPE ShallowWaterAMD

CONTAINS

  SUBROUTINE Build_ShallowWaterAMD(myDGSEM, params)
    CLASS(ShallowWaterAMD), INTENT(inout) :: myDGSEM
    TYPE(SWParamsAMD), INTENT(in)         :: params
    INTEGER(C_INT) :: N, nE, nEd, px, py, nNd, gDeg
    INTEGER(C_INT), ALLOCATABLE :: elemNodes(:,:), nodeType(:)
    REAL(C_DOUBLE), ALLOCATABLE :: nodeX(:), nodeY(:)
    CHARACTER(KIND=C_CHAR, LEN=257) :: path
    REAL(C_DOUBLE), ALLOCATABLE :: dMatS(:,:), dMatP(:,:), lW(:), lE(:), lS(:), lN(:)
    REAL(C_DOUBLE), ALLOCATABLE :: dxds(:,:,:), dxdp(:,:,:), dyds(:,:,:), dydp(:,:,:), J(:,:,:)
    REAL(C_DOUBLE), ALLOCATABLE :: nHat(:,:,:,:), nLen(:,:,:)
    INTEGER(C_INT), ALLOCATABLE :: edges(:,:)
    TYPE(self_sw2d_desc) :: desc

    myDGSEM % params = params
    N = params % polyDeg
    myDGSEM % nS = N
    myDGSEM % nP = N
    ALLOCATE(myDGSEM % nodes(0:N), myDGSEM % qWeight(0:N), dMatS(0:N,0:N), dMatP(0:N,0:N))
    ALLOCATE(lW(0:N), lE(0:N), lS(0:N), lN(0:N))
    CALL self_check(self_nodal2d_build(N, SELF_GAUSS, SELF_DG, INT(params % quirkDMatP, C_INT), myDGSEM % nodes, &
                                       myDGSEM % qWeight, &
                                       dMatS, dMatP, lW, lE, lS, lN), 'self_nodal2d_build')
    IF (LEN_TRIM(params % SpecMeshFile) > 0) THEN
      ! BuildQuadMesh (:289-304): ReadSpecMeshFile, then ScaleTheMesh
      path = TRIM(params % SpecMeshFile)//C_NULL_CHAR
      CALL self_check(self_specmesh_counts(path, nNd, nEd, nE, gDeg), 'self_specmesh_counts')
      CALL alloc_geometry()
      ALLOCATE(elemNodes(4, nE), nodeType(nNd), nodeX(nNd), nodeY(nNd))
      CALL self_check(self_specmesh_read(path, N, myDGSEM % nodes, edges, elemNodes, nodeType, nodeX, nodeY, &
                                         myDGSEM % x, myDGSEM % y, dxds, dxdp, dyds, dydp, J, nHat, nLen), &
                      'self_specmesh_read')
      CALL self_check(self_quadmesh_scale(N, myDGSEM % nodes, nE, nNd, params % xScale, params % yScale, nodeX, &
                                          nodeY, myDGSEM % x, myDGSEM % y, dxds, dxdp, dyds, dydp, J, nHat, nLen), &
                      'self_quadmesh_scale')
    ELSE
      px = MERGE(1, 0, params % periodicX)
      py = MERGE(1, 0, params % periodicY)
      CALL self_check(self_quadmesh_counts(params % nXElem, params % nYElem, px, py, nE, nEd), &
                      'self_quadmesh_counts')
      CALL alloc_geometry()
      CALL self_check(self_quadmesh_structured(params % nXElem, params % nYElem, N, px, py, params % xScale, &
                                               params % yScale, params % warp, myDGSEM % nodes, edges, &
                                               myDGSEM % x, myDGSEM % y, dxds, dxdp, dyds, dydp, J, nHat, nLen), &
                      'self_quadmesh_structured')
    ENDIF
    desc % N = N
    desc % nElems = nE
    desc % nEdges = nEd
    desc % precision = params % precision
    desc % form = params % form
    desc % g = params % g
    desc % dt = params % dt
    desc % linearDrag = params % linearDrag
    CALL self_check(self_sw2d_create(desc, myDGSEM % handle), 'self_sw2d_create')
    CALL self_check(self_sw2d_set_nodal(myDGSEM % handle, dMatS, dMatP, myDGSEM % qWeight, myDGSEM % qWeight, &
                                        lW, lE, lS, lN), 'self_sw2d_set_nodal')
    CALL self_check(self_sw2d_set_geometry(myDGSEM % handle, edges, dxds, dxdp, dyds, dydp, J, nHat, nLen), &
                    'self_sw2d_set_geometry')
    myDGSEM % dxds = dxds
    myDGSEM % dxdp = dxdp
    myDGSEM % dyds = dyds
    myDGSEM % dydp = dydp
    myDGSEM % J = J
    ALLOCATE(myDGSEM % solution(0:N,0:N,3,nE), myDGSEM % tendency(0:N,0:N,3,nE))
    ALLOCATE(myDGSEM % bathymetry(0:N,0:N,3,nE), myDGSEM % vorticity(0:N,0:N,nE))
    ! Build_ShallowWater defaults: depth hDefault = 1 with zero gradient,
    ! planetary vorticity f0 + betaX x + betaY y, zero state
    myDGSEM % solution = 0.0_C_DOUBLE
    myDGSEM % tendency = 0.0_C_DOUBLE
    myDGSEM % bathymetry = 0.0_C_DOUBLE
    myDGSEM % bathymetry(:,:,1,:) = 1.0_C_DOUBLE
    myDGSEM % vorticity = params % f0 + params % betaX * myDGSEM % x + params % betaY * myDGSEM % y
    myDGSEM % hostSolutionNewer = .TRUE.
    myDGSEM % hostBathymetryNewer = .TRUE.
    myDGSEM % hostVorticityNewer = .TRUE.

  CONTAINS

    SUBROUTINE alloc_geometry()
      myDGSEM % nElems = nE
      myDGSEM % nEdges = nEd
      ALLOCATE(edges(6, nEd), myDGSEM % x(0:N,0:N,nE), myDGSEM % y(0:N,0:N,nE))
      ALLOCATE(dxds(0:N,0:N,nE), dxdp(0:N,0:N,nE), dyds(0:N,0:N,nE), dydp(0:N,0:N,nE), J(0:N,0:N,nE))
      ALLOCATE(nHat(2,0:N,4,nE), nLen(0:N,4,nE))
    END SUBROUTINE alloc_geometry
  END SUBROUTINE Build_ShallowWaterAMD

  SUBROUTINE Trash_ShallowWaterAMD(myDGSEM)
    CLASS(ShallowWaterAMD), INTENT(inout) :: myDGSEM
    IF (C_ASSOCIATED(myDGSEM % handle)) THEN
      CALL self_check(self_sw2d_destroy(myDGSEM % handle), 'self_sw2d_destroy')
      myDGSEM % handle = C_NULL_PTR
    ENDIF
    IF (ALLOCATED(myDGSEM % solution)) DEALLOCATE(myDGSEM % solution, myDGSEM % tendency, &
                                                  myDGSEM % bathymetry, myDGSEM % vorticity, myDGSEM % nodes, &
                                                  myDGSEM % qWeight, myDGSEM % x, myDGSEM % y)
    IF (ALLOCATED(myDGSEM % J)) DEALLOCATE(myDGSEM % dxds, myDGSEM % dxdp, myDGSEM % dyds, myDGSEM % dydp, &
                                           myDGSEM % J)
  END SUBROUTINE Trash_ShallowWaterAMD

  ! Push host-side changes to the device before a device operation.
  SUBROUTINE Upload_ShallowWaterAMD(myDGSEM)
    CLASS(ShallowWaterAMD), INTENT(inout) :: myDGSEM
    IF (myDGSEM % hostBathymetryNewer) THEN
      CALL self_check(self_sw2d_set_field(myDGSEM % handle, SELF_SW_BATHYMETRY, myDGSEM % bathymetry), &
                      'SetBathymetry')
      myDGSEM % hostBathymetryNewer = .FALSE.
    ENDIF
    IF (myDGSEM % hostVorticityNewer) THEN
      CALL self_check(self_sw2d_set_field(myDGSEM % handle, SELF_SW_PLANETARY_VORTICITY, myDGSEM % vorticity), &
                      'SetPlanetaryVorticity')
      myDGSEM % hostVorticityNewer = .FALSE.
    ENDIF
    IF (myDGSEM % hostSolutionNewer) THEN
      CALL self_check(self_sw2d_set_field(myDGSEM % handle, SELF_SW_SOLUTION, myDGSEM % solution), 'SetSolution')
      myDGSEM % hostSolutionNewer = .FALSE.
    ENDIF
  END SUBROUTINE Upload_ShallowWaterAMD

  SUBROUTINE GetSolution_ShallowWaterAMD(myDGSEM, iEl, theSolution)
    CLASS(ShallowWaterAMD), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in)                   :: iEl
    REAL(C_DOUBLE), INTENT(out)           :: theSolution(0:myDGSEM % nS, 0:myDGSEM % nP, 1:3)
    IF (myDGSEM % deviceSolutionNewer) THEN
      CALL self_check(self_sw2d_get_field(myDGSEM % handle, SELF_SW_SOLUTION, myDGSEM % solution), 'GetSolution')
      myDGSEM % deviceSolutionNewer = .FALSE.
    ENDIF
    theSolution = myDGSEM % solution(:,:,:,iEl)
  END SUBROUTINE GetSolution_ShallowWaterAMD

  SUBROUTINE SetSolution_ShallowWaterAMD(myDGSEM, iEl, theSolution)
    CLASS(ShallowWaterAMD), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in)                   :: iEl
    REAL(C_DOUBLE), INTENT(in)            :: theSolution(0:myDGSEM % nS, 0:myDGSEM % nP, 1:3)
    IF (myDGSEM % deviceSolutionNewer) THEN
      CALL self_check(self_sw2d_get_field(myDGSEM % handle, SELF_SW_SOLUTION, myDGSEM % solution), 'GetSolution')
      myDGSEM % deviceSolutionNewer = .FALSE.
    ENDIF
    myDGSEM % solution(:,:,:,iEl) = theSolution
    myDGSEM % hostSolutionNewer = .TRUE.
  END SUBROUTINE SetSolution_ShallowWaterAMD

  SUBROUTINE GetTendency_ShallowWaterAMD(myDGSEM, iEl, theTend)
    CLASS(ShallowWaterAMD), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in)                   :: iEl
    REAL(C_DOUBLE), INTENT(out)           :: theTend(0:myDGSEM % nS, 0:myDGSEM % nP, 1:3)
    CALL self_check(self_sw2d_get_field(myDGSEM % handle, SELF_SW_TENDENCY, myDGSEM % tendency), 'GetTendency')
    theTend = myDGSEM % tendency(:,:,:,iEl)
  END SUBROUTINE GetTendency_ShallowWaterAMD

  SUBROUTINE GetBathymetry_ShallowWaterAMD(myDGSEM, iEl, theBathymetry)
    CLASS(ShallowWaterAMD), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in)                   :: iEl
    REAL(C_DOUBLE), INTENT(out)           :: theBathymetry(0:myDGSEM % nS, 0:myDGSEM % nP, 1:3)
    IF (myDGSEM % deviceBathymetryNewer) THEN
      CALL self_check(self_sw2d_get_field(myDGSEM % handle, SELF_SW_BATHYMETRY, myDGSEM % bathymetry), &
                      'GetBathymetry')
      myDGSEM % deviceBathymetryNewer = .FALSE.
    ENDIF
    theBathymetry = myDGSEM % bathymetry(:,:,:,iEl)
  END SUBROUTINE GetBathymetry_ShallowWaterAMD

  SUBROUTINE SetBathymetry_ShallowWaterAMD(myDGSEM, iEl, theBathymetry)
    CLASS(ShallowWaterAMD), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in)                   :: iEl
    REAL(C_DOUBLE), INTENT(in)            :: theBathymetry(0:myDGSEM % nS, 0:myDGSEM % nP, 1:3)
    IF (myDGSEM % deviceBathymetryNewer) THEN
      CALL self_check(self_sw2d_get_field(myDGSEM % handle, SELF_SW_BATHYMETRY, myDGSEM % bathymetry), &
                      'GetBathymetry')
      myDGSEM % deviceBathymetryNewer = .FALSE.
    ENDIF
    myDGSEM % bathymetry(:,:,:,iEl) = theBathymetry
    myDGSEM % hostBathymetryNewer = .TRUE.
  END SUBROUTINE SetBathymetry_ShallowWaterAMD

  SUBROUTINE SetPlanetaryVorticity_ShallowWaterAMD(myDGSEM, iEl, theVorticity)
    CLASS(ShallowWaterAMD), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in)                   :: iEl
    REAL(C_DOUBLE), INTENT(in)            :: theVorticity(0:myDGSEM % nS, 0:myDGSEM % nP)
    myDGSEM % vorticity(:,:,iEl) = theVorticity
    myDGSEM % hostVorticityNewer = .TRUE.
  END SUBROUTINE SetPlanetaryVorticity_ShallowWaterAMD

  ! QuadMesh%GetPositionAtNode analogue for building initial conditions
  SUBROUTINE GetPositionAtNode_ShallowWaterAMD(myDGSEM, iEl, x, y, i, j)
    CLASS(ShallowWaterAMD), INTENT(in) :: myDGSEM
    INTEGER, INTENT(in)                :: iEl, i, j
    REAL(C_DOUBLE), INTENT(out)        :: x, y
    x = myDGSEM % x(i, j, iEl)
    y = myDGSEM % y(i, j, iEl)
  END SUBROUTINE GetPositionAtNode_ShallowWaterAMD

  SUBROUTINE CalculateBathymetryGradient_ShallowWaterAMD(myDGSEM)
    CLASS(ShallowWaterAMD), INTENT(inout) :: myDGSEM
    CALL myDGSEM % Upload()
    CALL self_check(self_sw2d_calculate_bathymetry_gradient(myDGSEM % handle), 'CalculateBathymetryGradient')
    myDGSEM % deviceBathymetryNewer = .TRUE.
  END SUBROUTINE CalculateBathymetryGradient_ShallowWaterAMD

  SUBROUTINE GlobalTimeDerivative_ShallowWaterAMD(myDGSEM, tn)
    CLASS(ShallowWaterAMD), INTENT(inout) :: myDGSEM
    REAL(C_DOUBLE), INTENT(in)            :: tn
    CALL myDGSEM % Upload()
    CALL self_check(self_sw2d_global_time_derivative(myDGSEM % handle, tn), 'GlobalTimeDerivative')
  END SUBROUTINE GlobalTimeDerivative_ShallowWaterAMD

  SUBROUTINE ForwardStepRK3_ShallowWaterAMD(myDGSEM, tn)
    CLASS(ShallowWaterAMD), INTENT(inout) :: myDGSEM
    REAL(C_DOUBLE), INTENT(in)            :: tn
    CALL myDGSEM % Upload()
    CALL self_check(self_sw2d_forward_step_rk3(myDGSEM % handle, tn, 1_C_INT), 'ForwardStepRK3')
    myDGSEM % deviceSolutionNewer = .TRUE.
  END SUBROUTINE ForwardStepRK3_ShallowWaterAMD

  ! WritePickup (:2013-2065): ShallowWater.<iter, 10 digits>.pickup in the
  ! working directory, the reference's big-endian direct-access records
  SUBROUTINE WritePickup_ShallowWaterAMD(myDGSEM, iter)
    CLASS(ShallowWaterAMD), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in)                   :: iter
    CHARACTER(KIND=C_CHAR, LEN=64) :: fname
    CALL myDGSEM % Upload()
    CALL self_check(self_pickup_filename('ShallowWater'//C_NULL_CHAR, INT(iter, C_INT), fname, 64_C_INT), &
                    'WritePickup')
    CALL self_check(self_sw2d_write_pickup(myDGSEM % handle, fname), 'WritePickup')
  END SUBROUTINE WritePickup_ShallowWaterAMD

  ! ReadPickup (:2069-2125) and the bathymetry gradient Build computes after it
  ! (:224-228); the host mirrors follow the device
  SUBROUTINE ReadPickup_ShallowWaterAMD(myDGSEM, iter)
    CLASS(ShallowWaterAMD), INTENT(inout) :: myDGSEM
    INTEGER, INTENT(in)                   :: iter
    CHARACTER(KIND=C_CHAR, LEN=64) :: fname
    CALL self_check(self_pickup_filename('ShallowWater'//C_NULL_CHAR, INT(iter, C_INT), fname, 64_C_INT), &
                    'ReadPickup')
    CALL self_check(self_sw2d_read_pickup(myDGSEM % handle, fname), 'ReadPickup')
    CALL self_check(self_sw2d_get_field(myDGSEM % handle, SELF_SW_PLANETARY_VORTICITY, myDGSEM % vorticity), &
                    'ReadPickup')
    myDGSEM % hostSolutionNewer = .FALSE.
    myDGSEM % hostBathymetryNewer = .FALSE.
    myDGSEM % hostVorticityNewer = .FALSE.
    myDGSEM % deviceSolutionNewer = .TRUE.
    myDGSEM % deviceBathymetryNewer = .TRUE.
  END SUBROUTINE ReadPickup_ShallowWaterAMD

END MODULE ShallowWaterAMD_Class
