! cgsem_dropin.f90 -- the CG spectral-element elliptic solver under the
! reference's own module, type and procedure names, so that a reference
! driver program compiles unchanged against it and solves on the GPU.
!
! Modules (each mirrors the reference module of the same name, as far as a
! CGSEM driver uses it; ModelPrecision, ConstantsDictionary and QuadMeshClass
! come from sw_dropin.f90):
!   CGsemParams_Class       src/cgsem/CGsemParams_Class.f90 (:69-144): the
!                           runtime.params namelists and their defaults
!   CGsemElliptic_2D_Class  src/cgsem/CGsemElliptic_2D_Class.f90: TYPE
!                           CGsemElliptic_2D (:57-93) with Build (:99-196),
!                           Trash, Mask, MatrixAction, Residual,
!                           SetDirichletBoundaryCondition, Solve (the CGSEM
!                           PCG of :1114-1207), CoarseToFine, WriteTecplot,
!                           WriteResidual and ConstructMatrix (:1215-1453),
!                           and its component preconditioner (TYPE
!                           CGsemPreconditioner_2D) with ConstructPCMatrix
!                           (src/cgsem/CGsemPreconditioner_2D_Class.f90:750-851).
! The solver itself is the C ABI's (self_cgsem2d_*): operator, assembly,
! FEM preconditioner and PCG run on the device; this layer holds the
! reference's public arrays (solution, source, fluxCoeff, resi) and writes
! the reference's output files.
MODULE CGsemParams_Class
  USE ModelPrecision
  USE ConstantsDictionary
  IMPLICIT NONE
  CHARACTER(1), PARAMETER, PRIVATE :: blank = ' '

  TYPE CGsemParams
    ! SolverCriteria
    INTEGER    :: MaximumIterates, pcIterates, pcDegree
    REAL(prec) :: tolerance, pcTolerance
    ! SpaceManagement
    CHARACTER(50) :: SpecMeshFile
    INTEGER       :: polyDeg, nXElem, nYElem, nPlot
    REAL(prec)    :: dxPlot, xScale
  CONTAINS
    PROCEDURE :: Build => BuildParams_CGsem
  END TYPE CGsemParams

CONTAINS

  ! BuildParams (:69-144): defaults, then the namelists SolverCriteria and
  ! SpaceManagement of ./runtime.params, echoed to standard output.
  SUBROUTINE BuildParams_CGsem(thisParam)
    CLASS(CGsemParams), INTENT(out) :: thisParam
    INTEGER       :: nUnit
    INTEGER       :: MaximumIterates, pcIterates, pcDegree
    REAL(prec)    :: tolerance, pcTolerance
    CHARACTER(50) :: SpecMeshFile
    INTEGER       :: polyDeg, nXElem, nYElem, nPlot
    REAL(prec)    :: xScale
    NAMELIST / SolverCriteria / MaximumIterates, pcIterates, tolerance, pcTolerance, pcDegree
    NAMELIST / SpaceManagement / SpecMeshFile, polyDeg, nXElem, nYElem, nPlot, xScale

    MaximumIterates = 500
    pcIterates = 50
    tolerance = 10.0_prec**(-8)
    pcTolerance = 10.0_prec**(-4)
    pcDegree = 1
    SpecMeshFile = blank
    polyDeg = 5
    nXElem = 5
    nYElem = 5
    nPlot = 10
    xScale = ONE

    OPEN(NEWUNIT=nUnit, FILE='runtime.params', STATUS='old', ACTION='read')
    READ(UNIT=nUnit, NML=SolverCriteria)
    READ(UNIT=nUnit, NML=SpaceManagement)
    CLOSE(UNIT=nUnit)
    WRITE(UNIT=*, NML=SolverCriteria)
    WRITE(UNIT=*, NML=SpaceManagement)

    thisParam % MaximumIterates = MaximumIterates
    thisParam % pcIterates = pcIterates
    thisParam % tolerance = tolerance
    thisParam % pcTolerance = pcTolerance
    thisParam % pcDegree = 1
    thisParam % SpecMeshFile = SpecMeshFile
    thisParam % polyDeg = polyDeg
    thisParam % nXElem = nXElem
    thisParam % nYElem = nYElem
    thisParam % nPlot = nPlot
    thisParam % dxPlot = 2.0_prec / REAL(nPlot, prec)
    thisParam % xScale = xScale
  END SUBROUTINE BuildParams_CGsem

END MODULE CGsemParams_Class

MODULE CGsemElliptic_2D_Class
  USE ISO_C_BINDING
  USE ModelPrecision
  USE ConstantsDictionary
  USE QuadMeshClass
  USE CGsemParams_Class
  USE SELF_AMD_CABI
  IMPLICIT NONE

  ! The preconditioner as a driver sees it: CGsemPreconditioner_2D's degree
  ! (1, the linear sub-mesh through the Gauss-Lobatto nodes), its iteration
  ! limit and tolerance; the operator itself lives with the solver on the
  ! device (self_cgsem2d_set_preconditioner).
  TYPE CGsemPreconditioner_2D
    INTEGER    :: nS = 1, nP = 1, maxIters = 0
    REAL(prec) :: tolerance = ZERO
    TYPE(C_PTR) :: handle = C_NULL_PTR   ! the solver's
    INTEGER :: nOrig = 0, nElemsOrig = 0
    REAL(prec), ALLOCATABLE :: x(:,:,:), y(:,:,:)  ! the original mesh's node positions
  CONTAINS
    PROCEDURE :: ConstructPCMatrix
  END TYPE CGsemPreconditioner_2D

  TYPE CGsemElliptic_2D
    INTEGER                        :: maxIters, nPlot, nS, nP
    REAL(prec)                     :: tolerance
    TYPE(CGsemPreconditioner_2D)   :: preconditioner
    TYPE(QuadMesh)                 :: mesh
    REAL(prec), ALLOCATABLE        :: solution(:,:,:)
    REAL(prec), ALLOCATABLE        :: source(:,:,:)
    REAL(prec), ALLOCATABLE        :: fluxCoeff(:,:,:)
    REAL(prec), ALLOCATABLE        :: resi(:)
    REAL(prec), ALLOCATABLE        :: plMatS(:,:), plMatP(:,:)
    TYPE(C_PTR) :: handle = C_NULL_PTR
    INTEGER     :: pcIterates = 0
  CONTAINS
    PROCEDURE :: Build => Build_CGsemElliptic_2D
    PROCEDURE :: Trash => Trash_CGsemElliptic_2D
    PROCEDURE :: Mask => Mask_CGsemElliptic_2D
    PROCEDURE :: MatrixAction => MatrixAction_CGsemElliptic_2D
    PROCEDURE :: Residual => Residual_CGsemElliptic_2D
    PROCEDURE :: SetDirichletBoundaryCondition => SetDirichletBoundaryCondition_CGsemElliptic_2D
    PROCEDURE :: Solve => Solve_PCConjugateGradient
    PROCEDURE :: CoarseToFine => CoarseToFine_CGsemElliptic_2D
    PROCEDURE :: WriteTecplot => WriteTecplot_CGsemElliptic_2D
    PROCEDURE :: WriteResidual => WriteResidual_CGsemElliptic_2D
    PROCEDURE :: ConstructMatrix
  END TYPE CGsemElliptic_2D

CONTAINS

  ! Build (:99-154) with BuildQuadMesh (:158-196): the default mesh
  ! (LoadDefaultMesh: nXElem x nYElem elements on the unit square, every
  ! boundary edge Dirichlet) or the ISM-V2 file SpecMeshFile, on the
  ! Gauss-Lobatto nodes of degree polyDeg; solution = source = 0,
  ! fluxCoeff = 1; the plotting matrices to nPlot + 1 uniform points; the
  ! preconditioner with pcIterates inner iterations to pcTolerance on the
  ! coefficient the reference maps at Build (ONE).
  SUBROUTINE Build_CGsemElliptic_2D(myCGSEM, params)
    CLASS(CGsemElliptic_2D), INTENT(inout) :: myCGSEM
    TYPE(CGsemParams), INTENT(in)          :: params
    INTEGER(C_INT) :: N, nE, nEd, nNd, gDeg, iX, iY, e, n0
    INTEGER(C_INT), ALLOCATABLE :: edges(:), elemNodes(:,:), nodeType(:)
    REAL(C_DOUBLE), ALLOCATABLE :: nodes(:), w(:), bw(:), sNew(:), nHat(:), nLen(:), nodeX(:), nodeY(:)
    CHARACTER(KIND=C_CHAR, LEN=256) :: path
    INTEGER :: i

    N = params % polyDeg
    myCGSEM % nS = N
    myCGSEM % nP = N
    myCGSEM % nPlot = params % nPlot
    myCGSEM % tolerance = params % tolerance
    myCGSEM % maxIters = params % MaximumIterates
    ALLOCATE(nodes(0:N), w(0:N), bw(0:N))
    CALL self_check(self_legendre_quadrature(N, SELF_GAUSS_LOBATTO, nodes, w), 'Build')

    IF (LEN_TRIM(params % SpecMeshFile) == 0) THEN
      PRINT *, ' Loading default 2-D mesh.'
      CALL self_check(self_quadmesh_counts(params % nXElem, params % nYElem, 0_C_INT, 0_C_INT, nE, nEd), 'Build')
      nNd = (params % nXElem + 1) * (params % nYElem + 1)
      CALL alloc_mesh()
      CALL self_check(self_quadmesh_structured(params % nXElem, params % nYElem, N, 0_C_INT, 0_C_INT, ONE, ONE, &
                                               ZERO, nodes, edges, myCGSEM % mesh % x, myCGSEM % mesh % y, &
                                               myCGSEM % mesh % dxds, myCGSEM % mesh % dxdp, &
                                               myCGSEM % mesh % dyds, myCGSEM % mesh % dydp, myCGSEM % mesh % J, &
                                               nHat, nLen), 'Build')
      ! corner node ids of LoadDefaultMesh: SW, SE, NE, NW
      DO iY = 1, params % nYElem
        DO iX = 1, params % nXElem
          e = iX + (iY - 1) * params % nXElem
          n0 = iX + (iY - 1) * (params % nXElem + 1)
          elemNodes(:, e) = (/ n0, n0 + 1, n0 + 1 + params % nXElem + 1, n0 + params % nXElem + 1 /)
        ENDDO
      ENDDO
    ELSE
      PRINT *, 'Reading 2-D mesh from '//TRIM(params % SpecMeshFile)//'.'
      path = TRIM(params % SpecMeshFile)//C_NULL_CHAR
      CALL self_check(self_specmesh_counts(path, nNd, nEd, nE, gDeg), 'Build')
      CALL alloc_mesh()
      ALLOCATE(nodeType(nNd), nodeX(nNd), nodeY(nNd))
      CALL self_check(self_specmesh_read(path, N, nodes, edges, elemNodes, nodeType, nodeX, nodeY, &
                                         myCGSEM % mesh % x, myCGSEM % mesh % y, myCGSEM % mesh % dxds, &
                                         myCGSEM % mesh % dxdp, myCGSEM % mesh % dyds, myCGSEM % mesh % dydp, &
                                         myCGSEM % mesh % J, nHat, nLen), 'Build')
    ENDIF
    myCGSEM % mesh % nElems = nE
    myCGSEM % mesh % nS = N
    myCGSEM % mesh % nP = N
    CALL self_check(self_cgsem2d_create(N, nE, nEd, nNd, edges, elemNodes, myCGSEM % mesh % dxds, &
                                        myCGSEM % mesh % dxdp, myCGSEM % mesh % dyds, myCGSEM % mesh % dydp, &
                                        myCGSEM % mesh % J, myCGSEM % handle), 'Build')

    ALLOCATE(myCGSEM % solution(0:N, 0:N, nE), myCGSEM % source(0:N, 0:N, nE), myCGSEM % fluxCoeff(0:N, 0:N, nE))
    ALLOCATE(myCGSEM % resi(0:myCGSEM % maxIters))
    myCGSEM % solution = ZERO
    myCGSEM % source = ZERO
    myCGSEM % fluxCoeff = ONE

    ALLOCATE(sNew(0:params % nPlot), myCGSEM % plMatS(0:params % nPlot, 0:N), myCGSEM % plMatP(0:params % nPlot, 0:N))
    DO i = 0, params % nPlot  ! UniformPoints(-1, 1, nPlot), src/common/CommonRoutines.f90:460-482
      sNew(i) = -ONE + (TWO / REAL(params % nPlot, prec)) * REAL(i, prec)
    ENDDO
    CALL self_check(self_barycentric_weights(N, nodes, bw), 'Build')
    CALL self_check(self_interpolation_matrix(N, nodes, bw, INT(params % nPlot, C_INT), sNew, myCGSEM % plMatS), &
                    'Build')
    myCGSEM % plMatP = myCGSEM % plMatS

    myCGSEM % pcIterates = params % pcIterates
    myCGSEM % preconditioner % maxIters = params % pcIterates
    myCGSEM % preconditioner % tolerance = params % pcTolerance
    myCGSEM % preconditioner % handle = myCGSEM % handle
    myCGSEM % preconditioner % nOrig = N
    myCGSEM % preconditioner % nElemsOrig = nE
    myCGSEM % preconditioner % x = myCGSEM % mesh % x
    myCGSEM % preconditioner % y = myCGSEM % mesh % y
    CALL self_check(self_cgsem2d_set_preconditioner(myCGSEM % handle, myCGSEM % mesh % x, myCGSEM % mesh % y, &
                                                    INT(params % pcIterates, C_INT), params % pcTolerance, 0_C_INT), &
                    'Build')
  CONTAINS
    SUBROUTINE alloc_mesh()
      ALLOCATE(edges(6 * nEd), elemNodes(4, nE), nHat(nE * 4 * (N + 1) * 2), nLen(nE * 4 * (N + 1)))
      ALLOCATE(myCGSEM % mesh % x(0:N, 0:N, nE), myCGSEM % mesh % y(0:N, 0:N, nE), &
               myCGSEM % mesh % dxds(0:N, 0:N, nE), myCGSEM % mesh % dxdp(0:N, 0:N, nE), &
               myCGSEM % mesh % dyds(0:N, 0:N, nE), myCGSEM % mesh % dydp(0:N, 0:N, nE), &
               myCGSEM % mesh % J(0:N, 0:N, nE))
    END SUBROUTINE alloc_mesh
  END SUBROUTINE Build_CGsemElliptic_2D

  SUBROUTINE Trash_CGsemElliptic_2D(myCGSEM)
    CLASS(CGsemElliptic_2D), INTENT(inout) :: myCGSEM
    IF (C_ASSOCIATED(myCGSEM % handle)) CALL self_check(self_cgsem2d_destroy(myCGSEM % handle), 'Trash')
    myCGSEM % handle = C_NULL_PTR
    myCGSEM % preconditioner % handle = C_NULL_PTR
    IF (ALLOCATED(myCGSEM % solution)) DEALLOCATE(myCGSEM % solution, myCGSEM % source, myCGSEM % fluxCoeff, &
                                                  myCGSEM % resi, myCGSEM % plMatS, myCGSEM % plMatP)
    IF (ALLOCATED(myCGSEM % mesh % x)) DEALLOCATE(myCGSEM % mesh % x, myCGSEM % mesh % y, myCGSEM % mesh % dxds, &
                                                  myCGSEM % mesh % dxdp, myCGSEM % mesh % dyds, &
                                                  myCGSEM % mesh % dydp, myCGSEM % mesh % J)
    IF (ALLOCATED(myCGSEM % preconditioner % x)) DEALLOCATE(myCGSEM % preconditioner % x, &
                                                            myCGSEM % preconditioner % y)
  END SUBROUTINE Trash_CGsemElliptic_2D

  ! Mask (:350-430): secondary copies of shared nodes and Dirichlet nodes to 0
  SUBROUTINE Mask_CGsemElliptic_2D(myCGSEM, u)
    CLASS(CGsemElliptic_2D), INTENT(inout) :: myCGSEM
    REAL(prec), INTENT(inout) :: u(0:myCGSEM % nS, 0:myCGSEM % nP, 1:myCGSEM % mesh % nElems)
    CALL self_check(self_cgsem2d_mask(myCGSEM % handle, u), 'Mask')
  END SUBROUTINE Mask_CGsemElliptic_2D

  ! MatrixAction (:930-976) of a masked array (fluxCoeff as it stands)
  SUBROUTINE MatrixAction_CGsemElliptic_2D(myCGSEM, u, Au)
    CLASS(CGsemElliptic_2D), INTENT(inout) :: myCGSEM
    REAL(prec), INTENT(in)  :: u(0:myCGSEM % nS, 0:myCGSEM % nP, 1:myCGSEM % mesh % nElems)
    REAL(prec), INTENT(out) :: Au(0:myCGSEM % nS, 0:myCGSEM % nP, 1:myCGSEM % mesh % nElems)
    CALL self_check(self_cgsem2d_set_flux_coefficient(myCGSEM % handle, myCGSEM % fluxCoeff), 'MatrixAction')
    CALL self_check(self_cgsem2d_matrix_action(myCGSEM % handle, u, Au), 'MatrixAction')
  END SUBROUTINE MatrixAction_CGsemElliptic_2D

  ! Residual (:980-1068) of the current solution with the Dirichlet values
  ! dirF(x, y), masked; the solution then holds them (:1062-1066).
  SUBROUTINE Residual_CGsemElliptic_2D(myCGSEM, dirF, r)
    CLASS(CGsemElliptic_2D), INTENT(inout) :: myCGSEM
    REAL(prec), EXTERNAL    :: dirF
    REAL(prec), INTENT(out) :: r(0:myCGSEM % nS, 0:myCGSEM % nP, 1:myCGSEM % mesh % nElems)
    INTERFACE
      INTEGER(C_INT) FUNCTION residual_c(h, r) BIND(C, name="self_cgsem2d_residual")
        IMPORT :: C_INT, C_PTR, C_DOUBLE
        TYPE(C_PTR), VALUE :: h
        REAL(C_DOUBLE), INTENT(out) :: r(*)
      END FUNCTION residual_c
    END INTERFACE
    CALL upload(myCGSEM, dirF)
    CALL self_check(residual_c(myCGSEM % handle, r), 'Residual')
    CALL myCGSEM % Mask(myCGSEM % solution)
    CALL myCGSEM % SetDirichletBoundaryCondition(dirF)
  END SUBROUTINE Residual_CGsemElliptic_2D

  ! SetDirichletBoundaryCondition (:234-300): the solution takes dirF(x, y)
  ! at every Dirichlet node (all copies)
  SUBROUTINE SetDirichletBoundaryCondition_CGsemElliptic_2D(myCGSEM, dirF)
    CLASS(CGsemElliptic_2D), INTENT(inout) :: myCGSEM
    REAL(prec), EXTERNAL :: dirF
    INTEGER(C_INT), ALLOCATABLE :: gid(:,:,:)
    INTEGER :: i, j, k
    ALLOCATE(gid(0:myCGSEM % nS, 0:myCGSEM % nP, myCGSEM % mesh % nElems))
    CALL self_check(self_cgsem2d_dof_map(myCGSEM % handle, gid), 'SetDirichletBoundaryCondition')
    DO k = 1, myCGSEM % mesh % nElems
      DO j = 0, myCGSEM % nP
        DO i = 0, myCGSEM % nS
          IF (gid(i, j, k) < 0) myCGSEM % solution(i, j, k) = dirF(myCGSEM % mesh % x(i, j, k), &
                                                                    myCGSEM % mesh % y(i, j, k))
        ENDDO
      ENDDO
    ENDDO
  END SUBROUTINE SetDirichletBoundaryCondition_CGsemElliptic_2D

  ! the host arrays to the device: coefficient, source, solution, and the
  ! Dirichlet values dirF sampled at every node (only Dirichlet nodes read)
  SUBROUTINE upload(myCGSEM, dirF)
    CLASS(CGsemElliptic_2D), INTENT(inout) :: myCGSEM
    REAL(prec), EXTERNAL :: dirF
    REAL(prec), ALLOCATABLE :: d(:,:,:)
    INTEGER :: i, j, k
    ALLOCATE(d(0:myCGSEM % nS, 0:myCGSEM % nP, myCGSEM % mesh % nElems))
    DO k = 1, myCGSEM % mesh % nElems
      DO j = 0, myCGSEM % nP
        DO i = 0, myCGSEM % nS
          d(i, j, k) = dirF(myCGSEM % mesh % x(i, j, k), myCGSEM % mesh % y(i, j, k))
        ENDDO
      ENDDO
    ENDDO
    CALL self_check(self_cgsem2d_set_flux_coefficient(myCGSEM % handle, myCGSEM % fluxCoeff), 'upload')
    CALL self_check(self_cgsem2d_set_source(myCGSEM % handle, myCGSEM % source), 'upload')
    CALL self_check(self_cgsem2d_set_dirichlet(myCGSEM % handle, d), 'upload')
    CALL self_check(self_cgsem2d_set_solution(myCGSEM % handle, myCGSEM % solution), 'upload')
  END SUBROUTINE upload

  ! Solve (:1114-1207): the CGSEM PCG from the current solution, with the
  ! preconditioner of Build (identity when pcIterates = 0), to
  ! |r| / |r0| < tolerance in at most maxIters iterations; resi(0:maxIters)
  ! the residual history.  ioerr as the reference leaves it: -2 when the
  ! loop converged (the reference's ioerr = 0 sits after its EXIT), -1 when
  ! it did not.  The solution is unmasked and holds the Dirichlet values.
  SUBROUTINE Solve_PCConjugateGradient(myCGSEM, dirF, ioerr)
    CLASS(CGsemElliptic_2D), INTENT(inout) :: myCGSEM
    REAL(prec), EXTERNAL :: dirF
    INTEGER, INTENT(out) :: ioerr
    INTEGER(C_INT) :: io, its, pc
    CALL upload(myCGSEM, dirF)
    pc = MERGE(2_C_INT, 0_C_INT, myCGSEM % pcIterates > 0)
    myCGSEM % resi = ZERO
    CALL self_check(self_cgsem2d_solve(myCGSEM % handle, SELF_KRYLOV_CGSEM_PCG, pc, myCGSEM % tolerance, &
                                       INT(myCGSEM % maxIters, C_INT), 0_C_INT, 0_C_INT, io, its, myCGSEM % resi, &
                                       INT(myCGSEM % maxIters + 1, C_INT), C_NULL_PTR), 'Solve')
    ioerr = io
    IF (ioerr == -1) THEN
      PRINT *, 'MODULE IterativeSolvers : ConjugateGradient failed to converge '
      PRINT *, 'Last L-2 residual : ', myCGSEM % resi(its)
    ENDIF
    CALL self_check(self_cgsem2d_get_solution(myCGSEM % handle, myCGSEM % solution), 'Solve')
    CALL myCGSEM % SetDirichletBoundaryCondition(dirF)
  END SUBROUTINE Solve_PCConjugateGradient

  ! CoarseToFine (:1215-1261): positions, solution and source of element iEl
  ! at the plotting points
  SUBROUTINE CoarseToFine_CGsemElliptic_2D(myCGSEM, iEl, x, y, sol, source)
    CLASS(CGsemElliptic_2D), INTENT(in) :: myCGSEM
    INTEGER, INTENT(in)                 :: iEl
    REAL(prec), INTENT(out), DIMENSION(0:myCGSEM % nPlot, 0:myCGSEM % nPlot) :: x, y, sol, source
    CALL c2f(myCGSEM % mesh % x(:, :, iEl), x)
    CALL c2f(myCGSEM % mesh % y(:, :, iEl), y)
    CALL c2f(myCGSEM % solution(:, :, iEl), sol)
    CALL c2f(myCGSEM % source(:, :, iEl), source)
  CONTAINS
    SUBROUTINE c2f(fOld, fNew)
      REAL(prec), INTENT(in)  :: fOld(0:myCGSEM % nS, 0:myCGSEM % nP)
      REAL(prec), INTENT(out) :: fNew(0:myCGSEM % nPlot, 0:myCGSEM % nPlot)
      REAL(prec) :: a(0:myCGSEM % nS, 0:myCGSEM % nP), b(0:myCGSEM % nPlot, 0:myCGSEM % nPlot)
      a = fOld
      CALL self_check(self_coarse_to_fine_2d(INT(myCGSEM % nS, C_INT), INT(myCGSEM % nP, C_INT), 1_C_SIZE_T, &
                                             myCGSEM % plMatS, myCGSEM % plMatP, INT(myCGSEM % nPlot, C_INT), &
                                             INT(myCGSEM % nPlot, C_INT), a, b), 'CoarseToFine')
      fNew = b
    END SUBROUTINE c2f
  END SUBROUTINE CoarseToFine_CGsemElliptic_2D

  ! WriteTecplot (:1265-1312): <filename>.tec (or CGsemSol.tec), per element
  ! a zone of X, Y, sol, source at the plotting points; the whole mesh goes
  ! through one batched CoarseToFine per field.
  SUBROUTINE WriteTecplot_CGsemElliptic_2D(myCGSEM, filename)
    CLASS(CGsemElliptic_2D), INTENT(inout) :: myCGSEM
    CHARACTER(*), INTENT(in), OPTIONAL     :: filename
    INTEGER :: iX, iY, iEl, nS, nE, fUnit
    CHARACTER(len=5) :: zoneID
    REAL(prec), ALLOCATABLE :: xf(:,:,:), yf(:,:,:), sf(:,:,:), qf(:,:,:)
    nS = myCGSEM % nPlot
    nE = myCGSEM % mesh % nElems
    ALLOCATE(xf(0:nS, 0:nS, nE), yf(0:nS, 0:nS, nE), sf(0:nS, 0:nS, nE), qf(0:nS, 0:nS, nE))
    CALL c2f_all(myCGSEM % mesh % x, xf)
    CALL c2f_all(myCGSEM % mesh % y, yf)
    CALL c2f_all(myCGSEM % solution, sf)
    CALL c2f_all(myCGSEM % source, qf)
    IF (PRESENT(filename)) THEN
      OPEN(NEWUNIT=fUnit, FILE=TRIM(filename)//'.tec', FORM='FORMATTED')
    ELSE
      OPEN(NEWUNIT=fUnit, FILE='CGsemSol.tec', FORM='FORMATTED')
    ENDIF
    WRITE(fUnit, *) 'VARIABLES = "X", "Y", "sol", "source"'
    DO iEl = 1, nE
      WRITE(zoneID, '(I5.5)') iEl
      WRITE(fUnit, *) 'ZONE T="el'//TRIM(zoneID)//'", I=', nS + 1, ', J=', nS + 1, ',F=POINT'
      DO iY = 0, nS
        DO iX = 0, nS
          WRITE(fUnit, *) xf(iX, iY, iEl), yf(iX, iY, iEl), sf(iX, iY, iEl), qf(iX, iY, iEl)
        ENDDO
      ENDDO
    ENDDO
    CLOSE(fUnit)
  CONTAINS
    SUBROUTINE c2f_all(fOld, fNew)
      REAL(prec), INTENT(in)  :: fOld(0:myCGSEM % nS, 0:myCGSEM % nP, nE)
      REAL(prec), INTENT(out) :: fNew(0:nS, 0:nS, nE)
      REAL(prec), ALLOCATABLE :: a(:,:,:)
      ALLOCATE(a(0:myCGSEM % nS, 0:myCGSEM % nP, nE))
      a = fOld
      CALL self_check(self_coarse_to_fine_2d(INT(myCGSEM % nS, C_INT), INT(myCGSEM % nP, C_INT), &
                                             INT(nE, C_SIZE_T), myCGSEM % plMatS, myCGSEM % plMatP, &
                                             INT(nS, C_INT), INT(nS, C_INT), a, fNew), 'WriteTecplot')
    END SUBROUTINE c2f_all
  END SUBROUTINE WriteTecplot_CGsemElliptic_2D

  ! WriteResidual (:1316-1346): rFile (or residual.curve), resi(0:maxIters)
  SUBROUTINE WriteResidual_CGsemElliptic_2D(myCGSEM, rFile)
    CLASS(CGsemElliptic_2D), INTENT(in) :: myCGSEM
    CHARACTER(*), INTENT(in), OPTIONAL  :: rFile
    INTEGER :: i, fUnit
    CHARACTER(100) :: localFile
    IF (PRESENT(rFile)) THEN
      localFile = rFile
    ELSE
      localFile = 'residual.curve'
    ENDIF
    OPEN(NEWUNIT=fUnit, FILE=TRIM(localFile), FORM='FORMATTED')
    WRITE(fUnit, *) '#residual'
    DO i = 0, myCGSEM % maxIters
      WRITE(fUnit, '(I4,1x,E17.8)') i, myCGSEM % resi(i)
    ENDDO
    CLOSE(UNIT=fUnit)
  END SUBROUTINE WriteResidual_CGsemElliptic_2D

  ! ConstructMatrix (:1350-1453): the operator on the free nodes (mask 1),
  ! row by row from MatrixAction on unit vectors, to Omat.txt; the mask and
  ! node positions to omask.tec.
  SUBROUTINE ConstructMatrix(myPC)
    CLASS(CGsemElliptic_2D), INTENT(inout) :: myPC
    INTEGER :: iEl, iS, iP, jEl, jS, jP, nEl, nS, nP, row, col, nfree, fUnit
    REAL(prec), ALLOCATABLE :: ei(:,:,:), Aei(:,:,:), thismask(:,:,:), A(:,:)
    CHARACTER(5) :: zoneID
    nS = myPC % nS
    nP = myPC % nP
    nEl = myPC % mesh % nElems
    ALLOCATE(ei(0:nS, 0:nP, nEl), Aei(0:nS, 0:nP, nEl), thismask(0:nS, 0:nP, nEl))
    thismask = ONE
    CALL myPC % Mask(thismask)
    nfree = NINT(SUM(thismask))
    ALLOCATE(A(nfree, nfree))
    row = 0
    DO iEl = 1, nEl
      DO iP = 0, nP
        DO iS = 0, nS
          IF (thismask(iS, iP, iEl) == ONE) THEN
            ei = ZERO
            ei(iS, iP, iEl) = ONE
            row = row + 1
            CALL myPC % MatrixAction(ei, Aei)
            col = 0
            DO jEl = 1, nEl
              DO jP = 0, nP
                DO jS = 0, nS
                  IF (thismask(jS, jP, jEl) == ONE) THEN
                    col = col + 1
                    A(row, col) = Aei(jS, jP, jEl)
                  ENDIF
                ENDDO
              ENDDO
            ENDDO
          ENDIF
        ENDDO
      ENDDO
    ENDDO
    OPEN(NEWUNIT=fUnit, FILE='Omat.txt', FORM='FORMATTED')
    DO row = 1, nfree
      WRITE(fUnit, *) A(row, :)
    ENDDO
    CLOSE(fUnit)
    OPEN(NEWUNIT=fUnit, FILE='omask.tec', FORM='FORMATTED')
    WRITE(fUnit, *) 'VARIABLES = "X", "Y", "mask"'
    DO iEl = 1, nEl
      WRITE(zoneID, '(I5.5)') iEl
      WRITE(fUnit, *) 'ZONE T="el'//TRIM(zoneID)//'", I=', nS + 1, ', J=', nP + 1, ',F=POINT'
      DO iP = 0, nS
        DO iS = 0, nS
          WRITE(fUnit, *) myPC % mesh % x(iS, iP, iEl), myPC % mesh % y(iS, iP, iEl), thismask(iS, iP, iEl)
        ENDDO
      ENDDO
    ENDDO
    CLOSE(fUnit)
  END SUBROUTINE ConstructMatrix

  ! ConstructPCMatrix (CGsemPreconditioner_2D_Class.f90:750-851): the
  ! preconditioner's operator on the free nodes of the linear sub-mesh to
  ! PCmat.txt, its mask and corner positions to mask.tec.  The sub-mesh
  ! (ConstructLinearMeshFromOther, src/geom/QuadMeshClass.f90:1863-2096)
  ! numbers sub-element (iS, iP) of element iEl as iS + nS (iP - 1 + nP
  ! (iEl - 1)), its corners (0,0), (1,0), (0,1), (1,1) being the original
  ! nodes (iS-1, iP-1), (iS, iP-1), (iS-1, iP), (iS, iP); a node keeps its
  ! value in the first element of its list (elements join the lists in
  ! ascending order, :1631-1658) unless it is Dirichlet (Mask, :230-282).
  ! The operator's rows come assembled from the device
  ! (self_cgsem2d_pc_matrix) over the solver's global DOFs, which are
  ! exactly the sub-mesh's free nodes.
  SUBROUTINE ConstructPCMatrix(myPC)
    CLASS(CGsemPreconditioner_2D), INTENT(inout) :: myPC
    INTEGER :: N, nEl, iEl, iS, iP, jS, jP, eID, nSub, row, col, q, fUnit, nfree
    INTEGER(C_INT) :: width, nG
    INTEGER(C_INT), ALLOCATABLE, TARGET :: cols(:,:)
    REAL(C_DOUBLE), ALLOCATABLE, TARGET :: vals(:,:)
    INTEGER(C_INT), ALLOCATABLE :: gid(:,:,:)
    INTEGER, ALLOCATABLE :: order(:), pos(:)
    LOGICAL, ALLOCATABLE :: seen(:)
    REAL(prec), ALLOCATABLE :: thismask(:,:,:), A(:,:)
    CHARACTER(5) :: zoneID
    N = myPC % nOrig
    nEl = myPC % nElemsOrig
    nSub = nEl * N * N
    ALLOCATE(gid(0:N, 0:N, nEl), thismask(0:1, 0:1, nSub))
    CALL self_check(self_cgsem2d_dof_map(myPC % handle, gid), 'ConstructPCMatrix')
    CALL self_check(self_cgsem2d_global_dofs(myPC % handle, nG), 'ConstructPCMatrix')
    ALLOCATE(seen(0:nG - 1), order(nG), pos(0:nG - 1))
    seen = .FALSE.
    nfree = 0
    DO iEl = 1, nEl
      DO iP = 1, N
        DO iS = 1, N
          eID = iS + N * ((iP - 1) + (iEl - 1) * N)
          DO jP = 0, 1
            DO jS = 0, 1
              q = gid(iS - 1 + jS, iP - 1 + jP, iEl)
              thismask(jS, jP, eID) = ZERO
              IF (q >= 0) THEN
                IF (.NOT. seen(q)) THEN
                  seen(q) = .TRUE.
                  nfree = nfree + 1
                  order(nfree) = q
                  pos(q) = nfree
                  thismask(jS, jP, eID) = ONE
                ENDIF
              ENDIF
            ENDDO
          ENDDO
        ENDDO
      ENDDO
    ENDDO
    CALL self_check(self_cgsem2d_pc_matrix(myPC % handle, width, C_NULL_PTR, C_NULL_PTR), 'ConstructPCMatrix')
    ALLOCATE(cols(width, 0:nG - 1), vals(width, 0:nG - 1), A(nfree, nfree))
    CALL self_check(self_cgsem2d_pc_matrix(myPC % handle, width, C_LOC(cols), C_LOC(vals)), 'ConstructPCMatrix')
    A = ZERO
    DO row = 1, nfree
      DO q = 1, width
        col = cols(q, order(row))
        IF (vals(q, order(row)) /= ZERO) A(row, pos(col)) = A(row, pos(col)) + vals(q, order(row))
      ENDDO
    ENDDO
    OPEN(NEWUNIT=fUnit, FILE='PCmat.txt', FORM='FORMATTED')
    DO row = 1, nfree
      WRITE(fUnit, *) A(row, :)
    ENDDO
    CLOSE(fUnit)
    OPEN(NEWUNIT=fUnit, FILE='mask.tec', FORM='FORMATTED')
    WRITE(fUnit, *) 'VARIABLES = "X", "Y", "mask"'
    DO iEl = 1, nEl
      DO iP = 1, N
        DO iS = 1, N
          eID = iS + N * ((iP - 1) + (iEl - 1) * N)
          WRITE(zoneID, '(I5.5)') eID
          WRITE(fUnit, *) 'ZONE T="el'//TRIM(zoneID)//'", I=', myPC % nS + 1, ', J=', myPC % nP + 1, ',F=POINT'
          DO jP = 0, 1
            DO jS = 0, 1
              WRITE(fUnit, *) myPC % x(iS - 1 + jS, iP - 1 + jP, iEl), myPC % y(iS - 1 + jS, iP - 1 + jP, iEl), &
                              thismask(jS, jP, eID)
            ENDDO
          ENDDO
        ENDDO
      ENDDO
    ENDDO
    CLOSE(fUnit)
  END SUBROUTINE ConstructPCMatrix

END MODULE CGsemElliptic_2D_Class
