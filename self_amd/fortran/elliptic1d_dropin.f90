! elliptic1d_dropin.f90 -- the 1-D continuous-Galerkin elliptic example
! under the reference's own module, type and procedure names, so that its
! driver program (examples/elliptic1d/parabola/localmods/Elliptic1D_Driver.f90)
! compiles unchanged against it and runs its conjugate-gradient solve on the
! GPU.
!
! Modules (ModelPrecision and ConstantsDictionary come from sw_dropin.f90):
!   Elliptic1DParams_Class  src/cgsem/Elliptic1DParams_Class.f90 (:64-125):
!                           the runtime.params namelists and their defaults
!   Elliptic1D_Class        examples/elliptic1d/parabola/localmods/
!                           Elliptic1D_Class.f90: TYPE Elliptic1D (:63-95),
!                           the IterativeSolver components and accessors it
!                           inherits (src/iterativesolve/
!                           IterativeSolvers_Class.f90:55-78, :234-320),
!                           Build, Trash, CopyTo, CopyFrom, UnMaskSolution,
!                           Solve (the ConjugateGradient loop, :328-405),
!                           WriteTecplot and WriteResidual (:830-860).
! The operator (the element second-derivative matrices of Build :111-176,
! global sum, masking and Dirichlet ends) and the CG loop are the C ABI's
! (self_linop_elliptic1d_*, self_krylov_solve), on the device; this layer
! holds the reference's public arrays (sol, resi, cgsol, source) and writes
! its output files.
MODULE Elliptic1DParams_Class
  USE ModelPrecision
  USE ConstantsDictionary
  IMPLICIT NONE

  TYPE Elliptic1DParams
    INTEGER    :: MaximumIterates
    REAL(prec) :: tolerance
    INTEGER    :: polyDeg
    INTEGER    :: nElems
    INTEGER    :: nPlot
    REAL(prec) :: dxPlot
    REAL(prec) :: xScale
  CONTAINS
    PROCEDURE :: Build => BuildParams_Elliptic1D
  END TYPE Elliptic1DParams

CONTAINS

  ! :64-125: defaults, then the SolverCriteria and SpaceManagement
  ! namelists of runtime.params, echoed to standard output
  SUBROUTINE BuildParams_Elliptic1D(thisParam)
    CLASS(Elliptic1DParams), INTENT(out) :: thisParam
    INTEGER    :: nUnit
    INTEGER    :: MaximumIterates, polyDeg, nElems, nPlot
    REAL(prec) :: tolerance, xScale
    NAMELIST / SolverCriteria / MaximumIterates, tolerance
    NAMELIST / SpaceManagement / polyDeg, nElems, nPlot, xScale

    MaximumIterates = 1000
    tolerance = 10.0_prec**(-8)
    polyDeg = 5
    nElems = 5
    nPlot = 10
    xScale = ONE

    OPEN(NEWUNIT=nUnit, FILE='runtime.params')
    READ(UNIT=nUnit, NML=SolverCriteria)
    READ(UNIT=nUnit, NML=SpaceManagement)
    CLOSE(UNIT=nUnit)

    WRITE(UNIT=*, NML=SolverCriteria)
    WRITE(UNIT=*, NML=SpaceManagement)

    thisParam % MaximumIterates = MaximumIterates
    thisParam % tolerance = tolerance
    thisParam % polyDeg = polyDeg
    thisParam % nElems = nElems
    thisParam % nPlot = nPlot
    thisParam % dxPlot = 2.0_prec / REAL(nPlot, prec)
    thisParam % xScale = xScale
  END SUBROUTINE BuildParams_Elliptic1D

END MODULE Elliptic1DParams_Class

MODULE Elliptic1D_Class
  USE ISO_C_BINDING
  USE ModelPrecision
  USE ConstantsDictionary
  USE Elliptic1DParams_Class
  USE SELF_AMD_CABI
  IMPLICIT NONE

  TYPE Elliptic1D
    ! IterativeSolver (IterativeSolvers_Class.f90:55-60)
    INTEGER                 :: maxIters
    INTEGER                 :: nDOF
    REAL(prec)              :: tolerance
    REAL(prec), ALLOCATABLE :: resi(:)
    REAL(prec), ALLOCATABLE :: sol(:)
    ! Elliptic1D (Elliptic1D_Class.f90:63-72); x holds the node positions
    ! the reference reads from its SegmentMesh
    INTEGER                 :: nS, nElems
    REAL(prec), ALLOCATABLE :: x(:,:)
    REAL(prec), ALLOCATABLE :: cgsol(:,:)
    REAL(prec), ALLOCATABLE :: source(:,:)
    REAL(prec)              :: boundaryFlux(1:2)
    REAL(prec)              :: dirichletMask(1:2)
    TYPE(Elliptic1DParams)  :: params
    TYPE(C_PTR)             :: handle = C_NULL_PTR
  CONTAINS
    PROCEDURE :: Build => Build_Elliptic1D
    PROCEDURE :: Trash => Trash_Elliptic1D
    PROCEDURE :: SetMaxIters => SetMaxIters_Elliptic1D
    PROCEDURE :: GetMaxIters => GetMaxIters_Elliptic1D
    PROCEDURE :: SetNDOF => SetNDOF_Elliptic1D
    PROCEDURE :: GetNDOF => GetNDOF_Elliptic1D
    PROCEDURE :: SetTolerance => SetTolerance_Elliptic1D
    PROCEDURE :: GetTolerance => GetTolerance_Elliptic1D
    PROCEDURE :: CopyTo => CopyTo_Elliptic1D
    PROCEDURE :: CopyFrom => CopyFrom_Elliptic1D
    PROCEDURE :: UnMaskSolution => UnMaskSolution_Elliptic1D
    PROCEDURE :: Solve => Solve_Elliptic1D
    PROCEDURE :: WriteTecplot => WriteTecplot_Elliptic1D
    PROCEDURE :: WriteResidual => WriteResidual_Elliptic1D
  END TYPE Elliptic1D

CONTAINS

  ! Build (:105-177): parameters, the degree-polyDeg Gauss-Lobatto operator
  ! on nElems elements of [0, xScale] with stiffness weight 1, source 1 and
  ! both ends Dirichlet (dirichletMask = 0), solution and residual history
  ! zero.
  SUBROUTINE Build_Elliptic1D(this)
    CLASS(Elliptic1D), INTENT(inout) :: this
    INTEGER :: nS, nEl
    REAL(prec), ALLOCATABLE :: zero_(:), u(:)

    CALL this % params % Build()
    nS = this % params % polyDeg
    nEl = this % params % nElems
    this % nS = nS
    this % nElems = nEl
    this % maxIters = this % params % MaximumIterates
    this % nDOF = (nS + 1) * nEl
    this % tolerance = this % params % tolerance

    CALL self_check(self_linop_elliptic1d_create(INT(nS, C_INT), INT(nEl, C_INT), this % params % xScale, &
                                                 this % handle), 'Build')
    ALLOCATE(this % sol(1:this % nDOF), this % resi(0:this % maxIters))
    ALLOCATE(this % cgsol(0:nS, 1:nEl), this % source(0:nS, 1:nEl), this % x(0:nS, 1:nEl))
    this % sol = ZERO
    this % resi = ZERO
    this % cgsol = ZERO
    this % source = ONE                      ! FillSource :209-230
    this % boundaryFlux = ZERO
    this % dirichletMask = ZERO

    ALLOCATE(zero_(this % nDOF), u(this % nDOF))
    zero_ = ZERO
    CALL self_check(self_linop_elliptic1d_unmask(this % handle, zero_, u, this % x), 'Build')
    DEALLOCATE(zero_, u)
  END SUBROUTINE Build_Elliptic1D

  SUBROUTINE Trash_Elliptic1D(this)
    CLASS(Elliptic1D), INTENT(inout) :: this
    IF (C_ASSOCIATED(this % handle)) CALL self_check(self_linop_destroy(this % handle), 'Trash')
    this % handle = C_NULL_PTR
    IF (ALLOCATED(this % sol)) DEALLOCATE(this % sol)
    IF (ALLOCATED(this % resi)) DEALLOCATE(this % resi)
    IF (ALLOCATED(this % cgsol)) DEALLOCATE(this % cgsol)
    IF (ALLOCATED(this % source)) DEALLOCATE(this % source)
    IF (ALLOCATED(this % x)) DEALLOCATE(this % x)
  END SUBROUTINE Trash_Elliptic1D

  ! Accessors (IterativeSolvers_Class.f90:234-320)
  SUBROUTINE SetMaxIters_Elliptic1D(this, mI)
    CLASS(Elliptic1D), INTENT(inout) :: this
    INTEGER, INTENT(in) :: mI
    this % maxIters = mI
  END SUBROUTINE SetMaxIters_Elliptic1D

  FUNCTION GetMaxIters_Elliptic1D(this) RESULT(mI)
    CLASS(Elliptic1D) :: this
    INTEGER :: mI
    mI = this % maxIters
  END FUNCTION GetMaxIters_Elliptic1D

  SUBROUTINE SetNDOF_Elliptic1D(this, n)
    CLASS(Elliptic1D), INTENT(inout) :: this
    INTEGER, INTENT(in) :: n
    this % nDOF = n
  END SUBROUTINE SetNDOF_Elliptic1D

  FUNCTION GetNDOF_Elliptic1D(this) RESULT(n)
    CLASS(Elliptic1D) :: this
    INTEGER :: n
    n = this % nDOF
  END FUNCTION GetNDOF_Elliptic1D

  SUBROUTINE SetTolerance_Elliptic1D(this, tol)
    CLASS(Elliptic1D), INTENT(inout) :: this
    REAL(prec), INTENT(in) :: tol
    this % tolerance = tol
  END SUBROUTINE SetTolerance_Elliptic1D

  FUNCTION GetTolerance_Elliptic1D(this) RESULT(tol)
    CLASS(Elliptic1D) :: this
    REAL(prec) :: tol
    tol = this % tolerance
  END FUNCTION GetTolerance_Elliptic1D

  ! CopyTo / CopyFrom (:507-555): element-local (0:nS, nElems) <-> the flat
  ! vector k = iS + 1 + (nS+1)*(iEl-1), which is column-major order
  SUBROUTINE CopyTo_Elliptic1D(this, sol)
    CLASS(Elliptic1D), INTENT(in) :: this
    REAL(prec), INTENT(out) :: sol(1:this % nDOF)
    sol = RESHAPE(this % cgsol, [this % nDOF])
  END SUBROUTINE CopyTo_Elliptic1D

  SUBROUTINE CopyFrom_Elliptic1D(this, sol)
    CLASS(Elliptic1D), INTENT(inout) :: this
    REAL(prec), INTENT(in) :: sol(1:this % nDOF)
    this % cgsol = RESHAPE(sol, [this % nS + 1, this % nElems])
  END SUBROUTINE CopyFrom_Elliptic1D

  ! UnMaskSolution (:340-356): shared nodes take the left element's value,
  ! Dirichlet ends are zero
  SUBROUTINE UnMaskSolution_Elliptic1D(this)
    CLASS(Elliptic1D), INTENT(inout) :: this
    REAL(prec) :: s(this % nDOF), u(this % nDOF), xs(this % nDOF)
    CALL this % CopyTo(s)
    CALL self_check(self_linop_elliptic1d_unmask(this % handle, s, u, xs), 'UnMaskSolution')
    CALL this % CopyFrom(u)
  END SUBROUTINE UnMaskSolution_Elliptic1D

  ! Solve (ConjugateGradient, IterativeSolvers_Class.f90:328-405): CG on the
  ! device from the current cgsol to sqrt(|r|)/|r0| < tolerance in at most
  ! maxIters iterations; resi(0:maxIters) the residual history, zero after
  ! the last iterate.  ioerr as the reference leaves it: -2 when the loop
  ! ended with sqrt(|r|) <= tolerance (its ioerr = 0 sits after its EXIT),
  ! -1 otherwise, with the reference's message.
  SUBROUTINE Solve_Elliptic1D(this, ioerr)
    CLASS(Elliptic1D), INTENT(inout) :: this
    INTEGER, INTENT(out) :: ioerr
    INTEGER(C_INT) :: io, its
    CALL this % CopyTo(this % sol)
    IF (SIZE(this % resi) /= this % maxIters + 1) THEN
      DEALLOCATE(this % resi)
      ALLOCATE(this % resi(0:this % maxIters))
    ENDIF
    this % resi = ZERO
    CALL self_check(self_krylov_solve(this % handle, SELF_KRYLOV_CG, this % tolerance, INT(this % maxIters, C_INT), &
                                      0_C_INT, 0_C_INT, this % sol, this % resi, INT(this % maxIters + 1, C_INT), &
                                      its, io), 'Solve')
    ioerr = io
    CALL this % CopyFrom(this % sol)
    IF (ioerr == -1) THEN
      PRINT *, 'MODULE IterativeSolvers : ConjugateGradient failed to converge '
      PRINT *, 'Last L-2 residual : ', SQRT(this % resi(its))
    ENDIF
  END SUBROUTINE Solve_Elliptic1D

  ! WriteTecplot (:563-601): elliptic.curve, '#solution' then x and the
  ! unmasked solution, '#source' then x and the source, list-directed
  SUBROUTINE WriteTecplot_Elliptic1D(this)
    CLASS(Elliptic1D), INTENT(inout) :: this
    INTEGER :: iS, iEl, fUnit
    CALL this % UnMaskSolution()
    OPEN(NEWUNIT=fUnit, FILE='elliptic.curve', FORM='FORMATTED')
    WRITE(fUnit, *) '#solution'
    DO iEl = 1, this % nElems
      DO iS = 0, this % nS
        WRITE(fUnit, *) this % x(iS, iEl), this % cgsol(iS, iEl)
      ENDDO
    ENDDO
    WRITE(fUnit, *) '#source'
    DO iEl = 1, this % nElems
      DO iS = 0, this % nS
        WRITE(fUnit, *) this % x(iS, iEl), this % source(iS, iEl)
      ENDDO
    ENDDO
    CLOSE(fUnit)
  END SUBROUTINE WriteTecplot_Elliptic1D

  ! WriteResidual (IterativeSolvers_Class.f90:830-860): resi(0:maxIters) as
  ! '(I4,1x,E17.8)' under '#residual', to rFile or residual.curve
  SUBROUTINE WriteResidual_Elliptic1D(this, rFile)
    CLASS(Elliptic1D), INTENT(in) :: this
    CHARACTER(*), INTENT(in), OPTIONAL :: rFile
    INTEGER :: i, fUnit
    CHARACTER(100) :: localFile
    localFile = 'residual.curve'
    IF (PRESENT(rFile)) localFile = rFile
    OPEN(NEWUNIT=fUnit, FILE=TRIM(localFile), FORM='FORMATTED')
    WRITE(fUnit, *) '#residual'
    DO i = 0, this % GetMaxIters()
      WRITE(fUnit, '(I4,1x,E17.8)') i, this % resi(i)
    ENDDO
    CLOSE(UNIT=fUnit)
  END SUBROUTINE WriteResidual_Elliptic1D

END MODULE Elliptic1D_Class
