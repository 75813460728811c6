! ref_krylov.f90 -- test infrastructure only (a checker, never shipped or timed).
!
! Runs the reference's own Krylov loops, compiled unmodified from
! src/iterativesolve/IterativeSolvers_Class.f90 (recipe: oracle/ref.mk):
! Solve_ConjugateGradient, Solve_PCConjugateGradient, Solve_GMRES and
! Solve_PCGMRES, on a dense system, so that the C restatement
! (orc_krylov.c) and the HIP solver can be compared with the reference.
!
! The deferred procedures are this checker's own: MatrixAction is the dense
! product summed over columns in ascending order, Residual = b - A s,
! PreConditionInvert = s / diag(A), CopyTo hands over the initial guess and
! CopyFrom keeps the result (the same operator as orc_dense_matvec,
! orc_dense_residual and orc_dense_jacobi).
!
! Run in a directory holding ref_in.bin (stream): int32 n, method (1 CG,
! 2 PCG, 3 GMRES, 4 PCGMRES), maxIters, mInner; float64 tolerance, then
! float64 A(n,n), b(n), x0(n).  It writes ref_out.bin (stream): int32 ioerr,
! int32 resiLen, float64 resi(0:resiLen-1), float64 x(n).  resi is allocated
! with maxIters*mInner+1 entries for the GMRES forms (their history index
! runs over inner iterations) and maxIters+1 for the CG forms.
MODULE ref_krylov_ops
 USE ModelPrecision
 USE IterativeSolvers_Class
 IMPLICIT NONE
 INTEGER :: nD
 REAL(prec), ALLOCATABLE :: Am(:,:), bv(:), x0(:), xOut(:)

 TYPE, EXTENDS(ConjugateGradient) :: DenseCG
 CONTAINS
   PROCEDURE :: MatrixAction => MatVecCG
   PROCEDURE :: Residual => ResidCG
   PROCEDURE :: CopyTo => CopyToCG
   PROCEDURE :: CopyFrom => CopyFromCG
 END TYPE DenseCG

 TYPE, EXTENDS(PCConjugateGradient) :: DensePCG
 CONTAINS
   PROCEDURE :: MatrixAction => MatVecPCG
   PROCEDURE :: Residual => ResidPCG
   PROCEDURE :: CopyTo => CopyToPCG
   PROCEDURE :: CopyFrom => CopyFromPCG
   PROCEDURE :: PreConditionInvert => JacobiPCG
 END TYPE DensePCG

 TYPE, EXTENDS(GMRES) :: DenseGMRES
 CONTAINS
   PROCEDURE :: MatrixAction => MatVecGM
   PROCEDURE :: Residual => ResidGM
   PROCEDURE :: CopyTo => CopyToGM
   PROCEDURE :: CopyFrom => CopyFromGM
 END TYPE DenseGMRES

 TYPE, EXTENDS(PCGMRES) :: DensePCGMRES
 CONTAINS
   PROCEDURE :: MatrixAction => MatVecPGM
   PROCEDURE :: Residual => ResidPGM
   PROCEDURE :: CopyTo => CopyToPGM
   PROCEDURE :: CopyFrom => CopyFromPGM
   PROCEDURE :: PreConditionInvert => JacobiPGM
 END TYPE DensePCGMRES

CONTAINS

 FUNCTION MatVec( s ) RESULT( As )
   REAL(prec) :: s(1:nD), As(1:nD)
   INTEGER :: i, j
   DO i = 1, nD
      As(i) = 0.0_prec
      DO j = 1, nD
         As(i) = As(i) + Am(i,j)*s(j)
      ENDDO
   ENDDO
 END FUNCTION MatVec

 FUNCTION Resid( s ) RESULT( r )
   REAL(prec) :: s(1:nD), r(1:nD)
   INTEGER :: i
   r = MatVec( s )
   DO i = 1, nD
      r(i) = bv(i) - r(i)
   ENDDO
 END FUNCTION Resid

 FUNCTION Jacobi( s ) RESULT( w )
   REAL(prec) :: s(1:nD), w(1:nD)
   INTEGER :: i
   DO i = 1, nD
      w(i) = s(i)/Am(i,i)
   ENDDO
 END FUNCTION Jacobi

 FUNCTION MatVecCG( this, s ) RESULT( As )
   CLASS(DenseCG) :: this
   REAL(prec) :: s(1:this % nDOF), As(1:this % nDOF)
   As = MatVec( s )
 END FUNCTION MatVecCG
 FUNCTION ResidCG( this, s ) RESULT( r )
   CLASS(DenseCG) :: this
   REAL(prec) :: s(1:this % nDOF), r(1:this % nDOF)
   r = Resid( s )
 END FUNCTION ResidCG
 SUBROUTINE CopyToCG( this, sol )
   CLASS(DenseCG) :: this
   REAL(prec) :: sol(1:this % nDOF)
   sol = x0
 END SUBROUTINE CopyToCG
 SUBROUTINE CopyFromCG( this, sol )
   CLASS(DenseCG) :: this
   REAL(prec) :: sol(1:this % nDOF)
   xOut = sol
 END SUBROUTINE CopyFromCG

 FUNCTION MatVecPCG( this, s ) RESULT( As )
   CLASS(DensePCG) :: this
   REAL(prec) :: s(1:this % nDOF), As(1:this % nDOF)
   As = MatVec( s )
 END FUNCTION MatVecPCG
 FUNCTION ResidPCG( this, s ) RESULT( r )
   CLASS(DensePCG) :: this
   REAL(prec) :: s(1:this % nDOF), r(1:this % nDOF)
   r = Resid( s )
 END FUNCTION ResidPCG
 SUBROUTINE CopyToPCG( this, sol )
   CLASS(DensePCG) :: this
   REAL(prec) :: sol(1:this % nDOF)
   sol = x0
 END SUBROUTINE CopyToPCG
 SUBROUTINE CopyFromPCG( this, sol )
   CLASS(DensePCG) :: this
   REAL(prec) :: sol(1:this % nDOF)
   xOut = sol
 END SUBROUTINE CopyFromPCG
 FUNCTION JacobiPCG( this, s ) RESULT( w )
   CLASS(DensePCG) :: this
   REAL(prec) :: s(1:this % nDOF), w(1:this % nDOF)
   w = Jacobi( s )
 END FUNCTION JacobiPCG

 FUNCTION MatVecGM( this, s ) RESULT( As )
   CLASS(DenseGMRES) :: this
   REAL(prec) :: s(1:this % nDOF), As(1:this % nDOF)
   As = MatVec( s )
 END FUNCTION MatVecGM
 FUNCTION ResidGM( this, s ) RESULT( r )
   CLASS(DenseGMRES) :: this
   REAL(prec) :: s(1:this % nDOF), r(1:this % nDOF)
   r = Resid( s )
 END FUNCTION ResidGM
 SUBROUTINE CopyToGM( this, sol )
   CLASS(DenseGMRES) :: this
   REAL(prec) :: sol(1:this % nDOF)
   sol = x0
 END SUBROUTINE CopyToGM
 SUBROUTINE CopyFromGM( this, sol )
   CLASS(DenseGMRES) :: this
   REAL(prec) :: sol(1:this % nDOF)
   xOut = sol
 END SUBROUTINE CopyFromGM

 FUNCTION MatVecPGM( this, s ) RESULT( As )
   CLASS(DensePCGMRES) :: this
   REAL(prec) :: s(1:this % nDOF), As(1:this % nDOF)
   As = MatVec( s )
 END FUNCTION MatVecPGM
 FUNCTION ResidPGM( this, s ) RESULT( r )
   CLASS(DensePCGMRES) :: this
   REAL(prec) :: s(1:this % nDOF), r(1:this % nDOF)
   r = Resid( s )
 END FUNCTION ResidPGM
 SUBROUTINE CopyToPGM( this, sol )
   CLASS(DensePCGMRES) :: this
   REAL(prec) :: sol(1:this % nDOF)
   sol = x0
 END SUBROUTINE CopyToPGM
 SUBROUTINE CopyFromPGM( this, sol )
   CLASS(DensePCGMRES) :: this
   REAL(prec) :: sol(1:this % nDOF)
   xOut = sol
 END SUBROUTINE CopyFromPGM
 FUNCTION JacobiPGM( this, s ) RESULT( w )
   CLASS(DensePCGMRES) :: this
   REAL(prec) :: s(1:this % nDOF), w(1:this % nDOF)
   w = Jacobi( s )
 END FUNCTION JacobiPGM

 SUBROUTINE Setup( s, n, maxIters, solverTol, resiLen )
   CLASS(IterativeSolver), INTENT(inout) :: s
   INTEGER, INTENT(in)    :: n, maxIters, resiLen
   REAL(prec), INTENT(in) :: solverTol
   s % nDOF      = n
   s % maxIters  = maxIters
   s % tolerance = solverTol
   ALLOCATE( s % sol(1:n), s % resi(0:resiLen-1) )
   s % sol  = 0.0_prec
   s % resi = 0.0_prec
 END SUBROUTINE Setup

END MODULE ref_krylov_ops

PROGRAM ref_krylov
 USE ModelPrecision
 USE ref_krylov_ops
 IMPLICIT NONE
 TYPE(DenseCG)      :: cgS
 TYPE(DensePCG)     :: pcgS
 TYPE(DenseGMRES)   :: gmS
 TYPE(DensePCGMRES) :: pgmS
 INTEGER    :: method, maxIters, mInner, ioerr, resiLen, uIn, uOut
 REAL(prec) :: solverTol

   OPEN( NEWUNIT = uIn, FILE = 'ref_in.bin', ACCESS = 'STREAM', FORM = 'UNFORMATTED', STATUS = 'OLD' )
   READ( uIn ) nD, method, maxIters, mInner
   READ( uIn ) solverTol
   ALLOCATE( Am(1:nD,1:nD), bv(1:nD), x0(1:nD), xOut(1:nD) )
   READ( uIn ) Am, bv, x0
   CLOSE( uIn )
   xOut = 0.0_prec

   OPEN( NEWUNIT = uOut, FILE = 'ref_out.bin', ACCESS = 'STREAM', FORM = 'UNFORMATTED', STATUS = 'REPLACE' )
   SELECT CASE( method )
   CASE( 1 )
      resiLen = maxIters + 1
      CALL Setup( cgS, nD, maxIters, solverTol, resiLen )
      CALL cgS % Solve( ioerr )
      WRITE( uOut ) ioerr, resiLen
      WRITE( uOut ) cgS % resi
   CASE( 2 )
      resiLen = maxIters + 1
      CALL Setup( pcgS, nD, maxIters, solverTol, resiLen )
      pcgS % mInnerIters = mInner
      CALL pcgS % Solve( ioerr )
      WRITE( uOut ) ioerr, resiLen
      WRITE( uOut ) pcgS % resi
   CASE( 3 )
      resiLen = maxIters*mInner + 1
      CALL Setup( gmS, nD, maxIters, solverTol, resiLen )
      gmS % mInnerIters = mInner
      CALL gmS % Solve( ioerr )
      WRITE( uOut ) ioerr, resiLen
      WRITE( uOut ) gmS % resi
   CASE( 4 )
      resiLen = maxIters*mInner + 1
      CALL Setup( pgmS, nD, maxIters, solverTol, resiLen )
      pgmS % mInnerIters = mInner
      CALL pgmS % Solve( ioerr )
      WRITE( uOut ) ioerr, resiLen
      WRITE( uOut ) pgmS % resi
   END SELECT
   WRITE( uOut ) xOut
   CLOSE( uOut )

END PROGRAM ref_krylov
