! ref_cgsem2d.f90 -- test infrastructure only (a checker, never shipped or timed).
!
! Drives the reference's own CG spectral-element elliptic solver, compiled
! unmodified from src/cgsem/CGsemElliptic_2D_Class.f90 and the modules it
! uses (recipe: oracle/ref.mk), on one case and writes what it computed, so
! that the C restatement (orc_cgsem2d.c) and the HIP path can be compared
! with the reference itself.
!
! Run in a directory holding
!   runtime.params  the reference's namelist file (CGsemParams_Class.f90:96-120);
!                   pcIterates = 0 makes the inner-CG preconditioner the
!                   identity (CGsemPreconditioner_2D_Class.f90:740-742)
!   ref_in.bin      stream file: float64 dirValue (the Dirichlet boundary
!                   value), then float64 fluxCoeff, source, solution, u, v
!                   (0:N,0:N,nElems each)
! and it writes ref_out.bin (stream):
!   int32   nElems, nEdges, nNodes
!   float64 dMatS(0:N,0:N), qWeightS(0:N)
!   int32   edges(6,nEdges) = e1, s1, e2, s2, start, increment; nodeType(nNodes)
!   float64 x, y, dxds, dxdp, dyds, dydp, J (0:N,0:N,nElems each)
!   float64 Mask(v), UnMask(Mask(v)), GlobalSum(v)
!   float64 Au and u after MatrixAction(u) (which masks u in place)
!   float64 DotProduct(u, v) (u, v fresh copies of the inputs)
!   float64 r = Residual and the solution after it, with fluxCoeff, source
!           and solution set from the inputs
!   int32   ioerr of Solve from the input solution; float64 resi(0:maxIters),
!           the solution after Solve
!
! The mesh is the reference's default mesh with Dirichlet boundaries
! (BuildQuadMesh_CGsemElliptic_2D, :160-196).
MODULE ref_cgsem2d_bc
 USE ModelPrecision
 IMPLICIT NONE
 REAL(prec) :: dirValue = 0.0_prec
CONTAINS
 FUNCTION dirF( x, y ) RESULT( f )
   REAL(prec) :: x, y, f
   f = dirValue
 END FUNCTION dirF
END MODULE ref_cgsem2d_bc

PROGRAM ref_cgsem2d
 USE ModelPrecision
 USE CGsemParams_Class
 USE CGsemElliptic_2D_Class
 USE ref_cgsem2d_bc
 IMPLICIT NONE
 TYPE(CGsemParams)      :: params
 TYPE(CGsemElliptic_2D) :: ell
 INTEGER :: N, nE, nEd, nNo, iEl, iEdge, iNode, i, j, b, ioerr, uIn, uOut
 INTEGER, ALLOCATABLE :: edges(:,:), nodeType(:)
 REAL(prec) :: dotUV
 REAL(prec), ALLOCATABLE :: k(:,:,:), s(:,:,:), sol0(:,:,:), u(:,:,:), v(:,:,:), w1(:,:,:), w2(:,:,:)
 REAL(prec), ALLOCATABLE :: dMS(:,:), dMP(:,:), wS(:), wP(:), g(:,:,:,:)

   CALL params % Build( )
   CALL ell % Build( params )
   N   = ell % nS
   nE  = ell % mesh % nElems
   nEd = ell % mesh % nEdges
   nNo = ell % mesh % nNodes

   ALLOCATE( k(0:N,0:N,1:nE), s(0:N,0:N,1:nE), sol0(0:N,0:N,1:nE), u(0:N,0:N,1:nE), v(0:N,0:N,1:nE) )
   ALLOCATE( w1(0:N,0:N,1:nE), w2(0:N,0:N,1:nE), dMS(0:N,0:N), dMP(0:N,0:N), wS(0:N), wP(0:N) )
   ALLOCATE( edges(1:6,1:nEd), nodeType(1:nNo), g(0:N,0:N,1:nE,1:7) )

   OPEN( NEWUNIT = uIn, FILE = 'ref_in.bin', ACCESS = 'STREAM', FORM = 'UNFORMATTED', STATUS = 'OLD' )
   READ( uIn ) dirValue
   READ( uIn ) k, s, sol0, u, v
   CLOSE( uIn )

   OPEN( NEWUNIT = uOut, FILE = 'ref_out.bin', ACCESS = 'STREAM', FORM = 'UNFORMATTED', STATUS = 'REPLACE' )
   WRITE( uOut ) nE, nEd, nNo
   CALL ell % SpectralOps % GetDerivativeMatrix( dMS, dMP )
   CALL ell % SpectralOps % GetQuadratureWeights( wS, wP )
   WRITE( uOut ) dMS, wS

   DO iEdge = 1, nEd
      CALL ell % mesh % GetEdgePrimaryElementID( iEdge, edges(1,iEdge) )
      CALL ell % mesh % GetEdgePrimaryElementSide( iEdge, edges(2,iEdge) )
      CALL ell % mesh % GetEdgeSecondaryElementID( iEdge, edges(3,iEdge) )
      CALL ell % mesh % GetEdgeSecondaryElementSide( iEdge, edges(4,iEdge) )
      CALL ell % mesh % GetEdgeStart( iEdge, edges(5,iEdge) )
      CALL ell % mesh % GetEdgeIncrement( iEdge, edges(6,iEdge) )
   ENDDO
   DO iNode = 1, nNo
      CALL ell % mesh % GetNodeType( iNode, nodeType(iNode) )
   ENDDO
   WRITE( uOut ) edges, nodeType

   DO iEl = 1, nE
      DO j = 0, N
         DO i = 0, N
            CALL ell % mesh % GetPositionAtNode( iEl, g(i,j,iEl,1), g(i,j,iEl,2), i, j )
            CALL ell % mesh % GetCovariantMetricsAtNode( iEl, g(i,j,iEl,3), g(i,j,iEl,4), &
                                                        g(i,j,iEl,5), g(i,j,iEl,6), i, j )
            CALL ell % mesh % GetJacobianAtNode( iEl, g(i,j,iEl,7), i, j )
         ENDDO
      ENDDO
   ENDDO
   DO b = 1, 7
      WRITE( uOut ) g(:,:,:,b)
   ENDDO

   ell % fluxCoeff = k
   ell % source    = s

   w1 = v
   CALL ell % Mask( w1 )
   WRITE( uOut ) w1
   CALL ell % UnMask( w1 )
   WRITE( uOut ) w1
   w1 = v
   CALL ell % GlobalSum( w1 )
   WRITE( uOut ) w1

   w1 = u
   CALL ell % MatrixAction( w1, w2 )
   WRITE( uOut ) w2, w1

   w1 = u
   w2 = v
   dotUV = ell % DotProduct( w1, w2 )
   WRITE( uOut ) dotUV

   ell % solution = sol0
   CALL ell % Residual( dirF, w1 )
   WRITE( uOut ) w1, ell % solution

   ell % solution = sol0
   CALL ell % Solve( dirF, ioerr )
   WRITE( uOut ) ioerr
   WRITE( uOut ) ell % resi, ell % solution
   CLOSE( uOut )

END PROGRAM ref_cgsem2d
