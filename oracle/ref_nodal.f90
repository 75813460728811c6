! ref_nodal.f90 -- test infrastructure only (a checker, never shipped or timed).
!
! Runs the reference's own quadrature, nodal storage and interpolation
! routines, compiled unmodified (src/interp/Legendre.f90,
! Lagrange_2D_Class.f90, Lagrange_3D_Class.f90, src/nodal/
! NodalStorage_2D_Class.f90, NodalStorage_3D_Class.f90; recipe oracle/ref.mk)
! on one polynomial degree and writes what they computed.
!
! Run in a directory holding ref_in.bin (stream): int32 N, quadrature
! (1 Gauss, -1 Gauss-Lobatto), form (2000 DG, 2001 CG), nNew; float64
! f2(0:N,0:N), f3(0:N,0:N,0:N).  It writes ref_out.bin (stream), float64:
!   nodes(0:N), weights(0:N)                       GenerateLegendreQuadrature
!   2-D storage: dMatS, dMatP (0:N,0:N), qWeightS, lagWest, lagEast,
!                lagSouth, lagNorth (0:N)
!   ApplyDerivativeMatrix_2D(f2): dFds, dFdp;  CalculateAtBoundaries_2D(f2):
!                fWest, fEast, fSouth, fNorth
!   3-D storage: dMatS (0:N,0:N);  ApplyDerivativeMatrix_3D(f3): dFds,
!                dFdp, dFdq;  CalculateAtBoundaries_3D(f3): fWest, fEast,
!                fSouth, fNorth, fBottom, fTop (0:N,0:N each)
!   CalculateInterpolationMatrix_2D to nNew+1 uniform points: Ts (0:nNew,0:N);
!   CoarseToFine_2D(f2): fNew(0:nNew,0:nNew); CoarseToFine_3D(f3):
!                fNew(0:nNew,0:nNew,0:nNew)
PROGRAM ref_nodal
 USE ModelPrecision
 USE CommonRoutines
 USE Legendre
 USE NodalStorage_2D_Class
 USE NodalStorage_3D_Class
 IMPLICIT NONE
 TYPE(NodalStorage_2D) :: ns2
 TYPE(NodalStorage_3D) :: ns3
 INTEGER :: N, quad, form, nNew, uIn, uOut
 REAL(prec), ALLOCATABLE :: x(:), w(:), f2(:,:), f3(:,:,:), a2(:,:), b2(:,:), a3(:,:,:), b3(:,:,:), c3(:,:,:)
 REAL(prec), ALLOCATABLE :: dS(:,:), dQ(:,:), dR(:,:), wS(:), wQ(:), wR(:), l1(:), l2(:), l3(:), l4(:)
 REAL(prec), ALLOCATABLE :: g1(:,:), g2(:,:), g3(:,:), g4(:,:), g5(:,:), g6(:,:)
 REAL(prec), ALLOCATABLE :: sNew(:), Ts(:,:), Tq(:,:), Tr(:,:), fN2(:,:), fN3(:,:,:)

   OPEN( NEWUNIT = uIn, FILE = 'ref_in.bin', ACCESS = 'STREAM', FORM = 'UNFORMATTED', STATUS = 'OLD' )
   READ( uIn ) N, quad, form, nNew
   ALLOCATE( x(0:N), w(0:N), f2(0:N,0:N), f3(0:N,0:N,0:N), a2(0:N,0:N), b2(0:N,0:N) )
   ALLOCATE( a3(0:N,0:N,0:N), b3(0:N,0:N,0:N), c3(0:N,0:N,0:N) )
   ALLOCATE( dS(0:N,0:N), dQ(0:N,0:N), dR(0:N,0:N), wS(0:N), wQ(0:N), wR(0:N) )
   ALLOCATE( l1(0:N), l2(0:N), l3(0:N), l4(0:N) )
   ALLOCATE( g1(0:N,0:N), g2(0:N,0:N), g3(0:N,0:N), g4(0:N,0:N), g5(0:N,0:N), g6(0:N,0:N) )
   ALLOCATE( sNew(0:nNew), Ts(0:nNew,0:N), Tq(0:nNew,0:N), Tr(0:nNew,0:N), fN2(0:nNew,0:nNew) )
   ALLOCATE( fN3(0:nNew,0:nNew,0:nNew) )
   READ( uIn ) f2, f3
   CLOSE( uIn )

   OPEN( NEWUNIT = uOut, FILE = 'ref_out.bin', ACCESS = 'STREAM', FORM = 'UNFORMATTED', STATUS = 'REPLACE' )
   CALL GenerateLegendreQuadrature( N, x, w, quad )
   WRITE( uOut ) x, w

   CALL ns2 % Build( N, N, quad, form )
   CALL ns2 % GetDerivativeMatrix( dS, dQ )
   CALL ns2 % GetQuadratureWeights( wS, wQ )
   WRITE( uOut ) dS, dQ, wS
   CALL ns2 % GetWesternInterpolants( l1 )
   CALL ns2 % GetEasternInterpolants( l2 )
   CALL ns2 % GetSouthernInterpolants( l3 )
   CALL ns2 % GetNorthernInterpolants( l4 )
   WRITE( uOut ) l1, l2, l3, l4
   CALL ns2 % ApplyDerivativeMatrix( f2, a2, b2 )
   WRITE( uOut ) a2, b2
   CALL ns2 % CalculateAtBoundaries( f2, l1, l2, l3, l4 )
   WRITE( uOut ) l1, l2, l3, l4

   CALL ns3 % Build( N, N, N, quad, form )
   CALL ns3 % GetDerivativeMatrix( dS, dQ, dR )
   WRITE( uOut ) dS
   CALL ns3 % ApplyDerivativeMatrix( f3, a3, b3, c3 )
   WRITE( uOut ) a3, b3, c3
   CALL ns3 % CalculateAtBoundaries( f3, g1, g2, g3, g4, g5, g6 )
   WRITE( uOut ) g1, g2, g3, g4, g5, g6

   sNew = UniformPoints( -1.0_prec, 1.0_prec, nNew )
   CALL ns2 % interp % CalculateInterpolationMatrix( nNew, nNew, sNew, sNew, Ts, Tq )
   WRITE( uOut ) Ts
   CALL ns2 % interp % CoarseToFine( f2, Ts, Tq, nNew, nNew, fN2 )
   WRITE( uOut ) fN2
   CALL ns3 % interp % CalculateInterpolationMatrix( nNew, nNew, nNew, sNew, sNew, sNew, Ts, Tq, Tr )
   CALL ns3 % interp % CoarseToFine( f3, Ts, Tq, Tr, nNew, nNew, nNew, fN3 )
   WRITE( uOut ) fN3
   CLOSE( uOut )

END PROGRAM ref_nodal
