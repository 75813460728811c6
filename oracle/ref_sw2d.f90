! ref_sw2d.f90 -- test infrastructure only (a checker, never shipped or timed).
!
! Drives the reference's conservative shallow-water type compiled unmodified
! from examples/shallowwater/planewave/localmods/ConservativeShallowWaterClass.f90
! and its SWParamsClass.f90 (recipe: oracle/ref.mk).  That class is the
! src/highend/shallowwater one with three local changes: the Rusanov wave
! speed max|(u_n +- sqrt(u_n^2 + 2 g H^3))/(2H)| (:1555-1558), a Source of
! Coriolis and g H grad(h) only (:1643-1656) and a bathymetry gradient
! grad(h^2/2)/h (:1754-1825).  The src class itself does not compile (an
! undeclared g in its RiemannSolver and an unterminated DO in WriteTecplot).
!
! Run in a directory holding
!   shallowwater.params  the reference's namelist file (SWParamsClass.f90)
!   ref_in.bin           stream file: int32 nSteps, float64 t0, then float64
!                        sol(0:N,0:N,3,nElems), h(0:N,0:N,nElems),
!                        f(0:N,0:N,nElems)
! and it writes ref_out.bin (stream):
!   int32 nElems, nEdges
!   float64 dMatS(0:N,0:N), dMatP(0:N,0:N), qWeightS, qWeightP, lagWest,
!           lagEast, lagSouth, lagNorth (0:N each)
!   int32   edges(6,nEdges) = e1, s1, e2, s2, start, increment
!   float64 x, y, dxds, dxdp, dyds, dydp, J (0:N,0:N,nElems each),
!           nHat(2,0:N,4,nElems), nLen(0:N,4,nElems)
!   float64 bathymetry(0:N,0:N,3,nElems) after CalculateBathymetryGradient
!   float64 tendency(0:N,0:N,3,nElems) and boundary flux (0:N,3,4,nElems)
!           after GlobalTimeDerivative(t0)
!   float64 solution(0:N,0:N,3,nElems) after nSteps ForwardStepRK3 calls
!
! Timing mode (nSteps < 0 in ref_in.bin): after the same set-up it runs
! -nSteps ForwardStepRK3 calls and writes only int32 -nSteps and float64 the
! elapsed wall seconds of those calls (SYSTEM_CLOCK), for the CPU baseline.
!
! Build is called with forceNoPickup (no pickup file is read); the mesh is
! the reference's default walled mesh (NO_NORMAL_FLOW sides) scaled by
! xScale, yScale.  The bathymetry traces are computed as ReadPickup does
! (:2079) before the gradient, as Build does after it (:207-211).
PROGRAM ref_sw2d
 USE ModelPrecision
 USE ConservativeShallowWaterClass
 IMPLICIT NONE
 TYPE(ShallowWater) :: sw
 INTEGER :: N, nE, nEd, iEl, iEdge, i, j, b, iStep, nSteps, uIn, uOut
 INTEGER(8) :: c0, c1, cRate
 INTEGER, ALLOCATABLE :: edges(:,:)
 REAL(prec) :: t0, dt, nrm(1:2), length
 REAL(prec), ALLOCATABLE :: s(:,:,:,:), h(:,:,:), f(:,:,:), bath(:,:,:,:), o1(:,:,:,:), bf(:,:,:,:)
 REAL(prec), ALLOCATABLE :: dMatS(:,:), dMatP(:,:), wS(:), wP(:), lW(:), lE(:), lS(:), lN(:)
 REAL(prec), ALLOCATABLE :: g(:,:,:,:), nHat(:,:,:,:), nLen(:,:,:)

   CALL sw % Build( 1 )
   N   = sw % nS
   nE  = sw % mesh % nElems
   nEd = sw % mesh % nEdges
   dt  = sw % params % dt

   ALLOCATE( s(0:N,0:N,1:3,1:nE), h(0:N,0:N,1:nE), f(0:N,0:N,1:nE), bath(0:N,0:N,1:3,1:nE) )
   ALLOCATE( o1(0:N,0:N,1:3,1:nE), bf(0:N,1:3,1:4,1:nE) )
   ALLOCATE( dMatS(0:N,0:N), dMatP(0:N,0:N), wS(0:N), wP(0:N), lW(0:N), lE(0:N), lS(0:N), lN(0:N) )
   ALLOCATE( edges(1:6,1:nEd), g(0:N,0:N,1:nE,1:7), nHat(1:2,0:N,1:4,1:nE), nLen(0:N,1:4,1:nE) )

   OPEN( NEWUNIT = uIn, FILE = 'ref_in.bin', ACCESS = 'STREAM', FORM = 'UNFORMATTED', STATUS = 'OLD' )
   READ( uIn ) nSteps
   READ( uIn ) t0
   READ( uIn ) s, h, f
   CLOSE( uIn )

   bath = 0.0_prec
   bath(:,:,1,:) = h
   DO iEl = 1, nE
      CALL sw % SetSolution( iEl, s(:,:,:,iEl) )
      CALL sw % SetBathymetry( iEl, bath(:,:,:,iEl) )
      CALL sw % SetPlanetaryVorticity( iEl, f(:,:,iEl) )
   ENDDO
   DO iEl = 1, nE
      CALL sw % CalculateBathymetryAtBoundaries( iEl )
   ENDDO
   DO iEl = 1, nE
      CALL sw % CalculateBathymetryGradient( iEl )
   ENDDO
   DO iEl = 1, nE
      CALL sw % GetBathymetry( iEl, bath(:,:,:,iEl) )
   ENDDO

   IF( nSteps < 0 )THEN
      CALL SYSTEM_CLOCK( c0, cRate )
      DO iStep = 0, -nSteps - 1
         CALL sw % ForwardStepRK3( t0 + REAL(iStep,prec)*dt )
      ENDDO
      CALL SYSTEM_CLOCK( c1 )
      OPEN( NEWUNIT = uOut, FILE = 'ref_out.bin', ACCESS = 'STREAM', FORM = 'UNFORMATTED', STATUS = 'REPLACE' )
      WRITE( uOut ) -nSteps
      WRITE( uOut ) REAL(c1 - c0, prec)/REAL(cRate, prec)
      CLOSE( uOut )
      CALL sw % Trash( )
      STOP
   ENDIF

   CALL sw % dGStorage % GetDerivativeMatrix( dMatS, dMatP )
   CALL sw % dGStorage % GetQuadratureWeights( wS, wP )
   CALL sw % dGStorage % GetWesternInterpolants( lW )
   CALL sw % dGStorage % GetEasternInterpolants( lE )
   CALL sw % dGStorage % GetSouthernInterpolants( lS )
   CALL sw % dGStorage % GetNorthernInterpolants( lN )

   DO iEdge = 1, nEd
      CALL sw % mesh % GetEdgePrimaryElementID( iEdge, edges(1,iEdge) )
      CALL sw % mesh % GetEdgePrimaryElementSide( iEdge, edges(2,iEdge) )
      CALL sw % mesh % GetEdgeSecondaryElementID( iEdge, edges(3,iEdge) )
      CALL sw % mesh % GetEdgeSecondaryElementSide( iEdge, edges(4,iEdge) )
      CALL sw % mesh % GetEdgeStart( iEdge, edges(5,iEdge) )
      CALL sw % mesh % GetEdgeIncrement( iEdge, edges(6,iEdge) )
   ENDDO

   DO iEl = 1, nE
      DO j = 0, N
         DO i = 0, N
            CALL sw % mesh % GetPositionAtNode( iEl, g(i,j,iEl,1), g(i,j,iEl,2), i, j )
            CALL sw % mesh % GetCovariantMetricsAtNode( iEl, g(i,j,iEl,3), g(i,j,iEl,4), &
                                                        g(i,j,iEl,5), g(i,j,iEl,6), i, j )
            CALL sw % mesh % GetJacobianAtNode( iEl, g(i,j,iEl,7), i, j )
         ENDDO
      ENDDO
      DO b = 1, 4
         DO i = 0, N
            CALL sw % mesh % GetBoundaryNormalAtNode( iEl, nrm, length, i, b )
            nHat(1:2,i,b,iEl) = nrm
            nLen(i,b,iEl) = length
         ENDDO
      ENDDO
   ENDDO

   OPEN( NEWUNIT = uOut, FILE = 'ref_out.bin', ACCESS = 'STREAM', FORM = 'UNFORMATTED', STATUS = 'REPLACE' )
   WRITE( uOut ) nE, nEd
   WRITE( uOut ) dMatS, dMatP, wS, wP, lW, lE, lS, lN
   WRITE( uOut ) edges
   DO b = 1, 7
      WRITE( uOut ) g(:,:,:,b)
   ENDDO
   WRITE( uOut ) nHat, nLen
   WRITE( uOut ) bath

   CALL sw % GlobalTimeDerivative( t0 )
   DO iEl = 1, nE
      CALL sw % GetTendency( iEl, o1(:,:,:,iEl) )
      DO b = 1, 4
         CALL sw % GetBoundaryFluxAtBoundary( iEl, b, bf(:,:,b,iEl) )
      ENDDO
   ENDDO
   WRITE( uOut ) o1, bf

   DO iStep = 0, nSteps - 1
      CALL sw % ForwardStepRK3( t0 + REAL(iStep,prec)*dt )
   ENDDO
   DO iEl = 1, nE
      CALL sw % GetSolution( iEl, o1(:,:,:,iEl) )
   ENDDO
   WRITE( uOut ) o1
   CLOSE( uOut )

   CALL sw % Trash( )

END PROGRAM ref_sw2d
