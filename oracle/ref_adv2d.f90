! ref_adv2d.f90 -- test infrastructure only (a checker, never shipped or timed).
!
! Drives the reference's own Advection type, compiled unmodified from
! src/highend/advection-diffusion2d/AdvectionClass.f90 and the modules it
! uses (recipe: oracle/ref.mk), on one case and writes what it computed, so
! that the C restatement (orc_adv2d.c) and the HIP path can be compared with
! the reference itself.
!
! Run in a directory holding
!   advection.params  the reference's namelist file (AdvectionParamsClass.f90:86-113)
!   ref_in.bin        stream file: int32 nSteps, float64 t0, then float64
!                     c(0:N,0:N,nTracers,nElems), u(0:N,0:N,nElems),
!                     v(0:N,0:N,nElems), relax and rFac shaped as c
! and it writes ref_out.bin (stream):
!   int32 nElems, nEdges
!   float64 dMatS(0:N,0:N), dMatP(0:N,0:N), qWeightS, qWeightP, lagWest,
!           lagEast, lagSouth, lagNorth (0:N each)
!   int32   edges(6,nEdges) = e1, s1, e2, s2, start, increment
!   float64 x, y, dxds, dxdp, dyds, dydp, J (0:N,0:N,nElems each),
!           nHat(2,0:N,4,nElems), nLen(0:N,4,nElems)
!   float64 tendency, dcdx, dcdy (shaped as c) after GlobalTimeDerivative(t0)
!   float64 solution (shaped as c) after nSteps ForwardStepRK3 calls from t0
!
! Build_Advection is called with forceNoPickup so no pickup file is read
! (AdvectionClass.f90:232-238); the mesh is the reference's default walled
! mesh with its inflow flags (BuildQuadMesh_Advection, :292-337).
PROGRAM ref_adv2d
 USE ModelPrecision
 USE AdvectionClass
 IMPLICIT NONE
 TYPE(Advection) :: adv
 INTEGER :: N, nEq, nE, nEd, iEl, iEdge, i, j, b, iStep, nSteps, uIn, uOut
 INTEGER, ALLOCATABLE :: edges(:,:)
 REAL(prec) :: t0, dt, nrm(1:2), length
 REAL(prec), ALLOCATABLE :: c(:,:,:,:), u(:,:,:), v(:,:,:), rl(:,:,:,:), rf(:,:,:,:)
 REAL(prec), ALLOCATABLE :: o1(:,:,:,:), o2(:,:,:,:), o3(:,:,:,:)
 REAL(prec), ALLOCATABLE :: dMatS(:,:), dMatP(:,:), wS(:), wP(:), lW(:), lE(:), lS(:), lN(:)
 REAL(prec), ALLOCATABLE :: g(:,:,:,:), nHat(:,:,:,:), nLen(:,:,:)

   CALL adv % Build( 1 )
   N   = adv % nS
   nEq = adv % nEq
   nE  = adv % mesh % nElems
   nEd = adv % mesh % nEdges
   dt  = adv % params % dt

   ALLOCATE( c(0:N,0:N,1:nEq,1:nE), rl(0:N,0:N,1:nEq,1:nE), rf(0:N,0:N,1:nEq,1:nE) )
   ALLOCATE( u(0:N,0:N,1:nE), v(0:N,0:N,1:nE) )
   ALLOCATE( o1(0:N,0:N,1:nEq,1:nE), o2(0:N,0:N,1:nEq,1:nE), o3(0:N,0:N,1:nEq,1:nE) )
   ALLOCATE( dMatS(0:N,0:N), dMatP(0:N,0:N), wS(0:N), wP(0:N), lW(0:N), lE(0:N), lS(0:N), lN(0:N) )
   ALLOCATE( edges(1:6,1:nEd), g(0:N,0:N,1:nE,1:7), nHat(1:2,0:N,1:4,1:nE), nLen(0:N,1:4,1:nE) )

   OPEN( NEWUNIT = uIn, FILE = 'ref_in.bin', ACCESS = 'STREAM', FORM = 'UNFORMATTED', STATUS = 'OLD' )
   READ( uIn ) nSteps
   READ( uIn ) t0
   READ( uIn ) c, u, v, rl, rf
   CLOSE( uIn )

   DO iEl = 1, nE
      CALL adv % SetSolution( iEl, c(:,:,:,iEl) )
      CALL adv % SetVelocity( iEl, u(:,:,iEl), v(:,:,iEl) )
      CALL adv % SetRelaxationField( iEl, rl(:,:,:,iEl) )
      CALL adv % SetRelaxationFactor( iEl, rf(:,:,:,iEl) )
   ENDDO

   CALL adv % dGStorage % GetDerivativeMatrix( dMatS, dMatP )
   CALL adv % dGStorage % GetQuadratureWeights( wS, wP )
   CALL adv % dGStorage % GetWesternInterpolants( lW )
   CALL adv % dGStorage % GetEasternInterpolants( lE )
   CALL adv % dGStorage % GetSouthernInterpolants( lS )
   CALL adv % dGStorage % GetNorthernInterpolants( lN )

   DO iEdge = 1, nEd
      CALL adv % mesh % GetEdgePrimaryElementID( iEdge, edges(1,iEdge) )
      CALL adv % mesh % GetEdgePrimaryElementSide( iEdge, edges(2,iEdge) )
      CALL adv % mesh % GetEdgeSecondaryElementID( iEdge, edges(3,iEdge) )
      CALL adv % mesh % GetEdgeSecondaryElementSide( iEdge, edges(4,iEdge) )
      CALL adv % mesh % GetEdgeStart( iEdge, edges(5,iEdge) )
      CALL adv % mesh % GetEdgeIncrement( iEdge, edges(6,iEdge) )
   ENDDO

   DO iEl = 1, nE
      DO j = 0, N
         DO i = 0, N
            CALL adv % mesh % GetPositionAtNode( iEl, g(i,j,iEl,1), g(i,j,iEl,2), i, j )
            CALL adv % mesh % GetCovariantMetricsAtNode( iEl, g(i,j,iEl,3), g(i,j,iEl,4), &
                                                         g(i,j,iEl,5), g(i,j,iEl,6), i, j )
            CALL adv % mesh % GetJacobianAtNode( iEl, g(i,j,iEl,7), i, j )
         ENDDO
      ENDDO
      DO b = 1, 4
         DO i = 0, N
            CALL adv % mesh % GetBoundaryNormalAtNode( iEl, nrm, length, i, b )
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

   CALL adv % GlobalTimeDerivative( t0 )
   DO iEl = 1, nE
      CALL adv % GetTendency( iEl, o1(:,:,:,iEl) )
      CALL adv % GetGradient( iEl, o2(:,:,:,iEl), o3(:,:,:,iEl) )
   ENDDO
   WRITE( uOut ) o1, o2, o3

   DO iStep = 0, nSteps - 1
      CALL adv % ForwardStepRK3( t0 + REAL(iStep,prec)*dt )
   ENDDO
   DO iEl = 1, nE
      CALL adv % GetSolution( iEl, o1(:,:,:,iEl) )
   ENDDO
   WRITE( uOut ) o1
   CLOSE( uOut )

   CALL adv % Trash( )

END PROGRAM ref_adv2d
