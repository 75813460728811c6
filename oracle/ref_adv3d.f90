! ref_adv3d.f90 -- test infrastructure only (a checker, never shipped or timed).
!
! Drives the reference's own 3-D Advection type, compiled unmodified from
! src/highend/advection3d/Advection3DClass.f90 and the modules it uses
! (recipe: oracle/ref.mk), on one case and writes what it computed, so that
! the C restatement (orc_adv3d.c) and the HIP path can be compared with the
! reference itself.
!
! Run in a directory holding
!   advection.params  the reference's namelist file (Advection3DParamsClass.f90:103-132)
!   ref_in.bin        stream file: int32 nSteps, float64 t0, then float64
!                     c(0:N,0:N,0:N,nTracers,nElems), u, v, w (0:N,0:N,0:N,nElems),
!                     relax and rFac shaped as c
! and it writes ref_out.bin (stream):
!   int32   nElems, nFaces
!   float64 dMatS, dMatP, dMatQ (0:N,0:N), qWeightS, qWeightP, qWeightQ,
!           lagWest, lagEast, lagSouth, lagNorth, lagBottom, lagTop (0:N each),
!           filterMat(nDOF,nDOF) of the roll-off filter
!   int32   faces(9,nFaces) = e1, s1, e2, s2, iStart, jStart, iInc, jInc, swap
!   float64 x, y, z, J (0:N,0:N,0:N,nElems each), cov(9,0:N,0:N,0:N,nElems)
!           (dxds dxdp dxdq dyds dydp dydq dzds dzdp dzdq),
!           nHat(3,0:N,0:N,6,nElems), nLen(0:N,0:N,6,nElems), bpos(3,0:N,0:N,6,nElems)
!   after GlobalTimeDerivative(t0, 1) (which applies the adaptive filter first):
!   float64 tendency, solution (shaped as c), E1, E2 (nElems,nTracers),
!           lim(nTracers,nElems,3)
!   after nSteps more ForwardStepRK3 calls from t0:
!   float64 solution, E1, E2, lim
!
! Build_Advection is called with forceNoPickup so no pickup file is read
! (Advection3DClass.f90:241-248); the mesh is the reference's default mesh
! (BuildHexMesh_Advection, :318-340).
PROGRAM ref_adv3d
 USE ModelPrecision
 USE Advection3DClass
 IMPLICIT NONE
 TYPE(Advection) :: adv
 INTEGER :: N, nEq, nE, nF, iEl, iF, i, j, k, b, iStep, nSteps, uIn, uOut
 INTEGER, ALLOCATABLE :: faces(:,:)
 REAL(prec) :: t0, dt, nrm(1:3), length
 REAL(prec), ALLOCATABLE :: c(:,:,:,:,:), u(:,:,:,:), v(:,:,:,:), w(:,:,:,:)
 REAL(prec), ALLOCATABLE :: rl(:,:,:,:,:), rf(:,:,:,:,:), o1(:,:,:,:,:)
 REAL(prec), ALLOCATABLE :: dMS(:,:), dMP(:,:), dMQ(:,:), wS(:), wP(:), wQ(:), l1(:)
 REAL(prec), ALLOCATABLE :: g(:,:,:,:,:), cov(:,:,:,:,:), nHat(:,:,:,:,:), nLen(:,:,:,:), bpos(:,:,:,:,:)

   CALL adv % Build( 1 )
   N   = adv % nS
   nEq = adv % nEq
   nE  = adv % mesh % nElems
   nF  = adv % mesh % nFaces
   dt  = adv % params % dt

   ALLOCATE( c(0:N,0:N,0:N,1:nEq,1:nE), rl(0:N,0:N,0:N,1:nEq,1:nE), rf(0:N,0:N,0:N,1:nEq,1:nE) )
   ALLOCATE( o1(0:N,0:N,0:N,1:nEq,1:nE) )
   ALLOCATE( u(0:N,0:N,0:N,1:nE), v(0:N,0:N,0:N,1:nE), w(0:N,0:N,0:N,1:nE) )
   ALLOCATE( dMS(0:N,0:N), dMP(0:N,0:N), dMQ(0:N,0:N), wS(0:N), wP(0:N), wQ(0:N), l1(0:N) )
   ALLOCATE( faces(1:9,1:nF), g(0:N,0:N,0:N,1:nE,1:4), cov(1:9,0:N,0:N,0:N,1:nE) )
   ALLOCATE( nHat(1:3,0:N,0:N,1:6,1:nE), nLen(0:N,0:N,1:6,1:nE), bpos(1:3,0:N,0:N,1:6,1:nE) )

   OPEN( NEWUNIT = uIn, FILE = 'ref_in.bin', ACCESS = 'STREAM', FORM = 'UNFORMATTED', STATUS = 'OLD' )
   READ( uIn ) nSteps
   READ( uIn ) t0
   READ( uIn ) c, u, v, w, rl, rf
   CLOSE( uIn )

   DO iEl = 1, nE
      CALL adv % SetSolution( iEl, c(:,:,:,:,iEl) )
      CALL adv % SetVelocity( iEl, u(:,:,:,iEl), v(:,:,:,iEl), w(:,:,:,iEl) )
      CALL adv % SetRelaxationField( iEl, rl(:,:,:,:,iEl) )
      CALL adv % SetRelaxationFactor( iEl, rf(:,:,:,:,iEl) )
   ENDDO

   OPEN( NEWUNIT = uOut, FILE = 'ref_out.bin', ACCESS = 'STREAM', FORM = 'UNFORMATTED', STATUS = 'REPLACE' )
   WRITE( uOut ) nE, nF

   CALL adv % dGStorage % GetDerivativeMatrix( dMS, dMP, dMQ )
   CALL adv % dGStorage % GetQuadratureWeights( wS, wP, wQ )
   WRITE( uOut ) dMS, dMP, dMQ, wS, wP, wQ
   CALL adv % dGStorage % GetWesternInterpolants( l1 )
   WRITE( uOut ) l1
   CALL adv % dGStorage % GetEasternInterpolants( l1 )
   WRITE( uOut ) l1
   CALL adv % dGStorage % GetSouthernInterpolants( l1 )
   WRITE( uOut ) l1
   CALL adv % dGStorage % GetNorthernInterpolants( l1 )
   WRITE( uOut ) l1
   CALL adv % dGStorage % GetBottomInterpolants( l1 )
   WRITE( uOut ) l1
   CALL adv % dGStorage % GetTopInterpolants( l1 )
   WRITE( uOut ) l1
   WRITE( uOut ) adv % modalFilter % filterMat

   DO iF = 1, nF
      CALL adv % mesh % GetFacePrimaryElementID( iF, faces(1,iF) )
      CALL adv % mesh % GetFacePrimaryElementSide( iF, faces(2,iF) )
      CALL adv % mesh % GetFaceSecondaryElementID( iF, faces(3,iF) )
      CALL adv % mesh % GetFaceSecondaryElementSide( iF, faces(4,iF) )
      CALL adv % mesh % GetFaceStart( iF, faces(5,iF), faces(6,iF) )
      CALL adv % mesh % GetFaceIncrement( iF, faces(7,iF), faces(8,iF) )
      CALL adv % mesh % GetSwapDimensions( iF, faces(9,iF) )
   ENDDO
   WRITE( uOut ) faces

   DO iEl = 1, nE
      DO k = 0, N
         DO j = 0, N
            DO i = 0, N
               CALL adv % mesh % GetPositionAtNode( iEl, g(i,j,k,iEl,1), g(i,j,k,iEl,2), g(i,j,k,iEl,3), i, j, k )
               CALL adv % mesh % GetJacobianAtNode( iEl, g(i,j,k,iEl,4), i, j, k )
               CALL adv % mesh % GetCovariantMetricsAtNode( iEl, cov(1,i,j,k,iEl), cov(2,i,j,k,iEl), &
                                                            cov(3,i,j,k,iEl), cov(4,i,j,k,iEl), &
                                                            cov(5,i,j,k,iEl), cov(6,i,j,k,iEl), &
                                                            cov(7,i,j,k,iEl), cov(8,i,j,k,iEl), &
                                                            cov(9,i,j,k,iEl), i, j, k )
            ENDDO
         ENDDO
      ENDDO
      DO b = 1, 6
         DO j = 0, N
            DO i = 0, N
               CALL adv % mesh % GetBoundaryNormalAtNode( iEl, nrm, length, i, j, b )
               nHat(1:3,i,j,b,iEl) = nrm
               nLen(i,j,b,iEl) = length
               CALL adv % mesh % GetBoundaryLocationAtNode( iEl, bpos(1,i,j,b,iEl), bpos(2,i,j,b,iEl), &
                                                            bpos(3,i,j,b,iEl), i, j, b )
            ENDDO
         ENDDO
      ENDDO
   ENDDO
   DO b = 1, 4
      WRITE( uOut ) g(:,:,:,:,b)
   ENDDO
   WRITE( uOut ) cov, nHat, nLen, bpos

   CALL adv % GlobalTimeDerivative( t0, 1 )
   DO iEl = 1, nE
      CALL adv % GetTendency( iEl, o1(:,:,:,:,iEl) )
   ENDDO
   WRITE( uOut ) o1
   DO iEl = 1, nE
      CALL adv % GetSolution( iEl, o1(:,:,:,:,iEl) )
   ENDDO
   WRITE( uOut ) o1, adv % E1, adv % E2, adv % lim

   DO iStep = 0, nSteps - 1
      CALL adv % ForwardStepRK3( t0 + REAL(iStep,prec)*dt )
   ENDDO
   DO iEl = 1, nE
      CALL adv % GetSolution( iEl, o1(:,:,:,:,iEl) )
   ENDDO
   WRITE( uOut ) o1, adv % E1, adv % E2, adv % lim
   CLOSE( uOut )

END PROGRAM ref_adv3d
