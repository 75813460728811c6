! ref_pickup.f90 -- test infrastructure only (a checker, never shipped or timed).
!
! Runs the reference's own WritePickup_Advection and ReadPickup_Advection
! (src/highend/advection-diffusion2d/AdvectionClass.f90:2040-2171, compiled
! unmodified; recipe oracle/ref.mk) so that the pickup reader and writer of
! self_amd can be compared with the reference's files byte for byte.
!
! Run in a directory holding advection.params (the reference's namelist)
! and ref_in.bin (stream): int32 mode, iter, then for mode 1 float64
! c(0:N,0:N,nTracers,nElems), u(0:N,0:N,nElems), v(0:N,0:N,nElems),
! relax and rFac shaped as c.
!   mode 1: Build (no pickup read), set the fields, WritePickup(iter), which
!           writes Advection.<iter, 10 digits>.pickup here;
!   mode 2: Build (no pickup read), ReadPickup(iter), then write ref_out.bin
!           (stream): c, u, v, relax, rFac as read back through the Get
!           routines.
PROGRAM ref_pickup
 USE ModelPrecision
 USE AdvectionClass
 IMPLICIT NONE
 TYPE(Advection) :: adv
 INTEGER :: mode, iter, N, nEq, nE, iEl, uIn, uOut
 REAL(prec), ALLOCATABLE :: c(:,:,:,:), u(:,:,:), v(:,:,:), rl(:,:,:,:), rf(:,:,:,:)

   CALL adv % Build( 1 )
   N   = adv % nS
   nEq = adv % nEq
   nE  = adv % mesh % nElems
   ALLOCATE( c(0:N,0:N,1:nEq,1:nE), rl(0:N,0:N,1:nEq,1:nE), rf(0:N,0:N,1:nEq,1:nE) )
   ALLOCATE( u(0:N,0:N,1:nE), v(0:N,0:N,1:nE) )

   OPEN( NEWUNIT = uIn, FILE = 'ref_in.bin', ACCESS = 'STREAM', FORM = 'UNFORMATTED', STATUS = 'OLD' )
   READ( uIn ) mode, iter
   IF( mode == 1 )THEN
      READ( uIn ) c, u, v, rl, rf
   ENDIF
   CLOSE( uIn )

   IF( mode == 1 )THEN
      DO iEl = 1, nE
         CALL adv % SetSolution( iEl, c(:,:,:,iEl) )
         CALL adv % SetVelocity( iEl, u(:,:,iEl), v(:,:,iEl) )
         CALL adv % SetRelaxationField( iEl, rl(:,:,:,iEl) )
         CALL adv % SetRelaxationFactor( iEl, rf(:,:,:,iEl) )
      ENDDO
      CALL adv % WritePickup( iter )
   ELSE
      CALL adv % ReadPickup( iter )
      DO iEl = 1, nE
         CALL adv % GetSolution( iEl, c(:,:,:,iEl) )
         CALL adv % GetVelocity( iEl, u(:,:,iEl), v(:,:,iEl) )
         CALL adv % GetRelaxationField( iEl, rl(:,:,:,iEl) )
         CALL adv % GetRelaxationFactor( iEl, rf(:,:,:,iEl) )
      ENDDO
      OPEN( NEWUNIT = uOut, FILE = 'ref_out.bin', ACCESS = 'STREAM', FORM = 'UNFORMATTED', STATUS = 'REPLACE' )
      WRITE( uOut ) c, u, v, rl, rf
      CLOSE( uOut )
   ENDIF

END PROGRAM ref_pickup
