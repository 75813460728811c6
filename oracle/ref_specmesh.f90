! ref_specmesh.f90 -- test infrastructure only (a checker, never shipped or timed).
!
! Runs the reference's own ISM-V2 SpecMesh reader and element geometry
! (src/geom/QuadMeshClass.f90:2100-2465 ReadSpecMeshFile_QuadMesh, with
! ConstructEdges :1489-1627 and MappedGeometry_2D's transfinite
! interpolation, MappedGeometryClass_2D.f90:536-710; compiled unmodified,
! recipe oracle/ref.mk) on a mesh file and writes what it built.
!
! Run in a directory holding ref_in.bin (stream): int32 N, quadrature
! (1 Gauss, -1 Gauss-Lobatto), name length, the mesh file name, then
! int32 doScale and float64 xScale, yScale: with doScale = 1 the mesh is
! then scaled by the reference's ScaleTheMesh (QuadMeshClass.f90:1691-1714),
! as the shallow-water model does after reading it.
! It writes ref_out.bin (stream):
!   int32 nElems, nEdges, nNodes
!   int32 edges(6,nEdges) = e1, s1, e2, s2, start, increment
!   int32 elemNodes(4,nElems), nodeType(nNodes)
!   float64 nodeX(nNodes), nodeY(nNodes)
!   float64 x, y, dxds, dxdp, dyds, dydp, J (0:N,0:N,nElems each),
!           nHat(2,0:N,4,nElems), nLen(0:N,4,nElems)
PROGRAM ref_specmesh
 USE ModelPrecision
 USE ConstantsDictionary
 USE ModelFlags
 USE NodalStorage_2D_Class
 USE QuadMeshClass
 IMPLICIT NONE
 TYPE(NodalStorage_2D) :: ns
 TYPE(QuadMesh) :: mesh
 INTEGER :: N, quad, nameLen, doScale, nE, nEd, nNd, iEl, iEdge, iNode, i, j, b, k, uIn, uOut
 INTEGER, ALLOCATABLE :: edges(:,:), en(:,:), nt(:)
 REAL(prec) :: nrm(1:2), length, xs, ys
 REAL(prec), ALLOCATABLE :: g(:,:,:,:), nHat(:,:,:,:), nLen(:,:,:), nx(:), ny(:)
 CHARACTER(LEN=512) :: fname

   OPEN( NEWUNIT = uIn, FILE = 'ref_in.bin', ACCESS = 'STREAM', FORM = 'UNFORMATTED', STATUS = 'OLD' )
   READ( uIn ) N, quad, nameLen
   fname = ' '
   READ( uIn ) fname(1:nameLen)
   READ( uIn ) doScale, xs, ys
   CLOSE( uIn )

   CALL ns % Build( N, N, quad, DG )
   CALL mesh % ReadSpecMeshFile( ns % interp, TRIM(fname) )
   IF( doScale == 1 ) CALL mesh % ScaleTheMesh( ns % interp, xs, ys )
   nE  = mesh % nElems
   nEd = mesh % nEdges
   nNd = mesh % nNodes
   ALLOCATE( edges(1:6,1:nEd), en(1:4,1:nE), nt(1:nNd), nx(1:nNd), ny(1:nNd) )
   ALLOCATE( g(0:N,0:N,1:nE,1:7), nHat(1:2,0:N,1:4,1:nE), nLen(0:N,1:4,1:nE) )

   DO iEdge = 1, nEd
      CALL mesh % GetEdgePrimaryElementID( iEdge, edges(1,iEdge) )
      CALL mesh % GetEdgePrimaryElementSide( iEdge, edges(2,iEdge) )
      CALL mesh % GetEdgeSecondaryElementID( iEdge, edges(3,iEdge) )
      CALL mesh % GetEdgeSecondaryElementSide( iEdge, edges(4,iEdge) )
      CALL mesh % GetEdgeStart( iEdge, edges(5,iEdge) )
      CALL mesh % GetEdgeIncrement( iEdge, edges(6,iEdge) )
   ENDDO
   DO iEl = 1, nE
      DO k = 1, 4
         CALL mesh % GetElementNodeID( iEl, k, en(k,iEl) )
      ENDDO
   ENDDO
   DO iNode = 1, nNd
      CALL mesh % GetNodeType( iNode, nt(iNode) )
      CALL mesh % GetNodePosition( iNode, nx(iNode), ny(iNode) )
   ENDDO
   DO iEl = 1, nE
      DO j = 0, N
         DO i = 0, N
            CALL mesh % GetPositionAtNode( iEl, g(i,j,iEl,1), g(i,j,iEl,2), i, j )
            CALL mesh % GetCovariantMetricsAtNode( iEl, g(i,j,iEl,3), g(i,j,iEl,4), &
                                                   g(i,j,iEl,5), g(i,j,iEl,6), i, j )
            CALL mesh % GetJacobianAtNode( iEl, g(i,j,iEl,7), i, j )
         ENDDO
      ENDDO
      DO b = 1, 4
         DO i = 0, N
            CALL mesh % GetBoundaryNormalAtNode( iEl, nrm, length, i, b )
            nHat(1:2,i,b,iEl) = nrm
            nLen(i,b,iEl) = length
         ENDDO
      ENDDO
   ENDDO

   OPEN( NEWUNIT = uOut, FILE = 'ref_out.bin', ACCESS = 'STREAM', FORM = 'UNFORMATTED', STATUS = 'REPLACE' )
   WRITE( uOut ) nE, nEd, nNd
   WRITE( uOut ) edges, en, nt
   WRITE( uOut ) nx, ny
   DO b = 1, 7
      WRITE( uOut ) g(:,:,:,b)
   ENDDO
   WRITE( uOut ) nHat, nLen
   CLOSE( uOut )

END PROGRAM ref_specmesh
