! self_amd_cabi.f90 -- iso_c_binding interfaces to libself_amd.so.
!
! One interface per C entry point of include/self_amd.h, with the same name
! and argument meaning; scalars by VALUE, arrays by reference in the
! reference's column-major layouts, handles as TYPE(C_PTR).  Status codes:
! SELF_OK = 0, -1 / -2 as the reference solvers' ioerr, further negative
! codes for HIP, memory, argument and state errors.
MODULE SELF_AMD_CABI
  USE ISO_C_BINDING
  IMPLICIT NONE

  INTEGER(C_INT), PARAMETER :: SELF_OK = 0, SELF_NOT_CONVERGED = -1, SELF_ERR_OTHER = -2
  INTEGER(C_INT), PARAMETER :: SELF_ERR_HIP = -3, SELF_ERR_OOM = -4, SELF_ERR_ARG = -5, SELF_ERR_STATE = -6
  ! src/common/ModelFlags.f90 values
  INTEGER(C_INT), PARAMETER :: SELF_GAUSS = 1, SELF_GAUSS_LOBATTO = -1, SELF_DG = 2000, SELF_CG = 2001
  INTEGER(C_INT), PARAMETER :: SELF_NO_NORMAL_FLOW = -100, SELF_RADIATION = -101
  INTEGER(C_INT), PARAMETER :: SELF_INFLOW_ONE = -103, SELF_INFLOW_TWO = -104
  INTEGER(C_INT), PARAMETER :: SELF_FP64 = 0, SELF_MIXED = 1, SELF_PLAIN = 0, SELF_FUSED = 1
  INTEGER(C_INT), PARAMETER :: SELF_SW_SOLUTION = 0, SELF_SW_TENDENCY = 1, SELF_SW_BATHYMETRY = 2
  INTEGER(C_INT), PARAMETER :: SELF_SW_PLANETARY_VORTICITY = 3, SELF_SW_RELAXATION = 4
  INTEGER(C_INT), PARAMETER :: SELF_SW_BOUNDARY_SOLUTION = 5, SELF_SW_BOUNDARY_FLUX = 6
  INTEGER(C_INT), PARAMETER :: SELF_KRYLOV_CG = 0, SELF_KRYLOV_PCG = 1, SELF_KRYLOV_GMRES = 2
  INTEGER(C_INT), PARAMETER :: SELF_KRYLOV_PCGMRES = 3, SELF_KRYLOV_CGSEM_PCG = 4

  TYPE, BIND(C) :: self_sw2d_desc
    INTEGER(C_INT) :: N, nElems, nEdges, precision, form
    REAL(C_DOUBLE) :: g, dt, linearDrag
  END TYPE self_sw2d_desc

  INTERFACE
    ! ------------------------------------------------------------ runtime
    FUNCTION self_last_error() BIND(C, name="self_last_error")
      IMPORT :: C_PTR
      TYPE(C_PTR) :: self_last_error
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_version() BIND(C, name="self_version")
      IMPORT :: C_INT
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_init(device) BIND(C, name="self_init")
      IMPORT :: C_INT
      INTEGER(C_INT), VALUE :: device
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_device_count(n) BIND(C, name="self_device_count")
      IMPORT :: C_INT
      INTEGER(C_INT), INTENT(out) :: n
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_mem_info(freeBytes, totalBytes) BIND(C, name="self_mem_info")
      IMPORT :: C_INT, C_SIZE_T
      INTEGER(C_SIZE_T), INTENT(out) :: freeBytes, totalBytes
    END FUNCTION

    ! ------------------------------------------------------------ host set-up
    INTEGER(C_INT) FUNCTION self_legendre_quadrature(N, quadType, x, w) BIND(C, name="self_legendre_quadrature")
      IMPORT :: C_INT, C_DOUBLE
      INTEGER(C_INT), VALUE :: N, quadType
      REAL(C_DOUBLE), INTENT(out) :: x(*), w(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_derivative_matrix(N, x, bw, D) BIND(C, name="self_derivative_matrix")
      IMPORT :: C_INT, C_DOUBLE
      INTEGER(C_INT), VALUE :: N
      REAL(C_DOUBLE), INTENT(in) :: x(*), bw(*)
      REAL(C_DOUBLE), INTENT(out) :: D(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_barycentric_weights(N, x, bw) BIND(C, name="self_barycentric_weights")
      IMPORT :: C_INT, C_DOUBLE
      INTEGER(C_INT), VALUE :: N
      REAL(C_DOUBLE) :: x(*), bw(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_interpolation_matrix(N, x, bw, nNew, zNew, T) &
        BIND(C, name="self_interpolation_matrix")
      IMPORT :: C_INT, C_DOUBLE
      INTEGER(C_INT), VALUE :: N, nNew
      REAL(C_DOUBLE) :: x(*), bw(*), zNew(*), T(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_coarse_to_fine_2d(nS, nP, nElems, Ts, Tp, nNewS, nNewP, fOld, fNew) &
        BIND(C, name="self_coarse_to_fine_2d")
      IMPORT :: C_INT, C_DOUBLE, C_SIZE_T
      INTEGER(C_INT), VALUE :: nS, nP, nNewS, nNewP
      INTEGER(C_SIZE_T), VALUE :: nElems
      REAL(C_DOUBLE) :: Ts(*), Tp(*), fOld(*), fNew(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_nodal2d_build(N, quadType, form, quirk, nodes, qWeight, dMatS, dMatP, &
                                               lagWest, lagEast, lagSouth, lagNorth) &
                                               BIND(C, name="self_nodal2d_build")
      IMPORT :: C_INT, C_DOUBLE
      INTEGER(C_INT), VALUE :: N, quadType, form, quirk
      REAL(C_DOUBLE), INTENT(out) :: nodes(*), qWeight(*), dMatS(*), dMatP(*)
      REAL(C_DOUBLE), INTENT(out) :: lagWest(*), lagEast(*), lagSouth(*), lagNorth(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_quadmesh_counts(nX, nY, periodicX, periodicY, nElems, nEdges) &
                                                 BIND(C, name="self_quadmesh_counts")
      IMPORT :: C_INT
      INTEGER(C_INT), VALUE :: nX, nY, periodicX, periodicY
      INTEGER(C_INT), INTENT(out) :: nElems, nEdges
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_quadmesh_structured(nX, nY, N, periodicX, periodicY, xScale, yScale, warp, &
                                                     sNodes, edges, x, y, dxds, dxdp, dyds, dydp, J, nHat, nLen) &
                                                     BIND(C, name="self_quadmesh_structured")
      IMPORT :: C_INT, C_DOUBLE
      INTEGER(C_INT), VALUE :: nX, nY, N, periodicX, periodicY
      REAL(C_DOUBLE), VALUE :: xScale, yScale, warp
      REAL(C_DOUBLE), INTENT(in) :: sNodes(*)
      INTEGER(C_INT), INTENT(out) :: edges(*)
      REAL(C_DOUBLE), INTENT(out) :: x(*), y(*), dxds(*), dxdp(*), dyds(*), dydp(*), J(*), nHat(*), nLen(*)
    END FUNCTION

    ! ------------------------------------------------------------ 3-D set-up and advection
    INTEGER(C_INT) FUNCTION self_coarse_to_fine_3d(N, nElems, T, nNew, fOld, fNew) BIND(C, name="self_coarse_to_fine_3d")
      IMPORT :: C_INT, C_DOUBLE, C_SIZE_T
      INTEGER(C_INT), VALUE :: N, nNew
      INTEGER(C_SIZE_T), VALUE :: nElems
      REAL(C_DOUBLE) :: T(*), fOld(*), fNew(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_nodal3d_build(N, quadType, form, nodes, qWeight, dMat, lagLeft, lagRight) &
                                               BIND(C, name="self_nodal3d_build")
      IMPORT :: C_INT, C_DOUBLE
      INTEGER(C_INT), VALUE :: N, quadType, form
      REAL(C_DOUBLE), INTENT(out) :: nodes(*), qWeight(*), dMat(*), lagLeft(*), lagRight(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_rolloff_filter1d(N, nCutoff, nodes, w, F1) BIND(C, name="self_rolloff_filter1d")
      IMPORT :: C_INT, C_DOUBLE
      INTEGER(C_INT), VALUE :: N, nCutoff
      REAL(C_DOUBLE), INTENT(in) :: nodes(*), w(*)
      REAL(C_DOUBLE), INTENT(out) :: F1(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_hexmesh_counts(nX, nY, nZ, periodicX, periodicY, periodicZ, nElems, nFaces) &
                                                BIND(C, name="self_hexmesh_counts")
      IMPORT :: C_INT
      INTEGER(C_INT), VALUE :: nX, nY, nZ, periodicX, periodicY, periodicZ
      INTEGER(C_INT), INTENT(out) :: nElems, nFaces
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_hexmesh_structured(nX, nY, nZ, N, periodicX, periodicY, periodicZ, xScale, yScale, &
                                                    zScale, warp, warpSign, nodes, faces, x, y, z, cov, J, nHat, &
                                                    nLen, bpos) BIND(C, name="self_hexmesh_structured")
      IMPORT :: C_INT, C_DOUBLE
      INTEGER(C_INT), VALUE :: nX, nY, nZ, N, periodicX, periodicY, periodicZ
      REAL(C_DOUBLE), VALUE :: xScale, yScale, zScale, warp
      INTEGER(C_INT), INTENT(in) :: warpSign(*)
      REAL(C_DOUBLE), INTENT(in) :: nodes(*)
      INTEGER(C_INT), INTENT(out) :: faces(*)
      REAL(C_DOUBLE), INTENT(out) :: x(*), y(*), z(*), cov(*), J(*), nHat(*), nLen(*), bpos(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_adv3d_create(N, nTracers, nElems, nFaces, faces, cov, J, nHat, nLen, bpos, dMat, &
                                              qWeight, lagLeft, lagRight, filter1d, useFilter, dt, h) &
                                              BIND(C, name="self_adv3d_create")
      IMPORT :: C_INT, C_DOUBLE, C_PTR
      INTEGER(C_INT), VALUE :: N, nTracers, nElems, nFaces, useFilter
      INTEGER(C_INT), INTENT(in) :: faces(*)
      REAL(C_DOUBLE), INTENT(in) :: cov(*), J(*), nHat(*), nLen(*), bpos(*), dMat(*), qWeight(*), lagLeft(*), &
                                    lagRight(*), filter1d(*)
      REAL(C_DOUBLE), VALUE :: dt
      TYPE(C_PTR), INTENT(out) :: h
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_adv3d_destroy(h) BIND(C, name="self_adv3d_destroy")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: h
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_adv3d_set_field(h, field, dat) BIND(C, name="self_adv3d_set_field")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: field
      REAL(C_DOUBLE), INTENT(in) :: dat(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_adv3d_get_field(h, field, dat) BIND(C, name="self_adv3d_get_field")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: field
      REAL(C_DOUBLE), INTENT(out) :: dat(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_adv3d_global_time_derivative(h, t, stage) &
                                                             BIND(C, name="self_adv3d_global_time_derivative")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), VALUE :: t
      INTEGER(C_INT), VALUE :: stage
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_adv3d_forward_step_rk3(h, tn, nSteps, elapsedMs) &
                                                       BIND(C, name="self_adv3d_forward_step_rk3")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), VALUE :: tn
      INTEGER(C_INT), VALUE :: nSteps
      TYPE(C_PTR), VALUE :: elapsedMs
    END FUNCTION

    ! ------------------------------------------------------------ shallow water 2-D
    INTEGER(C_INT) FUNCTION self_sw2d_create(desc, h) BIND(C, name="self_sw2d_create")
      IMPORT :: C_INT, C_PTR, self_sw2d_desc
      TYPE(self_sw2d_desc), INTENT(in) :: desc
      TYPE(C_PTR), INTENT(out) :: h
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_sw2d_destroy(h) BIND(C, name="self_sw2d_destroy")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: h
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_sw2d_set_nodal(h, dMatS, dMatP, qWeightS, qWeightP, lagWest, lagEast, &
                                                lagSouth, lagNorth) BIND(C, name="self_sw2d_set_nodal")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), INTENT(in) :: dMatS(*), dMatP(*), qWeightS(*), qWeightP(*)
      REAL(C_DOUBLE), INTENT(in) :: lagWest(*), lagEast(*), lagSouth(*), lagNorth(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_sw2d_set_geometry(h, edges, dxds, dxdp, dyds, dydp, J, nHat, nLen) &
                                                   BIND(C, name="self_sw2d_set_geometry")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), INTENT(in) :: edges(*)
      REAL(C_DOUBLE), INTENT(in) :: dxds(*), dxdp(*), dyds(*), dydp(*), J(*), nHat(*), nLen(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_sw2d_set_field(h, field, dat) BIND(C, name="self_sw2d_set_field")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: field
      REAL(C_DOUBLE), INTENT(in) :: dat(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_sw2d_get_field(h, field, dat) BIND(C, name="self_sw2d_get_field")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: field
      REAL(C_DOUBLE), INTENT(out) :: dat(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_sw2d_set_params(h, g, dt, linearDrag) BIND(C, name="self_sw2d_set_params")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), VALUE :: g, dt, linearDrag
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_sw2d_global_time_derivative(h, t) BIND(C, name="self_sw2d_global_time_derivative")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), VALUE :: t
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_sw2d_forward_step_rk3(h, tn, nSteps) BIND(C, name="self_sw2d_forward_step_rk3")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), VALUE :: tn
      INTEGER(C_INT), VALUE :: nSteps
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_sw2d_calculate_bathymetry_gradient(h) &
                                                                   BIND(C, name="self_sw2d_calculate_bathymetry_gradient")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: h
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_sw2d_synchronize(h) BIND(C, name="self_sw2d_synchronize")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: h
    END FUNCTION

    ! ------------------------------------------------------------ advection-diffusion 2-D
    INTEGER(C_INT) FUNCTION self_adv2d_create(N, nTracers, nElems, nEdges, edges, dxds, dxdp, dyds, dydp, J, &
                                              nHat, nLen, dMatS, dMatP, qWeight, lagWest, lagEast, lagSouth, &
                                              lagNorth, kappa, dt, h) BIND(C, name="self_adv2d_create")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      INTEGER(C_INT), VALUE :: N, nTracers, nElems, nEdges
      INTEGER(C_INT), INTENT(in) :: edges(*)
      REAL(C_DOUBLE), INTENT(in) :: dxds(*), dxdp(*), dyds(*), dydp(*), J(*), nHat(*), nLen(*)
      REAL(C_DOUBLE), INTENT(in) :: dMatS(*), dMatP(*), qWeight(*), lagWest(*), lagEast(*), lagSouth(*), lagNorth(*)
      REAL(C_DOUBLE), VALUE :: kappa, dt
      TYPE(C_PTR), INTENT(out) :: h
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_adv2d_destroy(h) BIND(C, name="self_adv2d_destroy")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: h
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_adv2d_set_field(h, field, dat) BIND(C, name="self_adv2d_set_field")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: field
      REAL(C_DOUBLE), INTENT(in) :: dat(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_adv2d_get_field(h, field, dat) BIND(C, name="self_adv2d_get_field")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: field
      REAL(C_DOUBLE), INTENT(out) :: dat(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_adv2d_global_time_derivative(h, t) &
                                                             BIND(C, name="self_adv2d_global_time_derivative")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), VALUE :: t
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_adv2d_forward_step_rk3(h, tn, nSteps, elapsedMs) &
                                                       BIND(C, name="self_adv2d_forward_step_rk3")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), VALUE :: tn
      INTEGER(C_INT), VALUE :: nSteps
      TYPE(C_PTR), VALUE :: elapsedMs
    END FUNCTION

    ! ------------------------------------------------------------ pickup files and ensembles
    ! path / model: NUL-terminated (pass TRIM(name)//C_NULL_CHAR)
    ! ISM-V2 SpecMesh files and ScaleTheMesh (QuadMeshClass.f90:2100-2465, :1691-1714)
    INTEGER(C_INT) FUNCTION self_specmesh_counts(path, nNodes, nEdges, nElems, gPolyDeg) &
                                                 BIND(C, name="self_specmesh_counts")
      IMPORT :: C_INT, C_CHAR
      CHARACTER(KIND=C_CHAR), INTENT(in) :: path(*)
      INTEGER(C_INT), INTENT(out) :: nNodes, nEdges, nElems, gPolyDeg
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_specmesh_read(path, N, sNodes, edges, elemNodes, nodeType, nodeX, nodeY, x, y, &
                                               dxds, dxdp, dyds, dydp, J, nHat, nLen) &
                                               BIND(C, name="self_specmesh_read")
      IMPORT :: C_INT, C_DOUBLE, C_CHAR
      CHARACTER(KIND=C_CHAR), INTENT(in) :: path(*)
      INTEGER(C_INT), VALUE :: N
      REAL(C_DOUBLE), INTENT(in) :: sNodes(*)
      INTEGER(C_INT), INTENT(out) :: edges(*), elemNodes(*), nodeType(*)
      REAL(C_DOUBLE), INTENT(out) :: nodeX(*), nodeY(*), x(*), y(*), dxds(*), dxdp(*), dyds(*), dydp(*), J(*)
      REAL(C_DOUBLE), INTENT(out) :: nHat(*), nLen(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_quadmesh_scale(N, sNodes, nElems, nNodes, xScale, yScale, nodeX, nodeY, x, y, &
                                                dxds, dxdp, dyds, dydp, J, nHat, nLen) &
                                                BIND(C, name="self_quadmesh_scale")
      IMPORT :: C_INT, C_DOUBLE
      INTEGER(C_INT), VALUE :: N, nElems, nNodes
      REAL(C_DOUBLE), VALUE :: xScale, yScale
      REAL(C_DOUBLE), INTENT(in) :: sNodes(*)
      REAL(C_DOUBLE), INTENT(inout) :: nodeX(*), nodeY(*), x(*), y(*), dxds(*), dxdp(*), dyds(*), dydp(*), J(*)
      REAL(C_DOUBLE), INTENT(inout) :: nHat(*), nLen(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_pickup_filename(model, iter, out, outLen) BIND(C, name="self_pickup_filename")
      IMPORT :: C_INT, C_CHAR
      CHARACTER(KIND=C_CHAR), INTENT(in) :: model(*)
      INTEGER(C_INT), VALUE :: iter, outLen
      CHARACTER(KIND=C_CHAR), INTENT(out) :: out(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_sw2d_write_pickup(h, path) BIND(C, name="self_sw2d_write_pickup")
      IMPORT :: C_INT, C_PTR, C_CHAR
      TYPE(C_PTR), VALUE :: h
      CHARACTER(KIND=C_CHAR), INTENT(in) :: path(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_sw2d_read_pickup(h, path) BIND(C, name="self_sw2d_read_pickup")
      IMPORT :: C_INT, C_PTR, C_CHAR
      TYPE(C_PTR), VALUE :: h
      CHARACTER(KIND=C_CHAR), INTENT(in) :: path(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_adv2d_write_pickup(h, path) BIND(C, name="self_adv2d_write_pickup")
      IMPORT :: C_INT, C_PTR, C_CHAR
      TYPE(C_PTR), VALUE :: h
      CHARACTER(KIND=C_CHAR), INTENT(in) :: path(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_adv2d_read_pickup(h, path) BIND(C, name="self_adv2d_read_pickup")
      IMPORT :: C_INT, C_PTR, C_CHAR
      TYPE(C_PTR), VALUE :: h
      CHARACTER(KIND=C_CHAR), INTENT(in) :: path(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_sw2d_set_members(h, members) BIND(C, name="self_sw2d_set_members")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: members
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_sw2d_set_member_solution(h, member, dat) &
                                                         BIND(C, name="self_sw2d_set_member_solution")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: member
      REAL(C_DOUBLE), INTENT(in) :: dat(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_sw2d_get_member_solution(h, member, dat) &
                                                         BIND(C, name="self_sw2d_get_member_solution")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: member
      REAL(C_DOUBLE), INTENT(out) :: dat(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_sw2d_set_member_time_step(h, member, dt) &
                                                          BIND(C, name="self_sw2d_set_member_time_step")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: member
      REAL(C_DOUBLE), VALUE :: dt
    END FUNCTION

    ! ------------------------------------------------------------ Krylov and CGSEM
    INTEGER(C_INT) FUNCTION self_linop_dense_create(n, A, b, op) BIND(C, name="self_linop_dense_create")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      INTEGER(C_INT), VALUE :: n
      REAL(C_DOUBLE), INTENT(in) :: A(*), b(*)
      TYPE(C_PTR), INTENT(out) :: op
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_linop_elliptic1d_create(N, nElems, xScale, op) &
        BIND(C, name="self_linop_elliptic1d_create")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      INTEGER(C_INT), VALUE :: N, nElems
      REAL(C_DOUBLE), VALUE :: xScale
      TYPE(C_PTR), INTENT(out) :: op
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_linop_elliptic1d_unmask(op, sol, u, x) BIND(C, name="self_linop_elliptic1d_unmask")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: op
      REAL(C_DOUBLE), INTENT(in) :: sol(*)
      REAL(C_DOUBLE), INTENT(out) :: u(*), x(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_linop_destroy(op) BIND(C, name="self_linop_destroy")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: op
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_krylov_solve(op, method, tol, maxIters, mInner, ortho, sol, resi, resiLen, &
                                              iterations, ioerr) BIND(C, name="self_krylov_solve")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: op
      INTEGER(C_INT), VALUE :: method, maxIters, mInner, ortho, resiLen
      REAL(C_DOUBLE), VALUE :: tol
      REAL(C_DOUBLE), INTENT(inout) :: sol(*)
      REAL(C_DOUBLE), INTENT(out) :: resi(*)
      INTEGER(C_INT), INTENT(out) :: iterations, ioerr
    END FUNCTION
    ! CGsemElliptic_2D (CGsemElliptic_2D_Class.f90): element-local arrays (0:N,0:N,nElems)
    INTEGER(C_INT) FUNCTION self_cgsem2d_create(N, nElems, nEdges, nNodes, edges, elemNodes, dxds, dxdp, dyds, &
                                                dydp, J, h) BIND(C, name="self_cgsem2d_create")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      INTEGER(C_INT), VALUE :: N, nElems, nEdges, nNodes
      INTEGER(C_INT), INTENT(in) :: edges(*), elemNodes(*)
      REAL(C_DOUBLE), INTENT(in) :: dxds(*), dxdp(*), dyds(*), dydp(*), J(*)
      TYPE(C_PTR), INTENT(out) :: h
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_cgsem2d_destroy(h) BIND(C, name="self_cgsem2d_destroy")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: h
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_cgsem2d_set_flux_coefficient(h, k) BIND(C, name="self_cgsem2d_set_flux_coefficient")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), INTENT(in) :: k(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_cgsem2d_set_source(h, src) BIND(C, name="self_cgsem2d_set_source")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), INTENT(in) :: src(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_cgsem2d_set_dirichlet(h, vals) BIND(C, name="self_cgsem2d_set_dirichlet")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), INTENT(in) :: vals(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_cgsem2d_set_solution(h, u) BIND(C, name="self_cgsem2d_set_solution")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), INTENT(in) :: u(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_cgsem2d_get_solution(h, u) BIND(C, name="self_cgsem2d_get_solution")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), INTENT(out) :: u(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_cgsem2d_solve(h, method, precond, tol, maxIters, mInner, ortho, ioerr, &
                                               iterations, resi, resiLen, elapsedMs) BIND(C, name="self_cgsem2d_solve")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: method, precond, maxIters, mInner, ortho, resiLen
      REAL(C_DOUBLE), VALUE :: tol
      INTEGER(C_INT), INTENT(out) :: ioerr, iterations
      REAL(C_DOUBLE), INTENT(out) :: resi(*)
      TYPE(C_PTR), VALUE :: elapsedMs
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_cgsem2d_matrix_action(h, u, Au) BIND(C, name="self_cgsem2d_matrix_action")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), INTENT(in) :: u(*)
      REAL(C_DOUBLE), INTENT(out) :: Au(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_cgsem2d_mask(h, u) BIND(C, name="self_cgsem2d_mask")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), INTENT(inout) :: u(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_cgsem2d_dof_map(h, gid) BIND(C, name="self_cgsem2d_dof_map")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), INTENT(out) :: gid(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_cgsem2d_global_dofs(h, nGlobal) BIND(C, name="self_cgsem2d_global_dofs")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), INTENT(out) :: nGlobal
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_cgsem2d_set_preconditioner(h, x, y, pcIterates, pcTolerance, coefficient) &
                                                            BIND(C, name="self_cgsem2d_set_preconditioner")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), INTENT(in) :: x(*), y(*)
      INTEGER(C_INT), VALUE :: pcIterates, coefficient
      REAL(C_DOUBLE), VALUE :: pcTolerance
    END FUNCTION
    ! cols / vals: TYPE(C_PTR) so that a width query can pass C_NULL_PTR
    INTEGER(C_INT) FUNCTION self_cgsem2d_pc_matrix(h, width, cols, vals) BIND(C, name="self_cgsem2d_pc_matrix")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: h, cols, vals
      INTEGER(C_INT), INTENT(out) :: width
    END FUNCTION
    INTEGER(C_INT) FUNCTION self_cgsem2d_solve_reliable(h, precond, tol, delta, maxIters, maxUpdates, ioerr, &
                                                        updates, iterations, relResidual, elapsedMs) &
                                                        BIND(C, name="self_cgsem2d_solve_reliable")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: precond, maxIters, maxUpdates
      REAL(C_DOUBLE), VALUE :: tol, delta
      INTEGER(C_INT), INTENT(out) :: ioerr, updates, iterations
      REAL(C_DOUBLE), INTENT(out) :: relResidual
      TYPE(C_PTR), VALUE :: elapsedMs
    END FUNCTION
  END INTERFACE

CONTAINS

  ! Message of the last failed call on this thread, as a Fortran string.
  FUNCTION self_error_message() RESULT(msg)
    CHARACTER(LEN=:), ALLOCATABLE :: msg
    CHARACTER(KIND=C_CHAR), POINTER :: c(:)
    INTEGER :: n
    INTERFACE
      FUNCTION c_strlen(s) BIND(C, name="strlen")
        IMPORT :: C_PTR, C_SIZE_T
        TYPE(C_PTR), VALUE :: s
        INTEGER(C_SIZE_T) :: c_strlen
      END FUNCTION
    END INTERFACE
    TYPE(C_PTR) :: p
    p = self_last_error()
    n = INT(c_strlen(p))
    CALL C_F_POINTER(p, c, [n])
    ALLOCATE(CHARACTER(LEN=n) :: msg)
    msg = TRANSFER(c(1:n), msg)
  END FUNCTION self_error_message

  ! The reference stops on invalid input (PRINT + STOP); the shim does the
  ! same on any failed call so drivers keep their behaviour.
  SUBROUTINE self_check(rc, where)
    INTEGER(C_INT), INTENT(in) :: rc
    CHARACTER(LEN=*), INTENT(in) :: where
    IF (rc /= SELF_OK) THEN
      PRINT *, 'self_amd: ', where, ' failed (', rc, '): ', self_error_message()
      STOP 1
    ENDIF
  END SUBROUTINE self_check

END MODULE SELF_AMD_CABI
