/*
 * self_amd.h -- C ABI of the MI355X-native SELF spectral-element hot path.
 *
 * Plain C types only (int, double, float, pointers, sizes): every entry
 * point is callable from Fortran through iso_c_binding, from C/C++, or from
 * Python through ctypes.  Each declaration names the reference routine it
 * replaces (paths relative to the lileipku00/SELF repository root).
 *
 * Conventions
 *  - Return value: SELF_OK (0) or a negative status.  -1 and -2 keep the
 *    reference solvers' ioerr meaning (IterativeSolvers_Class.f90:337-340);
 *    the rest are new.  self_last_error() gives a message for the calling
 *    thread.  Nothing here aborts the process.
 *  - Host arrays use the reference's Fortran column-major layouts with its
 *    0-based node ranges.  For per-element storage the element index is the
 *    slowest dimension, i.e. the reference's per-element arrays
 *    sol(iEl)%solution(0:N,0:N,1:nEq) stacked for iEl = 1..nElems.
 *  - Element ids in edge tables are 1-based and sides 1..4 are
 *    south, east, north, west, as in QuadMeshClass.f90.
 *  - A handle owns its device buffers and one HIP stream (default priority);
 *    state stays resident between calls.  One host thread uses a handle at a
 *    time.  Calls on a handle are asynchronous on its stream unless stated;
 *    Get/Set calls and *_synchronize return after the stream has drained.
 */
#ifndef SELF_AMD_H
#define SELF_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SELF_OK 0
#define SELF_NOT_CONVERGED (-1)   /* reference ioerr -1 */
#define SELF_ERR_OTHER (-2)       /* reference ioerr -2 */
#define SELF_ERR_HIP (-3)
#define SELF_ERR_OOM (-4)
#define SELF_ERR_ARG (-5)
#define SELF_ERR_STATE (-6)

/* Flags with the values of src/common/ModelFlags.f90 */
#define SELF_GAUSS 1
#define SELF_GAUSS_LOBATTO (-1)
#define SELF_DG 2000
#define SELF_CG 2001
#define SELF_NO_NORMAL_FLOW (-100)
#define SELF_RADIATION (-101)
#define SELF_INFLOW_ONE (-103)
#define SELF_INFLOW_TWO (-104)

/* Precision of a model's tendency arithmetic */
#define SELF_FP64 0
#define SELF_MIXED 1   /* FP32 volume-flux contractions, FP64 state/metrics/RK */

/* Execution form of a model's stage */
#define SELF_PLAIN 0   /* one kernel per reference phase, fields round-trip HBM */
#define SELF_FUSED 1   /* traces + Riemann flux + volume derivative + lift + RK3 in one kernel */

/* ------------------------------------------------------------------ runtime */
const char *self_last_error(void);
int self_version(void);
/* Select the HIP device used by handles created afterwards on this thread. */
int self_init(int device);
int self_device_count(int *n);
/* Free device memory in bytes, as used for admission decisions. */
int self_mem_info(size_t *free_bytes, size_t *total_bytes);

/* STREAM-style copy of `bytes` between two device buffers, iters times:
 * attainable HBM bandwidth (read + write) and mean ms per copy.  Measurement
 * support (bandwidth ceiling and byte-counter calibration); no reference
 * counterpart. */
int self_stream_copy(size_t bytes, int iters, double *gbps, double *ms);

/* ------------------------------------------------------- host setup (CPU)
 * Quadrature, interpolants and derivative matrices are computed on the host
 * once, exactly as the reference's Build routines do, then uploaded. */

/* GenerateLegendreQuadrature: src/interp/Legendre.f90:180-212
 * (LegendreGauss :216-308, LegendreGaussLobatto :312-405). */
int self_legendre_quadrature(int N, int quadType, double *x, double *w);
/* CalculateBarycentricWeights_1D: src/interp/Lagrange_1D_Class.f90:349-392 */
int self_barycentric_weights(int N, const double *x, double *bw);
/* EvaluateLagrangePolynomial_1D: src/interp/Lagrange_1D_Class.f90:448-515 */
int self_lagrange_polynomials(int N, const double *x, const double *bw, double xEval, double *l);
/* CalculateDerivativeMatrix_1D: src/interp/Lagrange_1D_Class.f90:612-689; D(k,j) at k+(N+1)*j */
int self_derivative_matrix(int N, const double *x, const double *bw, double *D);
/* CalculateInterpolationMatrix_1D: src/interp/Lagrange_1D_Class.f90:693-764; T(k,j) at k+(nNew+1)*j */
int self_interpolation_matrix(int N, const double *x, const double *bw, int nNew,
                              const double *zNew, double *T);
/* BuildNodalStorage_2D: src/nodal/NodalStorage_2D_Class.f90:101-230.
 * quirk_dmatp_transposed = 1 stores dMatP exactly as the reference's DG
 * branch does (:209-218, the transpose of the weak operator); 0 stores the
 * weak operator -W^-1 D^T W in both directions, as the reference's 1-D and
 * 3-D storages do.  Arrays: (N+1) or (N+1)^2 doubles. */
int self_nodal2d_build(int N, int quadType, int form, int quirk_dmatp_transposed,
                       double *nodes, double *qWeight, double *dMatS, double *dMatP,
                       double *lagWest, double *lagEast, double *lagSouth, double *lagNorth);

/* Structured quad mesh with the node/element numbering of
 * LoadDefaultMesh_QuadMesh (src/geom/QuadMeshClass.f90:1718-1859), the edge
 * table of ConstructEdges_QuadMesh (:1489-1627) and its start/inc
 * convention (:1830-1846), optional periodic pairing per direction (new:
 * the reference has no periodic boundaries), and the covariant metrics,
 * Jacobian and boundary normals of MappedGeometry_2D (src/geom/
 * MappedGeometryClass_2D.f90:611-710, scaled as :921-948) for the bilinear
 * element map composed with the unit-square warp
 *   X = a + warp sin(2 pi a) sin(2 pi b),  Y = b + warp sin(2 pi a) sin(2 pi b).
 * edges: 6 ints per edge (e1, s1, e2, s2, start, inc); e2 = NO_NORMAL_FLOW on
 * walls.  Per-element arrays (N+1)^2 per element; nHat 2*(N+1)*4 and nLen
 * (N+1)*4 per element, ordered [node][side] with side 1..4. */
int self_quadmesh_counts(int nX, int nY, int periodicX, int periodicY, int *nElems, int *nEdges);
int self_quadmesh_structured(int nX, int nY, int N, int periodicX, int periodicY,
                             double xScale, double yScale, double warp, const double *sNodes,
                             int *edges, double *x, double *y, double *dxds, double *dxdp,
                             double *dyds, double *dydp, double *J, double *nHat, double *nLen);

/* ISM-V2 SpecMesh files (ReadSpecMeshFile_QuadMesh, src/geom/
 * QuadMeshClass.f90:2100-2465): unstructured quads, side curves of degree
 * gPolyDeg on Chebyshev-Gauss-Lobatto points, side names setting the
 * boundary flags.  counts: corners, edges (as rebuilt by ConstructEdges,
 * :1489-1627), elements, curve degree.  read: the edge table (6 ints per
 * edge as self_quadmesh_structured; e2 = the side's boundary flag on a
 * named boundary side, NO_NORMAL_FLOW otherwise; start/inc :2421-2431),
 * elemNodes (4 corners per element, 1-based), the corner node types
 * (INTERIOR = 1, or the flag of a named side through the corner, :2339-2419),
 * corner positions, and the element geometry at the interpolation nodes
 * sNodes by transfinite interpolation of the side curves
 * (MappedGeometryClass_2D.f90:536-710), in the layouts of
 * self_quadmesh_structured.  Files without the ISM-V2 version line, with
 * side flags other than 0/1, or whose edge count disagrees with their
 * elements are rejected (SELF_ERR_ARG). */
int self_specmesh_counts(const char *path, int *nNodes, int *nEdges, int *nElems, int *gPolyDeg);
/* ScaleTheMesh_QuadMesh (QuadMeshClass.f90:1691-1714) with
 * ScaleGeometry_MappedGeometry_2D (MappedGeometryClass_2D.f90:921-948), in
 * place on arrays in the layouts above: corner positions, x, y and the
 * metric terms scaled per direction, J by xScale*yScale, and the boundary
 * normals and lengths recomputed from the derivative of each element's
 * position interpolant on the nodes sNodes (GenerateBoundaryMetrics,
 * :714-788).  The shallow-water model applies it to every mesh it reads
 * (ConservativeShallowWaterClass.f90:302-304).  nNodes may be 0 (no corners). */
int self_quadmesh_scale(int N, const double *sNodes, int nElems, int nNodes, double xScale, double yScale,
                        double *nodeX, double *nodeY, double *x, double *y, double *dxds, double *dxdp, double *dyds,
                        double *dydp, double *J, double *nHat, double *nLen);
int self_specmesh_read(const char *path, int N, const double *sNodes, int *edges, int *elemNodes, int *nodeType,
                       double *nodeX, double *nodeY, double *x, double *y, double *dxds, double *dxdp, double *dyds,
                       double *dydp, double *J, double *nHat, double *nLen);

/* ------------------------------------------------- nodal library operators
 * Batched over nElems elements stored one after the other, each in the
 * reference's column-major per-element layout.  The plain forms take host
 * arrays (copied in and out, synchronous); the _dev forms take device
 * pointers and enqueue on `stream` (a hipStream_t, NULL = default stream).
 *
 * ApplyDerivativeMatrix_NodalStorage_2D (src/nodal/NodalStorage_2D_Class.f90
 * :620-646): dFds(:,j) = dMatS f(:,j), dFdp(i,:) = dMatP f(i,:). */
int self_nodal2d_apply_derivative(int N, size_t nElems, const double *dMatS, const double *dMatP, const double *f,
                                  double *dFds, double *dFdp);
int self_nodal2d_apply_derivative_dev(int N, size_t nElems, const double *dMatS, const double *dMatP,
                                      const double *f, double *dFds, double *dFdp, void *stream);
/* CalculateAtBoundaries_NodalStorage_2D (:650-680): fWest/fEast along s
 * lines, fSouth/fNorth along p lines; outputs (0:N) per element. */
int self_nodal2d_boundary_interp(int N, size_t nElems, const double *lW, const double *lE, const double *lS,
                                 const double *lN, const double *f, double *fW, double *fE, double *fS, double *fN);
int self_nodal2d_boundary_interp_dev(int N, size_t nElems, const double *lW, const double *lE, const double *lS,
                                     const double *lN, const double *f, double *fW, double *fE, double *fS,
                                     double *fN, void *stream);
/* ApplyDerivativeMatrix_NodalStorage_3D (src/nodal/NodalStorage_3D_Class.f90
 * :737-774), the same matrix in s, p and q. */
int self_nodal3d_apply_derivative(int N, size_t nElems, const double *dMat, const double *f, double *dFds,
                                  double *dFdp, double *dFdq);
int self_nodal3d_apply_derivative_dev(int N, size_t nElems, const double *dMat, const double *f, double *dFds,
                                      double *dFdp, double *dFdq, void *stream);
/* CalculateAtBoundaries_NodalStorage_3D (:778-823): faces (0:N,0:N,1:6)
 * per element in the order south, east, north, west, bottom, top. */
int self_nodal3d_boundary_interp(int N, size_t nElems, const double *lagLeft, const double *lagRight,
                                 const double *f, double *faces);
int self_nodal3d_boundary_interp_dev(int N, size_t nElems, const double *lagLeft, const double *lagRight,
                                     const double *f, double *faces, void *stream);
/* CoarseToFine_2D (src/interp/Lagrange_2D_Class.f90:488-528): fOld
 * (0:nS,0:nP) -> fNew (0:nNewP,0:nNewS), the new p index first as the
 * reference returns it; Ts (0:nNewS,0:nS), Tp (0:nNewP,0:nP). */
int self_coarse_to_fine_2d(int nS, int nP, size_t nElems, const double *Ts, const double *Tp, int nNewS, int nNewP,
                           const double *fOld, double *fNew);
int self_coarse_to_fine_2d_dev(int nS, int nP, size_t nElems, const double *Ts, const double *Tp, int nNewS,
                               int nNewP, const double *fOld, double *fNew, void *stream);
/* No counterpart in the reference: self_coarse_to_fine_2d for callers that
 * cannot allocate device memory (code under HIP graph capture, output loops
 * of a serving process).  Same arguments and results (host arrays, bitwise
 * the same values); the device staging lives in a 16 MiB array that the
 * library holds on every device it runs on until the process ends, so it
 * calls no hipMalloc/hipFree.  Rejects (-5) calls whose matrices, input and
 * output need more than 2^21 doubles together.  Calls on one device share
 * the array and are serialized inside the process. */
int self_coarse_to_fine_2d_noalloc(int nS, int nP, size_t nElems, const double *Ts, const double *Tp, int nNewS,
                                   int nNewP, const double *fOld, double *fNew);
/* CoarseToFine_3D (src/interp/Lagrange_3D_Class.f90:545-604): (0:N)^3 ->
 * (0:nNew)^3 with T (0:nNew,0:N) in each direction. */
int self_coarse_to_fine_3d(int N, size_t nElems, const double *T, int nNew, const double *fOld, double *fNew);
int self_coarse_to_fine_3d_dev(int N, size_t nElems, const double *T, int nNew, const double *fOld, double *fNew,
                               void *stream);

/* ------------------------------------------- conservative shallow water 2-D
 * Replaces the ShallowWater type of
 * src/highend/shallowwater/ConservativeShallowWaterClass.f90 (fields :60-151).
 */
typedef struct self_sw2d_s *self_sw2d_t;

typedef struct {
    int N;            /* polyDeg (nS = nP = N), 1 <= N <= 11 */
    int nElems;
    int nEdges;
    int precision;    /* SELF_FP64 | SELF_MIXED */
    int form;         /* SELF_PLAIN | SELF_FUSED */
    double g;         /* SWParams%g */
    double dt;        /* SWParams%dt */
    double linearDrag;/* SWParams%linearDrag */
} self_sw2d_desc;

/* Field selectors for Set/Get (host layout per element, element slowest). */
#define SELF_SW_SOLUTION 0        /* (0:N,0:N,1:3)  U, V, H                    (sol%solution) */
#define SELF_SW_TENDENCY 1        /* (0:N,0:N,1:3)                             (sol%tendency), get only */
#define SELF_SW_BATHYMETRY 2      /* (0:N,0:N,1:3)  h, dh/dx, dh/dy            (bathymetry%solution) */
#define SELF_SW_PLANETARY_VORTICITY 3 /* (0:N,0:N)  f                          (vorticity var 1) */
#define SELF_SW_RELAXATION 4      /* (0:N,0:N,1:4)  U_r, V_r, H_r, 1/timescale (relax%solution) */
#define SELF_SW_BOUNDARY_SOLUTION 5 /* (0:N,1:3,1:4)                           (sol%boundarySolution), get only */
#define SELF_SW_BOUNDARY_FLUX 6   /* (0:N,1:3,1:4)                             (sol%boundaryFlux), get only */

/* Build_ShallowWater (:163-245): allocates device state for the descriptor. */
int self_sw2d_create(const self_sw2d_desc *desc, self_sw2d_t *out);
/* Trash_ShallowWater (:249-268) */
int self_sw2d_destroy(self_sw2d_t h);
/* NodalStorage_2D arrays used by the model (GetDerivativeMatrix and friends,
 * NodalStorage_2D_Class.f90:403-418). */
int self_sw2d_set_nodal(self_sw2d_t h, const double *dMatS, const double *dMatP,
                        const double *qWeightS, const double *qWeightP,
                        const double *lagWest, const double *lagEast,
                        const double *lagSouth, const double *lagNorth);
/* Mesh inputs: edge table as self_quadmesh_structured, metrics per element,
 * boundary normals [node][side] per element. */
int self_sw2d_set_geometry(self_sw2d_t h, const int *edges, const double *dxds,
                           const double *dxdp, const double *dyds, const double *dydp,
                           const double *J, const double *nHat, const double *nLen);
/* SetSolution/SetBathymetry/SetPlanetaryVorticity/SetRelaxationField
 * (:300-900 accessors) and their Get counterparts, whole mesh at once. */
int self_sw2d_set_field(self_sw2d_t h, int field, const double *data);
int self_sw2d_get_field(self_sw2d_t h, int field, double *data);
int self_sw2d_set_params(self_sw2d_t h, double g, double dt, double linearDrag);
/* CalculateSolutionAtBoundaries for every element (:1146-1156) */
int self_sw2d_calculate_solution_at_boundaries(self_sw2d_t h);
/* EdgeFlux for every edge (:1266-1360) */
int self_sw2d_edge_flux(self_sw2d_t h, double t);
/* MappedTimeDerivative for every element (:1364-1522) */
int self_sw2d_mapped_time_derivative(self_sw2d_t h, double t);
/* GlobalTimeDerivative (:1222-1262): the three phases above. */
int self_sw2d_global_time_derivative(self_sw2d_t h, double t);
/* Guard of the mixed form (no counterpart: the reference computes in FP64
 * only).  For a mixed single-state model: the momentum tendency (U, V) of
 * the current state through one mixed fused stage, against the FP64
 * GlobalTimeDerivative of the same state: relErr = |T_mixed - T_FP64|_2 /
 * |T_FP64|_2.  Where the FP32 contractions lose the tendency to
 * cancellation (deep, steep bathymetry, where g H^2/2 and the bathymetry
 * source are many orders larger than their sum) relErr exceeds tol and the
 * model switches to the FP64 fused form (*switched = 1).  The state is not
 * changed; FP64 models return relErr = 0. */
int self_sw2d_check_mixed(self_sw2d_t h, double tol, double *relErr, int *switched);
/* ForwardStepRK3 (:1173-1218) applied nSteps times from tn. */
int self_sw2d_forward_step_rk3(self_sw2d_t h, double tn, int nSteps);
/* Same, bracketed by HIP events on the handle's stream; *elapsed_ms is the
 * device time of the nSteps steps. */
int self_sw2d_forward_step_rk3_timed(self_sw2d_t h, double tn, int nSteps, double *elapsed_ms);
/* Runs nSteps RK3 steps with one HIP event pair around every kernel launch
 * and returns the mean device time per launch of each kernel of the stage:
 * FUSED -> 1 kernel (the stage kernel); PLAIN -> 4 (traces, edge flux,
 * mapped tendency, RK update).  Measurement support, no reference counterpart. */
int self_sw2d_kernel_times(self_sw2d_t h, double tn, int nSteps, double *avg_ms, int *nKernels);
/* CalculateBathymetryAtBoundaries + CalculateBathymetryGradient for every
 * element (:1160-1168, :1733-1868): sets bathymetry vars 2 and 3. */
int self_sw2d_calculate_bathymetry_gradient(self_sw2d_t h);
/* Which shallow-water class the handle follows: 0 (default) the src one
 * (src/highend/shallowwater/ConservativeShallowWaterClass.f90, with g for the
 * undeclared name in its RiemannSolver); 1 the class as the examples ship it
 * (examples/shallowwater/planewave/localmods/ConservativeShallowWaterClass.f90):
 * Rusanov wave speed max|(u_n +- sqrt(u_n^2 + 2 g H^3))/(2H)| (:1555-1558),
 * Source of Coriolis and g H grad(h) only (:1649-1656; linearDrag and the
 * relaxation field are ignored), bathymetry gradient grad(h^2/2)/h
 * (:1754-1825).  -5 for any other value. */
int self_sw2d_set_variant(self_sw2d_t h, int variant);
int self_sw2d_synchronize(self_sw2d_t h);
/* Element groups per workgroup of the fused stage kernel: 1 (one round of
 * waves), 2 (each workgroup prefetches its second group while computing the
 * first), 0 = default (1).  No reference counterpart. */
int self_sw2d_set_groups(self_sw2d_t h, int groups);
/* Fused form: run pairs of RK3 steps as one captured HIP graph (six stage
 * launches), re-captured whenever a launch it holds would change (default
 * on); 0 launches every stage from the host.  No reference counterpart. */
int self_sw2d_set_graph(self_sw2d_t h, int enable);
/* ForwardStepRK3 nSteps times for nModels fused-form models of any degrees
 * (1..11, one precision), every RK stage of all of them in one persistent
 * launch whose workgroups walk the models' element groups; the same
 * results as stepping each model alone.  Blocking; elapsed_ms (device time
 * of the launches) nullable.  No reference counterpart (batching). */
int self_sw2d_forward_step_rk3_multi(self_sw2d_t *models, int nModels, int nSteps, double *elapsed_ms);
/* Ensembles (no reference counterpart: the reference runs one state per
 * model object).  `members` states on the same mesh, nodal storage,
 * bathymetry, vorticity and relaxation advance together, one
 * fused launch per RK stage for all of them (fused form only).  Setting
 * the count reallocates the states (zeroed); self_sw2d_set_field /
 * get_field (SOLUTION) address member 0. */
int self_sw2d_set_members(self_sw2d_t h, int members);
int self_sw2d_set_member_solution(self_sw2d_t h, int member, const double *data);
/* Time step of one member (default: the handle's dt; self_sw2d_set_params
 * resets every member to its dt).  The stage update of member k uses
 * g_m * dt_k, the same product a separate run forms. */
int self_sw2d_set_member_time_step(self_sw2d_t h, int member, double dt);
int self_sw2d_get_member_solution(self_sw2d_t h, int member, double *data);
/* Autotuner: element groups per workgroup (1, 2) x graph replay (off, on).
 * Takes the process-wide table's entry for this degree, arithmetic and size
 * class when there is one; otherwise times nSteps RK3 steps with each
 * setting on a copy of the current state (the state is restored), keeps
 * the fastest, records it in the table, and returns its groups and time
 * per step. */
int self_sw2d_autotune(self_sw2d_t h, int nSteps, int *groups, double *ms_per_step);
/* Bytes of HBM state the handle holds. */
int self_sw2d_device_bytes(self_sw2d_t h, size_t *bytes);
/* WritePickup (ConservativeShallowWaterClass.f90:2013-2065) of the state,
 * bathymetry and vorticity (member 0) to `path`; ReadPickup (:2069-2125)
 * from it, followed by the bathymetry gradient as Build does (:224-228). */
int self_sw2d_write_pickup(self_sw2d_t h, const char *path);
int self_sw2d_read_pickup(self_sw2d_t h, const char *path);

/* --------------------------------------------------------- pickup files
 * The reference's restart files: unformatted direct access, big-endian,
 * RECL = 8 (N+1)^2 bytes, one record per (0:N,0:N) element field, elements
 * in order.  Shallow water (ConservativeShallowWaterClass.f90:2013-2125):
 * U, V, H, h, f per element; reading leaves the bathymetry gradient zero
 * (bathy = ZERO, :2088).  Advection (AdvectionClass.f90:2040-2171):
 * c(1..nTracers), relaxation field(1..nTracers), relaxation
 * factor(1..nTracers), u, v per element.  Host arrays in the layouts of
 * the set/get calls above (element slowest).  A file whose size is not
 * that of nElems elements is rejected. */
/* '<model>.<iter as 10 digits>.pickup' (WRITE(iterChar,'(I10.10)')), e.g. model "ShallowWater" */
int self_pickup_filename(const char *model, int iter, char *out, int outLen);
int self_sw2d_pickup_write(const char *path, int N, int nElems, const double *sol, const double *bathy,
                           const double *plv);
int self_sw2d_pickup_read(const char *path, int N, int nElems, double *sol, double *bathy, double *plv);
int self_adv2d_pickup_write(const char *path, int N, int nElems, int nTracers, const double *c, const double *relax,
                            const double *rFac, const double *vel);
int self_adv2d_pickup_read(const char *path, int N, int nElems, int nTracers, double *c, double *relax, double *rFac,
                           double *vel);

/* ------------------------------------------------------------ Krylov solvers
 * The reference's plug-in point is the abstract IterativeSolver type with
 * deferred MatrixAction / Residual / PreConditionInvert procedures
 * (src/iterativesolve/IterativeSolvers_Class.f90:55-186).  Here an operator
 * handle plays that role: it owns device data and applies itself on the
 * device, and the solver keeps every vector and scalar on the device.
 * Methods keep the reference's control flow, residual history resi(0:) and
 * stop tests:
 *   SELF_KRYLOV_CG      Solve_ConjugateGradient    :328-405
 *   SELF_KRYLOV_PCG     Solve_PCConjugateGradient  :409-489  (Jacobi)
 *   SELF_KRYLOV_GMRES   Solve_GMRES                :493-652
 *   SELF_KRYLOV_PCGMRES Solve_PCGMRES              :656-822  (Jacobi, flexible)
 * ortho: 0 = modified Gram-Schmidt (the reference), 1 = classical
 * Gram-Schmidt applied twice (two sweeps over the basis per step).
 * ioerr: 0 converged, -1 not converged, -2 as the reference leaves it. */
#define SELF_KRYLOV_CG 0
#define SELF_KRYLOV_PCG 1
#define SELF_KRYLOV_GMRES 2
#define SELF_KRYLOV_PCGMRES 3
#define SELF_KRYLOV_CGSEM_PCG 4  /* CGsemElliptic_2D's own PCG loop (:1114-1207) */

typedef struct self_linop_s *self_linop_t;

/* Dense n x n operator (column-major A), residual b - A x, Jacobi diag(A)^-1. */
int self_linop_dense_create(int n, const double *A, const double *b, self_linop_t *out);
/* Elliptic1D (examples/elliptic1d/parabola/localmods/Elliptic1D_Class.f90):
 * Build :111-176 on nElems elements of degree N on [0,1], k = 1, source 1,
 * homogeneous Dirichlet ends; flat vectors of (N+1)*nElems entries in the
 * reference's masked element-local layout. */
int self_linop_elliptic1d_create(int N, int nElems, double xScale, self_linop_t *out);
/* UnMaskSolution (:268-276): the continuous solution at every local node,
 * and the node positions x (nullable). */
int self_linop_elliptic1d_unmask(self_linop_t op, const double *sol, double *u, double *x);
int self_linop_size(self_linop_t op, int *n);
int self_linop_destroy(self_linop_t op);
/* IterativeSolver%Solve(ioerr): sol (host, nDOF) is the initial guess in and
 * the solution out; resi (host, resiLen, nullable) receives resi(0:). */
int self_krylov_solve(self_linop_t op, int method, double tol, int maxIters, int mInner, int ortho,
                      double *sol, double *resi, int resiLen, int *iterations, int *ioerr);

/* ------------------------------------------- advection-diffusion 2-D (DGSEM)
 * Replaces the Advection type of
 * src/highend/advection-diffusion2d/AdvectionClass.f90:57-165: tracers
 * c(0:N,0:N,1:nTracers) carried by a static velocity, diffusivity kappa,
 * BR1-type central gradient, upwind advective flux, optional relaxation
 * source.  Boundary sides take the external state of GetExternalState
 * (:1851-1886): tanh(t) on tracer 1 at SELF_INFLOW_ONE, on tracer 2 at
 * SELF_INFLOW_TWO, 0 elsewhere.  FP64 throughout. */
typedef struct self_adv2d_s *self_adv2d_t;

/* Build_Advection (:176-252) for a mesh from self_quadmesh_structured (or
 * any edge table of the same form) and the nodal arrays of
 * self_nodal2d_build; N 1..8, nTracers 1..4. */
int self_adv2d_create(int N, int nTracers, int nElems, int nEdges, const int *edges, const double *dxds,
                      const double *dxdp, const double *dyds, const double *dydp, const double *J,
                      const double *nHat, const double *nLen, const double *dMatS, const double *dMatP,
                      const double *qWeight, const double *lagWest, const double *lagEast,
                      const double *lagSouth, const double *lagNorth, double kappa, double dt,
                      self_adv2d_t *out);
/* Trash_Advection (:256-288) */
int self_adv2d_destroy(self_adv2d_t h);
/* field 0: SetSolution (:362-375), (0:N,0:N,1:nTracers) per element;
 * 1: SetVelocity (:434-449), (0:N,0:N,1:2) per element, also runs
 * CalculateVelocityAtBoundaries (:1108-1119) once since velocity is static;
 * 2: SetRelaxationField (:954-967), 3: SetRelaxationFactor (:1022-1035),
 * tracer-shaped.  Element
 * slowest in every host array. */
int self_adv2d_set_field(self_adv2d_t h, int field, const double *data);
/* field 0: solution, 1: tendency (after self_adv2d_global_time_derivative),
 * 2: dc/dx, 3: dc/dy (CalculateConcentrationGradient :1123-1264),
 * 4: velocity (0:N,0:N,1:2) (GetVelocity :415-430), 5: relaxation field,
 * 6: relaxation factor (GetRelaxationField :937-950, GetRelaxationFactor
 * :1005-1018; zeros if never set). */
int self_adv2d_get_field(self_adv2d_t h, int field, double *data);
/* GlobalTimeDerivative (:1316-1378) at time t: tendency of the current state. */
int self_adv2d_global_time_derivative(self_adv2d_t h, double t);
/* ForwardStepRK3 (:1268-1312) applied nSteps times from tn; elapsed_ms
 * (nullable) is the device time of the steps. */
int self_adv2d_forward_step_rk3(self_adv2d_t h, double tn, int nSteps, double *elapsed_ms);
int self_adv2d_synchronize(self_adv2d_t h);
/* WritePickup (:2040-2103) / ReadPickup (:2107-2171) of the handle's
 * fields to / from `path` (see self_adv2d_pickup_write). */
int self_adv2d_write_pickup(self_adv2d_t h, const char *path);
int self_adv2d_read_pickup(self_adv2d_t h, const char *path);

/* ------------------------------------------------------ 3-D set-up (CPU)
 * BuildNodalStorage_3D (src/nodal/NodalStorage_3D_Class.f90:107-264) with
 * equal degree in s, p, q: nodes, weights, derivative matrix (SELF_DG: the
 * weak operator -W^-1 D^T W, :223-251; SELF_CG: D), Lagrange vectors at -1
 * (west, south, bottom) and +1 (east, north, top). */
int self_nodal3d_build(int N, int quadType, int form, double *nodes, double *qWeight, double *dMat,
                       double *lagLeft, double *lagRight);
/* Direction factor F1 ((N+1)^2, column-major) of RollOffFilter3D
 * (src/filters/RollOffFilter3D_Class.f90:87-180): the reference's dense
 * (N+1)^3 filter matrix equals F1 (x) F1 (x) F1, since the Legendre modes,
 * the quadrature weights and exp(-ln2 (a^2+b^2+c^2)/nc^2) all factor over
 * the three directions. */
int self_rolloff_filter1d(int N, int nCutoff, const double *nodes, const double *w, double *F1);
/* Structured hex mesh of the unit cube, elements iX fastest then iY, iZ;
 * sides 1..6 = south, east, north, west, bottom, top (ModelFlags.f90:36-41);
 * optional periodic pairing per direction (the reference's default mesh has
 * none) and the warp X = a + A s1 S, Y = b + A s2 S, Z = c + A s3 S with
 * S = sin(2 pi a) sin(2 pi b) sin(2 pi c), scaled by (xScale, yScale, zScale).
 * faces: 9 ints per face (e1, s1, e2, s2, iStart, jStart, iInc, jInc, swap)
 * as the reference's face loop uses them (Advection3DClass.f90:1286-1296);
 * e2 = NO_NORMAL_FLOW on boundaries.  Per element: x, y, z, J ((N+1)^3),
 * cov (9 per node: dxds dxdp dxdq dyds dydp dydq dzds dzdp dzdq), and per
 * face node [side][b][a]: nHat (3), nLen, bpos (3, the node position);
 * face node (a, b) is (p, q) on west/east, (s, q) on south/north, (s, p) on
 * bottom/top. */
int self_hexmesh_counts(int nX, int nY, int nZ, int periodicX, int periodicY, int periodicZ, int *nElems,
                        int *nFaces);
int self_hexmesh_structured(int nX, int nY, int nZ, int N, int periodicX, int periodicY, int periodicZ,
                            double xScale, double yScale, double zScale, double warp, const int *warpSign,
                            const double *nodes, int *faces, double *x, double *y, double *z, double *cov,
                            double *J, double *nHat, double *nLen, double *bpos);

/* ---------------------------------------------------------- autotuning
 * The process-wide table of tuned launch configurations, keyed by model
 * kind ("sw2d", "sw2d_ensemble"), dimension, degree N, arithmetic (0 FP64,
 * 1 mixed) and size class floor(log2(elements)); entries hold four
 * configuration integers and the time per step that chose them.  No
 * reference counterpart. */
int self_tune_count(int *n);
/* entry index: kind (NUL-terminated, kindLen bytes), key4 = dim, N, dtype,
 * sizeClass; cfg4; ms per step */
int self_tune_entry(int index, char *kind, int kindLen, int *key4, int *cfg4, double *ms);
int self_tune_clear(void);
/* text file, one entry a line: kind dim N dtype sizeClass c0 c1 c2 c3 ms */
int self_tune_save(const char *path);
int self_tune_load(const char *path);

/* ---------------------------------------------------- 3-D advection (DGSEM)
 * Replaces the Advection type of src/highend/advection3d/Advection3DClass.f90
 * (:57-160): tracers c(0:N,0:N,0:N,1:nTracers) carried by a static velocity,
 * upwind flux, relaxation source, the adaptive roll-off filter
 * (DoTheAdaptiveFiltering :1125-1233) applied at the start of every stage.
 * External states as GetExternalState (:1698-1735): tanh(t) on tracer 1
 * (SELF_INFLOW_ONE) or 2 (SELF_INFLOW_TWO), the moving Gaussian
 * exp(-|x_b - t|^2 / 0.04) at SELF_NO_NORMAL_FLOW, 0 otherwise.  FP64. */
typedef struct self_adv3d_s *self_adv3d_t;

/* Build_Advection (:167-272) on a mesh of self_hexmesh_structured's form;
 * nodal arrays from self_nodal3d_build; filter1d from self_rolloff_filter1d
 * (nullable when useFilter = 0); N 1..7, nTracers 1..2. */
int self_adv3d_create(int N, int nTracers, int nElems, int nFaces, const int *faces, const double *cov,
                      const double *J, const double *nHat, const double *nLen, const double *bpos,
                      const double *dMat, const double *qWeight, const double *lagLeft, const double *lagRight,
                      const double *filter1d, int useFilter, double dt, self_adv3d_t *out);
/* Trash_Advection (:276-312) */
int self_adv3d_destroy(self_adv3d_t h);
/* field 0: SetSolution (:369-382) (0:N,0:N,0:N,1:nTracers) per element;
 * 1: SetVelocity (:477-494) (0:N,0:N,0:N,1:3) per element, with the face
 * velocities of CalculateVelocityAtBoundaries computed once; 2: relaxation
 * field, 3: relaxation factor (tracer-shaped). */
int self_adv3d_set_field(self_adv3d_t h, int field, const double *data);
/* field 0: solution, 1: tendency, 2: filter state E1 (nTracers x nElems),
 * E2 (same), limiter (3 x nTracers x nElems), in that order. */
int self_adv3d_get_field(self_adv3d_t h, int field, double *data);
/* GlobalTimeDerivative(t, stage) (:1049-1121): filters the state in place
 * (when enabled), then computes the tendency. */
int self_adv3d_global_time_derivative(self_adv3d_t h, double t, int stage);
/* ForwardStepRK3 (:1000-1045) nSteps times from tn; elapsed_ms nullable. */
int self_adv3d_forward_step_rk3(self_adv3d_t h, double tn, int nSteps, double *elapsed_ms);
/* Element-tracer pairs the filter replaced since the last reset
 * (measurement support, no reference counterpart). */
int self_adv3d_filter_count(self_adv3d_t h, int reset, unsigned *count);
/* Arithmetic of the volume contractions of MappedTimeDerivative
 * (:1345-1562): 0 FP64 (default, the reference's), 1 FP32 on the matrix
 * cores, 2 split-BF16 (three bf16 products, FP32 accumulation) on the
 * matrix cores; the line deviations enter the narrow arithmetic and the
 * line values are added back in FP64.  State, RK register, geometry, face
 * fluxes and the filter stay FP64.  No reference counterpart (precision
 * selection). */
int self_adv3d_set_precision(self_adv3d_t h, int precision);
int self_adv3d_synchronize(self_adv3d_t h);

/* --------------------------------- hydrostatic primitive equations (DGSEM)
 * Replaces the HydrostaticPrimitive type of src/highend/hydrostaticprimitives/
 * HydrostaticPrimitivesClass.f90 (:80-210): u, v, b prognostic, w and p
 * diagnosed (variables 1..5 as :205-209), on a stacked hex mesh whose
 * element e = hE + nHElems * layer (StackedHexMeshClass.f90:337); bottom
 * boundary faces are SEA_FLOOR, other boundary faces NO_NORMAL_FLOW
 * (StackedHexMeshClass.f90:617-621).  The reference class does not compile;
 * the time derivative follows GlobalTimeDerivative (:1428-1506) with divU
 * from CalculateDivU (HydrostaticPrimitivesClass_implicit.f90:2520-2691)
 * and DiagnoseP inside it when diagnoseP is set.  FP64. */
typedef struct self_hpe_s *self_hpe_t;

/* Build (:219-317) on a mesh of self_hexmesh_structured's form with nZ =
 * nLayers (no periodicity in z); nodal arrays from self_nodal3d_build (Gauss,
 * DG); N 1..7.  The column quadrature degree is (N+1)/2 + 1 (:264). */
int self_hpe_create(int N, int nHElems, int nLayers, int nFaces, const int *faces, const double *cov, const double *J,
                    const double *nHat, const double *nLen, const double *bpos, const double *nodes,
                    const double *qWeight, const double *dMat, const double *lagLeft, const double *lagRight,
                    int diagnoseP, double dt, self_hpe_t *out);
/* Trash (:321-355) */
int self_hpe_destroy(self_hpe_t h);
/* field 0: SetSolution (:420-433) (0:N,0:N,0:N,1:5) per element;
 * 1: SetBackgroundState (:601-617), same shape (its buoyancy and vertical
 * velocity enter the fluxes); 2: SetPreviousTendency (:1090-1107)
 * (0:N,0:N,0:N,1:3), the AB2 history. */
int self_hpe_set_field(self_hpe_t h, int field, const double *data);
/* field 0: solution (5 variables), 1: GetTendency (:1111-1128) (3),
 * 2: divergence of the lateral velocity of the last time derivative (1),
 * 3: face values of b, w, p of the last time derivative, [e][side][var][node]. */
int self_hpe_get_field(self_hpe_t h, int field, double *data);
/* GlobalTimeDerivative (:1428-1506): diagnoses p (when enabled) and w into
 * the solution and computes the tendency. */
int self_hpe_global_time_derivative(self_hpe_t h);
/* method 0: ForwardStepRK3 (HydrostaticPrimitivesClass_implicit.f90:
 * 1422-1471) without the pressure correction; 1: ForwardStepAB2
 * (:1378-1424) without the barotropic step.  nSteps steps; elapsed_ms
 * nullable. */
int self_hpe_forward_step(self_hpe_t h, int method, int nSteps, double *elapsed_ms);
/* Average duration in ms of the column kernel (avg_ms[0]) and the stage
 * kernel (avg_ms[1]) over nSteps RK3 steps (measurement support, no
 * reference counterpart). */
int self_hpe_kernel_times(self_hpe_t h, int nSteps, double *avg_ms);
int self_hpe_synchronize(self_hpe_t h);

/* ---------------------------------------- CG spectral-element elliptic 2-D
 * Replaces CGsemElliptic_2D (src/cgsem/CGsemElliptic_2D_Class.f90:57-93):
 * div(k grad u) = s on GLL nodes of degree N, Dirichlet on every boundary
 * edge.  Arrays use the reference's element-local layout (0:N,0:N,nElems);
 * the device keeps global-DOF vectors (one entry per shared node). */
typedef struct self_cgsem2d_s *self_cgsem2d_t;

/* Build :99-154 with the mesh of self_quadmesh_structured (GLL sNodes);
 * elemNodes: 4 corner node ids per element (SW, SE, NE, NW), 1-based;
 * boundary edges (e2 <= 0) become Dirichlet (BuildQuadMesh :197-215). */
int self_cgsem2d_create(int N, int nElems, int nEdges, int nNodes, const int *edges, const int *elemNodes,
                        const double *dxds, const double *dxdp, const double *dyds, const double *dydp,
                        const double *J, self_cgsem2d_t *out);
int self_cgsem2d_destroy(self_cgsem2d_t h);
int self_cgsem2d_global_dofs(self_cgsem2d_t h, int *nGlobal);
/* fluxCoeff (:64), source (:63), Dirichlet values at boundary nodes (dirF of
 * SetDirichletBoundaryCondition :234-300, sampled per local node) */
int self_cgsem2d_set_flux_coefficient(self_cgsem2d_t h, const double *k);
int self_cgsem2d_set_source(self_cgsem2d_t h, const double *s);
int self_cgsem2d_set_dirichlet(self_cgsem2d_t h, const double *values);
int self_cgsem2d_set_solution(self_cgsem2d_t h, const double *u);
/* solution after UnMask + SetDirichletBoundaryCondition */
int self_cgsem2d_get_solution(self_cgsem2d_t h, double *u);
/* MatrixAction (:930-976) of a masked local array; Au returned masked */
int self_cgsem2d_matrix_action(self_cgsem2d_t h, const double *u, double *Au);
/* Mask (:350-430) in place on an element-local array: the owner copy of
 * each global DOF keeps its value (edge: the primary element's copy;
 * corner: the first element of the node's element list), other copies and
 * Dirichlet nodes become 0. */
int self_cgsem2d_mask(self_cgsem2d_t h, double *u);
/* Local node -> index of its global DOF in the solver's vectors (0-based),
 * -1 on Dirichlet nodes; element-local layout (no reference counterpart:
 * the reference keeps masked local arrays). */
int self_cgsem2d_dof_map(self_cgsem2d_t h, int *gid);
/* The preconditioner's assembled N = 1 operator (MatrixAction of
 * CGsemPreconditioner_2D, CGsemPreconditioner_2D_Class.f90:564-610, on the
 * sub-mesh's free nodes, which are the solver's global DOFs): row g holds
 * *width (column, value) pairs, unused slots (g, 0.0); cols / vals
 * (nGlobal x width, row-major) may be null to ask for the width.  What
 * ConstructPCMatrix (:750-851) writes out. */
int self_cgsem2d_pc_matrix(self_cgsem2d_t h, int *width, int *cols, double *vals);
/* Residual (:980-1068) of the current solution, masked local layout */
int self_cgsem2d_residual(self_cgsem2d_t h, double *r);
/* CGsemPreconditioner_2D % Build (src/cgsem/CGsemPreconditioner_2D_Class.f90
 * :88-126, called by Build_CGsemElliptic_2D :146-152): the operator at N = 1
 * on the linear mesh through the Gauss-Lobatto nodes x, y (element-local
 * layout as the metric terms), inverted by pcIterates iterations of CG to
 * the relative tolerance pcTolerance (Solve :664-746).  coefficient 0: k = 1,
 * which is what the reference maps at Build (fluxCoeff is still ONE there);
 * 1: the flux coefficient last given to self_cgsem2d_set_flux_coefficient.
 * Selected by precond = 2 in the solves below. */
int self_cgsem2d_set_preconditioner(self_cgsem2d_t h, const double *x, const double *y, int pcIterates,
                                    double pcTolerance, int coefficient);
/* CGsemPreconditioner_2D % Solve (:664-746): x = M^-1 b for a host array in
 * the element-local layout (owner copies read; every copy written, 0 on
 * Dirichlet nodes); iterations (nullable) = inner CG iterations. */
int self_cgsem2d_apply_preconditioner(self_cgsem2d_t h, const double *b, double *x, int *iterations);
/* Solve (:1114-1207 for SELF_KRYLOV_CGSEM_PCG; the generic methods above
 * applied to the same operator otherwise).  precond 0 = identity (the
 * reference's pcIterates = 0), 1 = Jacobi (assembled diagonal), 2 = the
 * preconditioner set above (CGSEM PCG and PCGMRES, the flexible form
 * Solve_PCGMRES of IterativeSolvers_Class.f90:656-822).
 * elapsed_ms (nullable): device time of the solve. */
int self_cgsem2d_solve(self_cgsem2d_t h, int method, int precond, double tol, int maxIters, int mInner,
                       int ortho, int *ioerr, int *iterations, double *resi, int resiLen, double *elapsed_ms);
/* Batched CGSEM PCG (the reference's loop, CGsemElliptic_2D_Class.f90:
 * 1114-1207, run for many problems at once; no counterpart for the
 * batching): problems[0..nProblems) -- distinct models on one device, each
 * with its own mesh, degree, flux coefficient, source, Dirichlet values
 * and initial guess -- are solved together, every CG iteration being two
 * launches for all problems still iterating.  precond 0 identity, 1 Jacobi.
 * Each problem's solution, ioerr[p] and iterations[p] are bitwise those of
 * self_cgsem2d_solve(problems[p], SELF_KRYLOV_CGSEM_PCG, precond, tol,
 * maxIters, ...).  Work is queued on problems[0]'s stream; elapsed_ms
 * (nullable): device time of the whole batch. */
int self_cgsem2d_solve_batched(self_cgsem2d_t *problems, int nProblems, int precond, double tol, int maxIters,
                               int *ioerr, int *iterations, double *elapsed_ms);
/* Restarted FGMRES(mInner) (IterativeSolvers_Class.f90:656-822 control
 * flow) whose operator and preconditioner applications run in FP32 while
 * the basis, Gram-Schmidt (ortho 0 modified, 1 classical twice), Givens
 * rotations, the update and the true residual of every restart stay FP64,
 * so each cycle refines the FP64 solution; up to maxCycles cycles to
 * |r|/|r0| <= tol.  precond as above.  No reference counterpart beyond the
 * loop it follows. */
int self_cgsem2d_solve_gmres_mixed(self_cgsem2d_t h, int precond, double tol, int maxCycles, int mInner, int ortho,
                                   int *ioerr, int *iterations, double *relResidual, double *elapsed_ms);
/* Mixed precision: FP32 PCG inner solves to innerTol, FP64 residual and
 * iterative refinement until |b - A x| <= tol |b - A x0|. */
int self_cgsem2d_solve_mixed(self_cgsem2d_t h, int precond, double tol, double innerTol, int maxInner,
                             int maxOuter, int *ioerr, int *outerIters, int *innerIters, double *relResidual,
                             double *elapsed_ms);
/* Mixed precision with reliable updates: one FP32 PCG sequence that is
 * never restarted; whenever its residual has dropped by delta (0 < delta
 * < 1) the FP32 iterate is added to the FP64 solution and the residual is
 * recomputed in FP64, until |b - A x| <= tol |b - A x0|.  maxIters bounds
 * the FP32 iterations, maxUpdates the refinements; ioerr 0 converged, -1
 * not.  (No reference counterpart: the reference solves in FP64 only.) */
int self_cgsem2d_solve_reliable(self_cgsem2d_t h, int precond, double tol, double delta, int maxIters,
                                int maxUpdates, int *ioerr, int *updates, int *iterations, double *relResidual,
                                double *elapsed_ms);
/* The same with an FP64 PCG sequence: the CGSEM PCG's iteration
 * (CGsemElliptic_2D_Class.f90:1114-1207) whose residual is replaced by the
 * true FP64 residual whenever it has dropped by delta, so that the solve
 * stops on |b - A x| <= tol |b - A x0| itself, where the reference's loop
 * stops on its recursively updated residual (which drifts from the true
 * one by rounding). */
int self_cgsem2d_solve_reliable_fp64(self_cgsem2d_t h, int precond, double tol, double delta, int maxIters,
                                     int maxUpdates, int *ioerr, int *updates, int *iterations, double *relResidual,
                                     double *elapsed_ms);

/* --------------------------------------------------------------- serving
 * Request loop (no reference counterpart: the reference's drivers run one
 * model at a time).  Shallow-water integrations and CGSEM elliptic solves
 * of different mesh sizes and degrees are admitted as they arrive against
 * free HBM, run concurrently on their own HIP streams from a pool of worker
 * threads, and report queue time, run time and latency.  Results are
 * bitwise the same alone and in any mix (every kernel is deterministic). */
#define SELF_REQ_SW2D 1
#define SELF_REQ_CGSEM2D 2

typedef struct {
    int kind;        /* SELF_REQ_SW2D | SELF_REQ_CGSEM2D */
    int nX;          /* nX x nX elements (SW: doubly periodic square of side L; CGSEM: walled unit square) */
    int N;           /* degree: SW on Gauss nodes, CGSEM on Gauss-Lobatto nodes */
    int precision;   /* SELF_FP64 | SELF_MIXED (SW: mixed tendency; CGSEM: FP32 inner PCG + FP64 refinement) */
    int steps;       /* SW: RK3 steps */
    int precond;     /* CGSEM: 0 identity, 1 Jacobi */
    int maxIters;    /* CGSEM: iteration limit */
    double L, g, f0, dt;  /* SW: domain side, gravity, Coriolis parameter, time step */
    double tol;      /* CGSEM: relative tolerance */
    const double *in0;    /* SW: solution (0:N,0:N,1:3) per element; CGSEM: flux coefficient (0:N,0:N) */
    const double *in1;    /* SW: bathymetry (h, 0, 0) per element; CGSEM: source */
    const double *in2;    /* CGSEM: Dirichlet values (nullable = 0) */
    double *out;          /* SW: final solution; CGSEM: solution */
} self_request;

typedef struct {
    int status;           /* SELF_OK or the failing call's status */
    int ioerr, iterations;/* CGSEM */
    double queue_ms, run_ms, latency_ms;  /* arrival -> start, start -> end, arrival -> end */
    size_t bytes;         /* device memory reserved for the request */
    int members;          /* SW: requests that ran in the same ensemble launch (1 = alone) */
    int models;           /* SW: models (ensembles of any mesh and degree) advanced in the same
                             persistent launches (1 = its ensemble alone) */
} self_request_result;

typedef struct self_server_s *self_server_t;

/* workers: concurrent requests (1..64); memFraction: share of the free HBM
 * (at creation) that admitted requests may reserve. */
int self_serve_create(int workers, double memFraction, self_server_t *out);
/* Enqueue a request (arrival time = now); its arrays must stay valid until
 * self_serve_wait returns for it. */
int self_serve_submit(self_server_t s, const self_request *req, long long *id);
/* Block until the request finished; fills res and forgets the request. */
int self_serve_wait(self_server_t s, long long id, self_request_result *res);
int self_serve_stats(self_server_t s, int *queued, int *running, size_t *reservedBytes, size_t *budgetBytes);
/* Runs what is queued, then joins the workers. */
int self_serve_destroy(self_server_t s);

#ifdef __cplusplus
}
#endif
#endif
