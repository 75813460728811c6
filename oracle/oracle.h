/*
 * oracle.h -- FP64 CPU restatement of the SELF reference algorithms.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in self_amd/: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never links it.
 *
 * Every routine restates one reference routine (file:line cited beside it,
 * paths relative to the reference repository root) in plain C99, keeping the
 * reference's loop order and operation order so that results agree to the
 * last few ulps.  Arrays follow the reference's Fortran column-major layout
 * with the reference's 0-based node ranges; element and side indices are
 * spelled out at each declaration.
 *
 * Pinning (how each function is checked, see tests/test_oracle_*.py):
 *   quadrature/Lagrange/nodal : closed forms, numpy.polynomial.legendre,
 *                               polynomial exactness, D*1 = 0.
 *   krylov CG                 : reference fixture examples/elliptic1d/
 *                               parabola/run/{residual,elliptic}.curve.
 *   krylov PCG/GMRES/FGMRES   : brute force vs numpy.linalg.solve on small
 *                               SPD / nonsymmetric systems; GMRES residual
 *                               monotone within a cycle.
 *   sw2d                      : Riemann-flux consistency F*(q,q)=F(q).n and
 *                               antisymmetry; free-stream (lake at rest,
 *                               constant state) preservation; exact discrete
 *                               conservation of sum J w H under periodic BCs;
 *                               derivative exactness on polynomials; RK3
 *                               third-order convergence on y'=lambda y.
 *   adv2d / adv3d / cgsem2d   : see their headers below.
 */
#ifndef SELF_ORACLE_H
#define SELF_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Flags: src/common/ModelFlags.f90 */
#define ORC_GAUSS 1
#define ORC_GAUSS_LOBATTO (-1)
#define ORC_DG 2000
#define ORC_CG 2001
#define ORC_NO_NORMAL_FLOW (-100)
#define ORC_RADIATION (-101)
#define ORC_PRESCRIBED (-102)
#define ORC_INFLOW_ONE (-103)
#define ORC_INFLOW_TWO (-104)
#define ORC_DIRICHLET (-200)

/* ---------------- L1: Legendre quadrature (src/interp/Legendre.f90) ------- */
void orc_legendre_polynomial(int nP, double x, double *lAtX, double *dLdxAtX); /* :44-114 */
void orc_legendre_q_and_l(int nP, double x, double *q, double *qprime, double *lN); /* :118-174 */
void orc_legendre_gauss(int nP, double *xNodes, double *weights);          /* :216-308 */
void orc_legendre_gauss_lobatto(int nP, double *xNodes, double *weights);  /* :312-405 */
int  orc_generate_legendre_quadrature(int nP, double *xNodes, double *weights,
                                      int quadType);                        /* :180-212 */

/* ---------------- L1: Lagrange interpolation (src/interp/Lagrange_1D_Class.f90) */
int    orc_almost_equal(double a, double b);                 /* src/common/CommonRoutines.f90:58-97 */
void   orc_barycentric_weights(int nP, const double *nodes, double *bw);        /* :349-392 */
void   orc_lagrange_polynomials_at(int nP, const double *nodes, const double *bw,
                                   double xEval, double *lAtX);                  /* :448-515 */
double orc_lagrange_interpolate(int nP, const double *nodes, const double *bw,
                                double xEval, const double *f);                  /* EvaluateLagrange_1D */
double orc_lagrange_derivative(int nP, const double *nodes, const double *bw,
                               double xEval, const double *f);                   /* :519-606 */
void   orc_derivative_matrix(int nP, const double *nodes, const double *bw,
                             double *dMat);                                      /* :612-689, dMat(k,j) at [k+(nP+1)*j] */
void   orc_interpolation_matrix(int nP, const double *nodes, const double *bw,
                                int nNew, const double *zNew, double *T);        /* :693-764, T(k,j) at [k+(nNew+1)*j] */
/* Lagrange_2D CoarseToFine (src/interp/Lagrange_2D_Class.f90:488-528):
 * fNew(0:nNewS,0:nNewP) = Ts * fOld * Tp^T for one element. */
void   orc_coarse_to_fine_2d(int nS, int nP, const double *fOld, int nNewS, int nNewP,
                             const double *Ts, const double *Tp, double *fNew);
/* Lagrange_3D CoarseToFine (src/interp/Lagrange_3D_Class.f90, same sum factorisation in 3 directions) */
void   orc_coarse_to_fine_3d(int n, const double *fOld, int nNew, const double *T, double *fNew);

/* ---------------- L2: nodal storage (src/nodal/NodalStorage_2D_Class.f90) -- */
typedef struct {
    int N;                 /* nS = nP = N */
    int quadrature;        /* ORC_GAUSS | ORC_GAUSS_LOBATTO */
    int form;              /* ORC_DG | ORC_CG */
    double *nodes;         /* (0:N) */
    double *qWeight;       /* (0:N) */
    double *bWeights;      /* barycentric weights (0:N) */
    double *dMatS;         /* (0:N,0:N) column-major */
    double *dMatP;         /* (0:N,0:N) column-major; for DG this is the reference's
                              transposed operator (quirk Q1), see orc_nodal2d_build */
    double *dMatCorrect;   /* (0:N,0:N) the un-transposed DG weak operator (== dMatS) */
    double *lagWest, *lagEast, *lagSouth, *lagNorth; /* (0:N) */
} orc_nodal2d;

/* BuildNodalStorage_2D :101-230.  All arrays are caller-allocated (N+1 or (N+1)^2). */
int  orc_nodal2d_build(orc_nodal2d *ns);
/* ApplyDerivativeMatrix_NodalStorage_2D :620-646 with an explicit pair of matrices */
void orc_apply_derivative_2d(int N, const double *dMatS, const double *dMatP,
                             const double *f, double *dFds, double *dFdp);
/* CalculateAtBoundaries_NodalStorage_2D :650-680 */
void orc_boundary_interp_2d(int N, const double *lW, const double *lE,
                            const double *lS, const double *lN, const double *f,
                            double *fW, double *fE, double *fS, double *fN);
/* NodalStorage_3D_Class.f90:737-774 ApplyDerivativeMatrix (3 directions) */
void orc_apply_derivative_3d(int N, const double *dMat, const double *f,
                             double *dFds, double *dFdp, double *dFdq);
/* NodalStorage_3D_Class.f90:778-823 CalculateAtBoundaries (6 faces, order
 * south,east,north,west,bottom,top; face arrays (0:N,0:N)) */
void orc_boundary_interp_3d(int N, const double *lLeft, const double *lRight,
                            const double *f, double *faces /* 6*(N+1)^2 */);

/* ---------------- L3: structured quad mesh + geometry ---------------------
 * LoadDefaultMesh_QuadMesh (src/geom/QuadMeshClass.f90:1718-1859) node and
 * element numbering, ConstructEdges_QuadMesh (:1489-1627) edge discovery and
 * start/inc convention, MappedGeometry_2D metrics (src/geom/
 * MappedGeometryClass_2D.f90:611-710) evaluated for a bilinear element map
 * (the reference's transfinite interpolation of straight sides) optionally
 * composed with the smooth warp  X = a + A sin(2 pi a) sin(2 pi b),
 * Y = b + A sin(2 pi a) sin(2 pi b)  of the unit square, then scaled by
 * (xScale, yScale) as ScaleGeometry_MappedGeometry_2D (:921-948).
 * Periodic identification of the last node column/row (not in the
 * reference, which has no periodic boundaries) is optional per direction.
 */
typedef struct {
    int nX, nY, N;
    int periodicX, periodicY;
    double xScale, yScale, warp;
    int nElems, nEdges, nNodes;   /* filled by orc_quadmesh_counts */
} orc_quadmesh_desc;

void orc_quadmesh_counts(orc_quadmesh_desc *d);
/* edges[6*iEdge + {0..5}] = e1, s1, e2, s2, start, inc  (element ids 1-based,
 * sides 1..4 = south,east,north,west; e2 = NO_NORMAL_FLOW on boundaries,
 * s2 < 0 for opposite orientation).  Geometry arrays per element e (0-based):
 * x,y,dxds,dxdp,dyds,dydp,J at [i + (N+1)*(j + (N+1)*e)];
 * nHat at [c + 2*(node + (N+1)*(side + 4*e))], nLen at [node + (N+1)*(side+4*e)],
 * side 0..3 = south,east,north,west. */
int orc_quadmesh_build(const orc_quadmesh_desc *d, const double *sNodes,
                       int *elemNodes, int *edges,
                       double *x, double *y, double *dxds, double *dxdp,
                       double *dyds, double *dydp, double *J,
                       double *nHat, double *nLen);

/* ---------------- L6: conservative shallow water 2-D ----------------------
 * src/highend/shallowwater/ConservativeShallowWaterClass.f90 (with the
 * one-token fix of quirk Q2: the undeclared g in RiemannSolver is locG).
 * Per-element arrays, element e 0-based, fields f 0..2 = (U, V, H):
 *   sol, tend      [i + n1*(j + n1*(f + 3*e))]
 *   bathy          [i + n1*(j + n1*(k + 3*e))]   k = h, dh/dx, dh/dy
 *   relax          [i + n1*(j + n1*(k + 4*e))]   k = U_r, V_r, H_r, 1/timescale
 *   plv            [i + n1*(j + n1*e)]           planetary vorticity f
 *   bsol, bflux    [node + n1*(f + 3*(side + 4*e))]
 *   bbathy         [node + n1*(k + 3*(side + 4*e))]
 */
typedef struct {
    int N, nElems, nEdges;
    double g, dt, linearDrag;
    const double *dMatS, *dMatP, *qWeightS, *qWeightP;
    const double *lagWest, *lagEast, *lagSouth, *lagNorth;
    const int *edges;
    const double *dxds, *dxdp, *dyds, *dydp, *J, *nHat, *nLen;
    double *sol, *tend, *bsol, *bflux;
    double *bathy, *bbathy, *relax, *plv;
    double *G;        /* RK3 register, same layout as sol (ForwardStepRK3 G2D) */
    /* 0: src/highend/shallowwater/ConservativeShallowWaterClass.f90;
     * 1: the class as the examples ship it
     *    (examples/shallowwater/planewave/localmods/ConservativeShallowWaterClass.f90):
     *    wave speed max|(u_n +- sqrt(u_n^2 + 2 g H^3))/(2H)| (:1555-1558),
     *    Source = g H grad(h) + Coriolis only (:1649-1656; drag and
     *    relaxation dropped), bathymetry gradient grad(h^2/2)/h (:1754-1825). */
    int form;
} orc_sw2d;

void orc_sw2d_riemann(const double *inS, const double *outS, const double *nHat,
                      double g, double *numFlux);                               /* :1555-1604 */
void orc_sw2d_riemann_form(const double *inS, const double *outS, const double *nHat,
                           double g, int form, double *numFlux);

void orc_sw2d_external_state(int bcFlag, const double *nHat, const double *inS,
                             double *exS);                                      /* :1698-1729 */
void orc_sw2d_solution_at_boundaries(orc_sw2d *m);   /* DGSEMSolutionStorageClass_2D.f90:749-802 */
void orc_sw2d_edge_flux(orc_sw2d *m, double tn);     /* :1266-1360 */
void orc_sw2d_mapped_time_derivative(orc_sw2d *m, double tn);  /* :1364-1522 */
void orc_sw2d_global_time_derivative(orc_sw2d *m, double tn);  /* :1222-1262 */
void orc_sw2d_forward_step_rk3(orc_sw2d *m, double tn);        /* :1173-1218 */
/* several RK3 steps with nThreads OpenMP threads over elements/edges (the
 * reference's !$OMP DO regions, ConservativeShallowWaterClass.f90:1236-1259) */
void orc_sw2d_integrate(orc_sw2d *m, double t0, int nSteps, int nThreads);
void orc_sw2d_bathymetry_gradient(orc_sw2d *m);  /* :1733-1868 (+ CalculateBathymetryAtBoundaries) */

/* ---------------- L5: Krylov solvers (src/iterativesolve/IterativeSolvers_Class.f90)
 * The deferred procedures are callbacks on flat vectors s(1:nDOF). */
typedef void (*orc_vec_op)(void *ctx, const double *in, double *out);
typedef struct {
    int nDOF, maxIters, mInnerIters;
    double tolerance;
    void *ctx;
    orc_vec_op matrixAction, residual, precondition;
    double *sol;     /* (1:nDOF), initial guess in, solution out */
    double *resi;    /* (0:resiLen-1) residual history */
    int resiLen;
    int iterations;  /* out: iterations performed (CG/PCG), total inner (GMRES) */
} orc_krylov;
int orc_solve_cg(orc_krylov *k);       /* :328-405 ; returns ioerr */
int orc_solve_pcg(orc_krylov *k);      /* :409-489 */
int orc_solve_gmres(orc_krylov *k);    /* :493-652 */
int orc_solve_pcgmres(orc_krylov *k);  /* :656-822 */

/* Dense-matrix operator helpers for brute-force tests: A column-major n x n,
 * residual = b - A s, preconditioner = diag(A)^-1. */
typedef struct { int n; const double *A; const double *b; } orc_dense_ctx;
void orc_dense_matvec(void *ctx, const double *in, double *out);
void orc_dense_residual(void *ctx, const double *in, double *out);
void orc_dense_jacobi(void *ctx, const double *in, double *out);

/* ---------------- Elliptic1D (examples/elliptic1d/parabola/localmods/Elliptic1D_Class.f90)
 * CG spectral-element d/dx(k du/dx) = s on [0,1], k = 1, s = 1, GLL nodes,
 * homogeneous Dirichlet, solved with Solve_ConjugateGradient.  Reproduces
 * quirk Q10 (Residual masks without GlobalSum, :457-503). */
int orc_elliptic1d_solve(int N, int nElems, int maxIters, double tol,
                         double *x /* (N+1)*nElems */, double *u /* unmasked solution */,
                         double *resi /* 0:maxIters */, int *ioerr);

#ifdef __cplusplus
}
#endif
#endif
