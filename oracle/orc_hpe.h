/* orc_hpe.h -- oracle (test infrastructure only): the hydrostatic primitive
 * equations of src/highend/hydrostaticprimitives on a stacked hex mesh.
 * See orc_hpe.c for the routines restated and the gaps of the reference
 * class that the restatement fills. */
#ifndef SELF_ORC_HPE_H
#define SELF_ORC_HPE_H

#define ORC_SEA_FLOOR (-105) /* src/common/ModelFlags.f90:85 */

/* Variables: u 0, v 1, b 2 (prognostic, nPrognostic = 3), w 3, p 4
 * (diagnosed), HydrostaticPrimitivesClass.f90:205-209.
 *
 * Mesh: the structured hex mesh of orc_hexmesh_build with nZ = nLayers and
 * no periodicity in z; element e = hE + nHElems * layer is the reference's
 * stackMap(hE, layer) (StackedHexMeshClass.f90:337).  Boundary faces
 * whose primary side is the bottom (5) are SEA_FLOOR, every other boundary
 * face NO_NORMAL_FLOW (ConstructFaces, StackedHexMeshClass.f90:617-621).
 *
 * Layouts: sol and bg [e][var][k][j][i] (5 variables), traces and boundary
 * fluxes [e][side][var][node] with node = a + (N+1) b as orc_hexmesh_build
 * numbers face nodes, tend and G [e][eq][k][j][i] (3 equations), divU
 * [e][k][j][i], bottomW [hE][j][i]. */
typedef struct {
    int N, M;            /* DG degree; column-quadrature degree M = (N+1)/2 + 1 */
    int nHElems, nLayers, nElems, nFaces;
    int diagnoseP;       /* 1: DiagnoseP runs inside GlobalTimeDerivative */
    double dt;
    const double *nodes, *bWeights, *dMat, *qWeight, *lagLeft, *lagRight; /* Gauss DG storage */
    const double *intNodes, *intWeights;                                   /* Gauss-Lobatto (0:M) */
    const int *faces;
    const double *cov, *J, *nHat, *nLen, *bpos;
    double *sol, *bg, *bsol, *bflux, *tend, *G, *divU, *bottomW;
} orc_hpe;

/* CalculateSolutionAtBoundaries (:1363-1374), all five variables */
void orc_hpe_traces(orc_hpe *m);
/* CalculateDivU (HydrostaticPrimitivesClass_implicit.f90:2520-2691) */
void orc_hpe_divu(orc_hpe *m);
/* DiagnoseP (:1739-1822) on every stack */
void orc_hpe_diagnose_p(orc_hpe *m);
/* CalculateBottomVerticalVelocity (:1916-1945) + DiagnoseW (:1826-1912) on every stack */
void orc_hpe_diagnose_w(orc_hpe *m);
/* EdgeFlux (:2088-2183) with RiemannSolver (:2738-2778) and GetExternalState (:2856-2898) */
void orc_hpe_edge_flux(orc_hpe *m);
/* MappedTimeDerivative (:2187-2400) with X/Y/ZFlux (:2782-2832) and Source (:2836-2852) */
void orc_hpe_mapped(orc_hpe *m);
/* GlobalTimeDerivative (:1428-1506): traces, divU, [P], W, traces, EdgeFlux, MappedTimeDerivative */
void orc_hpe_global_time_derivative(orc_hpe *m);
/* ForwardStepRK3 (HydrostaticPrimitivesClass_implicit.f90:1422-1471) without the pressure correction */
void orc_hpe_forward_step_rk3(orc_hpe *m);
/* ForwardStepAB2 (:1378-1424) without the barotropic step; G holds the previous tendency */
void orc_hpe_forward_step_ab2(orc_hpe *m);

#endif
