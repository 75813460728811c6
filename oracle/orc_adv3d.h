/* orc_adv3d.h -- oracle (test infrastructure only): structured hex mesh,
 * 3-D nodal storage, the roll-off filter and 3-D advection
 * (src/highend/advection3d/Advection3DClass.f90).  See orc_adv3d.c. */
#ifndef SELF_ORC_ADV3D_H
#define SELF_ORC_ADV3D_H

/* Structured hex mesh on the unit cube, nX x nY x nZ elements, each face of
 * the reference element numbered as src/common/ModelFlags.f90:36-41
 * (south 1 = p-, east 2 = s+, north 3 = p+, west 4 = s-, bottom 5 = q-,
 * top 6 = q+), optional periodic pairing per direction, and the warp
 *   X = a + A s1 S, Y = b + A s2 S, Z = c + A s3 S,
 *   S = sin(2 pi a) sin(2 pi b) sin(2 pi c)
 * of the unit cube (s1, s2, s3 = +-1), then scaled by (xScale, yScale, zScale). */
typedef struct {
    int nX, nY, nZ, N, periodicX, periodicY, periodicZ;
    double xScale, yScale, zScale, warp;
    int warpSign[3];
    int nElems, nFaces; /* filled by orc_hexmesh_counts */
} orc_hexmesh_desc;

void orc_hexmesh_counts(orc_hexmesh_desc *d);
/* faces: 9 ints per face (e1, s1, e2, s2, iStart, jStart, iInc, jInc, swap),
 * elements 1-based, e2 = NO_NORMAL_FLOW on boundaries.  For the secondary
 * the node (k, l) matching the primary's (i, j) is
 *   swap = 0: (iStart + i iInc, jStart + j jInc)
 *   swap = 1: (jStart + j jInc, iStart + i iInc)
 * (the loop of Advection3DClass.f90:1286-1296).
 * Volume arrays (N+1)^3 per element, i fastest; cov: 9 covariant metric
 * terms per node in the order dxds dxdp dxdq dyds dydp dydq dzds dzdp dzdq;
 * face arrays per element [side][node] with node = a + (N+1) b, where
 * (a, b) = (p, q) on west/east, (s, q) on south/north, (s, p) on
 * bottom/top; nHat 3 components per face node, nLen, and bpos = (x, y, z)
 * of each face node. */
int orc_hexmesh_build(const orc_hexmesh_desc *d, const double *nodes, int *faces, double *x, double *y,
                      double *z, double *cov, double *J, double *nHat, double *nLen, double *bpos);

/* RollOffFilter3D (src/filters/RollOffFilter3D_Class.f90:87-180): the dense
 * (N+1)^3 x (N+1)^3 filter matrix, column-major, built as the reference
 * does from Legendre modes on the Gauss nodes with cutoff nc. */
void orc_rolloff_filter3d_build(int N, int nc, const double *nodes, const double *w, double *F);
/* ApplyFilter (:273-292): fSol = F sol over the Map3Dto1D ordering */
void orc_rolloff_filter3d_apply(int N, const double *F, const double *sol, double *fsol);

typedef struct {
    int N, nq, nElems, nFaces, useFilter;
    double dt;
    const double *dMat, *qWeight, *lagLeft, *lagRight, *F;
    const int *faces;
    const double *cov, *J, *nHat, *nLen, *bpos;
    double *c, *tend, *G, *u, *v, *w, *relax, *rFac;
    double *bc, *fc, *bu, *bv, *bw;      /* traces: [node + n1^2 (f + nq (side + 6 e))] */
    double *E1, *E2, *lim;               /* E[e nq + f], lim[(m-1) nq nE + e nq + f] */
    int nFiltered;                       /* elements x tracers replaced by the filter, last call */
} orc_adv3d;

/* DoTheAdaptiveFiltering (:1125-1233) on every element, stage m (1..3) */
void orc_adv3d_adaptive_filter(orc_adv3d *m, int stage);
/* GlobalTimeDerivative (:1049-1121): filter, traces, EdgeFlux, MappedTimeDerivative */
void orc_adv3d_global_time_derivative(orc_adv3d *m, double t, int stage);
/* ForwardStepRK3 (:1000-1045) */
void orc_adv3d_forward_step_rk3(orc_adv3d *m, double tn);

#endif
