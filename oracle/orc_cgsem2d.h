/* orc_cgsem2d.h -- oracle (test infrastructure only): CG-SEM 2-D elliptic
 * solver of src/cgsem/CGsemElliptic_2D_Class.f90.  See orc_cgsem2d.c. */
#ifndef SELF_ORC_CGSEM2D_H
#define SELF_ORC_CGSEM2D_H

typedef struct {
    int N, nElems, nEdges, nNodes;
    int *edges;          /* 6 per edge, e2 = ORC_DIRICHLET on boundaries (set by connectivity) */
    int *nodeElemStart;  /* nNodes+1 */
    int *nodeElem;       /* element (0-based) of each list entry, increasing order per node */
    int *nodeLocal;      /* local corner 0..3 (SW, SE, NE, NW) */
    int *nodeType;       /* ORC_DIRICHLET or 1 */
    const double *D, *w; /* GLL derivative matrix (CG form) and weights */
    const double *dxds, *dxdp, *dyds, *dydp, *J, *fluxCoeff;
    double *solution, *source;
    double dirichletValue;
    double tolerance;
    int maxIters;
    double *resi;
    int resiLen;
    /* Q6 (Mask tests an uninitialised element ID, CGsemElliptic_2D_Class.f90:387):
     * 0 masks the secondary sides of interior edges (the evident intent);
     * 1 leaves them (what the reference does when compiled with amdflang -O2,
     * observed on its own build: the test never succeeds). */
    int q6AsBuilt;
} orc_cgsem2d;

/* CGsemPreconditioner_2D (src/cgsem/CGsemPreconditioner_2D_Class.f90) on the
 * linear mesh of ConstructLinearMeshFromOther_QuadMesh (src/geom/
 * QuadMeshClass.f90:1863-2096): nS*nP bilinear elements per element of the
 * original mesh with the original Gauss-Lobatto nodes as corners.  `lin` is
 * an N = 1 model on that mesh whose mask, unmask and global sum are the
 * preconditioner's (corner nodes only at N = 1). */
typedef struct {
    int nS, nEl, nLin, nNodes, maxIters;
    double tolerance;
    orc_cgsem2d lin;
    int *pcToOrig;   /* [(i + 2 j + 4 k) * 3 + c]: iS, iP, iEl (0-based) of linear element k's corner (i, j) */
    int *origToPC;   /* [((iS + n1 (iP + n1 e)) * 3 + c]: i, j, k */
    double *D1, *w1; /* N = 1 Gauss-Lobatto derivative matrix and weights */
    double *dxds, *dxdp, *dyds, *dydp, *J, *fluxCoeff;
    int *edges1, *nodeElemStart, *nodeElem, *nodeLocal, *nodeType;
    double *solution1, *source1;
} orc_cgsem2d_pc;

/* Build_CGsemPreconditioner_2D (:88-126) and the fluxCoeff map of
 * Build_CGsemElliptic_2D (:146-152); x, y are the original node positions
 * and fluxCoeff the original coefficient (the reference maps it at Build,
 * when it is still ONE). */
int orc_cgsem2d_pc_build(orc_cgsem2d_pc *pc, int N, int nEl, const double *x, const double *y,
                         const double *fluxCoeff, double tolerance, int maxIters);
void orc_cgsem2d_pc_free(orc_cgsem2d_pc *pc);
/* Solve_CGsemPreconditioner_2D (:664-746): x = (inner CG on the linear mesh)(b),
 * original element-local layout; returns ioerr and the inner iterations. */
int orc_cgsem2d_pc_solve(orc_cgsem2d_pc *pc, double *x, const double *b, int *iterations);
#include <stddef.h>
size_t orc_cgsem2d_pc_size(void);

int orc_cgsem2d_connectivity(orc_cgsem2d *c, const int *elemNodes, int nNodes);
void orc_cgsem2d_mask(const orc_cgsem2d *c, double *u);
void orc_cgsem2d_unmask(const orc_cgsem2d *c, double *u);
void orc_cgsem2d_global_sum(const orc_cgsem2d *c, double *u);
void orc_cgsem2d_matrix_action(const orc_cgsem2d *c, double *u, double *Au);
double orc_cgsem2d_dot(const orc_cgsem2d *c, double *u, double *v);
void orc_cgsem2d_residual(const orc_cgsem2d *c, double *r);
void orc_cgsem2d_set_dirichlet(const orc_cgsem2d *c);
void orc_cgsem2d_diagonal(const orc_cgsem2d *c, double *diag);
int orc_cgsem2d_solve_pcg(orc_cgsem2d *c, int precond, int *iterations);
/* precond 2: the preconditioner above (pc must be built) */
int orc_cgsem2d_solve_pcg_fem(orc_cgsem2d *c, orc_cgsem2d_pc *pc, int *iterations, int *pcIterations);

#endif
