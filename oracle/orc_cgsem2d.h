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

#endif
