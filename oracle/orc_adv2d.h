/* orc_adv2d.h -- oracle (test infrastructure only): 2-D advection-diffusion
 * (src/highend/advection-diffusion2d/AdvectionClass.f90).  See orc_adv2d.c. */
#ifndef SELF_ORC_ADV2D_H
#define SELF_ORC_ADV2D_H

typedef struct {
    int N, nq, nElems, nEdges;
    double kappa, dt;
    const double *dMatS, *dMatP, *qWeight, *lagWest, *lagEast, *lagSouth, *lagNorth;
    const int *edges;
    const double *dxds, *dxdp, *dyds, *dydp, *J, *nHat, *nLen;
    double *c, *tend, *G, *dcdx, *dcdy, *u, *v, *relax, *rFac;
    double *bc, *bdx, *bdy, *fc, *fgx, *fgy, *bu, *bv;
} orc_adv2d;

void orc_adv2d_global_time_derivative(orc_adv2d *m, double t);
void orc_adv2d_forward_step_rk3(orc_adv2d *m, double tn);

#endif
