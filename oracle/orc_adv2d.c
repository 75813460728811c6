/*
 * orc_adv2d.c -- oracle (test infrastructure only): 2-D advection-diffusion
 * DGSEM with a BR1-type gradient, restated from
 *   src/highend/advection-diffusion2d/AdvectionClass.f90
 *     CalculateConcentrationGradient :1123-1264, ForwardStepRK3 :1268-1312,
 *     GlobalTimeDerivative :1316-1378, EdgeFlux :1382-1478,
 *     EdgeGradientFlux :1482-1563, MappedTimeDerivative :1567-1723,
 *     RiemannSolver :1756-1787, XFlux/YFlux :1791-1825, Source :1829-1847,
 *     GetExternalState :1851-1886.
 * Velocity is a separate 2-field array (the reference keeps it as vars 1,2
 * of a tracer-sized storage, quirk Q4, which needs nTracers >= 2).
 * Layouts (element e 0-based, tracer f):
 *   c, tend, G, relax, rFac, dcdx, dcdy [i + n1*(j + n1*(f + nq*e))]
 *   u, v                               [i + n1*(j + n1*e)]
 *   bc, bdx, bdy (traces)              [node + n1*(f + nq*(side + 4*e))]
 *   fc (flux), fgx, fgy (gradient fluxes) same as traces
 *   bu, bv (velocity traces)           [node + n1*(side + 4*e)]
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"
#include "orc_adv2d.h"

static const double rk3_a[3] = {0.0, -5.0 / 9.0, -153.0 / 128.0};
static const double rk3_b[3] = {0.0, 1.0 / 3.0, 3.0 / 4.0};
static const double rk3_g[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};

#define ND(i, j, f, e) ((i) + n1 * ((j) + n1 * ((f) + nq * (e))))
#define TR(k, f, s, e) ((k) + n1 * ((f) + nq * ((s) + 4 * (e))))

static void traces(int N, int nq, int nE, const double *lW, const double *lE, const double *lS, const double *lN,
                   const double *q, double *out)
{
    int n1 = N + 1;
    for (int e = 0; e < nE; ++e)
        for (int f = 0; f < nq; ++f) {
            for (int k = 0; k <= N; ++k) {
                double w = 0, ea = 0, s = 0, no = 0;
                for (int m = 0; m <= N; ++m) {
                    w += lW[m] * q[ND(m, k, f, e)];
                    ea += lE[m] * q[ND(m, k, f, e)];
                    s += lS[m] * q[ND(k, m, f, e)];
                    no += lN[m] * q[ND(k, m, f, e)];
                }
                out[TR(k, f, 3, e)] = w;
                out[TR(k, f, 1, e)] = ea;
                out[TR(k, f, 0, e)] = s;
                out[TR(k, f, 2, e)] = no;
            }
        }
}

/* GetExternalState :1851-1886 (time t) */
static double external(int bcFlag, int f, double t)
{
    if (bcFlag == ORC_INFLOW_ONE && f == 0) return tanh(t);
    if (bcFlag == ORC_INFLOW_TWO && f == 1) return tanh(t);
    return 0.0;
}

static void dgsd(int N, const double *dMat, const double *qW, double lFlux, double rFlux, const double *intFlux,
                 const double *lL, const double *lR, double *tend)
{
    int n1 = N + 1;
    for (int k = 0; k <= N; ++k) {
        double acc = 0.0;
        for (int j = 0; j <= N; ++j) acc += dMat[k + n1 * j] * intFlux[j];
        tend[k] = acc + (rFlux * lR[k] + lFlux * lL[k]) / qW[k];
    }
}

/* EdgeGradientFlux :1482-1563 */
static void edge_gradient_flux(orc_adv2d *m, double t)
{
    int N = m->N, n1 = N + 1, nq = m->nq;
    for (int ie = 0; ie < m->nEdges; ++ie) {
        const int *ed = m->edges + 6 * ie;
        int e1 = ed[0] - 1, s1 = ed[1] - 1, e2 = ed[2], s2 = abs(ed[3]) - 1, start = ed[4], inc = ed[5];
        int k = start - inc;
        for (int iNode = 0; iNode <= N; ++iNode) {
            int nb = iNode + n1 * (s1 + 4 * e1);
            double nx = m->nHat[2 * nb], ny = m->nHat[2 * nb + 1], len = m->nLen[nb];
            for (int f = 0; f < nq; ++f) {
                double in = m->bc[TR(iNode, f, s1, e1)];
                double ex = (e2 > 0) ? m->bc[TR(k, f, s2, e2 - 1)] : external(e2, f, t);
                double fx = 0.5 * (in + ex) * nx * len;
                double fy = 0.5 * (in + ex) * ny * len;
                m->fgx[TR(iNode, f, s1, e1)] = fx;
                m->fgy[TR(iNode, f, s1, e1)] = fy;
                if (e2 > 0) {
                    m->fgx[TR(k, f, s2, e2 - 1)] = -fx;
                    m->fgy[TR(k, f, s2, e2 - 1)] = -fy;
                }
            }
            if (e2 > 0) k += inc;
        }
    }
}

/* CalculateConcentrationGradient :1123-1264 */
static void gradient(orc_adv2d *m)
{
    int N = m->N, n1 = N + 1, nq = m->nq;
    double sC[2][32], pC[2][32], der[32];
    for (int e = 0; e < m->nElems; ++e)
        for (int f = 0; f < nq; ++f) {
            for (int iP = 0; iP <= N; ++iP) {
                for (int iS = 0; iS <= N; ++iS) {
                    int g = iS + n1 * (iP + n1 * e);
                    double c = m->c[ND(iS, iP, f, e)];
                    sC[0][iS] = m->dydp[g] * c - m->dxdp[g] * 0.0;
                    sC[1][iS] = m->dydp[g] * 0.0 - m->dxdp[g] * c;
                }
                double fR[2] = {m->fgx[TR(iP, f, 1, e)], m->fgy[TR(iP, f, 1, e)]};
                double fL[2] = {m->fgx[TR(iP, f, 3, e)], m->fgy[TR(iP, f, 3, e)]};
                for (int d = 0; d < 2; ++d) {
                    dgsd(N, m->dMatS, m->qWeight, fL[d], fR[d], sC[d], m->lagWest, m->lagEast, der);
                    for (int iS = 0; iS <= N; ++iS) (d ? m->dcdy : m->dcdx)[ND(iS, iP, f, e)] = der[iS];
                }
            }
            for (int iS = 0; iS <= N; ++iS) {
                for (int iP = 0; iP <= N; ++iP) {
                    int g = iS + n1 * (iP + n1 * e);
                    double c = m->c[ND(iS, iP, f, e)];
                    pC[0][iP] = -m->dyds[g] * c + m->dxds[g] * 0.0;
                    pC[1][iP] = -m->dyds[g] * 0.0 + m->dxds[g] * c;
                }
                double fR[2] = {m->fgx[TR(iS, f, 2, e)], m->fgy[TR(iS, f, 2, e)]};
                double fL[2] = {m->fgx[TR(iS, f, 0, e)], m->fgy[TR(iS, f, 0, e)]};
                for (int d = 0; d < 2; ++d) {
                    dgsd(N, m->dMatP, m->qWeight, fL[d], fR[d], pC[d], m->lagSouth, m->lagNorth, der);
                    for (int iP = 0; iP <= N; ++iP) {
                        double J = m->J[iS + n1 * (iP + n1 * e)];
                        double *a = d ? m->dcdy : m->dcdx;
                        a[ND(iS, iP, f, e)] = (a[ND(iS, iP, f, e)] + der[iP]) / J;
                    }
                }
            }
        }
}

/* EdgeFlux :1382-1478 with RiemannSolver :1756-1787 */
static void edge_flux(orc_adv2d *m, double t)
{
    int N = m->N, n1 = N + 1, nq = m->nq;
    for (int ie = 0; ie < m->nEdges; ++ie) {
        const int *ed = m->edges + 6 * ie;
        int e1 = ed[0] - 1, s1 = ed[1] - 1, e2 = ed[2], s2 = abs(ed[3]) - 1, start = ed[4], inc = ed[5];
        int k = start - inc;
        for (int iNode = 0; iNode <= N; ++iNode) {
            int nb = iNode + n1 * (s1 + 4 * e1);
            double nx = m->nHat[2 * nb], ny = m->nHat[2 * nb + 1], len = m->nLen[nb];
            double u = m->bu[iNode + n1 * (s1 + 4 * e1)], v = m->bv[iNode + n1 * (s1 + 4 * e1)];
            double uN = u * nx + v * ny;
            for (int f = 0; f < nq; ++f) {
                double in = m->bc[TR(iNode, f, s1, e1)];
                double dx = m->bdx[TR(iNode, f, s1, e1)], dy = m->bdy[TR(iNode, f, s1, e1)];
                double ex;
                if (e2 > 0) {
                    ex = m->bc[TR(k, f, s2, e2 - 1)];
                    dx = (dx + m->bdx[TR(k, f, s2, e2 - 1)]) * 0.5;
                    dy = (dy + m->bdy[TR(k, f, s2, e2 - 1)]) * 0.5;
                } else {
                    ex = external(e2, f, t);
                }
                double jump = ex - in;
                double fl = 0.5 * (uN * (ex + in) - fabs(uN) * jump);
                double dcdn = dx * nx + dy * ny;
                fl = (fl - m->kappa * dcdn) * len;
                m->fc[TR(iNode, f, s1, e1)] = fl;
                if (e2 > 0) m->fc[TR(k, f, s2, e2 - 1)] = -fl;
            }
            if (e2 > 0) k += inc;
        }
    }
}

/* MappedTimeDerivative :1567-1723 */
static void mapped(orc_adv2d *m)
{
    int N = m->N, n1 = N + 1, nq = m->nq;
    double sC[32], pC[32], der[32];
    for (int e = 0; e < m->nElems; ++e)
        for (int f = 0; f < nq; ++f) {
            for (int iP = 0; iP <= N; ++iP) {
                for (int iS = 0; iS <= N; ++iS) {
                    int g = iS + n1 * (iP + n1 * e);
                    double c = m->c[ND(iS, iP, f, e)];
                    double xF = c * m->u[g] - m->kappa * m->dcdx[ND(iS, iP, f, e)];
                    double yF = c * m->v[g] - m->kappa * m->dcdy[ND(iS, iP, f, e)];
                    sC[iS] = m->dydp[g] * xF - m->dxdp[g] * yF;
                }
                dgsd(N, m->dMatS, m->qWeight, m->fc[TR(iP, f, 3, e)], m->fc[TR(iP, f, 1, e)], sC, m->lagWest,
                     m->lagEast, der);
                for (int iS = 0; iS <= N; ++iS) m->tend[ND(iS, iP, f, e)] = -der[iS];
            }
            for (int iS = 0; iS <= N; ++iS) {
                for (int iP = 0; iP <= N; ++iP) {
                    int g = iS + n1 * (iP + n1 * e);
                    double c = m->c[ND(iS, iP, f, e)];
                    double xF = c * m->u[g] - m->kappa * m->dcdx[ND(iS, iP, f, e)];
                    double yF = c * m->v[g] - m->kappa * m->dcdy[ND(iS, iP, f, e)];
                    pC[iP] = -m->dyds[g] * xF + m->dxds[g] * yF;
                }
                dgsd(N, m->dMatP, m->qWeight, m->fc[TR(iS, f, 0, e)], m->fc[TR(iS, f, 2, e)], pC, m->lagSouth,
                     m->lagNorth, der);
                for (int iP = 0; iP <= N; ++iP) {
                    double J = m->J[iS + n1 * (iP + n1 * e)];
                    m->tend[ND(iS, iP, f, e)] = (m->tend[ND(iS, iP, f, e)] - der[iP]) / J;
                }
            }
            if (m->relax)
                for (int iS = 0; iS <= N; ++iS)
                    for (int iP = 0; iP <= N; ++iP) {
                        int id = ND(iS, iP, f, e);
                        m->tend[id] = m->tend[id] + (m->relax[id] - m->c[id]) * m->rFac[id];
                    }
        }
}

void orc_adv2d_global_time_derivative(orc_adv2d *m, double t)
{
    int N = m->N, n1 = N + 1;
    traces(N, m->nq, m->nElems, m->lagWest, m->lagEast, m->lagSouth, m->lagNorth, m->c, m->bc);
    /* velocity traces (CalculateVelocityAtBoundaries) */
    for (int e = 0; e < m->nElems; ++e)
        for (int k = 0; k <= N; ++k) {
            double w[2] = {0, 0}, ea[2] = {0, 0}, s[2] = {0, 0}, no[2] = {0, 0};
            for (int mm = 0; mm <= N; ++mm) {
                const double *vv[2] = {m->u, m->v};
                for (int d = 0; d < 2; ++d) {
                    w[d] += m->lagWest[mm] * vv[d][mm + n1 * (k + n1 * e)];
                    ea[d] += m->lagEast[mm] * vv[d][mm + n1 * (k + n1 * e)];
                    s[d] += m->lagSouth[mm] * vv[d][k + n1 * (mm + n1 * e)];
                    no[d] += m->lagNorth[mm] * vv[d][k + n1 * (mm + n1 * e)];
                }
            }
            double *bb[2] = {m->bu, m->bv};
            for (int d = 0; d < 2; ++d) {
                bb[d][k + n1 * (3 + 4 * e)] = w[d];
                bb[d][k + n1 * (1 + 4 * e)] = ea[d];
                bb[d][k + n1 * (0 + 4 * e)] = s[d];
                bb[d][k + n1 * (2 + 4 * e)] = no[d];
            }
        }
    edge_gradient_flux(m, t);
    gradient(m);
    traces(N, m->nq, m->nElems, m->lagWest, m->lagEast, m->lagSouth, m->lagNorth, m->dcdx, m->bdx);
    traces(N, m->nq, m->nElems, m->lagWest, m->lagEast, m->lagSouth, m->lagNorth, m->dcdy, m->bdy);
    edge_flux(m, t);
    mapped(m);
}

void orc_adv2d_forward_step_rk3(orc_adv2d *m, double tn)
{
    long n = (long)(m->N + 1) * (m->N + 1) * m->nq * m->nElems;
    memset(m->G, 0, sizeof(double) * n);
    for (int st = 0; st < 3; ++st) {
        orc_adv2d_global_time_derivative(m, tn + rk3_b[st] * m->dt);
        for (long i = 0; i < n; ++i) {
            m->G[i] = rk3_a[st] * m->G[i] + m->tend[i];
            m->c[i] = m->c[i] + rk3_g[st] * m->dt * m->G[i];
        }
    }
}
