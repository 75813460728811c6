/*
 * orc_sw2d.c -- oracle (test infrastructure only): conservative 2-D shallow
 * water DGSEM restated from
 *   src/highend/shallowwater/ConservativeShallowWaterClass.f90
 *   src/dgsem/DGSEMSolutionStorageClass_2D.f90:749-802
 *   src/common/ConstantsDictionary.f90:66-68 (low-storage RK3 coefficients)
 * with the one-token fix of quirk Q2 (RiemannSolver's undeclared g -> locG),
 * and, behind m->form = 1, the class as the examples ship it
 *   examples/shallowwater/planewave/localmods/ConservativeShallowWaterClass.f90
 * (the form the compiled reference runs: the src copy does not compile).
 * Layouts are documented in oracle.h.  Fortran evaluates a*b*c*d left to
 * right; the C below keeps that association explicitly.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "oracle.h"

static const double rk3_a[3] = {0.0, -5.0 / 9.0, -153.0 / 128.0};
static const double rk3_b[3] = {0.0, 1.0 / 3.0, 3.0 / 4.0};
static const double rk3_g[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};

#define NQ 3

/* :1555-1604; form 1: examples/shallowwater/planewave/localmods/
 * ConservativeShallowWaterClass.f90:1523-1566 (only the wave speed differs) */
void orc_sw2d_riemann_form(const double *inS, const double *outS, const double *nHat,
                           double locG, int form, double *numFlux)
{
    double jump[NQ], aS[NQ];
    for (int k = 0; k < NQ; ++k) jump[k] = outS[k] - inS[k];
    double uL = inS[0], vL = inS[1], pL = inS[2];
    double uR = outS[0], vR = outS[1], pR = outS[2];
    double uOut = uR * nHat[0] + vR * nHat[1];
    double uIn = uL * nHat[0] + vL * nHat[1];
    double fac, t;
    if (form == 1) {
        double sR = sqrt(uOut * uOut + 2.0 * locG * (pR * pR * pR));
        double sL = sqrt(uIn * uIn + 2.0 * locG * (pL * pL * pL));
        fac = fabs(0.5 * (uOut + sR) / pR);
        t = fabs(0.5 * (uOut - sR) / pR); if (t > fac) fac = t;
        t = fabs(0.5 * (uIn + sL) / pL); if (t > fac) fac = t;
        t = fabs(0.5 * (uIn - sL) / pL); if (t > fac) fac = t;
    } else {
        double cR = sqrt(locG * pR), cL = sqrt(locG * pL);
        fac = fabs(uOut / pR + cR);
        t = fabs(uOut / pR - cR); if (t > fac) fac = t;
        t = fabs(uIn / pL + cL); if (t > fac) fac = t;
        t = fabs(uIn / pL - cL); if (t > fac) fac = t;
    }
    aS[0] = (uOut / pR) * uR + 0.5 * locG * pR * pR * nHat[0] + (uIn / pL) * uL + 0.5 * locG * pL * pL * nHat[0];
    aS[1] = (uOut / pR) * vR + 0.5 * locG * pR * pR * nHat[1] + (uIn / pL) * vL + 0.5 * locG * pL * pL * nHat[1];
    aS[2] = uOut + uIn;
    for (int k = 0; k < NQ; ++k) numFlux[k] = 0.5 * (aS[k] - fac * jump[k]);
}

void orc_sw2d_riemann(const double *inS, const double *outS, const double *nHat,
                      double locG, double *numFlux)
{
    orc_sw2d_riemann_form(inS, outS, nHat, locG, 0, numFlux);
}

/* :1698-1729 */
void orc_sw2d_external_state(int bcFlag, const double *n, const double *in, double *ex)
{
    if (bcFlag == ORC_NO_NORMAL_FLOW) {
        ex[0] = (n[1] * n[1] - n[0] * n[0]) * in[0] - 2.0 * n[0] * n[1] * in[1];
        ex[1] = (n[0] * n[0] - n[1] * n[1]) * in[1] - 2.0 * n[0] * n[1] * in[0];
        ex[2] = in[2];
    } else {
        ex[0] = ex[1] = ex[2] = 0.0; /* RADIATION and unknown flags */
    }
}

static inline int IDX(int n1, int i, int j, int f, int nf, int e) { return i + n1 * (j + n1 * (f + nf * e)); }
static inline int BDX(int n1, int node, int f, int nf, int side, int e) { return node + n1 * (f + nf * (side + 4 * e)); }

static void boundaries_one(int n1, const double *lW, const double *lE, const double *lS,
                           const double *lN, const double *sol, double *bs, int nf, int e)
{
    int N = n1 - 1;
    for (int iP = 0; iP <= N; ++iP)
        for (int f = 0; f < nf; ++f) {
            double w = 0.0, ea = 0.0;
            for (int k = 0; k <= N; ++k) {
                w += lW[k] * sol[IDX(n1, k, iP, f, nf, e)];
                ea += lE[k] * sol[IDX(n1, k, iP, f, nf, e)];
            }
            bs[BDX(n1, iP, f, nf, 3, e)] = w;
            bs[BDX(n1, iP, f, nf, 1, e)] = ea;
        }
    for (int iS = 0; iS <= N; ++iS)
        for (int f = 0; f < nf; ++f) {
            double s = 0.0, no = 0.0;
            for (int k = 0; k <= N; ++k) {
                s += lS[k] * sol[IDX(n1, iS, k, f, nf, e)];
                no += lN[k] * sol[IDX(n1, iS, k, f, nf, e)];
            }
            bs[BDX(n1, iS, f, nf, 0, e)] = s;
            bs[BDX(n1, iS, f, nf, 2, e)] = no;
        }
}

void orc_sw2d_solution_at_boundaries(orc_sw2d *m)
{
    int n1 = m->N + 1;
    for (int e = 0; e < m->nElems; ++e)
        boundaries_one(n1, m->lagWest, m->lagEast, m->lagSouth, m->lagNorth, m->sol, m->bsol, NQ, e);
}

static void edge_flux_one(orc_sw2d *m, int iEdge, double tn)
{
    (void)tn;
    int n1 = m->N + 1, N = m->N;
    const int *ed = m->edges + 6 * iEdge;
    int e1 = ed[0] - 1, s1 = ed[1] - 1, e2 = ed[2], s2 = abs(ed[3]) - 1;
    int start = ed[4], inc = ed[5];
    double inS[NQ], exS[NQ], flux[NQ], nh[2];
    if (e2 > 0) {
        e2 -= 1;
        int k = start - inc;
        for (int iNode = 0; iNode <= N; ++iNode) {
            for (int f = 0; f < NQ; ++f) {
                inS[f] = m->bsol[BDX(n1, iNode, f, NQ, s1, e1)];
                exS[f] = m->bsol[BDX(n1, k, f, NQ, s2, e2)];
            }
            int nb = iNode + n1 * (s1 + 4 * e1);
            nh[0] = m->nHat[2 * nb]; nh[1] = m->nHat[2 * nb + 1];
            double len = m->nLen[nb];
            orc_sw2d_riemann_form(inS, exS, nh, m->g, m->form, flux);
            for (int f = 0; f < NQ; ++f) {
                double fl = flux[f] * len;
                m->bflux[BDX(n1, iNode, f, NQ, s1, e1)] = fl;
                m->bflux[BDX(n1, k, f, NQ, s2, e2)] = -fl;
            }
            k = k + inc;
        }
    } else {
        for (int iNode = 0; iNode <= N; ++iNode) {
            int nb = iNode + n1 * (s1 + 4 * e1);
            nh[0] = m->nHat[2 * nb]; nh[1] = m->nHat[2 * nb + 1];
            double len = m->nLen[nb];
            for (int f = 0; f < NQ; ++f) inS[f] = m->bsol[BDX(n1, iNode, f, NQ, s1, e1)];
            orc_sw2d_external_state(e2, nh, inS, exS);
            orc_sw2d_riemann_form(inS, exS, nh, m->g, m->form, flux);
            for (int f = 0; f < NQ; ++f) m->bflux[BDX(n1, iNode, f, NQ, s1, e1)] = flux[f] * len;
        }
    }
}

void orc_sw2d_edge_flux(orc_sw2d *m, double tn)
{
    for (int iEdge = 0; iEdge < m->nEdges; ++iEdge) edge_flux_one(m, iEdge, tn);
}

/* :1526-1551 DGSystemDerivative */
static void dg_system_derivative(int N, const double *dMat, const double *qW, double lFlux,
                                 double rFlux, const double *intFlux, const double *lL,
                                 const double *lR, double *tend)
{
    int n1 = N + 1;
    for (int k = 0; k <= N; ++k) {
        double acc = 0.0;
        for (int j = 0; j <= N; ++j) acc += dMat[k + n1 * j] * intFlux[j];
        tend[k] = acc;
    }
    for (int k = 0; k <= N; ++k) tend[k] = tend[k] + (rFlux * lR[k] + lFlux * lL[k]) / qW[k];
}

static void xflux(double g, const double *s, double *fx) /* :1608-1631 */
{
    fx[0] = s[0] * s[0] / s[2] + g * 0.5 * s[2] * s[2];
    fx[1] = s[0] * s[1] / s[2];
    fx[2] = s[0];
}
static void yflux(double g, const double *s, double *fy) /* :1635-1658 */
{
    fy[0] = s[1] * s[0] / s[2];
    fy[1] = s[1] * s[1] / s[2] + g * 0.5 * s[2] * s[2];
    fy[2] = s[1];
}

static void mapped_one(orc_sw2d *m, int e, double tn)
{
    (void)tn;
    int N = m->N, n1 = N + 1;
    double sC[16][NQ], pC[16][NQ], der[16], line[16];
    double *tend = m->tend;
    double s[NQ], xF[NQ], yF[NQ], fL[NQ], fR[NQ];
    for (int iP = 0; iP <= N; ++iP) {
        for (int iS = 0; iS <= N; ++iS) {
            int g0 = iS + n1 * (iP + n1 * e);
            for (int f = 0; f < NQ; ++f) s[f] = m->sol[IDX(n1, iS, iP, f, NQ, e)];
            xflux(m->g, s, xF); yflux(m->g, s, yF);
            for (int f = 0; f < NQ; ++f) sC[iS][f] = m->dydp[g0] * xF[f] - m->dxdp[g0] * yF[f];
        }
        for (int f = 0; f < NQ; ++f) {
            fL[f] = m->bflux[BDX(n1, iP, f, NQ, 3, e)];
            fR[f] = m->bflux[BDX(n1, iP, f, NQ, 1, e)];
        }
        for (int f = 0; f < NQ; ++f) {
            for (int k = 0; k <= N; ++k) line[k] = sC[k][f];
            dg_system_derivative(N, m->dMatS, m->qWeightS, fL[f], fR[f], line, m->lagWest, m->lagEast, der);
            for (int iS = 0; iS <= N; ++iS) tend[IDX(n1, iS, iP, f, NQ, e)] = -der[iS];
        }
    }
    for (int iS = 0; iS <= N; ++iS) {
        for (int iP = 0; iP <= N; ++iP) {
            int g0 = iS + n1 * (iP + n1 * e);
            for (int f = 0; f < NQ; ++f) s[f] = m->sol[IDX(n1, iS, iP, f, NQ, e)];
            xflux(m->g, s, xF); yflux(m->g, s, yF);
            for (int f = 0; f < NQ; ++f) pC[iP][f] = -m->dyds[g0] * xF[f] + m->dxds[g0] * yF[f];
        }
        for (int f = 0; f < NQ; ++f) {
            fL[f] = m->bflux[BDX(n1, iS, f, NQ, 0, e)];
            fR[f] = m->bflux[BDX(n1, iS, f, NQ, 2, e)];
        }
        for (int f = 0; f < NQ; ++f) {
            for (int k = 0; k <= N; ++k) line[k] = pC[k][f];
            dg_system_derivative(N, m->dMatP, m->qWeightP, fL[f], fR[f], line, m->lagSouth, m->lagNorth, der);
            for (int iP = 0; iP <= N; ++iP) {
                double Jv = m->J[iS + n1 * (iP + n1 * e)];
                int id = IDX(n1, iS, iP, f, NQ, e);
                tend[id] = (tend[id] - der[iP]) / Jv;
            }
        }
    }
    /* Source :1662-1694 */
    for (int iS = 0; iS <= N; ++iS)
        for (int iP = 0; iP <= N; ++iP) {
            for (int f = 0; f < NQ; ++f) s[f] = m->sol[IDX(n1, iS, iP, f, NQ, e)];
            double r[NQ] = {0, 0, 0}, ts = 0.0;
            if (m->relax) {
                for (int f = 0; f < NQ; ++f) r[f] = m->relax[IDX(n1, iS, iP, f, 4, e)];
                ts = m->relax[IDX(n1, iS, iP, 3, 4, e)];
            }
            double vort = m->plv[iS + n1 * (iP + n1 * e)];
            double h = m->bathy[IDX(n1, iS, iP, 0, 3, e)];
            double dhdx = m->bathy[IDX(n1, iS, iP, 1, 3, e)];
            double dhdy = m->bathy[IDX(n1, iS, iP, 2, 3, e)];
            double lDrag = m->linearDrag, g = m->g;
            double q0, q1, q2;
            if (m->form == 1) { /* localmods :1649-1655 */
                q0 = g * s[2] * dhdx + vort * s[1];
                q1 = g * s[2] * dhdy - vort * s[0];
                q2 = 0.0;
            } else {
                q0 = vort * s[1] + g * s[2] * dhdx - lDrag * s[0] / h - (s[0] - r[0]) * ts;
                q1 = -vort * s[0] + g * s[2] * dhdy - lDrag * s[1] / h - (s[1] - r[1]) * ts;
                q2 = -(s[2] - r[2]) * ts;
            }
            tend[IDX(n1, iS, iP, 0, NQ, e)] += q0;
            tend[IDX(n1, iS, iP, 1, NQ, e)] += q1;
            tend[IDX(n1, iS, iP, 2, NQ, e)] += q2;
        }
}

void orc_sw2d_mapped_time_derivative(orc_sw2d *m, double tn)
{
    for (int e = 0; e < m->nElems; ++e) mapped_one(m, e, tn);
}

void orc_sw2d_global_time_derivative(orc_sw2d *m, double tn)
{
    orc_sw2d_solution_at_boundaries(m);
    orc_sw2d_edge_flux(m, tn);
    orc_sw2d_mapped_time_derivative(m, tn);
}

static void gtd_threads(orc_sw2d *m, double t, int nThreads)
{
    int n1 = m->N + 1;
    (void)n1;
#pragma omp parallel num_threads(nThreads)
    {
#pragma omp for schedule(static)
        for (int e = 0; e < m->nElems; ++e)
            boundaries_one(m->N + 1, m->lagWest, m->lagEast, m->lagSouth, m->lagNorth, m->sol, m->bsol, NQ, e);
#pragma omp for schedule(static)
        for (int ie = 0; ie < m->nEdges; ++ie) edge_flux_one(m, ie, t);
#pragma omp for schedule(static)
        for (int e = 0; e < m->nElems; ++e) mapped_one(m, e, t);
    }
}

static void rk_update(orc_sw2d *m, int stage, int nThreads)
{
    long n = (long)(m->N + 1) * (m->N + 1) * NQ * m->nElems;
    double a = rk3_a[stage], gdt = rk3_g[stage] * m->dt;
#pragma omp parallel for num_threads(nThreads) schedule(static)
    for (long i = 0; i < n; ++i) {
        m->G[i] = a * m->G[i] + m->tend[i];
        m->sol[i] = m->sol[i] + gdt * m->G[i];
    }
}

/* :1173-1218 */
void orc_sw2d_forward_step_rk3(orc_sw2d *m, double tn)
{
    long n = (long)(m->N + 1) * (m->N + 1) * NQ * m->nElems;
    memset(m->G, 0, sizeof(double) * n);
    for (int st = 0; st < 3; ++st) {
        double t = tn + rk3_b[st] * m->dt;
        orc_sw2d_global_time_derivative(m, t);
        rk_update(m, st, 1);
    }
}

void orc_sw2d_integrate(orc_sw2d *m, double t0, int nSteps, int nThreads)
{
    long n = (long)(m->N + 1) * (m->N + 1) * NQ * m->nElems;
    if (nThreads < 1) nThreads = 1;
    for (int it = 0; it < nSteps; ++it) {
        double tn = t0 + it * m->dt;
        memset(m->G, 0, sizeof(double) * n);
        for (int st = 0; st < 3; ++st) {
            double t = tn + rk3_b[st] * m->dt;
            if (nThreads == 1) orc_sw2d_global_time_derivative(m, t);
            else gtd_threads(m, t, nThreads);
            rk_update(m, st, nThreads);
        }
    }
}

/* :1733-1868 CalculateBathymetryGradient, preceded by the bathymetry traces
 * (CalculateBathymetryAtBoundaries, :1160-1168). */
void orc_sw2d_bathymetry_gradient(orc_sw2d *m)
{
    int N = m->N, n1 = N + 1;
    const int form1 = (m->form == 1);
    double sC[16][2], pC[16][2], der[16], line[16], fL[2], fR[2];
    for (int e = 0; e < m->nElems; ++e)
        boundaries_one(n1, m->lagWest, m->lagEast, m->lagSouth, m->lagNorth, m->bathy, m->bbathy, 3, e);
    double *tend = (double *)malloc(sizeof(double) * n1 * n1 * 2);
    for (int e = 0; e < m->nElems; ++e) {
        for (int iP = 0; iP <= N; ++iP) {
            for (int iS = 0; iS <= N; ++iS) {
                int g0 = iS + n1 * (iP + n1 * e);
                double h = m->bathy[IDX(n1, iS, iP, 0, 3, e)];
                /* form 1 (localmods :1754-1755): the flux of h^2/2 */
                double hf = form1 ? (h * h) * 0.5 : h, z = form1 ? 0.0 * 0.5 : 0.0;
                double xF[2] = {hf, z}, yF[2] = {z, hf};
                for (int q = 0; q < 2; ++q) sC[iS][q] = m->dydp[g0] * xF[q] - m->dxdp[g0] * yF[q];
            }
            int nbE = iP + n1 * (1 + 4 * e), nbW = iP + n1 * (3 + 4 * e);
            double hE = m->bbathy[BDX(n1, iP, 0, 3, 1, e)], hW = m->bbathy[BDX(n1, iP, 0, 3, 3, e)];
            if (form1) { hE = (0.5 * hE) * hE; hW = (0.5 * hW) * hW; } /* :1764,1768 */
            for (int q = 0; q < 2; ++q) {
                fR[q] = (hE * m->nHat[2 * nbE + q]) * m->nLen[nbE];
                fL[q] = (hW * m->nHat[2 * nbW + q]) * m->nLen[nbW];
            }
            for (int q = 0; q < 2; ++q) {
                for (int k = 0; k <= N; ++k) line[k] = sC[k][q];
                dg_system_derivative(N, m->dMatS, m->qWeightS, fL[q], fR[q], line, m->lagWest, m->lagEast, der);
                for (int iS = 0; iS <= N; ++iS) tend[iS + n1 * (iP + n1 * q)] = der[iS];
            }
        }
        for (int iS = 0; iS <= N; ++iS) {
            for (int iP = 0; iP <= N; ++iP) {
                int g0 = iS + n1 * (iP + n1 * e);
                double h = m->bathy[IDX(n1, iS, iP, 0, 3, e)];
                double hf = form1 ? (h * h) * 0.5 : h, z = form1 ? 0.0 * 0.5 : 0.0;
                double xF[2] = {hf, z}, yF[2] = {z, hf};
                for (int q = 0; q < 2; ++q) pC[iP][q] = -m->dyds[g0] * xF[q] + m->dxds[g0] * yF[q];
            }
            int nbN = iS + n1 * (2 + 4 * e), nbS = iS + n1 * (0 + 4 * e);
            double hN = m->bbathy[BDX(n1, iS, 0, 3, 2, e)], hS = m->bbathy[BDX(n1, iS, 0, 3, 0, e)];
            if (form1) { hN = (0.5 * hN) * hN; hS = (0.5 * hS) * hS; } /* :1807,1811 */
            for (int q = 0; q < 2; ++q) {
                fR[q] = (hN * m->nHat[2 * nbN + q]) * m->nLen[nbN];
                fL[q] = (hS * m->nHat[2 * nbS + q]) * m->nLen[nbS];
            }
            for (int q = 0; q < 2; ++q) {
                for (int k = 0; k <= N; ++k) line[k] = pC[k][q];
                dg_system_derivative(N, m->dMatP, m->qWeightP, fL[q], fR[q], line, m->lagSouth, m->lagNorth, der);
                for (int iP = 0; iP <= N; ++iP) {
                    double Jv = m->J[iS + n1 * (iP + n1 * e)];
                    int id = iS + n1 * (iP + n1 * q);
                    if (form1) /* :1823-1825, divided by J h */
                        tend[id] = (tend[id] + der[iP]) / (Jv * m->bathy[IDX(n1, iS, iP, 0, 3, e)]);
                    else
                        tend[id] = (tend[id] + der[iP]) / Jv;
                }
            }
        }
        for (int iP = 0; iP <= N; ++iP)
            for (int iS = 0; iS <= N; ++iS) {
                m->bathy[IDX(n1, iS, iP, 1, 3, e)] = tend[iS + n1 * (iP + n1 * 0)];
                m->bathy[IDX(n1, iS, iP, 2, 3, e)] = tend[iS + n1 * (iP + n1 * 1)];
            }
    }
    free(tend);
}
