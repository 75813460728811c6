/*
 * orc_hpe.c -- oracle (test infrastructure only): the hydrostatic primitive
 * equations (linear, Boussinesq, rigid lid) of
 *   src/highend/hydrostaticprimitives/HydrostaticPrimitivesClass.f90
 *     ForwardStepAB2 :1378-1424, GlobalTimeDerivative :1428-1506,
 *     DiagnoseP :1739-1822, DiagnoseW :1826-1912,
 *     CalculateBottomVerticalVelocity :1916-1945, EdgeFlux :2088-2183,
 *     MappedTimeDerivative :2187-2400, DGSystemDerivative :2709-2734,
 *     RiemannSolver :2738-2778, X/Y/ZFlux :2782-2832, Source :2836-2852,
 *     GetExternalState :2856-2898
 *   src/highend/hydrostaticprimitives/HydrostaticPrimitivesClass_implicit.f90
 *     CalculateDivU :2520-2691, ForwardStepRK3 :1422-1471
 *
 * The class does not compile (it uses the undeclared G2D and tm, calls the
 * undefined ForwardStepBarotropic and names timers that do not exist), so
 * this restatement fills its gaps as follows, each one a choice of this
 * restatement and not something the reference computes:
 *   - divU, read by DiagnoseW, is not a component of the explicit class and
 *     nothing computes it there; it is CalculateDivU of the implicit variant
 *     of the class, called before DiagnoseW as that variant's
 *     GlobalTimeDerivative does.
 *   - DiagnoseP is commented out inside the explicit GlobalTimeDerivative
 *     (:1458) and called by CalculateBaroclinicFields and by the implicit
 *     variant's GlobalTimeDerivative; the switch diagnoseP selects either.
 *   - DiagnoseP and DiagnoseW evaluate the Gauss-node column through the
 *     interpolant of the Gauss nodes (the class passes the N+1 values to the
 *     interpolant of its M+1 point Gauss-Lobatto column storage, whose
 *     nodes do not match them); the column quadrature itself, M = (N+1)/2+1
 *     Gauss-Lobatto points on each sub-interval between consecutive nodes
 *     (Build :264), is the class's.
 *   - The background state's face values are interpolated from its volume
 *     values like the solution's (the class never computes them, so they
 *     keep their Build value of zero).
 *   - The time steppers leave out the barotropic step and the pressure
 *     correction (ForwardStepBarotropic is undefined; PressureCorrection
 *     needs the free-surface CG solver of the implicit variant).
 * Positions enter only through differences (the column height ztop - zbot
 * of DiagnoseP/W) because Source is identically zero.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"
#include "orc_hpe.h"

static const double rk3_a[3] = {0.0, -5.0 / 9.0, -153.0 / 128.0};
static const double rk3_g[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};

#define NV 5
#define NP 3
#define VOLV(i, j, k, f, e) ((i) + n1 * ((j) + n1 * ((k) + n1 * ((f) + NV * (size_t)(e)))))
#define VOLP(i, j, k, f, e) ((i) + n1 * ((j) + n1 * ((k) + n1 * ((f) + NP * (size_t)(e)))))
#define GEO(i, j, k, e) ((i) + n1 * ((j) + n1 * ((k) + n1 * (size_t)(e))))
#define TRV(a, f, s, e) ((a) + n2 * ((f) + NV * ((s) + 6 * (size_t)(e))))
#define TRP(a, f, s, e) ((a) + n2 * ((f) + NP * ((s) + 6 * (size_t)(e))))
#define FB(a, s, e) ((a) + n2 * ((s) + 6 * (size_t)(e)))

/* cofactor row of the covariant metric for direction dir, as the contravariant
 * flux is written at MappedTimeDerivative :2229-2231, :2261-2263, :2300-2302 */
static void cofactor(const double *m, int dir, double out[3])
{
    double dxds = m[0], dxdp = m[1], dxdq = m[2], dyds = m[3], dydp = m[4], dydq = m[5];
    double dzds = m[6], dzdp = m[7], dzdq = m[8];
    if (dir == 0) {
        out[0] = dydp * dzdq - dzdp * dydq;
        out[1] = dxdq * dzdp - dxdp * dzdq;
        out[2] = dxdp * dydq - dxdq * dydp;
    } else if (dir == 1) {
        out[0] = dydq * dzds - dzdq * dyds;
        out[1] = dxds * dzdq - dxdq * dzds;
        out[2] = dxdq * dyds - dxds * dydq;
    } else {
        out[0] = dyds * dzdp - dzds * dydp;
        out[1] = dxdp * dzds - dxds * dzdp;
        out[2] = dxds * dydp - dxdp * dyds;
    }
}

/* DGSystemDerivative :2709-2734 */
static void dgsd(int N, const double *dMat, const double *qW, double lFlux, double rFlux, const double *intFlux,
                 const double *lL, const double *lR, double *out)
{
    int n1 = N + 1;
    for (int k = 0; k <= N; ++k) {
        double acc = 0.0;
        for (int j = 0; j <= N; ++j) acc += dMat[k + n1 * j] * intFlux[j];
        out[k] = acc + (rFlux * lR[k] + lFlux * lL[k]) / qW[k];
    }
}

/* face values of nv variables of one volume array (DGSEMSolutionStorageClass_3D
 * CalculateSolutionAtBoundaries, sides south, east, north, west, bottom, top) */
static void traces(int N, int nv, int nE, const double *lL, const double *lR, const double *q, double *out)
{
    int n1 = N + 1, n2 = n1 * n1;
#define QV(i, j, k, f, e) q[(i) + n1 * ((j) + n1 * ((k) + n1 * ((f) + nv * (size_t)(e))))]
#define OT(a, f, s, e) out[(a) + n2 * ((f) + nv * ((s) + 6 * (size_t)(e)))]
    for (int e = 0; e < nE; ++e)
        for (int f = 0; f < nv; ++f)
            for (int b = 0; b <= N; ++b)
                for (int a = 0; a <= N; ++a) {
                    double w = 0, ea = 0, s = 0, no = 0, bo = 0, to = 0;
                    for (int k = 0; k <= N; ++k) {
                        w += lL[k] * QV(k, a, b, f, e);
                        ea += lR[k] * QV(k, a, b, f, e);
                        s += lL[k] * QV(a, k, b, f, e);
                        no += lR[k] * QV(a, k, b, f, e);
                        bo += lL[k] * QV(a, b, k, f, e);
                        to += lR[k] * QV(a, b, k, f, e);
                    }
                    int node = a + n1 * b;
                    OT(node, f, 0, e) = s;
                    OT(node, f, 1, e) = ea;
                    OT(node, f, 2, e) = no;
                    OT(node, f, 3, e) = w;
                    OT(node, f, 4, e) = bo;
                    OT(node, f, 5, e) = to;
                }
#undef QV
#undef OT
}

void orc_hpe_traces(orc_hpe *m)
{
    traces(m->N, NV, m->nElems, m->lagLeft, m->lagRight, m->sol, m->bsol);
}

void orc_hpe_divu(orc_hpe *m)
{
    int N = m->N, n1 = N + 1, n2 = n1 * n1;
    double cf[32], der[32], cs[3];
    double *t = (double *)malloc(sizeof(double) * n1 * n2);
    for (int e = 0; e < m->nElems; ++e) {
        const double *u = m->sol + VOLV(0, 0, 0, 0, e), *v = m->sol + VOLV(0, 0, 0, 1, e);
#define UB(node, side) m->bsol[TRV(node, 0, side, e)]
#define VB(node, side) m->bsol[TRV(node, 1, side, e)]
#define NF(node, side)                                                                                      \
    ((UB(node, side) * m->nHat[3 * FB(node, side, e)] + VB(node, side) * m->nHat[3 * FB(node, side, e) + 1]) * \
     m->nLen[FB(node, side, e)])
        for (int iQ = 0; iQ <= N; ++iQ) {
            for (int iP = 0; iP <= N; ++iP) {
                for (int iS = 0; iS <= N; ++iS) {
                    int id = iS + n1 * (iP + n1 * iQ);
                    cofactor(m->cov + 9 * GEO(iS, iP, iQ, e), 0, cs);
                    cf[iS] = cs[0] * u[id] + cs[1] * v[id];
                }
                int node = iP + n1 * iQ;
                dgsd(N, m->dMat, m->qWeight, NF(node, 3), NF(node, 1), cf, m->lagLeft, m->lagRight, der);
                for (int iS = 0; iS <= N; ++iS) t[iS + n1 * (iP + n1 * iQ)] = der[iS];
            }
            for (int iS = 0; iS <= N; ++iS) {
                for (int iP = 0; iP <= N; ++iP) {
                    int id = iS + n1 * (iP + n1 * iQ);
                    cofactor(m->cov + 9 * GEO(iS, iP, iQ, e), 1, cs);
                    cf[iP] = cs[0] * u[id] + cs[1] * v[id];
                }
                int node = iS + n1 * iQ;
                dgsd(N, m->dMat, m->qWeight, NF(node, 0), NF(node, 2), cf, m->lagLeft, m->lagRight, der);
                for (int iP = 0; iP <= N; ++iP) t[iS + n1 * (iP + n1 * iQ)] = t[iS + n1 * (iP + n1 * iQ)] + der[iP];
            }
        }
        for (int iP = 0; iP <= N; ++iP)
            for (int iS = 0; iS <= N; ++iS) {
                for (int iQ = 0; iQ <= N; ++iQ) {
                    int id = iS + n1 * (iP + n1 * iQ);
                    cofactor(m->cov + 9 * GEO(iS, iP, iQ, e), 2, cs);
                    cf[iQ] = cs[0] * u[id] + cs[1] * v[id];
                }
                int node = iS + n1 * iP;
                dgsd(N, m->dMat, m->qWeight, NF(node, 4), NF(node, 5), cf, m->lagLeft, m->lagRight, der);
                for (int iQ = 0; iQ <= N; ++iQ) {
                    int id = iS + n1 * (iP + n1 * iQ);
                    m->divU[GEO(iS, iP, iQ, e)] = (t[id] + der[iQ]) / m->J[GEO(iS, iP, iQ, e)];
                }
            }
#undef UB
#undef VB
#undef NF
    }
    free(t);
}

/* one column sweep shared by DiagnoseP (down, integrand b) and DiagnoseW (up,
 * integrand -divU): for each node iQ in sweep order the increment over the
 * sub-interval from the previous node (or the element face) is the M+1 point
 * Gauss-Lobatto rule on the column interpolant, scaled by the element's
 * column height; the face value carried to the next element is the
 * interpolant of the new column at the far face. */
static void column_sweep(orc_hpe *m, int hE, int down)
{
    int N = m->N, n1 = N + 1, n2 = n1 * n1, M = m->M;
    int var = down ? 4 : 3;
    double line[32];
    double *base = (double *)calloc(n2, sizeof(double));
    if (!down) memcpy(base, m->bottomW + (size_t)n2 * hE, sizeof(double) * n2);
    const double *lFar = down ? m->lagLeft : m->lagRight;
    for (int li = 0; li < m->nLayers; ++li) {
        int layer = down ? m->nLayers - 1 - li : li;
        int e = hE + m->nHElems * layer;
        double xiI = down ? 1.0 : -1.0;
        for (int qi = 0; qi <= N; ++qi) {
            int iQ = down ? N - qi : qi;
            double xiIp1 = m->nodes[iQ];
            for (int iP = 0; iP <= N; ++iP)
                for (int iS = 0; iS <= N; ++iS) {
                    int node = iS + n1 * iP;
                    for (int k = 0; k <= N; ++k)
                        line[k] = down ? m->sol[VOLV(iS, iP, k, 2, e)] : m->divU[GEO(iS, iP, k, e)];
                    double acc = 0.0;
                    for (int k = 0; k <= M; ++k) {
                        double xiK = 0.5 * (xiIp1 - xiI) * (m->intNodes[k] + 1.0) + xiI;
                        double val = orc_lagrange_interpolate(N, m->nodes, m->bWeights, xiK, line);
                        acc = down ? acc + val * m->intWeights[k] : acc - val * m->intWeights[k];
                    }
                    double ztop = m->bpos[3 * FB(node, 5, e) + 2], zbot = m->bpos[3 * FB(node, 4, e) + 2];
                    double next = base[node] + acc * 0.5 * (ztop - zbot) * 0.5 * (xiIp1 - xiI);
                    m->sol[VOLV(iS, iP, iQ, var, e)] = next;
                    base[node] = next;
                }
            xiI = xiIp1;
        }
        for (int iP = 0; iP <= N; ++iP)
            for (int iS = 0; iS <= N; ++iS) {
                double acc = 0.0;
                for (int k = 0; k <= N; ++k) acc += lFar[k] * m->sol[VOLV(iS, iP, k, var, e)];
                base[iS + n1 * iP] = acc;
            }
    }
    free(base);
}

void orc_hpe_diagnose_p(orc_hpe *m)
{
    for (int hE = 0; hE < m->nHElems; ++hE) column_sweep(m, hE, 1);
}

void orc_hpe_diagnose_w(orc_hpe *m)
{
    int N = m->N, n1 = N + 1, n2 = n1 * n1;
    for (int hE = 0; hE < m->nHElems; ++hE) {
        int e = hE; /* bottom layer */
        for (int node = 0; node < n2; ++node) {
            size_t fb = FB(node, 4, e);
            double ub = m->bsol[TRV(node, 0, 4, e)], vb = m->bsol[TRV(node, 1, 4, e)];
            m->bottomW[(size_t)n2 * hE + node] =
                -(ub * m->nHat[3 * fb] + vb * m->nHat[3 * fb + 1]) / m->nHat[3 * fb + 2];
        }
        column_sweep(m, hE, 0);
    }
}

static void external_state(int bc, const double *n, const double *in, double *ex)
{
    for (int f = 0; f < NV; ++f) ex[f] = 0.0;
    if (bc == ORC_NO_NORMAL_FLOW) {
        ex[0] = (n[1] * n[1] - n[0] * n[0]) * in[0] - 2.0 * n[0] * n[1] * in[1];
        ex[1] = (n[0] * n[0] - n[1] * n[1]) * in[1] - 2.0 * n[0] * n[1] * in[0];
        ex[4] = in[4];
    } else if (bc == ORC_SEA_FLOOR) {
        ex[0] = in[0];
        ex[1] = in[1];
        ex[3] = -(in[3] + 2.0 * (in[0] * n[0] + in[1] * n[1])) / n[2];
        ex[4] = in[4];
    }
}

static void riemann(const double *in, const double *ex, const double *bg, const double *n, double *flux)
{
    double jump = ex[2] - in[2];
    double fIn = (bg[2] * in[3]) * n[2], fEx = (bg[2] * ex[3]) * n[2];
    double uIn = (in[3] + bg[3]) * n[2], uEx = (ex[3] + bg[3]) * n[2];
    double uN = fmax(fabs(uIn), fabs(uEx));
    flux[0] = 0.5 * (in[4] + ex[4]) * n[0];
    flux[1] = 0.5 * (in[4] + ex[4]) * n[1];
    flux[2] = 0.5 * (fEx + fIn - uN * jump);
}

void orc_hpe_edge_flux(orc_hpe *m)
{
    int N = m->N, n1 = N + 1, n2 = n1 * n1;
    double *bgt = (double *)malloc(sizeof(double) * (size_t)n2 * NV * 6 * m->nElems);
    traces(N, NV, m->nElems, m->lagLeft, m->lagRight, m->bg, bgt);
    for (int iF = 0; iF < m->nFaces; ++iF) {
        const int *fc = m->faces + 9 * iF;
        int e1 = fc[0] - 1, s1 = fc[1] - 1, e2 = fc[2], s2 = abs(fc[3]) - 1;
        int iS = fc[4], jS = fc[5], iI = fc[6], jI = fc[7], sw = fc[8];
        int bc = e2;
        if (e2 < 0 && s1 == 4) bc = ORC_SEA_FLOOR;
        for (int j = 0; j <= N; ++j)
            for (int i = 0; i <= N; ++i) {
                int node = i + n1 * j;
                size_t fb = FB(node, s1, e1);
                const double *n = m->nHat + 3 * fb;
                double in[NV], ex[NV], bg[NV], flux[NP];
                for (int f = 0; f < NV; ++f) {
                    in[f] = m->bsol[TRV(node, f, s1, e1)];
                    bg[f] = bgt[TRV(node, f, s1, e1)];
                }
                int k = sw ? jS + j * jI : iS + i * iI;
                int l = sw ? iS + i * iI : jS + j * jI;
                if (e2 > 0)
                    for (int f = 0; f < NV; ++f) ex[f] = m->bsol[TRV(k + n1 * l, f, s2, e2 - 1)];
                else
                    external_state(bc, n, in, ex);
                riemann(in, ex, bg, n, flux);
                for (int f = 0; f < NP; ++f) {
                    double fl = flux[f] * m->nLen[fb];
                    m->bflux[TRP(node, f, s1, e1)] = fl;
                    if (e2 > 0) m->bflux[TRP(k + n1 * l, f, s2, e2 - 1)] = -fl;
                }
            }
    }
    free(bgt);
}

void orc_hpe_mapped(orc_hpe *m)
{
    int N = m->N, n1 = N + 1, n2 = n1 * n1;
    double cf[3][32], der[32], cs[3];
    for (int e = 0; e < m->nElems; ++e) {
        const double *p = m->sol + VOLV(0, 0, 0, 4, e), *w = m->sol + VOLV(0, 0, 0, 3, e);
        const double *B = m->bg + VOLV(0, 0, 0, 2, e);
        double *tend = m->tend + VOLP(0, 0, 0, 0, e);
#define TD(i, j, k, f) tend[(i) + n1 * ((j) + n1 * ((k) + n1 * (f)))]
        for (int iQ = 0; iQ <= N; ++iQ) {
            for (int iP = 0; iP <= N; ++iP) {
                for (int iS = 0; iS <= N; ++iS) {
                    int id = iS + n1 * (iP + n1 * iQ);
                    cofactor(m->cov + 9 * GEO(iS, iP, iQ, e), 0, cs);
                    cf[0][iS] = cs[0] * p[id];
                    cf[1][iS] = cs[1] * p[id];
                    cf[2][iS] = cs[2] * (B[id] * w[id]);
                }
                int node = iP + n1 * iQ;
                for (int f = 0; f < NP; ++f) {
                    dgsd(N, m->dMat, m->qWeight, m->bflux[TRP(node, f, 3, e)], m->bflux[TRP(node, f, 1, e)], cf[f],
                         m->lagLeft, m->lagRight, der);
                    for (int iS = 0; iS <= N; ++iS) TD(iS, iP, iQ, f) = -der[iS];
                }
            }
            for (int iS = 0; iS <= N; ++iS) {
                for (int iP = 0; iP <= N; ++iP) {
                    int id = iS + n1 * (iP + n1 * iQ);
                    cofactor(m->cov + 9 * GEO(iS, iP, iQ, e), 1, cs);
                    cf[0][iP] = cs[0] * p[id];
                    cf[1][iP] = cs[1] * p[id];
                    cf[2][iP] = cs[2] * (B[id] * w[id]);
                }
                int node = iS + n1 * iQ;
                for (int f = 0; f < NP; ++f) {
                    dgsd(N, m->dMat, m->qWeight, m->bflux[TRP(node, f, 0, e)], m->bflux[TRP(node, f, 2, e)], cf[f],
                         m->lagLeft, m->lagRight, der);
                    for (int iP = 0; iP <= N; ++iP) TD(iS, iP, iQ, f) = TD(iS, iP, iQ, f) - der[iP];
                }
            }
        }
        for (int iP = 0; iP <= N; ++iP)
            for (int iS = 0; iS <= N; ++iS) {
                for (int iQ = 0; iQ <= N; ++iQ) {
                    int id = iS + n1 * (iP + n1 * iQ);
                    cofactor(m->cov + 9 * GEO(iS, iP, iQ, e), 2, cs);
                    cf[0][iQ] = cs[0] * p[id];
                    cf[1][iQ] = cs[1] * p[id];
                    cf[2][iQ] = cs[2] * (B[id] * w[id]);
                }
                int node = iS + n1 * iP;
                for (int f = 0; f < NP; ++f) {
                    /* the class passes qWeightY in the q direction (as Advection3D, quirk Q11); equal weights */
                    dgsd(N, m->dMat, m->qWeight, m->bflux[TRP(node, f, 4, e)], m->bflux[TRP(node, f, 5, e)], cf[f],
                         m->lagLeft, m->lagRight, der);
                    for (int iQ = 0; iQ <= N; ++iQ)
                        TD(iS, iP, iQ, f) = (TD(iS, iP, iQ, f) - der[iQ]) / m->J[GEO(iS, iP, iQ, e)];
                }
            }
        /* Source :2836-2852 is identically zero */
#undef TD
    }
}

void orc_hpe_global_time_derivative(orc_hpe *m)
{
    orc_hpe_traces(m);
    orc_hpe_divu(m);
    if (m->diagnoseP) orc_hpe_diagnose_p(m);
    orc_hpe_diagnose_w(m);
    orc_hpe_traces(m);
    orc_hpe_edge_flux(m);
    orc_hpe_mapped(m);
}

void orc_hpe_forward_step_rk3(orc_hpe *m)
{
    int n1 = m->N + 1;
    size_t n3 = (size_t)n1 * n1 * n1;
    memset(m->G, 0, sizeof(double) * n3 * NP * m->nElems);
    for (int st = 0; st < 3; ++st) {
        orc_hpe_global_time_derivative(m);
        for (int e = 0; e < m->nElems; ++e)
            for (int f = 0; f < NP; ++f)
                for (size_t i = 0; i < n3; ++i) {
                    size_t g = i + n3 * (f + NP * (size_t)e), s = i + n3 * (f + NV * (size_t)e);
                    m->G[g] = rk3_a[st] * m->G[g] + m->tend[g];
                    m->sol[s] = m->sol[s] + rk3_g[st] * m->dt * m->G[g];
                }
    }
}

void orc_hpe_forward_step_ab2(orc_hpe *m)
{
    int n1 = m->N + 1;
    size_t n3 = (size_t)n1 * n1 * n1;
    orc_hpe_global_time_derivative(m);
    for (int e = 0; e < m->nElems; ++e)
        for (int f = 0; f < NP; ++f)
            for (size_t i = 0; i < n3; ++i) {
                size_t g = i + n3 * (f + NP * (size_t)e), s = i + n3 * (f + NV * (size_t)e);
                double Gn = 0.5 * (3.0 * m->tend[g] - m->G[g]);
                m->sol[s] = m->sol[s] + m->dt * Gn;
                m->G[g] = m->tend[g];
            }
}
