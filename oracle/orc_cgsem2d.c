/*
 * orc_cgsem2d.c -- oracle (test infrastructure only): the continuous-Galerkin
 * spectral-element elliptic solver of src/cgsem/CGsemElliptic_2D_Class.f90,
 * restated on arrays:
 *   SetDirichletBoundaryCondition :234-300, Mask :350-430, MaskSide :434-466,
 *   UnMask :470-613, GlobalSum :617-702, SumSide :706-787,
 *   CalculateGradient :791-850, FluxDivergence :854-926,
 *   MatrixAction :930-976, Residual :980-1068, DotProduct :1072-1110,
 *   Solve_PCConjugateGradient :1114-1207.
 * Node-to-element lists are in increasing element order (LinkedList
 * AddToList appends, src/common/LinkedListClass.f90:239-312;
 * GetNodeToElementConnectivity_QuadMesh adds elements in order).
 * Quirk Q6 (Mask tests an uninitialised eID, :387): by default resolved to
 * the intended behaviour (secondary sides of interior edges are masked);
 * q6AsBuilt = 1 leaves them, as the reference compiled with amdflang -O2
 * does.
 * The preconditioner is the identity (the reference's pcIterates = 0
 * branch, CGsemPreconditioner_2D_Class.f90:740-742) or, as an addition,
 * the inverse of the assembled diagonal (Jacobi).
 * Layout: u(0:N,0:N,1:nEl) at [i + n1*(j + n1*e)].
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"
#include "orc_cgsem2d.h"

static const int cornerI[4] = {0, 1, 1, 0}; /* cornerMap(1,:) / N : SW, SE, NE, NW */
static const int cornerJ[4] = {0, 0, 1, 1};

#define U3(a, i, j, e) (a)[(i) + n1 * ((j) + n1 * (e))]

static void mask_side(const orc_cgsem2d *c, double *u, int e, int s)
{
    int N = c->N, n1 = N + 1;
    s = abs(s);
    if (s == 2 || s == 4) {
        int iS = (s == 2) ? N : 0;
        for (int iP = 1; iP <= N - 1; ++iP) U3(u, iS, iP, e) = 0.0;
    } else {
        int iP = (s == 3) ? N : 0;
        for (int iS = 1; iS <= N - 1; ++iS) U3(u, iS, iP, e) = 0.0;
    }
}

/* Mask :350-430 */
void orc_cgsem2d_mask(const orc_cgsem2d *c, double *u)
{
    int N = c->N, n1 = N + 1;
    for (int ie = 0; ie < c->nEdges; ++ie) {
        const int *ed = c->edges + 6 * ie;
        if (ed[2] == ORC_DIRICHLET) mask_side(c, u, ed[0] - 1, ed[1]);
        else if (ed[2] > 0 && !c->q6AsBuilt) mask_side(c, u, ed[2] - 1, ed[3]);
    }
    for (int n = 0; n < c->nNodes; ++n) {
        int start = c->nodeType[n] == ORC_DIRICHLET ? 0 : 1;
        for (int k = c->nodeElemStart[n] + start; k < c->nodeElemStart[n + 1]; ++k) {
            int e = c->nodeElem[k], lc = c->nodeLocal[k];
            U3(u, cornerI[lc] * N, cornerJ[lc] * N, e) = 0.0;
        }
    }
}

static void side_line(const orc_cgsem2d *c, int s, int *iS0, int *iP0, int *dS, int *dP)
{
    int N = c->N;
    if (s == 2 || s == 4) { *iS0 = (s == 2) ? N : 0; *iP0 = 0; *dS = 0; *dP = 1; }
    else { *iS0 = 0; *iP0 = (s == 3) ? N : 0; *dS = 1; *dP = 0; }
}

/* UnMask :470-613 */
void orc_cgsem2d_unmask(const orc_cgsem2d *c, double *u)
{
    int N = c->N, n1 = N + 1;
    double temp[32];
    for (int ie = 0; ie < c->nEdges; ++ie) {
        const int *ed = c->edges + 6 * ie;
        if (ed[2] <= 0) continue;
        int i0, j0, di, dj;
        side_line(c, ed[1], &i0, &j0, &di, &dj);
        for (int k = 1; k <= N - 1; ++k) temp[k] = U3(u, i0 + di * k, j0 + dj * k, ed[0] - 1);
        side_line(c, abs(ed[3]), &i0, &j0, &di, &dj);
        int jj = ed[4], inc = ed[5];
        for (int k = 1; k <= N - 1; ++k) {
            U3(u, i0 + di * jj, j0 + dj * jj, ed[2] - 1) = temp[k];
            jj += inc;
        }
    }
    for (int n = 0; n < c->nNodes; ++n) {
        if (c->nodeType[n] == ORC_DIRICHLET) continue;
        int k0 = c->nodeElemStart[n];
        int e0 = c->nodeElem[k0], l0 = c->nodeLocal[k0];
        double v = U3(u, cornerI[l0] * N, cornerJ[l0] * N, e0);
        for (int k = k0 + 1; k < c->nodeElemStart[n + 1]; ++k)
            U3(u, cornerI[c->nodeLocal[k]] * N, cornerJ[c->nodeLocal[k]] * N, c->nodeElem[k]) = v;
    }
}

/* GlobalSum :617-702 with SumSide :706-787 */
void orc_cgsem2d_global_sum(const orc_cgsem2d *c, double *u)
{
    int N = c->N, n1 = N + 1;
    double t1[32], t2[32];
    for (int ie = 0; ie < c->nEdges; ++ie) {
        const int *ed = c->edges + 6 * ie;
        if (ed[2] <= 0) continue;
        int e[2] = {ed[0] - 1, ed[2] - 1}, s[2] = {ed[1], abs(ed[3])};
        int i0[2], j0[2], di[2], dj[2];
        for (int q = 0; q < 2; ++q) side_line(c, s[q], &i0[q], &j0[q], &di[q], &dj[q]);
        for (int k = 1; k <= N - 1; ++k) {
            t1[k] = U3(u, i0[0] + di[0] * k, j0[0] + dj[0] * k, e[0]);
            t2[k] = U3(u, i0[1] + di[1] * k, j0[1] + dj[1] * k, e[1]);
        }
        int nn = ed[4], inc = ed[5];
        for (int k = 1; k <= N - 1; ++k) {
            double sum = t1[k] + t2[nn];
            t1[k] = sum;
            t2[nn] = sum;
            nn += inc;
        }
        for (int k = 1; k <= N - 1; ++k) {
            U3(u, i0[0] + di[0] * k, j0[0] + dj[0] * k, e[0]) = t1[k];
            U3(u, i0[1] + di[1] * k, j0[1] + dj[1] * k, e[1]) = t2[k];
        }
    }
    for (int n = 0; n < c->nNodes; ++n) {
        if (c->nodeType[n] == ORC_DIRICHLET) continue;
        double sum = 0.0;
        for (int k = c->nodeElemStart[n]; k < c->nodeElemStart[n + 1]; ++k)
            sum = sum + U3(u, cornerI[c->nodeLocal[k]] * N, cornerJ[c->nodeLocal[k]] * N, c->nodeElem[k]);
        for (int k = c->nodeElemStart[n]; k < c->nodeElemStart[n + 1]; ++k)
            U3(u, cornerI[c->nodeLocal[k]] * N, cornerJ[c->nodeLocal[k]] * N, c->nodeElem[k]) = sum;
    }
}

/* CalculateGradient :791-850 + FluxDivergence :854-926 for one element */
static void flux_divergence(const orc_cgsem2d *c, int e, const double *u, double *Lu)
{
    int N = c->N, n1 = N + 1;
    const double *D = c->D, *w = c->w;
    double duds[256], dudp[256], F1[256], F2[256];
    for (int iP = 0; iP <= N; ++iP)
        for (int k = 0; k <= N; ++k) {
            double acc = 0.0;
            for (int m = 0; m <= N; ++m) acc += D[k + n1 * m] * u[m + n1 * iP];
            duds[k + n1 * iP] = acc;
        }
    for (int iS = 0; iS <= N; ++iS)
        for (int k = 0; k <= N; ++k) {
            double acc = 0.0;
            for (int m = 0; m <= N; ++m) acc += D[k + n1 * m] * u[iS + n1 * m];
            dudp[iS + n1 * k] = acc;
        }
    for (int iP = 0; iP <= N; ++iP)
        for (int iS = 0; iS <= N; ++iS) {
            int g = iS + n1 * (iP + n1 * e), l = iS + n1 * iP;
            double J = c->J[g], dxds = c->dxds[g], dxdp = c->dxdp[g], dyds = c->dyds[g], dydp = c->dydp[g];
            double dudx = (dydp * duds[l] - dyds * dudp[l]) / J;
            double dudy = (dxds * dudp[l] - dxdp * duds[l]) / J;
            double k = c->fluxCoeff[g];
            F1[l] = k * (dudx * dydp - dudy * dxdp) * w[iS];
            F2[l] = k * (dudy * dxds - dudx * dyds) * w[iP];
        }
    /* dF1ds = MATMUL(TRANSPOSE(D), F1(:,iP)); dF2dp = MATMUL(TRANSPOSE(D), F2(iS,:)) */
    for (int iP = 0; iP <= N; ++iP)
        for (int iS = 0; iS <= N; ++iS) {
            double a = 0.0, b = 0.0;
            for (int m = 0; m <= N; ++m) {
                a += D[m + n1 * iS] * F1[m + n1 * iP];
                b += D[m + n1 * iP] * F2[iS + n1 * m];
            }
            Lu[iS + n1 * iP] = -(a * w[iP] + b * w[iS]);
        }
}

/* MatrixAction :930-976 (u is unmasked in place and masked again) */
void orc_cgsem2d_matrix_action(const orc_cgsem2d *c, double *u, double *Au)
{
    int np = (c->N + 1) * (c->N + 1);
    orc_cgsem2d_unmask(c, u);
    for (int e = 0; e < c->nElems; ++e) flux_divergence(c, e, u + (size_t)np * e, Au + (size_t)np * e);
    orc_cgsem2d_global_sum(c, Au);
    orc_cgsem2d_mask(c, Au);
    orc_cgsem2d_mask(c, u);
}

/* DotProduct :1072-1110 (masks both arguments in place, quirk Q7) */
double orc_cgsem2d_dot(const orc_cgsem2d *c, double *u, double *v)
{
    orc_cgsem2d_mask(c, u);
    orc_cgsem2d_mask(c, v);
    int n1 = c->N + 1;
    double s = 0.0;
    for (int e = 0; e < c->nElems; ++e)
        for (int iP = 0; iP <= c->N; ++iP) {
            double d = 0.0;
            for (int iS = 0; iS <= c->N; ++iS) d += U3(u, iS, iP, e) * U3(v, iS, iP, e);
            s = s + d;
        }
    return s;
}

/* Residual :980-1068 (homogeneous Dirichlet values: SetDirichlet writes dirichletValue) */
void orc_cgsem2d_residual(const orc_cgsem2d *c, double *r)
{
    int N = c->N, n1 = N + 1;
    size_t nl = (size_t)n1 * n1 * c->nElems;
    double *Au = (double *)malloc(sizeof(double) * nl);
    orc_cgsem2d_matrix_action(c, c->solution, Au);
    for (int e = 0; e < c->nElems; ++e)
        for (int iP = 0; iP <= N; ++iP)
            for (int iS = 0; iS <= N; ++iS) {
                size_t g = iS + n1 * (iP + (size_t)n1 * e);
                r[g] = c->source[g] * c->J[g] * c->w[iS] * c->w[iP] - Au[g];
            }
    orc_cgsem2d_global_sum(c, r);
    orc_cgsem2d_mask(c, r);
    orc_cgsem2d_mask(c, c->solution);
    orc_cgsem2d_set_dirichlet(c);
    free(Au);
}

/* SetDirichletBoundaryCondition :234-300 with a constant boundary value */
void orc_cgsem2d_set_dirichlet(const orc_cgsem2d *c)
{
    int N = c->N, n1 = N + 1;
    double *u = c->solution;
    for (int ie = 0; ie < c->nEdges; ++ie) {
        const int *ed = c->edges + 6 * ie;
        if (ed[2] != ORC_DIRICHLET) continue;
        int i0, j0, di, dj;
        side_line(c, ed[1], &i0, &j0, &di, &dj);
        for (int k = 1; k <= N - 1; ++k) U3(u, i0 + di * k, j0 + dj * k, ed[0] - 1) = c->dirichletValue;
    }
    for (int n = 0; n < c->nNodes; ++n) {
        if (c->nodeType[n] != ORC_DIRICHLET) continue;
        for (int k = c->nodeElemStart[n]; k < c->nodeElemStart[n + 1]; ++k)
            U3(u, cornerI[c->nodeLocal[k]] * N, cornerJ[c->nodeLocal[k]] * N, c->nodeElem[k]) = c->dirichletValue;
    }
}

/* assembled diagonal of MatrixAction (masked layout), by unit vectors per element */
void orc_cgsem2d_diagonal(const orc_cgsem2d *c, double *diag)
{
    int n1 = c->N + 1, np = n1 * n1;
    double *unit = (double *)calloc(np, sizeof(double)), *Lu = (double *)malloc(sizeof(double) * np);
    for (int e = 0; e < c->nElems; ++e)
        for (int l = 0; l < np; ++l) {
            memset(unit, 0, sizeof(double) * np);
            unit[l] = 1.0;
            flux_divergence(c, e, unit, Lu);
            diag[(size_t)np * e + l] = Lu[l];
        }
    orc_cgsem2d_global_sum(c, diag);
    free(unit); free(Lu);
}

/* Solve_PCConjugateGradient :1114-1207; precond 0 = identity, 1 = Jacobi */
int orc_cgsem2d_solve_pcg(orc_cgsem2d *c, int precond, int *iterations)
{
    int N = c->N, n1 = N + 1;
    size_t nl = (size_t)n1 * n1 * c->nElems;
    double *u = (double *)malloc(sizeof(double) * nl), *r = (double *)malloc(sizeof(double) * nl);
    double *v = (double *)malloc(sizeof(double) * nl), *z = (double *)malloc(sizeof(double) * nl);
    double *w = (double *)malloc(sizeof(double) * nl), *dinv = NULL;
    if (precond == 1) {
        dinv = (double *)malloc(sizeof(double) * nl);
        orc_cgsem2d_diagonal(c, dinv);
        for (size_t k = 0; k < nl; ++k) dinv[k] = (dinv[k] != 0.0) ? 1.0 / dinv[k] : 0.0;
    }
    int ioerr = -2, iter;
    double a, b, num, den = 0.0, r0, rNorm;
    memcpy(u, c->solution, sizeof(double) * nl);
    orc_cgsem2d_residual(c, r);
    r0 = sqrt(orc_cgsem2d_dot(c, r, r));
    for (int k = 0; k < c->resiLen; ++k) c->resi[k] = 0.0;
    c->resi[0] = r0;
    rNorm = r0;
    *iterations = 0;
    if (r0 < c->tolerance) goto done;
    for (iter = 1; iter <= c->maxIters; ++iter) {
        orc_cgsem2d_unmask(c, r);
        if (precond == 1) for (size_t k = 0; k < nl; ++k) w[k] = dinv[k] * r[k];
        else memcpy(w, r, sizeof(double) * nl);
        num = orc_cgsem2d_dot(c, r, w);
        if (iter == 1) memcpy(v, w, sizeof(double) * nl);
        else { b = num / den; for (size_t k = 0; k < nl; ++k) v[k] = w[k] + b * v[k]; }
        orc_cgsem2d_matrix_action(c, v, z);
        a = num / orc_cgsem2d_dot(c, v, z);
        for (size_t k = 0; k < nl; ++k) u[k] = u[k] + a * v[k];
        for (size_t k = 0; k < nl; ++k) r[k] = r[k] - a * z[k];
        den = num;
        rNorm = sqrt(orc_cgsem2d_dot(c, r, r));
        if (iter < c->resiLen) c->resi[iter] = rNorm;
        *iterations = iter;
        if (rNorm / r0 < c->tolerance) break;
    }
    if (rNorm / r0 > c->tolerance) ioerr = -1;
    orc_cgsem2d_unmask(c, u);
    memcpy(c->solution, u, sizeof(double) * nl);
    orc_cgsem2d_set_dirichlet(c);
done:
    free(u); free(r); free(v); free(z); free(w); free(dinv);
    return ioerr;
}

/* node-to-element lists and node types for a structured mesh */
int orc_cgsem2d_connectivity(orc_cgsem2d *c, const int *elemNodes, int nNodes)
{
    c->nNodes = nNodes;
    int *cnt = (int *)calloc(nNodes + 1, sizeof(int));
    for (int e = 0; e < c->nElems; ++e)
        for (int k = 0; k < 4; ++k) cnt[elemNodes[4 * e + k]]++;  /* 1-based node ids */
    c->nodeElemStart[0] = 0;
    for (int n = 0; n < nNodes; ++n) c->nodeElemStart[n + 1] = c->nodeElemStart[n] + cnt[n + 1];
    int *fill = (int *)calloc(nNodes, sizeof(int));
    for (int e = 0; e < c->nElems; ++e)
        for (int k = 0; k < 4; ++k) {
            int n = elemNodes[4 * e + k] - 1;
            int pos = c->nodeElemStart[n] + fill[n]++;
            c->nodeElem[pos] = e;
            c->nodeLocal[pos] = k;
        }
    for (int n = 0; n < nNodes; ++n) c->nodeType[n] = 1; /* interior */
    for (int ie = 0; ie < c->nEdges; ++ie) {
        int *ed = c->edges + 6 * ie;
        if (ed[2] == ORC_NO_NORMAL_FLOW || ed[2] == ORC_DIRICHLET) {
            ed[2] = ORC_DIRICHLET;  /* BuildQuadMesh_CGsemElliptic_2D :207-214 */
            int e = ed[0] - 1, s = ed[1];
            static const int em[2][4] = {{1, 2, 4, 1}, {2, 3, 3, 4}};
            c->nodeType[elemNodes[4 * e + em[0][s - 1] - 1] - 1] = ORC_DIRICHLET;
            c->nodeType[elemNodes[4 * e + em[1][s - 1] - 1] - 1] = ORC_DIRICHLET;
        }
    }
    free(cnt); free(fill);
    return 0;
}

/* ------------------------------------------------------------------------
 * CGsemPreconditioner_2D (src/cgsem/CGsemPreconditioner_2D_Class.f90) on the
 * linear mesh of ConstructLinearMeshFromOther_QuadMesh
 * (src/geom/QuadMeshClass.f90:1863-2096).
 */

/* Curve_2D of two nodes at s = -1, 1 (Lagrange_1D, N = 1): value and slope */
static void curve1_eval(const double *cx, const double *cy, const double *s, const double *bw, double a,
                        double *P)
{
    P[0] = orc_lagrange_interpolate(1, s, bw, a, cx);
    P[1] = orc_lagrange_interpolate(1, s, bw, a, cy);
}

static void curve1_slope(const double *cx, const double *cy, const double *s, const double *bw, double a,
                         double *P)
{
    P[0] = orc_lagrange_derivative(1, s, bw, a, cx);
    P[1] = orc_lagrange_derivative(1, s, bw, a, cy);
}

/* TransfiniteInterpolationMetrics_2D (GeometrySupportRoutines.f90:109-158) for
 * four two-node curves cx/cy[side][node], sides south, east, north, west */
static void tfi_metrics1(double cx[4][2], double cy[4][2], const double *s, const double *bw, double a,
                         double b, double covT[2][2])
{
    double w[2], e[2], so[2], no[2], dW[2], dE[2], dS[2], dN[2], c1[2], c2[2], c3[2], c4[2];
    curve1_eval(cx[3], cy[3], s, bw, b, w);
    curve1_eval(cx[1], cy[1], s, bw, b, e);
    curve1_eval(cx[0], cy[0], s, bw, a, so);
    curve1_eval(cx[2], cy[2], s, bw, a, no);
    curve1_slope(cx[3], cy[3], s, bw, b, dW);
    curve1_slope(cx[1], cy[1], s, bw, b, dE);
    curve1_slope(cx[0], cy[0], s, bw, a, dS);
    curve1_slope(cx[2], cy[2], s, bw, a, dN);
    curve1_eval(cx[0], cy[0], s, bw, -1.0, c1);
    curve1_eval(cx[0], cy[0], s, bw, 1.0, c2);
    curve1_eval(cx[2], cy[2], s, bw, -1.0, c3);
    curve1_eval(cx[2], cy[2], s, bw, 1.0, c4);
    for (int k = 0; k < 2; ++k) {
        double P1 = 0.5 * (e[k] - w[k]), Ps = 0.5 * (c2[k] - c1[k]), Pn = 0.5 * (c4[k] - c3[k]);
        double L = dS[k] - Ps, R = dN[k] - Pn;
        covT[k][0] = P1 + 0.5 * ((1.0 - b) * L + (1.0 + b) * R);
        double P2 = 0.5 * (no[k] - so[k]), Pw = 0.5 * (c3[k] - c1[k]), Pe = 0.5 * (c4[k] - c2[k]);
        L = dW[k] - Pw;
        R = dE[k] - Pe;
        covT[k][1] = P2 + 0.5 * ((1.0 - a) * L + (1.0 + a) * R);
    }
}

static int cmpll(const void *p, const void *q)
{
    long long a = *(const long long *)p, b = *(const long long *)q;
    return (a > b) - (a < b);
}

int orc_cgsem2d_pc_build(orc_cgsem2d_pc *pc, int N, int nEl, const double *x, const double *y,
                         const double *fluxCoeff, double tolerance, int maxIters)
{
    memset(pc, 0, sizeof(*pc));
    const int n1 = N + 1, nLin = nEl * N * N, nLinNodes = nEl * n1 * n1;
    pc->nS = N;
    pc->nEl = nEl;
    pc->nLin = nLin;
    pc->tolerance = tolerance;
    pc->maxIters = maxIters;
    /* unique nodes by AlmostEqual positions, first occurrence numbering (:1900-1925) */
    int *uniq = (int *)malloc(sizeof(int) * nLinNodes);
    int nUnique = 0;
    for (int i = 0; i < nLinNodes; ++i) {
        int inc = 1;
        for (int j = 0; j < i; ++j)
            if (orc_almost_equal(x[i], x[j]) && orc_almost_equal(y[i], y[j])) { uniq[i] = uniq[j]; inc = 0; break; }
        nUnique += inc;
        if (inc) uniq[i] = nUnique;
    }
    pc->nNodes = nUnique;
    /* node positions: the last copy visited wins (SetNodePosition, :1933-1944) */
    double *nx = (double *)malloc(sizeof(double) * nUnique), *ny = (double *)malloc(sizeof(double) * nUnique);
    for (int i = 0; i < nLinNodes; ++i) { nx[uniq[i] - 1] = x[i]; ny[uniq[i] - 1] = y[i]; }
    pc->pcToOrig = (int *)calloc((size_t)12 * nLin, sizeof(int));
    pc->origToPC = (int *)calloc((size_t)3 * nLinNodes, sizeof(int));
    int *elemNodes = (int *)malloc(sizeof(int) * 4 * nLin);

    const size_t nl1 = (size_t)4 * nLin;
    pc->dxds = (double *)malloc(sizeof(double) * nl1); pc->dxdp = (double *)malloc(sizeof(double) * nl1);
    pc->dyds = (double *)malloc(sizeof(double) * nl1); pc->dydp = (double *)malloc(sizeof(double) * nl1);
    pc->J = (double *)malloc(sizeof(double) * nl1); pc->fluxCoeff = (double *)malloc(sizeof(double) * nl1);
    pc->D1 = (double *)malloc(sizeof(double) * 4);
    pc->w1 = (double *)malloc(sizeof(double) * 2);
    double s1[2], bw1[2], wq[2];
    orc_generate_legendre_quadrature(1, s1, wq, ORC_GAUSS_LOBATTO);
    double sL[2] = {-1.0, 1.0};  /* the linear mesh's interpolant nodes (:1929-1931) */
    orc_barycentric_weights(1, sL, bw1);
    {
        double bwq[2];
        orc_barycentric_weights(1, s1, bwq);
        orc_derivative_matrix(1, s1, bwq, pc->D1);
        pc->w1[0] = wq[0]; pc->w1[1] = wq[1];
    }
    static const int em[2][4] = {{1, 2, 4, 1}, {2, 3, 3, 4}};  /* edgeMap */
    for (int iEl = 0; iEl < nEl; ++iEl)
        for (int iP = 1; iP <= N; ++iP)
            for (int iS = 1; iS <= N; ++iS) {
                const int eID = (iS - 1) + N * ((iP - 1) + iEl * N);  /* 0-based (:1960) */
                const int ci[4] = {iS - 1, iS, iS, iS - 1}, cj[4] = {iP - 1, iP - 1, iP, iP};
                /* linToOther (0,0), (1,0), (1,1), (0,1) and otherToLin, in the reference's order */
                const int li[4] = {0, 1, 1, 0}, lj[4] = {0, 0, 1, 1};
                for (int q = 0; q < 4; ++q) {
                    int *t = pc->pcToOrig + 3 * (li[q] + 2 * lj[q] + 4 * eID);
                    t[0] = ci[q]; t[1] = cj[q]; t[2] = iEl;
                    int *o = pc->origToPC + 3 * (ci[q] + n1 * (cj[q] + n1 * iEl));
                    o[0] = li[q]; o[1] = lj[q]; o[2] = eID;
                    elemNodes[4 * eID + q] = uniq[ci[q] + n1 * (cj[q] + n1 * iEl)];
                    pc->fluxCoeff[li[q] + 2 * lj[q] + 4 * eID] = fluxCoeff[ci[q] + n1 * (cj[q] + n1 * iEl)];
                }
                /* straight boundary curves between corner nodes (:1986-1997) */
                double cx[4][2], cy[4][2];
                for (int side = 0; side < 4; ++side) {
                    const int a = elemNodes[4 * eID + em[0][side] - 1] - 1, b = elemNodes[4 * eID + em[1][side] - 1] - 1;
                    const double x1 = nx[a], y1 = ny[a], x2 = nx[b], y2 = ny[b];
                    for (int nd = 0; nd < 2; ++nd) {
                        cx[side][nd] = x1 + (x2 - x1) * 0.5 * (sL[nd] + 1.0);
                        cy[side][nd] = y1 + (y2 - y1) * 0.5 * (sL[nd] + 1.0);
                    }
                }
                /* GenerateMetrics_MappedGeometry_2D at the corner nodes */
                for (int i = 0; i < 2; ++i)
                    for (int j = 0; j < 2; ++j) {
                        double covT[2][2];
                        tfi_metrics1(cx, cy, sL, bw1, sL[i], sL[j], covT);
                        const int g = i + 2 * j + 4 * eID;
                        pc->dxds[g] = covT[0][0]; pc->dxdp[g] = covT[0][1];
                        pc->dyds[g] = covT[1][0]; pc->dydp[g] = covT[1][1];
                        pc->J[g] = covT[0][0] * covT[1][1] - covT[0][1] * covT[1][0];
                    }
            }
    /* edges: element sides shared by two linear elements are interior; the
     * others are boundary and their nodes Dirichlet (:2047-2063 and :112-118) */
    int *cnt = (int *)calloc(nUnique, sizeof(int));
    pc->nodeType = (int *)malloc(sizeof(int) * nUnique);
    for (int n = 0; n < nUnique; ++n) pc->nodeType[n] = 1;
    {
        /* count side occurrences by (min, max) node pair with a hash of sorted pairs */
        long long *keys = (long long *)malloc(sizeof(long long) * 4 * nLin);
        for (int e = 0; e < nLin; ++e)
            for (int side = 0; side < 4; ++side) {
                long long a = elemNodes[4 * e + em[0][side] - 1], b = elemNodes[4 * e + em[1][side] - 1];
                if (a > b) { long long t = a; a = b; b = t; }
                keys[4 * e + side] = a * (long long)(nUnique + 1) + b;
            }
        /* sort copy to find unique sides */
        long long *sorted = (long long *)malloc(sizeof(long long) * 4 * nLin);
        memcpy(sorted, keys, sizeof(long long) * 4 * nLin);
        qsort(sorted, 4 * nLin, sizeof(long long), cmpll);
        for (int q = 0; q < 4 * nLin; ++q) {
            int m = 0;
            long long key = sorted[q];
            while (q + m < 4 * nLin && sorted[q + m] == key) ++m;
            if (m == 1) {
                int a = (int)(key / (nUnique + 1)), b = (int)(key % (nUnique + 1));
                pc->nodeType[a - 1] = ORC_DIRICHLET;
                pc->nodeType[b - 1] = ORC_DIRICHLET;
            }
            q += m - 1;
        }
        free(keys); free(sorted);
    }
    /* node -> element lists in AddToList order (GetNodeToElementConnectivity) */
    for (int e = 0; e < nLin; ++e)
        for (int k = 0; k < 4; ++k) cnt[elemNodes[4 * e + k] - 1]++;
    pc->nodeElemStart = (int *)malloc(sizeof(int) * (nUnique + 1));
    pc->nodeElem = (int *)malloc(sizeof(int) * 4 * nLin);
    pc->nodeLocal = (int *)malloc(sizeof(int) * 4 * nLin);
    pc->nodeElemStart[0] = 0;
    for (int n = 0; n < nUnique; ++n) pc->nodeElemStart[n + 1] = pc->nodeElemStart[n] + cnt[n];
    memset(cnt, 0, sizeof(int) * nUnique);
    for (int e = 0; e < nLin; ++e)
        for (int k = 0; k < 4; ++k) {
            int n = elemNodes[4 * e + k] - 1, pos = pc->nodeElemStart[n] + cnt[n]++;
            pc->nodeElem[pos] = e;
            pc->nodeLocal[pos] = k;
        }
    pc->edges1 = (int *)malloc(sizeof(int) * 6);  /* no edge loops at N = 1 */
    pc->solution1 = (double *)calloc(nl1, sizeof(double));
    pc->source1 = (double *)calloc(nl1, sizeof(double));
    orc_cgsem2d *l = &pc->lin;
    memset(l, 0, sizeof(*l));
    l->N = 1; l->nElems = nLin; l->nEdges = 0; l->nNodes = nUnique;
    l->edges = pc->edges1; l->nodeElemStart = pc->nodeElemStart; l->nodeElem = pc->nodeElem;
    l->nodeLocal = pc->nodeLocal; l->nodeType = pc->nodeType; l->D = pc->D1; l->w = pc->w1;
    l->dxds = pc->dxds; l->dxdp = pc->dxdp; l->dyds = pc->dyds; l->dydp = pc->dydp; l->J = pc->J;
    l->fluxCoeff = pc->fluxCoeff; l->solution = pc->solution1; l->source = pc->source1;
    free(uniq); free(nx); free(ny); free(elemNodes); free(cnt);
    return 0;
}

void orc_cgsem2d_pc_free(orc_cgsem2d_pc *pc)
{
    void *ptrs[] = {pc->pcToOrig, pc->origToPC, pc->D1, pc->w1, pc->dxds, pc->dxdp, pc->dyds, pc->dydp, pc->J,
                    pc->fluxCoeff, pc->edges1, pc->nodeElemStart, pc->nodeElem, pc->nodeLocal, pc->nodeType,
                    pc->solution1, pc->source1};
    for (size_t k = 0; k < sizeof(ptrs) / sizeof(ptrs[0]); ++k) free(ptrs[k]);
    memset(pc, 0, sizeof(*pc));
}

/* DotProduct_CGsemPreconditioner_2D (:600-660): masked copies, row sums */
static double pc_dot(orc_cgsem2d_pc *pc, const double *u, const double *v, double *uu, double *vv)
{
    size_t nl = (size_t)4 * pc->nLin;
    memcpy(uu, u, sizeof(double) * nl);
    memcpy(vv, v, sizeof(double) * nl);
    orc_cgsem2d_mask(&pc->lin, uu);
    orc_cgsem2d_mask(&pc->lin, vv);
    double s = 0.0;
    for (int e = 0; e < pc->nLin; ++e)
        for (int iP = 0; iP <= 1; ++iP) {
            double d = 0.0;
            for (int iS = 0; iS <= 1; ++iS) d += uu[iS + 2 * (iP + 2 * e)] * vv[iS + 2 * (iP + 2 * e)];
            s = s + d;
        }
    return s;
}

/* Solve_CGsemPreconditioner_2D (:664-746) */
int orc_cgsem2d_pc_solve(orc_cgsem2d_pc *pc, double *x, const double *b, int *iterations)
{
    const int nS = pc->nS, n1 = nS + 1;
    const size_t nl = (size_t)4 * pc->nLin;
    double *u = (double *)calloc(nl, sizeof(double)), *r = (double *)malloc(sizeof(double) * nl);
    double *v = (double *)calloc(nl, sizeof(double)), *z = (double *)malloc(sizeof(double) * nl);
    double *w = (double *)malloc(sizeof(double) * nl), *t1 = (double *)malloc(sizeof(double) * nl);
    double *t2 = (double *)malloc(sizeof(double) * nl);
    int ioerr = -2, iter, nIt = pc->maxIters;
    double TOL = pc->tolerance, rNorm = 0.0, a, c, num, den = 0.0, r0;
    for (size_t q = 0; q < nl; ++q) {  /* MapFromOrigToPC */
        const int *t = pc->pcToOrig + 3 * q;
        r[q] = b[t[0] + n1 * (t[1] + n1 * t[2])];
    }
    r0 = sqrt(pc_dot(pc, r, r, t1, t2));
    *iterations = 0;
    for (iter = 1; iter <= nIt; ++iter) {
        memcpy(w, r, sizeof(double) * nl);
        num = pc_dot(pc, r, w, t1, t2);
        if (iter == 1) memcpy(v, w, sizeof(double) * nl);
        else { c = num / den; for (size_t q = 0; q < nl; ++q) v[q] = w[q] + c * v[q]; }
        orc_cgsem2d_matrix_action(&pc->lin, v, z);
        a = num / pc_dot(pc, v, z, t1, t2);
        for (size_t q = 0; q < nl; ++q) u[q] = u[q] + a * v[q];
        for (size_t q = 0; q < nl; ++q) r[q] = r[q] - a * z[q];
        den = num;
        rNorm = sqrt(pc_dot(pc, r, r, t1, t2));
        *iterations = iter;
        if (rNorm / r0 < TOL) break;
    }
    if (rNorm / r0 > TOL) ioerr = -1;
    if (nIt == 0) memcpy(u, r, sizeof(double) * nl);
    orc_cgsem2d_unmask(&pc->lin, u);
    for (int e = 0; e < pc->nEl; ++e)  /* MapFromPCToOrig */
        for (int iP = 0; iP <= nS; ++iP)
            for (int iS = 0; iS <= nS; ++iS) {
                const int *o = pc->origToPC + 3 * (iS + n1 * (iP + n1 * e));
                x[iS + n1 * (iP + n1 * e)] = u[o[0] + 2 * (o[1] + 2 * o[2])];
            }
    free(u); free(r); free(v); free(z); free(w); free(t1); free(t2);
    return ioerr;
}

/* Solve_PCConjugateGradient (:1114-1207) with the preconditioner's Solve (:1169) */
int orc_cgsem2d_solve_pcg_fem(orc_cgsem2d *c, orc_cgsem2d_pc *pc, int *iterations, int *pcIterations)
{
    int N = c->N, n1 = N + 1;
    size_t nl = (size_t)n1 * n1 * c->nElems;
    double *u = (double *)malloc(sizeof(double) * nl), *r = (double *)malloc(sizeof(double) * nl);
    double *v = (double *)malloc(sizeof(double) * nl), *z = (double *)malloc(sizeof(double) * nl);
    double *w = (double *)malloc(sizeof(double) * nl);
    int ioerr = -2, iter, pcIt = 0;
    double a, b, num, den = 0.0, r0, rNorm;
    memcpy(u, c->solution, sizeof(double) * nl);
    orc_cgsem2d_residual(c, r);
    r0 = sqrt(orc_cgsem2d_dot(c, r, r));
    for (int k = 0; k < c->resiLen; ++k) c->resi[k] = 0.0;
    c->resi[0] = r0;
    rNorm = r0;
    *iterations = 0;
    *pcIterations = 0;
    if (r0 < c->tolerance) goto done;
    for (iter = 1; iter <= c->maxIters; ++iter) {
        orc_cgsem2d_unmask(c, r);
        orc_cgsem2d_pc_solve(pc, w, r, &pcIt);
        *pcIterations += pcIt;
        num = orc_cgsem2d_dot(c, r, w);
        if (iter == 1) memcpy(v, w, sizeof(double) * nl);
        else { b = num / den; for (size_t k = 0; k < nl; ++k) v[k] = w[k] + b * v[k]; }
        orc_cgsem2d_matrix_action(c, v, z);
        a = num / orc_cgsem2d_dot(c, v, z);
        for (size_t k = 0; k < nl; ++k) u[k] = u[k] + a * v[k];
        for (size_t k = 0; k < nl; ++k) r[k] = r[k] - a * z[k];
        den = num;
        rNorm = sqrt(orc_cgsem2d_dot(c, r, r));
        if (iter < c->resiLen) c->resi[iter] = rNorm;
        *iterations = iter;
        if (rNorm / r0 < c->tolerance) break;
    }
    if (rNorm / r0 > c->tolerance) ioerr = -1;
    orc_cgsem2d_unmask(c, u);
    memcpy(c->solution, u, sizeof(double) * nl);
    orc_cgsem2d_set_dirichlet(c);
done:
    free(u); free(r); free(v); free(z); free(w);
    return ioerr;
}

size_t orc_cgsem2d_pc_size(void) { return sizeof(orc_cgsem2d_pc); }
