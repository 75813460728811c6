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
