/*
 * orc_interp.c -- oracle (test infrastructure only): Legendre quadrature,
 * barycentric Lagrange interpolation and nodal storage, restated from
 *   src/interp/Legendre.f90, src/interp/Lagrange_1D_Class.f90,
 *   src/interp/Lagrange_2D_Class.f90, src/interp/Lagrange_3D_Class.f90,
 *   src/nodal/NodalStorage_2D_Class.f90, src/nodal/NodalStorage_3D_Class.f90,
 *   src/common/CommonRoutines.f90.
 * Loop order and operation order follow the Fortran source.
 */
#include <math.h>
#include <float.h>
#include <string.h>
#include <stdlib.h>
#include "oracle.h"

#define PI_ (4.0 * atan(1.0)) /* ConstantsDictionary.f90: pi = 4*atan(1) */
#define TOL_ DBL_EPSILON      /* ConstantsDictionary.f90: TOL = epsilon(1.0_prec) */
#define KITMAX 50             /* ConstantsDictionary.f90: kItMax */

/* Legendre.f90:44-114 */
void orc_legendre_polynomial(int nP, double x, double *lAtX, double *dLdxAtX)
{
    if (nP == 0) { *lAtX = 1.0; *dLdxAtX = 0.0; return; }
    if (nP == 1) { *lAtX = x; *dLdxAtX = 1.0; return; }
    double lNm2 = 1.0, lNm1 = x, dlNm2 = 0.0, dlNm1 = 1.0, l = 0.0, dl = 0.0;
    for (int iP = 2; iP <= nP; ++iP) {
        double r = (double)iP;
        /* ((2*iP - 1)*x*lnM1 - (iP - 1.0)*lnM2)/iP */
        l = ((2.0 * r - 1.0) * x * lNm1 - (r - 1.0) * lNm2) / r;
        dl = dlNm2 + (2.0 * r - 1.0) * lNm1;
        lNm2 = lNm1; lNm1 = l; dlNm2 = dlNm1; dlNm1 = dl;
    }
    *lAtX = l; *dLdxAtX = dl;
}

/* Legendre.f90:118-174 */
void orc_legendre_q_and_l(int nP, double x, double *q, double *qprime, double *lN)
{
    double lNm2 = 1.0, lNm1 = x, dlNm2 = 0.0, dlNm1 = 1.0, l = 0.0, dl;
    for (int iP = 2; iP <= nP; ++iP) {
        double r = (double)iP;
        /* (TWO*iP - ONE)/REAL(iP)*x*lNm1 - (REAL(iP) - ONE)/REAL(iP)*lNm2 */
        l = (2.0 * iP - 1.0) / r * x * lNm1 - (r - 1.0) / r * lNm2;
        dl = dlNm2 + (2.0 * r - 1.0) * lNm1;
        lNm2 = lNm1; lNm1 = l; dlNm2 = dlNm1; dlNm1 = dl;
    }
    int iP = nP + 1;
    double r = (double)iP;
    double lNp1 = (2.0 * iP - 1.0) / r * x * l - (r - 1.0) / r * lNm2;
    double dlNp1 = dlNm2 + (2.0 * r - 1.0) * lNm1;
    *q = lNp1 - lNm2;
    *qprime = dlNp1 - dlNm2;
    *lN = l;
}

/* Legendre.f90:216-308.  Note the root loop runs jX = 0..(nP+1)/2 inclusive,
 * and the Newton stop test abs(delta) <= TOL*x never fires for x < 0 (Q14). */
void orc_legendre_gauss(int nP, double *x, double *w)
{
    double l, dl, delta;
    if (nP == 0) { x[0] = 0.0; w[0] = 2.0; }
    else if (nP == 1) {
        x[0] = -sqrt(1.0 / 3.0); w[0] = 1.0;
        x[1] = -x[0]; w[1] = w[0];
    } else {
        for (int jX = 0; jX <= (nP + 1) / 2; ++jX) {
            x[jX] = -cos((2.0 * (double)jX + 1.0) * PI_ / (2.0 * (double)nP + 1.0));
            for (int kIt = 1; kIt <= KITMAX; ++kIt) {
                orc_legendre_polynomial(nP + 1, x[jX], &l, &dl);
                delta = -l / dl;
                x[jX] = x[jX] + delta;
                if (fabs(delta) <= TOL_ * x[jX]) break;
            }
            orc_legendre_polynomial(nP + 1, x[jX], &l, &dl);
            w[jX] = 2.0 / ((1.0 - x[jX] * x[jX]) * dl * dl);
            w[nP - jX] = w[jX];
            x[nP - jX] = -x[jX];
        }
    }
    if (fmod((double)nP, 2.0) == 0.0) {
        orc_legendre_polynomial(nP + 1, 0.0, &l, &dl);
        x[nP / 2] = 0.0;
        w[nP / 2] = 2.0 / (dl * dl);
    }
}

/* Legendre.f90:312-405 */
void orc_legendre_gauss_lobatto(int nP, double *x, double *w)
{
    double q, qp, lN, delta;
    if (nP == 1) {
        x[0] = -1.0; w[0] = 1.0; x[1] = 1.0; w[1] = 1.0;
    } else {
        x[0] = -1.0;
        w[0] = 2.0 / ((double)nP * ((double)nP + 1.0));
        x[nP] = 1.0;
        w[nP] = w[0];
        for (int jX = 1; jX <= (nP + 1) / 2 - 1; ++jX) {
            x[jX] = -cos(((double)jX + 0.25) * PI_ / (double)nP -
                         3.0 / (8.0 * (double)nP * PI_ * ((double)jX + 0.25)));
            for (int kIt = 1; kIt <= KITMAX; ++kIt) {
                orc_legendre_q_and_l(nP, x[jX], &q, &qp, &lN);
                delta = -q / qp;
                x[jX] = x[jX] + delta;
                if (fabs(delta) <= TOL_ * x[jX]) break;
            }
            orc_legendre_q_and_l(nP, x[jX], &q, &qp, &lN);
            w[jX] = 2.0 / ((double)nP * ((double)nP + 1.0) * lN * lN);
            w[nP - jX] = w[jX];
            x[nP - jX] = -x[jX];
        }
    }
    if (fmod((double)nP, 2.0) == 0.0) {
        orc_legendre_q_and_l(nP, 0.0, &q, &qp, &lN);
        x[nP / 2] = 0.0;
        w[nP / 2] = 2.0 / ((double)nP * ((double)nP + 1.0) * lN * lN);
    }
}

/* Legendre.f90:180-212 */
int orc_generate_legendre_quadrature(int nP, double *x, double *w, int quadType)
{
    if (quadType == ORC_GAUSS_LOBATTO) { orc_legendre_gauss_lobatto(nP, x, w); return 0; }
    if (quadType == ORC_GAUSS) { orc_legendre_gauss(nP, x, w); return 0; }
    return -2;
}

/* CommonRoutines.f90:58-97 */
int orc_almost_equal(double a, double b)
{
    if (a == 0.0 || b == 0.0) return fabs(a - b) <= DBL_EPSILON;
    return (fabs(a - b) <= DBL_EPSILON * fabs(a)) || (fabs(a - b) <= DBL_EPSILON * fabs(b));
}

/* Lagrange_1D_Class.f90:349-392 */
void orc_barycentric_weights(int nP, const double *nodes, double *bw)
{
    for (int j = 0; j <= nP; ++j) bw[j] = 1.0;
    for (int jP = 1; jP <= nP; ++jP) {
        for (int kP = 0; kP <= jP - 1; ++kP) {
            bw[kP] = bw[kP] * (nodes[kP] - nodes[jP]);
            bw[jP] = bw[jP] * (nodes[jP] - nodes[kP]);
        }
    }
    for (int jP = 0; jP <= nP; ++jP) bw[jP] = 1.0 / bw[jP];
}

/* Lagrange_1D_Class.f90:448-515 */
void orc_lagrange_polynomials_at(int nP, const double *nodes, const double *bw,
                                 double xEval, double *lAtX)
{
    int match = 0;
    for (int jP = 0; jP <= nP; ++jP) {
        lAtX[jP] = 0.0;
        if (orc_almost_equal(xEval, nodes[jP])) { lAtX[jP] = 1.0; match = 1; }
    }
    if (match) return;
    double temp1 = 0.0;
    for (int jP = 0; jP <= nP; ++jP) {
        double temp2 = bw[jP] / (xEval - nodes[jP]);
        lAtX[jP] = temp2;
        temp1 = temp1 + temp2;
    }
    for (int jP = 0; jP <= nP; ++jP) lAtX[jP] = lAtX[jP] / temp1;
}

/* EvaluateLagrange_1D (Lagrange_1D_Class.f90, barycentric formula of the 2nd kind) */
double orc_lagrange_interpolate(int nP, const double *nodes, const double *bw,
                                double xEval, const double *f)
{
    double num = 0.0, den = 0.0;
    for (int jP = 0; jP <= nP; ++jP) {
        if (orc_almost_equal(xEval, nodes[jP])) return f[jP];
        double t = bw[jP] / (xEval - nodes[jP]);
        num = num + t * f[jP];
        den = den + t;
    }
    return num / den;
}

/* Lagrange_1D_Class.f90:519-606 */
double orc_lagrange_derivative(int nP, const double *nodes, const double *bw,
                               double xEval, const double *f)
{
    double num = 0.0, den = 0.0, p = 0.0;
    int atNode = 0, jS = -1;
    for (int jP = 0; jP <= nP; ++jP) {
        if (orc_almost_equal(xEval, nodes[jP])) {
            atNode = 1; p = f[jP]; den = -bw[jP]; jS = jP;
        }
    }
    if (atNode) {
        for (int jP = 0; jP <= nP; ++jP) {
            if (jP != jS) num = num + bw[jP] * (p - f[jP]) / (xEval - nodes[jP]);
        }
    } else {
        den = 0.0;
        p = orc_lagrange_interpolate(nP, nodes, bw, xEval, f);
        for (int jP = 0; jP <= nP; ++jP) {
            double t = bw[jP] / (xEval - nodes[jP]);
            num = num + t * (p - f[jP]) / (xEval - nodes[jP]);
            den = den + t;
        }
    }
    return num / den;
}

/* Lagrange_1D_Class.f90:612-689 (native nodes branch) */
void orc_derivative_matrix(int nP, const double *x, const double *w, double *dMat)
{
    int n1 = nP + 1;
    for (int kP = 0; kP <= nP; ++kP) {
        dMat[kP + n1 * kP] = 0.0;
        for (int jP = 0; jP <= nP; ++jP) {
            if (jP != kP) {
                dMat[kP + n1 * jP] = w[jP] / (w[kP] * (x[kP] - x[jP]));
                dMat[kP + n1 * kP] = dMat[kP + n1 * kP] - dMat[kP + n1 * jP];
            }
        }
    }
}

/* Lagrange_1D_Class.f90:693-764 */
void orc_interpolation_matrix(int nP, const double *x, const double *w,
                              int nNew, const double *zNew, double *T)
{
    int m1 = nNew + 1;
    for (int kP = 0; kP <= nNew; ++kP) {
        int match = 0;
        for (int jP = 0; jP <= nP; ++jP) {
            T[kP + m1 * jP] = 0.0;
            if (orc_almost_equal(zNew[kP], x[jP])) { match = 1; T[kP + m1 * jP] = 1.0; }
        }
        if (!match) {
            double temp1 = 0.0;
            for (int jP = 0; jP <= nP; ++jP) {
                double temp2 = w[jP] / (zNew[kP] - x[jP]);
                T[kP + m1 * jP] = temp2;
                temp1 = temp1 + temp2;
            }
            for (int jP = 0; jP <= nP; ++jP) T[kP + m1 * jP] = T[kP + m1 * jP] / temp1;
        }
    }
}

/* Lagrange_2D_Class.f90:488-528:  fInt = MATMUL(Ts,fOld); fIntT = TRANSPOSE(fInt);
 * fNew = MATMUL(Tp,fIntT).  The product is (0:nNewP,0:nNewS) and is stored
 * into fNew(0:nNewS,0:nNewP) element-for-element, so the first index of the
 * result runs over the NEW p-points (quirk: the output is transposed with
 * respect to fOld; it is square in every reference caller). */
void orc_coarse_to_fine_2d(int nS, int nP, const double *fOld, int nNewS, int nNewP,
                           const double *Ts, const double *Tp, double *fNew)
{
    int s1 = nS + 1, ms = nNewS + 1, mp = nNewP + 1;
    double *fInt = (double *)malloc(sizeof(double) * ms * (nP + 1));
    for (int j = 0; j <= nP; ++j)
        for (int a = 0; a <= nNewS; ++a) {
            double acc = 0.0;
            for (int k = 0; k <= nS; ++k) acc += Ts[a + ms * k] * fOld[k + s1 * j];
            fInt[a + ms * j] = acc;
        }
    /* fNew(a,b) = sum_j Tp(a,j) * fInt(b,j), a over new p, b over new s */
    for (int b = 0; b <= nNewS; ++b)
        for (int a = 0; a <= nNewP; ++a) {
            double acc = 0.0;
            for (int j = 0; j <= nP; ++j) acc += Tp[a + mp * j] * fInt[b + ms * j];
            fNew[a + mp * b] = acc;
        }
    free(fInt);
}

/* Lagrange_3D_Class.f90:545-604 (s, then p, then q 1-D CoarseToFine) */
void orc_coarse_to_fine_3d(int n, const double *fOld, int nNew, const double *T, double *fNew)
{
    int n1 = n + 1, m1 = nNew + 1;
    double *f1 = (double *)malloc(sizeof(double) * m1 * n1 * n1);
    double *f2 = (double *)malloc(sizeof(double) * m1 * m1 * n1);
    for (int jP = 0; jP <= n; ++jP)
        for (int jQ = 0; jQ <= n; ++jQ)
            for (int a = 0; a <= nNew; ++a) {
                double acc = 0.0;
                for (int k = 0; k <= n; ++k) acc += T[a + m1 * k] * fOld[k + n1 * (jP + n1 * jQ)];
                f1[a + m1 * (jP + n1 * jQ)] = acc;
            }
    for (int jS = 0; jS <= nNew; ++jS)
        for (int jQ = 0; jQ <= n; ++jQ)
            for (int a = 0; a <= nNew; ++a) {
                double acc = 0.0;
                for (int k = 0; k <= n; ++k) acc += T[a + m1 * k] * f1[jS + m1 * (k + n1 * jQ)];
                f2[jS + m1 * (a + m1 * jQ)] = acc;
            }
    for (int jS = 0; jS <= nNew; ++jS)
        for (int jP = 0; jP <= nNew; ++jP)
            for (int a = 0; a <= nNew; ++a) {
                double acc = 0.0;
                for (int k = 0; k <= n; ++k) acc += T[a + m1 * k] * f2[jS + m1 * (jP + m1 * k)];
                fNew[jS + m1 * (jP + m1 * a)] = acc;
            }
    free(f1); free(f2);
}

/* NodalStorage_2D_Class.f90:101-230 (nS = nP = N).  The DG branch stores
 *   dMatS(i,j) = -D(j,i) w(j)/w(i)   (:199-207)
 *   dMatP(j,i) = -D(j,i) w(j)/w(i)   (:209-218)  == dMatS^T  (quirk Q1)
 * dMatCorrect receives the weak operator used in BOTH directions by the
 * reference's 1-D and 3-D storages. */
int orc_nodal2d_build(orc_nodal2d *ns)
{
    int N = ns->N, n1 = N + 1;
    if (orc_generate_legendre_quadrature(N, ns->nodes, ns->qWeight, ns->quadrature) != 0) return -2;
    orc_barycentric_weights(N, ns->nodes, ns->bWeights);
    orc_lagrange_polynomials_at(N, ns->nodes, ns->bWeights, -1.0, ns->lagWest);
    orc_lagrange_polynomials_at(N, ns->nodes, ns->bWeights, 1.0, ns->lagEast);
    orc_lagrange_polynomials_at(N, ns->nodes, ns->bWeights, -1.0, ns->lagSouth);
    orc_lagrange_polynomials_at(N, ns->nodes, ns->bWeights, 1.0, ns->lagNorth);
    double *D = (double *)malloc(sizeof(double) * n1 * n1);
    orc_derivative_matrix(N, ns->nodes, ns->bWeights, D);
    if (ns->form == ORC_CG) {
        memcpy(ns->dMatS, D, sizeof(double) * n1 * n1);
        memcpy(ns->dMatP, D, sizeof(double) * n1 * n1);
        if (ns->dMatCorrect) memcpy(ns->dMatCorrect, D, sizeof(double) * n1 * n1);
    } else if (ns->form == ORC_DG) {
        const double *w = ns->qWeight;
        for (int jP = 0; jP <= N; ++jP)
            for (int iP = 0; iP <= N; ++iP)
                ns->dMatS[iP + n1 * jP] = -D[jP + n1 * iP] * w[jP] / w[iP];
        for (int jP = 0; jP <= N; ++jP)
            for (int iP = 0; iP <= N; ++iP)
                ns->dMatP[jP + n1 * iP] = -D[jP + n1 * iP] * w[jP] / w[iP];
        if (ns->dMatCorrect) memcpy(ns->dMatCorrect, ns->dMatS, sizeof(double) * n1 * n1);
    } else {
        free(D);
        return -2;
    }
    free(D);
    return 0;
}

/* NodalStorage_2D_Class.f90:620-646 */
void orc_apply_derivative_2d(int N, const double *dMatS, const double *dMatP,
                             const double *f, double *dFds, double *dFdp)
{
    int n1 = N + 1;
    for (int i = 0; i <= N; ++i)
        for (int k = 0; k <= N; ++k) {
            double acc = 0.0;
            for (int j = 0; j <= N; ++j) acc += dMatS[k + n1 * j] * f[j + n1 * i];
            dFds[k + n1 * i] = acc;
        }
    for (int i = 0; i <= N; ++i)
        for (int k = 0; k <= N; ++k) {
            double acc = 0.0;
            for (int j = 0; j <= N; ++j) acc += dMatP[k + n1 * j] * f[i + n1 * j];
            dFdp[i + n1 * k] = acc;
        }
}

/* NodalStorage_2D_Class.f90:650-680 */
void orc_boundary_interp_2d(int N, const double *lW, const double *lE,
                            const double *lS, const double *lN, const double *f,
                            double *fW, double *fE, double *fS, double *fN)
{
    int n1 = N + 1;
    for (int i = 0; i <= N; ++i) {
        double w = 0.0, e = 0.0;
        for (int k = 0; k <= N; ++k) { w += lW[k] * f[k + n1 * i]; e += lE[k] * f[k + n1 * i]; }
        fW[i] = w; fE[i] = e;
    }
    for (int i = 0; i <= N; ++i) {
        double s = 0.0, n = 0.0;
        for (int k = 0; k <= N; ++k) { s += lS[k] * f[i + n1 * k]; n += lN[k] * f[i + n1 * k]; }
        fS[i] = s; fN[i] = n;
    }
}

/* NodalStorage_3D_Class.f90:737-774 (same matrix in the three directions) */
void orc_apply_derivative_3d(int N, const double *dMat, const double *f,
                             double *dFds, double *dFdp, double *dFdq)
{
    int n1 = N + 1;
#define F3(a, b, c) f[(a) + n1 * ((b) + n1 * (c))]
    for (int i = 0; i <= N; ++i)
        for (int j = 0; j <= N; ++j)
            for (int k = 0; k <= N; ++k) {
                double acc = 0.0;
                for (int m = 0; m <= N; ++m) acc += dMat[k + n1 * m] * F3(m, i, j);
                dFds[k + n1 * (i + n1 * j)] = acc;
            }
    for (int i = 0; i <= N; ++i)
        for (int j = 0; j <= N; ++j)
            for (int k = 0; k <= N; ++k) {
                double acc = 0.0;
                for (int m = 0; m <= N; ++m) acc += dMat[k + n1 * m] * F3(i, m, j);
                dFdp[i + n1 * (k + n1 * j)] = acc;
            }
    for (int i = 0; i <= N; ++i)
        for (int j = 0; j <= N; ++j)
            for (int k = 0; k <= N; ++k) {
                double acc = 0.0;
                for (int m = 0; m <= N; ++m) acc += dMat[k + n1 * m] * F3(i, j, m);
                dFdq[i + n1 * (j + n1 * k)] = acc;
            }
#undef F3
}

/* NodalStorage_3D_Class.f90:778-823; faces written south, east, north, west,
 * bottom, top, each (0:N,0:N) over the two in-face indices in ascending
 * direction order (s before p before q). */
void orc_boundary_interp_3d(int N, const double *lL, const double *lR,
                            const double *f, double *faces)
{
    int n1 = N + 1, nf = n1 * n1;
    double *fS = faces, *fE = faces + nf, *fN = faces + 2 * nf, *fW = faces + 3 * nf;
    double *fB = faces + 4 * nf, *fT = faces + 5 * nf;
#define F3(a, b, c) f[(a) + n1 * ((b) + n1 * (c))]
    for (int i = 0; i <= N; ++i)
        for (int j = 0; j <= N; ++j) {
            double w = 0.0, e = 0.0;
            for (int k = 0; k <= N; ++k) { w += lL[k] * F3(k, i, j); e += lR[k] * F3(k, i, j); }
            fW[i + n1 * j] = w; fE[i + n1 * j] = e;
        }
    for (int i = 0; i <= N; ++i)
        for (int j = 0; j <= N; ++j) {
            double s = 0.0, n = 0.0;
            for (int k = 0; k <= N; ++k) { s += lL[k] * F3(i, k, j); n += lR[k] * F3(i, k, j); }
            fS[i + n1 * j] = s; fN[i + n1 * j] = n;
        }
    for (int i = 0; i <= N; ++i)
        for (int j = 0; j <= N; ++j) {
            double b = 0.0, t = 0.0;
            for (int k = 0; k <= N; ++k) { b += lL[k] * F3(i, j, k); t += lR[k] * F3(i, j, k); }
            fB[i + n1 * j] = b; fT[i + n1 * j] = t;
        }
#undef F3
}
