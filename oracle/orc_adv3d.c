/*
 * orc_adv3d.c -- oracle (test infrastructure only): structured hex mesh and
 * 3-D linear advection DGSEM with the adaptive roll-off filter, restated from
 *   src/highend/advection3d/Advection3DClass.f90
 *     ForwardStepRK3 :1000-1045, GlobalTimeDerivative :1049-1121,
 *     DoTheAdaptiveFiltering :1125-1233, EdgeFlux :1237-1341,
 *     MappedTimeDerivative :1345-1562, DGSystemDerivative :1566-1591,
 *     RiemannSolver :1595-1620, X/Y/ZFlux :1624-1672, Source :1676-1694,
 *     GetExternalState :1698-1735
 *   src/filters/RollOffFilter3D_Class.f90 Build :87-180, ApplyFilter :273-292
 *   src/dgsem/DGSEMSolutionStorageClass_3D.f90:833-910 (traces)
 * The hex mesh is this build's structured generator (the reference's
 * HexMesh default mesh has no periodic faces, quirk Q3); geometry uses the
 * analytic derivatives of the warp map, and face normals are the
 * cofactor rows of the covariant metric at the face (outward, |n| = area
 * element), the quantities MappedGeometry_3D provides.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"
#include "orc_adv3d.h"

static const double rk3_a[3] = {0.0, -5.0 / 9.0, -153.0 / 128.0};
static const double rk3_b[3] = {0.0, 1.0 / 3.0, 3.0 / 4.0};
static const double rk3_g[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};

/* ------------------------------------------------------------------ mesh */
static int nfaces_dir(int n, int periodic, int a, int b) { return (periodic ? n : n + 1) * a * b; }

void orc_hexmesh_counts(orc_hexmesh_desc *d)
{
    d->nElems = d->nX * d->nY * d->nZ;
    d->nFaces = nfaces_dir(d->nX, d->periodicX, d->nY, d->nZ) + nfaces_dir(d->nY, d->periodicY, d->nX, d->nZ) +
                nfaces_dir(d->nZ, d->periodicZ, d->nX, d->nY);
}

/* position and d(x,y,z)/d(s,p,q) of reference point (s,p,q) of element (iX,iY,iZ) */
static void map_point(const orc_hexmesh_desc *d, int iX, int iY, int iZ, double s, double p, double q, double X[3],
                      double cov[9])
{
    const double tp = 8.0 * atan(1.0);
    double a = (iX + 0.5 * (s + 1.0)) / d->nX, b = (iY + 0.5 * (p + 1.0)) / d->nY, c = (iZ + 0.5 * (q + 1.0)) / d->nZ;
    double dads = 0.5 / d->nX, dbdp = 0.5 / d->nY, dcdq = 0.5 / d->nZ;
    double sa = sin(tp * a), sb = sin(tp * b), sc = sin(tp * c);
    double ca = cos(tp * a), cb = cos(tp * b), cc = cos(tp * c);
    double S = sa * sb * sc, Sa = tp * ca * sb * sc, Sb = tp * sa * cb * sc, Sc = tp * sa * sb * cc;
    double A = d->warp;
    double base[3] = {a, b, c};
    double scale[3] = {d->xScale, d->yScale, d->zScale};
    double dref[3] = {dads, dbdp, dcdq};
    for (int r = 0; r < 3; ++r) {
        double g = A * d->warpSign[r];
        X[r] = scale[r] * (base[r] + g * S);
        double dX[3] = {g * Sa, g * Sb, g * Sc};
        dX[r] += 1.0;
        for (int k = 0; k < 3; ++k) cov[3 * r + k] = scale[r] * dX[k] * dref[k];
    }
}

/* cofactor row for direction dir (0: s, 1: p, 2: q) as the reference writes
 * the contravariant flux (MappedTimeDerivative :1412-1414, :1460-1462, :1510-1512) */
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

static const int opposite[7] = {0, 3, 4, 1, 2, 6, 5};

int orc_hexmesh_build(const orc_hexmesh_desc *d, const double *nodes, int *faces, double *x, double *y,
                      double *z, double *cov, double *J, double *nHat, double *nLen, double *bpos)
{
    int N = d->N, n1 = N + 1, n2 = n1 * n1, n3 = n2 * n1;
    int nE = d->nX * d->nY * d->nZ;
    if (N < 1 || d->nX < 1 || d->nY < 1 || d->nZ < 1) return -1;
    /* volume geometry */
    for (int iZ = 0; iZ < d->nZ; ++iZ)
        for (int iY = 0; iY < d->nY; ++iY)
            for (int iX = 0; iX < d->nX; ++iX) {
                int e = iX + d->nX * (iY + d->nY * iZ);
                for (int k = 0; k < n1; ++k)
                    for (int j = 0; j < n1; ++j)
                        for (int i = 0; i < n1; ++i) {
                            int id = e * n3 + i + n1 * (j + n1 * k);
                            double X[3], m[9], cs[3];
                            map_point(d, iX, iY, iZ, nodes[i], nodes[j], nodes[k], X, m);
                            x[id] = X[0];
                            y[id] = X[1];
                            z[id] = X[2];
                            memcpy(cov + 9 * (size_t)id, m, sizeof(m));
                            cofactor(m, 0, cs);
                            J[id] = m[0] * cs[0] + m[3] * cs[1] + m[6] * cs[2];
                        }
                /* faces: node (a, b) of each side */
                for (int side = 1; side <= 6; ++side)
                    for (int bb = 0; bb < n1; ++bb)
                        for (int aa = 0; aa < n1; ++aa) {
                            double s = 0, p = 0, q = 0, sign = 1.0;
                            int dir = 0;
                            switch (side) {
                            case 1: s = nodes[aa]; p = -1.0; q = nodes[bb]; dir = 1; sign = -1.0; break;
                            case 2: s = 1.0; p = nodes[aa]; q = nodes[bb]; dir = 0; break;
                            case 3: s = nodes[aa]; p = 1.0; q = nodes[bb]; dir = 1; break;
                            case 4: s = -1.0; p = nodes[aa]; q = nodes[bb]; dir = 0; sign = -1.0; break;
                            case 5: s = nodes[aa]; p = nodes[bb]; q = -1.0; dir = 2; sign = -1.0; break;
                            case 6: s = nodes[aa]; p = nodes[bb]; q = 1.0; dir = 2; break;
                            }
                            double X[3], m[9], cf[3];
                            map_point(d, iX, iY, iZ, s, p, q, X, m);
                            cofactor(m, dir, cf);
                            double nx = sign * cf[0], ny = sign * cf[1], nz = sign * cf[2];
                            double len = sqrt(nx * nx + ny * ny + nz * nz);
                            size_t fb = (size_t)aa + n1 * bb + (size_t)n2 * ((side - 1) + 6 * (size_t)e);
                            nHat[3 * fb] = nx / len;
                            nHat[3 * fb + 1] = ny / len;
                            nHat[3 * fb + 2] = nz / len;
                            nLen[fb] = len;
                            bpos[3 * fb] = X[0];
                            bpos[3 * fb + 1] = X[1];
                            bpos[3 * fb + 2] = X[2];
                        }
            }
    /* face table: elements in order, sides 1..6, first visit is primary */
    int *seen = (int *)calloc((size_t)nE * 7, sizeof(int));
    int nf = 0;
    for (int iZ = 0; iZ < d->nZ; ++iZ)
        for (int iY = 0; iY < d->nY; ++iY)
            for (int iX = 0; iX < d->nX; ++iX) {
                int e = iX + d->nX * (iY + d->nY * iZ);
                for (int side = 1; side <= 6; ++side) {
                    if (seen[7 * e + side]) continue;
                    int jX = iX, jY = iY, jZ = iZ, ok = 1;
                    switch (side) {
                    case 1: jY--; if (jY < 0) { if (d->periodicY) jY += d->nY; else ok = 0; } break;
                    case 2: jX++; if (jX >= d->nX) { if (d->periodicX) jX -= d->nX; else ok = 0; } break;
                    case 3: jY++; if (jY >= d->nY) { if (d->periodicY) jY -= d->nY; else ok = 0; } break;
                    case 4: jX--; if (jX < 0) { if (d->periodicX) jX += d->nX; else ok = 0; } break;
                    case 5: jZ--; if (jZ < 0) { if (d->periodicZ) jZ += d->nZ; else ok = 0; } break;
                    case 6: jZ++; if (jZ >= d->nZ) { if (d->periodicZ) jZ -= d->nZ; else ok = 0; } break;
                    }
                    int *f = faces + 9 * nf;
                    f[0] = e + 1;
                    f[1] = side;
                    f[4] = 0; f[5] = 0; f[6] = 1; f[7] = 1; f[8] = 0;
                    seen[7 * e + side] = 1;
                    if (ok) {
                        int e2 = jX + d->nX * (jY + d->nY * jZ);
                        f[2] = e2 + 1;
                        f[3] = opposite[side];
                        seen[7 * e2 + opposite[side]] = 1;
                    } else {
                        f[2] = ORC_NO_NORMAL_FLOW;
                        f[3] = 0;
                    }
                    nf++;
                }
            }
    free(seen);
    return nf == d->nFaces ? 0 : -2;
}

/* ------------------------------------------------------------------ filter */
void orc_rolloff_filter3d_build(int N, int nc, const double *nodes, const double *w, double *F)
{
    int n1 = N + 1, nD = n1 * n1 * n1;
    double *V = (double *)calloc((size_t)nD * nD, sizeof(double));
    double *VI = (double *)calloc((size_t)nD * nD, sizeof(double));
    double *P = (double *)calloc((size_t)nD, sizeof(double));
    double sc = (double)nc, alpha = log(2.0) / (sc * sc);
    for (int c = 0; c <= N; ++c)
        for (int b = 0; b <= N; ++b)
            for (int a = 0; a <= N; ++a) {
                int row = a + (b + c * n1) * n1;
                double Lnorm = 0.0, r = (double)(a * a + b * b + c * c), Ls;
                P[row] = exp(-alpha * r);
                for (int k = 0; k <= N; ++k) {
                    double Lk, Lj, Li;
                    orc_legendre_polynomial(c, nodes[k], &Lk, &Ls);
                    for (int j = 0; j <= N; ++j) {
                        orc_legendre_polynomial(b, nodes[j], &Lj, &Ls);
                        for (int i = 0; i <= N; ++i) {
                            int col = i + (j + k * n1) * n1;
                            orc_legendre_polynomial(a, nodes[i], &Li, &Ls);
                            double L3 = Li * Lj * Lk, w3 = w[i] * w[j] * w[k];
                            Lnorm = Lnorm + (L3 * L3) * w3;
                            V[row + (size_t)nD * col] = L3 * w3;
                            VI[col + (size_t)nD * row] = L3;
                        }
                    }
                }
                for (int col = 0; col < nD; ++col) V[row + (size_t)nD * col] = V[row + (size_t)nD * col] / Lnorm;
            }
    /* Pfilt = diag(P) V (the zero products of MATMUL add nothing), then
     * filterMat = VInv Pfilt with the inner index ascending */
    for (int col = 0; col < nD; ++col)
        for (int row = 0; row < nD; ++row) V[row + (size_t)nD * col] = P[row] * V[row + (size_t)nD * col];
    for (int col = 0; col < nD; ++col)
        for (int row = 0; row < nD; ++row) {
            double acc = 0.0;
            for (int m = 0; m < nD; ++m) acc += VI[row + (size_t)nD * m] * V[m + (size_t)nD * col];
            F[row + (size_t)nD * col] = acc;
        }
    free(V);
    free(VI);
    free(P);
}

void orc_rolloff_filter3d_apply(int N, const double *F, const double *sol, double *fsol)
{
    int n1 = N + 1, nD = n1 * n1 * n1;
    for (int row = 0; row < nD; ++row) {
        double acc = 0.0;
        for (int col = 0; col < nD; ++col) acc += F[row + (size_t)nD * col] * sol[col];
        fsol[row] = acc;
    }
}

/* ------------------------------------------------------------------ model */
#define VOL(i, j, k, f, e) ((i) + n1 * ((j) + n1 * ((k) + n1 * ((f) + nq * (size_t)(e)))))
#define GEO(i, j, k, e) ((i) + n1 * ((j) + n1 * ((k) + n1 * (size_t)(e))))
#define TR(a, f, s, e) ((a) + n2 * ((f) + nq * ((s) + 6 * (size_t)(e))))
#define FB(a, s, e) ((a) + n2 * ((s) + 6 * (size_t)(e)))

void orc_adv3d_adaptive_filter(orc_adv3d *m, int stage)
{
    int N = m->N, n1 = N + 1, nq = m->nq, n3 = n1 * n1 * n1, nE = m->nElems;
    double *solf = (double *)malloc(sizeof(double) * n3);
    double *c1 = (double *)malloc(sizeof(double) * n3), *c2 = (double *)malloc(sizeof(double) * n3);
    const double *w = m->qWeight;
    m->nFiltered = 0;
    for (int e = 0; e < nE; ++e)
        for (int f = 0; f < nq; ++f) {
            double *sol = m->c + VOL(0, 0, 0, f, e);
            orc_rolloff_filter3d_apply(N, m->F, sol, solf);
            for (int id = 0; id < n3; ++id) {
                double dlt = sol[id] - solf[id];
                c2[id] = dlt * dlt;
                c1[id] = solf[id] * solf[id];
            }
            double E1prior = m->E1[e * nq + f], E2prior = m->E2[e * nq + f];
            double E1 = 0.0, E2 = 0.0;
            for (int id = 0; id < n3; ++id) {
                double Jv = m->J[(size_t)e * n3 + id];
                c1[id] = c1[id] * Jv;
                c2[id] = c2[id] * Jv;
            }
            for (int iQ = 0; iQ <= N; ++iQ)
                for (int iP = 0; iP <= N; ++iP) {
                    double d1 = 0.0, d2 = 0.0;
                    for (int iS = 0; iS <= N; ++iS) {
                        d1 += c1[iS + n1 * (iP + n1 * iQ)] * w[iS];
                        d2 += c2[iS + n1 * (iP + n1 * iQ)] * w[iS];
                    }
                    E1 = E1 + d1 * w[iP] * w[iQ];
                    E2 = E2 + d2 * w[iP] * w[iQ];
                }
            m->E1[e * nq + f] = E1;
            m->E2[e * nq + f] = E2;
            double dE1 = E1 - E1prior, dE2 = E2 - E2prior, xi = E2 / E1;
            double *lim = m->lim + (size_t)nq * nE * (stage - 1) + e * nq + f;
            if (dE2 > 0.0 && fabs(dE1) / dE2 > 1.0) {
                *lim = xi;
            } else if (dE2 < 0.0) {
                *lim = xi;
            } else if (stage > 1) {
                *lim = *(lim - (size_t)nq * nE);
            }
            if (xi > 1.05 * (*lim)) {
                memcpy(sol, solf, sizeof(double) * n3);
                m->nFiltered++;
            }
        }
    free(solf);
    free(c1);
    free(c2);
}

static void traces3(int N, int nq, int nE, const double *lL, const double *lR, const double *q, double *out)
{
    int n1 = N + 1, n2 = n1 * n1;
    for (int e = 0; e < nE; ++e)
        for (int f = 0; f < nq; ++f) {
            for (int b = 0; b <= N; ++b)
                for (int a = 0; a <= N; ++a) {
                    double w = 0, ea = 0, s = 0, no = 0, bo = 0, to = 0;
                    for (int k = 0; k <= N; ++k) {
                        w += lL[k] * q[VOL(k, a, b, f, e)];
                        ea += lR[k] * q[VOL(k, a, b, f, e)];
                        s += lL[k] * q[VOL(a, k, b, f, e)];
                        no += lR[k] * q[VOL(a, k, b, f, e)];
                        bo += lL[k] * q[VOL(a, b, k, f, e)];
                        to += lR[k] * q[VOL(a, b, k, f, e)];
                    }
                    int node = a + n1 * b;
                    out[TR(node, f, 3, e)] = w;
                    out[TR(node, f, 1, e)] = ea;
                    out[TR(node, f, 0, e)] = s;
                    out[TR(node, f, 2, e)] = no;
                    out[TR(node, f, 4, e)] = bo;
                    out[TR(node, f, 5, e)] = to;
                }
        }
}

static double external3(int bc, int f, double x, double y, double z, double t)
{
    if (bc == ORC_INFLOW_ONE) return f == 0 ? tanh(t) : 0.0;
    if (bc == ORC_INFLOW_TWO) return f == 1 ? tanh(t) : 0.0;
    if (bc == ORC_NO_NORMAL_FLOW)
        return exp(-((x - t) * (x - t) + (y - t) * (y - t) + (z - t) * (z - t)) / (0.2 * 0.2));
    return 0.0;
}

static void edge_flux3(orc_adv3d *m, double t)
{
    int N = m->N, n1 = N + 1, n2 = n1 * n1, nq = m->nq;
    for (int iF = 0; iF < m->nFaces; ++iF) {
        const int *fc = m->faces + 9 * iF;
        int e1 = fc[0] - 1, s1 = fc[1] - 1, e2 = fc[2], s2 = abs(fc[3]) - 1;
        int iS = fc[4], jS = fc[5], iI = fc[6], jI = fc[7], sw = fc[8];
        for (int j = 0; j <= N; ++j)
            for (int i = 0; i <= N; ++i) {
                int node = i + n1 * j;
                size_t fb = FB(node, s1, e1);
                double nx = m->nHat[3 * fb], ny = m->nHat[3 * fb + 1], nz = m->nHat[3 * fb + 2], len = m->nLen[fb];
                double uN = m->bu[fb] * nx + m->bv[fb] * ny + m->bw[fb] * nz;
                int k = sw ? jS + j * jI : iS + i * iI;
                int l = sw ? iS + i * iI : jS + j * jI;
                for (int f = 0; f < nq; ++f) {
                    double in = m->bc[TR(node, f, s1, e1)], ex;
                    if (e2 > 0) ex = m->bc[TR(k + n1 * l, f, s2, e2 - 1)];
                    else ex = external3(e2, f, m->bpos[3 * fb], m->bpos[3 * fb + 1], m->bpos[3 * fb + 2], t);
                    double jump = ex - in;
                    double fl = 0.5 * (uN * (ex + in) - fabs(uN) * jump) * len;
                    m->fc[TR(node, f, s1, e1)] = fl;
                    if (e2 > 0) m->fc[TR(k + n1 * l, f, s2, e2 - 1)] = -fl;
                }
            }
    }
}

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

static void mapped3(orc_adv3d *m)
{
    int N = m->N, n1 = N + 1, n2 = n1 * n1, nq = m->nq;
    double cf[16], der[16], cs[3];
    for (int e = 0; e < m->nElems; ++e)
        for (int f = 0; f < nq; ++f) {
            for (int iQ = 0; iQ <= N; ++iQ)
                for (int iP = 0; iP <= N; ++iP) {
                    for (int iS = 0; iS <= N; ++iS) {
                        size_t g = GEO(iS, iP, iQ, e);
                        double c = m->c[VOL(iS, iP, iQ, f, e)];
                        cofactor(m->cov + 9 * g, 0, cs);
                        cf[iS] = cs[0] * (c * m->u[g]) + cs[1] * (c * m->v[g]) + cs[2] * (c * m->w[g]);
                    }
                    int node = iP + n1 * iQ;
                    dgsd(N, m->dMat, m->qWeight, m->fc[TR(node, f, 3, e)], m->fc[TR(node, f, 1, e)], cf,
                         m->lagLeft, m->lagRight, der);
                    for (int iS = 0; iS <= N; ++iS) m->tend[VOL(iS, iP, iQ, f, e)] = -der[iS];
                }
            for (int iQ = 0; iQ <= N; ++iQ)
                for (int iS = 0; iS <= N; ++iS) {
                    for (int iP = 0; iP <= N; ++iP) {
                        size_t g = GEO(iS, iP, iQ, e);
                        double c = m->c[VOL(iS, iP, iQ, f, e)];
                        cofactor(m->cov + 9 * g, 1, cs);
                        cf[iP] = cs[0] * (c * m->u[g]) + cs[1] * (c * m->v[g]) + cs[2] * (c * m->w[g]);
                    }
                    int node = iS + n1 * iQ;
                    dgsd(N, m->dMat, m->qWeight, m->fc[TR(node, f, 0, e)], m->fc[TR(node, f, 2, e)], cf,
                         m->lagLeft, m->lagRight, der);
                    for (int iP = 0; iP <= N; ++iP)
                        m->tend[VOL(iS, iP, iQ, f, e)] = m->tend[VOL(iS, iP, iQ, f, e)] - der[iP];
                }
            for (int iP = 0; iP <= N; ++iP)
                for (int iS = 0; iS <= N; ++iS) {
                    for (int iQ = 0; iQ <= N; ++iQ) {
                        size_t g = GEO(iS, iP, iQ, e);
                        double c = m->c[VOL(iS, iP, iQ, f, e)];
                        cofactor(m->cov + 9 * g, 2, cs);
                        cf[iQ] = cs[0] * (c * m->u[g]) + cs[1] * (c * m->v[g]) + cs[2] * (c * m->w[g]);
                    }
                    int node = iS + n1 * iP;
                    /* the reference passes qWeightY here (Q11); the weights are equal */
                    dgsd(N, m->dMat, m->qWeight, m->fc[TR(node, f, 4, e)], m->fc[TR(node, f, 5, e)], cf,
                         m->lagLeft, m->lagRight, der);
                    for (int iQ = 0; iQ <= N; ++iQ) {
                        double Jv = m->J[GEO(iS, iP, iQ, e)];
                        m->tend[VOL(iS, iP, iQ, f, e)] = (m->tend[VOL(iS, iP, iQ, f, e)] - der[iQ]) / Jv;
                    }
                }
            if (m->relax)
                for (int id = 0; id < n1 * n2; ++id) {
                    size_t v = VOL(0, 0, 0, f, e) + id;
                    m->tend[v] = m->tend[v] + (m->relax[v] - m->c[v]) * m->rFac[v];
                }
        }
}

void orc_adv3d_global_time_derivative(orc_adv3d *m, double t, int stage)
{
    int N = m->N, n1 = N + 1, n2 = n1 * n1, nE = m->nElems;
    if (m->useFilter) orc_adv3d_adaptive_filter(m, stage);
    traces3(N, m->nq, nE, m->lagLeft, m->lagRight, m->c, m->bc);
    /* CalculateVelocityAtBoundaries: velocity traces in face order */
    double *tmp = (double *)malloc(sizeof(double) * (size_t)n2 * 6 * nE);
    const double *vel[3] = {m->u, m->v, m->w};
    double *bv[3] = {m->bu, m->bv, m->bw};
    for (int d = 0; d < 3; ++d) {
        traces3(N, 1, nE, m->lagLeft, m->lagRight, vel[d], tmp);
        memcpy(bv[d], tmp, sizeof(double) * (size_t)n2 * 6 * nE);
    }
    free(tmp);
    edge_flux3(m, t);
    mapped3(m);
}

void orc_adv3d_forward_step_rk3(orc_adv3d *m, double tn)
{
    size_t n = (size_t)(m->N + 1) * (m->N + 1) * (m->N + 1) * m->nq * m->nElems;
    memset(m->G, 0, sizeof(double) * n);
    for (int st = 0; st < 3; ++st) {
        orc_adv3d_global_time_derivative(m, tn + rk3_b[st] * m->dt, st + 1);
        for (size_t i = 0; i < n; ++i) {
            m->G[i] = rk3_a[st] * m->G[i] + m->tend[i];
            m->c[i] = m->c[i] + rk3_g[st] * m->dt * m->G[i];
        }
    }
}
