/*
 * orc_mesh2d.c -- oracle (test infrastructure only): structured quad mesh,
 * edge connectivity and element geometry.
 *   node/element numbering : src/geom/QuadMeshClass.f90:1718-1859 (LoadDefaultMesh)
 *   edge discovery         : src/geom/QuadMeshClass.f90:1489-1627 (ConstructEdges)
 *   start/inc convention   : src/geom/QuadMeshClass.f90:1830-1846
 *   covariant metrics, J,
 *   boundary normals       : src/geom/MappedGeometryClass_2D.f90:611-710
 *   scaling                : src/geom/MappedGeometryClass_2D.f90:921-948
 * Periodic pairing is an addition (the reference has none, quirk Q3): the
 * last node column/row is identified with the first, so the wrap edges are
 * found by the same edge-discovery loop as interior edges.
 */
#include <math.h>
#include <stdlib.h>
#include "oracle.h"

static const int edgeMap[2][4] = {{1, 2, 4, 1}, {2, 3, 3, 4}}; /* QuadMeshClass.f90:194-195 */

void orc_quadmesh_counts(orc_quadmesh_desc *d)
{
    int nxn = d->periodicX ? d->nX : d->nX + 1;
    int nyn = d->periodicY ? d->nY : d->nY + 1;
    d->nNodes = nxn * nyn;
    d->nElems = d->nX * d->nY;
    int ex = d->periodicX ? d->nX : d->nX + 1; /* vertical edges per row */
    int ey = d->periodicY ? d->nY : d->nY + 1; /* horizontal edges per column */
    d->nEdges = ex * d->nY + ey * d->nX;
}

static int node_id(const orc_quadmesh_desc *d, int iX, int iY) /* 1-based in, 1-based out */
{
    int nxn = d->periodicX ? d->nX : d->nX + 1;
    int nyn = d->periodicY ? d->nY : d->nY + 1;
    int ix = (iX - 1) % nxn, iy = (iY - 1) % nyn;
    return ix + 1 + iy * nxn;
}

/* unit-square warp used for curvilinear tests */
static void warp_map(double A, double a, double b, double *X, double *Y,
                     double *Xa, double *Xb, double *Ya, double *Yb)
{
    double tp = 2.0 * 4.0 * atan(1.0);
    double sa = sin(tp * a), sb = sin(tp * b), ca = cos(tp * a), cb = cos(tp * b);
    double w = A * sa * sb;
    *X = a + w; *Y = b + w;
    double wa = A * tp * ca * sb, wb = A * tp * sa * cb;
    *Xa = 1.0 + wa; *Xb = wb; *Ya = wa; *Yb = 1.0 + wb;
}

/* covariant metric tensor covT = [[dxds, dxdp],[dyds, dydp]] at (s,p) for the
 * bilinear element with unit-square corners c[4][2] (SW,SE,NE,NW) */
static void element_metrics(const orc_quadmesh_desc *d, const double c[4][2], double s, double p,
                            double *x, double *y, double covT[2][2])
{
    /* bilinear map (transfinite interpolation of straight sides) */
    double a = 0.25 * ((1 - s) * (1 - p) * c[0][0] + (1 + s) * (1 - p) * c[1][0] +
                       (1 + s) * (1 + p) * c[2][0] + (1 - s) * (1 + p) * c[3][0]);
    double b = 0.25 * ((1 - s) * (1 - p) * c[0][1] + (1 + s) * (1 - p) * c[1][1] +
                       (1 + s) * (1 + p) * c[2][1] + (1 - s) * (1 + p) * c[3][1]);
    double as = 0.25 * (-(1 - p) * c[0][0] + (1 - p) * c[1][0] + (1 + p) * c[2][0] - (1 + p) * c[3][0]);
    double ap = 0.25 * (-(1 - s) * c[0][0] - (1 + s) * c[1][0] + (1 + s) * c[2][0] + (1 - s) * c[3][0]);
    double bs = 0.25 * (-(1 - p) * c[0][1] + (1 - p) * c[1][1] + (1 + p) * c[2][1] - (1 + p) * c[3][1]);
    double bp = 0.25 * (-(1 - s) * c[0][1] - (1 + s) * c[1][1] + (1 + s) * c[2][1] + (1 - s) * c[3][1]);
    double X = a, Y = b, Xa = 1, Xb = 0, Ya = 0, Yb = 1;
    if (d->warp != 0.0) warp_map(d->warp, a, b, &X, &Y, &Xa, &Xb, &Ya, &Yb);
    *x = d->xScale * X;
    *y = d->yScale * Y;
    covT[0][0] = d->xScale * (Xa * as + Xb * bs);
    covT[0][1] = d->xScale * (Xa * ap + Xb * bp);
    covT[1][0] = d->yScale * (Ya * as + Yb * bs);
    covT[1][1] = d->yScale * (Ya * ap + Yb * bp);
}

typedef struct { int key2, edge; } kv;

int orc_quadmesh_build(const orc_quadmesh_desc *d, const double *sn,
                       int *elemNodes, int *edges,
                       double *x, double *y, double *dxds, double *dxdp,
                       double *dyds, double *dydp, double *J,
                       double *nHat, double *nLen)
{
    int nX = d->nX, nY = d->nY, N = d->N, n1 = N + 1;
    /* with fewer than 3 elements across a periodic direction two distinct
     * edges share the same node pair and edge discovery is ambiguous */
    if ((d->periodicX && nX < 3) || (d->periodicY && nY < 3)) return -5;
    orc_quadmesh_desc dd = *d;
    orc_quadmesh_counts(&dd);
    int nE = dd.nElems;
    /* element nodes (1-based ids) SW, SE, NE, NW */
    for (int iY = 1; iY <= nY; ++iY)
        for (int iX = 1; iX <= nX; ++iX) {
            int e = iX + (iY - 1) * nX - 1;
            elemNodes[4 * e + 0] = node_id(d, iX, iY);
            elemNodes[4 * e + 1] = node_id(d, iX + 1, iY);
            elemNodes[4 * e + 2] = node_id(d, iX + 1, iY + 1);
            elemNodes[4 * e + 3] = node_id(d, iX, iY + 1);
        }
    /* ConstructEdges: per-node lists keyed by (min,max) node ids */
    int nNodes = dd.nNodes;
    int *cnt = (int *)calloc(nNodes + 1, sizeof(int));
    kv *tab = (kv *)malloc(sizeof(kv) * (nNodes + 1) * 8);
    int nEdges = 0;
    for (int e = 0; e < nE; ++e) {
        for (int k = 1; k <= 4; ++k) {
            int startID = elemNodes[4 * e + edgeMap[0][k - 1] - 1];
            int endID = elemNodes[4 * e + edgeMap[1][k - 1] - 1];
            int key1 = startID < endID ? startID : endID;
            int key2 = startID < endID ? endID : startID;
            int found = -1;
            for (int q = 0; q < cnt[key1]; ++q)
                if (tab[key1 * 8 + q].key2 == key2) { found = tab[key1 * 8 + q].edge; break; }
            if (found >= 0) {
                int *ed = edges + 6 * found;
                int e1 = ed[0], s1 = ed[1];
                int n1id = elemNodes[4 * (e1 - 1) + edgeMap[0][s1 - 1] - 1];
                ed[2] = e + 1;
                ed[3] = (startID == n1id) ? k : -k;
            } else {
                int *ed = edges + 6 * nEdges;
                ed[0] = e + 1; ed[1] = k; ed[2] = ORC_NO_NORMAL_FLOW; ed[3] = 0;
                tab[key1 * 8 + cnt[key1]].key2 = key2;
                tab[key1 * 8 + cnt[key1]].edge = nEdges;
                cnt[key1]++;
                nEdges++;
            }
        }
    }
    free(cnt); free(tab);
    if (nEdges != dd.nEdges) return -2;
    for (int ie = 0; ie < nEdges; ++ie) {
        int *ed = edges + 6 * ie;
        if (ed[3] < 0) { ed[4] = N - 1; ed[5] = -1; }
        else { ed[4] = 1; ed[5] = 1; }
    }
    /* geometry */
    double dxE = 1.0 / nX, dyE = 1.0 / nY;
    for (int iY = 1; iY <= nY; ++iY)
        for (int iX = 1; iX <= nX; ++iX) {
            int e = iX + (iY - 1) * nX - 1;
            double c[4][2];
            c[0][0] = dxE * (iX - 1); c[0][1] = dyE * (iY - 1);
            c[1][0] = dxE * iX;       c[1][1] = dyE * (iY - 1);
            c[2][0] = dxE * iX;       c[2][1] = dyE * iY;
            c[3][0] = dxE * (iX - 1); c[3][1] = dyE * iY;
            for (int j = 0; j <= N; ++j)
                for (int i = 0; i <= N; ++i) {
                    double covT[2][2], xx, yy;
                    element_metrics(d, c, sn[i], sn[j], &xx, &yy, covT);
                    int id = i + n1 * (j + n1 * e);
                    x[id] = xx; y[id] = yy;
                    dxds[id] = covT[0][0]; dxdp[id] = covT[0][1];
                    dyds[id] = covT[1][0]; dydp[id] = covT[1][1];
                    J[id] = covT[0][0] * covT[1][1] - covT[0][1] * covT[1][0];
                }
            /* boundary normals (MappedGeometryClass_2D.f90:652-706) */
            for (int side = 0; side < 4; ++side)
                for (int i = 0; i <= N; ++i) {
                    double s, p, sgn;
                    int useP; /* 0: normal from ds column (south/north), 1: from dp column */
                    switch (side) {
                    case 0: s = sn[i]; p = -1.0; sgn = -1.0; useP = 0; break; /* south */
                    case 1: s = 1.0; p = sn[i]; sgn = 1.0; useP = 1; break;   /* east */
                    case 2: s = sn[i]; p = 1.0; sgn = 1.0; useP = 0; break;   /* north */
                    default: s = -1.0; p = sn[i]; sgn = -1.0; useP = 1; break; /* west */
                    }
                    double covT[2][2], xx, yy;
                    element_metrics(d, c, s, p, &xx, &yy, covT);
                    double Jb = covT[0][0] * covT[1][1] - covT[0][1] * covT[1][0];
                    double signJ = fabs(Jb) / Jb;
                    double dir0, dir1;
                    if (!useP) { dir0 = sgn * signJ * (-covT[1][0]); dir1 = sgn * signJ * covT[0][0]; }
                    else { dir0 = sgn * signJ * covT[1][1]; dir1 = sgn * signJ * (-covT[0][1]); }
                    double len = sqrt(dir0 * dir0 + dir1 * dir1);
                    int nb = i + n1 * (side + 4 * e);
                    nLen[nb] = len;
                    double inv = sqrt(dir0 * dir0 + dir1 * dir1);
                    nHat[2 * nb + 0] = dir0 / inv;
                    nHat[2 * nb + 1] = dir1 / inv;
                }
        }
    return 0;
}
