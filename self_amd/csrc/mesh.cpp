// mesh.cpp -- structured quad meshes for the 2-D models (host set-up).
//
// Node and element numbering follow LoadDefaultMesh_QuadMesh
// (src/geom/QuadMeshClass.f90:1718-1859): element iEl = iX + (iY-1)*nX,
// corner nodes SW, SE, NE, NW.  The edge table is discovered the way
// ConstructEdges_QuadMesh does it (:1489-1627): elements in order, sides
// 1..4 in order, the first visit of a node pair creates the edge with that
// element as primary, the second visit records the secondary element and
// the sign of its side (+ when both traverse the edge in the same
// direction).  start/inc follow :1830-1846.  Periodic pairing identifies the
// last node column (row) with the first, so wrap-around edges come out of
// the same discovery loop.
//
// Geometry: covariant metrics (dx/ds, dx/dp, dy/ds, dy/dp), Jacobian and
// outward unit normals with their lengths, as GenerateMetrics_MappedGeometry_2D
// (src/geom/MappedGeometryClass_2D.f90:611-710) computes them, here for the
// bilinear element map (the transfinite interpolation of straight sides)
// composed with an optional smooth warp of the unit square, then scaled by
// (xScale, yScale) (:921-948).

#include <cmath>
#include <unordered_map>
#include <vector>

#include "common.hpp"

namespace {

constexpr int kEdgeMap[2][4] = {{1, 2, 4, 1}, {2, 3, 3, 4}};  // QuadMeshClass.f90:194-195

struct Desc {
    int nX, nY, N, px, py;
    double xs, ys, warp;
};

void counts(int nX, int nY, int px, int py, int &nE, int &nEdges, int &nNodes)
{
    const int nxn = px ? nX : nX + 1, nyn = py ? nY : nY + 1;
    nNodes = nxn * nyn;
    nE = nX * nY;
    nEdges = nxn * nY + nyn * nX;
}

// position and covariant metric tensor at reference point (s,p) of the
// element whose unit-square corners are c (SW, SE, NE, NW)
void metrics_at(const Desc &d, const double c[4][2], double s, double p, double &x, double &y,
                double cov[2][2])
{
    const double w0 = (1 - s) * (1 - p), w1 = (1 + s) * (1 - p), w2 = (1 + s) * (1 + p), w3 = (1 - s) * (1 + p);
    const double a = 0.25 * (w0 * c[0][0] + w1 * c[1][0] + w2 * c[2][0] + w3 * c[3][0]);
    const double b = 0.25 * (w0 * c[0][1] + w1 * c[1][1] + w2 * c[2][1] + w3 * c[3][1]);
    const double as = 0.25 * (-(1 - p) * c[0][0] + (1 - p) * c[1][0] + (1 + p) * c[2][0] - (1 + p) * c[3][0]);
    const double ap = 0.25 * (-(1 - s) * c[0][0] - (1 + s) * c[1][0] + (1 + s) * c[2][0] + (1 - s) * c[3][0]);
    const double bs = 0.25 * (-(1 - p) * c[0][1] + (1 - p) * c[1][1] + (1 + p) * c[2][1] - (1 + p) * c[3][1]);
    const double bp = 0.25 * (-(1 - s) * c[0][1] - (1 + s) * c[1][1] + (1 + s) * c[2][1] + (1 - s) * c[3][1]);
    double X = a, Y = b, Xa = 1, Xb = 0, Ya = 0, Yb = 1;
    if (d.warp != 0.0) {
        const double tp = 2.0 * 4.0 * std::atan(1.0);
        const double sa = std::sin(tp * a), sb = std::sin(tp * b);
        const double wv = d.warp * sa * sb;
        const double wa = d.warp * tp * std::cos(tp * a) * sb, wb = d.warp * tp * sa * std::cos(tp * b);
        X = a + wv; Y = b + wv;
        Xa = 1.0 + wa; Xb = wb; Ya = wa; Yb = 1.0 + wb;
    }
    x = d.xs * X;
    y = d.ys * Y;
    cov[0][0] = d.xs * (Xa * as + Xb * bs);
    cov[0][1] = d.xs * (Xa * ap + Xb * bp);
    cov[1][0] = d.ys * (Ya * as + Yb * bs);
    cov[1][1] = d.ys * (Ya * ap + Yb * bp);
}

}  // namespace

extern "C" {

int self_quadmesh_counts(int nX, int nY, int periodicX, int periodicY, int *nElems, int *nEdges)
{
    SELF_REQUIRE(nX >= 1 && nY >= 1 && nElems && nEdges, SELF_ERR_ARG, "self_quadmesh_counts: bad arguments");
    int nn;
    counts(nX, nY, periodicX, periodicY, *nElems, *nEdges, nn);
    return SELF_OK;
}

int self_quadmesh_structured(int nX, int nY, int N, int periodicX, int periodicY, double xScale,
                             double yScale, double warp, const double *sn, int *edges, double *x,
                             double *y, double *dxds, double *dxdp, double *dyds, double *dydp,
                             double *J, double *nHat, double *nLen)
try {
    SELF_REQUIRE(nX >= 1 && nY >= 1 && N >= 1 && sn && edges && x && y && dxds && dxdp && dyds &&
                     dydp && J && nHat && nLen,
                 SELF_ERR_ARG, "self_quadmesh_structured: bad arguments");
    // fewer than 3 elements across a periodic direction makes two distinct
    // edges share one node pair
    SELF_REQUIRE(!(periodicX && nX < 3) && !(periodicY && nY < 3), SELF_ERR_ARG,
                 "periodic directions need at least 3 elements");
    const Desc d{nX, nY, N, periodicX, periodicY, xScale, yScale, warp};
    int nE, nEd, nNodes;
    counts(nX, nY, periodicX, periodicY, nE, nEd, nNodes);
    const int nxn = periodicX ? nX : nX + 1, nyn = periodicY ? nY : nY + 1;
    auto node = [&](int iX, int iY) { return (iX - 1) % nxn + 1 + ((iY - 1) % nyn) * nxn; };
    std::vector<int> en(4 * nE);
    for (int iY = 1; iY <= nY; ++iY)
        for (int iX = 1; iX <= nX; ++iX) {
            const int e = iX + (iY - 1) * nX - 1;
            en[4 * e + 0] = node(iX, iY);
            en[4 * e + 1] = node(iX + 1, iY);
            en[4 * e + 2] = node(iX + 1, iY + 1);
            en[4 * e + 3] = node(iX, iY + 1);
        }
    std::unordered_map<long long, int> table;
    table.reserve(2 * nEd);
    int ne = 0;
    for (int e = 0; e < nE; ++e)
        for (int k = 1; k <= 4; ++k) {
            const int a = en[4 * e + kEdgeMap[0][k - 1] - 1], b = en[4 * e + kEdgeMap[1][k - 1] - 1];
            const long long key = (long long)std::min(a, b) * (nNodes + 1LL) + std::max(a, b);
            auto it = table.find(key);
            if (it == table.end()) {
                int *ed = edges + 6 * ne;
                ed[0] = e + 1; ed[1] = k; ed[2] = SELF_NO_NORMAL_FLOW; ed[3] = 0;
                table.emplace(key, ne++);
            } else {
                int *ed = edges + 6 * it->second;
                const int first = en[4 * (ed[0] - 1) + kEdgeMap[0][ed[1] - 1] - 1];
                ed[2] = e + 1;
                ed[3] = (a == first) ? k : -k;
            }
        }
    SELF_REQUIRE(ne == nEd, SELF_ERR_OTHER, "edge discovery found %d edges, expected %d", ne, nEd);
    for (int i = 0; i < ne; ++i) {
        int *ed = edges + 6 * i;
        if (ed[3] < 0) { ed[4] = N - 1; ed[5] = -1; }
        else { ed[4] = 1; ed[5] = 1; }
    }
    const int n1 = N + 1;
    const double dxE = 1.0 / nX, dyE = 1.0 / nY;
    for (int iY = 1; iY <= nY; ++iY)
        for (int iX = 1; iX <= nX; ++iX) {
            const int e = iX + (iY - 1) * nX - 1;
            const double c[4][2] = {{dxE * (iX - 1), dyE * (iY - 1)},
                                    {dxE * iX, dyE * (iY - 1)},
                                    {dxE * iX, dyE * iY},
                                    {dxE * (iX - 1), dyE * iY}};
            for (int j = 0; j <= N; ++j)
                for (int i = 0; i <= N; ++i) {
                    double cov[2][2];
                    const int id = i + n1 * (j + n1 * e);
                    metrics_at(d, c, sn[i], sn[j], x[id], y[id], cov);
                    dxds[id] = cov[0][0]; dxdp[id] = cov[0][1];
                    dyds[id] = cov[1][0]; dydp[id] = cov[1][1];
                    J[id] = cov[0][0] * cov[1][1] - cov[0][1] * cov[1][0];
                }
            for (int side = 0; side < 4; ++side)
                for (int i = 0; i <= N; ++i) {
                    // south/north normals come from the s-tangent, east/west from the p-tangent
                    const double s = (side == 1) ? 1.0 : (side == 3) ? -1.0 : sn[i];
                    const double p = (side == 0) ? -1.0 : (side == 2) ? 1.0 : sn[i];
                    const double sgn = (side == 0 || side == 3) ? -1.0 : 1.0;
                    double cov[2][2], xx, yy;
                    metrics_at(d, c, s, p, xx, yy, cov);
                    const double Jb = cov[0][0] * cov[1][1] - cov[0][1] * cov[1][0];
                    const double signJ = std::fabs(Jb) / Jb;
                    double d0, d1;
                    if (side == 0 || side == 2) { d0 = sgn * signJ * (-cov[1][0]); d1 = sgn * signJ * cov[0][0]; }
                    else { d0 = sgn * signJ * cov[1][1]; d1 = sgn * signJ * (-cov[0][1]); }
                    const double len = std::sqrt(d0 * d0 + d1 * d1);
                    const int nb = i + n1 * (side + 4 * e);
                    nLen[nb] = len;
                    nHat[2 * nb + 0] = d0 / len;
                    nHat[2 * nb + 1] = d1 / len;
                }
        }
    return SELF_OK;
}
SELF_CATCH_ALL

}  // extern "C"
