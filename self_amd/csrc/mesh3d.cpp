// mesh3d.cpp -- structured hex meshes for the 3-D models (host set-up).
//
// Elements of the unit cube are numbered iX fastest, then iY, then iZ.
// Faces of the reference element follow src/common/ModelFlags.f90:36-41
// (south 1 = p-, east 2 = s+, north 3 = p+, west 4 = s-, bottom 5 = q-,
// top 6 = q+).  The face table lists each shared face once: elements in
// order, sides 1..6 in order, the first visit makes that element primary.
// Face records carry the node correspondence of the reference's face loop
// (Advection3DClass.f90:1286-1296): iStart, jStart, iInc, jInc, swap; on a
// structured mesh neighbours see a shared face with the same two reference
// coordinates, so the correspondence is the identity.  Periodic pairing
// (absent from the reference's default mesh) wraps each direction.
//
// Geometry comes from the warp map of the unit cube
//   X = a + A s1 S, Y = b + A s2 S, Z = c + A s3 S,
//   S = sin(2 pi a) sin(2 pi b) sin(2 pi c),
// scaled by (xScale, yScale, zScale): positions, the nine covariant metric
// terms d(x,y,z)/d(s,p,q), the Jacobian, and at every face node the outward
// normal (the cofactor row of the face's reference direction, whose length
// is the face area element), its length and the node position.

#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"

namespace {

struct Desc3 {
    int nX, nY, nZ, N, per[3];
    double scale[3], warp;
    int sign[3];
};

int faces_in(int n, int periodic, int a, int b) { return (periodic ? n : n + 1) * a * b; }

void point(const Desc3 &d, int iX, int iY, int iZ, double s, double p, double q, double X[3], double cov[9])
{
    const double tp = 8.0 * std::atan(1.0);
    const double ref[3] = {(iX + 0.5 * (s + 1.0)) / d.nX, (iY + 0.5 * (p + 1.0)) / d.nY,
                           (iZ + 0.5 * (q + 1.0)) / d.nZ};
    const double dref[3] = {0.5 / d.nX, 0.5 / d.nY, 0.5 / d.nZ};
    const double sa = std::sin(tp * ref[0]), sb = std::sin(tp * ref[1]), sc = std::sin(tp * ref[2]);
    const double S = sa * sb * sc;
    const double dS[3] = {tp * std::cos(tp * ref[0]) * sb * sc, tp * sa * std::cos(tp * ref[1]) * sc,
                          tp * sa * sb * std::cos(tp * ref[2])};
    for (int r = 0; r < 3; ++r) {
        const double g = d.warp * d.sign[r];
        X[r] = d.scale[r] * (ref[r] + g * S);
        for (int k = 0; k < 3; ++k) {
            const double dXk = g * dS[k] + (k == r ? 1.0 : 0.0);
            cov[3 * r + k] = d.scale[r] * dXk * dref[k];
        }
    }
}

}  // namespace

namespace self {

// cofactor row of direction dir (0 s, 1 p, 2 q) of the covariant metric m,
// in the grouping of Advection3DClass.f90:1412-1414, 1460-1462, 1510-1512
void cofactor3(const double *m, int dir, double out[3])
{
    const double dxds = m[0], dxdp = m[1], dxdq = m[2], dyds = m[3], dydp = m[4], dydq = m[5];
    const double dzds = m[6], dzdp = m[7], dzdq = m[8];
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

}  // namespace self

extern "C" {

int self_hexmesh_counts(int nX, int nY, int nZ, int periodicX, int periodicY, int periodicZ, int *nElems,
                        int *nFaces)
{
    SELF_REQUIRE(nElems && nFaces && nX >= 1 && nY >= 1 && nZ >= 1, SELF_ERR_ARG, "self_hexmesh_counts: bad arguments");
    *nElems = nX * nY * nZ;
    *nFaces = faces_in(nX, periodicX, nY, nZ) + faces_in(nY, periodicY, nX, nZ) + faces_in(nZ, periodicZ, nX, nY);
    return SELF_OK;
}

int self_hexmesh_structured(int nX, int nY, int nZ, int N, int periodicX, int periodicY, int periodicZ,
                            double xScale, double yScale, double zScale, double warp, const int *warpSign,
                            const double *nodes, int *faces, double *x, double *y, double *z, double *cov,
                            double *J, double *nHat, double *nLen, double *bpos)
{
    SELF_REQUIRE(nodes && faces && x && y && z && cov && J && nHat && nLen && bpos, SELF_ERR_ARG,
                 "self_hexmesh_structured: null argument");
    SELF_REQUIRE(N >= 1 && nX >= 1 && nY >= 1 && nZ >= 1, SELF_ERR_ARG, "self_hexmesh_structured: bad sizes");
    Desc3 d{nX, nY, nZ, N, {periodicX, periodicY, periodicZ}, {xScale, yScale, zScale}, warp,
            {warpSign ? warpSign[0] : 1, warpSign ? warpSign[1] : 1, warpSign ? warpSign[2] : 1}};
    const int n1 = N + 1, n2 = n1 * n1;
    const size_t n3 = (size_t)n2 * n1;
    const int nE = nX * nY * nZ;
    for (int iZ = 0; iZ < nZ; ++iZ)
        for (int iY = 0; iY < nY; ++iY)
            for (int iX = 0; iX < nX; ++iX) {
                const size_t e = iX + (size_t)nX * (iY + (size_t)nY * iZ);
                for (int k = 0; k < n1; ++k)
                    for (int j = 0; j < n1; ++j)
                        for (int i = 0; i < n1; ++i) {
                            const size_t id = e * n3 + i + n1 * (j + (size_t)n1 * k);
                            double X[3], m[9], cs[3];
                            point(d, iX, iY, iZ, nodes[i], nodes[j], nodes[k], X, m);
                            x[id] = X[0];
                            y[id] = X[1];
                            z[id] = X[2];
                            std::memcpy(cov + 9 * id, m, sizeof(m));
                            self::cofactor3(m, 0, cs);
                            J[id] = m[0] * cs[0] + m[3] * cs[1] + m[6] * cs[2];
                        }
                for (int side = 1; side <= 6; ++side)
                    for (int b = 0; b < n1; ++b)
                        for (int a = 0; a < n1; ++a) {
                            // face node (a, b) and the reference direction normal to the face
                            double s = nodes[a], p = nodes[b], q = 0.0, sgn = 1.0;
                            int dir = 0;
                            switch (side) {
                            case 1: s = nodes[a]; p = -1.0; q = nodes[b]; dir = 1; sgn = -1.0; break;
                            case 2: s = 1.0; p = nodes[a]; q = nodes[b]; dir = 0; break;
                            case 3: s = nodes[a]; p = 1.0; q = nodes[b]; dir = 1; break;
                            case 4: s = -1.0; p = nodes[a]; q = nodes[b]; dir = 0; sgn = -1.0; break;
                            case 5: s = nodes[a]; p = nodes[b]; q = -1.0; dir = 2; sgn = -1.0; break;
                            default: s = nodes[a]; p = nodes[b]; q = 1.0; dir = 2; break;
                            }
                            double X[3], m[9], cf[3];
                            point(d, iX, iY, iZ, s, p, q, X, m);
                            self::cofactor3(m, dir, cf);
                            const double nx = sgn * cf[0], ny = sgn * cf[1], nz = sgn * cf[2];
                            const double len = std::sqrt(nx * nx + ny * ny + nz * nz);
                            const size_t fb = a + (size_t)n1 * b + (size_t)n2 * ((side - 1) + 6 * e);
                            nHat[3 * fb] = nx / len;
                            nHat[3 * fb + 1] = ny / len;
                            nHat[3 * fb + 2] = nz / len;
                            nLen[fb] = len;
                            bpos[3 * fb] = X[0];
                            bpos[3 * fb + 1] = X[1];
                            bpos[3 * fb + 2] = X[2];
                        }
            }
    static const int opp[7] = {0, 3, 4, 1, 2, 6, 5};
    std::vector<char> seen((size_t)nE * 7, 0);
    int nf = 0;
    for (int iZ = 0; iZ < nZ; ++iZ)
        for (int iY = 0; iY < nY; ++iY)
            for (int iX = 0; iX < nX; ++iX) {
                const int e = iX + nX * (iY + nY * iZ);
                for (int side = 1; side <= 6; ++side) {
                    if (seen[7 * (size_t)e + side]) continue;
                    int c[3] = {iX, iY, iZ};
                    const int dim = (side == 2 || side == 4) ? 0 : (side == 1 || side == 3) ? 1 : 2;
                    const int step = (side == 2 || side == 3 || side == 6) ? 1 : -1;
                    const int n[3] = {nX, nY, nZ};
                    c[dim] += step;
                    bool ok = true;
                    if (c[dim] < 0 || c[dim] >= n[dim]) {
                        if (d.per[dim]) c[dim] = (c[dim] + n[dim]) % n[dim];
                        else ok = false;
                    }
                    int *f = faces + 9 * (size_t)nf;
                    f[0] = e + 1;
                    f[1] = side;
                    f[4] = 0; f[5] = 0; f[6] = 1; f[7] = 1; f[8] = 0;
                    seen[7 * (size_t)e + side] = 1;
                    if (ok) {
                        const int e2 = c[0] + nX * (c[1] + nY * c[2]);
                        f[2] = e2 + 1;
                        f[3] = opp[side];
                        seen[7 * (size_t)e2 + opp[side]] = 1;
                    } else {
                        f[2] = -100;  // NO_NORMAL_FLOW
                        f[3] = 0;
                    }
                    ++nf;
                }
            }
    int nfExpect = 0, nEdummy = 0;
    self_hexmesh_counts(nX, nY, nZ, periodicX, periodicY, periodicZ, &nEdummy, &nfExpect);
    SELF_REQUIRE(nf == nfExpect, SELF_ERR_OTHER, "self_hexmesh_structured: %d faces, expected %d", nf, nfExpect);
    return SELF_OK;
}

}  // extern "C"
