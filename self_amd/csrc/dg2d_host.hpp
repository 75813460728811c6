// dg2d_host.hpp -- host-side preparation of 2-D DGSEM mesh inputs shared by
// the models: SoA metric arrays, per-side normals, the "flux normals" of
// each side (the edge's primary-element normal at the matching node, so a
// side computes exactly the flux EdgeFlux stores for it) and the packed side
// descriptors of dg2d_device.hpp.
#pragma once

#include <cstring>
#include <vector>

#include "common.hpp"

namespace self {

struct Geo2D {
    std::vector<double> metric;  // [5][nE][np]: dxds, dxdp, dyds, dydp, J
    std::vector<double> nrm;     // [nE][4][3][n1]: nx, ny, |n| of the side itself
    std::vector<double> fnrm;    // [nE][4][3][n1]: primary-oriented normals
    std::vector<int> sideX, sideY;  // packed side descriptors (int2 x / y)
    // for each (e, side, k): the (e', side', k') whose trace is the primary's
    // value at this node (itself when primary or boundary)
    std::vector<int> primaryNode;   // [nE][4][n1] -> flat index (e'*4 + s')*n1 + k'
};

inline int build_geo2d(int N, size_t nE, int nEdges, const int *edges, const double *dxds, const double *dxdp,
                       const double *dyds, const double *dydp, const double *J, const double *nHat,
                       const double *nLen, Geo2D &G)
{
    const int n1 = N + 1, np = n1 * n1;
    G.metric.resize(5 * nE * np);
    const double *src[5] = {dxds, dxdp, dyds, dydp, J};
    for (int k = 0; k < 5; ++k) std::memcpy(G.metric.data() + k * nE * np, src[k], sizeof(double) * nE * np);
    for (size_t id = 0; id < nE * np; ++id)
        SELF_REQUIRE(J[id] != 0.0, SELF_ERR_ARG, "zero Jacobian at element %zu", id / np + 1);
    G.nrm.resize(nE * 12 * n1);
    for (size_t e = 0; e < nE; ++e)
        for (int s = 0; s < 4; ++s)
            for (int k = 0; k < n1; ++k) {
                const size_t hb = k + (size_t)n1 * (s + 4 * e);
                const size_t db = ((e * 4 + s) * 3) * n1 + k;
                G.nrm[db] = nHat[2 * hb];
                G.nrm[db + n1] = nHat[2 * hb + 1];
                G.nrm[db + 2 * n1] = nLen[hb];
            }
    G.fnrm = G.nrm;
    G.sideX.assign(nE * 4, 0);
    G.sideY.assign(nE * 4, 0);
    G.primaryNode.resize(nE * 4 * n1);
    for (size_t q = 0; q < nE * 4 * (size_t)n1; ++q) G.primaryNode[q] = (int)q;
    std::vector<int> seen(nE * 4, 0);
    for (int ie = 0; ie < nEdges; ++ie) {
        const int *ed = edges + 6 * ie;
        const int e1 = ed[0] - 1, s1 = ed[1] - 1, e2 = ed[2], s2 = std::abs(ed[3]) - 1, inc = ed[5];
        SELF_REQUIRE(e1 >= 0 && (size_t)e1 < nE && s1 >= 0 && s1 < 4, SELF_ERR_ARG, "edge %d: bad primary", ie + 1);
        SELF_REQUIRE(inc == 1 || inc == -1, SELF_ERR_ARG, "edge %d: increment %d", ie + 1, inc);
        if (e2 > 0) {
            SELF_REQUIRE((size_t)(e2 - 1) < nE && s2 >= 0 && s2 < 4, SELF_ERR_ARG, "edge %d: bad secondary", ie + 1);
            SELF_REQUIRE(ed[4] - inc == (inc > 0 ? 0 : N), SELF_ERR_ARG,
                         "edge %d: start/inc must traverse the whole side", ie + 1);
            const int flip = inc < 0 ? 1 : 0;
            G.sideX[(size_t)e1 * 4 + s1] = e2 - 1;
            G.sideY[(size_t)e1 * 4 + s1] = s2 | (flip << 2) | (1 << 3);
            G.sideX[(size_t)(e2 - 1) * 4 + s2] = e1;
            G.sideY[(size_t)(e2 - 1) * 4 + s2] = s1 | (flip << 2);
            seen[(size_t)e1 * 4 + s1]++;
            seen[(size_t)(e2 - 1) * 4 + s2]++;
            for (int k = 0; k < n1; ++k) {
                const int kp = flip ? N - k : k;
                const size_t dsec = (((size_t)(e2 - 1) * 4 + s2) * 3) * n1 + k;
                const size_t dpri = (((size_t)e1 * 4 + s1) * 3) * n1 + kp;
                G.fnrm[dsec] = G.nrm[dpri];
                G.fnrm[dsec + n1] = G.nrm[dpri + n1];
                G.fnrm[dsec + 2 * n1] = G.nrm[dpri + 2 * n1];
                G.primaryNode[((size_t)(e2 - 1) * 4 + s2) * n1 + k] = (int)(((size_t)e1 * 4 + s1) * n1 + kp);
            }
        } else {
            G.sideX[(size_t)e1 * 4 + s1] = e2;
            G.sideY[(size_t)e1 * 4 + s1] = 1 << 3;
            seen[(size_t)e1 * 4 + s1]++;
        }
    }
    for (size_t k = 0; k < nE * 4; ++k)
        SELF_REQUIRE(seen[k] == 1, SELF_ERR_ARG, "element %zu side %zu is on %d edges", k / 4 + 1, k % 4 + 1, seen[k]);
    return SELF_OK;
}

}  // namespace self
