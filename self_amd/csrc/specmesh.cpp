// specmesh.cpp -- ISM-V2 SpecMesh files (unstructured quad meshes with
// curved boundaries) and their element geometry, as the reference builds
// them (host set-up).
//
// ReadSpecMeshFile_QuadMesh (src/geom/QuadMeshClass.f90:2100-2465):
//   line 1 the version; line 2 nNodes nEdges nElems gPolyDeg; nNodes lines
//   x y z; nEdges lines of edge data (not used: the edges are rebuilt);
//   per element a line of 4 corner nodes, a line of 4 side flags (0 straight
//   side between its corners, 1 a boundary curve of gPolyDeg+1 lines x y z
//   follows), and a line whose columns 2..41 hold four blank-separated side
//   names.  Boundary curves are Lagrange interpolants on the Chebyshev-
//   Gauss-Lobatto points -cos(j pi / gPolyDeg) (Chebyshev.f90:95-137);
//   straight sides are sampled at those points.
// Element geometry (MappedGeometryClass_2D.f90:536-710): positions and the
// covariant metric tensor at each interpolation node by transfinite
// interpolation of the four side curves and its derivative
// (GeometrySupportRoutines.f90:60-158, Unidirectional :160-190), the
// Jacobian as det, boundary normals from the tangents with the sign of J,
// normalized, with their lengths.  The curves are evaluated in barycentric
// form (EvaluateLagrange_1D, Lagrange_1D_Class.f90:396-444;
// EvaluateDerivative_1D :519-608).
// Edges: ConstructEdges_QuadMesh (:1489-1627) -- elements in order, sides
// in order, the first visit of a corner pair creates the edge (secondary
// NO_NORMAL_FLOW), the second records the element and the signed side;
// then a boundary side named DIRICHLET, ROBIN, ... takes that flag as its
// secondary element (:2339-2419), and start/inc follow the side sign
// (:2421-2431).  The reference evaluates each expression left to right with
// no fused multiply-adds; this file is compiled with -ffp-contract=off.

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "common.hpp"
#include "self_amd.h"

namespace {

using namespace self;

constexpr int kEdgeMapS[2][4] = {{1, 2, 4, 1}, {2, 3, 3, 4}};  // QuadMesh edgeMap: side k runs corner a -> b

// AlmostEqual (CommonRoutines.f90)
bool almost_equal(double a, double b)
{
    const double eps = 2.220446049250313e-16;
    if (a == 0.0 || b == 0.0) return std::fabs(a - b) <= eps;
    return std::fabs(a - b) <= eps * std::fabs(a) || std::fabs(a - b) <= eps * std::fabs(b);
}

struct Curve {
    int g = 1;
    std::vector<double> s, w, x, y;  // parameter nodes, barycentric weights, node positions
};

// EvaluateLagrange_1D
double eval_interp(const Curve &c, const std::vector<double> &f, double xe)
{
    double num = 0.0, den = 0.0;
    for (int j = 0; j <= c.g; ++j) {
        if (almost_equal(xe, c.s[j])) return f[j];
        const double t = c.w[j] / (xe - c.s[j]);
        num = num + t * f[j];
        den = den + t;
    }
    return num / den;
}

// EvaluateDerivative_1D
double eval_deriv(const Curve &c, const std::vector<double> &f, double xe)
{
    double num = 0.0, den = 0.0, p = 0.0;
    int jS = -1;
    for (int j = 0; j <= c.g; ++j)
        if (almost_equal(xe, c.s[j])) {
            p = f[j];
            den = -c.w[j];
            jS = j;
        }
    if (jS >= 0) {
        for (int j = 0; j <= c.g; ++j)
            if (j != jS) num = num + c.w[j] * (p - f[j]) / (xe - c.s[j]);
    } else {
        den = 0.0;
        p = eval_interp(c, f, xe);
        for (int j = 0; j <= c.g; ++j) {
            const double t = c.w[j] / (xe - c.s[j]);
            num = num + t * (p - f[j]) / (xe - c.s[j]);
            den = den + t;
        }
    }
    return num / den;
}

void evaluate(const Curve &c, double s, double P[2])
{
    P[0] = eval_interp(c, c.x, s);
    P[1] = eval_interp(c, c.y, s);
}

void slope(const Curve &c, double s, double P[2])
{
    P[0] = eval_deriv(c, c.x, s);
    P[1] = eval_deriv(c, c.y, s);
}

// Unidirectional_2D, UnidirectionalDerivative_2D
void uni(const double L[2], const double R[2], double a, double P[2])
{
    for (int k = 0; k < 2; ++k) P[k] = 0.5 * ((1.0 - a) * L[k] + (1.0 + a) * R[k]);
}

void uni_d(const double L[2], const double R[2], double P[2])
{
    for (int k = 0; k < 2; ++k) P[k] = 0.5 * (R[k] - L[k]);
}

// TransfiniteInterpolation_2D (curves: south, east, north, west)
void tfi(const Curve cv[4], double a, double b, double P[2])
{
    double c1[2], c2[2], c3[2], c4[2], l[2], r[2], P1[2], P2[2], Ps[2], Pn[2], Q[2];
    evaluate(cv[0], -1.0, c1);
    evaluate(cv[0], 1.0, c2);
    evaluate(cv[2], -1.0, c3);
    evaluate(cv[2], 1.0, c4);
    evaluate(cv[3], b, l);
    evaluate(cv[1], b, r);
    uni(l, r, a, P1);
    evaluate(cv[0], a, l);
    evaluate(cv[2], a, r);
    uni(l, r, b, P2);
    uni(c1, c2, a, Ps);
    uni(c3, c4, a, Pn);
    uni(Ps, Pn, b, Q);
    for (int k = 0; k < 2; ++k) P[k] = P1[k] + P2[k] - Q[k];
}

// TransfiniteInterpolationMetrics_2D: m[r][c] with r = (x, y), c = (s, p)
void tfi_metrics(const Curve cv[4], double a, double b, double m[2][2])
{
    double w[2], e[2], so[2], no[2], dW[2], dE[2], dS[2], dN[2], c1[2], c2[2], c3[2], c4[2];
    evaluate(cv[3], b, w);
    evaluate(cv[1], b, e);
    evaluate(cv[0], a, so);
    evaluate(cv[2], a, no);
    slope(cv[3], b, dW);
    slope(cv[1], b, dE);
    slope(cv[0], a, dS);
    slope(cv[2], a, dN);
    evaluate(cv[0], -1.0, c1);
    evaluate(cv[0], 1.0, c2);
    evaluate(cv[2], -1.0, c3);
    evaluate(cv[2], 1.0, c4);
    double P1[2], Ps[2], Pn[2], L[2], R[2], U[2];
    uni_d(w, e, P1);
    uni_d(c1, c2, Ps);
    uni_d(c3, c4, Pn);
    for (int k = 0; k < 2; ++k) {
        L[k] = dS[k] - Ps[k];
        R[k] = dN[k] - Pn[k];
    }
    uni(L, R, b, U);
    for (int k = 0; k < 2; ++k) m[k][0] = P1[k] + U[k];
    double P2[2], Pw[2], Pe[2];
    uni_d(so, no, P2);
    uni_d(c1, c3, Pw);
    uni_d(c2, c4, Pe);
    for (int k = 0; k < 2; ++k) {
        L[k] = dW[k] - Pw[k];
        R[k] = dE[k] - Pe[k];
    }
    uni(L, R, a, U);
    for (int k = 0; k < 2; ++k) m[k][1] = P2[k] + U[k];
}

struct Mesh {
    int nNodes = 0, nEdgesFile = 0, nElems = 0, g = 1;
    std::vector<double> nx, ny;
    std::vector<int> en;                // 4 corner nodes per element (1-based)
    std::vector<int> bflag;             // 4 side flags per element
    std::vector<std::vector<double>> cx, cy;  // boundary curves read: [4*e + side]
    std::vector<int> sideFlag;          // 4 per element: INTERIOR or a boundary type
};

int side_type(const std::string &t)
{
    if (t == "DIRICHLET") return -200;
    if (t == "ROBIN") return -202;
    if (t == "ROBIN_FORCED") return -204;
    if (t == "HOMOGENEOUS_NEUMANN") return -201;
    if (t == "NEUMANN_WALL") return -206;
    if (t == "NEUMANN") return -100;  // = NO_NORMAL_FLOW (ModelFlags.f90:96)
    if (t == "NO_NORMAL_FLOW") return -100;
    if (t == "PRESCRIBED") return -102;
    if (t == "RADIATION") return -101;
    if (t == "InflowOne") return -103;
    if (t == "InflowTwo") return -104;
    return 1;  // INTERIOR: the flag stays as initialized
}

int parse(const char *path, Mesh &m)
{
    std::ifstream in(path);
    SELF_REQUIRE(in.good(), SELF_ERR_ARG, "SpecMesh: cannot open %s", path);
    std::string line;
    auto next = [&](const char *what) -> int {
        SELF_REQUIRE(std::getline(in, line), SELF_ERR_ARG, "SpecMesh %s: file ends before %s", path, what);
        if (!line.empty() && line.back() == '\r') line.pop_back();  // CRLF files
        return SELF_OK;
    };
    // list-directed READ also takes Fortran's D exponent (1.5D+02)
    auto fortran_reals = [&]() {
        for (char &ch : line)
            if (ch == 'D' || ch == 'd') ch = 'E';
    };
    int rc;
    if ((rc = next("the version line"))) return rc;
    SELF_REQUIRE(line.find("ISM-V2") != std::string::npos, SELF_ERR_ARG,
                 "SpecMesh %s: first line is not 'ISM-V2' (only that version is read)", path);
    if ((rc = next("the counts"))) return rc;
    {
        std::istringstream ss(line);
        SELF_REQUIRE(ss >> m.nNodes >> m.nEdgesFile >> m.nElems >> m.g, SELF_ERR_ARG,
                     "SpecMesh %s: bad counts line", path);
        SELF_REQUIRE(m.nNodes >= 4 && m.nElems >= 1 && m.g >= 1 && m.g <= 64, SELF_ERR_ARG,
                     "SpecMesh %s: nNodes %d nElems %d gPolyDeg %d", path, m.nNodes, m.nElems, m.g);
    }
    m.nx.resize(m.nNodes);
    m.ny.resize(m.nNodes);
    for (int i = 0; i < m.nNodes; ++i) {
        if ((rc = next("a node"))) return rc;
        fortran_reals();
        std::istringstream ss(line);
        SELF_REQUIRE(ss >> m.nx[i] >> m.ny[i], SELF_ERR_ARG, "SpecMesh %s: bad node line %d", path, i + 1);
    }
    for (int i = 0; i < m.nEdgesFile; ++i)
        if ((rc = next("the edge list"))) return rc;
    m.en.resize(4 * (size_t)m.nElems);
    m.bflag.resize(4 * (size_t)m.nElems);
    m.cx.assign(4 * (size_t)m.nElems, {});
    m.cy.assign(4 * (size_t)m.nElems, {});
    m.sideFlag.assign(4 * (size_t)m.nElems, 1);
    for (int e = 0; e < m.nElems; ++e) {
        if ((rc = next("an element's nodes"))) return rc;
        {
            std::istringstream ss(line);
            for (int k = 0; k < 4; ++k) {
                SELF_REQUIRE(ss >> m.en[4 * e + k], SELF_ERR_ARG, "SpecMesh %s: element %d nodes", path, e + 1);
                SELF_REQUIRE(m.en[4 * e + k] >= 1 && m.en[4 * e + k] <= m.nNodes, SELF_ERR_ARG,
                             "SpecMesh %s: element %d node %d out of range", path, e + 1, m.en[4 * e + k]);
            }
        }
        if ((rc = next("an element's side flags"))) return rc;
        {
            std::istringstream ss(line);
            for (int k = 0; k < 4; ++k)
                SELF_REQUIRE(ss >> m.bflag[4 * e + k], SELF_ERR_ARG, "SpecMesh %s: element %d flags", path, e + 1);
        }
        for (int k = 0; k < 4; ++k) {
            const int f = m.bflag[4 * e + k];
            SELF_REQUIRE(f == 0 || f == 1, SELF_ERR_ARG, "SpecMesh %s: element %d side %d flag %d", path, e + 1,
                         k + 1, f);
            if (f == 1) {
                auto &X = m.cx[4 * e + k], &Y = m.cy[4 * e + k];
                X.resize(m.g + 1);
                Y.resize(m.g + 1);
                for (int j = 0; j <= m.g; ++j) {
                    if ((rc = next("a boundary curve"))) return rc;
                    fortran_reals();
                    std::istringstream ss(line);
                    SELF_REQUIRE(ss >> X[j] >> Y[j], SELF_ERR_ARG, "SpecMesh %s: element %d curve point", path,
                                 e + 1);
                }
            }
        }
        if ((rc = next("an element's side names"))) return rc;
        // READ '(1x, A40)': columns 2..41, blank padded; names split at each blank
        std::string field = line.size() > 1 ? line.substr(1, 40) : std::string();
        field.resize(40, ' ');
        int side = 0;
        std::string tok;
        for (size_t ic = 0; side < 4; ++ic) {
            const char ch = ic < field.size() ? field[ic] : ' ';
            if (ch == ' ') {
                const int t = side_type(tok);
                if (t != 1) m.sideFlag[4 * e + side] = t;
                tok.clear();
                ++side;
            } else {
                tok.push_back(ch);
            }
        }
    }
    return SELF_OK;
}

// edges (6 per edge: e1, s1, e2, s2, start, increment) and the node types
// the named boundary sides give their corners (INTERIOR = 1 elsewhere)
int build_edges(const Mesh &m, int N, std::vector<int> &edges, std::vector<int> &nodeType)
{
    std::unordered_map<long long, int> table;
    table.reserve(4 * (size_t)m.nElems);
    edges.clear();
    for (int e = 0; e < m.nElems; ++e)
        for (int k = 1; k <= 4; ++k) {
            const int a = m.en[4 * e + kEdgeMapS[0][k - 1] - 1], b = m.en[4 * e + kEdgeMapS[1][k - 1] - 1];
            const long long key = (long long)std::min(a, b) * (m.nNodes + 1LL) + std::max(a, b);
            auto it = table.find(key);
            if (it == table.end()) {
                table.emplace(key, (int)(edges.size() / 6));
                edges.insert(edges.end(), {e + 1, k, -100, 0, 0, 0});
            } else {
                int *ed = edges.data() + 6 * it->second;
                SELF_REQUIRE(ed[3] == 0 && ed[2] < 0, SELF_ERR_ARG, "SpecMesh: corners %d-%d shared by more than two sides",
                             a, b);
                const int first = m.en[4 * (ed[0] - 1) + kEdgeMapS[0][ed[1] - 1] - 1];
                ed[2] = e + 1;
                ed[3] = (a == first) ? k : -k;
            }
        }
    const int nEd = (int)(edges.size() / 6);
    // the file's edge count bounds the reference's flag loop (:2339): a file
    // whose count disagrees with its elements is rejected, not half-flagged
    SELF_REQUIRE(nEd == m.nEdgesFile, SELF_ERR_ARG,
                 "SpecMesh: the file lists %d edges, its elements have %d", m.nEdgesFile, nEd);
    nodeType.assign(m.nNodes, 1);
    for (int i = 0; i < nEd; ++i) {
        int *ed = edges.data() + 6 * i;
        const int f = m.sideFlag[4 * (ed[0] - 1) + ed[1] - 1];
        if (f != 1) {  // a named boundary side (an interior side keeps its neighbour)
            ed[2] = f;
            const int *en = m.en.data() + 4 * (ed[0] - 1);
            nodeType[en[kEdgeMapS[0][ed[1] - 1] - 1] - 1] = f;
            nodeType[en[kEdgeMapS[1][ed[1] - 1] - 1] - 1] = f;
        }
        if (ed[3] < 0) { ed[4] = N - 1; ed[5] = -1; }
        else { ed[4] = 1; ed[5] = 1; }
    }
    return SELF_OK;
}

// EvaluateLagrangePolynomial_1D (Lagrange_1D_Class.f90:448-515)
void lagrange_values(const Curve &c, double xe, double *l)
{
    bool hit = false;
    for (int j = 0; j <= c.g; ++j) {
        l[j] = 0.0;
        if (almost_equal(xe, c.s[j])) {
            l[j] = 1.0;
            hit = true;
        }
    }
    if (hit) return;
    double t1 = 0.0;
    for (int j = 0; j <= c.g; ++j) {
        const double t2 = c.w[j] / (xe - c.s[j]);
        l[j] = t2;
        t1 = t1 + t2;
    }
    for (int j = 0; j <= c.g; ++j) l[j] = l[j] / t1;
}

// EvaluateDerivative_2D (Lagrange_2D_Class.f90:396-440) of the nodal
// array f (i fastest) at (s, p): (df/ds, df/dp)
void gradient_2d(const Curve &c, const double *f, double s, double p, double g[2])
{
    const int n1 = c.g + 1;
    std::vector<double> lS(n1), lP(n1), row(n1);
    lagrange_values(c, s, lS.data());
    lagrange_values(c, p, lP.data());
    double dfds = 0.0, dfdp = 0.0;
    for (int jP = 0; jP < n1; ++jP) {
        for (int i = 0; i < n1; ++i) row[i] = f[i + n1 * jP];
        dfds = dfds + eval_deriv(c, row, s) * lP[jP];
    }
    for (int jS = 0; jS < n1; ++jS) {
        for (int j = 0; j < n1; ++j) row[j] = f[jS + n1 * j];
        dfdp = dfdp + eval_deriv(c, row, p) * lS[jS];
    }
    g[0] = dfds;
    g[1] = dfdp;
}

}  // namespace

extern "C" {

int self_quadmesh_scale(int N, const double *sNodes, int nElems, int nNodes, double xScale, double yScale,
                        double *nodeX, double *nodeY, double *x, double *y, double *dxds, double *dxdp, double *dyds,
                        double *dydp, double *J, double *nHat, double *nLen)
try {
    SELF_REQUIRE(sNodes && x && y && dxds && dxdp && dyds && dydp && J && nHat && nLen && N >= 1 && N <= 64 &&
                     nElems >= 1 && nNodes >= 0 && (nNodes == 0 || (nodeX && nodeY)),
                 SELF_ERR_ARG, "self_quadmesh_scale: bad arguments");
    const int n1 = N + 1;
    const size_t np = (size_t)n1 * n1;
    for (int i = 0; i < nNodes; ++i) {
        nodeX[i] = xScale * nodeX[i];
        nodeY[i] = yScale * nodeY[i];
    }
    const double xy = xScale * yScale;
    for (size_t k = 0; k < np * nElems; ++k) {
        x[k] = xScale * x[k];
        y[k] = yScale * y[k];
        dxds[k] = xScale * dxds[k];
        dxdp[k] = xScale * dxdp[k];
        dyds[k] = yScale * dyds[k];
        dydp[k] = yScale * dydp[k];
        J[k] = xy * J[k];
    }
    // GenerateBoundaryMetrics: normals from the derivative of the element's
    // (scaled) position interpolant at the boundary points
    Curve c;
    c.g = N;
    c.s.assign(sNodes, sNodes + n1);
    c.w.resize(n1);
    if (int rc = self_barycentric_weights(N, c.s.data(), c.w.data())) return rc;
    for (int e = 0; e < nElems; ++e) {
        const double *xe = x + np * e, *ye = y + np * e;
        for (int side = 0; side < 4; ++side)
            for (int i = 0; i <= N; ++i) {
                const double sa = (side == 1) ? 1.0 : (side == 3) ? -1.0 : sNodes[i];
                const double pa = (side == 0) ? -1.0 : (side == 2) ? 1.0 : sNodes[i];
                const double sgn = (side == 0 || side == 3) ? -1.0 : 1.0;
                double gx[2], gy[2];
                gradient_2d(c, xe, sa, pa, gx);
                gradient_2d(c, ye, sa, pa, gy);
                const double Jb = gx[0] * gy[1] - gx[1] * gy[0];
                const double signJ = std::fabs(Jb) / Jb;
                double d0, d1;
                if (side == 0 || side == 2) { d0 = sgn * signJ * (-gy[0]); d1 = sgn * signJ * gx[0]; }
                else { d0 = sgn * signJ * gy[1]; d1 = sgn * signJ * (-gx[1]); }
                const double len = std::sqrt(d0 * d0 + d1 * d1);
                const size_t nb = i + (size_t)n1 * (side + 4 * (size_t)e);
                nLen[nb] = len;
                nHat[2 * nb + 0] = d0 / len;
                nHat[2 * nb + 1] = d1 / len;
            }
    }
    return SELF_OK;
}
SELF_CATCH_ALL

int self_specmesh_counts(const char *path, int *nNodes, int *nEdges, int *nElems, int *gPolyDeg)
try {
    SELF_REQUIRE(path && nNodes && nEdges && nElems && gPolyDeg, SELF_ERR_ARG, "self_specmesh_counts: null argument");
    Mesh m;
    int rc = parse(path, m);
    if (rc) return rc;
    std::vector<int> edges, types;
    if ((rc = build_edges(m, 2, edges, types))) return rc;
    *nNodes = m.nNodes;
    *nEdges = (int)(edges.size() / 6);
    *nElems = m.nElems;
    *gPolyDeg = m.g;
    return SELF_OK;
}
SELF_CATCH_ALL

int self_specmesh_read(const char *path, int N, const double *sNodes, int *edges, int *elemNodes, int *nodeType,
                       double *nodeX, double *nodeY, double *x, double *y, double *dxds, double *dxdp, double *dyds,
                       double *dydp, double *J, double *nHat, double *nLen)
try {
    SELF_REQUIRE(path && sNodes && edges && elemNodes && nodeType && nodeX && nodeY && x && y && dxds && dxdp &&
                     dyds && dydp && J && nHat && nLen && N >= 1 && N <= 64,
                 SELF_ERR_ARG, "self_specmesh_read: bad arguments");
    Mesh m;
    int rc = parse(path, m);
    if (rc) return rc;
    std::vector<int> ed, types;
    if ((rc = build_edges(m, N, ed, types))) return rc;
    std::memcpy(edges, ed.data(), sizeof(int) * ed.size());
    std::memcpy(elemNodes, m.en.data(), sizeof(int) * m.en.size());
    std::memcpy(nodeType, types.data(), sizeof(int) * types.size());
    std::memcpy(nodeX, m.nx.data(), sizeof(double) * m.nNodes);
    std::memcpy(nodeY, m.ny.data(), sizeof(double) * m.nNodes);
    // Chebyshev-Gauss-Lobatto curve parameters and their barycentric weights
    const int g = m.g;
    std::vector<double> s(g + 1), w(g + 1);
    for (int j = 0; j <= g; ++j) {
        const double wt = M_PI / (double)g;
        s[j] = -std::cos((double)j * wt);
    }
    if ((rc = self_barycentric_weights(g, s.data(), w.data()))) return rc;
    const int n1 = N + 1;
    for (int e = 0; e < m.nElems; ++e) {
        Curve cv[4];
        for (int k = 0; k < 4; ++k) {
            Curve &c = cv[k];
            c.g = g;
            c.s = s;
            c.w = w;
            if (m.bflag[4 * e + k] == 0) {  // straight side between its corners
                const int a = m.en[4 * e + kEdgeMapS[0][k] - 1] - 1, b = m.en[4 * e + kEdgeMapS[1][k] - 1] - 1;
                const double x1 = m.nx[a], y1 = m.ny[a], x2 = m.nx[b], y2 = m.ny[b];
                c.x.resize(g + 1);
                c.y.resize(g + 1);
                for (int j = 0; j <= g; ++j) {
                    c.x[j] = x1 + (x2 - x1) * 0.5 * (s[j] + 1.0);
                    c.y[j] = y1 + (y2 - y1) * 0.5 * (s[j] + 1.0);
                }
            } else {
                c.x = m.cx[4 * e + k];
                c.y = m.cy[4 * e + k];
            }
        }
        for (int j = 0; j <= N; ++j)
            for (int i = 0; i <= N; ++i) {
                const size_t id = i + (size_t)n1 * (j + (size_t)n1 * e);
                double P[2], cov[2][2];
                tfi(cv, sNodes[i], sNodes[j], P);
                x[id] = P[0];
                y[id] = P[1];
                tfi_metrics(cv, sNodes[i], sNodes[j], cov);
                dxds[id] = cov[0][0];
                dxdp[id] = cov[0][1];
                dyds[id] = cov[1][0];
                dydp[id] = cov[1][1];
                J[id] = cov[0][0] * cov[1][1] - cov[0][1] * cov[1][0];
            }
        for (int side = 0; side < 4; ++side)
            for (int i = 0; i <= N; ++i) {
                // south/north normals come from the s-tangent, east/west from the p-tangent
                const double sa = (side == 1) ? 1.0 : (side == 3) ? -1.0 : sNodes[i];
                const double pa = (side == 0) ? -1.0 : (side == 2) ? 1.0 : sNodes[i];
                const double sgn = (side == 0 || side == 3) ? -1.0 : 1.0;
                double cov[2][2];
                tfi_metrics(cv, sa, pa, cov);
                const double Jb = cov[0][0] * cov[1][1] - cov[0][1] * cov[1][0];
                const double signJ = std::fabs(Jb) / Jb;
                double d0, d1;
                if (side == 0 || side == 2) { d0 = sgn * signJ * (-cov[1][0]); d1 = sgn * signJ * cov[0][0]; }
                else { d0 = sgn * signJ * cov[1][1]; d1 = sgn * signJ * (-cov[0][1]); }
                const double len = std::sqrt(d0 * d0 + d1 * d1);
                const size_t nb = i + (size_t)n1 * (side + 4 * (size_t)e);
                nLen[nb] = len;
                nHat[2 * nb + 0] = d0 / len;
                nHat[2 * nb + 1] = d1 / len;
            }
    }
    return SELF_OK;
}
SELF_CATCH_ALL

}  // extern "C"
