// setup.cpp -- host-side set-up of the spectral-element operators and of
// structured quad meshes.  These run once per model on the CPU, as the
// reference's Build routines do; the results are uploaded to the device.
//
//   Legendre quadrature      src/interp/Legendre.f90:44-405
//   barycentric Lagrange     src/interp/Lagrange_1D_Class.f90:349-764
//   2-D nodal storage        src/nodal/NodalStorage_2D_Class.f90:101-230
//   structured quad mesh     src/geom/QuadMeshClass.f90:1489-1627,1718-1859
//   element metrics/normals  src/geom/MappedGeometryClass_2D.f90:611-710,921-948
//
// The quadrature keeps the reference's Newton iteration verbatim (initial
// guesses, 50 iterations, the one-sided stop test and the root loop bound
// (nP+1)/2) because the node ORDER it produces is part of the data layout a
// drop-in exchanges with the host program: for nP = 3 the reference returns
// the interior Gauss nodes in the order +x, -x.

#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "common.hpp"

namespace self {

static thread_local std::string g_err;

void set_error(const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
}

const char *last_error() { return g_err.c_str(); }

namespace {

constexpr int kNewtonIters = 50;  // ConstantsDictionary.f90: kItMax

// P_n(x) and P_n'(x) by the three-term recurrence (Legendre.f90:44-114)
void legendre_p(int n, double x, double &p, double &dp)
{
    if (n == 0) { p = 1.0; dp = 0.0; return; }
    if (n == 1) { p = x; dp = 1.0; return; }
    double pm2 = 1.0, pm1 = x, dpm2 = 0.0, dpm1 = 1.0;
    for (int k = 2; k <= n; ++k) {
        const double rk = k;
        p = ((2.0 * rk - 1.0) * x * pm1 - (rk - 1.0) * pm2) / rk;
        dp = dpm2 + (2.0 * rk - 1.0) * pm1;
        pm2 = pm1; pm1 = p; dpm2 = dpm1; dpm1 = dp;
    }
}

// q = P_{n+1} - P_{n-1}, q', and P_n (Legendre.f90:118-174)
void legendre_q(int n, double x, double &q, double &dq, double &pn)
{
    double pm2 = 1.0, pm1 = x, dpm2 = 0.0, dpm1 = 1.0, p = 0.0, dp;
    for (int k = 2; k <= n; ++k) {
        const double rk = k;
        p = (2.0 * k - 1.0) / rk * x * pm1 - (rk - 1.0) / rk * pm2;
        dp = dpm2 + (2.0 * rk - 1.0) * pm1;
        pm2 = pm1; pm1 = p; dpm2 = dpm1; dpm1 = dp;
    }
    const int k = n + 1;
    const double rk = k;
    const double pp1 = (2.0 * k - 1.0) / rk * x * p - (rk - 1.0) / rk * pm2;
    const double dpp1 = dpm2 + (2.0 * rk - 1.0) * pm1;
    q = pp1 - pm2;
    dq = dpp1 - dpm2;
    pn = p;
}

void gauss(int n, double *x, double *w)
{
    const double pi = 4.0 * std::atan(1.0);
    double p, dp;
    if (n == 0) { x[0] = 0.0; w[0] = 2.0; }
    else if (n == 1) { x[0] = -std::sqrt(1.0 / 3.0); w[0] = 1.0; x[1] = -x[0]; w[1] = 1.0; }
    else {
        for (int r = 0; r <= (n + 1) / 2; ++r) {
            double xr = -std::cos((2.0 * r + 1.0) * pi / (2.0 * n + 1.0));
            for (int it = 0; it < kNewtonIters; ++it) {
                legendre_p(n + 1, xr, p, dp);
                const double d = -p / dp;
                xr += d;
                if (std::fabs(d) <= DBL_EPSILON * xr) break;
            }
            legendre_p(n + 1, xr, p, dp);
            x[r] = xr;
            w[r] = 2.0 / ((1.0 - xr * xr) * dp * dp);
            w[n - r] = w[r];
            x[n - r] = -xr;
        }
    }
    if (n % 2 == 0) {
        legendre_p(n + 1, 0.0, p, dp);
        x[n / 2] = 0.0;
        w[n / 2] = 2.0 / (dp * dp);
    }
}

void gauss_lobatto(int n, double *x, double *w)
{
    const double pi = 4.0 * std::atan(1.0);
    double q, dq, pn;
    const double rn = n;
    if (n == 1) { x[0] = -1.0; w[0] = 1.0; x[1] = 1.0; w[1] = 1.0; }
    else {
        x[0] = -1.0; w[0] = 2.0 / (rn * (rn + 1.0));
        x[n] = 1.0; w[n] = w[0];
        for (int r = 1; r <= (n + 1) / 2 - 1; ++r) {
            double xr = -std::cos((r + 0.25) * pi / rn - 3.0 / (8.0 * rn * pi * (r + 0.25)));
            for (int it = 0; it < kNewtonIters; ++it) {
                legendre_q(n, xr, q, dq, pn);
                const double d = -q / dq;
                xr += d;
                if (std::fabs(d) <= DBL_EPSILON * xr) break;
            }
            legendre_q(n, xr, q, dq, pn);
            x[r] = xr;
            w[r] = 2.0 / (rn * (rn + 1.0) * pn * pn);
            w[n - r] = w[r];
            x[n - r] = -xr;
        }
    }
    if (n % 2 == 0) {
        legendre_q(n, 0.0, q, dq, pn);
        x[n / 2] = 0.0;
        w[n / 2] = 2.0 / (rn * (rn + 1.0) * pn * pn);
    }
}

// AlmostEqual (src/common/CommonRoutines.f90:58-97)
bool almost_equal(double a, double b)
{
    const double d = std::fabs(a - b);
    if (a == 0.0 || b == 0.0) return d <= DBL_EPSILON;
    return d <= DBL_EPSILON * std::fabs(a) || d <= DBL_EPSILON * std::fabs(b);
}

}  // namespace

int legendre_quadrature(int n, int type, double *x, double *w)
{
    if (n < 0 || !x || !w) { set_error("self_legendre_quadrature: bad arguments"); return SELF_ERR_ARG; }
    if (type == SELF_GAUSS) gauss(n, x, w);
    else if (type == SELF_GAUSS_LOBATTO) {
        if (n < 1) { set_error("Gauss-Lobatto needs N >= 1"); return SELF_ERR_ARG; }
        gauss_lobatto(n, x, w);
    } else {
        set_error("GenerateLegendreQuadrature: invalid form %d", type);
        return SELF_ERR_ARG;
    }
    return SELF_OK;
}

void barycentric(int n, const double *x, double *bw)
{
    for (int j = 0; j <= n; ++j) bw[j] = 1.0;
    for (int j = 1; j <= n; ++j)
        for (int k = 0; k < j; ++k) {
            bw[k] *= (x[k] - x[j]);
            bw[j] *= (x[j] - x[k]);
        }
    for (int j = 0; j <= n; ++j) bw[j] = 1.0 / bw[j];
}

void lagrange_at(int n, const double *x, const double *bw, double xe, double *l)
{
    bool hit = false;
    for (int j = 0; j <= n; ++j) {
        l[j] = 0.0;
        if (almost_equal(xe, x[j])) { l[j] = 1.0; hit = true; }
    }
    if (hit) return;
    double sum = 0.0;
    for (int j = 0; j <= n; ++j) {
        l[j] = bw[j] / (xe - x[j]);
        sum += l[j];
    }
    for (int j = 0; j <= n; ++j) l[j] /= sum;
}

void derivative_matrix(int n, const double *x, const double *bw, double *D)
{
    const int n1 = n + 1;
    for (int k = 0; k <= n; ++k) {
        double diag = 0.0;
        for (int j = 0; j <= n; ++j) {
            if (j == k) continue;
            const double v = bw[j] / (bw[k] * (x[k] - x[j]));
            D[k + n1 * j] = v;
            diag -= v;  // negative-sum trick, accumulated in column order
        }
        D[k + n1 * k] = diag;
    }
}

void interpolation_matrix(int n, const double *x, const double *bw, int m, const double *z, double *T)
{
    const int m1 = m + 1;
    for (int k = 0; k <= m; ++k) {
        bool hit = false;
        for (int j = 0; j <= n; ++j) {
            T[k + m1 * j] = 0.0;
            if (almost_equal(z[k], x[j])) { T[k + m1 * j] = 1.0; hit = true; }
        }
        if (hit) continue;
        double sum = 0.0;
        for (int j = 0; j <= n; ++j) {
            T[k + m1 * j] = bw[j] / (z[k] - x[j]);
            sum += T[k + m1 * j];
        }
        for (int j = 0; j <= n; ++j) T[k + m1 * j] /= sum;
    }
}

}  // namespace self

// ------------------------------------------------------------------ C ABI

using namespace self;

extern "C" {

const char *self_last_error(void) { return self::last_error(); }
int self_version(void) { return 100; }

int self_legendre_quadrature(int N, int quadType, double *x, double *w)
{
    return legendre_quadrature(N, quadType, x, w);
}

int self_barycentric_weights(int N, const double *x, double *bw)
{
    SELF_REQUIRE(N >= 0 && x && bw, SELF_ERR_ARG, "self_barycentric_weights: bad arguments");
    barycentric(N, x, bw);
    return SELF_OK;
}

int self_lagrange_polynomials(int N, const double *x, const double *bw, double xEval, double *l)
{
    SELF_REQUIRE(N >= 0 && x && bw && l, SELF_ERR_ARG, "self_lagrange_polynomials: bad arguments");
    lagrange_at(N, x, bw, xEval, l);
    return SELF_OK;
}

int self_derivative_matrix(int N, const double *x, const double *bw, double *D)
{
    SELF_REQUIRE(N >= 0 && x && bw && D, SELF_ERR_ARG, "self_derivative_matrix: bad arguments");
    derivative_matrix(N, x, bw, D);
    return SELF_OK;
}

int self_interpolation_matrix(int N, const double *x, const double *bw, int nNew,
                              const double *zNew, double *T)
{
    SELF_REQUIRE(N >= 0 && nNew >= 0 && x && bw && zNew && T, SELF_ERR_ARG,
                 "self_interpolation_matrix: bad arguments");
    interpolation_matrix(N, x, bw, nNew, zNew, T);
    return SELF_OK;
}

int self_nodal2d_build(int N, int quadType, int form, int quirk, double *nodes, double *qw,
                       double *dMatS, double *dMatP, double *lW, double *lE, double *lS, double *lN)
{
    SELF_REQUIRE(N >= 1 && nodes && qw && dMatS && dMatP && lW && lE && lS && lN, SELF_ERR_ARG,
                 "self_nodal2d_build: bad arguments");
    SELF_REQUIRE(form == SELF_DG || form == SELF_CG, SELF_ERR_ARG,
                 "BuildNodalStorage_2D: invalid SEM form %d", form);
    int rc = legendre_quadrature(N, quadType, nodes, qw);
    if (rc) return rc;
    const int n1 = N + 1;
    std::vector<double> bw(n1), D(n1 * n1);
    barycentric(N, nodes, bw.data());
    lagrange_at(N, nodes, bw.data(), -1.0, lW);
    lagrange_at(N, nodes, bw.data(), 1.0, lE);
    lagrange_at(N, nodes, bw.data(), -1.0, lS);
    lagrange_at(N, nodes, bw.data(), 1.0, lN);
    derivative_matrix(N, nodes, bw.data(), D.data());
    if (form == SELF_CG) {
        std::memcpy(dMatS, D.data(), sizeof(double) * n1 * n1);
        std::memcpy(dMatP, D.data(), sizeof(double) * n1 * n1);
        return SELF_OK;
    }
    // weak-form operator  Dhat(i,j) = -D(j,i) w_j / w_i
    for (int j = 0; j <= N; ++j)
        for (int i = 0; i <= N; ++i) dMatS[i + n1 * j] = -D[j + n1 * i] * qw[j] / qw[i];
    for (int j = 0; j <= N; ++j)
        for (int i = 0; i <= N; ++i)
            dMatP[i + n1 * j] = quirk ? dMatS[j + n1 * i] : dMatS[i + n1 * j];
    return SELF_OK;
}

// BuildNodalStorage_3D (src/nodal/NodalStorage_3D_Class.f90:107-264) with
// equal degree in the three directions: nodes, weights, the derivative
// matrix (weak DG operator -W^-1 D^T W for SELF_DG, :223-251) and the
// Lagrange vectors at -1 (west/south/bottom) and +1 (east/north/top).
int self_nodal3d_build(int N, int quadType, int form, double *nodes, double *qw, double *dMat, double *lagLeft,
                       double *lagRight)
{
    SELF_REQUIRE(N >= 1 && nodes && qw && dMat && lagLeft && lagRight, SELF_ERR_ARG, "self_nodal3d_build: bad arguments");
    SELF_REQUIRE(form == SELF_DG || form == SELF_CG, SELF_ERR_ARG, "BuildNodalStorage_3D: invalid SEM form %d", form);
    int rc = legendre_quadrature(N, quadType, nodes, qw);
    if (rc) return rc;
    const int n1 = N + 1;
    std::vector<double> bw(n1), D(n1 * n1);
    barycentric(N, nodes, bw.data());
    lagrange_at(N, nodes, bw.data(), -1.0, lagLeft);
    lagrange_at(N, nodes, bw.data(), 1.0, lagRight);
    derivative_matrix(N, nodes, bw.data(), D.data());
    for (int j = 0; j <= N; ++j)
        for (int i = 0; i <= N; ++i)
            dMat[i + n1 * j] = (form == SELF_CG) ? D[i + n1 * j] : -D[j + n1 * i] * qw[j] / qw[i];
    return SELF_OK;
}

// One-direction factor of RollOffFilter3D (src/filters/RollOffFilter3D_Class.f90
// :87-180).  The reference's dense filter is VInv diag(exp(-alpha r)) V with
// r = a^2 + b^2 + c^2 over Legendre tensor modes, alpha = ln 2 / nc^2; the
// modes, the weights and exp(-alpha r) all factor over the three directions,
// so the dense matrix equals F1 (x) F1 (x) F1 with
//   F1(i, m) = sum_a L_a(x_i) exp(-alpha a^2) L_a(x_m) w_m / n_a,
//   n_a = sum_i L_a(x_i)^2 w_i,
// which is applied as three (N+1) x (N+1) passes.  Column-major F1[i + (N+1) m].
int self_rolloff_filter1d(int N, int nCutoff, const double *nodes, const double *w, double *F1)
{
    SELF_REQUIRE(N >= 1 && nCutoff >= 1 && nodes && w && F1, SELF_ERR_ARG, "self_rolloff_filter1d: bad arguments");
    const int n1 = N + 1;
    std::vector<double> L(n1 * n1), nrm(n1);
    for (int a = 0; a <= N; ++a) {
        double acc = 0.0;
        for (int i = 0; i <= N; ++i) {
            double p, dp;
            legendre_p(a, nodes[i], p, dp);
            L[a + n1 * i] = p;
            acc += p * p * w[i];
        }
        nrm[a] = acc;
    }
    const double sc = (double)nCutoff, alpha = std::log(2.0) / (sc * sc);
    for (int m = 0; m <= N; ++m)
        for (int i = 0; i <= N; ++i) {
            double acc = 0.0;
            for (int a = 0; a <= N; ++a)
                acc += L[a + n1 * i] * std::exp(-alpha * (double)(a * a)) * L[a + n1 * m] * w[m] / nrm[a];
            F1[i + n1 * m] = acc;
        }
    return SELF_OK;
}

}  // extern "C"
