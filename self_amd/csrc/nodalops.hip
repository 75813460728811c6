// nodalops.hip -- the nodal storage library operators as batched device
// kernels: derivative matrix application, boundary interpolation and
// coarse-to-fine interpolation, for many elements per launch.
//
// Reference:
//   ApplyDerivativeMatrix_NodalStorage_2D  src/nodal/NodalStorage_2D_Class.f90:620-646
//   CalculateAtBoundaries_NodalStorage_2D  :650-680
//   ApplyDerivativeMatrix_NodalStorage_3D  src/nodal/NodalStorage_3D_Class.f90:737-774
//   CalculateAtBoundaries_NodalStorage_3D  :778-823
//   CoarseToFine_2D  src/interp/Lagrange_2D_Class.f90:488-528 (returns the
//     interpolant with the new p index first: MATMUL(Ts, f), transposed,
//     then MATMUL(Tp, .))
//   CoarseToFine_3D  src/interp/Lagrange_3D_Class.f90:545-604
// The reference applies each to one element; here nElems elements are laid
// out one after the other (element slowest), each in the reference's
// column-major per-element layout.  One thread computes one output value
// from a line of its element; the lines of one element stay in L1/L2
// across the threads that share them.  Bound: HBM (read f once, write the
// outputs once).

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace {

__global__ void k_deriv2d(int n1, size_t nE, const double *__restrict__ dS, const double *__restrict__ dP,
                          const double *__restrict__ f, double *__restrict__ ds, double *__restrict__ dp)
{
    const size_t np = (size_t)n1 * n1;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < nE * np; q += (size_t)gridDim.x * blockDim.x) {
        const size_t e = q / np;
        const int r = (int)(q - e * np), a = r % n1, b = r / n1;
        const double *fe = f + e * np;
        double s = 0.0, p = 0.0;
        for (int j = 0; j < n1; ++j) {
            s += dS[a + n1 * j] * fe[j + n1 * b];  // dFds(a, b) = sum_j dMatS(a, j) f(j, b)
            p += dP[b + n1 * j] * fe[a + n1 * j];  // dFdp(a, b) = sum_j dMatP(b, j) f(a, j)
        }
        ds[q] = s;
        dp[q] = p;
    }
}

__global__ void k_bound2d(int n1, size_t nE, const double *__restrict__ lW, const double *__restrict__ lE,
                          const double *__restrict__ lS, const double *__restrict__ lN, const double *__restrict__ f,
                          double *__restrict__ fW, double *__restrict__ fE, double *__restrict__ fS,
                          double *__restrict__ fN)
{
    const size_t np = (size_t)n1 * n1;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < nE * n1; q += (size_t)gridDim.x * blockDim.x) {
        const size_t e = q / n1;
        const int i = (int)(q - e * n1);
        const double *fe = f + e * np;
        double w = 0.0, ea = 0.0, s = 0.0, no = 0.0;
        for (int k = 0; k < n1; ++k) {
            w += lW[k] * fe[k + n1 * i];
            ea += lE[k] * fe[k + n1 * i];
            s += lS[k] * fe[i + n1 * k];
            no += lN[k] * fe[i + n1 * k];
        }
        fW[q] = w;
        fE[q] = ea;
        fS[q] = s;
        fN[q] = no;
    }
}

__global__ void k_deriv3d(int n1, size_t nE, const double *__restrict__ D, const double *__restrict__ f,
                          double *__restrict__ ds, double *__restrict__ dp, double *__restrict__ dq)
{
    const size_t np = (size_t)n1 * n1 * n1;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < nE * np; q += (size_t)gridDim.x * blockDim.x) {
        const size_t e = q / np;
        const int r = (int)(q - e * np), a = r % n1, b = (r / n1) % n1, c = r / (n1 * n1);
        const double *fe = f + e * np;
        double s = 0.0, p = 0.0, z = 0.0;
        for (int m = 0; m < n1; ++m) {
            s += D[a + n1 * m] * fe[m + n1 * (b + n1 * c)];
            p += D[b + n1 * m] * fe[a + n1 * (m + n1 * c)];
            z += D[c + n1 * m] * fe[a + n1 * (b + n1 * m)];
        }
        ds[q] = s;
        dp[q] = p;
        dq[q] = z;
    }
}

// faces per element: south, east, north, west, bottom, top, each (0:N,0:N)
__global__ void k_bound3d(int n1, size_t nE, const double *__restrict__ lL, const double *__restrict__ lR,
                          const double *__restrict__ f, double *__restrict__ faces)
{
    const size_t nf = (size_t)n1 * n1, np = nf * n1;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < nE * nf; q += (size_t)gridDim.x * blockDim.x) {
        const size_t e = q / nf;
        const int r = (int)(q - e * nf), a = r % n1, b = r / n1;
        const double *fe = f + e * np;
        double w = 0, ea = 0, s = 0, no = 0, bo = 0, to = 0;
        for (int k = 0; k < n1; ++k) {
            w += lL[k] * fe[k + n1 * (a + n1 * b)];
            ea += lR[k] * fe[k + n1 * (a + n1 * b)];
            s += lL[k] * fe[a + n1 * (k + n1 * b)];
            no += lR[k] * fe[a + n1 * (k + n1 * b)];
            bo += lL[k] * fe[a + n1 * (b + n1 * k)];
            to += lR[k] * fe[a + n1 * (b + n1 * k)];
        }
        double *out = faces + e * 6 * nf + r;
        out[0 * nf] = s;
        out[1 * nf] = ea;
        out[2 * nf] = no;
        out[3 * nf] = w;
        out[4 * nf] = bo;
        out[5 * nf] = to;
    }
}

// fNew(a, b), a over the new p nodes and b over the new s nodes:
// sum_j Tp(a, j) sum_k Ts(b, k) f(k, j)
__global__ void k_c2f_2d(int s1, int p1, int ms, int mp, size_t nE, const double *__restrict__ Ts,
                         const double *__restrict__ Tp, const double *__restrict__ f, double *__restrict__ fNew)
{
    const size_t no = (size_t)mp * ms, ni = (size_t)s1 * p1;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < nE * no; q += (size_t)gridDim.x * blockDim.x) {
        const size_t e = q / no;
        const int r = (int)(q - e * no), a = r % mp, b = r / mp;
        const double *fe = f + e * ni;
        double acc = 0.0;
        for (int j = 0; j < p1; ++j) {
            double t = 0.0;
            for (int k = 0; k < s1; ++k) t += Ts[b + ms * k] * fe[k + s1 * j];
            acc += Tp[a + mp * j] * t;
        }
        fNew[q] = acc;
    }
}

// fNew(a, b, c) = sum T(c, k) sum T(b, j) sum T(a, i) f(i, j, k)
__global__ void k_c2f_3d(int n1, int m1, size_t nE, const double *__restrict__ T, const double *__restrict__ f,
                         double *__restrict__ fNew)
{
    const size_t no = (size_t)m1 * m1 * m1, ni = (size_t)n1 * n1 * n1;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < nE * no; q += (size_t)gridDim.x * blockDim.x) {
        const size_t e = q / no;
        const int r = (int)(q - e * no), a = r % m1, b = (r / m1) % m1, c = r / (m1 * m1);
        const double *fe = f + e * ni;
        double acc = 0.0;
        for (int k = 0; k < n1; ++k) {
            double tj = 0.0;
            for (int j = 0; j < n1; ++j) {
                double ti = 0.0;
                for (int i = 0; i < n1; ++i) ti += T[a + m1 * i] * fe[i + n1 * (j + n1 * k)];
                tj += T[b + m1 * j] * ti;
            }
            acc += T[c + m1 * k] * tj;
        }
        fNew[q] = acc;
    }
}

int grid_for(size_t work)
{
    const size_t g = (work + 255) / 256;
    return (int)(g < 2048 ? (g > 0 ? g : 1) : 2048);
}

// device buffers of one host-array call
struct Staging {
    double *p[12] = {};
    int n = 0;
    hipError_t e = hipSuccess;
    double *alloc(size_t count)
    {
        double *q = nullptr;
        if (e == hipSuccess) e = hipMalloc(&q, sizeof(double) * (count ? count : 1));
        p[n++] = q;
        return q;
    }
    double *upload(const double *h, size_t count)
    {
        double *q = alloc(count);
        if (e == hipSuccess) e = hipMemcpy(q, h, sizeof(double) * count, hipMemcpyHostToDevice);
        return q;
    }
    void download(double *h, const double *d, size_t count)
    {
        if (e == hipSuccess) e = hipMemcpy(h, d, sizeof(double) * count, hipMemcpyDeviceToHost);
    }
    ~Staging()
    {
        for (int k = 0; k < n; ++k)
            if (p[k]) (void)hipFree(p[k]);
    }
};

}  // namespace

extern "C" {

// ------------------------------------------------------------ device-pointer forms
int self_nodal2d_apply_derivative_dev(int N, size_t nElems, const double *dMatS, const double *dMatP,
                                      const double *f, double *dFds, double *dFdp, void *stream)
{
    SELF_REQUIRE(N >= 0 && dMatS && dMatP && f && dFds && dFdp, SELF_ERR_ARG, "apply_derivative_2d: bad arguments");
    if (!nElems) return SELF_OK;
    const size_t w = nElems * (size_t)(N + 1) * (N + 1);
    hipLaunchKernelGGL(k_deriv2d, dim3(grid_for(w)), dim3(256), 0, (hipStream_t)stream, N + 1, nElems, dMatS, dMatP,
                       f, dFds, dFdp);
    SELF_HIP(hipGetLastError());
    return SELF_OK;
}

int self_nodal2d_boundary_interp_dev(int N, size_t nElems, const double *lW, const double *lE, const double *lS,
                                     const double *lN, const double *f, double *fW, double *fE, double *fS,
                                     double *fN, void *stream)
{
    SELF_REQUIRE(N >= 0 && lW && lE && lS && lN && f && fW && fE && fS && fN, SELF_ERR_ARG,
                 "boundary_interp_2d: bad arguments");
    if (!nElems) return SELF_OK;
    hipLaunchKernelGGL(k_bound2d, dim3(grid_for(nElems * (N + 1))), dim3(256), 0, (hipStream_t)stream, N + 1, nElems,
                       lW, lE, lS, lN, f, fW, fE, fS, fN);
    SELF_HIP(hipGetLastError());
    return SELF_OK;
}

int self_nodal3d_apply_derivative_dev(int N, size_t nElems, const double *dMat, const double *f, double *dFds,
                                      double *dFdp, double *dFdq, void *stream)
{
    SELF_REQUIRE(N >= 0 && dMat && f && dFds && dFdp && dFdq, SELF_ERR_ARG, "apply_derivative_3d: bad arguments");
    if (!nElems) return SELF_OK;
    const size_t w = nElems * (size_t)(N + 1) * (N + 1) * (N + 1);
    hipLaunchKernelGGL(k_deriv3d, dim3(grid_for(w)), dim3(256), 0, (hipStream_t)stream, N + 1, nElems, dMat, f, dFds,
                       dFdp, dFdq);
    SELF_HIP(hipGetLastError());
    return SELF_OK;
}

int self_nodal3d_boundary_interp_dev(int N, size_t nElems, const double *lagLeft, const double *lagRight,
                                     const double *f, double *faces, void *stream)
{
    SELF_REQUIRE(N >= 0 && lagLeft && lagRight && f && faces, SELF_ERR_ARG, "boundary_interp_3d: bad arguments");
    if (!nElems) return SELF_OK;
    hipLaunchKernelGGL(k_bound3d, dim3(grid_for(nElems * (N + 1) * (N + 1))), dim3(256), 0, (hipStream_t)stream,
                       N + 1, nElems, lagLeft, lagRight, f, faces);
    SELF_HIP(hipGetLastError());
    return SELF_OK;
}

int self_coarse_to_fine_2d_dev(int nS, int nP, size_t nElems, const double *Ts, const double *Tp, int nNewS,
                               int nNewP, const double *fOld, double *fNew, void *stream)
{
    SELF_REQUIRE(nS >= 0 && nP >= 0 && nNewS >= 0 && nNewP >= 0 && Ts && Tp && fOld && fNew, SELF_ERR_ARG,
                 "coarse_to_fine_2d: bad arguments");
    if (!nElems) return SELF_OK;
    hipLaunchKernelGGL(k_c2f_2d, dim3(grid_for(nElems * (nNewS + 1) * (nNewP + 1))), dim3(256), 0,
                       (hipStream_t)stream, nS + 1, nP + 1, nNewS + 1, nNewP + 1, nElems, Ts, Tp, fOld, fNew);
    SELF_HIP(hipGetLastError());
    return SELF_OK;
}

int self_coarse_to_fine_3d_dev(int N, size_t nElems, const double *T, int nNew, const double *fOld, double *fNew,
                               void *stream)
{
    SELF_REQUIRE(N >= 0 && nNew >= 0 && T && fOld && fNew, SELF_ERR_ARG, "coarse_to_fine_3d: bad arguments");
    if (!nElems) return SELF_OK;
    const size_t w = nElems * (size_t)(nNew + 1) * (nNew + 1) * (nNew + 1);
    hipLaunchKernelGGL(k_c2f_3d, dim3(grid_for(w)), dim3(256), 0, (hipStream_t)stream, N + 1, nNew + 1, nElems, T,
                       fOld, fNew);
    SELF_HIP(hipGetLastError());
    return SELF_OK;
}

// ------------------------------------------------------------ host-array forms
int self_nodal2d_apply_derivative(int N, size_t nElems, const double *dMatS, const double *dMatP, const double *f,
                                  double *dFds, double *dFdp)
try {
    SELF_REQUIRE(N >= 0 && dMatS && dMatP && f && dFds && dFdp, SELF_ERR_ARG, "apply_derivative_2d: bad arguments");
    const size_t n1 = N + 1, cnt = nElems * n1 * n1;
    Staging S;
    double *a = S.upload(dMatS, n1 * n1), *b = S.upload(dMatP, n1 * n1), *df = S.upload(f, cnt);
    double *o1 = S.alloc(cnt), *o2 = S.alloc(cnt);
    SELF_HIP(S.e);
    int rc = self_nodal2d_apply_derivative_dev(N, nElems, a, b, df, o1, o2, nullptr);
    if (rc) return rc;
    S.download(dFds, o1, cnt);
    S.download(dFdp, o2, cnt);
    SELF_HIP(S.e);
    return SELF_OK;
}
SELF_CATCH_ALL

int self_nodal2d_boundary_interp(int N, size_t nElems, const double *lW, const double *lE, const double *lS,
                                 const double *lN, const double *f, double *fW, double *fE, double *fS, double *fN)
try {
    SELF_REQUIRE(N >= 0 && lW && lE && lS && lN && f && fW && fE && fS && fN, SELF_ERR_ARG,
                 "boundary_interp_2d: bad arguments");
    const size_t n1 = N + 1, cnt = nElems * n1 * n1, nb = nElems * n1;
    Staging S;
    double *a = S.upload(lW, n1), *b = S.upload(lE, n1), *c = S.upload(lS, n1), *d = S.upload(lN, n1);
    double *df = S.upload(f, cnt), *o[4];
    for (auto &q : o) q = S.alloc(nb);
    SELF_HIP(S.e);
    int rc = self_nodal2d_boundary_interp_dev(N, nElems, a, b, c, d, df, o[0], o[1], o[2], o[3], nullptr);
    if (rc) return rc;
    S.download(fW, o[0], nb);
    S.download(fE, o[1], nb);
    S.download(fS, o[2], nb);
    S.download(fN, o[3], nb);
    SELF_HIP(S.e);
    return SELF_OK;
}
SELF_CATCH_ALL

int self_nodal3d_apply_derivative(int N, size_t nElems, const double *dMat, const double *f, double *dFds,
                                  double *dFdp, double *dFdq)
try {
    SELF_REQUIRE(N >= 0 && dMat && f && dFds && dFdp && dFdq, SELF_ERR_ARG, "apply_derivative_3d: bad arguments");
    const size_t n1 = N + 1, cnt = nElems * n1 * n1 * n1;
    Staging S;
    double *a = S.upload(dMat, n1 * n1), *df = S.upload(f, cnt);
    double *o1 = S.alloc(cnt), *o2 = S.alloc(cnt), *o3 = S.alloc(cnt);
    SELF_HIP(S.e);
    int rc = self_nodal3d_apply_derivative_dev(N, nElems, a, df, o1, o2, o3, nullptr);
    if (rc) return rc;
    S.download(dFds, o1, cnt);
    S.download(dFdp, o2, cnt);
    S.download(dFdq, o3, cnt);
    SELF_HIP(S.e);
    return SELF_OK;
}
SELF_CATCH_ALL

int self_nodal3d_boundary_interp(int N, size_t nElems, const double *lagLeft, const double *lagRight,
                                 const double *f, double *faces)
try {
    SELF_REQUIRE(N >= 0 && lagLeft && lagRight && f && faces, SELF_ERR_ARG, "boundary_interp_3d: bad arguments");
    const size_t n1 = N + 1, cnt = nElems * n1 * n1 * n1, nf = nElems * 6 * n1 * n1;
    Staging S;
    double *a = S.upload(lagLeft, n1), *b = S.upload(lagRight, n1), *df = S.upload(f, cnt), *o = S.alloc(nf);
    SELF_HIP(S.e);
    int rc = self_nodal3d_boundary_interp_dev(N, nElems, a, b, df, o, nullptr);
    if (rc) return rc;
    S.download(faces, o, nf);
    SELF_HIP(S.e);
    return SELF_OK;
}
SELF_CATCH_ALL

int self_coarse_to_fine_2d(int nS, int nP, size_t nElems, const double *Ts, const double *Tp, int nNewS, int nNewP,
                           const double *fOld, double *fNew)
try {
    SELF_REQUIRE(nS >= 0 && nP >= 0 && nNewS >= 0 && nNewP >= 0 && Ts && Tp && fOld && fNew, SELF_ERR_ARG,
                 "coarse_to_fine_2d: bad arguments");
    const size_t ci = nElems * (nS + 1) * (nP + 1), co = nElems * (nNewS + 1) * (nNewP + 1);
    Staging S;
    double *a = S.upload(Ts, (size_t)(nNewS + 1) * (nS + 1)), *b = S.upload(Tp, (size_t)(nNewP + 1) * (nP + 1));
    double *df = S.upload(fOld, ci), *o = S.alloc(co);
    SELF_HIP(S.e);
    int rc = self_coarse_to_fine_2d_dev(nS, nP, nElems, a, b, nNewS, nNewP, df, o, nullptr);
    if (rc) return rc;
    S.download(fNew, o, co);
    SELF_HIP(S.e);
    return SELF_OK;
}
SELF_CATCH_ALL

int self_coarse_to_fine_3d(int N, size_t nElems, const double *T, int nNew, const double *fOld, double *fNew)
try {
    SELF_REQUIRE(N >= 0 && nNew >= 0 && T && fOld && fNew, SELF_ERR_ARG, "coarse_to_fine_3d: bad arguments");
    const size_t ci = nElems * (size_t)(N + 1) * (N + 1) * (N + 1);
    const size_t co = nElems * (size_t)(nNew + 1) * (nNew + 1) * (nNew + 1);
    Staging S;
    double *a = S.upload(T, (size_t)(nNew + 1) * (N + 1)), *df = S.upload(fOld, ci), *o = S.alloc(co);
    SELF_HIP(S.e);
    int rc = self_coarse_to_fine_3d_dev(N, nElems, a, nNew, df, o, nullptr);
    if (rc) return rc;
    S.download(fNew, o, co);
    SELF_HIP(S.e);
    return SELF_OK;
}
SELF_CATCH_ALL

}  // extern "C"
