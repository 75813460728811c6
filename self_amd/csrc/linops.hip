// linops.hip -- operators for the generic Krylov solvers (the concrete
// types a reference model would derive from IterativeSolver):
//   Dense:      y = A x, residual b - A x, Jacobi = diag(A)^-1 (test and
//               small-system operator).
//   Elliptic1D: the 1-D CG spectral-element operator of
//               examples/elliptic1d/parabola/localmods/Elliptic1D_Class.f90
//               (Build d2Mat :150-172, MatrixAction :419-453 with
//               CopyFrom/UnMaskSolution/ApplyOperator/GlobalSum/Mask,
//               Residual :457-503 which masks without a GlobalSum -- quirk
//               Q10 kept, it is what produced the reference fixture).
// Both work on the flat vectors s(1:nDOF) the reference solvers see.

#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "krylov.hpp"

namespace self {
namespace {

__global__ __launch_bounds__(256) void k_dense_matvec(const double *__restrict__ A, const double *__restrict__ x,
                                                      double *__restrict__ y, const double *__restrict__ b, int n,
                                                      int residual, const SolverState *st)
{
    if (st && (st->done || (!residual && st->innerExit))) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double acc = 0.0;
        for (int j = 0; j < n; ++j) acc += A[(size_t)i + (size_t)n * j] * x[j];
        y[i] = residual ? b[i] - acc : acc;
    }
}

// One workgroup applies the whole 1-D operator (nDOF is small).
__global__ __launch_bounds__(256) void k_e1d_apply(const double *__restrict__ s, double *__restrict__ out,
                                                   const double *__restrict__ d2, const double *__restrict__ sJw,
                                                   int N, int nEl, double mL, double mR, double bfL, double bfR,
                                                   int residual, const SolverState *st)
{
    if (st && (st->done || (!residual && st->innerExit))) return;
    extern __shared__ double sh[];
    const int n1 = N + 1, nd = n1 * nEl;
    double *u = sh, *Au = sh + nd;
    for (int k = threadIdx.x; k < nd; k += blockDim.x) u[k] = s[k];  // CopyFrom
    __syncthreads();
    if (threadIdx.x == 0) {  // UnMaskSolution: UnMask + DirichletConditions
        for (int e = 1; e < nEl; ++e) u[n1 * e] = u[N + n1 * (e - 1)];
        if (mL == 0.0) u[0] = 0.0;
        if (mR == 0.0) u[N + n1 * (nEl - 1)] = 0.0;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < nd; k += blockDim.x) {  // ApplyOperator
        const int e = k / n1, row = k - e * n1;
        double acc = 0.0;
        for (int col = 0; col <= N; ++col) acc += d2[row + n1 * (col + n1 * e)] * u[col + n1 * e];
        Au[k] = acc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int e = 1; e < nEl; ++e) {  // GlobalSum
            const double t = Au[n1 * e] + Au[N + n1 * (e - 1)];
            Au[n1 * e] = t;
            Au[N + n1 * (e - 1)] = t;
        }
        for (int e = 1; e < nEl; ++e) Au[n1 * e] = 0.0;  // Mask
        Au[0] = Au[0] * mL;
        Au[N + n1 * (nEl - 1)] = Au[N + n1 * (nEl - 1)] * mR;
        if (residual) {  // r = source*J*w - Au, boundary fluxes, Mask (no GlobalSum)
            for (int k = 0; k < nd; ++k) Au[k] = sJw[k] - Au[k];
            Au[0] = Au[0] + bfL * mL;
            Au[N + n1 * (nEl - 1)] = Au[N + n1 * (nEl - 1)] - bfR * mR;
            for (int e = 1; e < nEl; ++e) Au[n1 * e] = 0.0;
            Au[0] = Au[0] * mL;
            Au[N + n1 * (nEl - 1)] = Au[N + n1 * (nEl - 1)] * mR;
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < nd; k += blockDim.x) out[k] = Au[k];
}

template <typename T, bool FIN>
__global__ __launch_bounds__(kRedThreads) void k_op_dot(const T *__restrict__ a, const T *__restrict__ b, size_t n,
                                                        SolverState *st, int fin, RedSlot slot, int residual)
{
    if (st->done || (!residual && st->innerExit)) return;
    __shared__ double lds4[4];
    if constexpr (FIN) {
        const double S = final_sum(slot, 0, lds4);
        if (threadIdx.x == 0) finalize_scalar(st, fin, S);
    } else {
        double acc = 0.0;
        for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (size_t)gridDim.x * blockDim.x)
            acc += (double)a[g] * (double)b[g];
        store_partial(slot, 0, acc, lds4);
    }
}

int fused_dot(const double *a, const double *b, size_t n, SolverState *st, int fin, const RedSlot &slot, int residual,
              hipStream_t s)
{
    if (fin == kFinNone || !st || !a) return SELF_OK;
    SELF_RED_LAUNCH((k_op_dot<double, false>), (k_op_dot<double, true>), s, a, b, n, st, fin, slot, residual);
    SELF_HIP(hipGetLastError());
    return SELF_OK;
}

}  // namespace

struct DenseOp final : DevOp {
    int nn = 0;
    double *A = nullptr, *b = nullptr, *dinv = nullptr;
    ~DenseOp() override
    {
        if (A) (void)hipFree(A);
        if (b) (void)hipFree(b);
        if (dinv) (void)hipFree(dinv);
    }
    const double *inv_diag() const override { return dinv; }
    int matvec(const double *x, double *y, const double *dotWith, SolverState *st, int fin, const RedSlot &slot,
               hipStream_t s) override
    {
        hipLaunchKernelGGL(k_dense_matvec, dim3((nn + 255) / 256), dim3(256), 0, s, A, x, y, b, nn, 0, st);
        SELF_HIP(hipGetLastError());
        return fused_dot(dotWith, y, n, st, fin, slot, 0, s);
    }
    int residual(const double *x, double *r, SolverState *st, int fin, const RedSlot &slot, hipStream_t s) override
    {
        hipLaunchKernelGGL(k_dense_matvec, dim3((nn + 255) / 256), dim3(256), 0, s, A, x, r, b, nn, 1, st);
        SELF_HIP(hipGetLastError());
        return fused_dot(r, r, n, st, fin, slot, 1, s);
    }
};

struct Elliptic1DOp final : DevOp {
    int N = 0, nEl = 0;
    double mL = 0.0, mR = 0.0, bfL = 0.0, bfR = 0.0;
    double *d2 = nullptr, *sJw = nullptr;
    std::vector<double> x;  // node positions
    ~Elliptic1DOp() override
    {
        if (d2) (void)hipFree(d2);
        if (sJw) (void)hipFree(sJw);
    }
    int launch(const double *in, double *out, int residual, SolverState *st, hipStream_t s)
    {
        const size_t sh = sizeof(double) * 2 * n;
        hipLaunchKernelGGL(k_e1d_apply, dim3(1), dim3(256), sh, s, in, out, d2, sJw, N, nEl, mL, mR, bfL, bfR,
                           residual, st);
        SELF_HIP(hipGetLastError());
        return SELF_OK;
    }
    int matvec(const double *xin, double *y, const double *dotWith, SolverState *st, int fin, const RedSlot &slot,
               hipStream_t s) override
    {
        int rc = launch(xin, y, 0, st, s);
        return rc ? rc : fused_dot(dotWith, y, n, st, fin, slot, 0, s);
    }
    int residual(const double *xin, double *r, SolverState *st, int fin, const RedSlot &slot, hipStream_t s) override
    {
        int rc = launch(xin, r, 1, st, s);
        return rc ? rc : fused_dot(r, r, n, st, fin, slot, 1, s);
    }
};

}  // namespace self

using namespace self;

struct self_linop_s {
    DevOp *op = nullptr;
    Elliptic1DOp *e1d = nullptr;
    KrylovWork wk;
    hipStream_t stream = nullptr;
    int device = 0;
    ~self_linop_s()
    {
        wk.release();
        delete op;
        if (stream) (void)hipStreamDestroy(stream);
    }
};

extern "C" {

int self_linop_dense_create(int n, const double *A, const double *b, self_linop_t *out)
try {
    SELF_REQUIRE(n >= 1 && A && b && out, SELF_ERR_ARG, "self_linop_dense_create: bad arguments");
    auto *h = new self_linop_s();
    auto *op = new DenseOp();
    h->op = op;
    op->n = (size_t)n;
    op->nn = n;
    std::vector<double> dinv(n);
    for (int i = 0; i < n; ++i) dinv[i] = 1.0 / A[(size_t)i + (size_t)n * i];
    hipError_t e = hipGetDevice(&h->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&op->A, sizeof(double) * n * n);
    if (e == hipSuccess) e = hipMalloc(&op->b, sizeof(double) * n);
    if (e == hipSuccess) e = hipMalloc(&op->dinv, sizeof(double) * n);
    if (e == hipSuccess) e = hipMemcpy(op->A, A, sizeof(double) * n * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(op->b, b, sizeof(double) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(op->dinv, dinv.data(), sizeof(double) * n, hipMemcpyHostToDevice);
    if (e != hipSuccess) { delete h; return hip_fail(e, "dense operator", __FILE__, __LINE__); }
    *out = h;
    return SELF_OK;
}
SELF_CATCH_ALL

// Build_Elliptic1D on the default segment mesh of [0, xScale] with k = 1,
// source = 1 and homogeneous Dirichlet ends (dirichletMask = 0).
int self_linop_elliptic1d_create(int N, int nElems, double xScale, self_linop_t *out)
try {
    SELF_REQUIRE(N >= 1 && N <= 32 && nElems >= 1 && out, SELF_ERR_ARG, "self_linop_elliptic1d_create: bad arguments");
    // one workgroup holds the operand and the result in LDS (2 x 8 B per DOF)
    SELF_REQUIRE((N + 1) * nElems <= 4096, SELF_ERR_ARG, "Elliptic1D operator limited to 4096 DOF");
    const int n1 = N + 1, nd = n1 * nElems;
    std::vector<double> s(n1), w(n1), bw(n1), D(n1 * n1);
    int rc = self_legendre_quadrature(N, SELF_GAUSS_LOBATTO, s.data(), w.data());
    if (!rc) rc = self_barycentric_weights(N, s.data(), bw.data());
    if (!rc) rc = self_derivative_matrix(N, s.data(), bw.data(), D.data());
    if (rc) return rc;
    auto *h = new self_linop_s();
    auto *op = new Elliptic1DOp();
    h->op = op;
    h->e1d = op;
    op->N = N;
    op->nEl = nElems;
    op->n = (size_t)nd;
    op->x.resize(nd);
    std::vector<double> J(nd), d2((size_t)n1 * n1 * nElems), sJw(nd);
    const double dx = 1.0 / (double)nElems;  // SegmentMeshClass.f90:406 (unit interval)
    for (int e = 0; e < nElems; ++e) {
        const double x1 = (double)e * dx, x2 = x1 + dx;
        std::vector<double> xe(n1);
        for (int i = 0; i < n1; ++i) xe[i] = 0.5 * (x2 - x1) * (s[i] + 1.0) + x1;
        for (int i = 0; i < n1; ++i) {
            // J = dx/ds at the node by the barycentric derivative formula at a node
            double num = 0.0;
            for (int j = 0; j < n1; ++j)
                if (j != i) num += bw[j] * (xe[i] - xe[j]) / (s[i] - s[j]);
            J[e * n1 + i] = num / (-bw[i]);
            op->x[e * n1 + i] = xe[i] * xScale;
            sJw[e * n1 + i] = 1.0 * J[e * n1 + i] * w[i];  // source * J * w
        }
        for (int col = 0; col <= N; ++col)
            for (int row = 0; row <= N; ++row) {
                double acc = 0.0;
                for (int m = 0; m <= N; ++m)
                    acc = acc - (1.0 * w[m] / J[e * n1 + m]) * D[m + n1 * row] * D[m + n1 * col];
                d2[row + n1 * (col + (size_t)n1 * e)] = acc;
            }
    }
    hipError_t e = hipGetDevice(&h->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&op->d2, sizeof(double) * d2.size());
    if (e == hipSuccess) e = hipMalloc(&op->sJw, sizeof(double) * nd);
    if (e == hipSuccess) e = hipMemcpy(op->d2, d2.data(), sizeof(double) * d2.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(op->sJw, sJw.data(), sizeof(double) * nd, hipMemcpyHostToDevice);
    if (e != hipSuccess) { delete h; return hip_fail(e, "elliptic1d operator", __FILE__, __LINE__); }
    *out = h;
    return SELF_OK;
}
SELF_CATCH_ALL

int self_linop_destroy(self_linop_t h)
{
    if (!h) return SELF_OK;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    delete h;
    return SELF_OK;
}

int self_linop_size(self_linop_t h, int *n)
{
    SELF_REQUIRE(h && n, SELF_ERR_ARG, "null argument");
    *n = (int)h->op->n;
    return SELF_OK;
}

// UnMaskSolution of the 1-D operator (Elliptic1D_Class.f90 UnMask + DirichletConditions)
int self_linop_elliptic1d_unmask(self_linop_t h, const double *sol, double *u, double *x)
{
    SELF_REQUIRE(h && h->e1d && sol && u, SELF_ERR_ARG, "not an Elliptic1D operator");
    const Elliptic1DOp &op = *h->e1d;
    const int n1 = op.N + 1;
    std::memcpy(u, sol, sizeof(double) * op.n);
    for (int e = 1; e < op.nEl; ++e) u[n1 * e] = u[op.N + n1 * (e - 1)];
    if (op.mL == 0.0) u[0] = 0.0;
    if (op.mR == 0.0) u[op.N + n1 * (op.nEl - 1)] = 0.0;
    if (x) std::memcpy(x, op.x.data(), sizeof(double) * op.n);
    return SELF_OK;
}

int self_krylov_solve(self_linop_t h, int method, double tol, int maxIters, int mInner, int ortho, double *sol,
                      double *resi, int resiLen, int *iterations, int *ioerr)
try {
    SELF_REQUIRE(h && sol && iterations && ioerr, SELF_ERR_ARG, "self_krylov_solve: null argument");
    SELF_REQUIRE(method >= 0 && method <= 3, SELF_ERR_ARG, "method %d (CG 0, PCG 1, GMRES 2, PCGMRES 3)", method);
    SELF_REQUIRE(maxIters >= 1 && tol > 0.0, SELF_ERR_ARG, "maxIters/tolerance");
    SELF_REQUIRE(method < 2 || (mInner >= 1 && mInner <= 200), SELF_ERR_ARG, "GMRES m = %d", mInner);
    SELF_HIP(hipSetDevice(h->device));
    const size_t n = h->op->n;
    double *d = nullptr;
    SELF_HIP(hipMalloc(&d, sizeof(double) * n));
    hipError_t e = hipMemcpy(d, sol, sizeof(double) * n, hipMemcpyHostToDevice);
    if (e != hipSuccess) { (void)hipFree(d); return hip_fail(e, "copy", __FILE__, __LINE__); }
    KrylovResult res;
    int rc = krylov_solve(*h->op, h->wk, method, method == 1 || method == 3, tol, maxIters, mInner, ortho, d, resi,
                          resiLen, res, h->stream);
    if (!rc) {
        e = hipStreamSynchronize(h->stream);
        if (e == hipSuccess) e = hipMemcpy(sol, d, sizeof(double) * n, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = hip_fail(e, "copy back", __FILE__, __LINE__);
    }
    (void)hipFree(d);
    if (rc) return rc;
    *iterations = res.iterations;
    *ioerr = res.ioerr;
    return SELF_OK;
}
SELF_CATCH_ALL

}  // extern "C"
