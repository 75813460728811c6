// krylov.hip -- device-resident Krylov solvers with the reference's control
// flow and stop tests:
//   Solve_ConjugateGradient   src/iterativesolve/IterativeSolvers_Class.f90:328-405
//   Solve_PCConjugateGradient                                            :409-489
//   Solve_GMRES (MGS, progressive Givens, back substitution)             :493-652
//   Solve_PCGMRES (flexible, right preconditioned)                       :656-822
//   CGsemElliptic_2D PCG       src/cgsem/CGsemElliptic_2D_Class.f90:1114-1207
// Every vector and scalar lives on the device.  Dot products are fused into
// the vector kernels that produce their operands and reduced
// deterministically (reduce.hpp); the scalar recurrences (alpha, beta,
// Givens rotations, the stop tests) run in the one-workgroup final launch of
// a reduction or in a one-thread kernel.  A CG iteration ends with one sweep
// that also starts the next (preconditioner and numerator); for operators
// that gather their input the direction update is folded into the gather
// and each final sum into the kernel that consumes it, so an iteration is
// two launches of the operator and one sweep.  A `done` flag in the device state
// turns the kernels of iterations after convergence into no-ops, so the
// host launches iterations in chunks and reads the flag between chunks.

#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "krylov.hpp"

namespace self {

namespace {

__device__ __forceinline__ bool st_done(const SolverState *st) { return st->done != 0; }

// FIN = false: store this workgroup's partial sums; FIN = true (one
// workgroup, after the partial launch): sum them and update the state
template <int K, bool FIN>
__device__ __forceinline__ void reduce_rows(const RedSlot &slot, double (&acc)[K], SolverState *st,
                                            const int (&fin)[K])
{
    __shared__ double lds4[4];
    #pragma unroll
    for (int k = 0; k < K; ++k) {
        if constexpr (FIN) {
            const double S = final_sum(slot, k, lds4);
            if (threadIdx.x == 0) finalize_scalar(st, fin[k], S);
        } else {
            store_partial(slot, k, acc[k], lds4);
        }
    }
}

// a.b with finalize
template <typename T, bool FIN>
__global__ __launch_bounds__(kRedThreads) void k_dot(const T *__restrict__ a, const T *__restrict__ b, size_t n,
                                                     SolverState *st, int fin, RedSlot slot)
{
    if (st_done(st)) return;
    double acc[1] = {0.0};
    if constexpr (!FIN)
        for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (size_t)gridDim.x * blockDim.x)
            acc[0] += (double)a[g] * (double)b[g];
    const int f[1] = {fin};
    reduce_rows<1, FIN>(slot, acc, st, f);
}

// w = P r (Jacobi or identity), num = r.w
template <typename T, bool FIN>
__global__ __launch_bounds__(kRedThreads) void k_precond_dot(const T *__restrict__ r, const T *__restrict__ dinv,
                                                             T *__restrict__ w, size_t n, SolverState *st, int fin,
                                                             RedSlot slot)
{
    if (st_done(st)) return;
    double acc[1] = {0.0};
    if constexpr (!FIN)
        for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (size_t)gridDim.x * blockDim.x) {
            const T rv = r[g];
            const T wv = dinv ? dinv[g] * rv : rv;
            w[g] = wv;
            acc[0] += (double)rv * (double)wv;
        }
    const int f[1] = {fin};
    reduce_rows<1, FIN>(slot, acc, st, f);
}

// v = w (first iteration) or w + beta v
template <typename T>
__global__ __launch_bounds__(256) void k_update_v(const T *__restrict__ w, T *__restrict__ v, size_t n,
                                                  const SolverState *st)
{
    if (st_done(st)) return;
    const bool first = st->iter == 1;
    const T beta = (T)st->beta;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (size_t)gridDim.x * blockDim.x)
        v[g] = first ? w[g] : w[g] + beta * v[g];
}

// The end of one CG iteration and the start of the next in one sweep:
// u += alpha v; r -= alpha z; |r|^2 -> rNorm and the stop test (row 0);
// w = P r and num = r.w, beta = num/den for the next iteration (row 1).
// Without a preconditioner w is r itself and row 1 repeats row 0.  Each
// thread visits the same indices in the same order as the separate
// precondition/dot and update kernels did, so the sums are the same.
template <typename T, bool FIN, bool PRE>
__global__ __launch_bounds__(kRedThreads) void k_update_urw(T *__restrict__ u, const T *__restrict__ v,
                                                            T *__restrict__ r, const T *__restrict__ z,
                                                            const T *__restrict__ dinv, T *__restrict__ w, size_t n,
                                                            SolverState *st, RedSlot slot)
{
    if (st_done(st)) return;
    double acc[2] = {0.0, 0.0};
    if constexpr (!FIN) {
        const T alpha = (T)st->alpha;
        for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (size_t)gridDim.x * blockDim.x) {
            u[g] = u[g] + alpha * v[g];
            const T rv = r[g] - alpha * z[g];
            r[g] = rv;
            acc[0] += (double)rv * (double)rv;
            if constexpr (PRE) {
                const T wv = dinv[g] * rv;
                w[g] = wv;
                acc[1] += (double)rv * (double)wv;
            }
        }
        if constexpr (!PRE) acc[1] = acc[0];
    }
    const int f[2] = {kFinRnorm, kFinNumNext};
    reduce_rows<2, FIN>(slot, acc, st, f);
}

// end of the last iteration of a chunk (one wavefront)
__global__ __launch_bounds__(64) void k_cg_fold_end(SolverState *st, RedSlot slot, int it)
{
    cg_fold_end_wave(st, slot, it);
}

// ---------------------------------------------------------------- GMRES
#define GH(st, a, b) (st)->H[((a)-1) + ((st)->m + 1) * ((b)-1)]
#define GRHO(st, a, b) (st)->rho[((a)-1) + (st)->m * ((b)-1)]

__device__ __forceinline__ bool almost_equal_dev(double a, double b)
{
    const double d = fabs(a - b), eps = 2.220446049250313e-16;
    if (a == 0.0 || b == 0.0) return d <= eps;
    return d <= eps * fabs(a) || d <= eps * fabs(b);
}

__device__ __forceinline__ bool inner_off(const SolverState *st) { return st->done != 0 || st->innerExit != 0; }

// b = |r|, v1 = r / b happens in k_gmres_scale after this finalize
__global__ void k_gmres_cycle_start(SolverState *st)
{
    if (threadIdx.x != 0 || st->done) return;
    st->innerExit = 0;
    st->b = st->bnorm;
    st->bhat[0] = st->b;
}

template <typename T>
__global__ __launch_bounds__(256) void k_scale(const T *__restrict__ src, T *__restrict__ dst, size_t n,
                                               const SolverState *st, int which, int col)
{
    if (inner_off(st)) return;
    // which 0: dst = src / b ; which 1: dst = src / H(col+1, col)
    const double d = which == 0 ? st->b : GH(st, col + 1, col);
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (size_t)gridDim.x * blockDim.x)
        dst[g] = (T)((double)src[g] / d);
}

__global__ void k_gmres_begin(SolverState *st, int i)
{
    if (threadIdx.x != 0 || inner_off(st)) return;
    st->l += 1;
    st->nr = i;
}

// H(k,i) = v_k . w
template <bool FIN>
__global__ __launch_bounds__(kRedThreads) void k_gmres_h(const double *__restrict__ vk, const double *__restrict__ w,
                                                         size_t n, SolverState *st, int k, int i, RedSlot slot)
{
    if (inner_off(st)) return;
    __shared__ double lds4[4];
    if constexpr (!FIN) {
        double acc = 0.0;
        for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (size_t)gridDim.x * blockDim.x)
            acc += vk[g] * w[g];
        store_partial(slot, 0, acc, lds4);
    } else {
        const double S = final_sum(slot, 0, lds4);
        if (threadIdx.x == 0) {
            if (k <= i) GH(st, k, i) = S;
            else GH(st, i + 1, i) = sqrt(S);  // k = i+1: the norm of w
        }
    }
}

// w -= H(k,i) v_k
__global__ __launch_bounds__(256) void k_gmres_axpy(double *__restrict__ w, const double *__restrict__ vk, size_t n,
                                                    const SolverState *st, int k, int i)
{
    if (inner_off(st)) return;
    const double h = GH(st, k, i);
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (size_t)gridDim.x * blockDim.x)
        w[g] = w[g] - h * vk[g];
}

// CGS2 pass: h_k = v_k . w for k = 1..i in one sweep (partials per k)
template <bool FIN>
__global__ __launch_bounds__(kRedThreads) void k_gmres_multidot(const double *__restrict__ V, const double *__restrict__ w,
                                                                size_t n, int i, SolverState *st, double *hOut,
                                                                RedSlot slot)
{
    if (inner_off(st)) return;
    __shared__ double lds4[4];
    for (int k = 0; k < i; ++k) {
        if constexpr (!FIN) {
            double acc = 0.0;
            const double *vk = V + (size_t)k * n;
            for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (size_t)gridDim.x * blockDim.x)
                acc += vk[g] * w[g];
            store_partial(slot, k, acc, lds4);
        } else {
            const double S = final_sum(slot, k, lds4);
            if (threadIdx.x == 0) hOut[k] = S;
        }
    }
}

// w -= V h  (CGS2), and accumulate h into H(:,i) when acc != 0
__global__ __launch_bounds__(256) void k_gmres_multiaxpy(double *__restrict__ w, const double *__restrict__ V,
                                                         size_t n, int i, SolverState *st, const double *h,
                                                         int accumulate)
{
    if (inner_off(st)) return;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (size_t)gridDim.x * blockDim.x) {
        double acc = w[g];
        for (int k = 0; k < i; ++k) acc -= h[k] * V[(size_t)k * n + g];
        w[g] = acc;
    }
    if (accumulate && blockIdx.x == 0 && threadIdx.x < i) {
        const int k = threadIdx.x;
        GH(st, k + 1, i) = (accumulate == 1) ? h[k] : GH(st, k + 1, i) + h[k];
    }
}

// breakdown test, progressive Givens rotation and the inner stop test
__global__ void k_gmres_givens(SolverState *st, int i)
{
    if (threadIdx.x != 0 || inner_off(st)) return;
    if (almost_equal_dev(GH(st, i + 1, i), 0.0)) {  // :568-570
        st->innerExit = 2;
        return;
    }
    GRHO(st, 1, i) = GH(st, 1, i);
    for (int k = 2; k <= i; ++k) {
        const double g = st->c[k - 2] * GRHO(st, k - 1, i) + st->s[k - 2] * GH(st, k, i);
        GRHO(st, k, i) = -st->s[k - 2] * GRHO(st, k - 1, i) + st->c[k - 2] * GH(st, k, i);
        GRHO(st, k - 1, i) = g;
    }
    const double d = sqrt(GRHO(st, i, i) * GRHO(st, i, i) + GH(st, i + 1, i) * GH(st, i + 1, i));
    st->c[i - 1] = GRHO(st, i, i) / d;
    st->s[i - 1] = GH(st, i + 1, i) / d;
    GRHO(st, i, i) = st->c[i - 1] * GRHO(st, i, i) + st->s[i - 1] * GH(st, i + 1, i);
    st->bhat[i] = -st->s[i - 1] * st->bhat[i - 1];
    st->bhat[i - 1] = st->c[i - 1] * st->bhat[i - 1];
    st->rc = fabs(st->bhat[i]);
    if (st->l < st->resiLen) st->resi[st->l] = st->rc;
    if (st->rc / st->r0 <= st->tol) st->innerExit = 1;
}

// v_{i+1} = w / H(i+1,i)
__global__ __launch_bounds__(256) void k_gmres_newv(const double *__restrict__ w, double *__restrict__ vnext, size_t n,
                                                    const SolverState *st, int i)
{
    if (st->done || st->innerExit == 2) return;  // breakdown exits before this (:568-572)
    if (st->innerExit == 1) return;                // converged: v_{i+1} is never used
    const double d = GH(st, i + 1, i);
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (size_t)gridDim.x * blockDim.x)
        vnext[g] = w[g] / d;
}

// back substitution (:606-626)
__global__ void k_gmres_backsub(SolverState *st)
{
    if (threadIdx.x != 0 || st->done) return;
    if (st->rc / st->r0 > st->tol) st->nr = st->m;
    const int nr = st->nr;
    st->y[nr - 1] = st->bhat[nr - 1] / GRHO(st, nr, nr);
    for (int k = nr - 1; k >= 1; --k) {
        double yk = st->bhat[k - 1];
        for (int i = k + 1; i <= nr; ++i) yk = yk - GRHO(st, k, i) * st->y[i - 1];
        st->y[k - 1] = yk / GRHO(st, k, k);
    }
}

// sol += basis(:,1:nr) y ; sets the convergence flag afterwards
__global__ __launch_bounds__(256) void k_gmres_update(double *__restrict__ sol, const double *__restrict__ basis,
                                                      size_t n, SolverState *st)
{
    if (st->done) return;
    const int nr = st->nr;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (size_t)gridDim.x * blockDim.x) {
        double acc = 0.0;
        for (int k = 0; k < nr; ++k) acc += basis[(size_t)k * n + g] * st->y[k];
        sol[g] = sol[g] + acc;
    }
}

__global__ void k_gmres_cycle_end(SolverState *st, int j, int nIt)
{
    if (threadIdx.x != 0 || st->done) return;
    if (st->rc / st->r0 <= st->tol) { st->ioerr = 0; st->done = 1; return; }
    if (j >= nIt) st->done = 2;
}

// Mixed form: the cycle's stop test on the FP64 true residual |b - A x|
// (bnorm, just computed), not on the Givens estimate, which describes the
// FP32 operator the cycle applied.
__global__ void k_gmres_cycle_end_true(SolverState *st, int j, int nIt)
{
    if (threadIdx.x != 0 || st->done) return;
    st->rc = st->bnorm;
    if (st->rc / st->r0 <= st->tol) { st->ioerr = 0; st->done = 1; return; }
    if (j >= nIt) st->done = 2;
}

template <typename T>
__global__ __launch_bounds__(256) void k_jacobi(const T *__restrict__ v, const T *__restrict__ dinv, T *__restrict__ z,
                                                size_t n, const SolverState *st)
{
    if (inner_off(st)) return;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (size_t)gridDim.x * blockDim.x)
        z[g] = dinv ? dinv[g] * v[g] : v[g];
}

template <typename A, typename B>
__global__ __launch_bounds__(256) void k_convert(const A *__restrict__ a, B *__restrict__ b, size_t n, double scale)
{
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (size_t)gridDim.x * blockDim.x)
        b[g] = (B)((double)a[g] * scale);
}

template <typename T>
__global__ __launch_bounds__(256) void k_add_scaled(double *__restrict__ x, const T *__restrict__ e, size_t n,
                                                    double scale)
{
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (size_t)gridDim.x * blockDim.x)
        x[g] = x[g] + scale * (double)e[g];
}

template <typename T>
__global__ __launch_bounds__(256) void k_fill(T *__restrict__ a, size_t n, T v)
{
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (size_t)gridDim.x * blockDim.x) a[g] = v;
}

// reliable updates (krylov_solve_reliable): let the FP64 residual kernels run
__global__ void k_reliable_begin(SolverState *st)
{
    if (threadIdx.x == 0) st->done = 0;
}

// after an update: stop (4 converged, 5 out of updates or iterations) or
// continue the FP32 sequence at the next iteration, its stop test relative
// to the new residual
__global__ void k_reliable_next(SolverState *st, double r0, double scale, double tol, double delta, int maxUpdates)
{
    if (threadIdx.x != 0) return;
    const double rn = st->bnorm;
    st->l += 1;
    st->rc = rn;
    if (rn <= tol * r0) { st->done = 4; return; }
    if (st->l >= maxUpdates || st->iter >= st->maxIters) { st->done = 5; return; }
    st->iter += 1;
    st->r0 = rn / scale;
    st->tol = fmax(delta, tol * r0 / rn);
}

inline int vgrid(size_t n) { return (int)std::min<size_t>(4096, (n + 255) / 256 > 0 ? (n + 255) / 256 : 1); }

}  // namespace

// ------------------------------------------------------------- workspace
int KrylovWork::ensure(size_t nn, int mm, int rl, bool f32)
{
    if (nn == n && mm <= m && rl <= resiLen && (!f32 || r32) && st) return SELF_OK;
    release();
    n = nn;
    m = mm;
    resiLen = rl;
    auto A = [&](void **p, size_t bytesReq) -> int {
        if (bytesReq == 0) { *p = nullptr; return SELF_OK; }
        SELF_HIP(hipMalloc(p, bytesReq));
        bytes += bytesReq;
        return SELF_OK;
    };
    int rc = SELF_OK;
    const size_t vb = sizeof(double) * n;
    if ((rc = A((void **)&r, vb)) || (rc = A((void **)&v, vb)) || (rc = A((void **)&z, vb)) ||
        (rc = A((void **)&w, vb)) || (rc = A((void **)&u, vb)))
        return rc;
    if ((rc = A((void **)&vB, vb))) return rc;
    if (f32) {
        const size_t fb = sizeof(float) * n;
        if ((rc = A((void **)&r32, fb)) || (rc = A((void **)&v32, fb)) || (rc = A((void **)&z32, fb)) ||
            (rc = A((void **)&w32, fb)) || (rc = A((void **)&u32, fb)) || (rc = A((void **)&vB32, fb)))
            return rc;
    }
    if (m > 0) {
        if ((rc = A((void **)&V, vb * (m + 1))) || (rc = A((void **)&Z, vb * (m + 1)))) return rc;
    }
    const int mm1 = std::max(m, 1);
    const size_t smallN = (size_t)(mm1 + 1) * mm1 + (size_t)mm1 * mm1 + 2 * mm1 + 2 * (mm1 + 1) + 64;
    if ((rc = A((void **)&small, sizeof(double) * smallN))) return rc;
    if ((rc = A((void **)&st, sizeof(SolverState)))) return rc;
    if ((rc = A((void **)&resi, sizeof(double) * std::max(resiLen, 1)))) return rc;
    if ((rc = A((void **)&slot.partials, sizeof(double) * kRedBlocks * std::max(32, m + 2)))) return rc;

    SELF_HIP(hipMemset(small, 0, sizeof(double) * smallN));
    return SELF_OK;
}

void KrylovWork::release()
{
    void *p[] = {r, v, z, w, u, vB, r32, v32, z32, w32, u32, vB32, V, Z, small, st, resi, slot.partials};
    for (void *q : p)
        if (q) (void)hipFree(q);
    r = v = z = w = u = vB = nullptr;
    r32 = v32 = z32 = w32 = u32 = vB32 = nullptr;
    V = Z = small = nullptr;
    st = nullptr;
    resi = nullptr;
    slot = RedSlot{};
    n = 0;
    m = 0;
    resiLen = 0;
    bytes = 0;
}

namespace {

int init_state(KrylovWork &wk, double tol, int maxIters, int stopRule, int m, hipStream_t s)
{
    SolverState h{};
    h.tol = tol;
    h.iter = 1;
    h.maxIters = maxIters;
    h.stopRule = stopRule;
    h.ioerr = -2;
    h.m = m;
    h.resi = wk.resi;
    h.resiLen = wk.resiLen;
    const int mm1 = std::max(m, 1);
    double *p = wk.small;
    h.H = p; p += (size_t)(mm1 + 1) * mm1;
    h.rho = p; p += (size_t)mm1 * mm1;
    h.c = p; p += mm1;
    h.s = p; p += mm1;
    h.bhat = p; p += mm1 + 1;
    h.y = p;
    SELF_HIP(hipMemcpyAsync(wk.st, &h, sizeof(SolverState), hipMemcpyHostToDevice, s));
    SELF_HIP(hipMemsetAsync(wk.resi, 0, sizeof(double) * wk.resiLen, s));
    return SELF_OK;
}

int read_state(KrylovWork &wk, SolverState &h, hipStream_t s)
{
    SELF_HIP(hipMemcpyAsync(&h, wk.st, sizeof(SolverState), hipMemcpyDeviceToHost, s));
    SELF_HIP(hipStreamSynchronize(s));
    return SELF_OK;
}

constexpr int kChunk = 32;  // iterations launched between flag reads

inline int op_matvec(DevOp &op, const double *x, double *y, SolverState *st, const RedSlot &slot, hipStream_t s)
{
    return op.matvec(x, y, x, st, kFinAlpha, slot, s);
}
inline int op_matvec(DevOp &op, const float *x, float *y, SolverState *st, const RedSlot &slot, hipStream_t s)
{
    return op.matvec32(x, y, x, st, kFinAlpha, slot, s);
}
inline int op_cg_iteration(DevOp &op, const double *dinv, double *u, double *r, double *w, const double *vOld,
                           double *vNew, SolverState *st, const RedSlot &slot, int it, bool pending, hipStream_t s)
{
    return op.cg_iteration(dinv, u, r, w, vOld, vNew, st, slot, it, pending, s);
}
inline int op_cg_iteration(DevOp &op, const float *dinv, float *u, float *r, float *w, const float *vOld,
                           float *vNew, SolverState *st, const RedSlot &slot, int it, bool pending, hipStream_t s)
{
    return op.cg_iteration32(dinv, u, r, w, vOld, vNew, st, slot, it, pending, s);
}

// CG iterations it0..it1 (the loop body of the reference's CG/PCG):
//   v = w + beta v, z = A v, alpha = num / v.z
//   u += alpha v, r -= alpha z, rNorm, stop test, w = P r, num, beta
// Operators that gather their input run each iteration folded
// (DevOp::cg_iteration, no final launches): the bookkeeping of an
// iteration's end is done by the next iteration's first kernel, and by
// k_cg_fold_end after the chunk's last; v alternates between vA (odd
// iterations) and vB.  Other operators: in-place direction update,
// operator, update sweep, each reduction with its final launch.
template <typename T>
int cg_chunk(DevOp &op, T *u, T *r, T *w, T *vA, T *vB, T *z, const T *dinv, size_t n, SolverState *st,
             const RedSlot &slot, int it0, int it1, hipStream_t s)
{
    if (op.fuses_direction()) {
        for (int it = it0; it <= it1; ++it) {
            const T *vOld = (it & 1) ? vB : vA;
            T *v = (it & 1) ? vA : vB;
            int rc = op_cg_iteration(op, dinv, u, r, w, vOld, v, st, slot, it, it > it0, s);
            if (rc) return rc;
        }
        hipLaunchKernelGGL(k_cg_fold_end, dim3(1), dim3(64), 0, s, st, slot, it1);
        SELF_HIP(hipGetLastError());
        return SELF_OK;
    }
    const T *wsrc = dinv ? w : r;
    const int gv = vgrid(n);
    for (int it = it0; it <= it1; ++it) {
        hipLaunchKernelGGL(k_update_v<T>, dim3(gv), dim3(256), 0, s, wsrc, vA, n, st);
        int rc = op_matvec(op, vA, z, st, slot, s);
        if (rc) return rc;
        if (dinv)
            SELF_RED_LAUNCH((k_update_urw<T, false, true>), (k_update_urw<T, true, true>), s, u, (const T *)vA, r,
                               (const T *)z, dinv, w, n, st, slot);
        else
            SELF_RED_LAUNCH((k_update_urw<T, false, false>), (k_update_urw<T, true, false>), s, u, (const T *)vA, r,
                               (const T *)z, dinv, w, n, st, slot);
    }
    SELF_HIP(hipGetLastError());
    return SELF_OK;
}

}  // namespace

int krylov_cg_begin(DevOp &op, KrylovWork &wk, bool precond, int stopRule, double tol, int maxIters,
                    const double *sol, hipStream_t s)
{
    const size_t n = op.n;
    int rc = init_state(wk, tol, maxIters, stopRule, 0, s);
    if (rc) return rc;
    const int fin0 = stopRule == kStopNormRel ? kFinR0 : kFinR0Plain;
    if ((rc = op.residual(sol, wk.r, wk.st, fin0, wk.slot, s))) return rc;
    SELF_HIP(hipMemcpyAsync(wk.u, sol, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    const double *dinv = precond ? op.inv_diag() : nullptr;
    // iteration 1's w = P r and num = r.w; later iterations get theirs from
    // the update sweep of the iteration before
    SELF_RED_LAUNCH((k_precond_dot<double, false>), (k_precond_dot<double, true>), s, wk.r, dinv, wk.w, n, wk.st,
                       (int)kFinNum, wk.slot);
    SELF_HIP(hipGetLastError());
    return SELF_OK;
}

int krylov_cg_finish(DevOp &op, KrylovWork &wk, int stopRule, double tol, double *sol, KrylovResult &res,
                     hipStream_t s)
{
    SolverState h{};
    int rc = read_state(wk, h, s);
    if (rc) return rc;
    SELF_HIP(hipMemcpyAsync(sol, wk.u, sizeof(double) * op.n, hipMemcpyDeviceToDevice, s));
    res.r0 = h.r0;
    res.rNorm = h.rNorm;
    res.iterations = (h.done == 3) ? 0 : h.iter;
    res.ioerr = -2;
    if (stopRule == kStopSqrtNormRel) {
        if (sqrt(h.rNorm) > tol) res.ioerr = -1;  // :399-403 (absolute test)
    } else if (h.done != 3) {
        if (h.rNorm / h.r0 > tol) res.ioerr = -1;  // CGsemElliptic_2D :1196-1200
    }
    return SELF_OK;
}

namespace {

// CG / PCG of IterativeSolvers (stopRule sqrt-norm) and the CGSEM PCG (norm rule)
int run_cg(DevOp &op, KrylovWork &wk, bool precond, int stopRule, double tol, int maxIters, double *sol,
           KrylovResult &res, hipStream_t s)
{
    int rc = krylov_cg_begin(op, wk, precond, stopRule, tol, maxIters, sol, s);
    if (rc) return rc;
    const double *dinv = precond ? op.inv_diag() : nullptr;
    for (int it0 = 1; it0 <= maxIters; it0 += kChunk) {
        if ((rc = cg_chunk<double>(op, wk.u, wk.r, wk.w, wk.v, wk.vB, wk.z, dinv, op.n, wk.st, wk.slot, it0,
                                   std::min(it0 + kChunk - 1, maxIters), s)))
            return rc;
        SolverState h{};
        if ((rc = read_state(wk, h, s))) return rc;
        if (h.done) break;
    }
    return krylov_cg_finish(op, wk, stopRule, tol, sol, res, s);
}

// CGSEM PCG (CGsemElliptic_2D_Class.f90:1114-1207) with the operator's own
// preconditioner (DevOp::precondition, e.g. CGsemPreconditioner_2D % Solve
// at :1169): w = M^-1 r, num = r.w, v = w + (num/den) v, z = A v with v.z,
// u += a v, r -= a z, rNorm and the stop test.  Unfused: the
// preconditioner is a solve of its own between the sweeps.
int run_cg_pc(DevOp &op, KrylovWork &wk, double tol, int maxIters, double *sol, KrylovResult &res, hipStream_t s)
{
    const size_t n = op.n;
    int rc = init_state(wk, tol, maxIters, kStopNormRel, 0, s);
    if (rc) return rc;
    if ((rc = op.residual(sol, wk.r, wk.st, kFinR0, wk.slot, s))) return rc;
    SELF_HIP(hipMemcpyAsync(wk.u, sol, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    const int gv = vgrid(n);
    constexpr int kChunkPc = 8;  // outer iterations between host reads of the state
    for (int it0 = 1; it0 <= maxIters; it0 += kChunkPc) {
        const int it1 = std::min(it0 + kChunkPc - 1, maxIters);
        for (int it = it0; it <= it1; ++it) {
            if ((rc = op.precondition(wk.r, wk.w, wk.st, s))) return rc;
            SELF_RED_LAUNCH((k_dot<double, false>), (k_dot<double, true>), s, (const double *)wk.r,
                               (const double *)wk.w, n, wk.st, (int)kFinNum, wk.slot);
            hipLaunchKernelGGL(k_update_v<double>, dim3(gv), dim3(256), 0, s, (const double *)wk.w, wk.v, n,
                               (const SolverState *)wk.st);
            if ((rc = op.matvec(wk.v, wk.z, wk.v, wk.st, kFinAlpha, wk.slot, s))) return rc;
            SELF_RED_LAUNCH((k_update_urw<double, false, false>), (k_update_urw<double, true, false>), s, wk.u,
                               (const double *)wk.v, wk.r, (const double *)wk.z, (const double *)nullptr, wk.w, n,
                               wk.st, wk.slot);
        }
        SELF_HIP(hipGetLastError());
        SolverState h{};
        if ((rc = read_state(wk, h, s))) return rc;
        if (h.done) break;
    }
    SolverState h{};
    if ((rc = read_state(wk, h, s))) return rc;
    SELF_HIP(hipMemcpyAsync(sol, wk.u, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    res.r0 = h.r0;
    res.rNorm = h.rNorm;
    res.iterations = (h.done == 3) ? 0 : h.iter;
    res.ioerr = -2;
    if (h.done != 3 && h.rNorm / h.r0 > tol) res.ioerr = -1;  // :1196-1200
    return SELF_OK;
}

// pc: kPcNone, kPcJacobi or kPcOperator (the operator's preconditioner);
// mixed: operator and preconditioner applications in FP32 (the vectors are
// converted on the way in and out), everything else FP64.
int run_gmres(DevOp &op, KrylovWork &wk, int pc, bool mixed, double tol, int nIt, int m, int ortho, double *sol,
              KrylovResult &res, hipStream_t s)
{
    const size_t n = op.n;
    int rc = init_state(wk, tol, nIt, kStopNormRel, m, s);
    if (rc) return rc;
    if ((rc = op.residual(sol, wk.r, wk.st, kFinR0Plain, wk.slot, s))) return rc;
    const bool precond = pc != kPcNone;
    const double *dinv = (pc == kPcJacobi) ? op.inv_diag() : nullptr;
    const int gv = vgrid(n);
    double *hTmp = wk.small + (size_t)(m + 1) * m + (size_t)m * m + 2 * m + 2 * (m + 1);  // scratch for CGS2
    for (int j = 1; j <= nIt; ++j) {
        if (!mixed || j == 1)  // mixed: |r| of later cycles comes with their stop test
            SELF_RED_LAUNCH((k_dot<double, false>), (k_dot<double, true>), s, wk.r, wk.r, n, wk.st,
                               (int)kFinBnorm, wk.slot);
        hipLaunchKernelGGL(k_gmres_cycle_start, dim3(1), dim3(64), 0, s, wk.st);
        hipLaunchKernelGGL(k_scale<double>, dim3(gv), dim3(256), 0, s, wk.r, wk.V, n, wk.st, 0, 0);
        for (int i = 1; i <= m; ++i) {
            double *vi = wk.V + (size_t)(i - 1) * n;
            hipLaunchKernelGGL(k_gmres_begin, dim3(1), dim3(64), 0, s, wk.st, i);
            const double *xin = vi;
            if (precond) {
                double *zi = wk.Z + (size_t)(i - 1) * n;
                if (pc == kPcOperator && mixed) {
                    hipLaunchKernelGGL((k_convert<double, float>), dim3(gv), dim3(256), 0, s, (const double *)vi,
                                       wk.r32, n, 1.0);
                    if ((rc = op.precondition32(wk.r32, wk.z32, wk.st, s))) return rc;
                    hipLaunchKernelGGL((k_convert<float, double>), dim3(gv), dim3(256), 0, s, (const float *)wk.z32,
                                       zi, n, 1.0);
                } else if (pc == kPcOperator) {
                    if ((rc = op.precondition(vi, zi, wk.st, s))) return rc;
                } else {
                    hipLaunchKernelGGL(k_jacobi<double>, dim3(gv), dim3(256), 0, s, vi, dinv, zi, n, wk.st);
                }
                xin = zi;
            }
            // w = A x ; the matvec kernels honour done but not innerExit, so
            // guard: they run and their result is ignored after an exit.
            if (mixed) {
                hipLaunchKernelGGL((k_convert<double, float>), dim3(gv), dim3(256), 0, s, xin, wk.v32, n, 1.0);
                if ((rc = op.matvec32(wk.v32, wk.w32, nullptr, wk.st, kFinNone, wk.slot, s))) return rc;
                hipLaunchKernelGGL((k_convert<float, double>), dim3(gv), dim3(256), 0, s, (const float *)wk.w32,
                                   wk.w, n, 1.0);
            } else {
                if ((rc = op.matvec(xin, wk.w, nullptr, wk.st, kFinNone, wk.slot, s))) return rc;
            }
            if (ortho == 0) {  // modified Gram-Schmidt, as the reference
                for (int k = 1; k <= i; ++k) {
                    SELF_RED_LAUNCH((k_gmres_h<false>), (k_gmres_h<true>), s,
                                       (const double *)(wk.V + (size_t)(k - 1) * n), (const double *)wk.w, n, wk.st,
                                       k, i, wk.slot);
                    hipLaunchKernelGGL(k_gmres_axpy, dim3(gv), dim3(256), 0, s, wk.w,
                                       (const double *)(wk.V + (size_t)(k - 1) * n), n, (const SolverState *)wk.st,
                                       k, i);
                }
            } else {  // classical Gram-Schmidt applied twice: 2 sweeps over V
                for (int pass = 0; pass < 2; ++pass) {
                    SELF_RED_LAUNCH((k_gmres_multidot<false>), (k_gmres_multidot<true>), s,
                                       (const double *)wk.V, (const double *)wk.w, n, i, wk.st, hTmp, wk.slot);
                    hipLaunchKernelGGL(k_gmres_multiaxpy, dim3(gv), dim3(256), 0, s, wk.w, (const double *)wk.V, n,
                                       i, wk.st, (const double *)hTmp, pass == 0 ? 1 : 2);
                }
            }
            SELF_RED_LAUNCH((k_gmres_h<false>), (k_gmres_h<true>), s, (const double *)wk.w,
                               (const double *)wk.w, n, wk.st, i + 1, i, wk.slot);
            hipLaunchKernelGGL(k_gmres_givens, dim3(1), dim3(64), 0, s, wk.st, i);
            if (i < m + 1)
                hipLaunchKernelGGL(k_gmres_newv, dim3(gv), dim3(256), 0, s, (const double *)wk.w,
                                   wk.V + (size_t)i * n, n, (const SolverState *)wk.st, i);
        }
        hipLaunchKernelGGL(k_gmres_backsub, dim3(1), dim3(64), 0, s, wk.st);
        hipLaunchKernelGGL(k_gmres_update, dim3(gv), dim3(256), 0, s, sol, (const double *)(precond ? wk.Z : wk.V), n,
                           wk.st);
        if (mixed) {  // FP64 true residual, its norm and the stop test on it: one refinement step
            if ((rc = op.residual(sol, wk.r, wk.st, kFinNone, wk.slot, s))) return rc;
            SELF_RED_LAUNCH((k_dot<double, false>), (k_dot<double, true>), s, wk.r, wk.r, n, wk.st,
                               (int)kFinBnorm, wk.slot);
            hipLaunchKernelGGL(k_gmres_cycle_end_true, dim3(1), dim3(64), 0, s, wk.st, j, nIt);
        } else {
            hipLaunchKernelGGL(k_gmres_cycle_end, dim3(1), dim3(64), 0, s, wk.st, j, nIt);
        }
        SELF_HIP(hipGetLastError());
        SolverState h{};
        if ((rc = read_state(wk, h, s))) return rc;
        if (h.done) break;
        if (!mixed && (rc = op.residual(sol, wk.r, wk.st, kFinNone, wk.slot, s))) return rc;
    }
    SolverState h{};
    if ((rc = read_state(wk, h, s))) return rc;
    res.r0 = h.r0;
    res.rNorm = h.rc;
    res.iterations = h.l;
    res.ioerr = (h.rc / h.r0 <= tol) ? 0 : -1;
    return SELF_OK;
}

}  // namespace

int krylov_solve(DevOp &op, KrylovWork &wk, int method, int precond, double tol, int maxIters, int m, int ortho,
                 double *sol, double *resiHost, int resiLen, KrylovResult &res, hipStream_t s)
{
    const bool gm = (method == kGMRES || method == kPCGMRES);
    const int rl = gm ? maxIters * m + 1 : maxIters + 1;
    int rc = wk.ensure(op.n, gm ? m : 0, std::max(rl, resiLen), false);
    if (rc) return rc;
    switch (method) {
    case kCG: rc = run_cg(op, wk, false, kStopSqrtNormRel, tol, maxIters, sol, res, s); break;
    case kPCG: rc = run_cg(op, wk, true, kStopSqrtNormRel, tol, maxIters, sol, res, s); break;
    case kCGSEM_PCG:
        if (precond == kPcOperator) {
            if (!op.has_pc()) { set_error("no preconditioner set on this operator"); return SELF_ERR_STATE; }
            rc = run_cg_pc(op, wk, tol, maxIters, sol, res, s);
        } else {
            rc = run_cg(op, wk, precond != 0, kStopNormRel, tol, maxIters, sol, res, s);
        }
        break;
    case kGMRES: rc = run_gmres(op, wk, kPcNone, false, tol, maxIters, m, ortho, sol, res, s); break;
    case kPCGMRES:
        if (precond == kPcOperator && !op.has_pc()) { set_error("no preconditioner set on this operator"); return SELF_ERR_STATE; }
        rc = run_gmres(op, wk, precond == kPcOperator ? kPcOperator : kPcJacobi, false, tol, maxIters, m, ortho, sol,
                       res, s);
        break;
    default: set_error("unknown Krylov method %d", method); return SELF_ERR_ARG;
    }
    if (rc) return rc;
    if (resiHost && resiLen > 0) {
        SELF_HIP(hipMemcpyAsync(resiHost, wk.resi, sizeof(double) * std::min(resiLen, wk.resiLen),
                                hipMemcpyDeviceToHost, s));
        SELF_HIP(hipStreamSynchronize(s));
    }
    return SELF_OK;
}

int krylov_solve_gmres_mixed(DevOp &op, KrylovWork &wk, int precond, double tol, int maxCycles, int m, int ortho,
                             double *sol, KrylovResult &res, hipStream_t s)
{
    if (!op.has_f32()) { set_error("operator has no FP32 form"); return SELF_ERR_STATE; }
    if (precond == kPcOperator && !op.has_pc()) { set_error("no preconditioner set on this operator"); return SELF_ERR_STATE; }
    int rc = wk.ensure(op.n, m, maxCycles * m + 1, true);
    if (rc) return rc;
    return run_gmres(op, wk, precond, true, tol, maxCycles, m, ortho, sol, res, s);
}

// FP32 PCG inner solve of A e = rhs (e0 = 0) to innerTol relative, then
// FP64 refinement x += |r| e.
int krylov_solve_ir(DevOp &op, KrylovWork &wk, int precond, double tol, double innerTol, int maxInner,
                    int maxOuter, double *x, KrylovResult &res, int &innerTotal, hipStream_t s)
{
    if (!op.has_f32()) { set_error("operator has no FP32 form"); return SELF_ERR_STATE; }
    const size_t n = op.n;
    int rc = wk.ensure(n, 0, maxInner + 1, true);
    if (rc) return rc;
    const int gv = vgrid(n);
    innerTotal = 0;
    const float *dinv = precond ? op.inv_diag32() : nullptr;
    double r0 = -1.0;
    for (int outer = 0; outer <= maxOuter; ++outer) {
        // FP64 residual and its norm
        if ((rc = init_state(wk, tol, 1, kStopNormRel, 0, s))) return rc;
        if ((rc = op.residual(x, wk.r, wk.st, kFinBnorm, wk.slot, s))) return rc;
        SolverState h{};
        if ((rc = read_state(wk, h, s))) return rc;
        const double rn = h.bnorm;
        if (r0 < 0) r0 = rn;
        res.r0 = r0;
        res.rNorm = rn;
        if (rn <= tol * r0 || rn == 0.0) { res.ioerr = 0; res.iterations = outer; return SELF_OK; }
        if (outer == maxOuter) break;
        // inner FP32 PCG on A e = r / |r|
        hipLaunchKernelGGL((k_convert<double, float>), dim3(gv), dim3(256), 0, s, wk.r, wk.r32, n, 1.0 / rn);
        hipLaunchKernelGGL(k_fill<float>, dim3(gv), dim3(256), 0, s, wk.u32, n, 0.0f);
        if ((rc = init_state(wk, innerTol, maxInner, kStopNormRel, 0, s))) return rc;
        // r0 of the inner problem is |r/|r|| = 1
        SELF_RED_LAUNCH((k_dot<float, false>), (k_dot<float, true>), s, wk.r32, wk.r32, n, wk.st,
                           (int)kFinR0Plain, wk.slot);
        SELF_RED_LAUNCH((k_precond_dot<float, false>), (k_precond_dot<float, true>), s, wk.r32, dinv, wk.w32, n,
                           wk.st, (int)kFinNum, wk.slot);
        for (int it0 = 1; it0 <= maxInner; it0 += kChunk) {
            if ((rc = cg_chunk<float>(op, wk.u32, wk.r32, wk.w32, wk.v32, wk.vB32, wk.z32, dinv, n, wk.st, wk.slot,
                                      it0, std::min(it0 + kChunk - 1, maxInner), s)))
                return rc;
            if ((rc = read_state(wk, h, s))) return rc;
            if (h.done) break;
        }
        if ((rc = read_state(wk, h, s))) return rc;
        innerTotal += h.iter;
        hipLaunchKernelGGL(k_add_scaled<float>, dim3(gv), dim3(256), 0, s, x, (const float *)wk.u32, n, rn);
        SELF_HIP(hipGetLastError());
    }
    res.ioerr = -1;
    res.iterations = maxOuter;
    return SELF_OK;
}


// Reliable updates (the solver of Clark et al., "Solving lattice QCD
// systems of equations using mixed precision solvers on GPUs", 2010, applied
// to PCG): one PCG sequence on A e = r / |r0| in the arithmetic T that is
// never restarted.  Whenever its recursive residual has dropped by `delta`
// since the last update (or reaches the overall target), the iterate is
// added to the FP64 solution, the true residual is recomputed in FP64 and
// replaces the recursive one, and the iteration continues with the same
// search direction (beta from the new residual).  T = float: every vector
// update and operator application runs in FP32 (the mixed solver); T =
// double: an FP64 PCG whose stop test sees the true residual, not only the
// recursively updated one.  Solution, residual and convergence test are
// FP64.  The update and the decision to continue run on the device
// (k_reliable_next), so an update costs the host no extra wait.  maxIters
// bounds the inner iterations, maxUpdates the refinements.
// State: done 1 = segment reached its target, 2 = maxIters, 4 = converged
// after an update, 5 = maxUpdates reached; l counts updates, rc holds the
// last true residual norm.
template <typename T>
int solve_reliable(DevOp &op, KrylovWork &wk, int precond, double tol, double delta, int maxIters, int maxUpdates,
                   double *x, KrylovResult &res, int &iters, int &updates, hipStream_t s)
{
    constexpr bool F32 = std::is_same<T, float>::value;
    if (F32 && !op.has_f32()) { set_error("operator has no FP32 form"); return SELF_ERR_STATE; }
    if (!(delta > 0.0 && delta < 1.0)) { set_error("reliable-update factor %g outside (0, 1)", delta); return SELF_ERR_ARG; }
    const size_t n = op.n;
    // FP64 inner sequence: its residual needs a buffer of its own beside the
    // true residual wk.r -- wk.z, which operators that gather their input
    // never use, else the first GMRES basis vector
    const bool own = !F32 && !op.fuses_direction();
    int rc = wk.ensure(n, own ? 1 : 0, maxIters + 1, F32);
    if (rc) return rc;
    T *uI, *rI, *wI, *vA, *vB, *zI;
    const T *dinv;
    if constexpr (F32) {
        uI = wk.u32; rI = wk.r32; wI = wk.w32; vA = wk.v32; vB = wk.vB32; zI = wk.z32;
        dinv = precond ? op.inv_diag32() : nullptr;
    } else {
        uI = wk.u; rI = own ? wk.V : wk.z; wI = wk.w; vA = wk.v; vB = wk.vB; zI = wk.z;
        dinv = precond ? op.inv_diag() : nullptr;
    }
    const int gv = vgrid(n);
    iters = 0;
    updates = 0;
    // FP64 residual of the initial guess; it fixes the scaling
    if ((rc = init_state(wk, tol, 1, kStopNormRel, 0, s))) return rc;
    if ((rc = op.residual(x, wk.r, wk.st, kFinBnorm, wk.slot, s))) return rc;
    SolverState h{};
    if ((rc = read_state(wk, h, s))) return rc;
    const double r0 = h.bnorm, scale = r0;
    res.r0 = r0;
    res.rNorm = r0;
    res.ioerr = 0;
    res.iterations = 0;
    if (r0 == 0.0) return SELF_OK;
    hipLaunchKernelGGL((k_convert<double, T>), dim3(gv), dim3(256), 0, s, wk.r, rI, n, 1.0 / scale);
    hipLaunchKernelGGL(k_fill<T>, dim3(gv), dim3(256), 0, s, uI, n, (T)0);
    // inner state: iteration 1, stop test relative to |r / r0| = 1
    if ((rc = init_state(wk, std::max(delta, tol), maxIters, kStopNormRel, 0, s))) return rc;
    SELF_RED_LAUNCH((k_dot<T, false>), (k_dot<T, true>), s, (const T *)rI, (const T *)rI, n, wk.st,
                       (int)kFinR0Plain, wk.slot);
    SELF_RED_LAUNCH((k_precond_dot<T, false>), (k_precond_dot<T, true>), s, (const T *)rI, dinv, wI, n, wk.st,
                       (int)kFinNum, wk.slot);
    // chunks sized from the last segment's length, so that the host sees a
    // segment's end soon after it happens (iterations launched after it
    // are no-ops)
    int it0 = 1, seg0 = 1, chunk = kChunk;
    for (;;) {
        // inner iterations until the state reports a stop
        for (;;) {
            const int it1 = std::min(it0 + chunk - 1, maxIters);
            if ((rc = cg_chunk<T>(op, uI, rI, wI, vA, vB, zI, dinv, n, wk.st, wk.slot, it0, it1, s))) return rc;
            if ((rc = read_state(wk, h, s))) return rc;
            if (h.done) break;
            it0 = it1 + 1;
        }
        if (h.done == 4 || h.done == 5) break;
        chunk = std::min(kChunk, std::max(4, h.iter - seg0 + 3));
        seg0 = h.iter + 1;
        // done 1 or 2: reliable update, x += scale e and r = b - A x in FP64
        hipLaunchKernelGGL(k_reliable_begin, dim3(1), dim3(64), 0, s, wk.st);
        hipLaunchKernelGGL(k_add_scaled<T>, dim3(gv), dim3(256), 0, s, x, (const T *)uI, n, scale);
        if ((rc = op.residual(x, wk.r, wk.st, kFinBnorm, wk.slot, s))) return rc;
        hipLaunchKernelGGL(k_reliable_next, dim3(1), dim3(64), 0, s, wk.st, r0, scale, tol, delta, maxUpdates);
        if (h.done == 2) {  // inner iterations exhausted: this update is the last
            if ((rc = read_state(wk, h, s))) return rc;
            break;
        }
        // continue from the true residual with the same direction (no-ops
        // once k_reliable_next has stopped the solve)
        hipLaunchKernelGGL((k_convert<double, T>), dim3(gv), dim3(256), 0, s, wk.r, rI, n, 1.0 / scale);
        hipLaunchKernelGGL(k_fill<T>, dim3(gv), dim3(256), 0, s, uI, n, (T)0);
        SELF_RED_LAUNCH((k_precond_dot<T, false>), (k_precond_dot<T, true>), s, (const T *)rI, dinv, wI, n, wk.st,
                           (int)kFinNum, wk.slot);
        it0 = h.iter + 1;
        if (it0 > maxIters) {  // k_reliable_next stops the solve (done 5)
            if ((rc = read_state(wk, h, s))) return rc;
            break;
        }
    }
    SELF_HIP(hipGetLastError());
    iters = h.iter;
    updates = h.l;
    res.iterations = h.l;
    res.rNorm = h.l > 0 ? h.rc : r0;
    res.ioerr = (h.done == 4 || res.rNorm <= tol * r0) ? 0 : -1;
    return SELF_OK;
}

int krylov_solve_reliable(DevOp &op, KrylovWork &wk, int precond, double tol, double delta, int maxIters,
                          int maxUpdates, double *x, KrylovResult &res, int &iters, int &updates, hipStream_t s,
                          bool innerFP64)
{
    return innerFP64 ? solve_reliable<double>(op, wk, precond, tol, delta, maxIters, maxUpdates, x, res, iters,
                                              updates, s)
                     : solve_reliable<float>(op, wk, precond, tol, delta, maxIters, maxUpdates, x, res, iters,
                                             updates, s);
}

}  // namespace self
