// krylov.hpp -- device-resident Krylov solvers and the operators they drive.
//
// The reference's Krylov framework (src/iterativesolve/
// IterativeSolvers_Class.f90:55-186) is an abstract type whose deferred
// MatrixAction / Residual / PreConditionInvert procedures a model overrides.
// Here the same plug-in point is the DevOp interface: an operator that owns
// its device data and launches its own kernels on a stream.  The solvers
// (CG, PCG, GMRES, FGMRES with the reference's control flow and stop tests,
// the CG-SEM PCG of CGsemElliptic_2D_Class.f90:1114-1207, and a mixed
// FP32-inner / FP64-refinement driver) keep every vector and scalar on the
// device; the host only launches and, at chunk boundaries, reads a flag.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "common.hpp"
#include "reduce.hpp"

namespace self {

// Scalars of one solve, in device memory.
struct SolverState {
    double num, den, alpha, beta, rNorm, r0, tol, vz, bnorm;
    double numP[2];  // CG numerator of iteration k in numP[k & 1]
    int iter, done, maxIters, stopRule, ioerr, l, nr, innerExit, m;
    double *resi;
    int resiLen;
    // GMRES work (device pointers, sized for m)
    double *H, *rho, *c, *s, *bhat, *y;
    double rc, b;
};

enum StopRule {
    kStopSqrtNormRel = 0,  // IterativeSolvers CG/PCG: sqrt(rNorm)/r0 < TOL (quirk Q5)
    kStopNormRel = 1,      // CGsemElliptic_2D PCG and GMRES: rNorm/r0 < TOL
};

// Finalize actions a fused reduction can perform in its last block.
enum Finalize {
    kFinNone = 0,
    kFinR0 = 1,      // r0 = rNorm = sqrt(S); resi[0] = r0; done if r0 < tol (CGSEM form)
    kFinNum = 2,     // num = S; beta = (iter == 1) ? 0 : num / den
    kFinAlpha = 3,   // vz = S; alpha = num / vz
    kFinRnorm = 4,   // rNorm = sqrt(S); resi[iter] = rNorm; den = num; stop test; ++iter
    kFinR0Plain = 5, // r0 = sqrt(S); resi[0] = r0 (IterativeSolvers form: no early return)
    kFinNumCG = 6,   // num = S (r.r); beta = num/den unless iter == 1
    kFinBnorm = 7,   // bnorm = sqrt(S)
    kFinNumNext = 8, // after kFinRnorm in the same epilogue: num = S, beta = num/den unless done
};

// Scalar recurrences run by the last workgroup of a fused reduction.
__device__ inline void finalize_scalar(SolverState *st, int kind, double S)
{
    switch (kind) {
    case kFinR0:
        st->r0 = sqrt(S);
        st->rNorm = st->r0;
        if (st->resiLen > 0) st->resi[0] = st->r0;
        if (st->r0 < st->tol) st->done = 3;  // CGSEM PCG: IF( r0 < TOL ) RETURN
        break;
    case kFinR0Plain:
        st->r0 = sqrt(S);
        st->rNorm = 0.0;
        if (st->resiLen > 0) st->resi[0] = st->r0;
        break;
    case kFinNum:
    case kFinNumCG:
        st->num = S;
        st->numP[st->iter & 1] = S;
        st->beta = (st->iter == 1) ? 0.0 : S / st->den;
        break;
    case kFinAlpha:
        st->vz = S;
        st->alpha = st->num / S;
        break;
    case kFinRnorm: {
        const double rn = sqrt(S);
        st->rNorm = rn;
        if (st->iter < st->resiLen) st->resi[st->iter] = rn;
        st->den = st->num;
        const bool stop = (st->stopRule == kStopSqrtNormRel) ? (sqrt(rn) / st->r0 < st->tol)
                                                               : (rn / st->r0 < st->tol);
        if (stop) st->done = 1;
        else if (st->iter >= st->maxIters) st->done = 2;
        else st->iter += 1;
        break;
    }
    case kFinBnorm:
        st->bnorm = sqrt(S);
        break;
    case kFinNumNext:
        if (!st->done) {
            st->num = S;
            st->beta = S / st->den;
        }
        break;
    default:
        break;
    }
}


// ---------------------------------------------------- folded CG reductions
// In the folded CG iteration no reduction has a final launch: the kernel
// that needs a scalar sums the producer's partials itself, every workgroup
// in the same fixed order (so every workgroup has the same bits), and
// workgroup 0 records the scalar bookkeeping in the state.  Partial rows of
// RedSlot: 0 = v.Av (one per workgroup of the operator kernel), 1 = r.r,
// 2 = r.w (kFoldParts each).
constexpr int kFoldParts = 512;    // workgroups of the producing kernels
constexpr int kFoldThreads = 1024; // their size
static_assert(kFoldParts % 64 == 0 && kFoldParts <= kRedBlocks, "fold partials: whole wavefront loads, one row");

// fixed-order sum of one partial row by one wavefront, valid in every lane
__device__ __forceinline__ double wave_fold_sum(const double *row)
{
    const int lane = threadIdx.x & 63;
    double v = 0.0;
    #pragma unroll
    for (int b = 0; b < kFoldParts / 64; ++b) v += row[lane + 64 * b];
    v = wave_sum(v);
    return __shfl(v, 0, 64);
}

__device__ __forceinline__ bool cg_stop(const SolverState *st, double rn)
{
    return (st->stopRule == kStopSqrtNormRel) ? (sqrt(rn) / st->r0 < st->tol) : (rn / st->r0 < st->tol);
}

// End of CG iteration it from S0 = r.r and S1 = r.w: rNorm, resi(it),
// den = num and the stop test, then num and beta of iteration it + 1 (the
// kFinRnorm and kFinNumNext epilogues).  One thread.
__device__ inline void cg_end_apply(SolverState *st, int it, double S0, double S1)
{
    const double rn = sqrt(S0), den = st->numP[it & 1];
    st->rNorm = rn;
    if (it < st->resiLen) st->resi[it] = rn;
    st->den = den;
    if (cg_stop(st, rn)) {
        st->done = 1;
    } else if (it >= st->maxIters) {
        st->done = 2;
    } else {
        st->iter = it + 1;
        st->num = S1;
        st->numP[(it + 1) & 1] = S1;
        st->beta = S1 / den;
    }
}

// The end of a chunk of folded CG iterations whose last is `it`: the
// pending end of iteration it from partial rows 1 and 2 (one wavefront).
__device__ __forceinline__ void cg_fold_end_wave(SolverState *st, const RedSlot &slot, int it)
{
    if (st->done) return;
    const double S0 = wave_fold_sum(slot.partials + kRedBlocks);
    const double S1 = wave_fold_sum(slot.partials + 2 * kRedBlocks);
    if (threadIdx.x == 0) cg_end_apply(st, it, S0, S1);
}

// Operator plug-in (the deferred procedures of IterativeSolver).
struct DevOp {
    size_t n = 0;                   // vector length the solver sees
    virtual ~DevOp() {}
    // y = A x.  When dotWith != nullptr the reduction dot(dotWith, y) is
    // fused and finalized with `fin` into *st.
    virtual int matvec(const double *x, double *y, const double *dotWith, SolverState *st, int fin,
                       const RedSlot &slot, hipStream_t s) = 0;
    // FP32 operator for mixed-precision inner solves (optional).
    virtual bool has_f32() const { return false; }
    virtual int matvec32(const float *x, float *y, const float *dotWith, SolverState *st, int fin,
                         const RedSlot &slot, hipStream_t s) { return SELF_ERR_STATE; }
    // One folded CG iteration `it` for operators that gather their input
    // (fuses_direction() true), no final launches:
    //  - when `pending`, the end of iteration it-1 from partial rows 1 and
    //    2 (cg_end_apply in workgroup 0; every workgroup exits on the stop
    //    test), otherwise beta from the state;
    //  - vNew = w (it == 1) or w + beta vOld formed in the input gather
    //    (w is r without a preconditioner), the operator, v.Av;
    //  - alpha = num / v.Av, u += alpha vNew, r -= alpha A vNew, w = P r,
    //    r.r and r.w as kFoldParts partials in rows 1 and 2.
    // vNew and vOld are distinct buffers (other elements still gather vOld
    // while vNew is written).  Operators without it leave
    // fuses_direction() false and the solver updates v in place.
    virtual bool fuses_direction() const { return false; }
    virtual int cg_iteration(const double *dinv, double *u, double *r, double *w, const double *vOld, double *vNew,
                             SolverState *st, const RedSlot &slot, int it, bool pending, hipStream_t s)
    {
        return SELF_ERR_STATE;
    }
    virtual int cg_iteration32(const float *dinv, float *u, float *r, float *w, const float *vOld, float *vNew,
                               SolverState *st, const RedSlot &slot, int it, bool pending, hipStream_t s)
    {
        return SELF_ERR_STATE;
    }
    // r = b - A x (the model's Residual); |r|^2 fused and finalized with fin.
    virtual int residual(const double *x, double *r, SolverState *st, int fin, const RedSlot &slot,
                         hipStream_t s) = 0;
    // inverse diagonal for Jacobi preconditioning (nullptr: identity)
    virtual const double *inv_diag() const { return nullptr; }
    virtual const float *inv_diag32() const { return nullptr; }
    // A general preconditioner z = M^-1 r (the reference's PreConditionInvert,
    // IterativeSolvers_Class.f90:168-186, or CGsemPreconditioner_2D % Solve),
    // in FP64 or FP32 arithmetic; st is the calling solver's state (work
    // after its done or innerExit flag is skipped).
    virtual bool has_pc() const { return false; }
    virtual int precondition(const double *r, double *z, const SolverState *st, hipStream_t s)
    {
        return SELF_ERR_STATE;
    }
    virtual int precondition32(const float *r, float *z, const SolverState *st, hipStream_t s)
    {
        return SELF_ERR_STATE;
    }
};

// Device workspace of a solver instance.
struct KrylovWork {
    size_t n = 0;
    int m = 0;
    double *r = nullptr, *v = nullptr, *z = nullptr, *w = nullptr, *u = nullptr;
    float *r32 = nullptr, *v32 = nullptr, *z32 = nullptr, *w32 = nullptr, *u32 = nullptr;
    double *vB = nullptr;  // second search-direction buffer (fused direction update)
    float *vB32 = nullptr;
    double *V = nullptr, *Z = nullptr;  // GMRES bases (n x (m+1))
    double *small = nullptr;            // GMRES small matrices
    SolverState *st = nullptr;          // device
    double *resi = nullptr;             // device residual history
    int resiLen = 0;
    RedSlot slot{};
    size_t bytes = 0;
    int ensure(size_t n, int m, int resiLen, bool f32);
    void release();
};

enum KrylovMethod { kCG = 0, kPCG = 1, kGMRES = 2, kPCGMRES = 3, kCGSEM_PCG = 4 };

// preconditioner selectors of the solvers
enum KrylovPrecond { kPcNone = 0, kPcJacobi = 1, kPcOperator = 2 };

struct KrylovResult {
    int ioerr = -2;
    int iterations = 0;
    double r0 = 0, rNorm = 0;
};

// Solve with the reference's control flow; sol is a device vector (initial
// guess in, solution out); resi (host, resiLen) receives the history.
int krylov_solve(DevOp &op, KrylovWork &wk, int method, int precond, double tol, int maxIters, int m,
                 int ortho, double *sol, double *resiHost, int resiLen, KrylovResult &res, hipStream_t s);

// Restarted FGMRES(m) whose operator and preconditioner applications run
// in FP32 while the Krylov basis, Gram-Schmidt, Givens rotations, the
// solution update and the true residual at every restart stay FP64: each
// cycle is one step of FP64 iterative refinement with an FP32 inner solve.
int krylov_solve_gmres_mixed(DevOp &op, KrylovWork &wk, int precond, double tol, int maxCycles, int m, int ortho,
                             double *sol, KrylovResult &res, hipStream_t s);

// The CG / PCG of krylov_solve in parts, for drivers that interleave the
// iterations of several solves (the batched CGSEM solve): the set-up
// before iteration 1 (state, r = b - A x, u = x, w = P r and num = r.w;
// the workspace must hold maxIters + 1 residuals) and, after the
// iterations, the result and sol = u.  precond: Jacobi, else identity.
int krylov_cg_begin(DevOp &op, KrylovWork &wk, bool precond, int stopRule, double tol, int maxIters,
                    const double *sol, hipStream_t s);
int krylov_cg_finish(DevOp &op, KrylovWork &wk, int stopRule, double tol, double *sol, KrylovResult &res,
                     hipStream_t s);

// Mixed precision: FP32 PCG inner solves, FP64 residual and iterative
// refinement to |b - A x| / |b - A x0| <= tol.
int krylov_solve_ir(DevOp &op, KrylovWork &wk, int precond, double tol, double innerTol, int maxInner,
                    int maxOuter, double *sol, KrylovResult &res, int &innerTotal, hipStream_t s);

// Reliable updates: one PCG sequence (FP32, or FP64 with innerFP64), never
// restarted; its iterate is folded into the FP64 solution and the residual
// recomputed in FP64 whenever the recursive residual has dropped by delta.
// res.iterations counts the updates, iters the inner iterations.
int krylov_solve_reliable(DevOp &op, KrylovWork &wk, int precond, double tol, double delta, int maxIters,
                          int maxUpdates, double *sol, KrylovResult &res, int &iters, int &updates, hipStream_t s,
                          bool innerFP64 = false);

}  // namespace self
