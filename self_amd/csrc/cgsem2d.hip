// cgsem2d.hip -- continuous-Galerkin spectral-element elliptic operator
// div(k grad u) = s on GLL nodes, as src/cgsem/CGsemElliptic_2D_Class.f90:
//   CalculateGradient :791-850, FluxDivergence :854-926,
//   MatrixAction :930-976 (UnMask, element operator, GlobalSum, Mask),
//   Residual :980-1068, DotProduct :1072-1110.
//
// Layout.  The reference keeps element-local arrays u(0:N,0:N,nEl) and a
// mask convention (one "owner" copy per shared node, the others zero),
// implemented with edge loops and node-to-element linked lists.  Here the
// Krylov vectors are global-DOF vectors -- exactly the owner copies, in the
// order the owners appear element by element -- so vector operations touch
// each degree of freedom once (0.77x the local count at N = 7) and
// DotProduct is a plain sum.  The operator gathers x through a local->global
// map, applies the element operator with LDS-staged tensor contractions,
// and assembles each global DOF as the sum of its copies in the reference's
// GlobalSum order (edge: primary + secondary; corner: node-list order) --
// a deterministic gather, no atomics.
//
// The element operator uses three precomputed geometric factors per node,
//   G11 = k ws wp (dydp^2 + dxdp^2)/J,  G12 = -k ws wp (dyds dydp + dxds dxdp)/J,
//   G22 = k ws wp (dxds^2 + dyds^2)/J,
// the algebraic expansion of the reference's gradient/flux sequence:
//   F1' = G11 du/ds + G12 du/dp,  F2' = G12 du/ds + G22 du/dp,
//   Lu(i,j) = -( sum_m D(m,i) F1'(m,j) + sum_m D(m,j) F2'(i,m) ).
// (3 doubles per node instead of the 6 metric/coefficient arrays.)

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "common.hpp"
#include "dg2d_device.hpp"
#include "fempc.hpp"
#include "krylov.hpp"

namespace self {
namespace {

template <int N, typename T>
struct CgLDS {
    static constexpr int n1 = N + 1;
    T D[n1 * n1];
};

// F1' = G11 du/ds + G12 du/dp and F2' = G12 du/ds + G22 du/dp at node
// (i, j) of one element's gathered values ue
template <int N, typename T>
__device__ __forceinline__ void cg_flux(const T *D, const T *ue, int i, int j, T g11, T g12, T g22, T *f1, T *f2)
{
    constexpr int n1 = N + 1;
    // du/ds = D u along s (row j), du/dp = D u along p (column i)
    T duds = D[i] * ue[n1 * j];
    T dudp = D[j] * ue[i];
    #pragma unroll
    for (int m = 1; m < n1; ++m) {
        duds = fma(D[i + n1 * m], ue[m + n1 * j], duds);
        dudp = fma(D[j + n1 * m], ue[i + n1 * m], dudp);
    }
    f1[i + n1 * j] = g11 * duds + g12 * dudp;
    f2[i + n1 * j] = g12 * duds + g22 * dudp;
}

// Lu(i, j) = -( sum_m D(m,i) F1'(m,j) + sum_m D(m,j) F2'(i,m) )
template <int N, typename T>
__device__ __forceinline__ T cg_divergence(const T *D, const T *a, const T *b, int i, int j)
{
    constexpr int n1 = N + 1;
    T s1 = D[n1 * i] * a[n1 * j];  // D(0,i) F1'(0,j)
    T s2 = D[n1 * j] * b[i];       // D(0,j) F2'(i,0)
    #pragma unroll
    for (int m = 1; m < n1; ++m) {
        s1 = fma(D[m + n1 * i], a[m + n1 * j], s1);
        s2 = fma(D[m + n1 * j], b[i + n1 * m], s2);
    }
    return -(s1 + s2);
}

// Element operator on gathered values -> Lbuf (local copies).
// RESID: Dirichlet nodes take the prescribed values uD (Residual of the
// solution); otherwise they are zero (search directions are masked there).
template <int N, typename T, bool RESID>
__global__ __launch_bounds__(256) void k_cg_local(const T *__restrict__ x, const int *__restrict__ gid,
                                                  const double *__restrict__ uD, const T *__restrict__ G3,
                                                  const T *__restrict__ Dg, T *__restrict__ Lbuf, int nE,
                                                  const SolverState *st)
{
    using S = Shape<N>;
    constexpr int n1 = S::n1, np = S::np, epb = S::epb;
    if (st && (st->done || (!RESID && st->innerExit))) return;
    __shared__ CgLDS<N, T> L;
    __shared__ T us[epb * np], f1[epb * np], f2[epb * np];
    static_assert(n1 * n1 <= 256, "one thread per entry of the nodal matrix");
    // every load at a clamped valid address, no thread-dependent branch
    // around them: the nodal matrix, the node's DOF index and geometry, then
    // the gathered value (a loaded value behind a branch is waited for at
    // the join, one load after the other)
    const T dk = Dg[threadIdx.x < n1 * n1 ? threadIdx.x : 0];
    const int t = threadIdx.x, le = t / np, nd = t - le * np;
    const int e = blockIdx.x * epb + le;
    const bool act = le < epb && e < nE;
    const int i = nd % n1, j = nd / n1;
    const size_t lid = act ? (size_t)e * np + nd : 0, stride = (size_t)nE * np;
    const int g = gid[lid];
    const T g11 = G3[lid], g12 = G3[stride + lid], g22 = G3[2 * stride + lid];
    T xv = x[g >= 0 ? g : 0];
    asm volatile("" : "+v"(xv));
    if (threadIdx.x < n1 * n1) L.D[threadIdx.x] = dk;
    if (act) {
        T u;
        if (g >= 0) u = xv;
        else u = RESID ? (T)uD[lid] : (T)0;
        us[le * np + nd] = u;
    }
    __syncthreads();
    if (act) cg_flux<N>(L.D, us + le * np, i, j, g11, g12, g22, f1 + le * np, f2 + le * np);
    __syncthreads();
    if (act) Lbuf[lid] = cg_divergence<N>(L.D, f1 + le * np, f2 + le * np, i, j);
}

// LDS of one direction-step workgroup
template <int N, typename T>
struct CgDirLds {
    CgLDS<N, T> L;
    T us[Shape<N>::epb * Shape<N>::np], f1[Shape<N>::epb * Shape<N>::np], f2[Shape<N>::epb * Shape<N>::np];
    double lds4[4];
    double sFold[2];
    int sDone;
};

// The folded CG direction step of iteration `it`, as workgroup b of nB
// (grid-stride over groups of epb elements, nB <= kRedBlocks).  When
// `pending`, the end of iteration it-1 comes first: every workgroup sums
// the update sweep's partial rows 1 and 2 itself (same order, same bits),
// workgroup 0 records it in the state, and all exit if the loop stops;
// otherwise beta is in the state.  The gathered value is the search
// direction formed on the fly, v = x (iteration 1) or x + beta vOld; every
// copy also stores it to vNew[g] (the copies of a DOF store the same bits).
// After the element operator, v.Av is accumulated as the sum over elements
// of v_e . A_e v_e (Dirichlet copies hold v = 0), one partial per
// workgroup in row 0 at entry b: the same number as the dot product of v
// with the assembled A v, so the assembly can be fused with the vector
// update that follows.
// A node's DOF index and geometry for one group of a direction step.  The
// loads are unconditional at a clamped index (node 0 of element 0 for a
// thread without a node, whose values are then discarded), so that no branch
// separates them and they stay in flight together.
template <typename T>
struct CgNodeLoads {
    int g;
    T g11, g12, g22;
};

template <int N, typename T>
__device__ __forceinline__ CgNodeLoads<T> cg_node_loads(const int *__restrict__ gid, const T *__restrict__ G3,
                                                        int nE, size_t stride, int grp, int le, int nd)
{
    constexpr int np = Shape<N>::np, epb = Shape<N>::epb;
    const int e = grp * epb + le;
    const bool act = grp >= 0 && le < epb && e < nE;
    const size_t lid = act ? (size_t)e * np + nd : 0;
    CgNodeLoads<T> o;
    const int g = gid[lid];
    const T a = G3[lid], bb = G3[stride + lid], c = G3[2 * stride + lid];
    o.g = act ? g : -1;
    o.g11 = act ? a : (T)0;
    o.g12 = act ? bb : (T)0;
    o.g22 = act ? c : (T)0;
    return o;
}

template <int N, typename T>
__device__ __forceinline__ void cg_direction_wg(const T *__restrict__ x, const T *__restrict__ vOld,
                                                T *__restrict__ vNew, const int *__restrict__ gid,
                                                const T *__restrict__ G3, const T *__restrict__ Dg,
                                                T *__restrict__ Lbuf, int nE, SolverState *st, const RedSlot &slot,
                                                int it, int pending, int b, int nB, CgDirLds<N, T> &sm)
{
    using S = Shape<N>;
    constexpr int n1 = S::n1, np = S::np, epb = S::epb;
    static_assert(n1 * n1 <= 256, "one thread per entry of the nodal matrix");
    const int t = threadIdx.x, le = t / np, nd = t - le * np;
    const int i = nd % n1, j = nd / n1;
    const size_t stride = (size_t)nE * np;
    const int nGroups = (nE + epb - 1) / epb;
    // What does not depend on beta is loaded first, so that its latency
    // overlaps the fold of the previous iteration's sums: the nodal matrix
    // and the first group's DOF indices and geometry.
    const T dk = Dg[t < n1 * n1 ? t : 0];
    CgNodeLoads<T> cur = cg_node_loads<N, T>(gid, G3, nE, stride, b < nGroups ? b : -1, le, nd);
    T beta;
    if (pending) {
        // workgroup 0 may set `done` during this kernel, so each workgroup
        // reads it once (thread 0) and decides as a whole
        if (threadIdx.x < 128) {  // wavefront q sums row 1 + q
            const int q = threadIdx.x >> 6;
            const double Sq = wave_fold_sum(slot.partials + (1 + q) * kRedBlocks);
            if ((threadIdx.x & 63) == 0) sm.sFold[q] = Sq;
        }
        if (threadIdx.x == 0) sm.sDone = st->done;
        __syncthreads();
        if (sm.sDone) return;
        if (b == 0 && threadIdx.x == 0) cg_end_apply(st, it - 1, sm.sFold[0], sm.sFold[1]);
        if (cg_stop(st, sqrt(sm.sFold[0]))) return;  // uniform over the workgroup
        beta = (T)(sm.sFold[1] / st->numP[(it - 1) & 1]);
    } else {
        if (st->done) return;  // nothing writes it during this kernel
        beta = (T)st->beta;
    }
    if (t < n1 * n1) sm.L.D[t] = dk;
    // iteration 1 has no previous direction: it reads x twice instead
    const T *__restrict__ vPrev = it == 1 ? x : vOld;
    T *us = sm.us, *f1 = sm.f1, *f2 = sm.f2;
    double acc = 0.0;
    for (int grp = b; grp < nGroups; grp += nB) {
        const int e = grp * epb + le;
        const bool act = le < epb && e < nE;
        const size_t lid = (size_t)e * np + nd;
        const int g = cur.g;
        const int gc = g >= 0 ? g : 0;  // a masked node reads DOF 0 and keeps u = 0
        T wv = x[gc], vo = vPrev[gc];
        // an empty statement that takes the gathered values, so the next
        // group's loads below are issued after them and stay in flight
        // while this group computes (3 % per iteration at 128x128, N = 7)
        asm volatile("" : "+v"(wv), "+v"(vo));
        const CgNodeLoads<T> nxt = cg_node_loads<N, T>(gid, G3, nE, stride, grp + nB < nGroups ? grp + nB : -1, le, nd);
        T u = 0;
        if (g >= 0) {
            u = (it == 1) ? wv : wv + beta * vo;
            vNew[g] = u;
        }
        if (act) us[le * np + nd] = u;
        __syncthreads();
        if (act) cg_flux<N>(sm.L.D, us + le * np, i, j, cur.g11, cur.g12, cur.g22, f1 + le * np, f2 + le * np);
        __syncthreads();
        if (act) {
            const T Lv = cg_divergence<N>(sm.L.D, f1 + le * np, f2 + le * np, i, j);
            Lbuf[lid] = Lv;
            acc += (double)u * (double)Lv;
        }
        cur = nxt;
    }
    store_partial_at(slot, 0, acc, sm.lds4, b);
}

template <int N, typename T>
__global__ __launch_bounds__(256) void k_cg_direction(const T *__restrict__ x, const T *__restrict__ vOld,
                                                      T *__restrict__ vNew, const int *__restrict__ gid,
                                                      const T *__restrict__ G3, const T *__restrict__ Dg,
                                                      T *__restrict__ Lbuf, int nE, SolverState *st, RedSlot slot,
                                                      int it, int pending)
{
    __shared__ CgDirLds<N, T> sm;
    cg_direction_wg<N, T>(x, vOld, vNew, gid, G3, Dg, Lbuf, nE, st, slot, it, pending, blockIdx.x, gridDim.x, sm);
}

// One DOF's own values and copy range for an update sweep (nG >= 1, so the
// clamped index g is valid): independent loads, issued together.
template <typename T>
struct CgDofLoads {
    int c0, c1;
    T u, v, r, d;
};

template <typename T, bool PRE>
__device__ __forceinline__ CgDofLoads<T> cg_dof_loads(const int *__restrict__ cs, const T *__restrict__ u,
                                                      const T *__restrict__ v, const T *__restrict__ r,
                                                      const T *__restrict__ dinv, size_t g)
{
    CgDofLoads<T> o;
    o.c0 = cs[g];
    o.c1 = cs[g + 1];
    o.u = u[g];
    o.v = v[g];
    o.r = r[g];
    o.d = (T)0;
    if constexpr (PRE) o.d = dinv[g];
    return o;
}

// Sum of the copies c0 .. c1-1 of one DOF in GlobalSum order (c1 > c0): the
// first four are loaded together at indices clamped to the DOF's range (a
// loop would wait for each copy before loading the next), any further ones
// after them; the additions are the loop's, in the loop's order.
template <typename T>
__device__ __forceinline__ T cg_copies_sum(const T *__restrict__ Lbuf, const int *__restrict__ ci, int c0, int c1)
{
    const int last = c1 - 1;
    const int i0 = ci[c0], i1 = ci[min(c0 + 1, last)], i2 = ci[min(c0 + 2, last)], i3 = ci[min(c0 + 3, last)];
    T l0 = Lbuf[i0], l1 = Lbuf[i1], l2 = Lbuf[i2], l3 = Lbuf[i3];
    // an empty statement that takes the values: without it the compiler
    // moves each load under the test that uses it, one after the other
    asm volatile("" : "+v"(l0), "+v"(l1), "+v"(l2), "+v"(l3));
    const int n = c1 - c0;
    T z = l0;
    if (n > 1) z = z + l1;
    if (n > 2) z = z + l2;
    if (n > 3) z = z + l3;
    for (int c = c0 + 4; c < c1; ++c) z = z + Lbuf[ci[c]];
    return z;
}

// The rest of the folded CG iteration in one sweep over the global DOFs, as
// workgroup b of kFoldParts (kFoldThreads threads each): alpha = num / v.Av
// (each workgroup sums the direction step's nPart partials itself), z = A v
// assembled from the copies in GlobalSum order and used at once:
// u += alpha v, r -= alpha z, w = P r; r.r and r.w partials in rows 1 and 2
// at entry b.  z is never stored.  A workgroup b >= ceil(nG / kFoldThreads)
// has no DOF: its partials are zero.  A DOF's own values and copy range do
// not depend on alpha, so the first DOF's are loaded before the state check
// and the fold, and each next DOF's while the current one is summed.
template <typename T, bool PRE>
__device__ __forceinline__ void cg_update_wg(const T *__restrict__ Lbuf, const int *__restrict__ cs,
                                             const int *__restrict__ ci, size_t nG, T *__restrict__ u,
                                             const T *__restrict__ v, T *__restrict__ r, const T *__restrict__ dinv,
                                             T *__restrict__ w, SolverState *st, const RedSlot &slot, int it,
                                             int nPart, int b)
{
    const size_t step = (size_t)kFoldParts * kFoldThreads;
    size_t g = (size_t)b * kFoldThreads + threadIdx.x;
    CgDofLoads<T> cur = cg_dof_loads<T, PRE>(cs, u, v, r, dinv, g < nG ? g : nG - 1);
    if (st->done) return;
    const double vz = block_fold_sum<kFoldThreads>(slot.partials, nPart);
    const double a = st->numP[it & 1] / vz;
    if (b == 0 && threadIdx.x == 0) {
        st->vz = vz;
        st->alpha = a;
    }
    const T alpha = (T)a;
    double acc0 = 0.0, acc1 = 0.0;
    for (; g < nG; g += step) {
        const size_t gn = g + step;
        const CgDofLoads<T> nxt = cg_dof_loads<T, PRE>(cs, u, v, r, dinv, gn < nG ? gn : nG - 1);
        const T z = cg_copies_sum(Lbuf, ci, cur.c0, cur.c1);
        u[g] = cur.u + alpha * cur.v;
        const T rv = cur.r - alpha * z;
        r[g] = rv;
        acc0 += (double)rv * (double)rv;
        if constexpr (PRE) {
            const T wv = cur.d * rv;
            w[g] = wv;
            acc1 += (double)rv * (double)wv;
        }
        cur = nxt;
    }
    if constexpr (!PRE) acc1 = acc0;
    store_partials2_at<kFoldThreads>(slot, 1, acc0, 2, acc1, b);
}

template <typename T, bool PRE>
__global__ __launch_bounds__(kFoldThreads) void k_cg_assemble_update(
    const T *__restrict__ Lbuf, const int *__restrict__ cs, const int *__restrict__ ci, size_t nG, T *__restrict__ u,
    const T *__restrict__ v, T *__restrict__ r, const T *__restrict__ dinv, T *__restrict__ w, SolverState *st,
    RedSlot slot, int it, int nPart)
{
    cg_update_wg<T, PRE>(Lbuf, cs, ci, nG, u, v, r, dinv, w, st, slot, it, nPart, blockIdx.x);
}

// y[g] = sum over the copies of g (GlobalSum order); optional fused dot.
// MODE 0: matvec.  MODE 1: residual  y = (b[c0] - sum L) + sum_{c>=1} b[c],
// with |y|^2 reduced.
template <typename T, int MODE, bool FIN, int TPB = kRedThreads>
__global__ __launch_bounds__(TPB) void k_cg_assemble(const T *__restrict__ Lbuf, const int *__restrict__ cs,
                                                             const int *__restrict__ ci, const double *__restrict__ bl,
                                                             size_t nG, T *__restrict__ y, const T *__restrict__ dotWith,
                                                             SolverState *st, int fin, RedSlot slot)
{
    if (st && (st->done || (MODE == 0 && st->innerExit))) return;
    __shared__ double lds4[4];
    if constexpr (FIN) {  // one workgroup after the partial launch
        const double S = final_sum(slot, 0, lds4);
        if (threadIdx.x == 0) finalize_scalar(st, fin, S);
        return;
    }
    double acc = 0.0;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < nG; g += (size_t)gridDim.x * blockDim.x) {
        const int c0 = cs[g], c1 = cs[g + 1];
        const T sum = cg_copies_sum(Lbuf, ci, c0, c1);
        T out;
        if (MODE == 1) {
            double r = bl[ci[c0]] - (double)sum;
            for (int c = c0 + 1; c < c1; ++c) r = r + bl[ci[c]];
            out = (T)r;
            acc += r * r;
        } else {
            out = sum;
            if (dotWith) acc += (double)dotWith[g] * (double)out;
        }
        y[g] = out;
    }
    if (fin == kFinNone || !st) return;
    if constexpr (TPB == kRedThreads) store_partial(slot, 0, acc, lds4);
    else store_partial1<TPB>(slot, 0, acc);
}

#define CG_DISPATCH_N(N_, CALL)                          \
    switch (N_) {                                        \
    case 1: { constexpr int NN = 1; CALL; } break;       \
    case 2: { constexpr int NN = 2; CALL; } break;       \
    case 3: { constexpr int NN = 3; CALL; } break;       \
    case 4: { constexpr int NN = 4; CALL; } break;       \
    case 5: { constexpr int NN = 5; CALL; } break;       \
    case 6: { constexpr int NN = 6; CALL; } break;       \
    case 7: { constexpr int NN = 7; CALL; } break;       \
    case 8: { constexpr int NN = 8; CALL; } break;       \
    case 9: { constexpr int NN = 9; CALL; } break;       \
    case 10: { constexpr int NN = 10; CALL; } break;     \
    case 11: { constexpr int NN = 11; CALL; } break;     \
    default: break;                                      \
    }

template <int N>
int cg_grid(size_t nE) { return (int)((nE + Shape<N>::epb - 1) / Shape<N>::epb); }

}  // namespace

// ------------------------------------------------------------ operator
struct CgsemOp final : DevOp {
    int N = 0, n1 = 0, np = 0;
    size_t nE = 0, nG = 0, nLocal = 0;
    hipStream_t stream = nullptr;
    // host maps
    std::vector<int> gidH, csH, ciH;
    std::vector<double> wH, DH;
    std::vector<double> met[5];  // dxds, dxdp, dyds, dydp, J (local)
    std::vector<int> dirH;       // 1 at Dirichlet local nodes
    // device
    int *gid = nullptr, *cs = nullptr, *ci = nullptr;
    double *G3 = nullptr, *D = nullptr, *bl = nullptr, *uD = nullptr, *dinv = nullptr, *Lbuf = nullptr;
    double *metD = nullptr, *kD = nullptr, *wD = nullptr;  // metric terms (5 x local), flux coefficient, weights
    float *G3f = nullptr, *Df = nullptr, *dinvf = nullptr, *Lbuf32 = nullptr;
    double *xg = nullptr;  // global solution vector
    size_t bytes = 0;
    std::vector<double> kH;         // flux coefficient last set (host, local layout; empty = ONE)
    std::unique_ptr<FemPC> pc;      // CGsemPreconditioner_2D (self_cgsem2d_set_preconditioner)

    ~CgsemOp() override
    {
        void *p[] = {gid, cs, ci, G3, D, bl, uD, dinv, Lbuf, G3f, Df, dinvf, Lbuf32, xg, metD, kD, wD};
        for (void *q : p)
            if (q) (void)hipFree(q);
        if (stream) (void)hipStreamDestroy(stream);
    }

    template <typename T>
    int dev(T **p, size_t cnt)
    {
        SELF_HIP(hipMalloc(reinterpret_cast<void **>(p), sizeof(T) * std::max<size_t>(cnt, 1)));
        bytes += sizeof(T) * cnt;
        return SELF_OK;
    }

    bool has_f32() const override { return true; }
    bool has_pc() const override { return pc != nullptr; }
    int precondition(const double *r, double *z, const SolverState *st, hipStream_t s) override
    {
        return pc ? pc->apply(r, z, st, s) : SELF_ERR_STATE;
    }
    int precondition32(const float *r, float *z, const SolverState *st, hipStream_t s) override
    {
        return pc ? pc->apply32(r, z, st, s) : SELF_ERR_STATE;
    }
    const double *inv_diag() const override { return dinv; }
    const float *inv_diag32() const override { return dinvf; }

    int matvec(const double *x, double *y, const double *dotWith, SolverState *st, int fin, const RedSlot &slot,
               hipStream_t s) override
    {
        CG_DISPATCH_N(N, hipLaunchKernelGGL((k_cg_local<NN, double, false>), dim3(cg_grid<NN>(nE)), dim3(256), 0, s,
                                            x, gid, uD, G3, D, Lbuf, (int)nE, st));
        hipLaunchKernelGGL((k_cg_assemble<double, 0, false>), dim3(kRedBlocks), dim3(kRedThreads), 0, s, Lbuf, cs,
                           ci, (const double *)nullptr, nG, y, dotWith, st, fin, slot);
        if (fin != kFinNone && st)
            hipLaunchKernelGGL((k_cg_assemble<double, 0, true>), dim3(1), dim3(kRedThreads), 0, s, Lbuf, cs, ci,
                               (const double *)nullptr, nG, y, dotWith, st, fin, slot);
        SELF_HIP(hipGetLastError());
        return SELF_OK;
    }

    int matvec32(const float *x, float *y, const float *dotWith, SolverState *st, int fin, const RedSlot &slot,
                 hipStream_t s) override
    {
        CG_DISPATCH_N(N, hipLaunchKernelGGL((k_cg_local<NN, float, false>), dim3(cg_grid<NN>(nE)), dim3(256), 0, s,
                                            x, gid, uD, G3f, Df, Lbuf32, (int)nE, st));
        hipLaunchKernelGGL((k_cg_assemble<float, 0, false>), dim3(kRedBlocks), dim3(kRedThreads), 0, s, Lbuf32, cs,
                           ci, (const double *)nullptr, nG, y, dotWith, st, fin, slot);
        if (fin != kFinNone && st)
            hipLaunchKernelGGL((k_cg_assemble<float, 0, true>), dim3(1), dim3(kRedThreads), 0, s, Lbuf32, cs, ci,
                               (const double *)nullptr, nG, y, dotWith, st, fin, slot);
        SELF_HIP(hipGetLastError());
        return SELF_OK;
    }

    bool fuses_direction() const override { return true; }

    template <typename T>
    int cg_iter(const T *dinvT, T *u, T *r, T *w, const T *vOld, T *vNew, const T *G3T, const T *DT, T *LbufT,
                SolverState *st, const RedSlot &slot, int it, bool pending, hipStream_t s)
    {
        const T *src = dinvT ? w : r;
        int nPart = 0;
        CG_DISPATCH_N(N, nPart = std::min(cg_grid<NN>(nE), kRedBlocks);
                      hipLaunchKernelGGL((k_cg_direction<NN, T>), dim3(nPart), dim3(256), 0, s, src, vOld, vNew, gid,
                                         G3T, DT, LbufT, (int)nE, st, slot, it, (int)pending));
        if (dinvT)
            hipLaunchKernelGGL((k_cg_assemble_update<T, true>), dim3(kFoldParts), dim3(kFoldThreads), 0, s,
                               (const T *)LbufT, cs, ci, nG, u, (const T *)vNew, r, dinvT, w, st, slot, it, nPart);
        else
            hipLaunchKernelGGL((k_cg_assemble_update<T, false>), dim3(kFoldParts), dim3(kFoldThreads), 0, s,
                               (const T *)LbufT, cs, ci, nG, u, (const T *)vNew, r, dinvT, w, st, slot, it, nPart);
        SELF_HIP(hipGetLastError());
        return SELF_OK;
    }

    int cg_iteration(const double *dinvA, double *u, double *r, double *w, const double *vOld, double *vNew,
                     SolverState *st, const RedSlot &slot, int it, bool pending, hipStream_t s) override
    {
        return cg_iter<double>(dinvA, u, r, w, vOld, vNew, G3, D, Lbuf, st, slot, it, pending, s);
    }

    int cg_iteration32(const float *dinvA, float *u, float *r, float *w, const float *vOld, float *vNew,
                       SolverState *st, const RedSlot &slot, int it, bool pending, hipStream_t s) override
    {
        return cg_iter<float>(dinvA, u, r, w, vOld, vNew, G3f, Df, Lbuf32, st, slot, it, pending, s);
    }

    int residual(const double *x, double *r, SolverState *st, int fin, const RedSlot &slot, hipStream_t s) override
    {
        CG_DISPATCH_N(N, hipLaunchKernelGGL((k_cg_local<NN, double, true>), dim3(cg_grid<NN>(nE)), dim3(256), 0, s,
                                            x, gid, uD, G3, D, Lbuf, (int)nE, st));
        hipLaunchKernelGGL((k_cg_assemble<double, 1, false>), dim3(kRedBlocks), dim3(kRedThreads), 0, s, Lbuf, cs,
                           ci, bl, nG, r, (const double *)nullptr, st, fin, slot);
        if (fin != kFinNone && st)
            hipLaunchKernelGGL((k_cg_assemble<double, 1, true>), dim3(1), dim3(kRedThreads), 0, s, Lbuf, cs, ci, bl,
                               nG, r, (const double *)nullptr, st, fin, slot);
        SELF_HIP(hipGetLastError());
        return SELF_OK;
    }
};

namespace {

// side s (1..4) position k (0..N) -> local (i, j)
inline void side_ij(int s, int k, int N, int &i, int &j)
{
    switch (s) {
    case 1: i = k; j = 0; break;
    case 2: i = N; j = k; break;
    case 3: i = k; j = N; break;
    default: i = 0; j = k; break;
    }
}

constexpr int kEdgeMapC[2][4] = {{1, 2, 4, 1}, {2, 3, 3, 4}};
constexpr int kCornerI[4] = {0, 1, 1, 0}, kCornerJ[4] = {0, 0, 1, 1};

// Build gid/copy maps with the reference's ownership and GlobalSum order.
int build_maps(CgsemOp &op, int nEdges, int nNodes, const int *edges, const int *elemNodes)
{
    const int N = op.N, n1 = op.n1, np = op.np;
    const size_t nE = op.nE;
    // per DOF (creation order): its local copies in summation order, CSR
    std::vector<int> cpStart{0}, cpList;
    cpStart.reserve(nE * np + 1);
    cpList.reserve(nE * np);
    op.dirH.assign(nE * np, 0);
    // node -> element list (increasing element order) and Dirichlet nodes
    std::vector<std::vector<std::pair<int, int>>> nodeList(nNodes);
    for (size_t e = 0; e < nE; ++e)
        for (int k = 0; k < 4; ++k) {
            const int nid = elemNodes[4 * e + k];
            SELF_REQUIRE(nid >= 1 && nid <= nNodes, SELF_ERR_ARG, "element %zu corner %d: node %d", e + 1, k + 1, nid);
            nodeList[nid - 1].push_back({(int)e, k});
        }
    std::vector<char> dirNode(nNodes, 0);
    std::vector<int> sideEdge(nE * 4, -1);
    for (int ie = 0; ie < nEdges; ++ie) {
        const int *ed = edges + 6 * ie;
        const int e1 = ed[0] - 1, s1 = ed[1];
        SELF_REQUIRE(e1 >= 0 && (size_t)e1 < nE && s1 >= 1 && s1 <= 4, SELF_ERR_ARG, "edge %d: bad primary", ie + 1);
        sideEdge[e1 * 4 + s1 - 1] = ie;
        if (ed[2] > 0) {
            SELF_REQUIRE((size_t)(ed[2] - 1) < nE && std::abs(ed[3]) >= 1 && std::abs(ed[3]) <= 4, SELF_ERR_ARG,
                         "edge %d: bad secondary", ie + 1);
            SELF_REQUIRE(ed[5] == 1 || ed[5] == -1, SELF_ERR_ARG, "edge %d: increment %d", ie + 1, ed[5]);
            sideEdge[(ed[2] - 1) * 4 + std::abs(ed[3]) - 1] = ie;
        } else {  // boundary edge: Dirichlet (BuildQuadMesh_CGsemElliptic_2D :207-214)
            dirNode[elemNodes[4 * e1 + kEdgeMapC[0][s1 - 1] - 1] - 1] = 1;
            dirNode[elemNodes[4 * e1 + kEdgeMapC[1][s1 - 1] - 1] - 1] = 1;
        }
    }
    auto new_dof = [&](std::initializer_list<int> cp) {
        const int d = (int)cpStart.size() - 1;
        for (int l : cp) cpList.push_back(l);
        cpStart.push_back((int)cpList.size());
        return d;
    };
    // assign DOF identities (not yet numbered): edges
    std::vector<int> dofKey(nE * np, -1);
    for (size_t e = 0; e < nE; ++e)
        for (int j = 1; j < N; ++j)
            for (int i = 1; i < N; ++i) {
                const int l = (int)e * np + i + n1 * j;
                dofKey[l] = new_dof({l});
            }
    for (int ie = 0; ie < nEdges; ++ie) {
        const int *ed = edges + 6 * ie;
        const int e1 = ed[0] - 1, s1 = ed[1];
        if (ed[2] <= 0) {
            for (int k = 1; k < N; ++k) {
                int i, j;
                side_ij(s1, k, N, i, j);
                op.dirH[e1 * np + i + n1 * j] = 1;
            }
            continue;
        }
        const int e2 = ed[2] - 1, s2 = std::abs(ed[3]);
        int nn = ed[4];
        for (int k = 1; k < N; ++k) {
            int i1, j1, i2, j2;
            side_ij(s1, k, N, i1, j1);
            SELF_REQUIRE(nn >= 1 && nn <= N - 1, SELF_ERR_ARG, "edge %d: start/inc leave the side", ie + 1);
            side_ij(s2, nn, N, i2, j2);
            const int l1 = e1 * np + i1 + n1 * j1, l2 = e2 * np + i2 + n1 * j2;
            const int d = new_dof({l1, l2});
            dofKey[l1] = d;
            dofKey[l2] = d;
            nn += ed[5];
        }
    }
    for (int nd = 0; nd < nNodes; ++nd) {
        if (nodeList[nd].empty()) continue;
        if (dirNode[nd]) {
            for (auto &pe : nodeList[nd])
                op.dirH[pe.first * np + kCornerI[pe.second] * N + n1 * (kCornerJ[pe.second] * N)] = 1;
            continue;
        }
        const int d = (int)cpStart.size() - 1;
        for (auto &pe : nodeList[nd]) {
            const int l = pe.first * np + kCornerI[pe.second] * N + n1 * (kCornerJ[pe.second] * N);
            cpList.push_back(l);
            dofKey[l] = d;
        }
        cpStart.push_back((int)cpList.size());
    }
    for (size_t l = 0; l < nE * (size_t)np; ++l)
        SELF_REQUIRE(dofKey[l] >= 0 || op.dirH[l], SELF_ERR_ARG,
                     "local node %zu of element %zu is on no edge (mesh tables incomplete)", l % np, l / np + 1);
    // number DOFs in the order their owner copies appear, element by element
    const size_t nDof = cpStart.size() - 1;
    std::vector<int> number(nDof, -1);
    int next = 0;
    for (size_t l = 0; l < nE * (size_t)np; ++l) {
        const int d = dofKey[l];
        if (d >= 0 && cpList[cpStart[d]] == (int)l) number[d] = next++;
    }
    op.nG = (size_t)next;
    // interior nodes are never Dirichlet nodes, so N >= 2 leaves at least one
    SELF_REQUIRE(op.nG >= 1, SELF_ERR_ARG, "no global degree of freedom");
    op.gidH.assign(nE * np, -1);
    for (size_t l = 0; l < nE * (size_t)np; ++l)
        if (dofKey[l] >= 0) op.gidH[l] = number[dofKey[l]];
    std::vector<int> order(nDof);
    for (size_t d = 0; d < nDof; ++d) order[number[d]] = (int)d;
    op.csH.assign(op.nG + 1, 0);
    op.ciH.clear();
    op.ciH.reserve(cpList.size());
    for (size_t g = 0; g < op.nG; ++g) {
        const int d = order[g];
        for (int c = cpStart[d]; c < cpStart[d + 1]; ++c) op.ciH.push_back(cpList[c]);
        op.csH[g + 1] = (int)op.ciH.size();
    }
    return SELF_OK;
}

}  // namespace
}  // namespace self

using namespace self;

struct self_cgsem2d_s {
    CgsemOp op;
    KrylovWork wk;
    int device = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<double> resi;
};

namespace {

// Geometric factors, element diagonal and Jacobi inverse, on the device,
// with the rounding of the plain expressions (no contraction):
//   c = k ws wp / J, G11 = c (dydp^2 + dxdp^2), G12 = -c (dyds dydp + dxds dxdp),
//   G22 = c (dxds^2 + dyds^2);
//   A(l,l) = -( sum_m G11(m,j) D(m,i)^2 + 2 G12(i,j) D(i,i) D(j,j) + sum_m G22(i,m) D(m,j)^2 ),
//   summed over the copies of each global DOF in GlobalSum order.
__global__ __launch_bounds__(256) void k_cg_geometry(const double *__restrict__ met, const double *__restrict__ k,
                                                     const double *__restrict__ w, int n1, size_t nl,
                                                     double *__restrict__ G3, float *__restrict__ G3f)
{
#pragma clang fp contract(off)
    const int np = n1 * n1;
    for (size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x; l < nl; l += (size_t)gridDim.x * blockDim.x) {
        const int nd = (int)(l % np), i = nd % n1, j = nd / n1;
        const double dxds = met[l], dxdp = met[nl + l], dyds = met[2 * nl + l], dydp = met[3 * nl + l];
        const double J = met[4 * nl + l];
        const double c = k[l] * w[i] * w[j] / J;
        const double g11 = c * (dydp * dydp + dxdp * dxdp);
        const double g12 = -c * (dyds * dydp + dxds * dxdp);
        const double g22 = c * (dxds * dxds + dyds * dyds);
        G3[l] = g11;
        G3[nl + l] = g12;
        G3[2 * nl + l] = g22;
        G3f[l] = (float)g11;
        G3f[nl + l] = (float)g12;
        G3f[2 * nl + l] = (float)g22;
    }
}

__global__ __launch_bounds__(256) void k_cg_diagonal(const double *__restrict__ G3, const double *__restrict__ D,
                                                     int n1, size_t nl, double *__restrict__ diag)
{
#pragma clang fp contract(off)
    const int np = n1 * n1;
    for (size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x; l < nl; l += (size_t)gridDim.x * blockDim.x) {
        const size_t e0 = l - l % np;
        const int nd = (int)(l % np), i = nd % n1, j = nd / n1;
        double a = 0.0, b = 0.0;
        for (int m = 0; m < n1; ++m) {
            const double dmi = D[m + n1 * i], dmj = D[m + n1 * j];
            a += G3[e0 + m + n1 * j] * dmi * dmi;
            b += G3[2 * nl + e0 + i + n1 * m] * dmj * dmj;
        }
        diag[l] = -(a + 2.0 * G3[nl + l] * D[i + n1 * i] * D[j + n1 * j] + b);
    }
}

__global__ __launch_bounds__(256) void k_cg_jacobi_inverse(const double *__restrict__ diag, const int *__restrict__ cs,
                                                           const int *__restrict__ ci, size_t nG,
                                                           double *__restrict__ dinv, float *__restrict__ dinvf)
{
#pragma clang fp contract(off)
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < nG; g += (size_t)gridDim.x * blockDim.x) {
        double sum = diag[ci[cs[g]]];
        for (int c = cs[g] + 1; c < cs[g + 1]; ++c) sum += diag[ci[c]];
        const double v = 1.0 / sum;
        dinv[g] = v;
        dinvf[g] = (float)v;
    }
}

__global__ __launch_bounds__(256) void k_fill_f64(double *__restrict__ a, size_t n, double v)
{
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) a[q] = v;
}

// G3 and the Jacobi inverse from the device copies of the metric terms and
// flux coefficient (the element diagonal goes through Lbuf, free here)
int device_geometry(CgsemOp &op)
{
    const size_t nl = op.nLocal;
    const int grid = (int)std::min<size_t>(8192, (nl + 255) / 256);
    hipLaunchKernelGGL(k_cg_geometry, dim3(grid), dim3(256), 0, op.stream, (const double *)op.metD,
                       (const double *)op.kD, (const double *)op.wD, op.n1, nl, op.G3, op.G3f);
    hipLaunchKernelGGL(k_cg_diagonal, dim3(grid), dim3(256), 0, op.stream, (const double *)op.G3,
                       (const double *)op.D, op.n1, nl, op.Lbuf);
    const int gridG = (int)std::min<size_t>(8192, (op.nG + 255) / 256);
    hipLaunchKernelGGL(k_cg_jacobi_inverse, dim3(gridG), dim3(256), 0, op.stream, (const double *)op.Lbuf,
                       (const int *)op.cs, (const int *)op.ci, op.nG, op.dinv, op.dinvf);
    SELF_HIP(hipGetLastError());
    SELF_HIP(hipStreamSynchronize(op.stream));
    return SELF_OK;
}

// ------------------------------------------------------ batched CGSEM PCG
// Many CGSEM problems (any mesh, degree, coefficient and source) solved
// together: every CG iteration is two launches for all unfinished
// problems.  Problem p owns nPart(p) = min(element groups, kRedBlocks)
// direction workgroups and nUpd(p) = min(kFoldParts, ceil(nG / kFoldThreads))
// update workgroups (the update workgroups of its single solve that have
// DOFs; the others would only write zero partials, which stay zero in rows
// 1 and 2 from the start).  A workgroup finds its problem and its index b
// in that problem's grid from prefix sums over the problems still running,
// and runs the single solve's workgroup body with that b: each problem's
// arithmetic, reduction order included, is its single solve's, so the
// results are bitwise those of self_cgsem2d_solve.
struct CgBatchProb {
    const int *gid, *cs, *ci;
    const double *G3, *D, *dinv;  // dinv null: no preconditioner
    double *Lbuf, *u, *r, *w, *vA, *vB;
    SolverState *st;
    RedSlot slot;
    size_t nG;
    int N, nE, nPart, nUpd;
};

union CgDirLdsAny {
    CgDirLds<2, double> n2;
    CgDirLds<3, double> n3;
    CgDirLds<4, double> n4;
    CgDirLds<5, double> n5;
    CgDirLds<6, double> n6;
    CgDirLds<7, double> n7;
    CgDirLds<8, double> n8;
    CgDirLds<9, double> n9;
    CgDirLds<10, double> n10;
    CgDirLds<11, double> n11;
};

// the k with start[k] <= wg < start[k + 1] (start[0] = 0, start[nAct] = grid)
__device__ __forceinline__ int cgb_find(const int *start, int nAct, int wg)
{
    int lo = 0, hi = nAct;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (start[mid] <= wg) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_cgb_direction(const CgBatchProb *__restrict__ probs,
                                                       const int *__restrict__ act, const int *__restrict__ start,
                                                       int nAct, int it, int pending)
{
    __shared__ CgDirLdsAny sm;
    const int k = cgb_find(start, nAct, blockIdx.x);
    const CgBatchProb &P = probs[act[k]];
    const int b = blockIdx.x - start[k];
    const double *src = P.dinv ? P.w : P.r;
    const double *vOld = (it & 1) ? P.vB : P.vA;
    double *vNew = (it & 1) ? P.vA : P.vB;
#define CGB_DIR(NN, F)                                                                                          \
    case NN:                                                                                                    \
        cg_direction_wg<NN, double>(src, vOld, vNew, P.gid, P.G3, P.D, P.Lbuf, P.nE, P.st, P.slot, it, pending, \
                                    b, P.nPart, sm.F);                                                          \
        break;
    switch (P.N) {
        CGB_DIR(2, n2) CGB_DIR(3, n3) CGB_DIR(4, n4) CGB_DIR(5, n5) CGB_DIR(6, n6)
        CGB_DIR(7, n7) CGB_DIR(8, n8) CGB_DIR(9, n9) CGB_DIR(10, n10) CGB_DIR(11, n11)
    default: break;
    }
#undef CGB_DIR
}

__global__ __launch_bounds__(kFoldThreads) void k_cgb_update(const CgBatchProb *__restrict__ probs,
                                                             const int *__restrict__ act,
                                                             const int *__restrict__ start, int nAct, int it)
{
    const int k = cgb_find(start, nAct, blockIdx.x);
    const CgBatchProb &P = probs[act[k]];
    const int b = blockIdx.x - start[k];
    const double *v = (it & 1) ? P.vA : P.vB;
    if (P.dinv)
        cg_update_wg<double, true>(P.Lbuf, P.cs, P.ci, P.nG, P.u, v, P.r, P.dinv, P.w, P.st, P.slot, it, P.nPart, b);
    else
        cg_update_wg<double, false>(P.Lbuf, P.cs, P.ci, P.nG, P.u, v, P.r, P.dinv, P.w, P.st, P.slot, it, P.nPart,
                                    b);
}

// one wavefront per running problem: the pending end of iteration it, then
// its done flag for the host
__global__ __launch_bounds__(64) void k_cgb_fold_end(const CgBatchProb *__restrict__ probs,
                                                     const int *__restrict__ act, int it, int *__restrict__ flags)
{
    const CgBatchProb &P = probs[act[blockIdx.x]];
    cg_fold_end_wave(P.st, P.slot, it);
    if (threadIdx.x == 0) flags[blockIdx.x] = P.st->done;
}

constexpr int kCgbChunk = 32;        // iterations between reads of the done flags
constexpr int kCgbMaxProblems = 4096;

}  // namespace

extern "C" {

int self_cgsem2d_create(int N, int nElems, int nEdges, int nNodes, const int *edges, const int *elemNodes,
                        const double *dxds, const double *dxdp, const double *dyds, const double *dydp,
                        const double *J, self_cgsem2d_t *out)
try {
    SELF_REQUIRE(out && edges && elemNodes && dxds && dxdp && dyds && dydp && J, SELF_ERR_ARG,
                 "self_cgsem2d_create: null argument");
    SELF_REQUIRE(N >= 2 && N <= 11 && nElems >= 1 && nEdges >= 1 && nNodes >= 1, SELF_ERR_ARG,
                 "self_cgsem2d_create: N=%d nElems=%d", N, nElems);
    auto *h = new self_cgsem2d_s();
    CgsemOp &op = h->op;
    op.N = N;
    op.n1 = N + 1;
    op.np = op.n1 * op.n1;
    op.nE = (size_t)nElems;
    op.nLocal = op.nE * op.np;
    auto fail = [&](int rc) { delete h; return rc; };
    hipError_t he = hipGetDevice(&h->device);
    if (he != hipSuccess) return fail(hip_fail(he, "hipGetDevice", __FILE__, __LINE__));
    he = hipStreamCreateWithFlags(&op.stream, hipStreamNonBlocking);
    if (he != hipSuccess) return fail(hip_fail(he, "hipStreamCreate", __FILE__, __LINE__));
    // GLL nodes/weights and the CG derivative matrix (NodalStorage_2D, CG form)
    std::vector<double> x(op.n1), bw(op.n1);
    op.wH.resize(op.n1);
    op.DH.resize(op.np);
    int rc = self_legendre_quadrature(N, SELF_GAUSS_LOBATTO, x.data(), op.wH.data());
    if (!rc) rc = self_barycentric_weights(N, x.data(), bw.data());
    if (!rc) rc = self_derivative_matrix(N, x.data(), bw.data(), op.DH.data());
    if (rc) return fail(rc);
    const double *src[5] = {dxds, dxdp, dyds, dydp, J};
    for (int k = 0; k < 5; ++k) op.met[k].assign(src[k], src[k] + op.nLocal);
    if ((rc = build_maps(op, nEdges, nNodes, edges, elemNodes))) return fail(rc);
    if ((rc = op.dev(&op.gid, op.nLocal)) || (rc = op.dev(&op.cs, op.nG + 1)) || (rc = op.dev(&op.ci, op.ciH.size())) ||
        (rc = op.dev(&op.G3, 3 * op.nLocal)) || (rc = op.dev(&op.G3f, 3 * op.nLocal)) ||
        (rc = op.dev(&op.D, op.np)) || (rc = op.dev(&op.Df, op.np)) || (rc = op.dev(&op.bl, op.nLocal)) ||
        (rc = op.dev(&op.uD, op.nLocal)) || (rc = op.dev(&op.dinv, op.nG)) || (rc = op.dev(&op.dinvf, op.nG)) ||
        (rc = op.dev(&op.Lbuf, op.nLocal)) || (rc = op.dev(&op.Lbuf32, op.nLocal)) || (rc = op.dev(&op.xg, op.nG)) ||
        (rc = op.dev(&op.metD, 5 * op.nLocal)) || (rc = op.dev(&op.kD, op.nLocal)) || (rc = op.dev(&op.wD, op.n1)))
        return fail(rc);
    op.n = op.nG;
    std::vector<float> Df(op.np);
    for (int k = 0; k < op.np; ++k) Df[k] = (float)op.DH[k];
    he = hipMemcpy(op.gid, op.gidH.data(), sizeof(int) * op.nLocal, hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipMemcpy(op.cs, op.csH.data(), sizeof(int) * (op.nG + 1), hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipMemcpy(op.ci, op.ciH.data(), sizeof(int) * op.ciH.size(), hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipMemcpy(op.D, op.DH.data(), sizeof(double) * op.np, hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipMemcpy(op.Df, Df.data(), sizeof(float) * op.np, hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipMemset(op.bl, 0, sizeof(double) * op.nLocal);
    if (he == hipSuccess) he = hipMemset(op.uD, 0, sizeof(double) * op.nLocal);
    if (he == hipSuccess) he = hipMemset(op.xg, 0, sizeof(double) * op.nG);
    for (int k = 0; k < 5 && he == hipSuccess; ++k)
        he = hipMemcpy(op.metD + k * op.nLocal, src[k], sizeof(double) * op.nLocal, hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipMemcpy(op.wD, op.wH.data(), sizeof(double) * op.n1, hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipEventCreate(&h->ev0);
    if (he == hipSuccess) he = hipEventCreate(&h->ev1);
    if (he != hipSuccess) return fail(hip_fail(he, "cgsem2d upload", __FILE__, __LINE__));
    // Build_CGsemElliptic_2D: fluxCoeff = ONE
    hipLaunchKernelGGL(k_fill_f64, dim3((int)std::min<size_t>(8192, (op.nLocal + 255) / 256)), dim3(256), 0,
                       op.stream, op.kD, op.nLocal, 1.0);
    if ((rc = device_geometry(op))) return fail(rc);
    *out = h;
    return SELF_OK;
}
SELF_CATCH_ALL

int self_cgsem2d_destroy(self_cgsem2d_t h)
{
    if (!h) return SELF_OK;
    (void)hipSetDevice(h->device);
    if (h->op.stream) (void)hipStreamSynchronize(h->op.stream);
    h->wk.release();
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    delete h;
    return SELF_OK;
}

int self_cgsem2d_global_dofs(self_cgsem2d_t h, int *nGlobal)
{
    SELF_REQUIRE(h && nGlobal, SELF_ERR_ARG, "null argument");
    *nGlobal = (int)h->op.nG;
    return SELF_OK;
}

int self_cgsem2d_set_flux_coefficient(self_cgsem2d_t h, const double *k)
try {
    SELF_REQUIRE(h && k, SELF_ERR_ARG, "null argument");
    SELF_HIP(hipSetDevice(h->device));
    SELF_HIP(hipMemcpy(h->op.kD, k, sizeof(double) * h->op.nLocal, hipMemcpyHostToDevice));
    h->op.kH.assign(k, k + h->op.nLocal);
    return device_geometry(h->op);
}
SELF_CATCH_ALL

int self_cgsem2d_set_preconditioner(self_cgsem2d_t h, const double *x, const double *y, int pcIterates,
                                    double pcTolerance, int coefficient)
try {
    SELF_REQUIRE(h && x && y, SELF_ERR_ARG, "null argument");
    SELF_REQUIRE(pcIterates >= 0 && pcTolerance > 0.0, SELF_ERR_ARG, "pcIterates %d, pcTolerance %g", pcIterates,
                 pcTolerance);
    SELF_REQUIRE(coefficient == 0 || coefficient == 1, SELF_ERR_ARG, "coefficient selector %d", coefficient);
    SELF_HIP(hipSetDevice(h->device));
    CgsemOp &op = h->op;
    SELF_HIP(hipStreamSynchronize(op.stream));
    auto pc = std::make_unique<FemPC>();
    const double *k = (coefficient == 1 && !op.kH.empty()) ? op.kH.data() : nullptr;
    int rc = pc->build(op.N, op.nE, x, y, k, op.gidH.data(), op.nG, pcIterates, pcTolerance);
    if (rc) return rc;
    op.bytes += pc->bytes;
    op.pc = std::move(pc);
    return SELF_OK;
}
SELF_CATCH_ALL

// CGsemPreconditioner_2D % Solve (:664-746) on a host array in the
// element-local layout: owner copies in, every copy out (UnMask +
// MapFromPCToOrig), Dirichlet nodes 0.  Returns the inner iterations.
int self_cgsem2d_apply_preconditioner(self_cgsem2d_t h, const double *b, double *x, int *iterations)
try {
    SELF_REQUIRE(h && b && x, SELF_ERR_ARG, "null argument");
    CgsemOp &op = h->op;
    SELF_REQUIRE(op.pc, SELF_ERR_STATE, "no preconditioner: call self_cgsem2d_set_preconditioner first");
    SELF_HIP(hipSetDevice(h->device));
    int rc = h->wk.ensure(op.nG, 0, 1, false);
    if (rc) return rc;
    std::vector<double> bg(op.nG), xg(op.nG);
    for (size_t g = 0; g < op.nG; ++g) bg[g] = b[op.ciH[op.csH[g]]];
    SELF_HIP(hipMemcpyAsync(h->wk.r, bg.data(), sizeof(double) * op.nG, hipMemcpyHostToDevice, op.stream));
    if ((rc = op.pc->apply(h->wk.r, h->wk.w, nullptr, op.stream))) return rc;
    SELF_HIP(hipMemcpyAsync(xg.data(), h->wk.w, sizeof(double) * op.nG, hipMemcpyDeviceToHost, op.stream));
    int its = 0;
    if ((rc = op.pc->last_iterations(op.stream, its))) return rc;
    for (size_t l = 0; l < op.nLocal; ++l) x[l] = op.gidH[l] >= 0 ? xg[op.gidH[l]] : 0.0;
    if (iterations) *iterations = its;
    return SELF_OK;
}
SELF_CATCH_ALL

int self_cgsem2d_solve_gmres_mixed(self_cgsem2d_t h, int precond, double tol, int maxCycles, int mInner, int ortho,
                                   int *ioerr, int *iterations, double *relResidual, double *elapsed_ms)
try {
    SELF_REQUIRE(h && ioerr && iterations, SELF_ERR_ARG, "null argument");
    SELF_REQUIRE(maxCycles >= 1 && tol > 0 && mInner >= 1 && mInner <= 200, SELF_ERR_ARG,
                 "maxCycles %d, tol %g, m %d", maxCycles, tol, mInner);
    SELF_REQUIRE(precond >= 0 && precond <= 2, SELF_ERR_ARG, "precond %d", precond);
    SELF_HIP(hipSetDevice(h->device));
    CgsemOp &op = h->op;
    KrylovResult res;
    SELF_HIP(hipEventRecord(h->ev0, op.stream));
    int rc = krylov_solve_gmres_mixed(op, h->wk, precond, tol, maxCycles, mInner, ortho, op.xg, res, op.stream);
    if (rc) return rc;
    SELF_HIP(hipEventRecord(h->ev1, op.stream));
    SELF_HIP(hipEventSynchronize(h->ev1));
    if (elapsed_ms) {
        float ms = 0;
        SELF_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
        *elapsed_ms = ms;
    }
    *ioerr = res.ioerr;
    *iterations = res.iterations;
    if (relResidual) *relResidual = res.r0 > 0 ? res.rNorm / res.r0 : 0.0;
    return SELF_OK;
}
SELF_CATCH_ALL

int self_cgsem2d_set_source(self_cgsem2d_t h, const double *s)
try {
    SELF_REQUIRE(h && s, SELF_ERR_ARG, "null argument");
    SELF_HIP(hipSetDevice(h->device));
    CgsemOp &op = h->op;
    std::vector<double> b(op.nLocal);
    for (size_t e = 0; e < op.nE; ++e)
        for (int j = 0; j < op.n1; ++j)
            for (int i = 0; i < op.n1; ++i) {
                const size_t l = e * op.np + i + op.n1 * j;
                b[l] = s[l] * op.met[4][l] * op.wH[i] * op.wH[j];  // source*J*ws*wp (:1052)
            }
    SELF_HIP(hipMemcpy(op.bl, b.data(), sizeof(double) * op.nLocal, hipMemcpyHostToDevice));
    return SELF_OK;
}
SELF_CATCH_ALL

int self_cgsem2d_set_dirichlet(self_cgsem2d_t h, const double *values)
try {
    SELF_REQUIRE(h && values, SELF_ERR_ARG, "null argument");
    SELF_HIP(hipSetDevice(h->device));
    CgsemOp &op = h->op;
    std::vector<double> u(op.nLocal, 0.0);
    for (size_t l = 0; l < op.nLocal; ++l)
        if (op.dirH[l]) u[l] = values[l];
    SELF_HIP(hipMemcpy(op.uD, u.data(), sizeof(double) * op.nLocal, hipMemcpyHostToDevice));
    return SELF_OK;
}
SELF_CATCH_ALL

// solution in the reference's local layout: owner copies are taken
int self_cgsem2d_set_solution(self_cgsem2d_t h, const double *u)
try {
    SELF_REQUIRE(h && u, SELF_ERR_ARG, "null argument");
    SELF_HIP(hipSetDevice(h->device));
    CgsemOp &op = h->op;
    std::vector<double> x(op.nG);
    for (size_t g = 0; g < op.nG; ++g) x[g] = u[op.ciH[op.csH[g]]];
    SELF_HIP(hipMemcpy(op.xg, x.data(), sizeof(double) * op.nG, hipMemcpyHostToDevice));
    return SELF_OK;
}
SELF_CATCH_ALL

// UnMask + SetDirichletBoundaryCondition: every copy holds its DOF's value
int self_cgsem2d_get_solution(self_cgsem2d_t h, double *u)
try {
    SELF_REQUIRE(h && u, SELF_ERR_ARG, "null argument");
    SELF_HIP(hipSetDevice(h->device));
    CgsemOp &op = h->op;
    SELF_HIP(hipStreamSynchronize(op.stream));
    std::vector<double> x(op.nG), d(op.nLocal);
    SELF_HIP(hipMemcpy(x.data(), op.xg, sizeof(double) * op.nG, hipMemcpyDeviceToHost));
    SELF_HIP(hipMemcpy(d.data(), op.uD, sizeof(double) * op.nLocal, hipMemcpyDeviceToHost));
    for (size_t l = 0; l < op.nLocal; ++l) u[l] = op.gidH[l] >= 0 ? x[op.gidH[l]] : d[l];
    return SELF_OK;
}
SELF_CATCH_ALL

// MatrixAction on a masked local array (host in, masked local out).
int self_cgsem2d_matrix_action(self_cgsem2d_t h, const double *u, double *Au)
try {
    SELF_REQUIRE(h && u && Au, SELF_ERR_ARG, "null argument");
    SELF_HIP(hipSetDevice(h->device));
    CgsemOp &op = h->op;
    int rc = h->wk.ensure(op.nG, 0, 1, false);
    if (rc) return rc;
    std::vector<double> x(op.nG), y(op.nG);
    for (size_t g = 0; g < op.nG; ++g) x[g] = u[op.ciH[op.csH[g]]];
    SELF_HIP(hipMemcpy(h->wk.v, x.data(), sizeof(double) * op.nG, hipMemcpyHostToDevice));
    if ((rc = op.matvec(h->wk.v, h->wk.z, nullptr, nullptr, kFinNone, h->wk.slot, op.stream))) return rc;
    SELF_HIP(hipStreamSynchronize(op.stream));
    SELF_HIP(hipMemcpy(y.data(), h->wk.z, sizeof(double) * op.nG, hipMemcpyDeviceToHost));
    std::memset(Au, 0, sizeof(double) * op.nLocal);
    for (size_t g = 0; g < op.nG; ++g) Au[op.ciH[op.csH[g]]] = y[g];
    return SELF_OK;
}
SELF_CATCH_ALL

// Mask (:350-430) on an element-local array: only the owner copy of each
// global DOF (edge: the primary element's; corner: the first element of the
// node's list) keeps its value; other copies and Dirichlet nodes become 0.
int self_cgsem2d_mask(self_cgsem2d_t h, double *u)
try {
    SELF_REQUIRE(h && u, SELF_ERR_ARG, "null argument");
    const CgsemOp &op = h->op;
    std::vector<char> keep(op.nLocal, 0);
    for (size_t g = 0; g < op.nG; ++g) keep[op.ciH[op.csH[g]]] = 1;
    for (size_t l = 0; l < op.nLocal; ++l)
        if (!keep[l]) u[l] = 0.0;
    return SELF_OK;
}
SELF_CATCH_ALL

int self_cgsem2d_dof_map(self_cgsem2d_t h, int *gid)
try {
    SELF_REQUIRE(h && gid, SELF_ERR_ARG, "null argument");
    std::memcpy(gid, h->op.gidH.data(), sizeof(int) * h->op.nLocal);
    return SELF_OK;
}
SELF_CATCH_ALL

int self_cgsem2d_pc_matrix(self_cgsem2d_t h, int *width, int *cols, double *vals)
try {
    SELF_REQUIRE(h && width, SELF_ERR_ARG, "null argument");
    SELF_REQUIRE(h->op.pc, SELF_ERR_STATE, "no preconditioner set (self_cgsem2d_set_preconditioner)");
    const FemPC &pc = *h->op.pc;
    *width = pc.width;
    if (!cols || !vals) return SELF_OK;
    SELF_HIP(hipSetDevice(h->device));
    const size_t ne = (size_t)pc.width * pc.n;
    std::vector<int> c(ne);
    std::vector<double> v(ne);
    SELF_HIP(hipMemcpy(c.data(), pc.col, sizeof(int) * ne, hipMemcpyDeviceToHost));
    SELF_HIP(hipMemcpy(v.data(), pc.val, sizeof(double) * ne, hipMemcpyDeviceToHost));
    for (size_t g = 0; g < pc.n; ++g)  // device slot-major -> row-major
        for (int q = 0; q < pc.width; ++q) {
            cols[g * pc.width + q] = c[(size_t)q * pc.n + g];
            vals[g * pc.width + q] = v[(size_t)q * pc.n + g];
        }
    return SELF_OK;
}
SELF_CATCH_ALL

// Residual of the current solution, masked local layout.
int self_cgsem2d_residual(self_cgsem2d_t h, double *r)
try {
    SELF_REQUIRE(h && r, SELF_ERR_ARG, "null argument");
    SELF_HIP(hipSetDevice(h->device));
    CgsemOp &op = h->op;
    int rc = h->wk.ensure(op.nG, 0, 1, false);
    if (rc) return rc;
    if ((rc = op.residual(op.xg, h->wk.r, nullptr, kFinNone, h->wk.slot, op.stream))) return rc;
    SELF_HIP(hipStreamSynchronize(op.stream));
    std::vector<double> y(op.nG);
    SELF_HIP(hipMemcpy(y.data(), h->wk.r, sizeof(double) * op.nG, hipMemcpyDeviceToHost));
    std::memset(r, 0, sizeof(double) * op.nLocal);
    for (size_t g = 0; g < op.nG; ++g) r[op.ciH[op.csH[g]]] = y[g];
    return SELF_OK;
}
SELF_CATCH_ALL

int self_cgsem2d_solve(self_cgsem2d_t h, int method, int precond, double tol, int maxIters, int mInner, int ortho,
                       int *ioerr, int *iterations, double *resi, int resiLen, double *elapsed_ms)
try {
    SELF_REQUIRE(h && ioerr && iterations, SELF_ERR_ARG, "null argument");
    SELF_REQUIRE(maxIters >= 1 && tol > 0, SELF_ERR_ARG, "maxIters/tolerance");
    SELF_REQUIRE(method >= 0 && method <= 4, SELF_ERR_ARG, "unknown method %d", method);
    SELF_REQUIRE(!(method == 2 || method == 3) || (mInner >= 1 && mInner <= 200), SELF_ERR_ARG, "GMRES m = %d", mInner);
    SELF_HIP(hipSetDevice(h->device));
    CgsemOp &op = h->op;
    KrylovResult res;
    SELF_HIP(hipEventRecord(h->ev0, op.stream));
    int rc = krylov_solve(op, h->wk, method, precond, tol, maxIters, mInner, ortho, op.xg, resi, resiLen, res,
                          op.stream);
    if (rc) return rc;
    SELF_HIP(hipEventRecord(h->ev1, op.stream));
    SELF_HIP(hipEventSynchronize(h->ev1));
    if (elapsed_ms) {
        float ms = 0;
        SELF_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
        *elapsed_ms = ms;
    }
    *ioerr = res.ioerr;
    *iterations = res.iterations;
    return SELF_OK;
}
SELF_CATCH_ALL

int self_cgsem2d_solve_batched(self_cgsem2d_t *problems, int nProblems, int precond, double tol, int maxIters,
                               int *ioerr, int *iterations, double *elapsed_ms)
try {
    SELF_REQUIRE(problems && ioerr && iterations, SELF_ERR_ARG, "null argument");
    SELF_REQUIRE(nProblems >= 1 && nProblems <= kCgbMaxProblems, SELF_ERR_ARG, "nProblems = %d (1..%d)", nProblems,
                 kCgbMaxProblems);
    SELF_REQUIRE(maxIters >= 1 && tol > 0, SELF_ERR_ARG, "maxIters/tolerance");
    SELF_REQUIRE(precond == 0 || precond == 1, SELF_ERR_ARG, "precond = %d (0 none, 1 Jacobi)", precond);
    std::vector<self_cgsem2d_t> hs(problems, problems + nProblems);
    for (int p = 0; p < nProblems; ++p) {
        SELF_REQUIRE(hs[p], SELF_ERR_ARG, "problem %d is null", p);
        SELF_REQUIRE(hs[p]->device == hs[0]->device, SELF_ERR_ARG, "problem %d is on another device", p);
    }
    {
        std::vector<self_cgsem2d_t> sorted(hs);
        std::sort(sorted.begin(), sorted.end());
        SELF_REQUIRE(std::adjacent_find(sorted.begin(), sorted.end()) == sorted.end(), SELF_ERR_ARG,
                     "a problem is listed twice");
    }
    SELF_HIP(hipSetDevice(hs[0]->device));
    self_cgsem2d_s *h0 = hs[0];
    const hipStream_t s = h0->op.stream;
    for (auto h : hs) SELF_HIP(hipStreamSynchronize(h->op.stream));
    // device tables: descriptors, then (per chunk) the running problems and
    // the prefix sums of their direction and update workgroups, done flags
    const size_t descBytes = sizeof(CgBatchProb) * nProblems;
    const size_t tblInts = (size_t)3 * nProblems + 2 + nProblems;
    char *tables = nullptr;
    SELF_HIP(hipMalloc(&tables, descBytes + sizeof(int) * tblInts));
    struct Release {
        char *p;
        ~Release() { (void)hipFree(p); }
    } release{tables};
    CgBatchProb *probsD = reinterpret_cast<CgBatchProb *>(tables);
    int *tblD = reinterpret_cast<int *>(tables + descBytes);
    int *flagsD = tblD + 3 * nProblems + 2;
    SELF_HIP(hipEventRecord(h0->ev0, s));
    std::vector<CgBatchProb> P(nProblems);
    for (int p = 0; p < nProblems; ++p) {
        self_cgsem2d_s *h = hs[p];
        CgsemOp &op = h->op;
        int rc = h->wk.ensure(op.n, 0, maxIters + 1, false);
        if (!rc) rc = krylov_cg_begin(op, h->wk, precond == 1, kStopNormRel, tol, maxIters, op.xg, s);
        if (rc) return rc;
        // update partials of workgroups without DOFs: zero throughout
        hipLaunchKernelGGL(k_fill_f64, dim3(8), dim3(256), 0, s, h->wk.slot.partials + kRedBlocks,
                           (size_t)2 * kRedBlocks, 0.0);
        CgBatchProb &q = P[p];
        q.gid = op.gid;
        q.cs = op.cs;
        q.ci = op.ci;
        q.G3 = op.G3;
        q.D = op.D;
        q.dinv = precond == 1 ? op.dinv : nullptr;
        q.Lbuf = op.Lbuf;
        q.u = h->wk.u;
        q.r = h->wk.r;
        q.w = h->wk.w;
        q.vA = h->wk.v;
        q.vB = h->wk.vB;
        q.st = h->wk.st;
        q.slot = h->wk.slot;
        q.nG = op.nG;
        q.N = op.N;
        q.nE = (int)op.nE;
        q.nPart = 0;
        CG_DISPATCH_N(op.N, q.nPart = std::min(cg_grid<NN>(op.nE), kRedBlocks));
        q.nUpd = (int)std::min<size_t>(kFoldParts, (op.nG + kFoldThreads - 1) / kFoldThreads);
        SELF_REQUIRE(q.nPart > 0, SELF_ERR_ARG, "problem %d: N = %d", p, op.N);
    }
    SELF_HIP(hipGetLastError());
    SELF_HIP(hipMemcpyAsync(probsD, P.data(), descBytes, hipMemcpyHostToDevice, s));
    std::vector<int> act(nProblems), flags(nProblems);
    for (int p = 0; p < nProblems; ++p) act[p] = p;
    int nAct = nProblems;
    for (int it0 = 1; it0 <= maxIters && nAct > 0; it0 += kCgbChunk) {
        const int it1 = std::min(it0 + kCgbChunk - 1, maxIters);
        // running list | direction prefix (nAct + 1) | update prefix (nAct + 1)
        std::vector<int> tb((size_t)3 * nAct + 2);
        int *dirS = tb.data() + nAct, *updS = dirS + nAct + 1;
        dirS[0] = updS[0] = 0;
        for (int k = 0; k < nAct; ++k) {
            tb[k] = act[k];
            dirS[k + 1] = dirS[k] + P[act[k]].nPart;
            updS[k + 1] = updS[k] + P[act[k]].nUpd;
        }
        SELF_HIP(hipMemcpyAsync(tblD, tb.data(), sizeof(int) * tb.size(), hipMemcpyHostToDevice, s));
        const int *actD = tblD, *dirD = tblD + nAct, *updD = dirD + nAct + 1;
        for (int it = it0; it <= it1; ++it) {
            hipLaunchKernelGGL(k_cgb_direction, dim3(dirS[nAct]), dim3(256), 0, s, (const CgBatchProb *)probsD, actD,
                               dirD, nAct, it, (int)(it > it0));
            hipLaunchKernelGGL(k_cgb_update, dim3(updS[nAct]), dim3(kFoldThreads), 0, s,
                               (const CgBatchProb *)probsD, actD, updD, nAct, it);
        }
        hipLaunchKernelGGL(k_cgb_fold_end, dim3(nAct), dim3(64), 0, s, (const CgBatchProb *)probsD, actD, it1,
                           flagsD);
        SELF_HIP(hipGetLastError());
        SELF_HIP(hipMemcpyAsync(flags.data(), flagsD, sizeof(int) * nAct, hipMemcpyDeviceToHost, s));
        SELF_HIP(hipStreamSynchronize(s));
        int m = 0;
        for (int k = 0; k < nAct; ++k)
            if (!flags[k]) act[m++] = act[k];
        nAct = m;
    }
    for (int p = 0; p < nProblems; ++p) {
        KrylovResult res;
        const int rc = krylov_cg_finish(hs[p]->op, hs[p]->wk, kStopNormRel, tol, hs[p]->op.xg, res, s);
        if (rc) return rc;
        ioerr[p] = res.ioerr;
        iterations[p] = res.iterations;
    }
    SELF_HIP(hipEventRecord(h0->ev1, s));
    SELF_HIP(hipEventSynchronize(h0->ev1));
    if (elapsed_ms) {
        float ms = 0;
        SELF_HIP(hipEventElapsedTime(&ms, h0->ev0, h0->ev1));
        *elapsed_ms = ms;
    }
    return SELF_OK;
}
SELF_CATCH_ALL

int self_cgsem2d_solve_mixed(self_cgsem2d_t h, int precond, double tol, double innerTol, int maxInner, int maxOuter,
                             int *ioerr, int *outerIters, int *innerIters, double *relResidual, double *elapsed_ms)
try {
    SELF_REQUIRE(h && ioerr && outerIters && innerIters, SELF_ERR_ARG, "null argument");
    SELF_HIP(hipSetDevice(h->device));
    CgsemOp &op = h->op;
    KrylovResult res;
    int inner = 0;
    SELF_HIP(hipEventRecord(h->ev0, op.stream));
    int rc = krylov_solve_ir(op, h->wk, precond, tol, innerTol, maxInner, maxOuter, op.xg, res, inner, op.stream);
    if (rc) return rc;
    SELF_HIP(hipEventRecord(h->ev1, op.stream));
    SELF_HIP(hipEventSynchronize(h->ev1));
    if (elapsed_ms) {
        float ms = 0;
        SELF_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
        *elapsed_ms = ms;
    }
    *ioerr = res.ioerr;
    *outerIters = res.iterations;
    *innerIters = inner;
    if (relResidual) *relResidual = res.r0 > 0 ? res.rNorm / res.r0 : 0.0;
    return SELF_OK;
}
SELF_CATCH_ALL

namespace {
int solve_reliable_entry(self_cgsem2d_t h, int precond, double tol, double delta, int maxIters, int maxUpdates,
                         int *ioerr, int *updates, int *iterations, double *relResidual, double *elapsed_ms,
                         bool innerFP64);
}

int self_cgsem2d_solve_reliable(self_cgsem2d_t h, int precond, double tol, double delta, int maxIters,
                                int maxUpdates, int *ioerr, int *updates, int *iterations, double *relResidual,
                                double *elapsed_ms)
{
    return solve_reliable_entry(h, precond, tol, delta, maxIters, maxUpdates, ioerr, updates, iterations,
                                relResidual, elapsed_ms, false);
}

int self_cgsem2d_solve_reliable_fp64(self_cgsem2d_t h, int precond, double tol, double delta, int maxIters,
                                     int maxUpdates, int *ioerr, int *updates, int *iterations, double *relResidual,
                                     double *elapsed_ms)
{
    return solve_reliable_entry(h, precond, tol, delta, maxIters, maxUpdates, ioerr, updates, iterations,
                                relResidual, elapsed_ms, true);
}

}  // extern "C"

namespace {
int solve_reliable_entry(self_cgsem2d_t h, int precond, double tol, double delta, int maxIters, int maxUpdates,
                         int *ioerr, int *updates, int *iterations, double *relResidual, double *elapsed_ms,
                         bool innerFP64)
try {
    SELF_REQUIRE(h && ioerr && updates && iterations, SELF_ERR_ARG, "null argument");
    SELF_REQUIRE(maxIters >= 1 && maxUpdates >= 1 && tol > 0.0, SELF_ERR_ARG,
                 "need maxIters >= 1, maxUpdates >= 1 and tol > 0");
    SELF_HIP(hipSetDevice(h->device));
    CgsemOp &op = h->op;
    KrylovResult res;
    int its = 0, ups = 0;
    SELF_HIP(hipEventRecord(h->ev0, op.stream));
    int rc = krylov_solve_reliable(op, h->wk, precond, tol, delta, maxIters, maxUpdates, op.xg, res, its, ups,
                                   op.stream, innerFP64);
    if (rc) return rc;
    SELF_HIP(hipEventRecord(h->ev1, op.stream));
    SELF_HIP(hipEventSynchronize(h->ev1));
    if (elapsed_ms) {
        float ms = 0;
        SELF_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
        *elapsed_ms = ms;
    }
    *ioerr = res.ioerr;
    *updates = ups;
    *iterations = its;
    if (relResidual) *relResidual = res.r0 > 0 ? res.rNorm / res.r0 : 0.0;
    return SELF_OK;
}
SELF_CATCH_ALL

}  // namespace
