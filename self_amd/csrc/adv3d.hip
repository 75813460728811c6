// adv3d.hip -- 3-D linear advection DGSEM with the adaptive roll-off filter
// on gfx950.
//
// Reference: src/highend/advection3d/Advection3DClass.f90
//   ForwardStepRK3 :1000-1045, GlobalTimeDerivative :1049-1121,
//   DoTheAdaptiveFiltering :1125-1233, EdgeFlux :1237-1341,
//   MappedTimeDerivative :1345-1562, RiemannSolver :1595-1620,
//   Source :1676-1694, GetExternalState :1698-1735;
//   src/filters/RollOffFilter3D_Class.f90 (filter), and
//   src/dgsem/DGSEMSolutionStorageClass_3D.f90:833-910 (traces).
//
// One workgroup holds one element at N = 7 (512 nodes, one node per
// thread) or several smaller ones.  A stage is one kernel:
//   volume fluxes c u projected on the cofactor rows -> LDS
//   upwind flux of every face node (own trace from LDS, neighbour trace
//     from the previous stage's trace buffer, u.n and |n| of the face's
//     primary side precomputed because velocity is static) -> LDS
//   weak divergence in s, p, q + lift, relaxation source, RK3 update
//   adaptive filter for the next stage (three 1-D passes of the roll-off
//     filter's direction factor, energies of the resolved and marginal
//     parts, the reference's limiter branches) when the next stage follows
//     inside the same call
//   traces of the new state for the next stage.
// The filter the reference applies at the start of a stage's
// GlobalTimeDerivative is applied at the end of the stage before it, so
// the next stage reads a filtered state and its traces without another
// pass.  The first stage of a call runs a prologue kernel (filter +
// traces) instead.

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "dg2d_device.hpp"
#include "dg3d_device.hpp"
#include "mfma.hpp"

namespace self {
void cofactor3(const double *m, int dir, double out[3]);

namespace {

constexpr double kRkA[3] = {0.0, -5.0 / 9.0, -153.0 / 128.0};
constexpr double kRkB[3] = {0.0, 1.0 / 3.0, 3.0 / 4.0};
constexpr double kRkG[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};

template <int N>
struct Shape3 {
    static constexpr int n1 = N + 1, n2 = n1 * n1, np = n2 * n1;
    static constexpr int epb = (512 / np) > 0 ? 512 / np : 1;
    static constexpr int threads = ((epb * np + 63) / 64) * 64;
};

// nodal constants: weak derivative matrix (column-major), 1/w, w, l(-1), l(+1),
// the filter's direction factor (column-major)
template <int N>
struct Nodal3 {
    static constexpr int n1 = N + 1;
    double D[n1 * n1], iw[n1], w[n1], lL[n1], lR[n1], F1[n1 * n1];
};

// global nodal buffer layout: D, w, lL, lR, F1
template <int N>
__device__ __forceinline__ void stage_nodal3(Nodal3<N> &L, const double *__restrict__ g)
{
    constexpr int n1 = N + 1;
    for (int k = threadIdx.x; k < n1 * n1; k += blockDim.x) {
        L.D[k] = g[k];
        L.F1[k] = g[n1 * n1 + 3 * n1 + k];
    }
    for (int k = threadIdx.x; k < n1; k += blockDim.x) {
        const double w = g[n1 * n1 + k];
        L.w[k] = w;
        L.iw[k] = 1.0 / w;
        L.lL[k] = g[n1 * n1 + n1 + k];
        L.lR[k] = g[n1 * n1 + 2 * n1 + k];
    }
}

// The same constants in two steps around a kernel's element loads:
// nodal3_fetch issues this thread's loads (clamped indices, no branch)
// before them, nodal3_put writes LDS after them, so the write waits for these
// loads only (the counter of outstanding loads is in order) and not for the
// element loads.  LANES threads, lane = 0 .. LANES-1.
template <int N, int LANES>
struct Nodal3Regs {
    static constexpr int n1 = N + 1, nv = 2 * n1 * n1 + 3 * n1;  // D, F1, then w, lL, lR
    static constexpr int K = (nv + LANES - 1) / LANES;
    double v[K];
};

template <int N, int LANES>
__device__ __forceinline__ void nodal3_fetch(Nodal3Regs<N, LANES> &R, const double *__restrict__ g, int lane)
{
    using Rg = Nodal3Regs<N, LANES>;
    constexpr int n1 = Rg::n1, n2 = n1 * n1;
    #pragma unroll
    for (int q = 0; q < Rg::K; ++q) {
        const int k = min(lane + q * LANES, Rg::nv - 1);
        // buffer: D (n2), w, lL, lR (n1 each), F1 (n2); here D, F1, w, lL, lR
        const int src = k < n2 ? k : (k < 2 * n2 ? k + 2 * n1 + n1 : k - n2);
        R.v[q] = g[src];
    }
}

template <int N, int LANES>
__device__ __forceinline__ void nodal3_put(Nodal3<N> &L, const Nodal3Regs<N, LANES> &R, int lane)
{
    using Rg = Nodal3Regs<N, LANES>;
    constexpr int n1 = Rg::n1, n2 = n1 * n1;
    #pragma unroll
    for (int q = 0; q < Rg::K; ++q) {
        const int k = lane + q * LANES;
        const double v = R.v[q];
        if (k < n2) L.D[k] = v;
        else if (k < 2 * n2) L.F1[k - n2] = v;
        else if (k < 2 * n2 + n1) { L.w[k - 2 * n2] = v; L.iw[k - 2 * n2] = 1.0 / v; }
        else if (k < 2 * n2 + 2 * n1) L.lL[k - 2 * n2 - n1] = v;
        else if (k < Rg::nv) L.lR[k - 2 * n2 - 2 * n1] = v;
    }
}

// trace of one element's nodal field q (LDS, i fastest) at side s, face node (a, b)
template <int N>
__device__ __forceinline__ double trace3(const Nodal3<N> &L, const double *q, int s, int a, int b)
{
    return face_trace3<N>(L.lL, L.lR, q, s, a, b);
}

struct Adv3Args {
    int nE, mode, useFilter, hasRelax;
    int stage;       // stage whose filter runs in the prologue (1..3)
    int filterNext;  // stage index of the filter applied after the update (0: none)
    double t, a, gdt;
    const double *nodal;
    double *c, *G, *tend;         // [nq][nE][np]
    const double *vel;            // [3][nE][np]
    const double *acoef;          // [3][nE][np]: u . cofactor row s, p, q (the static velocity folded in)
    const double *geo;            // [10][nE][np]: 9 cofactors (s row, p row, q row), J
    const double *fdat;           // [nE][6][2][n2]: u.n and |n| of the face's primary side
    const double *bpos;           // [slot][3][n2]
    const int2 *side;             // [nE][6]
    const double *trIn;
    double *trOut;                // [nE][6][nq][n2]
    const double *relax, *rFac;
    double *E1, *E2, *lim;        // [nq][nE], [3][nq][nE]
    unsigned *nFiltered;
};

__device__ __forceinline__ double adv3_external(int bc, int f, const double *bp, int n2, int node, double t)
{
    if (bc == -103) return f == 0 ? tanh(t) : 0.0;   // InflowOne
    if (bc == -104) return f == 1 ? tanh(t) : 0.0;   // InflowTwo
    if (bc == -100) {                                // NO_NORMAL_FLOW
        const double x = bp[node] - t, y = bp[n2 + node] - t, z = bp[2 * n2 + node] - t;
        return exp(-(x * x + y * y + z * z) / (0.2 * 0.2));
    }
    return 0.0;
}

// The filter's stored energies and limiters of one element (ec) for the
// filter of `stage` (0: none), loaded by every thread of the element (one
// address per wavefront) before the filter starts, so the thread that
// updates them does not wait for them between the filter's barriers.
// Only this element's workgroup writes them, after reading.
template <int NQ>
struct Adv3FilterPre {
    double E1[NQ], E2[NQ], lim[NQ], limPrev[NQ];
};

template <int NQ>
__device__ __forceinline__ void adv3_filter_prefetch(const Adv3Args &A, int ec, int stage, Adv3FilterPre<NQ> &P)
{
    #pragma unroll
    for (int f = 0; f < NQ; ++f) { P.E1[f] = 0.0; P.E2[f] = 0.0; P.lim[f] = 0.0; P.limPrev[f] = 0.0; }
    if (stage == 0) return;
    #pragma unroll
    for (int f = 0; f < NQ; ++f) {
        const size_t ei = (size_t)f * A.nE + ec;
        const size_t li = ((size_t)(stage - 1) * NQ + f) * A.nE + ec;
        P.E1[f] = A.E1[ei];
        P.E2[f] = A.E2[ei];
        P.lim[f] = A.lim[li];
        if (stage > 1) P.limPrev[f] = A.lim[li - (size_t)NQ * A.nE];
    }
}

// DoTheAdaptiveFiltering for the element group's state in cs (LDS) at the
// given stage; t1, t2, t3 are LDS scratch of the size of cs; red holds
// 2 n2 per element.  Replaces cs (and the register copy c) when the
// limiter asks for it.  Every thread of the block must call it.  pre: the
// element's stored energies and limiters loaded in advance (or null).
template <int N, int NQ>
__device__ void adaptive_filter(const Adv3Args &A, const Nodal3<N> &L, double *cs, double *t1, double *t2,
                                double *t3, double *red, int *flag, int le, int nd, int e, bool act, double J,
                                double (&c)[NQ], int stage, const Adv3FilterPre<NQ> *pre = nullptr)
{
    using S = Shape3<N>;
    constexpr int n1 = S::n1, n2 = S::n2, np = S::np;
    const int i = nd % n1, j = (nd / n1) % n1, k = nd / n2;
    for (int f = 0; f < NQ; ++f) {
        const int o = (le * NQ + f) * np;
        double cf = 0.0;
        if (act) {
            double acc = 0.0;
            #pragma unroll
            for (int m = 0; m < n1; ++m) acc = __builtin_fma(L.F1[i + n1 * m], cs[o + m + n1 * (j + n1 * k)], acc);
            t1[o + nd] = acc;
        }
        __syncthreads();
        if (act) {
            double acc = 0.0;
            #pragma unroll
            for (int m = 0; m < n1; ++m) acc = __builtin_fma(L.F1[j + n1 * m], t1[o + i + n1 * (m + n1 * k)], acc);
            t2[o + nd] = acc;
        }
        __syncthreads();
        if (act) {
            double acc = 0.0;
            #pragma unroll
            for (int m = 0; m < n1; ++m) acc = __builtin_fma(L.F1[k + n1 * m], t2[o + i + n1 * (j + n1 * m)], acc);
            cf = acc;
            const double d = cs[o + nd] - cf;
            t1[o + nd] = (cf * cf) * J;  // resolved energy density
            t3[o + nd] = (d * d) * J;    // marginally resolved energy density
        }
        __syncthreads();
        // line integrals in s, weighted by w_p w_q (the reference's order)
        if (act && nd < n2) {
            const int jj = nd % n1, kk = nd / n1;
            double d1 = 0.0, d2 = 0.0;
            #pragma unroll
            for (int m = 0; m < n1; ++m) {
                d1 = __builtin_fma(t1[o + m + n1 * nd], L.w[m], d1);
                d2 = __builtin_fma(t3[o + m + n1 * nd], L.w[m], d2);
            }
            red[le * 2 * n2 + nd] = d1 * L.w[jj] * L.w[kk];
            red[le * 2 * n2 + n2 + nd] = d2 * L.w[jj] * L.w[kk];
        }
        __syncthreads();
        if (act && nd == 0) {
            double E1 = 0.0, E2 = 0.0;
            for (int q = 0; q < n2; ++q) {
                E1 = E1 + red[le * 2 * n2 + q];
                E2 = E2 + red[le * 2 * n2 + n2 + q];
            }
            const size_t ei = (size_t)f * A.nE + e;
            const double E1o = pre ? pre->E1[f] : A.E1[ei], E2o = pre ? pre->E2[f] : A.E2[ei];
            const double dE1 = E1 - E1o, dE2 = E2 - E2o, xi = E2 / E1;
            const size_t li = ((size_t)(stage - 1) * NQ + f) * A.nE + e;
            double lim = pre ? pre->lim[f] : A.lim[li];
            if (dE2 > 0.0 && fabs(dE1) / dE2 > 1.0) lim = xi;
            else if (dE2 < 0.0) lim = xi;
            else if (stage > 1) lim = pre ? pre->limPrev[f] : A.lim[li - (size_t)NQ * A.nE];
            A.lim[li] = lim;
            A.E1[ei] = E1;
            A.E2[ei] = E2;
            const int fl = (xi > 1.05 * lim) ? 1 : 0;
            flag[le * NQ + f] = fl;
            if (fl && A.nFiltered) atomicAdd(A.nFiltered, 1u);
        }
        __syncthreads();
        if (act && flag[le * NQ + f]) {
            cs[o + nd] = cf;
            c[f] = cf;
        }
    }
    __syncthreads();
}

// traces of the group's states (LDS) to trOut
template <int N, int NQ>
__device__ __forceinline__ void write_traces(const Adv3Args &A, const Nodal3<N> &L, const double *cs)
{
    using S = Shape3<N>;
    constexpr int n1 = S::n1, n2 = S::n2, np = S::np, epb = S::epb;
    for (int w = threadIdx.x; w < epb * 6 * n2; w += blockDim.x) {
        const int le = w / (6 * n2), rem = w - le * 6 * n2, sd = rem / n2, node = rem - sd * n2;
        const int e = blockIdx.x * epb + le;
        if (e >= A.nE) continue;
        const int a = node % n1, b = node / n1;
        for (int f = 0; f < NQ; ++f)
            A.trOut[(((size_t)e * 6 + sd) * NQ + f) * n2 + node] = trace3<N>(L, cs + (le * NQ + f) * np, sd, a, b);
    }
}

template <int N, int NQ>
__global__ __launch_bounds__(Shape3<N>::threads) void k_adv3d_traces(const double *__restrict__ q,
                                                                     const double *__restrict__ nodal,
                                                                     double *__restrict__ out, int nE)
{
    using S = Shape3<N>;
    constexpr int np = S::np, epb = S::epb;
    __shared__ Nodal3<N> L;
    __shared__ double qs[epb * NQ * np];
    stage_nodal3<N>(L, nodal);
    const int t = threadIdx.x, le = t / np, nd = t - le * np;
    const int e = blockIdx.x * epb + le;
    if (le < epb && e < nE)
        for (int f = 0; f < NQ; ++f) qs[(le * NQ + f) * np + nd] = q[((size_t)f * nE + e) * np + nd];
    __syncthreads();
    Adv3Args A{};
    A.nE = nE;
    A.trOut = out;
    write_traces<N, NQ>(A, L, qs);
}

// prologue: adaptive filter of stage A.stage (when enabled) + traces
template <int N, int NQ>
__global__ __launch_bounds__(Shape3<N>::threads) void k_adv3d_prologue(Adv3Args A)
{
    using S = Shape3<N>;
    constexpr int n2 = S::n2, np = S::np, epb = S::epb;
    __shared__ Nodal3<N> L;
    __shared__ double cs[epb * NQ * np], t1[epb * NQ * np], t2[epb * NQ * np], t3[epb * NQ * np];
    __shared__ double red[epb * 2 * n2];
    __shared__ int flag[epb * NQ];
    stage_nodal3<N>(L, A.nodal);
    const int t = threadIdx.x, le = t / np, nd = t - le * np;
    const int e = blockIdx.x * epb + le;
    const bool act = le < epb && e < A.nE;
    const size_t ms = (size_t)A.nE * np, base = (size_t)e * np + nd;
    double c[NQ], J = 1.0;
    if (act) {
        for (int f = 0; f < NQ; ++f) {
            c[f] = A.c[f * ms + base];
            cs[(le * NQ + f) * np + nd] = c[f];
        }
        J = A.geo[9 * ms + base];
    }
    __syncthreads();
    if (A.useFilter) {
        adaptive_filter<N, NQ>(A, L, cs, t1, t2, t3, red, flag, le, nd, e, act, J, c, A.stage);
        if (act)
            for (int f = 0; f < NQ; ++f) A.c[f * ms + base] = c[f];
    }
    write_traces<N, NQ>(A, L, cs);
}

// PREC 0: FP64 contractions; 1: FP32 matrix-core contractions; 2: split-BF16
// matrix-core contractions (mfma.hpp).  State, register, geometry, face
// fluxes, lift, Jacobian and the filter are FP64 in every form.
template <int N, int NQ, int PREC>
__global__ __launch_bounds__(Shape3<N>::threads, 4) void k_adv3d_stage(Adv3Args A)
{
    using S = Shape3<N>;
    constexpr int n1 = S::n1, n2 = S::n2, np = S::np, epb = S::epb;
    constexpr int nT = PREC ? epb * NQ * 3 * np : 1;
    __shared__ Nodal3<N> L;
    __shared__ double cs[epb * NQ * np], Fs[epb * NQ * np], Fp[epb * NQ * np], Fq[epb * NQ * np];
    __shared__ double bf[epb * 6 * NQ * n2];
    __shared__ double red[epb * 2 * n2];
    __shared__ int flag[epb * NQ];
    __shared__ float Tm[nT];
    __shared__ float Df[64];
    __shared__ double rs[n1];
    constexpr int nF = epb * 6 * n2;                       // face nodes of the element group
    constexpr int FQ = (nF + S::threads - 1) / S::threads;  // per thread
    const int t = threadIdx.x, le = t / np, nd = t - le * np;
    const int e = blockIdx.x * epb + le;
    const bool act = le < epb && e < A.nE;
    const int i = nd % n1, j = (nd / n1) % n1, k = nd / n2;
    const size_t ms = (size_t)A.nE * np, base = act ? (size_t)e * np + nd : 0;
    const int ec = act ? e : 0;
    // Loads first, in the order they are used, with no thread-dependent
    // branch around them (a thread without a node or face node loads from a
    // clamped valid address and ignores the value): the nodal constants,
    // the face data, the element's state and geometry, the neighbours'
    // traces (addressed by the face data), the filter's energies.
    Nodal3Regs<N, S::threads> R;
    nodal3_fetch<N, S::threads>(R, A.nodal, t);
    int2 finfo[FQ];
    double fuN[FQ], flen[FQ], fnbr[FQ][NQ];
    #pragma unroll
    for (int q = 0; q < FQ; ++q) {
        const int w = t + q * S::threads;
        const int wc = w < nF ? w : 0;
        const int l2 = wc / (6 * n2), rem = wc - l2 * 6 * n2, sd = rem / n2, node = rem - sd * n2;
        const int e2 = blockIdx.x * epb + l2, e2c = e2 < A.nE ? e2 : 0;
        finfo[q] = A.side[(size_t)e2c * 6 + sd];
        const size_t fb = (((size_t)e2c * 6 + sd) * 2) * n2 + node;
        fuN[q] = A.fdat[fb];
        flen[q] = A.fdat[fb + n2];
    }
    double c[NQ], gold[NQ], rl[NQ], rf[NQ];
    const double J = A.geo[9 * ms + base];
    // X/Y/ZFlux (:1624-1672) projected on the cofactor rows, with the
    // static velocity folded into the rows once
    const double as = A.acoef[base], ap = A.acoef[ms + base], aq = A.acoef[2 * ms + base];
    #pragma unroll
    for (int f = 0; f < NQ; ++f) c[f] = A.c[f * ms + base];
    #pragma unroll
    for (int f = 0; f < NQ; ++f) { gold[f] = 0.0; rl[f] = 0.0; rf[f] = 0.0; }
    asm volatile("" ::: "memory");  // the optional fields after the others
    if (A.mode == 0 && A.a != 0.0) {
        #pragma unroll
        for (int f = 0; f < NQ; ++f) gold[f] = A.G[f * ms + base];
    }
    if (A.hasRelax) {
        #pragma unroll
        for (int f = 0; f < NQ; ++f) { rl[f] = A.relax[f * ms + base]; rf[f] = A.rFac[f * ms + base]; }
    }
    #pragma unroll
    for (int q = 0; q < FQ; ++q) {
        const int w = t + q * S::threads;
        const int wc = w < nF ? w : 0;
        const int rem = wc % (6 * n2), node = rem % n2;
        const int2 info = finfo[q];
        int na = node % n1, nb = node / n1;
        if (info.x >= 0) s3_map<N>(info.y, node % n1, node / n1, na, nb);
        // a boundary face node reads element 0's slot and uses its external state instead
        const size_t nbase = info.x >= 0 ? ((size_t)info.x * 6 + s3_nbr_side(info.y)) * NQ : 0;
        #pragma unroll
        for (int f = 0; f < NQ; ++f) fnbr[q][f] = A.trIn[(nbase + f) * n2 + na + n1 * nb];
    }
    nodal3_put<N, S::threads>(L, R, t);
    if (PREC) {
        for (int q = threadIdx.x; q < 64; q += blockDim.x) {
            const int r = q >> 3, m = q & 7;
            Df[q] = (r < n1 && m < n1) ? (float)A.nodal[r + n1 * m] : 0.0f;
        }
        for (int r = threadIdx.x; r < n1; r += blockDim.x) {
            double acc = 0.0;
            for (int m = 0; m < n1; ++m) acc += A.nodal[r + n1 * m];
            rs[r] = acc;
        }
    }
    if (act) {
        for (int f = 0; f < NQ; ++f) {
            const int o = (le * NQ + f) * np + nd;
            cs[o] = c[f];
            Fs[o] = as * c[f];
            Fp[o] = ap * c[f];
            Fq[o] = aq * c[f];
        }
    }
    __syncthreads();
    if (PREC) {  // matrix-core contractions of the volume fluxes, one 16x16 tile per wavefront
        const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6, nTiles = epb * NQ * 3 * 2;
        for (int tile = wave; tile < nTiles; tile += nw) {
            const int lf = tile / 6, dir = (tile % 6) >> 1, t2 = tile & 1;
            if (blockIdx.x * epb + lf / NQ >= A.nE || 32 * t2 >= n2) continue;  // wave-uniform
            const double *X = (dir == 0 ? Fs : dir == 1 ? Fp : Fq) + lf * np;
            mf_tile<N, PREC>(Df, X, dir, t2, Tm + (lf * 3 + dir) * np);
        }
    }
    // face fluxes (EdgeFlux :1237-1341 with RiemannSolver :1595-1620)
    #pragma unroll
    for (int q = 0; q < FQ; ++q) {
        const int w = t + q * S::threads;
        if (w >= nF) continue;
        const int l2 = w / (6 * n2), rem = w - l2 * 6 * n2, sd = rem / n2, node = rem - sd * n2;
        const int e2 = blockIdx.x * epb + l2;
        if (e2 >= A.nE) continue;
        const int a = node % n1, b = node / n1;
        const int2 info = finfo[q];
        const double uN = fuN[q], len = flen[q];
        for (int f = 0; f < NQ; ++f) {
            const double own = trace3<N>(L, cs + (l2 * NQ + f) * np, sd, a, b);
            double in, ex, sign = 1.0;
            if (info.x < 0) {
                in = own;
                ex = adv3_external(info.x, f, A.bpos + (size_t)s3_slot(info.y) * 3 * n2, n2, node, A.t);
            } else {
                const double nbr = fnbr[q][f];
                if (s3_primary(info.y)) { in = own; ex = nbr; }
                else { in = nbr; ex = own; sign = -1.0; }
            }
            const double jump = ex - in;
            const double fl = 0.5 * (uN * (ex + in) - fabs(uN) * jump) * len;
            bf[((l2 * 6 + sd) * NQ + f) * n2 + node] = sign * fl;
        }
    }
    __syncthreads();
    // weak divergence + lift (MappedTimeDerivative :1345-1562), source, RK3
    if (act) {
        const double *bfE = bf + le * 6 * NQ * n2;
        for (int f = 0; f < NQ; ++f) {
            const int o = (le * NQ + f) * np;
            double ts = 0.0, tp = 0.0, tq = 0.0;
            if (PREC == 0) {
                #pragma unroll
                for (int m = 0; m < n1; ++m) {
                    ts = __builtin_fma(L.D[i + n1 * m], Fs[o + m + n1 * (j + n1 * k)], ts);
                    tp = __builtin_fma(L.D[j + n1 * m], Fp[o + i + n1 * (m + n1 * k)], tp);
                    tq = __builtin_fma(L.D[k + n1 * m], Fq[o + i + n1 * (j + n1 * m)], tq);
                }
            } else {  // narrow contraction of the deviations + FP64 line value times the row sum
                const float *T = Tm + (le * NQ + f) * 3 * np;
                ts = (double)T[nd] + Fs[o + n1 * (j + n1 * k)] * rs[i];
                tp = (double)T[np + j + n1 * (i + n1 * k)] + Fp[o + i + n2 * k] * rs[j];
                tq = (double)T[2 * np + k + n1 * (i + n1 * j)] + Fq[o + i + n1 * j] * rs[k];
            }
            const double bE = bfE[(1 * NQ + f) * n2 + j + n1 * k], bW = bfE[(3 * NQ + f) * n2 + j + n1 * k];
            const double bN = bfE[(2 * NQ + f) * n2 + i + n1 * k], bS = bfE[(0 * NQ + f) * n2 + i + n1 * k];
            const double bT = bfE[(5 * NQ + f) * n2 + i + n1 * j], bB = bfE[(4 * NQ + f) * n2 + i + n1 * j];
            ts = ts + (bE * L.lR[i] + bW * L.lL[i]) * L.iw[i];
            tp = tp + (bN * L.lR[j] + bS * L.lL[j]) * L.iw[j];
            tq = tq + (bT * L.lR[k] + bB * L.lL[k]) * L.iw[k];
            double tend = ((-ts - tp) - tq) / J;
            if (A.hasRelax) tend = tend + (rl[f] - c[f]) * rf[f];
            if (A.mode == 1) {
                A.tend[f * ms + base] = tend;
            } else {
                const double gv = A.a * gold[f] + tend;
                A.G[f * ms + base] = gv;
                c[f] = c[f] + A.gdt * gv;
            }
        }
    }
    if (A.mode == 1) return;
    // the filter's stored energies: in flight during the next barriers and
    // the filter's first passes
    Adv3FilterPre<NQ> fp;
    adv3_filter_prefetch<NQ>(A, ec, A.filterNext, fp);
    __syncthreads();  // all reads of cs / Fs / Fp / Fq for the divergence are done
    if (act)
        for (int f = 0; f < NQ; ++f) cs[(le * NQ + f) * np + nd] = c[f];
    __syncthreads();
    if (A.filterNext)
        adaptive_filter<N, NQ>(A, L, cs, Fs, Fp, Fq, red, flag, le, nd, e, act, J, c, A.filterNext, &fp);
    if (act)
        for (int f = 0; f < NQ; ++f) A.c[f * ms + base] = c[f];
    write_traces<N, NQ>(A, L, cs);
}

// u.n and |n| of each face node in its face's primary orientation, from the
// velocity traces (CalculateVelocityAtBoundaries) and the primary's normals
template <int N>
__global__ void k_adv3d_face_un(const double *__restrict__ velTr, const double *__restrict__ nrm,
                                const int2 *__restrict__ side, double *__restrict__ fdat, int nE)
{
    constexpr int n1 = N + 1, n2 = n1 * n1;
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)nE * 6 * n2) return;
    const int es = (int)(q / n2), node = (int)(q - (size_t)es * n2);
    const int e = es / 6, sd = es - 6 * e;
    const int2 info = side[es];
    int pe = e, ps = sd, pa = node % n1, pb = node / n1;
    if (info.x >= 0 && !s3_primary(info.y)) {
        s3_map<N>(info.y, node % n1, node / n1, pa, pb);
        pe = info.x;
        ps = s3_nbr_side(info.y);
    }
    const size_t pes = (size_t)pe * 6 + ps;
    const int pn = pa + n1 * pb;
    const double u = velTr[(pes * 3 + 0) * n2 + pn], v = velTr[(pes * 3 + 1) * n2 + pn];
    const double w = velTr[(pes * 3 + 2) * n2 + pn];
    const double nx = nrm[(pes * 4 + 0) * n2 + pn], ny = nrm[(pes * 4 + 1) * n2 + pn];
    const double nz = nrm[(pes * 4 + 2) * n2 + pn], len = nrm[(pes * 4 + 3) * n2 + pn];
    fdat[((size_t)es * 2) * n2 + node] = u * nx + v * ny + w * nz;
    fdat[((size_t)es * 2 + 1) * n2 + node] = len;
}

// the static velocity folded into the cofactor rows: a_d = (u, v, w) . row d
__global__ void k_adv3d_fold(const double *__restrict__ vel, const double *__restrict__ geo,
                             double *__restrict__ acoef, size_t n)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const double u = vel[q], v = vel[n + q], w = vel[2 * n + q];
    for (int d = 0; d < 3; ++d)
        acoef[d * n + q] = geo[(3 * d) * n + q] * u + geo[(3 * d + 1) * n + q] * v + geo[(3 * d + 2) * n + q] * w;
}

}  // namespace
}  // namespace self

using namespace self;

struct self_adv3d_s {
    int N = 0, n1 = 0, n2 = 0, np = 0, nq = 0, device = 0, useFilter = 0;
    size_t nE = 0, nSlots = 0;
    double dt = 0.0;
    hipStream_t stream = nullptr;
    double *nodal = nullptr, *c = nullptr, *G = nullptr, *tend = nullptr, *vel = nullptr, *geo = nullptr;
    double *fdat = nullptr, *nrm = nullptr, *bpos = nullptr, *tr[2] = {nullptr, nullptr}, *velTr = nullptr;
    double *relax = nullptr, *rFac = nullptr, *E1 = nullptr, *E2 = nullptr, *lim = nullptr;
    double *acoef = nullptr;
    int2 *side = nullptr;
    unsigned *nFiltered = nullptr;
    int cur = 0, prec = 0;
    bool trValid = false, hasRelax = false;
};

namespace {

#define ADV3_DISPATCH(N_, NQ_, CALL)                                                        \
    switch ((N_) * 10 + (NQ_)) {                                                            \
    case 11: { constexpr int NN = 1, QQ = 1; CALL; } break;                                 \
    case 21: { constexpr int NN = 2, QQ = 1; CALL; } break;                                 \
    case 31: { constexpr int NN = 3, QQ = 1; CALL; } break;                                 \
    case 41: { constexpr int NN = 4, QQ = 1; CALL; } break;                                 \
    case 51: { constexpr int NN = 5, QQ = 1; CALL; } break;                                 \
    case 61: { constexpr int NN = 6, QQ = 1; CALL; } break;                                 \
    case 71: { constexpr int NN = 7, QQ = 1; CALL; } break;                                 \
    case 12: { constexpr int NN = 1, QQ = 2; CALL; } break;                                 \
    case 22: { constexpr int NN = 2, QQ = 2; CALL; } break;                                 \
    case 32: { constexpr int NN = 3, QQ = 2; CALL; } break;                                 \
    case 42: { constexpr int NN = 4, QQ = 2; CALL; } break;                                 \
    case 52: { constexpr int NN = 5, QQ = 2; CALL; } break;                                 \
    case 62: { constexpr int NN = 6, QQ = 2; CALL; } break;                                 \
    case 72: { constexpr int NN = 7, QQ = 2; CALL; } break;                                 \
    default: break;                                                                         \
    }

#define ADV3_DISPATCH_N(N_, CALL)                                                           \
    switch (N_) {                                                                           \
    case 1: { constexpr int NN = 1; CALL; } break;                                          \
    case 2: { constexpr int NN = 2; CALL; } break;                                          \
    case 3: { constexpr int NN = 3; CALL; } break;                                          \
    case 4: { constexpr int NN = 4; CALL; } break;                                          \
    case 5: { constexpr int NN = 5; CALL; } break;                                          \
    case 6: { constexpr int NN = 6; CALL; } break;                                          \
    case 7: { constexpr int NN = 7; CALL; } break;                                          \
    default: break;                                                                         \
    }

template <int N>
int grid3(size_t nE) { return (int)((nE + Shape3<N>::epb - 1) / Shape3<N>::epb); }

template <int N>
int threads3() { return Shape3<N>::threads; }

Adv3Args args3(self_adv3d_s *h)
{
    Adv3Args A{};
    A.nE = (int)h->nE;
    A.useFilter = h->useFilter;
    A.hasRelax = h->hasRelax;
    A.nodal = h->nodal;
    A.c = h->c;
    A.G = h->G;
    A.tend = h->tend;
    A.vel = h->vel;
    A.acoef = h->acoef;
    A.geo = h->geo;
    A.fdat = h->fdat;
    A.bpos = h->bpos;
    A.side = h->side;
    A.relax = h->relax;
    A.rFac = h->rFac;
    A.E1 = h->E1;
    A.E2 = h->E2;
    A.lim = h->lim;
    A.nFiltered = h->nFiltered;
    return A;
}

// filter (when enabled) at `stage` and traces into tr[cur]
int prologue(self_adv3d_s *h, int stage)
{
    Adv3Args A = args3(h);
    A.stage = stage;
    A.trOut = h->tr[h->cur];
    ADV3_DISPATCH(h->N, h->nq, hipLaunchKernelGGL((k_adv3d_prologue<NN, QQ>), dim3(grid3<NN>(h->nE)),
                                                  dim3(threads3<NN>()), 0, h->stream, A));
    SELF_HIP(hipGetLastError());
    h->trValid = true;
    return SELF_OK;
}

int stage3(self_adv3d_s *h, double t, int m, int mode, int filterNext)
{
    Adv3Args A = args3(h);
    A.t = t;
    A.mode = mode;
    A.a = m >= 0 ? kRkA[m] : 0.0;
    A.gdt = m >= 0 ? kRkG[m] * h->dt : 0.0;
    A.filterNext = h->useFilter ? filterNext : 0;
    A.trIn = h->tr[h->cur];
    A.trOut = h->tr[1 - h->cur];
    switch (h->prec) {
    case 1:
        ADV3_DISPATCH(h->N, h->nq, hipLaunchKernelGGL((k_adv3d_stage<NN, QQ, 1>), dim3(grid3<NN>(h->nE)),
                                                      dim3(threads3<NN>()), 0, h->stream, A));
        break;
    case 2:
        ADV3_DISPATCH(h->N, h->nq, hipLaunchKernelGGL((k_adv3d_stage<NN, QQ, 2>), dim3(grid3<NN>(h->nE)),
                                                      dim3(threads3<NN>()), 0, h->stream, A));
        break;
    default:
        ADV3_DISPATCH(h->N, h->nq, hipLaunchKernelGGL((k_adv3d_stage<NN, QQ, 0>), dim3(grid3<NN>(h->nE)),
                                                      dim3(threads3<NN>()), 0, h->stream, A));
    }
    SELF_HIP(hipGetLastError());
    if (mode == 0) h->cur = 1 - h->cur;
    return SELF_OK;
}

void h2d_fields(const double *src, double *dst, size_t nE, int nf, int np)
{
    for (size_t e = 0; e < nE; ++e)
        for (int f = 0; f < nf; ++f)
            std::memcpy(dst + ((size_t)f * nE + e) * np, src + (e * nf + f) * np, sizeof(double) * np);
}

void d2h_fields(const double *src, double *dst, size_t nE, int nf, int np)
{
    for (size_t e = 0; e < nE; ++e)
        for (int f = 0; f < nf; ++f)
            std::memcpy(dst + (e * nf + f) * np, src + ((size_t)f * nE + e) * np, sizeof(double) * np);
}

void free_all(self_adv3d_s *h)
{
    void *p[] = {h->nodal, h->c, h->G, h->tend, h->vel, h->geo, h->fdat, h->nrm, h->bpos, h->tr[0], h->tr[1],
                 h->velTr, h->relax, h->rFac, h->E1, h->E2, h->lim, h->acoef, h->side, h->nFiltered};
    for (void *q : p)
        if (q) (void)hipFree(q);
    if (h->stream) (void)hipStreamDestroy(h->stream);
}

}  // namespace

extern "C" {

int self_adv3d_create(int N, int nTracers, int nElems, int nFaces, const int *faces, const double *cov,
                      const double *J, const double *nHat, const double *nLen, const double *bpos,
                      const double *dMat, const double *qWeight, const double *lagLeft, const double *lagRight,
                      const double *filter1d, int useFilter, double dt, self_adv3d_t *out)
try {
    SELF_REQUIRE(out && faces && cov && J && nHat && nLen && bpos && dMat && qWeight && lagLeft && lagRight,
                 SELF_ERR_ARG, "self_adv3d_create: null argument");
    SELF_REQUIRE(N >= 1 && N <= 7 && nTracers >= 1 && nTracers <= 2 && nElems >= 1 && nFaces >= 1, SELF_ERR_ARG,
                 "self_adv3d_create: N=%d (1..7), nTracers=%d (1..2)", N, nTracers);
    SELF_REQUIRE(!useFilter || filter1d, SELF_ERR_ARG, "self_adv3d_create: the filter needs its 1-D factor");
    const int n1 = N + 1, n2 = n1 * n1, np = n2 * n1;
    const size_t nE = (size_t)nElems;
    // side descriptors and boundary slots from the face table
    std::vector<int2> sd(nE * 6, make_int2(0, 0));
    std::vector<int> seen(nE * 6, 0);
    std::vector<size_t> slotSide;
    for (int f = 0; f < nFaces; ++f) {
        const int *r = faces + 9 * (size_t)f;
        const int e1 = r[0] - 1, s1 = r[1] - 1, e2 = r[2], s2 = std::abs(r[3]) - 1;
        const int iS = r[4], jS = r[5], iI = r[6], jI = r[7], sw = r[8];
        SELF_REQUIRE(e1 >= 0 && (size_t)e1 < nE && s1 >= 0 && s1 < 6, SELF_ERR_ARG, "face %d: bad primary", f + 1);
        if (e2 > 0) {
            SELF_REQUIRE((size_t)(e2 - 1) < nE && s2 >= 0 && s2 < 6, SELF_ERR_ARG, "face %d: bad secondary", f + 1);
            SELF_REQUIRE((iI == 1 || iI == -1) && (jI == 1 || jI == -1) && (sw == 0 || sw == 1), SELF_ERR_ARG,
                         "face %d: increments/swap", f + 1);
            SELF_REQUIRE(iS == (iI > 0 ? 0 : N) && jS == (jI > 0 ? 0 : N), SELF_ERR_ARG,
                         "face %d: start must be 0 (inc 1) or N (inc -1)", f + 1);
            const int bits = (sw << 4) | ((iI < 0) << 5) | ((jI < 0) << 6);
            sd[(size_t)e1 * 6 + s1] = make_int2(e2 - 1, s2 | (1 << 3) | bits);
            sd[(size_t)(e2 - 1) * 6 + s2] = make_int2(e1, s1 | bits);
            seen[(size_t)e1 * 6 + s1]++;
            seen[(size_t)(e2 - 1) * 6 + s2]++;
        } else {
            SELF_REQUIRE(e2 < 0, SELF_ERR_ARG, "face %d: secondary element 0", f + 1);
            sd[(size_t)e1 * 6 + s1] = make_int2(e2, (1 << 3) | ((int)slotSide.size() << 8));
            slotSide.push_back((size_t)e1 * 6 + s1);
            seen[(size_t)e1 * 6 + s1]++;
        }
    }
    for (size_t k = 0; k < nE * 6; ++k)
        SELF_REQUIRE(seen[k] == 1, SELF_ERR_ARG, "element %zu side %zu is on %d faces", k / 6 + 1, k % 6 + 1,
                     seen[k]);
    // geometry: cofactor rows (the reference's contravariant projection) and J
    std::vector<double> geo(10 * nE * np);
    for (size_t id = 0; id < nE * np; ++id) {
        SELF_REQUIRE(J[id] != 0.0, SELF_ERR_ARG, "zero Jacobian at element %zu", id / np + 1);
        for (int dir = 0; dir < 3; ++dir) {
            double cf[3];
            cofactor3(cov + 9 * id, dir, cf);
            for (int r = 0; r < 3; ++r) geo[(3 * dir + r) * nE * np + id] = cf[r];
        }
        geo[9 * nE * np + id] = J[id];
    }
    std::vector<double> nrm(nE * 6 * 4 * n2), bp(std::max<size_t>(1, slotSide.size()) * 3 * n2, 0.0);
    for (size_t es = 0; es < nE * 6; ++es)
        for (int q = 0; q < n2; ++q) {
            const size_t hb = q + (size_t)n2 * es;
            for (int r = 0; r < 3; ++r) nrm[(es * 4 + r) * n2 + q] = nHat[3 * hb + r];
            nrm[(es * 4 + 3) * n2 + q] = nLen[hb];
        }
    for (size_t s = 0; s < slotSide.size(); ++s)
        for (int q = 0; q < n2; ++q)
            for (int r = 0; r < 3; ++r) bp[(s * 3 + r) * n2 + q] = bpos[3 * (q + (size_t)n2 * slotSide[s]) + r];
    std::vector<double> nod(2 * n1 * n1 + 3 * n1, 0.0);
    std::memcpy(nod.data(), dMat, sizeof(double) * n1 * n1);
    std::memcpy(nod.data() + n1 * n1, qWeight, sizeof(double) * n1);
    std::memcpy(nod.data() + n1 * n1 + n1, lagLeft, sizeof(double) * n1);
    std::memcpy(nod.data() + n1 * n1 + 2 * n1, lagRight, sizeof(double) * n1);
    if (filter1d) std::memcpy(nod.data() + n1 * n1 + 3 * n1, filter1d, sizeof(double) * n1 * n1);

    auto *h = new self_adv3d_s();
    h->N = N; h->n1 = n1; h->n2 = n2; h->np = np; h->nq = nTracers; h->nE = nE; h->dt = dt;
    h->useFilter = useFilter ? 1 : 0;
    h->nSlots = slotSide.size();
    const size_t nq = h->nq;
    hipError_t e = hipGetDevice(&h->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    auto M = [&](double **p, size_t n) { if (e == hipSuccess) e = hipMalloc(p, sizeof(double) * n); };
    M(&h->nodal, nod.size());
    M(&h->c, nq * nE * np);
    M(&h->G, nq * nE * np);
    M(&h->tend, nq * nE * np);
    M(&h->vel, 3 * nE * np);
    M(&h->acoef, 3 * nE * np);
    M(&h->geo, geo.size());
    M(&h->fdat, nE * 6 * 2 * n2);
    M(&h->nrm, nrm.size());
    M(&h->bpos, bp.size());
    M(&h->tr[0], nE * 6 * nq * n2);
    M(&h->tr[1], nE * 6 * nq * n2);
    M(&h->velTr, nE * 6 * 3 * n2);
    M(&h->E1, nq * nE);
    M(&h->E2, nq * nE);
    M(&h->lim, 3 * nq * nE);
    if (e == hipSuccess) e = hipMalloc(&h->side, sizeof(int2) * sd.size());
    if (e == hipSuccess) e = hipMalloc(&h->nFiltered, sizeof(unsigned));
    if (e == hipSuccess) e = hipMemcpy(h->nodal, nod.data(), sizeof(double) * nod.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h->geo, geo.data(), sizeof(double) * geo.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h->nrm, nrm.data(), sizeof(double) * nrm.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h->bpos, bp.data(), sizeof(double) * bp.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h->side, sd.data(), sizeof(int2) * sd.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(h->c, 0, sizeof(double) * nq * nE * np);
    if (e == hipSuccess) e = hipMemset(h->vel, 0, sizeof(double) * 3 * nE * np);
    if (e == hipSuccess) e = hipMemset(h->acoef, 0, sizeof(double) * 3 * nE * np);
    if (e == hipSuccess) e = hipMemset(h->fdat, 0, sizeof(double) * nE * 6 * 2 * n2);
    if (e == hipSuccess) e = hipMemset(h->E1, 0, sizeof(double) * nq * nE);
    if (e == hipSuccess) e = hipMemset(h->E2, 0, sizeof(double) * nq * nE);
    if (e == hipSuccess) e = hipMemset(h->lim, 0, sizeof(double) * 3 * nq * nE);
    if (e == hipSuccess) e = hipMemset(h->nFiltered, 0, sizeof(unsigned));
    if (e != hipSuccess) {
        free_all(h);
        delete h;
        return hip_fail(e, "self_adv3d_create", __FILE__, __LINE__);
    }
    *out = h;
    return SELF_OK;
}
SELF_CATCH_ALL

int self_adv3d_destroy(self_adv3d_t h)
{
    if (!h) return SELF_OK;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    free_all(h);
    delete h;
    return SELF_OK;
}

// field 0: solution (0:N,0:N,0:N,1:nTracers) per element; 1: velocity
// (0:N,0:N,0:N,1:3); 2: relaxation field; 3: relaxation factor
int self_adv3d_set_field(self_adv3d_t h, int field, const double *data)
try {
    SELF_REQUIRE(h && data, SELF_ERR_ARG, "null argument");
    SELF_REQUIRE(field >= 0 && field <= 3, SELF_ERR_ARG, "field %d", field);
    SELF_HIP(hipSetDevice(h->device));
    SELF_HIP(hipStreamSynchronize(h->stream));
    const size_t nE = h->nE, np = h->np;
    const int nf = field == 1 ? 3 : h->nq;
    std::vector<double> buf((size_t)nf * nE * np);
    h2d_fields(data, buf.data(), nE, nf, (int)np);
    double **dst = field == 0 ? &h->c : field == 1 ? &h->vel : field == 2 ? &h->relax : &h->rFac;
    if (!*dst) SELF_HIP(hipMalloc(dst, sizeof(double) * buf.size()));
    SELF_HIP(hipMemcpy(*dst, buf.data(), sizeof(double) * buf.size(), hipMemcpyHostToDevice));
    if (field == 0) h->trValid = false;
    if (field >= 2) {
        for (double **p : {&h->relax, &h->rFac})
            if (!*p) {
                SELF_HIP(hipMalloc(p, sizeof(double) * buf.size()));
                SELF_HIP(hipMemset(*p, 0, sizeof(double) * buf.size()));
            }
        h->hasRelax = true;
    }
    if (field == 1) {  // CalculateVelocityAtBoundaries once: velocity is static
        const int nsn = (int)(nE * 6 * h->n2);
        ADV3_DISPATCH_N(h->N, {
            hipLaunchKernelGGL((k_adv3d_traces<NN, 3>), dim3(grid3<NN>(nE)), dim3(threads3<NN>()), 0, h->stream,
                               h->vel, h->nodal, h->velTr, (int)nE);
            hipLaunchKernelGGL((k_adv3d_face_un<NN>), dim3((nsn + 255) / 256), dim3(256), 0, h->stream, h->velTr,
                               h->nrm, h->side, h->fdat, (int)nE);
        });
        const size_t nn = nE * np;
        hipLaunchKernelGGL(k_adv3d_fold, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, h->stream, h->vel,
                           h->geo, h->acoef, nn);
        SELF_HIP(hipGetLastError());
        SELF_HIP(hipStreamSynchronize(h->stream));
    }
    return SELF_OK;
}
SELF_CATCH_ALL

// field 0: solution, 1: tendency; 2: filter energies E1, E2 and limiter
// (nTracers x nElems each for E1, E2, then 3 x nTracers x nElems limiter)
int self_adv3d_get_field(self_adv3d_t h, int field, double *data)
try {
    SELF_REQUIRE(h && data, SELF_ERR_ARG, "null argument");
    SELF_REQUIRE(field >= 0 && field <= 2, SELF_ERR_ARG, "field %d", field);
    SELF_HIP(hipSetDevice(h->device));
    SELF_HIP(hipStreamSynchronize(h->stream));
    const size_t nE = h->nE, np = h->np, nq = h->nq;
    if (field == 2) {
        SELF_HIP(hipMemcpy(data, h->E1, sizeof(double) * nq * nE, hipMemcpyDeviceToHost));
        SELF_HIP(hipMemcpy(data + nq * nE, h->E2, sizeof(double) * nq * nE, hipMemcpyDeviceToHost));
        SELF_HIP(hipMemcpy(data + 2 * nq * nE, h->lim, sizeof(double) * 3 * nq * nE, hipMemcpyDeviceToHost));
        return SELF_OK;
    }
    std::vector<double> buf(nq * nE * np);
    SELF_HIP(hipMemcpy(buf.data(), field == 0 ? h->c : h->tend, sizeof(double) * buf.size(), hipMemcpyDeviceToHost));
    d2h_fields(buf.data(), data, nE, (int)nq, (int)np);
    return SELF_OK;
}
SELF_CATCH_ALL

// GlobalTimeDerivative(t, m) (:1049-1121): filter at stage m (when enabled),
// traces, face fluxes and tendency; the state is filtered in place as the
// reference does
int self_adv3d_global_time_derivative(self_adv3d_t h, double t, int stage)
{
    SELF_REQUIRE(h && stage >= 1 && stage <= 3, SELF_ERR_ARG, "bad arguments");
    SELF_HIP(hipSetDevice(h->device));
    int rc = SELF_OK;
    if (h->useFilter || !h->trValid) rc = prologue(h, stage);
    if (rc) return rc;
    return stage3(h, t, -1, 1, 0);
}

int self_adv3d_forward_step_rk3(self_adv3d_t h, double tn, int nSteps, double *elapsed_ms)
{
    SELF_REQUIRE(h && nSteps >= 0, SELF_ERR_ARG, "bad arguments");
    SELF_HIP(hipSetDevice(h->device));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (elapsed_ms) {
        SELF_HIP(hipEventCreate(&e0));
        SELF_HIP(hipEventCreate(&e1));
        SELF_HIP(hipEventRecord(e0, h->stream));
    }
    int rc = SELF_OK;
    if (nSteps > 0 && (h->useFilter || !h->trValid)) rc = prologue(h, 1);
    for (int s = 0; s < nSteps && !rc; ++s) {
        const double t0 = tn + s * h->dt;
        for (int m = 0; m < 3 && !rc; ++m) {
            const bool last = (s == nSteps - 1) && (m == 2);
            // the filter of the next stage (1..3) runs here unless the call ends
            rc = stage3(h, t0 + kRkB[m] * h->dt, m, 0, last ? 0 : (m + 1) % 3 + 1);
        }
    }
    // without the filter the traces written by the last stage are current
    h->trValid = !h->useFilter;
    if (elapsed_ms) {
        SELF_HIP(hipEventRecord(e1, h->stream));
        SELF_HIP(hipEventSynchronize(e1));
        float ms = 0;
        SELF_HIP(hipEventElapsedTime(&ms, e0, e1));
        *elapsed_ms = ms;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    return rc;
}

// number of element-tracer pairs the filter replaced since the last reset
int self_adv3d_filter_count(self_adv3d_t h, int reset, unsigned *count)
try {
    SELF_REQUIRE(h && count, SELF_ERR_ARG, "null argument");
    SELF_HIP(hipSetDevice(h->device));
    SELF_HIP(hipStreamSynchronize(h->stream));
    SELF_HIP(hipMemcpy(count, h->nFiltered, sizeof(unsigned), hipMemcpyDeviceToHost));
    if (reset) SELF_HIP(hipMemset(h->nFiltered, 0, sizeof(unsigned)));
    return SELF_OK;
}
SELF_CATCH_ALL

// 0: FP64 contractions; 1: FP32 matrix-core contractions; 2: split-BF16
// matrix-core contractions (the volume weak divergence only)
int self_adv3d_set_precision(self_adv3d_t h, int precision)
{
    SELF_REQUIRE(h, SELF_ERR_ARG, "null handle");
    SELF_REQUIRE(precision >= 0 && precision <= 2, SELF_ERR_ARG, "precision %d (0 fp64, 1 fp32, 2 bf16x3)",
                 precision);
    h->prec = precision;
    return SELF_OK;
}

int self_adv3d_synchronize(self_adv3d_t h)
{
    SELF_REQUIRE(h, SELF_ERR_ARG, "null handle");
    SELF_HIP(hipStreamSynchronize(h->stream));
    return SELF_OK;
}

}  // extern "C"
