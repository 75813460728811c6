// fempc.hip -- CGsemPreconditioner_2D on the device (see fempc.hpp).

#include <algorithm>
#include <cmath>
#include <cstring>

#include "fempc.hpp"
#include "krylov.hpp"

namespace self {

namespace {

constexpr int kPcThreads = 256;
constexpr int kMaxWidth = 32;  // ELL slots per row (a node shared by up to 8 elements of the linear mesh)
constexpr int kPcChunk = 8;    // ELL slots of a row loaded together

__device__ __forceinline__ bool pc_skip(const PcState *st, const SolverState *outer)
{
    return st->done || (outer && (outer->done || outer->innerExit));
}

// block sum of a double, result in thread 0 (fixed tree)
__device__ __forceinline__ double pc_block_sum(double v)
{
    __shared__ double lw[kPcThreads / 64];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) lw[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int q = 0; q < kPcThreads / 64; ++q) s += lw[q];
    return s;
}

// one workgroup: fixed-order sum of the kPcParts partials
__device__ __forceinline__ double pc_final_sum(const double *partials)
{
    double v = 0.0;
    for (int b = threadIdx.x; b < kPcParts; b += kPcThreads) v += partials[b];
    return pc_block_sum(v);
}

// Solve_CGsemPreconditioner_2D (:704-711): r = b (MapFromOrigToPC), u = 0, |r|^2
template <typename T>
__global__ __launch_bounds__(kPcThreads) void k_pc_start(const T *__restrict__ b, T *__restrict__ r,
                                                         T *__restrict__ x, size_t n, double *partials,
                                                         const PcState *st, const SolverState *outer)
{
    if (outer && (outer->done || outer->innerExit)) return;
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const T bv = b[i];
        r[i] = bv;
        x[i] = (T)0;
        acc += (double)bv * (double)bv;
    }
    const double p = pc_block_sum(acc);
    if (threadIdx.x == 0) partials[blockIdx.x] = p;
}

__global__ __launch_bounds__(kPcThreads) void k_pc_final_start(const double *partials, PcState *st,
                                                               const SolverState *outer, int maxIters, double tol)
{
    if (outer && (outer->done || outer->innerExit)) return;
    const double S = pc_final_sum(partials);
    if (threadIdx.x == 0) {
        st->r0 = sqrt(S);
        st->rNorm = st->r0;
        st->num = S;  // num = r.w with w = r (:716-717)
        st->numP[1] = S;
        st->den = 0.0;
        st->beta = 0.0;
        st->iter = 1;
        st->iterations = 0;
        st->maxIters = maxIters;
        st->tol = tol;
        st->done = (maxIters <= 0 || S == 0.0) ? 1 : 0;
    }
}

// v = w + c v (w = r; c = num/den, none on the first iteration), z = A v,
// v.z; the direction of neighbouring rows is formed on the fly from the
// same inputs with the same fused multiply-add, so it equals what their
// own rows store.
// The end of iteration it-1 (rNorm from the update's r.r partials, the
// stop test rNorm/r0 < TOL, num and c = num/den, :728-732) comes first:
// every workgroup sums the partials itself in one fixed order, so all take
// the same decision from the same bits; workgroup 0 records it.
template <typename T>
__global__ __launch_bounds__(kPcThreads) void k_pc_spmv(const T *__restrict__ r, const T *__restrict__ vOld,
                                                        T *__restrict__ vNew, T *__restrict__ z,
                                                        const int *__restrict__ col, const T *__restrict__ val,
                                                        int width, size_t n, double *partials, PcState *st,
                                                        const SolverState *outer, int it)
{
    if (pc_skip(st, outer)) return;  // decided by an earlier launch: the same for every workgroup
    const bool first = it == 1;
    T beta = (T)0;
    if (!first) {
        __shared__ double sS;
        const double S0 = pc_final_sum(partials + kPcParts);
        if (threadIdx.x == 0) sS = S0;
        __syncthreads();
        const double S = sS, rNorm = sqrt(S), den = st->numP[(it - 1) & 1];
        const bool stop = rNorm / st->r0 < st->tol || it - 1 >= st->maxIters;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            st->rNorm = rNorm;
            st->den = den;
            st->num = S;
            st->numP[it & 1] = S;
            st->beta = S / den;
            st->iterations = it - 1;
            if (stop) st->done = 1;
            else st->iter = it;
        }
        if (stop) return;
        beta = (T)(S / den);
    }
    // iteration 1 has no previous direction: it reads r in its place (and
    // ignores it), so the same loads serve every iteration
    const T *__restrict__ vPrev = first ? r : vOld;
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        T ri = r[i], vi = vPrev[i];
        T zi = (T)0;
        // the row's slots kPcChunk at a time: their column indices and values
        // first, then the gathered entries, all in flight together, then the
        // multiply-adds in slot order (a loop that loads a slot's column and
        // then its entry waits twice per slot); slots past the width repeat
        // the last one and are not added
        for (int k0 = 0; k0 < width; k0 += kPcChunk) {
            int cc[kPcChunk];
            T vv[kPcChunk], rc[kPcChunk], vc[kPcChunk];
            #pragma unroll
            for (int u = 0; u < kPcChunk; ++u) {
                const size_t kk = (size_t)min(k0 + u, width - 1) * n + i;
                cc[u] = col[kk];
                vv[u] = val[kk];
            }
            #pragma unroll
            for (int u = 0; u < kPcChunk; ++u) {
                rc[u] = r[cc[u]];
                vc[u] = vPrev[cc[u]];
            }
            #pragma unroll
            for (int u = 0; u < kPcChunk; ++u) asm volatile("" : "+v"(rc[u]), "+v"(vc[u]));
            #pragma unroll
            for (int u = 0; u < kPcChunk; ++u) {
                const T x = first ? rc[u] : __builtin_fma(beta, vc[u], rc[u]);
                if (k0 + u < width) zi = __builtin_fma(vv[u], x, zi);
            }
        }
        asm volatile("" : "+v"(ri), "+v"(vi));
        const T vn = first ? ri : __builtin_fma(beta, vi, ri);
        vNew[i] = vn;
        z[i] = zi;
        acc += (double)vn * (double)zi;
    }
    const double p = pc_block_sum(acc);
    if (threadIdx.x == 0) partials[blockIdx.x] = p;
}

// a = num / (v.z) (:723), each workgroup summing the product step's
// partials itself; u += a v, r -= a z, |r|^2 (:724-727)
template <typename T>
__global__ __launch_bounds__(kPcThreads) void k_pc_update(T *__restrict__ x, const T *__restrict__ v,
                                                          T *__restrict__ r, const T *__restrict__ z, size_t n,
                                                          double *partials, PcState *st,
                                                          const SolverState *outer, int it)
{
    if (pc_skip(st, outer)) return;
    __shared__ double sS;
    const double S0 = pc_final_sum(partials);
    if (threadIdx.x == 0) sS = S0;
    __syncthreads();
    const double alpha = st->numP[it & 1] / sS;
    if (blockIdx.x == 0 && threadIdx.x == 0) st->alpha = alpha;
    const T a = (T)alpha;
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        x[i] = x[i] + a * v[i];
        const T rv = r[i] - a * z[i];
        r[i] = rv;
        acc += (double)rv * (double)rv;
    }
    const double p = pc_block_sum(acc);
    if (threadIdx.x == 0) partials[kPcParts + blockIdx.x] = p;
}

// the end of the last iteration issued (it = maxIters): rNorm and the
// count; the loop ends there whatever the stop test says (:728-732)
__global__ __launch_bounds__(kPcThreads) void k_pc_final_rnorm(const double *partials, PcState *st,
                                                               const SolverState *outer, int it)
{
    if (pc_skip(st, outer)) return;
    const double S = pc_final_sum(partials + kPcParts);
    if (threadIdx.x == 0) {
        const double den = st->numP[it & 1];
        st->rNorm = sqrt(S);
        st->den = den;
        st->num = S;
        st->beta = S / den;
        st->iterations = it;
        st->done = 1;
    }
}

// nIt = 0: the preconditioner returns r itself (:738-740)
template <typename T>
__global__ __launch_bounds__(kPcThreads) void k_pc_finish(T *__restrict__ x, const T *__restrict__ r, size_t n,
                                                          const PcState *st, const SolverState *outer)
{
    if (outer && (outer->done || outer->innerExit)) return;
    if (st->maxIters > 0) return;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        x[i] = r[i];
}

template <typename T>
int pc_apply(FemPC &pc, const T *b, T *x, T *r, T *z, T *v0, T *v1, const T *val, const SolverState *outer,
             hipStream_t s)
{
    const size_t n = pc.n;
    const dim3 g(kPcParts), t(kPcThreads);
    hipLaunchKernelGGL(k_pc_start<T>, g, t, 0, s, b, r, x, n, pc.partials, (const PcState *)pc.st, outer);
    hipLaunchKernelGGL(k_pc_final_start, dim3(1), t, 0, s, (const double *)pc.partials, pc.st, outer, pc.maxIters,
                       pc.tol);
    T *v[2] = {v0, v1};
    for (int it = 1; it <= pc.maxIters; ++it) {
        T *vNew = v[it & 1];
        const T *vOld = v[(it + 1) & 1];
        hipLaunchKernelGGL(k_pc_spmv<T>, g, t, 0, s, (const T *)r, vOld, vNew, z, (const int *)pc.col, val,
                           pc.width, n, pc.partials, pc.st, outer, it);
        hipLaunchKernelGGL(k_pc_update<T>, g, t, 0, s, x, (const T *)vNew, r, (const T *)z, n, pc.partials, pc.st,
                           outer, it);
    }
    if (pc.maxIters > 0)
        hipLaunchKernelGGL(k_pc_final_rnorm, dim3(1), t, 0, s, (const double *)pc.partials, pc.st, outer,
                           pc.maxIters);
    hipLaunchKernelGGL(k_pc_finish<T>, g, t, 0, s, x, (const T *)r, n, (const PcState *)pc.st, outer);
    SELF_HIP(hipGetLastError());
    return SELF_OK;
}

}  // namespace

FemPC::~FemPC()
{
    void *p[] = {col, val, val32, r, z, v[0], v[1], r32, z32, v32[0], v32[1], partials, st};
    for (void *q : p)
        if (q) (void)hipFree(q);
}

int FemPC::build(int N, size_t nE, const double *x, const double *y, const double *k, const int *gid, size_t nG,
                 int pcIterates, double pcTolerance)
{
    SELF_REQUIRE(N >= 1 && nG >= 1 && pcIterates >= 0 && pcTolerance > 0.0, SELF_ERR_ARG,
                 "FEM preconditioner: N=%d, %zu unknowns, pcIterates=%d, pcTolerance=%g", N, nG, pcIterates,
                 pcTolerance);
    n = nG;
    maxIters = pcIterates;
    tol = pcTolerance;
    const int n1 = N + 1, np = n1 * n1;
    // the N = 1 Gauss-Lobatto derivative matrix (CG form) and weights, D(k, m) at [k + 2 m]
    double s1[2], w1[2], bw[2], D[4];
    int rc = self_legendre_quadrature(1, SELF_GAUSS_LOBATTO, s1, w1);
    if (!rc) rc = self_barycentric_weights(1, s1, bw);
    if (!rc) rc = self_derivative_matrix(1, s1, bw, D);
    if (rc) return rc;
    std::vector<int> cnt(nG, 0), cols(nG * kMaxWidth, -1);
    std::vector<double> vals(nG * kMaxWidth, 0.0);
    auto add = [&](int row, int c, double a) -> bool {
        int *rc_ = cols.data() + (size_t)row * kMaxWidth;
        double *rv = vals.data() + (size_t)row * kMaxWidth;
        for (int q = 0; q < cnt[row]; ++q)
            if (rc_[q] == c) { rv[q] += a; return true; }
        if (cnt[row] == kMaxWidth) return false;
        rc_[cnt[row]] = c;
        rv[cnt[row]] = a;
        cnt[row]++;
        return true;
    };
    for (size_t e = 0; e < nE; ++e)
        for (int iP = 1; iP <= N; ++iP)
            for (int iS = 1; iS <= N; ++iS) {
                // corners (i, j) of the sub-element: SW (0,0), SE (1,0), NW (0,1), NE (1,1)
                size_t l[2][2];
                for (int j = 0; j < 2; ++j)
                    for (int i = 0; i < 2; ++i) l[j][i] = e * np + (iS - 1 + i) + n1 * (iP - 1 + j);
                // metric terms of the straight-sided bilinear element at its corners
                double xs[2][2], xp[2][2], ys[2][2], yp[2][2], Jc[2][2], kc[2][2];
                for (int j = 0; j < 2; ++j)
                    for (int i = 0; i < 2; ++i) {
                        xs[j][i] = 0.5 * (x[l[j][1]] - x[l[j][0]]);
                        ys[j][i] = 0.5 * (y[l[j][1]] - y[l[j][0]]);
                        xp[j][i] = 0.5 * (x[l[1][i]] - x[l[0][i]]);
                        yp[j][i] = 0.5 * (y[l[1][i]] - y[l[0][i]]);
                        Jc[j][i] = xs[j][i] * yp[j][i] - xp[j][i] * ys[j][i];
                        kc[j][i] = k ? k[l[j][i]] : 1.0;
                    }
                // FluxDivergence at N = 1 (CGsemPreconditioner_2D_Class.f90:360-440) on unit vectors
                for (int bj = 0; bj < 2; ++bj)
                    for (int bi = 0; bi < 2; ++bi) {
                        const int gb = gid[l[bj][bi]];
                        if (gb < 0) continue;  // Dirichlet column: masked to zero
                        double u[2][2] = {{0, 0}, {0, 0}};
                        u[bj][bi] = 1.0;
                        double F1[2][2], F2[2][2];
                        for (int j = 0; j < 2; ++j)
                            for (int i = 0; i < 2; ++i) {
                                const double us = D[i] * u[j][0] + D[i + 2] * u[j][1];
                                const double up = D[j] * u[0][i] + D[j + 2] * u[1][i];
                                const double ux = (yp[j][i] * us - ys[j][i] * up) / Jc[j][i];
                                const double uy = (xs[j][i] * up - xp[j][i] * us) / Jc[j][i];
                                F1[j][i] = kc[j][i] * (ux * yp[j][i] - uy * xp[j][i]) * w1[i];
                                F2[j][i] = kc[j][i] * (uy * xs[j][i] - ux * ys[j][i]) * w1[j];
                            }
                        for (int j = 0; j < 2; ++j)
                            for (int i = 0; i < 2; ++i) {
                                const int ga = gid[l[j][i]];
                                if (ga < 0) continue;
                                const double d1 = D[2 * i] * F1[j][0] + D[2 * i + 1] * F1[j][1];
                                const double d2 = D[2 * j] * F2[0][i] + D[2 * j + 1] * F2[1][i];
                                const double Lu = -(d1 * w1[j] + d2 * w1[i]);
                                SELF_REQUIRE(add(ga, gb, Lu), SELF_ERR_ARG,
                                             "FEM preconditioner: a node couples to more than %d others", kMaxWidth);
                            }
                    }
            }
    width = 1;
    for (size_t g = 0; g < nG; ++g) width = std::max(width, cnt[g]);
    std::vector<int> ellc((size_t)width * nG);
    std::vector<double> ellv((size_t)width * nG);
    std::vector<float> ellf((size_t)width * nG);
    for (size_t g = 0; g < nG; ++g)
        for (int q = 0; q < width; ++q) {
            const bool used = q < cnt[g];
            ellc[(size_t)q * nG + g] = used ? cols[g * kMaxWidth + q] : (int)g;
            ellv[(size_t)q * nG + g] = used ? vals[g * kMaxWidth + q] : 0.0;
            ellf[(size_t)q * nG + g] = (float)ellv[(size_t)q * nG + g];
        }
    auto dev = [&](auto **p, size_t cnt_) -> int {
        SELF_HIP(hipMalloc(reinterpret_cast<void **>(p), sizeof(**p) * std::max<size_t>(cnt_, 1)));
        bytes += sizeof(**p) * cnt_;
        return SELF_OK;
    };
    const size_t ne = (size_t)width * nG;
    if ((rc = dev(&col, ne)) || (rc = dev(&val, ne)) || (rc = dev(&val32, ne)) || (rc = dev(&r, nG)) ||
        (rc = dev(&z, nG)) || (rc = dev(&v[0], nG)) || (rc = dev(&v[1], nG)) || (rc = dev(&r32, nG)) ||
        (rc = dev(&z32, nG)) || (rc = dev(&v32[0], nG)) || (rc = dev(&v32[1], nG)) ||
        (rc = dev(&partials, 2 * kPcParts)) || (rc = dev(&st, 1)))
        return rc;
    SELF_HIP(hipMemcpy(col, ellc.data(), sizeof(int) * ne, hipMemcpyHostToDevice));
    SELF_HIP(hipMemcpy(val, ellv.data(), sizeof(double) * ne, hipMemcpyHostToDevice));
    SELF_HIP(hipMemcpy(val32, ellf.data(), sizeof(float) * ne, hipMemcpyHostToDevice));
    SELF_HIP(hipMemset(st, 0, sizeof(PcState)));
    return SELF_OK;
}

int FemPC::apply(const double *b, double *x, const SolverState *outer, hipStream_t s)
{
    return pc_apply<double>(*this, b, x, r, z, v[0], v[1], val, outer, s);
}

int FemPC::apply32(const float *b, float *x, const SolverState *outer, hipStream_t s)
{
    return pc_apply<float>(*this, b, x, r32, z32, v32[0], v32[1], val32, outer, s);
}

int FemPC::last_iterations(hipStream_t s, int &its)
{
    PcState h{};
    SELF_HIP(hipMemcpyAsync(&h, st, sizeof(PcState), hipMemcpyDeviceToHost, s));
    SELF_HIP(hipStreamSynchronize(s));
    its = h.iterations;
    return SELF_OK;
}

}  // namespace self
