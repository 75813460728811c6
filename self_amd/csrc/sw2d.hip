// sw2d.hip -- conservative 2-D shallow-water DGSEM on gfx950.
//
// Reference: src/highend/shallowwater/ConservativeShallowWaterClass.f90
//   ForwardStepRK3 :1173-1218, GlobalTimeDerivative :1222-1262,
//   EdgeFlux :1266-1360, MappedTimeDerivative :1364-1522,
//   DGSystemDerivative :1526-1551, RiemannSolver :1555-1604,
//   XFlux/YFlux :1608-1658, Source :1662-1694, GetExternalState :1698-1729,
//   CalculateBathymetryGradient :1733-1868;
//   CalculateSolutionAtBoundaries src/dgsem/DGSEMSolutionStorageClass_2D.f90:749-802.
//
// Two execution forms:
//  PLAIN  one kernel per reference phase (traces, edge fluxes, mapped
//         tendency, RK3 update); traces, fluxes and the tendency go through
//         HBM as in the reference.  This is the FP64 drop-in.
//  FUSED  one kernel per RK stage.  Each element owns a workgroup slice:
//         it loads its state once, evaluates its own face traces from LDS,
//         reads the neighbours' traces written by the previous stage,
//         computes the Rusanov flux of each side node (taking the edge's
//         primary orientation so the value is the one EdgeFlux would store),
//         applies the tensor-product weak derivative and lift, adds the
//         source, does the low-storage RK3 update in place and writes the
//         traces of the new state for the next stage (ping-pong buffers).
// The MIXED precision form is FUSED with FP32 volume-flux contractions
// (applied to the deviation of each flux line from its FP64 mean; the mean's
// derivative is carried exactly by the FP64 lift, using
// sum_k Dhat(i,k) = -(l_E(i) - l_W(i))/w_i) and FP32 storage of the RK
// tendency register, bathymetry gradient and Coriolis parameter.  State,
// traces, metrics, normals and the RK accumulation stay FP64.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "dg2d_device.hpp"

namespace self {

namespace {

constexpr double kRkA[3] = {0.0, -5.0 / 9.0, -153.0 / 128.0};
constexpr double kRkG[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};

struct SWKernelArgs {
    int nE;
    int variant;               // 1: the examples' class (wave speed, no drag/relaxation, grad(h^2/2)/h)
    int members;               // ensemble members (fused form): q, G and traces per member
    double g, drag, a, gdt;
    double gcoef;              // RK stage coefficient g_m: member k steps with gcoef * dtM[k]
    const double *dtM;         // per-member time steps (members > 1)
    int hasRelax;
    // nodal storage
    const double *dS, *dP;     // (N+1)^2 column-major
    const float *dS32, *dP32;  // FP32 copies (mixed)
    const double *qwS, *qwP;   // (N+1)
    const double *lag;         // [4][N+1] west, east, south, north
    // state
    double *q;                 // [3][nE][np]
    double *G64;               // [3][nE][np] (FP64 form)
    float *G32;                // [3][nE][np] (mixed form)
    double *tend;              // [3][nE][np] (plain form)
    const double *metric;      // [5][nE][np] dxds, dxdp, dyds, dydp, J
    const double *metM;        // [nE][5] per-element reference values (mixed)
    const float *metD;         // [5][nE][np] FP32 deviations from them (mixed)
    const double *bathy;       // [3][nE][np] h, hx, hy
    const float *bathy32;      // [2][nE][np] hx, hy (mixed)
    const double *plv;         // [nE][np]
    const float *plv32;        // [nE][np] (mixed)
    const double *relax;       // [4][nE][np]
    const double *nrm;         // [nE][4][3][n1] own normals (nx, ny, |n|)
    const double *fnrm;        // [nE][4][3][n1] edge-primary normals seen from this side
    const double *fnM;         // [nE][4][3] per-side reference values (mixed)
    const float *fnD;          // [nE][4][3][n1] FP32 deviations from them (mixed)
    const int2 *side;          // [nE][4]
    const double *trIn;        // [nE][4][3][n1]
    double *trOut;
    const double *bflux;       // [nE][4][3][n1] (plain form)
};

// ---------------------------------------------------------------- helpers
template <int N>
__device__ __forceinline__ void load_lag(const double *__restrict__ lag, double lW[N + 1], double lE[N + 1],
                                         double lS[N + 1], double lN[N + 1])
{
    constexpr int n1 = N + 1;
    #pragma unroll
    for (int m = 0; m < n1; ++m) {
        lW[m] = lag[m];
        lE[m] = lag[n1 + m];
        lS[m] = lag[2 * n1 + m];
        lN[m] = lag[3 * n1 + m];
    }
}

// traces of the 3 fields of one element held in LDS (qs[f*np + node]) on
// side s at side node k
template <int N>
__device__ __forceinline__ void element_trace(const double *qs, int s, int k, const double *lW,
                                              const double *lE, const double *lS, const double *lN,
                                              double out[3])
{
    constexpr int n1 = N + 1, np = n1 * n1;
    #pragma unroll
    for (int f = 0; f < 3; ++f) {
        const double *base = qs + f * np;
        switch (s) {
        case 0: out[f] = line_dot<N>(lS, base + k, n1); break;       // south: along p at s = k
        case 1: out[f] = line_dot<N>(lE, base + n1 * k, 1); break;   // east:  along s at p = k
        case 2: out[f] = line_dot<N>(lN, base + k, n1); break;       // north
        default: out[f] = line_dot<N>(lW, base + n1 * k, 1); break;  // west
        }
    }
}

// Source, ConservativeShallowWaterClass.f90:1662-1694
// dragOverH = LinearDrag / h (0 without drag), so no division per call.
__device__ __forceinline__ void sw_source(const double s[3], double vort, double g, double hx, double hy,
                                          double dragOverH, bool relax, const double r[4], double q[3])
{
    q[0] = vort * s[1] + g * s[2] * hx;
    q[1] = -vort * s[0] + g * s[2] * hy;
    q[2] = 0.0;
    q[0] = q[0] - dragOverH * s[0];
    q[1] = q[1] - dragOverH * s[1];
    if (relax) {
        q[0] = q[0] - (s[0] - r[0]) * r[3];
        q[1] = q[1] - (s[1] - r[1]) * r[3];
        q[2] = -(s[2] - r[2]) * r[3];
    }
}

// ------------------------------------------------------------ PLAIN form
// CalculateSolutionAtBoundaries for all elements: out[e][side][f][k]
template <int N>
__global__ __launch_bounds__(256) void k_sw_traces(const double *__restrict__ q, const double *__restrict__ lag,
                                                   double *__restrict__ out, int nE, int nf)
{
    using S = Shape<N>;
    constexpr int n1 = S::n1, np = S::np, epb = S::epb;
    __shared__ double qs[epb * 3 * np];
    const int t = threadIdx.x, le = t / np, nd = t - le * np;
    const int e = blockIdx.x * epb + le;
    const bool act = le < epb && e < nE;
    double lW[n1], lE[n1], lS[n1], lN[n1];
    load_lag<N>(lag, lW, lE, lS, lN);
    if (act)
        for (int f = 0; f < nf; ++f) qs[(le * 3 + f) * np + nd] = q[((size_t)f * nE + e) * np + nd];
    __syncthreads();
    if (!act) return;
    for (int w = nd; w < 4 * n1; w += np) {
        const int s = w / n1, k = w - s * n1;
        double tr[3];
        element_trace<N>(qs + le * 3 * np, s, k, lW, lE, lS, lN, tr);
        for (int f = 0; f < nf; ++f) out[(((size_t)e * 4 + s) * nf + f) * n1 + k] = tr[f];
    }
}

// Traces of the fused form: each side node's trace is written into the
// slot of the matching node of the neighbour's side ([e'][side'][f][k'],
// k' flipped when the edge runs the other way), so a stage reads its
// neighbours' traces from its own slots without first loading the side
// descriptors.  Boundary sides write nothing.
template <int N>
__global__ __launch_bounds__(256) void k_sw_traces_push(const double *__restrict__ q, const double *__restrict__ lag,
                                                        const int2 *__restrict__ side, double *__restrict__ out,
                                                        int nE)
{
    using S = Shape<N>;
    constexpr int n1 = S::n1, np = S::np, epb = S::epb;
    __shared__ double qs[epb * 3 * np];
    const int t = threadIdx.x, le = t / np, nd = t - le * np;
    const int e = blockIdx.x * epb + le;
    const bool act = le < epb && e < nE;
    q += (size_t)blockIdx.y * 3 * nE * np;  // ensemble member
    out += (size_t)blockIdx.y * nE * 12 * n1;
    double lW[n1], lE[n1], lS[n1], lN[n1];
    load_lag<N>(lag, lW, lE, lS, lN);
    if (act)
        for (int f = 0; f < 3; ++f) qs[(le * 3 + f) * np + nd] = q[((size_t)f * nE + e) * np + nd];
    __syncthreads();
    if (!act) return;
    for (int w = nd; w < 4 * n1; w += np) {
        const int sd = w / n1, k = w - sd * n1;
        const int2 info = side[(size_t)e * 4 + sd];
        if (info.x < 0) continue;
        double tr[3];
        element_trace<N>(qs + le * 3 * np, sd, k, lW, lE, lS, lN, tr);
        const int kk = side_flip(info.y) ? N - k : k;
        double *o = out + (((size_t)info.x * 4 + side_nbr_side(info.y)) * 3) * n1 + kk;
        o[0] = tr[0]; o[n1] = tr[1]; o[2 * n1] = tr[2];
    }
}

// EdgeFlux for all edges; one thread per edge node.
template <int N>
__global__ __launch_bounds__(256) void k_sw_edge_flux(const int *__restrict__ edges, int nEdges,
                                                      const double *__restrict__ bsol,
                                                      const double *__restrict__ nrm, double g,
                                                      double *__restrict__ bflux, int variant)
{
    const bool ex = variant == 1;
    constexpr int n1 = N + 1;
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= nEdges * n1) return;
    const int ie = gid / n1, iNode = gid - ie * n1;
    const int *ed = edges + 6 * ie;
    const int e1 = ed[0] - 1, s1 = ed[1] - 1, e2 = ed[2], s2 = abs(ed[3]) - 1;
    const int start = ed[4], inc = ed[5];
    double in[3], ex_[3], fl[3];
    #pragma unroll
    for (int f = 0; f < 3; ++f) in[f] = bsol[(((size_t)e1 * 4 + s1) * 3 + f) * n1 + iNode];
    const size_t nb = ((size_t)e1 * 4 + s1) * 3 * n1 + iNode;
    const double nx = nrm[nb], ny = nrm[nb + n1], len = nrm[nb + 2 * n1];
    if (e2 > 0) {
        const int k = start - inc + iNode * inc;
        #pragma unroll
        for (int f = 0; f < 3; ++f) ex_[f] = bsol[(((size_t)(e2 - 1) * 4 + s2) * 3 + f) * n1 + k];
        sw_riemann(in, ex_, nx, ny, g, fl, ex);
        #pragma unroll
        for (int f = 0; f < 3; ++f) {
            const double v = fl[f] * len;
            bflux[(((size_t)e1 * 4 + s1) * 3 + f) * n1 + iNode] = v;
            bflux[(((size_t)(e2 - 1) * 4 + s2) * 3 + f) * n1 + k] = -v;
        }
    } else {
        sw_external(e2, nx, ny, in, ex_);
        sw_riemann(in, ex_, nx, ny, g, fl, ex);
        #pragma unroll
        for (int f = 0; f < 3; ++f) bflux[(((size_t)e1 * 4 + s1) * 3 + f) * n1 + iNode] = fl[f] * len;
    }
}

template <int N>
__device__ __forceinline__ void trace3(const NodalLDS<N> &L, const double *qs, int s, int k, double out[3])
{
    constexpr int n1 = N + 1, np = n1 * n1;
    const double *l = (s == 0) ? L.lS : (s == 1) ? L.lE : (s == 2) ? L.lN : L.lW;
    const int off = (s == 0 || s == 2) ? k : n1 * k;      // south/north: p-line at s = k
    const int stride = (s == 0 || s == 2) ? n1 : 1;       // east/west:   s-line at p = k
    #pragma unroll
    for (int f = 0; f < 3; ++f) {
        const double *v = qs + f * np + off;
        double acc = l[0] * v[0];
        #pragma unroll
        for (int m = 1; m < n1; ++m) acc = __builtin_fma(l[m], v[m * stride], acc);
        out[f] = acc;
    }
}

template <int N>
__device__ __forceinline__ void dg_divergence64_field(const NodalLDS<N> &L, const double *sC, const double *pC,
                                                      const double *bf, int i, int j, double invJ, int f,
                                                      double &tend)
{
    constexpr int n1 = N + 1, np = n1 * n1;
    const double *sl = sC + f * np + n1 * j;  // s-line at p = j
    const double *pl = pC + f * np + i;       // p-line at s = i
    double t1 = L.dS[i] * sl[0];
    double t2 = L.dP[j] * pl[0];
    #pragma unroll
    for (int m = 1; m < n1; ++m) {
        t1 = __builtin_fma(L.dS[i + n1 * m], sl[m], t1);
        t2 = __builtin_fma(L.dP[j + n1 * m], pl[n1 * m], t2);
    }
    const double a = t1 + (bf[(1 * 3 + f) * n1 + j] * L.lE[i] + bf[(3 * 3 + f) * n1 + j] * L.lW[i]) * L.iwS[i];
    const double b = t2 + (bf[(2 * 3 + f) * n1 + i] * L.lN[j] + bf[(0 * 3 + f) * n1 + i] * L.lS[j]) * L.iwP[j];
    tend = (-a - b) * invJ;
}

// Weak divergence + lift of one node (DGSystemDerivative in s then in p,
// MappedTimeDerivative :1364-1522): returns -(T1 + T2)/J per field.
// bf: [side][field][node] face fluxes (south, east, north, west).  UNROLL
// interleaves the three fields (more independent work per wave) where the
// registers allow it.
template <int N, bool UNROLL>
__device__ __forceinline__ void dg_divergence64(const NodalLDS<N> &L, const double *sC, const double *pC,
                                                const double *bf, int i, int j, double invJ, double tend[3])
{
    if constexpr (UNROLL) {
        #pragma unroll
        for (int f = 0; f < 3; ++f) dg_divergence64_field<N>(L, sC, pC, bf, i, j, invJ, f, tend[f]);
    } else {
        constexpr int n1 = N + 1, np = n1 * n1;
        const double lEi = L.lE[i], lWi = L.lW[i], lNj = L.lN[j], lSj = L.lS[j];
        const double iwS = L.iwS[i], iwP = L.iwP[j];
        #pragma unroll 1
        for (int f = 0; f < 3; ++f) {
            const double *sl = sC + f * np + n1 * j;
            const double *pl = pC + f * np + i;
            double t1 = L.dS[i] * sl[0];
            double t2 = L.dP[j] * pl[0];
            #pragma unroll
            for (int m = 1; m < n1; ++m) {
                t1 = __builtin_fma(L.dS[i + n1 * m], sl[m], t1);
                t2 = __builtin_fma(L.dP[j + n1 * m], pl[n1 * m], t2);
            }
            const double a = t1 + (bf[(1 * 3 + f) * n1 + j] * lEi + bf[(3 * 3 + f) * n1 + j] * lWi) * iwS;
            const double b = t2 + (bf[(2 * 3 + f) * n1 + i] * lNj + bf[(0 * 3 + f) * n1 + i] * lSj) * iwP;
            tend[f] = (-a - b) * invJ;
        }
    }
}

// Mixed form with the FP32 deviations prepared once per node in LDS:
// sD(m, j) = sC(m, j) - sC(0, j) along s-lines, pD(i, m) = pC(i, m) - pC(i, 0)
// along p-lines; the line's reference value re-enters in FP64 through the
// matrix row sums (D (r 1) = r sum_m D(i, m)), which holds for any operator
// the model is given, the reference's transposed dMatP included.
template <int N>
__device__ __forceinline__ void dg_divergence_dev32(const NodalLDS<N> &L, const double *sC, const double *pC,
                                                    const float *sD, const float *pD, const double *bf, int i, int j,
                                                    double invJ, double tend[3])
{
    constexpr int n1 = N + 1, np = n1 * n1;
    const double lEi = L.lE[i], lWi = L.lW[i], lNj = L.lN[j], lSj = L.lS[j];
    const double iwS = L.iwS[i], iwP = L.iwP[j], rsS = L.rsS[i], rsP = L.rsP[j];
    #pragma unroll
    for (int f = 0; f < 3; ++f) {
        const float *sl = sD + f * np + n1 * j;
        const float *pl = pD + f * np + i;
        float t1 = 0.f, t2 = 0.f;
        #pragma unroll
        for (int m = 0; m < n1; ++m) {
            t1 = __builtin_fmaf(L.dS32[i + n1 * m], sl[m], t1);
            t2 = __builtin_fmaf(L.dP32[j + n1 * m], pl[n1 * m], t2);
        }
        const double rs = sC[f * np + n1 * j], rp = pC[f * np + i];
        const double ta = (double)t1 + rs * rsS + (bf[(1 * 3 + f) * n1 + j] * lEi +
                                                    bf[(3 * 3 + f) * n1 + j] * lWi) * iwS;
        const double tb = (double)t2 + rp * rsP + (bf[(2 * 3 + f) * n1 + i] * lNj +
                                                    bf[(0 * 3 + f) * n1 + i] * lSj) * iwP;
        tend[f] = (-ta - tb) * invJ;
    }
}

// contravariant volume fluxes of one node into LDS (MappedTimeDerivative
// :1383-1398 and :1434-1449)
template <bool FAST = false>
__device__ __forceinline__ void contravariant(const double s[3], const double met[4], double g, double *sC,
                                              double *pC, int np)
{
    double xF[3], yF[3];
    sw_fluxes<FAST>(s, g, xF, yF);
    #pragma unroll
    for (int f = 0; f < 3; ++f) {
        sC[f * np] = met[3] * xF[f] - met[1] * yF[f];   //  dydp*f - dxdp*g
        pC[f * np] = -met[2] * xF[f] + met[0] * yF[f];  // -dyds*f + dxds*g
    }
}

// MappedTimeDerivative for all elements (PLAIN form, FP64): writes tend.
template <int N>
__global__ __launch_bounds__(256, 4) void k_sw_mapped(SWKernelArgs A)
{
    using S = Shape<N>;
    constexpr int n1 = S::n1, np = S::np, epb = S::epb;
    __shared__ NodalLDS<N> L;
    __shared__ double sC[epb * 3 * np], pC[epb * 3 * np], bfs[epb * 12 * n1];
    const int t = threadIdx.x, le = t / np, nd = t - le * np;
    const int e = xcd_block(blockIdx.x, gridDim.x) * epb + le;
    const bool act = le < epb && e < A.nE;
    const int i = nd % n1, j = nd / n1;
    const size_t mstride = (size_t)A.nE * np, base = (size_t)e * np + nd;
    stage_nodal<N, false>(L, A.dS, nullptr, nullptr);
    double s[3], met[4], invJ = 0, hx = 0, hy = 0, vort = 0, h = 0.0, r[4] = {0, 0, 0, 0};  // h: drag / depth
    if (act) {
        #pragma unroll
        for (int f = 0; f < 3; ++f) s[f] = A.q[f * mstride + base];
        #pragma unroll
        for (int k = 0; k < 4; ++k) met[k] = A.metric[k * mstride + base];
        invJ = 1.0 / A.metric[4 * mstride + base];
        hx = A.bathy[mstride + base];
        hy = A.bathy[2 * mstride + base];
        vort = A.plv[base];
        if (A.drag != 0.0) h = A.drag / A.bathy[base];
        if (A.hasRelax) {
            #pragma unroll
            for (int k = 0; k < 4; ++k) r[k] = A.relax[k * mstride + base];
        }
        for (int w = nd; w < 12 * n1; w += np) bfs[le * 12 * n1 + w] = A.bflux[(size_t)e * 12 * n1 + w];
        contravariant(s, met, A.g, sC + le * 3 * np + nd, pC + le * 3 * np + nd, np);
    }
    __syncthreads();
    if (!act) return;
    double tend[3], src[3];
    dg_divergence64<N, false>(L, sC + le * 3 * np, pC + le * 3 * np, bfs + le * 12 * n1, i, j, invJ, tend);
    sw_source(s, vort, A.g, hx, hy, h, A.hasRelax, r, src);
    #pragma unroll
    for (int f = 0; f < 3; ++f) A.tend[f * mstride + base] = tend[f] + src[f];
}

// RK3 low-storage update of ForwardStepRK3 (:1203-1211), element-wise.
__global__ __launch_bounds__(256) void k_sw_rk_update(double *__restrict__ q, double *__restrict__ G,
                                                      const double *__restrict__ tend, size_t n, double a,
                                                      double gdt)
{
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n) return;
    const double gold = (a != 0.0) ? G[gid] : 0.0;  // first stage: G was zeroed, a = 0
    const double gv = a * gold + tend[gid];
    G[gid] = gv;
    q[gid] = q[gid] + gdt * gv;
}

// ------------------------------------------------------------- FUSED form
// One RK stage of one group of elements (epb elements, one per 64-node
// slice of the workgroup).  sw_load issues every global load of the group,
// in the order the phases consume them: state, metrics and the side node's
// inputs (phases A, B) first, then the tendency-only inputs (phase C).  No
// load depends on another: neighbour traces sit in the element's own slots,
// written there by the neighbours in the previous stage.  sw_process runs
// the four phases on the loaded values.
// Per-thread values of one node (and of its side node): sw_load only issues
// the loads (raw values, no arithmetic, so no wait sits between them and
// every load of the phase is in flight at once); sw_finish derives the
// working values once the loads have all been issued.
struct SWPre {
    double s[3], met[4], invJ, hx, hy, vort, dragH, gold[3], r[4], fn0[3], oth0[3];
    int2 info0;
    // raw loads
    double Jv, hRaw, metM[5], fnM[3];
    float metD[5], fnD[3], hx32, hy32, vort32, g32[3];
};

// (nx, ny, |n|) of side node k of side sd of element e, in the edge's
// primary orientation: FP64, or (mixed) the side's FP64 reference value plus
// an FP32 deviation (exact for straight sides, where the deviation is 0).
template <int N, bool MIXED>
__device__ __forceinline__ void side_normal(const SWKernelArgs &A, size_t e, int sd, int k, double fn[3])
{
    constexpr int n1 = N + 1;
    const size_t es = e * 4 + sd;
    if constexpr (MIXED) {
        #pragma unroll
        for (int c = 0; c < 3; ++c) fn[c] = A.fnM[es * 3 + c] + (double)A.fnD[(es * 3 + c) * n1 + k];
    } else {
        const size_t nb = es * 3 * n1 + k;
        fn[0] = A.fnrm[nb]; fn[1] = A.fnrm[nb + n1]; fn[2] = A.fnrm[nb + 2 * n1];
    }
}

// Loads of one node (element e, node nd), and for nd < 4 (N+1) of that
// element's side node nd (side nd / (N+1), node nd mod (N+1)).  Every
// thread loads from an address that is valid whatever its node: node 0 of
// element 0 for a thread without a node (act false: it computes nothing
// from them), side node nd mod 4 (N+1) for a thread without a side node
// (its side values are never read).  No branch depends on the thread, so
// the loads stay in flight together and each phase waits only for the loads
// it uses: behind a thread-dependent branch the compiler waits for all
// outstanding loads where the branch joins, and no phase could start before
// the element's last load had arrived.  Optional fields (RK register, depth
// for the drag, relaxation) stay under uniform tests.
template <int N, bool MIXED>
__device__ __forceinline__ void sw_load(const SWKernelArgs &A, int e, int nd, bool act, SWPre &P)
{
    constexpr int n1 = N + 1, np = n1 * n1;
    const size_t mstride = (size_t)A.nE * np;
    const size_t base = act ? (size_t)e * np + nd : 0;
    const int ec = act ? e : 0;
    const int ns = nd % (4 * n1);
    const int sd0 = ns / n1, k0 = ns - sd0 * n1;
    const size_t es = (size_t)ec * 4 + sd0;
    P.Jv = 1.0; P.hRaw = 1.0;
    #pragma unroll
    for (int k = 0; k < 3; ++k) {
        P.s[k] = 0; P.gold[k] = 0; P.fn0[k] = 0; P.oth0[k] = 0; P.g32[k] = 0; P.fnD[k] = 0; P.fnM[k] = 0;
    }
    #pragma unroll
    for (int k = 0; k < 4; ++k) { P.r[k] = 0; P.met[k] = 0; }
    #pragma unroll
    for (int k = 0; k < 5; ++k) { P.metD[k] = 0; P.metM[k] = 0; }
    P.hx = 0; P.hy = 0; P.vort = 0; P.hx32 = 0; P.hy32 = 0; P.vort32 = 0;
    P.info0 = A.side[es];
    if constexpr (MIXED) {
        #pragma unroll
        for (int c = 0; c < 3; ++c) { P.fnM[c] = A.fnM[es * 3 + c]; P.fnD[c] = A.fnD[(es * 3 + c) * n1 + k0]; }
    } else {
        const size_t nb = es * 3 * n1 + k0;
        P.fn0[0] = A.fnrm[nb]; P.fn0[1] = A.fnrm[nb + n1]; P.fn0[2] = A.fnrm[nb + 2 * n1];
    }
    const double *tr = A.trIn + es * 3 * n1 + k0;
    P.oth0[0] = tr[0]; P.oth0[1] = tr[n1]; P.oth0[2] = tr[2 * n1];
    #pragma unroll
    for (int f = 0; f < 3; ++f) P.s[f] = A.q[f * mstride + base];
    if constexpr (MIXED) {
        #pragma unroll
        for (int k = 0; k < 5; ++k) { P.metM[k] = A.metM[(size_t)ec * 5 + k]; P.metD[k] = A.metD[k * mstride + base]; }
        P.hx32 = A.bathy32[base]; P.hy32 = A.bathy32[mstride + base]; P.vort32 = A.plv32[base];
        // the optional fields after the others (their uniform tests would
        // otherwise let the compiler issue them first)
        asm volatile("" ::: "memory");
        if (A.a != 0.0) {
            #pragma unroll
            for (int f = 0; f < 3; ++f) P.g32[f] = A.G32[f * mstride + base];
        }
    } else {
        #pragma unroll
        for (int k = 0; k < 4; ++k) P.met[k] = A.metric[k * mstride + base];
        P.Jv = A.metric[4 * mstride + base];
        P.hx = A.bathy[mstride + base]; P.hy = A.bathy[2 * mstride + base]; P.vort = A.plv[base];
        asm volatile("" ::: "memory");
        if (A.a != 0.0) {
            #pragma unroll
            for (int f = 0; f < 3; ++f) P.gold[f] = A.G64[f * mstride + base];
        }
    }
    if (A.drag != 0.0) P.hRaw = A.bathy[base];
    if (A.hasRelax) {
        #pragma unroll
        for (int k = 0; k < 4; ++k) P.r[k] = A.relax[k * mstride + base];
    }
}

// Working values from the raw loads (after every load has been issued).
template <bool MIXED>
__device__ __forceinline__ void sw_finish(const SWKernelArgs &A, bool act, SWPre &P)
{
    if constexpr (MIXED) {
        #pragma unroll
        for (int k = 0; k < 4; ++k) P.met[k] = P.metM[k] + (double)P.metD[k];
        P.Jv = P.metM[4] + (double)P.metD[4];
        #pragma unroll
        for (int c = 0; c < 3; ++c) P.fn0[c] = P.fnM[c] + (double)P.fnD[c];
        P.hx = P.hx32; P.hy = P.hy32; P.vort = P.vort32;
        #pragma unroll
        for (int f = 0; f < 3; ++f) P.gold[f] = P.g32[f];
    }
    P.invJ = act ? (MIXED ? rcp_refined(P.Jv) : 1.0 / P.Jv) : 0.0;
    P.dragH = 0.0;
    if (A.drag != 0.0) P.dragH = A.drag / P.hRaw;  // LinearDrag / depth
}

// Rusanov flux of one side node in the edge's primary orientation (so the
// value is the one EdgeFlux stores on both sides), scaled by |n| and signed
// for this side: one Riemann evaluation whichever way the side faces.
template <int N, bool MIXED>
__device__ __forceinline__ void side_flux(const SWKernelArgs &A, const NodalLDS<N> &L, const double *qsT, int sd,
                                          int k, int2 info, const double fn[3], double oth[3], double out[3])
{
    double own[3], lft[3], rgt[3], fl[3];
    trace3<N>(L, qsT, sd, k, own);
    const bool bnd = info.x < 0;
    if (bnd) sw_external(info.x, fn[0], fn[1], own, oth);
    const bool prim = bnd || side_primary(info.y);
    #pragma unroll
    for (int f = 0; f < 3; ++f) {
        lft[f] = prim ? own[f] : oth[f];
        rgt[f] = prim ? oth[f] : own[f];
    }
    if constexpr (MIXED) sw_riemann_mixed(lft, rgt, fn[0], fn[1], A.g, fl, A.variant == 1);
    else sw_riemann(lft, rgt, fn[0], fn[1], A.g, fl, A.variant == 1);
    const double sign = prim ? 1.0 : -1.0;
    #pragma unroll
    for (int f = 0; f < 3; ++f) out[f] = sign * (fl[f] * fn[2]);
}

// Synchronization between the phases of one element: the whole workgroup,
// or (WL) only the element's own wavefront, when every element lies inside
// one wavefront (64 % (N+1)^2 == 0) and the nodal constants are per wave.
template <bool WL>
__device__ __forceinline__ void phase_sync()
{
    if constexpr (WL) wave_sync();
    else __syncthreads();
}

// Every thread of the block (WL: of the wavefront) calls this.
template <int N, bool MIXED, bool WL>
__device__ __forceinline__ void sw_process(const SWKernelArgs &A, const NodalLDS<N> &L, double *qsE, double *sCE,
                                           double *pCE, float *sDE, float *pDE, double *bfE, int e, int nd, bool act,
                                           SWPre &P)
{
    sw_finish<MIXED>(A, act, P);
    constexpr int n1 = N + 1, np = n1 * n1;
    const int i = nd % n1, j = nd / n1;
    const size_t mstride = (size_t)A.nE * np, base = (size_t)e * np + nd;

    // ---- phase A: contravariant fluxes, state to LDS
    if (act) {
        #pragma unroll
        for (int f = 0; f < 3; ++f) qsE[f * np + nd] = P.s[f];
        contravariant<MIXED>(P.s, P.met, A.g, sCE + nd, pCE + nd, np);
    }
    phase_sync<WL>();

    // ---- phase B: FP32 deviations of the contravariant fluxes (mixed), and
    // the Rusanov flux of each side node in the edge's primary orientation
    if constexpr (MIXED) {
        if (act) {
            #pragma unroll
            for (int f = 0; f < 3; ++f) {
                sDE[f * np + nd] = (float)(sCE[f * np + nd] - sCE[f * np + n1 * j]);
                pDE[f * np + nd] = (float)(pCE[f * np + nd] - pCE[f * np + i]);
            }
        }
    }
    if (act) {
        for (int w = nd; w < 4 * n1; w += np) {
            const int sd = w / n1, k = w - sd * n1;
            int2 info;
            double fn[3], oth[3], fl[3];
            if (w == nd) {
                info = P.info0;
                fn[0] = P.fn0[0]; fn[1] = P.fn0[1]; fn[2] = P.fn0[2];
                oth[0] = P.oth0[0]; oth[1] = P.oth0[1]; oth[2] = P.oth0[2];
            } else {  // N <= 2: more side nodes than element nodes
                info = A.side[(size_t)e * 4 + sd];
                side_normal<N, MIXED>(A, e, sd, k, fn);
                const double *tr = A.trIn + (((size_t)e * 4 + sd) * 3) * n1 + k;
                oth[0] = tr[0]; oth[1] = tr[n1]; oth[2] = tr[2 * n1];
            }
            side_flux<N, MIXED>(A, L, qsE, sd, k, info, fn, oth, fl);
            #pragma unroll
            for (int f = 0; f < 3; ++f) bfE[(sd * 3 + f) * n1 + k] = fl[f];
        }
    }
    phase_sync<WL>();

    // ---- phase C: weak divergence + lift + source + RK3 update
    if (act) {
        double tend[3], src[3];
        if constexpr (MIXED) dg_divergence_dev32<N>(L, sCE, pCE, sDE, pDE, bfE, i, j, P.invJ, tend);
        else dg_divergence64<N, (N <= 5 || N % 2 == 1)>(L, sCE, pCE, bfE, i, j, P.invJ, tend);
        sw_source(P.s, P.vort, A.g, P.hx, P.hy, P.dragH, A.hasRelax, P.r, src);
        #pragma unroll
        for (int f = 0; f < 3; ++f) {
            // G = a*G + L; the first stage has a = 0 and ForwardStepRK3 zeroes
            // G before it, so that stage never reads G.
            double gv = A.a * P.gold[f] + (tend[f] + src[f]);
            if constexpr (MIXED) {
                gv = (double)(float)gv;
                A.G32[f * mstride + base] = (float)gv;
            } else {
                A.G64[f * mstride + base] = gv;
            }
            const double qn = P.s[f] + A.gdt * gv;
            A.q[f * mstride + base] = qn;
            qsE[f * np + nd] = qn;  // phase B finished reading qs before the barrier
        }
    }
    phase_sync<WL>();

    // ---- phase D: traces of the updated state, pushed into the
    // neighbours' slots for the next stage
    if (act) {
        for (int w = nd; w < 4 * n1; w += np) {
            const int sd = w / n1, k = w - sd * n1;
            const int2 info = (w == nd) ? P.info0 : A.side[(size_t)e * 4 + sd];
            if (info.x < 0) continue;
            double tr[3];
            trace3<N>(L, qsE, sd, k, tr);
            const int kk = side_flip(info.y) ? N - k : k;
            double *o = A.trOut + (((size_t)info.x * 4 + side_nbr_side(info.y)) * 3) * n1 + kk;
            o[0] = tr[0]; o[n1] = tr[1]; o[2 * n1] = tr[2];
        }
    }
}

// One RK stage for G element groups per workgroup.  G = 2 issues the
// second group's loads before the first group computes, so its memory
// traffic overlaps the first group's arithmetic (the stage is otherwise
// one round of waves that all load, then all compute).
//
// Ensembles: the grid holds `members` workgroups per element group, element
// group major, so the members of one group run side by side on one XCD and
// share its geometry in L2; each member has its own q, G and traces.
// LDS of one fused stage workgroup.  WL: each element lies inside one
// wavefront, so the phases synchronize per wavefront and the wavefronts of
// a workgroup never wait for each other (one that has its data computes
// while the others still load); each wavefront stages its own copy of the
// nodal constants.  Measured at N = 7 (64x64, A/B on one box): FP64 11.67
// -> 11.31 us per stage against workgroup barriers.
template <int N, bool MIXED, int G>
struct SwStageLds {
    using S = Shape<N>;
    static constexpr int np = S::np, epb = S::epb, n1 = S::n1;
    static constexpr bool WL = (G == 1) && (64 % np == 0);
    NodalLDS<N> Lw[WL ? 4 : 1];
    double qs[epb * 3 * np], sC[epb * 3 * np], pC[epb * 3 * np], bfs[epb * 12 * n1];
    float sD[MIXED ? epb * 3 * np : 1], pD[MIXED ? epb * 3 * np : 1];
};

// One RK stage of element group wg (G groups of epb elements; with
// ensembles, wg also selects the member) in the LDS block U.
template <int N, bool MIXED, int G>
__device__ __forceinline__ void sw_stage_group(const SWKernelArgs &A0, int wg, SwStageLds<N, MIXED, G> &U)
{
    using S = Shape<N>;
    using Lds = SwStageLds<N, MIXED, G>;
    constexpr int np = S::np, epb = S::epb, n1 = S::n1;
    constexpr bool WL = Lds::WL;
    SWKernelArgs A = A0;
    if (A0.members > 1) {  // uniform: a single state skips the member arithmetic
        const int mbr = wg % A0.members;
        wg /= A0.members;
        const size_t qOff = (size_t)mbr * 3 * A0.nE * np, tOff = (size_t)mbr * A0.nE * 12 * n1;
        A.q += qOff;
        if (MIXED) A.G32 += qOff;
        else A.G64 += qOff;
        A.trIn += tOff;
        A.trOut += tOff;
        A.gdt = A0.gcoef * A0.dtM[mbr];
    }
    NodalLDS<N> &L = U.Lw[WL ? threadIdx.x / 64 : 0];
    const int t = threadIdx.x, le = t / np, nd = t - le * np;
    constexpr int kLanes = WL ? 64 : 256;
    const int lane = WL ? threadIdx.x % 64 : threadIdx.x;
    NodalRegs<N, MIXED, kLanes> R;
    nodal_fetch<N, MIXED, kLanes>(R, A.dS, A.dS32, A.dP32, lane);
    SWPre P[G];
    int e[G];
    bool act[G];
    #pragma unroll
    for (int g = 0; g < G; ++g) {
        e[g] = (wg * G + g) * epb + le;
        act[g] = le < epb && e[g] < A.nE;
        sw_load<N, MIXED>(A, e[g], nd, act[g], P[g]);
    }
    nodal_put<N, MIXED, kLanes>(L, R, lane);
    if constexpr (WL) wave_sync();
    const int ls = le < epb ? le : 0;
    double *qsE = U.qs + ls * 3 * np, *sCE = U.sC + ls * 3 * np, *pCE = U.pC + ls * 3 * np;
    double *bfE = U.bfs + ls * 12 * n1;
    float *sDE = MIXED ? U.sD + ls * 3 * np : U.sD, *pDE = MIXED ? U.pD + ls * 3 * np : U.pD;
    #pragma unroll
    for (int g = 0; g < G; ++g) {
        if (g > 0) __syncthreads();  // the previous group's phase D read qs
        sw_process<N, MIXED, WL>(A, L, qsE, sCE, pCE, sDE, pDE, bfE, e[g], nd, act[g], P[g]);
    }
}

// One RK stage for G element groups per workgroup.  G = 2 issues the
// second group's loads before the first group computes, so its memory
// traffic overlaps the first group's arithmetic (the stage is otherwise
// one round of waves that all load, then all compute).
//
// Ensembles: the grid holds `members` workgroups per element group, element
// group major, so the members of one group run side by side on one XCD and
// share its geometry in L2; each member has its own q, G and traces.
template <int N, bool MIXED, int G>
__global__ __launch_bounds__(256, 4 / G) void k_sw_stage_fused(SWKernelArgs A0)
{
    __shared__ SwStageLds<N, MIXED, G> U;
    sw_stage_group<N, MIXED, G>(A0, xcd_block(blockIdx.x, gridDim.x), U);
}

// Models of uneven degree in one persistent launch: the grid (a few
// workgroups per CU) walks a list of tiles, each tile one element group (x
// member) of one model; the model's degree selects the code path and the
// shared LDS block is reinterpreted for it.  Every model advances one RK
// stage per launch.
constexpr int kMultiMax = 8;  // models per launch (the arguments travel as kernel arguments)

struct SWMultiArgs {
    int nItems, totalTiles;
    int N[kMultiMax], tileStart[kMultiMax];
    SWKernelArgs A[kMultiMax];
};

template <bool MIXED>
union SwMultiLds {
    SwStageLds<1, MIXED, 1> n1;
    SwStageLds<2, MIXED, 1> n2;
    SwStageLds<3, MIXED, 1> n3;
    SwStageLds<4, MIXED, 1> n4;
    SwStageLds<5, MIXED, 1> n5;
    SwStageLds<6, MIXED, 1> n6;
    SwStageLds<7, MIXED, 1> n7;
    SwStageLds<8, MIXED, 1> n8;
    SwStageLds<9, MIXED, 1> n9;
    SwStageLds<10, MIXED, 1> n10;
    SwStageLds<11, MIXED, 1> n11;
};

// MENU 0: degrees 3, 5 and 7 only (the serving mix); 1: every degree.
// One kernel has one register budget, so the fewer code paths it holds, the
// less the paths that need more registers than the others make it spill.
template <bool MIXED, int MENU>
__global__ __launch_bounds__(256, 4) void k_sw_stage_multi(SWMultiArgs M)
{
    __shared__ SwMultiLds<MIXED> U;
    for (int tile = blockIdx.x; tile < M.totalTiles; tile += gridDim.x) {
        int m = 0;
        while (m + 1 < M.nItems && M.tileStart[m + 1] <= tile) ++m;
        const int wg = tile - M.tileStart[m];
        if constexpr (MENU == 0) {
            switch (M.N[m]) {
            case 3: sw_stage_group<3, MIXED, 1>(M.A[m], wg, U.n3); break;
            case 5: sw_stage_group<5, MIXED, 1>(M.A[m], wg, U.n5); break;
            default: sw_stage_group<7, MIXED, 1>(M.A[m], wg, U.n7); break;
            }
        } else {
            switch (M.N[m]) {
            case 1: sw_stage_group<1, MIXED, 1>(M.A[m], wg, U.n1); break;
            case 2: sw_stage_group<2, MIXED, 1>(M.A[m], wg, U.n2); break;
            case 3: sw_stage_group<3, MIXED, 1>(M.A[m], wg, U.n3); break;
            case 4: sw_stage_group<4, MIXED, 1>(M.A[m], wg, U.n4); break;
            case 5: sw_stage_group<5, MIXED, 1>(M.A[m], wg, U.n5); break;
            case 6: sw_stage_group<6, MIXED, 1>(M.A[m], wg, U.n6); break;
            case 7: sw_stage_group<7, MIXED, 1>(M.A[m], wg, U.n7); break;
            case 8: sw_stage_group<8, MIXED, 1>(M.A[m], wg, U.n8); break;
            case 9: sw_stage_group<9, MIXED, 1>(M.A[m], wg, U.n9); break;
            case 10: sw_stage_group<10, MIXED, 1>(M.A[m], wg, U.n10); break;
            default: sw_stage_group<11, MIXED, 1>(M.A[m], wg, U.n11); break;
            }
        }
        __syncthreads();  // the next tile reuses the LDS block
    }
}

// CalculateBathymetryAtBoundaries + CalculateBathymetryGradient (setup)
template <int N>
__global__ __launch_bounds__(256) void k_sw_bathy_grad(SWKernelArgs A, double *bathy, float *bathy32)
{
    using S = Shape<N>;
    constexpr int n1 = S::n1, np = S::np, epb = S::epb;
    __shared__ double hs[epb * np], sC[epb * 2 * np], pC[epb * 2 * np], bfs[epb * 8 * n1];
    const int t = threadIdx.x, le = t / np, nd = t - le * np;
    const int e = blockIdx.x * epb + le;
    const bool act = le < epb && e < A.nE;
    const int i = nd % n1, j = nd / n1;
    const size_t mstride = (size_t)A.nE * np, base = (size_t)e * np + nd;
    double lW[n1], lE[n1], lS[n1], lN[n1];
    load_lag<N>(A.lag, lW, lE, lS, lN);
    // variant 1 (the examples' class, :1754-1825): the weak gradient of
    // h^2/2, divided by J h at the end
    const bool half2 = A.variant == 1;
    if (act) {
        const double h = bathy[base];
        hs[le * np + nd] = h;
        const double hf = half2 ? (h * h) * 0.5 : h;
        const double dxds = A.metric[base], dxdp = A.metric[mstride + base];
        const double dyds = A.metric[2 * mstride + base], dydp = A.metric[3 * mstride + base];
        // xF = (hf, 0), yF = (0, hf)
        sC[(le * 2 + 0) * np + nd] = dydp * hf - dxdp * 0.0;
        sC[(le * 2 + 1) * np + nd] = dydp * 0.0 - dxdp * hf;
        pC[(le * 2 + 0) * np + nd] = -dyds * hf + dxds * 0.0;
        pC[(le * 2 + 1) * np + nd] = -dyds * 0.0 + dxds * hf;
    }
    __syncthreads();
    if (act) {
        for (int w = nd; w < 4 * n1; w += np) {
            const int sd = w / n1, k = w - sd * n1;
            const double *hb = hs + le * np;
            double tr;
            switch (sd) {
            case 0: tr = line_dot<N>(lS, hb + k, n1); break;
            case 1: tr = line_dot<N>(lE, hb + n1 * k, 1); break;
            case 2: tr = line_dot<N>(lN, hb + k, n1); break;
            default: tr = line_dot<N>(lW, hb + n1 * k, 1); break;
            }
            if (half2) tr = (0.5 * tr) * tr;
            const size_t nb = ((size_t)e * 4 + sd) * 3 * n1 + k;
            const double nx = A.nrm[nb], ny = A.nrm[nb + n1], len = A.nrm[nb + 2 * n1];
            bfs[le * 8 * n1 + (sd * 2 + 0) * n1 + k] = (tr * nx) * len;
            bfs[le * 8 * n1 + (sd * 2 + 1) * n1 + k] = (tr * ny) * len;
        }
    }
    __syncthreads();
    if (!act) return;
    const double *bf = bfs + le * 8 * n1;
    const double Jv = half2 ? A.metric[4 * mstride + base] * hs[le * np + nd] : A.metric[4 * mstride + base];
    for (int qd = 0; qd < 2; ++qd) {
        const double *sl = sC + (le * 2 + qd) * np + n1 * j;
        double t1 = A.dS[i] * sl[0];
        for (int m = 1; m < n1; ++m) t1 = __builtin_fma(A.dS[i + n1 * m], sl[m], t1);
        t1 = t1 + (bf[(1 * 2 + qd) * n1 + j] * lE[i] + bf[(3 * 2 + qd) * n1 + j] * lW[i]) / A.qwS[i];
        const double *pl = pC + (le * 2 + qd) * np + i;
        double t2 = A.dP[j] * pl[0];
        for (int m = 1; m < n1; ++m) t2 = __builtin_fma(A.dP[j + n1 * m], pl[m * n1], t2);
        t2 = t2 + (bf[(2 * 2 + qd) * n1 + i] * lN[j] + bf[(0 * 2 + qd) * n1 + i] * lS[j]) / A.qwP[j];
        const double gval = (t1 + t2) / Jv;
        bathy[(1 + qd) * mstride + base] = gval;
        if (bathy32) bathy32[qd * mstride + base] = (float)gval;
    }
}

}  // namespace
}  // namespace self

// ==================================================================== host
using namespace self;

struct self_sw2d_s {
    self_sw2d_desc d{};
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int n1 = 0, np = 0;
    size_t nE = 0;
    // device buffers
    double *q = nullptr, *G64 = nullptr, *tend = nullptr, *metric = nullptr, *bathy = nullptr;
    float *G32 = nullptr, *bathy32 = nullptr, *plv32 = nullptr, *dS32 = nullptr, *dP32 = nullptr;
    double *plv = nullptr, *relax = nullptr, *nrm = nullptr, *fnrm = nullptr;
    double *metM = nullptr, *fnM = nullptr;  // mixed: reference values of the geometry
    float *metD = nullptr, *fnD = nullptr;   // mixed: FP32 deviations from them
    double *bsol = nullptr, *bflux = nullptr, *tr[2] = {nullptr, nullptr};
    double *nodal = nullptr;  // dS, dP, qwS, qwP, lag[4]
    int2 *side = nullptr;
    int *edges = nullptr;
    int trCur = 0;
    bool trValid = false, nodalSet = false, geomSet = false, hasRelax = false;
    int groups = 0;  // element groups per workgroup of the fused kernel (0: heuristic)
    int members = 1; // ensemble members (fused form)
    int variant = 0; // 1: the examples' class (self_sw2d_set_variant)
    double *dtM = nullptr;          // per-member time steps (device, members > 1)
    std::vector<double> dtH;        // host copy
    size_t bytes = 0;
    // two RK3 steps (six fused stages) captured as one graph, replayed while
    // the six launches it holds are exactly the ones the model would issue
    bool useGraph = true;
    hipGraphExec_t graphExec = nullptr;
    SWKernelArgs graphArgs[6];
    int graphDims[4] = {0, 0, 0, 0};
};

namespace {

template <typename T>
int dalloc(self_sw2d_s *h, T **p, size_t n)
{
    if (n == 0) { *p = nullptr; return SELF_OK; }
    SELF_HIP(hipMalloc(reinterpret_cast<void **>(p), n * sizeof(T)));
    h->bytes += n * sizeof(T);
    return SELF_OK;
}

void free_all(self_sw2d_s *h)
{
    void *ptrs[] = {h->q, h->G64, h->tend, h->metric, h->bathy, h->G32, h->bathy32, h->plv32,
                    h->dS32, h->dP32, h->plv, h->relax, h->nrm, h->fnrm, h->bsol, h->bflux,
                    h->tr[0], h->tr[1], h->nodal, h->side, h->edges, h->metM, h->fnM, h->metD, h->fnD,
                    h->dtM};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
}

// fills A member by member (callers that compare argument blocks bytewise
// zero the object first, so its padding stays zero)
void fill_args(self_sw2d_s *h, SWKernelArgs &A)
{
    const int n1 = h->n1;
    A.nE = (int)h->nE;
    A.members = h->members;
    A.dtM = h->dtM;
    A.g = h->d.g;
    A.variant = h->variant;
    // the examples' class has no drag and no relaxation in its Source
    A.drag = h->variant == 1 ? 0.0 : h->d.linearDrag;
    A.hasRelax = (h->hasRelax && h->variant != 1) ? 1 : 0;
    A.dS = h->nodal;
    A.dP = h->nodal + n1 * n1;
    A.qwS = h->nodal + 2 * n1 * n1;
    A.qwP = A.qwS + n1;
    A.lag = A.qwP + n1;
    A.dS32 = h->dS32;
    A.dP32 = h->dP32;
    A.q = h->q;
    A.G64 = h->G64;
    A.G32 = h->G32;
    A.tend = h->tend;
    A.metric = h->metric;
    A.metM = h->metM;
    A.metD = h->metD;
    A.fnM = h->fnM;
    A.fnD = h->fnD;
    A.bathy = h->bathy;
    A.bathy32 = h->bathy32;
    A.plv = h->plv;
    A.plv32 = h->plv32;
    A.relax = h->relax;
    A.nrm = h->nrm;
    A.fnrm = h->fnrm;
    A.side = h->side;
    A.bflux = h->bflux;
}

SWKernelArgs make_args(self_sw2d_s *h)
{
    SWKernelArgs A{};
    fill_args(h, A);
    return A;
}

#define SELF_DISPATCH_N(N_, CALL)                                                   \
    switch (N_) {                                                                   \
    case 1: { constexpr int NN = 1; CALL; } break;                                  \
    case 2: { constexpr int NN = 2; CALL; } break;                                  \
    case 3: { constexpr int NN = 3; CALL; } break;                                  \
    case 4: { constexpr int NN = 4; CALL; } break;                                  \
    case 5: { constexpr int NN = 5; CALL; } break;                                  \
    case 6: { constexpr int NN = 6; CALL; } break;                                  \
    case 7: { constexpr int NN = 7; CALL; } break;                                  \
    case 8: { constexpr int NN = 8; CALL; } break;                                  \
    case 9: { constexpr int NN = 9; CALL; } break;                                  \
    case 10: { constexpr int NN = 10; CALL; } break;                                \
    case 11: { constexpr int NN = 11; CALL; } break;                                \
    default: break;                                                                 \
    }

template <int N>
int grid_for(size_t nE)
{
    return (int)((nE + Shape<N>::epb - 1) / Shape<N>::epb);
}

// element groups per workgroup of the fused stage kernel: the handle's
// setting (self_sw2d_set_groups / self_sw2d_autotune), else one (halving the
// waves per SIMD to prefetch a second group costs more latency hiding than
// the overlap gains at 64x64, N = 7)
inline int fused_groups(int setting, size_t nE, int epb)
{
    (void)nE;
    (void)epb;
    return (setting == 2) ? 2 : 1;  // measured: one group is faster at 64x64, N = 7
}

template <int N>
int grid_fused(size_t nE, int G)
{
    const size_t per = (size_t)Shape<N>::epb * G;
    return (int)((nE + per - 1) / per);
}

int launch_traces(self_sw2d_s *h, const double *src, double *dst, int nf)
{
    const double *lag = h->nodal + 2 * h->n1 * h->n1 + 2 * h->n1;
    SELF_DISPATCH_N(h->d.N, hipLaunchKernelGGL(k_sw_traces<NN>, dim3(grid_for<NN>(h->nE)), dim3(256), 0,
                                               h->stream, src, lag, dst, (int)h->nE, nf));
    SELF_HIP(hipGetLastError());
    return SELF_OK;
}

int launch_traces_push(self_sw2d_s *h, double *dst)
{
    const double *lag = h->nodal + 2 * h->n1 * h->n1 + 2 * h->n1;
    SELF_DISPATCH_N(h->d.N, hipLaunchKernelGGL(k_sw_traces_push<NN>, dim3(grid_for<NN>(h->nE), h->members),
                                               dim3(256), 0, h->stream, h->q, lag, h->side, dst, (int)h->nE));
    SELF_HIP(hipGetLastError());
    return SELF_OK;
}

int launch_edge_flux(self_sw2d_s *h)
{
    const int total = h->d.nEdges * h->n1;
    const int grid = (total + 255) / 256;
    SELF_DISPATCH_N(h->d.N, hipLaunchKernelGGL(k_sw_edge_flux<NN>, dim3(grid), dim3(256), 0, h->stream,
                                               h->edges, h->d.nEdges, h->bsol, h->nrm, h->d.g, h->bflux,
                                               h->variant));
    SELF_HIP(hipGetLastError());
    return SELF_OK;
}

int launch_mapped(self_sw2d_s *h)
{
    SWKernelArgs A = make_args(h);
    SELF_DISPATCH_N(h->d.N, hipLaunchKernelGGL(k_sw_mapped<NN>, dim3(grid_for<NN>(h->nE)), dim3(256), 0,
                                               h->stream, A));
    SELF_HIP(hipGetLastError());
    return SELF_OK;
}

int gtd_plain(self_sw2d_s *h)
{
    int rc = launch_traces(h, h->q, h->bsol, 3);
    if (!rc) rc = launch_edge_flux(h);
    if (!rc) rc = launch_mapped(h);
    return rc;
}

int step_plain(self_sw2d_s *h)
{
    const size_t n = 3 * h->nE * h->np;
    for (int m = 0; m < 3; ++m) {
        int rc = gtd_plain(h);
        if (rc) return rc;
        hipLaunchKernelGGL(k_sw_rk_update, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, h->q,
                           h->G64, h->tend, n, kRkA[m], kRkG[m] * h->d.dt);
        SELF_HIP(hipGetLastError());
    }
    return SELF_OK;
}

// one fused stage for every ensemble member: grid_fused x members workgroups
int launch_stage_fused(self_sw2d_s *h, const SWKernelArgs &A)
{
    const bool two = fused_groups(h->groups, h->nE, 0) == 2;
    const unsigned M = (unsigned)h->members;
    if (h->d.precision == SELF_MIXED) {
        SELF_DISPATCH_N(h->d.N, if (two) { hipLaunchKernelGGL((k_sw_stage_fused<NN, true, 2>), dim3(grid_fused<NN>(h->nE, 2) * M), dim3(256), 0, h->stream, A); } else { hipLaunchKernelGGL((k_sw_stage_fused<NN, true, 1>), dim3(grid_fused<NN>(h->nE, 1) * M), dim3(256), 0, h->stream, A); });
    } else {
        SELF_DISPATCH_N(h->d.N, if (two) { hipLaunchKernelGGL((k_sw_stage_fused<NN, false, 2>), dim3(grid_fused<NN>(h->nE, 2) * M), dim3(256), 0, h->stream, A); } else { hipLaunchKernelGGL((k_sw_stage_fused<NN, false, 1>), dim3(grid_fused<NN>(h->nE, 1) * M), dim3(256), 0, h->stream, A); });
    }
    SELF_HIP(hipGetLastError());
    return SELF_OK;
}

int step_fused(self_sw2d_s *h)
{
    if (!h->trValid) {
        int rc = launch_traces_push(h, h->tr[h->trCur]);
        if (rc) return rc;
        h->trValid = true;
    }
    SWKernelArgs A = make_args(h);
    for (int m = 0; m < 3; ++m) {
        A.a = kRkA[m];
        A.gdt = kRkG[m] * h->d.dt;
        A.gcoef = kRkG[m];
        A.trIn = h->tr[h->trCur];
        A.trOut = h->tr[1 - h->trCur];
        if (int rc = launch_stage_fused(h, A)) return rc;
        h->trCur = 1 - h->trCur;
    }
    return SELF_OK;
}

// nPairs x two RK3 steps of the fused form through the captured graph
int steps_fused_graph(self_sw2d_s *h, int nPairs)
{
    if (!h->trValid) {
        int rc = launch_traces_push(h, h->tr[h->trCur]);
        if (rc) return rc;
        h->trValid = true;
    }
    SWKernelArgs cand[6];
    std::memset(static_cast<void *>(cand), 0, sizeof(cand));
    int cur = h->trCur;
    for (int k = 0; k < 6; ++k) {
        const int m = k % 3;
        fill_args(h, cand[k]);
        cand[k].a = kRkA[m];
        cand[k].gdt = kRkG[m] * h->d.dt;
        cand[k].gcoef = kRkG[m];
        cand[k].trIn = h->tr[cur];
        cand[k].trOut = h->tr[1 - cur];
        cur = 1 - cur;
    }
    const int dims[4] = {fused_groups(h->groups, h->nE, 0), h->members, h->d.N, h->d.precision};
    const bool same = h->graphExec && std::memcmp(cand, h->graphArgs, sizeof(cand)) == 0 &&
                      std::memcmp(dims, h->graphDims, sizeof(dims)) == 0;
    if (!same) {
        if (h->graphExec) {
            SELF_HIP(hipGraphExecDestroy(h->graphExec));
            h->graphExec = nullptr;
        }
        // a capture the runtime refuses leaves nothing launched: the handle
        // falls back to host launches for this and every later call
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal);
        int rc = SELF_OK;
        if (e == hipSuccess) {
            for (int k = 0; k < 6 && !rc; ++k) rc = launch_stage_fused(h, cand[k]);
            e = hipStreamEndCapture(h->stream, &g);
        }
        if (e == hipSuccess && !rc) e = hipGraphInstantiate(&h->graphExec, g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
        if (e != hipSuccess || rc) {
            (void)hipGetLastError();
            h->graphExec = nullptr;
            h->useGraph = false;
            for (int k = 0; k < 2 * nPairs; ++k)
                if (int r2 = step_fused(h)) return r2;
            return SELF_OK;
        }
        std::memcpy(static_cast<void *>(h->graphArgs), cand, sizeof(cand));
        std::memcpy(h->graphDims, dims, sizeof(dims));
    }
    for (int p = 0; p < nPairs; ++p) SELF_HIP(hipGraphLaunch(h->graphExec, h->stream));
    return SELF_OK;
}

// Partial sums of the mixed-form guard over the momentum fields: per
// workgroup, sum (G - T)^2 and sum T^2 (G the mixed stage's FP32 register
// after a first stage, T the FP64 tendency), fixed order.
__global__ __launch_bounds__(256) void k_sw_guard_sums(const float *__restrict__ gm, const double *__restrict__ t64,
                                                       size_t n, double *__restrict__ part)
{
    __shared__ double sd[256], st[256];
    double d2 = 0.0, t2 = 0.0;
    for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (size_t)gridDim.x * 256) {
        const double t = t64[k], d = (double)gm[k] - t;
        d2 += d * d;
        t2 += t * t;
    }
    sd[threadIdx.x] = d2;
    st[threadIdx.x] = t2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            sd[threadIdx.x] += sd[threadIdx.x + w];
            st[threadIdx.x] += st[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sd[0];
        part[2 * blockIdx.x + 1] = st[0];
    }
}

int check_ready(self_sw2d_s *h)
{
    SELF_REQUIRE(h, SELF_ERR_ARG, "null handle");
    SELF_REQUIRE(h->nodalSet && h->geomSet, SELF_ERR_STATE,
                 "shallow-water handle needs self_sw2d_set_nodal and self_sw2d_set_geometry first");
    SELF_HIP(hipSetDevice(h->device));
    return SELF_OK;
}

// host layout [e][f][node] <-> device [f][e][node]
void to_device_layout(const double *src, double *dst, size_t nE, int nf, int np)
{
    for (size_t e = 0; e < nE; ++e)
        for (int f = 0; f < nf; ++f)
            std::memcpy(dst + ((size_t)f * nE + e) * np, src + (e * nf + f) * np, sizeof(double) * np);
}

void from_device_layout(const double *src, double *dst, size_t nE, int nf, int np)
{
    for (size_t e = 0; e < nE; ++e)
        for (int f = 0; f < nf; ++f)
            std::memcpy(dst + (e * nf + f) * np, src + ((size_t)f * nE + e) * np, sizeof(double) * np);
}

}  // namespace

extern "C" {

int self_sw2d_create(const self_sw2d_desc *desc, self_sw2d_t *out)
try {
    SELF_REQUIRE(desc && out, SELF_ERR_ARG, "self_sw2d_create: null argument");
    SELF_REQUIRE(desc->N >= 1 && desc->N <= 11, SELF_ERR_ARG, "polyDeg %d outside 1..11", desc->N);
    SELF_REQUIRE(desc->nElems >= 1 && desc->nEdges >= 1, SELF_ERR_ARG, "empty mesh");
    SELF_REQUIRE(desc->precision == SELF_FP64 || desc->precision == SELF_MIXED, SELF_ERR_ARG,
                 "unknown precision %d", desc->precision);
    SELF_REQUIRE(desc->form == SELF_PLAIN || desc->form == SELF_FUSED, SELF_ERR_ARG, "unknown form %d",
                 desc->form);
    SELF_REQUIRE(!(desc->precision == SELF_MIXED && desc->form == SELF_PLAIN), SELF_ERR_ARG,
                 "mixed precision runs in the fused form only");
    auto *h = new self_sw2d_s();
    h->d = *desc;
    h->n1 = desc->N + 1;
    h->np = h->n1 * h->n1;
    h->nE = (size_t)desc->nElems;
    int rc = SELF_OK;
    auto fail = [&](int code) { free_all(h); if (h->stream) (void)hipStreamDestroy(h->stream); delete h; return code; };
    hipError_t e = hipGetDevice(&h->device);
    if (e != hipSuccess) return fail(hip_fail(e, "hipGetDevice", __FILE__, __LINE__));
    e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(hip_fail(e, "hipStreamCreate", __FILE__, __LINE__));
    const size_t nE = h->nE, np = h->np, n1 = h->n1;
    const size_t nfield = nE * np;
    if ((rc = dalloc(h, &h->q, 3 * nfield))) return fail(rc);
    if ((rc = dalloc(h, &h->tend, 3 * nfield))) return fail(rc);
    if ((rc = dalloc(h, &h->G64, 3 * nfield))) return fail(rc);
    if ((rc = dalloc(h, &h->metric, 5 * nfield))) return fail(rc);
    if ((rc = dalloc(h, &h->bathy, 3 * nfield))) return fail(rc);
    if ((rc = dalloc(h, &h->plv, nfield))) return fail(rc);
    if ((rc = dalloc(h, &h->nrm, nE * 12 * n1))) return fail(rc);
    if ((rc = dalloc(h, &h->fnrm, nE * 12 * n1))) return fail(rc);
    if ((rc = dalloc(h, &h->side, nE * 4))) return fail(rc);
    if ((rc = dalloc(h, &h->edges, (size_t)desc->nEdges * 6))) return fail(rc);
    if ((rc = dalloc(h, &h->bsol, nE * 12 * n1))) return fail(rc);
    if ((rc = dalloc(h, &h->bflux, nE * 12 * n1))) return fail(rc);
    if ((rc = dalloc(h, &h->tr[0], nE * 12 * n1))) return fail(rc);
    if ((rc = dalloc(h, &h->tr[1], nE * 12 * n1))) return fail(rc);
    if ((rc = dalloc(h, &h->nodal, nodal_buffer_size(n1)))) return fail(rc);
    if (desc->precision == SELF_MIXED) {
        if ((rc = dalloc(h, &h->G32, 3 * nfield))) return fail(rc);
        if ((rc = dalloc(h, &h->bathy32, 2 * nfield))) return fail(rc);
        if ((rc = dalloc(h, &h->plv32, nfield))) return fail(rc);
        if ((rc = dalloc(h, &h->dS32, n1 * n1))) return fail(rc);
        if ((rc = dalloc(h, &h->dP32, n1 * n1))) return fail(rc);
        if ((rc = dalloc(h, &h->metM, nE * 5))) return fail(rc);
        if ((rc = dalloc(h, &h->metD, 5 * nfield))) return fail(rc);
        if ((rc = dalloc(h, &h->fnM, nE * 12))) return fail(rc);
        if ((rc = dalloc(h, &h->fnD, nE * 12 * n1))) return fail(rc);
    }
    e = hipMemsetAsync(h->q, 0, 3 * nfield * sizeof(double), h->stream);
    if (e == hipSuccess) e = hipMemsetAsync(h->bathy, 0, 3 * nfield * sizeof(double), h->stream);
    if (e == hipSuccess) e = hipMemsetAsync(h->plv, 0, nfield * sizeof(double), h->stream);
    if (e == hipSuccess) e = hipMemsetAsync(h->tend, 0, 3 * nfield * sizeof(double), h->stream);
    // boundary slots of the pushed traces are never written
    if (e == hipSuccess) e = hipMemsetAsync(h->tr[0], 0, nE * 12 * n1 * sizeof(double), h->stream);
    if (e == hipSuccess) e = hipMemsetAsync(h->tr[1], 0, nE * 12 * n1 * sizeof(double), h->stream);
    if (e == hipSuccess && h->bathy32) e = hipMemsetAsync(h->bathy32, 0, 2 * nfield * sizeof(float), h->stream);
    if (e == hipSuccess && h->plv32) e = hipMemsetAsync(h->plv32, 0, nfield * sizeof(float), h->stream);
    if (e == hipSuccess) e = hipEventCreate(&h->ev0);
    if (e == hipSuccess) e = hipEventCreate(&h->ev1);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return fail(hip_fail(e, "create", __FILE__, __LINE__));
    *out = h;
    return SELF_OK;
}
SELF_CATCH_ALL

int self_sw2d_destroy(self_sw2d_t h)
{
    if (!h) return SELF_OK;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    if (h->graphExec) (void)hipGraphExecDestroy(h->graphExec);
    free_all(h);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return SELF_OK;
}

int self_sw2d_set_nodal(self_sw2d_t h, const double *dMatS, const double *dMatP, const double *qwS,
                        const double *qwP, const double *lW, const double *lE, const double *lS,
                        const double *lN)
try {
    SELF_REQUIRE(h && dMatS && dMatP && qwS && qwP && lW && lE && lS && lN, SELF_ERR_ARG,
                 "self_sw2d_set_nodal: null argument");
    SELF_HIP(hipSetDevice(h->device));
    const int n1 = h->n1;
    std::vector<double> buf(nodal_buffer_size(n1));
    std::memcpy(buf.data(), dMatS, sizeof(double) * n1 * n1);
    std::memcpy(buf.data() + n1 * n1, dMatP, sizeof(double) * n1 * n1);
    double *p = buf.data() + 2 * n1 * n1;
    std::memcpy(p, qwS, sizeof(double) * n1);
    std::memcpy(p + n1, qwP, sizeof(double) * n1);
    std::memcpy(p + 2 * n1, lW, sizeof(double) * n1);
    std::memcpy(p + 3 * n1, lE, sizeof(double) * n1);
    std::memcpy(p + 4 * n1, lS, sizeof(double) * n1);
    std::memcpy(p + 5 * n1, lN, sizeof(double) * n1);
    nodal_buffer_derived(buf.data(), n1);
    SELF_HIP(hipMemcpy(h->nodal, buf.data(), sizeof(double) * buf.size(), hipMemcpyHostToDevice));
    if (h->dS32) {
        std::vector<float> f(2 * n1 * n1);
        for (int k = 0; k < 2 * n1 * n1; ++k) f[k] = (float)buf[k];
        SELF_HIP(hipMemcpy(h->dS32, f.data(), sizeof(float) * n1 * n1, hipMemcpyHostToDevice));
        SELF_HIP(hipMemcpy(h->dP32, f.data() + n1 * n1, sizeof(float) * n1 * n1, hipMemcpyHostToDevice));
    }
    h->nodalSet = true;
    h->trValid = false;
    return SELF_OK;
}
SELF_CATCH_ALL

int self_sw2d_set_geometry(self_sw2d_t h, const int *edges, const double *dxds, const double *dxdp,
                           const double *dyds, const double *dydp, const double *J, const double *nHat,
                           const double *nLen)
try {
    SELF_REQUIRE(h && edges && dxds && dxdp && dyds && dydp && J && nHat && nLen, SELF_ERR_ARG,
                 "self_sw2d_set_geometry: null argument");
    SELF_HIP(hipSetDevice(h->device));
    const size_t nE = h->nE, np = h->np;
    const int n1 = h->n1, N = h->d.N, nEd = h->d.nEdges;
    // metrics [5][e][np]
    std::vector<double> met(5 * nE * np);
    const double *src[5] = {dxds, dxdp, dyds, dydp, J};
    for (int k = 0; k < 5; ++k) std::memcpy(met.data() + k * nE * np, src[k], sizeof(double) * nE * np);
    for (size_t id = 0; id < nE * np; ++id)
        SELF_REQUIRE(J[id] != 0.0, SELF_ERR_ARG, "zero Jacobian at element %zu", id / np + 1);
    // own normals [e][side][c][k]
    std::vector<double> nrm(nE * 12 * n1), fnrm(nE * 12 * n1);
    for (size_t e = 0; e < nE; ++e)
        for (int s = 0; s < 4; ++s)
            for (int k = 0; k < n1; ++k) {
                const size_t hb = k + (size_t)n1 * (s + 4 * e);
                const size_t db = ((e * 4 + s) * 3) * n1 + k;
                nrm[db] = nHat[2 * hb];
                nrm[db + n1] = nHat[2 * hb + 1];
                nrm[db + 2 * n1] = nLen[hb];
            }
    fnrm = nrm;
    // side descriptors from the edge table
    std::vector<int2> side(nE * 4);
    std::vector<int> seen(nE * 4, 0);
    for (int ie = 0; ie < nEd; ++ie) {
        const int *ed = edges + 6 * ie;
        const int e1 = ed[0] - 1, s1 = ed[1] - 1, e2 = ed[2], s2 = std::abs(ed[3]) - 1, inc = ed[5];
        SELF_REQUIRE(e1 >= 0 && (size_t)e1 < nE && s1 >= 0 && s1 < 4, SELF_ERR_ARG, "edge %d: bad primary", ie + 1);
        SELF_REQUIRE(inc == 1 || inc == -1, SELF_ERR_ARG, "edge %d: increment %d", ie + 1, inc);
        if (e2 > 0) {
            SELF_REQUIRE((size_t)(e2 - 1) < nE && s2 >= 0 && s2 < 4, SELF_ERR_ARG, "edge %d: bad secondary", ie + 1);
            SELF_REQUIRE(ed[4] - inc == (inc > 0 ? 0 : N), SELF_ERR_ARG,
                         "edge %d: start/inc must traverse the whole side", ie + 1);
            const int flip = inc < 0 ? 1 : 0;
            side[(size_t)e1 * 4 + s1] = make_int2(e2 - 1, s2 | (flip << 2) | (1 << 3));
            side[(size_t)(e2 - 1) * 4 + s2] = make_int2(e1, s1 | (flip << 2));
            seen[(size_t)e1 * 4 + s1]++;
            seen[(size_t)(e2 - 1) * 4 + s2]++;
            // the secondary side uses the primary's normal at the matching node
            for (int k = 0; k < n1; ++k) {
                const int kp = flip ? N - k : k;  // primary node seen from secondary node k
                const size_t dsec = (((size_t)(e2 - 1) * 4 + s2) * 3) * n1 + k;
                const size_t dpri = (((size_t)e1 * 4 + s1) * 3) * n1 + kp;
                fnrm[dsec] = nrm[dpri];
                fnrm[dsec + n1] = nrm[dpri + n1];
                fnrm[dsec + 2 * n1] = nrm[dpri + 2 * n1];
            }
        } else {
            side[(size_t)e1 * 4 + s1] = make_int2(e2, 1 << 3);
            seen[(size_t)e1 * 4 + s1]++;
        }
    }
    for (size_t k = 0; k < nE * 4; ++k)
        SELF_REQUIRE(seen[k] == 1, SELF_ERR_ARG, "element %zu side %zu is on %d edges", k / 4 + 1, k % 4 + 1, seen[k]);
    SELF_HIP(hipMemcpy(h->metric, met.data(), sizeof(double) * met.size(), hipMemcpyHostToDevice));
    SELF_HIP(hipMemcpy(h->nrm, nrm.data(), sizeof(double) * nrm.size(), hipMemcpyHostToDevice));
    SELF_HIP(hipMemcpy(h->fnrm, fnrm.data(), sizeof(double) * fnrm.size(), hipMemcpyHostToDevice));
    if (h->metD) {
        // mixed form: FP64 reference value per element (metrics, node 0) and
        // per side (normals, node 0) plus FP32 deviations; an affine element
        // or a straight side has zero deviations and is reproduced exactly
        std::vector<double> mM(nE * 5), fM(nE * 12);
        std::vector<float> mD(5 * nE * np), fD(nE * 12 * n1);
        for (int k = 0; k < 5; ++k)
            for (size_t e = 0; e < nE; ++e) {
                const double *m = met.data() + k * nE * np + e * np;
                mM[e * 5 + k] = m[0];
                for (size_t nd = 0; nd < np; ++nd) mD[k * nE * np + e * np + nd] = (float)(m[nd] - m[0]);
            }
        for (size_t ec = 0; ec < nE * 12; ++ec) {
            const double *f = fnrm.data() + ec * n1;
            fM[ec] = f[0];
            for (int k = 0; k < n1; ++k) fD[ec * n1 + k] = (float)(f[k] - f[0]);
        }
        SELF_HIP(hipMemcpy(h->metM, mM.data(), sizeof(double) * mM.size(), hipMemcpyHostToDevice));
        SELF_HIP(hipMemcpy(h->metD, mD.data(), sizeof(float) * mD.size(), hipMemcpyHostToDevice));
        SELF_HIP(hipMemcpy(h->fnM, fM.data(), sizeof(double) * fM.size(), hipMemcpyHostToDevice));
        SELF_HIP(hipMemcpy(h->fnD, fD.data(), sizeof(float) * fD.size(), hipMemcpyHostToDevice));
    }
    SELF_HIP(hipMemcpy(h->side, side.data(), sizeof(int2) * side.size(), hipMemcpyHostToDevice));
    SELF_HIP(hipMemcpy(h->edges, edges, sizeof(int) * 6 * nEd, hipMemcpyHostToDevice));
    h->geomSet = true;
    h->trValid = false;
    return SELF_OK;
}
SELF_CATCH_ALL

int self_sw2d_set_field(self_sw2d_t h, int field, const double *data)
try {
    SELF_REQUIRE(h && data, SELF_ERR_ARG, "self_sw2d_set_field: null argument");
    SELF_HIP(hipSetDevice(h->device));
    SELF_HIP(hipStreamSynchronize(h->stream));
    const size_t nE = h->nE, np = h->np;
    std::vector<double> buf;
    switch (field) {
    case SELF_SW_SOLUTION:
        buf.resize(3 * nE * np);
        to_device_layout(data, buf.data(), nE, 3, (int)np);
        SELF_HIP(hipMemcpy(h->q, buf.data(), sizeof(double) * buf.size(), hipMemcpyHostToDevice));
        h->trValid = false;
        break;
    case SELF_SW_BATHYMETRY: {
        buf.resize(3 * nE * np);
        to_device_layout(data, buf.data(), nE, 3, (int)np);
        SELF_HIP(hipMemcpy(h->bathy, buf.data(), sizeof(double) * buf.size(), hipMemcpyHostToDevice));
        if (h->bathy32) {
            std::vector<float> f(2 * nE * np);
            for (size_t k = 0; k < 2 * nE * np; ++k) f[k] = (float)buf[nE * np + k];
            SELF_HIP(hipMemcpy(h->bathy32, f.data(), sizeof(float) * f.size(), hipMemcpyHostToDevice));
        }
        break;
    }
    case SELF_SW_PLANETARY_VORTICITY: {
        SELF_HIP(hipMemcpy(h->plv, data, sizeof(double) * nE * np, hipMemcpyHostToDevice));
        if (h->plv32) {
            std::vector<float> f(nE * np);
            for (size_t k = 0; k < nE * np; ++k) f[k] = (float)data[k];
            SELF_HIP(hipMemcpy(h->plv32, f.data(), sizeof(float) * f.size(), hipMemcpyHostToDevice));
        }
        break;
    }
    case SELF_SW_RELAXATION: {
        buf.resize(4 * nE * np);
        to_device_layout(data, buf.data(), nE, 4, (int)np);
        bool any = false;
        for (size_t k = 0; k < nE * np; ++k) any |= buf[3 * nE * np + k] != 0.0;
        if (!any) {  // every timescale zero: the relaxation term vanishes identically
            h->hasRelax = false;
            break;
        }
        if (!h->relax) {
            int rc = dalloc(h, &h->relax, 4 * nE * np);
            if (rc) return rc;
        }
        SELF_HIP(hipMemcpy(h->relax, buf.data(), sizeof(double) * buf.size(), hipMemcpyHostToDevice));
        h->hasRelax = true;
        break;
    }
    default:
        SELF_REQUIRE(false, SELF_ERR_ARG, "field %d cannot be set", field);
    }
    return SELF_OK;
}
SELF_CATCH_ALL

int self_sw2d_get_field(self_sw2d_t h, int field, double *data)
try {
    SELF_REQUIRE(h && data, SELF_ERR_ARG, "self_sw2d_get_field: null argument");
    SELF_HIP(hipSetDevice(h->device));
    SELF_HIP(hipStreamSynchronize(h->stream));
    const size_t nE = h->nE, np = h->np, n1 = h->n1;
    std::vector<double> buf;
    auto get_nodal = [&](const double *dev, int nf) -> int {
        buf.resize((size_t)nf * nE * np);
        SELF_HIP(hipMemcpy(buf.data(), dev, sizeof(double) * buf.size(), hipMemcpyDeviceToHost));
        from_device_layout(buf.data(), data, nE, nf, (int)np);
        return SELF_OK;
    };
    switch (field) {
    case SELF_SW_SOLUTION: return get_nodal(h->q, 3);
    case SELF_SW_TENDENCY: return get_nodal(h->tend, 3);
    case SELF_SW_BATHYMETRY: return get_nodal(h->bathy, 3);
    case SELF_SW_PLANETARY_VORTICITY:
        SELF_HIP(hipMemcpy(data, h->plv, sizeof(double) * nE * np, hipMemcpyDeviceToHost));
        return SELF_OK;
    case SELF_SW_RELAXATION:
        if (!h->relax) { std::memset(data, 0, sizeof(double) * 4 * nE * np); return SELF_OK; }
        return get_nodal(h->relax, 4);
    case SELF_SW_BOUNDARY_SOLUTION:
    case SELF_SW_BOUNDARY_FLUX: {
        // device [e][side][f][k] -> host per element (0:N,1:3,1:4) = [e][side][f][k]: identical
        const double *dev = field == SELF_SW_BOUNDARY_SOLUTION ? h->bsol : h->bflux;
        SELF_HIP(hipMemcpy(data, dev, sizeof(double) * nE * 12 * n1, hipMemcpyDeviceToHost));
        return SELF_OK;
    }
    default:
        SELF_REQUIRE(false, SELF_ERR_ARG, "field %d cannot be read", field);
    }
    return SELF_OK;
}
SELF_CATCH_ALL

int self_sw2d_set_params(self_sw2d_t h, double g, double dt, double linearDrag)
{
    SELF_REQUIRE(h, SELF_ERR_ARG, "null handle");
    h->d.g = g;
    h->d.dt = dt;
    h->d.linearDrag = linearDrag;
    if (h->dtM) {  // every member takes the common step again
        SELF_HIP(hipSetDevice(h->device));
        SELF_HIP(hipStreamSynchronize(h->stream));
        h->dtH.assign(h->dtH.size(), dt);
        SELF_HIP(hipMemcpy(h->dtM, h->dtH.data(), sizeof(double) * h->dtH.size(), hipMemcpyHostToDevice));
    }
    return SELF_OK;
}

int self_sw2d_calculate_solution_at_boundaries(self_sw2d_t h)
{
    int rc = check_ready(h);
    if (rc) return rc;
    return launch_traces(h, h->q, h->bsol, 3);
}

int self_sw2d_edge_flux(self_sw2d_t h, double t)
{
    (void)t;  // no time-dependent boundary state in this model
    int rc = check_ready(h);
    if (rc) return rc;
    return launch_edge_flux(h);
}

int self_sw2d_mapped_time_derivative(self_sw2d_t h, double t)
{
    (void)t;
    int rc = check_ready(h);
    if (rc) return rc;
    return launch_mapped(h);
}

int self_sw2d_global_time_derivative(self_sw2d_t h, double t)
{
    (void)t;
    int rc = check_ready(h);
    if (rc) return rc;
    return gtd_plain(h);
}

int self_sw2d_check_mixed(self_sw2d_t h, double tol, double *relErr, int *switched)
{
    int rc = check_ready(h);
    if (rc) return rc;
    SELF_REQUIRE(relErr && switched, SELF_ERR_ARG, "null argument");
    *relErr = 0.0;
    *switched = 0;
    if (h->d.precision != SELF_MIXED) return SELF_OK;
    SELF_REQUIRE(h->members == 1, SELF_ERR_STATE, "the mixed-form check runs on a single state (members = %d)",
                 h->members);
    // the FP64 tendency of the current state
    if ((rc = gtd_plain(h))) return rc;
    // one mixed fused stage from the same state into scratch buffers: with
    // a = 0 its register holds the mixed tendency, rounded to FP32
    const size_t nf = h->nE * h->np, nt = h->nE * 12 * h->n1;
    constexpr int kParts = 256;
    double *qS = nullptr, *trA = nullptr, *trB = nullptr, *part = nullptr;
    float *gS = nullptr;
    struct Scratch {
        void *p[5];
        ~Scratch()
        {
            for (void *x : p)
                if (x) (void)hipFree(x);
        }
    } scratch{};
    SELF_HIP(hipMalloc(&qS, sizeof(double) * 3 * nf));
    scratch.p[0] = qS;
    SELF_HIP(hipMalloc(&gS, sizeof(float) * 3 * nf));
    scratch.p[1] = gS;
    SELF_HIP(hipMalloc(&trA, sizeof(double) * nt));
    scratch.p[2] = trA;
    SELF_HIP(hipMalloc(&trB, sizeof(double) * nt));
    scratch.p[3] = trB;
    SELF_HIP(hipMalloc(&part, sizeof(double) * 2 * kParts));
    scratch.p[4] = part;
    SELF_HIP(hipMemcpyAsync(qS, h->q, sizeof(double) * 3 * nf, hipMemcpyDeviceToDevice, h->stream));
    SELF_HIP(hipMemsetAsync(trA, 0, sizeof(double) * nt, h->stream));
    SELF_HIP(hipMemsetAsync(trB, 0, sizeof(double) * nt, h->stream));
    if ((rc = launch_traces_push(h, trA))) return rc;
    SWKernelArgs A = make_args(h);
    A.q = qS;
    A.G32 = gS;
    A.trIn = trA;
    A.trOut = trB;
    A.a = 0.0;
    A.gdt = 0.0;
    A.gcoef = 0.0;
    if ((rc = launch_stage_fused(h, A))) return rc;
    hipLaunchKernelGGL(k_sw_guard_sums, dim3(kParts), dim3(256), 0, h->stream, (const float *)gS,
                       (const double *)h->tend, 2 * nf, part);
    SELF_HIP(hipGetLastError());
    std::vector<double> ph(2 * kParts);
    SELF_HIP(hipMemcpyAsync(ph.data(), part, sizeof(double) * 2 * kParts, hipMemcpyDeviceToHost, h->stream));
    SELF_HIP(hipStreamSynchronize(h->stream));
    double d2 = 0.0, t2 = 0.0;
    for (int k = 0; k < kParts; ++k) {
        d2 += ph[2 * k];
        t2 += ph[2 * k + 1];
    }
    *relErr = t2 > 0.0 ? std::sqrt(d2 / t2) : (d2 > 0.0 ? INFINITY : 0.0);
    if (*relErr > tol) {
        // the FP64 fused form: the FP64 geometry, bathymetry and Coriolis
        // are set for every model; the FP64 register is allocated at creation
        // (and again here if self_sw2d_set_members replaced it)
        if (!h->G64 && (rc = dalloc(h, &h->G64, 3 * nf))) return rc;
        h->d.precision = SELF_FP64;
        *switched = 1;
    }
    return SELF_OK;
}

int self_sw2d_forward_step_rk3(self_sw2d_t h, double tn, int nSteps)
{
    (void)tn;
    int rc = check_ready(h);
    if (rc) return rc;
    SELF_REQUIRE(nSteps >= 0, SELF_ERR_ARG, "negative step count");
    int s = 0;
    if (h->d.form == SELF_FUSED && h->useGraph && nSteps >= 2) {
        rc = steps_fused_graph(h, nSteps / 2);
        s = 2 * (nSteps / 2);
    }
    for (; s < nSteps && !rc; ++s) rc = (h->d.form == SELF_FUSED) ? step_fused(h) : step_plain(h);
    return rc;
}

int self_sw2d_forward_step_rk3_timed(self_sw2d_t h, double tn, int nSteps, double *elapsed_ms)
{
    int rc = check_ready(h);
    if (rc) return rc;
    SELF_REQUIRE(elapsed_ms, SELF_ERR_ARG, "null elapsed_ms");
    if (h->d.form == SELF_FUSED && !h->trValid) {  // keep the initial traces out of the timed region
        rc = launch_traces_push(h, h->tr[h->trCur]);
        if (rc) return rc;
        h->trValid = true;
    }
    SELF_HIP(hipEventRecord(h->ev0, h->stream));
    rc = self_sw2d_forward_step_rk3(h, tn, nSteps);
    if (rc) return rc;
    SELF_HIP(hipEventRecord(h->ev1, h->stream));
    SELF_HIP(hipEventSynchronize(h->ev1));
    float ms = 0.0f;
    SELF_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    *elapsed_ms = ms;
    return SELF_OK;
}

int self_sw2d_kernel_times(self_sw2d_t h, double tn, int nSteps, double *avg_ms, int *nKernels)
try {
    (void)tn;
    int rc = check_ready(h);
    if (rc) return rc;
    SELF_REQUIRE(avg_ms && nKernels && nSteps > 0, SELF_ERR_ARG, "self_sw2d_kernel_times: bad arguments");
    const bool fused = h->d.form == SELF_FUSED;
    const int nk = fused ? 1 : 4;
    *nKernels = nk;
    if (fused && !h->trValid) {
        rc = launch_traces_push(h, h->tr[h->trCur]);
        if (rc) return rc;
        h->trValid = true;
    }
    // one event pair per kernel launch, on the model's stream
    const int nLaunch = 3 * nSteps * nk;
    std::vector<hipEvent_t> ev(2 * nLaunch);
    for (auto &e : ev) SELF_HIP(hipEventCreate(&e));
    int li = 0;
    auto mark = [&](int which) { return hipEventRecord(ev[2 * li + which], h->stream); };
    const size_t n = 3 * h->nE * h->np;
    for (int st = 0; st < nSteps && !rc; ++st) {
        if (fused) {
            SWKernelArgs A = make_args(h);
            for (int m = 0; m < 3; ++m) {
                A.a = kRkA[m];
                A.gdt = kRkG[m] * h->d.dt;
                A.gcoef = kRkG[m];
                A.trIn = h->tr[h->trCur];
                A.trOut = h->tr[1 - h->trCur];
                SELF_HIP(mark(0));
                if ((rc = launch_stage_fused(h, A))) break;
                SELF_HIP(mark(1));
                ++li;
                h->trCur = 1 - h->trCur;
            }
        } else {
            for (int m = 0; m < 3; ++m) {
                SELF_HIP(mark(0)); rc = launch_traces(h, h->q, h->bsol, 3); SELF_HIP(mark(1)); ++li;
                if (rc) break;
                SELF_HIP(mark(0)); rc = launch_edge_flux(h); SELF_HIP(mark(1)); ++li;
                if (rc) break;
                SELF_HIP(mark(0)); rc = launch_mapped(h); SELF_HIP(mark(1)); ++li;
                if (rc) break;
                SELF_HIP(mark(0));
                hipLaunchKernelGGL(k_sw_rk_update, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, h->q,
                                   h->G64, h->tend, n, kRkA[m], kRkG[m] * h->d.dt);
                SELF_HIP(hipGetLastError());
                SELF_HIP(mark(1)); ++li;
            }
        }
    }
    if (!rc) SELF_HIP(hipStreamSynchronize(h->stream));
    for (int k = 0; k < nk; ++k) avg_ms[k] = 0.0;
    for (int i = 0; i < li && !rc; ++i) {
        float ms = 0.0f;
        SELF_HIP(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
        avg_ms[i % nk] += ms;
    }
    for (int k = 0; k < nk; ++k) avg_ms[k] /= (double)(3 * nSteps);
    for (auto &e : ev) (void)hipEventDestroy(e);
    return rc;
}
SELF_CATCH_ALL

int self_sw2d_calculate_bathymetry_gradient(self_sw2d_t h)
{
    int rc = check_ready(h);
    if (rc) return rc;
    SWKernelArgs A = make_args(h);
    SELF_DISPATCH_N(h->d.N, hipLaunchKernelGGL(k_sw_bathy_grad<NN>, dim3(grid_for<NN>(h->nE)), dim3(256), 0,
                                               h->stream, A, h->bathy, h->bathy32));
    SELF_HIP(hipGetLastError());
    return SELF_OK;
}

int self_sw2d_set_members(self_sw2d_t h, int members)
try {
    SELF_REQUIRE(h, SELF_ERR_ARG, "null handle");
    SELF_REQUIRE(members >= 1 && members <= 4096, SELF_ERR_ARG, "members %d outside 1..4096", members);
    SELF_REQUIRE(h->d.form == SELF_FUSED || members == 1, SELF_ERR_ARG, "ensembles run in the fused form only");
    SELF_HIP(hipSetDevice(h->device));
    SELF_HIP(hipStreamSynchronize(h->stream));
    const size_t nf = 3 * h->nE * h->np, nt = h->nE * 12 * h->n1, M = (size_t)members;
    // reallocate the per-member state, RK register and trace buffers
    void *old[] = {h->q, h->G64, h->G32, h->tr[0], h->tr[1]};
    for (void *p : old)
        if (p) (void)hipFree(p);
    h->bytes -= h->members * (nf * sizeof(double) + nf * (h->G32 ? sizeof(float) : sizeof(double)) +
                              2 * nt * sizeof(double));
    h->q = nullptr; h->G64 = nullptr; h->G32 = nullptr; h->tr[0] = nullptr; h->tr[1] = nullptr;
    h->members = 1;
    int rc = dalloc(h, &h->q, M * nf);
    if (!rc) rc = (h->d.precision == SELF_MIXED) ? dalloc(h, &h->G32, M * nf) : dalloc(h, &h->G64, M * nf);
    if (!rc) rc = dalloc(h, &h->tr[0], M * nt);
    if (!rc) rc = dalloc(h, &h->tr[1], M * nt);
    if (rc) return rc;
    if (h->dtM) {
        (void)hipFree(h->dtM);
        h->dtM = nullptr;
        h->bytes -= h->dtH.size() * sizeof(double);
    }
    h->dtH.assign(M, h->d.dt);
    if (M > 1) {
        if ((rc = dalloc(h, &h->dtM, M))) return rc;
        SELF_HIP(hipMemcpy(h->dtM, h->dtH.data(), sizeof(double) * M, hipMemcpyHostToDevice));
    }
    SELF_HIP(hipMemsetAsync(h->q, 0, M * nf * sizeof(double), h->stream));
    // boundary slots of the pushed traces are never written
    SELF_HIP(hipMemsetAsync(h->tr[0], 0, M * nt * sizeof(double), h->stream));
    SELF_HIP(hipMemsetAsync(h->tr[1], 0, M * nt * sizeof(double), h->stream));
    SELF_HIP(hipStreamSynchronize(h->stream));
    h->members = members;
    h->trValid = false;
    return SELF_OK;
}
SELF_CATCH_ALL

int self_sw2d_set_member_solution(self_sw2d_t h, int member, const double *data)
try {
    SELF_REQUIRE(h && data, SELF_ERR_ARG, "null argument");
    SELF_REQUIRE(member >= 0 && member < h->members, SELF_ERR_ARG, "member %d outside 0..%d", member, h->members - 1);
    SELF_HIP(hipSetDevice(h->device));
    SELF_HIP(hipStreamSynchronize(h->stream));
    const size_t nE = h->nE, np = h->np;
    std::vector<double> buf(3 * nE * np);
    to_device_layout(data, buf.data(), nE, 3, (int)np);
    SELF_HIP(hipMemcpy(h->q + (size_t)member * buf.size(), buf.data(), sizeof(double) * buf.size(),
                       hipMemcpyHostToDevice));
    h->trValid = false;
    return SELF_OK;
}
SELF_CATCH_ALL

int self_sw2d_set_member_time_step(self_sw2d_t h, int member, double dt)
try {
    SELF_REQUIRE(h, SELF_ERR_ARG, "null handle");
    SELF_REQUIRE(member >= 0 && member < h->members, SELF_ERR_ARG, "member %d outside 0..%d", member, h->members - 1);
    SELF_REQUIRE(dt > 0.0, SELF_ERR_ARG, "time step %g", dt);
    if (h->members == 1) {  // a single state steps with the common dt
        h->d.dt = dt;
        return SELF_OK;
    }
    SELF_HIP(hipSetDevice(h->device));
    SELF_HIP(hipStreamSynchronize(h->stream));
    h->dtH[member] = dt;
    SELF_HIP(hipMemcpy(h->dtM + member, &dt, sizeof(double), hipMemcpyHostToDevice));
    return SELF_OK;
}
SELF_CATCH_ALL

int self_sw2d_get_member_solution(self_sw2d_t h, int member, double *data)
try {
    SELF_REQUIRE(h && data, SELF_ERR_ARG, "null argument");
    SELF_REQUIRE(member >= 0 && member < h->members, SELF_ERR_ARG, "member %d outside 0..%d", member, h->members - 1);
    SELF_HIP(hipSetDevice(h->device));
    SELF_HIP(hipStreamSynchronize(h->stream));
    const size_t nE = h->nE, np = h->np;
    std::vector<double> buf(3 * nE * np);
    SELF_HIP(hipMemcpy(buf.data(), h->q + (size_t)member * buf.size(), sizeof(double) * buf.size(),
                       hipMemcpyDeviceToHost));
    from_device_layout(buf.data(), data, nE, 3, (int)np);
    return SELF_OK;
}
SELF_CATCH_ALL

// WritePickup_ShallowWater (ConservativeShallowWaterClass.f90:2013-2065) of member 0
int self_sw2d_write_pickup(self_sw2d_t h, const char *path)
try {
    SELF_REQUIRE(h && path, SELF_ERR_ARG, "null argument");
    const size_t nv = h->nE * h->np;
    std::vector<double> sol(3 * nv), bathy(3 * nv), plv(nv);
    int rc = self_sw2d_get_field(h, SELF_SW_SOLUTION, sol.data());
    if (!rc) rc = self_sw2d_get_field(h, SELF_SW_BATHYMETRY, bathy.data());
    if (!rc) rc = self_sw2d_get_field(h, SELF_SW_PLANETARY_VORTICITY, plv.data());
    if (!rc) rc = self_sw2d_pickup_write(path, h->d.N, (int)h->nE, sol.data(), bathy.data(), plv.data());
    return rc;
}
SELF_CATCH_ALL

// ReadPickup_ShallowWater (:2069-2125) and the gradient Build computes after it (:224-228)
int self_sw2d_read_pickup(self_sw2d_t h, const char *path)
try {
    SELF_REQUIRE(h && path, SELF_ERR_ARG, "null argument");
    const size_t nv = h->nE * h->np;
    std::vector<double> sol(3 * nv), bathy(3 * nv), plv(nv);
    int rc = self_sw2d_pickup_read(path, h->d.N, (int)h->nE, sol.data(), bathy.data(), plv.data());
    if (!rc) rc = self_sw2d_set_field(h, SELF_SW_SOLUTION, sol.data());
    if (!rc) rc = self_sw2d_set_field(h, SELF_SW_BATHYMETRY, bathy.data());
    if (!rc) rc = self_sw2d_set_field(h, SELF_SW_PLANETARY_VORTICITY, plv.data());
    if (!rc) rc = self_sw2d_calculate_bathymetry_gradient(h);
    return rc;
}
SELF_CATCH_ALL

int self_sw2d_set_variant(self_sw2d_t h, int variant)
{
    SELF_REQUIRE(h, SELF_ERR_ARG, "null handle");
    SELF_REQUIRE(variant == 0 || variant == 1, SELF_ERR_ARG, "unknown shallow-water variant %d", variant);
    h->variant = variant;
    return SELF_OK;
}

// ForwardStepRK3 nSteps times for several fused-form models at once, of
// any degrees, one persistent launch per RK stage for all of them
int self_sw2d_forward_step_rk3_multi(self_sw2d_t *models, int nModels, int nSteps, double *elapsed_ms)
try {
    SELF_REQUIRE(models && nModels >= 1 && nSteps >= 0, SELF_ERR_ARG, "bad arguments");
    self_sw2d_s *h0 = models[0];
    for (int k = 0; k < nModels; ++k) {
        self_sw2d_s *h = models[k];
        int rc = check_ready(h);
        if (rc) return rc;
        SELF_REQUIRE(h->d.form == SELF_FUSED, SELF_ERR_ARG, "model %d is not in the fused form", k);
        SELF_REQUIRE(h->d.precision == h0->d.precision, SELF_ERR_ARG, "model %d: precisions differ", k);
        SELF_REQUIRE(h->device == h0->device, SELF_ERR_ARG, "model %d is on another device", k);
        SELF_REQUIRE(h->d.N >= 1 && h->d.N <= 11, SELF_ERR_ARG, "model %d: N=%d (1..11)", k, h->d.N);
        for (int j = 0; j < k; ++j) SELF_REQUIRE(models[j] != h, SELF_ERR_ARG, "model %d is listed twice", k);
        if (!h->trValid) {
            rc = launch_traces_push(h, h->tr[h->trCur]);
            if (rc) return rc;
            h->trValid = true;
        }
        SELF_HIP(hipStreamSynchronize(h->stream));
    }
    // six stage sets per batch of up to kMultiMax models (two steps bring
    // every model back to its trace parity)
    const int nB = (nModels + kMultiMax - 1) / kMultiMax;
    std::vector<SWMultiArgs> sets((size_t)6 * nB);
    std::memset(static_cast<void *>(sets.data()), 0, sizeof(SWMultiArgs) * sets.size());
    for (int b = 0; b < nB; ++b) {
        int total = 0;
        for (int k = b * kMultiMax; k < std::min(nModels, (b + 1) * kMultiMax); ++k) {
            self_sw2d_s *h = models[k];
            const int slot = k - b * kMultiMax;
            int tiles = 0;
            SELF_DISPATCH_N(h->d.N, tiles = grid_fused<NN>(h->nE, 1) * h->members);
            int cur = h->trCur;
            for (int st = 0; st < 6; ++st) {
                const int m = st % 3;
                SWMultiArgs &M = sets[(size_t)st * nB + b];
                M.nItems = slot + 1;
                M.N[slot] = h->d.N;
                M.tileStart[slot] = total;
                fill_args(h, M.A[slot]);
                M.A[slot].a = kRkA[m];
                M.A[slot].gdt = kRkG[m] * h->d.dt;
                M.A[slot].gcoef = kRkG[m];
                M.A[slot].trIn = h->tr[cur];
                M.A[slot].trOut = h->tr[1 - cur];
                cur = 1 - cur;
            }
            total += tiles;
        }
        for (int st = 0; st < 6; ++st) sets[(size_t)st * nB + b].totalTiles = total;
    }
    int cus = 0;
    SELF_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h0->device));
    hipError_t e = elapsed_ms ? hipEventRecord(h0->ev0, h0->stream) : hipSuccess;
    const bool mixed = h0->d.precision == SELF_MIXED;
    bool menu357 = true;
    for (int k = 0; k < nModels; ++k) {
        const int N = models[k]->d.N;
        menu357 = menu357 && (N == 3 || N == 5 || N == 7);
    }
    for (int s = 0; s < 3 * nSteps && e == hipSuccess; ++s)
        for (int b = 0; b < nB && e == hipSuccess; ++b) {
            const SWMultiArgs &M = sets[(size_t)(s % 6) * nB + b];
            const int grid = std::max(1, std::min(M.totalTiles, 4 * std::max(cus, 1)));
            if (mixed && menu357)
                hipLaunchKernelGGL((k_sw_stage_multi<true, 0>), dim3(grid), dim3(256), 0, h0->stream, M);
            else if (mixed)
                hipLaunchKernelGGL((k_sw_stage_multi<true, 1>), dim3(grid), dim3(256), 0, h0->stream, M);
            else if (menu357)
                hipLaunchKernelGGL((k_sw_stage_multi<false, 0>), dim3(grid), dim3(256), 0, h0->stream, M);
            else
                hipLaunchKernelGGL((k_sw_stage_multi<false, 1>), dim3(grid), dim3(256), 0, h0->stream, M);
            e = hipGetLastError();
        }
    if (e == hipSuccess && elapsed_ms) e = hipEventRecord(h0->ev1, h0->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h0->stream);
    SELF_HIP(e);
    if (elapsed_ms) {
        float ms = 0;
        SELF_HIP(hipEventElapsedTime(&ms, h0->ev0, h0->ev1));
        *elapsed_ms = ms;
    }
    if ((3 * nSteps) % 2)
        for (int k = 0; k < nModels; ++k) models[k]->trCur = 1 - models[k]->trCur;
    return SELF_OK;
}
SELF_CATCH_ALL

int self_sw2d_set_graph(self_sw2d_t h, int enable)
{
    SELF_REQUIRE(h, SELF_ERR_ARG, "null handle");
    h->useGraph = enable != 0;
    return SELF_OK;
}

int self_sw2d_set_groups(self_sw2d_t h, int groups)
{
    SELF_REQUIRE(h && groups >= 0 && groups <= 2, SELF_ERR_ARG, "groups must be 0 (heuristic), 1 or 2");
    h->groups = groups;
    return SELF_OK;
}

int self_sw2d_autotune(self_sw2d_t h, int nSteps, int *groups, double *ms_per_step)
try {
    int rc = check_ready(h);
    if (rc) return rc;
    SELF_REQUIRE(nSteps >= 1 && groups, SELF_ERR_ARG, "self_sw2d_autotune: bad arguments");
    if (h->d.form != SELF_FUSED) {
        *groups = 1;
        return SELF_OK;
    }
    // the table holds the choice for this degree, arithmetic and size
    // class once any model of that key has been tuned
    const char *kind = h->members > 1 ? "sw2d_ensemble" : "sw2d";
    const int dtype = h->d.precision == SELF_MIXED ? 1 : 0, sc = tune_size_class(h->nE * (size_t)h->members);
    int cfg[4];
    double known = 0.0;
    if (tune_lookup(kind, 2, h->d.N, dtype, sc, cfg, &known)) {
        h->groups = cfg[0];
        h->useGraph = cfg[1] != 0;
        *groups = cfg[0];
        if (ms_per_step) *ms_per_step = known;
        return SELF_OK;
    }
    // the trials advance a copy of the state: save and restore everything a step writes
    const size_t M = (size_t)h->members;
    const size_t nq = M * 3 * h->nE * h->np, nt = M * h->nE * 12 * h->n1;
    const size_t gb = h->G32 ? sizeof(float) * nq : sizeof(double) * nq;
    void *G = h->G32 ? (void *)h->G32 : (void *)h->G64;
    double *sq = nullptr, *st0 = nullptr, *st1 = nullptr;
    void *sg = nullptr;
    SELF_HIP(hipStreamSynchronize(h->stream));
    if (!h->trValid) {
        if ((rc = launch_traces_push(h, h->tr[h->trCur]))) return rc;
        h->trValid = true;
    }
    SELF_HIP(hipMalloc(&sq, sizeof(double) * nq));
    hipError_t e = hipMalloc(&sg, gb);
    if (e == hipSuccess) e = hipMalloc(&st0, sizeof(double) * nt);
    if (e == hipSuccess) e = hipMalloc(&st1, sizeof(double) * nt);
    if (e == hipSuccess) e = hipMemcpyAsync(sq, h->q, sizeof(double) * nq, hipMemcpyDeviceToDevice, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(sg, G, gb, hipMemcpyDeviceToDevice, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(st0, h->tr[0], sizeof(double) * nt, hipMemcpyDeviceToDevice, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(st1, h->tr[1], sizeof(double) * nt, hipMemcpyDeviceToDevice, h->stream);
    const int cur = h->trCur, saved = h->groups;
    const bool savedGraph = h->useGraph;
    double best = 1e300;
    int bestG = 1, bestGraph = 1;
    // candidates: element groups per workgroup x graph replay
    for (int trial = 0; trial < 4 && e == hipSuccess && !rc; ++trial) {
        const int g = 1 + (trial >> 1), gr = trial & 1;
        h->groups = g;
        h->useGraph = gr != 0;
        double ms = 0.0;
        rc = self_sw2d_forward_step_rk3(h, 0.0, 2);  // warm (captures the graph when on)
        if (!rc) rc = self_sw2d_forward_step_rk3_timed(h, 0.0, nSteps, &ms);
        if (!rc && ms < best) { best = ms; bestG = g; bestGraph = gr; }
        e = hipMemcpyAsync(h->q, sq, sizeof(double) * nq, hipMemcpyDeviceToDevice, h->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(G, sg, gb, hipMemcpyDeviceToDevice, h->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(h->tr[0], st0, sizeof(double) * nt, hipMemcpyDeviceToDevice, h->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(h->tr[1], st1, sizeof(double) * nt, hipMemcpyDeviceToDevice, h->stream);
        h->trCur = cur;
        h->trValid = true;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(sq); (void)hipFree(sg); (void)hipFree(st0); (void)hipFree(st1);
    if (e != hipSuccess || rc) {
        h->groups = saved;
        h->useGraph = savedGraph;
        return rc ? rc : hip_fail(e, "self_sw2d_autotune", __FILE__, __LINE__);
    }
    h->groups = bestG;
    h->useGraph = bestGraph != 0;
    *groups = bestG;
    if (ms_per_step) *ms_per_step = best / nSteps;
    const int chosen[4] = {bestG, bestGraph, 0, 0};
    tune_record(kind, 2, h->d.N, dtype, sc, chosen, best / nSteps);
    return SELF_OK;
}
SELF_CATCH_ALL

int self_sw2d_synchronize(self_sw2d_t h)
{
    SELF_REQUIRE(h, SELF_ERR_ARG, "null handle");
    SELF_HIP(hipStreamSynchronize(h->stream));
    return SELF_OK;
}

int self_sw2d_device_bytes(self_sw2d_t h, size_t *bytes)
{
    SELF_REQUIRE(h && bytes, SELF_ERR_ARG, "null argument");
    *bytes = h->bytes;
    return SELF_OK;
}

}  // extern "C"
