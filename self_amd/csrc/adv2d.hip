// adv2d.hip -- 2-D advection-diffusion DGSEM (BR1-type gradient) on gfx950.
//
// Reference: src/highend/advection-diffusion2d/AdvectionClass.f90
//   CalculateConcentrationGradient :1123-1264, ForwardStepRK3 :1268-1312,
//   GlobalTimeDerivative :1316-1378, EdgeFlux :1382-1478,
//   EdgeGradientFlux :1482-1563, MappedTimeDerivative :1567-1723,
//   RiemannSolver :1756-1787, XFlux/YFlux :1791-1825, Source :1829-1847,
//   GetExternalState :1851-1886.
//
// Each RK stage needs two neighbour exchanges (concentration traces for the
// central gradient flux, then gradient traces for the diffusive flux), so a
// stage is two kernels:
//   k_adv_gradient  own traces + neighbour traces -> central gradient flux
//                   -> weak gradient (dc/dx, dc/dy) -> gradient traces
//   k_adv_stage     upwind + diffusive flux of every side node, volume flux
//                   c u - kappa grad c, weak divergence + lift, relaxation
//                   source, low-storage RK3 update, traces of the new state
// Side fluxes are evaluated in the edge's primary orientation (primary
// normal, primary velocity trace, primary + secondary order in the averages)
// so each side holds exactly the value EdgeFlux/EdgeGradientFlux would store.
// Velocity is static: its traces are computed once when it is set (the
// reference recomputes them every stage).

#include <cstring>
#include <vector>

#include "common.hpp"
#include "dg2d_device.hpp"
#include "dg2d_host.hpp"

namespace self {
namespace {

constexpr double kRkA[3] = {0.0, -5.0 / 9.0, -153.0 / 128.0};
constexpr double kRkB[3] = {0.0, 1.0 / 3.0, 3.0 / 4.0};
constexpr double kRkG[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};

struct AdvArgs {
    int nE, mode;  // mode 0: RK update, 1: tendency only
    double kappa, t, a, gdt;
    int hasRelax;
    const double *nodal;
    double *c, *G, *tend, *grad;  // grad [2][nq][nE][np]
    const double *vel;            // [2][nE][np]
    const double *fvel;           // [nE][4][2][n1]
    const double *metric, *fnrm;
    const int2 *side;
    const double *trIn;
    double *trOut;                // c traces [nE][4][nq][n1]
    double *trG;                  // gradient traces [nE][4][2][nq][n1]
    const double *relax, *rFac;   // [nq][nE][np]
};

// GetExternalState :1851-1886
__device__ __forceinline__ double adv_external(int bc, int f, double t)
{
    if (bc == -103 && f == 0) return tanh(t);  // InflowOne
    if (bc == -104 && f == 1) return tanh(t);  // InflowTwo
    return 0.0;
}

template <int N>
__device__ __forceinline__ double trace1(const NodalLDS<N> &L, const double *q, int s, int k)
{
    constexpr int n1 = N + 1;
    const double *l = (s == 0) ? L.lS : (s == 1) ? L.lE : (s == 2) ? L.lN : L.lW;
    const int off = (s == 0 || s == 2) ? k : n1 * k;
    const int stride = (s == 0 || s == 2) ? n1 : 1;
    double acc = l[0] * q[off];
    #pragma unroll
    for (int m = 1; m < n1; ++m) acc = __builtin_fma(l[m], q[off + m * stride], acc);
    return acc;
}

// DGSystemDerivative in s and p of one node: sum_m D(i,m) sl[m] + lift_s,
// sum_m D(j,m) pl[m] + lift_p
template <int N>
__device__ __forceinline__ void dg_pair(const NodalLDS<N> &L, const double *sl, const double *pl, int i, int j,
                                        double bE, double bW, double bN, double bS, double &t1, double &t2)
{
    constexpr int n1 = N + 1;
    double a = L.dS[i] * sl[0], b = L.dP[j] * pl[0];
    #pragma unroll
    for (int m = 1; m < n1; ++m) {
        a = __builtin_fma(L.dS[i + n1 * m], sl[m], a);
        b = __builtin_fma(L.dP[j + n1 * m], pl[n1 * m], b);
    }
    t1 = a + (bE * L.lE[i] + bW * L.lW[i]) * L.iwS[i];
    t2 = b + (bN * L.lN[j] + bS * L.lS[j]) * L.iwP[j];
}

template <int N, int NQ>
__global__ __launch_bounds__(256) void k_adv_traces(const double *__restrict__ q, const double *__restrict__ nodal,
                                                    double *__restrict__ out, int nE)
{
    using S = Shape<N>;
    constexpr int n1 = S::n1, np = S::np, epb = S::epb;
    __shared__ NodalLDS<N> L;
    __shared__ double qs[epb * NQ * np];
    stage_nodal<N, false>(L, nodal, nullptr, nullptr);
    const int t = threadIdx.x, le = t / np, nd = t - le * np;
    const int e = blockIdx.x * epb + le;
    const bool act = le < epb && e < nE;
    if (act)
        for (int f = 0; f < NQ; ++f) qs[(le * NQ + f) * np + nd] = q[((size_t)f * nE + e) * np + nd];
    __syncthreads();
    if (!act) return;
    for (int w = nd; w < 4 * n1; w += np) {
        const int s = w / n1, k = w - s * n1;
        for (int f = 0; f < NQ; ++f)
            out[(((size_t)e * 4 + s) * NQ + f) * n1 + k] = trace1<N>(L, qs + (le * NQ + f) * np, s, k);
    }
}

__global__ void k_adv_fvel(const double *__restrict__ velTr, const int *__restrict__ pnode, double *__restrict__ fvel,
                           int nSideNodes, int n1)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nSideNodes) return;
    const int es = q / n1, k = q - es * n1;
    const int p = pnode[q], pes = p / n1, pk = p - pes * n1;
    // velTr/fvel layout [e*4+s][2][n1]
    fvel[((size_t)es * 2 + 0) * n1 + k] = velTr[((size_t)pes * 2 + 0) * n1 + pk];
    fvel[((size_t)es * 2 + 1) * n1 + k] = velTr[((size_t)pes * 2 + 1) * n1 + pk];
}

// ---- gradient kernel: EdgeGradientFlux + CalculateConcentrationGradient + gradient traces
template <int N, int NQ>
__global__ __launch_bounds__(256) void k_adv_gradient(AdvArgs A)
{
    using S = Shape<N>;
    constexpr int n1 = S::n1, np = S::np, epb = S::epb;
    __shared__ NodalLDS<N> L;
    __shared__ double cs[epb * NQ * np], sx[epb * NQ * np], sy[epb * NQ * np], px[epb * NQ * np],
        py[epb * NQ * np], bf[epb * 4 * 2 * NQ * n1];
    stage_nodal<N, false>(L, A.nodal, nullptr, nullptr);
    const int t = threadIdx.x, le = t / np, nd = t - le * np;
    const int e = blockIdx.x * epb + le;
    const bool act = le < epb && e < A.nE;
    const int i = nd % n1, j = nd / n1;
    const size_t mstride = (size_t)A.nE * np, base = (size_t)e * np + nd;
    const int lo = le * NQ * np;
    double *bfE = bf + le * 4 * 2 * NQ * n1;  // [side][dir][f][k]
    double J = 1.0;
    if (act) {
        const double dxds = A.metric[base], dxdp = A.metric[mstride + base];
        const double dyds = A.metric[2 * mstride + base], dydp = A.metric[3 * mstride + base];
        J = A.metric[4 * mstride + base];
        for (int f = 0; f < NQ; ++f) {
            const double c = A.c[f * mstride + base];
            cs[lo + f * np + nd] = c;
            // contravariant fluxes of (c, 0) and (0, c)
            sx[lo + f * np + nd] = dydp * c;
            sy[lo + f * np + nd] = -(dxdp * c);
            px[lo + f * np + nd] = -(dyds * c);
            py[lo + f * np + nd] = dxds * c;
        }
    }
    __syncthreads();
    if (act) {
        for (int w = nd; w < 4 * n1; w += np) {
            const int sd = w / n1, k = w - sd * n1;
            const int2 info = A.side[(size_t)e * 4 + sd];
            const size_t nb = ((size_t)e * 4 + sd) * 3 * n1 + k;
            const double nx = A.fnrm[nb], ny = A.fnrm[nb + n1], len = A.fnrm[nb + 2 * n1];
            for (int f = 0; f < NQ; ++f) {
                const double own = trace1<N>(L, cs + lo + f * np, sd, k);
                double pri, sec, sign = 1.0;
                if (info.x < 0) {
                    pri = own;
                    sec = adv_external(info.x, f, A.t);
                } else {
                    const int kk = side_flip(info.y) ? N - k : k;
                    const double nbr = A.trIn[(((size_t)info.x * 4 + side_nbr_side(info.y)) * NQ + f) * n1 + kk];
                    if (side_primary(info.y)) { pri = own; sec = nbr; }
                    else { pri = nbr; sec = own; sign = -1.0; }
                }
                const double avg = 0.5 * (pri + sec);
                bfE[((sd * 2 + 0) * NQ + f) * n1 + k] = sign * (avg * nx * len);
                bfE[((sd * 2 + 1) * NQ + f) * n1 + k] = sign * (avg * ny * len);
            }
        }
    }
    __syncthreads();
    double gxv[NQ], gyv[NQ];
    if (act) {
        for (int f = 0; f < NQ; ++f) {
            const int o = lo + f * np;
            double t1, t2, u1, u2;
            dg_pair<N>(L, sx + o + n1 * j, px + o + i, i, j, bfE[((1 * 2 + 0) * NQ + f) * n1 + j],
                       bfE[((3 * 2 + 0) * NQ + f) * n1 + j], bfE[((2 * 2 + 0) * NQ + f) * n1 + i],
                       bfE[((0 * 2 + 0) * NQ + f) * n1 + i], t1, t2);
            dg_pair<N>(L, sy + o + n1 * j, py + o + i, i, j, bfE[((1 * 2 + 1) * NQ + f) * n1 + j],
                       bfE[((3 * 2 + 1) * NQ + f) * n1 + j], bfE[((2 * 2 + 1) * NQ + f) * n1 + i],
                       bfE[((0 * 2 + 1) * NQ + f) * n1 + i], u1, u2);
            gxv[f] = (t1 + t2) / J;
            gyv[f] = (u1 + u2) / J;
            A.grad[((size_t)0 * NQ + f) * mstride + base] = gxv[f];
            A.grad[((size_t)1 * NQ + f) * mstride + base] = gyv[f];
        }
    }
    __syncthreads();  // every lane is done reading sx/sy before they hold the gradient
    if (act) {
        for (int f = 0; f < NQ; ++f) {
            sx[lo + f * np + nd] = gxv[f];
            sy[lo + f * np + nd] = gyv[f];
        }
    }
    __syncthreads();
    if (act) {
        for (int w = nd; w < 4 * n1; w += np) {
            const int sd = w / n1, k = w - sd * n1;
            for (int f = 0; f < NQ; ++f) {
                A.trG[((((size_t)e * 4 + sd) * 2 + 0) * NQ + f) * n1 + k] = trace1<N>(L, sx + lo + f * np, sd, k);
                A.trG[((((size_t)e * 4 + sd) * 2 + 1) * NQ + f) * n1 + k] = trace1<N>(L, sy + lo + f * np, sd, k);
            }
        }
    }
}

// ---- stage kernel: EdgeFlux + MappedTimeDerivative + Source + RK3 update
template <int N, int NQ>
__global__ __launch_bounds__(256) void k_adv_stage(AdvArgs A)
{
    using S = Shape<N>;
    constexpr int n1 = S::n1, np = S::np, epb = S::epb;
    __shared__ NodalLDS<N> L;
    __shared__ double cs[epb * NQ * np], gx[epb * NQ * np], gy[epb * NQ * np];
    __shared__ double sC[epb * NQ * np], pC[epb * NQ * np], bf[epb * 4 * NQ * n1];
    stage_nodal<N, false>(L, A.nodal, nullptr, nullptr);
    const int t = threadIdx.x, le = t / np, nd = t - le * np;
    const int e = blockIdx.x * epb + le;
    const bool act = le < epb && e < A.nE;
    const int i = nd % n1, j = nd / n1;
    const size_t mstride = (size_t)A.nE * np, base = (size_t)e * np + nd;
    double *cE = cs + le * NQ * np, *gxE = gx + le * NQ * np, *gyE = gy + le * NQ * np;
    double *sCE = sC + le * NQ * np, *pCE = pC + le * NQ * np, *bfE = bf + le * 4 * NQ * n1;
    double c[NQ], J = 1.0;
    if (act) {
        const double dxds = A.metric[base], dxdp = A.metric[mstride + base];
        const double dyds = A.metric[2 * mstride + base], dydp = A.metric[3 * mstride + base];
        J = A.metric[4 * mstride + base];
        const double u = A.vel[base], v = A.vel[mstride + base];
        for (int f = 0; f < NQ; ++f) {
            c[f] = A.c[f * mstride + base];
            const double dx = A.grad[((size_t)0 * NQ + f) * mstride + base];
            const double dy = A.grad[((size_t)1 * NQ + f) * mstride + base];
            cE[f * np + nd] = c[f];
            gxE[f * np + nd] = dx;
            gyE[f * np + nd] = dy;
            const double xF = c[f] * u - A.kappa * dx;  // XFlux :1791-1807
            const double yF = c[f] * v - A.kappa * dy;  // YFlux :1809-1825
            sCE[f * np + nd] = dydp * xF - dxdp * yF;
            pCE[f * np + nd] = -dyds * xF + dxds * yF;
        }
    }
    __syncthreads();
    if (act) {
        for (int w = nd; w < 4 * n1; w += np) {
            const int sd = w / n1, k = w - sd * n1;
            const int2 info = A.side[(size_t)e * 4 + sd];
            const size_t nb = ((size_t)e * 4 + sd) * 3 * n1 + k;
            const double nx = A.fnrm[nb], ny = A.fnrm[nb + n1], len = A.fnrm[nb + 2 * n1];
            const double u = A.fvel[(((size_t)e * 4 + sd) * 2 + 0) * n1 + k];
            const double v = A.fvel[(((size_t)e * 4 + sd) * 2 + 1) * n1 + k];
            const double uN = u * nx + v * ny;
            for (int f = 0; f < NQ; ++f) {
                const double ownC = trace1<N>(L, cE + f * np, sd, k);
                const double ownX = trace1<N>(L, gxE + f * np, sd, k);
                const double ownY = trace1<N>(L, gyE + f * np, sd, k);
                double in, ex, dx, dy, sign = 1.0;
                if (info.x < 0) {
                    in = ownC; ex = adv_external(info.x, f, A.t); dx = ownX; dy = ownY;
                } else {
                    const int kk = side_flip(info.y) ? N - k : k;
                    const size_t ns = (size_t)info.x * 4 + side_nbr_side(info.y);
                    const double nbC = A.trIn[(ns * NQ + f) * n1 + kk];
                    const double nbX = A.trG[((ns * 2 + 0) * NQ + f) * n1 + kk];
                    const double nbY = A.trG[((ns * 2 + 1) * NQ + f) * n1 + kk];
                    if (side_primary(info.y)) {
                        in = ownC; ex = nbC; dx = (ownX + nbX) * 0.5; dy = (ownY + nbY) * 0.5;
                    } else {
                        in = nbC; ex = ownC; dx = (nbX + ownX) * 0.5; dy = (nbY + ownY) * 0.5; sign = -1.0;
                    }
                }
                const double jump = ex - in;
                double fl = 0.5 * (uN * (ex + in) - fabs(uN) * jump);  // RiemannSolver :1756-1787
                const double dcdn = dx * nx + dy * ny;
                fl = fl - A.kappa * dcdn;
                bfE[(sd * NQ + f) * n1 + k] = sign * (fl * len);
            }
        }
    }
    __syncthreads();
    if (act) {
        for (int f = 0; f < NQ; ++f) {
            double t1, t2;
            dg_pair<N>(L, sCE + f * np + n1 * j, pCE + f * np + i, i, j, bfE[(1 * NQ + f) * n1 + j],
                       bfE[(3 * NQ + f) * n1 + j], bfE[(2 * NQ + f) * n1 + i], bfE[(0 * NQ + f) * n1 + i], t1, t2);
            double tend = (-t1 - t2) / J;
            if (A.hasRelax) tend = tend + (A.relax[f * mstride + base] - c[f]) * A.rFac[f * mstride + base];
            if (A.mode == 1) {
                A.tend[f * mstride + base] = tend;
            } else {
                const double gold = (A.a != 0.0) ? A.G[f * mstride + base] : 0.0;
                const double gv = A.a * gold + tend;
                A.G[f * mstride + base] = gv;
                const double qn = c[f] + A.gdt * gv;
                A.c[f * mstride + base] = qn;
                cE[f * np + nd] = qn;
            }
        }
    }
    if (A.mode == 1) return;
    __syncthreads();
    if (act) {
        for (int w = nd; w < 4 * n1; w += np) {
            const int sd = w / n1, k = w - sd * n1;
            for (int f = 0; f < NQ; ++f)
                A.trOut[(((size_t)e * 4 + sd) * NQ + f) * n1 + k] = trace1<N>(L, cE + f * np, sd, k);
        }
    }
}

}  // namespace
}  // namespace self

using namespace self;

struct self_adv2d_s {
    int N = 0, n1 = 0, np = 0, nq = 0, device = 0;
    size_t nE = 0;
    double kappa = 0.0, dt = 0.0;
    hipStream_t stream = nullptr;
    double *nodal = nullptr, *c = nullptr, *G = nullptr, *tend = nullptr, *grad = nullptr, *vel = nullptr;
    double *velTr = nullptr, *fvel = nullptr, *metric = nullptr, *fnrm = nullptr, *tr[2] = {nullptr, nullptr};
    double *trG = nullptr, *relax = nullptr, *rFac = nullptr;
    int2 *side = nullptr;
    int *pnode = nullptr;
    int cur = 0;
    bool trValid = false, ready = false, hasRelax = false;
};

namespace {

#define ADV_DISPATCH(N_, NQ_, CALL)                                                         \
    switch ((N_) * 10 + (NQ_)) {                                                            \
    case 11: { constexpr int NN = 1, QQ = 1; CALL; } break;                                 \
    case 21: { constexpr int NN = 2, QQ = 1; CALL; } break;                                 \
    case 31: { constexpr int NN = 3, QQ = 1; CALL; } break;                                 \
    case 41: { constexpr int NN = 4, QQ = 1; CALL; } break;                                 \
    case 51: { constexpr int NN = 5, QQ = 1; CALL; } break;                                 \
    case 61: { constexpr int NN = 6, QQ = 1; CALL; } break;                                 \
    case 71: { constexpr int NN = 7, QQ = 1; CALL; } break;                                 \
    case 81: { constexpr int NN = 8, QQ = 1; CALL; } break;                                 \
    case 12: { constexpr int NN = 1, QQ = 2; CALL; } break;                                 \
    case 22: { constexpr int NN = 2, QQ = 2; CALL; } break;                                 \
    case 32: { constexpr int NN = 3, QQ = 2; CALL; } break;                                 \
    case 42: { constexpr int NN = 4, QQ = 2; CALL; } break;                                 \
    case 52: { constexpr int NN = 5, QQ = 2; CALL; } break;                                 \
    case 62: { constexpr int NN = 6, QQ = 2; CALL; } break;                                 \
    case 72: { constexpr int NN = 7, QQ = 2; CALL; } break;                                 \
    case 82: { constexpr int NN = 8, QQ = 2; CALL; } break;                                 \
    case 13: { constexpr int NN = 1, QQ = 3; CALL; } break;                                 \
    case 23: { constexpr int NN = 2, QQ = 3; CALL; } break;                                 \
    case 33: { constexpr int NN = 3, QQ = 3; CALL; } break;                                 \
    case 43: { constexpr int NN = 4, QQ = 3; CALL; } break;                                 \
    case 53: { constexpr int NN = 5, QQ = 3; CALL; } break;                                 \
    case 63: { constexpr int NN = 6, QQ = 3; CALL; } break;                                 \
    case 73: { constexpr int NN = 7, QQ = 3; CALL; } break;                                 \
    case 83: { constexpr int NN = 8, QQ = 3; CALL; } break;                                 \
    case 14: { constexpr int NN = 1, QQ = 4; CALL; } break;                                 \
    case 24: { constexpr int NN = 2, QQ = 4; CALL; } break;                                 \
    case 34: { constexpr int NN = 3, QQ = 4; CALL; } break;                                 \
    case 44: { constexpr int NN = 4, QQ = 4; CALL; } break;                                 \
    case 54: { constexpr int NN = 5, QQ = 4; CALL; } break;                                 \
    case 64: { constexpr int NN = 6, QQ = 4; CALL; } break;                                 \
    case 74: { constexpr int NN = 7, QQ = 4; CALL; } break;                                 \
    case 84: { constexpr int NN = 8, QQ = 4; CALL; } break;                                 \
    default: break;                                                                         \
    }

template <int N>
int agrid(size_t nE) { return (int)((nE + Shape<N>::epb - 1) / Shape<N>::epb); }

AdvArgs adv_args(self_adv2d_s *h)
{
    AdvArgs A{};
    A.nE = (int)h->nE;
    A.kappa = h->kappa;
    A.hasRelax = h->hasRelax;
    A.nodal = h->nodal;
    A.c = h->c;
    A.G = h->G;
    A.tend = h->tend;
    A.grad = h->grad;
    A.vel = h->vel;
    A.fvel = h->fvel;
    A.metric = h->metric;
    A.fnrm = h->fnrm;
    A.side = h->side;
    A.trG = h->trG;
    A.relax = h->relax;
    A.rFac = h->rFac;
    return A;
}

int launch_c_traces(self_adv2d_s *h, double *dst)
{
    ADV_DISPATCH(h->N, h->nq, hipLaunchKernelGGL((k_adv_traces<NN, QQ>), dim3(agrid<NN>(h->nE)), dim3(256), 0,
                                                 h->stream, h->c, h->nodal, dst, (int)h->nE));
    SELF_HIP(hipGetLastError());
    return SELF_OK;
}

int stage(self_adv2d_s *h, double t, int m, int mode)
{
    AdvArgs A = adv_args(h);
    A.t = t;
    A.mode = mode;
    A.a = m >= 0 ? kRkA[m] : 0.0;
    A.gdt = m >= 0 ? kRkG[m] * h->dt : 0.0;
    A.trIn = h->tr[h->cur];
    A.trOut = h->tr[1 - h->cur];
    ADV_DISPATCH(h->N, h->nq, hipLaunchKernelGGL((k_adv_gradient<NN, QQ>), dim3(agrid<NN>(h->nE)), dim3(256), 0,
                                                 h->stream, A));
    ADV_DISPATCH(h->N, h->nq, hipLaunchKernelGGL((k_adv_stage<NN, QQ>), dim3(agrid<NN>(h->nE)), dim3(256), 0,
                                                 h->stream, A));
    SELF_HIP(hipGetLastError());
    if (mode == 0) h->cur = 1 - h->cur;
    return SELF_OK;
}

void from_dev(const double *src, double *dst, size_t nE, int nf, int np)
{
    for (size_t e = 0; e < nE; ++e)
        for (int f = 0; f < nf; ++f)
            std::memcpy(dst + (e * nf + f) * np, src + ((size_t)f * nE + e) * np, sizeof(double) * np);
}

void to_dev(const double *src, double *dst, size_t nE, int nf, int np)
{
    for (size_t e = 0; e < nE; ++e)
        for (int f = 0; f < nf; ++f)
            std::memcpy(dst + ((size_t)f * nE + e) * np, src + (e * nf + f) * np, sizeof(double) * np);
}

}  // namespace

extern "C" {

int self_adv2d_create(int N, int nTracers, int nElems, int nEdges, const int *edges, const double *dxds,
                      const double *dxdp, const double *dyds, const double *dydp, const double *J, const double *nHat,
                      const double *nLen, const double *dMatS, const double *dMatP, const double *qWeight,
                      const double *lagWest, const double *lagEast, const double *lagSouth, const double *lagNorth,
                      double kappa, double dt, self_adv2d_t *out)
try {
    SELF_REQUIRE(out && edges && dxds && dxdp && dyds && dydp && J && nHat && nLen && dMatS && dMatP && qWeight &&
                     lagWest && lagEast && lagSouth && lagNorth,
                 SELF_ERR_ARG, "self_adv2d_create: null argument");
    SELF_REQUIRE(N >= 1 && N <= 8 && nTracers >= 1 && nTracers <= 4 && nElems >= 1, SELF_ERR_ARG,
                 "self_adv2d_create: N=%d (1..8), nTracers=%d (1..4)", N, nTracers);
    Geo2D G;
    int rc = build_geo2d(N, (size_t)nElems, nEdges, edges, dxds, dxdp, dyds, dydp, J, nHat, nLen, G);
    if (rc) return rc;
    auto *h = new self_adv2d_s();
    h->N = N;
    h->n1 = N + 1;
    h->np = h->n1 * h->n1;
    h->nq = nTracers;
    h->nE = (size_t)nElems;
    h->kappa = kappa;
    h->dt = dt;
    const size_t nE = h->nE, np = h->np, n1 = h->n1, nq = h->nq;
    auto fail = [&](hipError_t e) {
        void *p[] = {h->nodal, h->c, h->G, h->tend, h->grad, h->vel, h->velTr, h->fvel, h->metric, h->fnrm,
                     h->tr[0], h->tr[1], h->trG, h->relax, h->rFac, h->side, h->pnode};
        for (void *q : p)
            if (q) (void)hipFree(q);
        if (h->stream) (void)hipStreamDestroy(h->stream);
        delete h;
        return hip_fail(e, "self_adv2d_create", __FILE__, __LINE__);
    };
    hipError_t e = hipGetDevice(&h->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    auto M = [&](double **p, size_t n) { if (e == hipSuccess) e = hipMalloc(p, sizeof(double) * n); };
    M(&h->nodal, nodal_buffer_size(n1));
    M(&h->c, nq * nE * np);
    M(&h->G, nq * nE * np);
    M(&h->tend, nq * nE * np);
    M(&h->grad, 2 * nq * nE * np);
    M(&h->vel, 2 * nE * np);
    M(&h->velTr, nE * 8 * n1);
    M(&h->fvel, nE * 8 * n1);
    M(&h->metric, 5 * nE * np);
    M(&h->fnrm, nE * 12 * n1);
    M(&h->tr[0], nE * 4 * nq * n1);
    M(&h->tr[1], nE * 4 * nq * n1);
    M(&h->trG, nE * 8 * nq * n1);
    if (e == hipSuccess) e = hipMalloc(&h->side, sizeof(int2) * nE * 4);
    if (e == hipSuccess) e = hipMalloc(&h->pnode, sizeof(int) * nE * 4 * n1);
    if (e != hipSuccess) return fail(e);
    std::vector<double> nod(nodal_buffer_size(n1));
    std::memcpy(nod.data(), dMatS, sizeof(double) * n1 * n1);
    std::memcpy(nod.data() + n1 * n1, dMatP, sizeof(double) * n1 * n1);
    double *p = nod.data() + 2 * n1 * n1;
    std::memcpy(p, qWeight, sizeof(double) * n1);
    std::memcpy(p + n1, qWeight, sizeof(double) * n1);
    std::memcpy(p + 2 * n1, lagWest, sizeof(double) * n1);
    std::memcpy(p + 3 * n1, lagEast, sizeof(double) * n1);
    std::memcpy(p + 4 * n1, lagSouth, sizeof(double) * n1);
    std::memcpy(p + 5 * n1, lagNorth, sizeof(double) * n1);
    nodal_buffer_derived(nod.data(), n1);
    std::vector<int2> sd(nE * 4);
    for (size_t k = 0; k < nE * 4; ++k) sd[k] = make_int2(G.sideX[k], G.sideY[k]);
    e = hipMemcpy(h->nodal, nod.data(), sizeof(double) * nod.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h->metric, G.metric.data(), sizeof(double) * G.metric.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h->fnrm, G.fnrm.data(), sizeof(double) * G.fnrm.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h->side, sd.data(), sizeof(int2) * sd.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h->pnode, G.primaryNode.data(), sizeof(int) * G.primaryNode.size(),
                                       hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(h->c, 0, sizeof(double) * nq * nE * np);
    if (e == hipSuccess) e = hipMemset(h->vel, 0, sizeof(double) * 2 * nE * np);
    if (e == hipSuccess) e = hipMemset(h->fvel, 0, sizeof(double) * nE * 8 * n1);
    if (e != hipSuccess) return fail(e);
    *out = h;
    return SELF_OK;
}
SELF_CATCH_ALL

int self_adv2d_destroy(self_adv2d_t h)
{
    if (!h) return SELF_OK;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    void *p[] = {h->nodal, h->c, h->G, h->tend, h->grad, h->vel, h->velTr, h->fvel, h->metric, h->fnrm,
                 h->tr[0], h->tr[1], h->trG, h->relax, h->rFac, h->side, h->pnode};
    for (void *q : p)
        if (q) (void)hipFree(q);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return SELF_OK;
}

// field: 0 concentration (0:N,0:N,1:nTracers), 1 velocity (0:N,0:N,1:2),
// 2 relaxation field, 3 relaxation factor (both (0:N,0:N,1:nTracers))
int self_adv2d_set_field(self_adv2d_t h, int field, const double *data)
try {
    SELF_REQUIRE(h && data, SELF_ERR_ARG, "null argument");
    SELF_HIP(hipSetDevice(h->device));
    SELF_HIP(hipStreamSynchronize(h->stream));
    const size_t nE = h->nE, np = h->np;
    const int nf = field == 1 ? 2 : h->nq;
    std::vector<double> buf((size_t)nf * nE * np);
    to_dev(data, buf.data(), nE, nf, (int)np);
    double **dst = nullptr;
    switch (field) {
    case 0: dst = &h->c; h->trValid = false; break;
    case 1: dst = &h->vel; break;
    case 2: dst = &h->relax; h->hasRelax = true; break;
    case 3: dst = &h->rFac; h->hasRelax = true; break;
    default: SELF_REQUIRE(false, SELF_ERR_ARG, "field %d", field);
    }
    if (!*dst) SELF_HIP(hipMalloc(dst, sizeof(double) * buf.size()));
    SELF_HIP(hipMemcpy(*dst, buf.data(), sizeof(double) * buf.size(), hipMemcpyHostToDevice));
    if (field == 2 || field == 3) {
        if (!h->relax) { SELF_HIP(hipMalloc(&h->relax, sizeof(double) * buf.size())); SELF_HIP(hipMemset(h->relax, 0, sizeof(double) * buf.size())); }
        if (!h->rFac) { SELF_HIP(hipMalloc(&h->rFac, sizeof(double) * buf.size())); SELF_HIP(hipMemset(h->rFac, 0, sizeof(double) * buf.size())); }
    }
    if (field == 1) {  // CalculateVelocityAtBoundaries, once: velocity is static
        const int N = h->N;
        ADV_DISPATCH(N, 2, hipLaunchKernelGGL((k_adv_traces<NN, QQ>), dim3(agrid<NN>(nE)), dim3(256), 0, h->stream,
                                              h->vel, h->nodal, h->velTr, (int)nE));
        const int nsn = (int)(nE * 4 * h->n1);
        hipLaunchKernelGGL(k_adv_fvel, dim3((nsn + 255) / 256), dim3(256), 0, h->stream, h->velTr, h->pnode, h->fvel,
                           nsn, h->n1);
        SELF_HIP(hipGetLastError());
        SELF_HIP(hipStreamSynchronize(h->stream));
    }
    return SELF_OK;
}
SELF_CATCH_ALL

// field: 0 concentration, 1 tendency, 2 dc/dx, 3 dc/dy, 4 velocity
// (0:N,0:N,1:2), 5 relaxation field, 6 relaxation factor (zero if never set)
int self_adv2d_get_field(self_adv2d_t h, int field, double *data)
try {
    SELF_REQUIRE(h && data, SELF_ERR_ARG, "null argument");
    SELF_HIP(hipSetDevice(h->device));
    SELF_HIP(hipStreamSynchronize(h->stream));
    const size_t nE = h->nE, np = h->np;
    size_t nq = h->nq;
    const double *src = nullptr;
    switch (field) {
    case 0: src = h->c; break;
    case 1: src = h->tend; break;
    case 2: src = h->grad; break;
    case 3: src = h->grad + nq * nE * np; break;
    case 4: src = h->vel; nq = 2; break;
    case 5: src = h->relax; break;
    case 6: src = h->rFac; break;
    default: SELF_REQUIRE(false, SELF_ERR_ARG, "field %d", field);
    }
    if (!src) {  // a field never set holds zeros, as the reference's Build leaves it
        std::memset(data, 0, sizeof(double) * nq * nE * np);
        return SELF_OK;
    }
    std::vector<double> buf(nq * nE * np);
    SELF_HIP(hipMemcpy(buf.data(), src, sizeof(double) * buf.size(), hipMemcpyDeviceToHost));
    from_dev(buf.data(), data, nE, (int)nq, (int)np);
    return SELF_OK;
}
SELF_CATCH_ALL

int self_adv2d_global_time_derivative(self_adv2d_t h, double t)
{
    SELF_REQUIRE(h, SELF_ERR_ARG, "null handle");
    SELF_HIP(hipSetDevice(h->device));
    int rc = launch_c_traces(h, h->tr[h->cur]);
    if (rc) return rc;
    h->trValid = true;
    return stage(h, t, -1, 1);
}

int self_adv2d_forward_step_rk3(self_adv2d_t h, double tn, int nSteps, double *elapsed_ms)
{
    SELF_REQUIRE(h && nSteps >= 0, SELF_ERR_ARG, "bad arguments");
    SELF_HIP(hipSetDevice(h->device));
    int rc = SELF_OK;
    if (!h->trValid) {
        if ((rc = launch_c_traces(h, h->tr[h->cur]))) return rc;
        h->trValid = true;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (elapsed_ms) {
        SELF_HIP(hipEventCreate(&e0));
        SELF_HIP(hipEventCreate(&e1));
        SELF_HIP(hipEventRecord(e0, h->stream));
    }
    for (int s = 0; s < nSteps && !rc; ++s) {
        const double t0 = tn + s * h->dt;
        for (int m = 0; m < 3 && !rc; ++m) rc = stage(h, t0 + kRkB[m] * h->dt, m, 0);
    }
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

// WritePickup_Advection (AdvectionClass.f90:2040-2103)
int self_adv2d_write_pickup(self_adv2d_t h, const char *path)
try {
    SELF_REQUIRE(h && path, SELF_ERR_ARG, "null argument");
    const size_t nv = h->nE * h->np, q = (size_t)h->nq;
    std::vector<double> c(q * nv), relax(q * nv), rFac(q * nv), vel(2 * nv);
    int rc = self_adv2d_get_field(h, 0, c.data());
    if (!rc) rc = self_adv2d_get_field(h, 5, relax.data());
    if (!rc) rc = self_adv2d_get_field(h, 6, rFac.data());
    if (!rc) rc = self_adv2d_get_field(h, 4, vel.data());
    if (!rc) rc = self_adv2d_pickup_write(path, h->N, (int)h->nE, h->nq, c.data(), relax.data(), rFac.data(), vel.data());
    return rc;
}
SELF_CATCH_ALL

// ReadPickup_Advection (:2107-2171): solution, relaxation field and factor, velocity
int self_adv2d_read_pickup(self_adv2d_t h, const char *path)
try {
    SELF_REQUIRE(h && path, SELF_ERR_ARG, "null argument");
    const size_t nv = h->nE * h->np, q = (size_t)h->nq;
    std::vector<double> c(q * nv), relax(q * nv), rFac(q * nv), vel(2 * nv);
    int rc = self_adv2d_pickup_read(path, h->N, (int)h->nE, h->nq, c.data(), relax.data(), rFac.data(), vel.data());
    if (!rc) rc = self_adv2d_set_field(h, 0, c.data());
    if (!rc) rc = self_adv2d_set_field(h, 2, relax.data());
    if (!rc) rc = self_adv2d_set_field(h, 3, rFac.data());
    if (!rc) rc = self_adv2d_set_field(h, 1, vel.data());
    return rc;
}
SELF_CATCH_ALL

int self_adv2d_synchronize(self_adv2d_t h)
{
    SELF_REQUIRE(h, SELF_ERR_ARG, "null handle");
    SELF_HIP(hipStreamSynchronize(h->stream));
    return SELF_OK;
}

}  // extern "C"
