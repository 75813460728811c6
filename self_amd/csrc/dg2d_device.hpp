// dg2d_device.hpp -- device building blocks shared by the 2-D DGSEM kernels.
//
// Element storage on the device is structure-of-arrays: a nodal field f of
// element e lives at  base[(f*nElems + e)*NP + node],  node = i + (N+1)*j,
// so the 64 nodes of an N=7 element are one contiguous 512-byte line per
// field and a wavefront reads them with one coalesced load.  Face traces
// and face fluxes are kept per element as [e][side][field][node] so one
// neighbouring side is a contiguous (N+1)*nEq run.
//
// Sums over nodes use explicit fused multiply-adds in a fixed order, so a
// trace computed in two different kernels (or twice in one) is bitwise the
// same value: neighbours read exactly what the owner used.
#pragma once

#include <hip/hip_runtime.h>

namespace self {

template <int N>
struct Shape {
    static constexpr int n1 = N + 1;
    static constexpr int np = n1 * n1;
    static constexpr int threads = 256;
    static constexpr int epb = (threads / np) > 0 ? threads / np : 1;  // elements per block
};

// dot of a Lagrange vector with a strided line of nodal values
template <int N, typename T>
__device__ __forceinline__ double line_dot(const double *__restrict__ l, const T *v, int stride)
{
    double acc = l[0] * (double)v[0];
    #pragma unroll
    for (int m = 1; m <= N; ++m) acc = __builtin_fma(l[m], (double)v[m * stride], acc);
    return acc;
}

// Wave speed of the shallow-water class as the examples ship it
// (examples/shallowwater/planewave/localmods/ConservativeShallowWaterClass.f90:1555-1558):
// max over both sides of |(u_n +- sqrt(u_n^2 + 2 g H^3)) / (2H)|.
__device__ __forceinline__ double sw_example_speed(double uOut, double pR, double uIn, double pL, double g)
{
    const double sR = sqrt(uOut * uOut + 2.0 * g * (pR * pR * pR));
    const double sL = sqrt(uIn * uIn + 2.0 * g * (pL * pL * pL));
    return fmax(fmax(fabs(0.5 * (uOut + sR) / pR), fabs(0.5 * (uOut - sR) / pR)),
                fmax(fabs(0.5 * (uIn + sL) / pL), fabs(0.5 * (uIn - sL) / pL)));
}

// Rusanov-type flux of the conservative shallow-water equations,
// RiemannSolver, ConservativeShallowWaterClass.f90:1555-1604 (with g in the
// wave speed, quirk Q2 fixed).  in/out are (U, V, H) on the two sides.
// example: the wave speed of the examples' class (sw_example_speed).
__device__ __forceinline__ void sw_riemann(const double in[3], const double out[3], double nx,
                                           double ny, double g, double f[3], bool example = false)
{
    const double uL = in[0], vL = in[1], pL = in[2];
    const double uR = out[0], vR = out[1], pR = out[2];
    const double uOut = uR * nx + vR * ny;
    const double uIn = uL * nx + vL * ny;
    const double cR = sqrt(g * pR), cL = sqrt(g * pL);
    const double wR = uOut / pR, wL = uIn / pL;
    const double fac = example ? sw_example_speed(uOut, pR, uIn, pL, g)
                               : fmax(fmax(fabs(wR + cR), fabs(wR - cR)), fmax(fabs(wL + cL), fabs(wL - cL)));
    const double hgR = 0.5 * g * pR * pR, hgL = 0.5 * g * pL * pL;
    const double a0 = wR * uR + hgR * nx + wL * uL + hgL * nx;
    const double a1 = wR * vR + hgR * ny + wL * vL + hgL * ny;
    const double a2 = uOut + uIn;
    f[0] = 0.5 * (a0 - fac * (out[0] - in[0]));
    f[1] = 0.5 * (a1 - fac * (out[1] - in[1]));
    f[2] = 0.5 * (a2 - fac * (out[2] - in[2]));
}

// 1/x from the FP32 reciprocal refined by two FP64 Newton steps
// (r <- r + r (1 - x r)): within an ulp or two of the IEEE quotient, for x
// inside the FP32 range; used by the mixed-precision form only.
__device__ __forceinline__ double rcp_refined(double x)
{
    double r = (double)__builtin_amdgcn_rcpf((float)x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}

// Rusanov flux with the mixed form's arithmetic: the central part in FP64
// (velocities from refined reciprocals), the dissipation coefficient
// max |u_n +- c| in FP32 (it multiplies the state jump only).
__device__ __forceinline__ void sw_riemann_mixed(const double in[3], const double out[3], double nx, double ny,
                                                 double g, double f[3], bool example = false)
{
    const double uL = in[0], vL = in[1], pL = in[2];
    const double uR = out[0], vR = out[1], pR = out[2];
    const double uOut = uR * nx + vR * ny;
    const double uIn = uL * nx + vL * ny;
    const double wR = uOut * rcp_refined(pR), wL = uIn * rcp_refined(pL);
    const float cR = __builtin_amdgcn_sqrtf((float)(g * pR)), cL = __builtin_amdgcn_sqrtf((float)(g * pL));
    const float wRf = (float)wR, wLf = (float)wL;
    const double fac = example ? sw_example_speed(uOut, pR, uIn, pL, g)
                               : (double)fmaxf(fmaxf(fabsf(wRf + cR), fabsf(wRf - cR)),
                                               fmaxf(fabsf(wLf + cL), fabsf(wLf - cL)));
    const double hgR = 0.5 * g * pR * pR, hgL = 0.5 * g * pL * pL;
    const double a0 = wR * uR + hgR * nx + wL * uL + hgL * nx;
    const double a1 = wR * vR + hgR * ny + wL * vL + hgL * ny;
    const double a2 = uOut + uIn;
    f[0] = 0.5 * (a0 - fac * (out[0] - in[0]));
    f[1] = 0.5 * (a1 - fac * (out[1] - in[1]));
    f[2] = 0.5 * (a2 - fac * (out[2] - in[2]));
}

// GetExternalState, ConservativeShallowWaterClass.f90:1698-1729
__device__ __forceinline__ void sw_external(int bc, double nx, double ny, const double in[3], double ex[3])
{
    if (bc == -100) {  // NO_NORMAL_FLOW: reflect the momentum about the wall
        ex[0] = (ny * ny - nx * nx) * in[0] - 2.0 * nx * ny * in[1];
        ex[1] = (nx * nx - ny * ny) * in[1] - 2.0 * nx * ny * in[0];
        ex[2] = in[2];
    } else {           // RADIATION and unknown flags: zero state
        ex[0] = 0.0; ex[1] = 0.0; ex[2] = 0.0;
    }
}

// XFlux / YFlux, ConservativeShallowWaterClass.f90:1608-1658
template <bool FAST = false>
__device__ __forceinline__ void sw_fluxes(const double s[3], double g, double xF[3], double yF[3])
{
    // one reciprocal instead of four divisions (differs from U*U/H by at most 1 ulp)
    const double invH = FAST ? rcp_refined(s[2]) : 1.0 / s[2];
    const double hp = g * 0.5 * s[2] * s[2];
    xF[0] = s[0] * s[0] * invH + hp;
    xF[1] = s[0] * s[1] * invH;
    xF[2] = s[0];
    yF[0] = s[1] * s[0] * invH;
    yF[1] = s[1] * s[1] * invH + hp;
    yF[2] = s[1];
}

// Side descriptor packed per element side (int2):
//   x >= 0 : neighbour element (0-based); x < 0 : boundary condition flag
//   y bits : 0-1 neighbour side (0..3), 2 flip (k -> N-k), 3 this side is the edge's primary
__device__ __forceinline__ int side_nbr_side(int y) { return y & 3; }
__device__ __forceinline__ bool side_flip(int y) { return (y >> 2) & 1; }
__device__ __forceinline__ bool side_primary(int y) { return (y >> 3) & 1; }

// XCD-aware block order: the hardware hands consecutive workgroups to the 8
// XCDs round-robin; remapping gives each XCD one contiguous run of elements,
// so neighbours share an L2 and DRAM pages.
__device__ __forceinline__ int xcd_block(int b, int nb)
{
    if ((nb & 7) != 0) return b;
    return (b & 7) * (nb >> 3) + (b >> 3);
}

// Nodal constants staged in LDS once per workgroup: weak derivative
// matrices (column-major, row i col m at i + n1*m), reciprocal quadrature
// weights and the four boundary Lagrange vectors.  Reading them from LDS
// (mostly broadcast reads) keeps them out of the VGPR budget.
// rsS, rsP: row sums sum_m D(i, m) of the two matrices (FP64, host), used by
// the mixed form to carry a line's reference value exactly: D (r 1) = r rs.
template <int N>
struct NodalLDS {
    static constexpr int n1 = N + 1;
    double dS[n1 * n1], dP[n1 * n1];
    double iwS[n1], iwP[n1];
    double lW[n1], lE[n1], lS[n1], lN[n1];
    double rsS[n1], rsP[n1];
    float dS32[n1 * n1], dP32[n1 * n1];
};

// global nodal buffer (doubles): dS, dP, qwS, qwP, lagW, lagE, lagS, lagN,
// then 1/qwS, 1/qwP (computed on the host: the same IEEE quotients), then
// the row sums of dS and dP
constexpr int nodal_buffer_size(int n1) { return 2 * n1 * n1 + 10 * n1; }

// Host side: fills the inverse weights and row sums of a nodal buffer whose
// first 2 n1^2 + 6 n1 entries (matrices, weights, Lagrange vectors) are set.
inline void nodal_buffer_derived(double *buf, int n1)
{
    double *p = buf + 2 * n1 * n1;
    for (int k = 0; k < n1; ++k) {
        p[6 * n1 + k] = 1.0 / p[k];
        p[7 * n1 + k] = 1.0 / p[n1 + k];
    }
    for (int i = 0; i < n1; ++i) {
        double a = 0.0, b = 0.0;
        for (int m = 0; m < n1; ++m) {
            a += buf[i + n1 * m];
            b += buf[n1 * n1 + i + n1 * m];
        }
        p[8 * n1 + i] = a;
        p[9 * n1 + i] = b;
    }
}

// lane / lanes: the threads that stage this copy (the whole workgroup by
// default, or one wavefront staging its own copy).
template <int N, bool F32>
__device__ __forceinline__ void stage_nodal(NodalLDS<N> &L, const double *__restrict__ g,
                                            const float *__restrict__ dS32, const float *__restrict__ dP32,
                                            int lane = -1, int lanes = 0)
{
    constexpr int n1 = N + 1, nd = 2 * n1 * n1 + 8 * n1;
    if (lane < 0) { lane = threadIdx.x; lanes = blockDim.x; }
    double *dst = L.dS;  // members up to rsP are contiguous doubles: dS, dP, iwS, iwP, lags, row sums
    for (int k = lane; k < nd; k += lanes) {
        const bool w = (k >= 2 * n1 * n1) && (k < 2 * n1 * n1 + 2 * n1);
        const bool r = k >= 2 * n1 * n1 + 6 * n1;
        dst[k] = g[w ? k + 6 * n1 : (r ? k + 2 * n1 : k)];
    }
    if constexpr (F32) {
        for (int k = lane; k < n1 * n1; k += lanes) {
            L.dS32[k] = dS32[k];
            L.dP32[k] = dP32[k];
        }
    }
}

// Nodal constants in two steps around a kernel's element loads: nodal_fetch
// issues this thread's loads of the nodal buffer (at clamped indices, so no
// branch separates them) before the element data, nodal_put writes them to
// LDS after it.  The counter of outstanding loads is in order, so writing
// them then waits for these loads only; staged after the element loads (as
// stage_nodal does) the write waits for every one of them, and no phase can
// start before the last element load has arrived.  LANES threads (the
// workgroup or one wavefront) share the copy, lane = 0 .. LANES-1.
template <int N, bool F32, int LANES>
struct NodalRegs {
    static constexpr int n1 = N + 1, nd = 2 * n1 * n1 + 8 * n1;
    static constexpr int K = (nd + LANES - 1) / LANES;
    static constexpr int KF = F32 ? (n1 * n1 + LANES - 1) / LANES : 1;
    double v[K];
    float s[KF], p[KF];
};

template <int N, bool F32, int LANES>
__device__ __forceinline__ void nodal_fetch(NodalRegs<N, F32, LANES> &R, const double *__restrict__ g,
                                            const float *__restrict__ dS32, const float *__restrict__ dP32, int lane)
{
    using Rg = NodalRegs<N, F32, LANES>;
    constexpr int n1 = Rg::n1, nd = Rg::nd;
    #pragma unroll
    for (int q = 0; q < Rg::K; ++q) {
        const int k = min(lane + q * LANES, nd - 1);
        const bool w = (k >= 2 * n1 * n1) && (k < 2 * n1 * n1 + 2 * n1);
        const bool r = k >= 2 * n1 * n1 + 6 * n1;
        R.v[q] = g[w ? k + 6 * n1 : (r ? k + 2 * n1 : k)];
    }
    if constexpr (F32) {
        #pragma unroll
        for (int q = 0; q < Rg::KF; ++q) {
            const int k = min(lane + q * LANES, n1 * n1 - 1);
            R.s[q] = dS32[k];
            R.p[q] = dP32[k];
        }
    }
}

template <int N, bool F32, int LANES>
__device__ __forceinline__ void nodal_put(NodalLDS<N> &L, const NodalRegs<N, F32, LANES> &R, int lane)
{
    using Rg = NodalRegs<N, F32, LANES>;
    constexpr int n1 = Rg::n1, nd = Rg::nd;
    double *dst = L.dS;  // members up to rsP are contiguous doubles
    #pragma unroll
    for (int q = 0; q < Rg::K; ++q) {
        const int k = lane + q * LANES;
        if (k < nd) dst[k] = R.v[q];
    }
    if constexpr (F32) {
        #pragma unroll
        for (int q = 0; q < Rg::KF; ++q) {
            const int k = lane + q * LANES;
            if (k < n1 * n1) {
                L.dS32[k] = R.s[q];
                L.dP32[k] = R.p[q];
            }
        }
    }
}

// Orders a wavefront's LDS writes before its following LDS reads of other
// lanes' data: the compiler keeps the order (fences, wave barrier) and the
// wavefront waits for its outstanding LDS operations (s_waitcnt lgkmcnt(0),
// gfx9 encoding with vmcnt and expcnt left at their maxima).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_s_waitcnt(0xc07f);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace self
