// reduce.hpp -- deterministic device reductions for the Krylov solvers.
//
// A reduction runs as two launches on one stream: the partial kernel
// (kRedBlocks workgroups) writes one partial sum per workgroup and row,
// then the same kernel compiled as its final form runs as a single
// workgroup that adds the partials in a fixed tree and applies the
// epilogue (scalar bookkeeping of the solver state).  The kernel boundary
// orders the two, so no workgroup needs a device-scope fence or ticket: on
// a multi-XCD part such a fence writes back the XCD's L2, and one per
// workgroup costs more than the reduction.  For a given vector length the
// summation order is fixed, so solves are bitwise reproducible whatever
// else runs on the device.
#pragma once

#include <hip/hip_runtime.h>

namespace self {

constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 2048;  // partial sums per row (8 workgroups per CU)

struct RedSlot {
    double *partials;  // [rows][kRedBlocks]
};

// sum over the 64 lanes of a wavefront
__device__ __forceinline__ double wave_sum(double v)
{
    #pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// sum over a 256-thread block (fixed tree), result valid in thread 0
__device__ __forceinline__ double block_sum(double v, double *lds4)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) lds4[wid] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) r = (lds4[0] + lds4[1]) + (lds4[2] + lds4[3]);
    __syncthreads();
    return r;
}

// partial kernel: store this workgroup's sum of row k at entry idx (every
// thread calls); store_partial: at the workgroup's index
__device__ __forceinline__ void store_partial_at(const RedSlot &slot, int k, double acc, double *lds4, int idx)
{
    const double p = block_sum(acc, lds4);
    if (threadIdx.x == 0) slot.partials[(size_t)k * kRedBlocks + idx] = p;
}

__device__ __forceinline__ void store_partial(const RedSlot &slot, int k, double acc, double *lds4)
{
    store_partial_at(slot, k, acc, lds4, blockIdx.x);
}

// partial kernel with TPB-thread workgroups: rows k0 and k1 of this
// workgroup (wavefront sums, then wavefronts in index order) at entry idx
template <int TPB>
__device__ __forceinline__ void store_partials2_at(const RedSlot &slot, int k0, double a0, int k1, double a1, int idx)
{
    constexpr int W = TPB / 64;
    __shared__ double l0[W], l1[W];
    a0 = wave_sum(a0);
    a1 = wave_sum(a1);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        l0[wid] = a0;
        l1[wid] = a1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double p0 = 0.0, p1 = 0.0;
        #pragma unroll
        for (int q = 0; q < W; ++q) {
            p0 += l0[q];
            p1 += l1[q];
        }
        slot.partials[(size_t)k0 * kRedBlocks + idx] = p0;
        slot.partials[(size_t)k1 * kRedBlocks + idx] = p1;
    }
}

template <int TPB>
__device__ __forceinline__ void store_partials2(const RedSlot &slot, int k0, double a0, int k1, double a1)
{
    store_partials2_at<TPB>(slot, k0, a0, k1, a1, blockIdx.x);
}

template <int TPB>
__device__ __forceinline__ void store_partial1(const RedSlot &slot, int k, double a)
{
    constexpr int W = TPB / 64;
    __shared__ double l0[W];
    a = wave_sum(a);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) l0[wid] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        double p = 0.0;
        #pragma unroll
        for (int q = 0; q < W; ++q) p += l0[q];
        slot.partials[(size_t)k * kRedBlocks + blockIdx.x] = p;
    }
}

// fixed-order sum of the first `count` entries of a partial row by a whole
// TPB-thread workgroup, valid in every thread (every workgroup that calls
// it gets the same bits)
template <int TPB>
__device__ __forceinline__ double block_fold_sum(const double *row, int count)
{
    constexpr int W = TPB / 64;
    __shared__ double lw[W];
    double v = 0.0;
    for (int b = threadIdx.x; b < count; b += TPB) v += row[b];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) lw[threadIdx.x >> 6] = v;
    __syncthreads();
    double sum = 0.0;
    #pragma unroll
    for (int q = 0; q < W; ++q) sum += lw[q];
    return sum;
}

// final kernel (one workgroup): fixed-order sum of row k, valid in thread 0
__device__ __forceinline__ double final_sum(const RedSlot &slot, int k, double *lds4)
{
    double v = 0.0;
    for (int b = threadIdx.x; b < kRedBlocks; b += blockDim.x) v += slot.partials[(size_t)k * kRedBlocks + b];
    return block_sum(v, lds4);
}

// the two launches of a reduction kernel KERNEL<..., FIN>: partial then final
#define SELF_RED_LAUNCH(KPART, KFIN, stream, ...)                                                   \
    do {                                                                                            \
        hipLaunchKernelGGL(KPART, dim3(::self::kRedBlocks), dim3(::self::kRedThreads), 0, stream, __VA_ARGS__); \
        hipLaunchKernelGGL(KFIN, dim3(1), dim3(::self::kRedThreads), 0, stream, __VA_ARGS__);      \
    } while (0)

}  // namespace self
