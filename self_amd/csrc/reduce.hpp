// reduce.hpp -- deterministic device reductions for the Krylov solvers.
//
// A reduction kernel runs on a fixed grid; each workgroup writes one partial
// sum, and the last workgroup to finish (atomic ticket) adds the partials
// in block order and publishes the result.  For a given vector length the
// summation tree is fixed, so a solve gives bitwise the same scalars
// whatever else runs on the device (batch invariance), and there is no
// second "finalize" launch.
#pragma once

#include <hip/hip_runtime.h>

namespace self {

constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 512;  // fixed partial count

struct RedSlot {
    double *partials;    // [kRedBlocks * 4] (up to 4 sums per reduction)
    unsigned *ticket;    // one counter, reset to 0 by the finishing block
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

// Publish this block's partials (K sums) and return true in the block that
// finishes last; that block then sees every partial.
template <int K>
__device__ __forceinline__ bool publish_partials(const RedSlot &slot, const double (&part)[K])
{
    __shared__ bool last;
    if (threadIdx.x == 0) {
        #pragma unroll
        for (int k = 0; k < K; ++k) slot.partials[k * kRedBlocks + blockIdx.x] = part[k];
        __threadfence();
        const unsigned t = atomicAdd(slot.ticket, 1u);
        last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (last) __threadfence();
    return last;
}

// In the last block: fixed-order sum of partial k over gridDim.x blocks.
__device__ __forceinline__ double final_sum(const RedSlot &slot, int k, double *lds4)
{
    double v = 0.0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x)
        v += __hip_atomic_load(&slot.partials[k * kRedBlocks + b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return block_sum(v, lds4);
}

}  // namespace self
