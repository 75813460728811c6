// dg3d_device.hpp -- pieces shared by the 3-D DG kernels (hexahedral
// elements, Gauss nodes, i fastest within an element, faces numbered as
// src/common/ModelFlags.f90:36-41 minus one: south 0 = p-, east 1 = s+,
// north 2 = p+, west 3 = s-, bottom 4 = q-, top 5 = q+).
#pragma once

#include <hip/hip_runtime.h>

namespace self {

// Side descriptor (int2) of one element side: x = neighbour element (>= 0)
// or boundary flag (< 0); y: bits 0-2 neighbour side, 3 primary, 4 swap,
// 5 i reversed, 6 j reversed, bits 8.. boundary slot.
__host__ __device__ __forceinline__ int s3_nbr_side(int y) { return y & 7; }
__host__ __device__ __forceinline__ bool s3_primary(int y) { return (y >> 3) & 1; }
__host__ __device__ __forceinline__ int s3_slot(int y) { return y >> 8; }

// node of the neighbour's face matching own face node (a, b): the forward
// map (i,j) -> (k,l) of the reference's face loop when this side is
// primary, its inverse otherwise
__host__ __device__ __forceinline__ void s3_map_n(int N, int y, int a, int b, int &na, int &nb)
{
    const int sw = (y >> 4) & 1, fi = (y >> 5) & 1, fj = (y >> 6) & 1;
    if (s3_primary(y)) {
        const int ii = fi ? N - a : a, jj = fj ? N - b : b;
        na = sw ? (fj ? N - b : b) : ii;
        nb = sw ? (fi ? N - a : a) : jj;
    } else if (!sw) {
        na = fi ? N - a : a;
        nb = fj ? N - b : b;
    } else {
        na = fi ? N - b : b;
        nb = fj ? N - a : a;
    }
}

template <int N>
__device__ __forceinline__ void s3_map(int y, int a, int b, int &na, int &nb)
{
    s3_map_n(N, y, a, b, na, nb);
}

// value at side s, face node (a, b) of one element's nodal field q (i
// fastest) through the interpolants to -1 (lL) and +1 (lR)
template <int N>
__device__ __forceinline__ double face_trace3(const double *lL, const double *lR, const double *q, int s, int a,
                                              int b)
{
    constexpr int n1 = N + 1;
    const double *l = (s == 1 || s == 2 || s == 5) ? lR : lL;
    int off, stride;
    if (s == 1 || s == 3) { off = n1 * (a + n1 * b); stride = 1; }          // east / west: (p, q)
    else if (s == 0 || s == 2) { off = a + n1 * n1 * b; stride = n1; }      // south / north: (s, q)
    else { off = a + n1 * b; stride = n1 * n1; }                            // bottom / top: (s, p)
    double acc = l[0] * q[off];
    #pragma unroll
    for (int m = 1; m < n1; ++m) acc = __builtin_fma(l[m], q[off + m * stride], acc);
    return acc;
}

}  // namespace self
