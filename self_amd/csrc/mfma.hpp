// mfma.hpp -- matrix-core (MFMA) form of the element-local 1-D contractions
// of the 3-D DG kernels on gfx950: for one direction, the weak derivative
// matrix D (n1 x n1, n1 <= 8) applied along that direction to every line of
// an element's nodal flux (n1^2 lines), i.e. one 8 x 8 by 8 x 64 product
// per element, direction and variable.
//
// One wavefront issues 16x16 output tiles holding two independent 8 x 16
// products: A = diag(D8, D8) (16 x 16, D8 = D zero-padded to 8 x 8), B = the
// two 8 x 16 column chunks stacked (16 x 16), C = [D8 B0; D8 B1].  The FP32
// form runs four v_mfma_f32_16x16x4_f32 (K = 4 each); the split-BF16 form
// writes every operand as hi + lo (hi = bf16(x), lo = bf16(x - hi)) and
// accumulates A_hi B_hi + A_hi B_lo + A_lo B_hi with three
// v_mfma_f32_16x16x16_bf16 (K = 16), about 16 significant bits per product,
// FP32 accumulation.  The line values enter as deviations from the line's
// first node (taken in FP64 before rounding), so what the narrow arithmetic
// sees is the variation along the line; the first node's value times the
// row sum of D is added back in FP64 by the caller.
//
// Lane layouts (CDNA3/CDNA4 ISA, 16x16 tiles, lane l, g = l / 16):
//   16x16x4 f32:   A[l % 16][4 s + g], B[4 s + g][l % 16] at k-step s
//   16x16x16 bf16: A[l % 16][4 g + q], B[4 g + q][l % 16], q = 0..3
//   C (both):      C[4 g + r][l % 16], r = 0..3
#pragma once

#include <hip/hip_runtime.h>

namespace self {

typedef float mf_v4f __attribute__((ext_vector_type(4)));
typedef short mf_v4s __attribute__((ext_vector_type(4)));

// round to nearest even bfloat16 (finite inputs)
__device__ __forceinline__ unsigned short mf_bf16(float x)
{
    unsigned u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float mf_bf16_to_float(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

__device__ __forceinline__ void mf_split(float x, short &hi, short &lo)
{
    const unsigned short h = mf_bf16(x);
    hi = (short)h;
    lo = (short)mf_bf16(x - mf_bf16_to_float(h));
}

// Value of the B operand: line `col` (of nLines), node m along the line, of
// direction dir's nodal array X (i fastest, n1 per direction), as the
// deviation from the line's first node; zero outside the element.
template <int N>
__device__ __forceinline__ float mf_line_dev(const double *X, int dir, int m, int col)
{
    constexpr int n1 = N + 1, n2 = n1 * n1;
    if (m >= n1 || col >= n2) return 0.0f;
    int at, at0;
    if (dir == 0) { at = m + n1 * col; at0 = n1 * col; }
    else if (dir == 1) { const int i = col % n1, k = col / n1; at = i + n1 * (m + n1 * k); at0 = i + n2 * k; }
    else { at = col + n2 * m; at0 = col; }
    return (float)(X[at] - X[at0]);
}

// One wavefront: tile t2 (0 or 1) of direction dir -- columns
// [32 t2, 32 t2 + 32) -- of T = D8 * dev(X), written to T as T[i + n1 col]
// (float; row i along the line, col the line).  Df: D8 row-major (i, m) as
// float.  PREC 1: FP32, 2: split BF16.
template <int N, int PREC>
__device__ __forceinline__ void mf_tile(const float *Df, const double *X, int dir, int t2, float *T)
{
    constexpr int n1 = N + 1, n2 = n1 * n1;
    const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15, ra = lane & 15;
    mf_v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
    if (PREC == 1) {
        #pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int ka = 4 * s + g;
            const float a = ((ra >> 3) == (ka >> 3)) ? Df[(ra & 7) * 8 + (ka & 7)] : 0.0f;
            const float b = mf_line_dev<N>(X, dir, ka & 7, 32 * t2 + 16 * (ka >> 3) + c);
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
        }
    } else {
        mf_v4s ahi, alo, bhi, blo;
        #pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int ka = 4 * g + q;
            const float a = ((ra >> 3) == (ka >> 3)) ? Df[(ra & 7) * 8 + (ka & 7)] : 0.0f;
            const float b = mf_line_dev<N>(X, dir, ka & 7, 32 * t2 + 16 * (ka >> 3) + c);
            short h, l;
            mf_split(a, h, l);
            ahi[q] = h;
            alo[q] = l;
            mf_split(b, h, l);
            bhi[q] = h;
            blo[q] = l;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ahi, bhi, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ahi, blo, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(alo, bhi, acc, 0, 0, 0);
    }
    #pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int row = 4 * g + r, i = row & 7, col = 32 * t2 + 16 * (row >> 3) + c;
        if (i < n1 && col < n2) T[i + n1 * col] = acc[r];
    }
}

}  // namespace self
