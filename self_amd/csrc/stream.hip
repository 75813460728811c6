// stream.hip -- STREAM-style copy used to measure the attainable HBM
// bandwidth of the device and to calibrate hardware byte counters
// (FETCH_SIZE / WRITE_SIZE of a copy of known size).  Measurement support;
// no counterpart in the reference.

#include "common.hpp"

namespace {

// 16 B per access, four loads in flight per thread before the stores.  The
// loads are unconditional selects (k < n ? a[k] : 0), not loads under an
// if: under an if the compiler waits for each load before the next one
// (one access in flight; 2.7x slower at 24 MiB).
constexpr int kCopyUnroll = 4;

__global__ __launch_bounds__(256) void k_stream_copy(const double2 *__restrict__ a, double2 *__restrict__ b, size_t n)
{
    const size_t stride = (size_t)gridDim.x * 256 * kCopyUnroll;
    for (size_t i = (size_t)blockIdx.x * 256 * kCopyUnroll + threadIdx.x; i < n; i += stride) {
        double2 v[kCopyUnroll];
        #pragma unroll
        for (int u = 0; u < kCopyUnroll; ++u) {
            const size_t k = i + (size_t)u * 256;
            v[u] = k < n ? a[k] : make_double2(0.0, 0.0);
        }
        #pragma unroll
        for (int u = 0; u < kCopyUnroll; ++u) {
            const size_t k = i + (size_t)u * 256;
            if (k < n) b[k] = v[u];
        }
    }
}

}  // namespace

extern "C" {

// Copies `bytes` (rounded down to 16) iters times between two fresh device
// buffers; *gbps = 2*bytes*iters / time (read + write), *ms = mean per copy.
int self_stream_copy(size_t bytes, int iters, double *gbps, double *ms)
{
    SELF_REQUIRE(bytes >= 16 && iters >= 1 && gbps && ms, SELF_ERR_ARG, "self_stream_copy: bad arguments");
    const size_t n = bytes / 16;
    double2 *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&a, n * 16);
    if (e == hipSuccess) e = hipMalloc(&b, n * 16);
    if (e == hipSuccess) e = hipMemset(a, 0, n * 16);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    const int grid = 256 * 16;
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, s, a, b, n);  // warm
        e = hipEventRecord(e0, s);
    }
    for (int k = 0; k < iters && e == hipSuccess; ++k) {
        hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, s, a, b, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (e != hipSuccess) return self::hip_fail(e, "stream copy", __FILE__, __LINE__);
    *ms = t / iters;
    *gbps = 2.0 * (double)(n * 16) * iters / (t * 1e-3) / 1e9;
    return SELF_OK;
}

}  // extern "C"
