// stream.hip -- STREAM-style copy used to measure the attainable HBM
// bandwidth of the device and to calibrate hardware byte counters
// (FETCH_SIZE / WRITE_SIZE of a copy of known size).  Measurement support;
// no counterpart in the reference.

#include "common.hpp"

namespace {

__global__ __launch_bounds__(256) void k_stream_copy(const double2 *__restrict__ a, double2 *__restrict__ b, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
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
    const int grid = 256 * 8 * 4;
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
