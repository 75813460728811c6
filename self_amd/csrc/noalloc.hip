// noalloc.hip -- entry points that allocate no device memory.
//
// self_coarse_to_fine_2d_noalloc is self_coarse_to_fine_2d (CoarseToFine_2D,
// src/interp/Lagrange_2D_Class.f90:488-528, batched over elements) with its
// device staging -- the two interpolation matrices, the input and the output
// fields -- in a file-scope device array instead of buffers from hipMalloc,
// for callers that cannot allocate (code under HIP graph capture, tight
// output loops of a serving process).  The array holds kScratchDoubles
// doubles (16 MiB) on every device the library runs on, until the process
// ends; larger calls are rejected.  Its device address is resolved once per
// device (self_init, or the first call on that device) with
// hipGetSymbolAddress, and only that address is used on the host side.
//
// Calls on one device share the array.  They are serialized by a per-device
// host mutex inside this process; the copies are synchronous, so a call has
// finished with the array when it returns.  Nothing in the array needs to
// start at zero: every call overwrites the parts it reads.

#include <mutex>

#include "common.hpp"

namespace self {

constexpr size_t kScratchDoubles = size_t(1) << 21;  // 16 MiB
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

std::mutex g_resolve_mu;
double *g_scratch[kMaxDevices] = {};
std::mutex g_call_mu[kMaxDevices];

// offsets inside the array stay 256-byte aligned
inline size_t round32(size_t n) { return (n + 31) / 32 * 32; }

}  // namespace

// Resolves the array's address on `device`, which must be current.
int noalloc_prepare(int device)
{
    SELF_REQUIRE(device >= 0 && device < kMaxDevices, SELF_ERR_ARG, "no-alloc scratch: device %d outside 0..%d",
                 device, kMaxDevices - 1);
    std::lock_guard<std::mutex> g(g_resolve_mu);
    if (g_scratch[device]) return SELF_OK;
    void *p = nullptr;
    SELF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    SELF_REQUIRE(p != nullptr, SELF_ERR_HIP, "no-alloc scratch: null device address on device %d", device);
    g_scratch[device] = static_cast<double *>(p);
    return SELF_OK;
}

}  // namespace self

using namespace self;

extern "C" {

int self_coarse_to_fine_2d_noalloc(int nS, int nP, size_t nElems, const double *Ts, const double *Tp, int nNewS,
                                   int nNewP, const double *fOld, double *fNew)
try {
    SELF_REQUIRE(nS >= 0 && nP >= 0 && nNewS >= 0 && nNewP >= 0 && Ts && Tp && fOld && fNew, SELF_ERR_ARG,
                 "coarse_to_fine_2d_noalloc: bad arguments");
    SELF_REQUIRE(nS < 1024 && nP < 1024 && nNewS < 1024 && nNewP < 1024, SELF_ERR_ARG,
                 "coarse_to_fine_2d_noalloc: degrees above 1023");
    const size_t nTs = (size_t)(nNewS + 1) * (nS + 1), nTp = (size_t)(nNewP + 1) * (nP + 1);
    const size_t perIn = (size_t)(nS + 1) * (nP + 1), perOut = (size_t)(nNewS + 1) * (nNewP + 1);
    const size_t fixed = round32(nTs) + round32(nTp) + 64;
    SELF_REQUIRE(fixed <= kScratchDoubles && nElems <= (kScratchDoubles - fixed) / (perIn + perOut), SELF_ERR_ARG,
                 "coarse_to_fine_2d_noalloc: %zu elements of %zu + %zu values need more than the %zu-double "
                 "(16 MiB) scratch array; use self_coarse_to_fine_2d",
                 nElems, perIn, perOut, kScratchDoubles);
    if (!nElems) return SELF_OK;
    const size_t ci = nElems * perIn, co = nElems * perOut;
    int dev = 0;
    SELF_HIP(hipGetDevice(&dev));
    if (int rc = noalloc_prepare(dev)) return rc;
    std::lock_guard<std::mutex> lk(g_call_mu[dev]);
    double *a = g_scratch[dev];
    double *b = a + round32(nTs);
    double *df = b + round32(nTp);
    double *o = df + round32(ci);
    SELF_HIP(hipMemcpy(a, Ts, sizeof(double) * nTs, hipMemcpyHostToDevice));
    SELF_HIP(hipMemcpy(b, Tp, sizeof(double) * nTp, hipMemcpyHostToDevice));
    SELF_HIP(hipMemcpy(df, fOld, sizeof(double) * ci, hipMemcpyHostToDevice));
    int rc = self_coarse_to_fine_2d_dev(nS, nP, nElems, a, b, nNewS, nNewP, df, o, nullptr);
    if (rc) return rc;
    SELF_HIP(hipMemcpy(fNew, o, sizeof(double) * co, hipMemcpyDeviceToHost));
    return SELF_OK;
}
SELF_CATCH_ALL

}  // extern "C"
