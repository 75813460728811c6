// runtime.cpp -- device selection and memory queries of the C ABI.

#include "common.hpp"

extern "C" {

int self_init(int device)
{
    int n = 0;
    SELF_HIP(hipGetDeviceCount(&n));
    SELF_REQUIRE(device >= 0 && device < n, SELF_ERR_ARG, "self_init: device %d of %d", device, n);
    SELF_HIP(hipSetDevice(device));
    return self::noalloc_prepare(device);
}

int self_device_count(int *n)
{
    SELF_REQUIRE(n, SELF_ERR_ARG, "self_device_count: null");
    SELF_HIP(hipGetDeviceCount(n));
    return SELF_OK;
}

int self_mem_info(size_t *free_bytes, size_t *total_bytes)
{
    SELF_REQUIRE(free_bytes && total_bytes, SELF_ERR_ARG, "self_mem_info: null");
    SELF_HIP(hipMemGetInfo(free_bytes, total_bytes));
    return SELF_OK;
}

}  // extern "C"
