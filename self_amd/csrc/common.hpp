// common.hpp -- shared runtime pieces of libself_amd: status/error
// reporting and HIP call checking.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "self_amd.h"

namespace self {

// Per-thread last error message (self_last_error).
void set_error(const char *fmt, ...);
const char *last_error();

// Autotuning table (tuning.cpp): size class of a mesh, lookup and record of
// a configuration by (kind, dimension, degree, arithmetic, size class).
int tune_size_class(size_t nElems);
bool tune_lookup(const char *kind, int dim, int N, int dtype, int sizeClass, int cfg[4], double *ms);
void tune_record(const char *kind, int dim, int N, int dtype, int sizeClass, const int cfg[4], double ms);

// Resolves the device address of the no-allocation entry points' scratch
// array on `device` (current); noalloc.hip.
int noalloc_prepare(int device);

inline int hip_fail(hipError_t e, const char *what, const char *file, int line)
{
    set_error("%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    return e == hipErrorOutOfMemory ? SELF_ERR_OOM : SELF_ERR_HIP;
}

}  // namespace self

#define SELF_HIP(call)                                                         \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess) return self::hip_fail(e_, #call, __FILE__, __LINE__); \
    } while (0)

#define SELF_REQUIRE(cond, code, ...)                                          \
    do {                                                                       \
        if (!(cond)) {                                                         \
            self::set_error(__VA_ARGS__);                                      \
            return (code);                                                     \
        }                                                                      \
    } while (0)

#define SELF_CATCH_ALL                                                         \
    catch (const std::bad_alloc &) {                                           \
        self::set_error("host allocation failed");                             \
        return SELF_ERR_OOM;                                                   \
    }                                                                          \
    catch (...) {                                                              \
        self::set_error("unexpected C++ exception");                           \
        return SELF_ERR_OTHER;                                                 \
    }
