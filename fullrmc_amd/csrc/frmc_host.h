// frmc_host.h -- host-side plumbing shared by the C-ABI translation units:
// thread-local error text, HIP error checks, argument checks, grid sizing.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "../../include/frmc.h"

namespace frmc {

void set_error(const char *fmt, ...);
void set_overflow(int64_t v);
bool timing_enabled();
void set_last_kernel_ms(float ms);

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

#define FRMC_CHECK_HIP(expr)                                                                     \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess) {                                                                  \
            ::frmc::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                              __LINE__);                                                         \
            return FRMC_EHIP;                                                                    \
        }                                                                                        \
    } while (0)

#define FRMC_CHECK_LAUNCH() FRMC_CHECK_HIP(hipGetLastError())

#define FRMC_REQUIRE(cond, ...)              \
    do {                                     \
        if (!(cond)) {                       \
            ::frmc::set_error(__VA_ARGS__);  \
            return FRMC_EINVAL;              \
        }                                    \
    } while (0)

inline unsigned blocks_for(int64_t n, int threads) {
    int64_t b = (n + threads - 1) / threads;
    return (unsigned)(b < 1 ? 1 : b);
}

// Smallest non-negative float s with RN(sqrt(s)) >= t, i.e. the exact squared
// norm threshold of the reference's distance comparison against t.
float sq_threshold(float t);

// Stream-ordered scratch allocation (hipMallocAsync from the device's default
// pool).  The pool keeps freed memory for reuse instead of returning it to
// the driver at each synchronisation, so per-call scratch costs no mapping.
hipError_t pool_malloc(void **p, size_t bytes, hipStream_t st);

// Basis / grid helpers (copy a device 3x3 to the host once per call).
struct Basis;
struct Grid;
int load_basis(const float *dbasis, Basis &B, bool &ortho, hipStream_t st);
Grid make_grid(float minD, float maxD, float bin, int32_t H);

}  // namespace frmc
