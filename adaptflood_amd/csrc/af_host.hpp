// af_host.hpp -- host-side helpers shared by the C ABI translation units:
// per-thread error text and HIP status checks.
#pragma once
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>

#include "../../include/adaptflood.h"

namespace af {

void set_error(const char* fmt, ...);

#define AF_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t af_e_ = (call);                                                     \
        if (af_e_ != hipSuccess) {                                                     \
            ::af::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(af_e_),  \
                            __FILE__, __LINE__);                                       \
            return AF_E_HIP;                                                           \
        }                                                                              \
    } while (0)

#define AF_REQUIRE(cond, code, ...)      \
    do {                                 \
        if (!(cond)) {                   \
            ::af::set_error(__VA_ARGS__); \
            return (code);               \
        }                                \
    } while (0)

inline hipStream_t to_stream(af_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned blocks_for(long long n, int threads) {
    long long b = (n + threads - 1) / threads;
    if (b < 1) b = 1;
    if (b > 1048576) b = 1048576;  // grid-stride loops cover the rest
    return (unsigned)b;
}

}  // namespace af
