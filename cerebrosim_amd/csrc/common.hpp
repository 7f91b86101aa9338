// Error reporting and launch helpers shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "cerebro_lbm.h"

namespace cs {

void set_error(const char* fmt, ...);

// Record a failing HIP call and return CS_E_HIP from the enclosing C-ABI function.
#define CS_HIP(call)                                                                             \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess) {                                                                  \
            ::cs::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
            return CS_E_HIP;                                                                     \
        }                                                                                        \
    } while (0)

#define CS_REQUIRE(cond, ...)               \
    do {                                    \
        if (!(cond)) {                      \
            ::cs::set_error(__VA_ARGS__);   \
            return CS_E_ARG;                \
        }                                   \
    } while (0)

// After a launch: report configuration errors (the launch itself is asynchronous).
#define CS_LAUNCHED(name)                                                                          \
    do {                                                                                           \
        hipError_t e_ = hipGetLastError();                                                         \
        if (e_ != hipSuccess) {                                                                    \
            ::cs::set_error("launch of %s failed: %s", name, hipGetErrorString(e_));               \
            return CS_E_HIP;                                                                       \
        }                                                                                          \
    } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace cs
