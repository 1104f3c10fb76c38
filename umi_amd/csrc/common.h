// Shared helpers for the libumi_hip translation units: status/last-error
// plumbing and small integer utilities.  Internal header, not installed.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdarg.h>
#include <stdio.h>
#include "../../include/umi_hip.h"

namespace umi {

// Records a formatted message as the thread's last error and returns `code`.
umi_status set_error(umi_status code, const char *fmt, ...);

// Checks a HIP return code; on failure records "<what>: <hip message>".
inline umi_status hip_check(hipError_t e, const char *what) {
    if (e == hipSuccess) return UMI_OK;
    return set_error(UMI_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

// Checks for a launch error right after a kernel launch.
inline umi_status launch_check(const char *kernel) {
    return hip_check(hipGetLastError(), kernel);
}

inline hipStream_t as_stream(umi_stream s) { return reinterpret_cast<hipStream_t>(s); }

__host__ __device__ constexpr int ceil_div(int a, int b) { return (a + b - 1) / b; }
__host__ __device__ constexpr int round_up(int a, int b) { return ceil_div(a, b) * b; }
__host__ __device__ constexpr int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Number of bits needed to hold values in [0, v].
inline int bits_for(uint32_t v) {
    int b = 0;
    while (b < 32 && (v >> b) != 0) ++b;
    return b;
}

}  // namespace umi

#define UMI_TRY(expr)                        \
    do {                                     \
        umi_status _st = (expr);             \
        if (_st != UMI_OK) return _st;       \
    } while (0)
