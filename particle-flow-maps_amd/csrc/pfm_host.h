// pfm_host.h -- host-side helpers of libpfm: error reporting, argument
// validation and launch geometry.
#pragma once

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include "../../include/pfm.h"
#include "pfm_device.h"

namespace pfm {

void set_error(const char* fmt, ...);

#define PFM_HIP_TRY(expr)                                                            \
    do {                                                                             \
        hipError_t _e = (expr);                                                      \
        if (_e != hipSuccess) {                                                      \
            ::pfm::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                             __FILE__, __LINE__);                                    \
            return PFM_ERR_HIP;                                                      \
        }                                                                            \
    } while (0)

#define PFM_CHECK_ARG(cond, ...)        \
    do {                                \
        if (!(cond)) {                  \
            ::pfm::set_error(__VA_ARGS__); \
            return PFM_ERR_ARG;         \
        }                               \
    } while (0)

// Catches launch-configuration errors of the kernel just launched.
#define PFM_LAUNCH_CHECK() PFM_HIP_TRY(hipGetLastError())

inline hipStream_t as_stream(pfm_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned grid_for(int64_t n, int block = kBlock) {
    int64_t b = (n + block - 1) / block;
    if (b < 1) b = 1;
    return (unsigned)b;
}

// Validates a pfm_grid and fills the device descriptor.
int make_grid(const pfm_grid* g, GridDev* out);

inline int64_t cell_count(const GridDev& g) {
    return (int64_t)g.res[0] * g.res[1] * g.res[2];
}

inline int64_t face_count_h(const GridDev& g, int a) {
    int64_t n = 1;
    for (int c = 0; c < g.dim; ++c) n *= (int64_t)(g.res[c] + (c == a ? 1 : 0));
    return n;
}

}  // namespace pfm
