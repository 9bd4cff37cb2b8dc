// abi.cpp -- version, thread-local error text and grid validation of libpfm.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include "pfm_host.h"

namespace pfm {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int make_grid(const pfm_grid* g, GridDev* out) {
    PFM_CHECK_ARG(g != nullptr, "grid descriptor is NULL");
    PFM_CHECK_ARG(g->dim == 2 || g->dim == 3, "dim must be 2 or 3 (got %d)", g->dim);
    for (int a = 0; a < g->dim; ++a)
        PFM_CHECK_ARG(g->res[a] >= 4 && g->res[a] <= (1 << 16), "res[%d]=%d out of range [4, 65536]",
                      a, g->res[a]);
    PFM_CHECK_ARG(g->dx > 0.f, "dx must be positive");
    out->dim = g->dim;
    for (int a = 0; a < 3; ++a) out->res[a] = a < g->dim ? g->res[a] : 1;
    out->dx = g->dx;
    out->inv_dx = (float)(1.0 / (double)g->dx);
    for (int a = 0; a < 3; ++a) {
        // interp_grad_2 clamp: float(extent - 2 - eps) computed in double and
        // rounded once, as Python evaluates `u_dim-2-eps` (3D/taichi_utils.py:282).
        out->hi_c[a] = (float)((double)out->res[a] - 2.0 - 1e-6);
        out->hi_f[a] = (float)((double)out->res[a] + 1.0 - 2.0 - 1e-6);
    }
    int64_t cells = (int64_t)out->res[0] * out->res[1] * out->res[2];
    PFM_CHECK_ARG(cells < (int64_t)INT32_MAX - 16, "grid too large for int32 cell indices");
    // face fields are addressed with 32-bit byte offsets
    for (int a = 0; a < out->dim; ++a) {
        int64_t faces = cells / out->res[a] * (out->res[a] + 1);
        PFM_CHECK_ARG(faces * 4 < (int64_t)INT32_MAX, "face field %d exceeds 2 GiB", a);
    }
    return PFM_OK;
}

}  // namespace pfm

extern "C" {

int pfm_abi_version(void) { return PFM_ABI_VERSION; }

const char* pfm_last_error(void) { return pfm::g_err; }

}  // extern "C"
