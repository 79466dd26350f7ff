// Error reporting of libfemgpu: a per-thread message buffer behind femgpu_last_error(). Host-only,
// so the host sources (structure, mesh, Delaunay, spectral) link without any device code.
#include <cstdarg>
#include <cstdio>

#include "common.h"

namespace fg {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int hip_fail(hipError_t e, const char *what, const char *file, int line)
{
    set_error("%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    return FEMGPU_ERR_HIP;
}

}  // namespace fg

extern "C" const char *femgpu_last_error(void) { return fg::g_err; }
extern "C" const char *femgpu_version(void) { return "femgpu 0.1 (gfx950)"; }
