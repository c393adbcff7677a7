// m25x_core.hip - the C ABI's error state and device set-up.
#include <cstdarg>
#include <mutex>
#include "m25x_common.h"

namespace m25x {

static thread_local char g_err[512];

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

int check_launch(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(M25X_EHIP, "%s: launch failed: %s", what, hipGetErrorString(e));
    return M25X_OK;
}

static std::mutex g_mu;
static int g_cus[64];

int num_cus()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_cus[dev] == 0) {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, dev) != hipSuccess) return 256;
        g_cus[dev] = p.multiProcessorCount;
    }
    return g_cus[dev];
}

}  // namespace m25x

extern "C" {

const char *m25x_last_error(void) { return m25x::g_err; }

int m25x_version(void) { return 1; }

int m25x_init(int device)
{
    M25X_HIP(hipSetDevice(device));
    hipDeviceProp_t p;
    M25X_HIP(hipGetDeviceProperties(&p, device));
    if (std::strncmp(p.gcnArchName, "gfx950", 6) != 0)
        return m25x::fail(M25X_EINVAL, "m25x_init: device %d is %s; libm25x is built for gfx950 (MI355X)", device,
                          p.gcnArchName);
    m25x::num_cus();
    m25x::set_error("");
    return M25X_OK;
}

}  // extern "C"
