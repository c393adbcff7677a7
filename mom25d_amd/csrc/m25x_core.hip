// m25x_core.hip - the C ABI's error state and device set-up.
#include <cstdarg>
#include <mutex>
#include <vector>
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

// ---- launch timing -------------------------------------------------------------------------------------------
namespace {
struct TimingRec {
    int kernel, dtype;
    double bytes, flops;
    hipEvent_t e0, e1;
    bool ended;
};
struct Timing {
    std::mutex mu;
    unsigned mask = 0;
    std::vector<hipEvent_t> pool;     // events kept across records (created once)
    size_t used = 0;
    std::vector<TimingRec> recs;
};
Timing g_tm;

hipEvent_t take_event()
{
    if (g_tm.used == g_tm.pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        g_tm.pool.push_back(e);
    }
    return g_tm.pool[g_tm.used++];
}
}  // namespace

TimedLaunch::TimedLaunch(hipStream_t s, int kernel, int dtype, double bytes, double flops) : s_(s), slot_(-1)
{
    std::lock_guard<std::mutex> lk(g_tm.mu);
    if (!(g_tm.mask & (1u << kernel))) return;
    hipEvent_t e0 = take_event(), e1 = take_event();
    if (!e0 || !e1 || hipEventRecord(e0, s) != hipSuccess) return;
    slot_ = (long)g_tm.recs.size();
    g_tm.recs.push_back(TimingRec{kernel, dtype, bytes, flops, e0, e1, false});
}

TimedLaunch::~TimedLaunch()
{
    if (slot_ < 0) return;
    std::lock_guard<std::mutex> lk(g_tm.mu);
    if (slot_ < (long)g_tm.recs.size() && hipEventRecord(g_tm.recs[slot_].e1, s_) == hipSuccess)
        g_tm.recs[slot_].ended = true;
}

}  // namespace m25x

extern "C" {

int m25x_timing(unsigned kernels)
{
    std::lock_guard<std::mutex> lk(m25x::g_tm.mu);
    if (kernels) {
        m25x::g_tm.used = 0;
        m25x::g_tm.recs.clear();
    }
    m25x::g_tm.mask = kernels;
    return M25X_OK;
}

long m25x_timing_records(int *kernel, int *dtype, float *ms, double *bytes, double *flops, long cap)
{
    std::lock_guard<std::mutex> lk(m25x::g_tm.mu);
    const long n = (long)m25x::g_tm.recs.size();
    for (long i = 0; i < n && i < cap; ++i) {
        const m25x::TimingRec &r = m25x::g_tm.recs[i];
        if (kernel) kernel[i] = r.kernel;
        if (dtype) dtype[i] = r.dtype;
        if (bytes) bytes[i] = r.bytes;
        if (flops) flops[i] = r.flops;
        if (ms) {
            if (!r.ended) return m25x::fail(M25X_EHIP, "timing_records: launch %ld has no end event", i);
            hipError_t e = hipEventSynchronize(r.e1);
            if (e == hipSuccess) e = hipEventElapsedTime(ms + i, r.e0, r.e1);
            if (e != hipSuccess) return m25x::fail(M25X_EHIP, "timing_records: %s", hipGetErrorString(e));
        }
    }
    return n;
}

const char *m25x_kernel_name(int kernel)
{
    static const char *names[] = {"k_fill_far", "k_near_moments", "k_near_scatter", "k_tslu", "k_apply_moves",
                                  "k_l21", "k_trsm", "k_gemm", "k_trsv", "k_zgemm_nk", "m25x_cgs2", "k_panel",
                                  "k_laswp"};
    if (kernel < 0 || kernel >= (int)(sizeof names / sizeof names[0])) return "";
    return names[kernel];
}

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
