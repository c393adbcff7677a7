// m25x_common.h - shared by the HIP sources of libm25x.so: complex helpers, the device views of the tables and
// the mesh, the error state of the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include "../../include/m25x.h"

namespace m25x {

// ---- error state (per host thread) -----------------------------------------------------------------------
void set_error(const char *fmt, ...);
int fail(int code, const char *fmt, ...);
int check_launch(const char *what);
int num_cus();

// ---- launch timing (benchmarks) ---------------------------------------------------------------------------
// While a record is open (m25x_timing with the kernel's bit in its mask), a TimedLaunch around a launch brackets
// it with two HIP events on its stream and notes the launch's algorithmic bytes and flops.
class TimedLaunch {
  public:
    TimedLaunch(hipStream_t s, int kernel, int dtype, double bytes, double flops);
    ~TimedLaunch();
    TimedLaunch(const TimedLaunch &) = delete;
    TimedLaunch &operator=(const TimedLaunch &) = delete;

  private:
    hipStream_t s_;
    long slot_;
};

#define M25X_HIP(call)                                                                                        \
    do {                                                                                                      \
        hipError_t e_ = (call);                                                                               \
        if (e_ != hipSuccess) return ::m25x::fail(M25X_EHIP, "%s: %s", #call, hipGetErrorString(e_));        \
    } while (0)

// ---- complex arithmetic on double2 / float2 --------------------------------------------------------------
__host__ __device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__host__ __device__ inline double2 cdiv(double2 a, double2 b)
{
    // Smith's algorithm (no overflow for the magnitudes met here)
    if (fabs(b.x) >= fabs(b.y)) {
        double r = b.y / b.x, den = b.x + b.y * r;
        return make_double2((a.x + a.y * r) / den, (a.y - a.x * r) / den);
    }
    double r = b.x / b.y, den = b.x * r + b.y;
    return make_double2((a.x * r + a.y) / den, (a.y * r - a.x) / den);
}
__host__ __device__ inline float2 caddf(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline float2 cmulf(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// ---- device views -----------------------------------------------------------------------------------------
struct DTab {
    int ntab, nfine, L;
    const double4 *AQ;                       // (Re A, Im A, Re Q, Im Q) per point
    const double *dr, *d, *dimg;
    const double2 *cA, *cq, *cAi, *cqi;
    const int *tab, *kind;
};

inline DTab dtab(const m25x_tables *t)
{
    DTab d;
    d.ntab = t->ntab; d.nfine = t->nfine; d.L = t->L;
    d.AQ = (const double4 *)t->AQ;
    d.dr = t->dr; d.d = t->d; d.dimg = t->dimg;
    d.cA = (const double2 *)t->cA; d.cq = (const double2 *)t->cq;
    d.cAi = (const double2 *)t->cAi; d.cqi = (const double2 *)t->cqi;
    d.tab = t->tab; d.kind = t->kind;
    return d;
}

// The table lookups and the kernels below follow the reference's operation order; the sources that use them are
// compiled without FP contraction (-ffp-contract=off), so the fill rounds as the reference's CPU build does.

// the smooth rests by linear lookup, u = r / dr clamped at nfine - 1.000001 (mom25d_core.cpp:70-78)
__device__ inline void tab_rest(const DTab &T, int t, double r, double2 &ga, double2 &gq)
{
    double u = r / T.dr[t];
    const double umax = T.nfine - 1.000001;
    if (!(u <= umax)) u = umax;            // (a NaN distance too: the index stays in the table)
    int i = (int)u;
    double f = u - i, g = 1 - f;
    const double4 *p = T.AQ + (size_t)t * T.nfine + i;
    double4 a = p[0], b = p[1];
    ga = make_double2(a.x * g + b.x * f, a.y * g + b.y * f);
    gq = make_double2(a.z * g + b.z * f, a.w * g + b.w * f);
}

// the kernel at a centroid distance with its extracted parts added back (mom25d_core.cpp:80-89)
__device__ inline void tab_kern(const DTab &T, int la, int lb, double r, double2 &ga, double2 &gq)
{
    int t = T.tab[la * T.L + lb];
    tab_rest(T, t, r, ga, gq);
    int kd = T.kind[t];
    if (kd == 0) {
        if (r > 0) {                       // complex / double, as the reference divides (no reciprocal)
            const double2 a = T.cA[t], q = T.cq[t];
            ga = cadd(ga, make_double2(a.x / r, a.y / r));
            gq = cadd(gq, make_double2(q.x / r, q.y / r));
        }
        double di = T.dimg[t];
        if (di > 0) {
            double ri = 1.0 / sqrt(r * r + di * di);
            ga = cadd(ga, cscale(T.cAi[t], ri));
            gq = cadd(gq, cscale(T.cqi[t], ri));
        }
    } else if (kd == 1) {
        double dd = T.d[t], ri = 1.0 / sqrt(r * r + dd * dd);
        ga = cadd(ga, cscale(T.cA[t], ri));
        gq = cadd(gq, cscale(T.cq[t], ri));
    }
}

}  // namespace m25x
