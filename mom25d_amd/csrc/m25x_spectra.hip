// m25x_spectra.hip - the layered stack's spectral Green's functions on a Sommerfeld contour: every spectrum that a
// frequency's Green's tables and via kernels need, in one launch (one thread per contour node and request), and
// the same functions at single points on the host (the extraction coefficients).
//
// The stack along z is a transmission-line network, one line section per dielectric layer j with
//   kz_j = -j sqrt(kr^2 - k0^2 eps_j)          (Im kz_j <= 0 on the contour)
//   Y_j  = kz_j / (w mu0)  (TE)   or   w eps0 eps_j / kz_j  (TM)
//   E_j  = exp(-2j kz_j h_j),  P_j = exp(-j kz_j h_j)   (|E_j|, |P_j| <= 1: nothing below overflows).
// A load Y_L seen through a section becomes Y_j (1 - G E_j) / (1 + G E_j) with G = (Y_j - Y_L) / (Y_j + Y_L) its
// reflection coefficient (a short: G = -1); the admittances seen looking down (Yd_k) and up (Yu_k) from every
// interface follow by stepping from the ground / cover or the half-spaces. A unit shunt current at interface s
// gives V_s = 1 / (Yu_s + Yd_s), and the voltage steps across a section toward a load of reflection G by
// P_j (1 + G) / (1 + G E_j). From the TE (h) and TM (e) voltages V(o|s):
//   G~_A(o, s) = V_h / (jw),  G~_q(o, s) = jw (V_e - V_h) / kr^2                          (tables)
//   D~(k, span) = sum_{j in span} dV_e(j | k) / kz2_j - sum_{ends e} div_e G~_q(k, e) / (jw)   (via -> interface)
//   R~(m, n) = sum_{i in m} sum_{j in n} kr^2 / (kz2_i kz2_j) (dV_e(i | j) - dV_e(i | j - 1))
//              - sum_{ends a of m, b of n} div_a div_b G~_q(a, b) / (jw)                   (via -> via)
// with kz2_j = k0^2 eps_j - kr^2, dV(i | s) = V(i | s) - V(i - 1 | s), zero at a PEC interface and for a PEC
// source. The tables' extracted parts c / sqrt(rho^2 + d^2) come off as 2 pi c exp(-kr d) / kr. Each output is
// written multiplied by kr dkr: the weights of the Hankel transform (m25x_hankel) that follows.
#include "m25x_common.h"

using namespace m25x;

namespace {

constexpr int ML = M25X_MAX_LAYERS;

__host__ __device__ inline double2 cneg(double2 a) { return make_double2(-a.x, -a.y); }

// principal square root (branch cut on the negative real axis, the sign of Im z kept there)
__host__ __device__ inline double2 csqrt_p(double2 z)
{
    if (z.x == 0.0 && z.y == 0.0) return make_double2(0.0, z.y);
    const double r = hypot(z.x, z.y);
    if (z.x >= 0.0) {
        const double t = sqrt(0.5 * (r + z.x));
        return make_double2(t, z.y / (2.0 * t));
    }
    const double t = sqrt(0.5 * (r - z.x));
    return make_double2(fabs(z.y) / (2.0 * t), copysign(t, z.y));
}

__host__ __device__ inline double2 cexp_(double2 z)
{
    const double m = exp(z.x);
    return make_double2(m * cos(z.y), m * sin(z.y));
}

// -j z
__host__ __device__ inline double2 mulnj(double2 z) { return make_double2(z.y, -z.x); }

// The network of one contour node: per layer kz, kz2, E, P and the line admittances of both modes; per interface
// the admittances looking down and up.
struct Net {
    int N, bot_pec, top_pec;
    double2 kz2[ML + 1], E[ML + 1], P[ML + 1];
    double2 Y[2][ML + 1], Yd[2][ML + 1], Yu[2][ML + 1];
};

// through a section of admittance y and two-way transfer e, a load of admittance yl
__host__ __device__ inline double2 seen_through(double2 y, double2 e, double2 yl)
{
    const double2 g = cmul(cdiv(csub(y, yl), cadd(y, yl)), e);
    return cmul(y, cdiv(make_double2(1.0 - g.x, -g.y), make_double2(1.0 + g.x, g.y)));
}

// a section shorted at its far end: y (1 + e) / (1 - e)
__host__ __device__ inline double2 shorted(double2 y, double2 e)
{
    return cmul(y, cdiv(make_double2(1.0 + e.x, e.y), make_double2(1.0 - e.x, -e.y)));
}

__host__ __device__ inline double2 half_space_y(const m25x_stack &st, double2 kr2, double2 eps, int mode)
{
    const double2 kz = mulnj(csqrt_p(csub(kr2, cscale(eps, st.k0 * st.k0))));
    return mode == 0 ? cscale(kz, 1.0 / st.wmu) : cdiv(cscale(eps, st.weps), kz);
}

__host__ __device__ void net_build(const m25x_stack &st, double2 kr, Net &n)
{
    const int N = st.N;
    n.N = N; n.bot_pec = st.bottom_pec; n.top_pec = st.top_pec;
    const double2 kr2 = cmul(kr, kr);
    for (int j = 1; j <= N; ++j) {
        const double2 eps = make_double2(st.eps[j - 1][0], st.eps[j - 1][1]);
        const double2 k2 = cscale(eps, st.k0 * st.k0);
        n.kz2[j] = csub(k2, kr2);
        const double2 kz = mulnj(csqrt_p(csub(kr2, k2)));
        const double2 th = cscale(kz, st.h[j - 1]);
        n.E[j] = cexp_(make_double2(2.0 * th.y, -2.0 * th.x));     // exp(-2j th)
        n.P[j] = cexp_(make_double2(th.y, -th.x));                 // exp(-j th)
        n.Y[0][j] = cscale(kz, 1.0 / st.wmu);                      // TE
        n.Y[1][j] = cdiv(cscale(eps, st.weps), kz);                // TM
    }
    for (int m = 0; m < 2; ++m) {
        int k0;
        if (st.bottom_pec) {
            n.Yd[m][1] = shorted(n.Y[m][1], n.E[1]);
            k0 = 2;
        } else {
            n.Yd[m][0] = half_space_y(st, kr2, make_double2(st.eps_bot[0], st.eps_bot[1]), m);
            k0 = 1;
        }
        for (int k = k0; k <= N; ++k) n.Yd[m][k] = seen_through(n.Y[m][k], n.E[k], n.Yd[m][k - 1]);
        int kt;
        if (st.top_pec) {
            n.Yu[m][N - 1] = shorted(n.Y[m][N], n.E[N]);
            kt = N - 1;
        } else {
            n.Yu[m][N] = half_space_y(st, kr2, make_double2(st.eps_top[0], st.eps_top[1]), m);
            kt = N;
        }
        const int kend = st.bottom_pec ? 2 : 1;
        for (int k = kt; k >= kend; --k) n.Yu[m][k - 1] = seen_through(n.Y[m][k], n.E[k], n.Yu[m][k]);
    }
}

__host__ __device__ inline bool is_pec(const Net &n, int k)
{
    return (k == 0 && n.bot_pec) || (k == n.N && n.top_pec);
}

// V(o | s) of mode m (0 TE, 1 TM): the voltage at interface o for a unit shunt current at interface s
__host__ __device__ double2 volt(const Net &n, int m, int o, int s)
{
    if (is_pec(n, o) || is_pec(n, s)) return make_double2(0.0, 0.0);
    double2 v = cdiv(make_double2(1.0, 0.0), cadd(n.Yu[m][s], n.Yd[m][s]));
    for (int k = s + 1; k <= o; ++k) {           // up through layer k toward the load above it
        const double2 g = cdiv(csub(n.Y[m][k], n.Yu[m][k]), cadd(n.Y[m][k], n.Yu[m][k]));
        v = cmul(v, cmul(n.P[k], cdiv(make_double2(1.0 + g.x, g.y), cadd(make_double2(1.0, 0.0), cmul(g, n.E[k])))));
    }
    for (int k = s; k > o; --k) {                // down through layer k toward the load below it
        const double2 g = cdiv(csub(n.Y[m][k], n.Yd[m][k - 1]), cadd(n.Y[m][k], n.Yd[m][k - 1]));
        v = cmul(v, cmul(n.P[k], cdiv(make_double2(1.0 + g.x, g.y), cadd(make_double2(1.0, 0.0), cmul(g, n.E[k])))));
    }
    return v;
}

// G~_A, G~_q between interfaces o and s
__host__ __device__ void spectral(const Net &n, double w, double2 kr, int o, int s, double2 &ga, double2 &gq)
{
    const double2 vh = volt(n, 0, o, s), ve = volt(n, 1, o, s);
    ga = cdiv(vh, make_double2(0.0, w));
    gq = cdiv(cmul(make_double2(0.0, w), csub(ve, vh)), cmul(kr, kr));
}

__host__ __device__ inline double2 gq_of(const Net &n, double w, double2 kr, int a, int b)
{
    if (is_pec(n, a) || is_pec(n, b)) return make_double2(0.0, 0.0);
    double2 ga, gq;
    spectral(n, w, kr, a < b ? a : b, a < b ? b : a, ga, gq);
    return gq;
}

// dV_e(i | s) = V_e(i | s) - V_e(i - 1 | s), zero for a PEC source
__host__ __device__ inline double2 dve(const Net &n, int i, int s)
{
    if (is_pec(n, s)) return make_double2(0.0, 0.0);
    return csub(volt(n, 1, i, s), volt(n, 1, i - 1, s));
}

// one request at one node: out[0] (and out[1] for a table)
__host__ __device__ void request_value(const Net &n, double w, double2 kr, const int *ri, const double *rd,
                                       double2 *out)
{
    const int type = ri[0];
    const double2 jw = make_double2(0.0, w);
    if (type == M25X_SPEC_TABLE) {
        const int o = ri[2], s = ri[3], nx = ri[4];
        double2 ga, gq;
        spectral(n, w, kr, o, s, ga, gq);
        for (int x = 0; x < nx; ++x) {              // c / sqrt(rho^2 + d^2)  <->  2 pi c exp(-kr d) / kr
            const double d = rd[5 * x];
            const double2 cA = make_double2(rd[5 * x + 1], rd[5 * x + 2]), cq = make_double2(rd[5 * x + 3], rd[5 * x + 4]);
            const double2 e = cdiv(cscale(cexp_(cscale(cneg(kr), d)), 2.0 * M_PI), kr);
            ga = csub(ga, cmul(cA, e));
            gq = csub(gq, cmul(cq, e));
        }
        out[0] = ga; out[1] = gq;
    } else if (type == M25X_SPEC_VIA_D) {
        const int k = ri[2], lo = ri[3], hi = ri[4];
        double2 v = {0.0, 0.0};
        for (int j = lo + 1; j <= hi; ++j) v = cadd(v, cdiv(dve(n, j, k), n.kz2[j]));
        v = csub(v, cdiv(gq_of(n, w, kr, k, lo), jw));        // the bottom end (divergence +1)
        v = cadd(v, cdiv(gq_of(n, w, kr, k, hi), jw));        // the top end (-1)
        out[0] = v;
    } else {                                                  // M25X_SPEC_VIA_R
        const int mlo = ri[2], mhi = ri[3], nlo = ri[4], nhi = ri[5];
        const double2 kr2 = cmul(kr, kr);
        double2 v = {0.0, 0.0};
        for (int i = mlo + 1; i <= mhi; ++i)
            for (int j = nlo + 1; j <= nhi; ++j)
                v = cadd(v, cmul(cdiv(kr2, cmul(n.kz2[i], n.kz2[j])), csub(dve(n, i, j), dve(n, i, j - 1))));
        const int ea[2] = {mlo, mhi}, eb[2] = {nlo, nhi};
        const double da[2] = {1.0, -1.0};
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                v = csub(v, cscale(cdiv(gq_of(n, w, kr, ea[a], eb[b]), jw), da[a] * da[b]));
        out[0] = v;
    }
}

// W[row + r, node] = value r of the request at the node, times kr dkr
__global__ __launch_bounds__(128) void k_spectra(m25x_stack st, double w, const double2 *kr, const double2 *dk, int npts,
                                                 const int *req_i, const double *req_d, double2 *W)
{
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= npts) return;
    const int r = blockIdx.y;
    const int *ri = req_i + M25X_SPEC_NI * r;
    const double *rd = req_d + M25X_SPEC_ND * r;
    const double2 k = kr[node];
    Net n;
    net_build(st, k, n);
    double2 out[2];
    request_value(n, w, k, ri, rd, out);
    const double2 wt = cmul(k, dk[node]);
    const int row = ri[1], nout = ri[0] == M25X_SPEC_TABLE ? 2 : 1;
    for (int q = 0; q < nout; ++q) W[(size_t)(row + q) * npts + node] = cmul(out[q], wt);
}

int check_stack(const m25x_stack *st)
{
    if (!st || st->N < 1 || st->N > ML) return fail(M25X_EINVAL, "spectra: 1 to %d layers", ML);
    if (st->top_pec && st->N < 2) return fail(M25X_EINVAL, "spectra: a PEC cover needs two layers or more");
    if (!(st->wmu > 0) || !(st->weps > 0) || !(st->k0 > 0)) return fail(M25X_EINVAL, "spectra: w > 0");
    return M25X_OK;
}

int check_requests(const m25x_stack *st, const int *ri, int nreq)
{
    for (int r = 0; r < nreq; ++r) {
        const int *q = ri + M25X_SPEC_NI * r;
        if (q[0] < M25X_SPEC_TABLE || q[0] > M25X_SPEC_VIA_R) return fail(M25X_EINVAL, "spectra: request type %d", q[0]);
        // the interfaces each type names: a table o, s; a via kernel k, lo, hi; a via reaction four span ends
        const int nif = q[0] == M25X_SPEC_TABLE ? 2 : (q[0] == M25X_SPEC_VIA_D ? 3 : 4);
        for (int c = 2; c < 2 + nif; ++c)
            if (q[c] < 0 || q[c] > st->N)
                return fail(M25X_EINVAL, "spectra: request %d names interface %d", r, q[c]);
        if (q[0] == M25X_SPEC_TABLE && (q[4] < 0 || q[4] > 2))
            return fail(M25X_EINVAL, "spectra: request %d extracts %d parts (0 to 2)", r, q[4]);
        if (q[1] < 0) return fail(M25X_EINVAL, "spectra: request %d: output row %d", r, q[1]);
    }
    return M25X_OK;
}

}  // namespace

extern "C" {

int m25x_spectra(const m25x_stack *st, double w, const double *kr, const double *dk, int npts, const int *req_i,
                 const double *req_d, int nreq, const int *req_i_host, double *W, long nrows, void *stream)
{
    if (int rc = check_stack(st)) return rc;
    if (npts < 0 || nreq < 0 || !(w > 0) || (npts && nreq && (!kr || !dk || !req_i || !req_d || !W || !req_i_host)))
        return fail(M25X_EINVAL, "spectra: bad arguments");
    if (npts == 0 || nreq == 0) return M25X_OK;
    if (int rc = check_requests(st, req_i_host, nreq)) return rc;
    for (int r = 0; r < nreq; ++r) {
        const int *q = req_i_host + M25X_SPEC_NI * r;
        if (q[1] + (q[0] == M25X_SPEC_TABLE ? 2 : 1) > nrows)
            return fail(M25X_EINVAL, "spectra: request %d writes past the %ld output rows", r, nrows);
    }
    hipLaunchKernelGGL(k_spectra, dim3((npts + 127) / 128, nreq), dim3(128), 0, (hipStream_t)stream, *st, w,
                       (const double2 *)kr, (const double2 *)dk, npts, req_i, req_d, (double2 *)W);
    return check_launch("spectra");
}

int m25x_spectral_point(const m25x_stack *st, double w, double kr_re, double kr_im, int o, int s, double *out)
{
    if (int rc = check_stack(st)) return rc;
    if (!out || o < 0 || s < 0 || o > st->N || s > st->N || !(w > 0)) return fail(M25X_EINVAL, "spectral_point: bad arguments");
    Net n;
    const double2 kr = make_double2(kr_re, kr_im);
    net_build(*st, kr, n);
    double2 ga, gq;
    spectral(n, w, kr, o, s, ga, gq);
    out[0] = ga.x; out[1] = ga.y; out[2] = gq.x; out[3] = gq.y;
    return M25X_OK;
}

}  // extern "C"
