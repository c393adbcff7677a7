// m25x_fill.hip - the impedance-matrix fill on gfx950: far field (centroid rule) in symmetric 32 x 32 tiles,
// the near-field correction as per-pair moments plus a deterministic per-entry gather, the via fan and radial
// sums, the Sommerfeld-Hankel transforms of the Green's tables, the near-pair search and the equilibration.
//
// Everything is FP64 (complex128); only the fill's store may narrow to complex64. The far fill is bound by its
// Z stores (16 B an entry, each tile written once and its mirror once through LDS); the kernel values come from
// the tables, read through L2 (a table is 8192 x 32 B).
#include "m25x_common.h"

using namespace m25x;

namespace {

struct DMesh {
    int Ne, Nt;
    const int *ep, *em, *lt, *te;
    const double *ap, *am, *vp, *vm, *cent, *area, *ta, *tv;
};

DMesh dmesh(const m25x_mesh *g)
{
    DMesh m;
    m.Ne = g->Ne; m.Nt = g->Nt;
    m.ep = g->ep; m.em = g->em; m.lt = g->lt; m.te = g->te;
    m.ap = g->ap; m.am = g->am; m.vp = g->vp; m.vm = g->vm; m.cent = g->cent; m.area = g->area;
    m.ta = g->ta; m.tv = g->tv;
    return m;
}

__device__ inline double cdist(const double *c, int p, int q)
{
    double dx = c[2 * p] - c[2 * q], dy = c[2 * p + 1] - c[2 * q + 1];
    return sqrt(dx * dx + dy * dy);
}

template <typename T> __device__ inline void store_c(T *Z, size_t i, double2 v);
template <> __device__ inline void store_c<double2>(double2 *Z, size_t i, double2 v) { Z[i] = v; }
template <> __device__ inline void store_c<float2>(float2 *Z, size_t i, double2 v)
{
    Z[i] = make_float2((float)v.x, (float)v.y);
}

// ---- far field -------------------------------------------------------------------------------------------
// One 256-thread workgroup per 32 x 32 tile of edge pairs in the upper triangle (tile row <= tile column):
// Z[m, n] for m in the row tile, n in the column tile, each the reference's sum over the halves (s, t) of
//   a_ms a_nt [jw (c_ms - v_ms).(c_nt - v_nt) G_A(r) + 4 G_q(r) / (jw)],  r = |c_ms - c_nt|,
// (mom25d_core.cpp:152-170), stored at (m, n) and, through LDS, at (n, m): Z is symmetric bit for bit.
struct Half {
    double a, dx, dy, cx, cy;
    int lay, pad;
};

constexpr int FT = 32;

template <typename TZ>
__global__ __launch_bounds__(256) void k_fill_far(TZ *__restrict__ Z, long ld, DMesh g, DTab T, double w)
{
    const int ti = blockIdx.y, tj = blockIdx.x;
    if (ti > tj) return;
    __shared__ Half hr[2][FT], hc[2][FT];
    __shared__ double2 tile[FT][FT + 1];
    const int tid = threadIdx.x;
    const int m0 = ti * FT, n0 = tj * FT, Ne = g.Ne;
    if (tid < 4 * FT) {
        const int side = tid / (2 * FT), k = tid % (2 * FT), s = k / FT, e = k % FT;
        const int edge = (side ? n0 : m0) + e;
        Half h;
        h.a = h.dx = h.dy = h.cx = h.cy = 0.0;
        h.lay = 0; h.pad = 0;
        if (edge < Ne) {
            const int tri = s ? g.em[edge] : g.ep[edge];
            const double alpha = s ? g.am[edge] : g.ap[edge];
            const double *v = (s ? g.vm : g.vp) + 2 * (size_t)edge;
            h.a = alpha * g.area[tri];
            h.cx = g.cent[2 * tri]; h.cy = g.cent[2 * tri + 1];
            h.dx = h.cx - v[0]; h.dy = h.cy - v[1];
            h.lay = g.lt[tri];
        }
        if (side) hc[s][e] = h; else hr[s][e] = h;
    }
    __syncthreads();
    const int c = tid % FT, rg = tid / FT;          // 8 row groups of 4 rows
    const double iw = 1.0 / w;
    const Half cn0 = hc[0][c], cn1 = hc[1][c];
#pragma unroll
    for (int k = 0; k < FT / 8; ++k) {
        const int r = rg + 8 * k;
        double vx = 0.0, vy = 0.0;
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const Half hm = hr[s][r];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const Half &hn = t ? cn1 : cn0;
                const double ddx = hm.cx - hn.cx, ddy = hm.cy - hn.cy;
                double2 ga, gq;
                tab_kern(T, hm.lay, hn.lay, sqrt(ddx * ddx + ddy * ddy), ga, gq);
                const double wd = w * (hm.dx * hn.dx + hm.dy * hn.dy);
                const double tx = -wd * ga.y + 4.0 * gq.y * iw, ty = wd * ga.x - 4.0 * gq.x * iw;
                const double f = hm.a * hn.a;
                vx += f * tx; vy += f * ty;
            }
        }
        tile[r][c] = make_double2(vx, vy);
    }
    __syncthreads();
    const bool diag = ti == tj;
#pragma unroll
    for (int k = 0; k < FT / 8; ++k) {
        const int r = rg + 8 * k, m = m0 + r, n = n0 + c;
        if (m < Ne && n < Ne) store_c(Z, (size_t)m * ld + n, (diag && r > c) ? tile[c][r] : tile[r][c]);
    }
    if (!diag) {
#pragma unroll
        for (int k = 0; k < FT / 8; ++k) {
            const int r = rg + 8 * k, n = n0 + r, m = m0 + c;
            if (n < Ne && m < Ne) store_c(Z, (size_t)n * ld + m, tile[c][r]);
        }
    }
}

// far field of arbitrary row / column edge lists, one thread an entry (mom25d_core.cpp:176-211)
__global__ void k_far_block(const int *R, int nr, const int *C, int nc, DMesh g, DTab T, double w, double2 *o)
{
    const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (idx >= (long)nr * nc) return;
    const int i = (int)(idx / nc), j = (int)(idx % nc);
    const int m = R[i], n = C[j];
    const int pm[2] = {g.ep[m], g.em[m]}, pn[2] = {g.ep[n], g.em[n]};
    const double am_[2] = {g.ap[m] * g.area[pm[0]], g.am[m] * g.area[pm[1]]};
    const double an[2] = {g.ap[n] * g.area[pn[0]], g.am[n] * g.area[pn[1]]};
    const double dx[2] = {g.cent[2 * pm[0]] - g.vp[2 * m], g.cent[2 * pm[1]] - g.vm[2 * m]};
    const double dy[2] = {g.cent[2 * pm[0] + 1] - g.vp[2 * m + 1], g.cent[2 * pm[1] + 1] - g.vm[2 * m + 1]};
    const double ex[2] = {g.cent[2 * pn[0]] - g.vp[2 * n], g.cent[2 * pn[1]] - g.vm[2 * n]};
    const double ey[2] = {g.cent[2 * pn[0] + 1] - g.vp[2 * n + 1], g.cent[2 * pn[1] + 1] - g.vm[2 * n + 1]};
    const double iw = 1.0 / w;
    double vx = 0, vy = 0;
    for (int s = 0; s < 2; ++s)
        for (int t = 0; t < 2; ++t) {
            double2 ga, gq;
            tab_kern(T, g.lt[pm[s]], g.lt[pn[t]], cdist(g.cent, pm[s], pn[t]), ga, gq);
            const double wd = w * (dx[s] * ex[t] + dy[s] * ey[t]);
            const double f = am_[s] * an[t];
            vx += f * (-wd * ga.y + 4.0 * gq.y * iw);
            vy += f * (wd * ga.x - 4.0 * gq.x * iw);
        }
    o[idx] = make_double2(vx, vy);
}

// ---- near field ------------------------------------------------------------------------------------------
// the rule over the smaller triangle of a pair, ties by index (mom25d_core.cpp:321-324)
__device__ inline bool rule_over_p(double ap, double aq, int p, int q)
{
    const double tie = 1e-9 * fmax(ap, aq);
    return fabs(ap - aq) <= tie ? p <= q : ap < aq;
}

// analytic inner integral of 1/sqrt(rho^2 + d^2) over triangle V, 6-point outer rule (mom25d_core.cpp:331-365)
__device__ void offplane_one(const double *V, const double *P6, const double *W6, double d, double &m0,
                             double mro[2], double mrs[2], double &mrr)
{
    const double ad = fabs(d);
    double L[3], tau[3][2], un[3][2];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int j = (i + 1) % 3;
        const double ex = V[2 * j] - V[2 * i], ey = V[2 * j + 1] - V[2 * i + 1];
        L[i] = sqrt(ex * ex + ey * ey);
        tau[i][0] = ex / L[i]; tau[i][1] = ey / L[i];
        un[i][0] = tau[i][1]; un[i][1] = -tau[i][0];
    }
    m0 = 0; mro[0] = mro[1] = 0; mrs[0] = mrs[1] = 0; mrr = 0;
    for (int k = 0; k < 6; ++k) {
        const double rx = P6[2 * k], ry = P6[2 * k + 1], wq = W6[k];
        double I1 = 0, IRx = 0, IRy = 0;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double dvx = V[2 * i] - rx, dvy = V[2 * i + 1] - ry;
            const double sm = dvx * tau[i][0] + dvy * tau[i][1];
            const double sp = sm + L[i];
            const double t0 = dvx * un[i][0] + dvy * un[i][1];
            const double R02 = t0 * t0 + d * d, R0 = sqrt(R02);
            const double Rm = sqrt(R02 + sm * sm), Rp = sqrt(R02 + sp * sp);
            const double lg = (R0 > 1e-14 * L[i]) ? asinh(sp / R0) - asinh(sm / R0) : 0.0;
            const double beta = atan2(t0 * sp, R02 + ad * Rp) - atan2(t0 * sm, R02 + ad * Rm);
            I1 += t0 * lg - ad * beta;
            const double eR = 0.5 * (R02 * lg + sp * Rp - sm * Rm);
            IRx += eR * un[i][0]; IRy += eR * un[i][1];
        }
        const double Wx = IRx + rx * I1, Wy = IRy + ry * I1;
        m0 += wq * I1;
        mro[0] += wq * rx * I1; mro[1] += wq * ry * I1;
        mrs[0] += wq * Wx; mrs[1] += wq * Wy;
        mrr += wq * (rx * Wx + ry * Wy);
    }
}

struct DQuad {
    const double *nodes;
    const int *tris;
    const double *P3, *W3, *P6, *W6;
};

// One thread a near pair: the smooth rest by 3 x 3 (same interface) or 6 x 6 (cross) quadrature, the extracted
// parts analytically (kept in stat across frequencies), less the centroid value the far fill put there
// (mom25d_core.cpp:417-476). M: n x 7 complex128 (M0, Mox, Moy, Msx, Msy, Mrr, Mq).
__global__ __launch_bounds__(256) void k_near_moments(DMesh g, DTab T, DQuad qd, const int *P, const int *Q, long n,
                                                      double *stat, int stat_ready, double2 *M, double2 *qcorr)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int p = P[i], q = Q[i];
    const int la = g.lt[p], lb = g.lt[q];
    const int t = T.tab[la * T.L + lb];
    const bool same = la == lb;
    const int np = same ? 3 : 6;
    const double *Pa = same ? qd.P3 + 6 * (size_t)p : qd.P6 + 12 * (size_t)p;
    const double *Pb = same ? qd.P3 + 6 * (size_t)q : qd.P6 + 12 * (size_t)q;
    const double *Wa = same ? qd.W3 + 3 * (size_t)p : qd.W6 + 6 * (size_t)p;
    const double *Wb = same ? qd.W3 + 3 * (size_t)q : qd.W6 + 6 * (size_t)q;
    double2 M0 = {0, 0}, Mox = {0, 0}, Moy = {0, 0}, Msx = {0, 0}, Msy = {0, 0}, Mrr = {0, 0}, Mq = {0, 0};
    for (int a = 0; a < np; ++a) {
        const double pax = Pa[2 * a], pay = Pa[2 * a + 1], wa = Wa[a];
        for (int b = 0; b < np; ++b) {
            const double pbx = Pb[2 * b], pby = Pb[2 * b + 1];
            const double dx = pax - pbx, dy = pay - pby;
            double2 ga, gq;
            tab_rest(T, t, sqrt(dx * dx + dy * dy), ga, gq);
            const double wt = wa * Wb[b];
            ga = cscale(ga, wt); gq = cscale(gq, wt);
            M0 = cadd(M0, ga);
            Mox = cadd(Mox, cscale(ga, pax)); Moy = cadd(Moy, cscale(ga, pay));
            Msx = cadd(Msx, cscale(ga, pbx)); Msy = cadd(Msy, cscale(ga, pby));
            Mrr = cadd(Mrr, cscale(ga, pax * pbx + pay * pby));
            Mq = cadd(Mq, gq);
        }
    }
    const bool keep = rule_over_p(g.area[p], g.area[q], p, q);
    const int po = keep ? p : q, qi = keep ? q : p;
    const int kd = T.kind[t];
    int next = 0;
    double dd[2] = {0, 0};
    double2 cAs[2], cqs[2];
    if (kd == 0) {
        dd[0] = 0.0; cAs[0] = T.cA[t]; cqs[0] = T.cq[t]; next = 1;
        if (T.dimg[t] > 0) { dd[1] = T.dimg[t]; cAs[1] = T.cAi[t]; cqs[1] = T.cqi[t]; next = 2; }
    } else if (kd == 1) {
        dd[0] = T.d[t]; cAs[0] = T.cA[t]; cqs[0] = T.cq[t]; next = 1;
    }
    for (int k = 0; k < next; ++k) {
        double m[6];
        double *sp = stat ? stat + 12 * (size_t)i + 6 * k : nullptr;
        if (sp && stat_ready) {
#pragma unroll
            for (int j = 0; j < 6; ++j) m[j] = sp[j];
        } else {
            double V[6];
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const int nd = qd.tris[3 * qi + j];
                V[2 * j] = qd.nodes[2 * nd]; V[2 * j + 1] = qd.nodes[2 * nd + 1];
            }
            double mro[2], mrs[2];
            offplane_one(V, qd.P6 + 12 * (size_t)po, qd.W6 + 6 * (size_t)po, dd[k], m[0], keep ? mro : mrs,
                         keep ? mrs : mro, m[5]);
            m[1] = mro[0]; m[2] = mro[1]; m[3] = mrs[0]; m[4] = mrs[1];
            if (sp) {
#pragma unroll
                for (int j = 0; j < 6; ++j) sp[j] = m[j];
            }
        }
        const double2 cA = cAs[k], cq = cqs[k];
        M0 = cadd(M0, cscale(cA, m[0])); Mox = cadd(Mox, cscale(cA, m[1])); Moy = cadd(Moy, cscale(cA, m[2]));
        Msx = cadd(Msx, cscale(cA, m[3])); Msy = cadd(Msy, cscale(cA, m[4])); Mrr = cadd(Mrr, cscale(cA, m[5]));
        Mq = cadd(Mq, cscale(cq, m[0]));
    }
    {
        const double r = cdist(g.cent, p, q);
        double2 ga, gq;
        tab_kern(T, la, lb, r, ga, gq);
        const double wgt = g.area[p] * g.area[q];
        const double cpx = g.cent[2 * p], cpy = g.cent[2 * p + 1], cqx = g.cent[2 * q], cqy = g.cent[2 * q + 1];
        ga = cscale(ga, wgt); gq = cscale(gq, wgt);
        M0 = csub(M0, ga);
        Mox = csub(Mox, cscale(ga, cpx)); Moy = csub(Moy, cscale(ga, cpy));
        Msx = csub(Msx, cscale(ga, cqx)); Msy = csub(Msy, cscale(ga, cqy));
        Mrr = csub(Mrr, cscale(ga, cpx * cqx + cpy * cqy));
        Mq = csub(Mq, gq);
    }
    double2 *o = M + 7 * (size_t)i;
    o[0] = M0; o[1] = Mox; o[2] = Moy; o[3] = Msx; o[4] = Msy; o[5] = Mrr; o[6] = Mq;
    if (qcorr) qcorr[i] = Mq;
}

// a pair's value for edge slots s (of p) and u (of q) (mom25d_core.cpp:483-488)
__device__ inline double2 slot_value(const DMesh &g, int p, int q, int s, int u, const double2 *Mi, double w,
                                     double iw)
{
    const double am_ = g.ta[3 * p + s], vx = g.tv[6 * p + 2 * s], vy = g.tv[6 * p + 2 * s + 1];
    const double an = g.ta[3 * q + u], wx = g.tv[6 * q + 2 * u], wy = g.tv[6 * q + 2 * u + 1];
    const double2 M0 = Mi[0], Mox = Mi[1], Moy = Mi[2], Msx = Mi[3], Msy = Mi[4], Mrr = Mi[5], Mq = Mi[6];
    const double vw = vx * wx + vy * wy;
    const double2 Ac = make_double2(Mrr.x - (wx * Mox.x + wy * Moy.x) - (vx * Msx.x + vy * Msy.x) + vw * M0.x,
                                    Mrr.y - (wx * Mox.y + wy * Moy.y) - (vx * Msx.y + vy * Msy.y) + vw * M0.y);
    // jw Ac + 4 Mq / (jw)
    const double2 t = make_double2(-w * Ac.y + 4.0 * Mq.y * iw, w * Ac.x - 4.0 * Mq.x * iw);
    return cscale(t, am_ * an);
}

// one thread a near entry: its contributions summed in a fixed order, then added to Z once
template <typename TZ>
__global__ __launch_bounds__(256) void k_near_scatter(TZ *Z, long ld, DMesh g, const int *P, const int *Q,
                                                      const double2 *M, long nentry, const int *erow,
                                                      const int *ecol, const int *estart, const int *contrib,
                                                      double w)
{
    const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (e >= nentry) return;
    const double iw = 1.0 / w;
    double2 acc = {0, 0};
    for (int c = estart[e]; c < estart[e + 1]; ++c) {
        const int code = contrib[c];
        const int i = code / 9, s = (code / 3) % 3, u = code % 3;
        acc = cadd(acc, slot_value(g, P[i], Q[i], s, u, M + 7 * (size_t)i, w, iw));
    }
    TZ *z = Z + (size_t)erow[e] * ld + ecol[e];
    if constexpr (sizeof(TZ) == 16) {
        double2 v = *(double2 *)z;
        *(double2 *)z = cadd(v, acc);
    } else {
        float2 v = *(float2 *)z;
        *(float2 *)z = make_float2(v.x + (float)acc.x, v.y + (float)acc.y);
    }
}

// the 9 COO triplets of each pair (mom25d_core.cpp:477-506, mode 2)
__global__ void k_near_coo(DMesh g, const int *P, const int *Q, const double2 *M, long n, double w, int *ri, int *ci,
                           double2 *vv)
{
    const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (idx >= 9 * n) return;
    const long i = idx / 9;
    const int s = (int)((idx / 3) % 3), u = (int)(idx % 3);
    const int p = P[i], q = Q[i];
    const int em_ = g.te[3 * p + s], en = g.te[3 * q + u];
    if (em_ < 0 || en < 0) {
        ri[idx] = -1; ci[idx] = 0; vv[idx] = make_double2(0, 0);
        return;
    }
    ri[idx] = em_; ci[idx] = en;
    vv[idx] = slot_value(g, p, q, s, u, M + 7 * (size_t)i, w, 1.0 / w);
}

// ---- near pairs: brute force, one workgroup per triangle p, q in order ------------------------------------
__device__ inline bool is_near(const double *c, const double *sz, const double *zz, int p, int q, double fac)
{
    const double dx = c[2 * p] - c[2 * q], dy = c[2 * p + 1] - c[2 * q + 1], dz = zz[p] - zz[q];
    const double R = fac * fmax(sz[p], sz[q]);
    return dx * dx + dy * dy + dz * dz < R * R;
}

__global__ __launch_bounds__(256) void k_near_count(const double *c, const double *sz, const double *zz, int Nt,
                                                    double fac, long long *counts)
{
    const int p = blockIdx.x;
    __shared__ int part[256];
    int k = 0;
    for (int q = threadIdx.x; q < Nt; q += 256) k += is_near(c, sz, zz, p, q, fac);
    part[threadIdx.x] = k;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) counts[p] = part[0];
}

__global__ __launch_bounds__(256) void k_near_write(const double *c, const double *sz, const double *zz, int Nt,
                                                    double fac, const long long *offs, int *P, int *Q)
{
    const int p = blockIdx.x;
    __shared__ int scan[256];
    __shared__ long long base;
    if (threadIdx.x == 0) base = offs[p];
    __syncthreads();
    for (int q0 = 0; q0 < Nt; q0 += 256) {
        const int q = q0 + threadIdx.x;
        const int f = (q < Nt) && is_near(c, sz, zz, p, q, fac);
        scan[threadIdx.x] = f;
        __syncthreads();
        for (int s = 1; s < 256; s <<= 1) {                 // inclusive Hillis-Steele scan
            const int v = (int)threadIdx.x >= s ? scan[threadIdx.x - s] : 0;
            __syncthreads();
            scan[threadIdx.x] += v;
            __syncthreads();
        }
        if (f) {
            const long long at = base + scan[threadIdx.x] - 1;
            P[at] = p; Q[at] = q;
        }
        __syncthreads();
        if (threadIdx.x == 255) base += scan[255];
        __syncthreads();
    }
}

// ---- fan columns and radial sums ---------------------------------------------------------------------------
__global__ void k_fan_q(DMesh g, DTab T, const int *F, int nf, double2 *o)
{
    const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (idx >= (long)g.Nt * nf) return;
    const int t = (int)(idx / nf), j = (int)(idx % nf);
    const int f = F[j];
    double2 ga, gq;
    tab_kern(T, g.lt[t], g.lt[f], cdist(g.cent, t, f), ga, gq);
    o[idx] = cscale(gq, g.area[t] * g.area[f]);
}

__global__ void k_radial(const double *x, const double *y, const double *W, long nt, long nq, const double *ax,
                         const double *ay, long nv, const double2 *d, long n, double dr, double2 *o)
{
    const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (idx >= nt * nv) return;
    const long t = idx / nv, v = idx % nv;
    const double top = (double)n - 1.000001;
    double2 acc = {0, 0};
    for (long q = 0; q < nq; ++q) {
        const double rx = x[t * nq + q] - ax[v], ry = y[t * nq + q] - ay[v];
        double u = hypot(rx, ry) / dr;
        if (!(u <= top)) u = top;
        const long i = (long)u;
        const double f = u - (double)i;
        const double2 a = d[i], b = d[i + 1];
        const double wt = W[t * nq + q];
        acc = cadd(acc, cscale(make_double2(a.x * (1.0 - f) + b.x * f, a.y * (1.0 - f) + b.y * f), wt));
    }
    o[idx] = acc;
}

// ---- Bessel functions of complex argument (power series / Hankel's asymptotic expansion, Abramowitz &
// Stegun 9.1.10 and 9.2.5; the reference's bessel_jn_complex1, mom25d_core.cpp:1515-1549) ---------------------
__device__ inline double c_abs(double2 z) { return hypot(z.x, z.y); }
__device__ inline double2 c_sqrt(double2 z)
{
    const double r = c_abs(z);
    if (r == 0.0) return make_double2(0.0, z.y);
    if (z.x >= 0) {
        const double t = sqrt(0.5 * (r + z.x));
        return make_double2(t, z.y / (2 * t));
    }
    const double t = sqrt(0.5 * (r - z.x));
    return make_double2(fabs(z.y) / (2 * t), copysign(t, z.y));
}
__device__ inline double2 c_cos(double2 z) { return make_double2(cos(z.x) * cosh(z.y), -sin(z.x) * sinh(z.y)); }
__device__ inline double2 c_sin(double2 z) { return make_double2(sin(z.x) * cosh(z.y), cos(z.x) * sinh(z.y)); }

__device__ double2 bessel_jn_c(int n, double2 z)
{
    const double PI = 3.14159265358979323846;
    const double az = c_abs(z);
    if (az <= 14.0) {
        const double2 h = cscale(z, 0.5);
        const double2 mh2 = cscale(cmul(h, h), -1.0);
        double2 term = n == 0 ? make_double2(1.0, 0.0) : n == 1 ? h : cscale(cmul(h, h), 0.5);
        double2 sum = term;
        double big = c_abs(term);
        for (int k = 1; k < 300; ++k) {
            term = cmul(term, cscale(mh2, 1.0 / ((double)k * (double)(k + n))));
            sum = cadd(sum, term);
            const double at = c_abs(term);
            big = fmax(big, at);
            if (k > az && at <= 1e-18 * big) break;
        }
        return sum;
    }
    const double mu = 4.0 * n * n;
    const double2 iz = cdiv(make_double2(1.0, 0.0), z);
    const double2 chi = make_double2(z.x - (0.5 * n + 0.25) * PI, z.y);
    double2 Pp = {1.0, 0.0}, Qq = {0.0, 0.0}, t = {1.0, 0.0};
    double prev = 1.0;
    for (int k = 1; k < 80; ++k) {
        const double m = 2.0 * k - 1.0;
        const double2 nt = cmul(cscale(t, (mu - m * m) / (8.0 * k)), iz);
        const double at = c_abs(nt);
        if (at >= prev) break;
        t = nt; prev = at;
        const double sg = ((k / 2) & 1) ? -1.0 : 1.0;
        if (k & 1) Qq = cadd(Qq, cscale(t, sg)); else Pp = cadd(Pp, cscale(t, sg));
        if (at < 1e-17) break;
    }
    const double2 pre = c_sqrt(cdiv(make_double2(2.0, 0.0), cscale(z, PI)));
    return cmul(pre, csub(cmul(Pp, c_cos(chi)), cmul(Qq, c_sin(chi))));
}

__global__ void k_bessel(int n, const double2 *z, long total, double2 *o)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < total) o[i] = bessel_jn_c(n, z[i]);
}

// out[s, i] = sum_k J0(rho_i kr_k) W[s, k]: one workgroup per rho; the real nodes and the lifted ones summed
// apart and added at the end, as the reference adds its two parts (via.py:97-101)
constexpr int HS = 4;
__global__ __launch_bounds__(256) void k_hankel(const double *rho, int nr, const double2 *kr, int nk,
                                                const double2 *W, int ns, double2 *out)
{
    const int i = blockIdx.x;
    const double r = rho[i];
    __shared__ double2 red[2][HS][256];
    for (int s0 = 0; s0 < ns; s0 += HS) {
        double2 aR[HS], aC[HS];
#pragma unroll
        for (int s = 0; s < HS; ++s) { aR[s] = make_double2(0, 0); aC[s] = make_double2(0, 0); }
        for (int k = threadIdx.x; k < nk; k += 256) {
            const double2 z = kr[k];
            if (fabs(z.y) < 1e-300) {
                const double j = j0(r * z.x);
#pragma unroll
                for (int s = 0; s < HS; ++s)
                    if (s0 + s < ns) aR[s] = cadd(aR[s], cscale(W[(size_t)(s0 + s) * nk + k], j));
            } else {
                const double2 J = bessel_jn_c(0, make_double2(r * z.x, r * z.y));
#pragma unroll
                for (int s = 0; s < HS; ++s)
                    if (s0 + s < ns) aC[s] = cadd(aC[s], cmul(J, W[(size_t)(s0 + s) * nk + k]));
            }
        }
#pragma unroll
        for (int s = 0; s < HS; ++s) { red[0][s][threadIdx.x] = aR[s]; red[1][s][threadIdx.x] = aC[s]; }
        __syncthreads();
        for (int st = 128; st > 0; st >>= 1) {
            if ((int)threadIdx.x < st) {
#pragma unroll
                for (int s = 0; s < HS; ++s) {
                    red[0][s][threadIdx.x] = cadd(red[0][s][threadIdx.x], red[0][s][threadIdx.x + st]);
                    red[1][s][threadIdx.x] = cadd(red[1][s][threadIdx.x], red[1][s][threadIdx.x + st]);
                }
            }
            __syncthreads();
        }
        if (threadIdx.x < HS && s0 + (int)threadIdx.x < ns)
            out[(size_t)(s0 + threadIdx.x) * nr + i] = cadd(red[0][threadIdx.x][0], red[1][threadIdx.x][0]);
        __syncthreads();
    }
}

// ---- equilibration -------------------------------------------------------------------------------------------
// Z <- (Z d_i) d_j in place (fast.py:300-303: rows, then columns) and its complex64 copy
__global__ void k_scale(double2 *Z, long n, long ld, const double *d, float2 *Z64, long ld64)
{
    const long j = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const long i = blockIdx.y;
    if (j >= n) return;
    const double di = d[i], dj = d[j];
    double2 z = Z[i * ld + j];
    z.x = (z.x * di) * dj; z.y = (z.y * di) * dj;
    Z[i * ld + j] = z;
    if (Z64) Z64[i * ld64 + j] = make_float2((float)z.x, (float)z.y);
}

__global__ void k_diag_scales(const double2 *Z, long n, long ld, double *d)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double2 z = Z[i * ld + i];
    d[i] = 1.0 / sqrt(hypot(z.x, z.y) + 1e-300);
}

inline unsigned blocks(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// ---- the far product without Z (matvec_far, matvec_far_multi) ---------------------------------------------
// Three passes: the sources of each triangle aggregated (per column c: s0 = sum a x_e, s1 = sum a x_e (c_t - v_e)
// over its edges, a = alpha_e area_t), one kernel evaluation per triangle pair for all k columns (each workgroup
// a block of observation triangles against one slice of the sources, partial sums kept per slice), and each
// edge's two halves from the slices summed in slice order. Deterministic: no atomics, fixed orders.
constexpr int MV_P = 128;      // observation triangles a workgroup, one a thread
constexpr int MV_Q = 64;       // source triangles staged in LDS a step
constexpr int MV_KMAX = 8;     // columns a call

// slices of the sources: enough workgroups to fill the chip (about 1024), each slice a multiple of MV_Q
inline void mv_slices(int Nt, int &nslice, int &qper)
{
    const int pblocks = (Nt + MV_P - 1) / MV_P;
    int want = (1024 + pblocks - 1) / pblocks;
    want = want < 1 ? 1 : (want > 64 ? 64 : want);
    qper = (Nt + want - 1) / want;
    qper = (qper + MV_Q - 1) / MV_Q * MV_Q;
    if (qper < MV_Q) qper = MV_Q;
    nslice = (Nt + qper - 1) / qper;
    if (nslice < 1) nslice = 1;
}

// S[(t k + c) 3 + {0, 1, 2}] = (s0, s1x, s1y); the triangle's edges in increasing index (the reference's edge loop
// adds them in that order); act[t]: some edge has t as a half
__global__ void k_mv_sources(DMesh g, const double2 *X, long ldx, int k, double2 *S, int *act)
{
    const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (idx >= (long)g.Nt * k) return;
    const int t = (int)(idx / k), c = (int)(idx % k);
    int e3[3], n = 0;
    for (int i = 0; i < 3; ++i) {
        const int e = g.te[3 * t + i];
        if (e >= 0) e3[n++] = e;
    }
    for (int i = 1; i < n; ++i)
        for (int j = i; j > 0 && e3[j - 1] > e3[j]; --j) {
            const int x = e3[j]; e3[j] = e3[j - 1]; e3[j - 1] = x;
        }
    double2 s0 = {0, 0}, sx = {0, 0}, sy = {0, 0};
    const double cx = g.cent[2 * t], cy = g.cent[2 * t + 1];
    for (int j = 0; j < n; ++j) {
        const int e = e3[j];
        const bool plus = g.ep[e] == t;
        const double f = (plus ? g.ap[e] : g.am[e]) * g.area[t];
        const double *v = (plus ? g.vp : g.vm) + 2 * (size_t)e;
        const double2 xe = X[(size_t)e * ldx + c];
        const double2 a = make_double2(f * xe.x, f * xe.y);
        s0 = cadd(s0, a);
        sx = cadd(sx, cscale(a, cx - v[0]));
        sy = cadd(sy, cscale(a, cy - v[1]));
    }
    S[3 * idx] = s0; S[3 * idx + 1] = sx; S[3 * idx + 2] = sy;
    if (c == 0) act[t] = n > 0;
}

// part[((slice Nt + p) k + c) 3 + {0, 1, 2}] = sum over the slice's active q of (G_A s1x, G_A s1y, G_q s0)
template <int KM>
__global__ __launch_bounds__(MV_P) void k_mv_pairs(DMesh g, DTab T, const double2 *S, const int *act, int k,
                                                   int qper, double2 *part)
{
    __shared__ double2 sS[MV_Q][3 * KM];
    __shared__ double2 sc[MV_Q];
    __shared__ int sl[MV_Q];
    const int p = blockIdx.x * MV_P + threadIdx.x;
    const int q0 = blockIdx.y * qper, q1 = min(g.Nt, q0 + qper);
    const bool live = p < g.Nt;
    const int lp = live ? g.lt[p] : 0;
    const double px = live ? g.cent[2 * p] : 0.0, py = live ? g.cent[2 * p + 1] : 0.0;
    double2 ax[KM], ay[KM], aq[KM];
#pragma unroll
    for (int c = 0; c < KM; ++c) ax[c] = ay[c] = aq[c] = make_double2(0, 0);
    for (int qb = q0; qb < q1; qb += MV_Q) {
        const int nq = min(MV_Q, q1 - qb);
        for (int i = threadIdx.x; i < nq * 3 * k; i += MV_P) {
            const int j = i / (3 * k), r = i % (3 * k);
            sS[j][r] = S[(size_t)(qb + j) * 3 * k + r];
        }
        for (int j = threadIdx.x; j < nq; j += MV_P) {
            sc[j] = make_double2(g.cent[2 * (qb + j)], g.cent[2 * (qb + j) + 1]);
            sl[j] = act[qb + j] ? g.lt[qb + j] : -1;
        }
        __syncthreads();
        if (live) {
            for (int j = 0; j < nq; ++j) {
                const int lq = sl[j];
                if (lq < 0) continue;
                const double dx = px - sc[j].x, dy = py - sc[j].y;
                double2 ga, gq;
                tab_kern(T, lp, lq, sqrt(dx * dx + dy * dy), ga, gq);
#pragma unroll
                for (int c = 0; c < KM; ++c) {
                    if (c < k) {
                        ax[c] = cadd(ax[c], cmul(ga, sS[j][3 * c + 1]));
                        ay[c] = cadd(ay[c], cmul(ga, sS[j][3 * c + 2]));
                        aq[c] = cadd(aq[c], cmul(gq, sS[j][3 * c]));
                    }
                }
            }
        }
        __syncthreads();
    }
    if (!live) return;
    double2 *o = part + ((size_t)blockIdx.y * g.Nt + p) * k * 3;
#pragma unroll
    for (int c = 0; c < KM; ++c)
        if (c < k) { o[3 * c] = ax[c]; o[3 * c + 1] = ay[c]; o[3 * c + 2] = aq[c]; }
}

// Y[e, c] = sum over the halves (t, alpha, v) of alpha area_t [jw ((c_t - v) . vA_t) + 4 wQ_t / (jw)]
// (mom25d_core.cpp:300-311), vA and wQ summed over the slices in order
__global__ void k_mv_out(DMesh g, const double2 *part, int nslice, int k, double w, double2 *Y, long ldy)
{
    const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (idx >= (long)g.Ne * k) return;
    const int e = (int)(idx / k), c = (int)(idx % k);
    const double iw = 1.0 / w;
    double2 y = {0, 0};
    for (int h = 0; h < 2; ++h) {
        const int t = h ? g.em[e] : g.ep[e];
        const double f = (h ? g.am[e] : g.ap[e]) * g.area[t];
        const double *v = (h ? g.vm : g.vp) + 2 * (size_t)e;
        const double dx = g.cent[2 * t] - v[0], dy = g.cent[2 * t + 1] - v[1];
        double2 vx = {0, 0}, vy = {0, 0}, wq = {0, 0};
        for (int sl = 0; sl < nslice; ++sl) {
            const double2 *o = part + (((size_t)sl * g.Nt + t) * k + c) * 3;
            vx = cadd(vx, o[0]); vy = cadd(vy, o[1]); wq = cadd(wq, o[2]);
        }
        const double2 u = make_double2(dx * vx.x + dy * vy.x, dx * vx.y + dy * vy.y);
        // jw u + 4 wQ / (jw)
        const double2 z = make_double2(-w * u.y + 4.0 * wq.y * iw, w * u.x - 4.0 * wq.x * iw);
        y = cadd(y, cscale(z, f));
    }
    Y[(size_t)e * ldy + c] = y;
}

// ---- the near correction as a sparse operator ------------------------------------------------------------
// vals[e] = the correction of near entry e (its contributions in their fixed order, as k_near_scatter adds them)
__global__ __launch_bounds__(256) void k_near_values(DMesh g, const int *P, const int *Q, const double2 *M,
                                                     long nentry, const int *estart, const int *contrib, double w,
                                                     double2 *vals)
{
    const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (e >= nentry) return;
    const double iw = 1.0 / w;
    double2 acc = {0, 0};
    for (int c = estart[e]; c < estart[e + 1]; ++c) {
        const int code = contrib[c];
        const int i = code / 9, s = (code / 3) % 3, u = code % 3;
        acc = cadd(acc, slot_value(g, P[i], Q[i], s, u, M + 7 * (size_t)i, w, iw));
    }
    vals[e] = acc;
}

// Y[r, c] (+)= sum over the entries of row r of vals[e] X[col[e], c]; one thread a (row, column)
__global__ void k_csr_matvec(const int *rowptr, const int *col, const double2 *vals, long nrow, const double2 *X,
                             int k, long ldx, double2 *Y, long ldy, int add)
{
    const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (idx >= nrow * k) return;
    const long r = idx / k;
    const int c = (int)(idx % k);
    double2 acc = add ? Y[r * ldy + c] : make_double2(0, 0);
    double2 sum = {0, 0};
    for (int e = rowptr[r]; e < rowptr[r + 1]; ++e) sum = cadd(sum, cmul(vals[e], X[(size_t)col[e] * ldx + c]));
    Y[r * ldy + c] = cadd(acc, sum);
}

}  // namespace

extern "C" {

int m25x_fill_far(void *Z, int zdtype, long ld, const m25x_mesh *g, const m25x_tables *tb, double w, void *stream)
{
    if (!Z || !g || !tb || ld < g->Ne || !(w > 0)) return fail(M25X_EINVAL, "fill_far: bad arguments");
    if (g->Ne == 0) return M25X_OK;
    const int nt = (g->Ne + FT - 1) / FT;
    dim3 grid(nt, nt);
    hipStream_t s = (hipStream_t)stream;
    TimedLaunch tl(s, M25X_K_FILL_FAR, zdtype, (zdtype == M25X_C128 ? 16.0 : 8.0) * g->Ne * (double)g->Ne, 0.0);
    if (zdtype == M25X_C128)
        hipLaunchKernelGGL(k_fill_far<double2>, grid, dim3(256), 0, s, (double2 *)Z, ld, dmesh(g), dtab(tb), w);
    else if (zdtype == M25X_C64)
        hipLaunchKernelGGL(k_fill_far<float2>, grid, dim3(256), 0, s, (float2 *)Z, ld, dmesh(g), dtab(tb), w);
    else
        return fail(M25X_EINVAL, "fill_far: zdtype %d", zdtype);
    return check_launch("fill_far");
}

int m25x_far_block(const int *rows, int nr, const int *cols, int nc, const m25x_mesh *g, const m25x_tables *tb,
                   double w, double *out, void *stream)
{
    if (nr < 0 || nc < 0 || !g || !tb) return fail(M25X_EINVAL, "far_block: bad arguments");
    if ((long)nr * nc == 0) return M25X_OK;
    hipLaunchKernelGGL(k_far_block, dim3(blocks((long)nr * nc, 256)), dim3(256), 0, (hipStream_t)stream, rows, nr,
                       cols, nc, dmesh(g), dtab(tb), w, (double2 *)out);
    return check_launch("far_block");
}

int m25x_near_moments(const m25x_mesh *g, const m25x_tables *tb, const m25x_quad *q, const int *P, const int *Q,
                      long n, double w, double *stat, int stat_ready, double *M, double *qcorr, void *stream)
{
    if (!g || !tb || !q || n < 0 || !M) return fail(M25X_EINVAL, "near_moments: bad arguments");
    if (n == 0) return M25X_OK;
    DQuad d{q->nodes, q->tris, q->P3, q->W3, q->P6, q->W6};
    TimedLaunch tl((hipStream_t)stream, M25X_K_NEAR_MOMENTS, M25X_C128, 7.0 * 16 * n, 0.0);
    hipLaunchKernelGGL(k_near_moments, dim3(blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, dmesh(g), dtab(tb),
                       d, P, Q, n, stat, stat_ready, (double2 *)M, (double2 *)qcorr);
    return check_launch("near_moments");
}

int m25x_near_scatter(void *Z, int zdtype, long ld, const m25x_mesh *g, const int *P, const int *Q, const double *M,
                      long nentry, const int *erow, const int *ecol, const int *estart, const int *contrib, double w,
                      void *stream)
{
    if (!Z || !g || nentry < 0) return fail(M25X_EINVAL, "near_scatter: bad arguments");
    if (nentry == 0) return M25X_OK;
    hipStream_t s = (hipStream_t)stream;
    TimedLaunch tl(s, M25X_K_NEAR_SCATTER, zdtype, (zdtype == M25X_C128 ? 32.0 : 16.0) * nentry, 0.0);
    if (zdtype == M25X_C128)
        hipLaunchKernelGGL(k_near_scatter<double2>, dim3(blocks(nentry, 256)), dim3(256), 0, s, (double2 *)Z, ld,
                           dmesh(g), P, Q, (const double2 *)M, nentry, erow, ecol, estart, contrib, w);
    else if (zdtype == M25X_C64)
        hipLaunchKernelGGL(k_near_scatter<float2>, dim3(blocks(nentry, 256)), dim3(256), 0, s, (float2 *)Z, ld,
                           dmesh(g), P, Q, (const double2 *)M, nentry, erow, ecol, estart, contrib, w);
    else
        return fail(M25X_EINVAL, "near_scatter: zdtype %d", zdtype);
    return check_launch("near_scatter");
}

size_t m25x_matvec_far_work_bytes(int Nt, int k)
{
    if (Nt <= 0 || k <= 0) return 256;
    int nslice, qper;
    mv_slices(Nt, nslice, qper);
    return sizeof(double2) * 3 * (size_t)Nt * k * (1 + (size_t)nslice) + sizeof(int) * (size_t)Nt + 256;
}

int m25x_matvec_far(const double *X, int k, long ldx, const m25x_mesh *g, const m25x_tables *tb, double w, double *Y,
                    long ldy, void *work, size_t work_bytes, void *stream)
{
    if (!X || !Y || !g || !tb || k < 1 || k > MV_KMAX || ldx < k || ldy < k || !(w > 0))
        return fail(M25X_EINVAL, "matvec_far: bad arguments (1 <= k <= %d, ldx, ldy >= k, w > 0)", MV_KMAX);
    if (g->Ne == 0) return M25X_OK;
    if (!work || work_bytes < m25x_matvec_far_work_bytes(g->Nt, k))
        return fail(M25X_EINVAL, "matvec_far: workspace of %zu bytes, %zu needed", work_bytes,
                    m25x_matvec_far_work_bytes(g->Nt, k));
    hipStream_t s = (hipStream_t)stream;
    int nslice, qper;
    mv_slices(g->Nt, nslice, qper);
    double2 *S = (double2 *)work, *part = S + 3 * (size_t)g->Nt * k;
    int *act = (int *)(part + 3 * (size_t)g->Nt * k * nslice);
    DMesh m = dmesh(g);
    hipLaunchKernelGGL(k_mv_sources, dim3(blocks((long)g->Nt * k, 256)), dim3(256), 0, s, m, (const double2 *)X, ldx,
                       k, S, act);
    if (int rc = check_launch("matvec_far")) return rc;
    const dim3 grid((g->Nt + MV_P - 1) / MV_P, nslice);
    if (k <= 1)
        hipLaunchKernelGGL(k_mv_pairs<1>, grid, dim3(MV_P), 0, s, m, dtab(tb), (const double2 *)S, (const int *)act,
                           k, qper, part);
    else if (k <= 4)
        hipLaunchKernelGGL(k_mv_pairs<4>, grid, dim3(MV_P), 0, s, m, dtab(tb), (const double2 *)S, (const int *)act,
                           k, qper, part);
    else
        hipLaunchKernelGGL(k_mv_pairs<MV_KMAX>, grid, dim3(MV_P), 0, s, m, dtab(tb), (const double2 *)S,
                           (const int *)act, k, qper, part);
    if (int rc = check_launch("matvec_far")) return rc;
    hipLaunchKernelGGL(k_mv_out, dim3(blocks((long)g->Ne * k, 256)), dim3(256), 0, s, m, (const double2 *)part,
                       nslice, k, w, (double2 *)Y, ldy);
    return check_launch("matvec_far");
}

int m25x_near_values(const m25x_mesh *g, const int *P, const int *Q, const double *M, long nentry, const int *estart,
                     const int *contrib, double w, double *vals, void *stream)
{
    if (!g || nentry < 0 || (nentry && (!P || !Q || !M || !estart || !contrib || !vals)))
        return fail(M25X_EINVAL, "near_values: bad arguments");
    if (nentry == 0) return M25X_OK;
    hipLaunchKernelGGL(k_near_values, dim3(blocks(nentry, 256)), dim3(256), 0, (hipStream_t)stream, dmesh(g), P, Q,
                       (const double2 *)M, nentry, estart, contrib, w, (double2 *)vals);
    return check_launch("near_values");
}

int m25x_csr_matvec(const int *rowptr, const int *col, const double *vals, long nrow, const double *X, int k, long ldx,
                    double *Y, long ldy, int add, void *stream)
{
    if (!rowptr || nrow < 0 || k < 1 || ldx < k || ldy < k || !X || !Y)
        return fail(M25X_EINVAL, "csr_matvec: bad arguments");
    if (nrow == 0) return M25X_OK;
    hipLaunchKernelGGL(k_csr_matvec, dim3(blocks(nrow * k, 256)), dim3(256), 0, (hipStream_t)stream, rowptr, col,
                       (const double2 *)vals, nrow, (const double2 *)X, k, ldx, (double2 *)Y, ldy, add);
    return check_launch("csr_matvec");
}

int m25x_near_coo(const m25x_mesh *g, const int *P, const int *Q, const double *M, long n, double w, int *ri, int *ci,
                  double *vv, void *stream)
{
    if (!g || n < 0) return fail(M25X_EINVAL, "near_coo: bad arguments");
    if (n == 0) return M25X_OK;
    hipLaunchKernelGGL(k_near_coo, dim3(blocks(9 * n, 256)), dim3(256), 0, (hipStream_t)stream, dmesh(g), P, Q,
                       (const double2 *)M, n, w, ri, ci, (double2 *)vv);
    return check_launch("near_coo");
}

int m25x_near_pairs(const double *cent, const double *size, const double *z, int Nt, double fac, long long *counts,
                    long *count, int *P, int *Q, void *stream)
{
    if (Nt < 0 || !counts || !count) return fail(M25X_EINVAL, "near_pairs: bad arguments");
    *count = 0;
    if (Nt == 0) return M25X_OK;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_near_count, dim3(Nt), dim3(256), 0, s, cent, size, z, Nt, fac, counts);
    if (int rc = check_launch("near_pairs")) return rc;
    long long *h = (long long *)malloc(sizeof(long long) * ((size_t)Nt + 1));
    if (!h) return fail(M25X_ENOMEM, "near_pairs: host memory");
    hipError_t e = hipMemcpyAsync(h, counts, sizeof(long long) * Nt, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) { free(h); return fail(M25X_EHIP, "near_pairs: %s", hipGetErrorString(e)); }
    long long run = 0;
    for (int p = 0; p < Nt; ++p) { const long long c = h[p]; h[p] = run; run += c; }
    h[Nt] = run;
    *count = (long)run;
    if (P && Q && run > 0) {
        e = hipMemcpyAsync(counts, h, sizeof(long long) * ((size_t)Nt + 1), hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        free(h);
        if (e != hipSuccess) return fail(M25X_EHIP, "near_pairs: %s", hipGetErrorString(e));
        hipLaunchKernelGGL(k_near_write, dim3(Nt), dim3(256), 0, s, cent, size, z, Nt, fac, counts, P, Q);
        return check_launch("near_pairs");
    }
    free(h);
    return M25X_OK;
}

int m25x_fan_q(const m25x_mesh *g, const m25x_tables *tb, const int *F, int nf, double *out, void *stream)
{
    if (!g || !tb || nf < 0) return fail(M25X_EINVAL, "fan_q: bad arguments");
    if ((long)g->Nt * nf == 0) return M25X_OK;
    hipLaunchKernelGGL(k_fan_q, dim3(blocks((long)g->Nt * nf, 256)), dim3(256), 0, (hipStream_t)stream, dmesh(g),
                       dtab(tb), F, nf, (double2 *)out);
    return check_launch("fan_q");
}

int m25x_radial_table_sum(const double *px, const double *py, const double *W, long nt, long nq, const double *vx,
                          const double *vy, long nv, const double *table, long n, double dr, double *out,
                          void *stream)
{
    if (n < 2 || !(dr > 0)) return fail(M25X_EINVAL, "radial_table_sum: a table of two points or more, dr > 0");
    if (nt < 0 || nq < 0 || nv < 0) return fail(M25X_EINVAL, "radial_table_sum: bad sizes");
    if (nt * nv == 0) return M25X_OK;
    hipLaunchKernelGGL(k_radial, dim3(blocks(nt * nv, 256)), dim3(256), 0, (hipStream_t)stream, px, py, W, nt, nq,
                       vx, vy, nv, (const double2 *)table, n, dr, (double2 *)out);
    return check_launch("radial_table_sum");
}

int m25x_hankel(const double *rho, int nr, const double *kr, int nk, const double *W, int ns, double *out,
                void *stream)
{
    if (nr < 0 || nk < 0 || ns < 0) return fail(M25X_EINVAL, "hankel: bad sizes");
    if ((long)nr * ns == 0) return M25X_OK;
    hipLaunchKernelGGL(k_hankel, dim3(nr), dim3(256), 0, (hipStream_t)stream, rho, nr, (const double2 *)kr, nk,
                       (const double2 *)W, ns, (double2 *)out);
    return check_launch("hankel");
}

int m25x_bessel_jn_complex(int n, const double *z, long total, double *out, void *stream)
{
    if (n < 0 || n > 2) return fail(M25X_EINVAL, "bessel_jn_complex: order 0, 1 or 2");
    if (total <= 0) return M25X_OK;
    hipLaunchKernelGGL(k_bessel, dim3(blocks(total, 256)), dim3(256), 0, (hipStream_t)stream, n,
                       (const double2 *)z, total, (double2 *)out);
    return check_launch("bessel_jn_complex");
}

int m25x_scale_c64(double *Z, long n, long ld, const double *d, float *Z64, long ld64, void *stream)
{
    if (!Z || !d || n < 0 || ld < n || (Z64 && ld64 < n)) return fail(M25X_EINVAL, "scale_c64: bad arguments");
    if (n == 0) return M25X_OK;
    hipLaunchKernelGGL(k_scale, dim3(blocks(n, 256), (unsigned)n), dim3(256), 0, (hipStream_t)stream, (double2 *)Z,
                       n, ld, d, (float2 *)Z64, ld64);
    return check_launch("scale_c64");
}

int m25x_diag_scales(const double *Z, long n, long ld, double *d, void *stream)
{
    if (!Z || !d || n < 0 || ld < n) return fail(M25X_EINVAL, "diag_scales: bad arguments");
    if (n == 0) return M25X_OK;
    hipLaunchKernelGGL(k_diag_scales, dim3(blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, (const double2 *)Z,
                       n, ld, d);
    return check_launch("diag_scales");
}

}  // extern "C"
