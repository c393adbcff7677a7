/*
 * m25_oracle.c - TEST INFRASTRUCTURE ONLY: a plain-C (C99 complex + OpenMP) restatement of the reference's
 * fill kernels, used as the parity checker of the HIP product path. Nothing in the product links, loads or
 * calls this file; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg do.
 *
 * Every function follows a reference function (paths relative to the reference root):
 *   orc_fill_far          src/mom25d/core/mom25d_core.cpp:132-173   (fill_far_t: centroid rule, symmetric)
 *   orc_far_block         src/mom25d/core/mom25d_core.cpp:176-211   (far_block)
 *   orc_matvec_far_multi  src/mom25d/core/mom25d_core.cpp:214-314   (matvec_far, matvec_far_multi: Z_far X
 *                                                                     triangle-aggregated, without Z)
 *   tab_rest / tab_kern   src/mom25d/core/mom25d_core.cpp:64-90     (Tables::rest, Tables::kern)
 *   orc_offplane_one      src/mom25d/core/mom25d_core.cpp:331-365   (offplane_one; sing.py:19-73)
 *   rule_over_p           src/mom25d/core/mom25d_core.cpp:321-324   (sing.py:76-85, RULE_TIE)
 *   orc_offplane_pairs    src/mom25d/core/mom25d_core.cpp:367-387
 *   orc_near              src/mom25d/core/mom25d_core.cpp:417-594   (near_core, modes dense / matvec / coo / q)
 *   orc_near_pairs        src/mom25d/core/mom25d_core.cpp:1216-1263 (grid search, sorted by (p, q))
 *   orc_radial_table_sum  src/mom25d/core/mom25d_core.cpp:1269-1298
 *   orc_hankel_real       src/mom25d/core/mom25d_core.cpp:1430-1454 (glibc j0)
 *   orc_bessel_jn_complex src/mom25d/core/mom25d_core.cpp:1510-1558 (power series / Hankel's expansion)
 *   orc_scale_c64         src/mom25d/core/mom25d_core.cpp:1563-1580
 *   orc_fan_q             src/mom25d/core/mom25d_core.cpp:1585-1603
 *
 * Error convention: functions return 0, or a negative code with a message in orc_last_error().
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cd;
typedef float complex cf;

static char g_err[256];
const char *orc_last_error(void) { return g_err; }
static int fail(const char *m) { snprintf(g_err, sizeof g_err, "%s", m); return -1; }

/* tables: A, Q (ntab x nfine complex), per-table dr, d, dimg (double), cA, cq, cAi, cqi (complex),
 * tab (L x L int: table of an interface pair), kind (0 same interface, 1 cross extracted, 2 cross plain) */
typedef struct {
    int ntab, nfine, L;
    const cd *A, *Q, *cA, *cq, *cAi, *cqi;
    const double *dr, *d, *dimg;
    const int *tab, *kind;
} orc_tables;

/* mesh: ep/em (edge -> T+, T-), lt (triangle layer index), te (Nt x 3 edge per slot or -1), ta (Nt x 3 signed
 * alpha), tv (Nt x 3 x 2 free vertex), ap/am (signed alpha per edge), vp/vm (Ne x 2), cent (Nt x 2), area (Nt) */
typedef struct {
    int Ne, Nt;
    const int *ep, *em, *lt, *te;
    const double *ap, *am, *vp, *vm, *cent, *area, *ta, *tv;
} orc_mesh;

/* linear lookup of the smooth rests: u = r / dr clamped at nfine - 1.000001 (mom25d_core.cpp:70-78) */
static inline void tab_rest(const orc_tables *tb, int t, double r, cd *ga, cd *gq)
{
    double u = r / tb->dr[t];
    const double umax = tb->nfine - 1.000001;
    if (u > umax) u = umax;
    int i = (int)u;
    double f = u - i;
    const cd *a = tb->A + (size_t)t * tb->nfine, *q = tb->Q + (size_t)t * tb->nfine;
    *ga = a[i] * (1 - f) + a[i + 1] * f;
    *gq = q[i] * (1 - f) + q[i + 1] * f;
}

/* the kernel at a centroid distance with the extracted parts added back (mom25d_core.cpp:80-89) */
static inline void tab_kern(const orc_tables *tb, int la, int lb, double r, cd *ga, cd *gq)
{
    int t = tb->tab[la * tb->L + lb];
    tab_rest(tb, t, r, ga, gq);
    if (tb->kind[t] == 0) {
        if (r > 0) { *ga += tb->cA[t] / r; *gq += tb->cq[t] / r; }
        if (tb->dimg[t] > 0) {
            double ri = 1.0 / sqrt(r * r + tb->dimg[t] * tb->dimg[t]);
            *ga += tb->cAi[t] * ri; *gq += tb->cqi[t] * ri;
        }
    } else if (tb->kind[t] == 1) {
        double ri = 1.0 / sqrt(r * r + tb->d[t] * tb->d[t]);
        *ga += tb->cA[t] * ri; *gq += tb->cq[t] * ri;
    }
}

static inline double cdist(const double *c, int p, int q)
{
    double dx = c[2 * p] - c[2 * q], dy = c[2 * p + 1] - c[2 * q + 1];
    return sqrt(dx * dx + dy * dy);
}

/* ---- far field: every edge pair by the centroid rule, Z[m, n] = Z[n, m] (mom25d_core.cpp:132-173) ---- */
int orc_fill_far(void *Zv, int single, long ld, const orc_mesh *g, const orc_tables *tb, double w)
{
    const cd jw = I * w, inv_jw = 1.0 / (I * w);
    const int Ne = g->Ne, Nt = g->Nt;
#pragma omp parallel
    {
        cd *GA = malloc(sizeof(cd) * 2 * (size_t)Nt), *GQ = malloc(sizeof(cd) * 2 * (size_t)Nt);
#pragma omp for schedule(dynamic, 8)
        for (int m = 0; m < Ne; ++m) {
            int pm[2] = {g->ep[m], g->em[m]};
            for (int s = 0; s < 2; ++s)
                for (int q = 0; q < Nt; ++q)
                    tab_kern(tb, g->lt[pm[s]], g->lt[q], cdist(g->cent, pm[s], q), &GA[s * Nt + q], &GQ[s * Nt + q]);
            double am_[2] = {g->ap[m] * g->area[pm[0]], g->am[m] * g->area[pm[1]]};
            double dx[2], dy[2];
            dx[0] = g->cent[2 * pm[0]] - g->vp[2 * m]; dy[0] = g->cent[2 * pm[0] + 1] - g->vp[2 * m + 1];
            dx[1] = g->cent[2 * pm[1]] - g->vm[2 * m]; dy[1] = g->cent[2 * pm[1] + 1] - g->vm[2 * m + 1];
            for (int n = m; n < Ne; ++n) {
                int pn[2] = {g->ep[n], g->em[n]};
                double an[2] = {g->ap[n] * g->area[pn[0]], g->am[n] * g->area[pn[1]]};
                double ex[2], ey[2];
                ex[0] = g->cent[2 * pn[0]] - g->vp[2 * n]; ey[0] = g->cent[2 * pn[0] + 1] - g->vp[2 * n + 1];
                ex[1] = g->cent[2 * pn[1]] - g->vm[2 * n]; ey[1] = g->cent[2 * pn[1] + 1] - g->vm[2 * n + 1];
                cd v = 0;
                for (int s = 0; s < 2; ++s)
                    for (int t = 0; t < 2; ++t) {
                        size_t k = s * (size_t)Nt + pn[t];
                        double dot = dx[s] * ex[t] + dy[s] * ey[t];
                        v += am_[s] * an[t] * (jw * dot * GA[k] + 4.0 * GQ[k] * inv_jw);
                    }
                if (single) {
                    ((cf *)Zv)[(size_t)m * ld + n] = (cf)v;
                    ((cf *)Zv)[(size_t)n * ld + m] = (cf)v;
                } else {
                    ((cd *)Zv)[(size_t)m * ld + n] = v;
                    ((cd *)Zv)[(size_t)n * ld + m] = v;
                }
            }
        }
        free(GA); free(GQ);
    }
    return 0;
}

/* ---- Y = Z_far X without Z (mom25d_core.cpp:214-314): the sources aggregated per triangle (s0 = sum a A x,
 * s1 = sum a A x (c - v), edges in order), one kernel evaluation per triangle pair for all k columns (X, Y: Ne x k,
 * row-major), then each edge's two halves. k = 1 is matvec_far (which skips a q whose three sums are zero, where
 * matvec_far_multi skips a q no edge touches: the same sums). ---- */
int orc_matvec_far_multi(const cd *X, int K, const orc_mesh *g, const orc_tables *tb, double w, cd *Y)
{
    const int Ne = g->Ne, Nt = g->Nt;
    const cd jw = I * w, inv_jw = 1.0 / (I * w);
    cd *s0 = calloc((size_t)Nt * K, sizeof(cd)), *s1x = calloc((size_t)Nt * K, sizeof(cd));
    cd *s1y = calloc((size_t)Nt * K, sizeof(cd));
    cd *vAx = malloc(sizeof(cd) * (size_t)Nt * K), *vAy = malloc(sizeof(cd) * (size_t)Nt * K);
    cd *wQ = malloc(sizeof(cd) * (size_t)Nt * K);
    char *act = calloc((size_t)Nt, 1);
    if (!s0 || !s1x || !s1y || !vAx || !vAy || !wQ || !act) {
        free(s0); free(s1x); free(s1y); free(vAx); free(vAy); free(wQ); free(act);
        return fail("matvec_far: out of memory");
    }
    for (int e = 0; e < Ne; ++e) {
        const int p = g->ep[e], q = g->em[e];
        for (int c = 0; c < K; ++c) {
            const cd xe = X[(size_t)e * K + c];
            const cd a = g->ap[e] * g->area[p] * xe, b = g->am[e] * g->area[q] * xe;
            const size_t ip = (size_t)p * K + c, iq = (size_t)q * K + c;
            s0[ip] += a; s1x[ip] += a * (g->cent[2 * p] - g->vp[2 * e]); s1y[ip] += a * (g->cent[2 * p + 1] - g->vp[2 * e + 1]);
            s0[iq] += b; s1x[iq] += b * (g->cent[2 * q] - g->vm[2 * e]); s1y[iq] += b * (g->cent[2 * q + 1] - g->vm[2 * e + 1]);
        }
        act[p] = act[q] = 1;
    }
#pragma omp parallel
    {
        cd *ax = malloc(sizeof(cd) * K), *ay = malloc(sizeof(cd) * K), *aq = malloc(sizeof(cd) * K);
#pragma omp for schedule(dynamic, 16)
        for (int p = 0; p < Nt; ++p) {
            for (int c = 0; c < K; ++c) ax[c] = ay[c] = aq[c] = 0;
            for (int q = 0; q < Nt; ++q) {
                if (!act[q]) continue;
                cd ga, gq;
                tab_kern(tb, g->lt[p], g->lt[q], cdist(g->cent, p, q), &ga, &gq);
                const cd *S0 = s0 + (size_t)q * K, *SX = s1x + (size_t)q * K, *SY = s1y + (size_t)q * K;
                for (int c = 0; c < K; ++c) { ax[c] += ga * SX[c]; ay[c] += ga * SY[c]; aq[c] += gq * S0[c]; }
            }
            for (int c = 0; c < K; ++c) {
                vAx[(size_t)p * K + c] = ax[c]; vAy[(size_t)p * K + c] = ay[c]; wQ[(size_t)p * K + c] = aq[c];
            }
        }
        free(ax); free(ay); free(aq);
    }
    for (int e = 0; e < Ne; ++e) {
        const int p = g->ep[e], q = g->em[e];
        const double dpx = g->cent[2 * p] - g->vp[2 * e], dpy = g->cent[2 * p + 1] - g->vp[2 * e + 1];
        const double dqx = g->cent[2 * q] - g->vm[2 * e], dqy = g->cent[2 * q + 1] - g->vm[2 * e + 1];
        for (int c = 0; c < K; ++c) {
            const size_t ip = (size_t)p * K + c, iq = (size_t)q * K + c;
            const cd yp = g->ap[e] * g->area[p] * (jw * (dpx * vAx[ip] + dpy * vAy[ip]) + 4.0 * wQ[ip] * inv_jw);
            const cd ym = g->am[e] * g->area[q] * (jw * (dqx * vAx[iq] + dqy * vAy[iq]) + 4.0 * wQ[iq] * inv_jw);
            Y[(size_t)e * K + c] = yp + ym;
        }
    }
    free(s0); free(s1x); free(s1y); free(vAx); free(vAy); free(wQ); free(act);
    return 0;
}

/* ---- far field for arbitrary row / column edge lists (mom25d_core.cpp:176-211) ---- */
int orc_far_block(const int *R, int nr, const int *C, int nc, const orc_mesh *g, const orc_tables *tb, double w,
                  cd *o)
{
    const cd jw = I * w, inv_jw = 1.0 / (I * w);
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < nr; ++i) {
        int m = R[i];
        int pm[2] = {g->ep[m], g->em[m]};
        double am_[2] = {g->ap[m] * g->area[pm[0]], g->am[m] * g->area[pm[1]]};
        double dx[2] = {g->cent[2 * pm[0]] - g->vp[2 * m], g->cent[2 * pm[1]] - g->vm[2 * m]};
        double dy[2] = {g->cent[2 * pm[0] + 1] - g->vp[2 * m + 1], g->cent[2 * pm[1] + 1] - g->vm[2 * m + 1]};
        for (int j = 0; j < nc; ++j) {
            int n = C[j];
            int pn[2] = {g->ep[n], g->em[n]};
            double an[2] = {g->ap[n] * g->area[pn[0]], g->am[n] * g->area[pn[1]]};
            double ex[2] = {g->cent[2 * pn[0]] - g->vp[2 * n], g->cent[2 * pn[1]] - g->vm[2 * n]};
            double ey[2] = {g->cent[2 * pn[0] + 1] - g->vp[2 * n + 1], g->cent[2 * pn[1] + 1] - g->vm[2 * n + 1]};
            cd v = 0;
            for (int s = 0; s < 2; ++s)
                for (int t = 0; t < 2; ++t) {
                    cd ga, gq;
                    tab_kern(tb, g->lt[pm[s]], g->lt[pn[t]], cdist(g->cent, pm[s], pn[t]), &ga, &gq);
                    v += am_[s] * an[t] * (jw * (dx[s] * ex[t] + dy[s] * ey[t]) * ga + 4.0 * gq * inv_jw);
                }
            o[(size_t)i * nc + j] = v;
        }
    }
    return 0;
}

/* ---- the rule over the smaller triangle of a pair, ties by index (mom25d_core.cpp:321-324, sing.py:76-85) ---- */
static inline int rule_over_p(double ap, double aq, int p, int q)
{
    const double tie = 1e-9 * (ap > aq ? ap : aq);
    return fabs(ap - aq) <= tie ? p <= q : ap < aq;
}

static inline double tri_area(const double *nd, const int *tr, int t)
{
    const double *a = nd + 2 * tr[3 * t], *b = nd + 2 * tr[3 * t + 1], *c = nd + 2 * tr[3 * t + 2];
    return 0.5 * fabs((b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0]));
}

/* analytic inner integral of 1/sqrt(rho^2 + d^2) (and of (r' - r) / R) over triangle V, a 6-point outer rule
 * (P6, W6): m0, mro (observation-weighted), mrs (source-weighted), mrr (mom25d_core.cpp:331-365) */
void orc_offplane_one(const double *V, const double *P6, const double *W6, double d, double *m0, double mro[2],
                      double mrs[2], double *mrr)
{
    double ad = fabs(d);
    double L[3], tau[3][2], un[3][2];
    for (int i = 0; i < 3; ++i) {
        int j = (i + 1) % 3;
        double ex = V[2 * j] - V[2 * i], ey = V[2 * j + 1] - V[2 * i + 1];
        L[i] = sqrt(ex * ex + ey * ey);
        tau[i][0] = ex / L[i]; tau[i][1] = ey / L[i];
        un[i][0] = tau[i][1]; un[i][1] = -tau[i][0];
    }
    *m0 = 0; mro[0] = mro[1] = 0; mrs[0] = mrs[1] = 0; *mrr = 0;
    for (int k = 0; k < 6; ++k) {
        double rx = P6[2 * k], ry = P6[2 * k + 1], wq = W6[k];
        double I1 = 0, IRx = 0, IRy = 0;
        for (int i = 0; i < 3; ++i) {
            double dvx = V[2 * i] - rx, dvy = V[2 * i + 1] - ry;
            double sm = dvx * tau[i][0] + dvy * tau[i][1];
            double sp = sm + L[i];
            double t0 = dvx * un[i][0] + dvy * un[i][1];
            double R02 = t0 * t0 + d * d, R0 = sqrt(R02);
            double Rm = sqrt(R02 + sm * sm), Rp = sqrt(R02 + sp * sp);
            double lg = (R0 > 1e-14 * L[i]) ? asinh(sp / R0) - asinh(sm / R0) : 0.0;
            double beta = atan2(t0 * sp, R02 + ad * Rp) - atan2(t0 * sm, R02 + ad * Rm);
            I1 += t0 * lg - ad * beta;
            double eR = 0.5 * (R02 * lg + sp * Rp - sm * Rm);
            IRx += eR * un[i][0]; IRy += eR * un[i][1];
        }
        double Wx = IRx + rx * I1, Wy = IRy + ry * I1;
        *m0 += wq * I1;
        mro[0] += wq * rx * I1; mro[1] += wq * ry * I1;
        mrs[0] += wq * Wx; mrs[1] += wq * Wy;
        *mrr += wq * (rx * Wx + ry * Wy);
    }
}

/* the pair moments, the rule over the smaller triangle (mom25d_core.cpp:367-387) */
int orc_offplane_pairs(const double *nd, const int *tr, const double *p6, const double *w6, const int *pp,
                       const int *qq, int n, double d, double *m0, double *mro, double *mrs, double *mrr)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        const int keep = rule_over_p(tri_area(nd, tr, pp[i]), tri_area(nd, tr, qq[i]), pp[i], qq[i]);
        const int po = keep ? pp[i] : qq[i], qi = keep ? qq[i] : pp[i];
        double V[6];
        for (int k = 0; k < 3; ++k) { V[2 * k] = nd[2 * tr[3 * qi + k]]; V[2 * k + 1] = nd[2 * tr[3 * qi + k] + 1]; }
        orc_offplane_one(V, p6 + 12 * (size_t)po, w6 + 6 * (size_t)po, d, &m0[i], keep ? mro + 2 * i : mrs + 2 * i,
                         keep ? mrs + 2 * i : mro + 2 * i, &mrr[i]);
    }
    return 0;
}

/* ---- near-field correction, accurate minus centroid per near pair (mom25d_core.cpp:393-507) ----
 * mode 0: Z (single: complex64, else complex128, leading dimension ld) += correction; mode 2: COO triplets
 * (9 a pair, row -1 for unused slots); mode 3: qcorr alone. stat: n x 12 kept analytic parts (or NULL). */
typedef struct {
    int mode;
    void *Z; long ld; int single;
    int *ri, *ci; cd *vv; cd *qcorr;
    double *stat; int stat_ready;
    const cd *x; cd *y;
} orc_nearout;

int orc_near(const orc_mesh *g, const orc_tables *tb, const double *nodes, const int *tris, const double *P3,
             const double *W3, const double *P6, const double *W6, const int *P, const int *Q, int n, double w,
             orc_nearout *o)
{
    const cd jw = I * w, inv_jw = 1.0 / (I * w);
    /* mode 0 adds with atomics in no fixed order, as the reference does (mom25d_core.cpp:402-415) */
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < n; ++i) {
        const int p = P[i], q = Q[i];
        const int la = g->lt[p], lb = g->lt[q];
        const int t = tb->tab[la * tb->L + lb];
        const int same = (la == lb);
        const int np = same ? 3 : 6;
        const double *Pa = same ? P3 + 6 * (size_t)p : P6 + 12 * (size_t)p;
        const double *Pb = same ? P3 + 6 * (size_t)q : P6 + 12 * (size_t)q;
        const double *Wa = same ? W3 + 3 * (size_t)p : W6 + 6 * (size_t)p;
        const double *Wb = same ? W3 + 3 * (size_t)q : W6 + 6 * (size_t)q;
        cd M0 = 0, Mox = 0, Moy = 0, Msx = 0, Msy = 0, Mrr = 0, Mq = 0;
        for (int a = 0; a < np; ++a)
            for (int b = 0; b < np; ++b) {
                double dx = Pa[2 * a] - Pb[2 * b], dy = Pa[2 * a + 1] - Pb[2 * b + 1];
                cd ga, gq;
                tab_rest(tb, t, sqrt(dx * dx + dy * dy), &ga, &gq);
                double wt = Wa[a] * Wb[b];
                ga *= wt; gq *= wt;
                M0 += ga; Mox += ga * Pa[2 * a]; Moy += ga * Pa[2 * a + 1];
                Msx += ga * Pb[2 * b]; Msy += ga * Pb[2 * b + 1];
                Mrr += ga * (Pa[2 * a] * Pb[2 * b] + Pa[2 * a + 1] * Pb[2 * b + 1]);
                Mq += gq;
            }
        const int keep = rule_over_p(g->area[p], g->area[q], p, q);
        const int po = keep ? p : q, qi = keep ? q : p;
        double V[6];
        for (int k = 0; k < 3; ++k) { V[2 * k] = nodes[2 * tris[3 * qi + k]]; V[2 * k + 1] = nodes[2 * tris[3 * qi + k] + 1]; }
        int next = 0; double dd[2]; cd cAs[2], cqs[2];
        if (tb->kind[t] == 0) {
            dd[next] = 0.0; cAs[next] = tb->cA[t]; cqs[next] = tb->cq[t]; ++next;
            if (tb->dimg[t] > 0) { dd[next] = tb->dimg[t]; cAs[next] = tb->cAi[t]; cqs[next] = tb->cqi[t]; ++next; }
        } else if (tb->kind[t] == 1) {
            dd[next] = tb->d[t]; cAs[next] = tb->cA[t]; cqs[next] = tb->cq[t]; ++next;
        }
        for (int k = 0; k < next; ++k) {
            double st[6], *m = o->stat ? o->stat + 12 * (size_t)i + 6 * k : st;
            if (!(o->stat && o->stat_ready)) {
                double mro[2], mrs[2];
                orc_offplane_one(V, P6 + 12 * (size_t)po, W6 + 6 * (size_t)po, dd[k], &m[0], keep ? mro : mrs,
                                 keep ? mrs : mro, &m[5]);
                m[1] = mro[0]; m[2] = mro[1]; m[3] = mrs[0]; m[4] = mrs[1];
            }
            M0 += cAs[k] * m[0]; Mox += cAs[k] * m[1]; Moy += cAs[k] * m[2]; Msx += cAs[k] * m[3];
            Msy += cAs[k] * m[4]; Mrr += cAs[k] * m[5]; Mq += cqs[k] * m[0];
        }
        {
            double r = cdist(g->cent, p, q);
            cd ga, gq;
            tab_kern(tb, la, lb, r, &ga, &gq);
            double wgt = g->area[p] * g->area[q];
            double cpx = g->cent[2 * p], cpy = g->cent[2 * p + 1], cqx = g->cent[2 * q], cqy = g->cent[2 * q + 1];
            ga *= wgt; gq *= wgt;
            M0 -= ga; Mox -= ga * cpx; Moy -= ga * cpy; Msx -= ga * cqx; Msy -= ga * cqy;
            Mrr -= ga * (cpx * cqx + cpy * cqy); Mq -= gq;
        }
        if (o->qcorr) o->qcorr[i] = Mq;
        if (o->mode == 3) continue;
        for (int s = 0; s < 3; ++s) {
            int em_ = g->te[3 * p + s];
            for (int u = 0; u < 3; ++u) {
                int slot = 9 * i + 3 * s + u;
                int en = g->te[3 * q + u];
                if (em_ < 0 || en < 0) { if (o->mode == 2) o->ri[slot] = -1; continue; }
                double am_ = g->ta[3 * p + s], vx = g->tv[6 * p + 2 * s], vy = g->tv[6 * p + 2 * s + 1];
                double an = g->ta[3 * q + u], wx = g->tv[6 * q + 2 * u], wy = g->tv[6 * q + 2 * u + 1];
                cd Acomb = Mrr - (wx * Mox + wy * Moy) - (vx * Msx + vy * Msy) + (vx * wx + vy * wy) * M0;
                cd val = am_ * an * (jw * Acomb + 4.0 * Mq * inv_jw);
                if (o->mode == 0) {
                    if (o->single) {
                        float *z = (float *)((cf *)o->Z + (size_t)em_ * o->ld + en);
#pragma omp atomic
                        z[0] += (float)creal(val);
#pragma omp atomic
                        z[1] += (float)cimag(val);
                    } else {
                        double *z = (double *)((cd *)o->Z + (size_t)em_ * o->ld + en);
#pragma omp atomic
                        z[0] += creal(val);
#pragma omp atomic
                        z[1] += cimag(val);
                    }
                } else if (o->mode == 1) {      /* y[em] += val x[en] (mom25d_core.cpp:539-556) */
                    cd t = val * o->x[en];
                    double *yy = (double *)(o->y + em_);
#pragma omp atomic
                    yy[0] += creal(t);
#pragma omp atomic
                    yy[1] += cimag(t);
                } else if (o->mode == 2) {
                    o->ri[slot] = em_; o->ci[slot] = en; o->vv[slot] = val;
                }
            }
        }
    }
    return 0;
}

/* ---- near triangle pairs by a grid search, sorted by (p, q) (mom25d_core.cpp:1216-1263) ----
 * two calls: out_P == NULL returns the count; then the pairs are written. */
static int cmp_int(const void *a, const void *b) { int x = *(const int *)a, y = *(const int *)b; return (x > y) - (x < y); }

long orc_near_pairs(const double *c, const double *sz, const double *zz, int Nt, double fac, int64_t *out_P,
                    int64_t *out_Q)
{
    double smax = 0, x0 = 1e300, y0 = 1e300, x1 = -1e300, y1 = -1e300;
    for (int t = 0; t < Nt; ++t) {
        if (sz[t] > smax) smax = sz[t];
        if (c[2 * t] < x0) x0 = c[2 * t];
        if (c[2 * t] > x1) x1 = c[2 * t];
        if (c[2 * t + 1] < y0) y0 = c[2 * t + 1];
        if (c[2 * t + 1] > y1) y1 = c[2 * t + 1];
    }
    double cell = fac * smax > 1e-12 ? fac * smax : 1e-12;
    int64_t nx = (int64_t)((x1 - x0) / cell) + 1, ny = (int64_t)((y1 - y0) / cell) + 1;
    if (nx < 1) nx = 1;
    if (ny < 1) ny = 1;
    while (nx * ny > (int64_t)4 * Nt + 16) { cell *= 1.5; nx = (int64_t)((x1 - x0) / cell) + 1; ny = (int64_t)((y1 - y0) / cell) + 1; }
    int64_t *cellof = malloc(sizeof(int64_t) * (size_t)Nt), *start = calloc((size_t)(nx * ny + 1), sizeof(int64_t));
    int *order = malloc(sizeof(int) * (size_t)Nt);
    for (int t = 0; t < Nt; ++t) {
        int64_t i = (int64_t)((c[2 * t] - x0) / cell), j = (int64_t)((c[2 * t + 1] - y0) / cell);
        if (i > nx - 1) i = nx - 1;
        if (j > ny - 1) j = ny - 1;
        cellof[t] = i * ny + j; start[cellof[t] + 1]++;
    }
    for (int64_t k = 0; k < nx * ny; ++k) start[k + 1] += start[k];
    {
        int64_t *fill = malloc(sizeof(int64_t) * (size_t)(nx * ny));
        memcpy(fill, start, sizeof(int64_t) * (size_t)(nx * ny));
        for (int t = 0; t < Nt; ++t) order[fill[cellof[t]]++] = t;
        free(fill);
    }
    const int64_t reach = (int64_t)ceil(fac * smax / cell);
    int **nb = malloc(sizeof(int *) * (size_t)Nt);
    int *cnt = calloc((size_t)Nt, sizeof(int));
#pragma omp parallel for schedule(dynamic, 64)
    for (int p = 0; p < Nt; ++p) {
        int64_t ci = cellof[p] / ny, cj = cellof[p] % ny;
        int cap = 64, k2 = 0;
        int *out = malloc(sizeof(int) * cap);
        for (int64_t i = (ci - reach > 0 ? ci - reach : 0); i <= (ci + reach < nx - 1 ? ci + reach : nx - 1); ++i)
            for (int64_t j = (cj - reach > 0 ? cj - reach : 0); j <= (cj + reach < ny - 1 ? cj + reach : ny - 1); ++j)
                for (int64_t k = start[i * ny + j]; k < start[i * ny + j + 1]; ++k) {
                    int q = order[k];
                    double dx = c[2 * p] - c[2 * q], dy = c[2 * p + 1] - c[2 * q + 1], dz = zz[p] - zz[q];
                    double R = fac * (sz[p] > sz[q] ? sz[p] : sz[q]);
                    if (dx * dx + dy * dy + dz * dz < R * R) {
                        if (k2 == cap) { cap *= 2; out = realloc(out, sizeof(int) * cap); }
                        out[k2++] = q;
                    }
                }
        qsort(out, (size_t)k2, sizeof(int), cmp_int);
        nb[p] = out; cnt[p] = k2;
    }
    long total = 0;
    for (int p = 0; p < Nt; ++p) total += cnt[p];
    if (out_P) {
        long k = 0;
        for (int p = 0; p < Nt; ++p)
            for (int j = 0; j < cnt[p]; ++j) { out_P[k] = p; out_Q[k] = nb[p][j]; ++k; }
    }
    for (int p = 0; p < Nt; ++p) free(nb[p]);
    free(nb); free(cnt); free(cellof); free(start); free(order);
    return total;
}

/* ---- out[t, v] = sum_q D(|P[t, q] - V[v]|) W[t, q], linear lookup (mom25d_core.cpp:1269-1298) ---- */
int orc_radial_table_sum(const double *x, const double *y, const double *w, long nt, long nq, const double *ax,
                         const double *ay, long nv, const cd *d, long n, double dr, cd *o)
{
    if (n < 2 || !(dr > 0)) return fail("radial_table_sum: a table of two points or more, dr > 0");
    const double top = (double)n - 1.000001;
#pragma omp parallel for schedule(static)
    for (long t = 0; t < nt; ++t)
        for (long v = 0; v < nv; ++v) {
            cd acc = 0;
            for (long q = 0; q < nq; ++q) {
                const double rx = x[t * nq + q] - ax[v], ry = y[t * nq + q] - ay[v];
                double u = hypot(rx, ry) / dr;
                if (u > top) u = top;
                const long i = (long)u;
                const double f = u - (double)i;
                acc += (d[i] * (1.0 - f) + d[i + 1] * f) * w[t * nq + q];
            }
            o[t * nv + v] = acc;
        }
    return 0;
}

/* ---- out[s, i] = sum_k J0(rho_i kr_k) W[s, k], real kr, glibc's j0 (mom25d_core.cpp:1430-1454) ---- */
int orc_hankel_real(const double *r, int nr, const double *k, int nk, const cd *W, int ns, cd *o)
{
#pragma omp parallel
    {
        cd *acc = malloc(sizeof(cd) * (size_t)(ns > 0 ? ns : 1));
#pragma omp for schedule(static)
        for (int i = 0; i < nr; ++i) {
            for (int s = 0; s < ns; ++s) acc[s] = 0;
            for (int m = 0; m < nk; ++m) {
                double j = j0(r[i] * k[m]);
                for (int s = 0; s < ns; ++s) acc[s] += j * W[(size_t)s * nk + m];
            }
            for (int s = 0; s < ns; ++s) o[(size_t)s * nr + i] = acc[s];
        }
        free(acc);
    }
    return 0;
}

/* ---- J_n(z), complex z, n = 0, 1, 2: power series to |z| = 14, Hankel's asymptotic expansion beyond, summed
 * to its smallest term (mom25d_core.cpp:1510-1558; Abramowitz & Stegun 9.1.10 and 9.2.5) ---- */
static cd bessel_jn_complex1(int n, cd z)
{
    const double PI = 3.14159265358979323846;
    const double az = cabs(z);
    if (az <= 14.0) {
        const cd h = 0.5 * z, mh2 = -h * h;
        cd term = n == 0 ? 1.0 : n == 1 ? h : 0.5 * h * h;
        cd sum = term;
        double big = cabs(term);
        for (int k = 1; k < 300; ++k) {
            term *= mh2 / ((double)k * (double)(k + n));
            sum += term;
            const double at = cabs(term);
            if (at > big) big = at;
            if (k > az && at <= 1e-18 * big) break;
        }
        return sum;
    }
    const double mu = 4.0 * n * n;
    const cd iz = 1.0 / z, chi = z - (0.5 * n + 0.25) * PI;
    cd Pp = 1.0, Qq = 0.0, t = 1.0;
    double prev = 1.0;
    for (int k = 1; k < 80; ++k) {
        const double m = 2.0 * k - 1.0;
        const cd nt = t * ((mu - m * m) / (8.0 * k)) * iz;
        const double at = cabs(nt);
        if (at >= prev) break;
        t = nt; prev = at;
        const double sg = ((k / 2) & 1) ? -1.0 : 1.0;
        if (k & 1) Qq += sg * t; else Pp += sg * t;
        if (at < 1e-17) break;
    }
    return csqrt(2.0 / (PI * z)) * (Pp * ccos(chi) - Qq * csin(chi));
}

int orc_bessel_jn_complex(int n, const cd *z, long total, cd *o)
{
    if (n < 0 || n > 2) return fail("bessel_jn_complex: order 0, 1 or 2");
#pragma omp parallel for schedule(static) if (total > 4096)
    for (long i = 0; i < total; ++i) o[i] = bessel_jn_complex1(n, z[i]);
    return 0;
}

/* ---- Z <- diag(d) Z diag(d) in place, and its complex64 copy (mom25d_core.cpp:1563-1580) ---- */
int orc_scale_c64(cd *z, long n, const double *dd, cf *o)
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
        cd *zr = z + (size_t)i * n; cf *orow = o + (size_t)i * n;
        for (long j = 0; j < n; ++j) { zr[j] *= dd[i] * dd[j]; orow[j] = (cf)zr[j]; }
    }
    return 0;
}

/* ---- G_q (centroid rule, extracted parts added) x area_t x area_f (mom25d_core.cpp:1585-1603) ---- */
int orc_fan_q(const orc_mesh *g, const orc_tables *tb, const int *f, int nf, cd *o)
{
#pragma omp parallel for schedule(static)
    for (int t = 0; t < g->Nt; ++t)
        for (int j = 0; j < nf; ++j) {
            cd ga, gq;
            tab_kern(tb, g->lt[t], g->lt[f[j]], cdist(g->cent, t, f[j]), &ga, &gq);
            o[(size_t)t * nf + j] = gq * (g->area[t] * g->area[f[j]]);
        }
    return 0;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
