/*
 * m25x.h - the C ABI of libm25x.so, the MI355X (gfx950) kernels of the per-frequency MoM solve.
 *
 * Plain pointers, sizes and a hipStream_t passed as void*; no torch or HIP types in the signatures. Every
 * pointer named "device" is HIP device memory; every function is asynchronous on the given stream unless it
 * says otherwise. Return value: 0, or a negative M25X_E* code with a message in m25x_last_error() (per thread).
 *
 * Each entry point names the reference interface it replaces (paths relative to the reference root). The
 * reference binds those through pybind11 (src/mom25d/core/mom25d_core.cpp:1658-1702) or through the GPU backend
 * contract (src/mom25d/gpu.py:28-48); the ctypes binding of this header is mom25d_amd/_lib.py.
 *
 * Complex numbers are interleaved (re, im): complex128 as two doubles, complex64 as two floats. Matrices are
 * row-major with a leading dimension (the reference's C-ordered Z), which is the column-major storage of Z^T:
 * the LU routines factor that column-major A = Z^T in place exactly as scipy.linalg.lu_factor(Z.T) lays it out.
 */
#ifndef M25X_H
#define M25X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define M25X_OK 0
#define M25X_EINVAL (-1)    /* bad argument (the reference's std::invalid_argument -> ValueError) */
#define M25X_EHIP (-2)      /* a HIP runtime error (RuntimeError) */
#define M25X_ENOMEM (-3)    /* device allocation failed (MemoryError) */
#define M25X_ESINGULAR (-4) /* the LU met an exactly zero pivot (LAPACK info > 0) */

#define M25X_C128 0
#define M25X_C64 1

/* Green's tables (device): AQ holds ntab x nfine points of (Re A, Im A, Re Q, Im Q); per table dr, d, dimg
 * (double) and cA, cq, cAi, cqi (complex128); tab (L x L int32: the table of an interface pair), kind (ntab int32:
 * 0 same interface, 1 cross extracted, 2 cross plain). The reference's tb dict (src/mom25d/fast.py:556-568). */
typedef struct m25x_tables {
    int ntab, nfine, L;
    const double *AQ;
    const double *dr, *d, *dimg;
    const double *cA, *cq, *cAi, *cqi;
    const int *tab, *kind;
} m25x_tables;

/* Mesh and RWG basis (device): the reference's g dict (src/mom25d/fast.py:389-398). ep/em: T+ and T- of each
 * edge; lt: triangle layer index; te: Nt x 3 edge of each triangle slot or -1; ta: Nt x 3 signed alpha; tv:
 * Nt x 3 x 2 free vertex; ap/am: signed alpha per edge; vp/vm: Ne x 2; cent: Nt x 2; area: Nt. */
typedef struct m25x_mesh {
    int Ne, Nt;
    const int *ep, *em, *lt, *te;
    const double *ap, *am, *vp, *vm, *cent, *area, *ta, *tv;
} m25x_mesh;

/* Quadrature geometry (device): the reference's q dict (src/mom25d/fast.py:399-400): nodes (nn x 2), tris
 * (Nt x 3 int32), P3 (Nt x 3 x 2), W3 (Nt x 3), P6 (Nt x 6 x 2), W6 (Nt x 6). */
typedef struct m25x_quad {
    const double *nodes;
    const int *tris;
    const double *P3, *W3, *P6, *W6;
} m25x_quad;

/* The layered stack at one angular frequency w (host struct, passed by value to the kernels): N dielectric layers
 * bottom to top (eps: complex permittivity er (1 - j tand) as (re, im); h: thickness), a PEC ground under layer 1
 * (bottom_pec = 1) or a half-space of eps_bot, a PEC cover on top of layer N (top_pec = 1) or a half-space of eps_top;
 * k0 = w / c0, wmu = w mu0, weps = w eps0. Interface k is the top of layer k (0: the bottom). The reference's
 * LayerStack (src/mom25d/multilayer.py:88-250) at one frequency. */
#define M25X_MAX_LAYERS 16
typedef struct m25x_stack {
    int N, bottom_pec, top_pec, pad;
    double eps[M25X_MAX_LAYERS][2];
    double h[M25X_MAX_LAYERS];
    double eps_bot[2], eps_top[2];
    double k0, wmu, weps;
} m25x_stack;

/* Spectral requests of m25x_spectra: M25X_SPEC_NI ints and M25X_SPEC_ND doubles each.
 *   ints: type, output row, then by type
 *     M25X_SPEC_TABLE  o, s, nx: G~_A and G~_q between interfaces o and s (two output rows) less nx (0..2) extracted
 *                      parts, doubles (d, Re cA, Im cA, Re cq, Im cq) per part: c / sqrt(rho^2 + d^2) taken out as
 *                      2 pi c exp(-kr d) / kr (src/mom25d/multilayer.py:253-318)
 *     M25X_SPEC_VIA_D  k, lo, hi: the via -> interface kernel of a chain from interface lo to hi seen on interface k
 *                      (src/mom25d/via.py:196-203)
 *     M25X_SPEC_VIA_R  m_lo, m_hi, n_lo, n_hi: the via -> via reaction remainder between two chains
 *                      (src/mom25d/via.py:208-213) */
#define M25X_SPEC_TABLE 0
#define M25X_SPEC_VIA_D 1
#define M25X_SPEC_VIA_R 2
#define M25X_SPEC_NI 8
#define M25X_SPEC_ND 10

const char *m25x_last_error(void);
int m25x_version(void);
/* Makes `device` current and caches its properties; called once per device before the others. */
int m25x_init(int device);

/* ---- fill ---------------------------------------------------------------------------------------------- */

/* Far-field (centroid-rule) entries of every edge pair into the Ne x Ne edge block of Z (row-major, leading
 * dimension ld, complex128 or complex64 per zdtype), symmetric. Replaces fill_far / fill_far_c64
 * (src/mom25d/core/mom25d_core.cpp:132-173). */
int m25x_fill_far(void *Z, int zdtype, long ld, const m25x_mesh *g, const m25x_tables *tb, double w, void *stream);

/* Centroid-rule entries of row x column edge lists into out (nr x nc complex128, row-major). Replaces far_block
 * (src/mom25d/core/mom25d_core.cpp:176-211). */
int m25x_far_block(const int *rows, int nr, const int *cols, int nc, const m25x_mesh *g, const m25x_tables *tb,
                   double w, double *out, void *stream);

/* The seven accurate-minus-centroid moments (M0, Mox, Moy, Msx, Msy, Mrr, Mq; complex128) of each near pair into
 * M (n x 7 complex128) and Mq alone into qcorr (n complex128, may be NULL). stat (n x 12 doubles, may be NULL):
 * the pairs' analytic (extracted-part) moments, written when stat_ready is 0 and read when it is 1. The
 * per-pair half of near_core (src/mom25d/core/mom25d_core.cpp:417-476). */
int m25x_near_moments(const m25x_mesh *g, const m25x_tables *tb, const m25x_quad *q, const int *P, const int *Q,
                      long n, double w, double *stat, int stat_ready, double *M, double *qcorr, void *stream);

/* Z[erow[e], ecol[e]] += sum over the contributions c of entry e (contrib[estart[e] .. estart[e+1]), each
 * 9 * pair + 3 * s + u) of the pair's moments combined for its edge slots s, u; deterministic, no atomics. The
 * scatter half of near_into_dense (src/mom25d/core/mom25d_core.cpp:477-507, 520-537). */
int m25x_near_scatter(void *Z, int zdtype, long ld, const m25x_mesh *g, const int *P, const int *Q,
                      const double *M, long nentry, const int *erow, const int *ecol, const int *estart,
                      const int *contrib, double w, void *stream);

/* 9 COO triplets a pair (ri, ci int32; vv complex128), row -1 for unused slots. Replaces near_coo's output
 * (src/mom25d/core/mom25d_core.cpp:559-575) from the moments of m25x_near_moments. */
int m25x_near_coo(const m25x_mesh *g, const int *P, const int *Q, const double *M, long n, double w, int *ri,
                  int *ci, double *vv, void *stream);

/* ---- matrix-free products (the refinement's FP64 operator without Z) ---------------------------------- */

/* Workspace bytes of m25x_matvec_far for Nt triangles and k columns. */
size_t m25x_matvec_far_work_bytes(int Nt, int k);

/* Y = Z_far X for the Ne x Ne edge block's centroid-rule far field, without forming Z: the sources aggregated per
 * triangle, one kernel evaluation per triangle pair for all columns. X, Y: Ne x k complex128 row-major (leading
 * dimensions ldx, ldy >= k), 1 <= k <= 8. Deterministic (fixed summation orders). Replaces matvec_far and
 * matvec_far_multi (src/mom25d/core/mom25d_core.cpp:214-314). */
int m25x_matvec_far(const double *X, int k, long ldx, const m25x_mesh *g, const m25x_tables *tb, double w, double *Y,
                    long ldy, void *work, size_t work_bytes, void *stream);

/* vals[e] (complex128) = the near-field correction of near entry e, its contributions (as for m25x_near_scatter)
 * summed in their fixed order: with the entries' rows and columns, the correction as a sparse matrix (the
 * reference's dZ_exact, src/mom25d/fast.py:571-586). */
int m25x_near_values(const m25x_mesh *g, const int *P, const int *Q, const double *M, long nentry, const int *estart,
                     const int *contrib, double w, double *vals, void *stream);

/* Y (+)= A X for a CSR matrix (rowptr: nrow + 1 int32, col int32, vals complex128), X and Y complex128 row-major
 * with k columns (add = 1 adds to Y, 0 overwrites it). With m25x_near_values' entries: near_matvec
 * (src/mom25d/core/mom25d_core.cpp:539-556), deterministic (one thread a row and column, the row's entries in
 * order). */
int m25x_csr_matvec(const int *rowptr, const int *col, const double *vals, long nrow, const double *X, int k, long ldx,
                    double *Y, long ldy, int add, void *stream);

/* Near triangle pairs rho^2 + dz^2 < (fac max(size_p, size_q))^2, sorted by (p, q). Two calls: with P == NULL
 * the count is written to *count (synchronous); then the pairs (int32) are written. Replaces near_pairs
 * (src/mom25d/core/mom25d_core.cpp:1216-1263). counts: Nt + 1 int64 of device scratch. */
int m25x_near_pairs(const double *cent, const double *size, const double *z, int Nt, double fac, long long *counts,
                    long *count, int *P, int *Q, void *stream);

/* G_q x area_t x area_f between every triangle t and the fan triangles F (Nt x nf complex128). Replaces fan_q
 * (src/mom25d/core/mom25d_core.cpp:1585-1603). */
int m25x_fan_q(const m25x_mesh *g, const m25x_tables *tb, const int *F, int nf, double *out, void *stream);

/* out[t, v] = sum_q D(|P[t, q] - V[v]|) W[t, q] over a radial table D (n complex128, step dr), linear lookup.
 * Replaces radial_table_sum (src/mom25d/core/mom25d_core.cpp:1269-1298). */
int m25x_radial_table_sum(const double *px, const double *py, const double *W, long nt, long nq, const double *vx,
                          const double *vy, long nv, const double *table, long n, double dr, double *out,
                          void *stream);

/* out[s, i] = sum_k J0(rho_i kr_k) W[s, k] over a Sommerfeld contour, kr complex (nk complex128), W (ns x nk
 * complex128), out (ns x nr complex128): real nodes through the real J0, lifted ones through the complex series
 * / asymptotic expansion. Replaces hankel_real + the lifted part of j0_sum (src/mom25d/core/mom25d_core.cpp:
 * 1430-1454; src/mom25d/via.py:90-101). */
int m25x_hankel(const double *rho, int nr, const double *kr, int nk, const double *W, int ns, double *out,
                void *stream);

/* The spectra of nreq requests at the npts nodes of a Sommerfeld contour (kr, dk: npts complex128 nodes and
 * weights, device), each times kr dk (the Hankel weights of m25x_hankel): W (nrows x npts complex128, device), a
 * request's values in its output row(s). req_i / req_d: the requests on the device; req_i_host: the same ints on the
 * host (checked before the launch). The stack's transmission-line network in the reflection-coefficient form,
 * overflow-free on the lifted contour. Replaces LayerStack._network / voltage / spectral, TMNetwork and the spectra
 * of GreenTableML and ViaKernels (src/mom25d/multilayer.py:183-250, 253-308; src/mom25d/via.py:127-213). */
int m25x_spectra(const m25x_stack *st, double w, const double *kr, const double *dk, int npts, const int *req_i,
                 const double *req_d, int nreq, const int *req_i_host, double *W, long nrows, void *stream);

/* G~_A and G~_q between interfaces o and s at one complex kr, on the host (synchronous): out = (Re G~_A, Im G~_A,
 * Re G~_q, Im G~_q). The extraction coefficients' single evaluations (src/mom25d/multilayer.py:279-282, 288-291). */
int m25x_spectral_point(const m25x_stack *st, double w, double kr_re, double kr_im, int o, int s, double *out);

/* J_n(z), complex z, n = 0, 1, 2, elementwise. Replaces bessel_jn_complex
 * (src/mom25d/core/mom25d_core.cpp:1510-1558). */
int m25x_bessel_jn_complex(int n, const double *z, long total, double *out, void *stream);

/* Z <- diag(d) Z diag(d) in place (n x n complex128, leading dimension ld) and, when Z64 is not NULL, its
 * complex64 copy (leading dimension ld64). Replaces scale_c64 and _equilibrate
 * (src/mom25d/core/mom25d_core.cpp:1563-1580; src/mom25d/fast.py:294-305). */
int m25x_scale_c64(double *Z, long n, long ld, const double *d, float *Z64, long ld64, void *stream);

/* d[i] = 1 / sqrt(|Z_ii| + 1e-300) (the equilibration scales, src/mom25d/fast.py:799). */
int m25x_diag_scales(const double *Z, long n, long ld, double *d, void *stream);

/* ---- dense complex LU ---------------------------------------------------------------------------------- */

/* Device workspace bytes m25x_getrf needs for an n x n matrix of the given dtype (either pivoting). */
size_t m25x_getrf_work_bytes(int dtype, long n);

#define M25X_PIVOT_PARTIAL 0      /* LAPACK's partial pivoting: the pivots of zgetrf / cgetrf exactly */
#define M25X_PIVOT_TOURNAMENT 1   /* tournament pivoting (CALU): a panel's pivots by a tree of GEPPs */

/* LU of the column-major n x n matrix A (leading dimension lda; the row-major Z seen as Z^T), in place: L unit
 * lower, U upper, ipiv 0-based LAPACK-style interchanges (int32, device), the layout of
 * scipy.linalg.lu_factor(Z.T). complex128: trailing updates on the FP64 matrix cores. complex64: with partial
 * pivoting on the FP32 matrix cores (FP32 accumulation); with tournament pivoting on the FP64 matrix cores (one
 * rounding to FP32 an entry). pivot: M25X_PIVOT_PARTIAL or M25X_PIVOT_TOURNAMENT. perm (device int[n], may be
 * NULL): the interchanges as one permutation, m25x_ipiv_perm's. info (device int, may be NULL): 0, 1 + the first
 * zero pivot, or -1 when a bounded wait between the workgroups of a partial-pivoting panel ran out (the factor is
 * then not valid; the tournament has no such wait). Replaces the GPU backend's factor (src/mom25d/gpu.py:69-74;
 * src/mom25d/gpu_vk.py:515-633) and LAPACK zgetrf/cgetrf (src/mom25d/fast.py:318-320). */
int m25x_getrf(int dtype, int pivot, void *A, long n, long lda, int *ipiv, int *perm, int *info, void *work,
               size_t work_bytes, void *stream);

/* The interchanges of ipiv as one permutation (perm = ipiv's swaps applied in order to 0..n-1), once per factor:
 * P^T b = b[perm], P z scatters z[i] to perm[i]. */
int m25x_ipiv_perm(const int *ipiv, long n, int *perm, void *stream);

/* Device workspace bytes m25x_getrs needs. */
size_t m25x_getrs_work_bytes(int dtype, long n, long ldb);

/* Solves with the factor of m25x_getrf: trans = 1: Z X = B (A^T X = B, lu_solve(..., trans=1)), trans = 0:
 * A X = B. B: n x nrhs row-major (leading dimension ldb >= nrhs, 1 <= nrhs <= 8), the factor's dtype, overwritten
 * with X. perm: m25x_getrf's (or m25x_ipiv_perm's, or NULL: computed here). Diagonal blocks by substitution, as
 * LAPACK's: the accuracy of a factor with small pivots is kept. Every wait between the solve's workgroups is
 * bounded: the int at work + (work_bytes - 256, rounded down to a multiple of 4), with the work_bytes passed, is
 * nonzero after a call in which one ran out (X is then not the solution), zero otherwise; it is zeroed once per
 * call, so it covers every right-hand side of the call. Replaces the backend's solve (src/mom25d/gpu.py:76-80; src/mom25d/gpu_vk.py:650-678)
 * and zgetrs / cgetrs (src/mom25d/fast.py:356-358). */
int m25x_getrs(int dtype, const void *LU, long n, long lda, const int *ipiv, const int *perm, void *B, long ldb,
               int nrhs, int trans, void *work, size_t work_bytes, void *stream);

/* The solve of m25x_getrs (the same kernels, the same results) with its workspace in a device array of this
 * library instead of a caller's buffer: it allocates nothing (no hipMalloc, no allocating library call), for
 * callers that cannot allocate - code under HIP graph capture, serving loops and harnesses that forbid it. No
 * counterpart in the reference. The array holds M25X_GETRS_SCRATCH_BYTES on each device the library runs on,
 * until the process ends; n is limited to what fits (m25x_getrs_work_bytes(dtype, n, ldb) - 256 bytes: about
 * n <= 41 000 complex128, 82 000 complex64 at ldb = 8), larger n is rejected with M25X_EINVAL. status (device int):
 * zeroed by the call, nonzero afterwards when a wait between the solve's workgroups ran out (X is then not the
 * solution). Every call on one device shares that device's array: calls that overlap on one device (two streams,
 * two threads) race and give wrong results; serialise them. */
#define M25X_GETRS_SCRATCH_BYTES (16u << 20)
int m25x_getrs_noalloc(int dtype, const void *LU, long n, long lda, const int *ipiv, const int *perm, void *B,
                       long ldb, int nrhs, int trans, int *status, void *stream);

/* The K-stage depth of the tournament LU's trailing update on the matrix cores (8, 16 or 32 columns of L21 and rows
 * of U12 staged in LDS between two barriers), for every later m25x_getrf / m25x_getrf_batched in the process (8 until
 * set); the results do not depend on it except by rounding order within a stage's sums. No counterpart in the
 * reference: the knob of this build's autotuner (mom25d_amd/tune.py), which times the choices on the device. */
int m25x_lu_set_gemm_stage(int ks);
int m25x_lu_gemm_stage(void);

/* Look-ahead of the tournament LU (m25x_getrf, one matrix), 1 on, 0 off (the default): the next panel's tournament
 * runs on a second stream of the library's (one per device, the default priority) beside the rest of the trailing
 * update on the caller's stream. The factor is the same bit for bit either way (the same kernels; each update
 * split by columns). No counterpart in the reference (its LU is LAPACK's or a vendor library's). */
int m25x_lu_set_lookahead(int on);
int m25x_lu_lookahead(void);

/* The matrix cores of the complex64 tournament LU's trailing update, for every later call in the process: 1 (the
 * default) FP32 with FP32 accumulation (twice the FP64 peak rate), 0 FP64 (products and sums in FP64, one rounding
 * to FP32 an update: a smaller backward error). The GPU backend's complex64 factor (src/mom25d/gpu.py:69-74) runs
 * complex FP32 throughout; this selects how much of it. */
int m25x_lu_set_update_precision(int fp32);

/* ---- batches: many systems in one launch per step ------------------------------------------------------
 * No counterpart in the reference, whose solver takes one frequency at a time (src/mom25d/run.py:55-64 calls
 * Simulation.sparams once a point; src/mom25d/fast.py:704-850 factors one Z). These run the same kernels as
 * m25x_getrf / m25x_getrs / m25x_zgemm_nk over a list of independent systems of any sizes, each launch covering
 * every system (a grid dimension over the items), so that small systems fill the GPU together and a sweep step is
 * a fixed sequence of launches (capturable in one HIP graph). Each item's results are those of the one-system call
 * on the same inputs. items_host: the item table on the host (sizes for the grids, checked); items_dev: the same
 * table in device memory. */

/* One LU of the batch: as m25x_getrf's arguments (tournament pivoting), work: m25x_getrf_work_bytes(dtype, n) bytes
 * of device memory of its own (work_bytes: its size). */
typedef struct m25x_lu_item {
    void *A;
    long n, lda;
    int *ipiv, *perm, *info;
    void *work;
    size_t work_bytes;
} m25x_lu_item;

/* The tournament-pivoting LUs of nitems matrices of the given dtype in one sequence of launches (each panel step
 * one launch per kernel for all items). Items sorted by n, largest first, make the later steps' grids smaller
 * (any order is correct). */
int m25x_getrf_batched(int dtype, int nitems, const m25x_lu_item *items_host, const m25x_lu_item *items_dev,
                       void *stream);

/* One solve of the batch: as m25x_getrs's arguments (perm required: m25x_getrf's), nrhs <= 4 (complex128) or 8
 * (complex64), work: m25x_getrs_work_bytes(dtype, n, nrhs) bytes of its own. */
typedef struct m25x_rs_item {
    const void *LU;
    long n, lda;
    const int *perm;
    void *B;
    long ldb;
    int nrhs, pad;
    void *work;
    size_t work_bytes;
} m25x_rs_item;

/* The solves of nitems factors (trans as m25x_getrs), each triangle one launch for all items: the chains of the
 * items' diagonal blocks side by side in one grid. status (device int): zeroed, then nonzero if a wait between
 * workgroups ran out in any item. */
int m25x_getrs_batched(int dtype, int nitems, const m25x_rs_item *items_host, const m25x_rs_item *items_dev, int trans,
                       int *status, void *stream);

/* One product of the batch: Y = Z X, Z n x n complex128 row-major (leading dimension ld), X and Y n x k (k <= 8)
 * with leading dimensions ldx, ldy. */
typedef struct m25x_mv_item {
    const void *Z;
    long n, ld;
    const void *X;
    void *Y;
    long ldx, ldy;
    int k, pad;
} m25x_mv_item;

/* The products of nitems items in one launch. */
int m25x_zgemm_nk_batched(int nitems, const m25x_mv_item *items_host, const m25x_mv_item *items_dev, void *stream);

/* ---- launch timing (benchmarks) ------------------------------------------------------------------------------
 * No counterpart in the reference (its fill and LU are CPU code, LAPACK or a vendor library). m25x_timing(mask)
 * opens a record (clearing the last) in which every launch of a kernel whose bit (1 << M25X_K_*) is in mask is
 * bracketed by two HIP events on its stream; m25x_timing(0) closes it. m25x_timing_records waits for the recorded
 * launches and writes, for up to cap of them in launch order, the kernel, the element type (M25X_C128 / M25X_C64),
 * the duration (ms), and the bytes and flops the launch moves and does by algorithm (0 where the kernel has no
 * such count); it returns the number recorded, or a negative error code. Thread-safe; one record per process. */
#define M25X_K_FILL_FAR 0       /* 16 B (8 B) per Z entry written */
#define M25X_K_NEAR_MOMENTS 1   /* per near pair: 7 complex moments out */
#define M25X_K_NEAR_SCATTER 2   /* per near entry: read-modify-write of Z, its contributions read */
#define M25X_K_TSLU 3           /* a 64-column panel's tournament: the panel read once; flops: its GEPPs */
#define M25X_K_APPLY_MOVES 4    /* the panel's row moves: 2 x element per moved element */
#define M25X_K_L21 5            /* L21 = A21 U11^-1: 8 m 64^2 / 2 flops */
#define M25X_K_TRSM 6           /* U12 = L11^-1 A12: 8 n 64^2 / 2 flops */
#define M25X_K_GEMM 7           /* trailing update: 8 M N K flops; bytes: C read and written, A and B read */
#define M25X_K_TRSV 8           /* a triangular solve: its triangle of the factor and the right-hand sides */
#define M25X_K_ZGEMM_NK 9       /* Y = Z X: 16 N^2 B of Z read */
#define M25X_K_CGS2 10          /* one Arnoldi step's orthogonalisation */
#define M25X_K_PANEL 11         /* a partial-pivoting panel */
#define M25X_K_LASWP 12         /* its row interchanges */
int m25x_timing(unsigned kernels);
long m25x_timing_records(int *kernel, int *dtype, float *ms, double *bytes, double *flops, long cap);
/* The kernel's name ("" for an unknown id). */
const char *m25x_kernel_name(int kernel);

/* ---- refinement products ------------------------------------------------------------------------------- */

/* Y = Z X: Z n x n complex128 row-major (leading dimension ld), X and Y n x k complex128 row-major (1 <= k <= 8).
 * The FP64 products of the GMRES refinement (src/mom25d/fast.py:1004-1007, A = Z @ V). */
int m25x_zgemm_nk(const double *Z, long n, long ld, const double *X, int k, double *Y, void *stream);

/* Workspace bytes of m25x_cgs2. */
size_t m25x_cgs2_work_bytes(void);

/* One Arnoldi step of the batched GMRES, per column c of the k <= 8 systems: classical Gram-Schmidt applied twice
 * to w (n x k complex128 row-major, overwritten) against the nv basis vectors V (nv x n x k, vector i at
 * V + 2 i n k doubles): hcol[i][c] = h1 + h2 for i < nv, hcol[nv][c] = |w| (real), vnext = w / |w|. Sums run over
 * fixed row blocks and then in block order (deterministic). Replaces the orthogonalisation of the reference's
 * gmres_batched (src/mom25d/fast.py:1437-1500, modified Gram-Schmidt: the same basis to rounding). */
int m25x_cgs2(const double *V, long n, int k, int nv, double *w, double *hcol, double *vnext, void *work,
              size_t work_bytes, void *stream);

#ifdef __cplusplus
}
#endif
#endif
