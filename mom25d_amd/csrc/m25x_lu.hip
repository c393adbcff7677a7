// m25x_lu.hip - dense complex LU with partial pivoting on gfx950, and the triangular solves with its factor.
//
// A is column-major (the row-major Z of the reference seen as A = Z^T, the layout of scipy's lu_factor(Z.T)).
// Blocked right-looking. Two pivoting strategies:
// Partial pivoting (LAPACK's pivots exactly), panels of NB columns:
//   panel   one cooperative launch per panel: each workgroup keeps a slab of the panel's rows in LDS; every column
//           costs one grid barrier (pivot search, row interchange, scaling and the panel's rank-1 update in
//           between, the next column's candidates published before the barrier); pivots by LAPACK's rule
//           (largest |Re| + |Im|, the first on ties)
//   laswp   the panel's interchanges on the columns left and right of it
//   trsm    U12 = L11^-1 A12 (unit lower, forward substitution, one workgroup per 16 columns)
//   gemm    A22 -= A21 U12 on the matrix cores: complex products as real ones through the real embedding
//           (Re C = [Re A, Im A] [Re B; -Im B], Im C = [Re A, Im A] [Im B; Re B]), v_mfma_f64_16x16x4 for
//           complex128 and v_mfma_f32_16x16x4 for complex64 (FP32 accumulation)
// Tournament pivoting (CALU), panels of TW = 64 columns: the panel's pivots by a tree of GEPPs in one launch with no
// grid barrier (k_tslu), its row moves on every column (k_apply_moves), then L21 = A21 U11^-1, U12 = L11^-1 A12 and
// A22 -= L21 U12, all three on the matrix cores.
// The solves (getrs) run one workgroup per 64-row block with ready flags between blocks (no launch per block), each
// block's diagonal solve a product with its inverse (k_trtri_diag, once per factor).
#include <algorithm>
#include <mutex>
#include "m25x_common.h"

using namespace m25x;

namespace {

typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

// ---- complex traits ----------------------------------------------------------------------------------------
template <typename T> struct CX;
template <> struct CX<double2> {
    typedef double R;
    typedef d4v acc4;
    __device__ static double2 make(double x, double y) { return make_double2(x, y); }
    __device__ static acc4 mfma(double a, double b, acc4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    // the output row of lane group kk (lane / 16), register r (measured on gfx950: kk + 4 r for the f64 form)
    __device__ static int out_row(int kk, int r) { return kk + 4 * r; }
};
template <> struct CX<float2> {
    typedef float R;
    typedef f4v acc4;
    __device__ static float2 make(float x, float y) { return make_float2(x, y); }
    __device__ static acc4 mfma(float a, float b, acc4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    // (4 kk + r for the f32 form)
    __device__ static int out_row(int kk, int r) { return 4 * kk + r; }
};

template <typename T> __device__ inline T tadd(T a, T b) { return CX<T>::make(a.x + b.x, a.y + b.y); }
template <typename T> __device__ inline T tsub(T a, T b) { return CX<T>::make(a.x - b.x, a.y - b.y); }
template <typename T> __device__ inline T tmul(T a, T b)
{
    return CX<T>::make(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// reciprocal by Smith's algorithm
template <typename T> __device__ inline T trecip(T b)
{
    typedef typename CX<T>::R R;
    if (fabs((double)b.x) >= fabs((double)b.y)) {
        R r = b.y / b.x, den = b.x + b.y * r;
        return CX<T>::make((R)1 / den, -r / den);
    }
    R r = b.x / b.y, den = b.x * r + b.y;
    return CX<T>::make(r / den, (R)-1 / den);
}
template <typename T> __device__ inline T tdiv(T a, T b)
{
    typedef typename CX<T>::R R;
    if (fabs((double)b.x) >= fabs((double)b.y)) {
        R r = b.y / b.x, den = b.x + b.y * r;
        return CX<T>::make((a.x + a.y * r) / den, (a.y - a.x * r) / den);
    }
    R r = b.x / b.y, den = b.x * r + b.y;
    return CX<T>::make((a.x * r + a.y) / den, (a.y * r - a.x) / den);
}
template <typename T> __device__ inline double cabs1(T a) { return fabs((double)a.x) + fabs((double)a.y); }

// ---- grid barrier (all workgroups of a cooperative launch co-resident) -------------------------------------
struct Bar {
    unsigned count, gen;
    int abort, pad;
};

// Sense by generation; release / acquire at agent scope orders the slot writes and reads across the XCDs.
// A bounded wait: past the limit the barrier sets abort and every later barrier returns at once, so the
// launch always drains (the host then reports the failure).
__device__ inline void grid_sync(Bar *b, unsigned nblocks, unsigned &gen)
{
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        if (!__hip_atomic_load(&b->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            const unsigned arrived = __hip_atomic_fetch_add(&b->count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            if (arrived == nblocks - 1) {
                __hip_atomic_store(&b->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_fetch_add(&b->gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                long spins = 0;
                while (__hip_atomic_load(&b->gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                    __builtin_amdgcn_s_sleep(1);
                    if (++spins > (1L << 26)) {
                        __hip_atomic_store(&b->abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                }
            }
        }
        gen += 1;
    }
    __syncthreads();
}

__global__ void k_zero_ints(int *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ---- panel --------------------------------------------------------------------------------------------------
constexpr int NB = 128;          // panel width
constexpr int PT = 256;          // panel threads

template <typename T> struct Slot {
    double val;
    int row, pad;
    T v[NB];
};

// argmax of |Re| + |Im| over (val, row) pairs: larger value, then smaller row
__device__ inline void better(double &v, int &r, double v2, int r2)
{
    if (v2 > v || (v2 == v && r2 < r && r2 >= 0) || (r < 0 && r2 >= 0)) { v = v2; r = r2; }
}

__device__ inline void block_argmax(double &v, int &r, double *sv, int *sr)
{
    for (int o = 32; o > 0; o >>= 1) {
        const double v2 = __shfl_xor(v, o);
        const int r2 = __shfl_xor(r, o);
        better(v, r, v2, r2);
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) { sv[wv] = v; sr[wv] = r; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < PT / 64; ++k) better(sv[0], sr[0], sv[k], sr[k]);
    }
    __syncthreads();
    v = sv[0]; r = sr[0];
    __syncthreads();
}

// Slab of workgroup g: global rows [r0, r0 + rn), panel columns [j0, j0 + jb), column-major in LDS (S[c * R + r]).
template <typename T>
__global__ __launch_bounds__(PT) void k_panel(T *A, long lda, long n, long j0, int jb, int R, int *ipiv, int *info,
                                              Slot<T> *slots, T *diag, Bar *bar)
{
    extern __shared__ char smem_raw[];
    T *S = (T *)smem_raw;
    __shared__ T urow[NB], drow[NB];
    __shared__ double sv[PT / 64];
    __shared__ int sr[PT / 64];
    __shared__ int piv_row;
    const unsigned G = gridDim.x;
    const int g = blockIdx.x, tid = threadIdx.x;
    const long r0 = j0 + (long)g * R;
    const int rn = (int)std::max(0L, std::min((long)R, n - r0));
    unsigned gen = 0;
    for (int idx = tid; idx < rn * jb; idx += PT) {
        const int c = idx / rn, r = idx % rn;
        S[c * R + r] = A[(j0 + c) * lda + r0 + r];
    }
    __syncthreads();
    // candidates of column k: rows >= j0 + k of my slab; publish (val, row, row values) in slots[buf][g], and the
    // diagonal row j0 + k in diag[buf] when it is mine
    auto publish = [&](int k, int buf) {
        double v = -1.0;
        int rr = -1;
        const long first = j0 + k;
        for (int r = tid; r < rn; r += PT)
            if (r0 + r >= first) better(v, rr, cabs1(S[k * R + r]), (int)(r0 + r));
        block_argmax(v, rr, sv, sr);
        Slot<T> *sl = slots + (size_t)buf * G + g;
        if (tid == 0) { sl->val = v; sl->row = rr; }
        if (rr >= 0)
            for (int c = tid; c < jb; c += PT) sl->v[c] = S[c * R + (rr - r0)];
        if (first >= r0 && first < r0 + rn)
            for (int c = tid; c < jb; c += PT) diag[(size_t)buf * NB + c] = S[c * R + (first - r0)];
    };
    publish(0, 0);
    grid_sync(bar, G, gen);
    for (int k = 0; k < jb; ++k) {
        const int buf = k & 1;
        const long dr = j0 + k;
        // the pivot: the best of the published candidates (every workgroup reads the same ones)
        if (tid < 64) {
            double v = -1.0;
            int rr = -1;
            for (unsigned q = tid; q < G; q += 64) {
                const Slot<T> *sl = slots + (size_t)buf * G + q;
                better(v, rr, sl->val, sl->row);
            }
            for (int o = 32; o > 0; o >>= 1) {
                const double v2 = __shfl_xor(v, o);
                const int r2 = __shfl_xor(rr, o);
                better(v, rr, v2, r2);
            }
            if (tid == 0) piv_row = rr < 0 ? (int)dr : rr;
        }
        __syncthreads();
        const int p = piv_row;
        const int gp = (int)((p - j0) / R);
        for (int c = tid; c < jb; c += PT) {
            urow[c] = slots[(size_t)buf * G + gp].v[c];
            drow[c] = diag[(size_t)buf * NB + c];
        }
        if (g == 0 && tid == 0) ipiv[dr] = p;
        __syncthreads();
        if (p != dr) {
            if (p >= r0 && p < r0 + rn)
                for (int c = tid; c < jb; c += PT) S[c * R + (p - r0)] = drow[c];
            if (dr >= r0 && dr < r0 + rn)
                for (int c = tid; c < jb; c += PT) S[c * R + (dr - r0)] = urow[c];
        }
        __syncthreads();
        const T pv = urow[k];
        const bool zero = pv.x == 0 && pv.y == 0;
        if (zero && g == 0 && tid == 0 && info && *info == 0) *info = (int)dr + 1;
        const T inv = zero ? CX<T>::make(0, 0) : trecip(pv);
        // multipliers, then the rank-1 update of the rest of the slab's panel columns
        const int rbeg = (int)std::max(0L, std::min((long)rn, dr + 1 - r0));
        if (!zero)
            for (int r = rbeg + tid; r < rn; r += PT) S[k * R + r] = tmul(S[k * R + r], inv);
        __syncthreads();
        const int nr = rn - rbeg, ncol = jb - k - 1;
        for (int idx = tid; idx < nr * ncol; idx += PT) {
            const int c = k + 1 + idx / nr, r = rbeg + idx % nr;
            S[c * R + r] = tsub(S[c * R + r], tmul(S[k * R + r], urow[c]));
        }
        __syncthreads();
        if (k + 1 < jb) {
            publish(k + 1, buf ^ 1);
            grid_sync(bar, G, gen);
        }
    }
    for (int idx = tid; idx < rn * jb; idx += PT) {
        const int c = idx / rn, r = idx % rn;
        A[(j0 + c) * lda + r0 + r] = S[c * R + r];
    }
}

// the panel's interchanges on columns [0, j0) and [j0 + jb, n): one thread a column, in order
template <typename T>
__global__ void k_laswp(T *A, long lda, long n, long j0, int jb, const int *ipiv)
{
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const long nleft = j0, nright = n - j0 - jb;
    if (t >= nleft + nright) return;
    const long c = t < nleft ? t : j0 + jb + (t - nleft);
    T *col = A + c * lda;
    for (int k = 0; k < jb; ++k) {
        const long r = j0 + k, p = ipiv[r];
        if (p != r) {
            const T a = col[r];
            col[r] = col[p];
            col[p] = a;
        }
    }
}

// U12 = L11^-1 A12: rows [j0, j0 + jb), columns [c0, n); 32 columns a workgroup, the block in LDS
constexpr int TC = 16;
template <typename T>
__global__ __launch_bounds__(256) void k_trsm(T *A, long lda, long n, long j0, int jb, long c0)
{
    __shared__ T X[NB][TC + 1];
    const long cb = c0 + (long)blockIdx.x * TC;
    const int nc = (int)std::min((long)TC, n - cb);
    const int tid = threadIdx.x;
    for (int idx = tid; idx < jb * TC; idx += 256) {
        const int c = idx / jb, r = idx % jb;
        if (c < nc) X[r][c] = A[(cb + c) * lda + j0 + r];
    }
    __syncthreads();
    const T *L = A + j0 * lda + j0;          // L11[i][k] = L[k * lda + i]
    for (int k = 0; k < jb - 1; ++k) {
        const int nr = jb - k - 1;
        for (int idx = tid; idx < nr * TC; idx += 256) {
            const int c = idx % TC, i = k + 1 + idx / TC;
            if (c < nc) X[i][c] = tsub(X[i][c], tmul(L[(long)k * lda + i], X[k][c]));
        }
        __syncthreads();
    }
    for (int idx = tid; idx < jb * TC; idx += 256) {
        const int c = idx / jb, r = idx % jb;
        if (c < nc) A[(cb + c) * lda + j0 + r] = X[r][c];
    }
}

// ---- C -= A B (SUB) or C = A B on the matrix cores -------------------------------------------------------
// C (M x N), A (M x K), B (K x N), column-major, complex. C = A B may overwrite A or B in place when a workgroup
// tile's output depends on nothing another workgroup writes (M <= 64 with C = B, or N <= 64 with C = A): every
// workgroup reads all of its K range before it stores. Workgroup tile 64 x 64, four waves of 32 x 32, each
// 2 x 2 blocks of 16 x 16 with a real and an imaginary accumulator. The MFMA's first operand is B (its rows are
// C's columns), its second A, so a lane's outputs run down C's columns (coalesced stores).
constexpr int GM = 64, GN = 64, GK = 8;

template <typename T, bool SUB>
__global__ __launch_bounds__(256) void k_gemm(T *C, long ldc, const T *A, long lda, const T *B, long ldb, long M,
                                              long N, long K)
{
    typedef typename CX<T>::R R;
    typedef typename CX<T>::acc4 acc4;
    __shared__ T As[GK][GM];         // As[p][i] = A[i][p]
    __shared__ T Bs[GK][GN + 1];     // Bs[p][j] = B[p][j]
    const long m0 = (long)blockIdx.x * GM, n0 = (long)blockIdx.y * GN;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wm = (wv & 1) * 32, wn = (wv >> 1) * 32;     // wave tile origin in the workgroup tile
    const int li = lane & 15, kk = lane >> 4, part = kk & 1, ph = kk >> 1;
    acc4 cr[2][2], ci[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) { cr[a][b][r] = 0; ci[a][b][r] = 0; }
    const T zero = CX<T>::make(0, 0);
    for (long k0 = 0; k0 < K; k0 += GK) {
        // stage A (GM x GK) and B (GK x GN); zero outside
        for (int idx = tid; idx < GM * GK; idx += 256) {
            const int i = idx % GM, p = idx / GM;
            const long gi = m0 + i, gp = k0 + p;
            As[p][i] = (gi < M && gp < K) ? A[gp * lda + gi] : zero;
        }
        for (int idx = tid; idx < GK * GN; idx += 256) {
            const int p = idx % GK, j = idx / GK;
            const long gp = k0 + p, gj = n0 + j;
            Bs[p][j] = (gp < K && gj < N) ? B[gj * ldb + gp] : zero;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < GK / 2; ++ks) {
            const int p = 2 * ks + ph;
            R av[2], br[2], bi[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const T x = As[p][wm + a * 16 + li];
                av[a] = part ? x.y : x.x;
            }
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const T y = Bs[p][wn + b * 16 + li];
                br[b] = part ? -y.y : y.x;
                bi[b] = part ? y.x : y.y;
            }
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int a = 0; a < 2; ++a) {
                    cr[b][a] = CX<T>::mfma(br[b], av[a], cr[b][a]);
                    ci[b][a] = CX<T>::mfma(bi[b], av[a], ci[b][a]);
                }
        }
        __syncthreads();
    }
    // lane (li, kk), register r: C row m0 + wm + a*16 + li, column n0 + wn + b*16 + out_row(kk, r)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const long gi = m0 + wm + a * 16 + li, gj = n0 + wn + b * 16 + CX<T>::out_row(kk, r);
                if (gi < M && gj < N) {
                    T *c = C + gj * ldc + gi;
                    if (SUB) {
                        const T v = *c;
                        *c = CX<T>::make(v.x - cr[b][a][r], v.y - ci[b][a][r]);
                    } else {
                        *c = CX<T>::make(cr[b][a][r], ci[b][a][r]);
                    }
                }
            }
}

// ---- the interchanges as a permutation: perm = the swaps of ipiv applied in order to 0..n-1 ------------------
__global__ void k_perm(const int *ipiv, long n, int *perm)
{
    // one workgroup: identity in parallel, then the swaps in order by one thread, ipiv read 8 at a time (the
    // loads are independent of the swaps; most pivots are on the diagonal)
    for (long i = threadIdx.x; i < n; i += blockDim.x) perm[i] = (int)i;
    __threadfence_block();
    __syncthreads();
    if (threadIdx.x == 0) {
        for (long i0 = 0; i0 < n; i0 += 8) {
            int pv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) pv[u] = i0 + u < n ? ipiv[i0 + u] : (int)(i0 + u);
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const long i = i0 + u;
                const int p = pv[u];
                if (i < n && p != i) {
                    const int t = perm[i];
                    perm[i] = perm[p];
                    perm[p] = t;
                }
            }
        }
    }
}

// ---- triangular solves, one workgroup a 64-row block, ready flags between blocks ---------------------------
// B: n x nrhs row-major (ldb). The blocks of the solution are published in order; block b adds the
// contributions of the blocks before it as their flags come up, then multiplies by the inverse of its diagonal
// block (inverted once per factor by k_trtri_diag): no serial substitution on the critical path.
constexpr int SB = 64;
constexpr int MAXRHS = 8;

// The inverse of the unit lower (which 0) or upper (which 1) triangle of the bn x bn block at LU (column-major,
// leading dimension lda), padded to 64 x 64 with the identity; out[i * SB + j] (row-major) or out[j * SB + i]
// (column-major). 64 threads, one column of the inverse each.
template <typename T>
__device__ void trtri64(const T *LU, long lda, int bn, int which, T (*M)[SB + 1], T (*X)[SB + 1], T *out,
                        bool colmajor)
{
    const int j = threadIdx.x;
    const T zero = CX<T>::make(0, 0), one = CX<T>::make(1, 0);
    // M[i][k]: L (which 0, strictly lower part, unit diagonal) or U (which 1, upper part) of the block
    for (int i = 0; i < SB; ++i) {
        T v;
        if (i < bn && j < bn) {
            const T a = LU[(long)j * lda + i];
            v = which == 0 ? (j < i ? a : (j == i ? one : zero)) : (j >= i ? a : zero);
        } else {
            v = (i == j) ? one : zero;
        }
        M[i][j] = v;
    }
    __syncthreads();
    if (which == 0) {               // forward: X[i][j] = -sum_{j <= k < i} L[i][k] X[k][j]
        for (int i = 0; i < SB; ++i) {
            T s = i == j ? one : zero;
            if (i > j) {
                for (int k = j; k < i; ++k) s = tsub(s, tmul(M[i][k], X[k][j]));
            }
            X[i][j] = i < j ? zero : s;
        }
    } else {                        // backward: X[i][j] = (delta_ij - sum_{i < k <= j} U[i][k] X[k][j]) / U[i][i]
        for (int i = SB - 1; i >= 0; --i) {
            T s = i == j ? one : zero;
            if (i <= j) {
                for (int k = i + 1; k <= j; ++k) s = tsub(s, tmul(M[i][k], X[k][j]));
                X[i][j] = tdiv(s, M[i][i]);
            } else {
                X[i][j] = zero;
            }
        }
    }
    __syncthreads();
    for (int i = 0; i < SB; ++i) {
        if (colmajor) out[(long)j * SB + i] = X[i][j];
        else out[(long)i * SB + j] = X[i][j];
    }
    __syncthreads();
}

// Per diagonal block b: Dinv[b][0] = L_bb^-1 (unit lower), Dinv[b][1] = U_bb^-1 (upper), 64 x 64 row-major;
// a last block shorter than 64 is padded with the identity. One workgroup a block, one column a thread.
template <typename T>
__global__ __launch_bounds__(64) void k_trtri_diag(const T *LU, long lda, long n, T *Dinv)
{
    __shared__ T M[SB][SB + 1];
    __shared__ T X[SB][SB + 1];
    const long i0 = (long)blockIdx.x * SB;
    const int bn = (int)std::min((long)SB, n - i0);
    for (int which = 0; which < 2; ++which)
        trtri64(LU + i0 * lda + i0, lda, bn, which, M, X, Dinv + ((size_t)blockIdx.x * 2 + which) * SB * SB, false);
}

// ---- tournament pivoting (communication-avoiding LU panels) -----------------------------------------------
// A panel's TW pivot rows are chosen by a tree of GEPPs (Grigori, Demmel and Xiang, "CALU: a communication
// optimal LU factorization algorithm", SIAM J. Matrix Anal. Appl. 32 (2011)): each leaf workgroup factors RL rows
// of the panel in LDS and nominates its pivot rows; a tree node factors its children's nominees (their original
// rows, stacked in child order) and nominates its own; the root's pivot rows are the panel's, and the root's
// factors of them are the panel's L11 \ U11. One launch a panel: a node is factored by whichever of its children
// arrives last (an atomic count per node), so no workgroup ever waits on another. The rows below take
// L21 = A21 U11^-1 and the trailing matrix the usual update, on the matrix cores.
constexpr int TW = 64;
template <typename T> struct Tour;
template <> struct Tour<double2> { static constexpr int RL = 128, AR = 2; };   // rows a leaf, children a node
template <> struct Tour<float2> { static constexpr int RL = 256, AR = 4; };

template <typename T> struct TNode {
    int id[TW];                 // the nominees' global rows, in their pivot order
    int count, pad[63];
    T v[TW * TW];               // their original panel rows, column-major: v[c * TW + i]
};

inline long tour_leaves(long m, int RL) { return (m + RL - 1) / RL; }
inline long tour_nodes(long n, int RL, int AR)
{
    long P = tour_leaves(n, RL), tot = 0;
    while (P > 1) { tot += P; P = (P + AR - 1) / AR; }
    return tot + 1;
}

// GEPP of the rows in slots [0, nr) of S (column-major, S[c * R + slot], the slots in the rows' stacking order)
// over columns [0, nc); thread t owns slot t and rows never move: a row's position in the elimination order is
// tracked instead, and ties in the pivot search go to the earliest position (LAPACK's rule). order[k] = the slot
// of the k-th pivot. On return a slot chosen at step k holds row k of L \ U, the other slots their multipliers
// and Schur complement. One barrier a step: the per-wave maxima are double-buffered.
template <typename T, int NT>
__device__ int lds_gepp(T *S, int R, int nr, int nc, int *order, double (*wv)[NT / 64], int (*wk)[NT / 64])
{
    constexpr int NW = NT / 64;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int npiv = nr < nc ? nr : nc;
    const T zero = CX<T>::make(0, 0);
    bool live = t < nr;
    int pos = t;
    for (int k = 0; k < npiv; ++k) {
        const T a = live ? S[k * R + t] : zero;
        double v = live ? cabs1(a) : -1.0;
        if (live && !(v >= 0)) v = 0;        // a NaN never outranks a live row, and a live row always wins over none
        int key = live ? ((pos << 16) | t) : 0x7fffffff;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const double v2 = __shfl_xor(v, o);
            const int k2 = __shfl_xor(key, o);
            if (v2 > v || (v2 == v && k2 < key)) { v = v2; key = k2; }
        }
        const int b = k & 1;
        if (lane == 0) { wv[b][w] = v; wk[b][w] = key; }
        __syncthreads();
        v = wv[b][0];
        key = wk[b][0];
#pragma unroll
        for (int q = 1; q < NW; ++q) {
            const double v2 = wv[b][q];
            const int k2 = wk[b][q];
            if (v2 > v || (v2 == v && k2 < key)) { v = v2; key = k2; }
        }
        const int ps = key & 0xffff, p = key >> 16;
        if (t == 0) order[k] = ps;
        if (t == ps) {
            live = false;
        } else if (live) {
            if (pos == k) pos = p;
            const T pv = S[k * R + ps];
            const T l = (pv.x == 0 && pv.y == 0) ? a : tdiv(a, pv);
            S[k * R + t] = l;
            for (int c = k + 1; c < nc; ++c) S[c * R + t] = tsub(S[c * R + t], tmul(l, S[c * R + ps]));
        }
    }
    __syncthreads();
    return npiv;
}

// The tournament of panel [j0, j0 + jb) over rows [j0, n). Leaves: blockIdx.x (P of them). The root writes the
// panel's L11 \ U11 (top, column-major TW x TW), the interchanges (ipiv[j0 .. j0 + jb), LAPACK's convention) and
// the row moves they make as (destination, source) pairs (pairs[0] = count), and info for a zero pivot.
template <typename T>
__global__ __launch_bounds__(Tour<T>::RL) void k_tslu(const T *A, long lda, long n, long j0, int jb, int P,
                                                     TNode<T> *nodes, int *arrive, T *top, int *ipiv, int *pairs,
                                                     int *info)
{
    constexpr int RL = Tour<T>::RL, AR = Tour<T>::AR, NT = RL, NW = NT / 64;
    extern __shared__ char smem_raw[];
    T *S = (T *)smem_raw;                   // RL x TW
    __shared__ int ids[RL];
    __shared__ int order[TW];
    __shared__ double wv[2][NW];
    __shared__ int wk[2][NW];
    __shared__ int go;
    const int t = threadIdx.x;
    int idx = blockIdx.x;
    const long r0 = j0 + (long)idx * RL;
    int nr = (int)std::min((long)RL, n - r0);
    for (int e = t; e < RL * jb; e += NT) {
        const int c = e / RL, r = e % RL;
        if (r < nr) S[c * RL + r] = A[(j0 + c) * lda + r0 + r];
    }
    if (t < nr) ids[t] = (int)(r0 + t);
    __syncthreads();
    int cnt = lds_gepp<T, NT>(S, RL, nr, jb, order, wv, wk);
    int lsize = P, loff = 0;
    while (lsize > 1) {
        const int parent = idx / AR, first = parent * AR;
        const int nchild = std::min(AR, lsize - first);
        const int poff = loff + lsize;
        if (nchild > 1) {
            // nominate: this node's pivot rows, their original values from A (untouched during the launch)
            TNode<T> *nd = nodes + loff + idx;
            if (t == 0) nd->count = cnt;
            if (t < cnt) nd->id[t] = ids[order[t]];
            for (int e = t; e < cnt * jb; e += NT) {
                const int c = e / cnt, i = e % cnt;
                nd->v[c * TW + i] = A[(j0 + c) * lda + ids[order[i]]];
            }
            __syncthreads();
            if (t == 0) {
                __threadfence();
                const int old = __hip_atomic_fetch_add(arrive + poff + parent, 1, __ATOMIC_ACQ_REL,
                                                       __HIP_MEMORY_SCOPE_AGENT);
                go = old == nchild - 1;
                if (go) __hip_atomic_store(arrive + poff + parent, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __threadfence();
            }
            __syncthreads();
            if (!go) return;
            __threadfence();
            // the last child to arrive factors the node: its children's nominees stacked in child order
            nr = 0;
            for (int ch = 0; ch < nchild; ++ch) {
                const TNode<T> *cn = nodes + loff + first + ch;
                const int cc = cn->count;
                for (int e = t; e < cc * jb; e += NT) {
                    const int c = e / cc, i = e % cc;
                    S[c * RL + nr + i] = cn->v[c * TW + i];
                }
                if (t < cc) ids[nr + t] = cn->id[t];
                nr += cc;
            }
            __syncthreads();
            cnt = lds_gepp<T, NT>(S, RL, nr, jb, order, wv, wk);
        }
        // a node with one child nominates the child's rows unchanged
        loff = poff;
        lsize = (lsize + AR - 1) / AR;
        idx = parent;
    }
    // the root: cnt == jb pivot rows
    for (int e = t; e < TW * TW; e += NT) {
        const int c = e / TW, i = e % TW;
        top[e] = (c < jb && i < jb) ? S[c * RL + order[i]] : CX<T>::make(i == c, 0);
    }
    if (t == 0 && info) {
        for (int k = 0; k < jb; ++k) {
            const T d = S[k * RL + order[k]];
            if (d.x == 0 && d.y == 0) {
                if (*info == 0) *info = (int)(j0 + k) + 1;
                break;
            }
        }
    }
    if (t < 64) {
        // LAPACK's interchanges: winner k to position j0 + k in order. Rows below the panel's first jb move only
        // when they win; lane i tracks loc (the position of the row first at j0 + i) and cont (the row at j0 + i).
        const int i = t;
        const int top_end = (int)j0 + jb;
        int loc = (int)j0 + i, cont = (int)j0 + i;
        const int win = i < jb ? ids[order[i]] : 0;
        for (int k = 0; k < jb; ++k) {
            const int w = __builtin_amdgcn_readlane(win, k);
            const int tpos = (int)j0 + k;
            const int pw = w < top_end ? __builtin_amdgcn_readlane(loc, w - (int)j0) : w;
            if (i == 0) ipiv[tpos] = pw;
            if (pw != tpos) {
                const int x = __builtin_amdgcn_readlane(cont, k);        // a row first in the panel's top
                if (i == k) cont = w;
                if (w < top_end && i == w - (int)j0) loc = tpos;
                if (pw < top_end && i == pw - (int)j0) cont = x;
                if (i == x - (int)j0) loc = pw;
            }
        }
        // the moves: position j0 + i takes row cont; a row of the top that ended below takes its place there
        const bool mv_top = i < jb && cont != (int)j0 + i;
        const bool mv_low = i < jb && loc >= top_end;
        const unsigned long long m1 = __ballot(mv_top), m2 = __ballot(mv_low);
        const unsigned long long below = (1ull << i) - 1;
        const int n1 = __popcll(m1);
        if (mv_top) {
            const int q = __popcll(m1 & below);
            pairs[1 + 2 * q] = (int)j0 + i;
            pairs[2 + 2 * q] = cont;
        }
        if (mv_low) {
            const int q = n1 + __popcll(m2 & below);
            pairs[1 + 2 * q] = loc;
            pairs[2 + 2 * q] = (int)j0 + i;
        }
        if (i == 0) pairs[0] = n1 + __popcll(m2);
    }
}

// The inverses of the panel's L11 (block 0) and U11 (block 1), column-major 64 x 64, for L21 = A21 U11^-1 and
// U12 = L11^-1 A12 on the matrix cores.
template <typename T>
__global__ __launch_bounds__(64) void k_trtri_top(const T *top, int jb, T *Linv, T *Uinv)
{
    __shared__ T M[SB][SB + 1];
    __shared__ T X[SB][SB + 1];
    const int which = blockIdx.x;
    trtri64(top, TW, jb, which, M, X, which ? Uinv : Linv, true);
}

// The panel's row moves on columns [c0, c0 + nc): a column per 128 threads, every source read before any write;
// the panel's own columns [j0, j0 + jb) then take the root's L11 \ U11 in rows [j0, j0 + jb).
template <typename T>
__global__ __launch_bounds__(256) void k_apply_moves(T *A, long lda, long c0, long nc, const int *pairs, long j0,
                                                     int jb, const T *top)
{
    const int t = threadIdx.x & 127, h = threadIdx.x >> 7;
    const int np = pairs[0];
    const int dst = t < np ? pairs[1 + 2 * t] : 0, src = t < np ? pairs[2 + 2 * t] : 0;
    for (long cb = (long)blockIdx.x * 8; cb < std::min(nc, (long)blockIdx.x * 8 + 8); cb += 2) {
        const long c = c0 + cb + h;
        const bool on = cb + h < nc;
        T *col = A + c * lda;
        T v = CX<T>::make(0, 0);
        if (on && t < np) v = col[src];
        __syncthreads();
        if (on && t < np) col[dst] = v;
        __syncthreads();
        if (on && c >= j0 && c < j0 + jb)
            for (int i = t; i < jb; i += 128) col[j0 + i] = top[(c - j0) * TW + i];
    }
}

// kind 0: U^T y = b (forward); 1: L^T z = y (backward); 2: L y = b (forward); 3: U x = y (backward).
// Operator entry (row i, column k): kinds 0, 1 read LU[i * lda + k] (U[k][i], L[k][i]); kinds 2, 3 LU[k * lda + i].
template <typename T>
__global__ __launch_bounds__(256) void k_trsv(const T *LU, long lda, long n, const T *Dinv, T *B, long ldb, int nrhs,
                                              int kind, int *flags, int *abort)
{
    __shared__ T Mi[SB][SB + 1];             // the diagonal block's inverse for this operator, Mi[i][k]
    __shared__ T acc[SB][MAXRHS];
    __shared__ T part[4][SB][MAXRHS];
    const int nb = (int)((n + SB - 1) / SB);
    const bool fwd = kind == 0 || kind == 2;
    const int b = fwd ? blockIdx.x : nb - 1 - blockIdx.x;     // publication order is blockIdx order
    const long i0 = (long)b * SB;
    const int bn = (int)std::min((long)SB, n - i0);
    const int tid = threadIdx.x;
    const bool rowmajor = kind <= 1;
    {
        // kind 0: (U^-1)^T, 1: (L^-1)^T, 2: L^-1, 3: U^-1
        const T *D = Dinv + ((size_t)b * 2 + ((kind == 0 || kind == 3) ? 1 : 0)) * SB * SB;
        const bool tr = kind <= 1;
        for (int idx = tid; idx < SB * SB; idx += 256) {
            const int i = idx / SB, k = idx % SB;
            Mi[i][k] = tr ? D[k * SB + i] : D[i * SB + k];
        }
    }
    for (int idx = tid; idx < SB * MAXRHS; idx += 256) {
        const int i = idx / MAXRHS, r = idx % MAXRHS;
        acc[i][r] = (i < bn && r < nrhs) ? B[(i0 + i) * ldb + r] : CX<T>::make(0, 0);
    }
    __syncthreads();
    const int row = tid & 63, kq = tid >> 6;                  // 4 slices of 16 columns
    for (int q = 0; q < (int)blockIdx.x; ++q) {
        const int a = fwd ? q : nb - 1 - q;
        const long k0 = (long)a * SB;
        const int kn = (int)std::min((long)SB, n - k0);
        // this thread's 16 operator entries of block (b, a), loaded before the wait
        T m[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int kx = kq * 16 + u;
            const bool in = row < bn && kx < kn;
            const long gi = i0 + row, gk = k0 + kx;
            m[u] = in ? (rowmajor ? LU[gi * lda + gk] : LU[gk * lda + gi]) : CX<T>::make(0, 0);
        }
        if (tid == 0) {
            long spins = 0;
            while (__hip_atomic_load(flags + a, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == 0) {
                __builtin_amdgcn_s_sleep(1);
                if (++spins > (1L << 26)) { __hip_atomic_store(abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); break; }
            }
        }
        __syncthreads();
        T s[MAXRHS];
#pragma unroll
        for (int r = 0; r < MAXRHS; ++r) s[r] = CX<T>::make(0, 0);
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int kx = kq * 16 + u;
            if (kx < kn) {
                const T *y = B + (k0 + kx) * ldb;
#pragma unroll
                for (int r = 0; r < MAXRHS; ++r)
                    if (r < nrhs) s[r] = tadd(s[r], tmul(m[u], y[r]));
            }
        }
#pragma unroll
        for (int r = 0; r < MAXRHS; ++r) part[kq][row][r] = s[r];
        __syncthreads();
        for (int idx = tid; idx < SB * nrhs; idx += 256) {
            const int i = idx / nrhs, r = idx % nrhs;
            acc[i][r] = tsub(acc[i][r], tadd(tadd(part[0][i][r], part[1][i][r]), tadd(part[2][i][r], part[3][i][r])));
        }
        __syncthreads();
    }
    // the diagonal block: y = Mi acc
    for (int idx = tid; idx < bn * nrhs; idx += 256) {
        const int i = idx / nrhs, r = idx % nrhs;
        T v = CX<T>::make(0, 0);
        for (int k = 0; k < bn; ++k) v = tadd(v, tmul(Mi[i][k], acc[k][r]));
        B[(i0 + i) * ldb + r] = v;
    }
    __syncthreads();
    if (tid == 0) {
        __threadfence();
        __hip_atomic_store(flags + b, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// out[perm[i]] = in[i] (scatter, trans) or out[i] = in[perm[i]] (gather)
template <typename T>
__global__ void k_permute(const T *in, T *out, long n, int nrhs, long ld, const int *perm, int scatter)
{
    const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (idx >= n * nrhs) return;
    const long i = idx / nrhs;
    const int r = (int)(idx % nrhs);
    if (scatter) out[(long)perm[i] * ld + r] = in[i * ld + r];
    else out[i * ld + r] = in[(long)perm[i] * ld + r];
}

inline unsigned blocks(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

template <typename T> size_t work_bytes(long n)
{
    const int cus = 256;
    return 2 * (size_t)cus * sizeof(Slot<T>) + 2 * NB * sizeof(T) + 256;
}

template <typename T> int max_rows()
{
    // LDS for the slab: at most 128 KB
    return (int)((128 * 1024) / (NB * sizeof(T)));
}

template <typename T>
int getrf_t(T *A, long n, long lda, int *ipiv, int *info, void *work, size_t wbytes, hipStream_t s)
{
    if (wbytes < work_bytes<T>(n)) return fail(M25X_EINVAL, "getrf: workspace of %zu bytes, %zu needed", wbytes, work_bytes<T>(n));
    const int cus = std::min(num_cus(), 256);
    char *w = (char *)work;
    Slot<T> *slots = (Slot<T> *)w;
    T *diag = (T *)(w + 2 * (size_t)256 * sizeof(Slot<T>));
    Bar *bar = (Bar *)(w + 2 * (size_t)256 * sizeof(Slot<T>) + 2 * NB * sizeof(T));
    const int Rmax = max_rows<T>();
    static bool attr_done[2] = {false, false};
    const size_t smem_max = (size_t)Rmax * NB * sizeof(T);
    {
        static std::mutex mu;
        std::lock_guard<std::mutex> lk(mu);
        if (!attr_done[sizeof(T) == 8]) {
            M25X_HIP(hipFuncSetAttribute((const void *)k_panel<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem_max));
            attr_done[sizeof(T) == 8] = true;
        }
    }
    if (info) {
        hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, s, info, 1);
        if (int rc = check_launch("getrf")) return rc;
    }
    for (long j0 = 0; j0 < n; j0 += NB) {
        const int jb = (int)std::min((long)NB, n - j0);
        const long m = n - j0;
        int G = (int)std::min((long)cus, (m + Rmax - 1) / Rmax);
        G = std::max(G, 1);
        int R = (int)((m + G - 1) / G);
        hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, s, (int *)bar, 4);
        if (int rc = check_launch("getrf")) return rc;
        const size_t smem = (size_t)R * jb * sizeof(T);
        void *args[] = {(void *)&A, (void *)&lda, (void *)&n, (void *)&j0, (void *)&jb, (void *)&R, (void *)&ipiv,
                        (void *)&info, (void *)&slots, (void *)&diag, (void *)&bar};
        M25X_HIP(hipLaunchCooperativeKernel((const void *)k_panel<T>, dim3(G), dim3(PT), args, (unsigned)smem, s));
        const long nsw = n - jb;
        if (nsw > 0) {
            hipLaunchKernelGGL(k_laswp<T>, dim3(blocks(nsw, 256)), dim3(256), 0, s, A, lda, n, j0, jb, ipiv);
            if (int rc = check_launch("getrf/laswp")) return rc;
        }
        const long c0 = j0 + jb;
        if (c0 < n) {
            hipLaunchKernelGGL(k_trsm<T>, dim3(blocks(n - c0, TC)), dim3(256), 0, s, A, lda, n, j0, jb, c0);
            if (int rc = check_launch("getrf/trsm")) return rc;
            const long M = n - c0;
            dim3 grid(blocks(M, GM), blocks(M, GN));
            hipLaunchKernelGGL((k_gemm<T, true>), grid, dim3(256), 0, s, A + c0 * lda + c0, lda, A + j0 * lda + c0, lda,
                               A + c0 * lda + j0, lda, M, M, (long)jb);
            if (int rc = check_launch("getrf/gemm")) return rc;
        }
    }
    return M25X_OK;
}

template <typename T> size_t tour_work_bytes(long n)
{
    const long nn = tour_nodes(n, Tour<T>::RL, Tour<T>::AR);
    return nn * sizeof(TNode<T>) + 3 * (size_t)TW * TW * sizeof(T) + sizeof(int) * ((size_t)nn + 4 * TW + 64);
}

template <typename T>
int getrf_tp(T *A, long n, long lda, int *ipiv, int *info, void *work, size_t wbytes, hipStream_t s)
{
    constexpr int RL = Tour<T>::RL, AR = Tour<T>::AR;
    if (wbytes < tour_work_bytes<T>(n))
        return fail(M25X_EINVAL, "getrf: workspace of %zu bytes, %zu needed", wbytes, tour_work_bytes<T>(n));
    const long nn = tour_nodes(n, RL, AR);
    char *w = (char *)work;
    TNode<T> *nodes = (TNode<T> *)w;
    w += nn * sizeof(TNode<T>);
    T *top = (T *)w, *Linv = top + TW * TW, *Uinv = Linv + TW * TW;
    w += 3 * (size_t)TW * TW * sizeof(T);
    int *arrive = (int *)w, *pairs = arrive + nn;
    const size_t smem = (size_t)RL * TW * sizeof(T);
    {
        static bool attr_done[2] = {false, false};
        static std::mutex mu;
        std::lock_guard<std::mutex> lk(mu);
        if (!attr_done[sizeof(T) == 8]) {
            M25X_HIP(hipFuncSetAttribute((const void *)k_tslu<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
            attr_done[sizeof(T) == 8] = true;
        }
    }
    hipLaunchKernelGGL(k_zero_ints, dim3(blocks(nn, 256)), dim3(256), 0, s, arrive, (int)nn);
    if (info) hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, s, info, 1);
    if (int rc = check_launch("getrf")) return rc;
    for (long j0 = 0; j0 < n; j0 += TW) {
        const int jb = (int)std::min((long)TW, n - j0);
        const long m = n - j0;
        const int P = (int)tour_leaves(m, RL);
        hipLaunchKernelGGL(k_tslu<T>, dim3(P), dim3(RL), (unsigned)smem, s, (const T *)A, lda, n, j0, jb, P, nodes,
                           arrive, top, ipiv, pairs, info);
        hipLaunchKernelGGL(k_trtri_top<T>, dim3(2), dim3(64), 0, s, (const T *)top, jb, Linv, Uinv);
        hipLaunchKernelGGL(k_apply_moves<T>, dim3(blocks(n, 8)), dim3(256), 0, s, A, lda, 0L, n, (const int *)pairs,
                           j0, jb, (const T *)top);
        if (int rc = check_launch("getrf/panel")) return rc;
        if (m > jb) {   // L21 = A21 U11^-1
            T *A21 = A + j0 * lda + j0 + jb;
            hipLaunchKernelGGL((k_gemm<T, false>), dim3(blocks(m - jb, GM), 1), dim3(256), 0, s, A21, lda,
                               (const T *)A21, lda, (const T *)Uinv, (long)TW, m - jb, (long)jb, (long)jb);
        }
        const long c0 = j0 + jb;
        if (c0 < n) {   // U12 = L11^-1 A12, A22 -= L21 U12
            T *A12 = A + c0 * lda + j0;
            hipLaunchKernelGGL((k_gemm<T, false>), dim3(1, blocks(n - c0, GN)), dim3(256), 0, s, A12, lda,
                               (const T *)Linv, (long)TW, (const T *)A12, lda, (long)jb, n - c0, (long)jb);
            const long M = n - c0;
            hipLaunchKernelGGL((k_gemm<T, true>), dim3(blocks(M, GM), blocks(M, GN)), dim3(256), 0, s,
                               A + c0 * lda + c0, lda, (const T *)(A + j0 * lda + c0), lda, (const T *)A12, lda, M, M,
                               (long)jb);
        }
        if (int rc = check_launch("getrf/update")) return rc;
    }
    return M25X_OK;
}

template <typename T> size_t dinv_bytes(long n)
{
    const long nb = (n + SB - 1) / SB;
    return (size_t)nb * 2 * SB * SB * sizeof(T);
}

template <typename T> size_t getrs_work(long n, long ldb)
{
    const long nb = (n + SB - 1) / SB;
    return (size_t)n * ldb * sizeof(T) + dinv_bytes<T>(n) + sizeof(int) * ((size_t)n + 2 * nb + 64);
}

template <typename T>
int getrs_t(const T *LU, long n, long lda, const int *ipiv, const int *perm_in, const T *dinv, T *B, long ldb,
            int nrhs, int trans, void *work, size_t wbytes, hipStream_t s)
{
    if (nrhs < 1 || nrhs > MAXRHS) return fail(M25X_EINVAL, "getrs: 1 to %d right-hand sides, not %d", MAXRHS, nrhs);
    if (ldb < nrhs) return fail(M25X_EINVAL, "getrs: ldb < nrhs");
    if (!work || wbytes < getrs_work<T>(n, ldb))
        return fail(M25X_EINVAL, "getrs: workspace of %zu bytes, %zu needed", wbytes, getrs_work<T>(n, ldb));
    const int nb = (int)((n + SB - 1) / SB);
    const size_t tmp_bytes = (size_t)n * ldb * sizeof(T);
    T *tmp = (T *)work;
    T *dwork = (T *)((char *)work + tmp_bytes);
    int *perm = (int *)((char *)dwork + dinv_bytes<T>(n));
    int *flags = perm + n;
    int *abort = flags + 2 * nb;
    hipLaunchKernelGGL(k_zero_ints, dim3(blocks(2 * nb + 1, 256)), dim3(256), 0, s, flags, 2 * nb + 1);
    if (!perm_in) {
        hipLaunchKernelGGL(k_perm, dim3(1), dim3(256), 0, s, ipiv, n, perm);
        perm_in = perm;
    }
    if (!dinv) {
        hipLaunchKernelGGL(k_trtri_diag<T>, dim3(nb), dim3(SB), 0, s, LU, lda, n, dwork);
        dinv = dwork;
    }
    if (int rc = check_launch("getrs")) return rc;
    if (trans) {
        hipLaunchKernelGGL(k_trsv<T>, dim3(nb), dim3(256), 0, s, LU, lda, n, dinv, B, ldb, nrhs, 0, flags, abort);
        hipLaunchKernelGGL(k_trsv<T>, dim3(nb), dim3(256), 0, s, LU, lda, n, dinv, B, ldb, nrhs, 1, flags + nb, abort);
        hipLaunchKernelGGL(k_permute<T>, dim3(blocks(n * nrhs, 256)), dim3(256), 0, s, (const T *)B, tmp, n, nrhs, ldb,
                           perm_in, 1);
    } else {
        hipLaunchKernelGGL(k_permute<T>, dim3(blocks(n * nrhs, 256)), dim3(256), 0, s, (const T *)B, tmp, n, nrhs, ldb,
                           perm_in, 0);
        hipLaunchKernelGGL(k_trsv<T>, dim3(nb), dim3(256), 0, s, LU, lda, n, dinv, tmp, ldb, nrhs, 2, flags, abort);
        hipLaunchKernelGGL(k_trsv<T>, dim3(nb), dim3(256), 0, s, LU, lda, n, dinv, tmp, ldb, nrhs, 3, flags + nb, abort);
    }
    if (int rc = check_launch("getrs")) return rc;
    M25X_HIP(hipMemcpyAsync(B, tmp, tmp_bytes, hipMemcpyDeviceToDevice, s));
    return M25X_OK;
}

}  // namespace

extern "C" {

size_t m25x_getrf_work_bytes(int dtype, long n)
{
    return dtype == M25X_C64 ? std::max(work_bytes<float2>(n), tour_work_bytes<float2>(n))
                             : std::max(work_bytes<double2>(n), tour_work_bytes<double2>(n));
}

int m25x_getrf(int dtype, int pivot, void *A, long n, long lda, int *ipiv, int *info, void *work, size_t wbytes,
               void *stream)
{
    if (!A || !ipiv || !work || n < 0 || lda < n) return fail(M25X_EINVAL, "getrf: bad arguments");
    if (pivot != M25X_PIVOT_PARTIAL && pivot != M25X_PIVOT_TOURNAMENT) return fail(M25X_EINVAL, "getrf: pivot %d", pivot);
    if (n == 0) return M25X_OK;
    if (n > (1L << 30)) return fail(M25X_EINVAL, "getrf: n too large");
    hipStream_t s = (hipStream_t)stream;
    const bool tp = pivot == M25X_PIVOT_TOURNAMENT;
    if (dtype == M25X_C128)
        return tp ? getrf_tp<double2>((double2 *)A, n, lda, ipiv, info, work, wbytes, s)
                  : getrf_t<double2>((double2 *)A, n, lda, ipiv, info, work, wbytes, s);
    if (dtype == M25X_C64)
        return tp ? getrf_tp<float2>((float2 *)A, n, lda, ipiv, info, work, wbytes, s)
                  : getrf_t<float2>((float2 *)A, n, lda, ipiv, info, work, wbytes, s);
    return fail(M25X_EINVAL, "getrf: dtype %d", dtype);
}

size_t m25x_getrs_work_bytes(int dtype, long n, long ldb)
{
    return dtype == M25X_C64 ? getrs_work<float2>(n, ldb) : getrs_work<double2>(n, ldb);
}

int m25x_ipiv_perm(const int *ipiv, long n, int *perm, void *stream)
{
    if (!ipiv || !perm || n < 0) return fail(M25X_EINVAL, "ipiv_perm: bad arguments");
    if (n == 0) return M25X_OK;
    hipLaunchKernelGGL(k_perm, dim3(1), dim3(256), 0, (hipStream_t)stream, ipiv, n, perm);
    return check_launch("ipiv_perm");
}

size_t m25x_dinv_bytes(int dtype, long n)
{
    return dtype == M25X_C64 ? dinv_bytes<float2>(n) : dinv_bytes<double2>(n);
}

int m25x_trtri_blocks(int dtype, const void *LU, long n, long lda, void *dinv, void *stream)
{
    if (!LU || !dinv || n < 0 || lda < n) return fail(M25X_EINVAL, "trtri_blocks: bad arguments");
    if (n == 0) return M25X_OK;
    const unsigned nb = (unsigned)((n + SB - 1) / SB);
    hipStream_t s = (hipStream_t)stream;
    if (dtype == M25X_C128)
        hipLaunchKernelGGL(k_trtri_diag<double2>, dim3(nb), dim3(SB), 0, s, (const double2 *)LU, lda, n, (double2 *)dinv);
    else if (dtype == M25X_C64)
        hipLaunchKernelGGL(k_trtri_diag<float2>, dim3(nb), dim3(SB), 0, s, (const float2 *)LU, lda, n, (float2 *)dinv);
    else
        return fail(M25X_EINVAL, "trtri_blocks: dtype %d", dtype);
    return check_launch("trtri_blocks");
}

int m25x_getrs(int dtype, const void *LU, long n, long lda, const int *ipiv, const int *perm, const void *dinv,
               void *B, long ldb, int nrhs, int trans, void *work, size_t work_bytes, void *stream)
{
    if (!LU || !ipiv || !B || n < 0 || lda < n) return fail(M25X_EINVAL, "getrs: bad arguments");
    if (n == 0) return M25X_OK;
    hipStream_t s = (hipStream_t)stream;
    if (dtype == M25X_C128)
        return getrs_t<double2>((const double2 *)LU, n, lda, ipiv, perm, (const double2 *)dinv, (double2 *)B, ldb, nrhs,
                                trans, work, work_bytes, s);
    if (dtype == M25X_C64)
        return getrs_t<float2>((const float2 *)LU, n, lda, ipiv, perm, (const float2 *)dinv, (float2 *)B, ldb, nrhs,
                               trans, work, work_bytes, s);
    return fail(M25X_EINVAL, "getrs: dtype %d", dtype);
}

}  // extern "C"
