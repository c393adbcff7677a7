// m25x_lu.hip - dense complex LU on gfx950, and the triangular solves with its factor.
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
// grid barrier (k_tslu), its row moves on every column (k_apply_moves), then L21 = A21 U11^-1 by substitution,
// U12 = L11^-1 A12 by substitution and A22 -= L21 U12 on the matrix cores (complex64 with FP64 accumulation: one
// rounding per update).
// The solves (getrs): one launch per triangle, one workgroup per 64-row diagonal block; each accumulates the
// products of its rows with the earlier blocks as their values appear, then solves its own block and hands its
// values on as self-marking words.
#include <atomic>
#include <mutex>
#include <vector>
#include "m25x_lu_dev.h"

using namespace m25x;
using namespace m25x::lu;

namespace {

template <typename T> constexpr int dtype_of() { return sizeof(T) == 16 ? M25X_C128 : M25X_C64; }

// ---- grid barrier (all workgroups of a cooperative launch co-resident) -------------------------------------
struct Bar {
    unsigned count, gen;
    int abort, pad;
};

// Sense by generation; release / acquire at agent scope orders the slot writes and reads across the XCDs.
// A bounded wait: past the limit the barrier sets abort and every later barrier returns at once, so the
// launch always drains. `aborted` becomes true in a workgroup that timed out or skipped a barrier for it; the
// panel kernel then writes -1 into the factor's info word, which the caller reads (the factor is not valid).
__device__ inline void grid_sync(Bar *b, unsigned nblocks, unsigned &gen, bool &aborted)
{
    __threadfence();
    __syncthreads();
    __shared__ int s_abort;
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
        s_abort = __hip_atomic_load(&b->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        gen += 1;
    }
    __syncthreads();
    aborted = aborted || s_abort != 0;
}

__global__ void k_zero_ints(int *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_iota(int *p, long n)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) p[i] = (int)i;
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
    bool aborted = false;
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
    grid_sync(bar, G, gen, aborted);
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
            grid_sync(bar, G, gen, aborted);
        }
    }
    for (int idx = tid; idx < rn * jb; idx += PT) {
        const int c = idx / rn, r = idx % rn;
        A[(j0 + c) * lda + r0 + r] = S[c * R + r];
    }
    if (aborted && tid == 0 && info) __hip_atomic_exchange(info, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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

// U12 = L11^-1 A12: rows [j0, j0 + jb), columns [c0, n); TC columns a workgroup, the block in LDS. JB: the largest
// panel width; with JB = TW (the tournament's panels) L11 is staged in LDS too (column k of L11 as row k of Lt, so a
// step's reads are consecutive), else read from memory at each step.
constexpr int TC = 16;
template <typename T, int JB>
__device__ __forceinline__ void trsm_body(T *A, long lda, long n, long j0, int jb, long c0, long bx)
{
    constexpr bool STAGE = JB <= TW;
    __shared__ T X[JB][TC + 1];
    __shared__ T Lt[STAGE ? TW : 1][TW];
    const long cb = c0 + bx * TC;
    const int nc = (int)std::min((long)TC, n - cb);
    const int tid = threadIdx.x;
    const T *L = A + j0 * lda + j0;          // L11[i][k] = L[k * lda + i]
    if constexpr (STAGE) {                   // jb <= TW: the loads of either block in flight before the stores
        constexpr int NX = TW * TC / 256, NL = TW * TW / 256;
        T vx[NX], vl[NL];
#pragma unroll
        for (int u = 0; u < NX; ++u) {
            const int idx = tid + u * 256, c = idx / jb, r = idx % jb;
            vx[u] = ldz(idx < jb * TC && c < nc, A + (cb + c) * lda + j0 + r);
        }
#pragma unroll
        for (int u = 0; u < NL; ++u) {
            const int idx = tid + u * 256, k = idx / jb, i = idx % jb;
            vl[u] = ldz(idx < jb * jb, L + (long)k * lda + i);
        }
#pragma unroll
        for (int u = 0; u < NX; ++u) {
            const int idx = tid + u * 256, c = idx / jb, r = idx % jb;
            if (idx < jb * TC && c < nc) X[r][c] = vx[u];
        }
#pragma unroll
        for (int u = 0; u < NL; ++u) {
            const int idx = tid + u * 256, k = idx / jb, i = idx % jb;
            if (idx < jb * jb) Lt[k][i] = vl[u];
        }
    } else {
        for (int idx = tid; idx < jb * TC; idx += 256) {
            const int c = idx / jb, r = idx % jb;
            if (c < nc) X[r][c] = A[(cb + c) * lda + j0 + r];
        }
    }
    __syncthreads();
    for (int k = 0; k < jb - 1; ++k) {
        const int nr = jb - k - 1;
        for (int idx = tid; idx < nr * TC; idx += 256) {
            const int c = idx % TC, i = k + 1 + idx / TC;
            T lik;
            if constexpr (STAGE) lik = Lt[k][i];
            else lik = L[(long)k * lda + i];
            if (c < nc) X[i][c] = tsub(X[i][c], tmul(lik, X[k][c]));
        }
        __syncthreads();
    }
    for (int idx = tid; idx < jb * TC; idx += 256) {
        const int c = idx / jb, r = idx % jb;
        if (c < nc) A[(cb + c) * lda + j0 + r] = X[r][c];
    }
}

template <typename T, int JB>
__global__ __launch_bounds__(256) void k_trsm(T *A, long lda, long n, long j0, int jb, long c0)
{
    trsm_body<T, JB>(A, lda, n, j0, jb, c0, blockIdx.x);
}

// ---- C -= A B (SUB) or C = A B on the matrix cores -------------------------------------------------------
// C (M x N), A (M x K), B (K x N), column-major, complex. C = A B may overwrite A or B in place when a workgroup
// tile's output depends on nothing another workgroup writes (M <= 64 with C = B, or N <= 64 with C = A): every
// workgroup reads all of its K range before it stores. Workgroup tile 64 x 64, four waves of 32 x 32, each
// 2 x 2 blocks of 16 x 16 with a real and an imaginary accumulator. The MFMA's first operand is B (its rows are
// C's columns), its second A, so a lane's outputs run down C's columns (coalesced stores).
constexpr int GM = 64, GN = 64, GK = 8;

template <typename T, bool SUB, typename AC = T, int KS = GK>
__device__ __forceinline__ void gemm_body(T *C, long ldc, const T *A, long lda, const T *B, long ldb, long M, long N,
                                          long K, long bx, long by)
{
    typedef typename CX<AC>::R R;           // the matrix cores' type: products and sums in it, one rounding to T
    typedef typename CX<T>::R TR;
    typedef typename CX<AC>::acc4 acc4;
    constexpr int LA = GM * KS / 256, LB = KS * GN / 256;   // a thread's elements of A and of B a stage
    __shared__ T As[KS][GM];         // As[p][i] = A[i][p]
    __shared__ T Bs[KS][GN + 1];     // Bs[p][j] = B[p][j]
    const long m0 = bx * GM, n0 = by * GN;
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
    // a stage's A (GM x KS) and B (KS x GN) elements of the thread, zero outside, read into registers one stage
    // ahead: the loads of stage s + 1 are in flight while stage s is multiplied
    T ra[LA], rb[LB];
    auto fetch = [&](long k0) __attribute__((always_inline)) {
#pragma unroll
        for (int it = 0; it < LA; ++it) {
            const int idx = tid + it * 256, i = idx % GM, p = idx / GM;
            const long gi = m0 + i, gp = k0 + p;
            ra[it] = ldz(gi < M && gp < K, A + gp * lda + gi);
        }
#pragma unroll
        for (int it = 0; it < LB; ++it) {
            const int idx = tid + it * 256, p = idx % KS, j = idx / KS;
            const long gp = k0 + p, gj = n0 + j;
            rb[it] = ldz(gp < K && gj < N, B + gj * ldb + gp);
        }
    };
    fetch(0);
    for (long k0 = 0; k0 < K; k0 += KS) {
#pragma unroll
        for (int it = 0; it < LA; ++it) {
            const int idx = tid + it * 256;
            As[idx / GM][idx % GM] = ra[it];
        }
#pragma unroll
        for (int it = 0; it < LB; ++it) {
            const int idx = tid + it * 256;
            Bs[idx % KS][idx / KS] = rb[it];
        }
        __syncthreads();
        if (k0 + KS < K) fetch(k0 + KS);
#pragma unroll
        for (int ks = 0; ks < KS / 2; ++ks) {
            const int p = 2 * ks + ph;
            R av[2], br[2], bi[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const T x = As[p][wm + a * 16 + li];
                av[a] = (R)(part ? x.y : x.x);
            }
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const T y = Bs[p][wn + b * 16 + li];
                br[b] = (R)(part ? -y.y : y.x);
                bi[b] = (R)(part ? y.x : y.y);
            }
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int a = 0; a < 2; ++a) {
                    cr[b][a] = CX<AC>::mfma(br[b], av[a], cr[b][a]);
                    ci[b][a] = CX<AC>::mfma(bi[b], av[a], ci[b][a]);
                }
        }
        __syncthreads();
    }
    // lane (li, kk), register r: C row m0 + wm + a*16 + li, column n0 + wn + b*16 + out_row(kk, r); for C -= A B
    // the thread's entries of C are all read first (in flight together), then written
    T cv[2][2][4];
    if (SUB) {
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const long gi = m0 + wm + a * 16 + li, gj = n0 + wn + b * 16 + CX<AC>::out_row(kk, r);
                    cv[b][a][r] = ldz(gi < M && gj < N, C + gj * ldc + gi);
                }
    }
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const long gi = m0 + wm + a * 16 + li, gj = n0 + wn + b * 16 + CX<AC>::out_row(kk, r);
                if (gi < M && gj < N) {
                    T *c = C + gj * ldc + gi;
                    if (SUB) {
                        const T v = cv[b][a][r];
                        *c = CX<T>::make((TR)((R)v.x - cr[b][a][r]), (TR)((R)v.y - ci[b][a][r]));
                    } else {
                        *c = CX<T>::make((TR)cr[b][a][r], (TR)ci[b][a][r]);
                    }
                }
            }
}

template <typename T, bool SUB, typename AC = T, int KS = GK>
__global__ __launch_bounds__(256) void k_gemm(T *C, long ldc, const T *A, long lda, const T *B, long ldb, long M,
                                              long N, long K)
{
    gemm_body<T, SUB, AC, KS>(C, ldc, A, lda, B, ldb, M, N, K, blockIdx.x, blockIdx.y);
}

// The depth of the K stages of the tournament LU's trailing update (8, 16 or 32: the number of columns of L21 and
// rows of U12 staged in LDS between two barriers), chosen at run time by m25x_lu_set_gemm_stage (an autotuner
// times the choices); 8 until set.
std::atomic<int> g_gemm_stage{8};
// The complex64 trailing update's matrix cores: 1 FP32 (the default: v_mfma_f32_16x16x4f32, FP32 accumulation, twice
// the FP64 peak rate), 0 FP64 (v_mfma_f64_16x16x4f64, products and sums in FP64, one rounding to FP32 per update:
// a smaller backward error); m25x_lu_set_update_precision. complex128 always updates in FP64.
std::atomic<int> g_update_fp32{1};

template <typename T, typename AC>
void launch_update_ac(hipStream_t s, dim3 grid, T *C, long ldc, const T *A, long lda, const T *B, long ldb, long M,
                      long N, long K)
{
    switch (g_gemm_stage.load(std::memory_order_relaxed)) {
    case 16: hipLaunchKernelGGL((k_gemm<T, true, AC, 16>), grid, dim3(256), 0, s, C, ldc, A, lda, B, ldb, M, N, K); break;
    case 32: hipLaunchKernelGGL((k_gemm<T, true, AC, 32>), grid, dim3(256), 0, s, C, ldc, A, lda, B, ldb, M, N, K); break;
    default: hipLaunchKernelGGL((k_gemm<T, true, AC, 8>), grid, dim3(256), 0, s, C, ldc, A, lda, B, ldb, M, N, K);
    }
}

template <typename T>
void launch_update(hipStream_t s, dim3 grid, T *C, long ldc, const T *A, long lda, const T *B, long ldb, long M,
                   long N, long K)
{
    if constexpr (sizeof(T) == 8) {
        if (g_update_fp32.load(std::memory_order_relaxed)) {
            launch_update_ac<T, float2>(s, grid, C, ldc, A, lda, B, ldb, M, N, K);
            return;
        }
    }
    launch_update_ac<T, double2>(s, grid, C, ldc, A, lda, B, ldb, M, N, K);
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
// B: n x nrhs row-major (ldb). The blocks of the solution are published in order; block b takes the contributions
// of the blocks before it as their flags come up, and solves its diagonal block by substitution (one wave, a row a
// lane): an explicit inverse of a diagonal block loses the accuracy of a factor with small pivots, which the
// refinement needs.
constexpr int SB = 64;
constexpr int MAXRHS = 8;

// L21 = A21 U11^-1 by substitution (x U11 = a for each row a of A21): a thread a row of [r0, r0 + m), U11 (the
// upper part of top) in LDS; the 64 columns in two halves of 32 so a row's half stays in registers for either
// element type.
template <typename T>
__device__ __forceinline__ void l21_body(T *A, long lda, long r0, long m, long j0, int jb, const T *top, long bx)
{
    __shared__ T Us[TW][TW + 1];
    __shared__ T rd[TW];
    const int tid = threadIdx.x;
    {
        T v[TW * TW / 256];                     // every load in flight before the first store
#pragma unroll
        for (int u = 0; u < TW * TW / 256; ++u) v[u] = top[tid + u * 256];
#pragma unroll
        for (int u = 0; u < TW * TW / 256; ++u) {
            const int e = tid + u * 256, c = e / TW, i = e % TW;
            Us[i][c] = v[u];
        }
    }
    __syncthreads();
    if (tid < TW) rd[tid] = tid < jb ? trecip(Us[tid][tid]) : CX<T>::make(1, 0);
    __syncthreads();
    const long r = r0 + bx * 256 + tid;
    if (r >= r0 + m) return;
    T *row = A + j0 * lda + r;                  // element (r, j0 + c) at row[c * lda]
    auto half = [&](auto hc) __attribute__((always_inline)) {
        constexpr int h = decltype(hc)::value;
        RowH<T> x;
        auto ld = [&](auto jc) __attribute__((always_inline)) {
            constexpr int j = decltype(jc)::value;
            x.set(j, ldz(32 * h + j < jb, row + (long)(32 * h + j) * lda));
        };
        Unroll<0, TH>::run(ld);
        if (h == 1) {                           // the first half's solution, final, from the previous pass
            for (int k0 = 0; k0 < 32 && k0 < jb; k0 += 8) {
                T xk[8];                        // eight of its values read together
#pragma unroll
                for (int u = 0; u < 8; ++u) xk[u] = ldz(k0 + u < jb, row + (long)(k0 + u) * lda);
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int k = k0 + u;
                    if (k >= jb) break;
                    auto up = [&](auto jc) __attribute__((always_inline)) {
                        constexpr int j = decltype(jc)::value;
                        x.set(j, tsub(x.get(j), tmul(xk[u], Us[k][32 + j])));
                    };
                    Unroll<0, TH>::run(up);
                }
            }
        }
        auto solve = [&](auto jc) __attribute__((always_inline)) {
            constexpr int j = decltype(jc)::value;
            if (32 * h + j < jb) {
                T v = x.get(j), v1 = CX<T>::make(0, 0);      // two partial sums: half the dependent chain
                auto sub = [&](auto kc) __attribute__((always_inline)) {
                    constexpr int k = decltype(kc)::value;
                    if constexpr (k % 2 == 0) v = tsub(v, tmul(x.get(k), Us[32 * h + k][32 * h + j]));
                    else v1 = tadd(v1, tmul(x.get(k), Us[32 * h + k][32 * h + j]));
                };
                Unroll<0, j>::run(sub);
                x.set(j, tmul(tsub(v, v1), rd[32 * h + j]));
            }
        };
        Unroll<0, TH>::run(solve);
        auto st = [&](auto jc) __attribute__((always_inline)) {
            constexpr int j = decltype(jc)::value;
            if (32 * h + j < jb) row[(long)(32 * h + j) * lda] = x.get(j);
        };
        Unroll<0, TH>::run(st);
    };
    Unroll<0, 2>::run(half);
}

template <typename T>
__global__ __launch_bounds__(256) void k_l21(T *A, long lda, long r0, long m, long j0, int jb, const T *top)
{
    l21_body<T>(A, lda, r0, m, j0, jb, top, blockIdx.x);
}

// The panel's row moves on columns [c0, c0 + nc): a column per 128 threads, every source read before any write;
// the panel's own columns [j0, j0 + jb) then take the root's L11 \ U11 in rows [j0, j0 + jb). With perm, one more
// workgroup moves the entries of the row permutation (position -> original row) the same way.
template <typename T>
__device__ __forceinline__ void apply_moves_body(T *A, long lda, long c0, long nc, const int *pairs, long j0, int jb,
                                                 const T *top, int *perm, long bx, long perm_block)
{
    const int t = threadIdx.x & 127, h = threadIdx.x >> 7;
    const int np = pairs[0];
    const int dst = t < np ? pairs[1 + 2 * t] : 0, src = t < np ? pairs[2 + 2 * t] : 0;
    if (perm && bx == perm_block) {                 // the extra workgroup: the same moves on the row permutation
        int v = 0;
        if (h == 0 && t < np) v = perm[src];
        __syncthreads();
        if (h == 0 && t < np) perm[dst] = v;
        return;
    }
    T v[4];                                         // the workgroup's 8 columns, two at a time: all read first
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const long cb = bx * 8 + 2 * u;
        v[u] = ldz(cb + h < nc && t < np, A + (c0 + cb + h) * lda + src);
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const long cb = bx * 8 + 2 * u, c = c0 + cb + h;
        if (cb + h < nc && t < np) A[c * lda + dst] = v[u];
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const long cb = bx * 8 + 2 * u, c = c0 + cb + h;
        if (cb + h < nc && c >= j0 && c < j0 + jb)
            for (int i = t; i < jb; i += 128) A[c * lda + j0 + i] = top[(c - j0) * TW + i];
    }
}

// grid: blocks(nc, 8) workgroups of columns, and with perm one more (the last) for the permutation
template <typename T>
__global__ __launch_bounds__(256) void k_apply_moves(T *A, long lda, long c0, long nc, const int *pairs, long j0,
                                                     int jb, const T *top, int *perm)
{
    apply_moves_body<T>(A, lda, c0, nc, pairs, j0, jb, top, perm, blockIdx.x, (long)gridDim.x - 1);
}

// kind 0: U^T y = b (forward); 1: L^T z = y (backward); 2: L y = b (forward); 3: U x = y (backward).
// Operator entry (row i, column k): kinds 0, 1 read LU[i * lda + k] (U[k][i], L[k][i]); kinds 2, 3 LU[k * lda + i].
__host__ __device__ inline bool kind_fwd(int kind) { return kind == 0 || kind == 2; }

__device__ inline float2 readlane_c(float2 v, int l)
{
    return make_float2(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.x), l)),
                       __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.y), l)));
}
__device__ inline double readlane_d(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ inline double2 readlane_c(double2 v, int l) { return make_double2(readlane_d(v.x, l), readlane_d(v.y, l)); }

// t - m y with four fused multiply-adds
__device__ inline float2 cfms(float2 t, float2 m, float2 y)
{
    return make_float2(fmaf(-m.x, y.x, fmaf(m.y, y.y, t.x)), fmaf(-m.y, y.x, fmaf(-m.x, y.y, t.y)));
}
__device__ inline double2 cfms(double2 t, double2 m, double2 y)
{
    return make_double2(fma(-m.x, y.x, fma(m.y, y.y, t.x)), fma(-m.y, y.x, fma(-m.x, y.y, t.y)));
}

// ---- values that carry their own ready mark ----------------------------------------------------------------
// The solves hand results between workgroups through buffers filled beforehand with all-ones 64-bit words (a NaN
// pattern no arithmetic produces): every 64-bit word is written and read with single-copy atomic accesses at
// device scope, so a reader that finds no all-ones word in a value has the value itself, with no flag and no
// second round trip. A computed word that happens to be all ones is written as the canonical NaN instead.
typedef unsigned long long u64;
constexpr u64 UNSET = ~0ull;

__device__ inline u64 mark_safe(u64 w, u64 nan) { return w == UNSET ? nan : w; }
__device__ inline void put(float2 *p, float2 v)
{
    const u64 w = ((u64)__float_as_uint(v.y) << 32) | __float_as_uint(v.x);
    __hip_atomic_store((u64 *)p, mark_safe(w, 0x7fc000007fc00000ull), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void put(double2 *p, double2 v)
{
    u64 *q = (u64 *)p;
    __hip_atomic_store(q, mark_safe((u64)__double_as_longlong(v.x), 0x7ff8000000000000ull), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(q + 1, mark_safe((u64)__double_as_longlong(v.y), 0x7ff8000000000000ull), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}
// the value at p (all-ones words if it has not been written yet), and whether it has been written
__device__ inline float2 fetch(const float2 *p)
{
    const u64 w = __hip_atomic_load((u64 *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return make_float2(__uint_as_float((unsigned)w), __uint_as_float((unsigned)(w >> 32)));
}
__device__ inline double2 fetch(const double2 *p)
{
    const u64 a = __hip_atomic_load((u64 *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const u64 b = __hip_atomic_load((u64 *)p + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return make_double2(__longlong_as_double((long long)a), __longlong_as_double((long long)b));
}
__device__ inline bool is_set(float2 v) { return (__float_as_uint(v.x) & __float_as_uint(v.y)) != ~0u; }
__device__ inline bool is_set(double2 v)
{
    return (u64)__double_as_longlong(v.x) != UNSET && (u64)__double_as_longlong(v.y) != UNSET;
}
template <typename T> __device__ inline bool take(const T *p, T &v)
{
    v = fetch(p);
    return is_set(v);
}
__device__ inline bool wave_all(bool c) { return __builtin_amdgcn_ballot_w64(!c) == 0; }

// ---- the triangular solve ----------------------------------------------------------------------------------
constexpr int TB = 512;                  // threads of a solve workgroup: 8 waves
constexpr int NWV = TB / 64;
constexpr int TE = SB * SB / TB;         // entries of a 64 x 64 tile a thread takes: 8
template <typename T> struct TrsvDepth { static constexpr int D = 3; };  // tiles a workgroup keeps in flight
template <> struct TrsvDepth<double2> { static constexpr int D = 2; };      // ahead of the values they meet
template <typename T> struct TrsvMaxRhs { static constexpr int NR = 8; };    // right-hand sides a launch takes
template <> struct TrsvMaxRhs<double2> { static constexpr int NR = 4; };

// Unit-lower (FWD) or unit-upper substitution of NW right-hand sides, lane i row i, with the block's operator Ms
// in LDS (rows pre-scaled to a unit diagonal; the diagonal, the other triangle and the padding zero): step k
// broadcasts y_k and every lane subtracts its multiple, the zeros leaving the finished rows alone. After each 8
// steps the 8 rows they finished (lanes [lo, lo + 8)) are final: pub(lo) hands them on while the rest go on. The
// row's operator entries are read from LDS one group of 8 steps ahead.
template <bool FWD, int NW, typename T, class Pub>
__device__ __attribute__((always_inline)) void subst(T (&t)[NW], const T (*Ms)[SB + 1], int lane, Pub &pub)
{
    T mk[2][8];
#pragma unroll
    for (int j = 0; j < 8; ++j) mk[0][j] = Ms[lane][FWD ? j : SB - 1 - j];
#pragma unroll
    for (int s0 = 0; s0 < SB; s0 += 8) {
        const int g = (s0 / 8) & 1;
        if (s0 + 8 < SB) {
#pragma unroll
            for (int j = 0; j < 8; ++j) mk[g ^ 1][j] = Ms[lane][FWD ? s0 + 8 + j : SB - 9 - s0 - j];
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int k = FWD ? s0 + j : SB - 1 - s0 - j;
#pragma unroll
            for (int w = 0; w < NW; ++w) t[w] = cfms(t[w], mk[g][j], readlane_c(t[w], k));
        }
        pub(FWD ? s0 : SB - 8 - s0);
    }
}

// One launch of nb workgroups of 512 threads solves one triangle for nrhs <= NR right-hand sides. Workgroup q
// owns the q-th 64-row diagonal block in solve order (the chain passes from workgroup to workgroup):
//  - at its start it stages its diagonal block's operator in LDS (strict triangle, rows scaled by 1 / U_ii) and
//    reads its right-hand sides;
//  - it then meets the earlier blocks a in order: thread (row lane, columns 8 wv + u of block a) holds its
//    entries of Op_qa, loaded TrsvDepth blocks ahead, and subtracts Op_qa y_a from its partial sums once y_a is
//    out (fused multiply-subtracts: four per complex term);
//  - after the last one (a = q - 1) the partial sums meet in LDS, waves 0-3 form the block's right-hand side
//    (wave w takes right-hand sides w and, for NR = 8, w + 4; lane i row i), substitute, and publish y_q eight
//    rows at a time as the substitution finishes them: the next workgroup's wave that takes those rows starts its
//    product with them while the substitution goes on, so the chain's hand-off overlaps the substitution.
// y travels as self-marking values (put / take) in dst: a reader spins on the values themselves, every lane on its
// own when the block is one of the two before its own (the chain waits on it), one lane with growing naps first
// otherwise. Workgroup q waits only on workgroups before it, which start no later, so the launch makes progress
// with any number of them resident.
//
// rdg: 1 / U_ii (read for the kinds with U only: L is unit). src: n x nrhs (leading dimension lds); dst: n x nrhs
// (leading dimension nrhs) filled with UNSET words beforehand. abort: raised if a wait exceeds its bound (the result
// is then garbage, not a hang). NR: 4 (nrhs <= 4) or 8.
template <typename T, int NR>
__device__ __forceinline__ void trsv_body(const T *LU, long lda, long n, const T *rdg, const T *src, long lds, T *dst,
                                          int nrhs, int kind, int *abort, int q)
{
    constexpr int NW = NR / 4;               // right-hand sides a solver wave takes
    constexpr int D = TrsvDepth<T>::D;
    __shared__ T Ms[SB][SB + 1];             // the diagonal block's operator
    __shared__ T red[NWV][NR][SB];           // the waves' partial sums
    __shared__ T yb[NWV][64];                // each wave's y values of the block it meets
    const int nb = (int)((n + SB - 1) / SB);
    if (q >= nb) return;
    const bool fwd = kind_fwd(kind), rowmajor = kind <= 1, unit = kind == 1 || kind == 2;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const T zero = CX<T>::make(0, 0), one = CX<T>::make(1, 0);
    auto order = [&](int p) { return fwd ? p : nb - 1 - p; };
    auto opp = [&](long gi, long gk) { return rowmajor ? LU + gi * lda + gk : LU + gk * lda + gi; };
    const long kstep = rowmajor ? 1 : lda;   // address step along a row of the operator
    const int b = order(q);
    const long r0 = (long)b * SB;
    const int m = (int)std::min((long)SB, n - r0);          // rows of this block inside the matrix

    // the diagonal block: thread (rows 8 wv + u, column lane) of a row-major operator, (row lane, columns 8 wv + u)
    // of a column-major one, so that every load reads whole lines; indices clamped into the block, the entries
    // outside it or outside the strict triangle zeroed when staged
    T md[TE], mr[TE];
#pragma unroll
    for (int u = 0; u < TE; ++u) {
        const int i = rowmajor ? wv * TE + u : lane, k = rowmajor ? lane : wv * TE + u;
        const long gi = r0 + std::min(i, m - 1), gk = r0 + std::min(k, m - 1);
        md[u] = *opp(gi, gk);
        mr[u] = rdg[gi];                     // read either way: a select of values, not of addresses
    }
    // the solver waves' right-hand sides and 1 / U_ii of the lane's row (rows past n: zeroed where used), parked in
    // LDS until the substitution: held in registers across the loop they are spilled to scratch for complex128, and
    // reloading them would put a memory round trip into every step of the chain
    __shared__ T bxs[NR][SB], rds[SB];
    int rr[NW];
#pragma unroll
    for (int w = 0; w < NW; ++w) rr[w] = (wv & 3) + 4 * w;
    if (wv < 4) {
        const long gi = r0 + std::min(lane, m - 1);
#pragma unroll
        for (int w = 0; w < NW; ++w) bxs[rr[w]][lane] = src[gi * lds + std::min(rr[w], nrhs - 1)];
        if (wv == 0) rds[lane] = csel(unit, one, rdg[gi]);
    }

    // the earlier blocks' entries: thread (row lane, columns 8 wv + u of block a); a row past the matrix reads the
    // last row (its sums are never used), columns past it meet zero y
    auto load_tile = [&](T (&mt)[TE], int a) __attribute__((always_inline)) {
        const long gi = r0 + std::min(lane, m - 1), k0 = (long)a * SB + wv * TE;
        const T *p = opp(gi, k0);
        if (k0 + TE <= n) {
#pragma unroll
            for (int u = 0; u < TE; ++u) mt[u] = p[u * kstep];
        } else {                             // columns from n on read column n - 1 (ul < 0 past the matrix)
            const int ul = (int)(n - 1 - k0);
#pragma unroll
            for (int u = 0; u < TE; ++u) mt[u] = p[(u < ul ? u : ul) * kstep];
        }
    };
    auto stage = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int u = 0; u < TE; ++u) {
            const int i = rowmajor ? wv * TE + u : lane, k = rowmajor ? lane : wv * TE + u;
            const bool in = i < m && k < m && (fwd ? k < i : k > i);
            Ms[i][k] = csel(in, csel(unit, md[u], tmul(md[u], mr[u])), zero);
        }
    };
    // complex128 stages the diagonal block before its first tiles are asked for (registers), complex64 after
    if constexpr (sizeof(T) == 16) stage();
    T m0[TE], m1[TE], m2[TE];
    if (q > 0) load_tile(m0, order(0));
    if (q > 1) load_tile(m1, order(1));
    if (D > 2 && q > 2) load_tile(m2, order(2));
    if constexpr (sizeof(T) != 16) stage();

    T s[NR];                                 // minus the wave's share of the products
#pragma unroll
    for (int r = 0; r < NR; ++r) s[r] = zero;
    // lane l < 8 NR of wave wv takes y of row 8 wv + l / NR of block a, right-hand side l % NR (zero for the
    // right-hand sides past nrhs)
    const int lu_ = lane / NR, lr = lane % NR;
    T *ybw = yb[wv];
    auto step = [&](int p, T (&mt)[TE]) __attribute__((always_inline)) {
        const int a = order(p);
        const long g0 = (long)a * SB + wv * TE, gk = g0 + lu_;
        const bool need = lane < TE * NR && lr < nrhs && gk < n;
        if (q - p > 2 && lane == 0 && g0 < n) {   // far from the chain: one lane watches the slice's first value
            const T *w0 = dst + g0 * nrhs;         // with growing naps first (a slice past n needs none)
            T v;
            long spins = 0;
            int nap = 1;
            while (!take(w0, v)) {
                for (int z = 0; z < nap; ++z) __builtin_amdgcn_s_sleep(2);
                nap = nap < 8 ? 2 * nap : 8;
                if (++spins > (1L << 22)) {
                    __hip_atomic_store(abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
            }
        }                                          // near it, every lane polls its own value from the start
        T yv = zero;
        bool ok = !need || take(dst + gk * nrhs + lr, yv);
        long spins = 0;
        while (!wave_all(ok)) {
            __builtin_amdgcn_s_sleep(1);
            if (++spins > (1L << 22)) {
                __hip_atomic_store(abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            if (!ok) ok = take(dst + gk * nrhs + lr, yv);
        }
        if (!need) yv = zero;
        ybw[lane] = yv;
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int u = 0; u < TE; ++u) {
#pragma unroll
            for (int r = 0; r < NR; ++r) s[r] = cfms(s[r], mt[u], ybw[u * NR + r]);
            if constexpr (sizeof(T) == 16) __builtin_amdgcn_sched_barrier(0);     // y read as it is used
        }
        __builtin_amdgcn_wave_barrier();
        if (p + D < q) load_tile(mt, order(p + D));
    };
    for (int p = 0; p < q; p += D) {
        step(p, m0);
        if (p + 1 < q) step(p + 1, m1);
        if (D > 2 && p + 2 < q) step(p + 2, m2);
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) red[wv][r][lane] = s[r];
    __syncthreads();
    if (wv >= 4) return;
    const bool inrow = lane < m;
    T t[NW];
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        T h[NWV];                            // the 8 waves' partial sums, added as a tree (depth 3)
#pragma unroll
        for (int sl = 0; sl < NWV; ++sl) h[sl] = red[sl][rr[w]][lane];
#pragma unroll
        for (int st = 1; st < NWV; st *= 2)
#pragma unroll
            for (int sl = 0; sl + st < NWV; sl += 2 * st) h[sl] = tadd(h[sl], h[sl + st]);
        const T v = h[0];
        t[w] = csel(inrow && rr[w] < nrhs, tmul(tadd(bxs[rr[w]][lane], v), rds[lane]), zero);
    }
    auto pub = [&](int lo) __attribute__((always_inline)) {
        if (lane >= lo && lane < lo + 8 && inrow) {
#pragma unroll
            for (int w = 0; w < NW; ++w)
                if (rr[w] < nrhs) put(dst + (r0 + lane) * nrhs + rr[w], t[w]);
        }
    };
    if (fwd) subst<true, NW>(t, Ms, lane, pub);
    else subst<false, NW>(t, Ms, lane, pub);
}

template <typename T, int NR>
__global__ __launch_bounds__(TB) void k_trsv(const T *LU, long lda, long n, const T *rdg, const T *src, long lds,
                                             T *dst, int nrhs, int kind, int *abort)
{
    trsv_body<T, NR>(LU, lda, n, rdg, src, lds, dst, nrhs, kind, abort, blockIdx.x);
}

// 1 / U_ii for the substitutions with U and U^T
template <typename T>
__global__ void k_rdiag(const T *LU, long lda, long n, T *rd)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) rd[i] = trecip(LU[i * lda + i]);
}

// out[perm[i]] = in[i] (scatter, trans) or out[i] = in[perm[i]] (gather); perm NULL: a copy. Leading dimensions
// ldi (in) and ldo (out).
template <typename T>
__global__ void k_permute(const T *in, long ldi, T *out, long ldo, long n, int nrhs, const int *perm, int scatter)
{
    const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (idx >= n * nrhs) return;
    const long i = idx / nrhs;
    const int r = (int)(idx % nrhs);
    const long p = perm ? (long)perm[i] : i;
    if (scatter) out[p * ldo + r] = in[i * ldi + r];
    else out[i * ldo + r] = in[p * ldi + r];
}

__global__ void k_fill_unset(unsigned long long *p, long n)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) p[i] = ~0ull;
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
        const int dt = dtype_of<T>();
        const double es = sizeof(T);
        {
            TimedLaunch tl(s, M25X_K_PANEL, dt, 2.0 * es * m * jb, 0.0);
            M25X_HIP(hipLaunchCooperativeKernel((const void *)k_panel<T>, dim3(G), dim3(PT), args, (unsigned)smem, s));
        }
        const long nsw = n - jb;
        if (nsw > 0) {
            TimedLaunch tl(s, M25X_K_LASWP, dt, 2.0 * es * 2 * jb * nsw, 0.0);
            hipLaunchKernelGGL(k_laswp<T>, dim3(blocks(nsw, 256)), dim3(256), 0, s, A, lda, n, j0, jb, ipiv);
            if (int rc = check_launch("getrf/laswp")) return rc;
        }
        const long c0 = j0 + jb;
        if (c0 < n) {
            {
                TimedLaunch tl(s, M25X_K_TRSM, dt, 2.0 * es * (n - c0) * jb, 4.0 * (n - c0) * jb * jb);
                hipLaunchKernelGGL((k_trsm<T, NB>), dim3(blocks(n - c0, TC)), dim3(256), 0, s, A, lda, n, j0, jb, c0);
            }
            if (int rc = check_launch("getrf/trsm")) return rc;
            const long M = n - c0;
            dim3 grid(blocks(M, GM), blocks(M, GN));
            {
                TimedLaunch tl(s, M25X_K_GEMM, dt, es * (2.0 * M * M + 2.0 * M * jb), 8.0 * M * M * jb);
                hipLaunchKernelGGL((k_gemm<T, true>), grid, dim3(256), 0, s, A + c0 * lda + c0, lda, A + j0 * lda + c0,
                                   lda, A + c0 * lda + j0, lda, M, M, (long)jb);
            }
            if (int rc = check_launch("getrf/gemm")) return rc;
        }
    }
    return M25X_OK;
}

// ---- look-ahead: the next panel's tournament beside the trailing update -----------------------------------
// The tournament of panel j + 1 needs only panel j + 1's columns updated, so the update of step j is split: the next
// panel's 64 columns first, then the rest, and the tournament runs on a second stream beside the rest (its waves
// raise their own issue priority on the SIMDs they share with the update, s_setprio; both streams keep the default
// queue priority). Off by default (m25x_lu_set_lookahead): one extra stream per device, created once.
std::atomic<int> g_lookahead{0};

struct LookAhead {
    hipStream_t p = nullptr;                  // the panel stream; the update runs on the caller's
    hipEvent_t ev[4] = {};
    bool ok = false;
};

LookAhead *lookahead_streams()
{
    static std::mutex mu;
    static LookAhead la[64];
    int dev = 0;
    if (!g_lookahead.load(std::memory_order_relaxed) || hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64)
        return nullptr;
    std::lock_guard<std::mutex> lk(mu);
    LookAhead &L = la[dev];
    if (L.ok) return &L;
    if (hipStreamCreateWithFlags(&L.p, hipStreamNonBlocking) != hipSuccess) return nullptr;
    for (auto &e : L.ev)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
    L.ok = true;
    return &L;
}

template <typename T> size_t tour_work_bytes(long n)
{
    const long nn = tour_nodes(n, Tour<T>::RL, Tour<T>::AR);
    return nn * sizeof(TNode<T>) + (size_t)TW * TW * sizeof(T) + sizeof(int) * ((size_t)nn + 4 * TW + 64);
}

// The tournament LU with one panel of look-ahead (see LookAhead): the same kernels and arithmetic as the plain loop
// (each trailing update split by columns into two launches), so the factor is the same bit for bit. The caller's
// stream runs the row moves, L21, U12 and the update, stream p the tournaments; events order them.
template <typename T>
int getrf_tp_lookahead(T *A, long n, long lda, int *ipiv, int *perm, int *info, TNode<T> *nodes, int *arrive, T *top,
                       int *pairs, LookAhead &la, hipStream_t s)
{
    constexpr int RL = Tour<T>::RL;
    const int dt = dtype_of<T>();
    const double es = sizeof(T);
    hipStream_t g = s, p = la.p;
    hipEvent_t start = la.ev[0], panel = la.ev[1], next = la.ev[2], done = la.ev[3];
    if (hipEventRecord(start, s) != hipSuccess || hipStreamWaitEvent(p, start, 0) != hipSuccess)
        return fail(M25X_EHIP, "getrf: look-ahead stream");
    auto tournament = [&](long j0) -> int {
        const int jb = (int)std::min((long)TW, n - j0);
        const long m = n - j0;
        TimedLaunch tl(p, M25X_K_TSLU, dt, es * m * jb, 0.0);
        if (int rc = launch_tslu((const T *)A, lda, n, j0, jb, (int)tour_leaves(m, RL), nodes, arrive, top, ipiv, pairs,
                                 info, p))
            return rc;
        return hipEventRecord(panel, p) == hipSuccess ? M25X_OK : fail(M25X_EHIP, "getrf: event");
    };
    auto update = [&](long j0, int jb, long c0, long c1, long c2) {       // A22[c0:, c1:c2] -= L21 U12[:, c1:c2]
        const long M = n - c0, N = c2 - c1;
        TimedLaunch tl(g, M25X_K_GEMM, dt, es * (2.0 * M * N + (M + N) * jb), 8.0 * M * N * jb);
        launch_update<T>(g, dim3(blocks(M, GM), blocks(N, GN)), A + c1 * lda + c0, lda,
                         (const T *)(A + j0 * lda + c0), lda, (const T *)(A + c1 * lda + j0), lda, M, N, (long)jb);
    };
    if (int rc = tournament(0)) return rc;
    for (long j0 = 0; j0 < n; j0 += TW) {
        const int jb = (int)std::min((long)TW, n - j0);
        const long m = n - j0;
        if (hipStreamWaitEvent(g, panel, 0) != hipSuccess) return fail(M25X_EHIP, "getrf: wait");
        {
            TimedLaunch tl(g, M25X_K_APPLY_MOVES, dt, 2.0 * es * 2 * jb * n, 0.0);
            hipLaunchKernelGGL(k_apply_moves<T>, dim3(blocks(n, 8) + (perm ? 1 : 0)), dim3(256), 0, g, A, lda, 0L, n,
                               (const int *)pairs, j0, jb, (const T *)top, perm);
        }
        if (m > jb) {
            TimedLaunch tl(g, M25X_K_L21, dt, 2.0 * es * (m - jb) * jb, 4.0 * (m - jb) * jb * jb);
            hipLaunchKernelGGL(k_l21<T>, dim3(blocks(m - jb, 256)), dim3(256), 0, g, A, lda, j0 + jb, m - jb, j0, jb,
                               (const T *)top);
        }
        const long c0 = j0 + jb;
        if (c0 < n) {
            {
                TimedLaunch tl(g, M25X_K_TRSM, dt, 2.0 * es * (n - c0) * jb, 4.0 * (n - c0) * jb * jb);
                hipLaunchKernelGGL((k_trsm<T, TW>), dim3(blocks(n - c0, TC)), dim3(256), 0, g, A, lda, n, j0, jb, c0);
            }
            const long c1 = std::min(n, c0 + TW);
            update(j0, jb, c0, c0, c1);                                   // the next panel's columns
            if (hipEventRecord(next, g) != hipSuccess || hipStreamWaitEvent(p, next, 0) != hipSuccess)
                return fail(M25X_EHIP, "getrf: event");
            if (int rc = tournament(c0)) return rc;                       // beside the rest of the update
            if (c1 < n) update(j0, jb, c0, c1, n);
        }
        if (int rc = check_launch("getrf/look-ahead")) return rc;
    }
    (void)done;                               // the last panel's tournament was waited for on the caller's stream
    return M25X_OK;
}

template <typename T>
int getrf_tp(T *A, long n, long lda, int *ipiv, int *perm, int *info, void *work, size_t wbytes, hipStream_t s)
{
    constexpr int RL = Tour<T>::RL, AR = Tour<T>::AR;
    if (wbytes < tour_work_bytes<T>(n))
        return fail(M25X_EINVAL, "getrf: workspace of %zu bytes, %zu needed", wbytes, tour_work_bytes<T>(n));
    const long nn = tour_nodes(n, RL, AR);
    char *w = (char *)work;
    TNode<T> *nodes = (TNode<T> *)w;
    w += nn * sizeof(TNode<T>);
    T *top = (T *)w;
    w += (size_t)TW * TW * sizeof(T);
    int *arrive = (int *)w, *pairs = arrive + nn;
    hipLaunchKernelGGL(k_zero_ints, dim3(blocks(nn, 256)), dim3(256), 0, s, arrive, (int)nn);
    if (info) hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, s, info, 1);
    if (perm) hipLaunchKernelGGL(k_iota, dim3(blocks(n, 256)), dim3(256), 0, s, perm, n);
    if (int rc = check_launch("getrf")) return rc;
    if (n > 2 * TW)
        if (LookAhead *la = lookahead_streams())
            return getrf_tp_lookahead<T>(A, n, lda, ipiv, perm, info, nodes, arrive, top, pairs, *la, s);
    for (long j0 = 0; j0 < n; j0 += TW) {
        const int jb = (int)std::min((long)TW, n - j0);
        const long m = n - j0;
        const int P = (int)tour_leaves(m, RL);
        const int dt = dtype_of<T>();
        const double es = sizeof(T);
        {
            TimedLaunch tl(s, M25X_K_TSLU, dt, es * m * jb, 0.0);
            if (int rc = launch_tslu((const T *)A, lda, n, j0, jb, P, nodes, arrive, top, ipiv, pairs, info, s))
                return rc;
        }
        {
            TimedLaunch tl(s, M25X_K_APPLY_MOVES, dt, 2.0 * es * 2 * jb * n, 0.0);
            hipLaunchKernelGGL(k_apply_moves<T>, dim3(blocks(n, 8) + (perm ? 1 : 0)), dim3(256), 0, s, A, lda, 0L, n,
                               (const int *)pairs, j0, jb, (const T *)top, perm);
        }
        if (int rc = check_launch("getrf/panel")) return rc;
        if (m > jb) {   // L21 = A21 U11^-1
            TimedLaunch tl(s, M25X_K_L21, dt, 2.0 * es * (m - jb) * jb, 4.0 * (m - jb) * jb * jb);
            hipLaunchKernelGGL(k_l21<T>, dim3(blocks(m - jb, 256)), dim3(256), 0, s, A, lda, j0 + jb, m - jb, j0, jb,
                               (const T *)top);
        }
        const long c0 = j0 + jb;
        if (c0 < n) {   // U12 = L11^-1 A12, A22 -= L21 U12
            T *A12 = A + c0 * lda + j0;
            {
                TimedLaunch tl(s, M25X_K_TRSM, dt, 2.0 * es * (n - c0) * jb, 4.0 * (n - c0) * jb * jb);
                hipLaunchKernelGGL((k_trsm<T, TW>), dim3(blocks(n - c0, TC)), dim3(256), 0, s, A, lda, n, j0, jb, c0);
            }
            const long M = n - c0;
            TimedLaunch tl(s, M25X_K_GEMM, dt, es * (2.0 * M * M + 2.0 * M * jb), 8.0 * M * M * jb);
            launch_update<T>(s, dim3(blocks(M, GM), blocks(M, GN)), A + c0 * lda + c0, lda, (const T *)(A + j0 * lda + c0),
                             lda, (const T *)A12, lda, M, M, (long)jb);
        }
        if (int rc = check_launch("getrf/update")) return rc;
    }
    return M25X_OK;
}

// Solve workspace: y of the two solves (W1, W2), written as self-marking values over UNSET words; the gathered
// right-hand sides (W0); 1 / U_ii; the permutation, the abort word (at the passed work_bytes - 256, rounded down
// to an int: callers read it there)
// and 63 spare ints.
template <typename T> size_t getrs_work(long n, long)
{
    return sizeof(T) * (3 * (size_t)n * MAXRHS + (size_t)n) + sizeof(int) * ((size_t)n + 64);
}

template <typename T>
int trsv_launch(hipStream_t s, int nb, const T *LU, long lda, long n, const T *rd, const T *src, long lds, T *dst,
                int nrhs, int kind, int *abort)
{
    const double tri = kind_fwd(kind) ? 0.5 * n * (n - 1) : 0.5 * n * (n + 1);
    TimedLaunch tl(s, M25X_K_TRSV, dtype_of<T>(), (tri + 2.0 * n * nrhs) * sizeof(T), 8.0 * tri * nrhs);
    if (nrhs <= 4)
        hipLaunchKernelGGL((k_trsv<T, 4>), dim3(nb), dim3(TB), 0, s, LU, lda, n, rd, src, lds, dst, nrhs, kind, abort);
    else
        hipLaunchKernelGGL((k_trsv<T, TrsvMaxRhs<T>::NR>), dim3(nb), dim3(TB), 0, s, LU, lda, n, rd, src, lds, dst,
                           nrhs, kind, abort);
    return check_launch("getrs/trsv");
}

// One solve of at most TrsvMaxRhs<T>::NR right-hand sides; the abort word is the caller's (zeroed once per call).
template <typename T>
int getrs_part(const T *LU, long n, long lda, const int *perm_in, T *B, long ldb, int nrhs, int trans, T *W1, T *W2,
               T *W0, const T *rd, int *abort, hipStream_t s)
{
    const int nb = (int)((n + SB - 1) / SB);
    const size_t wn = (size_t)n * MAXRHS;
    const long words = (long)(2 * wn * sizeof(T) / 8);
    hipLaunchKernelGGL(k_fill_unset, dim3(blocks(words, 256)), dim3(256), 0, s, (unsigned long long *)W1, words);
    if (int rc = check_launch("getrs")) return rc;
    const unsigned pg = blocks(n * nrhs, 256);
    int rc;
    if (trans) {
        if ((rc = trsv_launch<T>(s, nb, LU, lda, n, rd, B, ldb, W1, nrhs, 0, abort))) return rc;
        if ((rc = trsv_launch<T>(s, nb, LU, lda, n, rd, W1, nrhs, W2, nrhs, 1, abort))) return rc;
        hipLaunchKernelGGL(k_permute<T>, dim3(pg), dim3(256), 0, s, (const T *)W2, (long)nrhs, B, ldb, n, nrhs, perm_in, 1);
    } else {
        hipLaunchKernelGGL(k_permute<T>, dim3(pg), dim3(256), 0, s, (const T *)B, ldb, W0, (long)nrhs, n, nrhs, perm_in, 0);
        if ((rc = trsv_launch<T>(s, nb, LU, lda, n, rd, W0, nrhs, W1, nrhs, 2, abort))) return rc;
        if ((rc = trsv_launch<T>(s, nb, LU, lda, n, rd, W1, nrhs, W2, nrhs, 3, abort))) return rc;
        hipLaunchKernelGGL(k_permute<T>, dim3(pg), dim3(256), 0, s, (const T *)W2, (long)nrhs, B, ldb, n, nrhs,
                           (const int *)nullptr, 0);
    }
    return check_launch("getrs");
}

// The solve with the buffers at buf (getrs_work(n) - 256 bytes: W1, W2, W0, 1 / U_ii, the permutation) and the
// abort word given (not reset here): in parts of at most TrsvMaxRhs<T>::NR right-hand sides.
template <typename T>
int getrs_noreset(const T *LU, long n, long lda, const int *ipiv, const int *perm_in, T *B, long ldb, int nrhs,
                  int trans, char *buf, int *abort, hipStream_t s)
{
    const size_t wn = (size_t)n * MAXRHS;
    T *W1 = (T *)buf, *W2 = W1 + wn, *W0 = W2 + wn, *rd = W0 + wn;
    int *perm = (int *)(rd + n);
    hipLaunchKernelGGL(k_rdiag<T>, dim3(blocks(n, 256)), dim3(256), 0, s, LU, lda, n, rd);
    if (!perm_in) {
        hipLaunchKernelGGL(k_perm, dim3(1), dim3(256), 0, s, ipiv, n, perm);
        perm_in = perm;
    }
    if (int rc = check_launch("getrs")) return rc;
    for (int c0 = 0; c0 < nrhs; c0 += TrsvMaxRhs<T>::NR) {
        const int kc = std::min(TrsvMaxRhs<T>::NR, nrhs - c0);
        if (int rc = getrs_part<T>(LU, n, lda, perm_in, B + c0, ldb, kc, trans, W1, W2, W0, rd, abort, s)) return rc;
    }
    return M25X_OK;
}

// The solve: buffers carved from `work` (W1, W2, W0: n x MAXRHS each; 1 / U_ii; the permutation), the abort word
// at work + wbytes - 256 (rounded down to an int), zeroed once here and raised by any part whose wait ran out.
// More right-hand sides than the solve kernel holds are solved in parts that share the word.
template <typename T>
int getrs_t(const T *LU, long n, long lda, const int *ipiv, const int *perm_in, T *B, long ldb, int nrhs, int trans,
            void *work, size_t wbytes, hipStream_t s)
{
    if (nrhs < 1 || nrhs > MAXRHS) return fail(M25X_EINVAL, "getrs: 1 to %d right-hand sides, not %d", MAXRHS, nrhs);
    if (ldb < nrhs) return fail(M25X_EINVAL, "getrs: ldb < nrhs");
    if (!work || wbytes < getrs_work<T>(n, ldb))
        return fail(M25X_EINVAL, "getrs: workspace of %zu bytes, %zu needed", wbytes, getrs_work<T>(n, ldb));
    int *abort = (int *)((char *)work + ((wbytes - 256) & ~(size_t)3));
    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, s, abort, 1);
    return getrs_noreset<T>(LU, n, lda, ipiv, perm_in, B, ldb, nrhs, trans, (char *)work, abort, s);
}

// ---- batches ------------------------------------------------------------------------------------------------
// Every kernel of a panel step (or of a solve) once for all items of a batch: grid rows (or the last grid
// dimension) over the items, each workgroup reading its item from the item table; workgroups past an item's own
// extent, and items already done, leave at once. The items' results are those of the one-matrix calls.

// the tournament counters zeroed, info cleared, perm = 0 .. n - 1
template <typename T>
__global__ void k_lu_init_b(const m25x_lu_item *items)
{
    const m25x_lu_item it = items[blockIdx.y];
    const TourWork<T> w = tour_carve<T>(it.work, it.n);
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < w.nn) w.arrive[i] = 0;
    if (it.perm && i < it.n) it.perm[i] = (int)i;
    if (i == 0 && it.info) *it.info = 0;
}

template <typename T>
__global__ __launch_bounds__(256) void k_apply_moves_b(const m25x_lu_item *items, long j0)
{
    const m25x_lu_item it = items[blockIdx.y];
    const long n = it.n;
    if (j0 >= n) return;
    const long nb = (n + 7) / 8;
    if ((long)blockIdx.x > nb || ((long)blockIdx.x == nb && !it.perm)) return;
    const TourWork<T> w = tour_carve<T>(it.work, n);
    apply_moves_body<T>((T *)it.A, it.lda, 0, n, w.pairs, j0, (int)std::min((long)TW, n - j0), w.top, it.perm,
                        blockIdx.x, nb);
}

template <typename T>
__global__ __launch_bounds__(256) void k_l21_b(const m25x_lu_item *items, long j0)
{
    const m25x_lu_item it = items[blockIdx.y];
    const long n = it.n;
    if (j0 >= n) return;
    const int jb = (int)std::min((long)TW, n - j0);
    const long rows = n - j0 - jb;
    if (rows <= 0 || (long)blockIdx.x * 256 >= rows) return;
    const TourWork<T> w = tour_carve<T>(it.work, n);
    l21_body<T>((T *)it.A, it.lda, j0 + jb, rows, j0, jb, (const T *)w.top, blockIdx.x);
}

template <typename T>
__global__ __launch_bounds__(256) void k_trsm_b(const m25x_lu_item *items, long j0)
{
    const m25x_lu_item it = items[blockIdx.y];
    const long n = it.n;
    if (j0 >= n) return;
    const int jb = (int)std::min((long)TW, n - j0);
    const long c0 = j0 + jb;
    if (c0 >= n || (long)blockIdx.x * TC >= n - c0) return;
    trsm_body<T, TW>((T *)it.A, it.lda, n, j0, jb, c0, blockIdx.x);
}

template <typename T, int KS, typename AC = double2>
__global__ __launch_bounds__(256) void k_gemm_b(const m25x_lu_item *items, long j0)
{
    const m25x_lu_item it = items[blockIdx.z];
    const long n = it.n;
    if (j0 >= n) return;
    const int jb = (int)std::min((long)TW, n - j0);
    const long c0 = j0 + jb, M = n - c0;
    if (M <= 0 || (long)blockIdx.x * GM >= M || (long)blockIdx.y * GN >= M) return;
    T *A = (T *)it.A;
    const long lda = it.lda;
    gemm_body<T, true, AC, KS>(A + c0 * lda + c0, lda, A + j0 * lda + c0, lda, A + c0 * lda + j0, lda, M, M, jb,
                               blockIdx.x, blockIdx.y);
}

template <typename T>
int getrf_batched_t(int nitems, const m25x_lu_item *h, const m25x_lu_item *d, hipStream_t s)
{
    long nmax = 0, work_max = 1;
    bool sorted = true;
    for (int i = 0; i < nitems; ++i) {
        const m25x_lu_item &it = h[i];
        if (!it.A || !it.ipiv || !it.work || it.n < 0 || it.lda < it.n || it.n > (1L << 30))
            return fail(M25X_EINVAL, "getrf_batched: item %d: bad arguments", i);
        if (it.work_bytes < tour_work_bytes<T>(it.n))
            return fail(M25X_EINVAL, "getrf_batched: item %d: workspace of %zu bytes, %zu needed", i, it.work_bytes,
                        tour_work_bytes<T>(it.n));
        nmax = std::max(nmax, it.n);
        work_max = std::max(work_max, std::max(it.n, tour_nodes(it.n, Tour<T>::RL, Tour<T>::AR)));
        if (i && it.n > h[i - 1].n) sorted = false;
    }
    if (nitems == 0 || nmax == 0) return M25X_OK;
    const int dt = dtype_of<T>();
    const double es = sizeof(T);
    hipLaunchKernelGGL(k_lu_init_b<T>, dim3(blocks(work_max, 256), nitems), dim3(256), 0, s, d);
    if (int rc = check_launch("getrf_batched")) return rc;
    int na = nitems;
    for (long j0 = 0; j0 < nmax; j0 += TW) {
        // the items still in the factorisation: a prefix when sorted by size (the grids then shrink with them)
        while (sorted && na > 0 && h[na - 1].n <= j0) --na;
        const long nbig = sorted ? h[0].n : nmax, mbig = nbig - j0;
        double flops = 0, pbytes = 0;
        for (int i = 0; i < na; ++i) {
            const long n = h[i].n;
            if (j0 >= n) continue;
            const long jb = std::min((long)TW, n - j0), M = n - j0 - jb;
            flops += 8.0 * M * M * jb;
            pbytes += es * (n - j0) * jb;
        }
        {
            TimedLaunch tl(s, M25X_K_TSLU, dt, pbytes, 0.0);
            if (int rc = launch_tslu_b((const T *)nullptr, d, na, j0, (int)tour_leaves(mbig, Tour<T>::RL), s))
                return rc;
        }
        hipLaunchKernelGGL(k_apply_moves_b<T>, dim3(blocks(nbig, 8) + 1, na), dim3(256), 0, s, d, j0);
        if (mbig > TW) hipLaunchKernelGGL(k_l21_b<T>, dim3(blocks(mbig - TW, 256), na), dim3(256), 0, s, d, j0);
        if (int rc = check_launch("getrf_batched/panel")) return rc;
        if (mbig > TW) {
            hipLaunchKernelGGL((k_trsm_b<T>), dim3(blocks(mbig - TW, TC), na), dim3(256), 0, s, d, j0);
            TimedLaunch tl(s, M25X_K_GEMM, dt, 0.0, flops);
            const dim3 grid(blocks(mbig - TW, GM), blocks(mbig - TW, GN), na);
            const bool f32 = sizeof(T) == 8 && g_update_fp32.load(std::memory_order_relaxed);
            switch (g_gemm_stage.load(std::memory_order_relaxed)) {
            case 16:
                if (f32) hipLaunchKernelGGL((k_gemm_b<T, 16, float2>), grid, dim3(256), 0, s, d, j0);
                else hipLaunchKernelGGL((k_gemm_b<T, 16>), grid, dim3(256), 0, s, d, j0);
                break;
            case 32:
                if (f32) hipLaunchKernelGGL((k_gemm_b<T, 32, float2>), grid, dim3(256), 0, s, d, j0);
                else hipLaunchKernelGGL((k_gemm_b<T, 32>), grid, dim3(256), 0, s, d, j0);
                break;
            default:
                if (f32) hipLaunchKernelGGL((k_gemm_b<T, 8, float2>), grid, dim3(256), 0, s, d, j0);
                else hipLaunchKernelGGL((k_gemm_b<T, 8>), grid, dim3(256), 0, s, d, j0);
            }
        }
        if (int rc = check_launch("getrf_batched/update")) return rc;
    }
    return M25X_OK;
}

// the solve's buffers in an item's workspace (as getrs_noreset carves them)
template <typename T> struct RsWork {
    T *W1, *W2, *W0, *rd;
};
template <typename T> __device__ inline RsWork<T> rs_carve(void *work, long n)
{
    RsWork<T> w;
    const size_t wn = (size_t)n * MAXRHS;
    w.W1 = (T *)work;
    w.W2 = w.W1 + wn;
    w.W0 = w.W2 + wn;
    w.rd = w.W0 + wn;
    return w;
}

// W1, W2 filled with UNSET words; with rdiag, 1 / U_ii too
template <typename T>
__global__ void k_rs_prep_b(const m25x_rs_item *items, int c0, int rdiag)
{
    const m25x_rs_item it = items[blockIdx.y];
    if (c0 >= it.nrhs) return;
    const RsWork<T> w = rs_carve<T>(it.work, it.n);
    const long words = (long)(2 * (size_t)it.n * MAXRHS * sizeof(T) / 8);
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < words) ((unsigned long long *)w.W1)[i] = ~0ull;
    if (rdiag && i < it.n) {
        const T *LU = (const T *)it.LU;
        w.rd[i] = trecip(LU[i * it.lda + i]);
    }
}

// the right-hand sides of part c0 in or out through the permutation: mode 0 B -> W0 gathered by perm (trans 0),
// mode 1 W2 -> B scattered by perm (trans 1), mode 2 W2 -> B as is (trans 0)
template <typename T>
__global__ void k_rs_move_b(const m25x_rs_item *items, int c0, int NR, int mode)
{
    const m25x_rs_item it = items[blockIdx.y];
    if (c0 >= it.nrhs) return;
    const int kc = std::min(NR, it.nrhs - c0);
    const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (idx >= it.n * kc) return;
    const long i = idx / kc;
    const int r = (int)(idx % kc);
    const RsWork<T> w = rs_carve<T>(it.work, it.n);
    T *B = (T *)it.B + c0;
    if (mode == 0) w.W0[i * kc + r] = B[(long)it.perm[i] * it.ldb + r];
    else if (mode == 1) B[(long)it.perm[i] * it.ldb + r] = w.W2[i * kc + r];
    else B[i * it.ldb + r] = w.W2[i * kc + r];
}

// one triangle of every item's part c0: workgroup g is block g / nitems of item g % nitems (the chains side by
// side; a workgroup waits only on blocks of its own item with smaller indices, which have smaller g)
template <typename T, int NR>
__global__ __launch_bounds__(TB) void k_trsv_b(const m25x_rs_item *items, int nitems, int c0, int kind, int *abort)
{
    const int item = blockIdx.x % nitems, q = blockIdx.x / nitems;
    const m25x_rs_item it = items[item];
    if (c0 >= it.nrhs || q >= (int)((it.n + SB - 1) / SB)) return;
    const int kc = std::min(NR, it.nrhs - c0);
    const RsWork<T> w = rs_carve<T>(it.work, it.n);
    const T *src;
    long lds;
    T *dst;
    if (kind == 0) { src = (const T *)it.B + c0; lds = it.ldb; dst = w.W1; }
    else if (kind == 2) { src = w.W0; lds = kc; dst = w.W1; }
    else { src = w.W1; lds = kc; dst = w.W2; }
    trsv_body<T, NR>((const T *)it.LU, it.lda, it.n, w.rd, src, lds, dst, kc, kind, abort, q);
}

template <typename T>
int getrs_batched_t(int nitems, const m25x_rs_item *h, const m25x_rs_item *d, int trans, int *status, hipStream_t s)
{
    constexpr int NR = TrsvMaxRhs<T>::NR;
    long nmax = 0;
    int rmax = 0, nbmax = 0;
    for (int i = 0; i < nitems; ++i) {
        const m25x_rs_item &it = h[i];
        if (!it.LU || !it.perm || !it.B || !it.work || it.n < 0 || it.lda < it.n || it.nrhs < 1 ||
            it.nrhs > MAXRHS || it.ldb < it.nrhs)
            return fail(M25X_EINVAL, "getrs_batched: item %d: bad arguments (perm required, 1 <= nrhs <= %d)", i,
                        MAXRHS);
        if (it.work_bytes < getrs_work<T>(it.n, it.ldb))
            return fail(M25X_EINVAL, "getrs_batched: item %d: workspace of %zu bytes, %zu needed", i, it.work_bytes,
                        getrs_work<T>(it.n, it.ldb));
        nmax = std::max(nmax, it.n);
        rmax = std::max(rmax, it.nrhs);
        nbmax = std::max(nbmax, (int)((it.n + SB - 1) / SB));
    }
    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, s, status, 1);
    if (int rc = check_launch("getrs_batched")) return rc;
    if (nitems == 0 || nmax == 0) return M25X_OK;
    const long words = (long)(2 * (size_t)nmax * MAXRHS * sizeof(T) / 8);
    const int dt = dtype_of<T>();
    for (int c0 = 0; c0 < rmax; c0 += NR) {
        hipLaunchKernelGGL(k_rs_prep_b<T>, dim3(blocks(std::max(words, nmax), 256), nitems), dim3(256), 0, s, d, c0,
                           (int)(c0 == 0));
        const dim3 mg(blocks(nmax * std::min(NR, rmax - c0), 256), nitems);
        if (!trans) hipLaunchKernelGGL(k_rs_move_b<T>, mg, dim3(256), 0, s, d, c0, NR, 0);
        if (int rc = check_launch("getrs_batched")) return rc;
        const int kmax = std::min(NR, rmax - c0);
        for (int pass = 0; pass < 2; ++pass) {
            const int kind = trans ? pass : 2 + pass;
            double bytes = 0, flops = 0;
            for (int i = 0; i < nitems; ++i)
                if (c0 < h[i].nrhs) {
                    const double n = h[i].n, tri = kind_fwd(kind) ? 0.5 * n * (n - 1) : 0.5 * n * (n + 1);
                    const int kc = std::min(NR, h[i].nrhs - c0);
                    bytes += (tri + 2.0 * n * kc) * sizeof(T);
                    flops += 8.0 * tri * kc;
                }
            TimedLaunch tl(s, M25X_K_TRSV, dt, bytes, flops);
            if (kmax <= 4)
                hipLaunchKernelGGL((k_trsv_b<T, 4>), dim3(nbmax * nitems), dim3(TB), 0, s, d, nitems, c0, kind, status);
            else
                hipLaunchKernelGGL((k_trsv_b<T, NR>), dim3(nbmax * nitems), dim3(TB), 0, s, d, nitems, c0, kind, status);
            if (int rc = check_launch("getrs_batched/trsv")) return rc;
        }
        hipLaunchKernelGGL(k_rs_move_b<T>, mg, dim3(256), 0, s, d, c0, NR, trans ? 1 : 2);
        if (int rc = check_launch("getrs_batched")) return rc;
    }
    return M25X_OK;
}

// ---- the solve without allocation ------------------------------------------------------------------------
// The workspace of m25x_getrs_noalloc: one array per device, in the code object of this file, held until the process
// ends. Host code never passes the array's name to a kernel or a copy (on the host it is only a stand-in): the
// device address comes from hipGetSymbolAddress, once per device, and is kept in scratch_dev.
static __device__ __align__(256) unsigned char scratch[M25X_GETRS_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
void *scratch_dev[MAX_DEVICES];

// The current device's address of `scratch` (looked up at the first call on that device).
int scratch_of_current(char **out)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES)
        return fail(M25X_EHIP, "getrs_noalloc: no current device");
    std::lock_guard<std::mutex> g(scratch_mu);
    if (!scratch_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(M25X_EHIP, "getrs_noalloc: hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        scratch_dev[dev] = p;
    }
    *out = (char *)scratch_dev[dev];
    return M25X_OK;
}

template <typename T>
int getrs_noalloc_t(const T *LU, long n, long lda, const int *ipiv, const int *perm_in, T *B, long ldb, int nrhs,
                    int trans, int *status, hipStream_t s)
{
    if (nrhs < 1 || nrhs > MAXRHS) return fail(M25X_EINVAL, "getrs_noalloc: 1 to %d right-hand sides, not %d", MAXRHS, nrhs);
    if (ldb < nrhs) return fail(M25X_EINVAL, "getrs_noalloc: ldb < nrhs");
    // the buffers getrs_noreset carves (W1, W2, W0, 1 / U_ii, the permutation): getrs_work less its 256-byte tail
    const size_t need = getrs_work<T>(n, ldb) - 256;
    if (need > (size_t)M25X_GETRS_SCRATCH_BYTES)
        return fail(M25X_EINVAL, "getrs_noalloc: n = %ld needs %zu bytes of workspace, more than the %u held per device",
                    n, need, (unsigned)M25X_GETRS_SCRATCH_BYTES);
    char *buf = nullptr;
    if (int rc = scratch_of_current(&buf)) return rc;
    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, s, status, 1);
    if (int rc = check_launch("getrs_noalloc")) return rc;
    return getrs_noreset<T>(LU, n, lda, ipiv, perm_in, B, ldb, nrhs, trans, buf, status, s);
}

}  // namespace

extern "C" {

int m25x_lu_set_gemm_stage(int ks)
{
    if (ks != 8 && ks != 16 && ks != 32) return fail(M25X_EINVAL, "lu_set_gemm_stage: 8, 16 or 32, not %d", ks);
    g_gemm_stage.store(ks, std::memory_order_relaxed);
    return M25X_OK;
}

int m25x_lu_gemm_stage(void) { return g_gemm_stage.load(std::memory_order_relaxed); }

int m25x_lu_set_update_precision(int fp32)
{
    if (fp32 != 0 && fp32 != 1) return fail(M25X_EINVAL, "lu_set_update_precision: 0 (FP64) or 1 (FP32), not %d", fp32);
    g_update_fp32.store(fp32, std::memory_order_relaxed);
    return M25X_OK;
}

int m25x_getrf_batched(int dtype, int nitems, const m25x_lu_item *items_host, const m25x_lu_item *items_dev,
                       void *stream)
{
    if (nitems < 0 || (nitems && (!items_host || !items_dev))) return fail(M25X_EINVAL, "getrf_batched: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    if (dtype == M25X_C128) return getrf_batched_t<double2>(nitems, items_host, items_dev, s);
    if (dtype == M25X_C64) return getrf_batched_t<float2>(nitems, items_host, items_dev, s);
    return fail(M25X_EINVAL, "getrf_batched: dtype %d", dtype);
}

int m25x_getrs_batched(int dtype, int nitems, const m25x_rs_item *items_host, const m25x_rs_item *items_dev, int trans,
                       int *status, void *stream)
{
    if (nitems < 0 || !status || (nitems && (!items_host || !items_dev)))
        return fail(M25X_EINVAL, "getrs_batched: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    if (dtype == M25X_C128) return getrs_batched_t<double2>(nitems, items_host, items_dev, trans, status, s);
    if (dtype == M25X_C64) return getrs_batched_t<float2>(nitems, items_host, items_dev, trans, status, s);
    return fail(M25X_EINVAL, "getrs_batched: dtype %d", dtype);
}

int m25x_getrs_noalloc(int dtype, const void *LU, long n, long lda, const int *ipiv, const int *perm, void *B,
                       long ldb, int nrhs, int trans, int *status, void *stream)
{
    if (!LU || !ipiv || !B || !status || n < 0 || lda < n) return fail(M25X_EINVAL, "getrs_noalloc: bad arguments");
    if (dtype != M25X_C128 && dtype != M25X_C64) return fail(M25X_EINVAL, "getrs_noalloc: dtype %d", dtype);
    hipStream_t s = (hipStream_t)stream;
    if (n == 0) {
        hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, s, status, 1);
        return check_launch("getrs_noalloc");
    }
    if (dtype == M25X_C128)
        return getrs_noalloc_t<double2>((const double2 *)LU, n, lda, ipiv, perm, (double2 *)B, ldb, nrhs, trans,
                                        status, s);
    return getrs_noalloc_t<float2>((const float2 *)LU, n, lda, ipiv, perm, (float2 *)B, ldb, nrhs, trans, status, s);
}

size_t m25x_getrf_work_bytes(int dtype, long n)
{
    return dtype == M25X_C64 ? std::max(work_bytes<float2>(n), tour_work_bytes<float2>(n))
                             : std::max(work_bytes<double2>(n), tour_work_bytes<double2>(n));
}

int m25x_getrf(int dtype, int pivot, void *A, long n, long lda, int *ipiv, int *perm, int *info, void *work,
               size_t wbytes, void *stream)
{
    if (!A || !ipiv || !work || n < 0 || lda < n) return fail(M25X_EINVAL, "getrf: bad arguments");
    if (pivot != M25X_PIVOT_PARTIAL && pivot != M25X_PIVOT_TOURNAMENT) return fail(M25X_EINVAL, "getrf: pivot %d", pivot);
    if (n == 0) return M25X_OK;
    if (n > (1L << 30)) return fail(M25X_EINVAL, "getrf: n too large");
    hipStream_t s = (hipStream_t)stream;
    const bool tp = pivot == M25X_PIVOT_TOURNAMENT;
    int rc;
    if (dtype == M25X_C128)
        rc = tp ? getrf_tp<double2>((double2 *)A, n, lda, ipiv, perm, info, work, wbytes, s)
                : getrf_t<double2>((double2 *)A, n, lda, ipiv, info, work, wbytes, s);
    else if (dtype == M25X_C64)
        rc = tp ? getrf_tp<float2>((float2 *)A, n, lda, ipiv, perm, info, work, wbytes, s)
                : getrf_t<float2>((float2 *)A, n, lda, ipiv, info, work, wbytes, s);
    else
        return fail(M25X_EINVAL, "getrf: dtype %d", dtype);
    if (rc == M25X_OK && perm && !tp) {
        hipLaunchKernelGGL(k_perm, dim3(1), dim3(256), 0, s, (const int *)ipiv, n, perm);
        rc = check_launch("getrf/perm");
    }
    return rc;
}

int m25x_lu_set_lookahead(int on)
{
    if (on != 0 && on != 1) return fail(M25X_EINVAL, "lu_set_lookahead: 0 or 1, not %d", on);
    g_lookahead.store(on, std::memory_order_relaxed);
    return M25X_OK;
}

int m25x_lu_lookahead(void) { return g_lookahead.load(std::memory_order_relaxed); }

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

int m25x_getrs(int dtype, const void *LU, long n, long lda, const int *ipiv, const int *perm, void *B, long ldb,
               int nrhs, int trans, void *work, size_t work_bytes, void *stream)
{
    if (!LU || !ipiv || !B || n < 0 || lda < n) return fail(M25X_EINVAL, "getrs: bad arguments");
    if (n == 0) return M25X_OK;
    hipStream_t s = (hipStream_t)stream;
    if (dtype == M25X_C128)
        return getrs_t<double2>((const double2 *)LU, n, lda, ipiv, perm, (double2 *)B, ldb, nrhs, trans, work,
                                work_bytes, s);
    if (dtype == M25X_C64)
        return getrs_t<float2>((const float2 *)LU, n, lda, ipiv, perm, (float2 *)B, ldb, nrhs, trans, work, work_bytes,
                               s);
    return fail(M25X_EINVAL, "getrs: dtype %d", dtype);
}

}  // extern "C"
