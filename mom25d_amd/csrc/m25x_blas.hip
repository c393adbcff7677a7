// m25x_blas.hip - the FP64 kernels of the GMRES refinement: Y = Z X with Z (n x n complex128, row-major) read
// once from HBM per product and X (n x k, k <= 8) staged through LDS in chunks shared by 16 rows; one build per k.
#include <algorithm>
#include "m25x_common.h"

using namespace m25x;

namespace {

constexpr int RK = 8;          // right-hand sides at most
constexpr int CH = 256;        // columns a chunk
constexpr int UJ = CH / 64;    // columns a lane per chunk

// rows a wave: four, two from six right-hand sides on (registers)
template <int K> constexpr int rows_per_wave() { return K <= 5 ? 4 : 2; }

// Each wave takes RPW rows; per chunk of CH columns a lane loads its UJ columns of all RPW rows at once (RPW x UJ
// independent 16 B loads in flight), reads each column's K values of X once from LDS (laid out [K][CH], so a
// wave's reads are consecutive) and applies them to the RPW rows.
// X, Y: leading dimensions ldx, ldy; kk <= K columns in use (the others read as zero and are not written)
template <int K>
__device__ __forceinline__ void zgemm_body(const double2 *__restrict__ Z, long n, long ld,
                                           const double2 *__restrict__ X, long ldx, double2 *__restrict__ Y, long ldy,
                                           int kk, long bx)
{
    constexpr int RPW = rows_per_wave<K>();
    __shared__ double2 xs[K][CH];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const long i0 = bx * (4 * RPW) + wv * RPW;
    double2 acc[RPW][K];
#pragma unroll
    for (int r = 0; r < RPW; ++r)
#pragma unroll
        for (int c = 0; c < K; ++c) acc[r][c] = make_double2(0, 0);
    for (long j0 = 0; j0 < n; j0 += CH) {
        const long jn = n - j0 < CH ? n - j0 : CH;
        // the chunk's Z entries first (from HBM: the long wait), then its X values (K a thread, all in flight
        // before the first is stored to LDS)
        double2 z[RPW][UJ];
#pragma unroll
        for (int r = 0; r < RPW; ++r)
#pragma unroll
            for (int u = 0; u < UJ; ++u) {
                const long jj = lane + 64 * u;
                z[r][u] = (i0 + r < n && jj < jn) ? Z[(i0 + r) * ld + j0 + jj] : make_double2(0, 0);
            }
        double2 xv[CH * K / 256];
#pragma unroll
        for (int it = 0; it < CH * K / 256; ++it) {
            const int idx = tid + it * 256, t = idx / K, c = idx % K;
            xv[it] = (t < jn && c < kk) ? X[(j0 + t) * ldx + c] : make_double2(0, 0);
        }
#pragma unroll
        for (int it = 0; it < CH * K / 256; ++it) {
            const int idx = tid + it * 256;
            xs[idx % K][idx / K] = xv[it];
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < UJ; ++u) {
            double2 x[K];
#pragma unroll
            for (int c = 0; c < K; ++c) x[c] = xs[c][lane + 64 * u];
#pragma unroll
            for (int r = 0; r < RPW; ++r)
#pragma unroll
                for (int c = 0; c < K; ++c) {
                    acc[r][c].x += z[r][u].x * x[c].x - z[r][u].y * x[c].y;
                    acc[r][c].y += z[r][u].x * x[c].y + z[r][u].y * x[c].x;
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < RPW; ++r)
#pragma unroll
        for (int c = 0; c < K; ++c) {
            double x = acc[r][c].x, y = acc[r][c].y;
            for (int o = 32; o > 0; o >>= 1) { x += __shfl_xor(x, o); y += __shfl_xor(y, o); }
            if (lane == 0 && i0 + r < n && c < kk) Y[(i0 + r) * ldy + c] = make_double2(x, y);
        }
}

template <int K>
__global__ __launch_bounds__(256) void k_zgemm_nk(const double2 *__restrict__ Z, long n, long ld,
                                                  const double2 *__restrict__ X, double2 *__restrict__ Y)
{
    zgemm_body<K>(Z, n, ld, X, K, Y, K, K, blockIdx.x);
}

// one item a grid row (blockIdx.y); row blocks past the item's n leave at once
template <int K>
__global__ __launch_bounds__(256) void k_zgemm_nk_b(const m25x_mv_item *__restrict__ items)
{
    const m25x_mv_item it = items[blockIdx.y];
    if ((long)blockIdx.x * (4 * rows_per_wave<K>()) >= it.n) return;
    zgemm_body<K>((const double2 *)it.Z, it.n, it.ld, (const double2 *)it.X, it.ldx, (double2 *)it.Y, it.ldy, it.k,
                  blockIdx.x);
}

}  // namespace

extern "C" int m25x_zgemm_nk(const double *Z, long n, long ld, const double *X, int k, double *Y, void *stream)
{
    if (!Z || !X || !Y || n < 0 || ld < n || k < 1 || k > RK)
        return fail(M25X_EINVAL, "zgemm_nk: bad arguments (1 <= k <= %d)", RK);
    if (n == 0) return M25X_OK;
    hipStream_t s = (hipStream_t)stream;
    const double2 *Zc = (const double2 *)Z, *Xc = (const double2 *)X;
    double2 *Yc = (double2 *)Y;
    TimedLaunch tl(s, M25X_K_ZGEMM_NK, M25X_C128, 16.0 * n * (double)n + 32.0 * n * k, 8.0 * n * (double)n * k);
    switch (k) {
#define M25X_NK(K)                                                                                            \
    case K:                                                                                                   \
        hipLaunchKernelGGL(k_zgemm_nk<K>, dim3((unsigned)((n + 4 * rows_per_wave<K>() - 1) / (4 * rows_per_wave<K>()))), \
                           dim3(256), 0, s, Zc, n, ld, Xc, Yc);                                                \
        break;
        M25X_NK(1) M25X_NK(2) M25X_NK(3) M25X_NK(4) M25X_NK(5) M25X_NK(6) M25X_NK(7) M25X_NK(8)
#undef M25X_NK
    }
    return check_launch("zgemm_nk");
}

extern "C" int m25x_zgemm_nk_batched(int nitems, const m25x_mv_item *h, const m25x_mv_item *d, void *stream)
{
    if (nitems < 0 || (nitems && (!h || !d))) return fail(M25X_EINVAL, "zgemm_nk_batched: bad arguments");
    long nmax = 0;
    int kmax = 1;
    double bytes = 0, flops = 0;
    for (int i = 0; i < nitems; ++i) {
        const m25x_mv_item &it = h[i];
        if (!it.Z || !it.X || !it.Y || it.n < 0 || it.ld < it.n || it.k < 1 || it.k > RK || it.ldx < it.k ||
            it.ldy < it.k)
            return fail(M25X_EINVAL, "zgemm_nk_batched: item %d: bad arguments (1 <= k <= %d, ld >= n, ldx, ldy >= k)",
                        i, RK);
        nmax = std::max(nmax, it.n);
        kmax = std::max(kmax, it.k);
        bytes += 16.0 * it.n * (double)it.n + 32.0 * it.n * it.k;
        flops += 8.0 * it.n * (double)it.n * it.k;
    }
    if (nitems == 0 || nmax == 0) return M25X_OK;
    hipStream_t s = (hipStream_t)stream;
    TimedLaunch tl(s, M25X_K_ZGEMM_NK, M25X_C128, bytes, flops);
    switch (kmax) {
#define M25X_NKB(K)                                                                                           \
    case K:                                                                                                   \
        hipLaunchKernelGGL(k_zgemm_nk_b<K>, dim3((unsigned)((nmax + 4 * rows_per_wave<K>() - 1) / (4 * rows_per_wave<K>())), \
                                                 (unsigned)nitems), dim3(256), 0, s, d);                      \
        break;
        M25X_NKB(1) M25X_NKB(2) M25X_NKB(3) M25X_NKB(4) M25X_NKB(5) M25X_NKB(6) M25X_NKB(7) M25X_NKB(8)
#undef M25X_NKB
    }
    return check_launch("zgemm_nk_batched");
}

// ---- classical Gram-Schmidt, twice, for the batched GMRES ------------------------------------------------------
// V: (j + 1) basis vectors, each n x k complex128 row-major (vector i at V + i n k); w: n x k. Per column c:
// h1 = V^H w, w1 = w - V h1, h2 = V^H w1, w2 = w1 - V h2, h = h1 + h2, nrm = |w2|, V[j + 1] = w2 / nrm (nrm = 0: w2).
// Sums over rows in fixed blocks, then over blocks in order: the result does not depend on timing.
namespace {

constexpr int GS_B = 64;                 // row blocks
constexpr int GS_T = 256;                // threads a block: 8 columns x 32 row lanes
constexpr int GS_MAXV = 64;              // basis vectors at most

// partial dots of the rows of block blockIdx.x: P[b][i][c] = sum_r conj(V[i][r][c]) x[r][c], for i < nv; with
// upd: first x -= sum_i V[i] h[i] (h from H), stored back. The dots run over the basis in chunks of GS_CH vectors.
constexpr int GS_CH = 32;
template <bool UPD>
__global__ __launch_bounds__(GS_T) void k_gs_dots(const double2 *__restrict__ V, long n, int k, int nv,
                                                 double2 *__restrict__ x, const double2 *__restrict__ H,
                                                 double2 *__restrict__ P)
{
    __shared__ double2 red[GS_T / 64][GS_CH][8];
    const int c = threadIdx.x & 7, rl = threadIdx.x >> 3, wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long rows = (n + GS_B - 1) / GS_B, r0 = blockIdx.x * rows, r1 = r0 + rows < n ? r0 + rows : n;
    if (UPD && c < k)                              // each thread updates the rows it reads below
        for (long r = r0 + rl; r < r1; r += GS_T / 8) {
            double2 xv = x[r * k + c];
            for (int i = 0; i < nv; ++i) {
                const double2 v = V[(long)i * n * k + r * k + c], h = H[i * k + c];
                xv.x -= v.x * h.x - v.y * h.y;
                xv.y -= v.x * h.y + v.y * h.x;
            }
            x[r * k + c] = xv;
        }
    for (int i0 = 0; i0 < nv; i0 += GS_CH) {
        double2 acc[GS_CH];
#pragma unroll
        for (int i = 0; i < GS_CH; ++i) acc[i] = make_double2(0, 0);
        if (c < k)
            for (long r = r0 + rl; r < r1; r += GS_T / 8) {
                const double2 xv = x[r * k + c];
#pragma unroll
                for (int i = 0; i < GS_CH; ++i)
                    if (i0 + i < nv) {
                        const double2 v = V[(long)(i0 + i) * n * k + r * k + c];
                        acc[i].x += v.x * xv.x + v.y * xv.y;
                        acc[i].y += v.x * xv.y - v.y * xv.x;
                    }
            }
        // the eight row lanes of a wave (lane bits 3-5), then the four waves
#pragma unroll
        for (int i = 0; i < GS_CH; ++i)
            if (i0 + i < nv) {
                double sx = acc[i].x, sy = acc[i].y;
#pragma unroll
                for (int o = 8; o < 64; o <<= 1) { sx += __shfl_xor(sx, o); sy += __shfl_xor(sy, o); }
                if (lane < 8) red[wv][i][c] = make_double2(sx, sy);
            }
        __syncthreads();
        const int m = nv - i0 < GS_CH ? nv - i0 : GS_CH;
        for (int idx = threadIdx.x; idx < m * k; idx += GS_T) {
            const int i = idx / k, cc = idx % k;
            double2 s = make_double2(0, 0);
            for (int l = 0; l < GS_T / 64; ++l) { s.x += red[l][i][cc].x; s.y += red[l][i][cc].y; }
            P[((long)blockIdx.x * GS_MAXV + i0 + i) * 8 + cc] = s;
        }
        __syncthreads();
    }
}

// H[i][c] = sum_b P[b][i][c] (in block order); with acc: Hs[i][c] += H[i][c] as well
__global__ void k_gs_sum(const double2 *__restrict__ P, int k, int nv, double2 *__restrict__ H, double2 *Hs, int first)
{
    const int idx = threadIdx.x;
    if (idx >= nv * k) return;
    const int i = idx / k, c = idx % k;
    double2 s = make_double2(0, 0);
    for (int b = 0; b < GS_B; ++b) {
        const double2 p = P[((long)b * GS_MAXV + i) * 8 + c];
        s.x += p.x; s.y += p.y;
    }
    H[i * k + c] = s;
    if (first) Hs[i * k + c] = s;
    else { Hs[i * k + c].x += s.x; Hs[i * k + c].y += s.y; }
}

// x -= V h (h from H), stored back; Q[b][c] = partial |x|^2
__global__ __launch_bounds__(GS_T) void k_gs_final(const double2 *__restrict__ V, long n, int k, int nv,
                                                  double2 *__restrict__ x, const double2 *__restrict__ H,
                                                  double *__restrict__ Q)
{
    __shared__ double red[GS_T / 8][8];
    const int c = threadIdx.x & 7, rl = threadIdx.x >> 3;
    const long rows = (n + GS_B - 1) / GS_B, r0 = blockIdx.x * rows, r1 = r0 + rows < n ? r0 + rows : n;
    double q = 0;
    if (c < k)
        for (long r = r0 + rl; r < r1; r += GS_T / 8) {
            double2 xv = x[r * k + c];
            for (int i = 0; i < nv; ++i) {
                const double2 v = V[(long)i * n * k + r * k + c], h = H[i * k + c];
                xv.x -= v.x * h.x - v.y * h.y;
                xv.y -= v.x * h.y + v.y * h.x;
            }
            x[r * k + c] = xv;
            q += xv.x * xv.x + xv.y * xv.y;
        }
    red[rl][c] = q;
    __syncthreads();
    if (threadIdx.x < k) {
        double s = 0;
        for (int l = 0; l < GS_T / 8; ++l) s += red[l][threadIdx.x];
        Q[blockIdx.x * 8 + threadIdx.x] = s;
    }
}

// hcol[nv][c] = |x| (real); V[nv] = x / |x|
__global__ void k_gs_norm(const double *__restrict__ Q, int k, int nv, double2 *__restrict__ hcol)
{
    const int c = threadIdx.x;
    if (c >= k) return;
    double s = 0;
    for (int b = 0; b < GS_B; ++b) s += Q[b * 8 + c];
    hcol[nv * k + c] = make_double2(sqrt(s), 0);
}

__global__ void k_gs_scale(const double2 *__restrict__ x, long nk, int k, int nv, const double2 *__restrict__ hcol,
                           double2 *__restrict__ out)
{
    const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (idx >= nk) return;
    const double nr = hcol[nv * k + (int)(idx % k)].x;
    const double sc = nr > 0 ? 1.0 / nr : 1.0;
    out[idx] = make_double2(x[idx].x * sc, x[idx].y * sc);
}

}  // namespace

extern "C" size_t m25x_cgs2_work_bytes(void)
{
    return (size_t)GS_B * GS_MAXV * 8 * sizeof(double2) + GS_MAXV * 8 * sizeof(double2) + GS_B * 8 * sizeof(double);
}

extern "C" int m25x_cgs2(const double *V, long n, int k, int nv, double *w, double *hcol, double *vnext, void *work,
                         size_t work_bytes, void *stream)
{
    if (!V || !w || !hcol || !vnext || !work || n < 1 || k < 1 || k > 8 || nv < 1 || nv > GS_MAXV - 1)
        return fail(M25X_EINVAL, "cgs2: bad arguments (1 <= k <= 8, 1 <= nv <= %d)", GS_MAXV - 1);
    if (work_bytes < m25x_cgs2_work_bytes()) return fail(M25X_EINVAL, "cgs2: workspace too small");
    hipStream_t s = (hipStream_t)stream;
    const double2 *Vc = (const double2 *)V;
    double2 *x = (double2 *)w, *H = (double2 *)hcol;
    double2 *P = (double2 *)work, *h = P + (size_t)GS_B * GS_MAXV * 8;
    double *Q = (double *)(h + GS_MAXV * 8);
    TimedLaunch tl(s, M25X_K_CGS2, M25X_C128, 16.0 * n * k * (2.0 * nv + 3), 16.0 * n * k * nv);
    hipLaunchKernelGGL(k_gs_dots<false>, dim3(GS_B), dim3(GS_T), 0, s, Vc, n, k, nv, x, (const double2 *)nullptr, P);
    hipLaunchKernelGGL(k_gs_sum, dim3(1), dim3(GS_MAXV * 8), 0, s, (const double2 *)P, k, nv, h, H, 1);
    hipLaunchKernelGGL(k_gs_dots<true>, dim3(GS_B), dim3(GS_T), 0, s, Vc, n, k, nv, x, (const double2 *)h, P);
    hipLaunchKernelGGL(k_gs_sum, dim3(1), dim3(GS_MAXV * 8), 0, s, (const double2 *)P, k, nv, h, H, 0);
    hipLaunchKernelGGL(k_gs_final, dim3(GS_B), dim3(GS_T), 0, s, Vc, n, k, nv, x, (const double2 *)h, Q);
    hipLaunchKernelGGL(k_gs_norm, dim3(1), dim3(64), 0, s, (const double *)Q, k, nv, H);
    hipLaunchKernelGGL(k_gs_scale, dim3((unsigned)((n * k + 255) / 256)), dim3(256), 0, s, (const double2 *)x, n * k,
                       k, nv, (const double2 *)H, (double2 *)vnext);
    return check_launch("cgs2");
}
