// m25x_blas.hip - the FP64 products of the GMRES refinement: Y = Z X with Z (n x n complex128, row-major) read
// once from HBM per product and X (n x k, k <= 8) staged through LDS in chunks shared by 16 rows; one build per k.
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
template <int K>
__global__ __launch_bounds__(256) void k_zgemm_nk(const double2 *__restrict__ Z, long n, long ld,
                                                  const double2 *__restrict__ X, double2 *__restrict__ Y)
{
    constexpr int RPW = rows_per_wave<K>();
    __shared__ double2 xs[K][CH];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const long i0 = (long)blockIdx.x * (4 * RPW) + wv * RPW;
    double2 acc[RPW][K];
#pragma unroll
    for (int r = 0; r < RPW; ++r)
#pragma unroll
        for (int c = 0; c < K; ++c) acc[r][c] = make_double2(0, 0);
    for (long j0 = 0; j0 < n; j0 += CH) {
        const long jn = n - j0 < CH ? n - j0 : CH;
        for (int idx = tid; idx < CH * K; idx += 256) {
            const int t = idx / K, c = idx % K;
            xs[c][t] = t < jn ? X[(j0 + t) * K + c] : make_double2(0, 0);
        }
        double2 z[RPW][UJ];
#pragma unroll
        for (int r = 0; r < RPW; ++r)
#pragma unroll
            for (int u = 0; u < UJ; ++u) {
                const long jj = lane + 64 * u;
                z[r][u] = (i0 + r < n && jj < jn) ? Z[(i0 + r) * ld + j0 + jj] : make_double2(0, 0);
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
            if (lane == 0 && i0 + r < n) Y[(i0 + r) * K + c] = make_double2(x, y);
        }
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
