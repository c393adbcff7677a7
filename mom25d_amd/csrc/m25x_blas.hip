// m25x_blas.hip - the FP64 products of the GMRES refinement: Y = Z X with Z (n x n complex128, row-major) read
// once from HBM per product and X (n x k, k <= 8) staged through LDS in chunks shared by 16 rows.
#include "m25x_common.h"

using namespace m25x;

namespace {

constexpr int RK = 8;          // right-hand sides at most
constexpr int RPW = 4;         // rows a wave
constexpr int CH = 256;        // columns a chunk

__global__ __launch_bounds__(256) void k_zgemm_nk(const double2 *__restrict__ Z, long n, long ld,
                                                  const double2 *__restrict__ X, int k, double2 *__restrict__ Y)
{
    __shared__ double2 xs[CH][RK];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const long i0 = (long)blockIdx.x * (4 * RPW) + wv * RPW;
    double2 acc[RPW][RK];
#pragma unroll
    for (int r = 0; r < RPW; ++r)
#pragma unroll
        for (int c = 0; c < RK; ++c) acc[r][c] = make_double2(0, 0);
    for (long j0 = 0; j0 < n; j0 += CH) {
        const long jn = n - j0 < CH ? n - j0 : CH;
        for (int idx = tid; idx < CH * k; idx += 256) {
            const int t = idx / k, c = idx % k;
            xs[t][c] = t < jn ? X[(j0 + t) * k + c] : make_double2(0, 0);
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            const long i = i0 + r;
            if (i < n) {
                const double2 *zr = Z + i * ld + j0;
                for (int jj = lane; jj < jn; jj += 64) {
                    const double2 z = zr[jj];
#pragma unroll
                    for (int c = 0; c < RK; ++c)
                        if (c < k) {
                            const double2 x = xs[jj][c];
                            acc[r][c].x += z.x * x.x - z.y * x.y;
                            acc[r][c].y += z.x * x.y + z.y * x.x;
                        }
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < RPW; ++r)
#pragma unroll
        for (int c = 0; c < RK; ++c) {
            double x = acc[r][c].x, y = acc[r][c].y;
            for (int o = 32; o > 0; o >>= 1) { x += __shfl_xor(x, o); y += __shfl_xor(y, o); }
            if (lane == 0 && c < k && i0 + r < n) Y[(i0 + r) * k + c] = make_double2(x, y);
        }
}

}  // namespace

extern "C" int m25x_zgemm_nk(const double *Z, long n, long ld, const double *X, int k, double *Y, void *stream)
{
    if (!Z || !X || !Y || n < 0 || ld < n || k < 1 || k > RK)
        return fail(M25X_EINVAL, "zgemm_nk: bad arguments (1 <= k <= %d)", RK);
    if (n == 0) return M25X_OK;
    const unsigned nblk = (unsigned)((n + 4 * RPW - 1) / (4 * RPW));
    hipLaunchKernelGGL(k_zgemm_nk, dim3(nblk), dim3(256), 0, (hipStream_t)stream, (const double2 *)Z, n, ld,
                       (const double2 *)X, k, (double2 *)Y);
    return check_launch("zgemm_nk");
}
