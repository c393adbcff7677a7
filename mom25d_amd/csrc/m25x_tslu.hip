// m25x_tslu.hip - the tournament of a tournament-pivoting LU panel (k_tslu), built once per element type
// (M25X_TSLU_C64 set: complex64, else complex128) so the two fully unrolled instantiations compile in parallel.
#include "m25x_lu_dev.h"

namespace m25x {
namespace lu {
namespace {

// The pivot search orders candidates by (magnitude, then the smaller key): magnitudes |Re| + |Im| are non-negative,
// so their IEEE bit patterns order as unsigned integers (a dead row has bits 0 and the largest key, so any live row,
// even one with a zero entry, outranks it). Wave reductions in 32-bit passes - the largest magnitude word(s), then
// the smallest key among the lanes holding them - each within rows of 16 lanes by DPP exchanges (quad xor 1, quad
// xor 2, half-row mirror, row mirror), then over the four rows by readlane.
typedef unsigned long long u64;

// 32-bit wave reductions: four DPP rounds within rows of 16 lanes (each one instruction: the exchange is the
// operation's source modifier, the old value the operation's identity), then the four rows by readlane
__device__ inline unsigned wave_max_u32(unsigned v)
{
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false));
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false));
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, false));
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xF, 0xF, false));
    const unsigned a = max((unsigned)__builtin_amdgcn_readlane((int)v, 0), (unsigned)__builtin_amdgcn_readlane((int)v, 16));
    const unsigned b = max((unsigned)__builtin_amdgcn_readlane((int)v, 32), (unsigned)__builtin_amdgcn_readlane((int)v, 48));
    return max(a, b);
}
__device__ inline int wave_min_i32(int v)
{
    v = min(v, __builtin_amdgcn_update_dpp(0x7fffffff, v, 0xB1, 0xF, 0xF, false));
    v = min(v, __builtin_amdgcn_update_dpp(0x7fffffff, v, 0x4E, 0xF, 0xF, false));
    v = min(v, __builtin_amdgcn_update_dpp(0x7fffffff, v, 0x141, 0xF, 0xF, false));
    v = min(v, __builtin_amdgcn_update_dpp(0x7fffffff, v, 0x140, 0xF, 0xF, false));
    return min(min(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
               min(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}

__device__ inline u64 mag_bits(float2 a) { return (u64)__float_as_uint(fabsf(a.x) + fabsf(a.y)); }
__device__ inline u64 mag_bits(double2 a) { return (u64)__double_as_longlong(fabs(a.x) + fabs(a.y)); }

// GEPP of rows held in registers, two threads a row: in wave w, lanes L and L + 32 hold row 32 w + L % 32, the
// thread of half h = L / 32 columns 2 j + h in x[j] (j < TW / 2), so both halves have equal work at every step.
// Multipliers: the column times the pivot's reciprocal conj(p) / |p|^2 (LAPACK's getf2 scales by 1 / pivot).
// nr rows over columns [0, nc). Rows never move: a row's position in the elimination order is tracked instead
// (pos), and ties in the pivot search go to the earliest position (LAPACK's rule). Step k: every wave's best row
// publishes itself in its slot (LDS) with its (magnitude, key); after one barrier every thread picks the best of
// the slots, takes its multiplier and updates. order[k] = the k-th pivot row; a pivot row's threads keep its step in
// step_of. On return a pivot row holds its row of L \ U, the other rows their multipliers and Schur complement.
// One barrier a step: the slots are double-buffered.
// lane L's value to lane L ^ 32: v_permlane32_swap exchanges the upper half of its destination with the lower half of
// its source (a VALU operation, no LDS round trip); with both the same register, lane L < 32 finds lane L + 32's
// value in the second result and lane L >= 32 lane L - 32's in the first
__device__ inline unsigned swap32(unsigned v)
{
    const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return (threadIdx.x & 32) ? r[0] : r[1];
}
__device__ inline float swap32(float v) { return __uint_as_float(swap32(__float_as_uint(v))); }
__device__ inline double swap32(double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return __longlong_as_double((long long)(((unsigned long long)swap32((unsigned)(b >> 32)) << 32) | swap32((unsigned)b)));
}
template <typename T> __device__ inline T shfl_xor32(T v) { return CX<T>::make(swap32(v.x), swap32(v.y)); }

template <typename T, int NT> struct GeppLds {
    T rows[2][NT / 64][TW];
    u64 vb[2][NT / 64];
    int key[2][NT / 64];
};

template <typename T, int NT>
__device__ __attribute__((always_inline)) int reg_gepp(RowH<T> &x, int nr, int nc, int *order, GeppLds<T, NT> &sh,
                                                       int &step_of)
{
    constexpr int NW = NT / 64;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int row = w * 32 + (lane & 31), h = lane >> 5;
    const int npiv = nr < nc ? nr : nc;
    bool live = row < nr;
    int pos = row;
    step_of = -1;
    auto step = [&](auto kc) __attribute__((always_inline)) {
        constexpr int k = decltype(kc)::value, jk = k / 2, hk = k % 2;
        if (k >= npiv) return;
        const int b = k & 1;
        const T mine = x.get(jk);
        const T other = shfl_xor32(mine);                    // column k of this row, for the non-owning half
        const bool cand = live && h == hk;
        u64 vb = cand ? mag_bits(mine) : 0ull;
        if (cand && !(vb <= (sizeof(T) == 8 ? 0x7f800000ull : 0x7ff0000000000000ull))) vb = 0;   // NaN: as zero
        int key = cand ? ((pos << 16) | row) : 0x7fffffff;
        int gk;
        if constexpr (sizeof(T) == 8) {
            // complex64: the wave's largest magnitude (32 bits; a dead row 0), then the smallest key among the
            // lanes holding it (a dead row's key is the largest, so a live row with a zero entry wins over it). The
            // wave's slot packs the pair as the magnitude above the complemented key, so that one maximum over the
            // slots picks the largest magnitude and, among equals, the smallest key (0: no live row in the wave).
            const unsigned mx = wave_max_u32((unsigned)vb);
            key = wave_min_i32((unsigned)vb == mx ? key : 0x7fffffff);
            const u64 pk = key != 0x7fffffff ? ((u64)mx << 32) | (u64)(0xffffffffu - (unsigned)key) : 0ull;
            if ((key & 0xffff) == row && key != 0x7fffffff) {     // this wave's best row publishes itself
                auto pub = [&](auto jc) __attribute__((always_inline)) {
                    constexpr int j = decltype(jc)::value;
                    sh.rows[b][w][2 * j + h] = x.get(j);
                };
                Unroll<jk, TH>::run(pub);
            }
            if (lane == 0) sh.vb[b][w] = pk;
            __syncthreads();
            u64 g = sh.vb[b][0];
#pragma unroll
            for (int q = 1; q < NW; ++q) g = max(g, sh.vb[b][q]);
            gk = (int)(0xffffffffu - (unsigned)g);
        } else {
            // complex128: the largest 64-bit magnitude as (high word, then low word among the lanes holding the
            // largest high word), then the smallest key among the lanes holding both
            const unsigned hi = (unsigned)(vb >> 32), lo = (unsigned)vb;
            const unsigned mh = wave_max_u32(hi);
            const unsigned ml = wave_max_u32(hi == mh ? lo : 0u);
            key = wave_min_i32(hi == mh && lo == ml ? key : 0x7fffffff);
            vb = ((u64)mh << 32) | ml;
            if ((key & 0xffff) == row && key != 0x7fffffff) {     // this wave's best row publishes itself
                auto pub = [&](auto jc) __attribute__((always_inline)) {
                    constexpr int j = decltype(jc)::value;
                    sh.rows[b][w][2 * j + h] = x.get(j);
                };
                Unroll<jk, TH>::run(pub);
            }
            if (lane == 0) { sh.vb[b][w] = vb; sh.key[b][w] = key; }
            __syncthreads();
            u64 gv = sh.vb[b][0];
#pragma unroll
            for (int q = 1; q < NW; ++q) gv = max(gv, sh.vb[b][q]);
            gk = 0x7fffffff;
#pragma unroll
            for (int q = 0; q < NW; ++q) gk = sh.vb[b][q] == gv ? min(gk, sh.key[b][q]) : gk;
        }
        const int ps = gk & 0xffff, p = gk >> 16;
        const T *prow = sh.rows[b][ps >> 5];
        if (row == ps) {
            live = false;
            step_of = k;
        }
        if (t == 0) order[k] = ps;
        if (live) {
            const T a = csel(h == hk, mine, other);
            if (pos == k) pos = p;
            // the multiplier by the pivot's reciprocal conj(pv) / |pv|^2, as LAPACK's getf2 scales by 1 / pivot (a
            // zero pivot leaves the column as it is); formed here, where it overlaps the update's LDS reads (formed
            // by the pivot row's wave before the barrier it measured slower)
            const T pv = prow[k];
            typedef typename CX<T>::R R;
            const R d = pv.x * pv.x + pv.y * pv.y;
            const R inv = d != (R)0 ? (R)1 / d : (R)0;
            const T l = d != (R)0 ? tmul(a, CX<T>::make(pv.x * inv, -pv.y * inv)) : a;
            if (h == hk) x.set(jk, l);
            else if (hk == 0) x.set(jk, tfms(x.get(jk), l, prow[2 * jk + 1]));   // column k + 1 of half 1
            auto upd = [&](auto jc) __attribute__((always_inline)) {
                constexpr int j = decltype(jc)::value;
                x.set(j, tfms(x.get(j), l, prow[2 * j + h]));
            };
            Unroll<jk + 1, TH>::run(upd);
        }
    };
    Unroll<0, TW>::run(step);
    __syncthreads();
    return npiv;
}

// The tournament of panel [j0, j0 + jb) over rows [j0, n). Leaves: blockIdx.x (P of them) of RL rows, two threads
// a row. The root writes the panel's L11 \ U11 (top, column-major TW x TW), the interchanges (ipiv[j0 .. j0 + jb),
// LAPACK's convention) and the row moves they make as (destination, source) pairs (pairs[0] = count), and info for
// a zero pivot.
template <typename T>
__device__ __forceinline__ void tslu_body(const T *A, long lda, long n, long j0, int jb, int P, TNode<T> *nodes,
                                          int *arrive, T *top, int *ipiv, int *pairs, int *info, int leaf)
{
    constexpr int RL = Tour<T>::RL, AR = Tour<T>::AR, NT = 2 * RL, NW = NT / 64;
    __shared__ GeppLds<T, NT> gl;
    __shared__ int ids[RL];
    __shared__ int order[TW];
    __shared__ int go, counts[AR];
    const int t = threadIdx.x, lane = t & 63;
    const int row = (t >> 6) * 32 + (lane & 31), h = lane >> 5;
    const T zero = CX<T>::make(0, 0);
    int idx = leaf;
    const long r0 = j0 + (long)idx * RL;
    int nr = (int)std::min((long)RL, n - r0);
    RowH<T> x;
    {
        auto ld = [&](auto jc) __attribute__((always_inline)) {
            constexpr int j = decltype(jc)::value;
            x.set(j, ldz(2 * j + h < jb && row < nr, A + (j0 + 2 * j + h) * lda + r0 + row));
        };
        Unroll<0, TH>::run(ld);
    }
    if (t < nr) ids[t] = (int)(r0 + t);
    __syncthreads();
    int step_of;
    int cnt = reg_gepp<T, NT>(x, nr, jb, order, gl, step_of);
    const int li = t & 63, cg = t >> 6;          // nominee li, columns cg, cg + NW, ...
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
            {
                const int r = li < cnt ? ids[order[li]] : 0;
                for (int c0 = cg; c0 < jb; c0 += 8 * NW) {
                    T v[8];
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        const int c = c0 + j * NW;
                        v[j] = ldz(li < cnt && c < jb, A + (j0 + c) * lda + r);
                    }
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        const int c = c0 + j * NW;
                        if (li < cnt && c < jb) nd->v[c * TW + li] = v[j];
                    }
                }
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
            const TNode<T> *c0n = nodes + loff + first;
            if (t < nchild) counts[t] = c0n[t].count;
            __syncthreads();
            int ch = 0, i = row;
            nr = 0;
            for (int q = 0; q < nchild; ++q) nr += counts[q];
            while (ch < nchild - 1 && i >= counts[ch]) { i -= counts[ch]; ++ch; }
            const bool have = row < nr;
            const TNode<T> *cn = c0n + ch;
            auto ld = [&](auto jc) __attribute__((always_inline)) {
                constexpr int j = decltype(jc)::value;
                x.set(j, ldz(have && 2 * j + h < jb, cn->v + (2 * j + h) * TW + i));
            };
            Unroll<0, TH>::run(ld);
            if (have && h == 0) ids[row] = cn->id[i];
            __syncthreads();
            cnt = reg_gepp<T, NT>(x, nr, jb, order, gl, step_of);
        }
        // a node with one child nominates the child's rows unchanged
        loff = poff;
        lsize = (lsize + AR - 1) / AR;
        idx = parent;
    }
    // the root: cnt == jb pivot rows, the pivot row of step k holding row k of L11 \ U11
    if (step_of >= 0) {
        auto st = [&](auto jc) __attribute__((always_inline)) {
            constexpr int j = decltype(jc)::value;
            top[(2 * j + h) * TW + step_of] = csel(2 * j + h < jb, x.get(j), zero);
        };
        Unroll<0, TH>::run(st);
    }
    for (int e = t; e < TW * (TW - jb); e += NT) {          // identity padding of a short last panel
        const int i = jb + e / TW, c = e % TW;
        top[c * TW + i] = CX<T>::make(i == c, 0);
    }
    __syncthreads();
    if (t == 0 && info) {
        for (int k = 0; k < jb; ++k) {
            const T d = top[k * TW + k];
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

template <typename T>
__global__ __launch_bounds__(2 * Tour<T>::RL) void k_tslu(const T *A, long lda, long n, long j0, int jb, int P,
                                                         TNode<T> *nodes, int *arrive, T *top, int *ipiv, int *pairs,
                                                         int *info)
{
    __builtin_amdgcn_s_setprio(3);    // ahead of the trailing update's waves where the LU's look-ahead runs it beside
    tslu_body<T>(A, lda, n, j0, jb, P, nodes, arrive, top, ipiv, pairs, info, blockIdx.x);
}

// item blockIdx.y of a batch; leaves past its own count, and items already factored, leave at once
template <typename T>
__global__ __launch_bounds__(2 * Tour<T>::RL) void k_tslu_b(const m25x_lu_item *items, long j0)
{
    const m25x_lu_item it = items[blockIdx.y];
    const long n = it.n;
    if (j0 >= n) return;
    const int jb = (int)std::min((long)TW, n - j0);
    const int P = (int)tour_leaves(n - j0, Tour<T>::RL);
    if ((int)blockIdx.x >= P) return;
    const TourWork<T> w = tour_carve<T>(it.work, n);
    tslu_body<T>((const T *)it.A, it.lda, n, j0, jb, P, w.nodes, w.arrive, w.top, it.ipiv, w.pairs, it.info,
                 blockIdx.x);
}

}  // namespace

#ifdef M25X_TSLU_C64
typedef float2 TT;
#else
typedef double2 TT;
#endif

int launch_tslu(const TT *A, long lda, long n, long j0, int jb, int P, TNode<TT> *nodes, int *arrive, TT *top,
                int *ipiv, int *pairs, int *info, hipStream_t s)
{
    hipLaunchKernelGGL(k_tslu<TT>, dim3(P), dim3(2 * Tour<TT>::RL), 0, s, A, lda, n, j0, jb, P, nodes, arrive, top, ipiv,
                       pairs, info);
    return check_launch("getrf/tournament");
}

int launch_tslu_b(const TT *, const m25x_lu_item *items, int nitems, long j0, int Pmax, hipStream_t s)
{
    hipLaunchKernelGGL(k_tslu_b<TT>, dim3(Pmax, nitems), dim3(2 * Tour<TT>::RL), 0, s, items, j0);
    return check_launch("getrf_batched/tournament");
}

}  // namespace lu
}  // namespace m25x
