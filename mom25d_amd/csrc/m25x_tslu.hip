// m25x_tslu.hip - the tournament of a tournament-pivoting LU panel (k_tslu), built once per element type
// (M25X_TSLU_C64 set: complex64, else complex128) so the two fully unrolled instantiations compile in parallel.
#include "m25x_lu_dev.h"

namespace m25x {
namespace lu {
namespace {

// Wave argmax of (v, key): the larger v, then the smaller key; every lane ends with the result. Within each row
// of 16 lanes by DPP exchanges (quad xor 1, quad xor 2, half-row mirror, row mirror), then the four rows' results
// combined in a fixed order.
__device__ inline void better_key(double &v, int &key, double v2, int k2)
{
    if (v2 > v || (v2 == v && k2 < key)) { v = v2; key = k2; }
}

template <int CTRL> __device__ inline void dpp_better(double &v, int &key)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    const int k2 = __builtin_amdgcn_update_dpp(0, key, CTRL, 0xF, 0xF, false);
    better_key(v, key, __longlong_as_double(((long long)hi << 32) | (unsigned)lo), k2);
}

__device__ inline void wave_argmax(double &v, int &key)
{
    dpp_better<0xB1>(v, key);       // quad_perm [1, 0, 3, 2]
    dpp_better<0x4E>(v, key);       // quad_perm [2, 3, 0, 1]
    dpp_better<0x141>(v, key);      // row_half_mirror
    dpp_better<0x140>(v, key);      // row_mirror
    const long long b = __double_as_longlong(v);
    double bv = -2.0;
    int bk = 0x7fffffff;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int lo = __builtin_amdgcn_readlane((int)b, 16 * r), hi = __builtin_amdgcn_readlane((int)(b >> 32), 16 * r);
        better_key(bv, bk, __longlong_as_double(((long long)hi << 32) | (unsigned)lo),
                   __builtin_amdgcn_readlane(key, 16 * r));
    }
    v = bv;
    key = bk;
}

// Calls f(integral_constant<int, K>) for K = B .. E - 1: the steps of an elimination with every register index a
// compile-time constant (a row held in registers is never indexed at run time, so it stays in registers).
template <int B, int E> struct Unroll {
    template <class F> __device__ __attribute__((always_inline)) static void run(F &f)
    {
        f(std::integral_constant<int, B>{});
        Unroll<B + 1, E>::run(f);
    }
};
template <int E> struct Unroll<E, E> {
    template <class F> __device__ __attribute__((always_inline)) static void run(F &) {}
};

// GEPP of rows held in registers, two threads a row: in wave w, lanes L and L + 32 hold row 32 w + L % 32, the
// thread of half h = L / 32 columns 2 j + h in x[j] (j < TW / 2), so both halves have equal work at every step.
// nr rows over columns [0, nc). Rows never move: a row's position in the elimination order is tracked instead
// (pos), and ties in the pivot search go to the earliest position (LAPACK's rule). Step k: the pivot row publishes
// itself in prow (LDS), the live rows take their multiplier and update. order[k] = the k-th pivot row; a pivot
// row's threads keep its step in step_of. On return a pivot row holds its row of L \ U, the other rows their
// multipliers and Schur complement. Two barriers a step; the per-wave maxima double-buffered.
constexpr int TH = TW / 2;
template <typename T> __device__ inline T shfl_xor32(T v)
{
    return CX<T>::make(__shfl_xor(v.x, 32), __shfl_xor(v.y, 32));
}

// A row's half in registers as separate real and imaginary arrays (an array of the vector-typed complex is kept
// in scratch memory for complex128).
template <typename T> struct RowH {
    typename CX<T>::R re[TH], im[TH];
    __device__ T get(int j) const { return CX<T>::make(re[j], im[j]); }
    __device__ void set(int j, T v) { re[j] = v.x; im[j] = v.y; }
};

template <typename T, int NT>
__device__ __attribute__((always_inline)) int reg_gepp(RowH<T> &x, int nr, int nc, int *order, T *prow,
                                                       double (*wv)[NT / 64], int (*wk)[NT / 64], int &step_of)
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
        const T mine = x.get(jk), other = shfl_xor32(mine);
        const T a = h == hk ? mine : other;                 // column k of this row
        double v = (live && h == hk) ? cabs1(a) : -1.0;
        if (live && h == hk && !(v >= 0)) v = 0;           // a NaN never outranks a live row
        int key = (live && h == hk) ? ((pos << 16) | row) : 0x7fffffff;
        wave_argmax(v, key);
        const int b = k & 1;
        if (lane == 0) { wv[b][w] = v; wk[b][w] = key; }
        __syncthreads();
        v = wv[b][0];
        key = wk[b][0];
#pragma unroll
        for (int q = 1; q < NW; ++q) better_key(v, key, wv[b][q], wk[b][q]);
        const int ps = key & 0xffff, p = key >> 16;
        if (row == ps) {
            auto pub = [&](auto jc) __attribute__((always_inline)) {
                constexpr int j = decltype(jc)::value;
                prow[2 * j + h] = x.get(j);
            };
            Unroll<jk, TH>::run(pub);
            live = false;
            step_of = k;
        }
        if (t == 0) order[k] = ps;
        __syncthreads();
        if (live) {
            if (pos == k) pos = p;
            const T pv = prow[k];
            const T l = (pv.x == 0 && pv.y == 0) ? a : tdiv(a, pv);
            if (h == hk) x.set(jk, l);
            else if (hk == 0) x.set(jk, tsub(x.get(jk), tmul(l, prow[2 * jk + 1])));   // column k + 1 of half 1
            auto upd = [&](auto jc) __attribute__((always_inline)) {
                constexpr int j = decltype(jc)::value;
                x.set(j, tsub(x.get(j), tmul(l, prow[2 * j + h])));
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
__global__ __launch_bounds__(2 * Tour<T>::RL) void k_tslu(const T *A, long lda, long n, long j0, int jb, int P,
                                                         TNode<T> *nodes, int *arrive, T *top, int *ipiv, int *pairs,
                                                         int *info)
{
    constexpr int RL = Tour<T>::RL, AR = Tour<T>::AR, NT = 2 * RL, NW = NT / 64;
    __shared__ T prow[TW];
    __shared__ int ids[RL];
    __shared__ int order[TW];
    __shared__ double wv[2][NW];
    __shared__ int wk[2][NW];
    __shared__ int go, counts[AR];
    const int t = threadIdx.x, lane = t & 63;
    const int row = (t >> 6) * 32 + (lane & 31), h = lane >> 5;
    const T zero = CX<T>::make(0, 0);
    int idx = blockIdx.x;
    const long r0 = j0 + (long)idx * RL;
    int nr = (int)std::min((long)RL, n - r0);
    RowH<T> x;
    {
        auto ld = [&](auto jc) __attribute__((always_inline)) {
            constexpr int j = decltype(jc)::value;
            x.set(j, (2 * j + h < jb && row < nr) ? A[(j0 + 2 * j + h) * lda + r0 + row] : zero);
        };
        Unroll<0, TH>::run(ld);
    }
    if (t < nr) ids[t] = (int)(r0 + t);
    __syncthreads();
    int step_of;
    int cnt = reg_gepp<T, NT>(x, nr, jb, order, prow, wv, wk, step_of);
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
                        v[j] = (li < cnt && c < jb) ? A[(j0 + c) * lda + r] : zero;
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
                x.set(j, (have && 2 * j + h < jb) ? cn->v[(2 * j + h) * TW + i] : zero);
            };
            Unroll<0, TH>::run(ld);
            if (have && h == 0) ids[row] = cn->id[i];
            __syncthreads();
            cnt = reg_gepp<T, NT>(x, nr, jb, order, prow, wv, wk, step_of);
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
            top[(2 * j + h) * TW + step_of] = 2 * j + h < jb ? x.get(j) : zero;
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

}  // namespace lu
}  // namespace m25x
