// m25x_lu_dev.h - device arithmetic shared by the LU sources (m25x_lu.hip and the two builds of m25x_tslu.hip):
// complex traits over double2 / float2, the tournament's tree layout and its launchers.
#pragma once
#include <algorithm>
#include <type_traits>
#include "m25x_common.h"

namespace m25x {
namespace lu {

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
// t - m y in four fused multiply-adds
__device__ inline float2 tfms(float2 t, float2 m, float2 y)
{
    return make_float2(fmaf(-m.x, y.x, fmaf(m.y, y.y, t.x)), fmaf(-m.y, y.x, fmaf(-m.x, y.y, t.y)));
}
__device__ inline double2 tfms(double2 t, double2 m, double2 y)
{
    return make_double2(fma(-m.x, y.x, fma(m.y, y.y, t.x)), fma(-m.y, y.x, fma(-m.x, y.y, t.y)));
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
// *p where c holds, else zero: a guarded load (a select between the loaded value and a zero constant is lowered as a
// select of addresses, with the zero kept in scratch memory and the load through a generic pointer)
template <typename T> __device__ __attribute__((always_inline)) inline T ldz(bool c, const T *p)
{
    T v = CX<T>::make(0, 0);
    if (c) v = *p;
    return v;
}

// a ? x : y component by component (a select of the vector-typed complex itself is lowered through scratch memory
// for complex128)
template <typename T> __device__ inline T csel(bool c, T a, T b) { return CX<T>::make(c ? a.x : b.x, c ? a.y : b.y); }
template <typename T> __device__ inline double cabs1(T a) { return fabs((double)a.x) + fabs((double)a.y); }

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
template <> struct Tour<double2> { static constexpr int RL = 256, AR = 4; };   // rows a leaf, children a node
template <> struct Tour<float2> { static constexpr int RL = 256, AR = 4; };

template <typename T> struct TNode {
    int id[TW];                 // the nominees' global rows, in their pivot order
    int count, pad[63];
    T v[TW * TW];               // their original panel rows, column-major: v[c * TW + i]
};

__host__ __device__ inline long tour_leaves(long m, int RL) { return (m + RL - 1) / RL; }
__host__ __device__ inline long tour_nodes(long n, int RL, int AR)
{
    long P = tour_leaves(n, RL), tot = 0;
    while (P > 1) { tot += P; P = (P + AR - 1) / AR; }
    return tot + 1;
}

// A tournament LU's workspace, carved from one buffer: the tree nodes, the panel's top block, the nodes' arrival
// counts and the row moves.
template <typename T> struct TourWork {
    TNode<T> *nodes;
    T *top;
    int *arrive, *pairs;
    long nn;
};
template <typename T> __host__ __device__ inline TourWork<T> tour_carve(void *work, long n)
{
    TourWork<T> w;
    w.nn = tour_nodes(n, Tour<T>::RL, Tour<T>::AR);
    char *p = (char *)work;
    w.nodes = (TNode<T> *)p;
    p += w.nn * sizeof(TNode<T>);
    w.top = (T *)p;
    p += (size_t)TW * TW * sizeof(T);
    w.arrive = (int *)p;
    w.pairs = w.arrive + w.nn;
    return w;
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

constexpr int TH = TW / 2;

// A row's half in registers as separate real and imaginary arrays (an array of the vector-typed complex is kept
// in scratch memory for complex128).
template <typename T> struct RowH {
    typename CX<T>::R re[TH], im[TH];
    __device__ T get(int j) const { return CX<T>::make(re[j], im[j]); }
    __device__ void set(int j, T v) { re[j] = v.x; im[j] = v.y; }
};

// One tournament launch (k_tslu) for panel [j0, j0 + jb): P leaves; defined in m25x_tslu.hip, one build per type.
int launch_tslu(const double2 *A, long lda, long n, long j0, int jb, int P, TNode<double2> *nodes, int *arrive,
                double2 *top, int *ipiv, int *pairs, int *info, hipStream_t s);
int launch_tslu(const float2 *A, long lda, long n, long j0, int jb, int P, TNode<float2> *nodes, int *arrive,
                float2 *top, int *ipiv, int *pairs, int *info, hipStream_t s);
// The tournaments of panel step j0 of a batch (items[0 .. nitems)): grid (Pmax leaves, nitems); the element type
// is the tag's (a null pointer).
int launch_tslu_b(const double2 *tag, const m25x_lu_item *items, int nitems, long j0, int Pmax, hipStream_t s);
int launch_tslu_b(const float2 *tag, const m25x_lu_item *items, int nitems, long j0, int Pmax, hipStream_t s);

}  // namespace lu
}  // namespace m25x
