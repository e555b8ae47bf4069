// Lane-vector helpers for the row-marching kernels on pitched 2D arrays.
//
// A lane owns N adjacent cells of a row moved as one 16-byte vector (2 FP64
// or 4 FP32 cells), so a wave covers 64*N columns with dwordx4 accesses.
// Windows start on a multiple of N; cells of a lane outside the window's
// output range are masked per cell.
#pragma once

#include "tl_common.h"

namespace tl {

typedef double dvec2 __attribute__((ext_vector_type(2)));
typedef float fvec4 __attribute__((ext_vector_type(4)));

template <class T> struct Lane16;
template <> struct Lane16<double> { static constexpr int N = 2; using V = dvec2; };
template <> struct Lane16<float> { static constexpr int N = 4; using V = fvec4; };

// The N cells of one lane.
template <class T>
struct Cells {
    T v[Lane16<T>::N];
};

template <class T>
__device__ __forceinline__ Cells<T> ldc(const T* __restrict__ p, long i)
{
    return __builtin_bit_cast(Cells<T>, *reinterpret_cast<const typename Lane16<T>::V*>(p + i));
}

template <class T>
__device__ __forceinline__ void stc(T* __restrict__ p, long i, const Cells<T>& c)
{
    *reinterpret_cast<typename Lane16<T>::V*>(p + i) = __builtin_bit_cast(typename Lane16<T>::V, c);
}

template <class T>
__device__ __forceinline__ Cells<T> zeros()
{
    Cells<T> c;
#pragma unroll
    for (int i = 0; i < Lane16<T>::N; ++i) c.v[i] = T(0);
    return c;
}

// Row pitch X (a multiple of N), interior columns [hd, hd+nx), rows [hd, hd+ny).
struct PitchGrid {
    int X, hd, nx, ny;
};

// Strips of 64*N-column windows whose centre `w` columns are written.
struct StripTile {
    int w;     // output columns per wave (a multiple of N)
    int cw0;   // first column of strip 0's window (a multiple of N; may be < 0)
    int ns;    // strips along x
    int nbx;   // blocks along x (kWaves strips each)
    int nby;   // tiles along y
    int ty;    // output rows per tile
};

// Strip geometry for a window that must extend `margin` columns beyond its
// output columns on both sides.
template <class T>
inline StripTile strip_tile(const PitchGrid& g, int margin, int ty, int waves_per_block)
{
    constexpr int N = Lane16<T>::N;
    StripTile t;
    const int off = margin + (((g.hd - margin) % N) + N) % N;
    t.w = (TL_WAVE * N - off - margin) / N * N;
    t.cw0 = g.hd - off;
    t.ns = (int)tl_cdiv(g.nx, t.w);
    t.nbx = (int)tl_cdiv(t.ns, waves_per_block);
    t.ty = ty;
    t.nby = (int)tl_cdiv(g.ny, t.ty);
    return t;
}

// blockIdx -> tile so that each XCD (blocks are dealt round-robin to the 8
// XCDs) works on a contiguous run of tiles and shares their rows in its L2.
__device__ __forceinline__ int xcd_tile_of(int bid, int nb)
{
    const int q = nb >> 3, r = nb & 7, k = bid & 7, i = bid >> 3;
    return k * q + (k < r ? k : r) + i;
}

// A * a for the N cells of a lane (2d/shared.h:50-54, same association), with
// x neighbours outside the interior mirroring the cell (a depth-1 reflective
// halo).  up/dn are the rows above/below, already mirrored by the caller.
// The first lane's left neighbour is el and the last lane's right neighbour
// and kx are er, ek (they do not exist in the wave).
template <class T>
__device__ __forceinline__ Cells<T> apply_a_lane(const Cells<T>& cen, const Cells<T>& up,
                                                 const Cells<T>& dn, const Cells<T>& kxc,
                                                 const Cells<T>& kyu, const Cells<T>& kyc, int c,
                                                 int xin0, int xin1, T el, T er, T ek)
{
    constexpr int N = Lane16<T>::N;
    const T lft_edge = tl_lane_prev(cen.v[N - 1], el);
    const T rgt_edge = tl_lane_next(cen.v[0], er);
    const T kx_edge = tl_lane_next(kxc.v[0], ek);
    Cells<T> a;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        T lft = i == 0 ? lft_edge : cen.v[i - 1];
        T rgt = i == N - 1 ? rgt_edge : cen.v[i + 1];
        if (c + i == xin0) lft = cen.v[i];
        if (c + i == xin1 - 1) rgt = cen.v[i];
        const T kxr = i == N - 1 ? kx_edge : kxc.v[i + 1];
        const T d = (T(1) + (kxr + kxc.v[i])) + (kyu.v[i] + kyc.v[i]);
        a.v[i] = d * cen.v[i] - (kxr * rgt + kxc.v[i] * lft) - (kyu.v[i] * up.v[i] + kyc.v[i] * dn.v[i]);
    }
    return a;
}

// The same where the wave's end lanes never produce kept results (windows
// that overlap their neighbours): the edge neighbours are don't-cares.
template <class T>
__device__ __forceinline__ Cells<T> apply_a_lane(const Cells<T>& cen, const Cells<T>& up,
                                                 const Cells<T>& dn, const Cells<T>& kxc,
                                                 const Cells<T>& kyu, const Cells<T>& kyc, int c,
                                                 int xin0, int xin1)
{
    constexpr int N = Lane16<T>::N;
    return apply_a_lane(cen, up, dn, kxc, kyu, kyc, c, xin0, xin1, cen.v[0], cen.v[N - 1],
                        kxc.v[N - 1]);
}

// The coefficient parts of that product that do not depend on a: the
// diagonal (1 + (kx[i+1] + kx[i])) + (ky[i+x] + ky[i]) of each cell and the
// kx[i+1] of the lane's last cell (the next lane's first kx; the wave's end
// lane gets a don't-care).  A kernel that applies A to several fields on
// the same row forms them once; the values are the same bits either way.
template <class T>
__device__ __forceinline__ T kx_next_lane(const Cells<T>& kxc)
{
    constexpr int N = Lane16<T>::N;
    return tl_lane_next(kxc.v[0], kxc.v[N - 1]);
}

template <class T>
__device__ __forceinline__ Cells<T> diag_lane(const Cells<T>& kxc, T kxe, const Cells<T>& kyu,
                                              const Cells<T>& kyc)
{
    constexpr int N = Lane16<T>::N;
    Cells<T> d;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const T kxr = i == N - 1 ? kxe : kxc.v[i + 1];
        d.v[i] = (T(1) + (kxr + kxc.v[i])) + (kyu.v[i] + kyc.v[i]);
    }
    return d;
}

// kx of a lane's cells on one row and ky of two rows' faces, formed from
// the conductivity w (cg_init's w: density, or 1/density with the
// reciprocal coefficient; 2d/c_kernels/omp3/cg.c:45-53): kx from the row's
// w (wc) and the left neighbour of the wave's first cell (wl_edge), ky from
// the row below (wd) and the row (wc); kz likewise from two planes.
__device__ __forceinline__ Cells<double> kx_of_row(double rx, const Cells<double>& wc,
                                                   double wl_edge)
{
    Cells<double> k;
    k.v[0] = tl_face_coef(rx, tl_lane_prev(wc.v[1], wl_edge), wc.v[0]);
    k.v[1] = tl_face_coef(rx, wc.v[0], wc.v[1]);
    return k;
}

__device__ __forceinline__ Cells<double> ky_of_rows(double ry, const Cells<double>& wd,
                                                    const Cells<double>& wc)
{
    Cells<double> k;
    k.v[0] = tl_face_coef(ry, wd.v[0], wc.v[0]);
    k.v[1] = tl_face_coef(ry, wd.v[1], wc.v[1]);
    return k;
}

// a * b + c.  FP64: two roundings, the reference's arithmetic.  FP32 (the
// sweeps of the FP32-refined solve, which FP64 residual replacement
// corrects): one fused multiply-add, fewer instructions in the issue-bound
// fused loops.
__device__ __forceinline__ double mad(double a, double b, double c) { return a * b + c; }
__device__ __forceinline__ float mad(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// apply_a_lane (don't-care wave ends) with the diagonal and next-lane kx
// formed beforehand.  FP32 uses fused multiply-adds (mad): the x and y
// terms as mad(a, b, c*d), then mad(diag, sd, -x terms) - y terms.  XMIRROR: the lane's cells may include the first or
// last interior column (a wave whose window holds neither skips the tests).
template <bool XMIRROR, class T>
__device__ __forceinline__ Cells<T> apply_a_pre(const Cells<T>& cen, const Cells<T>& up,
                                                const Cells<T>& dn, const Cells<T>& kxc, T kxe,
                                                const Cells<T>& kyu, const Cells<T>& kyc,
                                                const Cells<T>& d, int c, int xin0, int xin1)
{
    constexpr int N = Lane16<T>::N;
    const T lft_edge = tl_lane_prev(cen.v[N - 1], cen.v[0]);
    const T rgt_edge = tl_lane_next(cen.v[0], cen.v[N - 1]);
    Cells<T> a;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        T lft = i == 0 ? lft_edge : cen.v[i - 1];
        T rgt = i == N - 1 ? rgt_edge : cen.v[i + 1];
        if constexpr (XMIRROR) {
            if (c + i == xin0) lft = cen.v[i];
            if (c + i == xin1 - 1) rgt = cen.v[i];
        }
        const T kxr = i == N - 1 ? kxe : kxc.v[i + 1];
        const T tx = mad(kxr, rgt, kxc.v[i] * lft), ty = mad(kyu.v[i], up.v[i], kyc.v[i] * dn.v[i]);
        a.v[i] = mad(d.v[i], cen.v[i], -tx) - ty;
    }
    return a;
}

}  // namespace tl
