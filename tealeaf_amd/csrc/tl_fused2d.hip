// Temporally blocked PPCG inner loop (2D): S consecutive inner steps of
// 2d/drivers/ppcg_driver.c:165-183 in one pass over memory, in FP64 (the
// reference's arithmetic) or FP32 (the inner sweeps of the mixed-precision
// solve).
//
// The reference runs, per inner step pp, a depth-1 halo update of sd
// followed by ppcg_inner_iteration (2d/c_kernels/omp3/ppcg.c:28-61):
//     r  -= A sd ;  u += sd ;  sd = alpha_pp sd + beta_pp r.
// With every face external the halo update is a mirror, so a neighbour
// outside the interior simply takes the cell's own value.  Every time level
// k of a cell depends only on level k-1 within one cell of it, so a wave can
// carry S levels at once:
//  * each lane holds N adjacent cells of a row (one 16-byte vector: 2 FP64
//    or 4 FP32 cells); the wave covers a window of 64*N columns;
//  * it marches up its rows keeping, for each level, the last three rows of
//    sd and the last row of r and u in registers; level k computes row
//    (j - k + 1) while level 1 computes row j;
//  * windows overlap their neighbours by at least S columns on each side and
//    a wave writes only its centre columns, where all S levels are exact;
//    the march starts S-1 rows below its tile for the same reason in y;
//  * r, u and sd are read from one buffer and written to another, because
//    the overlapped windows of other waves still need the level-0 values.
// Each level applies the reference's per-cell expression in the reference's
// order, so in FP64 the result is bit-for-bit that of S separate steps.  HBM
// traffic for S steps is that of one: read sd, kx, ky, r, u; write sd, r, u.
// Optionally sd at level 0 is ppcg_init's r/theta (ppcg.c:8-25), computed on
// the fly from r, and r, u at level 0 may first take cg_calc_ur's update
// u += alpha p, r -= alpha w (cg.c:116-143; the PPCG outer step that
// precedes the inner loop), also on the fly.
#include "tl_fused.h"
#include "tl_mixed.h"
#include "tl_loop.h"
#include "tealeaf_hip.h"

#include <algorithm>
#include <type_traits>

using namespace tl;
using namespace tl::fused;

namespace tl {
namespace fused {

int g_fused_ty = -1;       // output rows per tile (<= 0: default)
int g_fused_pf = -1;       // row prefetch (< 0: default)
int g_fused_variant = 0;   // 0: LDS-ring PPCG kernel; 1: register windows

// Compute units of the current device (cached; the runtime uses one device).
int device_cus()
{
    static int cus = 0;
    if (cus <= 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            cus <= 0)
            cus = 256;
    }
    return cus;
}

}  // namespace fused
}  // namespace tl

namespace {

// Level 0 of the loop: MODE 0 reads sd; 1 forms sd = r/theta (ppcg_init);
// 2 first applies cg_calc_ur to r and u, then forms sd = r/theta.
template <class T, int S, int MODE, int PF>
__global__ void __launch_bounds__(kBlock)
k_ppcg_fused2d(PitchGrid g, StripTile t, PpcgCoefs<T, S> cf, T theta, T alpha_arg,
               const double* __restrict__ alpha_dev,
               const T* __restrict__ sd_in, T* __restrict__ sd_out,
               const T* __restrict__ r_in, T* __restrict__ r_out,
               const T* __restrict__ u_in, T* __restrict__ u_out,
               const T* __restrict__ kx, const T* __restrict__ ky,
               const T* __restrict__ p_in, const T* __restrict__ w_in, TlGate gate)
{
    constexpr bool FROM_R = MODE >= 1;
    constexpr bool UR = MODE == 2;
    constexpr int N = Lane16<T>::N;
    using C = Cells<T>;
    if (gate.closed()) return;
    // cg_calc_ur's alpha: the argument, or the resident loop's device value
    const T alpha_ur = UR && alpha_dev ? (T)*alpha_dev : alpha_arg;
    const int lane = threadIdx.x & (TL_WAVE - 1), wave = threadIdx.x / TL_WAVE;
    const int tile = xcd_tile_of(blockIdx.x, t.nbx * t.nby);
    const int bx = tile % t.nbx, by = tile / t.nbx;
    const int sx = bx * kWaves + wave;
    if (sx >= t.ns) return;                      // uniform per wave; no block sync below

    const int xin0 = g.hd, xin1 = g.hd + g.nx;  // interior columns
    const int yin0 = g.hd, yin1 = g.hd + g.ny;  // interior rows
    const int o0 = xin0 + sx * t.w;
    const int o1 = min(o0 + t.w, xin1);
    const int c = t.cw0 + sx * t.w + N * lane;  // this lane's first column
    const bool ld_ok = c >= 0 && c + N <= g.X;
    const bool all_out = c >= o0 && c + N <= o1;
    const bool any_out = c + N > o0 && c < o1;

    const int j0 = yin0 + by * t.ty;
    const int j1 = min(j0 + t.ty, yin1);
    const int rs = j0 - S + 1;                  // first row of level 1
    const int niter = j1 - j0 + 2 * S - 2;
    const long X = g.X;

    // Field rows are read through buffer descriptors based at the first
    // row this tile touches.  A lane whose row is outside [lo, hi) or whose
    // cells are outside the array passes an offset beyond the descriptor's
    // range, and the hardware returns zeros: no load sits behind a branch,
    // so the compiler can keep prefetched rows in flight across iterations.
    const int base_row = max(rs - S, 0);
    const long span = (long)(g.ny + 2 * g.hd - base_row) * X * (long)sizeof(T);
    const int nrec = (int)min(span, (long)kBufferRange);
    auto rsrc = [&](const T* f) {
        return __builtin_amdgcn_make_buffer_rsrc((void*)(f + (long)base_row * X), (short)0, nrec,
                                                 kBufferWord3);
    };
    auto ld_field = [&](__amdgpu_buffer_rsrc_t f, int row, int lo, int hi) -> C {
        const bool ok = ld_ok && row >= lo && row < hi;
        const unsigned off =
            ok ? (unsigned)(((long)(row - base_row) * X + c) * (long)sizeof(T)) : kBufferOutside;
        return __builtin_bit_cast(C, __builtin_amdgcn_raw_buffer_load_b128(f, off, 0, 0));
    };
    const __amdgpu_buffer_rsrc_t b_sd = rsrc(FROM_R ? r_in : sd_in), b_r = rsrc(r_in),
                                 b_u = rsrc(u_in), b_kx = rsrc(kx), b_ky = rsrc(ky),
                                 b_p = rsrc(UR ? p_in : u_in), b_w = rsrc(UR ? w_in : u_in);
    auto over_theta = [&](C a) -> C {
#pragma unroll
        for (int i = 0; i < N; ++i) a.v[i] = a.v[i] / theta;
        return a;
    };

    // Windows (index 0 oldest).  sdw[k][*]: rows of level k for k < S.
    C sdw[S][3];
    C rprev[S], uprev[S];     // level k's r, u from the previous iteration
    C kxw[S], kyw[S + 1];     // kx rows rho1 - m; ky rows rho1 + 1 - m
#pragma unroll
    for (int k = 0; k < S; ++k) {
        sdw[k][0] = sdw[k][1] = sdw[k][2] = zeros<T>();
        rprev[k] = uprev[k] = zeros<T>();
        kxw[k] = zeros<T>();
    }
#pragma unroll
    for (int m = 0; m <= S; ++m) kyw[m] = zeros<T>();

    // r - alpha w and u + alpha p (cg.c:128-129, same association)
    auto minus_aw = [&](C a, const C& b) -> C {
#pragma unroll
        for (int i = 0; i < N; ++i) a.v[i] = a.v[i] - alpha_ur * b.v[i];
        return a;
    };
    auto plus_ap = [&](C a, const C& b) -> C {
#pragma unroll
        for (int i = 0; i < N; ++i) a.v[i] = a.v[i] + alpha_ur * b.v[i];
        return a;
    };
    auto r_level0 = [&](int row) -> C {
        C a = ld_field(b_r, row, yin0, yin1);
        if constexpr (UR) a = minus_aw(a, ld_field(b_w, row, yin0, yin1));
        return a;
    };

    // Prologue: level-0 sd rows rs-1, rs; kx rows below rs; ky rows up to rs.
    C rcarry = zeros<T>();    // FROM_R: level-0 r at row rho1, loaded one iteration ago
    if constexpr (FROM_R) {
        sdw[0][1] = over_theta(r_level0(rs - 1));
        rcarry = r_level0(rs);
        sdw[0][2] = over_theta(rcarry);
    } else {
        sdw[0][1] = ld_field(b_sd, rs - 1, yin0, yin1);
        sdw[0][2] = ld_field(b_sd, rs, yin0, yin1);
    }
#pragma unroll
    for (int m = 0; m + 1 < S; ++m) kxw[m] = ld_field(b_kx, rs - 1 - m, yin0, yin1);
#pragma unroll
    for (int m = 0; m < S; ++m) kyw[m] = ld_field(b_ky, rs - m, yin0, yin1 + 1);

    // Level-0 inputs of the iteration whose level 1 computes row rho1:
    // a = sd (or r for FROM_R) at rho1+1, b = r at rho1 (or, for UR, w at
    // rho1+1), u at rho1 (and p at rho1 for UR), kx at rho1, ky at rho1+1.
    // They are combined only when used, so that prefetched loads stay in
    // flight across the arithmetic.
    struct In {
        C a, b, u, p, kx, ky;
    };
    auto load_in = [&](int rho1) -> In {
        In v;
        v.a = ld_field(b_sd, rho1 + 1, yin0, yin1);
        if constexpr (!FROM_R) v.b = ld_field(b_r, rho1, yin0, yin1);
        if constexpr (UR) {
            v.b = ld_field(b_w, rho1 + 1, yin0, yin1);
            v.p = ld_field(b_p, rho1, yin0, yin1);
        }
        v.u = ld_field(b_u, rho1, yin0, yin1);
        v.kx = ld_field(b_kx, rho1, yin0, yin1);
        v.ky = ld_field(b_ky, rho1 + 1, yin0, yin1 + 1);
        return v;
    };
    // With PF > 0 the loads of the row PF iterations ahead are issued
    // before this iteration's arithmetic, which then hides their latency.
    In pre[PF > 0 ? PF : 1];
    if constexpr (PF > 0) {
#pragma unroll
        for (int q = 0; q < PF; ++q) pre[q] = load_in(rs + q);
    }

    // With PRE, the diagonal and next-lane kx of rows rho1 - m (m < S) are
    // formed once per row for all the levels that use it; deep loops (whose
    // windows already fill the registers) form them per level instead.
    constexpr bool PRE = S <= 6;
    constexpr int SW = PRE ? S : 1;
    C dgw[SW];
    T kxew[SW];
#pragma unroll
    for (int m = 0; m < SW; ++m) {
        dgw[m] = zeros<T>();
        kxew[m] = T(0);
    }
    if constexpr (PRE) {
#pragma unroll
        for (int m = 0; m + 1 < S; ++m) {  // rows rs-1-m, from the prologue's kx, ky
            kxew[m] = kx_next_lane(kxw[m]);
            dgw[m] = diag_lane(kxw[m], kxew[m], kyw[m], kyw[m + 1]);
        }
    }

    // A wave whose window holds neither the first nor the last interior
    // column never mirrors in x; an iteration none of whose levels is on
    // the first or last interior row never mirrors in y.  The loop body is
    // instantiated for each case so that the common one carries no tests.
    const int wc0 = t.cw0 + sx * t.w;
    const bool xedge = (xin0 >= wc0 && xin0 < wc0 + TL_WAVE * N) ||
                       (xin1 - 1 >= wc0 && xin1 - 1 < wc0 + TL_WAVE * N);

    auto body = [&](int it, auto xm, auto ym) __attribute__((always_inline)) {
        constexpr bool XM = decltype(xm)::value;
        constexpr bool YM = decltype(ym)::value;
        const int rho1 = rs + it;
        In cur;
        if constexpr (PF > 0) {
            cur = pre[0];
#pragma unroll
            for (int q = 0; q + 1 < PF; ++q) pre[q] = pre[q + 1];
            pre[PF - 1] = load_in(rho1 + PF);    // rows past the tile: loaded, unused
        } else {
            cur = load_in(rho1);
        }
        C sd_new, r0, u0 = cur.u;
        if constexpr (FROM_R) {
            r0 = rcarry;
            rcarry = UR ? minus_aw(cur.a, cur.b) : cur.a;
            sd_new = over_theta(rcarry);
        } else {
            r0 = cur.b;
            sd_new = cur.a;
        }
        if constexpr (UR) u0 = plus_ap(u0, cur.p);
        sdw[0][0] = sdw[0][1];
        sdw[0][1] = sdw[0][2];
        sdw[0][2] = sd_new;
#pragma unroll
        for (int m = S - 1; m > 0; --m) kxw[m] = kxw[m - 1];
        kxw[0] = cur.kx;
#pragma unroll
        for (int m = S; m > 0; --m) kyw[m] = kyw[m - 1];
        kyw[0] = cur.ky;
        if constexpr (PRE) {
#pragma unroll
            for (int m = S - 1; m > 0; --m) {
                kxew[m] = kxew[m - 1];
                dgw[m] = dgw[m - 1];
            }
            kxew[0] = kx_next_lane(kxw[0]);
            dgw[0] = diag_lane(kxw[0], kxew[0], kyw[0], kyw[1]);
        }

        C rcur[S], ucur[S];
#pragma unroll
        for (int k = 1; k <= S; ++k) {
            const int rho = rho1 - (k - 1);
            const C cen = sdw[k - 1][1];
            // y neighbours; outside the interior they mirror the cell
            C upv = sdw[k - 1][2], dnv = sdw[k - 1][0];
            if constexpr (YM) {
                if (rho == yin1 - 1) upv = cen;
                if (rho == yin0) dnv = cen;
            }
            const C rin = k == 1 ? r0 : rprev[k - 1];
            const C uin = k == 1 ? u0 : uprev[k - 1];
            // SMVP (2d/shared.h:50-54) and the ppcg.c:46-59 updates, same order
            T kxe;
            C dg;
            if constexpr (PRE) {
                kxe = kxew[k - 1];
                dg = dgw[k - 1];
            } else {
                kxe = kx_next_lane(kxw[k - 1]);
                dg = diag_lane(kxw[k - 1], kxe, kyw[k - 1], kyw[k]);
            }
            const C av = apply_a_pre<XM>(cen, upv, dnv, kxw[k - 1], kxe, kyw[k - 1], kyw[k], dg, c,
                                         xin0, xin1);
            C rn, un, sn;
#pragma unroll
            for (int i = 0; i < N; ++i) {
                rn.v[i] = rin.v[i] - av.v[i];
                un.v[i] = uin.v[i] + cen.v[i];
                sn.v[i] = cf.alpha[k - 1] * cen.v[i] + cf.beta[k - 1] * rn.v[i];
            }
            rcur[k - 1] = rn;
            ucur[k - 1] = un;
            if (k < S) {
                sdw[k][0] = sdw[k][1];
                sdw[k][1] = sdw[k][2];
                sdw[k][2] = sn;
            } else if (any_out && rho >= j0 && rho < j1) {
                const long i0 = (long)rho * X + c;
                if (all_out) {
                    stc(r_out, i0, rn);
                    stc(u_out, i0, un);
                    stc(sd_out, i0, sn);
                } else {
#pragma unroll
                    for (int i = 0; i < N; ++i)
                        if (c + i >= o0 && c + i < o1) {
                            r_out[i0 + i] = rn.v[i];
                            u_out[i0 + i] = un.v[i];
                            sd_out[i0 + i] = sn.v[i];
                        }
                }
            }
        }
#pragma unroll
        for (int k = 1; k < S; ++k) {
            rprev[k] = rcur[k - 1];
            uprev[k] = ucur[k - 1];
        }
    };

    using Yes = std::integral_constant<bool, true>;
    using No = std::integral_constant<bool, false>;
    for (int it = 0; it < niter; ++it) {
        // levels of this iteration cover rows rho1 - S + 1 .. rho1
        const int top = rs + it, bot = top - S + 1;
        const bool yedge = (yin0 >= bot && yin0 <= top) || (yin1 - 1 >= bot && yin1 - 1 <= top);
        if (xedge) {
            if (yedge) body(it, Yes{}, Yes{});
            else body(it, Yes{}, No{});
        } else {
            if (yedge) body(it, No{}, Yes{});
            else body(it, No{}, No{});
        }
    }
}

// Temporally blocked Chebyshev iterations (2D, FP64): S consecutive turns
// of the Chebyshev loop of 2d/drivers/cheby_driver.c (a depth-1 halo update
// of u, then cheby_iterate: w = A u, r = u0 - w, p = alpha_k p + beta_k r,
// then cheby_calc_u: u += p; 2d/c_kernels/omp3/cheby.c:8-24,57-85) in one
// pass.  The scheme is the PPCG loop's: each level needs only level k-1
// within one cell, the halo update is a mirror, a wave carries S levels
// while it marches up its rows and writes only its centre columns.  Per
// level it keeps three rows of u and the row of p the next level needs;
// u0, kx and ky rows are shared by all levels.  r and p of the last level
// and u are stored; w is not (nothing reads it before cg_init or cg_calc_w
// rewrite it).  Bit-exact with S single iterations.  FP32 (the sweeps of
// the FP32-refined solve) forms p with one fused multiply-add, as A u (mad).
// NORM: also the squared 2-norm of the r written, one FP64 partial per wave
// (slot blockIdx.x * kWaves + wave of rd.partials).
template <class T, int S, int PF, bool NORM>
__global__ void __launch_bounds__(kBlock)
k_cheby_fused2d(PitchGrid g, StripTile t, PpcgCoefs<T, S> cf,
                const T* __restrict__ u_in, T* __restrict__ u_out,
                const T* __restrict__ u0f, const T* __restrict__ p_in, T* __restrict__ p_out,
                T* __restrict__ r_out, const T* __restrict__ kx, const T* __restrict__ ky,
                TlReduce rd)
{
    constexpr int N = Lane16<T>::N;
    using C = Cells<T>;
    const int lane = threadIdx.x & (TL_WAVE - 1), wave = threadIdx.x / TL_WAVE;
    const int tile = xcd_tile_of(blockIdx.x, t.nbx * t.nby);
    const int bx = tile % t.nbx, by = tile / t.nbx;
    const int sx = bx * kWaves + wave;
    double* const part = NORM ? rd.partials + (size_t)blockIdx.x * kWaves + wave : nullptr;
    if (sx >= t.ns) {                            // uniform per wave; no block sync below
        if constexpr (NORM) {
            if (lane == 0) *part = 0.0;
        }
        return;
    }
    double racc = 0.0;

    const int xin0 = g.hd, xin1 = g.hd + g.nx;
    const int yin0 = g.hd, yin1 = g.hd + g.ny;
    const int o0 = xin0 + sx * t.w;
    const int o1 = min(o0 + t.w, xin1);
    const int c = t.cw0 + sx * t.w + N * lane;
    const bool ld_ok = c >= 0 && c + N <= g.X;
    const bool all_out = c >= o0 && c + N <= o1;
    const bool any_out = c + N > o0 && c < o1;

    const int j0 = yin0 + by * t.ty;
    const int j1 = min(j0 + t.ty, yin1);
    const int rs = j0 - S + 1;
    const int niter = j1 - j0 + 2 * S - 2;
    const long X = g.X;

    // Buffer-descriptor loads, as in the PPCG loop.
    const int base_row = max(rs - S, 0);
    const long span = (long)(g.ny + 2 * g.hd - base_row) * X * (long)sizeof(T);
    const int nrec = (int)min(span, (long)kBufferRange);
    auto rsrc = [&](const T* f) {
        return __builtin_amdgcn_make_buffer_rsrc((void*)(f + (long)base_row * X), (short)0, nrec,
                                                 kBufferWord3);
    };
    auto ld_field = [&](__amdgpu_buffer_rsrc_t f, int row, int lo, int hi) -> C {
        const bool ok = ld_ok && row >= lo && row < hi;
        const unsigned off =
            ok ? (unsigned)(((long)(row - base_row) * X + c) * (long)sizeof(T)) : kBufferOutside;
        return __builtin_bit_cast(C, __builtin_amdgcn_raw_buffer_load_b128(f, off, 0, 0));
    };
    const __amdgpu_buffer_rsrc_t b_u = rsrc(u_in), b_u0 = rsrc(u0f), b_p = rsrc(p_in),
                                 b_kx = rsrc(kx), b_ky = rsrc(ky);

    // Windows (index 0 oldest).  uw[k][*]: rows of u at level k for k < S;
    // pprev[k]: level k's p from the previous iteration; u0w, kxw: rows
    // rho1 - m; kyw: rows rho1 + 1 - m.
    C uw[S][3];
    C pprev[S];
    C u0w[S], kxw[S], kyw[S + 1];
#pragma unroll
    for (int k = 0; k < S; ++k) {
        uw[k][0] = uw[k][1] = uw[k][2] = zeros<T>();
        pprev[k] = u0w[k] = kxw[k] = zeros<T>();
    }
#pragma unroll
    for (int m = 0; m <= S; ++m) kyw[m] = zeros<T>();

    // Prologue: level-0 u rows rs-1, rs; u0 and kx rows below rs; ky rows up to rs.
    uw[0][1] = ld_field(b_u, rs - 1, yin0, yin1);
    uw[0][2] = ld_field(b_u, rs, yin0, yin1);
#pragma unroll
    for (int m = 0; m + 1 < S; ++m) {
        kxw[m] = ld_field(b_kx, rs - 1 - m, yin0, yin1);
        u0w[m] = ld_field(b_u0, rs - 1 - m, yin0, yin1);
    }
#pragma unroll
    for (int m = 0; m < S; ++m) kyw[m] = ld_field(b_ky, rs - m, yin0, yin1 + 1);

    // Level-0 inputs of the iteration whose level 1 computes row rho1:
    // u at rho1+1, p, u0 and kx at rho1, ky at rho1+1.
    struct In {
        C u, p, u0, kx, ky;
    };
    auto load_in = [&](int rho1) -> In {
        In v;
        v.u = ld_field(b_u, rho1 + 1, yin0, yin1);
        v.p = ld_field(b_p, rho1, yin0, yin1);
        v.u0 = ld_field(b_u0, rho1, yin0, yin1);
        v.kx = ld_field(b_kx, rho1, yin0, yin1);
        v.ky = ld_field(b_ky, rho1 + 1, yin0, yin1 + 1);
        return v;
    };
    In pre[PF > 0 ? PF : 1];
    if constexpr (PF > 0) {
#pragma unroll
        for (int q = 0; q < PF; ++q) pre[q] = load_in(rs + q);
    }

    constexpr bool PRE = S <= 6;
    constexpr int SW = PRE ? S : 1;
    C dgw[SW];
    T kxew[SW];
#pragma unroll
    for (int m = 0; m < SW; ++m) {
        dgw[m] = zeros<T>();
        kxew[m] = T(0);
    }
    if constexpr (PRE) {
#pragma unroll
        for (int m = 0; m + 1 < S; ++m) {
            kxew[m] = kx_next_lane(kxw[m]);
            dgw[m] = diag_lane(kxw[m], kxew[m], kyw[m], kyw[m + 1]);
        }
    }

    const int wc0 = t.cw0 + sx * t.w;
    const bool xedge = (xin0 >= wc0 && xin0 < wc0 + TL_WAVE * N) ||
                       (xin1 - 1 >= wc0 && xin1 - 1 < wc0 + TL_WAVE * N);

    auto body = [&](int it, auto xm, auto ym) __attribute__((always_inline)) {
        constexpr bool XM = decltype(xm)::value;
        constexpr bool YM = decltype(ym)::value;
        const int rho1 = rs + it;
        In cur;
        if constexpr (PF > 0) {
            cur = pre[0];
#pragma unroll
            for (int q = 0; q + 1 < PF; ++q) pre[q] = pre[q + 1];
            pre[PF - 1] = load_in(rho1 + PF);
        } else {
            cur = load_in(rho1);
        }
        uw[0][0] = uw[0][1];
        uw[0][1] = uw[0][2];
        uw[0][2] = cur.u;
#pragma unroll
        for (int m = S - 1; m > 0; --m) {
            kxw[m] = kxw[m - 1];
            u0w[m] = u0w[m - 1];
        }
        kxw[0] = cur.kx;
        u0w[0] = cur.u0;
#pragma unroll
        for (int m = S; m > 0; --m) kyw[m] = kyw[m - 1];
        kyw[0] = cur.ky;
        if constexpr (PRE) {
#pragma unroll
            for (int m = S - 1; m > 0; --m) {
                kxew[m] = kxew[m - 1];
                dgw[m] = dgw[m - 1];
            }
            kxew[0] = kx_next_lane(kxw[0]);
            dgw[0] = diag_lane(kxw[0], kxew[0], kyw[0], kyw[1]);
        }

        C pcur[S];
#pragma unroll
        for (int k = 1; k <= S; ++k) {
            const int rho = rho1 - (k - 1);
            const C cen = uw[k - 1][1];
            C upv = uw[k - 1][2], dnv = uw[k - 1][0];
            if constexpr (YM) {
                if (rho == yin1 - 1) upv = cen;
                if (rho == yin0) dnv = cen;
            }
            T kxe;
            C dg;
            if constexpr (PRE) {
                kxe = kxew[k - 1];
                dg = dgw[k - 1];
            } else {
                kxe = kx_next_lane(kxw[k - 1]);
                dg = diag_lane(kxw[k - 1], kxe, kyw[k - 1], kyw[k]);
            }
            const C pin = k == 1 ? cur.p : pprev[k - 1];
            // SMVP (2d/shared.h:50-54) and cheby.c:76-84, same order
            const C wv = apply_a_pre<XM>(cen, upv, dnv, kxw[k - 1], kxe, kyw[k - 1], kyw[k], dg,
                                         c, xin0, xin1);
            C rn, pn, un;
#pragma unroll
            for (int i = 0; i < N; ++i) {
                rn.v[i] = u0w[k - 1].v[i] - wv.v[i];
                pn.v[i] = mad(cf.alpha[k - 1], pin.v[i], cf.beta[k - 1] * rn.v[i]);
                un.v[i] = cen.v[i] + pn.v[i];
            }
            pcur[k - 1] = pn;
            if (k < S) {
                uw[k][0] = uw[k][1];
                uw[k][1] = uw[k][2];
                uw[k][2] = un;
            } else if (any_out && rho >= j0 && rho < j1) {
                const long i0 = (long)rho * X + c;
                if constexpr (NORM) {
#pragma unroll
                    for (int i = 0; i < N; ++i)
                        if (all_out || (c + i >= o0 && c + i < o1))
                            racc += (double)rn.v[i] * (double)rn.v[i];
                }
                if (all_out) {
                    stc(u_out, i0, un);
                    stc(p_out, i0, pn);
                    stc(r_out, i0, rn);
                } else {
#pragma unroll
                    for (int i = 0; i < N; ++i)
                        if (c + i >= o0 && c + i < o1) {
                            u_out[i0 + i] = un.v[i];
                            p_out[i0 + i] = pn.v[i];
                            r_out[i0 + i] = rn.v[i];
                        }
                }
            }
        }
#pragma unroll
        for (int k = 1; k < S; ++k) pprev[k] = pcur[k - 1];
    };

    using Yes = std::integral_constant<bool, true>;
    using No = std::integral_constant<bool, false>;
    for (int it = 0; it < niter; ++it) {
        const int top = rs + it, bot = top - S + 1;
        const bool yedge = (yin0 >= bot && yin0 <= top) || (yin1 - 1 >= bot && yin1 - 1 <= top);
        if (xedge) {
            if (yedge) body(it, Yes{}, Yes{});
            else body(it, Yes{}, No{});
        } else {
            if (yedge) body(it, No{}, Yes{});
            else body(it, No{}, No{});
        }
    }
    if constexpr (NORM) {
        const double tot = tl_wave_sum(racc);
        if (lane == 0) *part = tot;
    }
}

// w = A p, sum w*p (cg.c:88-113) with the depth-1 mirror built in; with
// UPDATE, p is first replaced by beta p + r (cg_calc_p, cg.c:146-164),
// computed at every cell the stencil reads and stored to a second buffer.
// Both are the reference's expressions in its order, so p_out and w equal
// cg_calc_p, a depth-1 halo update of p and cg_calc_w bit for bit; the dot
// product differs only by summation order.  Each wave marches up ty rows of
// its own strip of 64*N columns (strips start on a multiple of N; a block
// is kWaves strips side by side): the new p of rows j-1, j, j+1 and ky of
// rows j, j+1 stay in registers, so a row costs one load each of p, r, kx
// and ky, and the new p is formed once per cell; only the wave's end lanes
// read x neighbours from memory.  beta is the argument or, in a resident
// loop, the loop's device value.
template <class T, bool UPDATE>
__global__ void __launch_bounds__(kBlock)
k_calc_pw(PitchGrid g, int cw0, int ns, int nbx, int nby, int ty, T beta_arg,
          const double* __restrict__ beta_dev, const T* __restrict__ p_in, T* __restrict__ p_out,
          const T* __restrict__ r, T* __restrict__ w, const T* __restrict__ kx,
          const T* __restrict__ ky, TlReduce rd, TlGate gate)
{
    if (gate.closed()) return;
    const T beta = UPDATE && beta_dev ? (T)*beta_dev : beta_arg;
    constexpr int N = Lane16<T>::N;
    using C = Cells<T>;
    const int lane = threadIdx.x & (TL_WAVE - 1), wave = threadIdx.x / TL_WAVE;
    const int tile = xcd_tile_of(blockIdx.x, nbx * nby);
    const int sx = (tile % nbx) * kWaves + wave, by = tile / nbx;
    const int xin0 = g.hd, xin1 = g.hd + g.nx, yin0 = g.hd, yin1 = g.hd + g.ny;
    const int c = cw0 + sx * (TL_WAVE * N) + N * lane;
    const bool ld_ok = sx < ns && c + N <= g.X;
    const bool all_out = c >= xin0 && c + N <= xin1;
    const bool any_out = sx < ns && c + N > xin0 && c < xin1;
    const long X = g.X;
    auto ld = [&](const T* __restrict__ f, long i) -> C {
        return ld_ok ? ldc(f, i) : zeros<T>();
    };
    auto pnew = [&](long i) -> C {
        C a = ld(p_in, i);
        if constexpr (UPDATE) {
            const C b = ld(r, i);
#pragma unroll
            for (int k = 0; k < N; ++k) a.v[k] = beta * a.v[k] + b.v[k];
        }
        return a;
    };
    auto pnew1 = [&](long i) -> T {
        if constexpr (UPDATE) return beta * p_in[i] + r[i];
        return p_in[i];
    };
    double acc[1] = {0.0};
    const int j0 = yin0 + by * ty;
    const int j1 = min(j0 + ty, yin1);
    if (sx < ns && j0 < j1) {                    // uniform per wave
        long i0 = (long)j0 * X + c;
        C cen = pnew(i0);
        C dn = j0 == yin0 ? cen : pnew(i0 - X);  // y neighbours outside the interior mirror
        C kyc = ld(ky, i0);
        for (int j = j0; j < j1; ++j, i0 += X) {
            const C up = j == yin1 - 1 ? cen : pnew(i0 + X);
            const C kxc = ld(kx, i0), kyu = ld(ky, i0 + X);
            T el = T(0), er = T(0), ek = T(0);
            if (lane == 0 && c >= 1) el = pnew1(i0 - 1);
            if (lane == TL_WAVE - 1 && c + N < g.X) {
                er = pnew1(i0 + N);
                ek = kx[i0 + N];
            }
            const C av = apply_a_lane(cen, up, dn, kxc, kyu, kyc, c, xin0, xin1, el, er, ek);
            if (any_out) {
                if (all_out) {
                    stc(w, i0, av);
                    if constexpr (UPDATE) stc(p_out, i0, cen);
#pragma unroll
                    for (int k = 0; k < N; ++k) acc[0] += (double)av.v[k] * (double)cen.v[k];
                } else {
#pragma unroll
                    for (int k = 0; k < N; ++k)
                        if (c + k >= xin0 && c + k < xin1) {
                            w[i0 + k] = av.v[k];
                            if constexpr (UPDATE) p_out[i0 + k] = cen.v[k];
                            acc[0] += (double)av.v[k] * (double)cen.v[k];
                        }
                }
            }
            dn = cen;
            cen = up;
            kyc = kyu;
        }
    }
    tl_reduce_partial<1, kBlock>(acc, rd);
}

// The same pass laid out as the single-step stencils, for FP32 fields (4
// cells per lane), where it measures faster than the march (0.325 vs 0.35
// ms at 8192^2): a block is kWaves rows of one strip of 64*N columns, a
// wave one row; rows j-1, j+1 and the end lanes' x neighbours are read
// directly (L2 serves the overlap).  Same values and reduction terms.
template <class T, bool UPDATE>
__global__ void __launch_bounds__(kBlock)
k_calc_pw_rows(PitchGrid g, int cw0, int nbx, int nby, int ty, T beta, const T* __restrict__ p_in,
               T* __restrict__ p_out, const T* __restrict__ r, T* __restrict__ w,
               const T* __restrict__ kx, const T* __restrict__ ky, TlReduce rd, TlGate gate)
{
    if (gate.closed()) return;
    constexpr int N = Lane16<T>::N;
    using C = Cells<T>;
    const int lane = threadIdx.x & (TL_WAVE - 1), wave = threadIdx.x / TL_WAVE;
    const int tile = xcd_tile_of(blockIdx.x, nbx * nby);
    const int sx = tile % nbx, by = tile / nbx;
    const int xin0 = g.hd, xin1 = g.hd + g.nx, yin0 = g.hd, yin1 = g.hd + g.ny;
    const int c = cw0 + sx * (TL_WAVE * N) + N * lane;
    const bool ld_ok = c + N <= g.X;
    const bool all_out = c >= xin0 && c + N <= xin1;
    const bool any_out = c + N > xin0 && c < xin1;
    const long X = g.X;
    auto ld = [&](const T* __restrict__ f, long i) -> C {
        return ld_ok ? ldc(f, i) : zeros<T>();
    };
    auto pnew = [&](long i) -> C {
        C a = ld(p_in, i);
        if constexpr (UPDATE) {
            const C b = ld(r, i);
#pragma unroll
            for (int k = 0; k < N; ++k) a.v[k] = beta * a.v[k] + b.v[k];
        }
        return a;
    };
    auto pnew1 = [&](long i) -> T {
        if constexpr (UPDATE) return beta * p_in[i] + r[i];
        return p_in[i];
    };
    double acc[1] = {0.0};
    const int jend = min(yin0 + (by + 1) * ty, yin1);
    for (int j = yin0 + by * ty + wave; j < jend; j += kWaves) {   // uniform per wave
        const long i0 = (long)j * X + c;
        const C cen = pnew(i0);
        // y neighbours outside the interior mirror the cell
        const C up = j == yin1 - 1 ? cen : pnew(i0 + X);
        const C dn = j == yin0 ? cen : pnew(i0 - X);
        const C kxc = ld(kx, i0), kyc = ld(ky, i0), kyu = ld(ky, i0 + X);
        T el = T(0), er = T(0), ek = T(0);
        if (lane == 0 && c >= 1) el = pnew1(i0 - 1);
        if (lane == TL_WAVE - 1 && c + N < g.X) {
            er = pnew1(i0 + N);
            ek = kx[i0 + N];
        }
        const C av = apply_a_lane(cen, up, dn, kxc, kyu, kyc, c, xin0, xin1, el, er, ek);
        if (any_out) {
            if (all_out) {
                stc(w, i0, av);
                if constexpr (UPDATE) stc(p_out, i0, cen);
#pragma unroll
                for (int k = 0; k < N; ++k) acc[0] += (double)av.v[k] * (double)cen.v[k];
            } else {
#pragma unroll
                for (int k = 0; k < N; ++k)
                    if (c + k >= xin0 && c + k < xin1) {
                        w[i0 + k] = av.v[k];
                        if constexpr (UPDATE) p_out[i0 + k] = cen.v[k];
                        acc[0] += (double)av.v[k] * (double)cen.v[k];
                    }
            }
        }
    }
    tl_reduce_partial<1, kBlock>(acc, rd);
}


// The CG pair of passes with kx and ky formed from the cells' conductivity
// instead of loaded: cond holds cg_init's w and every face coefficient is
// cg_init's expression on the same operands (tl_face_coef), so the stencil
// sees kx and ky bit for bit while one 8-byte field per cell replaces two.
// The first pass is k_calc_pw's (FP64): p' = beta p + r with UPDATE, w = A p'
// and sum w p'; with WOUT false w stays in registers and only its dot
// product leaves the kernel.  The second (k_calc_ur_cond) forms the same w
// again for cg_calc_ur.  cond is read at rows and columns hd-1 .. n-hd,
// which cg_init's w covers for halo depth >= 2 (cg.c:45 starts at 1).
template <bool UPDATE, bool WOUT>
__global__ void __launch_bounds__(kBlock)
k_calc_pw_cond(PitchGrid g, int cw0, int ns, int nbx, int nby, int ty, double beta_arg,
               const double* __restrict__ beta_dev, const double* __restrict__ p_in,
               double* __restrict__ p_out, const double* __restrict__ r, double* __restrict__ w,
               const double* __restrict__ cond, double rx, double ry, TlReduce rd, TlGate gate)
{
    if (gate.closed()) return;
    const double beta = UPDATE && beta_dev ? *beta_dev : beta_arg;
    constexpr int N = 2;
    using C = Cells<double>;
    const int lane = threadIdx.x & (TL_WAVE - 1), wave = threadIdx.x / TL_WAVE;
    const int tile = xcd_tile_of(blockIdx.x, nbx * nby);
    const int sx = (tile % nbx) * kWaves + wave, by = tile / nbx;
    const int xin0 = g.hd, xin1 = g.hd + g.nx, yin0 = g.hd, yin1 = g.hd + g.ny;
    const int c = cw0 + sx * (TL_WAVE * N) + N * lane;
    const bool ld_ok = sx < ns && c + N <= g.X;
    const bool all_out = c >= xin0 && c + N <= xin1;
    const bool any_out = sx < ns && c + N > xin0 && c < xin1;
    const long X = g.X;
    auto ld = [&](const double* __restrict__ f, long i) -> C {
        return ld_ok ? ldc(f, i) : zeros<double>();
    };
    auto pnew = [&](long i) -> C {
        C a = ld(p_in, i);
        if constexpr (UPDATE) {
            const C b = ld(r, i);
#pragma unroll
            for (int k = 0; k < N; ++k) a.v[k] = beta * a.v[k] + b.v[k];
        }
        return a;
    };
    auto pnew1 = [&](long i) -> double {
        if constexpr (UPDATE) return beta * p_in[i] + r[i];
        return p_in[i];
    };
    double acc[1] = {0.0};
    const int j0 = yin0 + by * ty;
    const int j1 = min(j0 + ty, yin1);
    if (sx < ns && j0 < j1) {                    // uniform per wave
        long i0 = (long)j0 * X + c;
        C cen = pnew(i0);
        C dn = j0 == yin0 ? cen : pnew(i0 - X);  // y neighbours outside the interior mirror
        C wc = ld(cond, i0);
        C kyc = ky_of_rows(ry, ld(cond, i0 - X), wc);
        for (int j = j0; j < j1; ++j, i0 += X) {
            const C up = j == yin1 - 1 ? cen : pnew(i0 + X);
            const C wu = ld(cond, i0 + X);
            const C kyu = ky_of_rows(ry, wc, wu);
            double el = 0.0, er = 0.0, ek = 0.0, wl = 0.0;
            if (lane == 0 && c >= 1) {
                el = pnew1(i0 - 1);
                wl = cond[i0 - 1];
            }
            if (lane == TL_WAVE - 1 && c + N < g.X) {
                er = pnew1(i0 + N);
                ek = tl_face_coef(rx, wc.v[N - 1], cond[i0 + N]);
            }
            const C kxc = kx_of_row(rx, wc, wl);
            const C av = apply_a_lane(cen, up, dn, kxc, kyu, kyc, c, xin0, xin1, el, er, ek);
            if (any_out) {
                if (all_out) {
                    if constexpr (WOUT) stc(w, i0, av);
                    if constexpr (UPDATE) stc(p_out, i0, cen);
#pragma unroll
                    for (int k = 0; k < N; ++k) acc[0] += av.v[k] * cen.v[k];
                } else {
#pragma unroll
                    for (int k = 0; k < N; ++k)
                        if (c + k >= xin0 && c + k < xin1) {
                            if constexpr (WOUT) w[i0 + k] = av.v[k];
                            if constexpr (UPDATE) p_out[i0 + k] = cen.v[k];
                            acc[0] += av.v[k] * cen.v[k];
                        }
                }
            }
            dn = cen;
            cen = up;
            kyc = kyu;
            wc = wu;
        }
    }
    tl_reduce_partial<1, kBlock>(acc, rd);
}

// cg_calc_ur (2d/c_kernels/omp3/cg.c:116-143) with w = A p formed here from
// p and cond as k_calc_pw_cond forms it: u += alpha p, r -= alpha w,
// sum r r.  u and r are the reference's bit for bit; the sum differs only
// by summation order.  alpha is the argument or, in a resident loop, the
// loop's device value.  Same march and tiling as k_calc_pw_cond.
__global__ void __launch_bounds__(kBlock)
k_calc_ur_cond(PitchGrid g, int cw0, int ns, int nbx, int nby, int ty, double alpha_arg,
               const double* __restrict__ alpha_dev, double* __restrict__ u,
               const double* __restrict__ p, double* __restrict__ r,
               const double* __restrict__ cond, double rx, double ry, TlReduce rd, TlGate gate)
{
    if (gate.closed()) return;
    const double alpha = alpha_dev ? *alpha_dev : alpha_arg;
    constexpr int N = 2;
    using C = Cells<double>;
    const int lane = threadIdx.x & (TL_WAVE - 1), wave = threadIdx.x / TL_WAVE;
    const int tile = xcd_tile_of(blockIdx.x, nbx * nby);
    const int sx = (tile % nbx) * kWaves + wave, by = tile / nbx;
    const int xin0 = g.hd, xin1 = g.hd + g.nx, yin0 = g.hd, yin1 = g.hd + g.ny;
    const int c = cw0 + sx * (TL_WAVE * N) + N * lane;
    const bool ld_ok = sx < ns && c + N <= g.X;
    const bool all_out = c >= xin0 && c + N <= xin1;
    const bool any_out = sx < ns && c + N > xin0 && c < xin1;
    const long X = g.X;
    auto ld = [&](const double* __restrict__ f, long i) -> C {
        return ld_ok ? ldc(f, i) : zeros<double>();
    };
    double acc[1] = {0.0};
    const int j0 = yin0 + by * ty;
    const int j1 = min(j0 + ty, yin1);
    if (sx < ns && j0 < j1) {                    // uniform per wave
        long i0 = (long)j0 * X + c;
        C cen = ld(p, i0);
        C dn = j0 == yin0 ? cen : ld(p, i0 - X);
        C wc = ld(cond, i0);
        C kyc = ky_of_rows(ry, ld(cond, i0 - X), wc);
        for (int j = j0; j < j1; ++j, i0 += X) {
            const C up = j == yin1 - 1 ? cen : ld(p, i0 + X);
            const C wu = ld(cond, i0 + X);
            const C kyu = ky_of_rows(ry, wc, wu);
            const C uu = ld(u, i0), rr = ld(r, i0);
            double el = 0.0, er = 0.0, ek = 0.0, wl = 0.0;
            if (lane == 0 && c >= 1) {
                el = p[i0 - 1];
                wl = cond[i0 - 1];
            }
            if (lane == TL_WAVE - 1 && c + N < g.X) {
                er = p[i0 + N];
                ek = tl_face_coef(rx, wc.v[N - 1], cond[i0 + N]);
            }
            const C kxc = kx_of_row(rx, wc, wl);
            const C av = apply_a_lane(cen, up, dn, kxc, kyu, kyc, c, xin0, xin1, el, er, ek);
            C un, rn;
#pragma unroll
            for (int k = 0; k < N; ++k) {
                un.v[k] = uu.v[k] + alpha * cen.v[k];
                rn.v[k] = rr.v[k] - alpha * av.v[k];
            }
            if (any_out) {
                if (all_out) {
                    stc(u, i0, un);
                    stc(r, i0, rn);
#pragma unroll
                    for (int k = 0; k < N; ++k) acc[0] += rn.v[k] * rn.v[k];
                } else {
#pragma unroll
                    for (int k = 0; k < N; ++k)
                        if (c + k >= xin0 && c + k < xin1) {
                            u[i0 + k] = un.v[k];
                            r[i0 + k] = rn.v[k];
                            acc[0] += rn.v[k] * rn.v[k];
                        }
                }
            }
            dn = cen;
            cen = up;
            kyc = kyu;
            wc = wu;
        }
    }
    tl_reduce_partial<1, kBlock>(acc, rd);
}

// cond = cg_init's w (cg.c:45-53) over the whole grid: density, or
// 1/density with the reciprocal coefficient (coefficient 2).
__global__ void __launch_bounds__(kBlock)
k_cond2d(long n, int coef, const double* __restrict__ density, double* __restrict__ cond)
{
    for (long c = blockIdx.x * (long)kBlock + threadIdx.x; c < n; c += (long)gridDim.x * kBlock)
        cond[c] = coef == 1 ? density[c] : 1.0 / density[c];
}


constexpr int kPwRows = 32;   // rows per wave of k_calc_pw (measured best at 4096^2)

// Everything outside the interior of up to three fields, src -> dst.
template <class T>
struct Ring3 {
    const T* src[3];
    T* dst[3];
    int n;
};

template <class T>
__global__ void __launch_bounds__(kBlock)
k_copy_rings2d(PitchGrid g, int rows, Ring3<T> rg, TlGate gate)
{
    if (gate.closed()) return;
    // rows [0,hd) and [hd+ny,rows) whole; rows [hd,hd+ny) outside [hd,hd+nx)
    const int side = g.X - g.nx;
    const long nrow = (long)(rows - g.ny) * g.X;
    const long ncol = (long)side * g.ny;
    const long per = nrow + ncol;
    for (long t = blockIdx.x * (long)kBlock + threadIdx.x; t < per * rg.n;
         t += (long)gridDim.x * kBlock) {
        const int f = (int)(t / per);
        const long q = t % per;
        long cell;
        if (q < nrow) {
            const int rr = (int)(q / g.X), kk = (int)(q % g.X);
            const int jj = rr < g.hd ? rr : g.ny + rr;
            cell = (long)jj * g.X + kk;
        } else {
            const long v = q - nrow;
            const int jj = g.hd + (int)(v / side);
            const int p = (int)(v % side);
            const int kk = p < g.hd ? p : g.nx + p;
            cell = (long)jj * g.X + kk;
        }
        rg.dst[f][cell] = rg.src[f][cell];
    }
}

// Default tile height: as many row tiles as fill one round of the blocks
// the device holds at once.  A wave marches its strip down the tile, and
// the S-1 rows above and below it are recomputed by the neighbouring
// tile, so taller tiles waste less; but a block count beyond what is
// resident at once leaves a second, mostly idle round.  On MI355X
// (256 CUs) a sweep at 4096^2 put the best block count at ~0.88 of the
// resident capacity for every S (S <= 5: 2 blocks per CU; S >= 6, over
// 256 VGPRs: 1), e.g. S = 10 at 225 blocks (ty 164) ran 391 us and at
// 234 blocks (ty 158) 657 us.
template <class T, int S, int MODE, int PF>
int auto_tile_rows(const PitchGrid& g, int nbx)
{
    static int occ = 0;
    if (occ <= 0 &&
        (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_ppcg_fused2d<T, S, MODE, PF>, kBlock,
                                                      0) != hipSuccess ||
         occ <= 0))
        occ = 1;
    const long cap = (long)(0.88 * device_cus() * occ);
    const long nby = std::max(1L, cap / std::max(1, nbx));
    return (int)std::max(1L, ((long)g.ny + nby - 1) / nby);
}

template <class T, int S, int MODE>
int launch_fused(const FusedArgs<T>& a)
{
    // Windows reach S columns beyond their output columns on both sides.
    const int pf = g_fused_pf < 0 ? 1 : std::min(g_fused_pf, 3);
    StripTile t = strip_tile<T>(a.g, S, 1, kWaves);
    int ty = g_fused_ty;
    if (ty <= 0) {
        switch (pf) {
        case 0: ty = auto_tile_rows<T, S, MODE, 0>(a.g, t.nbx); break;
        case 1: ty = auto_tile_rows<T, S, MODE, 1>(a.g, t.nbx); break;
        case 2: ty = auto_tile_rows<T, S, MODE, 2>(a.g, t.nbx); break;
        default: ty = auto_tile_rows<T, S, MODE, 3>(a.g, t.nbx); break;
        }
    }
    t = strip_tile<T>(a.g, S, ty, kWaves);
    if ((long)t.nbx * t.nby > (1L << 30))
        return tl_fail(TLHIP_ERR_ARG, "%s: grid too large", a.what);
    // The rows one tile reads (ty + 2S - 2 marched, S below, PF + 1 above)
    // must lie within one buffer descriptor's range.
    if ((long)(t.ty + 3 * S + pf + 2) * a.g.X * (long)sizeof(T) > (long)kBufferRange)
        return tl_fail(TLHIP_ERR_ARG, "%s: %d-row tiles of %d-element rows exceed a buffer range",
                       a.what, t.ty, a.g.X);
    PpcgCoefs<T, S> cf;
    for (int k = 0; k < S; ++k) {
        cf.alpha[k] = (T)a.al[k];
        cf.beta[k] = (T)a.be[k];
    }
    const unsigned nb = (unsigned)(t.nbx * t.nby);
    const int slot = tl_timer_begin(a.what);
#define TL_FUSED_LAUNCH(PFD)                                                                   \
    hipLaunchKernelGGL((k_ppcg_fused2d<T, S, MODE, PFD>), dim3(nb), dim3(kBlock), 0, tl_stream(), \
                       a.g, t, cf, (T)a.theta, (T)a.alpha_ur, a.alpha_dev, a.sd_in, a.sd_out,     \
                       a.r_in, a.r_out, a.u_in, a.u_out, a.kx, a.ky, a.p, a.w, tl_gate())
    switch (pf) {
    case 0: TL_FUSED_LAUNCH(0); break;
    case 1: TL_FUSED_LAUNCH(1); break;
    case 2: TL_FUSED_LAUNCH(2); break;
    default: TL_FUSED_LAUNCH(3); break;
    }
#undef TL_FUSED_LAUNCH
    tl_timer_end(slot);
    return tl_check_launch(a.what);
}

template <class T, int MODE>
int dispatch_fused(int count, const FusedArgs<T>& a)
{
    // FP32 always takes the ring kernel (its FP32 arithmetic uses fused
    // multiply-adds; the register-window kernel's does not)
    if (g_fused_variant == 0 || std::is_same_v<T, float>) return launch_ppcg_ring<T>(count, MODE, a);
    switch (count) {
    case 1: return launch_fused<T, 1, MODE>(a);
    case 2: return launch_fused<T, 2, MODE>(a);
    case 3: return launch_fused<T, 3, MODE>(a);
    case 4: return launch_fused<T, 4, MODE>(a);
    case 5: return launch_fused<T, 5, MODE>(a);
    case 6: return launch_fused<T, 6, MODE>(a);
    case 7: return launch_fused<T, 7, MODE>(a);
    case 8: return launch_fused<T, 8, MODE>(a);
    case 9: return launch_fused<T, 9, MODE>(a);
    case 10: return launch_fused<T, 10, MODE>(a);
    default: return tl_fail(TLHIP_ERR_ARG, "%s: count must be 1..%d", a.what, kMaxFused);
    }
}

template <class T>
int inner_iterations(const char* what, int x, int y, int hd, int pitch, int count,
                     const double* alphas, const double* betas, double theta, int init_from_r,
                     const int* neighbours, const T* sd_in, T* sd_out, const T* r_in, T* r_out,
                     const T* u_in, T* u_out, const T* kx, const T* ky, double alpha_ur,
                     const T* p, const T* w, const double* alpha_dev = nullptr,
                     double* rrn = nullptr, bool copy_ring = true, unsigned* norm_nb = nullptr)
{
    constexpr int N = Lane16<T>::N;
    TL_TRY(tl_ensure_init());
    if ((rrn || norm_nb) && std::is_same_v<T, double> && g_fused_variant != 0)
        return tl_fail(TLHIP_ERR_ARG, "%s: the fused 2-norm needs the ring kernel", what);
    if (hd < 1 || x < 2 * hd + 1 || y < 2 * hd + 1 || pitch < x || pitch % N ||
        (long)pitch * y > (1L << 31) - 1)
        return tl_fail(TLHIP_ERR_ARG,
                       "%s: needs a row pitch that is a multiple of %d and >= x (x=%d pitch=%d "
                       "y=%d hd=%d)", what, N, x, pitch, y, hd);
    if (count < 1 || count > kMaxFused)
        return tl_fail(TLHIP_ERR_ARG, "%s: count must be 1..%d (got %d)", what, kMaxFused, count);
    TL_REQUIRE(alphas && betas && neighbours, what, "null coefficients/neighbours");
    TL_REQUIRE((init_from_r || sd_in) && sd_out && r_in && r_out && u_in && u_out && kx && ky, what,
               "null field");
    TL_REQUIRE(sd_in != sd_out && r_in != r_out && u_in != u_out, what,
               "input and output buffers must be distinct");
    TL_REQUIRE((p == nullptr) == (w == nullptr), what, "p and w go together");
    TL_REQUIRE(!p || init_from_r, what, "the cg_calc_ur update needs init_from_r");
    for (int f = 0; f < 4; ++f)
        TL_REQUIRE(neighbours[f] == -1, what, "every face must be external (single chunk)");
    TL_REQUIRE(all_aligned(sd_in, sd_out, r_in, r_out, u_in, u_out, kx, ky, p, w), what,
               "fields must be 16-byte aligned");
    if (init_from_r && theta == 0.0) return tl_fail(TLHIP_ERR_ARG, "%s: theta is zero", what);
    const PitchGrid g{pitch, hd, x - 2 * hd, y - 2 * hd};
    Ring3<T> rg{};
    rg.src[0] = r_in;  rg.dst[0] = r_out;
    rg.src[1] = u_in;  rg.dst[1] = u_out;
    rg.n = 2;
    if (sd_in) {
        rg.src[2] = sd_in;
        rg.dst[2] = sd_out;
        rg.n = 3;
    }
    const long ring = ((long)(y - g.ny) * pitch + (long)(pitch - g.nx) * g.ny) * rg.n;
    if (copy_ring) {
        hipLaunchKernelGGL(k_copy_rings2d<T>, dim3(flat_blocks(ring)), dim3(kBlock), 0,
                           tl_stream(), g, y, rg, tl_gate());
        TL_TRY(tl_check_launch(what));
    }
    const FusedArgs<T> a{g, alphas, betas, theta, alpha_ur, alpha_dev, sd_in, sd_out, r_in, r_out,
                         u_in, u_out, kx, ky, p, w, what, rrn, norm_nb};
    if (!init_from_r) return dispatch_fused<T, 0>(count, a);
    if (!p) return dispatch_fused<T, 1>(count, a);
    return dispatch_fused<T, 2>(count, a);
}

template <class T, int S, bool NORM>
int launch_cheby(const PitchGrid& g, const double* al, const double* be, const T* u_in,
                 T* u_out, const T* u0, const T* p_in, T* p_out, T* r, const T* kx, const T* ky,
                 double* rrn, const char* what)
{
    const int pf = g_fused_pf < 0 ? 1 : std::min(g_fused_pf, 1);
    StripTile t = strip_tile<T>(g, S, 1, kWaves);
    static int occ = 0;
    if (occ <= 0 &&
        (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_cheby_fused2d<T, S, 1, NORM>,
                                                      kBlock, 0) != hipSuccess ||
         occ <= 0))
        occ = 1;
    int ty = g_fused_ty;
    if (ty <= 0) {
        const long cap = (long)(0.88 * device_cus() * occ);
        const long nby = std::max(1L, cap / std::max(1, t.nbx));
        ty = (int)std::max(1L, ((long)g.ny + nby - 1) / nby);
    }
    t = strip_tile<T>(g, S, ty, kWaves);
    if ((long)t.nbx * t.nby > (NORM ? (long)TL_MAX_RED_BLOCKS / kWaves : (1L << 30)))
        return tl_fail(TLHIP_ERR_ARG, "%s: grid too large", what);
    if ((long)(t.ty + 3 * S + pf + 2) * g.X * (long)sizeof(T) > (long)kBufferRange)
        return tl_fail(TLHIP_ERR_ARG, "%s: %d-row tiles of %d-element rows exceed a buffer range",
                       what, t.ty, g.X);
    PpcgCoefs<T, S> cf;
    for (int k = 0; k < S; ++k) {
        cf.alpha[k] = (T)al[k];
        cf.beta[k] = (T)be[k];
    }
    const unsigned nb = (unsigned)(t.nbx * t.nby);
    const int slot = tl_timer_begin(what);
    if (pf)
        hipLaunchKernelGGL((k_cheby_fused2d<T, S, 1, NORM>), dim3(nb), dim3(kBlock), 0,
                           tl_stream(), g, t, cf, u_in, u_out, u0, p_in, p_out, r, kx, ky,
                           tl_reduce_slots());
    else
        hipLaunchKernelGGL((k_cheby_fused2d<T, S, 0, NORM>), dim3(nb), dim3(kBlock), 0,
                           tl_stream(), g, t, cf, u_in, u_out, u0, p_in, p_out, r, kx, ky,
                           tl_reduce_slots());
    tl_timer_end(slot);
    TL_TRY(tl_check_launch(what));
    if constexpr (NORM) {
        TL_TRY(tl_finish_reduce(1, nb * kWaves, what));
        return tl_fetch(1, rrn, what);
    }
    return 0;
}

// The fused Chebyshev loop: FP64 on x-pitched fields, or FP32 on fields of
// row pitch `pitch`; rrn (FP32 only) also forms r.r; copy_ring: carry the
// halo rings of u and p over to the outputs.
template <class T>
int cheby_iterations(const char* what, int x, int y, int hd, int pitch, int count,
                     const double* alphas, const double* betas, const int* neighbours,
                     const T* u_in, T* u_out, const T* u0, const T* p_in, T* p_out, T* r,
                     const T* kx, const T* ky, double* rrn = nullptr, bool copy_ring = true)
{
    constexpr int N = Lane16<T>::N;
    TL_TRY(tl_ensure_init());
    if (hd < 1 || x < 2 * hd + 1 || y < 2 * hd + 1 || pitch < x || pitch % N ||
        (long)pitch * y > (1L << 31) - 1)
        return tl_fail(TLHIP_ERR_ARG,
                       "%s: needs a row pitch that is a multiple of %d and >= x (x=%d pitch=%d "
                       "y=%d hd=%d)", what, N, x, pitch, y, hd);
    if (count < 1 || count > kMaxFused)
        return tl_fail(TLHIP_ERR_ARG, "%s: count must be 1..%d (got %d)", what, kMaxFused, count);
    if (rrn && !std::is_same_v<T, float>)
        return tl_fail(TLHIP_ERR_ARG, "%s: the fused 2-norm is FP32 only", what);
    TL_REQUIRE(alphas && betas && neighbours, what, "null coefficients/neighbours");
    TL_REQUIRE(u_in && u_out && u0 && p_in && p_out && r && kx && ky, what, "null field");
    TL_REQUIRE(u_in != u_out && p_in != p_out, what, "input and output buffers must be distinct");
    TL_REQUIRE(r != u_in && r != u_out && r != p_in && r != p_out && r != u0, what,
               "r must be distinct from the other fields");
    for (int f = 0; f < 4; ++f)
        TL_REQUIRE(neighbours[f] == -1, what, "every face must be external (single chunk)");
    TL_REQUIRE(all_aligned(u_in, u_out, u0, p_in, p_out, r, kx, ky), what,
               "fields must be 16-byte aligned");
    const PitchGrid g{pitch, hd, x - 2 * hd, y - 2 * hd};
    if (copy_ring) {
        Ring3<T> rg{};
        rg.src[0] = u_in;  rg.dst[0] = u_out;
        rg.src[1] = p_in;  rg.dst[1] = p_out;
        rg.n = 2;
        const long ring = ((long)(y - g.ny) * pitch + (long)(pitch - g.nx) * g.ny) * rg.n;
        hipLaunchKernelGGL(k_copy_rings2d<T>, dim3(flat_blocks(ring)), dim3(kBlock), 0,
                           tl_stream(), g, y, rg, tl_gate());
        TL_TRY(tl_check_launch(what));
    }
    if (g_fused_variant == 0 || std::is_same_v<T, float>) {
        // the LDS-ring kernel, Chebyshev mode: u in sd's role, u0 in r's, p in u's
        FusedArgs<T> a{g, alphas, betas, 1.0, 0.0, nullptr, u_in, u_out, u0, r, p_in, p_out,
                       kx, ky, nullptr, nullptr, what, rrn};
        return launch_ppcg_ring<T>(count, 3, a);
    }
    const bool norm = rrn != nullptr;
    switch (count) {
#define TL_CHEBY_CASE(S)                                                                      \
    case S:                                                                                   \
        if constexpr (std::is_same_v<T, float>) {                                             \
            if (norm)                                                                         \
                return launch_cheby<T, S, true>(g, alphas, betas, u_in, u_out, u0, p_in,      \
                                                p_out, r, kx, ky, rrn, what);                 \
        }                                                                                     \
        return launch_cheby<T, S, false>(g, alphas, betas, u_in, u_out, u0, p_in, p_out, r,   \
                                         kx, ky, nullptr, what);
    TL_CHEBY_CASE(1) TL_CHEBY_CASE(2) TL_CHEBY_CASE(3) TL_CHEBY_CASE(4) TL_CHEBY_CASE(5)
    TL_CHEBY_CASE(6) TL_CHEBY_CASE(7) TL_CHEBY_CASE(8) TL_CHEBY_CASE(9) TL_CHEBY_CASE(10)
#undef TL_CHEBY_CASE
    default: return tl_fail(TLHIP_ERR_ARG, "%s: count must be 1..%d", what, kMaxFused);
    }
}


// Checks, the ring copy p_in -> p_out (the halo the reference's p keeps;
// skipped with ring = false) and the k_calc_pw launch under the current
// gate; *nb: its partials.
template <class T>
int launch_pw(const char* what, int x, int y, int hd, int pitch, bool update, double beta,
              const double* beta_dev, const T* p_in, T* p_out, const T* r, T* w, const T* kx,
              const T* ky, unsigned* nb, bool ring = true)
{
    constexpr int N = Lane16<T>::N;
    TL_TRY(tl_ensure_init());
    if (hd < 1 || x < 2 * hd + 1 || y < 2 * hd + 1 || pitch < x || pitch % N ||
        (long)pitch * y > (1L << 31) - 1)
        return tl_fail(TLHIP_ERR_ARG,
                       "%s: needs a row pitch that is a multiple of %d and >= x (x=%d pitch=%d "
                       "y=%d hd=%d)", what, N, x, pitch, y, hd);
    TL_REQUIRE(p_in && w && kx && ky && (!update || (p_out && r)), what, "null argument");
    TL_REQUIRE(w != p_in && (!update || (p_out != p_in && p_out != w && p_out != r)), what,
               "p_out and w must be distinct from the inputs");
    TL_REQUIRE(all_aligned(p_in, p_out, r, w, kx, ky), what, "fields must be 16-byte aligned");
    const PitchGrid g{pitch, hd, x - 2 * hd, y - 2 * hd};
    const int cw0 = hd / N * N;
    constexpr bool kMarch = std::is_same<T, double>::value;
    const int ns = (int)tl_cdiv(hd + g.nx - cw0, TL_WAVE * N);
    const int nbx = kMarch ? (int)tl_cdiv(ns, kWaves) : ns;
    int ty = tl_tile_rows_override() > 0 ? tl_tile_rows_override() : (kMarch ? kPwRows : kWaves);
    if (!kMarch) ty = (int)tl_cdiv(ty, kWaves) * kWaves;
    while ((long)nbx * tl_cdiv(g.ny, ty) > TL_MAX_RED_BLOCKS) ty *= 2;
    const int nby = (int)tl_cdiv(g.ny, ty);
    *nb = (unsigned)(nbx * nby);
    if (update && ring) {
        Ring3<T> rg{};
        rg.src[0] = p_in;
        rg.dst[0] = p_out;
        rg.n = 1;
        const long ring = (long)(y - g.ny) * pitch + (long)(pitch - g.nx) * g.ny;
        hipLaunchKernelGGL(k_copy_rings2d<T>, dim3(flat_blocks(ring)), dim3(kBlock), 0,
                           tl_stream(), g, y, rg, tl_gate());
        TL_TRY(tl_check_launch(what));
    }
    const int slot = tl_timer_begin(what);
    if constexpr (kMarch) {
        if (update)
            hipLaunchKernelGGL((k_calc_pw<T, true>), dim3(*nb), dim3(kBlock), 0, tl_stream(), g,
                               cw0, ns, nbx, nby, ty, (T)beta, beta_dev, p_in, p_out, r, w, kx,
                               ky, tl_reduce_slots(), tl_gate());
        else
            hipLaunchKernelGGL((k_calc_pw<T, false>), dim3(*nb), dim3(kBlock), 0, tl_stream(), g,
                               cw0, ns, nbx, nby, ty, (T)beta, beta_dev, p_in, p_out, r, w, kx,
                               ky, tl_reduce_slots(), tl_gate());
    } else {
        if (beta_dev) return tl_fail(TLHIP_ERR_ARG, "%s: FP32 takes beta from the host", what);
        if (update)
            hipLaunchKernelGGL((k_calc_pw_rows<T, true>), dim3(*nb), dim3(kBlock), 0, tl_stream(),
                               g, cw0, nbx, nby, ty, (T)beta, p_in, p_out, r, w, kx, ky,
                               tl_reduce_slots(), tl_gate());
        else
            hipLaunchKernelGGL((k_calc_pw_rows<T, false>), dim3(*nb), dim3(kBlock), 0,
                               tl_stream(), g, cw0, nbx, nby, ty, (T)beta, p_in, p_out, r, w, kx,
                               ky, tl_reduce_slots(), tl_gate());
    }
    tl_timer_end(slot);
    return tl_check_launch(what);
}

// Tiling of the FP64 march kernels (k_calc_pw's, shared by the cond pair).
struct March {
    PitchGrid g;
    int cw0, ns, nbx, nby, ty;
    unsigned nb() const { return (unsigned)(nbx * nby); }
};

March march_geom(int x, int y, int hd)
{
    March m;
    m.g = PitchGrid{x, hd, x - 2 * hd, y - 2 * hd};
    m.cw0 = hd / 2 * 2;
    m.ns = (int)tl_cdiv(hd + m.g.nx - m.cw0, TL_WAVE * 2);
    m.nbx = (int)tl_cdiv(m.ns, kWaves);
    m.ty = tl_tile_rows_override() > 0 ? tl_tile_rows_override() : kPwRows;
    while ((long)m.nbx * tl_cdiv(m.g.ny, m.ty) > TL_MAX_RED_BLOCKS) m.ty *= 2;
    m.nby = (int)tl_cdiv(m.g.ny, m.ty);
    return m;
}

int check_cond_grid(const char* what, int x, int y, int hd)
{
    TL_TRY(tl_ensure_init());
    if (hd < 2 || x % 2 || x < 2 * hd + 1 || y < 2 * hd + 1 || (long)x * y > (1L << 31) - 1)
        return tl_fail(TLHIP_ERR_ARG,
                       "%s: needs halo depth >= 2 and an even x (x=%d y=%d hd=%d)", what, x, y,
                       hd);
    return 0;
}

// Checks, the ring copy p_in -> p_out (with update and ring) and the
// k_calc_pw_cond launch (w written unless null) under the current gate;
// *nb: its partials.
int launch_pw_cond(const char* what, int x, int y, int hd, bool update, double beta,
                   const double* beta_dev, const double* p_in, double* p_out, const double* r,
                   double* w, const double* cond, double rx, double ry, unsigned* nb, bool ring)
{
    TL_TRY(check_cond_grid(what, x, y, hd));
    TL_REQUIRE(p_in && cond && (!update || (p_out && r)), what, "null argument");
    TL_REQUIRE(w != p_in && w != cond && (!update || (p_out != p_in && p_out != w &&
                                                      p_out != r && p_out != cond)),
               what, "p_out and w must be distinct from the inputs");
    TL_REQUIRE(all_aligned(p_in, p_out, r, w, cond), what, "fields must be 16-byte aligned");
    const March m = march_geom(x, y, hd);
    *nb = m.nb();
    if (update && ring) {
        Ring3<double> rg{};
        rg.src[0] = p_in;
        rg.dst[0] = p_out;
        rg.n = 1;
        const long nring = (long)(y - m.g.ny) * x + (long)(x - m.g.nx) * m.g.ny;
        hipLaunchKernelGGL(k_copy_rings2d<double>, dim3(flat_blocks(nring)), dim3(kBlock), 0,
                           tl_stream(), m.g, y, rg, tl_gate());
        TL_TRY(tl_check_launch(what));
    }
    // kernel-timer label: the entry point's name when w is stored, with
    // ".no_w" appended when w stays in registers (another byte count)
    const int slot = tl_timer_begin(w ? "tlhip2d_cg_calc_pw_cond" : "tlhip2d_cg_calc_pw_cond.no_w");
#define TL_PWC(U, W)                                                                          \
    hipLaunchKernelGGL((k_calc_pw_cond<U, W>), dim3(*nb), dim3(kBlock), 0, tl_stream(), m.g,    \
                       m.cw0, m.ns, m.nbx, m.nby, m.ty, beta, beta_dev, p_in, p_out, r, w,      \
                       cond, rx, ry, tl_reduce_slots(), tl_gate())
    if (update) {
        if (w) TL_PWC(true, true);
        else TL_PWC(true, false);
    } else {
        if (w) TL_PWC(false, true);
        else TL_PWC(false, false);
    }
#undef TL_PWC
    tl_timer_end(slot);
    return tl_check_launch(what);
}

// Checks and the k_calc_ur_cond launch under the current gate; *nb: its
// partials.
int launch_ur_cond(const char* what, int x, int y, int hd, double alpha,
                   const double* alpha_dev, double* u, const double* p, double* r,
                   const double* cond, double rx, double ry, unsigned* nb)
{
    TL_TRY(check_cond_grid(what, x, y, hd));
    TL_REQUIRE(u && p && r && cond, what, "null argument");
    TL_REQUIRE(u != p && u != r && r != p && u != cond && r != cond, what,
               "u, p, r and cond must be distinct buffers");
    TL_REQUIRE(all_aligned(u, p, r, cond), what, "fields must be 16-byte aligned");
    const March m = march_geom(x, y, hd);
    *nb = m.nb();
    const int slot = tl_timer_begin(what);
    hipLaunchKernelGGL(k_calc_ur_cond, dim3(*nb), dim3(kBlock), 0, tl_stream(), m.g, m.cw0, m.ns,
                       m.nbx, m.nby, m.ty, alpha, alpha_dev, u, p, r, cond, rx, ry,
                       tl_reduce_slots(), tl_gate());
    tl_timer_end(slot);
    return tl_check_launch(what);
}

// The entry points' form: pw finished and fetched to the host (+=).
template <class T>
int calc_pw(const char* what, int x, int y, int hd, int pitch, bool update, double beta,
            double* pw, const T* p_in, T* p_out, const T* r, T* w, const T* kx, const T* ky)
{
    TL_REQUIRE(pw, what, "null argument");
    unsigned nb = 0;
    TL_TRY(launch_pw<T>(what, x, y, hd, pitch, update, beta, nullptr, p_in, p_out, r, w, kx, ky,
                        &nb));
    TL_TRY(tl_finish_reduce(1, nb, what));
    return fetch_add(pw, what);
}

}  // namespace

namespace {

// alpha = rro / pw, pw the grid total of the partials (the finisher's
// order), for a launch that takes alpha from the device.
__global__ void __launch_bounds__(256)
k_mix_alpha(unsigned nblocks, const double* __restrict__ partials, double rro,
            double* __restrict__ alpha, TlGate gate)
{
    if (gate.closed()) return;
    double pw[1];
    tl_sum_partials<1>(nblocks, partials, pw);
    if (threadIdx.x == 0) alpha[0] = rro / pw[0];
}

}  // namespace

namespace tl {
namespace loop2d {

// The fused inner loop of a resident PPCG iteration: gated, and the first
// launch (init_from_r with p, w) takes cg_calc_ur's alpha from the loop.
int ppcg_inner(int x, int y, int hd, int count, const double* alphas, const double* betas,
               double theta, int init_from_r, const int* neighbours, const double* sd_in,
               double* sd_out, const double* r_in, double* r_out, const double* u_in,
               double* u_out, const double* kx, const double* ky, const double* p,
               const double* w, unsigned* norm_nb)
{
    return inner_iterations<double>("tlhip2d_ppcg_inner_iterations", x, y, hd, x, count, alphas,
                                    betas, theta, init_from_r, neighbours, sd_in, sd_out, r_in,
                                    r_out, u_in, u_out, kx, ky, 0.0, p, w,
                                    p ? loop::alpha() : nullptr, nullptr, true, norm_nb);
}

bool ring_kernel() { return g_fused_variant == 0; }

// The same launch on the host-driven path, with the 2-norm of r fetched to
// *rrn when rrn is not null.
int ppcg_inner_norm(int x, int y, int hd, int count, const double* alphas, const double* betas,
                    double theta, int init_from_r, const int* neighbours, const double* sd_in,
                    double* sd_out, const double* r_in, double* r_out, const double* u_in,
                    double* u_out, const double* kx, const double* ky, double alpha,
                    const double* p, const double* w, double* rrn)
{
    return inner_iterations<double>("tlhip2d_ppcg_inner_iterations", x, y, hd, x, count, alphas,
                                    betas, theta, init_from_r, neighbours, sd_in, sd_out, r_in,
                                    r_out, u_in, u_out, kx, ky, alpha, p, w, nullptr, rrn);
}

// calc_p with the loop's beta fused into calc_w, pw finished into alpha.
// p_out's halo is left as it is: in the resident loop nothing reads p's
// halo until the loop's closing halo update refreshes it (the fused passes
// mirror, the deeper rings of p and its partner are the same zeros).
int calc_pw(int x, int y, int hd, const double* p_in, double* p_out, const double* r, double* w,
            const double* kx, const double* ky)
{
    const char* what = "tlhip2d_cg_calc_pw";
    if (x % 2) return tl_fail(TLHIP_ERR_ARG, "%s: needs an even x (got %d)", what, x);
    unsigned nb = 0;
    TL_TRY(launch_pw<double>(what, x, y, hd, x, true, 0.0, loop::beta(), p_in, p_out, r, w, kx,
                             ky, &nb, false));
    return loop::finish_alpha(nb, what);
}

// The cond pair in a resident loop: calc_pw_cond (beta from the loop with
// update; w written unless null) with pw finished into alpha, and
// calc_ur_cond (alpha from the loop) with rrn finished into beta and the
// stop rule.  p_out's halo is left as calc_pw leaves it.
int calc_pw_cond(int x, int y, int hd, bool update, const double* p_in, double* p_out,
                 const double* r, double* w, const double* cond, double rx, double ry)
{
    const char* what = "tlhip2d_cg_calc_pw_cond";
    unsigned nb = 0;
    TL_TRY(launch_pw_cond(what, x, y, hd, update, 0.0, update ? loop::beta() : nullptr, p_in,
                          p_out, r, w, cond, rx, ry, &nb, false));
    return loop::finish_alpha(nb, what);
}

int calc_ur_cond(int x, int y, int hd, double* u, const double* p, double* r, const double* cond,
                 double rx, double ry, const loop::StopRule& rule)
{
    const char* what = "tlhip2d_cg_calc_ur_cond";
    unsigned nb = 0;
    TL_TRY(launch_ur_cond(what, x, y, hd, 0.0, loop::alpha(), u, p, r, cond, rx, ry, &nb));
    return loop::finish_beta(nb, rule, what);
}

}  // namespace loop2d

namespace mix {

int calc_w32(int x, int y, int hd, int P, double* pw, const float* p, float* w, const float* kx,
             const float* ky)
{
    *pw = 0.0;
    return calc_pw<float>("tlhip2d_mix_calc_w32", x, y, hd, P, false, 0.0, pw, p, nullptr,
                          nullptr, w, kx, ky);
}

int calc_pw32(int x, int y, int hd, int P, double beta, double* pw, const float* p_in,
              float* p_out, const float* r, float* w, const float* kx, const float* ky)
{
    *pw = 0.0;
    return calc_pw<float>("tlhip2d_mix_calc_pw32", x, y, hd, P, true, beta, pw, p_in, p_out, r,
                          w, kx, ky);
}

int calc_pw32_alpha(int x, int y, int hd, int P, bool update, double beta, double rro,
                    double* alpha_dev, const float* p_in, float* p_out, const float* r, float* w,
                    const float* kx, const float* ky)
{
    const char* what = update ? "tlhip2d_mix_calc_pw32" : "tlhip2d_mix_calc_w32";
    TL_REQUIRE(alpha_dev, what, "null alpha");
    unsigned nb = 0;
    TL_TRY(launch_pw<float>(what, x, y, hd, P, update, beta, nullptr, p_in, p_out, r, w, kx, ky,
                            &nb, false));
    hipLaunchKernelGGL(k_mix_alpha, dim3(1), dim3(256), 0, tl_stream(), nb,
                       (const double*)tl_reduce_slots().partials, rro, alpha_dev, tl_gate());
    return tl_check_launch(what);
}

}  // namespace mix
}  // namespace tl

extern "C" int tlhip2d_cg_calc_pw(int x, int y, int hd, double beta, double* pw,
                                  const double* p_in, double* p_out, const double* r, double* w,
                                  const double* kx, const double* ky)
{
    if (x % 2)
        return tl_fail(TLHIP_ERR_ARG, "tlhip2d_cg_calc_pw: needs an even x (got %d)", x);
    return calc_pw<double>("tlhip2d_cg_calc_pw", x, y, hd, x, true, beta, pw, p_in, p_out, r, w,
                           kx, ky);
}

extern "C" int tlhip_set_fused_tuning(int tile_rows, int prefetch)
{
    g_fused_ty = tile_rows;
    g_fused_pf = prefetch;
    return 0;
}

extern "C" int tlhip_set_fused_variant(int variant)
{
    if (variant < 0 || variant > 1)
        return tl_fail(TLHIP_ERR_ARG, "tlhip_set_fused_variant: 0 or 1 (got %d)", variant);
    g_fused_variant = variant;
    return 0;
}

extern "C" int tlhip2d_ppcg_inner_iterations(int x, int y, int hd, int count,
                                             const double* alphas, const double* betas,
                                             double theta, int init_from_r,
                                             const int* neighbours, const double* sd_in,
                                             double* sd_out, const double* r_in, double* r_out,
                                             const double* u_in, double* u_out,
                                             const double* kx, const double* ky, double alpha,
                                             const double* p, const double* w)
{
    return inner_iterations<double>("tlhip2d_ppcg_inner_iterations", x, y, hd, x, count, alphas,
                                    betas, theta, init_from_r, neighbours, sd_in, sd_out, r_in,
                                    r_out, u_in, u_out, kx, ky, alpha, p, w);
}

extern "C" int tlhip2d_ppcg_inner_iterations_f32(int x, int y, int hd, int pitch, int count,
                                                 const double* alphas, const double* betas,
                                                 double theta, int init_from_r,
                                                 const int* neighbours, const float* sd_in,
                                                 float* sd_out, const float* r_in, float* r_out,
                                                 const float* u_in, float* u_out,
                                                 const float* kx, const float* ky, double alpha,
                                                 const float* p, const float* w)
{
    return inner_iterations<float>("tlhip2d_ppcg_inner_iterations_f32", x, y, hd, pitch, count,
                                   alphas, betas, theta, init_from_r, neighbours, sd_in, sd_out,
                                   r_in, r_out, u_in, u_out, kx, ky, alpha, p, w);
}

extern "C" int tlhip2d_ppcg_inner_iterations_f32_norm(
    int x, int y, int hd, int pitch, int count, const double* alphas, const double* betas,
    double theta, int init_from_r, const int* neighbours, const float* sd_in, float* sd_out,
    const float* r_in, float* r_out, const float* u_in, float* u_out, const float* kx,
    const float* ky, double alpha, const float* p, const float* w, double* rrn)
{
    if (!rrn) return tl_fail(TLHIP_ERR_ARG, "tlhip2d_ppcg_inner_iterations_f32_norm: null rrn");
    return inner_iterations<float>("tlhip2d_ppcg_inner_iterations_f32", x, y, hd, pitch, count,
                                   alphas, betas, theta, init_from_r, neighbours, sd_in, sd_out,
                                   r_in, r_out, u_in, u_out, kx, ky, alpha, p, w, nullptr, rrn);
}

namespace tl {
namespace mix {

int ppcg_inner32_norm(int x, int y, int hd, int pitch, int count, const double* alphas,
                      const double* betas, double theta, int init_from_r, const int* neighbours,
                      const float* sd_in, float* sd_out, const float* r_in, float* r_out,
                      const float* u_in, float* u_out, const float* kx, const float* ky,
                      double alpha, const float* p, const float* w, const double* alpha_dev,
                      double* rrn)
{
    return inner_iterations<float>("tlhip2d_ppcg_inner_iterations_f32", x, y, hd, pitch, count,
                                   alphas, betas, theta, init_from_r, neighbours, sd_in, sd_out,
                                   r_in, r_out, u_in, u_out, kx, ky, alpha, p, w, alpha_dev, rrn,
                                   false);
}

}  // namespace mix
}  // namespace tl

extern "C" int tlhip2d_cheby_iterations(int x, int y, int hd, int count, const double* alphas,
                                        const double* betas, const int* neighbours,
                                        const double* u_in, double* u_out, const double* u0,
                                        const double* p_in, double* p_out, double* r,
                                        const double* kx, const double* ky)
{
    if (x % 2)
        return tl_fail(TLHIP_ERR_ARG, "tlhip2d_cheby_iterations: needs an even x (got %d)", x);
    return cheby_iterations<double>("tlhip2d_cheby_iterations", x, y, hd, x, count, alphas, betas,
                                    neighbours, u_in, u_out, u0, p_in, p_out, r, kx, ky);
}

extern "C" int tlhip2d_cheby_iterations_f32(int x, int y, int hd, int pitch, int count,
                                            const double* alphas, const double* betas,
                                            const int* neighbours, const float* u_in,
                                            float* u_out, const float* u0, const float* p_in,
                                            float* p_out, float* r, const float* kx,
                                            const float* ky, double* rrn)
{
    return cheby_iterations<float>("tlhip2d_cheby_iterations_f32", x, y, hd, pitch, count, alphas,
                                   betas, neighbours, u_in, u_out, u0, p_in, p_out, r, kx, ky,
                                   rrn);
}

namespace tl {
namespace mix {

int cheby_iterations32(int x, int y, int hd, int P, int count, const double* alphas,
                       const double* betas, const int* neighbours, const float* d_in,
                       float* d_out, const float* b, const float* p_in, float* p_out, float* r,
                       const float* kx, const float* ky, double* rrn)
{
    return cheby_iterations<float>("tlhip2d_cheby_iterations_f32", x, y, hd, P, count, alphas,
                                   betas, neighbours, d_in, d_out, b, p_in, p_out, r, kx, ky, rrn,
                                   false);
}

}  // namespace mix
}  // namespace tl

// ------------------------------------------------ CG from the conductivity

extern "C" int tlhip2d_cg_cond(int x, int y, int hd, int coefficient, const double* density,
                               double* cond)
{
    const char* what = "tlhip2d_cg_cond";
    TL_TRY(tl_ensure_init());
    if (x < 1 || y < 1 || hd < 1) return tl_fail(TLHIP_ERR_ARG, "%s: bad grid", what);
    if (coefficient != 1 && coefficient != 2)
        return tl_fail(TLHIP_ERR_ARG, "%s: coefficient %d is not valid", what, coefficient);
    TL_REQUIRE(density && cond && density != cond, what, "bad arguments");
    const long n = (long)x * y;
    hipLaunchKernelGGL(k_cond2d, dim3(flat_blocks(n)), dim3(kBlock), 0, tl_stream(), n,
                       coefficient, density, cond);
    return tl_check_launch(what);
}

extern "C" int tlhip2d_cg_calc_pw_cond(int x, int y, int hd, int update, double beta, double rx,
                                       double ry, double* pw, const double* p_in, double* p_out,
                                       const double* r, double* w, const double* cond)
{
    const char* what = "tlhip2d_cg_calc_pw_cond";
    TL_REQUIRE(pw, what, "null argument");
    unsigned nb = 0;
    TL_TRY(launch_pw_cond(what, x, y, hd, update != 0, beta, nullptr, p_in, p_out, r, w, cond,
                          rx, ry, &nb, true));
    TL_TRY(tl_finish_reduce(1, nb, what));
    return fetch_add(pw, what);
}

extern "C" int tlhip2d_cg_calc_ur_cond(int x, int y, int hd, double alpha, double rx, double ry,
                                       double* rrn, double* u, const double* p, double* r,
                                       const double* cond)
{
    const char* what = "tlhip2d_cg_calc_ur_cond";
    TL_REQUIRE(rrn, what, "null argument");
    unsigned nb = 0;
    TL_TRY(launch_ur_cond(what, x, y, hd, alpha, nullptr, u, p, r, cond, rx, ry, &nb));
    TL_TRY(tl_finish_reduce(1, nb, what));
    return fetch_add(rrn, what);
}
