// Temporally blocked PPCG inner loop (2D) with the coefficient rows in LDS.
//
// The same S-level march as the register-window kernel in tl_fused2d.hip
// (S consecutive inner steps of 2d/drivers/ppcg_driver.c:165-183 -- a
// depth-1 mirror halo update of sd, then ppcg_inner_iteration,
// 2d/c_kernels/omp3/ppcg.c:28-61 -- in one pass; each lane holds one
// 16-byte vector of a row, a wave marches up its rows and level k computes
// row rho1 - k + 1 while level 1 computes row rho1), with the state laid
// out for the instruction budget of one wave per SIMD:
//
//  * Level k needs the diagonal (1 + (kx[i+1] + kx[i])) + (ky[i+x] + ky[i]),
//    kx (and the next lane's kx) and ky of its own row, and ky of the row
//    above.  Instead of S-row register windows of kx and ky, each wave keeps
//    a ring of S+1 row slots in LDS, each slot holding one row's diagonal,
//    kx and ky.  A row's diagonal is formed once, when the row is loaded,
//    not once per level.  Level k reads its row's slot (four LDS reads:
//    d, kx, the next lane's kx, ky); ky of the row above is the previous
//    level's ky, passed along in registers.  The slot roles rotate by one
//    slot per row (S+1 uniform offsets).
//  * The march is unrolled by three rows.  The three-row sd window of each
//    level and the one-row hand-off of r and u from level k to level k+1
//    rotate through three register sets whose roles are fixed per unrolled
//    row, so no values are copied to rotate them.
//  * Rows whose levels touch the first or last interior row, and waves
//    whose window holds the first or last interior column, take the body
//    with the mirror tests; the others the body without.
//
// Mode 3 runs S turns of the Chebyshev loop (2d/drivers/cheby_driver.c: a
// depth-1 mirror of u, then cheby_iterate and cheby_calc_u,
// 2d/c_kernels/omp3/cheby.c:8-24,57-85) the same way: u takes sd's place
// (three-row windows per level), and the per-level hand-off carries u0 and p
// of the row instead of r and u.
//
// The per-cell expressions are those of the reference, in its order and
// association (the diagonal included), so the FP64 result is bit-for-bit
// that of S separate steps; HBM traffic for S steps is that of one (read
// sd or r, kx, ky, r, u -- plus w and p for MODE 2; write sd, r, u).
// Rows outside the interior, and the levels' first rows below the tile,
// hold values that no kept cell depends on.
#include "tl_fused.h"
#include "tl_loop.h"
#include "tealeaf_hip.h"

#include <algorithm>
#include <type_traits>

// Scheduling regions of the level loop: 0 a barrier after every level (the
// default), 1 none, 2 after every second level.
#ifndef TL_RING_SCHED
#define TL_RING_SCHED 0
#endif

using namespace tl;
using namespace tl::fused;

namespace tl {
namespace fused {
double g_ring_fill = -1.0;   // tlhip_set_ring_fill
}
}

namespace {

// Per-wave LDS ring: R = S + 1 slots, each one row's d | kx | ky, one
// 16-byte vector per lane each.
template <int S>
struct RingGeom {
    static constexpr int R = S + 1;
    static constexpr int kRow = TL_WAVE * 16;
    static constexpr int kD = 0, kKx = kRow, kKy = 2 * kRow;
    static constexpr int kSlot = 3 * kRow;
    static constexpr int kCoef = R * kSlot;          // alpha_k, beta_k pairs (16 B each)
    static constexpr int kWave = kCoef + 16 * S;
};

template <class T>
__device__ __forceinline__ Cells<T> lds_ld(const unsigned char* p)
{
    return __builtin_bit_cast(Cells<T>, *reinterpret_cast<const typename Lane16<T>::V*>(p));
}

template <class T>
__device__ __forceinline__ void lds_st(unsigned char* p, const Cells<T>& c)
{
    *reinterpret_cast<typename Lane16<T>::V*>(p) = __builtin_bit_cast(typename Lane16<T>::V, c);
}

// Value of lane-1 / lane+1 (DPP wave_shr:1 / wave_shl:1).  The wave's end
// lanes receive zero: in this kernel their results are never kept (windows
// overlap their neighbours), so no old value has to be set up first.
__device__ __forceinline__ int shr1_i(int v)
{
    return __builtin_amdgcn_mov_dpp(v, 0x138, 0xf, 0xf, true);
}
__device__ __forceinline__ int shl1_i(int v)
{
    return __builtin_amdgcn_mov_dpp(v, 0x130, 0xf, 0xf, true);
}
__device__ __forceinline__ double shr1(double v)
{
    const int2 a = __builtin_bit_cast(int2, v);
    return __builtin_bit_cast(double, make_int2(shr1_i(a.x), shr1_i(a.y)));
}
__device__ __forceinline__ double shl1(double v)
{
    const int2 a = __builtin_bit_cast(int2, v);
    return __builtin_bit_cast(double, make_int2(shl1_i(a.x), shl1_i(a.y)));
}
__device__ __forceinline__ float shr1(float v)
{
    return __builtin_bit_cast(float, shr1_i(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ float shl1(float v)
{
    return __builtin_bit_cast(float, shl1_i(__builtin_bit_cast(int, v)));
}

// A * a for a lane's N cells (2d/shared.h:50-54, same association) with
// the diagonal and the next lane's kx formed beforehand, and x neighbours
// outside the interior mirroring the cell (XM 0: the wave's window holds
// neither the first nor the last interior column; 1: it does, and both
// lie on a lane's first / last cell, so only the values shifted in from
// the neighbouring lanes change; 2: anywhere).  lm / rm: this lane holds
// the first interior column as its first cell / the last as its last
// (XM 1).  The same operations as apply_a_pre, with don't-care values at
// the wave ends.
template <int XM, class T>
__device__ __forceinline__ Cells<T> apply_a_ring(const Cells<T>& cen, const Cells<T>& up,
                                                 const Cells<T>& dn, const Cells<T>& kxc, T kxe,
                                                 const Cells<T>& kyu, const Cells<T>& kyc,
                                                 const Cells<T>& d, int c, int xin0, int xin1,
                                                 bool lm, bool rm)
{
    constexpr int N = Lane16<T>::N;
    T lft_edge = shr1(cen.v[N - 1]);
    T rgt_edge = shl1(cen.v[0]);
    if constexpr (XM == 1) {
        if (lm) lft_edge = cen.v[0];
        if (rm) rgt_edge = cen.v[N - 1];
    }
    Cells<T> a;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        T lft = i == 0 ? lft_edge : cen.v[i - 1];
        T rgt = i == N - 1 ? rgt_edge : cen.v[i + 1];
        if constexpr (XM == 2) {
            if (c + i == xin0) lft = cen.v[i];
            if (c + i == xin1 - 1) rgt = cen.v[i];
        }
        const T kxr = i == N - 1 ? kxe : kxc.v[i + 1];
        const T tx = mad(kxr, rgt, kxc.v[i] * lft), ty = mad(kyu.v[i], up.v[i], kyc.v[i] * dn.v[i]);
        a.v[i] = mad(d.v[i], cen.v[i], -tx) - ty;
    }
    return a;
}

template <int P>
using Phase = std::integral_constant<int, P>;

// One wave per block: a block is one strip of one tile, so the launch has
// no idle waves for strip counts that are not a multiple of a block's
// waves, and the tile count can be sized to the SIMDs the device has.
constexpr int kRingBlock = TL_WAVE;

// NORM: also the squared 2-norm of the r written (the kept interior cells),
// as one FP64 partial per block into rd (summed by the finisher).
template <class T, int S, int MODE, bool NORM>
__global__ void __launch_bounds__(kRingBlock)
k_ppcg_ring2d(PitchGrid g, StripTile t, PpcgCoefs<T, S> cf, T theta, T alpha_arg,
              const double* __restrict__ alpha_dev,
              const T* __restrict__ sd_in, T* __restrict__ sd_out,
              const T* __restrict__ r_in, T* __restrict__ r_out,
              const T* __restrict__ u_in, T* __restrict__ u_out,
              const T* __restrict__ kx, const T* __restrict__ ky,
              const T* __restrict__ p_in, const T* __restrict__ w_in, TlGate gate, TlReduce rd)
{
    using RG = RingGeom<S>;
    constexpr int R = RG::R;
    __shared__ __attribute__((aligned(16))) unsigned char s_ring[RG::kWave];
    constexpr bool FROM_R = MODE == 1 || MODE == 2;
    constexpr bool UR = MODE == 2;
    constexpr bool CHEBY = MODE == 3;
    constexpr int N = Lane16<T>::N;
    using C = Cells<T>;
    if (gate.closed()) return;
    // cg_calc_ur's alpha: the argument, or the resident loop's device value
    const T alpha_ur = UR && alpha_dev ? (T)*alpha_dev : alpha_arg;
    const int lane = threadIdx.x;
    const int tile = xcd_tile_of(blockIdx.x, t.nbx * t.nby);
    const int sx = tile % t.nbx, by = tile / t.nbx;   // t.nbx == t.ns: one strip per block

    const int xin0 = g.hd, xin1 = g.hd + g.nx;  // interior columns
    const int yin0 = g.hd, yin1 = g.hd + g.ny;  // interior rows
    const int o0 = xin0 + sx * t.w;
    const int o1 = min(o0 + t.w, xin1);
    const int c = t.cw0 + sx * t.w + N * lane;  // this lane's first column
    const bool all_out = c >= o0 && c + N <= o1;
    const bool any_out = c + N > o0 && c < o1;
    const bool lm = c == xin0, rm = c + N == xin1;  // the lane's first / last cell is an interior edge

    const int j0 = yin0 + by * t.ty;
    const int j1 = min(j0 + t.ty, yin1);
    const int rs = j0 - S + 1;                  // level 1's row in the first iteration
    const int niter = j1 - j0 + 2 * S - 2;
    const long X = g.X;

    // Fields are read and written through buffer descriptors: the lane's
    // column as the per-lane offset (a lane left of column 0 passes a
    // negative, i.e. out-of-range, offset and reads zeros; writes are masked
    // to the output columns) and the row as a uniform offset, clamped into
    // the array.  Rows outside the interior and lanes outside the window's
    // output columns hold values that no kept cell depends on (the mirror
    // replaces every neighbour outside the interior), so clamping only keeps
    // every access inside the array.
    const int ylast = g.ny + 2 * g.hd - 1;
    const int row_bytes = (int)(X * (long)sizeof(T));
    const int lane_off = c * (int)sizeof(T);
    const int nrec = (int)min((long)(ylast + 1) * row_bytes, (long)0x7FFFFF00);
    auto rsrc = [&](const T* f) {
        return __builtin_amdgcn_make_buffer_rsrc((void*)f, (short)0, nrec, kBufferWord3);
    };
    const __amdgpu_buffer_rsrc_t b_a = rsrc(FROM_R ? r_in : sd_in), b_r = rsrc(r_in),
                                 b_u = rsrc(u_in), b_kx = rsrc(kx), b_ky = rsrc(ky),
                                 b_p = rsrc(UR ? p_in : u_in), b_w = rsrc(UR ? w_in : u_in),
                                 o_r = rsrc(r_out), o_u = rsrc(u_out), o_sd = rsrc(sd_out);
    auto row_of = [&](int row) { return min(max(row, 0), ylast) * row_bytes; };
    auto ld_field = [&](__amdgpu_buffer_rsrc_t f, int row) -> C {
        return __builtin_bit_cast(
            C, __builtin_amdgcn_raw_buffer_load_b128(f, lane_off, row_of(row), 0));
    };
    // sd = r / theta (ppcg.c:19-26); FP32 multiplies by 1/theta
    const T inv_theta = (T)1 / theta;
    auto over_theta = [&](C a) -> C {
#pragma unroll
        for (int i = 0; i < N; ++i) {
            if constexpr (std::is_same_v<T, double>) a.v[i] = a.v[i] / theta;
            else a.v[i] = a.v[i] * inv_theta;
        }
        return a;
    };
    // r - alpha w and u + alpha p (cg.c:128-129, same association)
    auto minus_aw = [&](C a, const C& b) -> C {
#pragma unroll
        for (int i = 0; i < N; ++i) a.v[i] = mad(-alpha_ur, b.v[i], a.v[i]);
        return a;
    };
    auto plus_ap = [&](C a, const C& b) -> C {
#pragma unroll
        for (int i = 0; i < N; ++i) a.v[i] = mad(alpha_ur, b.v[i], a.v[i]);
        return a;
    };
    auto r_level0 = [&](int row) -> C {
        C a = ld_field(b_r, row);
        if constexpr (UR) a = minus_aw(a, ld_field(b_w, row));
        return a;
    };

    // This lane's part of the wave's LDS ring; the slot of role j (row
    // rho1 + 1 - j, where rho1 is level 1's row) is at byte ao[j].  The
    // wave's copy of the step coefficients follows the ring: (alpha_k,
    // beta_k) at kCoef + 16 k, read by every lane at the same address.
    unsigned char* const wbase = s_ring;
    unsigned char* const ring = wbase + lane * 16;
#pragma unroll
    for (int q = 0; q < 3 * R; ++q) lds_st<T>(ring + q * RG::kRow, zeros<T>());
    if (lane < S) {
        T* ab = reinterpret_cast<T*>(wbase + RG::kCoef + 16 * lane);
        ab[0] = cf.alpha[lane];
        ab[1] = cf.beta[lane];
    }
    struct AB {
        T al, be;
    };
    auto coef = [&](int k) -> AB {    // (alpha, beta) of level k
        const T* ab = reinterpret_cast<const T*>(wbase + RG::kCoef + 16 * (k - 1));
        return AB{ab[0], ab[1]};
    };
    unsigned ao[R];
#pragma unroll
    for (int j = 0; j < R; ++j) ao[j] = (unsigned)(j * RG::kSlot);

    // Register state.  sdw[k][*]: three rows of level k's sd (k < S), in
    // roles that rotate with the row's phase (it % 3): in phase P, dn is
    // slot P, cen slot P+1, and the new row goes to slot P+2.  rc[P][k],
    // uc[P][k]: r and u of level k at the row level k+1 computes in phase P.
    C sdw[S][3];
    C rc[3][S], uc[3][S];
#pragma unroll
    for (int k = 0; k < S; ++k) {
        sdw[k][0] = sdw[k][1] = sdw[k][2] = zeros<T>();
#pragma unroll
        for (int q = 0; q < 3; ++q) rc[q][k] = uc[q][k] = zeros<T>();
    }

    // Prologue: level-0 sd rows rs-1 and rs (dn, cen of the first row).
    C rcarry = zeros<T>();    // FROM_R: level-0 r at row rho1, loaded one iteration ago
    if constexpr (FROM_R) {
        sdw[0][0] = over_theta(r_level0(rs - 1));
        rcarry = r_level0(rs);
        sdw[0][1] = over_theta(rcarry);
    } else {
        sdw[0][0] = ld_field(b_a, rs - 1);
        sdw[0][1] = ld_field(b_a, rs);
    }
    C kyprev = ld_field(b_ky, rs);               // ky of level 1's row
    lds_st<T>(ring + ao[1] + RG::kKy, kyprev);

    // Level-0 inputs of the row whose level 1 computes row rho1: a = sd (or
    // r for FROM_R) at rho1+1, b = r at rho1 (or, for UR, w at rho1+1), u
    // at rho1 (and p at rho1 for UR), kx at rho1, ky at rho1+1.  The loads
    // of the next row are issued before this row's arithmetic.
    struct In {
        C a, b, u, p, kx, ky;
    };
    auto load_in = [&](int rho1) -> In {
        In v;
        v.a = ld_field(b_a, rho1 + 1);
        if constexpr (!FROM_R) v.b = ld_field(b_r, rho1);
        if constexpr (UR) {
            v.b = ld_field(b_w, rho1 + 1);
            v.p = ld_field(b_p, rho1);
        }
        v.u = ld_field(b_u, rho1);
        v.kx = ld_field(b_kx, rho1);
        v.ky = ld_field(b_ky, rho1 + 1);
        return v;
    };
    In pre = load_in(rs);
    double racc = 0.0;    // NORM: this lane's sum of r^2 over the cells it stores

    // gen = 4 * XM + YM: YM the y-mirror tests, XM the x mirror (as
    // apply_a_ring)
    auto body = [&](int it, auto ph, auto gen) __attribute__((always_inline)) {
        constexpr int P = decltype(ph)::value;
        constexpr int P1 = (P + 1) % 3, P2 = (P + 2) % 3;
        constexpr bool G = (decltype(gen)::value & 1) != 0;   // y mirror
        constexpr int XG = decltype(gen)::value >> 2;          // x mirror
        const int rho1 = rs + it;
        const int dy0 = rho1 - yin0;              // level k is on row yin0 when dy0 == k - 1
        const int dy1 = rho1 - (yin1 - 1);        // ... on row yin1 - 1 when dy1 == k - 1
        const int dst = it - (2 * S - 2);         // level S's row - j0
        const In cur = pre;
        pre = load_in(rho1 + 1);                      // past the tile: loaded, unused
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
        sdw[0][P2] = sd_new;

        // Row rho1's coefficients: formed here for level 1, kept in the ring
        // for levels 2..S.
        const T kxe1 = shl1(cur.kx.v[0]);
        const C d1 = diag_lane(cur.kx, kxe1, cur.ky, kyprev);
        lds_st<T>(ring + ao[1] + RG::kD, d1);
        lds_st<T>(ring + ao[1] + RG::kKx, cur.kx);
        lds_st<T>(ring + ao[0] + RG::kKy, cur.ky);

        // Each level's ring reads are issued before the previous level's
        // arithmetic; the scheduling barriers keep the levels in order,
        // which bounds what is live at once.
        struct Co {
            C d, kx, ky;
            T kxe, al, be;
        };
        auto ring_ld = [&](int k) -> Co {
            Co v;
            const unsigned char* q = ring + ao[k];
            v.d = lds_ld<T>(q + RG::kD);
            v.kx = lds_ld<T>(q + RG::kKx);
            v.kxe = *reinterpret_cast<const T*>(q + RG::kKx + 16);
            v.ky = lds_ld<T>(q + RG::kKy);
            const AB ab = coef(k);
            v.al = ab.al;
            v.be = ab.be;
            return v;
        };
        Co nx;
        {
            const AB ab = coef(1);
            nx.d = d1;
            nx.kx = cur.kx;
            nx.kxe = kxe1;
            nx.ky = kyprev;
            nx.al = ab.al;
            nx.be = ab.be;
        }
        C kyu = cur.ky;                               // ky of the row above the level's row
#pragma unroll
        for (int k = 1; k <= S; ++k) {
            const int rho = rho1 - (k - 1);
            const Co co = nx;
            if (k < S) nx = ring_ld(k + 1);
            const C cen = sdw[k - 1][P1];
            // y neighbours; outside the interior they mirror the cell
            C upv = sdw[k - 1][P2], dnv = sdw[k - 1][P];
            if constexpr (G) {
                if (dy1 == k - 1) upv = cen;
                if (dy0 == k - 1) dnv = cen;
            }
            // PPCG: r and u of the row from level k-1; Chebyshev: u0 and p
            const C rin = k == 1 ? r0 : rc[P][k - 1];
            const C uin = k == 1 ? u0 : uc[P][k - 1];
            // SMVP (2d/shared.h:50-54) and the ppcg.c:46-59 (Chebyshev:
            // cheby.c:76-84 and cheby_calc_u) updates, same order
            const C av = apply_a_ring<XG>(cen, upv, dnv, co.kx, co.kxe, kyu, co.ky, co.d, c, xin0,
                                          xin1, lm, rm);
            // stored as r, u, sd (Chebyshev: r, p, u)
            C rn, un, sn;
#pragma unroll
            for (int i = 0; i < N; ++i) {
                rn.v[i] = rin.v[i] - av.v[i];
                if constexpr (CHEBY) {
                    un.v[i] = mad(co.al, uin.v[i], co.be * rn.v[i]);    // p
                    sn.v[i] = cen.v[i] + un.v[i];                       // u
                } else {
                    un.v[i] = uin.v[i] + cen.v[i];
                    sn.v[i] = mad(co.al, cen.v[i], co.be * rn.v[i]);
                }
            }
            if (k < S) {
                sdw[k][P2] = sn;
                rc[P1][k] = CHEBY ? rin : rn;
                uc[P1][k] = un;
            } else if (any_out && (unsigned)dst < (unsigned)(j1 - j0)) {
                const int ro = rho * row_bytes;          // rho is an interior row here
                if constexpr (NORM) {
#pragma unroll
                    for (int i = 0; i < N; ++i)
                        if (all_out || (c + i >= o0 && c + i < o1))
                            racc += (double)rn.v[i] * (double)rn.v[i];
                }
                if (all_out) {
                    using V = typename Lane16<T>::V;
                    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(V, rn), o_r,
                                                           lane_off, ro, 0);
                    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(V, un), o_u,
                                                           lane_off, ro, 0);
                    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(V, sn), o_sd,
                                                           lane_off, ro, 0);
                } else {
                    const long i0 = (long)rho * X + c;
#pragma unroll
                    for (int i = 0; i < N; ++i)
                        if (c + i >= o0 && c + i < o1) {
                            r_out[i0 + i] = rn.v[i];
                            u_out[i0 + i] = un.v[i];
                            sd_out[i0 + i] = sn.v[i];
                        }
                }
            }
            kyu = co.ky;
#if TL_RING_SCHED == 0
            __builtin_amdgcn_sched_barrier(0);
#elif TL_RING_SCHED == 2
            if (k % 2 == 0) __builtin_amdgcn_sched_barrier(0);
#endif
        }
        kyprev = cur.ky;
        // Roles move up one row: role j takes role j-1's slot, role 0 the
        // slot that role R-1 (level S's row) has finished with.
        const unsigned last = ao[R - 1];
#pragma unroll
        for (int j = R - 1; j > 0; --j) ao[j] = ao[j - 1];
        ao[0] = last;
    };

    // The march runs in trips of three rows (phases 0, 1, 2: the register
    // roles return to where they started), niter rounded up to a multiple
    // of 3 (the extra rows are above the tile and not stored).  Trips in
    // [m0, m1) take the body without mirror tests; the others, and every
    // trip of a wave whose window holds the first or last interior column,
    // the body with them.  Levels of iteration it cover rows
    // rs+it-S+1 .. rs+it.
    const int wc0 = t.cw0 + sx * t.w;
    const bool xedge = (xin0 >= wc0 && xin0 < wc0 + TL_WAVE * N) ||
                       (xin1 - 1 >= wc0 && xin1 - 1 < wc0 + TL_WAVE * N);
    const bool lane_edges = xin0 % N == 0 && xin1 % N == 0;
    const int niter3 = (niter + 2) / 3 * 3;
    const int e0 = max(0, yin0 - rs + S);           // first iteration above yin0's
    const int e1 = min(niter, yin1 - 1 - rs);       // first iteration touching yin1 - 1
    const int m0 = min((e0 + 2) / 3 * 3, niter3);
    const int m1 = e1 > m0 ? m0 + (e1 - m0) / 3 * 3 : m0;
    // Each path is its own loop with no branch inside a trip, so that the
    // register roles of one kind of trip are all that is live across it.
    auto trips = [&](int from, int to, auto gen) __attribute__((always_inline)) {
        for (int it = from; it < to; it += 3) {
            body(it, Phase<0>{}, gen);
            body(it + 1, Phase<1>{}, gen);
            body(it + 2, Phase<2>{}, gen);
        }
    };
    using Gen = std::integral_constant<int, 0>;     // no tests
    using GenY = std::integral_constant<int, 1>;    // y mirror
    using GenE = std::integral_constant<int, 4>;    // x mirror at lane edges
    using GenEY = std::integral_constant<int, 5>;
    using GenXY = std::integral_constant<int, 9>;   // x mirror anywhere, y mirror
    if (!xedge) {
        if (m0 == 0 && m1 == niter3) {
            trips(0, niter3, Gen{});                    // a tile clear of the first and last rows
        } else {
            trips(0, m0, GenY{});
            trips(m0, m1, Gen{});
            trips(m1, niter3, GenY{});
        }
    } else if (lane_edges) {
        trips(0, m0, GenEY{});
        trips(m0, m1, GenE{});
        trips(m1, niter3, GenEY{});
    } else {
        trips(0, niter3, GenXY{});
    }
    if constexpr (NORM) tl_reduce_partial<1, kRingBlock>(&racc, rd);
}

// Default tile height: as many tiles (one wave each) as fill a fraction of
// the waves the device holds at once -- one round; a second, mostly idle
// round is the cliff.
template <class T, int S, int MODE, bool NORM>
int ring_tile_rows(const PitchGrid& g, int ns)
{
    static int occ = 0;
    if (occ <= 0 &&
        (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_ppcg_ring2d<T, S, MODE, NORM>,
                                                      kRingBlock, 0) != hipSuccess ||
         occ <= 0))
        occ = 1;
    const double fill = g_ring_fill > 0.0 ? g_ring_fill : 0.95;
    const long cap = (long)(fill * device_cus() * occ);
    const long nby = std::max(1L, cap / std::max(1, ns));
    return (int)std::max(1L, ((long)g.ny + nby - 1) / nby);
}

template <class T, int S, int MODE, bool NORM>
int launch_ring(const FusedArgs<T>& a)
{
    // Windows reach S columns beyond their output columns on both sides.
    StripTile t = strip_tile<T>(a.g, S, 1, 1);
    const int ty = g_fused_ty > 0 ? g_fused_ty : ring_tile_rows<T, S, MODE, NORM>(a.g, t.ns);
    t = strip_tile<T>(a.g, S, ty, 1);
    if ((long)t.nbx * t.nby > (NORM ? (long)TL_MAX_RED_BLOCKS : (1L << 30)))
        return tl_fail(TLHIP_ERR_ARG, "%s: grid too large", a.what);
    PpcgCoefs<T, S> cf;
    for (int k = 0; k < S; ++k) {
        cf.alpha[k] = (T)a.al[k];
        cf.beta[k] = (T)a.be[k];
    }
    const unsigned nb = (unsigned)(t.nbx * t.nby);
    const int slot = tl_timer_begin(a.what);
    hipLaunchKernelGGL((k_ppcg_ring2d<T, S, MODE, NORM>), dim3(nb), dim3(kRingBlock), 0,
                       tl_stream(), a.g, t, cf, (T)a.theta, (T)a.alpha_ur, a.alpha_dev, a.sd_in,
                       a.sd_out, a.r_in, a.r_out, a.u_in, a.u_out, a.kx, a.ky, a.p, a.w,
                       tl_gate(), tl_reduce_slots());
    tl_timer_end(slot);
    TL_TRY(tl_check_launch(a.what));
    if constexpr (NORM) {
        if (a.norm_nb) {
            *a.norm_nb = nb;
            return 0;
        }
        TL_TRY(tl_finish_reduce(1, nb, a.what));
        return tl_fetch(1, a.rrn, a.what);
    }
    return 0;
}

template <class T, int MODE, bool NORM>
int dispatch_ring_n(int count, const FusedArgs<T>& a)
{
    switch (count) {
    case 1: return launch_ring<T, 1, MODE, NORM>(a);
    case 2: return launch_ring<T, 2, MODE, NORM>(a);
    case 3: return launch_ring<T, 3, MODE, NORM>(a);
    case 4: return launch_ring<T, 4, MODE, NORM>(a);
    case 5: return launch_ring<T, 5, MODE, NORM>(a);
    case 6: return launch_ring<T, 6, MODE, NORM>(a);
    case 7: return launch_ring<T, 7, MODE, NORM>(a);
    case 8: return launch_ring<T, 8, MODE, NORM>(a);
    case 9: return launch_ring<T, 9, MODE, NORM>(a);
    case 10: return launch_ring<T, 10, MODE, NORM>(a);
    default: return tl_fail(TLHIP_ERR_ARG, "%s: count must be 1..%d", a.what, kMaxFused);
    }
}

template <class T, int MODE>
int dispatch_ring(int count, const FusedArgs<T>& a)
{
    if (!a.rrn && !a.norm_nb) return dispatch_ring_n<T, MODE, false>(count, a);
    if constexpr (MODE == 0 || MODE == 2 || std::is_same_v<T, float>)
        return dispatch_ring_n<T, MODE, true>(count, a);
    return tl_fail(TLHIP_ERR_ARG, "%s: no fused 2-norm for this launch", a.what);
}

}  // namespace

namespace tl {
namespace fused {

template <class T>
int launch_ppcg_ring(int count, int mode, const FusedArgs<T>& a)
{
    switch (mode) {
    case 0: return dispatch_ring<T, 0>(count, a);
    case 1: return dispatch_ring<T, 1>(count, a);
    case 2: return dispatch_ring<T, 2>(count, a);
    default: return dispatch_ring<T, 3>(count, a);
    }
}

template int launch_ppcg_ring<double>(int, int, const FusedArgs<double>&);
template int launch_ppcg_ring<float>(int, int, const FusedArgs<float>&);

}  // namespace fused
}  // namespace tl

extern "C" int tlhip_set_ring_fill(double fill)
{
    if (!(fill <= 1.5)) return tl_fail(TLHIP_ERR_ARG, "tlhip_set_ring_fill: %g", fill);
    g_ring_fill = fill;
    return 0;
}
