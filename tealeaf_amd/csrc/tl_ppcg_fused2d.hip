// Temporally blocked PPCG inner loop (2D): S consecutive inner steps of
// 2d/drivers/ppcg_driver.c:165-183 in one pass over memory.
//
// The reference runs, per inner step pp, a depth-1 halo update of sd
// followed by ppcg_inner_iteration (2d/c_kernels/omp3/ppcg.c:28-61):
//     r  -= A sd ;  u += sd ;  sd = alpha_pp sd + beta_pp r.
// With every face external the halo update is a mirror, so a neighbour
// outside the interior simply takes the cell's own value.  Every time level
// k of a cell depends only on level k-1 within one cell of it, so a wave can
// carry S levels at once:
//  * it marches up its rows keeping, for each level, the last three rows of
//    sd and the last row of r and u in registers; level k computes row
//    (j - k + 1) while level 1 computes row j;
//  * its 128-column window overlaps the neighbouring windows by SP >= S
//    columns on each side and it writes only the centre 128 - 2*SP, where
//    all S levels are exact; the march starts S-1 rows below its tile for
//    the same reason in y;
//  * r, u and sd are read from one buffer and written to another, because
//    the overlapped windows of other waves still need the level-0 values.
// Each level applies the reference's per-cell expression in the reference's
// order, so the result is bit-for-bit that of S separate steps.  HBM traffic
// for S steps is that of one: read sd, kx, ky, r, u; write sd, r, u.
// Optionally sd at level 0 is ppcg_init's r/theta (ppcg.c:8-25), computed on
// the fly from r.
#include "tl_ops.h"
#include "tealeaf_hip.h"

using namespace tl;

namespace {

constexpr int kMaxFused = 5;

template <int S>
struct PpcgCoefs {
    double alpha[S], beta[S];
};

struct FusedTile {
    int w;     // output columns per wave
    int ns;    // strips along x
    int nbx;   // blocks along x
    int nby;   // tiles along y
    int ty;    // output rows per tile
};

__device__ __forceinline__ int xcd_tile(int bid, int nb)
{
    const int q = nb >> 3, r = nb & 7, k = bid & 7, i = bid >> 3;
    return k * q + (k < r ? k : r) + i;
}

__device__ __forceinline__ double2 ld2(const double* __restrict__ p, long i)
{
    return *reinterpret_cast<const double2*>(p + i);
}

template <int S, bool FROM_R>
__global__ void __launch_bounds__(kBlock)
k_ppcg_fused2d(TlGrid2 g, FusedTile t, PpcgCoefs<S> cf, double theta,
               const double* __restrict__ sd_in, double* __restrict__ sd_out,
               const double* __restrict__ r_in, double* __restrict__ r_out,
               const double* __restrict__ u_in, double* __restrict__ u_out,
               const double* __restrict__ kx, const double* __restrict__ ky)
{
    constexpr int SP = S + (S & 1);
    const int lane = threadIdx.x & (TL_WAVE - 1), wave = threadIdx.x / TL_WAVE;
    const int tile = xcd_tile(blockIdx.x, t.nbx * t.nby);
    const int bx = tile % t.nbx, by = tile / t.nbx;
    const int sx = bx * kWaves + wave;
    if (sx >= t.ns) return;                      // uniform per wave; no block sync below

    const int xin0 = g.hd, xin1 = g.x - g.hd;   // interior columns
    const int yin0 = g.hd, yin1 = g.y - g.hd;   // interior rows
    const int o0 = xin0 + sx * t.w;
    const int o1 = min(o0 + t.w, xin1);
    const int c = o0 - SP + 2 * lane;           // this lane's columns c, c+1
    const bool ld_ok = c >= 0 && c + 2 <= g.x;
    const bool out_col = c >= o0 && c < o1;     // both cells or neither
    // x mirror: a neighbour outside the interior is the cell itself.
    const bool mir_l0 = c == xin0, mir_l1 = c + 1 == xin0;
    const bool mir_r0 = c == xin1 - 1, mir_r1 = c + 1 == xin1 - 1;

    const int j0 = yin0 + by * t.ty;
    const int j1 = min(j0 + t.ty, yin1);
    const int rs = j0 - S + 1;                  // first row of level 1
    const int niter = j1 - j0 + 2 * S - 2;
    const long X = g.x;

    auto ld_field = [&](const double* __restrict__ f, int row, int lo, int hi) -> double2 {
        if (ld_ok && row >= lo && row < hi) return ld2(f, (long)row * X + c);
        return make_double2(0.0, 0.0);
    };

    // Windows (index 0 oldest).  sdw[k][*]: rows of level k for k < S.
    double2 sdw[S][3];
    double2 rprev[S], uprev[S];     // level k's r, u from the previous iteration
    double2 kxw[S], kyw[S + 1];     // kx rows rho1 - m; ky rows rho1 + 1 - m
#pragma unroll
    for (int k = 0; k < S; ++k) {
        sdw[k][0] = sdw[k][1] = sdw[k][2] = make_double2(0.0, 0.0);
        rprev[k] = uprev[k] = make_double2(0.0, 0.0);
        kxw[k] = make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int m = 0; m <= S; ++m) kyw[m] = make_double2(0.0, 0.0);

    // Prologue: level-0 sd rows rs-1, rs; kx rows below rs; ky rows up to rs.
    double2 rnext = make_double2(0.0, 0.0);     // r_in row rho1 (FROM_R carries it)
    if constexpr (FROM_R) {
        const double2 a = ld_field(r_in, rs - 1, yin0, yin1);
        const double2 b = ld_field(r_in, rs, yin0, yin1);
        sdw[0][1] = make_double2(a.x / theta, a.y / theta);
        sdw[0][2] = make_double2(b.x / theta, b.y / theta);
        rnext = b;
    } else {
        sdw[0][1] = ld_field(sd_in, rs - 1, yin0, yin1);
        sdw[0][2] = ld_field(sd_in, rs, yin0, yin1);
    }
#pragma unroll
    for (int m = 0; m + 1 < S; ++m) kxw[m] = ld_field(kx, rs - 1 - m, yin0, yin1);
#pragma unroll
    for (int m = 0; m < S; ++m) kyw[m] = ld_field(ky, rs - m, yin0, yin1 + 1);

    for (int it = 0; it < niter; ++it) {
        const int rho1 = rs + it;
        // level-0 inputs for this iteration
        double2 sd_new, r0;
        if constexpr (FROM_R) {
            r0 = rnext;
            rnext = ld_field(r_in, rho1 + 1, yin0, yin1);
            sd_new = make_double2(rnext.x / theta, rnext.y / theta);
        } else {
            r0 = ld_field(r_in, rho1, yin0, yin1);
            sd_new = ld_field(sd_in, rho1 + 1, yin0, yin1);
        }
        const double2 u0 = ld_field(u_in, rho1, yin0, yin1);
        const double2 kx_new = ld_field(kx, rho1, yin0, yin1);
        const double2 ky_new = ld_field(ky, rho1 + 1, yin0, yin1 + 1);
        sdw[0][0] = sdw[0][1];
        sdw[0][1] = sdw[0][2];
        sdw[0][2] = sd_new;
#pragma unroll
        for (int m = S - 1; m > 0; --m) kxw[m] = kxw[m - 1];
        kxw[0] = kx_new;
#pragma unroll
        for (int m = S; m > 0; --m) kyw[m] = kyw[m - 1];
        kyw[0] = ky_new;

        double2 rcur[S], ucur[S];
#pragma unroll
        for (int k = 1; k <= S; ++k) {
            const int rho = rho1 - (k - 1);
            const double2 cen = sdw[k - 1][1];
            double2 upv = sdw[k - 1][2], dnv = sdw[k - 1][0];
            if (rho == yin0) dnv = cen;                 // bottom mirror
            if (rho == yin1 - 1) upv = cen;             // top mirror
            // x neighbours: lane-1's second cell, lane+1's first cell
            double lft0 = tl_lane_prev(cen.y, cen.x);
            double rgt1 = tl_lane_next(cen.x, cen.y);
            double lft1 = cen.x, rgt0 = cen.y;
            if (mir_l0) lft0 = cen.x;
            if (mir_l1) lft1 = cen.y;
            if (mir_r0) rgt0 = cen.x;
            if (mir_r1) rgt1 = cen.y;
            const double2 kxc = kxw[k - 1];
            const double kxr0 = kxc.y;
            const double kxr1 = tl_lane_next(kxc.x, kxc.y);
            const double2 kyu = kyw[k - 1], kyc = kyw[k];
            const double2 rin = k == 1 ? r0 : rprev[k - 1];
            const double2 uin = k == 1 ? u0 : uprev[k - 1];
            // SMVP (2d/shared.h:50-54) and the ppcg.c:46-59 updates, same order
            const double d0 = (1.0 + (kxr0 + kxc.x)) + (kyu.x + kyc.x);
            const double d1 = (1.0 + (kxr1 + kxc.y)) + (kyu.y + kyc.y);
            const double a0 = d0 * cen.x - (kxr0 * rgt0 + kxc.x * lft0) - (kyu.x * upv.x + kyc.x * dnv.x);
            const double a1 = d1 * cen.y - (kxr1 * rgt1 + kxc.y * lft1) - (kyu.y * upv.y + kyc.y * dnv.y);
            double2 rn, un, sn;
            rn.x = rin.x - a0;
            rn.y = rin.y - a1;
            un.x = uin.x + cen.x;
            un.y = uin.y + cen.y;
            sn.x = cf.alpha[k - 1] * cen.x + cf.beta[k - 1] * rn.x;
            sn.y = cf.alpha[k - 1] * cen.y + cf.beta[k - 1] * rn.y;
            rcur[k - 1] = rn;
            ucur[k - 1] = un;
            if (k < S) {
                sdw[k][0] = sdw[k][1];
                sdw[k][1] = sdw[k][2];
                sdw[k][2] = sn;
            } else if (out_col && rho >= j0 && rho < j1) {
                const long i = (long)rho * X + c;
                *reinterpret_cast<double2*>(r_out + i) = rn;
                *reinterpret_cast<double2*>(u_out + i) = un;
                *reinterpret_cast<double2*>(sd_out + i) = sn;
            }
        }
#pragma unroll
        for (int k = 1; k < S; ++k) {
            rprev[k] = rcur[k - 1];
            uprev[k] = ucur[k - 1];
        }
    }
}

// Halo ring of up to three fields, src -> dst.
struct Ring3 {
    const double* src[3];
    double* dst[3];
    int n;
};

__global__ void __launch_bounds__(kBlock) k_copy_rings2d(TlGrid2 g, Ring3 rg)
{
    const long nrow = 2L * g.hd * g.x;
    const long ncol = 2L * g.hd * (g.y - 2 * g.hd);
    const long per = nrow + ncol;
    for (long t = blockIdx.x * (long)kBlock + threadIdx.x; t < per * rg.n;
         t += (long)gridDim.x * kBlock) {
        const int f = (int)(t / per);
        const long q = t % per;
        long c;
        if (q < nrow) {
            const int rr = (int)(q / g.x), kk = (int)(q % g.x);
            const int jj = rr < g.hd ? rr : g.y - 2 * g.hd + rr;
            c = (long)jj * g.x + kk;
        } else {
            const long u = q - nrow;
            const int jj = g.hd + (int)(u / (2 * g.hd));
            const int p = (int)(u % (2 * g.hd));
            const int kk = p < g.hd ? p : g.x - 2 * g.hd + p;
            c = (long)jj * g.x + kk;
        }
        rg.dst[f][c] = rg.src[f][c];
    }
}

int g_fused_ty = -1;

template <int S, bool FROM_R>
int launch_fused(const TlGrid2& g, const double* al, const double* be, double theta,
                 const double* sd_in, double* sd_out, const double* r_in, double* r_out,
                 const double* u_in, double* u_out, const double* kx, const double* ky,
                 const char* what)
{
    constexpr int SP = S + (S & 1);
    FusedTile t;
    const int nx = g.x - 2 * g.hd, ny = g.y - 2 * g.hd;
    t.w = TL_WAVE * 2 - 2 * SP;
    t.ns = (int)tl_cdiv(nx, t.w);
    t.nbx = (int)tl_cdiv(t.ns, kWaves);
    t.ty = g_fused_ty > 0 ? g_fused_ty : 32;
    t.nby = (int)tl_cdiv(ny, t.ty);
    PpcgCoefs<S> cf;
    for (int k = 0; k < S; ++k) {
        cf.alpha[k] = al[k];
        cf.beta[k] = be[k];
    }
    const unsigned nb = (unsigned)(t.nbx * t.nby);
    const int slot = tl_timer_begin(what);
    hipLaunchKernelGGL((k_ppcg_fused2d<S, FROM_R>), dim3(nb), dim3(kBlock), 0, tl_stream(), g, t,
                       cf, theta, sd_in, sd_out, r_in, r_out, u_in, u_out, kx, ky);
    tl_timer_end(slot);
    return tl_check_launch(what);
}

template <bool FROM_R>
int dispatch_fused(int count, const TlGrid2& g, const double* al, const double* be, double theta,
                   const double* sd_in, double* sd_out, const double* r_in, double* r_out,
                   const double* u_in, double* u_out, const double* kx, const double* ky,
                   const char* what)
{
    switch (count) {
    case 1: return launch_fused<1, FROM_R>(g, al, be, theta, sd_in, sd_out, r_in, r_out, u_in, u_out, kx, ky, what);
    case 2: return launch_fused<2, FROM_R>(g, al, be, theta, sd_in, sd_out, r_in, r_out, u_in, u_out, kx, ky, what);
    case 3: return launch_fused<3, FROM_R>(g, al, be, theta, sd_in, sd_out, r_in, r_out, u_in, u_out, kx, ky, what);
    case 4: return launch_fused<4, FROM_R>(g, al, be, theta, sd_in, sd_out, r_in, r_out, u_in, u_out, kx, ky, what);
    case 5: return launch_fused<5, FROM_R>(g, al, be, theta, sd_in, sd_out, r_in, r_out, u_in, u_out, kx, ky, what);
    default: return tl_fail(TLHIP_ERR_ARG, "%s: count must be 1..%d", what, kMaxFused);
    }
}

}  // namespace

extern "C" int tlhip_set_fused_tile_rows(int rows)
{
    g_fused_ty = rows;
    return 0;
}

extern "C" int tlhip2d_ppcg_inner_iterations(int x, int y, int hd, int count,
                                             const double* alphas, const double* betas,
                                             double theta, int init_from_r,
                                             const int* neighbours, const double* sd_in,
                                             double* sd_out, const double* r_in, double* r_out,
                                             const double* u_in, double* u_out,
                                             const double* kx, const double* ky)
{
    const char* what = "tlhip2d_ppcg_inner_iterations";
    TL_TRY(tl_ensure_init());
    const TlGrid2 g{x, y, hd};
    if (hd < 2 || hd % 2 || x % 2 || x < 2 * hd + 1 || y < 2 * hd + 1 || (long)x * y > (1L << 31) - 1)
        return tl_fail(TLHIP_ERR_ARG, "%s: needs even x, even halo_depth >= 2 (x=%d hd=%d)", what, x, hd);
    if (count < 1 || count > kMaxFused)
        return tl_fail(TLHIP_ERR_ARG, "%s: count must be 1..%d (got %d)", what, kMaxFused, count);
    TL_REQUIRE(alphas && betas && neighbours, what, "null coefficients/neighbours");
    TL_REQUIRE((init_from_r || sd_in) && sd_out && r_in && r_out && u_in && u_out && kx && ky, what,
               "null field");
    TL_REQUIRE(sd_in != sd_out && r_in != r_out && u_in != u_out, what,
               "input and output buffers must be distinct");
    for (int f = 0; f < 4; ++f)
        TL_REQUIRE(neighbours[f] == -1, what, "every face must be external (single chunk)");
    TL_REQUIRE(all_aligned(sd_in, sd_out, r_in, r_out, u_in, u_out, kx, ky), what,
               "fields must be 16-byte aligned");
    if (!init_from_r && theta == 0.0) theta = 1.0;
    Ring3 rg{};
    rg.src[0] = r_in;  rg.dst[0] = r_out;
    rg.src[1] = u_in;  rg.dst[1] = u_out;
    rg.n = 2;
    if (sd_in) {
        rg.src[2] = sd_in;
        rg.dst[2] = sd_out;
        rg.n = 3;
    }
    const long ring = (2L * hd * x + 2L * hd * (y - 2 * hd)) * rg.n;
    hipLaunchKernelGGL(k_copy_rings2d, dim3(flat_blocks(ring)), dim3(kBlock), 0, tl_stream(), g, rg);
    TL_TRY(tl_check_launch(what));
    if (init_from_r)
        return dispatch_fused<true>(count, g, alphas, betas, theta, sd_in, sd_out, r_in, r_out,
                                    u_in, u_out, kx, ky, what);
    return dispatch_fused<false>(count, g, alphas, betas, theta, sd_in, sd_out, r_in, r_out, u_in,
                                 u_out, kx, ky, what);
}
