// 2D kernels of the hip_cdna4 suite and their C-ABI entry points.
//
// Each tlhip2d_* function replaces one run_* entry point of the reference
// kernel interface (2d/kernel_interface.h); argument meaning follows the
// omp3 binding of that entry point (2d/c_kernels/omp3/c_kernels.h) with
// device pointers for fields and host pointers for scalar results.
//
// Two kernel skeletons carry all the work:
//  * k_stencil2d: the 5-point operator.  A wave owns a strip of 64*V
//    columns (V cells per lane, V=2 -> one 16-byte load per lane per row) and
//    marches up a tile of rows keeping rows j-1, j, j+1 of the stencil field
//    and rows j, j+1 of ky in registers, so every field element is read from
//    HBM once per sweep.  x-neighbours come from the adjacent lane by DPP;
//    only the two end lanes of a wave touch memory for them.  Tiles are
//    handed out XCD-contiguously so vertically adjacent tiles share an L2.
//  * k_interior2d: pointwise work over interior rows in 16-byte chunks.
// Vector (V=2) kernels are used when rows and pointers are 16-byte aligned
// and halo_depth >= 2; otherwise the same code runs with V=1.
#include "tl_ops.h"
#include "tl_loop.h"
#include "tealeaf_hip.h"

#include <stdint.h>
#include <stdlib.h>

int tl_staging(size_t doubles, double** out);

using namespace tl;

namespace {

// Stencil neighbourhood of the V cells a lane owns in row j.
template <int V>
struct Nb2 {
    DV<V> c, up, dn, l, r;      // field at (j), (j+1), (j-1), left, right
    DV<V> kxc, kxr, kyc, kyu;   // kx[i], kx[i+1], ky[i], ky[i+x]
};

// One row of the operator, association as 2d/shared.h:50-54 (SMVP).
template <int V>
__device__ __forceinline__ double apply_a(const Nb2<V>& n, int v)
{
    const double diag = (1.0 + (n.kxr.v[v] + n.kxc.v[v])) + (n.kyu.v[v] + n.kyc.v[v]);
    return diag * n.c.v[v] - (n.kxr.v[v] * n.r.v[v] + n.kxc.v[v] * n.l.v[v])
                           - (n.kyu.v[v] * n.up.v[v] + n.kyc.v[v] * n.dn.v[v]);
}

struct Tile2 {
    int ns;    // strips along x (64*V columns each)
    int nbx;   // blocks along x (one strip each)
    int nby;   // blocks along y
    int ty;    // rows per block (a multiple of kWaves)
};

// Blocks are dealt to the 8 XCDs round-robin; hand each XCD a contiguous
// run of tiles (a horizontal band) so the rows shared by vertically adjacent
// tiles are found in that XCD's L2.  A bijection on [0, nb).
__device__ __forceinline__ int xcd_tile(int bid, int nb)
{
    const int q = nb >> 3, r = nb & 7, k = bid & 7, i = bid >> 3;
    return k * q + (k < r ? k : r) + i;
}

// Stencil sweep: a block is kWaves consecutive rows of one strip of 64*V
// columns, a wave one row, a lane V cells (one 16-byte load per row when
// V = 2).  Each lane reads rows j-1, j, j+1 of the stencil field, ky at j
// and j+1, kx at j and the op's own inputs directly; the rows that adjacent
// waves and blocks share are served by L2 (blocks are dealt XCD-contiguously,
// so vertically adjacent blocks share an XCD's L2).  x-neighbours come from
// the adjacent lane by DPP; only the wave's end lanes read them from memory.
// On MI355X this beats marching a wave up a tile with rows held in
// registers: more independent rows in flight, more bytes in flight.  Halo
// values are read as stored, as the reference reads them.
template <int V, class Op>
__global__ void __launch_bounds__(kBlock)
k_stencil2d(TlGrid2 g, Tile2 t, const double* __restrict__ a,
            const double* __restrict__ kx, const double* __restrict__ ky, Op op, TlReduce rd,
            TlGate gate)
{
    if (gate.closed()) return;
    const int lane = threadIdx.x & (TL_WAVE - 1), wave = threadIdx.x / TL_WAVE;
    const int tile = xcd_tile(blockIdx.x, t.nbx * t.nby);
    const int sx = tile % t.nbx, by = tile / t.nbx;
    const int xend = g.x - g.hd;
    const int c0 = g.hd + sx * (TL_WAVE * V) + lane * V;
    int nval = xend - c0;
    nval = nval < 0 ? 0 : (nval > V ? V : nval);
    const bool ld_ok = c0 + V <= g.x;        // whole vector inside the row
    const bool edge_r = lane == TL_WAVE - 1 && c0 + V < g.x;
    const long X = g.x;
    constexpr int NR = Op::NRED > 0 ? Op::NRED : 1;
    double acc[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) acc[q] = 0.0;

    const int jend = min(g.hd + (by + 1) * t.ty, g.y - g.hd);
    for (int j = g.hd + by * t.ty + wave; j < jend; j += kWaves) {   // uniform per wave
        const long row = (long)j * X;
        Nb2<V> n;
        In4<V> in{};
        double el = 0.0, er = 0.0, ek = 0.0;
        if (ld_ok) {
            n.c = ldv<V>(a, row + c0);
            n.up = ldv<V>(a, row + X + c0);
            n.dn = ldv<V>(a, row - X + c0);
            n.kxc = ldv<V>(kx, row + c0);
            n.kyc = ldv<V>(ky, row + c0);
            n.kyu = ldv<V>(ky, row + X + c0);
            in = op.template load<V>(row + c0);
            if (lane == 0) el = a[row + c0 - 1];
        } else {
            n.c = n.up = n.dn = n.kxc = n.kyc = n.kyu = zdv<V>();
        }
        if (edge_r) {
            er = a[row + c0 + V];
            ek = kx[row + c0 + V];
        }
        const double al = tl_lane_prev(n.c.v[V - 1], el);
        const double ar = tl_lane_next(n.c.v[0], er);
        const double kr = tl_lane_next(n.kxc.v[0], ek);
#pragma unroll
        for (int v = 0; v < V; ++v) {
            n.l.v[v] = v == 0 ? al : n.c.v[v - 1];
            n.r.v[v] = v == V - 1 ? ar : n.c.v[v + 1];
            n.kxr.v[v] = v == V - 1 ? kr : n.kxc.v[v + 1];
        }
        if (nval > 0) op.template apply<V>(row + c0, nval, n, in, acc);
    }
    if constexpr (Op::NRED > 0) tl_reduce_partial<Op::NRED, kBlock>(acc, rd);
}

// Jacobi update from the copy r of u (jacobi.c:61-100); the stencil field
// is r.  err += |u_new - r|.
struct OpJacobi2 {
    static constexpr int NRED = 1;
    const double* __restrict__ u0;
    double* __restrict__ u;
    template <int V>
    __device__ In4<V> load(long i) const
    {
        In4<V> in{};
        in.a = ldv<V>(u0, i);
        return in;
    }
    template <int V, class N>
    __device__ void apply(long i, int nval, const N& n, const In4<V>& in, double* acc) const
    {
        const DV<V>& b = in.a;
        DV<V> uo;
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const double num = (b.v[v] + (n.kxr.v[v] * n.r.v[v] + n.kxc.v[v] * n.l.v[v]))
                             + (n.kyu.v[v] * n.up.v[v] + n.kyc.v[v] * n.dn.v[v]);
            const double den = (1.0 + (n.kxc.v[v] + n.kxr.v[v])) + (n.kyc.v[v] + n.kyu.v[v]);
            uo.v[v] = num / den;
            if (v < nval) acc[0] += fabs(uo.v[v] - n.c.v[v]);
        }
        stv<V>(u, i, uo, nval);
    }
};

// ------------------------------------------------------------ whole-array kernels

// cg_init passes 1-3 in one sweep over the whole array (cg.c:31-66):
// p = r = 0, u = e*rho everywhere; w = rho or 1/rho on [1,n-1); kx, ky on
// [hd,n-1) from the same w values recomputed from rho.  A neighbour outside
// [1,n-1) (only with halo_depth 1) is not written by this pass, so its w is
// read from memory as the reference reads it.
__global__ void __launch_bounds__(kBlock)
k_cg_init_fields(TlGrid2 g, int coef, double rx, double ry, const double* __restrict__ density,
                 const double* __restrict__ energy, double* __restrict__ u, double* __restrict__ p,
                 double* __restrict__ r, double* __restrict__ w, double* __restrict__ kx,
                 double* __restrict__ ky)
{
    const long n = (long)g.x * g.y;
    for (long c = blockIdx.x * (long)kBlock + threadIdx.x; c < n; c += (long)gridDim.x * kBlock) {
        const int jj = (int)(c / g.x), kk = (int)(c % g.x);
        const double rho = density[c];
        p[c] = 0.0;
        r[c] = 0.0;
        u[c] = energy[c] * rho;
        if (jj >= 1 && jj < g.y - 1 && kk >= 1 && kk < g.x - 1) {
            const double wc = coef == 1 ? rho : 1.0 / rho;
            w[c] = wc;
            if (jj >= g.hd && kk >= g.hd) {
                const double wl = kk >= 2 ? (coef == 1 ? density[c - 1] : 1.0 / density[c - 1])
                                          : w[c - 1];
                const double wd = jj >= 2 ? (coef == 1 ? density[c - g.x] : 1.0 / density[c - g.x])
                                          : w[c - g.x];
                kx[c] = tl_face_coef(rx, wl, wc);
                ky[c] = tl_face_coef(ry, wd, wc);
            }
        }
    }
}

// jacobi_init (jacobi.c:10-58): u0 = u = e*rho on [1,n-1); kx, ky on
// [hd,n-1); any coefficient other than 1 takes the reciprocal branch.
__global__ void __launch_bounds__(kBlock)
k_jacobi_init(TlGrid2 g, int coef, double rx, double ry, const double* __restrict__ density,
              const double* __restrict__ energy, double* __restrict__ u0, double* __restrict__ u,
              double* __restrict__ kx, double* __restrict__ ky)
{
    const long n = (long)g.x * g.y;
    for (long c = blockIdx.x * (long)kBlock + threadIdx.x; c < n; c += (long)gridDim.x * kBlock) {
        const int jj = (int)(c / g.x), kk = (int)(c % g.x);
        if (jj < 1 || jj >= g.y - 1 || kk < 1 || kk >= g.x - 1) continue;
        const double t = energy[c] * density[c];
        u0[c] = t;
        u[c] = t;
        if (jj >= g.hd && kk >= g.hd) {
            const bool cond = coef == 1;
            const double wc = cond ? density[c] : 1.0 / density[c];
            const double wl = cond ? density[c - 1] : 1.0 / density[c - 1];
            const double wd = cond ? density[c - g.x] : 1.0 / density[c - g.x];
            kx[c] = rx * (wl + wc) / (2.0 * wl * wc);
            ky[c] = ry * (wd + wc) / (2.0 * wd * wc);
        }
    }
}

// Copy the halo ring (every cell outside the interior) from src to dst; used
// after a ping-pong sweep so the output buffer equals the in-place result.
__global__ void __launch_bounds__(kBlock)
k_copy_ring2d(TlGrid2 g, const double* __restrict__ src, double* __restrict__ dst)
{
    const long nrow = 2L * g.hd * g.x;                    // hd rows at bottom and top
    const long ncol = 2L * g.hd * (g.y - 2 * g.hd);       // hd columns each side
    for (long t = blockIdx.x * (long)kBlock + threadIdx.x; t < nrow + ncol;
         t += (long)gridDim.x * kBlock) {
        long c;
        if (t < nrow) {
            const int rr = (int)(t / g.x), kk = (int)(t % g.x);
            const int jj = rr < g.hd ? rr : g.y - 2 * g.hd + rr;
            c = (long)jj * g.x + kk;
        } else {
            const long u = t - nrow;
            const int jj = g.hd + (int)(u / (2 * g.hd));
            const int q = (int)(u % (2 * g.hd));
            const int kk = q < g.hd ? q : g.x - 2 * g.hd + q;
            c = (long)jj * g.x + kk;
        }
        dst[c] = src[c];
    }
}

// set_chunk_data (set_chunk_data.c:10-52): vertices, centres, volume, areas.
__global__ void __launch_bounds__(kBlock)
k_set_chunk_data(TlGrid2 g, double x0, double y0, double dx, double dy, double* cell_x,
                 double* cell_y, double* vertex_x, double* vertex_y, double* volume,
                 double* x_area, double* y_area)
{
    const long n = (long)g.x * g.y;
    const long tot = n + (g.x + 1) + (g.y + 1);
    for (long t = blockIdx.x * (long)kBlock + threadIdx.x; t < tot; t += (long)gridDim.x * kBlock) {
        if (t < n) {
            volume[t] = dx * dy;
            x_area[t] = dy;
            y_area[t] = dx;
        } else if (t < n + g.x + 1) {
            const int i = (int)(t - n);
            const double vx = x0 + dx * (i - g.hd);
            vertex_x[i] = vx;
            if (i < g.x) cell_x[i] = 0.5 * (vx + (x0 + dx * (i + 1 - g.hd)));
        } else {
            const int j = (int)(t - n - g.x - 1);
            const double vy = y0 + dy * (j - g.hd);
            vertex_y[j] = vy;
            if (j < g.y) cell_y[j] = 0.5 * (vy + (y0 + dy * (j + 1 - g.hd)));
        }
    }
}

// set_chunk_state (set_chunk_state.c:10-84): the last state (in order)
// covering a cell wins, which equals applying the states in turn.
__global__ void __launch_bounds__(kBlock)
k_set_chunk_state(TlGrid2 g, const double* __restrict__ vertex_x, const double* __restrict__ vertex_y,
                  const double* __restrict__ cell_x, const double* __restrict__ cell_y,
                  double* __restrict__ density, double* __restrict__ energy0, double* __restrict__ u,
                  int nstates, const tlhip_state* __restrict__ st)
{
    const long n = (long)g.x * g.y;
    for (long c = blockIdx.x * (long)kBlock + threadIdx.x; c < n; c += (long)gridDim.x * kBlock) {
        const int jj = (int)(c / g.x), kk = (int)(c % g.x);
        double e = st[0].energy, d = st[0].density;
        for (int s = 1; s < nstates; ++s) {
            bool hit = false;
            if (st[s].geometry == 0) {
                hit = vertex_x[kk + 1] >= st[s].x_min && vertex_x[kk] < st[s].x_max &&
                      vertex_y[jj + 1] >= st[s].y_min && vertex_y[jj] < st[s].y_max;
            } else if (st[s].geometry == 1) {
                const double ddx = cell_x[kk] - st[s].x_min;
                const double ddy = cell_y[jj] - st[s].y_min;
                hit = sqrt(ddx * ddx + ddy * ddy) <= st[s].radius;
            } else if (st[s].geometry == 2) {
                hit = vertex_x[kk] == st[s].x_min && vertex_y[jj] == st[s].y_min;
            }
            if (hit) {
                e = st[s].energy;
                d = st[s].density;
            }
        }
        energy0[c] = e;
        density[c] = d;
        if (jj >= 1 && jj < g.y - 1 && kk >= 1 && kk < g.x - 1) u[c] = e * d;
    }
}

// All flagged fields, all external faces, one launch (local_halos.c:22-149).
struct HaloArgs2 {
    double* f[6];
    int nf;
    int faces;   // bit 0 left, 1 right, 2 bottom, 3 top
    int depth;
};

__global__ void __launch_bounds__(kBlock) k_local_halos2d(TlGrid2 g, HaloArgs2 h, TlGate gate)
{
    if (gate.closed()) return;
    double* f = h.f[blockIdx.y];
    const int ny = g.y - 2 * g.hd, nx = g.x - 2 * g.hd;
    const long nlr = (long)ny * h.depth, ntb = (long)nx * h.depth;
    const long tot = 4 * (nlr > ntb ? nlr : ntb);
    for (long t = blockIdx.x * (long)kBlock + threadIdx.x; t < tot; t += (long)gridDim.x * kBlock) {
        const int face = (int)(t & 3);
        const long u = t >> 2;
        if (!(h.faces & (1 << face))) continue;
        if (face < 2) {
            if (u >= nlr) continue;
            const int jj = g.hd + (int)(u / h.depth), k = (int)(u % h.depth);
            const long b = (long)jj * g.x;
            if (face == 0) f[b + (g.hd - k - 1)] = f[b + (g.hd + k)];
            else f[b + (g.x - g.hd + k)] = f[b + (g.x - g.hd - 1 - k)];
        } else {
            if (u >= ntb) continue;
            const int k = (int)(u / nx), kk = g.hd + (int)(u % nx);
            if (face == 2) f[kk + (long)(g.hd - k - 1) * g.x] = f[kk + (long)(g.hd + k) * g.x];
            else f[kk + (long)(g.y - g.hd + k) * g.x] = f[kk + (long)(g.y - g.hd - 1 - k) * g.x];
        }
    }
}

// Face box <-> contiguous buffer (pack_halos.c:5-173), x fastest.
__global__ void __launch_bounds__(kBlock)
k_box2d(TlGrid2 g, int k0, int j0, int ex, int ey, int pack, double* __restrict__ field,
        double* __restrict__ buf)
{
    const long tot = (long)ex * ey;
    for (long t = blockIdx.x * (long)kBlock + threadIdx.x; t < tot; t += (long)gridDim.x * kBlock) {
        const int dk = (int)(t % ex), dj = (int)(t / ex);
        const long c = (k0 + dk) + (long)(j0 + dj) * g.x;
        if (pack) buf[t] = field[c];
        else field[c] = buf[t];
    }
}

// ------------------------------------------------------------ host side

int g_ty_override = -1;

int check_grid(const TlGrid2& g, const char* what)
{
    if (g.hd < 1 || g.x < 2 * g.hd || g.y < 2 * g.hd)
        return tl_fail(TLHIP_ERR_ARG, "%s: invalid grid x=%d y=%d halo_depth=%d", what, g.x, g.y, g.hd);
    if ((long)g.x * g.y > (1L << 31) - 1)
        return tl_fail(TLHIP_ERR_ARG, "%s: grid too large (%d x %d)", what, g.x, g.y);
    return 0;
}

TlRows rows_of(const TlGrid2& g)
{
    return TlRows{g.x, g.hd, g.y - 2 * g.hd, 0L, 0, (long)(g.y - 2 * g.hd)};
}

// Vector width usable for these fields on this grid.
template <typename... P>
int pick_v(const TlGrid2& g, P... ps)
{
    return (g.x % 2 == 0 && g.hd % 2 == 0 && g.hd >= 2 && all_aligned(ps...)) ? 2 : 1;
}

// One block per t.ty rows of one strip (a multiple of kWaves, each wave
// every kWaves-th row): 16 by default, which at 4096^2 leaves a reducing
// sweep 8192 partials instead of 32768 (its finisher 4.5 us instead of 10)
// and measured 1-2% faster per CG iteration than 4; rows per block grow
// further only to keep the block count within the reduction slots.
Tile2 make_tile(const TlGrid2& g, int V)
{
    Tile2 t;
    const int nx = g.x - 2 * g.hd, ny = g.y - 2 * g.hd;
    t.ns = (int)tl_cdiv(nx, TL_WAVE * V);
    t.nbx = t.ns;
    int ty = g_ty_override > 0 ? (int)tl_cdiv(g_ty_override, kWaves) * kWaves : 4 * kWaves;
    while ((long)t.nbx * tl_cdiv(ny, ty) > TL_MAX_RED_BLOCKS) ty *= 2;
    t.ty = ty;
    t.nby = (int)tl_cdiv(ny, ty);
    return t;
}

// Under the current launch gate; with nb_out, a reducing op's partials are
// left for the caller's finisher (see launch_interior).
template <class Op>
int launch_stencil(const TlGrid2& g, int V, const double* a, const double* kx,
                   const double* ky, const Op& op, const char* what, unsigned* nb_out = nullptr)
{
    if (g.x - 2 * g.hd <= 0 || g.y - 2 * g.hd <= 0) return 0;
    if (tl_cdiv(g.x - 2 * g.hd, TL_WAVE * V) > TL_MAX_RED_BLOCKS)
        return tl_fail(TLHIP_ERR_ARG, "%s: rows too long (%d)", what, g.x);
    const Tile2 t = make_tile(g, V);
    const unsigned nb = (unsigned)(t.nbx * t.nby);
    const int slot = tl_timer_begin(what);
    if (V == 2)
        hipLaunchKernelGGL((k_stencil2d<2, Op>), dim3(nb), dim3(kBlock), 0, tl_stream(), g, t, a,
                           kx, ky, op, tl_reduce_slots(), tl_gate());
    else
        hipLaunchKernelGGL((k_stencil2d<1, Op>), dim3(nb), dim3(kBlock), 0, tl_stream(), g, t, a,
                           kx, ky, op, tl_reduce_slots(), tl_gate());
    tl_timer_end(slot);
    TL_TRY(tl_check_launch(what));
    if (nb_out) {
        *nb_out = nb;
        return 0;
    }
    if constexpr (Op::NRED > 0) return tl_finish_reduce(Op::NRED, nb, what);
    return 0;
}

template <class Op>
int launch_interior2(const TlGrid2& g, int V, const Op& op, const char* what,
                     unsigned* nb_out = nullptr)
{
    return launch_interior(rows_of(g), g.x - 2 * g.hd, V, op, what, nb_out);
}

bool interior_empty(const TlGrid2& g) { return g.x - 2 * g.hd <= 0 || g.y - 2 * g.hd <= 0; }

#define REQUIRE TL_REQUIRE

}  // namespace

int tl_tile_rows_override() { return g_ty_override; }

namespace {
int g_interior_blocks = 4096;
}
int tl_interior_blocks() { return g_interior_blocks; }

extern "C" int tlhip_set_interior_blocks(int blocks)
{
    if (blocks > TL_MAX_RED_BLOCKS)
        return tl_fail(TLHIP_ERR_ARG, "tlhip_set_interior_blocks: at most %d", TL_MAX_RED_BLOCKS);
    g_interior_blocks = blocks > 0 ? blocks : 4096;
    return 0;
}

extern "C" int tlhip_set_tile_rows(int rows)
{
    g_ty_override = rows;
    return 0;
}

// ---------------------------------------------------------------- init

extern "C" int tlhip2d_set_chunk_data(int x, int y, int hd, int left, int bottom,
                                      double grid_x_min, double grid_y_min, double dx, double dy,
                                      double* cell_x, double* cell_y, double* vertex_x,
                                      double* vertex_y, double* volume, double* x_area,
                                      double* y_area)
{
    const char* what = "tlhip2d_set_chunk_data";
    TL_TRY(tl_ensure_init());
    const TlGrid2 g{x, y, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(cell_x && cell_y && vertex_x && vertex_y && volume && x_area && y_area, what,
            "null field");
    const double x0 = grid_x_min + dx * (double)left;
    const double y0 = grid_y_min + dy * (double)bottom;
    const long tot = (long)x * y + x + y + 2;
    hipLaunchKernelGGL(k_set_chunk_data, dim3(flat_blocks(tot)), dim3(kBlock), 0, tl_stream(), g,
                       x0, y0, dx, dy, cell_x, cell_y, vertex_x, vertex_y, volume, x_area, y_area);
    return tl_check_launch(what);
}

extern "C" int tlhip2d_set_chunk_state(int x, int y, const double* vertex_x,
                                       const double* vertex_y, const double* cell_x,
                                       const double* cell_y, double* density, double* energy0,
                                       double* u, int num_states, const tlhip_state* states)
{
    const char* what = "tlhip2d_set_chunk_state";
    TL_TRY(tl_ensure_init());
    const TlGrid2 g{x, y, 1};
    REQUIRE(x >= 1 && y >= 1, what, "invalid grid");
    REQUIRE(num_states >= 1 && states, what, "need at least one state");
    REQUIRE(vertex_x && vertex_y && cell_x && cell_y && density && energy0 && u, what,
            "null field");
    void* dst = nullptr;
    TL_TRY(tlhip_alloc(sizeof(tlhip_state) * num_states, &dst));
    int rc = tlhip_copy_to_device(dst, states, sizeof(tlhip_state) * num_states);
    if (!rc) {
        hipLaunchKernelGGL(k_set_chunk_state, dim3(flat_blocks((long)x * y)), dim3(kBlock), 0,
                           tl_stream(), g, vertex_x, vertex_y, cell_x, cell_y, density, energy0,
                           u, num_states, (const tlhip_state*)dst);
        rc = tl_check_launch(what);
    }
    int rc2 = tlhip_free(dst);
    return rc ? rc : rc2;
}

// ---------------------------------------------------------------- halos

extern "C" int tlhip2d_local_halos(int x, int y, int depth, int hd, const int* neighbours,
                                   const bool* fields, double* density, double* energy0,
                                   double* energy, double* u, double* p, double* sd)
{
    const char* what = "tlhip2d_local_halos";
    TL_TRY(tl_ensure_init());
    const TlGrid2 g{x, y, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(neighbours && fields, what, "null neighbours/fields");
    REQUIRE(depth >= 0 && depth <= hd, what, "depth must be in [0, halo_depth]");
    HaloArgs2 h{};
    double* all[6] = {density, energy0, energy, u, p, sd};
    for (int i = 0; i < 6; ++i)
        if (fields[i]) {
            REQUIRE(all[i], what, "flagged field is null");
            h.f[h.nf++] = all[i];
        }
    // Face order of the flags: CHUNK_LEFT 0, RIGHT 1, BOTTOM 2, TOP 3.
    for (int f = 0; f < 4; ++f)
        if (neighbours[f] == -1) h.faces |= 1 << f;
    h.depth = depth;
    if (h.nf == 0 || h.faces == 0 || depth == 0 || interior_empty(g)) return 0;
    const long ny = y - 2 * hd, nx = x - 2 * hd;
    const long tot = 4 * (ny > nx ? ny : nx) * depth;
    hipLaunchKernelGGL(k_local_halos2d, dim3(flat_blocks(tot), h.nf), dim3(kBlock), 0,
                       tl_stream(), g, h, tl_gate());
    return tl_check_launch(what);
}

extern "C" int tlhip2d_pack_or_unpack(int x, int y, int depth, int hd, int face, bool pack,
                                      double* field, double* buffer)
{
    const char* what = "tlhip2d_pack_or_unpack";
    TL_TRY(tl_ensure_init());
    const TlGrid2 g{x, y, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(field && buffer, what, "null field/buffer");
    REQUIRE(depth >= 1 && depth <= hd, what, "depth must be in [1, halo_depth]");
    const int xi = x - 2 * hd, yi = y - 2 * hd;
    int k0 = hd, j0 = hd, ex = xi, ey = yi;
    switch (face) {
    case 0: k0 = pack ? hd : hd - depth; ex = depth; break;
    case 1: k0 = pack ? x - hd - depth : x - hd; ex = depth; break;
    case 2: j0 = pack ? hd : hd - depth; ey = depth; break;
    case 3: j0 = pack ? y - hd - depth : y - hd; ey = depth; break;
    default: return tl_fail(TLHIP_ERR_ARG, "%s: Incorrect face provided: %d.", what, face);
    }
    const long n = (long)ex * ey;
    if (n <= 0) return 0;
    double* stage = nullptr;
    TL_TRY(tl_staging((size_t)n, &stage));
    if (!pack) TL_TRY(tlhip_copy_to_device(stage, buffer, sizeof(double) * n));
    hipLaunchKernelGGL(k_box2d, dim3(flat_blocks(n)), dim3(kBlock), 0, tl_stream(), g, k0, j0, ex,
                       ey, pack ? 1 : 0, field, stage);
    TL_TRY(tl_check_launch(what));
    if (pack) TL_TRY(tlhip_copy_to_host(buffer, stage, sizeof(double) * n));
    return 0;
}

extern "C" int tlhip2d_store_energy(int x, int y, const double* energy0, double* energy)
{
    const char* what = "tlhip2d_store_energy";
    TL_TRY(tl_ensure_init());
    REQUIRE(x >= 1 && y >= 1 && energy0 && energy, what, "invalid arguments");
    const long n = (long)x * y;
    hipLaunchKernelGGL(k_copy_all, dim3(flat_blocks(n)), dim3(kBlock), 0, tl_stream(), n, energy0,
                       energy);
    return tl_check_launch(what);
}

extern "C" int tlhip2d_field_summary(int x, int y, int hd, const double* volume,
                                     const double* density, const double* energy0,
                                     const double* u, double* vol, double* mass, double* ie,
                                     double* temp)
{
    const char* what = "tlhip2d_field_summary";
    TL_TRY(tl_ensure_init());
    const TlGrid2 g{x, y, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(volume && density && energy0 && u && vol && mass && ie && temp, what, "null argument");
    if (interior_empty(g)) return 0;
    OpFieldSummary op{volume, density, energy0, u};
    TL_TRY(launch_interior2(g, pick_v(g, volume, density, energy0, u), op, what));
    double s[4];
    TL_TRY(tl_fetch(4, s, what));
    *vol += s[0];
    *mass += s[1];
    *ie += s[2];
    *temp += s[3];
    return 0;
}

// ---------------------------------------------------------------- CG

extern "C" int tlhip2d_cg_init(int x, int y, int hd, int coefficient, double rx, double ry,
                               double* rro, const double* density, const double* energy,
                               double* u, double* p, double* r, double* w, double* kx, double* ky)
{
    const char* what = "tlhip2d_cg_init";
    TL_TRY(tl_ensure_init());
    const TlGrid2 g{x, y, hd};
    TL_TRY(check_grid(g, what));
    if (coefficient != 1 && coefficient != 2)
        return tl_fail(TLHIP_ERR_ARG, "Coefficient %d is not valid.", coefficient);
    REQUIRE(rro && density && energy && u && p && r && w && kx && ky, what, "null argument");
    const long n = (long)x * y;
    hipLaunchKernelGGL(k_cg_init_fields, dim3(flat_blocks(n)), dim3(kBlock), 0, tl_stream(), g,
                       coefficient, rx, ry, density, energy, u, p, r, w, kx, ky);
    TL_TRY(tl_check_launch(what));
    if (interior_empty(g)) return 0;
    OpCgInit op{w, r, p};
    TL_TRY(launch_stencil(g, pick_v(g, u, kx, ky, w, r, p), u, kx, ky, op, what));
    return fetch_add(rro, what);
}

extern "C" int tlhip2d_cg_calc_w(int x, int y, int hd, double* pw, const double* p, double* w,
                                 const double* kx, const double* ky)
{
    const char* what = "tlhip2d_cg_calc_w";
    TL_TRY(tl_ensure_init());
    const TlGrid2 g{x, y, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(pw && p && w && kx && ky, what, "null argument");
    REQUIRE(p != w, what, "p and w must be distinct buffers");
    if (interior_empty(g)) return 0;
    OpCalcW op{w};
    TL_TRY(launch_stencil(g, pick_v(g, p, w, kx, ky), p, kx, ky, op, what));
    return fetch_add(pw, what);
}

extern "C" int tlhip2d_cg_calc_ur(int x, int y, int hd, double alpha, double* rrn, double* u,
                                  const double* p, double* r, const double* w)
{
    const char* what = "tlhip2d_cg_calc_ur";
    TL_TRY(tl_ensure_init());
    const TlGrid2 g{x, y, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(rrn && u && p && r && w, what, "null argument");
    if (interior_empty(g)) return 0;
    OpCalcUr op{u, p, r, w, alpha};
    TL_TRY(launch_interior2(g, pick_v(g, u, p, r, w), op, what));
    return fetch_add(rrn, what);
}

extern "C" int tlhip2d_cg_calc_p(int x, int y, int hd, double beta, double* p, const double* r)
{
    const char* what = "tlhip2d_cg_calc_p";
    TL_TRY(tl_ensure_init());
    const TlGrid2 g{x, y, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(p && r, what, "null argument");
    OpCalcP op{p, r, beta};
    return launch_interior2(g, pick_v(g, p, r), op, what);
}

// ---------------------------------------------------------------- Chebyshev

extern "C" int tlhip2d_cheby_init(int x, int y, int hd, double theta, double* u,
                                  const double* u0, double* p, double* r, double* w,
                                  const double* kx, const double* ky)
{
    const char* what = "tlhip2d_cheby_init";
    TL_TRY(tl_ensure_init());
    const TlGrid2 g{x, y, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(u && u0 && p && r && w && kx && ky, what, "null argument");
    OpChebyInit op{u0, w, r, p, theta};
    TL_TRY(launch_stencil(g, pick_v(g, u, u0, p, r, w, kx, ky), u, kx, ky, op, what));
    OpSimple up{u, p, 1, nullptr, 0.0};
    return launch_interior2(g, pick_v(g, u, p), up, what);
}

extern "C" int tlhip2d_cheby_iterate(int x, int y, int hd, double alpha, double beta,
                                     const double* u_in, double* u_out, const double* u0,
                                     double* p, double* r, double* w, const double* kx,
                                     const double* ky)
{
    const char* what = "tlhip2d_cheby_iterate";
    TL_TRY(tl_ensure_init());
    const TlGrid2 g{x, y, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(u_in && u_out && u0 && p && r && kx && ky, what, "null argument");
    REQUIRE(u_in != u_out, what, "u_in and u_out must be distinct buffers");
    OpChebyIter op{u0, p, r, w, u_out, alpha, beta};
    TL_TRY(launch_stencil(g, pick_v(g, u_in, u_out, u0, p, r, w, kx, ky), u_in, kx, ky, op, what));
    const long ring = 2L * hd * x + 2L * hd * (y - 2 * hd);
    hipLaunchKernelGGL(k_copy_ring2d, dim3(flat_blocks(ring)), dim3(kBlock), 0, tl_stream(), g,
                       u_in, u_out);
    return tl_check_launch(what);
}

// ---------------------------------------------------------------- Jacobi

extern "C" int tlhip2d_jacobi_init(int x, int y, int hd, int coefficient, double rx, double ry,
                                   const double* density, const double* energy, double* u0,
                                   double* u, double* kx, double* ky)
{
    const char* what = "tlhip2d_jacobi_init";
    TL_TRY(tl_ensure_init());
    const TlGrid2 g{x, y, hd};
    TL_TRY(check_grid(g, what));
    if (coefficient < 1) return tl_fail(TLHIP_ERR_ARG, "Coefficient %d is not valid.", coefficient);
    REQUIRE(density && energy && u0 && u && kx && ky, what, "null argument");
    hipLaunchKernelGGL(k_jacobi_init, dim3(flat_blocks((long)x * y)), dim3(kBlock), 0,
                       tl_stream(), g, coefficient, rx, ry, density, energy, u0, u, kx, ky);
    return tl_check_launch(what);
}

extern "C" int tlhip2d_jacobi_iterate(int x, int y, int hd, double* error, const double* kx,
                                      const double* ky, const double* u0, double* u, double* r)
{
    const char* what = "tlhip2d_jacobi_iterate";
    TL_TRY(tl_ensure_init());
    const TlGrid2 g{x, y, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(error && kx && ky && u0 && u && r, what, "null argument");
    REQUIRE(u != r, what, "u and r must be distinct buffers");
    const long n = (long)x * y;
    hipLaunchKernelGGL(k_copy_all, dim3(flat_blocks(n)), dim3(kBlock), 0, tl_stream(), n, u, r);
    TL_TRY(tl_check_launch(what));
    if (interior_empty(g)) {
        *error = 0.0;
        return 0;
    }
    OpJacobi2 op{u0, u};
    TL_TRY(launch_stencil(g, pick_v(g, kx, ky, u0, u, r), r, kx, ky, op, what));
    double v = 0.0;
    TL_TRY(tl_fetch(1, &v, what));
    *error = v;    // assigned, as jacobi.c:99
    return 0;
}

// ---------------------------------------------------------------- PPCG

extern "C" int tlhip2d_ppcg_init(int x, int y, int hd, double theta, const double* r, double* sd)
{
    const char* what = "tlhip2d_ppcg_init";
    TL_TRY(tl_ensure_init());
    const TlGrid2 g{x, y, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(r && sd, what, "null argument");
    OpSimple op{sd, r, 2, nullptr, theta};
    return launch_interior2(g, pick_v(g, r, sd), op, what);
}

extern "C" int tlhip2d_ppcg_inner_iteration(int x, int y, int hd, double alpha, double beta,
                                            double* u, double* r, const double* kx,
                                            const double* ky, const double* sd_in,
                                            double* sd_out)
{
    const char* what = "tlhip2d_ppcg_inner_iteration";
    TL_TRY(tl_ensure_init());
    const TlGrid2 g{x, y, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(u && r && kx && ky && sd_in && sd_out, what, "null argument");
    REQUIRE(sd_in != sd_out, what, "sd_in and sd_out must be distinct buffers");
    OpPpcgInner op{r, u, sd_out, alpha, beta};
    TL_TRY(launch_stencil(g, pick_v(g, u, r, kx, ky, sd_in, sd_out), sd_in, kx, ky, op, what));
    const long ring = 2L * hd * x + 2L * hd * (y - 2 * hd);
    hipLaunchKernelGGL(k_copy_ring2d, dim3(flat_blocks(ring)), dim3(kBlock), 0, tl_stream(), g,
                       sd_in, sd_out);
    return tl_check_launch(what);
}

// ---------------------------------------------------------------- shared

extern "C" int tlhip2d_copy_u(int x, int y, int hd, double* u0, const double* u)
{
    const char* what = "tlhip2d_copy_u";
    TL_TRY(tl_ensure_init());
    const TlGrid2 g{x, y, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(u0 && u, what, "null argument");
    OpSimple op{u0, u, 0, nullptr, 0.0};
    return launch_interior2(g, pick_v(g, u0, u), op, what);
}

extern "C" int tlhip2d_calculate_residual(int x, int y, int hd, const double* u,
                                          const double* u0, double* r, const double* kx,
                                          const double* ky)
{
    const char* what = "tlhip2d_calculate_residual";
    TL_TRY(tl_ensure_init());
    const TlGrid2 g{x, y, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(u && u0 && r && kx && ky, what, "null argument");
    REQUIRE(u != r, what, "u and r must be distinct buffers");
    OpResidual op{u0, r};
    return launch_stencil(g, pick_v(g, u, u0, r, kx, ky), u, kx, ky, op, what);
}

extern "C" int tlhip2d_calculate_2norm(int x, int y, int hd, const double* buffer, double* norm)
{
    const char* what = "tlhip2d_calculate_2norm";
    TL_TRY(tl_ensure_init());
    const TlGrid2 g{x, y, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(buffer && norm, what, "null argument");
    if (interior_empty(g)) return 0;
    OpNorm op{buffer};
    TL_TRY(launch_interior2(g, pick_v(g, buffer), op, what));
    return fetch_add(norm, what);
}

extern "C" int tlhip2d_finalise(int x, int y, int hd, double* energy, const double* density,
                                const double* u)
{
    const char* what = "tlhip2d_finalise";
    TL_TRY(tl_ensure_init());
    const TlGrid2 g{x, y, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(energy && density && u, what, "null argument");
    OpSimple op{energy, u, 3, density, 0.0};
    return launch_interior2(g, pick_v(g, energy, density, u), op, what);
}

// ---------------------------------------------------------------- resident loop
// The CG pieces of one iteration of a device-resident loop (tl_loop.h):
// the same kernels as the entry points above, gated, with alpha and beta
// taken from the loop's device state and the reductions finished by the
// loop's finishers.

namespace tl {
namespace loop2d {

int calc_w(int x, int y, int hd, const double* p, double* w, const double* kx, const double* ky)
{
    const char* what = "tlhip2d_cg_calc_w";
    const TlGrid2 g{x, y, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(p && w && kx && ky && p != w, what, "bad arguments");
    if (interior_empty(g)) return tl_fail(TLHIP_ERR_ARG, "%s: empty interior", what);
    unsigned nb = 0;
    TL_TRY(launch_stencil(g, pick_v(g, p, w, kx, ky), p, kx, ky, OpCalcW{w}, what, &nb));
    return loop::finish_alpha(nb, what);
}

int calc_ur(int x, int y, int hd, double* u, const double* p, double* r, const double* w,
            const loop::StopRule& rule)
{
    const char* what = "tlhip2d_cg_calc_ur";
    const TlGrid2 g{x, y, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(u && p && r && w, what, "null argument");
    if (interior_empty(g)) return tl_fail(TLHIP_ERR_ARG, "%s: empty interior", what);
    OpCalcUr op{u, p, r, w, 0.0, loop::alpha()};
    unsigned nb = 0;
    TL_TRY(launch_interior2(g, pick_v(g, u, p, r, w), op, what, &nb));
    return loop::finish_beta(nb, rule, what);
}

int calc_p(int x, int y, int hd, double* p, const double* r)
{
    const char* what = "tlhip2d_cg_calc_p";
    const TlGrid2 g{x, y, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(p && r, what, "null argument");
    OpCalcP op{p, r, 0.0, loop::beta()};
    return launch_interior2(g, pick_v(g, p, r), op, what);
}

int norm(int x, int y, int hd, const double* buf, const loop::StopRule& rule)
{
    const char* what = "tlhip2d_calculate_2norm";
    const TlGrid2 g{x, y, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(buf, what, "null argument");
    if (interior_empty(g)) return tl_fail(TLHIP_ERR_ARG, "%s: empty interior", what);
    unsigned nb = 0;
    TL_TRY(launch_interior2(g, pick_v(g, buf), OpNorm{buf}, what, &nb));
    return loop::finish_beta(nb, rule, what);
}

int jacobi(int x, int y, int hd, const double* kx, const double* ky, const double* u0,
           const double* u_in, double* u_out, loop::ErrMode mode, const loop::StopRule& rule)
{
    const char* what = "tlhip2d_jacobi_iterate";
    const TlGrid2 g{x, y, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(kx && ky && u0 && u_in && u_out && u_in != u_out, what, "bad arguments");
    if (interior_empty(g)) return tl_fail(TLHIP_ERR_ARG, "%s: empty interior", what);
    OpJacobi2 op{u0, u_out};
    unsigned nb = 0;
    TL_TRY(launch_stencil(g, pick_v(g, kx, ky, u0, u_in, u_out), u_in, kx, ky, op, what, &nb));
    return loop::finish_err(nb, mode, rule, what);
}

int norm_err(int x, int y, int hd, const double* buf, const loop::StopRule& rule)
{
    const char* what = "tlhip2d_calculate_2norm";
    const TlGrid2 g{x, y, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(buf, what, "null argument");
    if (interior_empty(g)) return tl_fail(TLHIP_ERR_ARG, "%s: empty interior", what);
    unsigned nb = 0;
    TL_TRY(launch_interior2(g, pick_v(g, buf), OpNorm{buf}, what, &nb));
    return loop::finish_err(nb, loop::kErrAddTest, rule, what);
}

}  // namespace loop2d
}  // namespace tl

// ------------------------------------------------- allocation-free cg_calc_w
// tlhip2d_cg_calc_w_noalloc: cg_calc_w (w = A p, pw += sum w*p) for callers
// that cannot allocate -- code under HIP graph capture, tight serving loops.
// The stencil sweep is the same kernel as tlhip2d_cg_calc_w; its per-block
// partial sums go to `scratch`, a file-scope device array, instead of the
// runtime's allocated reduction slots, and a one-block finisher adds the
// total, summed in the same order as the standard finisher, into a device
// double the caller passes.  The call itself allocates nothing: no
// hipMalloc/hipFree, no pinned mailbox, no timer events, no runtime set-up
// (tlhip_init must have run; it records the array's device address).
// Every call on a GPU shares that GPU's array; its contents are the last
// call's partials, each fully rewritten before it is read, so nothing needs
// clearing.  Calls that overlap on one GPU (other streams, other threads)
// race on the array.
namespace {

// One partial per stencil block: make_tile keeps the block count of every
// grid that passes check_grid and the row-length check within
// TL_MAX_RED_BLOCKS (it grows the rows per block), so this holds them all.
constexpr int kNoAllocPartials = TL_MAX_RED_BLOCKS;   // 512 KiB
constexpr int kNoAllocDevices = 64;
static __device__ __align__(256) double scratch[kNoAllocPartials];
double* g_scratch[kNoAllocDevices] = {};    // per device: the array's device address

__global__ void __launch_bounds__(256)
k_finish_into(unsigned nblocks, const double* __restrict__ partials, double* __restrict__ out)
{
    double tot[1];
    tl_sum_partials<1>(nblocks, partials, tot);
    if (threadIdx.x == 0) out[0] += tot[0];
}

}  // namespace

// Called by tlhip_init with `device` current: the array's address on it.
int tl_noalloc_setup(int device)
{
    if (device < 0 || device >= kNoAllocDevices)
        return tl_fail(TLHIP_ERR_ARG, "device %d: at most %d devices", device, kNoAllocDevices);
    if (g_scratch[device]) return 0;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return tl_fail(TLHIP_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch[device] = static_cast<double*>(ptr);
    return 0;
}

extern "C" int tlhip2d_cg_calc_w_noalloc(int x, int y, int hd, double* pw_dev, const double* p,
                                         double* w, const double* kx, const double* ky)
{
    const char* what = "tlhip2d_cg_calc_w_noalloc";
    const int dev = tlhip_device();
    if (dev < 0 || dev >= kNoAllocDevices || !g_scratch[dev])
        return tl_fail(TLHIP_ERR_ARG, "%s: call tlhip_init first (this entry point sets up "
                       "nothing itself)", what);
    const TlGrid2 g{x, y, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(pw_dev && p && w && kx && ky, what, "null argument");
    REQUIRE(p != w, what, "p and w must be distinct buffers");
    if (interior_empty(g)) return 0;
    const int V = pick_v(g, p, w, kx, ky);
    if (tl_cdiv(g.x - 2 * g.hd, TL_WAVE * V) > kNoAllocPartials)
        return tl_fail(TLHIP_ERR_ARG, "%s: rows too long (%d cells): more than %d strips", what,
                       x, kNoAllocPartials);
    const Tile2 t = make_tile(g, V);
    const long nb = (long)t.nbx * t.nby;
    if (nb > kNoAllocPartials)
        return tl_fail(TLHIP_ERR_ARG, "%s: %dx%d needs %ld reduction partials, more than the %d "
                       "its scratch array holds", what, x, y, nb, kNoAllocPartials);
    TlReduce rd{};
    rd.partials = g_scratch[dev];
    const OpCalcW op{w};
    if (V == 2)
        hipLaunchKernelGGL((k_stencil2d<2, OpCalcW>), dim3((unsigned)nb), dim3(kBlock), 0,
                           tl_stream(), g, t, p, kx, ky, op, rd, TlGate{});
    else
        hipLaunchKernelGGL((k_stencil2d<1, OpCalcW>), dim3((unsigned)nb), dim3(kBlock), 0,
                           tl_stream(), g, t, p, kx, ky, op, rd, TlGate{});
    TL_TRY(tl_check_launch(what));
    hipLaunchKernelGGL(k_finish_into, dim3(1), dim3(256), 0, tl_stream(), (unsigned)nb,
                       rd.partials, pw_dev);
    return tl_check_launch(what);
}
