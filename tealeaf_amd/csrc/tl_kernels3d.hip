// 3D kernels of the hip_cdna4 suite and their C-ABI entry points.
//
// Each tlhip3d_* function replaces one run_* entry point of
// 3d/kernel_interface.h; arguments follow the omp3 binding
// (3d/c_kernels/omp3/kernel_interface.c) with device pointers for fields and
// host pointers for scalar results.
//
// k_stencil3d applies the 7-point operator.  A wave owns a strip of 64*V
// x-cells in one y-row and marches along z keeping planes z-1, z, z+1 of the
// stencil field (and kz at z, z+1) in registers; x-neighbours come from the
// adjacent lane by DPP and y-neighbours are read from the cache lines the
// waves of the adjacent rows fetch (a block is 4 consecutive rows of one
// strip, and tiles are dealt out XCD-contiguously).
#include "tl_lanes.h"
#include "tl_ops.h"
#include "tl_loop.h"
#include "tl_mixed.h"
#include "tealeaf_hip.h"

#include <stdint.h>

int tl_staging(size_t doubles, double** out);

using namespace tl;

namespace {

template <int V>
struct Nb3 {
    DV<V> c, up, dn, no, so, l, r;        // k, k+1 (z), k-1, j+1 (y), j-1, x-1, x+1
    DV<V> kxc, kxr, kyc, kyn, kzc, kzu;   // kx[i], kx[i+1], ky[i], ky[i+x], kz[i], kz[i+xy]
};

// One row of the operator, association as 3d/shared.h:52-58 (SMVP).
template <int V>
__device__ __forceinline__ double apply_a(const Nb3<V>& n, int v)
{
    const double diag = ((1.0 + (n.kxr.v[v] + n.kxc.v[v])) + (n.kyn.v[v] + n.kyc.v[v]))
                      + (n.kzu.v[v] + n.kzc.v[v]);
    return diag * n.c.v[v] - (n.kxr.v[v] * n.r.v[v] + n.kxc.v[v] * n.l.v[v])
                           - (n.kyn.v[v] * n.no.v[v] + n.kyc.v[v] * n.so.v[v])
                           - (n.kzu.v[v] * n.up.v[v] + n.kzc.v[v] * n.dn.v[v]);
}

struct Tile3 {
    int ns;    // strips along x
    int nbj;   // blocks along y (kWaves rows each)
    int nbk;   // tiles along z
    int tz;    // planes per tile
};

__device__ __forceinline__ int xcd_tile(int bid, int nb)
{
    const int q = nb >> 3, r = nb & 7, k = bid & 7, i = bid >> 3;
    return k * q + (k < r ? k : r) + i;
}

// Stencil sweep: a block is kWaves consecutive rows of one 64*V-column
// strip in tz consecutive planes; a wave one row, a lane V cells.  Marching
// along z, a lane keeps planes k-1 and k of the field and kz at k from the
// previous plane and reads the field at plane k+1 and rows j-1, j+1 of
// plane k, kz at k+1, kx and ky at k (ky also at row j+1) and the op's
// inputs; what adjacent waves and blocks share is served by L2 (blocks are
// dealt XCD-contiguously, so each XCD sweeps a run of consecutive planes).
// x-neighbours come from the adjacent lane by DPP; only the wave's end lanes
// read them from memory.  Halo values are read as stored.
template <int V, class Op>
__global__ void __launch_bounds__(kBlock)
k_stencil3d(TlGrid3 g, Tile3 t, const double* __restrict__ a, const double* __restrict__ kx,
            const double* __restrict__ ky, const double* __restrict__ kz, Op op, TlReduce rd,
            TlGate gate)
{
    if (gate.closed()) return;
    const int lane = threadIdx.x & (TL_WAVE - 1), wave = threadIdx.x / TL_WAVE;
    const int tile = xcd_tile(blockIdx.x, t.ns * t.nbj * t.nbk);
    const int sx = tile % t.ns;
    const int rest = tile / t.ns;
    const int jb = rest % t.nbj, kb = rest / t.nbj;
    const int j = g.hd + jb * kWaves + wave;
    const int xend = g.x - g.hd;
    const int c0 = g.hd + sx * (TL_WAVE * V) + lane * V;
    int nval = xend - c0;
    nval = nval < 0 ? 0 : (nval > V ? V : nval);
    const bool ld_ok = c0 + V <= g.x;
    const bool edge_r = lane == TL_WAVE - 1 && c0 + V < g.x;
    const long X = g.x, XY = (long)g.x * g.y;
    constexpr int NR = Op::NRED > 0 ? Op::NRED : 1;
    double acc[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) acc[q] = 0.0;

    if (j < g.y - g.hd) {   // uniform across the wave
        const long jrow = (long)j * X + c0;
        const int k0 = g.hd + kb * t.tz;             // < g.z - g.hd: kb < nbk
        const int k1 = min(k0 + t.tz, g.z - g.hd);
        // planes k-1 and k of the field and kz at k, carried along z
        DV<V> a_dn = zdv<V>(), a_c = zdv<V>(), kz_c = zdv<V>();
        if (ld_ok) {
            a_dn = ldv<V>(a, (long)(k0 - 1) * XY + jrow);
            a_c = ldv<V>(a, (long)k0 * XY + jrow);
            kz_c = ldv<V>(kz, (long)k0 * XY + jrow);
        }
        for (int k = k0; k < k1; ++k) {
            const long b = (long)k * XY + jrow;
            Nb3<V> n;
            In4<V> in{};
            double el = 0.0, er = 0.0, ek = 0.0;
            n.c = a_c;
            n.dn = a_dn;
            n.kzc = kz_c;
            if (ld_ok) {
                n.up = ldv<V>(a, b + XY);
                n.no = ldv<V>(a, b + X);
                n.so = ldv<V>(a, b - X);
                n.kxc = ldv<V>(kx, b);
                n.kyc = ldv<V>(ky, b);
                n.kyn = ldv<V>(ky, b + X);
                n.kzu = ldv<V>(kz, b + XY);
                in = op.template load<V>(b);
                if (lane == 0) el = a[b - 1];
            } else {
                n.up = n.no = n.so = zdv<V>();
                n.kxc = n.kyc = n.kyn = n.kzu = zdv<V>();
            }
            a_dn = n.c;
            a_c = n.up;
            kz_c = n.kzu;
            if (edge_r) {
                er = a[b + V];
                ek = kx[b + V];
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
            if (nval > 0) op.template apply<V>(b, nval, n, in, acc);
        }
    }
    if constexpr (Op::NRED > 0) tl_reduce_partial<Op::NRED, kBlock>(acc, rd);
}

// 3D Jacobi update (jacobi_solver.c:72-116).
struct OpJacobi3 {
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
            const double num = ((b.v[v] + (n.kxr.v[v] * n.r.v[v] + n.kxc.v[v] * n.l.v[v]))
                              + (n.kyn.v[v] * n.no.v[v] + n.kyc.v[v] * n.so.v[v]))
                              + (n.kzu.v[v] * n.up.v[v] + n.kzc.v[v] * n.dn.v[v]);
            const double den = ((1.0 + (n.kxc.v[v] + n.kxr.v[v])) + (n.kyc.v[v] + n.kyn.v[v]))
                             + (n.kzc.v[v] + n.kzu.v[v]);
            uo.v[v] = num / den;
            if (v < nval) acc[0] += fabs(uo.v[v] - n.c.v[v]);
        }
        stv<V>(u, i, uo, nval);
    }
};

// ------------------------------------------------------------ whole-array kernels

__device__ __forceinline__ void decode3(long c, const TlGrid3& g, int& kk, int& jj, int& ii)
{
    const long xy = (long)g.x * g.y;
    ii = (int)(c / xy);
    const long rem = c - (long)ii * xy;
    jj = (int)(rem / g.x);
    kk = (int)(rem - (long)jj * g.x);
}

// cg_solver_init passes 1-3 in one sweep (cg_solver.c:31-83).
__global__ void __launch_bounds__(kBlock)
k_cg_init_fields3(TlGrid3 g, int coef, double rx, double ry, double rz,
                  const double* __restrict__ density, const double* __restrict__ energy,
                  double* __restrict__ u, double* __restrict__ p, double* __restrict__ r,
                  double* __restrict__ w, double* __restrict__ kx, double* __restrict__ ky,
                  double* __restrict__ kz)
{
    const long xy = (long)g.x * g.y, n = xy * g.z;
    for (long c = blockIdx.x * (long)kBlock + threadIdx.x; c < n; c += (long)gridDim.x * kBlock) {
        int kk, jj, ii;
        decode3(c, g, kk, jj, ii);
        const double rho = density[c];
        p[c] = 0.0;
        r[c] = 0.0;
        u[c] = energy[c] * rho;
        if (ii >= 1 && ii < g.z - 1 && jj >= 1 && jj < g.y - 1 && kk >= 1 && kk < g.x - 1) {
            const double wc = coef == 1 ? rho : 1.0 / rho;
            w[c] = wc;
            if (ii >= g.hd && jj >= g.hd && kk >= g.hd) {
                // A neighbour outside [1,n-1) (halo_depth 1 only) keeps the w
                // in memory, as the reference reads it.
                const double wl = kk >= 2 ? (coef == 1 ? density[c - 1] : 1.0 / density[c - 1])
                                          : w[c - 1];
                const double wd = jj >= 2 ? (coef == 1 ? density[c - g.x] : 1.0 / density[c - g.x])
                                          : w[c - g.x];
                const double wb = ii >= 2 ? (coef == 1 ? density[c - xy] : 1.0 / density[c - xy])
                                          : w[c - xy];
                kx[c] = rx * (wl + wc) / (2.0 * wl * wc);
                ky[c] = ry * (wd + wc) / (2.0 * wd * wc);
                kz[c] = rz * (wb + wc) / (2.0 * wb * wc);
            }
        }
    }
}

// jacobi_solver_init (jacobi_solver.c:9-69).
__global__ void __launch_bounds__(kBlock)
k_jacobi_init3(TlGrid3 g, int coef, double rx, double ry, double rz,
               const double* __restrict__ density, const double* __restrict__ energy,
               double* __restrict__ u0, double* __restrict__ u, double* __restrict__ kx,
               double* __restrict__ ky, double* __restrict__ kz)
{
    const long xy = (long)g.x * g.y, n = xy * g.z;
    for (long c = blockIdx.x * (long)kBlock + threadIdx.x; c < n; c += (long)gridDim.x * kBlock) {
        int kk, jj, ii;
        decode3(c, g, kk, jj, ii);
        if (ii < 1 || ii >= g.z - 1 || jj < 1 || jj >= g.y - 1 || kk < 1 || kk >= g.x - 1) continue;
        const double t = energy[c] * density[c];
        u0[c] = t;
        u[c] = t;
        if (ii >= g.hd && jj >= g.hd && kk >= g.hd) {
            const bool cond = coef == 1;
            const double wc = cond ? density[c] : 1.0 / density[c];
            const double wl = cond ? density[c - 1] : 1.0 / density[c - 1];
            const double wd = cond ? density[c - g.x] : 1.0 / density[c - g.x];
            const double wb = cond ? density[c - xy] : 1.0 / density[c - xy];
            kx[c] = rx * (wl + wc) / (2.0 * wl * wc);
            ky[c] = ry * (wd + wc) / (2.0 * wd * wc);
            kz[c] = rz * (wb + wc) / (2.0 * wb * wc);
        }
    }
}

// calc_p fused into the next calc_w: p' = beta p + r (cg_solver.c:172-193)
// formed at every point the 7-point stencil reads, with the depth-1 mirror
// built in (a neighbour outside the interior takes the cell's own p'), then
// w = A p' and pw += w p' (cg_solver.c:105-135); p' goes to a second
// buffer.  The same expressions in the same order as cg_calc_p, a depth-1
// halo update of p and cg_calc_w, so p' and w are theirs bit for bit.
// Layout and z carry as k_stencil3d.
template <int V>
__global__ void __launch_bounds__(kBlock)
k_calc_pw3d(TlGrid3 g, Tile3 t, double beta, const double* __restrict__ p, const double* __restrict__ r, double* __restrict__ p_out,
            double* __restrict__ w, const double* __restrict__ kx, const double* __restrict__ ky,
            const double* __restrict__ kz, TlReduce rd, TlGate gate)
{
    if (gate.closed()) return;
    const int lane = threadIdx.x & (TL_WAVE - 1), wave = threadIdx.x / TL_WAVE;
    const int tile = xcd_tile(blockIdx.x, t.ns * t.nbj * t.nbk);
    const int sx = tile % t.ns;
    const int rest = tile / t.ns;
    const int jb = rest % t.nbj, kb = rest / t.nbj;
    const int j = g.hd + jb * kWaves + wave;
    const int xend = g.x - g.hd;
    const int c0 = g.hd + sx * (TL_WAVE * V) + lane * V;
    int nval = xend - c0;
    nval = nval < 0 ? 0 : (nval > V ? V : nval);
    const bool ld_ok = c0 + V <= g.x;
    const bool edge_r = lane == TL_WAVE - 1 && c0 + V < g.x;
    const long X = g.x, XY = (long)g.x * g.y;
    const int ylast = g.y - g.hd - 1, zfirst = g.hd, zlast = g.z - g.hd - 1;
    auto pn = [&](long i) -> DV<V> {
        DV<V> a = ldv<V>(p, i);
        const DV<V> b = ldv<V>(r, i);
#pragma unroll
        for (int v = 0; v < V; ++v) a.v[v] = beta * a.v[v] + b.v[v];
        return a;
    };
    double acc[1] = {0.0};

    if (j < g.y - g.hd) {   // uniform across the wave
        const long jrow = (long)j * X + c0;
        const int k0 = g.hd + kb * t.tz;             // < g.z - g.hd: kb < nbk
        const int k1 = min(k0 + t.tz, g.z - g.hd);
        DV<V> a_dn = zdv<V>(), a_c = zdv<V>(), kz_c = zdv<V>();
        if (ld_ok) {
            a_c = pn((long)k0 * XY + jrow);
            a_dn = k0 == zfirst ? a_c : pn((long)(k0 - 1) * XY + jrow);
            kz_c = ldv<V>(kz, (long)k0 * XY + jrow);
        }
        for (int k = k0; k < k1; ++k) {
            const long b = (long)k * XY + jrow;
            Nb3<V> n;
            double el = 0.0, er = 0.0, ek = 0.0;
            n.c = a_c;
            n.dn = a_dn;
            n.kzc = kz_c;
            if (ld_ok) {
                n.up = k == zlast ? a_c : pn(b + XY);
                n.no = j == ylast ? a_c : pn(b + X);
                n.so = j == g.hd ? a_c : pn(b - X);
                n.kxc = ldv<V>(kx, b);
                n.kyc = ldv<V>(ky, b);
                n.kyn = ldv<V>(ky, b + X);
                n.kzu = ldv<V>(kz, b + XY);
                if (lane == 0) el = beta * p[b - 1] + r[b - 1];
            } else {
                n.up = n.no = n.so = zdv<V>();
                n.kxc = n.kyc = n.kyn = n.kzu = zdv<V>();
            }
            if (edge_r) {
                er = beta * p[b + V] + r[b + V];
                ek = kx[b + V];
            }
            const double al = tl_lane_prev(n.c.v[V - 1], el);
            const double ar = tl_lane_next(n.c.v[0], er);
            const double kr = tl_lane_next(n.kxc.v[0], ek);
#pragma unroll
            for (int v = 0; v < V; ++v) {
                n.l.v[v] = v == 0 ? al : n.c.v[v - 1];
                n.r.v[v] = v == V - 1 ? ar : n.c.v[v + 1];
                n.kxr.v[v] = v == V - 1 ? kr : n.kxc.v[v + 1];
                if (c0 + v == g.hd) n.l.v[v] = n.c.v[v];          // x mirror
                if (c0 + v == xend - 1) n.r.v[v] = n.c.v[v];
            }
            if (nval > 0) {
                DV<V> o;
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    o.v[v] = apply_a(n, v);
                    if (v < nval) acc[0] += o.v[v] * n.c.v[v];
                }
                stv<V>(w, b, o, nval);
                stv<V>(p_out, b, n.c, nval);
            }
            a_dn = n.c;
            a_c = n.up;
            kz_c = n.kzu;
        }
    }
    tl_reduce_partial<1, kBlock>(acc, rd);
}

// Halo ring (every cell outside the interior box) src -> dst.
__global__ void __launch_bounds__(kBlock)
k_copy_ring3d(TlGrid3 g, const double* __restrict__ src, double* __restrict__ dst, TlGate gate)
{
    if (gate.closed()) return;
    const long xy = (long)g.x * g.y;
    const int nz = g.z - 2 * g.hd, ny = g.y - 2 * g.hd;
    const long nplane = 2L * g.hd * xy;                 // hd planes front and back
    const long nrow = 2L * g.hd * g.x * nz;             // hd rows each side, interior planes
    const long ncol = 2L * g.hd * ny * nz;              // hd columns each side, interior rows
    for (long t = blockIdx.x * (long)kBlock + threadIdx.x; t < nplane + nrow + ncol;
         t += (long)gridDim.x * kBlock) {
        long c;
        if (t < nplane) {
            const long pl = t / xy, rem = t % xy;
            const long ii = pl < g.hd ? pl : g.z - 2 * g.hd + pl;
            c = ii * xy + rem;
        } else if (t < nplane + nrow) {
            const long u = t - nplane;
            const long per = 2L * g.hd * g.x;
            const long ii = g.hd + u / per, q = u % per;
            const long rr = q / g.x, kk = q % g.x;
            const long jj = rr < g.hd ? rr : g.y - 2 * g.hd + rr;
            c = ii * xy + jj * g.x + kk;
        } else {
            const long u = t - nplane - nrow;
            const long per = 2L * g.hd;
            const long line = u / per, q = u % per;
            const long ii = g.hd + line / ny, jj = g.hd + line % ny;
            const long kk = q < g.hd ? q : g.x - 2 * g.hd + q;
            c = ii * xy + jj * g.x + kk;
        }
        dst[c] = src[c];
    }
}

// set_chunk_data (3d set_chunk_data.c:9-67).
__global__ void __launch_bounds__(kBlock)
k_set_chunk_data3(TlGrid3 g, double x0, double y0, double z0, double dx, double dy, double dz,
                  double* cell_x, double* cell_y, double* cell_z, double* vertex_x,
                  double* vertex_y, double* vertex_z, double* volume, double* x_area,
                  double* y_area, double* z_area)
{
    const long n = (long)g.x * g.y * g.z;
    const long tot = n + (g.x + 1) + (g.y + 1) + (g.z + 1);
    for (long t = blockIdx.x * (long)kBlock + threadIdx.x; t < tot; t += (long)gridDim.x * kBlock) {
        if (t < n) {
            volume[t] = dx * dy * dz;
            x_area[t] = dy * dz;
            y_area[t] = dx * dz;
            z_area[t] = dx * dy;
        } else if (t < n + g.x + 1) {
            const int i = (int)(t - n);
            const double v = x0 + dx * (i - g.hd);
            vertex_x[i] = v;
            if (i < g.x) cell_x[i] = 0.5 * (v + (x0 + dx * (i + 1 - g.hd)));
        } else if (t < n + g.x + g.y + 2) {
            const int i = (int)(t - n - g.x - 1);
            const double v = y0 + dy * (i - g.hd);
            vertex_y[i] = v;
            if (i < g.y) cell_y[i] = 0.5 * (v + (y0 + dy * (i + 1 - g.hd)));
        } else {
            const int i = (int)(t - n - g.x - g.y - 2);
            const double v = z0 + dz * (i - g.hd);
            vertex_z[i] = v;
            if (i < g.z) cell_z[i] = 0.5 * (v + (z0 + dz * (i + 1 - g.hd)));
        }
    }
}

// set_chunk_state (3d set_chunk_state.c:10-97): last covering state wins.
__global__ void __launch_bounds__(kBlock)
k_set_chunk_state3(TlGrid3 g, const double* __restrict__ vx, const double* __restrict__ vy,
                   const double* __restrict__ vz, const double* __restrict__ cx,
                   const double* __restrict__ cy, const double* __restrict__ cz,
                   double* __restrict__ density, double* __restrict__ energy0,
                   double* __restrict__ u, int nstates, const tlhip_state* __restrict__ st)
{
    const long n = (long)g.x * g.y * g.z;
    for (long c = blockIdx.x * (long)kBlock + threadIdx.x; c < n; c += (long)gridDim.x * kBlock) {
        int kk, jj, ii;
        decode3(c, g, kk, jj, ii);
        double e = st[0].energy, d = st[0].density;
        for (int s = 1; s < nstates; ++s) {
            bool hit = false;
            if (st[s].geometry == 0) {
                hit = vx[kk + 1] >= st[s].x_min && vx[kk] < st[s].x_max &&
                      vy[jj + 1] >= st[s].y_min && vy[jj] < st[s].y_max &&
                      vz[ii + 1] >= st[s].z_min && vz[ii] < st[s].z_max;
            } else if (st[s].geometry == 1) {
                const double ddx = cx[kk] - st[s].x_min;
                const double ddy = cy[jj] - st[s].y_min;
                const double ddz = cz[ii] - st[s].z_min;
                hit = sqrt(ddx * ddx + ddy * ddy + ddz * ddz) <= st[s].radius;
            } else if (st[s].geometry == 2) {
                hit = vx[kk] == st[s].x_min && vy[jj] == st[s].y_min && vz[ii] == st[s].z_min;
            }
            if (hit) {
                e = st[s].energy;
                d = st[s].density;
            }
        }
        energy0[c] = e;
        density[c] = d;
        if (ii >= 1 && ii < g.z - 1 && jj >= 1 && jj < g.y - 1 && kk >= 1 && kk < g.x - 1)
            u[c] = e * d;
    }
}

// All flagged fields, all external faces, one launch (local_halos.c:26-222).
struct HaloArgs3 {
    double* f[6];
    int nf;
    int faces;   // bit f set when neighbours[f] is external (3d/shared.h:24-30)
    int depth;
};

__global__ void __launch_bounds__(kBlock) k_local_halos3d(TlGrid3 g, HaloArgs3 h, TlGate gate)
{
    if (gate.closed()) return;
    double* f = h.f[blockIdx.y];
    const int nx = g.x - 2 * g.hd, ny = g.y - 2 * g.hd, nz = g.z - 2 * g.hd;
    const long xy = (long)g.x * g.y;
    const long nlr = (long)ny * nz * h.depth, ntb = (long)nx * nz * h.depth,
               nfb = (long)nx * ny * h.depth;
    const long seg[6] = {nlr, nlr, ntb, ntb, nfb, nfb};
    const long tot = 2 * (nlr + ntb + nfb);
    for (long t = blockIdx.x * (long)kBlock + threadIdx.x; t < tot; t += (long)gridDim.x * kBlock) {
        int face = 0;
        long u = t;
        while (u >= seg[face]) {
            u -= seg[face];
            ++face;
        }
        if (!(h.faces & (1 << face))) continue;
        const int d = (int)(u % h.depth);
        const long q = u / h.depth;
        if (face < 2) {          // left / right: over (jj, ii)
            const long jj = g.hd + q % ny, ii = g.hd + q / ny;
            const long b = ii * xy + jj * g.x;
            if (face == 0) f[b + (g.hd - d - 1)] = f[b + (g.hd + d)];
            else f[b + (g.x - g.hd + d)] = f[b + (g.x - g.hd - 1 - d)];
        } else if (face < 4) {   // bottom / top: over (kk, ii)
            const long kk = g.hd + q % nx, ii = g.hd + q / nx;
            const long b = ii * xy + kk;
            if (face == 2) f[b + (long)(g.hd - d - 1) * g.x] = f[b + (long)(g.hd + d) * g.x];
            else f[b + (long)(g.y - g.hd + d) * g.x] = f[b + (long)(g.y - g.hd - 1 - d) * g.x];
        } else {                 // back / front: over (kk, jj)
            const long kk = g.hd + q % nx, jj = g.hd + q / nx;
            const long b = jj * g.x + kk;
            if (face == 4) f[b + (g.hd - d - 1) * xy] = f[b + (g.hd + d) * xy];
            else f[b + (g.z - g.hd + d) * xy] = f[b + (g.z - g.hd - 1 - d) * xy];
        }
    }
}

// Face box <-> contiguous buffer (3d pack_halos.c:5-363): x, then y, then z.
__global__ void __launch_bounds__(kBlock)
k_box3d(TlGrid3 g, int k0, int j0, int i0, int ex, int ey, int ez, int pack,
        double* __restrict__ field, double* __restrict__ buf)
{
    const long tot = (long)ex * ey * ez;
    for (long t = blockIdx.x * (long)kBlock + threadIdx.x; t < tot; t += (long)gridDim.x * kBlock) {
        const long dk = t % ex, dj = (t / ex) % ey, di = t / ((long)ex * ey);
        const long c = (k0 + dk) + (j0 + dj) * g.x + (i0 + di) * (long)g.x * g.y;
        if (pack) buf[t] = field[c];
        else field[c] = buf[t];
    }
}

// ------------------------------------------------------------ host side

int check_grid(const TlGrid3& g, const char* what)
{
    if (g.hd < 1 || g.x < 2 * g.hd || g.y < 2 * g.hd || g.z < 2 * g.hd)
        return tl_fail(TLHIP_ERR_ARG, "%s: invalid grid x=%d y=%d z=%d halo_depth=%d", what, g.x,
                       g.y, g.z, g.hd);
    if ((long)g.x * g.y * g.z > (1L << 31) - 1)
        return tl_fail(TLHIP_ERR_ARG, "%s: grid too large", what);
    return 0;
}

bool interior_empty(const TlGrid3& g)
{
    return g.x - 2 * g.hd <= 0 || g.y - 2 * g.hd <= 0 || g.z - 2 * g.hd <= 0;
}

TlRows rows_of(const TlGrid3& g)
{
    const int ny = g.y - 2 * g.hd, nz = g.z - 2 * g.hd;
    return TlRows{g.x, g.hd, ny, (long)g.x * g.y, g.hd, (long)ny * nz};
}

template <typename... P>
int pick_v(const TlGrid3& g, P... ps)
{
    return (g.x % 2 == 0 && g.hd % 2 == 0 && g.hd >= 2 && all_aligned(ps...)) ? 2 : 1;
}

int g_tz_override = -1;

Tile3 make_tile(const TlGrid3& g, int V)
{
    Tile3 t;
    const int nx = g.x - 2 * g.hd, ny = g.y - 2 * g.hd, nz = g.z - 2 * g.hd;
    t.ns = (int)tl_cdiv(nx, TL_WAVE * V);
    t.nbj = (int)tl_cdiv(ny, kWaves);
    int tz = g_tz_override > 0 ? g_tz_override : 1;
    while ((long)t.ns * t.nbj * tl_cdiv(nz, tz) > TL_MAX_RED_BLOCKS) tz *= 2;
    t.tz = tz;
    t.nbk = (int)tl_cdiv(nz, tz);
    return t;
}

// Under the current launch gate; with nb_out, a reducing op's partials are
// left for the caller's finisher (see launch_interior).
template <class Op>
int launch_stencil(const TlGrid3& g, int V, const double* a, const double* kx, const double* ky,
                   const double* kz, const Op& op, const char* what, unsigned* nb_out = nullptr)
{
    if (interior_empty(g)) return 0;
    if ((long)tl_cdiv(g.x - 2 * g.hd, TL_WAVE * V) * tl_cdiv(g.y - 2 * g.hd, kWaves) >
        TL_MAX_RED_BLOCKS)
        return tl_fail(TLHIP_ERR_ARG, "%s: planes too large (%d x %d)", what, g.x, g.y);
    const Tile3 t = make_tile(g, V);
    const unsigned nb = (unsigned)(t.ns * t.nbj * t.nbk);
    const int slot = tl_timer_begin(what);
    if (V == 2)
        hipLaunchKernelGGL((k_stencil3d<2, Op>), dim3(nb), dim3(kBlock), 0, tl_stream(), g, t, a,
                           kx, ky, kz, op, tl_reduce_slots(), tl_gate());
    else
        hipLaunchKernelGGL((k_stencil3d<1, Op>), dim3(nb), dim3(kBlock), 0, tl_stream(), g, t, a,
                           kx, ky, kz, op, tl_reduce_slots(), tl_gate());
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
int launch_interior3(const TlGrid3& g, int V, const Op& op, const char* what,
                     unsigned* nb_out = nullptr)
{
    if (interior_empty(g)) return 0;
    return launch_interior(rows_of(g), g.x - 2 * g.hd, V, op, what, nb_out);
}

int copy_ring(const TlGrid3& g, const double* src, double* dst, const char* what)
{
    const long n = (long)g.x * g.y * g.z;
    const long inner = (long)(g.x - 2 * g.hd) * (g.y - 2 * g.hd) * (g.z - 2 * g.hd);
    hipLaunchKernelGGL(k_copy_ring3d, dim3(flat_blocks(n - inner)), dim3(kBlock), 0, tl_stream(),
                       g, src, dst, tl_gate());
    return tl_check_launch(what);
}

// Ring copy p -> p_out and k_calc_pw3d under the current gate; *nb: its
// partials.
int launch_pw3d(const TlGrid3& g, double beta, const double* p,
                double* p_out, const double* r, double* w, const double* kx, const double* ky,
                const double* kz, const char* what, unsigned* nb)
{
    if (!(p && p_out && r && w && kx && ky && kz))
        return tl_fail(TLHIP_ERR_ARG, "%s: null argument", what);
    if (p_out == p || p_out == w || p_out == r || w == p)
        return tl_fail(TLHIP_ERR_ARG, "%s: p_out and w must be distinct from the inputs", what);
    if (interior_empty(g)) return tl_fail(TLHIP_ERR_ARG, "%s: empty interior", what);
    if ((long)tl_cdiv(g.x - 2 * g.hd, TL_WAVE * 2) * tl_cdiv(g.y - 2 * g.hd, kWaves) >
        TL_MAX_RED_BLOCKS)
        return tl_fail(TLHIP_ERR_ARG, "%s: planes too large (%d x %d)", what, g.x, g.y);
    TL_TRY(copy_ring(g, p, p_out, what));
    const int V = pick_v(g, p, p_out, r, w, kx, ky, kz);
    const Tile3 t = make_tile(g, V);
    *nb = (unsigned)(t.ns * t.nbj * t.nbk);
    const int slot = tl_timer_begin(what);
    if (V == 2)
        hipLaunchKernelGGL(k_calc_pw3d<2>, dim3(*nb), dim3(kBlock), 0, tl_stream(), g, t, beta,
                           p, r, p_out, w, kx, ky, kz, tl_reduce_slots(), tl_gate());
    else
        hipLaunchKernelGGL(k_calc_pw3d<1>, dim3(*nb), dim3(kBlock), 0, tl_stream(), g, t, beta,
                           p, r, p_out, w, kx, ky, kz, tl_reduce_slots(), tl_gate());
    tl_timer_end(slot);
    return tl_check_launch(what);
}

#define REQUIRE TL_REQUIRE

}  // namespace

extern "C" int tlhip_set_tile_planes(int planes)
{
    g_tz_override = planes;
    return 0;
}

// ---------------------------------------------------------------- init

extern "C" int tlhip3d_set_chunk_data(int x, int y, int z, int hd, int left, int bottom, int back,
                                      double grid_x_min, double grid_y_min, double grid_z_min,
                                      double dx, double dy, double dz, double* cell_x,
                                      double* cell_y, double* cell_z, double* vertex_x,
                                      double* vertex_y, double* vertex_z, double* volume,
                                      double* x_area, double* y_area, double* z_area)
{
    const char* what = "tlhip3d_set_chunk_data";
    TL_TRY(tl_ensure_init());
    const TlGrid3 g{x, y, z, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(cell_x && cell_y && cell_z && vertex_x && vertex_y && vertex_z && volume && x_area &&
                y_area && z_area,
            what, "null field");
    const double x0 = grid_x_min + dx * (double)left;
    const double y0 = grid_y_min + dy * (double)bottom;
    const double z0 = grid_z_min + dz * (double)back;
    const long tot = (long)x * y * z + x + y + z + 3;
    hipLaunchKernelGGL(k_set_chunk_data3, dim3(flat_blocks(tot)), dim3(kBlock), 0, tl_stream(), g,
                       x0, y0, z0, dx, dy, dz, cell_x, cell_y, cell_z, vertex_x, vertex_y,
                       vertex_z, volume, x_area, y_area, z_area);
    return tl_check_launch(what);
}

extern "C" int tlhip3d_set_chunk_state(int x, int y, int z, const double* vertex_x,
                                       const double* vertex_y, const double* vertex_z,
                                       const double* cell_x, const double* cell_y,
                                       const double* cell_z, double* density, double* energy0,
                                       double* u, int num_states, const tlhip_state* states)
{
    const char* what = "tlhip3d_set_chunk_state";
    TL_TRY(tl_ensure_init());
    REQUIRE(x >= 1 && y >= 1 && z >= 1, what, "invalid grid");
    REQUIRE(num_states >= 1 && states, what, "need at least one state");
    REQUIRE(vertex_x && vertex_y && vertex_z && cell_x && cell_y && cell_z && density &&
                energy0 && u,
            what, "null field");
    const TlGrid3 g{x, y, z, 1};
    void* dst = nullptr;
    TL_TRY(tlhip_alloc(sizeof(tlhip_state) * num_states, &dst));
    int rc = tlhip_copy_to_device(dst, states, sizeof(tlhip_state) * num_states);
    if (!rc) {
        hipLaunchKernelGGL(k_set_chunk_state3, dim3(flat_blocks((long)x * y * z)), dim3(kBlock), 0,
                           tl_stream(), g, vertex_x, vertex_y, vertex_z, cell_x, cell_y, cell_z,
                           density, energy0, u, num_states, (const tlhip_state*)dst);
        rc = tl_check_launch(what);
    }
    const int rc2 = tlhip_free(dst);
    return rc ? rc : rc2;
}

// ---------------------------------------------------------------- halos

extern "C" int tlhip3d_local_halos(int x, int y, int z, int depth, int hd, const int* neighbours,
                                   const bool* fields, double* density, double* energy0,
                                   double* energy, double* u, double* p, double* sd)
{
    const char* what = "tlhip3d_local_halos";
    TL_TRY(tl_ensure_init());
    const TlGrid3 g{x, y, z, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(neighbours && fields, what, "null neighbours/fields");
    REQUIRE(depth >= 0 && depth <= hd, what, "depth must be in [0, halo_depth]");
    HaloArgs3 h{};
    double* all[6] = {density, energy0, energy, u, p, sd};
    for (int i = 0; i < 6; ++i)
        if (fields[i]) {
            REQUIRE(all[i], what, "flagged field is null");
            h.f[h.nf++] = all[i];
        }
    for (int f = 0; f < 6; ++f)
        if (neighbours[f] == -1) h.faces |= 1 << f;
    h.depth = depth;
    if (h.nf == 0 || h.faces == 0 || depth == 0 || interior_empty(g)) return 0;
    const long nx = x - 2 * hd, ny = y - 2 * hd, nz = z - 2 * hd;
    const long tot = 2 * (ny * nz + nx * nz + nx * ny) * depth;
    hipLaunchKernelGGL(k_local_halos3d, dim3(flat_blocks(tot), h.nf), dim3(kBlock), 0,
                       tl_stream(), g, h, tl_gate());
    return tl_check_launch(what);
}

extern "C" int tlhip3d_pack_or_unpack(int x, int y, int z, int depth, int hd, int face, bool pack,
                                      double* field, double* buffer)
{
    const char* what = "tlhip3d_pack_or_unpack";
    TL_TRY(tl_ensure_init());
    const TlGrid3 g{x, y, z, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(field && buffer, what, "null field/buffer");
    REQUIRE(depth >= 1 && depth <= hd, what, "depth must be in [1, halo_depth]");
    int k0 = hd, j0 = hd, i0 = hd, ex = x - 2 * hd, ey = y - 2 * hd, ez = z - 2 * hd;
    switch (face) {
    case 0: k0 = pack ? hd : hd - depth; ex = depth; break;
    case 1: k0 = pack ? x - hd - depth : x - hd; ex = depth; break;
    case 2: j0 = pack ? hd : hd - depth; ey = depth; break;
    case 3: j0 = pack ? y - hd - depth : y - hd; ey = depth; break;
    case 4: i0 = pack ? hd : hd - depth; ez = depth; break;
    case 5: i0 = pack ? z - hd - depth : z - hd; ez = depth; break;
    default: return tl_fail(TLHIP_ERR_ARG, "%s: Incorrect face provided: %d.", what, face);
    }
    const long n = (long)ex * ey * ez;
    if (n <= 0) return 0;
    double* stage = nullptr;
    TL_TRY(tl_staging((size_t)n, &stage));
    if (!pack) TL_TRY(tlhip_copy_to_device(stage, buffer, sizeof(double) * n));
    hipLaunchKernelGGL(k_box3d, dim3(flat_blocks(n)), dim3(kBlock), 0, tl_stream(), g, k0, j0, i0,
                       ex, ey, ez, pack ? 1 : 0, field, stage);
    TL_TRY(tl_check_launch(what));
    if (pack) TL_TRY(tlhip_copy_to_host(buffer, stage, sizeof(double) * n));
    return 0;
}

extern "C" int tlhip3d_store_energy(int x, int y, int z, const double* energy0, double* energy)
{
    const char* what = "tlhip3d_store_energy";
    TL_TRY(tl_ensure_init());
    REQUIRE(x >= 1 && y >= 1 && z >= 1 && energy0 && energy, what, "invalid arguments");
    const long n = (long)x * y * z;
    hipLaunchKernelGGL(k_copy_all, dim3(flat_blocks(n)), dim3(kBlock), 0, tl_stream(), n, energy0,
                       energy);
    return tl_check_launch(what);
}

extern "C" int tlhip3d_field_summary(int x, int y, int z, int hd, const double* volume,
                                     const double* density, const double* energy0,
                                     const double* u, double* vol, double* mass, double* ie,
                                     double* temp)
{
    const char* what = "tlhip3d_field_summary";
    TL_TRY(tl_ensure_init());
    const TlGrid3 g{x, y, z, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(volume && density && energy0 && u && vol && mass && ie && temp, what, "null argument");
    if (interior_empty(g)) return 0;
    OpFieldSummary op{volume, density, energy0, u};
    TL_TRY(launch_interior3(g, pick_v(g, volume, density, energy0, u), op, what));
    double s[4];
    TL_TRY(tl_fetch(4, s, what));
    *vol += s[0];
    *mass += s[1];
    *ie += s[2];
    *temp += s[3];
    return 0;
}

// ---------------------------------------------------------------- CG

extern "C" int tlhip3d_cg_solver_init(int x, int y, int z, int hd, int coefficient, double rx,
                                      double ry, double rz, double* rro, const double* density,
                                      const double* energy, double* u, double* p, double* r,
                                      double* w, double* kx, double* ky, double* kz)
{
    const char* what = "tlhip3d_cg_solver_init";
    TL_TRY(tl_ensure_init());
    const TlGrid3 g{x, y, z, hd};
    TL_TRY(check_grid(g, what));
    if (coefficient != 1 && coefficient != 2)
        return tl_fail(TLHIP_ERR_ARG, "Coefficient %d is not valid.", coefficient);
    REQUIRE(rro && density && energy && u && p && r && w && kx && ky && kz, what, "null argument");
    const long n = (long)x * y * z;
    hipLaunchKernelGGL(k_cg_init_fields3, dim3(flat_blocks(n)), dim3(kBlock), 0, tl_stream(), g,
                       coefficient, rx, ry, rz, density, energy, u, p, r, w, kx, ky, kz);
    TL_TRY(tl_check_launch(what));
    if (interior_empty(g)) return 0;
    OpCgInit op{w, r, p};
    TL_TRY(launch_stencil(g, pick_v(g, u, kx, ky, kz, w, r, p), u, kx, ky, kz, op, what));
    return fetch_add(rro, what);
}

extern "C" int tlhip3d_cg_solver_calc_w(int x, int y, int z, int hd, double* pw, const double* p,
                                        double* w, const double* kx, const double* ky,
                                        const double* kz)
{
    const char* what = "tlhip3d_cg_solver_calc_w";
    TL_TRY(tl_ensure_init());
    const TlGrid3 g{x, y, z, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(pw && p && w && kx && ky && kz, what, "null argument");
    REQUIRE(p != w, what, "p and w must be distinct buffers");
    if (interior_empty(g)) return 0;
    OpCalcW op{w};
    TL_TRY(launch_stencil(g, pick_v(g, p, w, kx, ky, kz), p, kx, ky, kz, op, what));
    return fetch_add(pw, what);
}

extern "C" int tlhip3d_cg_solver_calc_pw(int x, int y, int z, int hd, double beta, double* pw,
                                         const double* p_in, double* p_out, const double* r,
                                         double* w, const double* kx, const double* ky,
                                         const double* kz)
{
    const char* what = "tlhip3d_cg_solver_calc_pw";
    TL_TRY(tl_ensure_init());
    const TlGrid3 g{x, y, z, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(pw, what, "null argument");
    unsigned nb = 0;
    TL_TRY(launch_pw3d(g, beta, p_in, p_out, r, w, kx, ky, kz, what, &nb));
    TL_TRY(tl_finish_reduce(1, nb, what));
    return fetch_add(pw, what);
}

extern "C" int tlhip3d_cg_solver_calc_ur(int x, int y, int z, int hd, double alpha, double* rrn,
                                         double* u, const double* p, double* r, const double* w)
{
    const char* what = "tlhip3d_cg_solver_calc_ur";
    TL_TRY(tl_ensure_init());
    const TlGrid3 g{x, y, z, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(rrn && u && p && r && w, what, "null argument");
    if (interior_empty(g)) return 0;
    OpCalcUr op{u, p, r, w, alpha};
    TL_TRY(launch_interior3(g, pick_v(g, u, p, r, w), op, what));
    return fetch_add(rrn, what);
}

extern "C" int tlhip3d_cg_solver_calc_p(int x, int y, int z, int hd, double beta, double* p,
                                        const double* r)
{
    const char* what = "tlhip3d_cg_solver_calc_p";
    TL_TRY(tl_ensure_init());
    const TlGrid3 g{x, y, z, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(p && r, what, "null argument");
    OpCalcP op{p, r, beta};
    return launch_interior3(g, pick_v(g, p, r), op, what);
}

// ---------------------------------------------------------------- Chebyshev

extern "C" int tlhip3d_cheby_solver_init(int x, int y, int z, int hd, double theta, double* u,
                                         const double* u0, double* p, double* r, double* w,
                                         const double* kx, const double* ky, const double* kz)
{
    const char* what = "tlhip3d_cheby_solver_init";
    TL_TRY(tl_ensure_init());
    const TlGrid3 g{x, y, z, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(u && u0 && p && r && w && kx && ky && kz, what, "null argument");
    OpChebyInit op{u0, w, r, p, theta};
    TL_TRY(launch_stencil(g, pick_v(g, u, u0, p, r, w, kx, ky, kz), u, kx, ky, kz, op, what));
    OpSimple up{u, p, 1, nullptr, 0.0};
    return launch_interior3(g, pick_v(g, u, p), up, what);
}

extern "C" int tlhip3d_cheby_solver_iterate(int x, int y, int z, int hd, double alpha,
                                            double beta, const double* u_in, double* u_out,
                                            const double* u0, double* p, double* r, double* w,
                                            const double* kx, const double* ky, const double* kz)
{
    const char* what = "tlhip3d_cheby_solver_iterate";
    TL_TRY(tl_ensure_init());
    const TlGrid3 g{x, y, z, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(u_in && u_out && u0 && p && r && kx && ky && kz, what, "null argument");
    REQUIRE(u_in != u_out, what, "u_in and u_out must be distinct buffers");
    OpChebyIter op{u0, p, r, w, u_out, alpha, beta};
    TL_TRY(launch_stencil(g, pick_v(g, u_in, u_out, u0, p, r, w, kx, ky, kz), u_in, kx, ky, kz,
                          op, what));
    return copy_ring(g, u_in, u_out, what);
}

// ---------------------------------------------------------------- Jacobi

extern "C" int tlhip3d_jacobi_solver_init(int x, int y, int z, int hd, int coefficient, double rx,
                                          double ry, double rz, const double* density,
                                          const double* energy, double* u0, double* u, double* kx,
                                          double* ky, double* kz)
{
    const char* what = "tlhip3d_jacobi_solver_init";
    TL_TRY(tl_ensure_init());
    const TlGrid3 g{x, y, z, hd};
    TL_TRY(check_grid(g, what));
    if (coefficient < 1) return tl_fail(TLHIP_ERR_ARG, "Coefficient %d is not valid.", coefficient);
    REQUIRE(density && energy && u0 && u && kx && ky && kz, what, "null argument");
    hipLaunchKernelGGL(k_jacobi_init3, dim3(flat_blocks((long)x * y * z)), dim3(kBlock), 0,
                       tl_stream(), g, coefficient, rx, ry, rz, density, energy, u0, u, kx, ky,
                       kz);
    return tl_check_launch(what);
}

extern "C" int tlhip3d_jacobi_solver_iterate(int x, int y, int z, int hd, double* error,
                                             const double* kx, const double* ky, const double* kz,
                                             const double* u0, double* u, double* r)
{
    const char* what = "tlhip3d_jacobi_solver_iterate";
    TL_TRY(tl_ensure_init());
    const TlGrid3 g{x, y, z, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(error && kx && ky && kz && u0 && u && r, what, "null argument");
    REQUIRE(u != r, what, "u and r must be distinct buffers");
    const long n = (long)x * y * z;
    hipLaunchKernelGGL(k_copy_all, dim3(flat_blocks(n)), dim3(kBlock), 0, tl_stream(), n, u, r);
    TL_TRY(tl_check_launch(what));
    if (interior_empty(g)) {
        *error = 0.0;
        return 0;
    }
    OpJacobi3 op{u0, u};
    TL_TRY(launch_stencil(g, pick_v(g, kx, ky, kz, u0, u, r), r, kx, ky, kz, op, what));
    double v = 0.0;
    TL_TRY(tl_fetch(1, &v, what));
    *error = v;   // assigned, as jacobi_solver.c:115
    return 0;
}

// ---------------------------------------------------------------- PPCG

extern "C" int tlhip3d_ppcg_init(int x, int y, int z, int hd, double theta, const double* r,
                                 double* sd)
{
    const char* what = "tlhip3d_ppcg_init";
    TL_TRY(tl_ensure_init());
    const TlGrid3 g{x, y, z, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(r && sd, what, "null argument");
    OpSimple op{sd, r, 2, nullptr, theta};
    return launch_interior3(g, pick_v(g, r, sd), op, what);
}

namespace {

// ---------------------------------------------------------------- two PPCG steps
// Two turns of the 3D PPCG inner loop (3d/drivers/ppcg_solver_driver.c:
// 170-183: a depth-1 halo update of sd, then ppcg_inner_iteration,
// 3d/c_kernels/omp3/ppcg_solver.c:32-73) in one pass.  On a single chunk the
// halo update is a mirror (a neighbour outside the interior is the cell
// itself), so level 2 of a cell needs level 1 only within one cell of it.
// A wave owns one output row j of a 128-column window and marches up its
// planes; it forms level 1 for rows j-1, j and j+1 itself (the y
// neighbours level 2 needs; the rows j-2..j+2 of sd it reads come from the
// cache lines the block's other waves fetch), carries planes k-1, k, k+1 of
// sd (five rows) and of level-1 sd (three rows) in registers, and writes
// level 2 of plane k-1 after forming level 1 of plane k.  x neighbours come
// from the adjacent lane; the window's end lanes have none, so a window
// writes 122 of its 128 columns and windows overlap.  r, u and sd are read
// from *_in and written to *_out.  Each cell's expressions are the
// reference's, in its order, so the result is bit-for-bit that of two
// separate steps.
constexpr int kP2Out = 122;    // output columns per 128-column window
int g_pair_rows = 8;           // output rows per block of the LDS variant (0: per-wave rows)
constexpr int kP2Rows = 4;     // output rows per block (one wave each)

using C2 = Cells<double>;

__device__ __forceinline__ C2 ld3(const double* __restrict__ f, const TlGrid3& g, int j, int k,
                                  int c)
{
    if (c < 0 || c + 2 > g.x || j < 0 || j >= g.y || k < 0 || k >= g.z) return zeros<double>();
    return ldc(f, ((long)k * g.y + j) * g.x + c);
}

// A a (3d/shared.h:52-58, same association) for a lane's two cells; x
// neighbours from the adjacent lanes (don't-care at the wave's ends), the
// cell itself where the neighbour is outside the interior.
__device__ __forceinline__ C2 apply7(const C2& cen, const C2& no, const C2& so, const C2& up,
                                     const C2& dn, const C2& kxc, const C2& kyc, const C2& kyn,
                                     const C2& kzc, const C2& kzu, int c, int xin0, int xin1)
{
    const double lft_e = tl_lane_prev(cen.v[1], cen.v[0]);
    const double rgt_e = tl_lane_next(cen.v[0], cen.v[1]);
    const double kxe = tl_lane_next(kxc.v[0], kxc.v[1]);
    C2 a;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        double l = i == 0 ? lft_e : cen.v[0];
        double r = i == 1 ? rgt_e : cen.v[1];
        if (c + i == xin0) l = cen.v[i];
        if (c + i == xin1 - 1) r = cen.v[i];
        const double kxr = i == 1 ? kxe : kxc.v[1];
        const double diag = ((1.0 + (kxr + kxc.v[i])) + (kyn.v[i] + kyc.v[i])) +
                            (kzu.v[i] + kzc.v[i]);
        a.v[i] = diag * cen.v[i] - (kxr * r + kxc.v[i] * l) - (kyn.v[i] * no.v[i] + kyc.v[i] * so.v[i]) -
                 (kzu.v[i] * up.v[i] + kzc.v[i] * dn.v[i]);
    }
    return a;
}

__global__ void __launch_bounds__(kBlock)
k_ppcg3d_pair(TlGrid3 g, int nwx, int nby, int nbz, int tz, double a1, double b1, double a2,
              double b2, const double* __restrict__ sd_in, double* __restrict__ sd_out,
              const double* __restrict__ r_in, double* __restrict__ r_out,
              const double* __restrict__ u_in, double* __restrict__ u_out,
              const double* __restrict__ kx, const double* __restrict__ ky,
              const double* __restrict__ kz, TlGate gate)
{
    if (gate.closed()) return;
    const int lane = threadIdx.x & (TL_WAVE - 1), wave = threadIdx.x / TL_WAVE;
    const int tile = xcd_tile(blockIdx.x, nwx * nby * nbz);
    const int wx = tile % nwx, rest = tile / nwx;
    const int by = rest % nby, bz = rest / nby;
    const int xin0 = g.hd, xin1 = g.x - g.hd, yin0 = g.hd, yin1 = g.y - g.hd;
    const int zin0 = g.hd, zin1 = g.z - g.hd;
    const int j = yin0 + by * kP2Rows + wave;
    if (j >= yin1) return;                     // uniform per wave; no block sync below
    const int o0 = xin0 + wx * kP2Out, o1 = min(o0 + kP2Out, xin1);
    const int cw = (o0 - 2) >= 0 ? (o0 - 2) / 2 * 2 : -((3 - o0) / 2 * 2);   // even, <= o0 - 2
    const int c = cw + 2 * lane;
    const bool st0 = c >= o0 && c < o1, st1 = c + 1 >= o0 && c + 1 < o1;
    const int k0 = zin0 + bz * tz, k1 = min(k0 + tz, zin1);
    const long X = g.x, XY = (long)g.x * g.y;

    // s0[r][p]: sd at row j-2+r, planes kk-1, kk, kk+1 of the level-1 plane kk
    // s1[r][p]: level-1 sd at row j-1+r, planes km-1, km, km+1 of the level-2 plane km
    C2 s0[5][3], s1[3][3];
#pragma unroll
    for (int r = 0; r < 5; ++r) {
        s0[r][0] = zeros<double>();
        s0[r][1] = ld3(sd_in, g, j - 2 + r, k0 - 2, c);
        s0[r][2] = ld3(sd_in, g, j - 2 + r, k0 - 1, c);
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) s1[r][0] = s1[r][1] = s1[r][2] = zeros<double>();
    // level 1's r and u of row j, and row j's coefficients, one plane back
    C2 r1p = zeros<double>(), u1p = zeros<double>();
    C2 kxp = zeros<double>(), kyp = zeros<double>(), kynp = zeros<double>();
    C2 kzp = zeros<double>(), kzup = zeros<double>();

    for (int kk = k0 - 1; kk <= k1; ++kk) {
#pragma unroll
        for (int r = 0; r < 5; ++r) {
            s0[r][0] = s0[r][1];
            s0[r][1] = s0[r][2];
            s0[r][2] = ld3(sd_in, g, j - 2 + r, kk + 1, c);
        }
        C2 kxr_[3], kyr_[4], kzr_[3], kzur_[3], r0_[3];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            kxr_[r] = ld3(kx, g, j - 1 + r, kk, c);
            kzr_[r] = ld3(kz, g, j - 1 + r, kk, c);
            kzur_[r] = ld3(kz, g, j - 1 + r, kk + 1, c);
            r0_[r] = ld3(r_in, g, j - 1 + r, kk, c);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) kyr_[r] = ld3(ky, g, j - 1 + r, kk, c);
        const C2 u0 = ld3(u_in, g, j, kk, c);
        // level 1 at plane kk, rows j-1..j+1 (a row or plane outside the
        // interior is never read: the mirrors below replace it)
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            s1[r][0] = s1[r][1];
            s1[r][1] = s1[r][2];
        }
        C2 r1 = zeros<double>(), u1 = zeros<double>();
        if (kk >= zin0 && kk < zin1) {
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                const int jj = j - 1 + r;
                const C2 cen = s0[r + 1][1];
                const C2 no = jj == yin1 - 1 ? cen : s0[r + 2][1];
                const C2 so = jj == yin0 ? cen : s0[r][1];
                const C2 up = kk == zin1 - 1 ? cen : s0[r + 1][2];
                const C2 dn = kk == zin0 ? cen : s0[r + 1][0];
                const C2 w = apply7(cen, no, so, up, dn, kxr_[r], kyr_[r], kyr_[r + 1], kzr_[r],
                                    kzur_[r], c, xin0, xin1);
                C2 rn, sn;
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    rn.v[i] = r0_[r].v[i] - w.v[i];
                    sn.v[i] = a1 * cen.v[i] + b1 * rn.v[i];
                }
                s1[r][2] = sn;
                if (r == 1) {
                    r1 = rn;
#pragma unroll
                    for (int i = 0; i < 2; ++i) u1.v[i] = u0.v[i] + cen.v[i];
                }
            }
        }
        // level 2 at plane km = kk - 1
        const int km = kk - 1;
        if (km >= k0 && km < k1) {
            const C2 cen = s1[1][1];
            const C2 no = j == yin1 - 1 ? cen : s1[2][1];
            const C2 so = j == yin0 ? cen : s1[0][1];
            const C2 up = km == zin1 - 1 ? cen : s1[1][2];
            const C2 dn = km == zin0 ? cen : s1[1][0];
            const C2 w = apply7(cen, no, so, up, dn, kxp, kyp, kynp, kzp, kzup, c, xin0, xin1);
            C2 rn, un, sn;
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                rn.v[i] = r1p.v[i] - w.v[i];
                un.v[i] = u1p.v[i] + cen.v[i];
                sn.v[i] = a2 * cen.v[i] + b2 * rn.v[i];
            }
            const long b = (long)km * XY + (long)j * X + c;
            if (st0 && st1) {
                stc(r_out, b, rn);
                stc(u_out, b, un);
                stc(sd_out, b, sn);
            } else {
#pragma unroll
                for (int i = 0; i < 2; ++i)
                    if (i == 0 ? st0 : st1) {
                        r_out[b + i] = rn.v[i];
                        u_out[b + i] = un.v[i];
                        sd_out[b + i] = sn.v[i];
                    }
            }
        }
        r1p = r1;
        u1p = u1;
        kxp = kxr_[1];
        kyp = kyr_[1];
        kynp = kyr_[2];
        kzp = kzr_[1];
        kzup = kzur_[1];
    }
}

// Two cells per lane in either precision: Cells<double> for FP64, F2 for
// the FP32 fields of the FP32-refined solve (rows of pitch >= x).
struct F2 {
    float v[2];
};
template <class T>
struct Pair2;
template <>
struct Pair2<double> {
    using type = C2;
};
template <>
struct Pair2<float> {
    using type = F2;
};
template <class T>
using P2 = typename Pair2<T>::type;

template <class T>
__device__ __forceinline__ P2<T> zeros2()
{
    P2<T> z;
    z.v[0] = z.v[1] = T(0);
    return z;
}
__device__ __forceinline__ C2 ld2(const double* __restrict__ f, long i) { return ldc(f, i); }
__device__ __forceinline__ F2 ld2(const float* __restrict__ f, long i)
{
    const float2 a = *reinterpret_cast<const float2*>(f + i);
    return F2{{a.x, a.y}};
}
__device__ __forceinline__ void st2(double* __restrict__ f, long i, const C2& v) { stc(f, i, v); }
__device__ __forceinline__ void st2(float* __restrict__ f, long i, const F2& v)
{
    float2 a;
    a.x = v.v[0];
    a.y = v.v[1];
    *reinterpret_cast<float2*>(f + i) = a;
}

// ld3 with a row pitch of its own (FP32 rows are padded)
template <class T>
__device__ __forceinline__ P2<T> ld3p(const T* __restrict__ f, const TlGrid3& g, int pitch,
                                      int j, int k, int c)
{
    if (c < 0 || c + 2 > g.x || j < 0 || j >= g.y || k < 0 || k >= g.z) return zeros2<T>();
    return ld2(f, ((long)k * g.y + j) * pitch + c);
}

// apply7 in either precision (the same association; FP32 as the FP32
// 7-point group kernel forms it)
template <class V>
__device__ __forceinline__ V apply7t(const V& cen, const V& no, const V& so, const V& up,
                                     const V& dn, const V& kxc, const V& kyc, const V& kyn,
                                     const V& kzc, const V& kzu, int c, int xin0, int xin1)
{
    using T = decltype(cen.v[0] + cen.v[0]);
    const T lft_e = tl_lane_prev(cen.v[1], cen.v[0]);
    const T rgt_e = tl_lane_next(cen.v[0], cen.v[1]);
    const T kxe = tl_lane_next(kxc.v[0], kxc.v[1]);
    V a;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        T l = i == 0 ? lft_e : cen.v[0];
        T r = i == 1 ? rgt_e : cen.v[1];
        if (c + i == xin0) l = cen.v[i];
        if (c + i == xin1 - 1) r = cen.v[i];
        const T kxr = i == 1 ? kxe : kxc.v[1];
        const T diag = ((T(1) + (kxr + kxc.v[i])) + (kyn.v[i] + kyc.v[i])) + (kzu.v[i] + kzc.v[i]);
        a.v[i] = diag * cen.v[i] - (kxr * r + kxc.v[i] * l) - (kyn.v[i] * no.v[i] + kyc.v[i] * so.v[i]) -
                 (kzu.v[i] * up.v[i] + kzc.v[i] * dn.v[i]);
    }
    return a;
}

// The same two steps with the block's rows sharing level 0 and level 1
// through LDS: a block is ROWS + 2 waves, one per row j0-1 .. j0+ROWS; each
// wave forms level 1 of its own row only, and the waves of rows j0 ..
// j0+ROWS-1 write level 2.  Per plane: every wave publishes its row of sd,
// then (barrier) forms level 1 from its neighbours' rows and publishes it,
// then (barrier) forms level 2 of the plane below from the level-1 rows of
// its neighbours (double-buffered by plane parity).  T = float runs the
// FP32 fields of the FP32-refined solve (rows of `pitch` floats; the u role
// is the correction d), with the FP32 7-point group kernel's expressions.
template <class T, int ROWS>
__global__ void __launch_bounds__((ROWS + 2) * TL_WAVE)
k_ppcg3d_pair_lds(TlGrid3 g, int pitch, int nwx, int nby, int nbz, int tz, T a1, T b1, T a2,
                  T b2, const T* __restrict__ sd_in, T* __restrict__ sd_out,
                  const T* __restrict__ r_in, T* __restrict__ r_out, const T* __restrict__ u_in,
                  T* __restrict__ u_out, const T* __restrict__ kx, const T* __restrict__ ky,
                  const T* __restrict__ kz, TlGate gate)
{
    using C2 = P2<T>;
    constexpr int W = ROWS + 2;
    __shared__ C2 s0_row[W][TL_WAVE];
    __shared__ C2 s1_row[2][W][TL_WAVE];
    if (gate.closed()) return;                   // uniform per launch
    const int lane = threadIdx.x & (TL_WAVE - 1), wave = threadIdx.x / TL_WAVE;
    const int tile = xcd_tile(blockIdx.x, nwx * nby * nbz);
    const int wx = tile % nwx, rest = tile / nwx;
    const int by = rest % nby, bz = rest / nby;
    const int xin0 = g.hd, xin1 = g.x - g.hd, yin0 = g.hd, yin1 = g.y - g.hd;
    const int zin0 = g.hd, zin1 = g.z - g.hd;
    const int jw = yin0 + by * ROWS - 1 + wave;  // this wave's row
    const bool l1_row = jw >= yin0 && jw < yin1;
    const bool out_row = wave >= 1 && wave <= ROWS && l1_row;
    const int o0 = xin0 + wx * kP2Out, o1 = min(o0 + kP2Out, xin1);
    const int cw = (o0 - 2) >= 0 ? (o0 - 2) / 2 * 2 : -((3 - o0) / 2 * 2);
    const int c = cw + 2 * lane;
    const bool st0 = c >= o0 && c < o1, st1 = c + 1 >= o0 && c + 1 < o1;
    const int k0 = zin0 + bz * tz, k1 = min(k0 + tz, zin1);
    const long X = pitch, XY = (long)pitch * g.y;

    // own row: sd at planes kk-1, kk, kk+1; level 1 at planes km-1, km, km+1
    C2 s0w[3], s1w[3];
    s0w[0] = zeros2<T>();
    s0w[1] = ld3p(sd_in, g, pitch, jw, k0 - 2, c);
    s0w[2] = ld3p(sd_in, g, pitch, jw, k0 - 1, c);
    s1w[0] = s1w[1] = s1w[2] = zeros2<T>();
    C2 r1p = zeros2<T>(), u1p = zeros2<T>();
    C2 kxp = zeros2<T>(), kyp = zeros2<T>(), kynp = zeros2<T>();
    C2 kzp = zeros2<T>(), kzup = zeros2<T>();
    C2 kz_c = ld3p(kz, g, pitch, jw, k0 - 1, c);         // kz of the level-1 plane

    // What level 1 of plane kk reads from memory, loaded one plane ahead:
    // sd at plane kk+1, kx, ky (rows j, j+1), r and u at plane kk, kz at
    // kk+1, and for the block's first and last wave the row beyond it.
    struct PlaneIn {
        C2 sdn, kxc, kyc, kyn, kzu, r0, u0, outer;
    };
    auto load_plane = [&](int kk) -> PlaneIn {
        PlaneIn v;
        v.sdn = ld3p(sd_in, g, pitch, jw, kk + 1, c);
        v.kxc = ld3p(kx, g, pitch, jw, kk, c);
        v.kyc = ld3p(ky, g, pitch, jw, kk, c);
        v.kyn = ld3p(ky, g, pitch, jw + 1, kk, c);
        v.kzu = ld3p(kz, g, pitch, jw, kk + 1, c);
        v.r0 = ld3p(r_in, g, pitch, jw, kk, c);
        v.u0 = out_row ? ld3p(u_in, g, pitch, jw, kk, c) : zeros2<T>();
        v.outer = wave == 0 ? ld3p(sd_in, g, pitch, jw - 1, kk, c)
                            : (wave == W - 1 ? ld3p(sd_in, g, pitch, jw + 1, kk, c) : zeros2<T>());
        return v;
    };
    PlaneIn nxt = load_plane(k0 - 1);

    for (int kk = k0 - 1; kk <= k1; ++kk) {
        const PlaneIn cur = nxt;
        if (kk < k1) nxt = load_plane(kk + 1);    // uniform
        s0w[0] = s0w[1];
        s0w[1] = s0w[2];
        s0w[2] = cur.sdn;
        const C2 kxc = cur.kxc, kyc = cur.kyc, kyn = cur.kyn, kzu = cur.kzu;
        const C2 r0 = cur.r0, u0 = cur.u0, outer = cur.outer;
        s0_row[wave][lane] = s0w[1];
        __syncthreads();
        s1w[0] = s1w[1];
        s1w[1] = s1w[2];
        C2 r1 = zeros2<T>(), u1 = zeros2<T>();
        if (l1_row && kk >= zin0 && kk < zin1) {
            const C2 cen = s0w[1];
            const C2 so = jw == yin0 ? cen : (wave == 0 ? outer : s0_row[wave - 1][lane]);
            const C2 no = jw == yin1 - 1 ? cen : (wave == W - 1 ? outer : s0_row[wave + 1][lane]);
            const C2 up = kk == zin1 - 1 ? cen : s0w[2];
            const C2 dn = kk == zin0 ? cen : s0w[0];
            const C2 w = apply7t(cen, no, so, up, dn, kxc, kyc, kyn, kz_c, kzu, c, xin0, xin1);
            C2 sn;
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                r1.v[i] = r0.v[i] - w.v[i];
                sn.v[i] = a1 * cen.v[i] + b1 * r1.v[i];
                u1.v[i] = u0.v[i] + cen.v[i];
            }
            s1w[2] = sn;
        }
        s1_row[kk & 1][wave][lane] = s1w[2];
        __syncthreads();
        // level 2 at plane km = kk - 1
        const int km = kk - 1;
        if (out_row && km >= k0 && km < k1) {
            const C2 cen = s1w[1];
            const C2 so = jw == yin0 ? cen : s1_row[km & 1][wave - 1][lane];
            const C2 no = jw == yin1 - 1 ? cen : s1_row[km & 1][wave + 1][lane];
            const C2 up = km == zin1 - 1 ? cen : s1w[2];
            const C2 dn = km == zin0 ? cen : s1w[0];
            const C2 w = apply7t(cen, no, so, up, dn, kxp, kyp, kynp, kzp, kzup, c, xin0, xin1);
            C2 rn, un, sn;
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                rn.v[i] = r1p.v[i] - w.v[i];
                un.v[i] = u1p.v[i] + cen.v[i];
                sn.v[i] = a2 * cen.v[i] + b2 * rn.v[i];
            }
            const long b = (long)km * XY + (long)jw * X + c;
            if (st0 && st1) {
                st2(r_out, b, rn);
                st2(u_out, b, un);
                st2(sd_out, b, sn);
            } else {
#pragma unroll
                for (int i = 0; i < 2; ++i)
                    if (i == 0 ? st0 : st1) {
                        r_out[b + i] = rn.v[i];
                        u_out[b + i] = un.v[i];
                        sd_out[b + i] = sn.v[i];
                    }
            }
        }
        r1p = r1;
        u1p = u1;
        kxp = kxc;
        kyp = kyc;
        kynp = kyn;
        kzp = kz_c;
        kzup = kzu;
        kz_c = kzu;
    }
}

// The CG pair in 3D with kx, ky and kz formed from cg_init's conductivity
// (cond: density, or 1/density with the reciprocal coefficient;
// 3d/c_kernels/omp3/cg_solver.c:47-75) instead of loaded, each face
// coefficient with cg_init's expression on the same operands (tl_face_coef):
// one 8-byte field replaces three.  A block is kC3Rows waves, one per row j0
// .. j0+kC3Rows-1 of a 128-column window (122 columns written, windows
// overlap as in the two-step kernel), marching up a tile of planes.  Per
// plane every wave publishes its row of the stencil field and of cond to LDS
// (double-buffered by plane parity: one barrier per plane); its y
// neighbours come from the adjacent waves (the block's first and last waves
// load the rows beyond the block), its z neighbours and kz from its own
// planes k-1, k, k+1 in registers.  Neighbours outside the interior are the
// cell (the depth-1 reflective halo update before each pass).
// UR false: calc_p fused into calc_w: the stencil field is p' = beta p + r
// with UPDATE (stored to p_out; cg_solver.c:172-194 then :105-135), else
// p; w = A p' stored unless w is null; sum w p'.
// UR true: cg_calc_ur (cg_solver.c:138-170) with w = A p formed here:
// u += alpha p, r -= alpha w, sum r r.
// Every cell's values are the reference's bits; the sums differ only by
// summation order.  Needs halo depth >= 2 (cond is read at hd-1 .. n-hd)
// and an even x.
constexpr int kC3Rows = 8;

template <bool UR, bool UPDATE>
__global__ void __launch_bounds__(kC3Rows * TL_WAVE)
k_cg3_cond(TlGrid3 g, int nwx, int nby, int nbz, int tz, double sc_arg,
           const double* __restrict__ sc_dev, const double* __restrict__ p,
           double* __restrict__ p_out, double* __restrict__ r, double* __restrict__ w,
           double* __restrict__ u, const double* __restrict__ cond, double rx, double ry,
           double rz, TlReduce rd, TlGate gate)
{
    constexpr int R = kC3Rows;
    __shared__ C2 s_f[2][R][TL_WAVE];
    __shared__ C2 s_w[2][R][TL_WAVE];
    if (gate.closed()) return;                   // uniform per launch
    const double sc = sc_dev ? *sc_dev : sc_arg; // beta (UPDATE) or alpha (UR)
    const int lane = threadIdx.x & (TL_WAVE - 1), wave = threadIdx.x / TL_WAVE;
    const int tile = xcd_tile(blockIdx.x, nwx * nby * nbz);
    const int wx = tile % nwx, rest = tile / nwx;
    const int by = rest % nby, bz = rest / nby;
    const int xin0 = g.hd, xin1 = g.x - g.hd, yin0 = g.hd, yin1 = g.y - g.hd;
    const int zin0 = g.hd, zin1 = g.z - g.hd;
    const int jw = yin0 + by * R + wave;         // this wave's row
    const bool row_ok = jw < yin1;
    const int o0 = xin0 + wx * kP2Out, o1 = min(o0 + kP2Out, xin1);
    const int cw = (o0 - 2) >= 0 ? (o0 - 2) / 2 * 2 : -((3 - o0) / 2 * 2);
    const int c = cw + 2 * lane;
    const bool st0 = c >= o0 && c < o1, st1 = c + 1 >= o0 && c + 1 < o1;
    const int k0 = zin0 + bz * tz, k1 = min(k0 + tz, zin1);
    const long X = g.x, XY = (long)g.x * g.y;
    // Loads without lane-divergent branches: a lane past the row's end
    // reads the row's last pair (it writes nothing, and no written cell
    // takes its values: the last interior cell mirrors), a row past the
    // grid the grid's last row (only for rows that write nothing); rows
    // jw-1 >= hd-1 and planes k0-1 .. k1+1 lie in the grid.
    const int cc = min(c, g.x - 2);
    auto ldr = [&](const double* __restrict__ f, int j, int k) -> C2 {
        return ldc(f, ((long)k * g.y + min(j, g.y - 1)) * X + cc);
    };
    // the stencil field from its raw loads: p' = beta p + r with UPDATE, p
    struct Raw {
        C2 p, r;
    };
    auto raw = [&](int j, int k) -> Raw {
        Raw a{ldr(p, j, k), zeros<double>()};
        if constexpr (UPDATE) a.r = ldr(r, j, k);
        return a;
    };
    auto comb = [&](const Raw& a) -> C2 {
        C2 o = a.p;
        if constexpr (UPDATE) {
#pragma unroll
            for (int i = 0; i < 2; ++i) o.v[i] = sc * a.p.v[i] + a.r.v[i];
        }
        return o;
    };
    auto pick = [](bool t, const C2& a, const C2& b) -> C2 {
        C2 o;
#pragma unroll
        for (int i = 0; i < 2; ++i) o.v[i] = t ? a.v[i] : b.v[i];
        return o;
    };
    const int wso = max(wave - 1, 0), wno = min(wave + 1, R - 1);
    const int jo = wave == 0 ? jw - 1 : jw + 1;  // the row beyond the block (first, last wave)
    const bool edge = wave == 0 || wave == R - 1;
    C2 fc = comb(raw(jw, k0));
    C2 fm = pick(k0 == zin0, fc, comb(raw(jw, k0 - 1)));
    C2 wc = ldr(cond, jw, k0);
    C2 kzc = ky_of_rows(rz, ldr(cond, jw, k0 - 1), wc);
    // loads run one plane ahead: the own row's plane kk+2 and the outer
    // row's plane kk+1 are in flight while plane kk is formed (planes up to
    // k1+1 <= z-hd+1 lie in the grid for halo depth >= 2)
    Raw rp = raw(jw, k0 + 1), ro = edge ? raw(jo, k0) : Raw{fc, zeros<double>()};
    C2 wp = ldr(cond, jw, k0 + 1), wo = edge ? ldr(cond, jo, k0) : wc;
    C2 uu = zeros<double>(), rr = zeros<double>();   // UR: u and r of plane kk, one ahead
    if constexpr (UR) {
        uu = ldr(u, jw, k0);
        rr = ldr(r, jw, k0);
    }
    double acc[1] = {0.0};
    for (int kk = k0; kk < k1; ++kk) {
        const int b = kk & 1;
        const Raw rp_k = rp, ro_k = ro;
        const C2 wp_k = wp, wo_k = wo, uu_k = uu, rr_k = rr;
        rp = raw(jw, kk + 2);
        wp = ldr(cond, jw, kk + 2);
        if constexpr (UR) {
            uu = ldr(u, jw, kk + 1);
            rr = ldr(r, jw, kk + 1);
        }
        if (edge) {                              // uniform per wave
            ro = raw(jo, kk + 1);
            wo = ldr(cond, jo, kk + 1);
        }
        s_f[b][wave][lane] = fc;
        s_w[b][wave][lane] = wc;
        __syncthreads();
        const C2 fs = s_f[b][wso][lane], fn = s_f[b][wno][lane];
        const C2 ws_l = s_w[b][wso][lane], wn_l = s_w[b][wno][lane];
        const C2 fp = comb(rp_k);
        if (row_ok) {                            // uniform per wave
            const C2 fo = comb(ro_k);
            const C2 so = pick(jw == yin0, fc, pick(wave == 0, fo, fs));
            const C2 no = pick(jw == yin1 - 1, fc, pick(wave == R - 1, fo, fn));
            const C2 ws = pick(wave == 0, wo_k, ws_l);
            const C2 wn = pick(wave == R - 1, wo_k, wn_l);
            const C2 up = pick(kk == zin1 - 1, fc, fp);
            const C2 dn = pick(kk == zin0, fc, fm);
            const C2 wpk = wp_k;
            const C2 kxc = kx_of_row(rx, wc, wc.v[0]);   // the wave's first lane writes nothing
            const C2 kyc = ky_of_rows(ry, ws, wc), kyn = ky_of_rows(ry, wc, wn);
            const C2 kzu = ky_of_rows(rz, wc, wpk);
            const C2 av = apply7(fc, no, so, up, dn, kxc, kyc, kyn, kzc, kzu, c, xin0, xin1);
            const long i0 = (long)kk * XY + (long)jw * X + c;
            if constexpr (!UR) {
                if (st0 && st1) {
                    if (w) stc(w, i0, av);
                    if constexpr (UPDATE) stc(p_out, i0, fc);
#pragma unroll
                    for (int i = 0; i < 2; ++i) acc[0] += av.v[i] * fc.v[i];
                } else {
#pragma unroll
                    for (int i = 0; i < 2; ++i)
                        if (i == 0 ? st0 : st1) {
                            if (w) w[i0 + i] = av.v[i];
                            if constexpr (UPDATE) p_out[i0 + i] = fc.v[i];
                            acc[0] += av.v[i] * fc.v[i];
                        }
                }
            } else {
                C2 un, rn;
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    un.v[i] = uu_k.v[i] + sc * fc.v[i];
                    rn.v[i] = rr_k.v[i] - sc * av.v[i];
                }
                if (st0 && st1) {
                    stc(u, i0, un);
                    stc(r, i0, rn);
#pragma unroll
                    for (int i = 0; i < 2; ++i) acc[0] += rn.v[i] * rn.v[i];
                } else {
#pragma unroll
                    for (int i = 0; i < 2; ++i)
                        if (i == 0 ? st0 : st1) {
                            u[i0 + i] = un.v[i];
                            r[i0 + i] = rn.v[i];
                            acc[0] += rn.v[i] * rn.v[i];
                        }
                }
            }
            kzc = kzu;
        }
        fm = fc;
        fc = fp;
        wc = wp_k;
    }
    tl_reduce_partial<1, kC3Rows * TL_WAVE>(acc, rd);
}

// cond = cg_init's w over the whole grid (cg_solver.c:47-59).
__global__ void __launch_bounds__(kBlock)
k_cond3d(long n, int coef, const double* __restrict__ density, double* __restrict__ cond)
{
    for (long c = blockIdx.x * (long)kBlock + threadIdx.x; c < n; c += (long)gridDim.x * kBlock)
        cond[c] = coef == 1 ? density[c] : 1.0 / density[c];
}

// Checks and the k_cg3_cond launch under the current gate (ring: p_in's
// halo ring copied to p_out first, as the entry point's form keeps it);
// *nb: its partials.
int launch_cg3_cond(const char* what, const TlGrid3& g, bool ur, bool update, double sc,
                    const double* sc_dev, const double* p, double* p_out, double* r, double* w,
                    double* u, const double* cond, double rx, double ry, double rz,
                    unsigned* nb, bool ring)
{
    TL_TRY(tl_ensure_init());
    TL_TRY(check_grid(g, what));
    if (g.hd < 2 || g.x % 2)
        return tl_fail(TLHIP_ERR_ARG, "%s: needs halo depth >= 2 and an even x (x=%d hd=%d)",
                       what, g.x, g.hd);
    if (interior_empty(g)) return tl_fail(TLHIP_ERR_ARG, "%s: empty interior", what);
    REQUIRE(p && cond, what, "null argument");
    if (ur) {
        REQUIRE(u && r, what, "null argument");
        REQUIRE(u != p && u != r && r != p && u != cond && r != cond, what,
                "u, p, r and cond must be distinct buffers");
    } else if (update) {
        REQUIRE(p_out && r, what, "null argument");
        REQUIRE(p_out != p && p_out != r && p_out != cond && p_out != w && w != p && w != r &&
                    w != cond, what, "p_out and w must be distinct from the inputs");
    } else {
        REQUIRE(w != p && w != cond, what, "w must be distinct from the inputs");
    }
    for (const double* f : {p, (const double*)p_out, (const double*)r, (const double*)w,
                            (const double*)u, cond})
        REQUIRE(((uintptr_t)f & 15) == 0, what, "fields must be 16-byte aligned");
    const int nx = g.x - 2 * g.hd, ny = g.y - 2 * g.hd, nz = g.z - 2 * g.hd;
    // planes per tile: tlhip_set_tile_planes when set, else 16 (256^3:
    // 0.160 + 0.189 ms for the pair against 0.175 + 0.195 at 8 planes and
    // 0.167 + 0.213 at 32)
    const int tz = g_tz_override > 0 ? g_tz_override : 16;
    const int nwx = (int)tl_cdiv(nx, kP2Out), nby = (int)tl_cdiv(ny, kC3Rows);
    const int nbz = (int)tl_cdiv(nz, tz);
    const long nbl = (long)nwx * nby * nbz;
    if (nbl > TL_MAX_RED_BLOCKS) return tl_fail(TLHIP_ERR_ARG, "%s: grid too large", what);
    *nb = (unsigned)nbl;
    if (!ur && update && ring) TL_TRY(copy_ring(g, p, p_out, what));
    const char* label = ur ? what : (w ? what : "tlhip3d_cg_calc_pw_cond.no_w");
    const int slot = tl_timer_begin(label);
#define TL_C3(U, P)                                                                            \
    hipLaunchKernelGGL((k_cg3_cond<U, P>), dim3(*nb), dim3(kC3Rows * TL_WAVE), 0, tl_stream(), \
                       g, nwx, nby, nbz, tz, sc, sc_dev, p, p_out, r, w, u, cond, rx, ry, rz,  \
                       tl_reduce_slots(), tl_gate())
    if (ur) TL_C3(true, false);
    else if (update) TL_C3(false, true);
    else TL_C3(false, false);
#undef TL_C3
    tl_timer_end(slot);
    return tl_check_launch(what);
}

}  // namespace

extern "C" int tlhip3d_ppcg_inner_iteration(int x, int y, int z, int hd, double alpha,
                                            double beta, double* u, double* r, const double* kx,
                                            const double* ky, const double* kz,
                                            const double* sd_in, double* sd_out)
{
    const char* what = "tlhip3d_ppcg_inner_iteration";
    TL_TRY(tl_ensure_init());
    const TlGrid3 g{x, y, z, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(u && r && kx && ky && kz && sd_in && sd_out, what, "null argument");
    REQUIRE(sd_in != sd_out, what, "sd_in and sd_out must be distinct buffers");
    OpPpcgInner op{r, u, sd_out, alpha, beta};
    TL_TRY(launch_stencil(g, pick_v(g, u, r, kx, ky, kz, sd_in, sd_out), sd_in, kx, ky, kz, op,
                          what));
    return copy_ring(g, sd_in, sd_out, what);
}

extern "C" int tlhip_set_pair_rows(int rows)
{
    if (rows != 0 && rows != 4 && rows != 6 && rows != 8 && rows != 14)
        return tl_fail(TLHIP_ERR_ARG, "tlhip_set_pair_rows: 0, 4, 6, 8 or 14 (got %d)", rows);
    g_pair_rows = rows;
    return 0;
}

// Two PPCG inner steps per launch; planes per z-tile: tlhip_set_tile_planes
// when set, else 32 (256^3: 0.384 ms against 0.40 at 16 and 0.46 at 64).
extern "C" int tlhip3d_ppcg_inner_iterations(int x, int y, int z, int hd, int count,
                                             const double* alphas, const double* betas,
                                             const int* neighbours, const double* sd_in,
                                             double* sd_out, const double* r_in, double* r_out,
                                             const double* u_in, double* u_out,
                                             const double* kx, const double* ky,
                                             const double* kz)
{
    const char* what = "tlhip3d_ppcg_inner_iterations";
    TL_TRY(tl_ensure_init());
    const TlGrid3 g{x, y, z, hd};
    TL_TRY(check_grid(g, what));
    if (count != 2) return tl_fail(TLHIP_ERR_ARG, "%s: count must be 2 (got %d)", what, count);
    if (x % 2) return tl_fail(TLHIP_ERR_ARG, "%s: needs an even x (got %d)", what, x);
    REQUIRE(alphas && betas && neighbours, what, "null coefficients/neighbours");
    REQUIRE(sd_in && sd_out && r_in && r_out && u_in && u_out && kx && ky && kz, what,
            "null argument");
    REQUIRE(sd_in != sd_out && r_in != r_out && u_in != u_out, what,
            "input and output buffers must be distinct");
    for (int f = 0; f < 6; ++f)
        REQUIRE(neighbours[f] == -1, what, "every face must be external (single chunk)");
    if (interior_empty(g)) return tl_fail(TLHIP_ERR_ARG, "%s: empty interior", what);
    for (const double* f : {sd_in, (const double*)sd_out, r_in, (const double*)r_out, u_in,
                            (const double*)u_out, kx, ky, kz})
        REQUIRE(((uintptr_t)f & 15) == 0, what, "fields must be 16-byte aligned");
    const int nx = x - 2 * hd, ny = y - 2 * hd, nz = z - 2 * hd;
    const int tz = g_tz_override > 0 ? g_tz_override : 32;
    const int rows = g_pair_rows;                 // 0: each wave forms its three level-1 rows
    const int nwx = (int)tl_cdiv(nx, kP2Out), nby = (int)tl_cdiv(ny, rows > 0 ? rows : kP2Rows);
    const int nbz = (int)tl_cdiv(nz, tz);
    const long nb = (long)nwx * nby * nbz;
    if (nb > (1L << 30)) return tl_fail(TLHIP_ERR_ARG, "%s: grid too large", what);
    TL_TRY(copy_ring(g, sd_in, sd_out, what));
    TL_TRY(copy_ring(g, r_in, r_out, what));
    TL_TRY(copy_ring(g, u_in, u_out, what));
    const int slot = tl_timer_begin(what);
#define TL_PAIR_ARGS nwx, nby, nbz, tz, alphas[0], betas[0], alphas[1], betas[1], sd_in, \
        sd_out, r_in, r_out, u_in, u_out, kx, ky, kz, tl_gate()
    switch (rows) {
    case 4: hipLaunchKernelGGL((k_ppcg3d_pair_lds<double, 4>), dim3((unsigned)nb),
                               dim3(6 * TL_WAVE), 0, tl_stream(), g, x, TL_PAIR_ARGS); break;
    case 6: hipLaunchKernelGGL((k_ppcg3d_pair_lds<double, 6>), dim3((unsigned)nb),
                               dim3(8 * TL_WAVE), 0, tl_stream(), g, x, TL_PAIR_ARGS); break;
    case 8: hipLaunchKernelGGL((k_ppcg3d_pair_lds<double, 8>), dim3((unsigned)nb),
                               dim3(10 * TL_WAVE), 0, tl_stream(), g, x, TL_PAIR_ARGS); break;
    case 14: hipLaunchKernelGGL((k_ppcg3d_pair_lds<double, 14>), dim3((unsigned)nb),
                                dim3(16 * TL_WAVE), 0, tl_stream(), g, x, TL_PAIR_ARGS); break;
    default: hipLaunchKernelGGL(k_ppcg3d_pair, dim3((unsigned)nb), dim3(kBlock), 0, tl_stream(),
                                g, TL_PAIR_ARGS); break;
    }
#undef TL_PAIR_ARGS
    tl_timer_end(slot);
    return tl_check_launch(what);
}

// ---------------------------------------------------------------- shared

extern "C" int tlhip3d_solver_copy_u(int x, int y, int z, int hd, double* u0, const double* u)
{
    const char* what = "tlhip3d_solver_copy_u";
    TL_TRY(tl_ensure_init());
    const TlGrid3 g{x, y, z, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(u0 && u, what, "null argument");
    OpSimple op{u0, u, 0, nullptr, 0.0};
    return launch_interior3(g, pick_v(g, u0, u), op, what);
}

extern "C" int tlhip3d_calculate_residual(int x, int y, int z, int hd, const double* u,
                                          const double* u0, double* r, const double* kx,
                                          const double* ky, const double* kz)
{
    const char* what = "tlhip3d_calculate_residual";
    TL_TRY(tl_ensure_init());
    const TlGrid3 g{x, y, z, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(u && u0 && r && kx && ky && kz, what, "null argument");
    REQUIRE(u != r, what, "u and r must be distinct buffers");
    OpResidual op{u0, r};
    return launch_stencil(g, pick_v(g, u, u0, r, kx, ky, kz), u, kx, ky, kz, op, what);
}

extern "C" int tlhip3d_calculate_2norm(int x, int y, int z, int hd, const double* buffer,
                                       double* norm)
{
    const char* what = "tlhip3d_calculate_2norm";
    TL_TRY(tl_ensure_init());
    const TlGrid3 g{x, y, z, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(buffer && norm, what, "null argument");
    if (interior_empty(g)) return 0;
    OpNorm op{buffer};
    TL_TRY(launch_interior3(g, pick_v(g, buffer), op, what));
    return fetch_add(norm, what);
}

extern "C" int tlhip3d_solver_finalise(int x, int y, int z, int hd, double* energy,
                                       const double* density, const double* u)
{
    const char* what = "tlhip3d_solver_finalise";
    TL_TRY(tl_ensure_init());
    const TlGrid3 g{x, y, z, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(energy && density && u, what, "null argument");
    OpSimple op{energy, u, 3, density, 0.0};
    return launch_interior3(g, pick_v(g, energy, density, u), op, what);
}

// ---------------------------------------------------------------- resident loop
// The CG pieces of one iteration of a device-resident loop (tl_loop.h), as
// in the 2D file.

namespace tl {
namespace loop3d {

int calc_w(int x, int y, int z, int hd, const double* p, double* w, const double* kx,
           const double* ky, const double* kz)
{
    const char* what = "tlhip3d_cg_solver_calc_w";
    const TlGrid3 g{x, y, z, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(p && w && kx && ky && kz && p != w, what, "bad arguments");
    if (interior_empty(g)) return tl_fail(TLHIP_ERR_ARG, "%s: empty interior", what);
    unsigned nb = 0;
    TL_TRY(launch_stencil(g, pick_v(g, p, w, kx, ky, kz), p, kx, ky, kz, OpCalcW{w}, what, &nb));
    return loop::finish_alpha(nb, what);
}

int calc_ur(int x, int y, int z, int hd, double* u, const double* p, double* r, const double* w,
            const loop::StopRule& rule)
{
    const char* what = "tlhip3d_cg_solver_calc_ur";
    const TlGrid3 g{x, y, z, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(u && p && r && w, what, "null argument");
    if (interior_empty(g)) return tl_fail(TLHIP_ERR_ARG, "%s: empty interior", what);
    OpCalcUr op{u, p, r, w, 0.0, loop::alpha()};
    unsigned nb = 0;
    TL_TRY(launch_interior3(g, pick_v(g, u, p, r, w), op, what, &nb));
    return loop::finish_beta(nb, rule, what);
}

// cg_calc_ur of a PPCG outer iteration, whose rrn the driver does not use
// (3d/drivers/ppcg_solver_driver.c:107-123): no finisher, the partials are
// left unread.
int calc_ur_only(int x, int y, int z, int hd, double* u, const double* p, double* r,
                 const double* w)
{
    const char* what = "tlhip3d_cg_solver_calc_ur";
    const TlGrid3 g{x, y, z, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(u && p && r && w, what, "null argument");
    if (interior_empty(g)) return tl_fail(TLHIP_ERR_ARG, "%s: empty interior", what);
    OpCalcUr op{u, p, r, w, 0.0, loop::alpha()};
    unsigned nb = 0;
    return launch_interior3(g, pick_v(g, u, p, r, w), op, what, &nb);
}

int norm(int x, int y, int z, int hd, const double* buf, const loop::StopRule& rule)
{
    const char* what = "tlhip3d_calculate_2norm";
    const TlGrid3 g{x, y, z, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(buf, what, "null argument");
    if (interior_empty(g)) return tl_fail(TLHIP_ERR_ARG, "%s: empty interior", what);
    unsigned nb = 0;
    TL_TRY(launch_interior3(g, pick_v(g, buf), OpNorm{buf}, what, &nb));
    return loop::finish_beta(nb, rule, what);
}

int calc_p(int x, int y, int z, int hd, double* p, const double* r)
{
    const char* what = "tlhip3d_cg_solver_calc_p";
    const TlGrid3 g{x, y, z, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(p && r, what, "null argument");
    OpCalcP op{p, r, 0.0, loop::beta()};
    return launch_interior3(g, pick_v(g, p, r), op, what);
}

int jacobi(int x, int y, int z, int hd, const double* kx, const double* ky, const double* kz,
           const double* u0, const double* u_in, double* u_out, loop::ErrMode mode,
           const loop::StopRule& rule)
{
    const char* what = "tlhip3d_jacobi_solver_iterate";
    const TlGrid3 g{x, y, z, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(kx && ky && kz && u0 && u_in && u_out && u_in != u_out, what, "bad arguments");
    if (interior_empty(g)) return tl_fail(TLHIP_ERR_ARG, "%s: empty interior", what);
    OpJacobi3 op{u0, u_out};
    unsigned nb = 0;
    TL_TRY(launch_stencil(g, pick_v(g, kx, ky, kz, u0, u_in, u_out), u_in, kx, ky, kz, op, what,
                          &nb));
    return loop::finish_err(nb, mode, rule, what);
}

int norm_err(int x, int y, int z, int hd, const double* buf, const loop::StopRule& rule)
{
    const char* what = "tlhip3d_calculate_2norm";
    const TlGrid3 g{x, y, z, hd};
    TL_TRY(check_grid(g, what));
    REQUIRE(buf, what, "null argument");
    if (interior_empty(g)) return tl_fail(TLHIP_ERR_ARG, "%s: empty interior", what);
    unsigned nb = 0;
    TL_TRY(launch_interior3(g, pick_v(g, buf), OpNorm{buf}, what, &nb));
    return loop::finish_err(nb, loop::kErrAddTest, rule, what);
}

// The cond pair in a resident loop (tl_loop.h).
int calc_pw_cond(int x, int y, int z, int hd, bool update, const double* p_in, double* p_out,
                 const double* r, double* w, const double* cond, double rx, double ry, double rz)
{
    const char* what = "tlhip3d_cg_calc_pw_cond";
    unsigned nb = 0;
    TL_TRY(launch_cg3_cond(what, TlGrid3{x, y, z, hd}, false, update, 0.0,
                           update ? loop::beta() : nullptr, p_in, p_out, (double*)r, w, nullptr,
                           cond, rx, ry, rz, &nb, false));
    return loop::finish_alpha(nb, what);
}

int calc_ur_cond(int x, int y, int z, int hd, double* u, const double* p, double* r,
                 const double* cond, double rx, double ry, double rz,
                 const loop::StopRule& rule)
{
    const char* what = "tlhip3d_cg_calc_ur_cond";
    unsigned nb = 0;
    TL_TRY(launch_cg3_cond(what, TlGrid3{x, y, z, hd}, true, false, 0.0, loop::alpha(), p,
                           nullptr, r, nullptr, u, cond, rx, ry, rz, &nb, false));
    return loop::finish_beta(nb, rule, what);
}

}  // namespace loop3d
}  // namespace tl

// ------------------------------------------------ CG from the conductivity

extern "C" int tlhip3d_cg_cond(int x, int y, int z, int hd, int coefficient,
                               const double* density, double* cond)
{
    const char* what = "tlhip3d_cg_cond";
    TL_TRY(tl_ensure_init());
    if (x < 1 || y < 1 || z < 1 || hd < 1) return tl_fail(TLHIP_ERR_ARG, "%s: bad grid", what);
    if (coefficient != 1 && coefficient != 2)
        return tl_fail(TLHIP_ERR_ARG, "%s: coefficient %d is not valid", what, coefficient);
    REQUIRE(density && cond && density != cond, what, "bad arguments");
    const long n = (long)x * y * z;
    hipLaunchKernelGGL(k_cond3d, dim3(flat_blocks(n)), dim3(kBlock), 0, tl_stream(), n,
                       coefficient, density, cond);
    return tl_check_launch(what);
}

extern "C" int tlhip3d_cg_calc_pw_cond(int x, int y, int z, int hd, int update, double beta,
                                       double rx, double ry, double rz, double* pw,
                                       const double* p_in, double* p_out, const double* r,
                                       double* w, const double* cond)
{
    const char* what = "tlhip3d_cg_calc_pw_cond";
    REQUIRE(pw, what, "null argument");
    unsigned nb = 0;
    TL_TRY(launch_cg3_cond(what, TlGrid3{x, y, z, hd}, false, update != 0, beta, nullptr, p_in,
                           p_out, (double*)r, w, nullptr, cond, rx, ry, rz, &nb, true));
    TL_TRY(tl_finish_reduce(1, nb, what));
    return fetch_add(pw, what);
}

extern "C" int tlhip3d_cg_calc_ur_cond(int x, int y, int z, int hd, double alpha, double rx,
                                       double ry, double rz, double* rrn, double* u,
                                       const double* p, double* r, const double* cond)
{
    const char* what = "tlhip3d_cg_calc_ur_cond";
    REQUIRE(rrn, what, "null argument");
    unsigned nb = 0;
    TL_TRY(launch_cg3_cond(what, TlGrid3{x, y, z, hd}, true, false, alpha, nullptr, p, nullptr,
                           r, nullptr, u, cond, rx, ry, rz, &nb, false));
    TL_TRY(tl_finish_reduce(1, nb, what));
    return fetch_add(rrn, what);
}

// ------------------------------------------------ FP32 two-step PPCG (3D)

namespace tl {
namespace mix {

// Two FP32 PPCG inner steps in one pass (tl_mixed.h): the 8-row LDS
// kernel on FP32 fields of row pitch P; r, d and sd from *_in to *_out
// (interiors only: no FP32 kernel reads a halo), bit-identical to two
// ppcg_inner3_32 turns.
int ppcg_pair3_32(int x, int y, int z, int hd, int P, const double* alphas,
                  const double* betas, const float* sd_in, float* sd_out, const float* r_in,
                  float* r_out, const float* d_in, float* d_out, const float* kx,
                  const float* ky, const float* kz)
{
    const char* what = "tlhip3d_mix_ppcg_pair32";
    const TlGrid3 g{x, y, z, hd};
    TL_TRY(check_grid(g, what));
    if (x % 2 || P < x || P % 4)
        return tl_fail(TLHIP_ERR_ARG, "%s: needs an even x and a pitch >= x, multiple of 4",
                       what);
    if (interior_empty(g)) return tl_fail(TLHIP_ERR_ARG, "%s: empty interior", what);
    REQUIRE(sd_in && sd_out && r_in && r_out && d_in && d_out && kx && ky && kz, what,
            "null argument");
    REQUIRE(sd_in != sd_out && r_in != r_out && d_in != d_out, what,
            "input and output buffers must be distinct");
    const int nx = x - 2 * hd, ny = y - 2 * hd, nz = z - 2 * hd;
    const int tz = g_tz_override > 0 ? g_tz_override : 32;
    const int nwx = (int)tl_cdiv(nx, kP2Out), nby = (int)tl_cdiv(ny, 8);
    const int nbz = (int)tl_cdiv(nz, tz);
    const long nb = (long)nwx * nby * nbz;
    if (nb > (1L << 30)) return tl_fail(TLHIP_ERR_ARG, "%s: grid too large", what);
    const int slot = tl_timer_begin(what);
    hipLaunchKernelGGL((k_ppcg3d_pair_lds<float, 8>), dim3((unsigned)nb), dim3(10 * TL_WAVE), 0,
                       tl_stream(), g, P, nwx, nby, nbz, tz, (float)alphas[0], (float)betas[0],
                       (float)alphas[1], (float)betas[1], sd_in, sd_out, r_in, r_out, d_in, d_out,
                       kx, ky, kz, tl_gate());
    tl_timer_end(slot);
    return tl_check_launch(what);
}

}  // namespace mix
}  // namespace tl
