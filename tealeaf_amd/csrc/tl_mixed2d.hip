// FP32 kernels of the mixed-precision solve (FP32 Krylov sweeps inside an
// FP64 residual-replacement loop), 2D and 3D.
//
// The reference solves in FP64 throughout (2d/drivers/cg_driver.c,
// ppcg_driver.c).  Here the CG and PPCG vector sweeps run on FP32 copies of
// kx, ky, p, w, r, sd and of the correction d accumulated since the last
// FP64 update, halving their HBM traffic; every dot product is accumulated
// in FP64; the solution u, the right-hand side u0 and the true residual
// u0 - A u stay FP64.
//
// FP32 arrays have a row pitch P >= x that is a multiple of 4, so that a lane
// moves four cells with one 16-byte access; interior cells are [hd, x-hd) x
// [hd, y-hd) as in the FP64 arrays.  No kernel here reads an FP32 halo: a
// stencil neighbour outside the interior takes the cell's own value, which
// is what the reference's depth-1 reflective halo update provides on a
// single chunk (2d/c_kernels/omp3/local_halos.c).
#include "tl_lanes.h"
#include "tl_mixed.h"
#include "tl_ops.h"
#include "tealeaf_hip.h"

using namespace tl;

namespace {

using C4 = Cells<float>;

// ---------------------------------------------------------------- pointwise
// Thread q of a row covers the 4-cell group [4q, 4q+4) of the pitched row;
// groups that meet the interior are visited, cells outside it are masked.
// 3D fields are rows of x cells too: row (k*y + j) holds plane k's row j,
// and the interior rows are those with j and k inside (z <= 1: 2D).
struct Groups {
    int hd, x, P;      // FP32 pitch P
    int q0, nq;        // first group, groups per row
    int y0, ny;        // interior rows of a plane
    int Y, z, z0, nz;  // rows per plane, planes, interior planes
    __host__ __device__ long total() const { return (long)nq * ny * nz; }
    __device__ int row_of(long r) const { return (z0 + (int)(r / ny)) * Y + y0 + (int)(r % ny); }
};

Groups groups_of(int x, int y, int z, int hd, int P)
{
    Groups G;
    G.hd = hd;
    G.x = x;
    G.P = P;
    G.q0 = hd / 4;
    G.nq = (x - hd + 3) / 4 - G.q0;
    G.y0 = hd;
    G.ny = y - 2 * hd;
    G.Y = y;
    G.z = z > 1 ? z : 1;
    G.z0 = z > 1 ? hd : 0;
    G.nz = z > 1 ? z - 2 * hd : 1;
    return G;
}

template <class Op>
__global__ void __launch_bounds__(kBlock) k_groups(Groups G, Op op, TlReduce rd)
{
    double acc[Op::NRED > 0 ? Op::NRED : 1] = {};
    for (long t = blockIdx.x * (long)kBlock + threadIdx.x; t < G.total();
         t += (long)gridDim.x * kBlock) {
        const int row = G.row_of(t / G.nq);
        const int c = 4 * (G.q0 + (int)(t % G.nq));
        unsigned mask = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (c + i >= G.hd && c + i < G.x - G.hd) mask |= 1u << i;
        op.apply(row, c, mask, acc);
    }
    if constexpr (Op::NRED > 0) tl_reduce_partial<Op::NRED, kBlock>(acc, rd);
}

// d += alpha p; r -= alpha w; rrn += r*r   (cg.c:116-143 in FP32, FP64 sum)
struct OpUr32 {
    static constexpr int NRED = 1;
    int P;
    float alpha;
    float* d;
    const float* p;
    float* r;
    const float* w;
    __device__ void apply(int row, int c, unsigned mask, double* acc) const
    {
        const long i0 = (long)row * P + c;
        C4 dv = ldc(d, i0), pv = ldc(p, i0), rv = ldc(r, i0), wv = ldc(w, i0);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (!(mask >> i & 1)) continue;
            dv.v[i] = dv.v[i] + alpha * pv.v[i];
            rv.v[i] = rv.v[i] - alpha * wv.v[i];
            acc[0] += (double)rv.v[i] * (double)rv.v[i];
        }
        if (mask == 15u) {
            stc(d, i0, dv);
            stc(r, i0, rv);
        } else {
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (mask >> i & 1) {
                    d[i0 + i] = dv.v[i];
                    r[i0 + i] = rv.v[i];
                }
        }
    }
};

// p = beta p + r (cg.c:146-164), or p = r when init
struct OpP32 {
    static constexpr int NRED = 0;
    int P;
    float beta;
    int init;
    float* p;
    const float* r;
    __device__ void apply(int row, int c, unsigned mask, double*) const
    {
        const long i0 = (long)row * P + c;
        C4 pv = ldc(p, i0);
        const C4 rv = ldc(r, i0);
#pragma unroll
        for (int i = 0; i < 4; ++i) pv.v[i] = init ? rv.v[i] : beta * pv.v[i] + rv.v[i];
        if (mask == 15u) {
            stc(p, i0, pv);
        } else {
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (mask >> i & 1) p[i0 + i] = pv.v[i];
        }
    }
};

// sum of squares, FP64
struct OpNorm32 {
    static constexpr int NRED = 1;
    int P;
    const float* a;
    __device__ void apply(int row, int c, unsigned mask, double* acc) const
    {
        const C4 v = ldc(a, (long)row * P + c);
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (mask >> i & 1) acc[0] += (double)v.v[i] * (double)v.v[i];
    }
};

// p = r / theta; d += p   (cheby.c:27-48 and cheby_calc_u, cheby.c:8-24, on
// the correction equation, whose residual r the caller has set)
struct OpChebyInit32 {
    static constexpr int NRED = 0;
    int P;
    float theta;
    float* p;
    const float* r;
    float* d;
    __device__ void apply(int row, int c, unsigned mask, double*) const
    {
        const long i0 = (long)row * P + c;
        const C4 rv = ldc(r, i0);
        C4 pv, dv = ldc(d, i0);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            pv.v[i] = rv.v[i] / theta;
            dv.v[i] = dv.v[i] + pv.v[i];
        }
        if (mask == 15u) {
            stc(p, i0, pv);
            stc(d, i0, dv);
        } else {
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (mask >> i & 1) {
                    p[i0 + i] = pv.v[i];
                    d[i0 + i] = dv.v[i];
                }
        }
    }
};

// u += d (FP64 update of the solution); d = 0
struct OpPromote {
    static constexpr int NRED = 0;
    int P, x;
    double* u;
    float* d;
    __device__ void apply(int row, int c, unsigned mask, double*) const
    {
        const long i0 = (long)row * P + c, j0 = (long)row * x + c;
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (mask >> i & 1) u[j0 + i] = u[j0 + i] + (double)d[i0 + i];
        if (mask == 15u) {
            stc(d, i0, zeros<float>());
        } else {
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (mask >> i & 1) d[i0 + i] = 0.0f;
        }
    }
};

// r = u0 - A u in FP64 (solver_methods.c:28-48, with u's x/y neighbours
// outside the interior mirrored), stored as FP32; rr += r*r in FP64.
struct OpResidual64 {
    static constexpr int NRED = 1;
    int P, x, y, hd;
    const double* u;
    const double* u0;
    const double* kx;
    const double* ky;
    float* r32;
    __device__ void apply(int row, int c, unsigned mask, double* acc) const
    {
        const int xin0 = hd, xin1 = x - hd, yin0 = hd, yin1 = y - hd;
        const long j0 = (long)row * x + c;
        if (mask == 15u && (x & 1) == 0) {
            // the group's four cells from 16-byte loads (j0 is even)
            using D2 = Cells<double>;
            auto ld4 = [&](const double* f, long j, double* v) {
                const D2 a = ldc(f, j), b = ldc(f, j + 2);
                v[0] = a.v[0]; v[1] = a.v[1]; v[2] = b.v[0]; v[3] = b.v[1];
            };
            double cu[6], dn[4], up[4], kxv[5], kyc[4], kyu[4], b0[4];
            ld4(u, j0, cu + 1);
            cu[0] = u[j0 - 1];
            cu[5] = u[j0 + 4];
            if (row != yin0) ld4(u, j0 - x, dn);
            if (row != yin1 - 1) ld4(u, j0 + x, up);
            ld4(kx, j0, kxv);
            kxv[4] = kx[j0 + 4];
            ld4(ky, j0, kyc);
            ld4(ky, j0 + x, kyu);
            ld4(u0, j0, b0);
            float out[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int k = c + i;
                const double cen = cu[i + 1];
                const double lft = k == xin0 ? cen : cu[i];
                const double rgt = k == xin1 - 1 ? cen : cu[i + 2];
                const double dv = row == yin0 ? cen : dn[i];
                const double uv = row == yin1 - 1 ? cen : up[i];
                const double d = (1.0 + (kxv[i + 1] + kxv[i])) + (kyu[i] + kyc[i]);
                const double a = d * cen - (kxv[i + 1] * rgt + kxv[i] * lft) -
                                 (kyu[i] * uv + kyc[i] * dv);
                const double rv = b0[i] - a;
                out[i] = (float)rv;
                acc[0] += rv * rv;
            }
            C4 o;
#pragma unroll
            for (int i = 0; i < 4; ++i) o.v[i] = out[i];
            stc(r32, (long)row * P + c, o);
            return;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (!(mask >> i & 1)) continue;
            const int k = c + i;
            const long j = j0 + i;
            const double cen = u[j];
            const double lft = k == xin0 ? cen : u[j - 1];
            const double rgt = k == xin1 - 1 ? cen : u[j + 1];
            const double dn = row == yin0 ? cen : u[j - x];
            const double up = row == yin1 - 1 ? cen : u[j + x];
            const double kxc = kx[j], kxr = kx[j + 1], kyc = ky[j], kyu = ky[j + x];
            const double d = (1.0 + (kxr + kxc)) + (kyu + kyc);
            const double a = d * cen - (kxr * rgt + kxc * lft) - (kyu * up + kyc * dn);
            const double rv = u0[j] - a;
            r32[(long)row * P + k] = (float)rv;
            acc[0] += rv * rv;
        }
    }
};

template <class Op>
int launch_groups(const Groups& G, const Op& op, const char* what, double* out)
{
    if (G.nq <= 0 || G.ny <= 0 || G.nz <= 0) {
        if (out) *out = 0.0;
        return 0;
    }
    const unsigned nb = flat_blocks(G.total());
    const int slot = tl_timer_begin(what);
    hipLaunchKernelGGL((k_groups<Op>), dim3(nb), dim3(kBlock), 0, tl_stream(), G, op,
                       tl_reduce_slots());
    tl_timer_end(slot);
    TL_TRY(tl_check_launch(what));
    if constexpr (Op::NRED > 0) {
        TL_TRY(tl_finish_reduce(Op::NRED, nb, what));
        return tl_fetch(1, out, what);
    }
    return 0;
}

// ---------------------------------------------------------------- demote
__global__ void __launch_bounds__(kBlock)
k_demote(const double* __restrict__ src, int x, int y, float* __restrict__ dst, int P)
{
    const long n = (long)P * y;
    for (long t = blockIdx.x * (long)kBlock + threadIdx.x; t < n; t += (long)gridDim.x * kBlock) {
        const int row = (int)(t / P), col = (int)(t % P);
        dst[t] = col < x ? (float)src[(long)row * x + col] : 0.0f;
    }
}

__global__ void __launch_bounds__(kBlock) k_zero32(float* __restrict__ a, long n)
{
    for (long t = blockIdx.x * (long)kBlock + threadIdx.x; t < n; t += (long)gridDim.x * kBlock)
        a[t] = 0.0f;
}

// ---------------------------------------------------------------- 3D
// A a (FP32, 7-point, 3d/shared.h:52-58 association) at a group's four
// cells of row `row` (plane row / Y, row row % Y); a neighbour outside the
// interior takes the cell's own value (the depth-1 reflective halo).
// Cells outside the interior (mask) get don't-care values; only in-bounds
// elements are read.
__device__ __forceinline__ void apply3_32(const Groups& G, const float* __restrict__ a,
                                          const float* __restrict__ kx,
                                          const float* __restrict__ ky,
                                          const float* __restrict__ kz, int row, int c, C4& cen,
                                          C4& av)
{
    const int j = row % G.Y, k = row / G.Y;
    const long P = G.P, PY = (long)G.P * G.Y;
    const long i0 = (long)row * P + c;
    cen = ldc(a, i0);
    const C4 so = j == G.hd ? cen : ldc(a, i0 - P);
    const C4 no = j == G.Y - G.hd - 1 ? cen : ldc(a, i0 + P);
    const C4 dn = k == G.hd ? cen : ldc(a, i0 - PY);
    const C4 up = k == G.z - G.hd - 1 ? cen : ldc(a, i0 + PY);
    const float el = c >= 1 ? a[i0 - 1] : 0.0f;
    const float er = c + 4 < G.x ? a[i0 + 4] : 0.0f;
    const C4 kxc = ldc(kx, i0);
    const float kxe = c + 4 < G.x ? kx[i0 + 4] : 0.0f;
    const C4 kyc = ldc(ky, i0), kyn = ldc(ky, i0 + P);
    const C4 kzc = ldc(kz, i0), kzu = ldc(kz, i0 + PY);
    const int xin0 = G.hd, xin1 = G.x - G.hd;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float cv = cen.v[i];
        const float lft = c + i == xin0 ? cv : (i == 0 ? el : cen.v[i - 1]);
        const float rgt = c + i == xin1 - 1 ? cv : (i == 3 ? er : cen.v[i + 1]);
        const float kxr = i == 3 ? kxe : kxc.v[i + 1];
        const float diag = ((1.0f + (kxr + kxc.v[i])) + (kyn.v[i] + kyc.v[i])) +
                           (kzu.v[i] + kzc.v[i]);
        av.v[i] = diag * cv - (kxr * rgt + kxc.v[i] * lft) - (kyn.v[i] * no.v[i] + kyc.v[i] * so.v[i]) -
                  (kzu.v[i] * up.v[i] + kzc.v[i] * dn.v[i]);
    }
}

__device__ __forceinline__ void store_masked(float* __restrict__ f, long i0, unsigned mask,
                                             const C4& v)
{
    if (mask == 15u) {
        stc(f, i0, v);
    } else {
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (mask >> i & 1) f[i0 + i] = v.v[i];
    }
}

// w = A p; pw += w*p   (3d/c_kernels/omp3/cg_solver.c calc_w, FP32, FP64 sum)
struct OpCalcW3_32 {
    static constexpr int NRED = 1;
    Groups G;
    const float* p;
    float* w;
    const float *kx, *ky, *kz;
    __device__ void apply(int row, int c, unsigned mask, double* acc) const
    {
        C4 cen, av;
        apply3_32(G, p, kx, ky, kz, row, c, cen, av);
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (mask >> i & 1) acc[0] += (double)av.v[i] * (double)cen.v[i];
        store_masked(w, (long)row * G.P + c, mask, av);
    }
};

// w = A sd; r -= w; d += sd; sd_out = alpha sd + beta r
// (3d/c_kernels/omp3/ppcg_solver.c inner iteration, FP32; d is the
// correction to u)
struct OpPpcgInner3_32 {
    static constexpr int NRED = 0;
    Groups G;
    float alpha, beta;
    const float* sd;
    float* sd_out;
    float* r;
    float* d;
    const float *kx, *ky, *kz;
    __device__ void apply(int row, int c, unsigned mask, double*) const
    {
        C4 cen, av;
        apply3_32(G, sd, kx, ky, kz, row, c, cen, av);
        const long i0 = (long)row * G.P + c;
        C4 rv = ldc(r, i0), dv = ldc(d, i0), sv;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            rv.v[i] = rv.v[i] - av.v[i];
            dv.v[i] = dv.v[i] + cen.v[i];
            sv.v[i] = alpha * cen.v[i] + beta * rv.v[i];
        }
        store_masked(r, i0, mask, rv);
        store_masked(d, i0, mask, dv);
        store_masked(sd_out, i0, mask, sv);
    }
};

// r = b - A d; p = alpha p + beta r; d_out = d + p; rr += r*r
// (3d/c_kernels/omp3/cheby_solver.c iterate and calc_u, FP32, on the
// correction equation A d = b)
template <bool NORM>
struct OpChebyIter3_32 {
    static constexpr int NRED = NORM ? 1 : 0;
    Groups G;
    float alpha, beta;
    const float* d;
    float* d_out;
    const float* b;
    float* p;
    float* r;
    const float *kx, *ky, *kz;
    __device__ void apply(int row, int c, unsigned mask, double* acc) const
    {
        C4 cen, av;
        apply3_32(G, d, kx, ky, kz, row, c, cen, av);
        const long i0 = (long)row * G.P + c;
        const C4 bv = ldc(b, i0);
        C4 pv = ldc(p, i0), rv, dv;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            rv.v[i] = bv.v[i] - av.v[i];
            pv.v[i] = alpha * pv.v[i] + beta * rv.v[i];
            dv.v[i] = cen.v[i] + pv.v[i];
            if constexpr (NORM) {
                if (mask >> i & 1) acc[0] += (double)rv.v[i] * (double)rv.v[i];
            }
        }
        store_masked(r, i0, mask, rv);
        store_masked(p, i0, mask, pv);
        store_masked(d_out, i0, mask, dv);
    }
};

// sd = r / theta   (3d ppcg_init)
struct OpScale32 {
    static constexpr int NRED = 0;
    int P;
    float theta;
    const float* r;
    float* sd;
    __device__ void apply(int row, int c, unsigned mask, double*) const
    {
        const long i0 = (long)row * P + c;
        C4 v = ldc(r, i0);
#pragma unroll
        for (int i = 0; i < 4; ++i) v.v[i] = v.v[i] / theta;
        store_masked(sd, i0, mask, v);
    }
};

// r = u0 - A u in FP64 (3d/c_kernels/omp3/solver_methods.c, neighbours
// outside the interior mirrored), stored as FP32; rr += r*r in FP64.
struct OpResidual3_64 {
    static constexpr int NRED = 1;
    Groups G;
    const double* u;
    const double* u0;
    const double *kx, *ky, *kz;
    float* r32;
    __device__ void apply(int row, int c, unsigned mask, double* acc) const
    {
        const int j = row % G.Y, k = row / G.Y;
        const long X = G.x, XY = (long)G.x * G.Y;
        const int xin0 = G.hd, xin1 = G.x - G.hd;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (!(mask >> i & 1)) continue;
            const int q = c + i;
            const long n = (long)row * X + q;
            const double cen = u[n];
            const double lft = q == xin0 ? cen : u[n - 1];
            const double rgt = q == xin1 - 1 ? cen : u[n + 1];
            const double so = j == G.hd ? cen : u[n - X];
            const double no = j == G.Y - G.hd - 1 ? cen : u[n + X];
            const double dn = k == G.hd ? cen : u[n - XY];
            const double up = k == G.z - G.hd - 1 ? cen : u[n + XY];
            const double kxc = kx[n], kxr = kx[n + 1], kyc = ky[n], kyn = ky[n + X];
            const double kzc = kz[n], kzu = kz[n + XY];
            const double diag = ((1.0 + (kxr + kxc)) + (kyn + kyc)) + (kzu + kzc);
            const double a = diag * cen - (kxr * rgt + kxc * lft) - (kyn * no + kyc * so) -
                             (kzu * up + kzc * dn);
            const double rv = u0[n] - a;
            r32[(long)row * G.P + q] = (float)rv;
            acc[0] += rv * rv;
        }
    }
};


}  // namespace

namespace tl {
namespace mix {

int zero32(float* a, long n)
{
    TL_TRY(tl_ensure_init());
    hipLaunchKernelGGL(k_zero32, dim3(flat_blocks(n)), dim3(kBlock), 0, tl_stream(), a, n);
    return tl_check_launch("tlhip_mix_zero32");
}

int demote(const double* src, int x, int y, float* dst, int P)
{
    const char* what = "tlhip_mix_demote";
    TL_TRY(tl_ensure_init());
    hipLaunchKernelGGL(k_demote, dim3(flat_blocks((long)P * y)), dim3(kBlock), 0, tl_stream(), src,
                       x, y, dst, P);
    return tl_check_launch(what);
}

int calc_ur32(int x, int y, int z, int hd, int P, double alpha, double* rrn, float* d,
              const float* p, float* r, const float* w)
{
    const OpUr32 op{P, (float)alpha, d, p, r, w};
    return launch_groups(groups_of(x, y, z, hd, P), op, "tlhip2d_mix_calc_ur32", rrn);
}

int calc_p32(int x, int y, int z, int hd, int P, double beta, bool init, float* p,
             const float* r)
{
    const OpP32 op{P, (float)beta, init ? 1 : 0, p, r};
    return launch_groups(groups_of(x, y, z, hd, P), op, "tlhip2d_mix_calc_p32", nullptr);
}

int norm32(int x, int y, int z, int hd, int P, const float* a, double* out)
{
    const OpNorm32 op{P, a};
    return launch_groups(groups_of(x, y, z, hd, P), op, "tlhip2d_mix_norm32", out);
}

int cheby_init32(int x, int y, int z, int hd, int P, double theta, float* p, const float* r,
                 float* d)
{
    const OpChebyInit32 op{P, (float)theta, p, r, d};
    return launch_groups(groups_of(x, y, z, hd, P), op, "tlhip2d_mix_cheby_init32", nullptr);
}

int promote(int x, int y, int z, int hd, int P, double* u, float* d)
{
    const OpPromote op{P, x, u, d};
    return launch_groups(groups_of(x, y, z, hd, P), op, "tlhip2d_mix_promote", nullptr);
}

int residual64(int x, int y, int hd, int P, const double* u, const double* u0, const double* kx,
               const double* ky, float* r32, double* rr)
{
    const OpResidual64 op{P, x, y, hd, u, u0, kx, ky, r32};
    return launch_groups(groups_of(x, y, 1, hd, P), op, "tlhip2d_mix_residual64", rr);
}

int calc_w3_32(int x, int y, int z, int hd, int P, double* pw, const float* p, float* w,
               const float* kx, const float* ky, const float* kz)
{
    const Groups G = groups_of(x, y, z, hd, P);
    const OpCalcW3_32 op{G, p, w, kx, ky, kz};
    *pw = 0.0;
    return launch_groups(G, op, "tlhip3d_mix_calc_w32", pw);
}

int ppcg_init3_32(int x, int y, int z, int hd, int P, double theta, const float* r, float* sd)
{
    const OpScale32 op{P, (float)theta, r, sd};
    return launch_groups(groups_of(x, y, z, hd, P), op, "tlhip3d_mix_ppcg_init32", nullptr);
}

int ppcg_inner3_32(int x, int y, int z, int hd, int P, double alpha, double beta,
                   const float* sd_in, float* sd_out, float* r, float* d, const float* kx,
                   const float* ky, const float* kz)
{
    const Groups G = groups_of(x, y, z, hd, P);
    const OpPpcgInner3_32 op{G, (float)alpha, (float)beta, sd_in, sd_out, r, d, kx, ky, kz};
    return launch_groups(G, op, "tlhip3d_mix_ppcg_inner32", nullptr);
}

int cheby_iter3_32(int x, int y, int z, int hd, int P, double alpha, double beta,
                   const float* d_in, float* d_out, const float* b, float* p, float* r,
                   const float* kx, const float* ky, const float* kz, double* rr)
{
    const Groups G = groups_of(x, y, z, hd, P);
    if (rr) {
        const OpChebyIter3_32<true> op{G, (float)alpha, (float)beta, d_in, d_out, b, p, r, kx,
                                       ky, kz};
        return launch_groups(G, op, "tlhip3d_mix_cheby_iter32", rr);
    }
    const OpChebyIter3_32<false> op{G, (float)alpha, (float)beta, d_in, d_out, b, p, r, kx, ky,
                                    kz};
    return launch_groups(G, op, "tlhip3d_mix_cheby_iter32", nullptr);
}

int residual3_64(int x, int y, int z, int hd, int P, const double* u, const double* u0,
                 const double* kx, const double* ky, const double* kz, float* r32, double* rr)
{
    const Groups G = groups_of(x, y, z, hd, P);
    const OpResidual3_64 op{G, u, u0, kx, ky, kz, r32};
    return launch_groups(G, op, "tlhip3d_mix_residual64", rr);
}

}  // namespace mix
}  // namespace tl
