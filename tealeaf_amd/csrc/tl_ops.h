// Per-cell operations shared by the 2D and 3D kernels.
//
// Every op is split in two phases so the kernels can software-pipeline:
//   load<V>(i)            issues the op's own reads for the V cells at i;
//   apply<V>(i, nval, ...) computes and stores (only the first nval cells).
// A kernel issues the loads of its next row (or segment) before it applies
// the current one, which keeps two rows of reads in flight per wavefront.
// Pointwise ops act on interior rows through k_interior; stencil ops also
// receive a neighbourhood `N` (Nb2 or Nb3) and call apply_a(n, v), which each
// dimension defines with the reference's association.
#pragma once

#include "tl_common.h"

namespace tl {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / TL_WAVE;

template <int V>
struct DV {
    double v[V];
};

template <int V>
__device__ __forceinline__ DV<V> ldv(const double* __restrict__ p, long i)
{
    DV<V> r;
    if constexpr (V == 2) {
        const double2 t = *reinterpret_cast<const double2*>(p + i);
        r.v[0] = t.x;
        r.v[1] = t.y;
    } else {
        r.v[0] = p[i];
    }
    return r;
}

template <int V>
__device__ __forceinline__ void stv(double* __restrict__ p, long i, const DV<V>& x, int nval)
{
    if constexpr (V == 2) {
        if (nval == 2)
            *reinterpret_cast<double2*>(p + i) = make_double2(x.v[0], x.v[1]);
        else if (nval == 1)
            p[i] = x.v[0];
    } else {
        if (nval >= 1) p[i] = x.v[0];
    }
}

template <int V>
__device__ __forceinline__ DV<V> zdv()
{
    DV<V> r;
#pragma unroll
    for (int v = 0; v < V; ++v) r.v[v] = 0.0;
    return r;
}

// Up to four op-private input vectors.
template <int V>
struct In4 {
    DV<V> a, b, c, d;
};

// Interior rows of a 2D or 3D chunk.  Row r (0-based over the interior) is
// at offset (hd + r % ny)*x + (zoff + r / ny)*xy; 2D uses zoff = xy = 0.
struct TlRows {
    int x, hd, ny;
    long xy;
    int zoff;
    long nrows;
};

struct Seg {
    long i;     // index of the lane's first cell
    int nval;   // interior cells it owns (0..V)
};

template <int V>
__device__ __forceinline__ Seg seg_of(const TlRows& g, int segs_per_row, long s)
{
    const long r = s / segs_per_row;
    const long base = (long)(g.hd + r % g.ny) * g.x + (long)(g.zoff + r / g.ny) * g.xy;
    const int c0 = g.hd + (int)(s % segs_per_row) * (kBlock * V) + threadIdx.x * V;
    int nval = g.x - g.hd - c0;
    nval = nval < 0 ? 0 : (nval > V ? V : nval);
    return Seg{base + c0, nval};
}

// Pointwise work over interior rows: block-strided row segments of kBlock*V
// columns, two segments in flight per thread.
template <int V, class Op>
__global__ void __launch_bounds__(kBlock)
k_interior(TlRows g, int segs_per_row, long nseg, Op op, TlReduce rd, TlGate gate)
{
    if (gate.closed()) return;
    constexpr int NR = Op::NRED > 0 ? Op::NRED : 1;
    double acc[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) acc[q] = 0.0;
    const long stride = gridDim.x;
    for (long s = blockIdx.x; s < nseg; s += 2 * stride) {
        const Seg a = seg_of<V>(g, segs_per_row, s);
        const bool has_b = s + stride < nseg;
        const Seg b = has_b ? seg_of<V>(g, segs_per_row, s + stride) : Seg{0, 0};
        In4<V> ia{}, ib{};
        if (a.nval > 0) ia = op.template load<V>(a.i);
        if (b.nval > 0) ib = op.template load<V>(b.i);
        if (a.nval > 0) op.template apply<V>(a.i, a.nval, ia, acc);
        if (b.nval > 0) op.template apply<V>(b.i, b.nval, ib, acc);
    }
    if constexpr (Op::NRED > 0) tl_reduce_partial<Op::NRED, kBlock>(acc, rd);
}

// ------------------------------------------------------------ pointwise ops

struct OpCalcUr {   // u += alpha p ; r -= alpha w ; rrn += r.r  (2d/c_kernels/omp3/cg.c:116-143)
    static constexpr int NRED = 1;
    double* __restrict__ u;
    const double* __restrict__ p;
    double* __restrict__ r;
    const double* __restrict__ w;
    double alpha;
    const double* __restrict__ alpha_dev = nullptr;   // when set, alpha is read from device memory
    template <int V>
    __device__ In4<V> load(long i) const
    {
        return In4<V>{ldv<V>(u, i), ldv<V>(p, i), ldv<V>(r, i), ldv<V>(w, i)};
    }
    template <int V>
    __device__ void apply(long i, int nval, const In4<V>& in, double* acc) const
    {
        const double a = alpha_dev ? *alpha_dev : alpha;
        DV<V> uu = in.a, rr = in.c;
#pragma unroll
        for (int v = 0; v < V; ++v) {
            uu.v[v] += a * in.b.v[v];
            rr.v[v] -= a * in.d.v[v];
            if (v < nval) acc[0] += rr.v[v] * rr.v[v];
        }
        stv<V>(u, i, uu, nval);
        stv<V>(r, i, rr, nval);
    }
};

struct OpCalcP {    // p = beta p + r  (2d/c_kernels/omp3/cg.c:146-164)
    static constexpr int NRED = 0;
    double* __restrict__ p;
    const double* __restrict__ r;
    double beta;
    const double* __restrict__ beta_dev = nullptr;    // when set, beta is read from device memory
    template <int V>
    __device__ In4<V> load(long i) const
    {
        In4<V> in{};
        in.a = ldv<V>(p, i);
        in.b = ldv<V>(r, i);
        return in;
    }
    template <int V>
    __device__ void apply(long i, int nval, const In4<V>& in, double*) const
    {
        const double b = beta_dev ? *beta_dev : beta;
        DV<V> pp;
#pragma unroll
        for (int v = 0; v < V; ++v) pp.v[v] = b * in.a.v[v] + in.b.v[v];
        stv<V>(p, i, pp, nval);
    }
};

// Small pointwise forms:
//   0: dst = src                    (copy_u, solver_methods.c:9-25)
//   1: dst = dst + src              (cheby_calc_u, cheby.c:8-24)
//   2: dst = src / s                (ppcg_init, ppcg.c:8-25)
//   3: dst = src / other            (finalise, solver_methods.c:74-91)
struct OpSimple {
    static constexpr int NRED = 0;
    double* __restrict__ dst;
    const double* __restrict__ src;
    int mode;
    const double* __restrict__ other;
    double s;
    template <int V>
    __device__ In4<V> load(long i) const
    {
        In4<V> in{};
        in.a = ldv<V>(src, i);
        if (mode == 1) in.b = ldv<V>(dst, i);
        if (mode == 3) in.c = ldv<V>(other, i);
        return in;
    }
    template <int V>
    __device__ void apply(long i, int nval, const In4<V>& in, double*) const
    {
        DV<V> o;
#pragma unroll
        for (int v = 0; v < V; ++v) {
            if (mode == 0) o.v[v] = in.a.v[v];
            else if (mode == 1) o.v[v] = in.b.v[v] + in.a.v[v];
            else if (mode == 2) o.v[v] = in.a.v[v] / s;
            else o.v[v] = in.a.v[v] / in.c.v[v];
        }
        stv<V>(dst, i, o, nval);
    }
};

struct OpNorm {     // norm += b.b  (solver_methods.c:51-71)
    static constexpr int NRED = 1;
    const double* __restrict__ b;
    template <int V>
    __device__ In4<V> load(long i) const
    {
        In4<V> in{};
        in.a = ldv<V>(b, i);
        return in;
    }
    template <int V>
    __device__ void apply(long, int nval, const In4<V>& in, double* acc) const
    {
#pragma unroll
        for (int v = 0; v < V; ++v)
            if (v < nval) acc[0] += in.a.v[v] * in.a.v[v];
    }
};

struct OpFieldSummary {   // field_summary.c:9-45
    static constexpr int NRED = 4;
    const double* __restrict__ volume;
    const double* __restrict__ density;
    const double* __restrict__ energy0;
    const double* __restrict__ u;
    template <int V>
    __device__ In4<V> load(long i) const
    {
        return In4<V>{ldv<V>(volume, i), ldv<V>(density, i), ldv<V>(energy0, i), ldv<V>(u, i)};
    }
    template <int V>
    __device__ void apply(long, int nval, const In4<V>& in, double* acc) const
    {
#pragma unroll
        for (int v = 0; v < V; ++v) {
            if (v < nval) {
                const double cm = in.a.v[v] * in.b.v[v];
                acc[0] += in.a.v[v];
                acc[1] += cm;
                acc[2] += cm * in.c.v[v];
                acc[3] += cm * in.d.v[v];
            }
        }
    }
};

// ------------------------------------------------------------ stencil ops
// N provides: c (field at the cell, DV<V>) and apply_a(n, v).

struct OpCalcW {   // w = A p ; pw += w.p  (cg.c:88-113)
    static constexpr int NRED = 1;
    double* __restrict__ w;
    template <int V>
    __device__ In4<V> load(long) const
    {
        return In4<V>{};
    }
    template <int V, class N>
    __device__ void apply(long i, int nval, const N& n, const In4<V>&, double* acc) const
    {
        DV<V> o;
#pragma unroll
        for (int v = 0; v < V; ++v) {
            o.v[v] = apply_a(n, v);
            if (v < nval) acc[0] += o.v[v] * n.c.v[v];
        }
        stv<V>(w, i, o, nval);
    }
};

struct OpCgInit {  // w = A u ; r = u - w ; p = r ; rro += r.p  (cg.c:71-84)
    static constexpr int NRED = 1;
    double* __restrict__ w;
    double* __restrict__ r;
    double* __restrict__ p;
    template <int V>
    __device__ In4<V> load(long) const
    {
        return In4<V>{};
    }
    template <int V, class N>
    __device__ void apply(long i, int nval, const N& n, const In4<V>&, double* acc) const
    {
        DV<V> wo, ro;
#pragma unroll
        for (int v = 0; v < V; ++v) {
            wo.v[v] = apply_a(n, v);
            ro.v[v] = n.c.v[v] - wo.v[v];
            if (v < nval) acc[0] += ro.v[v] * ro.v[v];
        }
        stv<V>(w, i, wo, nval);
        stv<V>(r, i, ro, nval);
        stv<V>(p, i, ro, nval);
    }
};

struct OpChebyInit {   // w = A u ; r = u0 - w ; p = r/theta  (cheby.c:27-54)
    static constexpr int NRED = 0;
    const double* __restrict__ u0;
    double* __restrict__ w;
    double* __restrict__ r;
    double* __restrict__ p;
    double theta;
    template <int V>
    __device__ In4<V> load(long i) const
    {
        In4<V> in{};
        in.a = ldv<V>(u0, i);
        return in;
    }
    template <int V, class N>
    __device__ void apply(long i, int nval, const N& n, const In4<V>& in, double*) const
    {
        DV<V> wo, ro, po;
#pragma unroll
        for (int v = 0; v < V; ++v) {
            wo.v[v] = apply_a(n, v);
            ro.v[v] = in.a.v[v] - wo.v[v];
            po.v[v] = ro.v[v] / theta;
        }
        stv<V>(w, i, wo, nval);
        stv<V>(r, i, ro, nval);
        stv<V>(p, i, po, nval);
    }
};

// Fused Chebyshev step (cheby.c:57-85): w = A u_in, r = u0 - w,
// p = alpha p + beta r, u_out = u_in + p; w written only if given.
struct OpChebyIter {
    static constexpr int NRED = 0;
    const double* __restrict__ u0;
    double* __restrict__ p;
    double* __restrict__ r;
    double* __restrict__ w;
    double* __restrict__ u_out;
    double alpha, beta;
    template <int V>
    __device__ In4<V> load(long i) const
    {
        In4<V> in{};
        in.a = ldv<V>(u0, i);
        in.b = ldv<V>(p, i);
        return in;
    }
    template <int V, class N>
    __device__ void apply(long i, int nval, const N& n, const In4<V>& in, double*) const
    {
        DV<V> wo, ro, po, uo;
#pragma unroll
        for (int v = 0; v < V; ++v) {
            wo.v[v] = apply_a(n, v);
            ro.v[v] = in.a.v[v] - wo.v[v];
            po.v[v] = alpha * in.b.v[v] + beta * ro.v[v];
            uo.v[v] = n.c.v[v] + po.v[v];
        }
        if (w) stv<V>(w, i, wo, nval);
        stv<V>(r, i, ro, nval);
        stv<V>(p, i, po, nval);
        stv<V>(u_out, i, uo, nval);
    }
};

// Fused PPCG inner step (ppcg.c:28-61): r -= A sd_in, u += sd_in,
// sd_out = alpha sd_in + beta r.
struct OpPpcgInner {
    static constexpr int NRED = 0;
    double* __restrict__ r;
    double* __restrict__ u;
    double* __restrict__ sd_out;
    double alpha, beta;
    template <int V>
    __device__ In4<V> load(long i) const
    {
        In4<V> in{};
        in.a = ldv<V>(r, i);
        in.b = ldv<V>(u, i);
        return in;
    }
    template <int V, class N>
    __device__ void apply(long i, int nval, const N& n, const In4<V>& in, double*) const
    {
        DV<V> ro, uo, so;
#pragma unroll
        for (int v = 0; v < V; ++v) {
            ro.v[v] = in.a.v[v] - apply_a(n, v);
            uo.v[v] = in.b.v[v] + n.c.v[v];
            so.v[v] = alpha * n.c.v[v] + beta * ro.v[v];
        }
        stv<V>(r, i, ro, nval);
        stv<V>(u, i, uo, nval);
        stv<V>(sd_out, i, so, nval);
    }
};

struct OpResidual {   // r = u0 - A u  (solver_methods.c:28-48)
    static constexpr int NRED = 0;
    const double* __restrict__ u0;
    double* __restrict__ r;
    template <int V>
    __device__ In4<V> load(long i) const
    {
        In4<V> in{};
        in.a = ldv<V>(u0, i);
        return in;
    }
    template <int V, class N>
    __device__ void apply(long i, int nval, const N& n, const In4<V>& in, double*) const
    {
        DV<V> ro;
#pragma unroll
        for (int v = 0; v < V; ++v) ro.v[v] = in.a.v[v] - apply_a(n, v);
        stv<V>(r, i, ro, nval);
    }
};

// ------------------------------------------------------------ whole-array

__global__ void __launch_bounds__(kBlock)
k_copy_all(long n, const double* __restrict__ src, double* __restrict__ dst);

static inline unsigned flat_blocks(long n)
{
    long b = (n + kBlock - 1) / kBlock;
    if (b < 1) b = 1;
    if (b > 8192) b = 8192;
    return (unsigned)b;
}

static inline bool aligned16(const void* p) { return p == nullptr || ((uintptr_t)p & 15) == 0; }

template <typename... P>
static inline bool all_aligned(P... ps)
{
    return (aligned16(ps) && ...);
}

// Launch a pointwise op over the interior rows, under the current launch
// gate.  With nb_out, a reducing op's partials are left for the caller's
// finisher and the block count is returned there.
template <class Op>
static int launch_interior(const TlRows& g, int nx, int V, const Op& op, const char* what,
                           unsigned* nb_out = nullptr)
{
    if (nx <= 0 || g.nrows <= 0) return 0;
    const int spr = (int)tl_cdiv(nx, kBlock * V);
    const long nseg = (long)spr * g.nrows;
    // ~16 waves per CU, each thread owning two segments at a time.
    const long want = (nseg + 1) / 2;
    const long cap = tl_interior_blocks();
    const unsigned nb = (unsigned)(want < cap ? want : cap);
    const int slot = tl_timer_begin(what);
    if (V == 2)
        hipLaunchKernelGGL((k_interior<2, Op>), dim3(nb), dim3(kBlock), 0, tl_stream(), g, spr,
                           nseg, op, tl_reduce_slots(), tl_gate());
    else
        hipLaunchKernelGGL((k_interior<1, Op>), dim3(nb), dim3(kBlock), 0, tl_stream(), g, spr,
                           nseg, op, tl_reduce_slots(), tl_gate());
    tl_timer_end(slot);
    TL_TRY(tl_check_launch(what));
    if (nb_out) {
        *nb_out = nb;
        return 0;
    }
    if constexpr (Op::NRED > 0) return tl_finish_reduce(Op::NRED, nb, what);
    return 0;
}

// Add the published reduction value to *dst (the reference's += contract).
static inline int fetch_add(double* dst, const char* what)
{
    if (!dst) return tl_fail(1, "%s: null scalar pointer", what);
    double v = 0.0;
    TL_TRY(tl_fetch(1, &v, what));
    *dst += v;
    return 0;
}

}  // namespace tl

#define TL_REQUIRE(cond, what, msg)                             \
    do {                                                        \
        if (!(cond)) return tl_fail(1, "%s: %s", what, msg);    \
    } while (0)
