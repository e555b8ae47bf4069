// Shared device/host helpers for the hip_cdna4 TeaLeaf kernels (gfx950).
//
// Conventions used by every kernel in this library:
//  * fields are row-major with the halo included: index = kk + jj*x
//    (+ ii*x*y in 3D), interior [hd, n-hd) on each axis;
//  * one in-order HIP stream per process (tl_stream()); no kernel ever waits
//    on another;
//  * scalar results use a deterministic two-kernel grid reduction: every
//    block writes its partial to a fixed slot, then a one-block finisher
//    sums the slots in index order and publishes the result to device
//    memory and to a pinned host mailbox.  The summation order depends only
//    on the launch shape, so results are bitwise reproducible run to run;
//  * a driver that runs iterations ahead of the host (tl_loop.h) gates the
//    kernels it launches: each takes a TlGate and does nothing once the
//    device has decided that its iteration is past the end of the loop.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define TL_WAVE 64

// ---------------------------------------------------------------- runtime
struct TlReduce {
    double* partials;         // [max_blocks * TL_MAX_RED], one slot per block
    double* result;           // [TL_MAX_RED] device copy of the last result
    double* mailbox;          // pinned, host-mapped [TL_MAIL_SLOTS * TL_MAX_RED]
};

#define TL_MAX_RED 4
#define TL_MAX_RED_BLOCKS 65536
#define TL_MAIL_SLOTS 8

// Pinned result slots are armed by the host with this signalling-NaN bit
// pattern before the kernel that fills them is launched; a slot is ready
// once its bits differ.  No arithmetic result is a signalling NaN, so a
// finisher's value can never be mistaken for the mark.  Each value is
// thereby its own ready flag, and the device publishes with plain
// system-scope stores, without the release (a write-back of the XCD's L2)
// that a separate ready flag would need.
#define TL_MAIL_EMPTY 0x7FF4DEADBEEF0001ULL

// Launch gate of the device-resident solver loops: a kernel of iteration
// `iter` returns at once when *last < iter, where *last is the last
// iteration the device decided to run (INT_MAX while the loop goes on).
// The value is written only by one-block finisher kernels that run
// strictly before or after the gated kernel on the same stream, so a plain
// load sees a value that is constant for the kernel's lifetime.
struct TlGate {
    const int* last;   // null: always run
    int iter;
    const int* base = nullptr;   // graph replays: iter counts from *base
    __device__ __forceinline__ bool closed() const
    {
        return last && iter + (base ? *base : 0) > *last;
    }
};
// The gate that launches issued now should carry (null outside a loop).
TlGate tl_gate();

hipStream_t tl_stream();
TlReduce tl_reduce_slots();
// Record a failure message (thread-local); returns the error code given.
int tl_fail(int code, const char* fmt, ...);
// Check the last launch; returns 0 or records the error.
int tl_check_launch(const char* what);
// Output rows per tile set by tlhip_set_tile_rows (<= 0: the kernel's default).
int tl_tile_rows_override();
// Most blocks a pointwise sweep launches (tlhip_set_interior_blocks).
int tl_interior_blocks();
// Wait for the stream and copy `n` published reduction values out.
int tl_fetch(int n, double* out, const char* what);
// Longest a host wait for a device result may take before it fails
// (seconds; tlhip_set_wait_limit).
double tl_wait_limit_s();
int tl_ensure_init();
// Records, for the device current at initialisation, the address of the
// allocation-free entry point's scratch array (tl_kernels2d.hip).
int tl_noalloc_setup(int device);
// Kernel timer hooks: begin returns a slot (or -1 when `what` is not being
// timed); end records the closing event for that slot.
int tl_timer_begin(const char* what);
// Whether the kernel timer is recording (tlhip_kernel_timer_start).
bool tl_timer_active();
void tl_timer_end(int slot);

#define TL_TRY(expr)                      \
    do {                                  \
        int _rc = (expr);                 \
        if (_rc) return _rc;              \
    } while (0)

// ---------------------------------------------------------------- lanes
// Value held by lane-1 / lane+1 of the same wavefront (DPP wave_shr:1 /
// wave_shl:1; the end lanes receive `edge`).  Doubles move as two dwords.
__device__ __forceinline__ int tl_lane_prev_i(int v, int edge)
{
    return __builtin_amdgcn_update_dpp(edge, v, 0x138, 0xf, 0xf, false);
}
__device__ __forceinline__ int tl_lane_next_i(int v, int edge)
{
    return __builtin_amdgcn_update_dpp(edge, v, 0x130, 0xf, 0xf, false);
}
__device__ __forceinline__ double tl_lane_prev(double v, double edge)
{
    const int2 a = __builtin_bit_cast(int2, v), e = __builtin_bit_cast(int2, edge);
    int2 r;
    r.x = tl_lane_prev_i(a.x, e.x);
    r.y = tl_lane_prev_i(a.y, e.y);
    return __builtin_bit_cast(double, r);
}
__device__ __forceinline__ double tl_lane_next(double v, double edge)
{
    const int2 a = __builtin_bit_cast(int2, v), e = __builtin_bit_cast(int2, edge);
    int2 r;
    r.x = tl_lane_next_i(a.x, e.x);
    r.y = tl_lane_next_i(a.y, e.y);
    return __builtin_bit_cast(double, r);
}
__device__ __forceinline__ float tl_lane_prev(float v, float edge)
{
    return __builtin_bit_cast(float, tl_lane_prev_i(__builtin_bit_cast(int, v),
                                                    __builtin_bit_cast(int, edge)));
}
__device__ __forceinline__ float tl_lane_next(float v, float edge)
{
    return __builtin_bit_cast(float, tl_lane_next_i(__builtin_bit_cast(int, v),
                                                    __builtin_bit_cast(int, edge)));
}

// Butterfly sum over the 64 lanes; every lane gets the total.  The order of
// additions is fixed, so the result is reproducible.
// A face's coefficient as cg_init forms kx and ky (2d/c_kernels/omp3/cg.c:55-64):
// with w_a the conductivity of the cell below / left of the face and w_b
// that of the cell itself, r (w_a + w_b) / (2 w_a w_b), operands in that order.
__device__ __forceinline__ double tl_face_coef(double r, double wa, double wb)
{
    return r * (wa + wb) / (2.0 * wa * wb);
}

__device__ __forceinline__ double tl_wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, TL_WAVE);
    return v;
}

// ---------------------------------------------------------------- reduce
// Block-level reduction of NV values: every block writes its partial sums
// to slot blockIdx (linearised) of rd.partials with plain stores.  Must be
// called by every thread of the block.  The grid total is then formed by
// a one-block finisher kernel launched right after (tl_finish_reduce), in
// a fixed order, so results are reproducible.  (A last-block-done ticket
// inside the same kernel would need an agent-scope release per block, which
// on a multi-XCD part writes back the whole L2 each time.)
template <int NV, int BLOCK>
__device__ __forceinline__ void tl_reduce_partial(const double* v, TlReduce rd)
{
    static_assert(BLOCK % TL_WAVE == 0, "block must be whole waves");
    constexpr int NW = BLOCK / TL_WAVE;
    __shared__ double s_part[NW][NV];
    const int tid = threadIdx.x;
    const int lane = tid & (TL_WAVE - 1), wave = tid / TL_WAVE;
    const unsigned bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const double s = tl_wave_sum(v[i]);
        if (lane == 0) s_part[wave][i] = s;
    }
    __syncthreads();
    if (tid < NV) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < NW; ++w) s += s_part[w][tid];
        rd.partials[(size_t)bid * NV + tid] = s;
    }
}

// Grid total of `nblocks` per-block partials of NV values, in a fixed
// order: thread t of a 256-thread block takes slots t, t+256, ... in turn,
// then the waves' butterfly sums, then the four waves in order.  Every
// finisher uses this, so a value is the same bits whichever finisher formed
// it.  Must be called by all 256 threads; thread 0 receives the totals.
template <int NV>
__device__ __forceinline__ void tl_sum_partials(unsigned nblocks, const double* partials,
                                                double* out)
{
    constexpr int NW = 256 / TL_WAVE;
    __shared__ double s_tot[NW][NV];
    const int tid = threadIdx.x, lane = tid & (TL_WAVE - 1), wave = tid / TL_WAVE;
    double acc[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) acc[i] = 0.0;
    // Loads of kBatch slots are issued together and then added in slot
    // order: the same additions in the same order as one slot at a time,
    // without a memory round trip per slot (a 4096^2 stencil sweep leaves
    // 32768 partials, 128 per thread).
    constexpr int kBatch = 16;
    for (unsigned b0 = tid; b0 < nblocks; b0 += 256u * kBatch) {
        double v[kBatch][NV];
#pragma unroll
        for (int q = 0; q < kBatch; ++q) {
            const unsigned b = b0 + 256u * q;
#pragma unroll
            for (int i = 0; i < NV; ++i) v[q][i] = b < nblocks ? partials[(size_t)b * NV + i] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < kBatch; ++q)
            if (b0 + 256u * q < nblocks) {
#pragma unroll
                for (int i = 0; i < NV; ++i) acc[i] += v[q][i];
            }
    }
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const double s = tl_wave_sum(acc[i]);
        if (lane == 0) s_tot[wave][i] = s;
    }
    __syncthreads();
    if (tid == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            double s = 0.0;
#pragma unroll
            for (int w = 0; w < NW; ++w) s += s_tot[w][i];
            out[i] = s;
        }
    }
}

// Launch the one-block finisher that sums `nblocks` partials of `nv` values
// into rd.result and the host mailbox.
int tl_finish_reduce(int nv, unsigned nblocks, const char* what);

// ---------------------------------------------------------------- shapes
struct TlGrid2 {
    int x, y, hd;
};
struct TlGrid3 {
    int x, y, z, hd;
};

static inline unsigned tl_cdiv(long a, long b) { return (unsigned)((a + b - 1) / b); }
