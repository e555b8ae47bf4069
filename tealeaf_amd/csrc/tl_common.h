// Shared device/host helpers for the hip_cdna4 TeaLeaf kernels (gfx950).
//
// Conventions used by every kernel in this library:
//  * fields are row-major with the halo included: index = kk + jj*x
//    (+ ii*x*y in 3D), interior [hd, n-hd) on each axis;
//  * one in-order HIP stream per process (tl_stream()); no kernel ever waits
//    on another;
//  * scalar results use a single-pass, deterministic grid reduction: every
//    block writes its partial to a fixed slot, the last block to arrive (an
//    agent-scope ticket) sums the slots in index order and publishes the
//    result to device memory and to a pinned host mailbox.  The summation
//    order depends only on the launch shape, so results are bitwise
//    reproducible run to run.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define TL_WAVE 64

// ---------------------------------------------------------------- runtime
struct TlReduce {
    double* partials;         // [max_blocks * TL_MAX_RED]
    unsigned int* ticket;     // zero between launches (last block resets it)
    double* result;           // [TL_MAX_RED] device copy of the last result
    double* mailbox;          // pinned, host-mapped [TL_MAX_RED]
};

#define TL_MAX_RED 4
#define TL_MAX_RED_BLOCKS 65536

hipStream_t tl_stream();
TlReduce tl_reduce_slots();
// Record a failure message (thread-local); returns the error code given.
int tl_fail(int code, const char* fmt, ...);
// Check the last launch; returns 0 or records the error.
int tl_check_launch(const char* what);
// Wait for the stream and copy `n` published reduction values out.
int tl_fetch(int n, double* out, const char* what);
int tl_ensure_init();

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
__device__ __forceinline__ double tl_wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, TL_WAVE);
    return v;
}

// ---------------------------------------------------------------- reduce
// Block-level reduction of NV values followed by the grid finish described
// at the top of this file.  Must be called by every thread of every block
// (no early return before it).  `nblocks` = gridDim.x*gridDim.y*gridDim.z.
template <int NV, int BLOCK>
__device__ __forceinline__ void tl_reduce_finish(const double* v, TlReduce rd)
{
    static_assert(BLOCK % TL_WAVE == 0, "block must be whole waves");
    constexpr int NW = BLOCK / TL_WAVE;
    __shared__ double s_part[NW][NV];
    __shared__ int s_last;
    const int tid = threadIdx.x;
    const int lane = tid & (TL_WAVE - 1), wave = tid / TL_WAVE;
    const unsigned nblocks = gridDim.x * gridDim.y * gridDim.z;
    const unsigned bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);

#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const double s = tl_wave_sum(v[i]);
        if (lane == 0) s_part[wave][i] = s;
    }
    __syncthreads();
    if (tid == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            double s = 0.0;
#pragma unroll
            for (int w = 0; w < NW; ++w) s += s_part[w][i];
            __hip_atomic_store(&rd.partials[(size_t)bid * NV + i], s,
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        const unsigned t = __hip_atomic_fetch_add(rd.ticket, 1u, __ATOMIC_ACQ_REL,
                                                  __HIP_MEMORY_SCOPE_AGENT);
        s_last = (t == nblocks - 1);
    }
    __syncthreads();
    if (!s_last) return;

    // Last block: fixed-order sum of all partials.
    double acc[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) acc[i] = 0.0;
    for (unsigned b = tid; b < nblocks; b += BLOCK) {
#pragma unroll
        for (int i = 0; i < NV; ++i)
            acc[i] += __hip_atomic_load(&rd.partials[(size_t)b * NV + i],
                                        __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();   // s_part reuse
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const double s = tl_wave_sum(acc[i]);
        if (lane == 0) s_part[wave][i] = s;
    }
    __syncthreads();
    if (tid == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            double s = 0.0;
#pragma unroll
            for (int w = 0; w < NW; ++w) s += s_part[w][i];
            rd.result[i] = s;
            __hip_atomic_store(&rd.mailbox[i], s, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_SYSTEM);
        }
        __hip_atomic_store(rd.ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ---------------------------------------------------------------- shapes
struct TlGrid2 {
    int x, y, hd;
};
struct TlGrid3 {
    int x, y, z, hd;
};

static inline unsigned tl_cdiv(long a, long b) { return (unsigned)((a + b - 1) / b); }
