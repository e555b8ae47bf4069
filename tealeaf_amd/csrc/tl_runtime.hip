// Process-wide runtime of the hip_cdna4 suite: device selection, the one
// in-order stream, reduction slots, the pinned result mailbox, error state,
// and the memory entry points of the C ABI.
//
// The reference kernel ABI is single-threaded and sequential (one chunk per
// rank, 2d/drivers/*_driver.c call run_* one after another), and so is this
// runtime: calls from several host threads at once are not supported.
#include "tl_ops.h"
#include "tl_loop.h"
#include "tealeaf_hip.h"

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <atomic>
#include <chrono>
#include <vector>

namespace {

struct Runtime {
    int initialised = 0;
    int device = -1;
    hipStream_t stream = nullptr;
    int own_stream = 0;
    TlReduce red{};
    unsigned long long seq = 0;     // finisher launches so far (see tl_fetch)
    double* staging = nullptr;      // device scratch for pack/unpack
    size_t staging_len = 0;
};

// Event pairs around every launch of one entry point's main kernel, for
// measuring average kernel duration on the suite's own stream.
// Up to kTimerNames entry points timed at once (tlhip_kernel_timer_start
// takes a comma-separated list).
constexpr int kTimerNames = 8;
struct KernelTimer {
    char name[kTimerNames][64] = {};
    int nnames = 0;
    int on = 0;
    std::vector<hipEvent_t> ev[kTimerNames];   // 2 per timed launch
    size_t used[kTimerNames] = {};             // launches recorded
};

Runtime g_rt;
KernelTimer g_timer;
// Longest the host waits for a result the device owes it (tlhip_set_wait_limit).
double g_wait_limit_s = 300.0;
thread_local char g_err[512] = "";

}  // namespace

int tl_fail(int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

static int hip_fail(hipError_t e, const char* what)
{
    return tl_fail(TLHIP_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

#define HIP_TRY(call, what)                    \
    do {                                       \
        hipError_t _e = (call);                \
        if (_e != hipSuccess) return hip_fail(_e, what); \
    } while (0)

extern "C" int tlhip_init(int device)
{
    if (g_rt.initialised) {
        if (device >= 0 && device != g_rt.device)
            return tl_fail(TLHIP_ERR_ARG, "tlhip_init: already initialised on device %d",
                           g_rt.device);
        return 0;
    }
    int count = 0;
    HIP_TRY(hipGetDeviceCount(&count), "hipGetDeviceCount");
    if (count < 1) return tl_fail(TLHIP_ERR_HIP, "tlhip_init: no HIP device");
    const int dev = device < 0 ? 0 : device % count;
    HIP_TRY(hipSetDevice(dev), "hipSetDevice");
    TL_TRY(tl_noalloc_setup(dev));
    // Default priority on purpose: streams of mixed priority are not used.
    HIP_TRY(hipStreamCreateWithFlags(&g_rt.stream, hipStreamNonBlocking), "hipStreamCreate");
    g_rt.own_stream = 1;
    HIP_TRY(hipMalloc(&g_rt.red.partials, sizeof(double) * TL_MAX_RED * TL_MAX_RED_BLOCKS),
            "hipMalloc(partials)");
    HIP_TRY(hipMalloc(&g_rt.red.result, sizeof(double) * TL_MAX_RED), "hipMalloc(result)");
    HIP_TRY(hipHostMalloc(&g_rt.red.mailbox, sizeof(double) * TL_MAIL_SLOTS * TL_MAX_RED,
                          hipHostMallocMapped | hipHostMallocCoherent),
            "hipHostMalloc(mailbox)");
    memset(g_rt.red.mailbox, 0, sizeof(double) * TL_MAIL_SLOTS * TL_MAX_RED);
    g_rt.device = dev;
    g_rt.initialised = 1;
    return 0;
}

int tl_ensure_init() { return g_rt.initialised ? 0 : tlhip_init(-1); }

hipStream_t tl_stream() { return g_rt.stream; }
TlReduce tl_reduce_slots() { return g_rt.red; }

int tl_check_launch(const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, what);
    return 0;
}

// Finisher launch `seq` writes its values to mailbox slot seq % TL_MAIL_SLOTS,
// which the host armed (TL_MAIL_EMPTY) just before the launch.  Spinning
// on the slot is a few microseconds quicker than a stream synchronisation,
// and the drivers wait for two scalars per CG iteration.  Every finisher
// is fetched before the next is launched, so a slot is rewritten only
// after its last value was read.  A finisher that does not report within
// 2 ms falls back to polling the stream, which also surfaces any launch
// error; a stream still busy after the wait limit is reported as an error
// instead of blocking the caller for ever.
static volatile unsigned long long* mail_slot(unsigned long long seq)
{
    return (volatile unsigned long long*)(g_rt.red.mailbox +
                                          (size_t)(seq % TL_MAIL_SLOTS) * TL_MAX_RED);
}

static bool mail_ready(volatile unsigned long long* m, int n)
{
    for (int i = 0; i < n; ++i)
        if (m[i] == TL_MAIL_EMPTY) return false;
    return true;
}

double tl_wait_limit_s() { return g_wait_limit_s; }

int tl_fetch(int n, double* out, const char* what)
{
    volatile unsigned long long* m = mail_slot(g_rt.seq);
    if (!mail_ready(m, n)) {
        const auto t0 = std::chrono::steady_clock::now();
        for (unsigned spin = 1; !mail_ready(m, n); ++spin) {
            if ((spin & 1023u) != 0) continue;
            const double waited =
                std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (waited < 2e-3) continue;
            const hipError_t e = hipStreamQuery(g_rt.stream);
            if (e == hipErrorNotReady) {
                if (waited > g_wait_limit_s)
                    return tl_fail(TLHIP_ERR_HIP,
                                   "%s: reduction result not published after %.0f s (stream "
                                   "still busy; wait limit %.0f s, tlhip_set_wait_limit)",
                                   what, waited, g_wait_limit_s);
                continue;
            }
            if (e != hipSuccess) return hip_fail(e, what);
            if (!mail_ready(m, n))
                return tl_fail(TLHIP_ERR_HIP, "%s: reduction result not published", what);
            break;
        }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    for (int i = 0; i < n; ++i) {
        const unsigned long long bits = m[i];
        memcpy(&out[i], &bits, sizeof(double));
    }
    return 0;
}

extern "C" int tlhip_set_wait_limit(double seconds)
{
    if (seconds != seconds) return tl_fail(TLHIP_ERR_ARG, "tlhip_set_wait_limit: NaN");
    g_wait_limit_s = seconds > 0.0 ? seconds : 300.0;
    return 0;
}

extern "C" const char* tlhip_last_error(void) { return g_err; }

extern "C" int tlhip_device(void) { return g_rt.initialised ? g_rt.device : -1; }

extern "C" int tlhip_set_stream(void* stream)
{
    TL_TRY(tl_ensure_init());
    HIP_TRY(hipStreamSynchronize(g_rt.stream), "hipStreamSynchronize");
    if (g_rt.own_stream) {
        HIP_TRY(hipStreamDestroy(g_rt.stream), "hipStreamDestroy");
        g_rt.own_stream = 0;
    }
    g_rt.stream = (hipStream_t)stream;
    return 0;
}

extern "C" void* tlhip_get_stream(void) { return (void*)g_rt.stream; }

extern "C" int tlhip_synchronize(void)
{
    TL_TRY(tl_ensure_init());
    HIP_TRY(hipStreamSynchronize(g_rt.stream), "hipStreamSynchronize");
    return 0;
}

extern "C" int tlhip_finalise(void)
{
    if (!g_rt.initialised) return 0;
    // Teardown: report the first failure but release everything regardless.
    hipError_t first = hipStreamSynchronize(g_rt.stream);
    tl::loop::release();
    hipError_t e[6] = {hipFree(g_rt.red.partials), hipSuccess,
                       hipFree(g_rt.red.result), hipHostFree(g_rt.red.mailbox),
                       g_rt.staging ? hipFree(g_rt.staging) : hipSuccess,
                       g_rt.own_stream ? hipStreamDestroy(g_rt.stream) : hipSuccess};
    for (hipError_t x : e)
        if (first == hipSuccess) first = x;
    g_rt = Runtime{};
    return first == hipSuccess ? 0 : hip_fail(first, "tlhip_finalise");
}

// ------------------------------------------------------------ memory ABI
extern "C" int tlhip_alloc(size_t bytes, void** out)
{
    TL_TRY(tl_ensure_init());
    *out = nullptr;
    void* p = nullptr;
    HIP_TRY(hipMalloc(&p, bytes ? bytes : 8), "hipMalloc");
    hipError_t e = hipMemsetAsync(p, 0, bytes ? bytes : 8, g_rt.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g_rt.stream);
    if (e != hipSuccess) {
        (void)hipFree(p);
        return hip_fail(e, "tlhip_alloc zero fill");
    }
    *out = p;
    return 0;
}

extern "C" int tlhip_free(void* ptr)
{
    if (!ptr) return 0;
    TL_TRY(tl_ensure_init());
    HIP_TRY(hipStreamSynchronize(g_rt.stream), "hipStreamSynchronize");
    HIP_TRY(hipFree(ptr), "hipFree");
    return 0;
}

extern "C" int tlhip_copy_to_device(void* dst, const void* src, size_t bytes)
{
    TL_TRY(tl_ensure_init());
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g_rt.stream), "H2D copy");
    HIP_TRY(hipStreamSynchronize(g_rt.stream), "H2D sync");
    return 0;
}

extern "C" int tlhip_copy_to_host(void* dst, const void* src, size_t bytes)
{
    TL_TRY(tl_ensure_init());
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g_rt.stream), "D2H copy");
    HIP_TRY(hipStreamSynchronize(g_rt.stream), "D2H sync");
    return 0;
}

extern "C" int tlhip_copy_device(void* dst, const void* src, size_t bytes)
{
    TL_TRY(tl_ensure_init());
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, g_rt.stream), "D2D copy");
    return 0;
}

// ------------------------------------------------------------ kernel timer
bool tl_timer_active() { return g_timer.on != 0; }

int tl_timer_begin(const char* what)
{
    if (!g_timer.on) return -1;
    int idx = 0;
    while (idx < g_timer.nnames && strcmp(what, g_timer.name[idx]) != 0) ++idx;
    if (idx == g_timer.nnames) return -1;
    std::vector<hipEvent_t>& ev = g_timer.ev[idx];
    const size_t k = g_timer.used[idx];
    if (k >= (1u << 24)) return -1;
    while (ev.size() < 2 * (k + 1)) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return -1;
        ev.push_back(e);
    }
    if (hipEventRecord(ev[2 * k], g_rt.stream) != hipSuccess) return -1;
    return (idx << 24) | (int)k;
}

void tl_timer_end(int code)
{
    if (code < 0) return;
    const int idx = code >> 24;
    const size_t k = (size_t)(code & 0xFFFFFF);
    if (hipEventRecord(g_timer.ev[idx][2 * k + 1], g_rt.stream) == hipSuccess)
        g_timer.used[idx] = k + 1;
}

extern "C" int tlhip_kernel_timer_start(const char* entry_points)
{
    TL_TRY(tl_ensure_init());
    if (!entry_points) return tl_fail(TLHIP_ERR_ARG, "tlhip_kernel_timer_start: null name");
    int n = 0;
    const char* p = entry_points;
    while (true) {
        const char* q = strchr(p, ',');
        const size_t len = q ? (size_t)(q - p) : strlen(p);
        if (len == 0 || len >= sizeof(g_timer.name[0]) || n == kTimerNames)
            return tl_fail(TLHIP_ERR_ARG, "tlhip_kernel_timer_start: bad name list '%s'",
                           entry_points);
        memcpy(g_timer.name[n], p, len);
        g_timer.name[n][len] = 0;
        ++n;
        if (!q) break;
        p = q + 1;
    }
    g_timer.nnames = n;
    for (int i = 0; i < kTimerNames; ++i) g_timer.used[i] = 0;
    g_timer.on = 1;
    return 0;
}

static int timer_sum(int idx, double* total_ms, long* launches)
{
    double sum = 0.0;
    for (size_t k = 0; k < g_timer.used[idx]; ++k) {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, g_timer.ev[idx][2 * k], g_timer.ev[idx][2 * k + 1]),
                "hipEventElapsedTime");
        sum += ms;
    }
    *total_ms = sum;
    *launches = (long)g_timer.used[idx];
    return 0;
}

extern "C" int tlhip_kernel_timer_read_entry(const char* entry_point, double* total_ms,
                                             long* launches)
{
    TL_TRY(tl_ensure_init());
    if (!entry_point || !total_ms || !launches)
        return tl_fail(TLHIP_ERR_ARG, "tlhip_kernel_timer_read_entry: null argument");
    HIP_TRY(hipStreamSynchronize(g_rt.stream), "timer sync");
    for (int i = 0; i < g_timer.nnames; ++i)
        if (strcmp(entry_point, g_timer.name[i]) == 0) return timer_sum(i, total_ms, launches);
    return tl_fail(TLHIP_ERR_ARG, "tlhip_kernel_timer_read_entry: '%s' is not being timed",
                   entry_point);
}

extern "C" int tlhip_kernel_timer_read(double* total_ms, long* launches)
{
    TL_TRY(tl_ensure_init());
    if (!total_ms || !launches) return tl_fail(TLHIP_ERR_ARG, "tlhip_kernel_timer_read: null argument");
    HIP_TRY(hipStreamSynchronize(g_rt.stream), "timer sync");
    if (g_timer.nnames > 0) {
        TL_TRY(timer_sum(0, total_ms, launches));
    } else {
        *total_ms = 0.0;
        *launches = 0;
    }
    g_timer.on = 0;
    for (int i = 0; i < kTimerNames; ++i) g_timer.used[i] = 0;
    return 0;
}

// A plain streaming copy kernel on the suite's stream: the reference point
// for scaling HBM byte counters (read n*8 bytes, write n*8 bytes).
extern "C" int tlhip_stream_copy(double* dst, const double* src, size_t n)
{
    TL_TRY(tl_ensure_init());
    if (!dst || !src) return tl_fail(TLHIP_ERR_ARG, "tlhip_stream_copy: null argument");
    const int t = tl_timer_begin("tlhip_stream_copy");
    hipLaunchKernelGGL(tl::k_copy_all, dim3(tl::flat_blocks((long)n)), dim3(tl::kBlock), 0,
                       g_rt.stream, (long)n, src, dst);
    tl_timer_end(t);
    return tl_check_launch("tlhip_stream_copy");
}

// Device staging area for halo pack/unpack (grown on demand, never shrunk).
int tl_staging(size_t doubles, double** out)
{
    TL_TRY(tl_ensure_init());
    if (doubles > g_rt.staging_len) {
        HIP_TRY(hipStreamSynchronize(g_rt.stream), "hipStreamSynchronize");
        if (g_rt.staging) HIP_TRY(hipFree(g_rt.staging), "hipFree(staging)");
        g_rt.staging = nullptr;
        g_rt.staging_len = 0;
        HIP_TRY(hipMalloc(&g_rt.staging, doubles * sizeof(double)), "hipMalloc(staging)");
        g_rt.staging_len = doubles;
    }
    *out = g_rt.staging;
    return 0;
}

// Grid total of the per-block partials (tl_sum_partials), published to
// rd.result and to the armed mailbox slot `mail` (plain system-scope
// stores: each value is its own ready flag).
template <int NV>
__global__ void __launch_bounds__(256)
k_finish_reduce(unsigned nblocks, TlReduce rd, double* mail)
{
    double tot[NV];
    tl_sum_partials<NV>(nblocks, rd.partials, tot);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            rd.result[i] = tot[i];
            __hip_atomic_store(&mail[i], tot[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

int tl_finish_reduce(int nv, unsigned nblocks, const char* what)
{
    if (nblocks > TL_MAX_RED_BLOCKS)
        return tl_fail(TLHIP_ERR_ARG, "%s: %u reduction blocks exceed %d", what, nblocks,
                       TL_MAX_RED_BLOCKS);
    if (nv != 1 && nv != 4)
        return tl_fail(TLHIP_ERR_ARG, "%s: unsupported reduction width %d", what, nv);
    const unsigned long long seq = g_rt.seq + 1;
    volatile unsigned long long* m = mail_slot(seq);
    for (int i = 0; i < TL_MAX_RED; ++i) m[i] = TL_MAIL_EMPTY;
    std::atomic_thread_fence(std::memory_order_seq_cst);
    double* mail = (double*)m;
    if (nv == 1)
        hipLaunchKernelGGL(k_finish_reduce<1>, dim3(1), dim3(256), 0, g_rt.stream, nblocks,
                           g_rt.red, mail);
    else
        hipLaunchKernelGGL(k_finish_reduce<4>, dim3(1), dim3(256), 0, g_rt.stream, nblocks,
                           g_rt.red, mail);
    TL_TRY(tl_check_launch(what));
    g_rt.seq = seq;
    return 0;
}

namespace tl {
__global__ void __launch_bounds__(kBlock)
k_copy_all(long n, const double* __restrict__ src, double* __restrict__ dst)
{
    for (long c = blockIdx.x * (long)kBlock + threadIdx.x; c < n; c += (long)gridDim.x * kBlock)
        dst[c] = src[c];
}
}  // namespace tl

// In-bounds check of the DPP lane-shift helpers: out[l] = in[l-1] (or -1 at
// lane 0) and out[64+l] = in[l+1] (or -1 at lane 63), for one wavefront.
__global__ void __launch_bounds__(64) k_lane_selftest(const double* in, double* out)
{
    const int l = threadIdx.x;
    const double v = in[l];
    out[l] = tl_lane_prev(v, -1.0);
    out[64 + l] = tl_lane_next(v, -1.0);
}

extern "C" int tlhip_selftest_lanes(const double* in_dev, double* out_dev)
{
    TL_TRY(tl_ensure_init());
    hipLaunchKernelGGL(k_lane_selftest, dim3(1), dim3(64), 0, g_rt.stream, in_dev, out_dev);
    TL_TRY(tl_check_launch("k_lane_selftest"));
    HIP_TRY(hipStreamSynchronize(g_rt.stream), "selftest sync");
    return 0;
}
