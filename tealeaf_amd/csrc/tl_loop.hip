// Device state, finishers and host side of the device-resident solver loops
// (tl_loop.h).
#include "tl_loop.h"
#include "tealeaf_hip.h"

#include <limits.h>
#include <math.h>

#include <string.h>

#include <atomic>
#include <chrono>

namespace tl {
namespace loop {
namespace {

struct State {
    double rro, alpha, beta;
    double err;          // the Jacobi error of the current iteration
    int last;            // last iteration to run; INT_MAX while the loop goes on
    int base;            // graph replays: the iteration their first one is
};

// One iteration's report in pinned host memory.  The host arms the three
// values with TL_MAIL_EMPTY before it enqueues the iteration's finisher;
// the report is complete once none of them holds the mark (each value is
// its own ready flag; the flags travel as a double).
struct Slot {
    double rrn, err, flags, pad;
};

constexpr int kRing = 8;
static_assert(kRing > kAhead + 1, "a slot must outlive the iterations in flight");

struct Host {
    State* st = nullptr;           // device
    double* alog = nullptr;        // device, [cap]
    double* blog = nullptr;
    int cap = 0;
    Slot* ring = nullptr;          // pinned, host-mapped
    bool active = false;
    int iter = 0;
    bool capturing = false;        // launches go into a graph (iter is relative)
    hipGraphExec_t exec = nullptr;
    int period = 0;
};

Host g_loop;

__global__ void __launch_bounds__(64) k_loop_begin(State* st, double rro)
{
    if (threadIdx.x == 0) {
        st->rro = rro;
        st->last = INT_MAX;
        st->base = 0;
    }
}

__global__ void __launch_bounds__(64) k_loop_base(State* st, int j0, int add)
{
    if (threadIdx.x == 0) st->base = add ? st->base + j0 : j0;
}

// The stop rule and the report of iteration j (thread 0 of a finisher).
__device__ void test_and_report(State* st, Slot* ring, int j, double rrn, const StopRule& rule)
{
    const double err = rule.sqrt_err ? sqrt(rrn) : rrn;
    int flags = 0;
    if (rule.rel ? rrn <= rule.rel_target : fabs(err) < rule.eps) flags |= kConverged;
    if (rule.sw == 1 && err < rule.eps_lim && j + 1 > 20) flags |= kSwitch;
    if (rule.sw == 2 && j + 1 > rule.presteps && err < 1.0) flags |= kSwitch;
    if (flags) st->last = j;
    Slot* s = ring + (j % kRing);
    __hip_atomic_store(&s->rrn, rrn, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&s->err, err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&s->flags, (double)flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// alpha = rro / pw of iteration j (cg_driver.c:88-90, ppcg_driver.c:102-104)
__global__ void __launch_bounds__(256)
k_loop_alpha(unsigned nblocks, const double* __restrict__ partials, State* st, double* alog,
             int j, int rel)
{
    if (rel) j += st->base;
    if (j > st->last) return;                 // uniform: the whole block returns
    double pw[1];
    tl_sum_partials<1>(nblocks, partials, pw);
    if (threadIdx.x == 0) {
        const double a = st->rro / pw[0];
        st->alpha = a;
        alog[j] = a;
    }
}

// beta = rrn / rro, rro = rrn, and the driver's tests of iteration j
// (cg_driver.c:92-99 and the loops of cg_driver.c:30-46, cheby_driver.c,
// ppcg_driver.c); the report goes to ring slot j % kRing.
__global__ void __launch_bounds__(256)
k_loop_beta(unsigned nblocks, const double* __restrict__ partials, State* st, double* blog,
            Slot* ring, int j, int rel, StopRule rule)
{
    if (rel) j += st->base;
    if (j > st->last) return;
    double tot[1];
    tl_sum_partials<1>(nblocks, partials, tot);
    if (threadIdx.x != 0) return;
    const double rrn = tot[0];
    const double b = rrn / st->rro;
    st->beta = b;
    blog[j] = b;
    st->rro = rrn;
    test_and_report(st, ring, j, rrn, rule);
}

// The Jacobi error of iteration j: assigned (jacobi.c:99) or, for the
// residual norm of every 50th iteration, added (calculate_2norm's +=).
__global__ void __launch_bounds__(256)
k_loop_err(unsigned nblocks, const double* __restrict__ partials, State* st, Slot* ring, int j,
           int rel, int mode, StopRule rule)
{
    if (rel) j += st->base;
    if (j > st->last) return;
    double tot[1];
    tl_sum_partials<1>(nblocks, partials, tot);
    if (threadIdx.x != 0) return;
    const double e = mode == kErrAddTest ? st->err + tot[0] : tot[0];
    st->err = e;
    if (mode != kErrSet) test_and_report(st, ring, j, e, rule);
}

volatile unsigned long long* slot_bits(int j)
{
    return (volatile unsigned long long*)(g_loop.ring + (j % kRing));
}

bool slot_ready(volatile unsigned long long* b)
{
    return b[0] != TL_MAIL_EMPTY && b[1] != TL_MAIL_EMPTY && b[2] != TL_MAIL_EMPTY;
}

int hip_fail(hipError_t e, const char* what)
{
    return tl_fail(TLHIP_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

#define LOOP_HIP(call, what)                                  \
    do {                                                      \
        hipError_t _e = (call);                               \
        if (_e != hipSuccess) return hip_fail(_e, what);      \
    } while (0)

}  // namespace

int begin(double rro, int max_iters)
{
    const char* what = "tl::loop::begin";
    TL_TRY(tl_ensure_init());
    Host& h = g_loop;
    if (!h.st) {
        LOOP_HIP(hipMalloc(&h.st, sizeof(State)), what);
        LOOP_HIP(hipHostMalloc(&h.ring, sizeof(Slot) * kRing,
                               hipHostMallocMapped | hipHostMallocCoherent), what);
        for (int i = 0; i < kRing; ++i) h.ring[i] = Slot{};
    }
    const int need = max_iters > 1 ? max_iters : 1;
    if (need > h.cap) {
        LOOP_HIP(hipStreamSynchronize(tl_stream()), what);
        if (h.alog) LOOP_HIP(hipFree(h.alog), what);
        if (h.blog) LOOP_HIP(hipFree(h.blog), what);
        h.alog = h.blog = nullptr;
        h.cap = 0;
        LOOP_HIP(hipMalloc(&h.alog, sizeof(double) * need), what);
        LOOP_HIP(hipMalloc(&h.blog, sizeof(double) * need), what);
        h.cap = need;
    }
    hipLaunchKernelGGL(k_loop_begin, dim3(1), dim3(64), 0, tl_stream(), h.st, rro);
    TL_TRY(tl_check_launch(what));
    h.active = true;
    h.iter = 0;
    return 0;
}

void set_iter(int j) { g_loop.iter = j; }

void end() { g_loop.active = false; }

const double* alpha() { return &g_loop.st->alpha; }
const double* beta() { return &g_loop.st->beta; }

namespace {
// A finisher may be enqueued: a loop is armed, its iteration is in range and
// the partials fit the reduction slots.
int check_finisher(unsigned nblocks, const char* what)
{
    const Host& h = g_loop;
    if (!h.active) return tl_fail(TLHIP_ERR_ARG, "%s: no resident loop is armed", what);
    if (h.iter < 0 || h.iter >= h.cap)
        return tl_fail(TLHIP_ERR_ARG, "%s: iteration %d outside the loop's %d", what, h.iter, h.cap);
    if (nblocks > TL_MAX_RED_BLOCKS)
        return tl_fail(TLHIP_ERR_ARG, "%s: %u reduction blocks exceed %d", what, nblocks,
                       TL_MAX_RED_BLOCKS);
    return 0;
}

// Arm iteration j's report slot.  Its previous report (iteration j - kRing)
// was read before this iteration could be enqueued.
void arm_slot(int j)
{
    volatile unsigned long long* b = slot_bits(j);
    b[0] = b[1] = b[2] = TL_MAIL_EMPTY;
    std::atomic_thread_fence(std::memory_order_seq_cst);
}
}  // namespace

int finish_alpha(unsigned nblocks, const char* what)
{
    Host& h = g_loop;
    TL_TRY(check_finisher(nblocks, what));
    hipLaunchKernelGGL(k_loop_alpha, dim3(1), dim3(256), 0, tl_stream(), nblocks,
                       tl_reduce_slots().partials, h.st, h.alog, h.iter, (int)h.capturing);
    return tl_check_launch(what);
}

int finish_beta(unsigned nblocks, const StopRule& rule, const char* what)
{
    Host& h = g_loop;
    TL_TRY(check_finisher(nblocks, what));
    if (!h.capturing) arm_slot(h.iter);       // a replay arms its own
    hipLaunchKernelGGL(k_loop_beta, dim3(1), dim3(256), 0, tl_stream(), nblocks,
                       tl_reduce_slots().partials, h.st, h.blog, h.ring, h.iter,
                       (int)h.capturing, rule);
    return tl_check_launch(what);
}

int finish_err(unsigned nblocks, ErrMode mode, const StopRule& rule, const char* what)
{
    Host& h = g_loop;
    TL_TRY(check_finisher(nblocks, what));
    if (mode != kErrSet && !h.capturing) arm_slot(h.iter);     // kErrSet does not report
    hipLaunchKernelGGL(k_loop_err, dim3(1), dim3(256), 0, tl_stream(), nblocks,
                       tl_reduce_slots().partials, h.st, h.ring, h.iter, (int)h.capturing,
                       (int)mode, rule);
    return tl_check_launch(what);
}

// Spin on the slot.  Every millisecond the stream is queried: an error
// surfaces there, and an idle stream with no report means the iteration
// was never run (its gate was closed), which the callers rule out by
// waiting only for iterations after the last one that reported.  A stream
// still busy after the wait limit (tlhip_set_wait_limit) fails the call
// with the iteration it waited for and the last report it saw, rather
// than spinning for ever on a kernel that does not finish.
int wait(int j, Report* out)
{
    const char* what = "tl::loop::wait";
    static int last_j = -1;
    static Report last{};
    volatile unsigned long long* b = slot_bits(j);
    if (!slot_ready(b)) {
        const auto start = std::chrono::steady_clock::now();
        auto t0 = start;
        for (unsigned spin = 1; !slot_ready(b); ++spin) {
            if ((spin & 1023u) != 0) continue;
            const auto now = std::chrono::steady_clock::now();
            if (now - t0 < std::chrono::milliseconds(1)) continue;
            t0 = now;
            const hipError_t e = hipStreamQuery(tl_stream());
            if (e == hipErrorNotReady) {
                const double waited = std::chrono::duration<double>(now - start).count();
                if (waited > tl_wait_limit_s())
                    return tl_fail(TLHIP_ERR_HIP,
                                   "%s: iteration %d has not reported after %.0f s (wait limit; "
                                   "last report: iteration %d, rrn %.17g, error %.17g)",
                                   what, j, waited, last_j, last.rrn, last.err);
                continue;
            }
            if (e != hipSuccess) return hip_fail(e, what);
            if (!slot_ready(b))
                return tl_fail(TLHIP_ERR_HIP, "%s: iteration %d did not report", what, j);
        }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    double v[3];
    for (int i = 0; i < 3; ++i) {
        const unsigned long long bits = b[i];
        memcpy(&v[i], &bits, sizeof(double));
    }
    out->rrn = v[0];
    out->err = v[1];
    out->flags = (int)v[2];
    last_j = j;
    last = *out;
    return 0;
}

int capture_begin()
{
    const char* what = "tl::loop::capture_begin";
    Host& h = g_loop;
    if (!h.active) return tl_fail(TLHIP_ERR_ARG, "%s: no resident loop is armed", what);
    drop_graph();
    LOOP_HIP(hipStreamBeginCapture(tl_stream(), hipStreamCaptureModeRelaxed), what);
    h.capturing = true;
    return 0;
}

int capture_end(int period)
{
    const char* what = "tl::loop::capture_end";
    Host& h = g_loop;
    if (!h.capturing) return tl_fail(TLHIP_ERR_ARG, "%s: no capture in progress", what);
    hipLaunchKernelGGL(k_loop_base, dim3(1), dim3(64), 0, tl_stream(), h.st, period, 1);
    const hipError_t le = hipGetLastError();
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(tl_stream(), &graph);
    h.capturing = false;
    if (le != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return hip_fail(le, what);
    }
    if (e != hipSuccess) return hip_fail(e, what);
    const hipError_t ie = hipGraphInstantiateWithFlags(&h.exec, graph, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) {
        h.exec = nullptr;
        return hip_fail(ie, what);
    }
    h.period = period;
    return 0;
}

int set_base(int j0)
{
    hipLaunchKernelGGL(k_loop_base, dim3(1), dim3(64), 0, tl_stream(), g_loop.st, j0, 0);
    return tl_check_launch("tl::loop::set_base");
}

int replay(int j)
{
    const char* what = "tl::loop::replay";
    Host& h = g_loop;
    if (!h.exec) return tl_fail(TLHIP_ERR_ARG, "%s: no graph", what);
    if (j < 0 || j + h.period > h.cap)
        return tl_fail(TLHIP_ERR_ARG, "%s: iterations %d.. outside the loop's %d", what, j, h.cap);
    for (int q = 0; q < h.period; ++q) arm_slot(j + q);
    LOOP_HIP(hipGraphLaunch(h.exec, tl_stream()), what);
    return 0;
}

void drop_graph()
{
    Host& h = g_loop;
    if (h.exec) {
        // replays of it may still be queued (gated iterations past the stop)
        (void)hipStreamSynchronize(tl_stream());
        (void)hipGraphExecDestroy(h.exec);
    }
    h.exec = nullptr;
    h.period = 0;
}

int logs(int j0, int n, double* alphas, double* betas)
{
    const char* what = "tl::loop::logs";
    Host& h = g_loop;
    if (n <= 0) return 0;
    if (j0 < 0 || j0 + n > h.cap) return tl_fail(TLHIP_ERR_ARG, "%s: range outside the logs", what);
    LOOP_HIP(hipMemcpyAsync(alphas, h.alog + j0, sizeof(double) * n, hipMemcpyDeviceToHost,
                            tl_stream()), what);
    LOOP_HIP(hipMemcpyAsync(betas, h.blog + j0, sizeof(double) * n, hipMemcpyDeviceToHost,
                            tl_stream()), what);
    LOOP_HIP(hipStreamSynchronize(tl_stream()), what);
    return 0;
}

void release()
{
    Host& h = g_loop;
    drop_graph();
    if (h.st) (void)hipFree(h.st);
    if (h.alog) (void)hipFree(h.alog);
    if (h.blog) (void)hipFree(h.blog);
    if (h.ring) (void)hipHostFree(h.ring);
    h = Host{};
}

}  // namespace loop
}  // namespace tl

TlGate tl_gate()
{
    const tl::loop::Host& h = tl::loop::g_loop;
    if (!h.active) return TlGate{nullptr, 0};
    if (h.capturing) return TlGate{&h.st->last, h.iter, &h.st->base};
    return TlGate{&h.st->last, h.iter};
}
