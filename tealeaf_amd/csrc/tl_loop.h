// Device-resident solver loops: the host enqueues CG / PPCG / Jacobi
// iterations ahead of the device and reads each iteration's outcome
// afterwards.
//
// The reference's drivers (2d/drivers/cg_driver.c, ppcg_driver.c) return
// two scalars to the host every iteration (pw, then rrn) and form
// alpha = rro / pw and beta = rrn / rro there; each is a round trip that
// idles the device.  Here the finishers of those two reductions form alpha
// and beta on the device, in the same FP64 operations, and later kernels
// of the iteration read them from device memory.  The second finisher also
// applies the driver's stopping rule (convergence, and the CG -> Chebyshev
// / PPCG switch of switch_now) and, when it fires, closes the launch gate:
// every kernel already enqueued for a later iteration then returns at once.
// The host meanwhile keeps kAhead iterations in flight and reads each
// iteration's report from a pinned ring.  Values, iteration counts and the
// order of every floating-point operation are those of the host-driven
// loop.  Jacobi (jacobi_driver.c) has one reduction, the error, whose
// finisher applies the same rule.
#pragma once

#include "tl_common.h"

namespace tl {
namespace loop {

// The stopping rule the second finisher of iteration j applies to rrn.
struct StopRule {
    double eps;          // reference test: |err| < eps
    double rel_target;   // relative test (rel != 0): rrn <= rel_target
    double eps_lim;      // switch rule 1: err < eps_lim && j + 1 > 20
    int sqrt_err;        // err = sqrt(rrn) (the 3D reference test), else err = rrn
    int rel;
    int sw;              // 0: no switch; 1: rule 1 (errswitch); 2: j + 1 > presteps && err < 1
    int presteps;
};

enum : int { kConverged = 1, kSwitch = 2 };

struct Report {
    double rrn, err;
    int flags;           // kConverged | kSwitch
};

constexpr int kAhead = 4;   // iterations in flight beyond the one the host waits for

// Arm a loop: the device's rro takes `rro`, the gate opens, and the
// alpha/beta logs hold at least `max_iters` entries.
int begin(double rro, int max_iters);
// Launches issued from now on belong to iteration j.
void set_iter(int j);
// Launches are ungated again (the device state stays as it is).
void end();
// Device scalars that gated kernels read.
const double* alpha();
const double* beta();
// Finisher of a pw reduction: alpha = rro / pw (cg_driver.c:88-90).
int finish_alpha(unsigned nblocks, const char* what);
// Finisher of an rrn reduction: beta = rrn / rro, rro = rrn, the stop rule,
// and the iteration's report.
int finish_beta(unsigned nblocks, const StopRule& rule, const char* what);
// Finisher of the Jacobi error (jacobi_driver.c:57-83): the sum of the
// partials is assigned to the error (kErrSet: no test yet, the 2-norm of the
// residual follows) or assigned / added to it, and then the stop rule and
// the iteration's report as in finish_beta.
enum ErrMode : int { kErrSet = 0, kErrSetTest = 1, kErrAddTest = 2 };
int finish_err(unsigned nblocks, ErrMode mode, const StopRule& rule, const char* what);
// The report of iteration j, once the device has published it.
int wait(int j, Report* out);
// alpha and beta of iterations j0 .. j0+n-1 into host arrays.
int logs(int j0, int n, double* alphas, double* betas);
// Graph replays of a periodic body (the same launches every `period`
// iterations, field pointers back where they started): capture_begin()
// records the launches issued until capture_end(period) into a graph whose
// iteration numbers count from a device base; set_base(j0) puts the base
// at the first replayed iteration, and replay(j) arms the reports of
// iterations j .. j+period-1 and launches the graph, which ends by moving
// the base on by `period`.  drop_graph() frees it.
int capture_begin();
int capture_end(int period);
int set_base(int j0);
int replay(int j);
void drop_graph();
// Free the loop's device and pinned memory (tlhip_finalise).
void release();

}  // namespace loop

// Gated single-iteration pieces (tl_kernels2d.hip, tl_kernels3d.hip,
// tl_fused2d.hip).  Each launches under tl_gate() and takes its scalar
// from the loop's device state.
namespace loop2d {
int calc_w(int x, int y, int hd, const double* p, double* w, const double* kx, const double* ky);
int calc_ur(int x, int y, int hd, double* u, const double* p, double* r, const double* w,
            const loop::StopRule& rule);
int calc_p(int x, int y, int hd, double* p, const double* r);
// calc_p with the loop's beta fused into the next calc_w (p_in -> p_out, w),
// pw finished into alpha (tlhip2d_cg_calc_pw's pass).
int calc_pw(int x, int y, int hd, const double* p_in, double* p_out, const double* r, double* w,
            const double* kx, const double* ky);
// The CG pair with kx and ky formed from cg_init's conductivity (cond):
// calc_p (beta from the loop, with update) fused into calc_w, w written
// unless null, pw finished into alpha; and cg_calc_ur forming w = A p
// again, rrn finished into beta and the stop rule.  Halo depth >= 2, even x.
int calc_pw_cond(int x, int y, int hd, bool update, const double* p_in, double* p_out,
                 const double* r, double* w, const double* cond, double rx, double ry);
int calc_ur_cond(int x, int y, int hd, double* u, const double* p, double* r, const double* cond,
                 double rx, double ry, const loop::StopRule& rule);
int norm(int x, int y, int hd, const double* buf, const loop::StopRule& rule);
// One Jacobi sweep from u_in into u_out (run_jacobi_iterate without the copy
// of u into r: the sweep reads u_in where the reference reads that copy).
int jacobi(int x, int y, int hd, const double* kx, const double* ky, const double* u0,
           const double* u_in, double* u_out, loop::ErrMode mode, const loop::StopRule& rule);
// The 2-norm of buf added to the Jacobi error.
int norm_err(int x, int y, int hd, const double* buf, const loop::StopRule& rule);
// The fused PPCG launch (calc_ur's alpha from the loop); with norm_nb, the
// launch also leaves the 2-norm of the r it writes as per-block partials,
// their count in *norm_nb, for finish_beta.
int ppcg_inner(int x, int y, int hd, int count, const double* alphas, const double* betas,
               double theta, int init_from_r, const int* neighbours, const double* sd_in,
               double* sd_out, const double* r_in, double* r_out, const double* u_in,
               double* u_out, const double* kx, const double* ky, const double* p,
               const double* w, unsigned* norm_nb = nullptr);
// Whether the FP64 fused PPCG loop runs the LDS-ring kernel (the one that
// can form the 2-norm in-launch; tlhip_set_fused_variant).
bool ring_kernel();
// The same on the host-driven path (alpha from the host), with the 2-norm
// of r fetched to *rrn when rrn is not null.
int ppcg_inner_norm(int x, int y, int hd, int count, const double* alphas, const double* betas,
                    double theta, int init_from_r, const int* neighbours, const double* sd_in,
                    double* sd_out, const double* r_in, double* r_out, const double* u_in,
                    double* u_out, const double* kx, const double* ky, double alpha,
                    const double* p, const double* w, double* rrn);
}  // namespace loop2d

namespace loop3d {
int calc_w(int x, int y, int z, int hd, const double* p, double* w, const double* kx,
           const double* ky, const double* kz);
int calc_ur(int x, int y, int z, int hd, double* u, const double* p, double* r, const double* w,
            const loop::StopRule& rule);
int calc_p(int x, int y, int z, int hd, double* p, const double* r);
// The 3D cond pair (loop2d's, with kz): halo depth >= 2, even x.
int calc_pw_cond(int x, int y, int z, int hd, bool update, const double* p_in, double* p_out,
                 const double* r, double* w, const double* cond, double rx, double ry, double rz);
int calc_ur_cond(int x, int y, int z, int hd, double* u, const double* p, double* r,
                 const double* cond, double rx, double ry, double rz,
                 const loop::StopRule& rule);
// cg_calc_ur with alpha from the device and no finisher (PPCG's outer step,
// whose rrn the driver replaces by the 2-norm after the inner loop).
int calc_ur_only(int x, int y, int z, int hd, double* u, const double* p, double* r,
                 const double* w);
int norm(int x, int y, int z, int hd, const double* buf, const loop::StopRule& rule);
int jacobi(int x, int y, int z, int hd, const double* kx, const double* ky, const double* kz,
           const double* u0, const double* u_in, double* u_out, loop::ErrMode mode,
           const loop::StopRule& rule);
int norm_err(int x, int y, int z, int hd, const double* buf, const loop::StopRule& rule);
}  // namespace loop3d

}  // namespace tl
