// Native timestep driver over the hip_cdna4 entry points.
//
// Follows the single-chunk, NO_MPI control flow of the reference host:
// 2d/diffuse.c:22-80 (solve), 2d/drivers/cg_driver.c, cheby_driver.c,
// ppcg_driver.c, jacobi_driver.c, eigenvalue_driver.c,
// solve_finished_driver.c, halo_update_driver.c (3D: 3d/diffuse.c and
// 3d/drivers/, which differ only by testing sqrt(error) for convergence).
// The driver state that changes results is kept as the reference keeps it:
// the fields-to-exchange flags persist between driver calls, the CG and
// Chebyshev coefficient arrays persist across timesteps, and alpha/beta and
// the eigenvalue estimate are computed on the host in FP64.
//
// Every field lives on the device.  The CG, PPCG and Jacobi iterations run
// device-resident (tl_loop.h): alpha, beta, the Jacobi error and the stopping rule are
// evaluated on the device, the host enqueues iterations ahead and reads one
// small report per iteration, and the results are those of the host-driven
// loop bit for bit (tlhip_set_resident(0) selects that loop).
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <vector>

#include "tealeaf_hip.h"
#include "tl_adjoint.h"
#include "tl_loop.h"
#include "tl_mixed.h"

int tl_fail(int code, const char* fmt, ...);

#define TRY(expr)                      \
    do {                               \
        int _rc = (expr);              \
        if (_rc) return _rc;           \
    } while (0)

namespace {

enum { F_DENSITY, F_ENERGY0, F_ENERGY1, F_U, F_P, F_SD };

// Inner steps per fused PPCG launch (2D, single chunk): the whole inner
// loop of the default deck (10 steps) in one launch.  At 4096^2 on MI355X
// 10 FP64 steps take ~0.43 ms in one launch against ~0.25 ms per 5 (one
// pass over memory either way; the deeper kernel holds 1 wave per SIMD
// instead of 2); in FP32 at 8192^2, 0.94 ms against 0.50 + 0.56 ms.
int g_ppcg_fusion = 10;
int g_ppcg_fused_norm = 1;   // tlhip_set_ppcg_fused_norm
// Chebyshev iterations per fused launch (2D, FP64, reference stop test).
int g_cheby_fusion = 10;
int g_cg_fusion = 1;     // fuse calc_p into the next calc_w (2D, tlhip2d_cg_calc_pw)
int g_cg_cond = 1;       // 2D CG passes read the conductivity for kx, ky (tlhip2d_cg_*_cond)
int g_resident = 1;      // device-resident CG / PPCG iterations (tl_loop.h)
// Replay the resident loop's iterations from a HIP graph on small grids
// (up to kGraphCells cells with halo, where an iteration is a handful of
// launches of a few microseconds).  Off by default: on MI355X with ROCm 7.2
// the replays measured 5-9% slower than launch-by-launch enqueueing (wall
// time per iteration of a timestep, setup included: 512^2 CG 0.0478 against
// 0.0455 ms, 1024^2 CG 0.0625 against 0.0584, 128^3 CG 0.0617 against
// 0.0580): the loop is bound by each kernel's dispatch and drain on the
// device, not by the host.
int g_graphs = 0;
constexpr long kGraphCells = 2500000;

double* dalloc(long n, int* rc)
{
    void* p = nullptr;
    if (*rc == 0) *rc = tlhip_alloc(sizeof(double) * (size_t)(n > 0 ? n : 1), &p);
    return (double*)p;
}

// ------------------------------------------------------------ eigenvalues
// Symmetric tridiagonal eigenvalues by implicit QL with Wilkinson shifts
// (Numerical Recipes in C, 2nd ed. sec. 11.3), as 2d/drivers/eigenvalue_driver.c:71-121
// uses it, including carrying on through an exact zero rotation radius.
// d: diagonal -> eigenvalues; e: sub-diagonal in e[1..n-1].
int ql_eigenvalues(std::vector<double>& d, std::vector<double>& e)
{
    const int n = (int)d.size();
    for (int i = 0; i + 1 < n; ++i) e[i] = e[i + 1];
    if (n) e[n - 1] = 0.0;
    auto sgn = [](double a, double b) { return b < 0 ? -std::fabs(a) : std::fabs(a); };
    for (int l = 0; l < n; ++l) {
        int sweeps = 0, m;
        do {
            for (m = l; m < n - 1; ++m) {
                const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) + dd == dd) break;
            }
            if (m == l) break;
            if (sweeps++ == 30) return tl_fail(TLHIP_ERR_SOLVER, "Too many iterations in TQLI routine");
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double rad = std::sqrt(g * g + 1.0);
            g = d[m] - d[l] + e[l] / (g + sgn(rad, g));
            double s = 1.0, c = 1.0, p = 0.0;
            for (int i = m - 1; i >= l; --i) {
                const double f = s * e[i], b = c * e[i];
                rad = std::sqrt(f * f + g * g);
                e[i + 1] = rad;
                if (rad == 0.0) {
                    d[i + 1] -= p;
                    e[m] = 0.0;
                    continue;
                }
                s = f / rad;
                c = g / rad;
                g = d[i + 1] - p;
                rad = (d[i] - g) * s + 2.0 * c * b;
                p = s * rad;
                d[i + 1] = g + p;
                g = c * rad - b;
            }
            d[l] = d[l] - p;
            e[l] = g;
            e[m] = 0.0;
        } while (m != l);
    }
    return 0;
}

}  // namespace

struct tlhip_problem {
    tlhip_settings s{};
    int dims = 2, x = 0, y = 0, z = 1, hd = 2;
    long n = 0;
    double dx = 0, dy = 0, dz = 0;
    double *density0 = nullptr, *density = nullptr, *energy0 = nullptr, *energy = nullptr;
    double *u = nullptr, *u0 = nullptr, *p = nullptr, *r = nullptr, *mi = nullptr, *w = nullptr;
    double *kx = nullptr, *ky = nullptr, *kz = nullptr, *sd = nullptr;
    double *u_alt = nullptr, *sd_alt = nullptr, *r_alt = nullptr, *p_alt = nullptr;
    double* cond = nullptr;   // cg_init's conductivity w, for the cond CG passes
    double rx = 0.0, ry = 0.0, rz = 0.0;
    double *volume = nullptr, *x_area = nullptr, *y_area = nullptr, *z_area = nullptr;
    double *cell_x = nullptr, *cell_y = nullptr, *cell_z = nullptr;
    double *vertex_x = nullptr, *vertex_y = nullptr, *vertex_z = nullptr;
    std::vector<double> cg_alphas, cg_betas, cheby_alphas, cheby_betas;
    double theta = 0, eigmin = 0, eigmax = 0;
    int neighbours[6] = {-1, -1, -1, -1, -1, -1};
    bool fields[6] = {};
    tlhip_step_report* rep = nullptr;
    // A problem over a chunk the caller allocated (tlhip_problem_adopt): the
    // caller's buffers of u, sd, r and p, which the partners swap with.
    bool adopted = false;
    double* home[4] = {};

    // How the solve runs (tlhip_problem_set_solve).
    int precision = 0;       // 0: FP64; 1: FP32 sweeps + FP64 residual replacement
    double rel_tol = 0.0;    // > 0: stop on ||u0 - A u|| <= rel_tol ||u0||
    double bb = 0.0;         // ||u0||^2 of the current solve
    // FP32 fields of the mixed-precision solve, row pitch P.
    int P = 0;
    float *kx32 = nullptr, *ky32 = nullptr, *p32 = nullptr, *w32 = nullptr, *r32 = nullptr,
          *d32 = nullptr, *sd32 = nullptr, *sd32b = nullptr, *r32b = nullptr, *d32b = nullptr,
          *p32b = nullptr;
    float* kz32 = nullptr;       // 3D
    double* alpha32 = nullptr;   // device: a PPCG step's cg_calc_ur alpha (rro / pw)
    // A CG/PPCG step's calc_p waits to be fused into the next step's calc_w
    // (tlhip2d_cg_calc_pw); flush_p() applies it where p is read otherwise.
    bool pending_p = false;
    double pending_beta = 0.0;

    // -------------------------------------------------------- kernels
    int local_halos(int depth)
    {
        if (dims == 2)
            return tlhip2d_local_halos(x, y, depth, hd, neighbours, fields, density, energy0,
                                       energy, u, p, sd);
        return tlhip3d_local_halos(x, y, z, depth, hd, neighbours, fields, density, energy0,
                                   energy, u, p, sd);
    }
    // halo_update_driver.c:6-25
    int halo_update(int depth)
    {
        for (bool f : fields)
            if (f) return local_halos(depth);
        return 0;
    }
    void reset_fields() { std::memset(fields, 0, sizeof(fields)); }
    int calc_w(double* pw)
    {
        return dims == 2 ? tlhip2d_cg_calc_w(x, y, hd, pw, p, w, kx, ky)
                         : tlhip3d_cg_solver_calc_w(x, y, z, hd, pw, p, w, kx, ky, kz);
    }
    int calc_ur(double alpha, double* rrn)
    {
        return dims == 2 ? tlhip2d_cg_calc_ur(x, y, hd, alpha, rrn, u, p, r, w)
                         : tlhip3d_cg_solver_calc_ur(x, y, z, hd, alpha, rrn, u, p, r, w);
    }
    int calc_p(double beta)
    {
        return dims == 2 ? tlhip2d_cg_calc_p(x, y, hd, beta, p, r)
                         : tlhip3d_cg_solver_calc_p(x, y, z, hd, beta, p, r);
    }
    int copy_u()
    {
        return dims == 2 ? tlhip2d_copy_u(x, y, hd, u0, u) : tlhip3d_solver_copy_u(x, y, z, hd, u0, u);
    }
    int residual()
    {
        return dims == 2 ? tlhip2d_calculate_residual(x, y, hd, u, u0, r, kx, ky)
                         : tlhip3d_calculate_residual(x, y, z, hd, u, u0, r, kx, ky, kz);
    }
    int norm2(const double* b, double* out)
    {
        return dims == 2 ? tlhip2d_calculate_2norm(x, y, hd, b, out)
                         : tlhip3d_calculate_2norm(x, y, z, hd, b, out);
    }
    int store_energy()
    {
        return dims == 2 ? tlhip2d_store_energy(x, y, energy0, energy)
                         : tlhip3d_store_energy(x, y, z, energy0, energy);
    }
    int field_summary(double* o)
    {
        o[0] = o[1] = o[2] = o[3] = 0.0;
        return dims == 2
                   ? tlhip2d_field_summary(x, y, hd, volume, density, energy0, u, &o[0], &o[1], &o[2], &o[3])
                   : tlhip3d_field_summary(x, y, z, hd, volume, density, energy0, u, &o[0], &o[1], &o[2], &o[3]);
    }

    // -------------------------------------------------------- CG (cg_driver.c)
    int cg_init_driver(double rx_, double ry_, double rz_, double* rro)
    {
        *rro = 0.0;
        rx = rx_;
        ry = ry_;
        rz = rz_;
        if (dims == 2) {
            TRY(tlhip2d_cg_init(x, y, hd, s.coefficient, rx, ry, rro, density, energy, u, p, r, w, kx, ky));
            if (cond_ok()) TRY(tlhip2d_cg_cond(x, y, hd, s.coefficient, density, cond));
        } else {
            TRY(tlhip3d_cg_solver_init(x, y, z, hd, s.coefficient, rx, ry, rz, rro, density,
                                       energy, u, p, r, w, kx, ky, kz));
            if (cond3_ok()) TRY(tlhip3d_cg_cond(x, y, z, hd, s.coefficient, density, cond));
        }
        reset_fields();
        fields[F_U] = fields[F_P] = true;
        TRY(halo_update(1));
        TRY(copy_u());
        return start_solve();
    }

    // -------------------------------------------------------- deferred calc_p
    // 2D: the p = beta p + r of one step and the w = A p of the next run as
    // one pass (p ping-pongs with p_alt); the reference's halo update of p in
    // between is the kernel's mirror.  The results are the reference's bit
    // for bit.
    // 2D only: in 3D the fused pass (tlhip3d_cg_solver_calc_pw) reads p and r
    // at six neighbours each and measured no faster inside a solve (128^3 CG
    // 0.058 ms per iteration either way, 160^3 0.125 against 0.101, 256^3
    // 0.232 against 0.217 ms per pass).
    bool can_defer_p() const { return g_cg_fusion && dims == 2 && x % 2 == 0 && p_alt; }

    // 2D CG steps (deferred calc_p, halo depth >= 2) run as the cond pair:
    // calc_pw reads cg_init's conductivity (cond) in place of kx and ky and
    // keeps w in registers; calc_ur forms w = A p again from p and cond.
    // u, r and p are the reference's values; rrn is summed in another order.
    // PPCG's calc_pw, whose w the fused inner loop reads, reads cond too and
    // stores w (bit for bit the kx/ky pass).
    bool cond_ok() const
    {
        return g_cg_cond && can_defer_p() && hd >= 2 && cond &&
               (s.coefficient == 1 || s.coefficient == 2);
    }

    // 3D: the CG steps as the 3D cond pair, calc_p deferred into the next
    // calc_pw (the pending calc_p applied as calc_p + halo update before
    // any other reader of p: PPCG's calc_w, Chebyshev's start).
    bool cond3_ok() const
    {
        return g_cg_cond && g_cg_fusion && dims == 3 && x % 2 == 0 && hd >= 2 && cond &&
               p_alt && (s.coefficient == 1 || s.coefficient == 2);
    }

    int cg_step_cond3(double* rro, double* alpha, double* rrn)
    {
        double pw = 0.0;
        if (pending_p) {
            pending_p = false;
            TRY(tlhip3d_cg_calc_pw_cond(x, y, z, hd, 1, pending_beta, rx, ry, rz, &pw, p, p_alt,
                                        r, nullptr, cond));
            std::swap(p, p_alt);
        } else {
            TRY(tlhip3d_cg_calc_pw_cond(x, y, z, hd, 0, 0.0, rx, ry, rz, &pw, p, nullptr,
                                        nullptr, nullptr, cond));
        }
        *alpha = *rro / pw;
        *rrn = 0.0;
        return tlhip3d_cg_calc_ur_cond(x, y, z, hd, *alpha, rx, ry, rz, rrn, u, p, r, cond);
    }

    // cg_main_step's calc_w (or pending calc_p + calc_w) and calc_ur as the
    // cond pair, alpha and beta on the host.
    int cg_step_cond(double* rro, double* alpha, double* rrn)
    {
        double pw = 0.0;
        if (pending_p) {
            pending_p = false;
            TRY(tlhip2d_cg_calc_pw_cond(x, y, hd, 1, pending_beta, rx, ry, &pw, p, p_alt, r,
                                        nullptr, cond));
            std::swap(p, p_alt);
        } else {
            TRY(tlhip2d_cg_calc_pw_cond(x, y, hd, 0, 0.0, rx, ry, &pw, p, nullptr, nullptr,
                                        nullptr, cond));
        }
        *alpha = *rro / pw;
        *rrn = 0.0;
        return tlhip2d_cg_calc_ur_cond(x, y, hd, *alpha, rx, ry, rrn, u, p, r, cond);
    }

    int calc_w_pending(double* pw)
    {
        if (!pending_p) return calc_w(pw);
        if (dims == 3) {
            TRY(flush_p());
            return calc_w(pw);
        }
        pending_p = false;
        if (cond_ok())
            TRY(tlhip2d_cg_calc_pw_cond(x, y, hd, 1, pending_beta, rx, ry, pw, p, p_alt, r, w,
                                        cond));
        else
            TRY(tlhip2d_cg_calc_pw(x, y, hd, pending_beta, pw, p, p_alt, r, w, kx, ky));
        std::swap(p, p_alt);
        return 0;
    }

    int calc_p_deferred(double beta)
    {
        if (!can_defer_p() && !cond3_ok()) return calc_p(beta);
        pending_p = true;
        pending_beta = beta;
        return 0;
    }

    // Apply a deferred calc_p and the depth-1 halo update of p that followed it.
    int flush_p()
    {
        if (!pending_p) return 0;
        pending_p = false;
        TRY(calc_p(pending_beta));
        bool keep[6];
        std::memcpy(keep, fields, sizeof(fields));
        reset_fields();
        fields[F_P] = true;
        const int rc = local_halos(1);
        std::memcpy(fields, keep, sizeof(fields));
        return rc;
    }

    int cg_main_step(int tt, double* rro, double* error)
    {
        double alpha = 0.0, rrn = 0.0;
        if (cond_ok()) {
            TRY(cg_step_cond(rro, &alpha, &rrn));
        } else if (cond3_ok()) {
            TRY(cg_step_cond3(rro, &alpha, &rrn));
        } else {
            double pw = 0.0;
            TRY(calc_w_pending(&pw));
            alpha = *rro / pw;
            TRY(calc_ur(alpha, &rrn));
        }
        cg_alphas[tt] = alpha;
        const double beta = rrn / *rro;
        cg_betas[tt] = beta;
        TRY(calc_p_deferred(beta));
        *error = rrn;
        *rro = rrn;
        rep->n_cg_steps++;
        return 0;
    }

    // 3D tests sqrt(error) (3d/drivers/cg_solver_driver.c:24-25), 2D error.
    bool converged(double* error)
    {
        if (dims == 3) *error = std::sqrt(*error);
        return std::fabs(*error) < s.eps;
    }

    // ||u0 - A u||^2 in FP64, with u's depth-1 halo refreshed first (PPCG
    // leaves it stale); w serves as scratch.
    int true_residual(double* rr)
    {
        bool keep[6];
        std::memcpy(keep, fields, sizeof(fields));
        reset_fields();
        fields[F_U] = true;
        int rc = local_halos(1);
        std::memcpy(fields, keep, sizeof(fields));
        TRY(rc);
        if (dims == 2) TRY(tlhip2d_calculate_residual(x, y, hd, u, u0, w, kx, ky));
        else TRY(tlhip3d_calculate_residual(x, y, z, hd, u, u0, w, kx, ky, kz));
        *rr = 0.0;
        TRY(norm2(w, rr));
        rep->n_true_residuals++;
        rep->rel_residual = bb > 0.0 ? std::sqrt(*rr / bb) : 0.0;
        return 0;
    }

    // The stopping test of the FP64 solvers: the reference's, or with
    // rel_tol set, a recurrence residual ||r||^2 = error at or below
    // rel_tol^2 ||u0||^2 that the true FP64 residual confirms.
    int stop(double* error, bool* done)
    {
        if (rel_tol <= 0.0) {
            *done = converged(error);
            return 0;
        }
        *done = false;
        const double target = rel_tol * rel_tol * bb;
        if (!(*error <= target)) return 0;
        double rr;
        TRY(true_residual(&rr));
        *done = rr <= target;
        return 0;
    }

    int start_solve()
    {
        bb = 0.0;
        if (rel_tol > 0.0 || precision == 1) TRY(norm2(u0, &bb));
        return 0;
    }

    // -------------------------------------------------------- resident loops
    // CG / PPCG iterations enqueued ahead of the host (tl_loop.h).  The
    // bodies launch the same kernels as cg_main_step / ppcg_main_step with
    // the same arguments, gated and with alpha and beta read on the device.
    bool interior() const
    {
        return x > 2 * hd && y > 2 * hd && (dims == 2 || z > 2 * hd);
    }
    bool resident_cg() const { return g_resident && interior(); }
    bool resident_ppcg() const
    {
        return resident_cg() && (fused_ppcg() || dims == 3) && s.ppcg_inner_steps > 0;
    }

    // The stop test of stop() and, with sw, the switch_now test of the next
    // iteration (sw 1: the error switch, 2: presteps), evaluated on the device.
    tl::loop::StopRule stop_rule(int sw) const
    {
        tl::loop::StopRule r{};
        r.eps = s.eps;
        r.rel = rel_tol > 0.0;
        r.rel_target = rel_tol * rel_tol * bb;
        r.sqrt_err = dims == 3 && !r.rel;
        r.sw = sw;
        r.eps_lim = s.eps_lim;
        r.presteps = s.presteps;
        return r;
    }

    // What the host would have decided from the same rrn; a device report
    // that disagrees is an error, not a silently different solve.
    int check_report(const tl::loop::Report& rp, int j, int sw) const
    {
        const double err = dims == 3 && rel_tol <= 0.0 ? std::sqrt(rp.rrn) : rp.rrn;
        int want = 0;
        const bool conv = rel_tol > 0.0 ? rp.rrn <= rel_tol * rel_tol * bb : std::fabs(err) < s.eps;
        if (conv) want |= tl::loop::kConverged;
        if (sw && switch_now(0, j + 1, err)) want |= tl::loop::kSwitch;
        const bool same_err = err == rp.err || (err != err && rp.err != rp.err);
        if (want != rp.flags || !same_err)
            return tl_fail(TLHIP_ERR_SOLVER,
                           "resident loop: device report of iteration %d differs from the host's "
                           "test (flags %d, host %d)", j, rp.flags, want);
        return 0;
    }

    // calc_w, or with a calc_p pending, calc_p(beta on the device) fused
    // into it (as calc_w_pending); alpha is finished on the device.
    int calc_w_resident()
    {
        if (!pending_p) return tl::loop2d::calc_w(x, y, hd, p, w, kx, ky);
        pending_p = false;
        if (cond_ok())
            TRY(tl::loop2d::calc_pw_cond(x, y, hd, true, p, p_alt, r, w, cond, rx, ry));
        else
            TRY(tl::loop2d::calc_pw(x, y, hd, p, p_alt, r, w, kx, ky));
        std::swap(p, p_alt);
        return 0;
    }

    int cg_body(const tl::loop::StopRule& rule)
    {
        if (cond_ok()) {
            if (pending_p) {
                pending_p = false;
                TRY(tl::loop2d::calc_pw_cond(x, y, hd, true, p, p_alt, r, nullptr, cond, rx, ry));
                std::swap(p, p_alt);
            } else {
                TRY(tl::loop2d::calc_pw_cond(x, y, hd, false, p, nullptr, nullptr, nullptr, cond,
                                             rx, ry));
            }
            TRY(tl::loop2d::calc_ur_cond(x, y, hd, u, p, r, cond, rx, ry, rule));
            pending_p = true;
            return 0;
        } else if (cond3_ok()) {
            if (pending_p) {
                pending_p = false;
                TRY(tl::loop3d::calc_pw_cond(x, y, z, hd, true, p, p_alt, r, nullptr, cond, rx, ry,
                                             rz));
                std::swap(p, p_alt);
            } else {
                TRY(tl::loop3d::calc_pw_cond(x, y, z, hd, false, p, nullptr, nullptr, nullptr,
                                             cond, rx, ry, rz));
            }
            TRY(tl::loop3d::calc_ur_cond(x, y, z, hd, u, p, r, cond, rx, ry, rz, rule));
            pending_p = true;
            return 0;            // p is mirrored, u read pointwise: halo at the end
        } else if (can_defer_p()) {
            TRY(calc_w_resident());
            TRY(tl::loop2d::calc_ur(x, y, hd, u, p, r, w, rule));
            pending_p = true;    // calc_p waits for the next calc_w (beta stays on the device)
            return 0;            // u and p are read pointwise or mirrored: halo at the end
        } else if (dims == 2) {
            TRY(tl::loop2d::calc_w(x, y, hd, p, w, kx, ky));
            TRY(tl::loop2d::calc_ur(x, y, hd, u, p, r, w, rule));
            TRY(tl::loop2d::calc_p(x, y, hd, p, r));
        } else {
            TRY(tl::loop3d::calc_w(x, y, z, hd, p, w, kx, ky, kz));
            TRY(tl::loop3d::calc_ur(x, y, z, hd, u, p, r, w, rule));
            TRY(tl::loop3d::calc_p(x, y, z, hd, p, r));
        }
        return halo_update(1);
    }

    // ppcg_main_step's fused path (2D) or its step-by-step path (3D), then
    // its halo update of p.
    int ppcg_body(const tl::loop::StopRule& rule)
    {
        if (dims == 3) return ppcg_body3(rule);
        if (can_defer_p()) TRY(calc_w_resident());
        else TRY(tl::loop2d::calc_w(x, y, hd, p, w, kx, ky));
        int done = 0;
        unsigned norm_nb = 0;
        while (done < s.ppcg_inner_steps) {
            const int cnt = std::min(g_ppcg_fusion, s.ppcg_inner_steps - done);
            const bool first = done == 0, last = done + cnt == s.ppcg_inner_steps;
            TRY(tl::loop2d::ppcg_inner(x, y, hd, cnt, &cheby_alphas[done], &cheby_betas[done],
                                       theta, first, neighbours, first ? nullptr : sd, sd_alt, r,
                                       r_alt, u, u_alt, kx, ky, first ? p : nullptr,
                                       first ? w : nullptr,
                                       last && fused_norm() ? &norm_nb : nullptr));
            std::swap(sd, sd_alt);
            std::swap(r, r_alt);
            std::swap(u, u_alt);
            done += cnt;
        }
        reset_fields();
        fields[F_P] = true;
        if (norm_nb > 0)      // the last launch formed r.r (ppcg_driver.c:121-131)
            TRY(tl::loop::finish_beta(norm_nb, rule, "tlhip2d_calculate_2norm"));
        else
            TRY(tl::loop2d::norm(x, y, hd, r, rule));
        if (can_defer_p()) {
            pending_p = true;
            return 0;            // p is read pointwise or mirrored: halo at the end
        }
        TRY(tl::loop2d::calc_p(x, y, hd, p, r));
        return halo_update(1);
    }

    // Iteration j of jacobi_driver.c's loop (jacobi_driver.c:22-40,
    // 3d/drivers/jacobi_solver_driver.c:15-23): the sweep reads u and writes
    // u_alt, and the two swap (the reference copies u into r and sweeps from
    // the copy into u: the same values, one copy fewer); every 50th
    // iteration the 2-norm of the residual is added to the sweep's error.
    int jacobi_body(const tl::loop::StopRule& rule, int j)
    {
        const bool res = j % 50 == 0;
        const tl::loop::ErrMode mode = res ? tl::loop::kErrSet : tl::loop::kErrSetTest;
        if (dims == 2) TRY(tl::loop2d::jacobi(x, y, hd, kx, ky, u0, u, u_alt, mode, rule));
        else TRY(tl::loop3d::jacobi(x, y, z, hd, kx, ky, kz, u0, u, u_alt, mode, rule));
        std::swap(u, u_alt);
        if (res) {
            TRY(halo_update(1));
            TRY(residual());
            if (dims == 2) TRY(tl::loop2d::norm_err(x, y, hd, r, rule));
            else TRY(tl::loop3d::norm_err(x, y, z, hd, r, rule));
        }
        return halo_update(1);
    }

    // 3D PPCG inner steps: two per launch (tlhip3d_ppcg_inner_iterations,
    // the halo updates of sd mirrored in the kernel), a last odd one after a
    // halo update; one per launch after a halo update when pairs are off.
    bool pair3d() const
    {
        for (int f : neighbours)
            if (f != -1) return false;
        return dims == 3 && g_ppcg_fusion > 1 && x % 2 == 0 && r_alt && u_alt;
    }

    int inner_steps3()
    {
        int pp = 0;
        if (pair3d()) {
            for (; pp + 2 <= s.ppcg_inner_steps; pp += 2) {
                TRY(tlhip3d_ppcg_inner_iterations(x, y, z, hd, 2, &cheby_alphas[pp],
                                                  &cheby_betas[pp], neighbours, sd, sd_alt, r,
                                                  r_alt, u, u_alt, kx, ky, kz));
                std::swap(sd, sd_alt);
                std::swap(r, r_alt);
                std::swap(u, u_alt);
            }
        }
        for (; pp < s.ppcg_inner_steps; ++pp) {
            TRY(halo_update(1));
            TRY(tlhip3d_ppcg_inner_iteration(x, y, z, hd, cheby_alphas[pp], cheby_betas[pp], u, r,
                                             kx, ky, kz, sd, sd_alt));
            std::swap(sd, sd_alt);
        }
        return 0;
    }

    // 3D: calc_w, calc_ur (its rrn unused), ppcg_init, the inner steps each
    // after a depth-1 halo update of sd, the 2-norm of r, calc_p
    // (3d/drivers/ppcg_solver_driver.c:88-140).
    int ppcg_body3(const tl::loop::StopRule& rule)
    {
        TRY(flush_p());          // the CG phase's last calc_p, when deferred
        TRY(tl::loop3d::calc_w(x, y, z, hd, p, w, kx, ky, kz));
        TRY(tl::loop3d::calc_ur_only(x, y, z, hd, u, p, r, w));
        TRY(tlhip3d_ppcg_init(x, y, z, hd, theta, r, sd));
        reset_fields();
        fields[F_SD] = true;
        TRY(inner_steps3());
        reset_fields();
        fields[F_P] = true;
        TRY(tl::loop3d::norm(x, y, z, hd, r, rule));
        TRY(tl::loop3d::calc_p(x, y, z, hd, p, r));
        return halo_update(1);
    }

    struct Ptrs {
        double *u, *u_alt, *r, *r_alt, *sd, *sd_alt, *p, *p_alt;
    };

    // Iterations t0, t0+1, ... of `body`, up to kAhead beyond the one whose
    // report the host waits for, until a report carries a flag or the
    // iteration max_iters - 1 has reported.  *last and *out: that
    // iteration and its report; the field pointers are restored to what it
    // left (later iterations were enqueued but their gate was closed), and
    // *rro / *error take its rrn / err.  CG runs also copy their alpha and
    // beta into cg_alphas / cg_betas for the eigenvalue estimate.
    Ptrs ptrs() const { return Ptrs{u, u_alt, r, r_alt, sd, sd_alt, p, p_alt}; }
    static bool same(const Ptrs& a, const Ptrs& b)
    {
        return a.u == b.u && a.u_alt == b.u_alt && a.r == b.r && a.r_alt == b.r_alt &&
               a.sd == b.sd && a.sd_alt == b.sd_alt && a.p == b.p && a.p_alt == b.p_alt;
    }

    // Capture two iterations of a periodic body (its launches are the same
    // every two iterations and the field pointers come back) into a graph
    // whose replays run iterations g0, g0+1, then g0+2, g0+3, ...  *mid:
    // the pointers after the first of the two.  Returns false (and leaves
    // eager enqueueing) if the body turns out not to be periodic.
    template <class Body>
    int capture_pair(int g0, const tl::loop::StopRule& rule, Body&& body, Ptrs* mid, bool* ok)
    {
        *ok = false;
        const Ptrs start = ptrs();
        const bool pending0 = pending_p;
        TRY(tl::loop::set_base(g0));
        TRY(tl::loop::capture_begin());
        tl::loop::set_iter(0);
        int rc = body(rule, g0);
        *mid = ptrs();
        if (!rc) {
            tl::loop::set_iter(1);
            rc = body(rule, g0 + 1);
        }
        const int rc2 = tl::loop::capture_end(2);
        TRY(rc);
        TRY(rc2);
        if (!same(ptrs(), start)) {       // not periodic: undo the host-side effects
            tl::loop::drop_graph();
            const Ptrs& q = start;
            u = q.u, u_alt = q.u_alt, r = q.r, r_alt = q.r_alt, sd = q.sd, sd_alt = q.sd_alt;
            p = q.p, p_alt = q.p_alt;
            pending_p = pending0;
            return 0;
        }
        *ok = true;
        return 0;
    }

    template <class Body>
    int run_resident(int t0, int sw, bool cg, Body&& body, double* rro, double* error, int* last,
                     tl::loop::Report* out, bool periodic = true, bool halo_after = false)
    {
        if (t0 < 0 || t0 >= s.max_iters)
            return tl_fail(TLHIP_ERR_ARG, "resident loop: no iteration to run from %d", t0);
        constexpr int kSnap = tl::loop::kAhead + 2;
        Ptrs snap[kSnap];
        TRY(tl::loop::begin(*rro, s.max_iters));
        struct Close {
            ~Close()
            {
                tl::loop::drop_graph();
                tl::loop::end();
            }
        } close;
        const tl::loop::StopRule rule = stop_rule(sw);
        // Small grids replay pairs of iterations from a graph after the
        // first iteration (which may differ: calc_w before calc_pw).
        bool graph = periodic && g_graphs && n <= kGraphCells && !tl_timer_active();
        int g0 = -1;
        Ptrs gmid{}, gstart{};
        int next = t0, chk = t0;
        tl::loop::Report rp{};
        for (;;) {
            while (next < s.max_iters && next <= chk + tl::loop::kAhead) {
                if (graph && g0 < 0 && next > t0 && next + 2 <= s.max_iters) {
                    bool ok;
                    gstart = ptrs();
                    TRY(capture_pair(next, rule, body, &gmid, &ok));
                    if (ok) g0 = next;
                    else graph = false;
                }
                if (g0 >= 0 && next + 2 <= s.max_iters && (next - g0) % 2 == 0) {
                    TRY(tl::loop::replay(next));
                    snap[next % kSnap] = gmid;
                    snap[(next + 1) % kSnap] = gstart;
                    next += 2;
                    continue;
                }
                tl::loop::set_iter(next);
                TRY(body(rule, next));
                snap[next % kSnap] = ptrs();
                ++next;
            }
            TRY(tl::loop::wait(chk, &rp));
            TRY(check_report(rp, chk, sw));
            if (rp.flags || chk + 1 >= s.max_iters) break;
            ++chk;
        }
        const Ptrs& q = snap[chk % kSnap];
        u = q.u, u_alt = q.u_alt, r = q.r, r_alt = q.r_alt, sd = q.sd, sd_alt = q.sd_alt;
        p = q.p, p_alt = q.p_alt;
        if (halo_after) {
            // the depth-1 halo update that each iteration of the reference
            // ends with, once: the values depend only on the last iteration
            tl::loop::end();
            TRY(halo_update(1));
        }
        if (cg) TRY(tl::loop::logs(t0, chk - t0 + 1, &cg_alphas[t0], &cg_betas[t0]));
        if (pending_p) {
            // the last iteration's calc_p waits with its beta (host copy for flush_p)
            double a;
            TRY(tl::loop::logs(chk, 1, &a, &pending_beta));
        }
        *last = chk;
        *out = rp;
        *rro = rp.rrn;
        *error = rp.err;
        return 0;
    }

    // The stop() decision for a flagged report (the true residual confirms a
    // relative-residual candidate).
    int confirm(const tl::loop::Report& rp, bool* done)
    {
        *done = false;
        if (!(rp.flags & tl::loop::kConverged)) return 0;
        if (rel_tol <= 0.0) {
            *done = true;
            return 0;
        }
        double rr;
        TRY(true_residual(&rr));
        *done = rr <= rel_tol * rel_tol * bb;
        return 0;
    }

    // CG iterations from t0 until the stop test passes (*done, *tt = that
    // iteration), the switch test of the next iteration passes (*tt = the
    // last CG iteration), or max_iters - 1 has run (*tt = max_iters - 1).
    int cg_resident(int t0, int sw, double* rro, double* error, int* tt, bool* done)
    {
        for (int from = t0;;) {
            int last;
            tl::loop::Report rp;
            TRY(run_resident(from, sw, true,
                             [&](const tl::loop::StopRule& rule, int) { return cg_body(rule); }, rro,
                             error, &last, &rp, true, can_defer_p() || cond3_ok()));
            rep->n_cg_steps += last - from + 1;
            *tt = last;
            TRY(confirm(rp, done));
            if (*done || (rp.flags & tl::loop::kSwitch) || last + 1 >= s.max_iters) return 0;
            from = last + 1;    // a relative-residual candidate the true residual refused
        }
    }

    int cg_driver(double rx, double ry, double rz, double* error)
    {
        double rro;
        TRY(cg_init_driver(rx, ry, rz, &rro));
        int tt = 0;
        if (resident_cg() && s.max_iters > 0) {
            bool done;
            TRY(cg_resident(0, 0, &rro, error, &tt, &done));
            if (!done) tt = s.max_iters;
        } else {
            for (tt = 0; tt < s.max_iters; ++tt) {
                TRY(cg_main_step(tt, &rro, error));
                TRY(halo_update(1));
                bool done;
                TRY(stop(error, &done));
                if (done) break;
            }
        }
        TRY(flush_p());
        rep->printed_cg_iters = tt;
        return 0;
    }

    bool switch_now(int started, int tt, double error) const
    {
        if (started) return true;
        if (s.error_switch) return error < s.eps_lim && tt > 20;   // CG_ITERS_FOR_EIGENVALUES
        return tt > s.presteps && error < 1.0;                     // ERROR_SWITCH_MAX
    }

    // eigenvalue_driver.c:11-67: Lanczos tridiagonal of the CG run so far.
    int eigenvalues(int ncg)
    {
        std::vector<double> d(ncg, 0.0), e(ncg, 0.0);
        for (int i = 0; i < ncg; ++i) {
            d[i] = 1.0 / cg_alphas[i];
            if (i > 0) d[i] += cg_betas[i - 1] / cg_alphas[i - 1];
            if (i < ncg - 1) e[i + 1] = std::sqrt(cg_betas[i]) / cg_alphas[i];
        }
        TRY(ql_eigenvalues(d, e));
        double lo = DBL_MAX, hi = DBL_MIN;
        for (double v : d) {
            lo = v < lo ? v : lo;
            hi = v > hi ? v : hi;
        }
        if (lo < 0.0 || hi < 0.0) return tl_fail(TLHIP_ERR_SOLVER, "Calculated negative eigenvalues.");
        eigmin = lo * 0.95;
        eigmax = hi * 1.05;
        rep->eigmin = eigmin;
        rep->eigmax = eigmax;
        return 0;
    }

    // cheby_driver.c:163-183
    void cheby_coefs(int count)
    {
        theta = (eigmax + eigmin) / 2.0;
        const double delta = (eigmax - eigmin) / 2.0;
        const double sigma = theta / delta;
        double rho_old = 1.0 / sigma;
        for (int i = 0; i < count && i < (int)cheby_alphas.size(); ++i) {
            const double rho_new = 1.0 / (2.0 * sigma - rho_old);
            cheby_alphas[i] = rho_new * rho_old;
            cheby_betas[i] = 2.0 * rho_new / delta;
            rho_old = rho_new;
        }
    }

    // cheby_driver.c:146-160, single-precision logs and rounding as there.
    int est_iterations(double error, double bb) const
    {
        const double cn = eigmax / eigmin;
        const double it_alpha = DBL_EPSILON * bb / (4.0 * error);
        const double gamm = (std::sqrt(cn) - 1.0) / (std::sqrt(cn) + 1.0);
        const float a = logf((float)it_alpha), b = logf((float)gamm);
        return (int)roundf((float)((double)a / (2.0 * (double)b)));
    }

    // -------------------------------------------------------- Chebyshev
    int cheby_step(int k, bool calc_norm, double* error)
    {
        double* uo = u_alt;
        if (dims == 2)
            TRY(tlhip2d_cheby_iterate(x, y, hd, cheby_alphas[k], cheby_betas[k], u, uo, u0, p, r,
                                      nullptr, kx, ky));
        else
            TRY(tlhip3d_cheby_solver_iterate(x, y, z, hd, cheby_alphas[k], cheby_betas[k], u, uo,
                                             u0, p, r, nullptr, kx, ky, kz));
        u_alt = u;
        u = uo;
        rep->n_cheby_steps++;
        if (calc_norm) {
            *error = 0.0;
            TRY(norm2(r, error));
        }
        return 0;
    }

    int cheby_driver(double rx, double ry, double rz, double* error)
    {
        double rro;
        TRY(cg_init_driver(rx, ry, rz, &rro));
        int est = 0, ncheby = 0, tt;
        for (tt = 0; tt < s.max_iters; ++tt) {
            if (!switch_now(ncheby, tt, *error) && resident_cg()) {
                // the CG phase up to the switch; the loop goes on at tt + 1
                bool done;
                TRY(cg_resident(tt, s.error_switch ? 1 : 2, &rro, error, &tt, &done));
                if (done) break;
                continue;
            }
            if (!switch_now(ncheby, tt, *error)) {
                TRY(cg_main_step(tt, &rro, error));
            } else if (++ncheby == 1) {
                TRY(flush_p());
                double bb = 0.0;
                TRY(eigenvalues(tt));
                cheby_coefs(s.max_iters - tt);
                TRY(norm2(u0, &bb));
                if (dims == 2)
                    TRY(tlhip2d_cheby_init(x, y, hd, theta, u, u0, p, r, w, kx, ky));
                else
                    TRY(tlhip3d_cheby_solver_init(x, y, z, hd, theta, u, u0, p, r, w, kx, ky, kz));
                reset_fields();
                fields[F_U] = true;
                TRY(halo_update(1));
                TRY(cheby_step(ncheby, true, error));
                est = est_iterations(*error, bb);
            } else if (fused_cheby()) {
                // Iterations tt, tt+1, ... in one pass, up to and including
                // the next one that computes the norm (the only ones after
                // which the stop test can change its answer), or max_iters;
                // the depth-1 halo updates of u in between are the kernel's
                // mirror, the last one is the loop's own below.
                auto norm_at = [&](int t, int k) { return k >= est && (t + 1) % 10 == 0; };
                int cnt = 1;
                while (!norm_at(tt + cnt - 1, ncheby + cnt - 1) && tt + cnt < s.max_iters &&
                       cnt < g_cheby_fusion)
                    ++cnt;
                TRY(tlhip2d_cheby_iterations(x, y, hd, cnt, &cheby_alphas[ncheby],
                                             &cheby_betas[ncheby], neighbours, u, u_alt, u0, p,
                                             p_alt, r, kx, ky));
                std::swap(u, u_alt);
                std::swap(p, p_alt);
                rep->n_cheby_steps += cnt;
                const bool calc = norm_at(tt + cnt - 1, ncheby + cnt - 1);
                ncheby += cnt - 1;
                tt += cnt - 1;
                if (calc) {
                    *error = 0.0;
                    TRY(norm2(r, error));
                }
            } else {
                TRY(cheby_step(ncheby, ncheby >= est && (tt + 1) % 10 == 0, error));
            }
            TRY(halo_update(1));
            bool done;
            TRY(stop(error, &done));
            if (done) break;
        }
        TRY(flush_p());
        rep->printed_cg_iters = tt - ncheby + 1;
        rep->printed_inner_iters = ncheby;
        rep->est_iters = est;
        return 0;
    }

    // Several Chebyshev iterations per pass (2D).  With the relative stop
    // the reference-order loop re-tests the true residual every iteration
    // after a refused candidate, so batching is only exact for the
    // reference's own test.
    bool fused_cheby() const
    {
        return dims == 2 && g_cheby_fusion > 1 && x % 2 == 0 && p_alt && u_alt && rel_tol <= 0.0;
    }

    // -------------------------------------------------------- PPCG
    // The temporally blocked inner loop needs the 16-byte row layout and
    // all faces external (always true for this single-chunk driver).
    // The 2-norm of r formed in the last fused launch of a PPCG step (2D
    // FP64 ring kernel; tlhip_set_ppcg_fused_norm)
    bool fused_norm() const { return g_ppcg_fused_norm != 0 && tl::loop2d::ring_kernel(); }

    bool fused_ppcg() const
    {
        return dims == 2 && g_ppcg_fusion > 1 && hd >= 2 && hd % 2 == 0 && x % 2 == 0 && r_alt;
    }

    int ppcg_main_step(double* rro, double* error)
    {
        double pw = 0.0;
        TRY(calc_w_pending(&pw));
        const double alpha = *rro / pw;
        double rrn = 0.0;
        bool norm_done = false;    // the fused launch formed rrn

        if (fused_ppcg() && s.ppcg_inner_steps > 0) {
            // calc_ur, ppcg_init and the inner loop, several steps per pass;
            // the depth-1 reflections of sd happen inside the kernel.  (The
            // rrn of calc_ur is not used by PPCG: ppcg_driver.c:121-130.)
            int done = 0;
            while (done < s.ppcg_inner_steps) {
                const int cnt = std::min(g_ppcg_fusion, s.ppcg_inner_steps - done);
                const bool first = done == 0, last = done + cnt == s.ppcg_inner_steps;
                TRY(tl::loop2d::ppcg_inner_norm(x, y, hd, cnt, &cheby_alphas[done],
                                                &cheby_betas[done], theta, first, neighbours,
                                                first ? nullptr : sd, sd_alt, r, r_alt, u, u_alt,
                                                kx, ky, alpha, first ? p : nullptr,
                                                first ? w : nullptr,
                                                last && fused_norm() ? &rrn : nullptr));
                std::swap(sd, sd_alt);
                std::swap(r, r_alt);
                std::swap(u, u_alt);
                done += cnt;
                rep->n_ppcg_inner += cnt;
            }
            norm_done = fused_norm();
        } else if (fused_ppcg()) {
            TRY(calc_ur(alpha, &rrn));
            TRY(tlhip2d_ppcg_init(x, y, hd, theta, r, sd));
        } else {
            TRY(calc_ur(alpha, &rrn));
            if (dims == 2) TRY(tlhip2d_ppcg_init(x, y, hd, theta, r, sd));
            else TRY(tlhip3d_ppcg_init(x, y, z, hd, theta, r, sd));
            reset_fields();
            fields[F_SD] = true;
            if (dims == 3) {
                TRY(inner_steps3());
                rep->n_ppcg_inner += s.ppcg_inner_steps;
            }
            for (int pp = 0; dims == 2 && pp < s.ppcg_inner_steps; ++pp) {
                TRY(halo_update(1));
                double* so = sd_alt;
                TRY(tlhip2d_ppcg_inner_iteration(x, y, hd, cheby_alphas[pp], cheby_betas[pp], u, r,
                                                 kx, ky, sd, so));
                sd_alt = sd;
                sd = so;
                rep->n_ppcg_inner++;
            }
        }
        reset_fields();
        fields[F_P] = true;

        if (!norm_done) {
            rrn = 0.0;
            TRY(norm2(r, &rrn));
        }
        const double beta = rrn / *rro;
        TRY(calc_p_deferred(beta));
        *error = rrn;
        *rro = rrn;
        rep->n_ppcg_outer++;
        return 0;
    }

    int ppcg_driver(double rx, double ry, double rz, double* error)
    {
        double rro;
        TRY(cg_init_driver(rx, ry, rz, &rro));
        int nppcg = 0, tt;
        for (tt = 0; tt < s.max_iters; ++tt) {
            if (!switch_now(nppcg, tt, *error) && resident_cg()) {
                bool done;
                TRY(cg_resident(tt, s.error_switch ? 1 : 2, &rro, error, &tt, &done));
                if (done) break;
                continue;
            }
            if (!switch_now(nppcg, tt, *error)) {
                TRY(cg_main_step(tt, &rro, error));
            } else {
                if (++nppcg == 1) {
                    TRY(flush_p());             // residual() overwrites r, which calc_p reads
                    TRY(eigenvalues(tt));
                    cheby_coefs(s.ppcg_inner_steps);
                    TRY(residual());
                    reset_fields();
                    fields[F_P] = true;
                    TRY(halo_update(1));
                }
                if (resident_ppcg()) {
                    // PPCG iterations tt .. last; the loop goes on at last + 1
                    int last;
                    tl::loop::Report rp;
                    TRY(run_resident(tt, 0, false,
                                     [&](const tl::loop::StopRule& rule, int) {
                                         return ppcg_body(rule);
                                     },
                                     &rro, error, &last, &rp, true, dims == 2 && can_defer_p()));
                    const int ran = last - tt + 1;
                    rep->n_ppcg_outer += ran;
                    rep->n_ppcg_inner += ran * s.ppcg_inner_steps;
                    nppcg += ran - 1;
                    tt = last;
                    bool done;
                    TRY(confirm(rp, &done));
                    if (done) break;
                    continue;
                }
                TRY(ppcg_main_step(&rro, error));
            }
            TRY(halo_update(1));
            bool done;
            TRY(stop(error, &done));
            if (done) break;
        }
        TRY(flush_p());
        rep->printed_cg_iters = tt - nppcg + 1;
        rep->printed_inner_iters = nppcg;
        return 0;
    }

    // -------------------------------------------------------- mixed precision
    // CG / PPCG with FP32 sweeps (kx, ky, p, w, r, sd and d, the correction
    // to u since the last FP64 update) and FP64-accumulated dots, inside an
    // FP64 residual-replacement loop ("reliable updates", as in Sleijpen &
    // van der Vorst, Computing 56 (1996)): whenever the FP32 recurrence
    // residual has fallen kDrop below the last true one, or claims
    // convergence, u += d in FP64 and r is replaced by u0 - A u computed in
    // FP64.  The CG/PPCG control flow, the switch rule, the Lanczos
    // eigenvalue estimate and the Chebyshev coefficients are the reference's
    // (cg_driver.c, ppcg_driver.c), in FP64 on the host.  Convergence is
    // only ever declared on a true FP64 residual.
    static constexpr double kDrop = 0.1;

    // FP32 rows of the chunk (3D: planes of rows)
    long rows32() const { return dims == 3 ? (long)y * z : (long)y; }
    int zz() const { return dims == 3 ? z : 1; }

    int mixed_alloc()
    {
        if (kx32) return 0;
        P = (x + 3) / 4 * 4;
        const size_t bytes = sizeof(float) * (size_t)P * (size_t)rows32();
        float** all[] = {&kx32, &ky32, &p32, &w32, &r32, &d32, &sd32, &sd32b, &r32b, &d32b, &p32b};
        for (float** f : all) TRY(tlhip_alloc(bytes, (void**)f));
        if (dims == 3) TRY(tlhip_alloc(bytes, (void**)&kz32));
        TRY(tlhip_alloc(sizeof(double), (void**)&alpha32));
        return 0;
    }

    // FP32 calc_w, with a deferred calc_p fused in (as calc_w_pending; 3D
    // applies it first)
    int calc_w32_pending(double* pw)
    {
        if (dims == 3) {
            TRY(flush_p32());
            return tl::mix::calc_w3_32(x, y, z, hd, P, pw, p32, w32, kx32, ky32, kz32);
        }
        if (!pending_p) return tl::mix::calc_w32(x, y, hd, P, pw, p32, w32, kx32, ky32);
        pending_p = false;
        TRY(tl::mix::calc_pw32(x, y, hd, P, pending_beta, pw, p32, p32b, r32, w32, kx32, ky32));
        std::swap(p32, p32b);
        return 0;
    }

    // The same, with alpha = rro / pw formed on the device into alpha32 and
    // no wait for pw: a PPCG step needs alpha only inside its first fused
    // launch.
    int calc_w32_pending_alpha(double rro)
    {
        if (!pending_p)
            return tl::mix::calc_pw32_alpha(x, y, hd, P, false, 0.0, rro, alpha32, p32, nullptr,
                                            nullptr, w32, kx32, ky32);
        pending_p = false;
        TRY(tl::mix::calc_pw32_alpha(x, y, hd, P, true, pending_beta, rro, alpha32, p32, p32b,
                                     r32, w32, kx32, ky32));
        std::swap(p32, p32b);
        return 0;
    }

    int flush_p32()
    {
        if (!pending_p) return 0;
        pending_p = false;
        return tl::mix::calc_p32(x, y, zz(), hd, P, pending_beta, false, p32, r32);
    }

    // u += d (FP64); r32 = u0 - A u computed in FP64; *rr = its squared norm.
    // A deferred calc_p is applied first, with the recurrence residual, as
    // the step that deferred it would have.
    int replace_residual(double* rr)
    {
        TRY(flush_p32());
        TRY(tl::mix::promote(x, y, zz(), hd, P, u, d32));
        *rr = 0.0;
        if (dims == 3)
            TRY(tl::mix::residual3_64(x, y, z, hd, P, u, u0, kx, ky, kz, r32, rr));
        else
            TRY(tl::mix::residual64(x, y, hd, P, u, u0, kx, ky, r32, rr));
        rep->n_true_residuals++;
        rep->rel_residual = bb > 0.0 ? std::sqrt(*rr / bb) : 0.0;
        return 0;
    }

    int mixed_cg_step(int tt, double* rro, double* rrn)
    {
        double pw = 0.0;
        TRY(calc_w32_pending(&pw));
        const double alpha = *rro / pw;
        cg_alphas[tt] = alpha;
        TRY(tl::mix::calc_ur32(x, y, zz(), hd, P, alpha, rrn, d32, p32, r32, w32));
        const double beta = *rrn / *rro;
        cg_betas[tt] = beta;
        pending_p = true;
        pending_beta = beta;
        *rro = *rrn;
        rep->n_cg_steps++;
        return 0;
    }

    // 3D: calc_w, calc_ur, ppcg_init and one launch per inner step
    // (ppcg_solver_driver.c), then the 2-norm of r.
    int mixed_ppcg_step3(double* rro, double* rrn)
    {
        double pw = 0.0, unused = 0.0;
        TRY(calc_w32_pending(&pw));
        TRY(tl::mix::calc_ur32(x, y, z, hd, P, *rro / pw, &unused, d32, p32, r32, w32));
        if (s.ppcg_inner_steps > 0) TRY(tl::mix::ppcg_init3_32(x, y, z, hd, P, theta, r32, sd32));
        int k = 0;
        if (g_ppcg_fusion > 1 && x % 2 == 0) {
            // two steps per pass, bit-identical to two single steps
            for (; k + 2 <= s.ppcg_inner_steps; k += 2) {
                TRY(tl::mix::ppcg_pair3_32(x, y, z, hd, P, &cheby_alphas[k], &cheby_betas[k],
                                           sd32, sd32b, r32, r32b, d32, d32b, kx32, ky32,
                                           kz32));
                std::swap(sd32, sd32b);
                std::swap(r32, r32b);
                std::swap(d32, d32b);
                rep->n_ppcg_inner += 2;
            }
        }
        for (; k < s.ppcg_inner_steps; ++k) {
            TRY(tl::mix::ppcg_inner3_32(x, y, z, hd, P, cheby_alphas[k], cheby_betas[k], sd32,
                                        sd32b, r32, d32, kx32, ky32, kz32));
            std::swap(sd32, sd32b);
            rep->n_ppcg_inner++;
        }
        *rrn = 0.0;
        TRY(tl::mix::norm32(x, y, z, hd, P, r32, rrn));
        pending_p = true;
        pending_beta = *rrn / *rro;
        *rro = *rrn;
        rep->n_ppcg_outer++;
        return 0;
    }

    int mixed_ppcg_step(double* rro, double* rrn)
    {
        if (dims == 3) return mixed_ppcg_step3(rro, rrn);
        if (s.ppcg_inner_steps == 0) {
            double pw = 0.0, unused = 0.0;
            TRY(calc_w32_pending(&pw));
            TRY(tl::mix::calc_ur32(x, y, 1, hd, P, *rro / pw, &unused, d32, p32, r32, w32));
        } else {
            TRY(calc_w32_pending_alpha(*rro));    // alpha stays on the device
        }
        const int per = std::max(1, g_ppcg_fusion);
        int done = 0;
        *rrn = 0.0;
        while (done < s.ppcg_inner_steps) {
            // the first launch also applies calc_ur (u is d here); the last
            // also forms rrn = r.r
            const int cnt = std::min(per, s.ppcg_inner_steps - done);
            const bool first = done == 0, last = done + cnt == s.ppcg_inner_steps;
            TRY(tl::mix::ppcg_inner32_norm(x, y, hd, P, cnt, &cheby_alphas[done],
                                           &cheby_betas[done], theta, first, neighbours,
                                           first ? nullptr : sd32, sd32b, r32, r32b, d32, d32b,
                                           kx32, ky32, 0.0, first ? p32 : nullptr,
                                           first ? w32 : nullptr, first ? alpha32 : nullptr,
                                           last ? rrn : nullptr));
            std::swap(sd32, sd32b);
            std::swap(r32, r32b);
            std::swap(d32, d32b);
            done += cnt;
            rep->n_ppcg_inner += cnt;
        }
        if (done == 0) TRY(tl::mix::norm32(x, y, 1, hd, P, r32, rrn));
        const double beta = *rrn / *rro;
        pending_p = true;
        pending_beta = beta;
        *rro = *rrn;
        rep->n_ppcg_outer++;
        return 0;
    }

    // One Chebyshev launch of cnt turns on A d = b (b = r32), r in r32b;
    // *rrn = the last turn's r.r.  2D: one fused launch; 3D: one per turn.
    int mixed_cheby_turns(int first, int cnt, double* rrn)
    {
        if (dims == 2) {
            TRY(tl::mix::cheby_iterations32(x, y, hd, P, cnt, &cheby_alphas[first],
                                            &cheby_betas[first], neighbours, d32, d32b, r32, p32,
                                            p32b, r32b, kx32, ky32, rrn));
            std::swap(d32, d32b);
            std::swap(p32, p32b);
            return 0;
        }
        for (int k = 0; k < cnt; ++k) {
            TRY(tl::mix::cheby_iter3_32(x, y, z, hd, P, cheby_alphas[first + k],
                                        cheby_betas[first + k], d32, d32b, r32, p32, r32b, kx32,
                                        ky32, kz32, k == cnt - 1 ? rrn : nullptr));
            std::swap(d32, d32b);
        }
        return 0;
    }

    int mixed_driver(double rx, double ry, double rz, double* error)
    {
        if (s.solver != 1 && s.solver != 2 && s.solver != 3)
            return tl_fail(TLHIP_ERR_ARG, "mixed precision: CG, Chebyshev or PPCG only (solver %d)",
                           s.solver);
        TRY(mixed_alloc());
        double rro;
        TRY(cg_init_driver(rx, ry, rz, &rro));    // FP64 kx, ky(, kz), u, r = u0 - A u, p = r
        const int ny = (int)rows32();             // demote treats planes as rows
        TRY(tl::mix::demote(kx, x, ny, kx32, P));
        TRY(tl::mix::demote(ky, x, ny, ky32, P));
        if (dims == 3) TRY(tl::mix::demote(kz, x, ny, kz32, P));
        TRY(tl::mix::demote(r, x, ny, r32, P));
        pending_p = false;
        TRY(tl::mix::calc_p32(x, y, zz(), hd, P, 0.0, true, p32, r32));
        TRY(tl::mix::zero32(d32, (long)P * rows32()));
        const double target = rel_tol > 0.0 ? rel_tol * rel_tol * bb : s.eps;
        double rr_true = rro, last_true = rro;
        *error = rro;
        int nppcg = 0, est = 0, tt;      // nppcg: PPCG outer or Chebyshev iterations
        for (tt = 0; tt < s.max_iters; ++tt) {
            double rrn = 0.0;
            if (s.solver == 2 && switch_now(nppcg, tt, *error)) {
                // Chebyshev on the correction equation A d = b, b = r32 (the
                // last true residual), r of each turn in r32b
                int cnt = 1;
                if (++nppcg == 1) {
                    TRY(flush_p32());
                    TRY(eigenvalues(tt));
                    cheby_coefs(s.max_iters - tt);
                    TRY(replace_residual(&rr_true));          // d = 0, b = u0 - A u
                    last_true = rr_true;
                    TRY(tl::mix::cheby_init32(x, y, zz(), hd, P, theta, p32, r32, d32));
                } else {
                    cnt = std::min({g_cheby_fusion > 1 ? g_cheby_fusion : 1, s.max_iters - tt,
                                    (int)cheby_alphas.size() - nppcg});
                }
                TRY(mixed_cheby_turns(nppcg, cnt, &rrn));
                rep->n_cheby_steps += cnt;
                if (nppcg == 1) est = est_iterations(rrn, bb);
                nppcg += cnt - 1;
                tt += cnt - 1;
                *error = rrn;
                if (rrn <= target || rrn < kDrop * kDrop * last_true) {
                    TRY(replace_residual(&rr_true));
                    last_true = rr_true;
                    *error = rr_true;
                    if (rr_true <= target) break;
                }
                continue;
            }
            if (s.solver == 3 && switch_now(nppcg, tt, *error)) {
                if (++nppcg == 1) {
                    TRY(eigenvalues(tt));
                    cheby_coefs(s.ppcg_inner_steps);
                    // ppcg_init_driver recomputes the residual (ppcg_driver.c:72-77)
                    TRY(replace_residual(&rr_true));
                    last_true = rr_true;
                }
                TRY(mixed_ppcg_step(&rro, &rrn));
            } else {
                TRY(mixed_cg_step(tt, &rro, &rrn));
            }
            *error = rrn;
            if (rrn <= target || rrn < kDrop * kDrop * last_true) {
                TRY(replace_residual(&rr_true));
                last_true = rr_true;
                rro = rr_true;        // continue the recurrence from the true residual
                *error = rr_true;
                if (rr_true <= target) break;
            }
        }
        if (tt == s.max_iters) TRY(replace_residual(&rr_true));   // fold the last correction
        TRY(flush_p32());
        if (s.solver == 1) {
            rep->printed_cg_iters = tt;                  // as cg_driver.c prints
        } else {
            rep->printed_cg_iters = tt - nppcg + 1;      // as ppcg_driver.c / cheby_driver.c print
            rep->printed_inner_iters = nppcg;
            if (s.solver == 2) rep->est_iters = est;
        }
        return 0;
    }

    // -------------------------------------------------------- Jacobi
    int jacobi_driver(double rx, double ry, double rz, double* error)
    {
        if (dims == 2)
            TRY(tlhip2d_jacobi_init(x, y, hd, s.coefficient, rx, ry, density, energy, u0, u, kx, ky));
        else
            TRY(tlhip3d_jacobi_solver_init(x, y, z, hd, s.coefficient, rx, ry, rz, density, energy,
                                           u0, u, kx, ky, kz));
        TRY(copy_u());
        TRY(start_solve());
        reset_fields();
        fields[F_U] = true;
        if (g_resident && interior()) return jacobi_resident(error);
        int tt;
        for (tt = 0; tt < s.max_iters; ++tt) {
            if (dims == 2) TRY(tlhip2d_jacobi_iterate(x, y, hd, error, kx, ky, u0, u, r));
            else TRY(tlhip3d_jacobi_solver_iterate(x, y, z, hd, error, kx, ky, kz, u0, u, r));
            rep->n_jacobi++;
            if (tt % 50 == 0) {
                TRY(halo_update(1));
                TRY(residual());
                TRY(norm2(r, error));
            }
            TRY(halo_update(1));
            bool done;
            TRY(stop(error, &done));
            if (done) break;
        }
        rep->printed_cg_iters = tt;
        return 0;
    }

    // The loop of jacobi_driver with jacobi_body's iterations enqueued ahead
    // of the host.  u_alt starts as a copy of u: sweeps write the interior
    // and halo updates depth 1 only, so the deeper halo of both buffers stays
    // the same and u ends as the host-driven loop leaves it.  r is then made
    // what the host-driven loop leaves: the last sweep's input, with the
    // residual in its interior when that iteration computed one.
    int jacobi_resident(double* error)
    {
        const size_t bytes = sizeof(double) * (size_t)n;
        TRY(tlhip_copy_device(u_alt, u, bytes));
        int last = 0;
        bool done = false;
        for (int from = 0;;) {
            double rro = 0.0;
            tl::loop::Report rp;
            TRY(run_resident(from, 0, false,
                             [&](const tl::loop::StopRule& rule, int j) {
                                 return jacobi_body(rule, j);
                             },
                             &rro, error, &last, &rp, false));
            rep->n_jacobi += last - from + 1;
            TRY(confirm(rp, &done));
            if (done || last + 1 >= s.max_iters) break;
            from = last + 1;    // a relative-residual candidate the true residual refused
        }
        TRY(tlhip_copy_device(r, u_alt, bytes));
        if (last % 50 == 0) TRY(residual());
        rep->printed_cg_iters = done ? last : s.max_iters;
        return 0;
    }

    // -------------------------------------------------------- adjoint
    // CG on A x = b from x = b (A is symmetric, so the adjoint of a solve is
    // a solve), leaving x in u.  Uses the solver scratch fields u, u0, r, p,
    // w and the forward solves' stopping test.
    int adjoint_solve(const double* b, tlhip_step_report* report)
    {
        rep = report;
        const size_t bytes = sizeof(double) * (size_t)n;
        TRY(tlhip_copy_device(u0, b, bytes));
        TRY(tlhip_copy_device(u, b, bytes));
        reset_fields();
        fields[F_U] = true;
        TRY(halo_update(1));
        TRY(residual());
        TRY(tlhip_copy_device(p, r, bytes));
        reset_fields();
        fields[F_U] = fields[F_P] = true;
        TRY(halo_update(1));
        double rro = 0.0;
        TRY(norm2(r, &rro));
        TRY(start_solve());
        double error = rro;
        int tt;
        for (tt = 0; tt < s.max_iters; ++tt) {
            TRY(cg_main_step(tt, &rro, &error));
            TRY(halo_update(1));
            bool done;
            TRY(stop(&error, &done));
            if (done) break;
        }
        TRY(flush_p());
        report->printed_cg_iters = tt;
        report->error = error;
        return 0;
    }

    // Gradient of L = sum_interior (u_N - target)^2 with respect to ln(rho)
    // per cell, u_N being u after `steps` timesteps from the current state.
    // The forward steps are ordinary steps (the state advances) and keep u
    // after each solve; the backward sweep solves A mu = lambda per step and
    // scatters the face terms onto the cells (tl_adjoint.hip).
    int adjoint(int steps, const double* target_h, double* loss, double* grad_h,
                tlhip_step_report* reports)
    {
        if (steps < 1) return tl_fail(TLHIP_ERR_ARG, "adjoint: steps must be >= 1");
        if (precision != 0) return tl_fail(TLHIP_ERR_ARG, "adjoint: needs the FP64 solve");
        const size_t bytes = sizeof(double) * (size_t)n;
        int rc = 0;
        std::vector<double*> ck(steps, nullptr);
        for (auto& c : ck) c = dalloc(n, &rc);
        double* target = dalloc(n, &rc);
        double* lambda = dalloc(n, &rc);
        double* grad = dalloc(n, &rc);
        double* b0 = dalloc(n, &rc);
        auto body = [&]() -> int {
            TRY(rc);
            TRY(tlhip_copy_to_device(target, target_h, bytes));
            for (int t = 0; t < steps; ++t) {
                tlhip_step_report tmp;
                TRY(step(t, reports ? &reports[t] : &tmp));
                if (t == 0) TRY(tlhip_copy_device(b0, u0, bytes));   // e_0 rho
                TRY(tlhip_copy_device(ck[t], u, bytes));            // u_{t+1}
            }
            const TlGrid3 g{x, y, z, hd};
            TRY(tl::adj::loss_seed(g, ck[steps - 1], target, lambda, loss));
            TRY(tl::adj::zero(grad, n));
            const double dt = s.dt_init;
            const double rx = dt / (dx * dx), ry = dt / (dy * dy);
            const double rz = dims == 3 ? dt / (dz * dz) : 0.0;
            for (int k = steps - 1; k >= 0; --k) {
                tlhip_step_report tmp{};
                TRY(adjoint_solve(lambda, &tmp));                   // mu_k in u
                TRY(tl::adj::face_grad(g, s.coefficient, rx, ry, rz, u, ck[k], density, grad));
                if (k == 0) TRY(tl::adj::rhs_grad(g, u, b0, grad));
                else TRY(tlhip_copy_device(lambda, u, bytes));
            }
            return tlhip_copy_to_host(grad_h, grad, bytes);
        };
        const int out = body();
        for (double* c : ck)
            if (c) (void)tlhip_free(c);
        for (double* a : {target, lambda, grad, b0})
            if (a) (void)tlhip_free(a);
        return out;
    }

    // -------------------------------------------------------- lifecycle
    // initialise.c:24-30: prime halos of density/energy0/energy, store energy.
    int prime()
    {
        reset_fields();
        fields[F_DENSITY] = fields[F_ENERGY0] = fields[F_ENERGY1] = true;
        TRY(halo_update(2));
        return store_energy();
    }

    int step(int tt, tlhip_step_report* report)
    {
        std::memset(report, 0, sizeof(*report));
        rep = report;
        pending_p = false;
        const double dt = s.dt_init;
        const double rx = dt / (dx * dx), ry = dt / (dy * dy);
        const double rz = dims == 3 ? dt / (dz * dz) : 0.0;
        reset_fields();
        fields[F_ENERGY1] = fields[F_DENSITY] = true;
        TRY(halo_update(2));
        double error = 1e+10;
        if (precision == 1) {
            TRY(mixed_driver(rx, ry, rz, &error));
        } else switch (s.solver) {
        case 0: TRY(jacobi_driver(rx, ry, rz, &error)); break;
        case 1: TRY(cg_driver(rx, ry, rz, &error)); break;
        case 2: TRY(cheby_driver(rx, ry, rz, &error)); break;
        case 3: TRY(ppcg_driver(rx, ry, rz, &error)); break;
        default: return tl_fail(TLHIP_ERR_ARG, "unknown solver %d", s.solver);
        }
        report->error = error;
        // solve_finished_driver.c:7-43
        double exact = 0.0;
        if (s.check_result) {
            TRY(residual());
            TRY(norm2(r, &exact));
        }
        report->exact_error = exact;
        if (dims == 2) TRY(tlhip2d_finalise(x, y, hd, energy, density, u));
        else TRY(tlhip3d_solver_finalise(x, y, z, hd, energy, density, u));
        fields[F_ENERGY1] = true;
        TRY(halo_update(1));
        if (s.summary_frequency > 0 && tt % s.summary_frequency == 0) {
            double o[4];
            TRY(field_summary(o));
            report->summary_done = 1;
            report->vol = o[0];
            report->mass = o[1];
            report->ie = o[2];
            report->temp = o[3];
        }
        return 0;
    }

    double* field(int id)
    {
        double* f[] = {density, energy0, energy, u, u0, p, r, w, kx, ky, kz, sd};
        return (id >= 0 && id < 12) ? f[id] : nullptr;
    }

    // An adopted chunk's u, sd, r and p back at the caller's pointers: a
    // field whose partner holds it is copied home and the pair swaps back.
    int settle()
    {
        if (!adopted) return 0;
        double** pairs[4][2] = {{&u, &u_alt}, {&sd, &sd_alt}, {&r, &r_alt}, {&p, &p_alt}};
        for (int i = 0; i < 4; ++i) {
            double*& a = *pairs[i][0];
            double*& b = *pairs[i][1];
            if (a == home[i]) continue;
            if (b != home[i])
                return tl_fail(TLHIP_ERR_ARG, "tlhip_problem_step: a field left its buffer pair");
            TRY(tlhip_copy_device(home[i], a, sizeof(double) * (size_t)n));
            std::swap(a, b);
        }
        return 0;
    }

    void release()
    {
        float* all32[] = {kx32, ky32, p32, w32, r32, d32, sd32, sd32b, r32b, d32b, p32b};
        for (float* a : all32)
            if (a) (void)tlhip_free(a);
        if (kz32) (void)tlhip_free(kz32);
        if (alpha32) (void)tlhip_free(alpha32);
        if (cond) (void)tlhip_free(cond);
        if (adopted) {      // the caller's fields stay; only the partners of r and p are ours
            for (double* a : {r == home[2] ? r_alt : r, p == home[3] ? p_alt : p})
                if (a) (void)tlhip_free(a);
            return;
        }
        double* all[] = {density0, density, energy0, energy, u, u0, p, r, mi, w, kx, ky, kz,
                         sd, u_alt, sd_alt, r_alt, p_alt, volume, x_area, y_area, z_area, cell_x,
                         cell_y, cell_z, vertex_x, vertex_y, vertex_z};
        for (double* a : all)
            if (a) (void)tlhip_free(a);
    }
};

extern "C" int tlhip_set_resident(int on)
{
    g_resident = on ? 1 : 0;
    return 0;
}

extern "C" int tlhip_set_graphs(int on)
{
    g_graphs = on ? 1 : 0;
    return 0;
}

extern "C" int tlhip_set_cheby_fusion(int steps)
{
    if (steps < 0 || steps > 10) return tl_fail(TLHIP_ERR_ARG, "tlhip_set_cheby_fusion: 0..10");
    g_cheby_fusion = steps;
    return 0;
}

extern "C" int tlhip_set_cg_fusion(int on)
{
    g_cg_fusion = on ? 1 : 0;
    return 0;
}

extern "C" int tlhip_set_cg_cond(int on)
{
    g_cg_cond = on ? 1 : 0;
    return 0;
}

extern "C" int tlhip_set_ppcg_fusion(int steps)
{
    if (steps < 0 || steps > 10) return tl_fail(TLHIP_ERR_ARG, "tlhip_set_ppcg_fusion: 0..10");
    g_ppcg_fusion = steps;
    return 0;
}

extern "C" int tlhip_set_ppcg_fused_norm(int on)
{
    g_ppcg_fused_norm = on ? 1 : 0;
    return 0;
}

extern "C" int tlhip_problem_set_solve(tlhip_problem* P, int precision, double rel_residual)
{
    if (!P) return tl_fail(TLHIP_ERR_ARG, "tlhip_problem_set_solve: null problem");
    if (precision != 0 && precision != 1)
        return tl_fail(TLHIP_ERR_ARG, "tlhip_problem_set_solve: precision must be 0 or 1");
    if (!(rel_residual >= 0.0))
        return tl_fail(TLHIP_ERR_ARG, "tlhip_problem_set_solve: rel_residual must be >= 0");
    if (precision == 1) {
        if (P->s.solver != 1 && P->s.solver != 2 && P->s.solver != 3)
            return tl_fail(TLHIP_ERR_ARG,
                           "tlhip_problem_set_solve: mixed precision needs CG, Chebyshev or PPCG");
        TRY(P->mixed_alloc());
    }
    P->precision = precision;
    P->rel_tol = rel_residual;
    return 0;
}

extern "C" int tlhip_problem_adjoint(tlhip_problem* P, int steps, const double* u_target,
                                     double* loss, double* grad_log_density,
                                     tlhip_step_report* reports)
{
    if (!P || !u_target || !loss || !grad_log_density)
        return tl_fail(TLHIP_ERR_ARG, "tlhip_problem_adjoint: null argument");
    *loss = 0.0;
    TRY(P->adjoint(steps, u_target, loss, grad_log_density, reports));
    TRY(P->settle());
    return tlhip_synchronize();
}

namespace {
int check_settings(const tlhip_settings* s, const char* what)
{
    if (s->dims != 2 && s->dims != 3) return tl_fail(TLHIP_ERR_ARG, "%s: dims must be 2 or 3", what);
    if (s->x_cells < 1 || s->y_cells < 1 || (s->dims == 3 && s->z_cells < 1) || s->halo_depth < 1 ||
        s->max_iters < 1)
        return tl_fail(TLHIP_ERR_ARG, "%s: invalid sizes", what);
    return 0;
}

// Sizes as 2d/chunk.c:7-8 and parse_config.c:188-191 form them.
tlhip_problem* new_problem(const tlhip_settings* s)
{
    auto* P = new tlhip_problem();
    P->s = *s;
    P->dims = s->dims;
    P->hd = s->halo_depth;
    P->x = s->x_cells + 2 * P->hd;
    P->y = s->y_cells + 2 * P->hd;
    P->z = s->dims == 3 ? s->z_cells + 2 * P->hd : 1;
    P->n = (long)P->x * P->y * P->z;
    P->dx = (s->xmax - s->xmin) / (double)s->x_cells;
    P->dy = (s->ymax - s->ymin) / (double)s->y_cells;
    P->dz = s->dims == 3 ? (s->zmax - s->zmin) / (double)s->z_cells : 0.0;
    // The reference sizes these max_iters (2d/c_kernels/omp3/kernel_initialise.c:68)
    // while PPCG writes ppcg_inner_steps Chebyshev coefficients into them; a
    // deck with more inner steps than max_iters overruns the reference's
    // arrays.  Here they hold both.
    const size_t mi = (size_t)std::max(s->max_iters, s->ppcg_inner_steps) + 1;
    P->cg_alphas.assign(mi, 0.0);
    P->cg_betas.assign(mi, 0.0);
    P->cheby_alphas.assign(mi, 0.0);
    P->cheby_betas.assign(mi, 0.0);
    return P;
}

// Partners the driver's fused paths write into.
void alloc_partners(tlhip_problem* P, int* rc)
{
    // cg_init's conductivity for the cond CG passes (CG, Chebyshev, PPCG)
    const bool cg = P->hd >= 2 && P->s.solver >= 1 && P->s.solver <= 3;
    if (cg) P->cond = dalloc(P->n, rc);
    // partner of p: the fused calc_p + calc_w (2D; 3D with cond) and the
    // 2D fused Chebyshev loop
    if (P->dims == 2 || cg) P->p_alt = dalloc(P->n, rc);
    // partner of r: the fused PPCG inner loops (2D ring, 3D two-step launch)
    if (P->s.solver == 3) P->r_alt = dalloc(P->n, rc);
}
}  // namespace

extern "C" int tlhip_problem_create(const tlhip_settings* s, const tlhip_state* states,
                                    int num_states, tlhip_problem** out)
{
    if (!s || !states || !out || num_states < 1)
        return tl_fail(TLHIP_ERR_ARG, "tlhip_problem_create: invalid arguments");
    TRY(check_settings(s, "tlhip_problem_create"));
    *out = nullptr;
    TRY(tlhip_init(-1));
    auto* P = new_problem(s);
    int rc = 0;
    const long n = P->n;
    double** big[] = {&P->density0, &P->density, &P->energy0, &P->energy, &P->u, &P->u0, &P->p,
                      &P->r, &P->mi, &P->w, &P->kx, &P->ky, &P->sd, &P->u_alt, &P->sd_alt,
                      &P->volume, &P->x_area, &P->y_area};
    for (double** b : big) *b = dalloc(n, &rc);
    if (P->dims == 3) {
        P->kz = dalloc(n, &rc);
        P->z_area = dalloc(n, &rc);
    }
    alloc_partners(P, &rc);
    P->cell_x = dalloc(P->x + 1, &rc);
    P->cell_y = dalloc(P->y + 1, &rc);
    P->cell_z = dalloc(P->z + 1, &rc);
    P->vertex_x = dalloc(P->x + 1, &rc);
    P->vertex_y = dalloc(P->y + 1, &rc);
    P->vertex_z = dalloc(P->z + 1, &rc);
    if (!rc) {
        if (P->dims == 2) {
            rc = tlhip2d_set_chunk_data(P->x, P->y, P->hd, 0, 0, s->xmin, s->ymin, P->dx, P->dy,
                                        P->cell_x, P->cell_y, P->vertex_x, P->vertex_y, P->volume,
                                        P->x_area, P->y_area);
            if (!rc)
                rc = tlhip2d_set_chunk_state(P->x, P->y, P->vertex_x, P->vertex_y, P->cell_x,
                                             P->cell_y, P->density, P->energy0, P->u, num_states,
                                             states);
        } else {
            rc = tlhip3d_set_chunk_data(P->x, P->y, P->z, P->hd, 0, 0, 0, s->xmin, s->ymin, s->zmin,
                                        P->dx, P->dy, P->dz, P->cell_x, P->cell_y, P->cell_z,
                                        P->vertex_x, P->vertex_y, P->vertex_z, P->volume,
                                        P->x_area, P->y_area, P->z_area);
            if (!rc)
                rc = tlhip3d_set_chunk_state(P->x, P->y, P->z, P->vertex_x, P->vertex_y,
                                             P->vertex_z, P->cell_x, P->cell_y, P->cell_z,
                                             P->density, P->energy0, P->u, num_states, states);
        }
    }
    if (!rc) rc = P->prime();
    if (!rc) rc = tlhip_synchronize();
    if (rc) {
        P->release();
        delete P;
        return rc;
    }
    *out = P;
    return 0;
}

namespace {
// u = e*rho on [1, n-1) of every axis (2d/c_kernels/omp3/set_chunk_state.c:76-83),
// computed on the host for an injected state and uploaded.
void host_u(const tlhip_problem* P, const double* d, const double* e, std::vector<double>& u)
{
    u.assign((size_t)P->n, 0.0);
    const int z0 = P->dims == 3 ? 1 : 0, z1 = P->dims == 3 ? P->z - 1 : 1;
    for (int ii = z0; ii < z1; ++ii)
        for (int jj = 1; jj < P->y - 1; ++jj)
            for (int kk = 1; kk < P->x - 1; ++kk) {
                const long c = kk + (long)jj * P->x + (long)ii * P->x * P->y;
                u[(size_t)c] = e[c] * d[c];
            }
}
}  // namespace

extern "C" int tlhip_problem_inject(tlhip_problem* P, const double* density, const double* energy0)
{
    if (!P || !density || !energy0) return tl_fail(TLHIP_ERR_ARG, "tlhip_problem_inject: null argument");
    const size_t bytes = sizeof(double) * (size_t)P->n;
    TRY(tlhip_copy_to_device(P->density, density, bytes));
    TRY(tlhip_copy_to_device(P->energy0, energy0, bytes));
    std::vector<double> u;
    host_u(P, density, energy0, u);
    TRY(tlhip_copy_to_device(P->u, u.data(), bytes));
    TRY(P->prime());
    return tlhip_synchronize();
}

extern "C" int tlhip_problem_adopt(const tlhip_settings* s, const tlhip_chunk_fields* f,
                                   tlhip_problem** out)
{
    const char* what = "tlhip_problem_adopt";
    if (!s || !f || !out) return tl_fail(TLHIP_ERR_ARG, "%s: null argument", what);
    TRY(check_settings(s, what));
    const bool d3 = s->dims == 3;
    const double* need[] = {f->density0, f->density, f->energy0, f->energy, f->u, f->u0, f->p,
                            f->r, f->mi, f->w, f->kx, f->ky, f->sd, f->u_alt, f->sd_alt,
                            f->volume, f->x_area, f->y_area, d3 ? f->kz : f->kx,
                            d3 ? f->z_area : f->x_area};
    for (const double* a : need)
        if (!a) return tl_fail(TLHIP_ERR_ARG, "%s: a field of the chunk is NULL", what);
    *out = nullptr;
    TRY(tlhip_init(-1));
    auto* P = new_problem(s);
    P->adopted = true;
    P->density0 = f->density0, P->density = f->density, P->energy0 = f->energy0;
    P->energy = f->energy, P->u = f->u, P->u0 = f->u0, P->p = f->p, P->r = f->r, P->mi = f->mi;
    P->w = f->w, P->kx = f->kx, P->ky = f->ky, P->kz = d3 ? f->kz : nullptr, P->sd = f->sd;
    P->u_alt = f->u_alt, P->sd_alt = f->sd_alt, P->volume = f->volume;
    P->x_area = f->x_area, P->y_area = f->y_area, P->z_area = d3 ? f->z_area : nullptr;
    P->home[0] = P->u, P->home[1] = P->sd, P->home[2] = P->r, P->home[3] = P->p;
    int rc = 0;
    alloc_partners(P, &rc);
    if (rc) {
        P->release();
        delete P;
        return rc;
    }
    *out = P;
    return 0;
}

extern "C" int tlhip_problem_step(tlhip_problem* P, int tt, tlhip_step_report* report)
{
    if (!P || !report) return tl_fail(TLHIP_ERR_ARG, "tlhip_problem_step: null argument");
    TRY(P->step(tt, report));
    TRY(P->settle());
    return tlhip_synchronize();
}

extern "C" int tlhip_problem_field_summary(tlhip_problem* P, double* out4)
{
    if (!P || !out4) return tl_fail(TLHIP_ERR_ARG, "tlhip_problem_field_summary: null argument");
    return P->field_summary(out4);
}

extern "C" int tlhip_problem_field_len(tlhip_problem* P, long* n)
{
    if (!P || !n) return tl_fail(TLHIP_ERR_ARG, "tlhip_problem_field_len: null argument");
    *n = P->n;
    return 0;
}

extern "C" int tlhip_problem_get_field(tlhip_problem* P, int id, double* host_out)
{
    if (!P || !host_out) return tl_fail(TLHIP_ERR_ARG, "tlhip_problem_get_field: null argument");
    double* f = P->field(id);
    if (!f) return tl_fail(TLHIP_ERR_ARG, "tlhip_problem_get_field: no field %d", id);
    return tlhip_copy_to_host(host_out, f, sizeof(double) * (size_t)P->n);
}

extern "C" int tlhip_problem_field_ptr(tlhip_problem* P, int id, void** dev_out)
{
    if (!P || !dev_out) return tl_fail(TLHIP_ERR_ARG, "tlhip_problem_field_ptr: null argument");
    double* f = P->field(id);
    if (!f) return tl_fail(TLHIP_ERR_ARG, "tlhip_problem_field_ptr: no field %d", id);
    *dev_out = f;
    return 0;
}

extern "C" int tlhip_problem_destroy(tlhip_problem* P)
{
    if (!P) return 0;
    const int rc = tlhip_synchronize();
    P->release();
    delete P;
    return rc;
}
