/*
 * TeaLeaf CPU ORACLE, host solver drivers -- test infrastructure (oracle.h).
 *
 * Restates the single-chunk, NO_MPI control flow of 2d/diffuse.c and
 * the 2d/drivers/ sources (3D: 3d/diffuse.c and 3d/drivers/) over the oracle kernels,
 * including the driver state that changes results: the persistent
 * fields-to-exchange flags, the persistent CG/Chebyshev coefficient arrays,
 * and the 3D sqrt(error) convergence test.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

enum { F_DENSITY, F_ENERGY0, F_ENERGY1, F_U, F_P, F_SD };

struct or_problem {
    tl_settings s;
    int x, y, z, hd;
    long n;
    double dx, dy, dz;
    double *density0, *density, *energy0, *energy, *u, *u0, *p, *r, *mi, *w;
    double *kx, *ky, *kz, *sd;
    double *volume, *x_area, *y_area, *z_area;
    double *cell_x, *cell_y, *cell_z, *vertex_x, *vertex_y, *vertex_z;
    double *cg_alphas, *cg_betas, *cheby_alphas, *cheby_betas;
    double theta, eigmin, eigmax;
    int neighbours[6];
    bool fields[6];
    tl_step_report* rep;
    double rel_tol;   /* > 0: relative true-residual stop (not in the reference) */
    double bb;        /* ||u0||^2 of the current solve when rel_tol > 0 */
};

/* ---------- eigenvalues (2d/drivers/eigenvalue_driver.c) ---------- */

static double sign_of(double a, double b) { return b < 0 ? -fabs(a) : fabs(a); }

/* Eigenvalues of a symmetric tridiagonal matrix by the implicit QL method
 * with Wilkinson shifts (Numerical Recipes in C, 2nd ed., sec. 11.3, tqli),
 * values only.  d: diagonal (overwritten with eigenvalues); e: sub-diagonal
 * with e[0] unused on entry.  Follows 2d/drivers/eigenvalue_driver.c:71-121,
 * including its handling of an exact zero rotation radius (the inner sweep
 * carries on rather than restarting).  Returns -2 after 30 sweeps. */
static int ql_implicit(double* d, double* e, int n)
{
    for (int i = 0; i + 1 < n; ++i) e[i] = e[i + 1];
    if (n > 0) e[n - 1] = 0.0;
    for (int l = 0; l < n; ++l) {
        int sweeps = 0;
        int m;
        do {
            for (m = l; m < n - 1; ++m) {
                const double dd = fabs(d[m]) + fabs(d[m + 1]);
                if (fabs(e[m]) + dd == dd) break;
            }
            if (m == l) break;
            if (sweeps++ == 30) return -2;
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double rad = sqrt(g * g + 1.0);
            g = d[m] - d[l] + e[l] / (g + sign_of(rad, g));
            double s = 1.0, c = 1.0, p = 0.0;
            for (int i = m - 1; i >= l; --i) {
                const double f = s * e[i];
                const double b = c * e[i];
                rad = sqrt(f * f + g * g);
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

/* Lanczos tridiagonal of the CG run: diag_i = 1/a_i + b_{i-1}/a_{i-1},
 * offdiag_{i+1} = sqrt(b_i)/a_i (2d/drivers/eigenvalue_driver.c:24-36). */
int or_eigenvalues(const double* al, const double* be, int n,
        double* eigmin, double* eigmax)
{
    double* d = (double*)calloc(n > 0 ? n : 1, sizeof(double));
    double* e = (double*)calloc(n > 0 ? n : 1, sizeof(double));
    for (int i = 0; i < n; ++i) {
        d[i] = 1.0 / al[i];
        if (i > 0) d[i] += be[i - 1] / al[i - 1];
        if (i < n - 1) e[i + 1] = sqrt(be[i]) / al[i];
    }
    int rc = ql_implicit(d, e, n);
    double lo = DBL_MAX, hi = DBL_MIN;
    for (int i = 0; i < n; ++i) {
        lo = d[i] < lo ? d[i] : lo;
        hi = d[i] > hi ? d[i] : hi;
    }
    free(d);
    free(e);
    if (rc) return rc;
    if (lo < 0.0 || hi < 0.0) return -1;
    *eigmin = lo * 0.95;
    *eigmax = hi * 1.05;
    return 0;
}

/* 2d/drivers/cheby_driver.c:163-183 */
void or_cheby_coefs(double eigmin, double eigmax, int n, double* theta,
        double* alphas, double* betas)
{
    *theta = (eigmax + eigmin) / 2.0;
    const double delta = (eigmax - eigmin) / 2.0;
    const double sigma = *theta / delta;
    double rho_old = 1.0 / sigma;
    for (int i = 0; i < n; ++i) {
        const double rho_new = 1.0 / (2.0 * sigma - rho_old);
        alphas[i] = rho_new * rho_old;
        betas[i] = 2.0 * rho_new / delta;
        rho_old = rho_new;
    }
}

/* 2d/drivers/cheby_driver.c:146-160: the logs and the rounding are single
 * precision in the reference, and that is kept. */
int or_cheby_est_iterations(double eigmin, double eigmax, double error,
        double bb)
{
    const double cn = eigmax / eigmin;
    const double it_alpha = DBL_EPSILON * bb / (4.0 * error);
    const double gamm = (sqrt(cn) - 1.0) / (sqrt(cn) + 1.0);
    const float num = logf((float)it_alpha);
    const float den = logf((float)gamm);
    return (int)roundf((float)((double)num / (2.0 * (double)den)));
}

/* ---------- kernel dispatch by dimension ---------- */

static void k_local_halos(or_problem* P, int depth)
{
    if (P->s.dims == 2)
        or2d_local_halos(P->x, P->y, depth, P->hd, P->neighbours, P->fields,
                P->density, P->energy0, P->energy, P->u, P->p, P->sd);
    else
        or3d_local_halos(P->x, P->y, P->z, depth, P->hd, P->neighbours,
                P->fields, P->density, P->energy0, P->energy, P->u, P->p, P->sd);
}

/* 2d/drivers/halo_update_driver.c:6-25 (remote exchange is a no-op here) */
static void halo_update(or_problem* P, int depth)
{
    int any = 0;
    for (int i = 0; i < 6; ++i) any |= P->fields[i];
    if (any) k_local_halos(P, depth);
}

static void reset_fields(or_problem* P)
{
    memset(P->fields, 0, sizeof(P->fields));
}

static void k_cg_calc_w(or_problem* P, double* pw)
{
    if (P->s.dims == 2) or2d_cg_calc_w(P->x, P->y, P->hd, pw, P->p, P->w, P->kx, P->ky);
    else or3d_cg_solver_calc_w(P->x, P->y, P->z, P->hd, pw, P->p, P->w, P->kx, P->ky, P->kz);
}
static void k_cg_calc_ur(or_problem* P, double alpha, double* rrn)
{
    if (P->s.dims == 2) or2d_cg_calc_ur(P->x, P->y, P->hd, alpha, rrn, P->u, P->p, P->r, P->w);
    else or3d_cg_solver_calc_ur(P->x, P->y, P->z, P->hd, alpha, rrn, P->u, P->p, P->r, P->w);
}
static void k_cg_calc_p(or_problem* P, double beta)
{
    if (P->s.dims == 2) or2d_cg_calc_p(P->x, P->y, P->hd, beta, P->p, P->r);
    else or3d_cg_solver_calc_p(P->x, P->y, P->z, P->hd, beta, P->p, P->r);
}
static void k_copy_u(or_problem* P)
{
    if (P->s.dims == 2) or2d_copy_u(P->x, P->y, P->hd, P->u0, P->u);
    else or3d_solver_copy_u(P->x, P->y, P->z, P->hd, P->u0, P->u);
}
static void k_residual(or_problem* P)
{
    if (P->s.dims == 2) or2d_calculate_residual(P->x, P->y, P->hd, P->u, P->u0, P->r, P->kx, P->ky);
    else or3d_calculate_residual(P->x, P->y, P->z, P->hd, P->u, P->u0, P->r, P->kx, P->ky, P->kz);
}
static void k_2norm(or_problem* P, const double* buf, double* norm)
{
    if (P->s.dims == 2) or2d_calculate_2norm(P->x, P->y, P->hd, buf, norm);
    else or3d_calculate_2norm(P->x, P->y, P->z, P->hd, buf, norm);
}

/* ---------- CG (2d/drivers/cg_driver.c) ---------- */

static void start_solve(or_problem* P);

static int cg_init_driver(or_problem* P, double rx, double ry, double rz,
        double* rro)
{
    int rc;
    *rro = 0.0;
    if (P->s.dims == 2)
        rc = or2d_cg_init(P->x, P->y, P->hd, P->s.coefficient, rx, ry, rro,
                P->density, P->energy, P->u, P->p, P->r, P->w, P->kx, P->ky);
    else
        rc = or3d_cg_solver_init(P->x, P->y, P->z, P->hd, P->s.coefficient,
                rx, ry, rz, rro, P->density, P->energy, P->u, P->p, P->r,
                P->w, P->kx, P->ky, P->kz);
    if (rc) return rc;
    reset_fields(P);
    P->fields[F_U] = true;
    P->fields[F_P] = true;
    halo_update(P, 1);
    k_copy_u(P);
    start_solve(P);
    return 0;
}

static void cg_main_step(or_problem* P, int tt, double* rro, double* error)
{
    double pw = 0.0;
    k_cg_calc_w(P, &pw);
    const double alpha = *rro / pw;
    double rrn = 0.0;
    P->cg_alphas[tt] = alpha;
    k_cg_calc_ur(P, alpha, &rrn);
    const double beta = rrn / *rro;
    P->cg_betas[tt] = beta;
    k_cg_calc_p(P, beta);
    *error = rrn;
    *rro = rrn;
    P->rep->n_cg_steps++;
}

/* ||u0 - A u||^2 with u's depth-1 halo refreshed first; w is scratch. */
static double true_residual(or_problem* P)
{
    bool keep[6];
    memcpy(keep, P->fields, sizeof(keep));
    reset_fields(P);
    P->fields[F_U] = true;
    k_local_halos(P, 1);
    memcpy(P->fields, keep, sizeof(keep));
    if (P->s.dims == 2) or2d_calculate_residual(P->x, P->y, P->hd, P->u, P->u0, P->w, P->kx, P->ky);
    else or3d_calculate_residual(P->x, P->y, P->z, P->hd, P->u, P->u0, P->w, P->kx, P->ky, P->kz);
    double rr = 0.0;
    k_2norm(P, P->w, &rr);
    P->rep->n_true_residuals++;
    P->rep->rel_residual = P->bb > 0.0 ? sqrt(rr / P->bb) : 0.0;
    return rr;
}

/* Convergence test shared by every solver loop: 3D compares sqrt(error)
 * (3d/drivers/cg_solver_driver.c:24-25), 2D the raw value.  With rel_tol
 * set (an extension, not in the reference): error = ||r||^2 at or below
 * rel_tol^2 ||u0||^2, confirmed by the true residual. */
static int converged(or_problem* P, double* error)
{
    if (P->rel_tol > 0.0) {
        const double target = P->rel_tol * P->rel_tol * P->bb;
        if (!(*error <= target)) return 0;
        return true_residual(P) <= target;
    }
    if (P->s.dims == 3) *error = sqrt(*error);
    return fabs(*error) < P->s.eps;
}

static void start_solve(or_problem* P)
{
    P->bb = 0.0;
    if (P->rel_tol > 0.0) k_2norm(P, P->u0, &P->bb);
}

static int cg_driver(or_problem* P, double rx, double ry, double rz,
        double* error)
{
    double rro;
    int rc = cg_init_driver(P, rx, ry, rz, &rro);
    if (rc) return rc;
    int tt;
    for (tt = 0; tt < P->s.max_iters; ++tt) {
        cg_main_step(P, tt, &rro, error);
        halo_update(P, 1);
        if (converged(P, error)) break;
    }
    P->rep->printed_cg_iters = tt;
    return 0;
}

/* ---------- switch rule shared by Chebyshev and PPCG ---------- */
static int switch_now(or_problem* P, int started, int tt, double error)
{
    if (started) return 1;
    if (P->s.error_switch) return error < P->s.eps_lim && tt > 20;
    return tt > P->s.presteps && error < 1.0;
}

static int eigen_init(or_problem* P, int num_cg)
{
    int rc = or_eigenvalues(P->cg_alphas, P->cg_betas, num_cg, &P->eigmin, &P->eigmax);
    P->rep->eigmin = P->eigmin;
    P->rep->eigmax = P->eigmax;
    return rc;
}

/* ---------- Chebyshev (2d/drivers/cheby_driver.c) ---------- */

static void cheby_main_step(or_problem* P, int k, int calc_norm, double* error)
{
    if (P->s.dims == 2)
        or2d_cheby_iterate(P->x, P->y, P->hd, P->cheby_alphas[k], P->cheby_betas[k],
                P->u, P->u0, P->p, P->r, P->w, P->kx, P->ky);
    else
        or3d_cheby_solver_iterate(P->x, P->y, P->z, P->hd, P->cheby_alphas[k],
                P->cheby_betas[k], P->u, P->u0, P->p, P->r, P->w, P->kx, P->ky, P->kz);
    P->rep->n_cheby_steps++;
    if (calc_norm) {
        *error = 0.0;
        k_2norm(P, P->r, error);
    }
}

static int cheby_driver(or_problem* P, double rx, double ry, double rz,
        double* error)
{
    double rro;
    int rc = cg_init_driver(P, rx, ry, rz, &rro);
    if (rc) return rc;
    int est = 0, ncheby = 0, tt;
    for (tt = 0; tt < P->s.max_iters; ++tt) {
        if (!switch_now(P, ncheby, tt, *error)) {
            cg_main_step(P, tt, &rro, error);
        } else {
            ncheby++;
            if (ncheby == 1) {
                double bb = 0.0;
                if ((rc = eigen_init(P, tt))) return rc;
                or_cheby_coefs(P->eigmin, P->eigmax, P->s.max_iters - tt,
                        &P->theta, P->cheby_alphas, P->cheby_betas);
                k_2norm(P, P->u0, &bb);
                if (P->s.dims == 2)
                    or2d_cheby_init(P->x, P->y, P->hd, P->theta, P->u, P->u0,
                            P->p, P->r, P->w, P->kx, P->ky);
                else
                    or3d_cheby_solver_init(P->x, P->y, P->z, P->hd, P->theta,
                            P->u, P->u0, P->p, P->r, P->w, P->kx, P->ky, P->kz);
                reset_fields(P);
                P->fields[F_U] = true;
                halo_update(P, 1);
                cheby_main_step(P, ncheby, 1, error);
                est = or_cheby_est_iterations(P->eigmin, P->eigmax, *error, bb);
            } else {
                const int calc = ncheby >= est && (tt + 1) % 10 == 0;
                cheby_main_step(P, ncheby, calc, error);
            }
        }
        halo_update(P, 1);
        if (converged(P, error)) break;
    }
    P->rep->printed_cg_iters = tt - ncheby + 1;
    P->rep->printed_inner_iters = ncheby;
    P->rep->est_iters = est;
    return 0;
}

/* ---------- PPCG (2d/drivers/ppcg_driver.c) ---------- */

static void ppcg_main_step(or_problem* P, double* rro, double* error)
{
    double pw = 0.0;
    k_cg_calc_w(P, &pw);
    const double alpha = *rro / pw;
    double rrn = 0.0;
    k_cg_calc_ur(P, alpha, &rrn);

    if (P->s.dims == 2) or2d_ppcg_init(P->x, P->y, P->hd, P->theta, P->r, P->sd);
    else or3d_ppcg_init(P->x, P->y, P->z, P->hd, P->theta, P->r, P->sd);
    reset_fields(P);
    P->fields[F_SD] = true;
    for (int pp = 0; pp < P->s.ppcg_inner_steps; ++pp) {
        halo_update(P, 1);
        if (P->s.dims == 2)
            or2d_ppcg_inner_iteration(P->x, P->y, P->hd, P->cheby_alphas[pp],
                    P->cheby_betas[pp], P->u, P->r, P->kx, P->ky, P->sd);
        else
            or3d_ppcg_inner_iteration(P->x, P->y, P->z, P->hd, P->cheby_alphas[pp],
                    P->cheby_betas[pp], P->u, P->r, P->kx, P->ky, P->kz, P->sd);
        P->rep->n_ppcg_inner++;
    }
    reset_fields(P);
    P->fields[F_P] = true;

    rrn = 0.0;
    k_2norm(P, P->r, &rrn);
    const double beta = rrn / *rro;
    k_cg_calc_p(P, beta);
    *error = rrn;
    *rro = rrn;
    P->rep->n_ppcg_outer++;
}

static int ppcg_driver(or_problem* P, double rx, double ry, double rz,
        double* error)
{
    double rro;
    int rc = cg_init_driver(P, rx, ry, rz, &rro);
    if (rc) return rc;
    int nppcg = 0, tt;
    for (tt = 0; tt < P->s.max_iters; ++tt) {
        if (!switch_now(P, nppcg, tt, *error)) {
            cg_main_step(P, tt, &rro, error);
        } else {
            nppcg++;
            if (nppcg == 1) {
                if ((rc = eigen_init(P, tt))) return rc;
                or_cheby_coefs(P->eigmin, P->eigmax, P->s.ppcg_inner_steps,
                        &P->theta, P->cheby_alphas, P->cheby_betas);
                k_residual(P);
                reset_fields(P);
                P->fields[F_P] = true;
                halo_update(P, 1);
            }
            ppcg_main_step(P, &rro, error);
        }
        halo_update(P, 1);
        if (converged(P, error)) break;
    }
    P->rep->printed_cg_iters = tt - nppcg + 1;
    P->rep->printed_inner_iters = nppcg;
    return 0;
}

/* ---------- Jacobi (2d/drivers/jacobi_driver.c) ---------- */

static int jacobi_driver(or_problem* P, double rx, double ry, double rz,
        double* error)
{
    int rc;
    if (P->s.dims == 2)
        rc = or2d_jacobi_init(P->x, P->y, P->hd, P->s.coefficient, rx, ry,
                P->density, P->energy, P->u0, P->u, P->kx, P->ky);
    else
        rc = or3d_jacobi_solver_init(P->x, P->y, P->z, P->hd, P->s.coefficient,
                rx, ry, rz, P->density, P->energy, P->u0, P->u, P->kx, P->ky, P->kz);
    if (rc) return rc;
    k_copy_u(P);
    start_solve(P);
    reset_fields(P);
    P->fields[F_U] = true;
    int tt;
    for (tt = 0; tt < P->s.max_iters; ++tt) {
        if (P->s.dims == 2)
            or2d_jacobi_iterate(P->x, P->y, P->hd, error, P->kx, P->ky, P->u0, P->u, P->r);
        else
            or3d_jacobi_solver_iterate(P->x, P->y, P->z, P->hd, error, P->kx, P->ky,
                    P->kz, P->u0, P->u, P->r);
        P->rep->n_jacobi++;
        if (tt % 50 == 0) {
            halo_update(P, 1);
            k_residual(P);
            k_2norm(P, P->r, error);
        }
        halo_update(P, 1);
        if (converged(P, error)) break;
    }
    P->rep->printed_cg_iters = tt;
    return 0;
}

/* ---------- problem lifecycle ---------- */

static double* zalloc(long n) { return (double*)calloc(n > 0 ? n : 1, sizeof(double)); }

void or_problem_field_summary(or_problem* P, double out[4])
{
    double v = 0, m = 0, ie = 0, t = 0;
    if (P->s.dims == 2)
        or2d_field_summary(P->x, P->y, P->hd, P->volume, P->density, P->energy0,
                P->u, &v, &m, &ie, &t);
    else
        or3d_field_summary(P->x, P->y, P->z, P->hd, P->volume, P->density,
                P->energy0, P->u, &v, &m, &ie, &t);
    out[0] = v; out[1] = m; out[2] = ie; out[3] = t;
}

/* Halo prime + store_energy after the state fill (2d/initialise.c:24-30). */
static void prime(or_problem* P)
{
    reset_fields(P);
    P->fields[F_DENSITY] = true;
    P->fields[F_ENERGY0] = true;
    P->fields[F_ENERGY1] = true;
    halo_update(P, 2);
    if (P->s.dims == 2) or2d_store_energy(P->x, P->y, P->energy0, P->energy);
    else or3d_store_energy(P->x, P->y, P->z, P->energy0, P->energy);
}

or_problem* or_problem_create(const tl_settings* s, const tl_state* states,
        int num_states)
{
    if (!s || (s->dims != 2 && s->dims != 3) || num_states < 1) return NULL;
    or_problem* P = (or_problem*)calloc(1, sizeof(or_problem));
    P->s = *s;
    P->hd = s->halo_depth;
    P->x = s->x_cells + 2 * P->hd;
    P->y = s->y_cells + 2 * P->hd;
    P->z = s->dims == 3 ? s->z_cells + 2 * P->hd : 1;
    P->n = (long)P->x * P->y * P->z;
    P->dx = (s->xmax - s->xmin) / (double)s->x_cells;
    P->dy = (s->ymax - s->ymin) / (double)s->y_cells;
    P->dz = s->dims == 3 ? (s->zmax - s->zmin) / (double)s->z_cells : 0.0;
    for (int f = 0; f < 6; ++f) P->neighbours[f] = -1;
    const long n = P->n;
    double** fields[] = {&P->density0, &P->density, &P->energy0, &P->energy,
        &P->u, &P->u0, &P->p, &P->r, &P->mi, &P->w, &P->kx, &P->ky, &P->kz,
        &P->sd, &P->volume, &P->x_area, &P->y_area, &P->z_area};
    for (size_t i = 0; i < sizeof(fields) / sizeof(fields[0]); ++i)
        *fields[i] = zalloc(n + 2L * P->x * P->y + 2L * P->x + 2);
    P->cell_x = zalloc(P->x + 1); P->vertex_x = zalloc(P->x + 1);
    P->cell_y = zalloc(P->y + 1); P->vertex_y = zalloc(P->y + 1);
    P->cell_z = zalloc(P->z + 1); P->vertex_z = zalloc(P->z + 1);
    /* The reference sizes these max_iters (2d/c_kernels/omp3/kernel_initialise.c:68)
     * while PPCG writes ppcg_inner_steps Chebyshev coefficients into them
     * (cheby_coef_driver(chunks, settings->ppcg_inner_steps)): a deck with
     * more inner steps than max_iters overruns the reference's arrays.  Here
     * they hold both. */
    const int mi = (s->max_iters > s->ppcg_inner_steps ? s->max_iters : s->ppcg_inner_steps) + 1;
    P->cg_alphas = zalloc(mi); P->cg_betas = zalloc(mi);
    P->cheby_alphas = zalloc(mi); P->cheby_betas = zalloc(mi);

    if (s->dims == 2) {
        or2d_set_chunk_data(P->x, P->y, P->hd, 0, 0, s->xmin, s->ymin, P->dx, P->dy,
                P->cell_x, P->cell_y, P->vertex_x, P->vertex_y, P->volume,
                P->x_area, P->y_area);
        or2d_set_chunk_state(P->x, P->y, P->vertex_x, P->vertex_y, P->cell_x,
                P->cell_y, P->density, P->energy0, P->u, num_states, states);
    } else {
        or3d_set_chunk_data(P->x, P->y, P->z, P->hd, 0, 0, 0, s->xmin, s->ymin,
                s->zmin, P->dx, P->dy, P->dz, P->cell_x, P->cell_y, P->cell_z,
                P->vertex_x, P->vertex_y, P->vertex_z, P->volume, P->x_area,
                P->y_area, P->z_area);
        or3d_set_chunk_state(P->x, P->y, P->z, P->vertex_x, P->vertex_y,
                P->vertex_z, P->cell_x, P->cell_y, P->cell_z, P->density,
                P->energy0, P->u, num_states, states);
    }
    prime(P);
    return P;
}

int or_problem_set_rel_tol(or_problem* P, double rel_tol)
{
    if (!(rel_tol >= 0.0)) return -1;
    P->rel_tol = rel_tol;
    return 0;
}

int or_problem_inject(or_problem* P, const double* density,
        const double* energy0)
{
    memcpy(P->density, density, sizeof(double) * P->n);
    memcpy(P->energy0, energy0, sizeof(double) * P->n);
    const int z0 = P->s.dims == 3 ? 1 : 0, z1 = P->s.dims == 3 ? P->z - 1 : 1;
    for (int ii = z0; ii < z1; ++ii)
        for (int jj = 1; jj < P->y - 1; ++jj)
            for (int kk = 1; kk < P->x - 1; ++kk) {
                const long c = kk + (long)jj * P->x + (long)ii * P->x * P->y;
                P->u[c] = P->energy0[c] * P->density[c];
            }
    prime(P);
    return 0;
}

/* One timestep: 2d/diffuse.c:22-80. */
int or_problem_step(or_problem* P, int tt, tl_step_report* rep)
{
    memset(rep, 0, sizeof(*rep));
    P->rep = rep;
    const double dt = P->s.dt_init;
    const double rx = dt / (P->dx * P->dx);
    const double ry = dt / (P->dy * P->dy);
    const double rz = P->s.dims == 3 ? dt / (P->dz * P->dz) : 0.0;

    reset_fields(P);
    P->fields[F_ENERGY1] = true;
    P->fields[F_DENSITY] = true;
    halo_update(P, 2);

    double error = 1e+10;
    int rc = 0;
    switch (P->s.solver) {
    case 0: rc = jacobi_driver(P, rx, ry, rz, &error); break;
    case 1: rc = cg_driver(P, rx, ry, rz, &error); break;
    case 2: rc = cheby_driver(P, rx, ry, rz, &error); break;
    case 3: rc = ppcg_driver(P, rx, ry, rz, &error); break;
    default: rc = -3;
    }
    if (rc) return rc;
    rep->error = error;

    /* solve_finished_driver.c:7-43 */
    double exact = 0.0;
    if (P->s.check_result) {
        k_residual(P);
        k_2norm(P, P->r, &exact);
    }
    rep->exact_error = exact;
    if (P->s.dims == 2) or2d_finalise(P->x, P->y, P->hd, P->energy, P->density, P->u);
    else or3d_solver_finalise(P->x, P->y, P->z, P->hd, P->energy, P->density, P->u);
    P->fields[F_ENERGY1] = true;
    halo_update(P, 1);

    if (tt % P->s.summary_frequency == 0) {
        double fs[4];
        or_problem_field_summary(P, fs);
        rep->summary_done = 1;
        rep->vol = fs[0]; rep->mass = fs[1]; rep->ie = fs[2]; rep->temp = fs[3];
    }
    return 0;
}

long or_problem_field_len(or_problem* P) { return P->n; }

int or_problem_get_field(or_problem* P, int id, double* out)
{
    double* f[] = {P->density, P->energy0, P->energy, P->u, P->u0, P->p,
        P->r, P->w, P->kx, P->ky, P->kz, P->sd};
    if (id < 0 || id >= (int)(sizeof(f) / sizeof(f[0]))) return -1;
    memcpy(out, f[id], sizeof(double) * P->n);
    return 0;
}

void or_problem_destroy(or_problem* P)
{
    if (!P) return;
    double* all[] = {P->density0, P->density, P->energy0, P->energy, P->u,
        P->u0, P->p, P->r, P->mi, P->w, P->kx, P->ky, P->kz, P->sd, P->volume,
        P->x_area, P->y_area, P->z_area, P->cell_x, P->cell_y, P->cell_z,
        P->vertex_x, P->vertex_y, P->vertex_z, P->cg_alphas, P->cg_betas,
        P->cheby_alphas, P->cheby_betas};
    for (size_t i = 0; i < sizeof(all) / sizeof(all[0]); ++i) free(all[i]);
    free(P);
}
