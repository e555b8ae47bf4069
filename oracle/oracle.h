/*
 * TeaLeaf CPU ORACLE -- test infrastructure, not product code.
 *
 * A serial, plain-C restatement of the reference's omp3 kernel suite and of
 * its host solver drivers (georgepaw/TeaLeaf, 2d/ and 3d/).  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker.  The product (libtealeaf_hip.so) never links it.
 *
 * Every kernel keeps the reference's per-cell expression order and loop
 * bounds; build with -ffp-contract=off so that, compiled the same way, it is
 * bit-identical to the reference omp3 kernels built without OpenMP.
 *
 * Layout (both dims): x, y (, z) include the halo; index = kk + jj*x
 * (+ ii*x*y).  Interior = [hd, n-hd) in every dimension.
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   - every kernel: bit-exact against the reference omp3 kernels compiled
 *     from /root/reference (oracle/_ref/libomp3_{2d,3d}.so) on seeded grids,
 *     and against the committed fixtures in tests/golden/;
 *   - drivers: end-to-end against the reference host+omp3 binaries
 *     (oracle/_ref/tealeaf{2d,3d}_omp3) and the reference's known answers
 *     (2d/tea.problems, 3d/tea.problems);
 *   - closed forms: A*const = const, symmetry <Ax,y> = <x,Ay>, uniform-density
 *     coefficients, the cosine-mode exact solve, dense numpy solve at 12^2.
 */
#ifndef TL_ORACLE_H
#define TL_ORACLE_H

#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Problem settings: the subset of the reference Settings (2d/settings.h:64-122)
 * that the solve path reads.  Field order is shared with the product header. */
typedef struct {
    int dims;                 /* 2 or 3 */
    int x_cells, y_cells, z_cells;
    double xmin, ymin, zmin, xmax, ymax, zmax;
    double dt_init;
    int end_step;
    int max_iters;
    double eps;
    int solver;               /* 0 Jacobi, 1 CG, 2 Chebyshev, 3 PPCG (2d/settings.h:48-54) */
    int coefficient;          /* 1 CONDUCTIVITY, 2 RECIP_CONDUCTIVITY (2d/shared.h:38-39) */
    int presteps;
    int ppcg_inner_steps;
    double eps_lim;
    int error_switch;
    int check_result;
    int summary_frequency;
    int halo_depth;
} tl_settings;

/* One material state (2d/settings.h:133-144, 3d/settings.h).  Extents are
 * already shrunk by d/100 as the reference parser does (2d/parse_config.c:222-260).
 * geometry: 0 rectangle/cuboid, 1 circular/spherical, 2 point. */
typedef struct {
    double density, energy;
    double x_min, y_min, z_min, x_max, y_max, z_max;
    double radius;
    int geometry;
    int defined;
} tl_state;

/* What one timestep reports (the reference prints these, 2d/diffuse.c:70-78,
 * 2d/drivers/cg_driver.c et al.). */
typedef struct {
    int printed_cg_iters;     /* "CG: N iterations" (or "Jacobi: N") */
    int printed_inner_iters;  /* Cheby / PPCG outer count, 0 for CG */
    int est_iters;            /* Chebyshev estimate, 0 otherwise */
    int n_cg_steps;           /* CG main steps actually executed */
    int n_cheby_steps;        /* cheby_iterate calls */
    int n_ppcg_outer;         /* PPCG outer iterations */
    int n_ppcg_inner;         /* ppcg_inner_iteration calls */
    int n_jacobi;             /* jacobi_iterate calls */
    double error;             /* final solver error as printed */
    double exact_error;       /* ||u0 - A u||^2 from solve_finished (if check_result) */
    double eigmin, eigmax;
    int summary_done;
    double vol, mass, ie, temp;
    /* Not printed by the reference: sqrt(||u0 - A u||^2 / ||u0||^2) in FP64
     * at the last true-residual evaluation (relative-residual stop or mixed
     * precision; 0 otherwise) and how many such evaluations the solve made. */
    double rel_residual;
    int n_true_residuals;
} tl_step_report;

/* ---------------- 2D kernels (2d/c_kernels/omp3) ---------------- */
void or2d_set_chunk_data(int x, int y, int hd, int left, int bottom,
        double grid_x_min, double grid_y_min, double dx, double dy,
        double* cell_x, double* cell_y, double* vertex_x, double* vertex_y,
        double* volume, double* x_area, double* y_area);
void or2d_set_chunk_state(int x, int y, const double* vertex_x,
        const double* vertex_y, const double* cell_x, const double* cell_y,
        double* density, double* energy0, double* u, int num_states,
        const tl_state* states);
void or2d_local_halos(int x, int y, int depth, int hd, const int* neighbours,
        const bool* fields, double* density, double* energy0, double* energy,
        double* u, double* p, double* sd);
int or2d_pack_or_unpack(int x, int y, int depth, int hd, int face, bool pack,
        double* field, double* buffer);
void or2d_store_energy(int x, int y, const double* energy0, double* energy);
void or2d_field_summary(int x, int y, int hd, const double* volume,
        const double* density, const double* energy0, const double* u,
        double* vol, double* mass, double* ie, double* temp);
int or2d_cg_init(int x, int y, int hd, int coefficient, double rx, double ry,
        double* rro, const double* density, const double* energy, double* u,
        double* p, double* r, double* w, double* kx, double* ky);
void or2d_cg_calc_w(int x, int y, int hd, double* pw, const double* p,
        double* w, const double* kx, const double* ky);
void or2d_cg_calc_ur(int x, int y, int hd, double alpha, double* rrn,
        double* u, const double* p, double* r, const double* w);
void or2d_cg_calc_p(int x, int y, int hd, double beta, double* p,
        const double* r);
void or2d_cheby_init(int x, int y, int hd, double theta, double* u,
        const double* u0, double* p, double* r, double* w, const double* kx,
        const double* ky);
void or2d_cheby_iterate(int x, int y, int hd, double alpha, double beta,
        double* u, const double* u0, double* p, double* r, double* w,
        const double* kx, const double* ky);
int or2d_jacobi_init(int x, int y, int hd, int coefficient, double rx,
        double ry, const double* density, const double* energy, double* u0,
        double* u, double* kx, double* ky);
void or2d_jacobi_iterate(int x, int y, int hd, double* error,
        const double* kx, const double* ky, const double* u0, double* u,
        double* r);
void or2d_ppcg_init(int x, int y, int hd, double theta, const double* r,
        double* sd);
void or2d_ppcg_inner_iteration(int x, int y, int hd, double alpha,
        double beta, double* u, double* r, const double* kx, const double* ky,
        double* sd);
void or2d_copy_u(int x, int y, int hd, double* u0, const double* u);
void or2d_calculate_residual(int x, int y, int hd, const double* u,
        const double* u0, double* r, const double* kx, const double* ky);
void or2d_calculate_2norm(int x, int y, int hd, const double* buffer,
        double* norm);
void or2d_finalise(int x, int y, int hd, double* energy,
        const double* density, const double* u);

/* ---------------- 3D kernels (3d/c_kernels/omp3) ---------------- */
void or3d_set_chunk_data(int x, int y, int z, int hd, int left, int bottom,
        int back, double grid_x_min, double grid_y_min, double grid_z_min,
        double dx, double dy, double dz, double* cell_x, double* cell_y,
        double* cell_z, double* vertex_x, double* vertex_y, double* vertex_z,
        double* volume, double* x_area, double* y_area, double* z_area);
void or3d_set_chunk_state(int x, int y, int z, const double* vertex_x,
        const double* vertex_y, const double* vertex_z, const double* cell_x,
        const double* cell_y, const double* cell_z, double* density,
        double* energy0, double* u, int num_states, const tl_state* states);
void or3d_local_halos(int x, int y, int z, int depth, int hd,
        const int* neighbours, const bool* fields, double* density,
        double* energy0, double* energy, double* u, double* p, double* sd);
int or3d_pack_or_unpack(int x, int y, int z, int depth, int hd, int face,
        bool pack, double* field, double* buffer);
void or3d_store_energy(int x, int y, int z, const double* energy0,
        double* energy);
void or3d_field_summary(int x, int y, int z, int hd, const double* volume,
        const double* density, const double* energy0, const double* u,
        double* vol, double* mass, double* ie, double* temp);
int or3d_cg_solver_init(int x, int y, int z, int hd, int coefficient,
        double rx, double ry, double rz, double* rro, const double* density,
        const double* energy, double* u, double* p, double* r, double* w,
        double* kx, double* ky, double* kz);
void or3d_cg_solver_calc_w(int x, int y, int z, int hd, double* pw,
        const double* p, double* w, const double* kx, const double* ky,
        const double* kz);
void or3d_cg_solver_calc_ur(int x, int y, int z, int hd, double alpha,
        double* rrn, double* u, const double* p, double* r, const double* w);
void or3d_cg_solver_calc_p(int x, int y, int z, int hd, double beta,
        double* p, const double* r);
void or3d_cheby_solver_init(int x, int y, int z, int hd, double theta,
        double* u, const double* u0, double* p, double* r, double* w,
        const double* kx, const double* ky, const double* kz);
void or3d_cheby_solver_iterate(int x, int y, int z, int hd, double alpha,
        double beta, double* u, const double* u0, double* p, double* r,
        double* w, const double* kx, const double* ky, const double* kz);
int or3d_jacobi_solver_init(int x, int y, int z, int hd, int coefficient,
        double rx, double ry, double rz, const double* density,
        const double* energy, double* u0, double* u, double* kx, double* ky,
        double* kz);
void or3d_jacobi_solver_iterate(int x, int y, int z, int hd, double* error,
        const double* kx, const double* ky, const double* kz,
        const double* u0, double* u, double* r);
void or3d_ppcg_init(int x, int y, int z, int hd, double theta,
        const double* r, double* sd);
void or3d_ppcg_inner_iteration(int x, int y, int z, int hd, double alpha,
        double beta, double* u, double* r, const double* kx, const double* ky,
        const double* kz, double* sd);
void or3d_solver_copy_u(int x, int y, int z, int hd, double* u0,
        const double* u);
void or3d_calculate_residual(int x, int y, int z, int hd, const double* u,
        const double* u0, double* r, const double* kx, const double* ky,
        const double* kz);
void or3d_calculate_2norm(int x, int y, int z, int hd, const double* buffer,
        double* norm);
void or3d_solver_finalise(int x, int y, int z, int hd, double* energy,
        const double* density, const double* u);

/* ---------------- host-side solver helpers (2d/drivers) ---------------- */
/* Lanczos tridiagonal from CG alphas/betas + implicit-QL eigenvalues
 * (2d/drivers/eigenvalue_driver.c:11-67,71-121).  Returns 0, or -1 on
 * negative eigenvalues / -2 when QL needs more than 30 sweeps. */
int or_eigenvalues(const double* cg_alphas, const double* cg_betas, int n,
        double* eigmin, double* eigmax);
/* Chebyshev coefficients (2d/drivers/cheby_driver.c:163-183). */
void or_cheby_coefs(double eigmin, double eigmax, int n, double* theta,
        double* alphas, double* betas);
/* Iteration estimate in float logf/roundf (2d/drivers/cheby_driver.c:146-160). */
int or_cheby_est_iterations(double eigmin, double eigmax, double error,
        double bb);

/* ---------------- whole-problem driver (2d/diffuse.c, drivers/) -------- */
typedef struct or_problem or_problem;
/* Builds the chunk as initialise_application does (2d/initialise.c:11-31):
 * allocate, set_chunk_data, set_chunk_state, halo(depth 2), store_energy. */
or_problem* or_problem_create(const tl_settings* s, const tl_state* states,
        int num_states);
/* Overwrite density/energy0 (full arrays incl. halo), then redo the halo
 * prime, u = e*rho and store_energy, as for a deck-defined state. */
int or_problem_inject(or_problem* p, const double* density,
        const double* energy0);
/* One timestep, as 2d/diffuse.c:22-80 (3D: 3d/diffuse.c). tt is 0-based. */
int or_problem_step(or_problem* p, int tt, tl_step_report* rep);
/* Extension (not in the reference): stop a solve once ||u0 - A u||_2 <=
 * rel_tol ||u0||_2, the recurrence residual being confirmed by the true one
 * (u's depth-1 halo refreshed first); 0 restores the eps test. */
int or_problem_set_rel_tol(or_problem* p, double rel_tol);
void or_problem_field_summary(or_problem* p, double out[4]);
/* Copy a field (full array incl. halo) out.  id: 0 density, 1 energy0,
 * 2 energy, 3 u, 4 u0, 5 p, 6 r, 7 w, 8 kx, 9 ky, 10 kz, 11 sd. */
int or_problem_get_field(or_problem* p, int id, double* out);
long or_problem_field_len(or_problem* p);
void or_problem_destroy(or_problem* p);

#ifdef __cplusplus
}
#endif
#endif
