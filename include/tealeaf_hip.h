/*
 * tealeaf_hip.h -- C ABI of libtealeaf_hip.so, the MI355X (gfx950) kernel
 * suite for TeaLeaf's implicit heat-conduction solve.
 *
 * Each tlhip2d_* / tlhip3d_* kernel entry point replaces one run_* entry
 * point of the reference kernel interface, cited below as
 * 2d/kernel_interface.h:LINE or 3d/kernel_interface.h:LINE.  Arguments are
 * the ones the reference's omp3 suite passes for that entry point
 * (2d/c_kernels/omp3/kernel_interface.c): grid sizes x, y (, z) INCLUDING the
 * halo, halo_depth, scalars, and field pointers.
 *
 *  - Field pointers are DEVICE pointers to x*y (x*y*z) doubles, row-major,
 *    index = kk + jj*x (+ ii*x*y), exactly the reference layout.
 *  - Scalar results are HOST pointers and are ACCUMULATED (+=), as the
 *    reference does; the Jacobi error is ASSIGNED (2d/c_kernels/omp3/jacobi.c:99).
 *    A scalar is final when the call returns.
 *  - The Chebyshev and PPCG sweeps fuse the reference's two passes into one
 *    and therefore read the stencil field from one buffer and write the
 *    updated field to another (u_in/u_out, sd_in/sd_out).  The output buffer
 *    ends up equal, halo included, to what the reference computes in place;
 *    callers swap the two pointers afterwards.
 *  - Every function returns 0 on success or a TLHIP_ERR_* code, with a
 *    message from tlhip_last_error().  (The reference calls die() instead;
 *    the hip_cdna4 suite glue does exactly that on a non-zero return.)
 *  - Work is queued on one in-order HIP stream.  Calls are sequential, as the
 *    reference's are; the library is not re-entrant.
 *
 * The struct layouts below are shared with the CPU oracle (oracle/oracle.h).
 */
#ifndef TEALEAF_HIP_H
#define TEALEAF_HIP_H

#include <stdbool.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TLHIP_OK 0
#define TLHIP_ERR_ARG 1     /* invalid argument or shape */
#define TLHIP_ERR_HIP 2     /* HIP runtime / launch failure */
#define TLHIP_ERR_SOLVER 3  /* solver failure (negative eigenvalues, QL) */

/* One material state, as 2d/settings.h:133-144 / 3d/settings.h:131-145.
 * geometry: 0 rectangle/cuboid, 1 circular/spherical, 2 point.  Extents
 * already include the reference parser's d/100 shrink (2d/parse_config.c:222-260). */
typedef struct {
    double density, energy;
    double x_min, y_min, z_min, x_max, y_max, z_max;
    double radius;
    int geometry;
    int defined;
} tlhip_state;

/* ------------------------------------------------------------ runtime */
/* Select the device (rank -> device, as the reference's CUDA suite does,
 * 2d/c_kernels/cuda/kernel_initialise.cu:57-66) and create the stream.
 * Called implicitly with device 0 by any other entry point. */
int tlhip_init(int device);
int tlhip_finalise(void);
int tlhip_device(void);
const char* tlhip_last_error(void);
int tlhip_synchronize(void);
/* Longest a host-side wait for a device result (a reduction scalar, a
 * resident loop's iteration report) may take before the call fails with
 * TLHIP_ERR_HIP and a message naming what it waited for; <= 0 restores
 * the default of 300 s.  No reference counterpart: the reference's
 * drivers block in their device synchronisation. */
int tlhip_set_wait_limit(double seconds);
/* Run the suite's work on a caller-owned stream (e.g. a framework's). */
int tlhip_set_stream(void* stream);
void* tlhip_get_stream(void);
/* Tuning knobs: rows (2D) / planes (3D) per stencil tile; <= 0 restores
 * the automatic choice. */
int tlhip_set_tile_rows(int rows);
int tlhip_set_tile_planes(int planes);
/* Most blocks a pointwise sweep (calc_ur, calc_p, 2-norm, ...) launches;
 * <= 0 restores the default (4096). */
int tlhip_set_interior_blocks(int blocks);

/* Device memory, zero-filled (replaces the allocate_buffer calls of
 * run_kernel_initialise, 2d/kernel_interface.h:17 / 3d/kernel_interface.h:13,
 * and run_kernel_finalise, 2d/kernel_interface.h:19). */
int tlhip_alloc(size_t bytes, void** out);
int tlhip_free(void* ptr);
int tlhip_copy_to_device(void* dst, const void* src, size_t bytes);
int tlhip_copy_to_host(void* dst, const void* src, size_t bytes);
int tlhip_copy_device(void* dst, const void* src, size_t bytes);
/* Kernel timer: after start, every launch of the main kernel of the named
 * entry points (one name, e.g. "tlhip2d_ppcg_inner_iteration", or a
 * comma-separated list of up to 8) is bracketed by HIP events on the
 * suite's stream.  read_entry synchronises and returns one entry point's
 * summed kernel time and launch count so far; read does the same for the
 * first name and stops timing. */
int tlhip_kernel_timer_start(const char* entry_points);
int tlhip_kernel_timer_read_entry(const char* entry_point, double* total_ms, long* launches);
int tlhip_kernel_timer_read(double* total_ms, long* launches);
/* A plain streaming copy (n doubles) on the suite's stream, timed under the
 * name "tlhip_stream_copy": the reference point for HBM byte counters. */
int tlhip_stream_copy(double* dst, const double* src, size_t n);
/* Lane-shift self test: out[l] = in[l-1] (lane 0: -1), out[64+l] = in[l+1]
 * (lane 63: -1); 64 doubles in, 128 out, device pointers. */
int tlhip_selftest_lanes(const double* in_dev, double* out_dev);

/* ------------------------------------------------------------ 2D */
/* 2d/kernel_interface.h:13 run_set_chunk_data */
int tlhip2d_set_chunk_data(int x, int y, int halo_depth, int left, int bottom,
        double grid_x_min, double grid_y_min, double dx, double dy,
        double* cell_x, double* cell_y, double* vertex_x, double* vertex_y,
        double* volume, double* x_area, double* y_area);
/* 2d/kernel_interface.h:15 run_set_chunk_state (states: host array) */
int tlhip2d_set_chunk_state(int x, int y, const double* vertex_x,
        const double* vertex_y, const double* cell_x, const double* cell_y,
        double* density, double* energy0, double* u, int num_states,
        const tlhip_state* states);
/* 2d/kernel_interface.h:23 run_local_halos.  neighbours[4] and fields[6]
 * are host arrays (faces 2d/shared.h:25-29, field ids :31-36); every
 * flagged field and external face is done in one launch. */
int tlhip2d_local_halos(int x, int y, int depth, int halo_depth,
        const int* neighbours, const bool* fields, double* density,
        double* energy0, double* energy, double* u, double* p, double* sd);
/* 2d/kernel_interface.h:25 run_pack_or_unpack (buffer: host memory) */
int tlhip2d_pack_or_unpack(int x, int y, int depth, int halo_depth, int face,
        bool pack, double* field, double* buffer);
/* 2d/kernel_interface.h:28 run_store_energy */
int tlhip2d_store_energy(int x, int y, const double* energy0, double* energy);
/* 2d/kernel_interface.h:30 run_field_summary */
int tlhip2d_field_summary(int x, int y, int halo_depth, const double* volume,
        const double* density, const double* energy0, const double* u,
        double* vol, double* mass, double* ie, double* temp);
/* 2d/kernel_interface.h:35 run_cg_init */
int tlhip2d_cg_init(int x, int y, int halo_depth, int coefficient, double rx,
        double ry, double* rro, const double* density, const double* energy,
        double* u, double* p, double* r, double* w, double* kx, double* ky);
/* 2d/kernel_interface.h:38 run_cg_calc_w */
int tlhip2d_cg_calc_w(int x, int y, int halo_depth, double* pw,
        const double* p, double* w, const double* kx, const double* ky);
/* No counterpart in the reference: tlhip2d_cg_calc_w for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops).  The
 * call allocates nothing: its per-block partial sums go to a fixed device
 * array of 65536 doubles (512 KiB, held by each GPU the library runs on
 * until the process ends; the sweep's rows per block grow so that every
 * accepted grid fits) and the dot product is added (+=) into *pw_dev, a
 * device double, on the suite's stream; w and pw_dev equal, bit for bit,
 * what tlhip2d_cg_calc_w gives.  tlhip_init must have been called (it
 * records the array's address).  Rejected: grids over 2^31 - 1 cells and
 * rows over 65536 strips of 64 lanes (more than 4 million cells).  Calls
 * that overlap on one GPU (other streams or threads) share the array and
 * race. */
int tlhip2d_cg_calc_w_noalloc(int x, int y, int halo_depth, double* pw_dev,
        const double* p, double* w, const double* kx, const double* ky);
/* 2d/kernel_interface.h:40 run_cg_calc_ur */
int tlhip2d_cg_calc_ur(int x, int y, int halo_depth, double alpha,
        double* rrn, double* u, const double* p, double* r, const double* w);
/* 2d/kernel_interface.h:42 run_cg_calc_p */
int tlhip2d_cg_calc_p(int x, int y, int halo_depth, double beta, double* p,
        const double* r);
/* 2d/kernel_interface.h:46 run_cheby_init (in place, as the reference) */
int tlhip2d_cheby_init(int x, int y, int halo_depth, double theta, double* u,
        const double* u0, double* p, double* r, double* w, const double* kx,
        const double* ky);
/* 2d/kernel_interface.h:48 run_cheby_iterate, fused; u_out != u_in.
 * w may be NULL: the reference never reads w before rewriting it. */
int tlhip2d_cheby_iterate(int x, int y, int halo_depth, double alpha,
        double beta, const double* u_in, double* u_out, const double* u0,
        double* p, double* r, double* w, const double* kx, const double* ky);
/* 2d/kernel_interface.h:52 run_jacobi_init */
int tlhip2d_jacobi_init(int x, int y, int halo_depth, int coefficient,
        double rx, double ry, const double* density, const double* energy,
        double* u0, double* u, double* kx, double* ky);
/* 2d/kernel_interface.h:54 run_jacobi_iterate (error is assigned) */
int tlhip2d_jacobi_iterate(int x, int y, int halo_depth, double* error,
        const double* kx, const double* ky, const double* u0, double* u,
        double* r);
/* 2d/kernel_interface.h:58 run_ppcg_init */
int tlhip2d_ppcg_init(int x, int y, int halo_depth, double theta,
        const double* r, double* sd);
/* 2d/kernel_interface.h:60 run_ppcg_inner_iteration, fused; sd_out != sd_in */
int tlhip2d_ppcg_inner_iteration(int x, int y, int halo_depth, double alpha,
        double beta, double* u, double* r, const double* kx, const double* ky,
        const double* sd_in, double* sd_out);
/* Temporally blocked PPCG inner loop, 1 <= count <= 10 steps in one pass.
 * No single counterpart in the kernel interface: it replaces `count` turns
 * of the loop in 2d/drivers/ppcg_driver.c:165-183 (a depth-1 halo update of
 * sd, then run_ppcg_inner_iteration with alphas[i], betas[i]) for a chunk
 * whose four faces are external (neighbours[4] all -1).  With init_from_r,
 * sd starts as r/theta (run_ppcg_init, 2d/kernel_interface.h:58) and sd_in
 * is not read; if p and w are also given, r and u first take run_cg_calc_ur's
 * update (u += alpha p, r -= alpha w; 2d/kernel_interface.h:40, which
 * ppcg_main_step_driver runs before the inner loop, ppcg_driver.c:106-116,
 * and whose rrn PPCG overwrites unused, ppcg_driver.c:121-130).  r, u and sd are read from *_in and written
 * to *_out; the interiors equal, bit for bit, the reference's in-place
 * results and the halos of r and u are carried over (sd's too unless
 * init_from_r).  Needs an even x and 16-byte aligned fields. */
int tlhip2d_ppcg_inner_iterations(int x, int y, int halo_depth, int count,
        const double* alphas, const double* betas, double theta,
        int init_from_r, const int* neighbours, const double* sd_in,
        double* sd_out, const double* r_in, double* r_out, const double* u_in,
        double* u_out, const double* kx, const double* ky, double alpha,
        const double* p, const double* w);
/* The same loop on FP32 fields (the inner sweeps of the mixed-precision
 * solve): rows of `pitch` floats (a multiple of 4, >= x), interior
 * [halo_depth, x-halo_depth) x [halo_depth, y-halo_depth).  Coefficients
 * and theta are rounded to FP32.  The reference's expressions in its order,
 * with each product that feeds an addition fused into one FP32
 * multiply-add: kxr*rgt + kxc*lft and kyu*up + kyc*dn as fma(a, b, c*d);
 * A sd = fma(diag, sd, -x terms) - y terms; sd' = fma(alpha, sd, beta*r');
 * r - alpha w and u + alpha p as one fma each; sd = r * (1/theta). */
int tlhip2d_ppcg_inner_iterations_f32(int x, int y, int halo_depth, int pitch,
        int count, const double* alphas, const double* betas, double theta,
        int init_from_r, const int* neighbours, const float* sd_in,
        float* sd_out, const float* r_in, float* r_out, const float* u_in,
        float* u_out, const float* kx, const float* ky, double alpha,
        const float* p, const float* w);
/* No counterpart in the reference (whose PPCG driver calls
 * calculate_2norm after the inner steps, 2d/drivers/ppcg_driver.c:121-131):
 * tlhip2d_ppcg_inner_iterations_f32 that also sets *rrn (host) to the
 * squared 2-norm, accumulated in FP64, of the interior of r_out, formed in
 * the same launch. */
int tlhip2d_ppcg_inner_iterations_f32_norm(int x, int y, int halo_depth, int pitch,
        int count, const double* alphas, const double* betas, double theta,
        int init_from_r, const int* neighbours, const float* sd_in,
        float* sd_out, const float* r_in, float* r_out, const float* u_in,
        float* u_out, const float* kx, const float* ky, double alpha,
        const float* p, const float* w, double* rrn);
/* run_cg_calc_p (2d/kernel_interface.h:42), a depth-1 halo update of p and
 * run_cg_calc_w (:38) in one pass, for a chunk whose faces are external:
 * p_out = beta p_in + r, w = A p_out, *pw += sum w p_out.  p_out and w equal
 * the reference's results bit for bit (p_out's halo ring is p_in's).  This
 * is one CG step's calc_p followed by the next step's calc_w
 * (2d/drivers/cg_driver.c:115 then :78).  Needs an even x. */
int tlhip2d_cg_calc_pw(int x, int y, int halo_depth, double beta, double* pw,
        const double* p_in, double* p_out, const double* r, double* w,
        const double* kx, const double* ky);
/* The CG pair of passes with kx and ky formed from the cells' conductivity
 * instead of loaded.  No counterpart in the reference: one 8-byte field
 * replaces kx and ky, and w = A p need not be stored between the passes.
 * cond is run_cg_init's intermediate w (2d/c_kernels/omp3/cg.c:45-53:
 * density, or 1/density for coefficient 2) over the whole grid, formed by
 * tlhip2d_cg_cond from the density cg_init reads; every face coefficient is
 * then cg_init's expression on the same operands (cg.c:55-64), so results
 * equal those from cg_init's kx and ky bit for bit.  Halo depth >= 2 (where
 * cg_init's kx reads no stale w) and an even x.
 * tlhip2d_cg_calc_pw_cond: with update, run_cg_calc_p (2d/kernel_interface.h:42)
 * and run_cg_calc_w (:38) as tlhip2d_cg_calc_pw; without, run_cg_calc_w on
 * p_in (p_out and r unused).  w is written unless NULL; *pw += sum w p.
 * tlhip2d_cg_calc_ur_cond: run_cg_calc_ur (:40) with w = A p formed from p
 * and cond in the pass: u += alpha p, r -= alpha w, *rrn += sum r r. */
int tlhip2d_cg_cond(int x, int y, int halo_depth, int coefficient, const double* density,
        double* cond);
int tlhip2d_cg_calc_pw_cond(int x, int y, int halo_depth, int update, double beta, double rx,
        double ry, double* pw, const double* p_in, double* p_out, const double* r,
        double* w, const double* cond);
int tlhip2d_cg_calc_ur_cond(int x, int y, int halo_depth, double alpha, double rx, double ry,
        double* rrn, double* u, const double* p, double* r, const double* cond);
/* The same three in 3D (3d/c_kernels/omp3/cg_solver.c:47-75 for cond and
 * the face coefficients; run_cg_solver_calc_p, _calc_w and _calc_ur,
 * 3d/kernel_interface.h:30-36): kx, ky and kz formed from cond, one field
 * in place of three.  Halo depth >= 2 and an even x. */
int tlhip3d_cg_cond(int x, int y, int z, int halo_depth, int coefficient,
        const double* density, double* cond);
int tlhip3d_cg_calc_pw_cond(int x, int y, int z, int halo_depth, int update, double beta,
        double rx, double ry, double rz, double* pw, const double* p_in, double* p_out,
        const double* r, double* w, const double* cond);
int tlhip3d_cg_calc_ur_cond(int x, int y, int z, int halo_depth, double alpha, double rx,
        double ry, double rz, double* rrn, double* u, const double* p, double* r,
        const double* cond);
/* Temporally blocked Chebyshev loop, 1 <= count <= 10 iterations in one
 * pass.  No single counterpart in the kernel interface: it replaces `count`
 * turns of the Chebyshev loop of 2d/drivers/cheby_driver.c (a depth-1 halo
 * update of u, then run_cheby_iterate, 2d/kernel_interface.h:48, with
 * alphas[i], betas[i]) for a chunk whose four faces are external.  u and p
 * are read from *_in and written to *_out (interiors bit-exact with the
 * reference's in-place results, halo rings carried over); r receives the
 * last iteration's r; w is not written (nothing reads it before cg_init or
 * cg_calc_w rewrite it).  Needs an even x and 16-byte aligned fields. */
int tlhip2d_cheby_iterations(int x, int y, int halo_depth, int count,
        const double* alphas, const double* betas, const int* neighbours,
        const double* u_in, double* u_out, const double* u0, const double* p_in,
        double* p_out, double* r, const double* kx, const double* ky);
/* The same loop on FP32 fields (the Chebyshev sweeps of the mixed-precision
 * solve, where u is the correction and u0 the residual it corrects): rows
 * of `pitch` floats (a multiple of 4, >= x).  Coefficients are rounded to
 * FP32; A u as in tlhip2d_ppcg_inner_iterations_f32 (fused multiply-adds),
 * p = fma(alpha, p, beta*r).  rrn: null, or (host) the squared 2-norm,
 * accumulated in FP64, of the interior of the r written, formed in the same
 * launch (no reference counterpart: the driver's calculate_2norm,
 * 2d/drivers/cheby_driver.c:128-139). */
int tlhip2d_cheby_iterations_f32(int x, int y, int halo_depth, int pitch,
        int count, const double* alphas, const double* betas,
        const int* neighbours, const float* u_in, float* u_out, const float* u0,
        const float* p_in, float* p_out, float* r, const float* kx,
        const float* ky, double* rrn);
/* Tuning of the fused PPCG kernels: output rows per tile (<= 0: default,
 * the height whose block count fills ~0.88 of what the device holds at
 * once, from the kernel's occupancy and the CU count) and prefetch distance
 * in rows (0-3; < 0: default 1; the Chebyshev loop uses 0 or 1). */
int tlhip_set_fused_tuning(int tile_rows, int prefetch);
/* Which kernel the FP64 fused PPCG and Chebyshev loops launch: 0 (default)
 * keeps each row's diagonal, kx and ky in a per-wave LDS ring (prefetch
 * fixed at one row); 1 keeps them in register windows (the prefetch setting
 * applies).  Results are bit-identical.  The FP32 loops always run the ring
 * kernel. */
int tlhip_set_fused_variant(int variant);
/* Tuning of the LDS-ring PPCG kernel: its tiles (one wave each) number
 * about fill x the waves the device holds at once (<= 0: default 0.95,
 * measured best at 4096^2 among 0.88, 0.95 and 1.0). */
int tlhip_set_ring_fill(double fill);
/* 2d/kernel_interface.h:64 run_copy_u */
int tlhip2d_copy_u(int x, int y, int halo_depth, double* u0, const double* u);
/* 2d/kernel_interface.h:66 run_calculate_residual */
int tlhip2d_calculate_residual(int x, int y, int halo_depth, const double* u,
        const double* u0, double* r, const double* kx, const double* ky);
/* 2d/kernel_interface.h:68 run_calculate_2norm */
int tlhip2d_calculate_2norm(int x, int y, int halo_depth,
        const double* buffer, double* norm);
/* 2d/kernel_interface.h:70 run_finalise */
int tlhip2d_finalise(int x, int y, int halo_depth, double* energy,
        const double* density, const double* u);

/* ------------------------------------------------------------ 3D */
/* 3d/kernel_interface.h:9 run_set_chunk_data */
int tlhip3d_set_chunk_data(int x, int y, int z, int halo_depth, int left,
        int bottom, int back, double grid_x_min, double grid_y_min,
        double grid_z_min, double dx, double dy, double dz, double* cell_x,
        double* cell_y, double* cell_z, double* vertex_x, double* vertex_y,
        double* vertex_z, double* volume, double* x_area, double* y_area,
        double* z_area);
/* 3d/kernel_interface.h:11 run_set_chunk_state */
int tlhip3d_set_chunk_state(int x, int y, int z, const double* vertex_x,
        const double* vertex_y, const double* vertex_z, const double* cell_x,
        const double* cell_y, const double* cell_z, double* density,
        double* energy0, double* u, int num_states, const tlhip_state* states);
/* 3d/kernel_interface.h:17 run_local_halos (neighbours[6], faces 3d/shared.h:24-30) */
int tlhip3d_local_halos(int x, int y, int z, int depth, int halo_depth,
        const int* neighbours, const bool* fields, double* density,
        double* energy0, double* energy, double* u, double* p, double* sd);
/* 3d/kernel_interface.h:19 run_pack_or_unpack (buffer: host memory) */
int tlhip3d_pack_or_unpack(int x, int y, int z, int depth, int halo_depth,
        int face, bool pack, double* field, double* buffer);
/* 3d/kernel_interface.h:22 run_store_energy */
int tlhip3d_store_energy(int x, int y, int z, const double* energy0,
        double* energy);
/* 3d/kernel_interface.h:24 run_field_summary */
int tlhip3d_field_summary(int x, int y, int z, int halo_depth,
        const double* volume, const double* density, const double* energy0,
        const double* u, double* vol, double* mass, double* ie, double* temp);
/* 3d/kernel_interface.h:29 run_cg_solver_init */
int tlhip3d_cg_solver_init(int x, int y, int z, int halo_depth,
        int coefficient, double rx, double ry, double rz, double* rro,
        const double* density, const double* energy, double* u, double* p,
        double* r, double* w, double* kx, double* ky, double* kz);
/* 3d/kernel_interface.h:32 run_cg_solver_calc_w */
int tlhip3d_cg_solver_calc_w(int x, int y, int z, int halo_depth, double* pw,
        const double* p, double* w, const double* kx, const double* ky,
        const double* kz);
/* No single counterpart: one CG step's calc_p (3d/kernel_interface.h:36
 * run_cg_solver_calc_p, called at 3d/drivers/cg_solver_driver.c:116), the
 * depth-1 halo update of p and the next step's calc_w (:32, called at
 * cg_solver_driver.c:79) in one pass, p_in -> p_out (the halo ring copied)
 * and w, bit-exact with the three; pw accumulated (+=).  The native driver
 * does not use it: inside a 3D solve it measured no faster than the two. */
int tlhip3d_cg_solver_calc_pw(int x, int y, int z, int halo_depth, double beta, double* pw,
        const double* p_in, double* p_out, const double* r, double* w, const double* kx,
        const double* ky, const double* kz);
/* 3d/kernel_interface.h:34 run_cg_solver_calc_ur */
int tlhip3d_cg_solver_calc_ur(int x, int y, int z, int halo_depth,
        double alpha, double* rrn, double* u, const double* p, double* r,
        const double* w);
/* 3d/kernel_interface.h:36 run_cg_solver_calc_p */
int tlhip3d_cg_solver_calc_p(int x, int y, int z, int halo_depth, double beta,
        double* p, const double* r);
/* 3d/kernel_interface.h:40 run_cheby_solver_init (in place) */
int tlhip3d_cheby_solver_init(int x, int y, int z, int halo_depth,
        double theta, double* u, const double* u0, double* p, double* r,
        double* w, const double* kx, const double* ky, const double* kz);
/* 3d/kernel_interface.h:42 run_cheby_solver_iterate, fused; u_out != u_in */
int tlhip3d_cheby_solver_iterate(int x, int y, int z, int halo_depth,
        double alpha, double beta, const double* u_in, double* u_out,
        const double* u0, double* p, double* r, double* w, const double* kx,
        const double* ky, const double* kz);
/* 3d/kernel_interface.h:46 run_jacobi_solver_init */
int tlhip3d_jacobi_solver_init(int x, int y, int z, int halo_depth,
        int coefficient, double rx, double ry, double rz,
        const double* density, const double* energy, double* u0, double* u,
        double* kx, double* ky, double* kz);
/* 3d/kernel_interface.h:48 run_jacobi_solver_iterate (error is assigned) */
int tlhip3d_jacobi_solver_iterate(int x, int y, int z, int halo_depth,
        double* error, const double* kx, const double* ky, const double* kz,
        const double* u0, double* u, double* r);
/* 3d/kernel_interface.h:52 run_ppcg_init */
int tlhip3d_ppcg_init(int x, int y, int z, int halo_depth, double theta,
        const double* r, double* sd);
/* 3d/kernel_interface.h:54 run_ppcg_inner_iteration, fused; sd_out != sd_in */
int tlhip3d_ppcg_inner_iteration(int x, int y, int z, int halo_depth,
        double alpha, double beta, double* u, double* r, const double* kx,
        const double* ky, const double* kz, const double* sd_in,
        double* sd_out);
/* Two PPCG inner steps in one pass (count must be 2).  No single
 * counterpart in the kernel interface: it replaces two turns of the inner
 * loop of 3d/drivers/ppcg_solver_driver.c:170-183 (a depth-1 halo update of
 * sd, then run_ppcg_inner_iteration, 3d/kernel_interface.h:54, with
 * alphas[i], betas[i]) for a chunk whose six faces are external.  sd, r
 * and u are read from *_in and written to *_out: interiors bit-exact with
 * the reference's in-place results, halo rings carried over.  Needs an
 * even x and 16-byte aligned fields. */
/* Output rows per block of that pass: 4, 6, 8 (default) or 14, the block's
 * rows sharing level 1 through LDS; 0 = each wave forms the three level-1
 * rows it needs itself.  Results are identical. */
int tlhip_set_pair_rows(int rows);
int tlhip3d_ppcg_inner_iterations(int x, int y, int z, int halo_depth, int count,
        const double* alphas, const double* betas, const int* neighbours,
        const double* sd_in, double* sd_out, const double* r_in, double* r_out,
        const double* u_in, double* u_out, const double* kx, const double* ky,
        const double* kz);
/* 3d/kernel_interface.h:58 run_solver_copy_u */
int tlhip3d_solver_copy_u(int x, int y, int z, int halo_depth, double* u0,
        const double* u);
/* 3d/kernel_interface.h:60 run_calculate_residual */
int tlhip3d_calculate_residual(int x, int y, int z, int halo_depth,
        const double* u, const double* u0, double* r, const double* kx,
        const double* ky, const double* kz);
/* 3d/kernel_interface.h:62 run_calculate_2norm */
int tlhip3d_calculate_2norm(int x, int y, int z, int halo_depth,
        const double* buffer, double* norm);
/* 3d/kernel_interface.h:64 run_solver_finalise */
int tlhip3d_solver_finalise(int x, int y, int z, int halo_depth,
        double* energy, const double* density, const double* u);

/* ------------------------------------------------------------ timestep driver
 * A native host driver over the entry points above, for callers without the
 * reference host (benchmarks, tests, Python).  It follows the single-chunk,
 * NO_MPI control flow of 2d/diffuse.c:22-80 and 2d/drivers/ (3D:
 * 3d/diffuse.c, 3d/drivers/), so that iteration counts and results match
 * the reference host driving the same kernels.  No counterpart in the
 * reference kernel interface. */

/* Same layout as oracle/oracle.h tl_settings. solver: 0 Jacobi, 1 CG,
 * 2 Chebyshev, 3 PPCG (2d/settings.h:48-54). */
typedef struct {
    int dims;
    int x_cells, y_cells, z_cells;
    double xmin, ymin, zmin, xmax, ymax, zmax;
    double dt_init;
    int end_step;
    int max_iters;
    double eps;
    int solver;
    int coefficient;
    int presteps;
    int ppcg_inner_steps;
    double eps_lim;
    int error_switch;
    int check_result;
    int summary_frequency;
    int halo_depth;
} tlhip_settings;

/* Per-timestep report: the counts the reference prints (2d/diffuse.c:70-78,
 * 2d/drivers/cg_driver.c:27) plus how many sweeps actually ran. */
typedef struct {
    int printed_cg_iters;
    int printed_inner_iters;
    int est_iters;
    int n_cg_steps;
    int n_cheby_steps;
    int n_ppcg_outer;
    int n_ppcg_inner;
    int n_jacobi;
    double error;
    double exact_error;
    double eigmin, eigmax;
    int summary_done;
    double vol, mass, ie, temp;
    /* Not printed by the reference: sqrt(||u0 - A u||^2 / ||u0||^2) in FP64
     * at the last true-residual evaluation (relative-residual stop or mixed
     * precision; 0 otherwise) and how many such evaluations the solve made. */
    double rel_residual;
    int n_true_residuals;
} tlhip_step_report;

typedef struct tlhip_problem tlhip_problem;

/* Whether the driver (2D) runs a CG/PPCG step's calc_p fused into the next
 * step's calc_w (tlhip2d_cg_calc_pw) or as calc_p + halo update + calc_w.
 * Results are identical either way.  Default on: at 4096^2 on MI355X the
 * fused pass (48 B/cell, 0.162 ms) takes 9% less than the two (0.177 ms). */
int tlhip_set_cg_fusion(int on);
/* Whether the driver's CG steps (CG, and the CG phase of Chebyshev and
 * PPCG; 2D and 3D, halo depth >= 2, even x, with tlhip_set_cg_fusion on)
 * run as the cond pair (tlhip2d/3d_cg_calc_pw_cond without storing w, then
 * tlhip2d/3d_cg_calc_ur_cond) and 2D PPCG's calc_p + calc_w reads cond in
 * place of kx and ky.  u, p and r
 * are the reference's expressions either way; rrn is summed in another
 * order, so results agree to rounding.  Default on: at 4096^2 on MI355X a
 * CG step's two passes take 0.284 ms against 0.333 ms (32 + 48 B/cell
 * against 48 + 48).  No counterpart in the reference. */
int tlhip_set_cg_cond(int on);
/* Whether the driver runs CG, Jacobi and PPCG outer iterations (2D, 3D)
 * device-resident: alpha = rro/pw, beta = rrn/rro, the
 * Jacobi error and the stop and switch tests of cg_driver.c /
 * cheby_driver.c / ppcg_driver.c / jacobi_driver.c are evaluated on the
 * device, the host keeps several iterations enqueued and reads one report
 * per iteration, and kernels of iterations past the stop return at once.
 * Jacobi then sweeps between u and a partner buffer instead of copying u
 * into r every iteration.  Results are identical either way (same values,
 * same operations in the same order); on = no host round trip between the
 * kernels of an iteration.  Default on.  No counterpart in the reference:
 * its drivers take pw and rrn to the host (2d/drivers/cg_driver.c:78-99). */
int tlhip_set_resident(int on);
/* Whether the resident loops on small grids (up to 2.5M cells with halo,
 * where an iteration is a few launches of a few microseconds) replay their
 * iterations from a HIP graph: after the first iteration of a run, two
 * iterations are captured once and replayed with one host call per two
 * iterations.  Results are identical either way.  Default off: on MI355X
 * (ROCm 7.2) the replays measured 5-9% slower than launch-by-launch
 * enqueueing.  No counterpart in the reference. */
int tlhip_set_graphs(int on);
/* Steps per fused PPCG launch used by the driver (1..10; 0 or 1 = one step
 * per launch through tlhip2d_ppcg_inner_iteration).  Default 10: the
 * reference deck's 10 inner steps in one launch (FP64, and FP32 in the
 * mixed-precision solve).  In 3D any value above 1 runs the inner steps two
 * per launch (tlhip3d_ppcg_inner_iterations, even x, every face external),
 * an odd last step on its own. */
int tlhip_set_ppcg_fusion(int steps);
/* Whether the 2D driver's fused PPCG launches form the 2-norm of r that
 * ends each outer iteration (2d/drivers/ppcg_driver.c:121-131) in the last
 * launch instead of a separate 2-norm pass (default 1; FP64 with the
 * LDS-ring kernel, tlhip_set_fused_variant(0)).  The sum then runs in
 * another order, so rrn, and through beta the fields, differ in the last
 * bits from the unfused path's; 0 restores bit-identity with it.  No
 * counterpart in the reference. */
int tlhip_set_ppcg_fused_norm(int on);
/* Chebyshev iterations per fused launch used by the 2D driver with the
 * reference stop test (0 or 1 = one tlhip2d_cheby_iterate per iteration).
 * Default 10; results are identical either way. */
int tlhip_set_cheby_fusion(int steps);

/* How tlhip_problem_step solves (no counterpart in the reference, which
 * always solves in FP64 and stops on tl_eps):
 *   precision 0  FP64 throughout, the reference's arithmetic (default);
 *   precision 1  FP32 CG / PPCG sweeps (FP32 kx, ky, p, w, r, sd and the
 *                correction to u) with FP64-accumulated dot products, inside
 *                an FP64 loop that adds the correction to u and replaces r by
 *                the true residual u0 - A u whenever ||r|| has fallen 10x;
 *                2D, CG or PPCG only;
 *   rel_residual > 0  stop once ||u0 - A u||_2 <= rel_residual ||u0||_2,
 *                confirmed by the FP64 residual, instead of the eps test.
 * With precision 1 and rel_residual 0 the eps test applies to the true
 * ||u0 - A u||^2. */
int tlhip_problem_set_solve(tlhip_problem* p, int precision, double rel_residual);
/* Reverse-mode gradient through `steps` implicit timesteps (no counterpart
 * in the reference, which only runs forward).  Runs the steps from the
 * current state as tlhip_problem_step does (FP64 solve; the state
 * advances; reports[i] filled when reports is not NULL), keeping u after
 * each solve, then returns
 *   *loss = sum over the interior of (u_N - u_target)^2 and
 *   grad_log_density[c] = dL/d ln(density[c]) (0 on the halo),
 * through the dependence of the conduction coefficients (2d/c_kernels/omp3/
 * cg.c:49-63, 3d .../cg_solver.c) and of u_0 = energy0 * density on the
 * density.  Each backward step solves A mu = lambda with the forward
 * solver and stopping test.  u_target and grad_log_density are host arrays
 * of the whole chunk; u, u0, r, p and w are overwritten. */
int tlhip_problem_adjoint(tlhip_problem* p, int steps, const double* u_target,
        double* loss, double* grad_log_density, tlhip_step_report* reports);
/* Allocate a chunk on the device and initialise it as
 * 2d/initialise.c:11-31 does (set_chunk_data, set_chunk_state, halo prime,
 * store_energy). */
int tlhip_problem_create(const tlhip_settings* settings, const tlhip_state* states,
        int num_states, tlhip_problem** out);
/* Replace density and energy0 (host arrays of the full chunk incl. halo) and
 * redo u = e*rho, the halo prime and store_energy. */
int tlhip_problem_inject(tlhip_problem* p, const double* density,
        const double* energy0);
/* Device fields of a chunk that the caller allocated and initialised (the
 * hip_cdna4 suite's run_kernel_initialise, run_set_chunk_data and
 * run_set_chunk_state, then initialise.c's halo prime and store_energy).
 * kz and z_area are used in 3D only.  u_alt and sd_alt are the partners of u
 * and sd (the suite's chunk extension). */
typedef struct {
    double *density0, *density, *energy0, *energy, *u, *u0, *p, *r, *mi, *w;
    double *kx, *ky, *kz, *sd, *u_alt, *sd_alt;
    double *volume, *x_area, *y_area, *z_area;
} tlhip_chunk_fields;
/* A problem over an existing single chunk whose faces are all external, for
 * the suite's diffuse_overload (the reference's DIFFUSE_OVERLOAD hook,
 * 2d/main.c:31-36, 3d/main.c:30-35), which runs the reference's timestep
 * loop with the native driver (device-resident loops, fused sweeps) in place
 * of 2d/drivers/.  No counterpart in the reference's kernel interface.  The
 * fields stay the caller's: tlhip_problem_destroy does not free them, and
 * after each tlhip_problem_step (and tlhip_problem_adjoint) u, sd, r and p
 * hold their values at the caller's pointers again (u_alt and sd_alt are
 * scratch).  The problem allocates the partners of r and p that the 2D
 * driver needs. */
int tlhip_problem_adopt(const tlhip_settings* settings, const tlhip_chunk_fields* fields,
        tlhip_problem** out);
/* One timestep (tt 0-based), as 2d/diffuse.c:22-80. */
int tlhip_problem_step(tlhip_problem* p, int tt, tlhip_step_report* report);
/* vol, mass, ie, temp of the current state (2d/drivers/field_summary_driver.c). */
int tlhip_problem_field_summary(tlhip_problem* p, double* out4);
/* Number of cells of one field (incl. halo). */
int tlhip_problem_field_len(tlhip_problem* p, long* n);
/* Copy a field to host memory.  id: 0 density, 1 energy0, 2 energy, 3 u,
 * 4 u0, 5 p, 6 r, 7 w, 8 kx, 9 ky, 10 kz, 11 sd. */
int tlhip_problem_get_field(tlhip_problem* p, int id, double* host_out);
/* Device pointer of a field (same ids), valid until the next step. */
int tlhip_problem_field_ptr(tlhip_problem* p, int id, void** dev_out);
int tlhip_problem_destroy(tlhip_problem* p);

#ifdef __cplusplus
}
#endif
#endif
