// hip_cdna4 kernel suite for TeaLeaf 3D: the reference kernel interface
// (3d/kernel_interface.h:9-65) implemented over libtealeaf_hip.so.
//
// Compiled inside the reference tree as 3d/c_kernels/hip_cdna4/ with the
// reference's own headers (its chunk_extension.h is copied to build/ by
// 3d/Makefile:38).  Fields are device pointers; the CG / Chebyshev
// coefficient arrays are host memory.  Failures call the reference's die().
#include <stdlib.h>
#include "kernel_interface.h"
#include "tealeaf_hip.h"

#define TL_CHECK(call)                                                   \
    do {                                                                 \
        if ((call) != 0) die(__LINE__, __FILE__, "%s\n", tlhip_last_error()); \
    } while (0)

static double* dev_buffer(long n)
{
    void* p = NULL;
    TL_CHECK(tlhip_alloc(sizeof(double) * (size_t)(n > 0 ? n : 1), &p));
    return (double*)p;
}

static double* host_buffer(int n)
{
    double* p = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    if (!p) die(__LINE__, __FILE__, "Error allocating host buffer\n");
    return p;
}

// Sizes as 3d/c_kernels/omp3/kernel_init_chunk.c:29-86.
void run_kernel_initialise(Chunk* chunk, Settings* settings)
{
    TL_CHECK(tlhip_init(settings->rank));
    const int x = chunk->x, y = chunk->y, z = chunk->z;
    const long n = (long)x * y * z;
    chunk->density0 = dev_buffer(n);
    chunk->density = dev_buffer(n);
    chunk->energy0 = dev_buffer(n);
    chunk->energy = dev_buffer(n);
    chunk->vec_u = dev_buffer(n);
    chunk->vec_u0 = dev_buffer(n);
    chunk->vec_p = dev_buffer(n);
    chunk->vec_r = dev_buffer(n);
    chunk->vec_mi = dev_buffer(n);
    chunk->vec_w = dev_buffer(n);
    chunk->vec_kx = dev_buffer(n);
    chunk->vec_ky = dev_buffer(n);
    chunk->vec_kz = dev_buffer(n);
    chunk->vec_sd = dev_buffer(n);
    chunk->vec_z = dev_buffer(n);
    chunk->volume = dev_buffer(n);
    chunk->x_area = dev_buffer((long)(x + 1) * y * z);
    chunk->y_area = dev_buffer((long)x * (y + 1) * z);
    chunk->z_area = dev_buffer((long)x * y * (z + 1));
    chunk->cell_x = dev_buffer(x);
    chunk->cell_y = dev_buffer(y);
    chunk->cell_z = dev_buffer(z);
    chunk->cell_dx = dev_buffer(x);
    chunk->cell_dy = dev_buffer(y);
    chunk->cell_dz = dev_buffer(z);
    chunk->vertex_dx = dev_buffer(x + 1);
    chunk->vertex_dy = dev_buffer(y + 1);
    chunk->vertex_dz = dev_buffer(z + 1);
    chunk->vertex_x = dev_buffer(x + 1);
    chunk->vertex_y = dev_buffer(y + 1);
    chunk->vertex_z = dev_buffer(z + 1);
    chunk->cg_alphas = host_buffer(settings->max_iters);
    chunk->cg_betas = host_buffer(settings->max_iters);
    /* PPCG writes ppcg_inner_steps Chebyshev coefficients here; the omp3
     * suite sizes these max_iters only (kernel_initialise.c:68). */
    chunk->cheby_alphas = host_buffer(settings->max_iters > settings->ppcg_inner_steps
                                          ? settings->max_iters : settings->ppcg_inner_steps);
    chunk->cheby_betas = host_buffer(settings->max_iters > settings->ppcg_inner_steps
                                         ? settings->max_iters : settings->ppcg_inner_steps);
    chunk->ext->u_alt = dev_buffer(n);
    chunk->ext->sd_alt = dev_buffer(n);
}

void run_set_chunk_data(Chunk* chunk, Settings* settings)
{
    TL_CHECK(tlhip3d_set_chunk_data(chunk->x, chunk->y, chunk->z,
            settings->halo_depth, chunk->left, chunk->bottom, chunk->back,
            settings->grid_x_min, settings->grid_y_min, settings->grid_z_min,
            settings->dx, settings->dy, settings->dz, chunk->cell_x,
            chunk->cell_y, chunk->cell_z, chunk->vertex_x, chunk->vertex_y,
            chunk->vertex_z, chunk->volume, chunk->x_area, chunk->y_area,
            chunk->z_area));
}

void run_set_chunk_state(Chunk* chunk, Settings* settings, State* states)
{
    const int n = settings->num_states;
    tlhip_state* st = (tlhip_state*)calloc((size_t)(n > 0 ? n : 1), sizeof(tlhip_state));
    for (int s = 0; s < n; ++s) {
        st[s].density = states[s].density;
        st[s].energy = states[s].energy;
        st[s].x_min = states[s].x_min;
        st[s].y_min = states[s].y_min;
        st[s].z_min = states[s].z_min;
        st[s].x_max = states[s].x_max;
        st[s].y_max = states[s].y_max;
        st[s].z_max = states[s].z_max;
        st[s].radius = states[s].radius;
        st[s].geometry = (int)states[s].geometry;
        st[s].defined = states[s].defined;
    }
    const int rc = tlhip3d_set_chunk_state(chunk->x, chunk->y, chunk->z,
            chunk->vertex_x, chunk->vertex_y, chunk->vertex_z, chunk->cell_x,
            chunk->cell_y, chunk->cell_z, chunk->density, chunk->energy0,
            chunk->vec_u, n, st);
    free(st);
    TL_CHECK(rc);
}

void run_local_halos(Chunk* chunk, Settings* settings, int depth)
{
    START_PROFILING(settings->kernel_profile);
    TL_CHECK(tlhip3d_local_halos(chunk->x, chunk->y, chunk->z, depth,
            settings->halo_depth, chunk->neighbours, settings->fields_to_exchange,
            chunk->density, chunk->energy0, chunk->energy, chunk->vec_u,
            chunk->vec_p, chunk->vec_sd));
    STOP_PROFILING(settings->kernel_profile, __func__);
}

void run_pack_or_unpack(Chunk* chunk, Settings* settings, int depth, int face,
        bool pack, double* field, double* buffer)
{
    START_PROFILING(settings->kernel_profile);
    TL_CHECK(tlhip3d_pack_or_unpack(chunk->x, chunk->y, chunk->z, depth,
            settings->halo_depth, face, pack, field, buffer));
    STOP_PROFILING(settings->kernel_profile, __func__);
}

void run_store_energy(Chunk* chunk, Settings* settings)
{
    START_PROFILING(settings->kernel_profile);
    TL_CHECK(tlhip3d_store_energy(chunk->x, chunk->y, chunk->z, chunk->energy0,
            chunk->energy));
    STOP_PROFILING(settings->kernel_profile, __func__);
}

void run_field_summary(Chunk* chunk, Settings* settings, double* vol,
        double* mass, double* ie, double* temp)
{
    START_PROFILING(settings->kernel_profile);
    TL_CHECK(tlhip3d_field_summary(chunk->x, chunk->y, chunk->z,
            settings->halo_depth, chunk->volume, chunk->density, chunk->energy0,
            chunk->vec_u, vol, mass, ie, temp));
    STOP_PROFILING(settings->kernel_profile, __func__);
}

void run_cg_solver_init(Chunk* chunk, Settings* settings, double rx, double ry,
        double rz, double* rro)
{
    START_PROFILING(settings->kernel_profile);
    TL_CHECK(tlhip3d_cg_solver_init(chunk->x, chunk->y, chunk->z,
            settings->halo_depth, settings->coefficient, rx, ry, rz, rro,
            chunk->density, chunk->energy, chunk->vec_u, chunk->vec_p,
            chunk->vec_r, chunk->vec_w, chunk->vec_kx, chunk->vec_ky,
            chunk->vec_kz));
    STOP_PROFILING(settings->kernel_profile, __func__);
}

void run_cg_solver_calc_w(Chunk* chunk, Settings* settings, double* pw)
{
    START_PROFILING(settings->kernel_profile);
    TL_CHECK(tlhip3d_cg_solver_calc_w(chunk->x, chunk->y, chunk->z,
            settings->halo_depth, pw, chunk->vec_p, chunk->vec_w, chunk->vec_kx,
            chunk->vec_ky, chunk->vec_kz));
    STOP_PROFILING(settings->kernel_profile, __func__);
}

void run_cg_solver_calc_ur(Chunk* chunk, Settings* settings, double alpha,
        double* rrn)
{
    START_PROFILING(settings->kernel_profile);
    TL_CHECK(tlhip3d_cg_solver_calc_ur(chunk->x, chunk->y, chunk->z,
            settings->halo_depth, alpha, rrn, chunk->vec_u, chunk->vec_p,
            chunk->vec_r, chunk->vec_w));
    STOP_PROFILING(settings->kernel_profile, __func__);
}

void run_cg_solver_calc_p(Chunk* chunk, Settings* settings, double beta)
{
    START_PROFILING(settings->kernel_profile);
    TL_CHECK(tlhip3d_cg_solver_calc_p(chunk->x, chunk->y, chunk->z,
            settings->halo_depth, beta, chunk->vec_p, chunk->vec_r));
    STOP_PROFILING(settings->kernel_profile, __func__);
}

void run_cheby_solver_init(Chunk* chunk, Settings* settings)
{
    START_PROFILING(settings->kernel_profile);
    TL_CHECK(tlhip3d_cheby_solver_init(chunk->x, chunk->y, chunk->z,
            settings->halo_depth, chunk->theta, chunk->vec_u, chunk->vec_u0,
            chunk->vec_p, chunk->vec_r, chunk->vec_w, chunk->vec_kx,
            chunk->vec_ky, chunk->vec_kz));
    STOP_PROFILING(settings->kernel_profile, __func__);
}

void run_cheby_solver_iterate(Chunk* chunk, Settings* settings, double alpha,
        double beta)
{
    START_PROFILING(settings->kernel_profile);
    TL_CHECK(tlhip3d_cheby_solver_iterate(chunk->x, chunk->y, chunk->z,
            settings->halo_depth, alpha, beta, chunk->vec_u, chunk->ext->u_alt,
            chunk->vec_u0, chunk->vec_p, chunk->vec_r, NULL, chunk->vec_kx,
            chunk->vec_ky, chunk->vec_kz));
    double* t = chunk->vec_u;
    chunk->vec_u = chunk->ext->u_alt;
    chunk->ext->u_alt = t;
    STOP_PROFILING(settings->kernel_profile, __func__);
}

void run_jacobi_solver_init(Chunk* chunk, Settings* settings, double rx,
        double ry, double rz)
{
    START_PROFILING(settings->kernel_profile);
    TL_CHECK(tlhip3d_jacobi_solver_init(chunk->x, chunk->y, chunk->z,
            settings->halo_depth, settings->coefficient, rx, ry, rz,
            chunk->density, chunk->energy, chunk->vec_u0, chunk->vec_u,
            chunk->vec_kx, chunk->vec_ky, chunk->vec_kz));
    STOP_PROFILING(settings->kernel_profile, __func__);
}

void run_jacobi_solver_iterate(Chunk* chunk, Settings* settings, double* error)
{
    START_PROFILING(settings->kernel_profile);
    TL_CHECK(tlhip3d_jacobi_solver_iterate(chunk->x, chunk->y, chunk->z,
            settings->halo_depth, error, chunk->vec_kx, chunk->vec_ky,
            chunk->vec_kz, chunk->vec_u0, chunk->vec_u, chunk->vec_r));
    STOP_PROFILING(settings->kernel_profile, __func__);
}

void run_ppcg_init(Chunk* chunk, Settings* settings)
{
    START_PROFILING(settings->kernel_profile);
    TL_CHECK(tlhip3d_ppcg_init(chunk->x, chunk->y, chunk->z, settings->halo_depth,
            chunk->theta, chunk->vec_r, chunk->vec_sd));
    STOP_PROFILING(settings->kernel_profile, __func__);
}

void run_ppcg_inner_iteration(Chunk* chunk, Settings* settings, double alpha,
        double beta)
{
    START_PROFILING(settings->kernel_profile);
    TL_CHECK(tlhip3d_ppcg_inner_iteration(chunk->x, chunk->y, chunk->z,
            settings->halo_depth, alpha, beta, chunk->vec_u, chunk->vec_r,
            chunk->vec_kx, chunk->vec_ky, chunk->vec_kz, chunk->vec_sd,
            chunk->ext->sd_alt));
    double* t = chunk->vec_sd;
    chunk->vec_sd = chunk->ext->sd_alt;
    chunk->ext->sd_alt = t;
    STOP_PROFILING(settings->kernel_profile, __func__);
}

void run_solver_copy_u(Chunk* chunk, Settings* settings)
{
    START_PROFILING(settings->kernel_profile);
    TL_CHECK(tlhip3d_solver_copy_u(chunk->x, chunk->y, chunk->z,
            settings->halo_depth, chunk->vec_u0, chunk->vec_u));
    STOP_PROFILING(settings->kernel_profile, __func__);
}

void run_calculate_residual(Chunk* chunk, Settings* settings)
{
    START_PROFILING(settings->kernel_profile);
    TL_CHECK(tlhip3d_calculate_residual(chunk->x, chunk->y, chunk->z,
            settings->halo_depth, chunk->vec_u, chunk->vec_u0, chunk->vec_r,
            chunk->vec_kx, chunk->vec_ky, chunk->vec_kz));
    STOP_PROFILING(settings->kernel_profile, __func__);
}

void run_calculate_2norm(Chunk* chunk, Settings* settings, double* buffer,
        double* norm)
{
    START_PROFILING(settings->kernel_profile);
    TL_CHECK(tlhip3d_calculate_2norm(chunk->x, chunk->y, chunk->z,
            settings->halo_depth, buffer, norm));
    STOP_PROFILING(settings->kernel_profile, __func__);
}

void run_solver_finalise(Chunk* chunk, Settings* settings)
{
    START_PROFILING(settings->kernel_profile);
    TL_CHECK(tlhip3d_solver_finalise(chunk->x, chunk->y, chunk->z,
            settings->halo_depth, chunk->energy, chunk->density, chunk->vec_u));
    STOP_PROFILING(settings->kernel_profile, __func__);
}
