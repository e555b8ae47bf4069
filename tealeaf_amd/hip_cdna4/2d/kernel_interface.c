// hip_cdna4 kernel suite for TeaLeaf 2D: the reference kernel interface
// (2d/kernel_interface.h:13-71) implemented over libtealeaf_hip.so.
//
// Compiled inside the reference tree as 2d/c_kernels/hip_cdna4/ with the
// reference's own headers; the host (diffuse.c, drivers/, chunk.c) is used
// unchanged.  Each run_* forwards the chunk's sizes, settings and field
// pointers (device memory) to the matching tlhip2d_* entry point and calls the
// reference's die() when it fails, which is the reference's error behaviour.
#include <stdlib.h>
#include "kernel_interface.h"
#include "tealeaf_hip.h"

#define TL_CHECK(call)                                                   \
    do {                                                                 \
        if ((call) != 0) die(__LINE__, __FILE__, "%s\n", tlhip_last_error()); \
    } while (0)

static double* dev_buffer(int n)
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

// Field ownership follows 2d/c_kernels/omp3/kernel_initialise.c:27-70: the
// suite allocates every field (here on the device) with the same sizes; the
// CG / Chebyshev coefficient arrays stay in host memory because the drivers
// write and read them directly (2d/drivers/cg_driver.c:93,111).
void run_kernel_initialise(Chunk* chunk, Settings* settings)
{
    TL_CHECK(tlhip_init(settings->rank));
    print_and_log(settings,
            "Performing this solve with the hip_cdna4 (gfx950) %s solver\n",
            settings->solver_name);
    const int x = chunk->x, y = chunk->y, n = x * y;
    chunk->density0 = dev_buffer(n);
    chunk->density = dev_buffer(n);
    chunk->energy0 = dev_buffer(n);
    chunk->energy = dev_buffer(n);
    chunk->u = dev_buffer(n);
    chunk->u0 = dev_buffer(n);
    chunk->p = dev_buffer(n);
    chunk->r = dev_buffer(n);
    chunk->mi = dev_buffer(n);
    chunk->w = dev_buffer(n);
    chunk->kx = dev_buffer(n);
    chunk->ky = dev_buffer(n);
    chunk->sd = dev_buffer(n);
    chunk->volume = dev_buffer(n);
    chunk->x_area = dev_buffer((x + 1) * y);
    chunk->y_area = dev_buffer(x * (y + 1));
    chunk->cell_x = dev_buffer(x);
    chunk->cell_y = dev_buffer(y);
    chunk->cell_dx = dev_buffer(x);
    chunk->cell_dy = dev_buffer(y);
    chunk->vertex_dx = dev_buffer(x + 1);
    chunk->vertex_dy = dev_buffer(y + 1);
    chunk->vertex_x = dev_buffer(x + 1);
    chunk->vertex_y = dev_buffer(y + 1);
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

void run_kernel_finalise(Chunk* chunk, Settings* settings)
{
    (void)settings;
    double* dev[] = {chunk->density0, chunk->density, chunk->energy0,
        chunk->energy, chunk->u, chunk->u0, chunk->p, chunk->r, chunk->mi,
        chunk->w, chunk->kx, chunk->ky, chunk->sd, chunk->volume,
        chunk->x_area, chunk->y_area, chunk->cell_x, chunk->cell_y,
        chunk->cell_dx, chunk->cell_dy, chunk->vertex_dx, chunk->vertex_dy,
        chunk->vertex_x, chunk->vertex_y, chunk->ext->u_alt, chunk->ext->sd_alt};
    for (size_t i = 0; i < sizeof(dev) / sizeof(dev[0]); ++i)
        TL_CHECK(tlhip_free(dev[i]));
    free(chunk->cg_alphas);
    free(chunk->cg_betas);
    free(chunk->cheby_alphas);
    free(chunk->cheby_betas);
}

void run_set_chunk_data(Chunk* chunk, Settings* settings)
{
    TL_CHECK(tlhip2d_set_chunk_data(chunk->x, chunk->y, settings->halo_depth,
            chunk->left, chunk->bottom, settings->grid_x_min,
            settings->grid_y_min, settings->dx, settings->dy, chunk->cell_x,
            chunk->cell_y, chunk->vertex_x, chunk->vertex_y, chunk->volume,
            chunk->x_area, chunk->y_area));
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
        st[s].x_max = states[s].x_max;
        st[s].y_max = states[s].y_max;
        st[s].radius = states[s].radius;
        st[s].geometry = (int)states[s].geometry;
        st[s].defined = states[s].defined;
    }
    const int rc = tlhip2d_set_chunk_state(chunk->x, chunk->y, chunk->vertex_x,
            chunk->vertex_y, chunk->cell_x, chunk->cell_y, chunk->density,
            chunk->energy0, chunk->u, n, st);
    free(st);
    TL_CHECK(rc);
}

void run_local_halos(Chunk* chunk, Settings* settings, int depth)
{
    START_PROFILING(settings->kernel_profile);
    TL_CHECK(tlhip2d_local_halos(chunk->x, chunk->y, depth, settings->halo_depth,
            chunk->neighbours, settings->fields_to_exchange, chunk->density,
            chunk->energy0, chunk->energy, chunk->u, chunk->p, chunk->sd));
    STOP_PROFILING(settings->kernel_profile, __func__);
}

void run_pack_or_unpack(Chunk* chunk, Settings* settings, int depth, int face,
        bool pack, FieldBufferType field, double* buffer)
{
    START_PROFILING(settings->kernel_profile);
    TL_CHECK(tlhip2d_pack_or_unpack(chunk->x, chunk->y, depth,
            settings->halo_depth, face, pack, field, buffer));
    STOP_PROFILING(settings->kernel_profile, __func__);
}

void run_store_energy(Chunk* chunk, Settings* settings)
{
    START_PROFILING(settings->kernel_profile);
    TL_CHECK(tlhip2d_store_energy(chunk->x, chunk->y, chunk->energy0, chunk->energy));
    STOP_PROFILING(settings->kernel_profile, __func__);
}

void run_field_summary(Chunk* chunk, Settings* settings, double* vol,
        double* mass, double* ie, double* temp)
{
    START_PROFILING(settings->kernel_profile);
    TL_CHECK(tlhip2d_field_summary(chunk->x, chunk->y, settings->halo_depth,
            chunk->volume, chunk->density, chunk->energy0, chunk->u, vol, mass,
            ie, temp));
    STOP_PROFILING(settings->kernel_profile, __func__);
}

void run_cg_init(Chunk* chunk, Settings* settings, double rx, double ry,
        double* rro)
{
    START_PROFILING(settings->kernel_profile);
    TL_CHECK(tlhip2d_cg_init(chunk->x, chunk->y, settings->halo_depth,
            settings->coefficient, rx, ry, rro, chunk->density, chunk->energy,
            chunk->u, chunk->p, chunk->r, chunk->w, chunk->kx, chunk->ky));
    STOP_PROFILING(settings->kernel_profile, __func__);
}

void run_cg_calc_w(Chunk* chunk, Settings* settings, double* pw)
{
    START_PROFILING(settings->kernel_profile);
    TL_CHECK(tlhip2d_cg_calc_w(chunk->x, chunk->y, settings->halo_depth, pw,
            chunk->p, chunk->w, chunk->kx, chunk->ky));
    STOP_PROFILING(settings->kernel_profile, __func__);
}

void run_cg_calc_ur(Chunk* chunk, Settings* settings, double alpha, double* rrn)
{
    START_PROFILING(settings->kernel_profile);
    TL_CHECK(tlhip2d_cg_calc_ur(chunk->x, chunk->y, settings->halo_depth, alpha,
            rrn, chunk->u, chunk->p, chunk->r, chunk->w));
    STOP_PROFILING(settings->kernel_profile, __func__);
}

void run_cg_calc_p(Chunk* chunk, Settings* settings, double beta)
{
    START_PROFILING(settings->kernel_profile);
    TL_CHECK(tlhip2d_cg_calc_p(chunk->x, chunk->y, settings->halo_depth, beta,
            chunk->p, chunk->r));
    STOP_PROFILING(settings->kernel_profile, __func__);
}

void run_cheby_init(Chunk* chunk, Settings* settings)
{
    START_PROFILING(settings->kernel_profile);
    TL_CHECK(tlhip2d_cheby_init(chunk->x, chunk->y, settings->halo_depth,
            chunk->theta, chunk->u, chunk->u0, chunk->p, chunk->r, chunk->w,
            chunk->kx, chunk->ky));
    STOP_PROFILING(settings->kernel_profile, __func__);
}

// One fused sweep; w is not written because nothing reads it before
// cg_init / cg_calc_w rewrite it.  u moves to the other buffer.
void run_cheby_iterate(Chunk* chunk, Settings* settings, double alpha, double beta)
{
    START_PROFILING(settings->kernel_profile);
    TL_CHECK(tlhip2d_cheby_iterate(chunk->x, chunk->y, settings->halo_depth,
            alpha, beta, chunk->u, chunk->ext->u_alt, chunk->u0, chunk->p,
            chunk->r, NULL, chunk->kx, chunk->ky));
    double* t = chunk->u;
    chunk->u = chunk->ext->u_alt;
    chunk->ext->u_alt = t;
    STOP_PROFILING(settings->kernel_profile, __func__);
}

void run_jacobi_init(Chunk* chunk, Settings* settings, double rx, double ry)
{
    START_PROFILING(settings->kernel_profile);
    TL_CHECK(tlhip2d_jacobi_init(chunk->x, chunk->y, settings->halo_depth,
            settings->coefficient, rx, ry, chunk->density, chunk->energy,
            chunk->u0, chunk->u, chunk->kx, chunk->ky));
    STOP_PROFILING(settings->kernel_profile, __func__);
}

void run_jacobi_iterate(Chunk* chunk, Settings* settings, double* error)
{
    START_PROFILING(settings->kernel_profile);
    TL_CHECK(tlhip2d_jacobi_iterate(chunk->x, chunk->y, settings->halo_depth,
            error, chunk->kx, chunk->ky, chunk->u0, chunk->u, chunk->r));
    STOP_PROFILING(settings->kernel_profile, __func__);
}

void run_ppcg_init(Chunk* chunk, Settings* settings)
{
    START_PROFILING(settings->kernel_profile);
    TL_CHECK(tlhip2d_ppcg_init(chunk->x, chunk->y, settings->halo_depth,
            chunk->theta, chunk->r, chunk->sd));
    STOP_PROFILING(settings->kernel_profile, __func__);
}

// One fused sweep; sd moves to the other buffer.
void run_ppcg_inner_iteration(Chunk* chunk, Settings* settings, double alpha,
        double beta)
{
    START_PROFILING(settings->kernel_profile);
    TL_CHECK(tlhip2d_ppcg_inner_iteration(chunk->x, chunk->y,
            settings->halo_depth, alpha, beta, chunk->u, chunk->r, chunk->kx,
            chunk->ky, chunk->sd, chunk->ext->sd_alt));
    double* t = chunk->sd;
    chunk->sd = chunk->ext->sd_alt;
    chunk->ext->sd_alt = t;
    STOP_PROFILING(settings->kernel_profile, __func__);
}

void run_copy_u(Chunk* chunk, Settings* settings)
{
    START_PROFILING(settings->kernel_profile);
    TL_CHECK(tlhip2d_copy_u(chunk->x, chunk->y, settings->halo_depth, chunk->u0,
            chunk->u));
    STOP_PROFILING(settings->kernel_profile, __func__);
}

void run_calculate_residual(Chunk* chunk, Settings* settings)
{
    START_PROFILING(settings->kernel_profile);
    TL_CHECK(tlhip2d_calculate_residual(chunk->x, chunk->y, settings->halo_depth,
            chunk->u, chunk->u0, chunk->r, chunk->kx, chunk->ky));
    STOP_PROFILING(settings->kernel_profile, __func__);
}

void run_calculate_2norm(Chunk* chunk, Settings* settings,
        FieldBufferType buffer, double* norm)
{
    START_PROFILING(settings->kernel_profile);
    TL_CHECK(tlhip2d_calculate_2norm(chunk->x, chunk->y, settings->halo_depth,
            buffer, norm));
    STOP_PROFILING(settings->kernel_profile, __func__);
}

void run_finalise(Chunk* chunk, Settings* settings)
{
    START_PROFILING(settings->kernel_profile);
    TL_CHECK(tlhip2d_finalise(chunk->x, chunk->y, settings->halo_depth,
            chunk->energy, chunk->density, chunk->u));
    STOP_PROFILING(settings->kernel_profile, __func__);
}
