/*
 * Report hook for the reference host binaries built under oracle/_ref/
 * (test infrastructure).  Compiled with -DDIFFUSE_OVERLOAD, which makes the
 * reference main call diffuse_overload() instead of diffuse()
 * (2d/main.c:31-36, 3d/main.c:30-35).  It runs the reference's own diffuse()
 * unchanged and then prints
 *
 *   TL_FIELD_SUMMARY  the suite's run_field_summary: all four values, which
 *                     the reference computes but only partly prints
 *                     (2d/drivers/field_summary_driver.c:40-41);
 *   TL_REFORDER_SUMMARY  the same four sums formed here on the host from the
 *                     chunk's final fields, one cell after another in the
 *                     loop order and with the operations of
 *                     2d/c_kernels/omp3/field_summary.c:27-39 (3D: 3d/...:28-40).
 *
 * The second line separates the solver from the summation: two suites whose
 * fields agree give the same TL_REFORDER_SUMMARY however each forms its own
 * sums (the omp3 suite sums serially; the hip_cdna4 suite with a
 * deterministic tree, whose result is close to the exactly rounded sum).
 * Built with -DTL_HOOK_DEVICE the fields are device memory (the hip_cdna4
 * suite) and are copied to the host first; with -DTL_HOOK_3D the 3D chunk.
 * The same hook is linked with the omp3 suite and with the hip_cdna4 suite.
 */
#include <stdio.h>
#include <stdlib.h>
#include "application.h"
#include "kernel_interface.h"
#ifdef TL_HOOK_DEVICE
#include "tealeaf_hip.h"
#endif

static double* host_copy(const double* f, size_t n)
{
#ifdef TL_HOOK_DEVICE
    double* h = (double*)malloc(n * sizeof(double));
    if (!h || tlhip_copy_to_host(h, f, n * sizeof(double)) != 0) {
        fprintf(stderr, "report hook: copy of a field failed\n");
        exit(1);
    }
    return h;
#else
    (void)n;
    return (double*)f;
#endif
}

static void host_free(double* h)
{
#ifdef TL_HOOK_DEVICE
    free(h);
#else
    (void)h;
#endif
}

void diffuse_overload(Chunk* chunks, Settings* settings)
{
    diffuse(chunks, settings);
    double vol = 0.0, mass = 0.0, ie = 0.0, temp = 0.0;
    run_field_summary(&chunks[0], settings, &vol, &mass, &ie, &temp);
    printf("TL_FIELD_SUMMARY %.17e %.17e %.17e %.17e\n", vol, mass, ie, temp);

    const Chunk* c = &chunks[0];
    const int hd = settings->halo_depth;
#ifdef TL_HOOK_3D
    const size_t n = (size_t)c->x * c->y * c->z;
    const double* u_field = c->vec_u;
    const int nz = c->z;
#else
    const size_t n = (size_t)c->x * c->y;
    const double* u_field = c->u;
    const int nz = 2 * hd + 1;
#endif
    double* volume = host_copy(c->volume, n);
    double* density = host_copy(c->density, n);
    double* energy0 = host_copy(c->energy0, n);
    double* u = host_copy(u_field, n);
    double rv = 0.0, rm = 0.0, ri = 0.0, rt = 0.0;
    for (int ii = hd; ii < nz - hd; ++ii)
        for (int jj = hd; jj < c->y - hd; ++jj)
            for (int kk = hd; kk < c->x - hd; ++kk) {
#ifdef TL_HOOK_3D
                const size_t index = kk + (size_t)jj * c->x + (size_t)ii * c->x * c->y;
#else
                const size_t index = kk + (size_t)jj * c->x;
#endif
                const double cellVol = volume[index];
                const double cellMass = cellVol * density[index];
                rv += cellVol;
                rm += cellMass;
                ri += cellMass * energy0[index];
                rt += cellMass * u[index];
            }
    host_free(volume);
    host_free(density);
    host_free(energy0);
    host_free(u);
    printf("TL_REFORDER_SUMMARY %.17e %.17e %.17e %.17e\n", rv, rm, ri, rt);
    fflush(stdout);
}
