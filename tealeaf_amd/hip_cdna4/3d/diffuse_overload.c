// hip_cdna4 suite for TeaLeaf 3D: the timestep loop of 3d/diffuse.c:9-81
// over the native driver, for a reference host built with -DDIFFUSE_OVERLOAD
// (3d/main.c:30-35 then calls diffuse_overload instead of diffuse).
//
// Each timestep prints what diffuse() and the solver drivers print, in the
// same order and formats (3d/diffuse.c:24,74-80, drivers/cg_solver_driver.c:28,
// cheby_solver_driver.c:74-77, ppcg_solver_driver.c:60-63,
// jacobi_solver_driver.c:25, eigenvalue_driver.c:61-63), and the field
// summaries go through the reference's own field_summary_driver.  The solve itself is
// tlhip_problem_step over the chunk's device fields: CG and Jacobi
// iterations run device-resident, where the reference's drivers would take
// every dot product to the host.  Values
// and iteration counts are those of the reference's drivers over the same
// kernels.  One chunk on one rank only, which is what -DNO_MPI builds run.
#include <stdlib.h>
#include "application.h"
#include "chunk.h"
#include "shared.h"
#include "drivers/drivers.h"
#include "tealeaf_hip.h"

#define TL_CHECK(call)                                                   \
    do {                                                                 \
        if ((call) != 0) die(__LINE__, __FILE__, "%s\n", tlhip_last_error()); \
    } while (0)

static void print_iterations(Settings* settings, const tlhip_step_report* rep)
{
    if (rep->eigmax > 0.0)      // the estimate ran (the switch to Chebyshev / PPCG)
        print_and_log(settings, "Min. eigenvalue: \t%.12e\nMax. eigenvalue: \t%.12e\n",
                rep->eigmin, rep->eigmax);
    switch (settings->solver) {
    case JACOBI_SOLVER:
        print_and_log(settings, "Jacobi: \t\t%d iterations\n", rep->printed_cg_iters);
        break;
    case CG_SOLVER:
        print_and_log(settings, "CG: \t\t\t%d iterations\n", rep->printed_cg_iters);
        break;
    case CHEBY_SOLVER:
        print_and_log(settings, "CG: \t\t\t%d iterations\n", rep->printed_cg_iters);
        print_and_log(settings, "Cheby: \t\t\t%d iterations (%d estimated)\n",
                rep->printed_inner_iters, rep->est_iters);
        break;
    case PPCG_SOLVER:
        print_and_log(settings, "CG: \t\t\t%d iterations\n", rep->printed_cg_iters);
        print_and_log(settings, "PPCG: \t\t\t%d iterations (%d inner iterations per)\n",
                rep->printed_inner_iters, settings->ppcg_inner_steps);
        break;
    }
}

void diffuse_overload(Chunk* chunks, Settings* settings)
{
    if (settings->num_chunks_per_rank != 1 || settings->num_ranks != 1)
        die(__LINE__, __FILE__, "hip_cdna4 diffuse_overload runs one chunk on one rank\n");
    Chunk* c = &chunks[0];
    for (int f = 0; f < NUM_FACES; ++f)
        if (c->neighbours[f] != EXTERNAL_FACE)
            die(__LINE__, __FILE__, "hip_cdna4 diffuse_overload needs every face external\n");

    // dt as solve() forms it (3d/diffuse.c:27-32: the smaller of the deck's
    // and the chunk's dt_init; one rank)
    double dt = settings->dt_init;
    if (c->dt_init < dt) dt = c->dt_init;

    tlhip_settings s = {0};
    s.dims = 3;
    s.x_cells = settings->grid_x_cells;
    s.y_cells = settings->grid_y_cells;
    s.z_cells = settings->grid_z_cells;
    s.xmin = settings->grid_x_min;
    s.ymin = settings->grid_y_min;
    s.zmin = settings->grid_z_min;
    s.xmax = settings->grid_x_max;
    s.ymax = settings->grid_y_max;
    s.zmax = settings->grid_z_max;
    s.dt_init = dt;
    s.end_step = settings->end_step;
    s.max_iters = settings->max_iters;
    s.eps = settings->eps;
    s.solver = (int)settings->solver;
    s.coefficient = settings->coefficient;
    s.presteps = settings->presteps;
    s.ppcg_inner_steps = settings->ppcg_inner_steps;
    s.eps_lim = settings->eps_lim;
    s.error_switch = settings->error_switch;
    s.check_result = settings->check_result;
    s.summary_frequency = 0;    // summaries go through field_summary_driver below
    s.halo_depth = settings->halo_depth;

    tlhip_chunk_fields f = {0};
    f.density0 = c->density0;
    f.density = c->density;
    f.energy0 = c->energy0;
    f.energy = c->energy;
    f.u = c->vec_u;
    f.u0 = c->vec_u0;
    f.p = c->vec_p;
    f.r = c->vec_r;
    f.mi = c->vec_mi;
    f.w = c->vec_w;
    f.kx = c->vec_kx;
    f.ky = c->vec_ky;
    f.kz = c->vec_kz;
    f.sd = c->vec_sd;
    f.u_alt = c->ext->u_alt;
    f.sd_alt = c->ext->sd_alt;
    f.volume = c->volume;
    f.x_area = c->x_area;
    f.y_area = c->y_area;
    f.z_area = c->z_area;

    tlhip_problem* problem = NULL;
    TL_CHECK(tlhip_problem_adopt(&s, &f, &problem));

    double wallclock_prev = 0.0;
    for (int tt = 0; tt < settings->end_step; ++tt) {
        print_and_log(settings, "\nTimestep %d\n", tt + 1);
        profiler_start_timer(settings->wallclock_profile);

        tlhip_step_report rep;
        TL_CHECK(tlhip_problem_step(problem, tt, &rep));
        print_iterations(settings, &rep);

        if (tt % settings->summary_frequency == 0) field_summary_driver(chunks, settings, false);

        profiler_end_timer(settings->wallclock_profile, "Wallclock");
        double wallclock = settings->wallclock_profile->profiler_entries[0].time;
        print_and_log(settings, "Wallclock: \t\t%.3lfs\n", wallclock);
        print_and_log(settings, "Avg. time per cell: \t%.6e\n",
                (wallclock - wallclock_prev) /
                (settings->grid_x_cells * settings->grid_y_cells * settings->grid_z_cells));
        print_and_log(settings, "Error: \t\t\t%.6e\n", rep.error);
    }

    field_summary_driver(chunks, settings, true);
    TL_CHECK(tlhip_problem_destroy(problem));
}
