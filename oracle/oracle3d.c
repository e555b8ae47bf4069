/*
 * TeaLeaf CPU ORACLE, 3D kernels -- test infrastructure (see oracle.h).
 *
 * Serial restatement of the 3d/c_kernels/omp3/ kernels.  Index = kk + jj*x + ii*x*y;
 * reductions run ii outer, jj, kk inner into a local, then add into the
 * caller's scalar.
 */
#include <math.h>
#include <stddef.h>
#include "oracle.h"

#define EXTERNAL_FACE (-1)

/* The 7-point operator row, 3d/shared.h:52-58 (SMVP), same association. */
static inline double apply_a(const double* a, const double* kx,
        const double* ky, const double* kz, long i, long x, long xy)
{
    const double kxc = kx[i], kxr = kx[i + 1];
    const double kyc = ky[i], kyu = ky[i + x];
    const double kzc = kz[i], kzf = kz[i + xy];
    const double diag = ((1.0 + (kxr + kxc)) + (kyu + kyc)) + (kzf + kzc);
    return diag * a[i] - (kxr * a[i + 1] + kxc * a[i - 1])
                       - (kyu * a[i + x] + kyc * a[i - x])
                       - (kzf * a[i + xy] + kzc * a[i - xy]);
}

#define FOR_RANGE3(lo0, hi0, lo1, hi1, lo2, hi2)                       \
    for (int ii = (lo0); ii < (hi0); ++ii)                            \
        for (int jj = (lo1); jj < (hi1); ++jj)                        \
            for (int kk = (lo2); kk < (hi2); ++kk)
#define FOR_INTERIOR3 FOR_RANGE3(hd, z - hd, hd, y - hd, hd, x - hd)
#define IDX3 ((long)kk + (long)jj * x + (long)ii * x * y)

/* 3d/c_kernels/omp3/set_chunk_data.c:9-67 */
void or3d_set_chunk_data(int x, int y, int z, int hd, int left, int bottom,
        int back, double grid_x_min, double grid_y_min, double grid_z_min,
        double dx, double dy, double dz, double* cell_x, double* cell_y,
        double* cell_z, double* vertex_x, double* vertex_y, double* vertex_z,
        double* volume, double* x_area, double* y_area, double* z_area)
{
    const double x0 = grid_x_min + dx * (double)left;
    const double y0 = grid_y_min + dy * (double)bottom;
    const double z0 = grid_z_min + dz * (double)back;
    for (int i = 0; i <= x; ++i) vertex_x[i] = x0 + dx * (i - hd);
    for (int j = 0; j <= y; ++j) vertex_y[j] = y0 + dy * (j - hd);
    for (int k = 0; k <= z; ++k) vertex_z[k] = z0 + dz * (k - hd);
    for (int i = 0; i < x; ++i) cell_x[i] = 0.5 * (vertex_x[i] + vertex_x[i + 1]);
    for (int j = 0; j < y; ++j) cell_y[j] = 0.5 * (vertex_y[j] + vertex_y[j + 1]);
    for (int k = 0; k < z; ++k) cell_z[k] = 0.5 * (vertex_z[k] + vertex_z[k + 1]);
    const long n = (long)x * y * z;
    for (long c = 0; c < n; ++c) {
        volume[c] = dx * dy * dz;
        x_area[c] = dy * dz;
        y_area[c] = dx * dz;
        z_area[c] = dx * dy;
    }
}

/* 3d/c_kernels/omp3/set_chunk_state.c:10-97 */
void or3d_set_chunk_state(int x, int y, int z, const double* vertex_x,
        const double* vertex_y, const double* vertex_z, const double* cell_x,
        const double* cell_y, const double* cell_z, double* density,
        double* energy0, double* u, int num_states, const tl_state* states)
{
    const long n = (long)x * y * z;
    for (long c = 0; c < n; ++c) {
        energy0[c] = states[0].energy;
        density[c] = states[0].density;
    }
    for (int s = 1; s < num_states; ++s) {
        const tl_state* st = &states[s];
        FOR_RANGE3(0, z, 0, y, 0, x) {
            int hit = 0;
            if (st->geometry == 0) {
                hit = vertex_x[kk + 1] >= st->x_min && vertex_x[kk] < st->x_max
                   && vertex_y[jj + 1] >= st->y_min && vertex_y[jj] < st->y_max
                   && vertex_z[ii + 1] >= st->z_min && vertex_z[ii] < st->z_max;
            } else if (st->geometry == 1) {
                const double ddx = cell_x[kk] - st->x_min;
                const double ddy = cell_y[jj] - st->y_min;
                const double ddz = cell_z[ii] - st->z_min;
                hit = sqrt(ddx * ddx + ddy * ddy + ddz * ddz) <= st->radius;
            } else if (st->geometry == 2) {
                hit = vertex_x[kk] == st->x_min && vertex_y[jj] == st->y_min
                   && vertex_z[ii] == st->z_min;
            }
            if (hit) {
                energy0[IDX3] = st->energy;
                density[IDX3] = st->density;
            }
        }
    }
    FOR_RANGE3(1, z - 1, 1, y - 1, 1, x - 1) u[IDX3] = energy0[IDX3] * density[IDX3];
}

/* 3d/c_kernels/omp3/local_halos.c:57-222: six faces, each over the interior
 * extent of the other two axes; edges and corners are never written. */
static void mirror_field(int x, int y, int z, int depth, int hd,
        const int* nb, double* f)
{
    if (!f) return;
    const long xy = (long)x * y;
    if (nb[0] == EXTERNAL_FACE)          /* left */
        FOR_RANGE3(hd, z - hd, hd, y - hd, 0, depth) {
            const long b = ii * xy + (long)jj * x;
            f[b + (hd - kk - 1)] = f[b + (hd + kk)];
        }
    if (nb[1] == EXTERNAL_FACE)          /* right */
        FOR_RANGE3(hd, z - hd, hd, y - hd, 0, depth) {
            const long b = ii * xy + (long)jj * x;
            f[b + (x - hd + kk)] = f[b + (x - hd - 1 - kk)];
        }
    if (nb[3] == EXTERNAL_FACE)          /* top */
        FOR_RANGE3(hd, z - hd, 0, depth, hd, x - hd) {
            const long b = ii * xy + kk;
            f[b + (long)(y - hd + jj) * x] = f[b + (long)(y - hd - 1 - jj) * x];
        }
    if (nb[2] == EXTERNAL_FACE)          /* bottom */
        FOR_RANGE3(hd, z - hd, 0, depth, hd, x - hd) {
            const long b = ii * xy + kk;
            f[b + (long)(hd - jj - 1) * x] = f[b + (long)(hd + jj) * x];
        }
    if (nb[5] == EXTERNAL_FACE)          /* front */
        FOR_RANGE3(0, depth, hd, y - hd, hd, x - hd) {
            const long b = (long)jj * x + kk;
            f[b + (z - hd + ii) * xy] = f[b + (z - hd - 1 - ii) * xy];
        }
    if (nb[4] == EXTERNAL_FACE)          /* back */
        FOR_RANGE3(0, depth, hd, y - hd, hd, x - hd) {
            const long b = (long)jj * x + kk;
            f[b + (hd - ii - 1) * xy] = f[b + (hd + ii) * xy];
        }
}

/* 3d/c_kernels/omp3/local_halos.c:26-55 */
void or3d_local_halos(int x, int y, int z, int depth, int hd,
        const int* neighbours, const bool* fields, double* density,
        double* energy0, double* energy, double* u, double* p, double* sd)
{
    double* f[6] = {density, energy0, energy, u, p, sd};
    for (int i = 0; i < 6; ++i)
        if (fields[i]) mirror_field(x, y, z, depth, hd, neighbours, f[i]);
}

/* 3d/c_kernels/omp3/pack_halos.c:5-363: a depth-thick box on the face,
 * x-fastest in the buffer, then y, then z. */
int or3d_pack_or_unpack(int x, int y, int z, int depth, int hd, int face,
        bool pack, double* field, double* buffer)
{
    const int xi = x - 2 * hd, yi = y - 2 * hd, zi = z - 2 * hd;
    int k0 = hd, j0 = hd, i0 = hd, ex = xi, ey = yi, ez = zi;
    switch (face) {
    case 0: k0 = pack ? hd : hd - depth;         ex = depth; break;
    case 1: k0 = pack ? x - hd - depth : x - hd; ex = depth; break;
    case 2: j0 = pack ? hd : hd - depth;         ey = depth; break;
    case 3: j0 = pack ? y - hd - depth : y - hd; ey = depth; break;
    case 4: i0 = pack ? hd : hd - depth;         ez = depth; break;
    case 5: i0 = pack ? z - hd - depth : z - hd; ez = depth; break;
    default: return -1;
    }
    for (int di = 0; di < ez; ++di)
        for (int dj = 0; dj < ey; ++dj)
            for (int dk = 0; dk < ex; ++dk) {
                const long c = (k0 + dk) + (long)(j0 + dj) * x + (long)(i0 + di) * x * y;
                const long b = dk + (long)dj * ex + (long)di * ex * ey;
                if (pack) buffer[b] = field[c]; else field[c] = buffer[b];
            }
    return 0;
}

/* 3d/c_kernels/omp3/store_energy.c:4-16 */
void or3d_store_energy(int x, int y, int z, const double* energy0,
        double* energy)
{
    const long n = (long)x * y * z;
    for (long c = 0; c < n; ++c) energy[c] = energy0[c];
}

/* 3d/c_kernels/omp3/field_summary.c:9-50 */
void or3d_field_summary(int x, int y, int z, int hd, const double* volume,
        const double* density, const double* energy0, const double* u,
        double* vol, double* mass, double* ie, double* temp)
{
    double sv = 0.0, sm = 0.0, si = 0.0, st = 0.0;
    FOR_INTERIOR3 {
        const long c = IDX3;
        const double cv = volume[c];
        const double cm = cv * density[c];
        sv += cv;
        sm += cm;
        si += cm * energy0[c];
        st += cm * u[c];
    }
    *vol += sv;
    *ie += si;
    *temp += st;
    *mass += sm;
}

/* 3d/c_kernels/omp3/cg_solver.c:9-102 */
int or3d_cg_solver_init(int x, int y, int z, int hd, int coefficient,
        double rx, double ry, double rz, double* rro, const double* density,
        const double* energy, double* u, double* p, double* r, double* w,
        double* kx, double* ky, double* kz)
{
    if (coefficient != 1 && coefficient != 2) return -1;
    const long xy = (long)x * y;
    FOR_RANGE3(0, z, 0, y, 0, x) {
        const long c = IDX3;
        p[c] = 0.0;
        r[c] = 0.0;
        u[c] = energy[c] * density[c];
    }
    FOR_RANGE3(1, z - 1, 1, y - 1, 1, x - 1) {
        const long c = IDX3;
        w[c] = coefficient == 1 ? density[c] : 1.0 / density[c];
    }
    FOR_RANGE3(hd, z - 1, hd, y - 1, hd, x - 1) {
        const long c = IDX3;
        kx[c] = rx * (w[c - 1] + w[c]) / (2.0 * w[c - 1] * w[c]);
        ky[c] = ry * (w[c - x] + w[c]) / (2.0 * w[c - x] * w[c]);
        kz[c] = rz * (w[c - xy] + w[c]) / (2.0 * w[c - xy] * w[c]);
    }
    double acc = 0.0;
    FOR_INTERIOR3 {
        const long c = IDX3;
        w[c] = apply_a(u, kx, ky, kz, c, x, xy);
        r[c] = u[c] - w[c];
        p[c] = r[c];
        acc += r[c] * p[c];
    }
    *rro += acc;
    return 0;
}

/* 3d/c_kernels/omp3/cg_solver.c:105-135 */
void or3d_cg_solver_calc_w(int x, int y, int z, int hd, double* pw,
        const double* p, double* w, const double* kx, const double* ky,
        const double* kz)
{
    const long xy = (long)x * y;
    double acc = 0.0;
    FOR_INTERIOR3 {
        const long c = IDX3;
        w[c] = apply_a(p, kx, ky, kz, c, x, xy);
        acc += w[c] * p[c];
    }
    *pw += acc;
}

/* 3d/c_kernels/omp3/cg_solver.c:138-169 */
void or3d_cg_solver_calc_ur(int x, int y, int z, int hd, double alpha,
        double* rrn, double* u, const double* p, double* r, const double* w)
{
    double acc = 0.0;
    FOR_INTERIOR3 {
        const long c = IDX3;
        u[c] += alpha * p[c];
        r[c] -= alpha * w[c];
        acc += r[c] * r[c];
    }
    *rrn += acc;
}

/* 3d/c_kernels/omp3/cg_solver.c:172-193 */
void or3d_cg_solver_calc_p(int x, int y, int z, int hd, double beta,
        double* p, const double* r)
{
    FOR_INTERIOR3 {
        const long c = IDX3;
        p[c] = beta * p[c] + r[c];
    }
}

/* 3d/c_kernels/omp3/cheby_solver.c:8-28 */
static void add_p_to_u3(int x, int y, int z, int hd, double* u,
        const double* p)
{
    FOR_INTERIOR3 {
        const long c = IDX3;
        u[c] += p[c];
    }
}

/* 3d/c_kernels/omp3/cheby_solver.c:31-63 */
void or3d_cheby_solver_init(int x, int y, int z, int hd, double theta,
        double* u, const double* u0, double* p, double* r, double* w,
        const double* kx, const double* ky, const double* kz)
{
    const long xy = (long)x * y;
    FOR_INTERIOR3 {
        const long c = IDX3;
        w[c] = apply_a(u, kx, ky, kz, c, x, xy);
        r[c] = u0[c] - w[c];
        p[c] = r[c] / theta;
    }
    add_p_to_u3(x, y, z, hd, u, p);
}

/* 3d/c_kernels/omp3/cheby_solver.c:66-99 */
void or3d_cheby_solver_iterate(int x, int y, int z, int hd, double alpha,
        double beta, double* u, const double* u0, double* p, double* r,
        double* w, const double* kx, const double* ky, const double* kz)
{
    const long xy = (long)x * y;
    FOR_INTERIOR3 {
        const long c = IDX3;
        w[c] = apply_a(u, kx, ky, kz, c, x, xy);
        r[c] = u0[c] - w[c];
        p[c] = alpha * p[c] + beta * r[c];
    }
    add_p_to_u3(x, y, z, hd, u, p);
}

/* 3d/c_kernels/omp3/jacobi_solver.c:9-69 */
int or3d_jacobi_solver_init(int x, int y, int z, int hd, int coefficient,
        double rx, double ry, double rz, const double* density,
        const double* energy, double* u0, double* u, double* kx, double* ky,
        double* kz)
{
    if (coefficient < 1) return -1;
    const long xy = (long)x * y;
    FOR_RANGE3(1, z - 1, 1, y - 1, 1, x - 1) {
        const long c = IDX3;
        const double t = energy[c] * density[c];
        u0[c] = t;
        u[c] = t;
    }
    FOR_RANGE3(hd, z - 1, hd, y - 1, hd, x - 1) {
        const long c = IDX3;
        const int cond = coefficient == 1;
        const double wc = cond ? density[c] : 1.0 / density[c];
        const double wl = cond ? density[c - 1] : 1.0 / density[c - 1];
        const double wd = cond ? density[c - x] : 1.0 / density[c - x];
        const double wb = cond ? density[c - xy] : 1.0 / density[c - xy];
        kx[c] = rx * (wl + wc) / (2.0 * wl * wc);
        ky[c] = ry * (wd + wc) / (2.0 * wd * wc);
        kz[c] = rz * (wb + wc) / (2.0 * wb * wc);
    }
    return 0;
}

/* 3d/c_kernels/omp3/jacobi_solver.c:72-116 (assigns *error). */
void or3d_jacobi_solver_iterate(int x, int y, int z, int hd, double* error,
        const double* kx, const double* ky, const double* kz,
        const double* u0, double* u, double* r)
{
    const long xy = (long)x * y;
    const long n = xy * z;
    for (long c = 0; c < n; ++c) r[c] = u[c];
    double err = 0.0;
    FOR_INTERIOR3 {
        const long c = IDX3;
        const double num = ((u0[c] + (kx[c + 1] * r[c + 1] + kx[c] * r[c - 1]))
                          + (ky[c + x] * r[c + x] + ky[c] * r[c - x]))
                          + (kz[c + xy] * r[c + xy] + kz[c] * r[c - xy]);
        const double den = ((1.0 + (kx[c] + kx[c + 1])) + (ky[c] + ky[c + x]))
                         + (kz[c] + kz[c + xy]);
        u[c] = num / den;
        err += fabs(u[c] - r[c]);
    }
    *error = err;
}

/* 3d/c_kernels/omp3/ppcg_solver.c:8-29 */
void or3d_ppcg_init(int x, int y, int z, int hd, double theta,
        const double* r, double* sd)
{
    FOR_INTERIOR3 {
        const long c = IDX3;
        sd[c] = r[c] / theta;
    }
}

/* 3d/c_kernels/omp3/ppcg_solver.c:32-73 */
void or3d_ppcg_inner_iteration(int x, int y, int z, int hd, double alpha,
        double beta, double* u, double* r, const double* kx, const double* ky,
        const double* kz, double* sd)
{
    const long xy = (long)x * y;
    FOR_INTERIOR3 {
        const long c = IDX3;
        const double a = apply_a(sd, kx, ky, kz, c, x, xy);
        r[c] -= a;
        u[c] += sd[c];
    }
    FOR_INTERIOR3 {
        const long c = IDX3;
        sd[c] = alpha * sd[c] + beta * r[c];
    }
}

/* 3d/c_kernels/omp3/solver_methods.c:9-29 */
void or3d_solver_copy_u(int x, int y, int z, int hd, double* u0,
        const double* u)
{
    FOR_INTERIOR3 u0[IDX3] = u[IDX3];
}

/* 3d/c_kernels/omp3/solver_methods.c:32-57 */
void or3d_calculate_residual(int x, int y, int z, int hd, const double* u,
        const double* u0, double* r, const double* kx, const double* ky,
        const double* kz)
{
    const long xy = (long)x * y;
    FOR_INTERIOR3 {
        const long c = IDX3;
        r[c] = u0[c] - apply_a(u, kx, ky, kz, c, x, xy);
    }
}

/* 3d/c_kernels/omp3/solver_methods.c:60-84 */
void or3d_calculate_2norm(int x, int y, int z, int hd, const double* buffer,
        double* norm)
{
    double acc = 0.0;
    FOR_INTERIOR3 acc += buffer[IDX3] * buffer[IDX3];
    *norm += acc;
}

/* 3d/c_kernels/omp3/solver_methods.c:87-108 */
void or3d_solver_finalise(int x, int y, int z, int hd, double* energy,
        const double* density, const double* u)
{
    FOR_INTERIOR3 energy[IDX3] = u[IDX3] / density[IDX3];
}
