/*
 * TeaLeaf CPU ORACLE, 2D kernels -- test infrastructure (see oracle.h).
 *
 * Serial restatement of the 2d/c_kernels/omp3/ kernels.  Each function cites the
 * reference function it restates.  Reductions accumulate into a local in
 * row-major order (jj outer, kk inner) and then add into the caller's
 * scalar, which is the reference's order when its OpenMP pragmas are off.
 */
#include <math.h>
#include <stddef.h>
#include "oracle.h"

#define EXTERNAL_FACE (-1)

/* The 5-point operator row, 2d/shared.h:50-54 (SMVP), same association:
 * ((1 + (kx_r + kx_c)) + (ky_u + ky_c))*a_c - (kx_r a_r + kx_c a_l)
 *                                          - (ky_u a_u + ky_c a_d)      */
static inline double apply_a(const double* a, const double* kx,
        const double* ky, long i, long x)
{
    const double kxc = kx[i], kxr = kx[i + 1];
    const double kyc = ky[i], kyu = ky[i + x];
    const double diag = (1.0 + (kxr + kxc)) + (kyu + kyc);
    return diag * a[i] - (kxr * a[i + 1] + kxc * a[i - 1])
                       - (kyu * a[i + x] + kyc * a[i - x]);
}

/* 2d/c_kernels/omp3/set_chunk_data.c:10-52 */
void or2d_set_chunk_data(int x, int y, int hd, int left, int bottom,
        double grid_x_min, double grid_y_min, double dx, double dy,
        double* cell_x, double* cell_y, double* vertex_x, double* vertex_y,
        double* volume, double* x_area, double* y_area)
{
    const double x0 = grid_x_min + dx * (double)left;
    const double y0 = grid_y_min + dy * (double)bottom;
    for (int i = 0; i <= x; ++i) vertex_x[i] = x0 + dx * (i - hd);
    for (int j = 0; j <= y; ++j) vertex_y[j] = y0 + dy * (j - hd);
    for (int i = 0; i < x; ++i) cell_x[i] = 0.5 * (vertex_x[i] + vertex_x[i + 1]);
    for (int j = 0; j < y; ++j) cell_y[j] = 0.5 * (vertex_y[j] + vertex_y[j + 1]);
    const long n = (long)x * y;
    for (long c = 0; c < n; ++c) {
        volume[c] = dx * dy;
        x_area[c] = dy;
        y_area[c] = dx;
    }
}

/* 2d/c_kernels/omp3/set_chunk_state.c:10-84.  State 0 fills everything,
 * later states overwrite the cells they cover, then u = e*rho on [1,n-1). */
void or2d_set_chunk_state(int x, int y, const double* vertex_x,
        const double* vertex_y, const double* cell_x, const double* cell_y,
        double* density, double* energy0, double* u, int num_states,
        const tl_state* states)
{
    const long n = (long)x * y;
    for (long c = 0; c < n; ++c) {
        energy0[c] = states[0].energy;
        density[c] = states[0].density;
    }
    for (int s = 1; s < num_states; ++s) {
        const tl_state* st = &states[s];
        for (int jj = 0; jj < y; ++jj) {
            for (int kk = 0; kk < x; ++kk) {
                int hit = 0;
                if (st->geometry == 0) {
                    hit = vertex_x[kk + 1] >= st->x_min && vertex_x[kk] < st->x_max
                       && vertex_y[jj + 1] >= st->y_min && vertex_y[jj] < st->y_max;
                } else if (st->geometry == 1) {
                    const double ddx = cell_x[kk] - st->x_min;
                    const double ddy = cell_y[jj] - st->y_min;
                    hit = sqrt(ddx * ddx + ddy * ddy) <= st->radius;
                } else if (st->geometry == 2) {
                    hit = vertex_x[kk] == st->x_min && vertex_y[jj] == st->y_min;
                }
                if (hit) {
                    const long c = kk + (long)jj * x;
                    energy0[c] = st->energy;
                    density[c] = st->density;
                }
            }
        }
    }
    for (int jj = 1; jj < y - 1; ++jj)
        for (int kk = 1; kk < x - 1; ++kk) {
            const long c = kk + (long)jj * x;
            u[c] = energy0[c] * density[c];
        }
}

/* Mirror `depth` halo layers of one field on every external face.
 * 2d/c_kernels/omp3/local_halos.c:52-149: left/right cover interior rows,
 * top/bottom interior columns; corners are never written. */
static void mirror_field(int x, int y, int depth, int hd,
        const int* nb, double* f)
{
    if (!f) return;
    if (nb[0] == EXTERNAL_FACE)          /* left */
        for (int jj = hd; jj < y - hd; ++jj)
            for (int k = 0; k < depth; ++k)
                f[(long)jj * x + (hd - k - 1)] = f[(long)jj * x + (hd + k)];
    if (nb[1] == EXTERNAL_FACE)          /* right */
        for (int jj = hd; jj < y - hd; ++jj)
            for (int k = 0; k < depth; ++k)
                f[(long)jj * x + (x - hd + k)] = f[(long)jj * x + (x - hd - 1 - k)];
    if (nb[3] == EXTERNAL_FACE)          /* top */
        for (int k = 0; k < depth; ++k)
            for (int kk = hd; kk < x - hd; ++kk)
                f[kk + (long)(y - hd + k) * x] = f[kk + (long)(y - hd - 1 - k) * x];
    if (nb[2] == EXTERNAL_FACE)          /* bottom */
        for (int k = 0; k < depth; ++k)
            for (int kk = hd; kk < x - hd; ++kk)
                f[kk + (long)(hd - k - 1) * x] = f[kk + (long)(hd + k) * x];
}

/* 2d/c_kernels/omp3/local_halos.c:22-50; field flags indexed as
 * 2d/shared.h:31-36 (density, energy0, energy1, u, p, sd). */
void or2d_local_halos(int x, int y, int depth, int hd, const int* neighbours,
        const bool* fields, double* density, double* energy0, double* energy,
        double* u, double* p, double* sd)
{
    double* f[6] = {density, energy0, energy, u, p, sd};
    for (int i = 0; i < 6; ++i)
        if (fields[i]) mirror_field(x, y, depth, hd, neighbours, f[i]);
}

/* Face box for pack/unpack, 2d/c_kernels/omp3/pack_halos.c:5-173.  Packing
 * reads the `depth` interior layers next to the face, unpacking writes the
 * `depth` halo layers; the buffer is x-fastest over the box. */
static int face_box2d(int x, int y, int depth, int hd, int face, bool pack,
        int* k0, int* j0, int* ex, int* ey)
{
    const int xi = x - 2 * hd, yi = y - 2 * hd;
    switch (face) {
    case 0: *k0 = pack ? hd : hd - depth;         *j0 = hd; *ex = depth; *ey = yi; return 0;
    case 1: *k0 = pack ? x - hd - depth : x - hd; *j0 = hd; *ex = depth; *ey = yi; return 0;
    case 2: *j0 = pack ? hd : hd - depth;         *k0 = hd; *ex = xi; *ey = depth; return 0;
    case 3: *j0 = pack ? y - hd - depth : y - hd; *k0 = hd; *ex = xi; *ey = depth; return 0;
    default: return -1;
    }
}

int or2d_pack_or_unpack(int x, int y, int depth, int hd, int face, bool pack,
        double* field, double* buffer)
{
    int k0, j0, ex, ey;
    if (face_box2d(x, y, depth, hd, face, pack, &k0, &j0, &ex, &ey)) return -1;
    for (int dj = 0; dj < ey; ++dj)
        for (int dk = 0; dk < ex; ++dk) {
            const long c = (k0 + dk) + (long)(j0 + dj) * x;
            const long b = dk + (long)dj * ex;
            if (pack) buffer[b] = field[c]; else field[c] = buffer[b];
        }
    return 0;
}

/* 2d/c_kernels/omp3/store_energy.c:4-16 */
void or2d_store_energy(int x, int y, const double* energy0, double* energy)
{
    const long n = (long)x * y;
    for (long c = 0; c < n; ++c) energy[c] = energy0[c];
}

/* 2d/c_kernels/omp3/field_summary.c:9-45 (serial in the reference too). */
void or2d_field_summary(int x, int y, int hd, const double* volume,
        const double* density, const double* energy0, const double* u,
        double* vol, double* mass, double* ie, double* temp)
{
    double sv = 0.0, sm = 0.0, si = 0.0, st = 0.0;
    for (int jj = hd; jj < y - hd; ++jj)
        for (int kk = hd; kk < x - hd; ++kk) {
            const long c = kk + (long)jj * x;
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

/* 2d/c_kernels/omp3/cg.c:9-85.  Returns -1 for an invalid coefficient
 * (the reference dies there). */
int or2d_cg_init(int x, int y, int hd, int coefficient, double rx, double ry,
        double* rro, const double* density, const double* energy, double* u,
        double* p, double* r, double* w, double* kx, double* ky)
{
    if (coefficient != 1 && coefficient != 2) return -1;
    const long n = (long)x * y;
    for (long c = 0; c < n; ++c) {
        p[c] = 0.0;
        r[c] = 0.0;
        u[c] = energy[c] * density[c];
    }
    for (int jj = 1; jj < y - 1; ++jj)
        for (int kk = 1; kk < x - 1; ++kk) {
            const long c = kk + (long)jj * x;
            w[c] = coefficient == 1 ? density[c] : 1.0 / density[c];
        }
    for (int jj = hd; jj < y - 1; ++jj)
        for (int kk = hd; kk < x - 1; ++kk) {
            const long c = kk + (long)jj * x;
            kx[c] = rx * (w[c - 1] + w[c]) / (2.0 * w[c - 1] * w[c]);
            ky[c] = ry * (w[c - x] + w[c]) / (2.0 * w[c - x] * w[c]);
        }
    double acc = 0.0;
    for (int jj = hd; jj < y - hd; ++jj)
        for (int kk = hd; kk < x - hd; ++kk) {
            const long c = kk + (long)jj * x;
            w[c] = apply_a(u, kx, ky, c, x);
            r[c] = u[c] - w[c];
            p[c] = r[c];
            acc += r[c] * p[c];
        }
    *rro += acc;
    return 0;
}

/* 2d/c_kernels/omp3/cg.c:88-113 */
void or2d_cg_calc_w(int x, int y, int hd, double* pw, const double* p,
        double* w, const double* kx, const double* ky)
{
    double acc = 0.0;
    for (int jj = hd; jj < y - hd; ++jj)
        for (int kk = hd; kk < x - hd; ++kk) {
            const long c = kk + (long)jj * x;
            w[c] = apply_a(p, kx, ky, c, x);
            acc += w[c] * p[c];
        }
    *pw += acc;
}

/* 2d/c_kernels/omp3/cg.c:116-143 */
void or2d_cg_calc_ur(int x, int y, int hd, double alpha, double* rrn,
        double* u, const double* p, double* r, const double* w)
{
    double acc = 0.0;
    for (int jj = hd; jj < y - hd; ++jj)
        for (int kk = hd; kk < x - hd; ++kk) {
            const long c = kk + (long)jj * x;
            u[c] += alpha * p[c];
            r[c] -= alpha * w[c];
            acc += r[c] * r[c];
        }
    *rrn += acc;
}

/* 2d/c_kernels/omp3/cg.c:146-164 */
void or2d_cg_calc_p(int x, int y, int hd, double beta, double* p,
        const double* r)
{
    for (int jj = hd; jj < y - hd; ++jj)
        for (int kk = hd; kk < x - hd; ++kk) {
            const long c = kk + (long)jj * x;
            p[c] = beta * p[c] + r[c];
        }
}

/* u += p on the interior, 2d/c_kernels/omp3/cheby.c:8-24 */
static void add_p_to_u(int x, int y, int hd, double* u, const double* p)
{
    for (int jj = hd; jj < y - hd; ++jj)
        for (int kk = hd; kk < x - hd; ++kk) {
            const long c = kk + (long)jj * x;
            u[c] += p[c];
        }
}

/* 2d/c_kernels/omp3/cheby.c:27-54 */
void or2d_cheby_init(int x, int y, int hd, double theta, double* u,
        const double* u0, double* p, double* r, double* w, const double* kx,
        const double* ky)
{
    for (int jj = hd; jj < y - hd; ++jj)
        for (int kk = hd; kk < x - hd; ++kk) {
            const long c = kk + (long)jj * x;
            w[c] = apply_a(u, kx, ky, c, x);
            r[c] = u0[c] - w[c];
            p[c] = r[c] / theta;
        }
    add_p_to_u(x, y, hd, u, p);
}

/* 2d/c_kernels/omp3/cheby.c:57-85 */
void or2d_cheby_iterate(int x, int y, int hd, double alpha, double beta,
        double* u, const double* u0, double* p, double* r, double* w,
        const double* kx, const double* ky)
{
    for (int jj = hd; jj < y - hd; ++jj)
        for (int kk = hd; kk < x - hd; ++kk) {
            const long c = kk + (long)jj * x;
            w[c] = apply_a(u, kx, ky, c, x);
            r[c] = u0[c] - w[c];
            p[c] = alpha * p[c] + beta * r[c];
        }
    add_p_to_u(x, y, hd, u, p);
}

/* 2d/c_kernels/omp3/jacobi.c:10-58.  Any coefficient other than 1 takes
 * the reciprocal branch; only coefficient < 1 is rejected. */
int or2d_jacobi_init(int x, int y, int hd, int coefficient, double rx,
        double ry, const double* density, const double* energy, double* u0,
        double* u, double* kx, double* ky)
{
    if (coefficient < 1) return -1;
    for (int jj = 1; jj < y - 1; ++jj)
        for (int kk = 1; kk < x - 1; ++kk) {
            const long c = kk + (long)jj * x;
            const double t = energy[c] * density[c];
            u0[c] = t;
            u[c] = t;
        }
    for (int jj = hd; jj < y - 1; ++jj)
        for (int kk = hd; kk < x - 1; ++kk) {
            const long c = kk + (long)jj * x;
            const int cond = coefficient == 1;
            const double wc = cond ? density[c] : 1.0 / density[c];
            const double wl = cond ? density[c - 1] : 1.0 / density[c - 1];
            const double wd = cond ? density[c - x] : 1.0 / density[c - x];
            kx[c] = rx * (wl + wc) / (2.0 * wl * wc);
            ky[c] = ry * (wd + wc) / (2.0 * wd * wc);
        }
    return 0;
}

/* 2d/c_kernels/omp3/jacobi.c:61-100.  *error is assigned, not accumulated. */
void or2d_jacobi_iterate(int x, int y, int hd, double* error,
        const double* kx, const double* ky, const double* u0, double* u,
        double* r)
{
    const long n = (long)x * y;
    for (long c = 0; c < n; ++c) r[c] = u[c];
    double err = 0.0;
    for (int jj = hd; jj < y - hd; ++jj)
        for (int kk = hd; kk < x - hd; ++kk) {
            const long c = kk + (long)jj * x;
            const double num = (u0[c] + (kx[c + 1] * r[c + 1] + kx[c] * r[c - 1]))
                             + (ky[c + x] * r[c + x] + ky[c] * r[c - x]);
            const double den = (1.0 + (kx[c] + kx[c + 1])) + (ky[c] + ky[c + x]);
            u[c] = num / den;
            err += fabs(u[c] - r[c]);
        }
    *error = err;
}

/* 2d/c_kernels/omp3/ppcg.c:8-25 */
void or2d_ppcg_init(int x, int y, int hd, double theta, const double* r,
        double* sd)
{
    for (int jj = hd; jj < y - hd; ++jj)
        for (int kk = hd; kk < x - hd; ++kk) {
            const long c = kk + (long)jj * x;
            sd[c] = r[c] / theta;
        }
}

/* 2d/c_kernels/omp3/ppcg.c:28-61: two sweeps, the second after the first
 * has finished everywhere. */
void or2d_ppcg_inner_iteration(int x, int y, int hd, double alpha,
        double beta, double* u, double* r, const double* kx, const double* ky,
        double* sd)
{
    for (int jj = hd; jj < y - hd; ++jj)
        for (int kk = hd; kk < x - hd; ++kk) {
            const long c = kk + (long)jj * x;
            const double a = apply_a(sd, kx, ky, c, x);
            r[c] -= a;
            u[c] += sd[c];
        }
    for (int jj = hd; jj < y - hd; ++jj)
        for (int kk = hd; kk < x - hd; ++kk) {
            const long c = kk + (long)jj * x;
            sd[c] = alpha * sd[c] + beta * r[c];
        }
}

/* 2d/c_kernels/omp3/solver_methods.c:9-25 */
void or2d_copy_u(int x, int y, int hd, double* u0, const double* u)
{
    for (int jj = hd; jj < y - hd; ++jj)
        for (int kk = hd; kk < x - hd; ++kk) {
            const long c = kk + (long)jj * x;
            u0[c] = u[c];
        }
}

/* 2d/c_kernels/omp3/solver_methods.c:28-48 */
void or2d_calculate_residual(int x, int y, int hd, const double* u,
        const double* u0, double* r, const double* kx, const double* ky)
{
    for (int jj = hd; jj < y - hd; ++jj)
        for (int kk = hd; kk < x - hd; ++kk) {
            const long c = kk + (long)jj * x;
            r[c] = u0[c] - apply_a(u, kx, ky, c, x);
        }
}

/* 2d/c_kernels/omp3/solver_methods.c:51-71 */
void or2d_calculate_2norm(int x, int y, int hd, const double* buffer,
        double* norm)
{
    double acc = 0.0;
    for (int jj = hd; jj < y - hd; ++jj)
        for (int kk = hd; kk < x - hd; ++kk) {
            const long c = kk + (long)jj * x;
            acc += buffer[c] * buffer[c];
        }
    *norm += acc;
}

/* 2d/c_kernels/omp3/solver_methods.c:74-91 */
void or2d_finalise(int x, int y, int hd, double* energy,
        const double* density, const double* u)
{
    for (int jj = hd; jj < y - hd; ++jj)
        for (int kk = hd; kk < x - hd; ++kk) {
            const long c = kk + (long)jj * x;
            energy[c] = u[c] / density[c];
        }
}
