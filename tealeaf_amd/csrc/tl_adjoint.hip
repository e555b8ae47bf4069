// Kernels of the adjoint of the implicit timestep (2D and 3D).
//
// The reference only runs forward: each timestep solves A(rho) u_{n+1} =
// u_n (2d/diffuse.c:22-80; cg_init builds A from the density, finalise
// and the next cg_init turn u_{n+1} into the next right-hand side).  For
// L = sum (u_N - target)^2 the adjoint runs backwards:
//     lambda_N = 2 (u_N - target)
//     A mu_n = lambda_{n+1}               (A is symmetric: the same solver)
//     dL/dk_f += -(mu_a - mu_b)(u_{n+1,a} - u_{n+1,b})   for each face f=(a,b)
//     lambda_n = mu_n
// and dL/ds_i (s = ln rho) collects dk_f/ds_i over the faces of cell i plus
// mu_0,i * u_0,i from u_0 = e_0 rho.  A face between two interior cells has
// k_f = (r/2)(1/w_a + 1/w_b) (cg.c:49-63) with w = rho (conductivity) or
// 1/rho (recip_conductivity), so dk_f/ds_a = -(r/2)/w_a or +(r/2)/w_a.
// Faces to the halo carry no gradient: the reflected halo makes u_a = u_b.
#include "tl_ops.h"
#include "tealeaf_hip.h"

using namespace tl;

namespace {

// lambda = 2 (u - target) on the interior, 0 elsewhere; loss += (u - target)^2
__global__ void __launch_bounds__(kBlock)
k_loss_seed(TlGrid3 g, const double* __restrict__ u, const double* __restrict__ target,
            double* __restrict__ lambda, TlReduce rd)
{
    const long n = (long)g.x * g.y * g.z;
    double acc[1] = {0.0};
    const int zlo = g.z > 1 ? g.hd : 0, zhi = g.z > 1 ? g.z - g.hd : 1;
    for (long t = blockIdx.x * (long)kBlock + threadIdx.x; t < n; t += (long)gridDim.x * kBlock) {
        const int kk = (int)(t % g.x), jj = (int)((t / g.x) % g.y), ii = (int)(t / ((long)g.x * g.y));
        const bool in = kk >= g.hd && kk < g.x - g.hd && jj >= g.hd && jj < g.y - g.hd &&
                        ii >= zlo && ii < zhi;
        double l = 0.0;
        if (in) {
            const double d = u[t] - target[t];
            acc[0] += d * d;
            l = 2.0 * d;
        }
        lambda[t] = l;
    }
    tl_reduce_partial<1, kBlock>(acc, rd);
}

// grad += sum over the interior faces of each cell of
//   -(mu_a - mu_b)(u_a - u_b) * dk_f/ds_cell
__global__ void __launch_bounds__(kBlock)
k_face_grad(TlGrid3 g, int coefficient, double rx, double ry, double rz,
            const double* __restrict__ mu, const double* __restrict__ u,
            const double* __restrict__ density, double* __restrict__ grad)
{
    const bool three = g.z > 1;
    const int zlo = three ? g.hd : 0, zhi = three ? g.z - g.hd : 1;
    const long nx = g.x - 2 * g.hd, ny = g.y - 2 * g.hd, nz = zhi - zlo;
    const long n = nx * ny * nz;
    const long sy = g.x, sz = (long)g.x * g.y;
    for (long t = blockIdx.x * (long)kBlock + threadIdx.x; t < n; t += (long)gridDim.x * kBlock) {
        const int kk = g.hd + (int)(t % nx);
        const int jj = g.hd + (int)((t / nx) % ny);
        const int ii = zlo + (int)(t / (nx * ny));
        const long c = kk + jj * sy + ii * sz;
        const double m = mu[c], v = u[c], rho = density[c];
        // dk/ds for this cell per unit r: -(1/2)/rho (w = rho) or +(1/2) rho (w = 1/rho)
        const double dks = coefficient == 1 ? -0.5 / rho : 0.5 * rho;   // 1: w = rho
        double s = 0.0;
        auto face = [&](long o, double r) {
            s += r * (-(m - mu[o]) * (v - u[o]));
        };
        if (kk > g.hd) face(c - 1, rx);
        if (kk < g.x - g.hd - 1) face(c + 1, rx);
        if (jj > g.hd) face(c - sy, ry);
        if (jj < g.y - g.hd - 1) face(c + sy, ry);
        if (three) {
            if (ii > g.hd) face(c - sz, rz);
            if (ii < g.z - g.hd - 1) face(c + sz, rz);
        }
        grad[c] += s * dks;
    }
}

__global__ void __launch_bounds__(kBlock) k_zero64(double* __restrict__ a, long n)
{
    for (long t = blockIdx.x * (long)kBlock + threadIdx.x; t < n; t += (long)gridDim.x * kBlock)
        a[t] = 0.0;
}

// grad += mu * b0 on the interior (the dependence of u_0 = e_0 rho on rho)
__global__ void __launch_bounds__(kBlock)
k_rhs_grad(TlGrid3 g, const double* __restrict__ mu, const double* __restrict__ b0,
           double* __restrict__ grad)
{
    const bool three = g.z > 1;
    const int zlo = three ? g.hd : 0, zhi = three ? g.z - g.hd : 1;
    const long nx = g.x - 2 * g.hd, ny = g.y - 2 * g.hd, nz = zhi - zlo;
    const long n = nx * ny * nz;
    for (long t = blockIdx.x * (long)kBlock + threadIdx.x; t < n; t += (long)gridDim.x * kBlock) {
        const long c = (g.hd + t % nx) + (g.hd + (t / nx) % ny) * (long)g.x +
                       (zlo + t / (nx * ny)) * (long)g.x * g.y;
        grad[c] += mu[c] * b0[c];
    }
}

}  // namespace

namespace tl {
namespace adj {

int zero(double* a, long n)
{
    hipLaunchKernelGGL(k_zero64, dim3(flat_blocks(n)), dim3(kBlock), 0, tl_stream(), a, n);
    return tl_check_launch("tlhip_adjoint_zero");
}

int loss_seed(const TlGrid3& g, const double* u, const double* target, double* lambda,
              double* loss)
{
    const char* what = "tlhip_adjoint_loss";
    const long n = (long)g.x * g.y * g.z;
    const unsigned nb = flat_blocks(n);
    hipLaunchKernelGGL(k_loss_seed, dim3(nb), dim3(kBlock), 0, tl_stream(), g, u, target, lambda,
                       tl_reduce_slots());
    TL_TRY(tl_check_launch(what));
    TL_TRY(tl_finish_reduce(1, nb, what));
    return tl_fetch(1, loss, what);
}

int face_grad(const TlGrid3& g, int coefficient, double rx, double ry, double rz, const double* mu,
              const double* u, const double* density, double* grad)
{
    const char* what = "tlhip_adjoint_face_grad";
    const long n = (long)(g.x - 2 * g.hd) * (g.y - 2 * g.hd) * (g.z > 1 ? g.z - 2 * g.hd : 1);
    const int slot = tl_timer_begin(what);
    hipLaunchKernelGGL(k_face_grad, dim3(flat_blocks(n)), dim3(kBlock), 0, tl_stream(), g,
                       coefficient, rx, ry, rz, mu, u, density, grad);
    tl_timer_end(slot);
    return tl_check_launch(what);
}

int rhs_grad(const TlGrid3& g, const double* mu, const double* b0, double* grad)
{
    const char* what = "tlhip_adjoint_rhs_grad";
    const long n = (long)(g.x - 2 * g.hd) * (g.y - 2 * g.hd) * (g.z > 1 ? g.z - 2 * g.hd : 1);
    hipLaunchKernelGGL(k_rhs_grad, dim3(flat_blocks(n)), dim3(kBlock), 0, tl_stream(), g, mu, b0,
                       grad);
    return tl_check_launch(what);
}

}  // namespace adj
}  // namespace tl
