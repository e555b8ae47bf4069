// Internal interface of the adjoint kernels used by the native driver.
// TlGrid3 with z = 1 describes a 2D chunk.
#pragma once

#include "tl_common.h"

namespace tl {
namespace adj {

// a[0..n) = 0
int zero(double* a, long n);
// lambda = 2 (u - target) on the interior (0 elsewhere); *loss = sum (u - target)^2
int loss_seed(const TlGrid3& g, const double* u, const double* target, double* lambda,
              double* loss);
// grad += sum over each cell's interior faces of -(mu_a - mu_b)(u_a - u_b) dk_f/ds
int face_grad(const TlGrid3& g, int coefficient, double rx, double ry, double rz, const double* mu,
              const double* u, const double* density, double* grad);
// grad += mu * b0 on the interior
int rhs_grad(const TlGrid3& g, const double* mu, const double* b0, double* grad);

}  // namespace adj
}  // namespace tl
