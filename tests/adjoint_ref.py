"""Test-harness pieces for the adjoint: an FP64 unrolled-CG forward in torch
(autodiff reference) and a torch.autograd.Function over the suite's C-ABI
adjoint.

The torch forward states the timestep from the mathematics, not from the
suite's code: A = I + sum over interior faces f=(a,b) of
k_f (e_a - e_b)(e_a - e_b)^T with k_f = (r/2)(1/w_a + 1/w_b), w = rho
(conductivity) or 1/rho (recip), which is the reference's operator
(2d/shared.h:50-54 with cg.c:49-63's coefficients) once the reflective halo
is taken into account; u_0 = e_0 rho; each step solves A u_{n+1} = u_n by
CG from the initial guess u_n (as cg_init does), unrolled for a fixed
number of iterations so that autograd differentiates through it.
"""
from __future__ import annotations

import numpy as np
import torch


def _coef(w_a, w_b, r):
    return 0.5 * r * (1.0 / w_a + 1.0 / w_b)


class Operator2D:
    def __init__(self, rho, rx, ry, coefficient):
        w = rho if coefficient == 1 else 1.0 / rho
        self.kx = _coef(w[:, :-1], w[:, 1:], rx)      # faces between columns k-1, k
        self.ky = _coef(w[:-1, :], w[1:, :], ry)      # faces between rows j-1, j

    def __call__(self, u):
        fx = self.kx * (u[:, 1:] - u[:, :-1])
        fy = self.ky * (u[1:, :] - u[:-1, :])
        au = u.clone()
        au[:, 1:] = au[:, 1:] + fx
        au[:, :-1] = au[:, :-1] - fx
        au[1:, :] = au[1:, :] + fy
        au[:-1, :] = au[:-1, :] - fy
        return au


class Operator3D:
    def __init__(self, rho, rx, ry, rz, coefficient):
        w = rho if coefficient == 1 else 1.0 / rho
        self.kx = _coef(w[:, :, :-1], w[:, :, 1:], rx)
        self.ky = _coef(w[:, :-1, :], w[:, 1:, :], ry)
        self.kz = _coef(w[:-1, :, :], w[1:, :, :], rz)

    def __call__(self, u):
        fx = self.kx * (u[:, :, 1:] - u[:, :, :-1])
        fy = self.ky * (u[:, 1:, :] - u[:, :-1, :])
        fz = self.kz * (u[1:, :, :] - u[:-1, :, :])
        au = u.clone()
        au[:, :, 1:] = au[:, :, 1:] + fx
        au[:, :, :-1] = au[:, :, :-1] - fx
        au[:, 1:, :] = au[:, 1:, :] + fy
        au[:, :-1, :] = au[:, :-1, :] - fy
        au[1:, :, :] = au[1:, :, :] + fz
        au[:-1, :, :] = au[:-1, :, :] - fz
        return au


def cg(apply_a, b, x0, iters):
    x = x0
    r = b - apply_a(x)
    p = r
    rr = torch.sum(r * r)
    for _ in range(iters):
        ap = apply_a(p)
        alpha = rr / torch.sum(p * ap)
        x = x + alpha * p
        r = r - alpha * ap
        rr_new = torch.sum(r * r)
        p = r + (rr_new / rr) * p
        rr = rr_new
    return x


def iterations_to(apply_a, b, x0, rel):
    """CG iterations needed for ||b - A x|| <= rel ||b|| (no autograd)."""
    with torch.no_grad():
        x = x0.clone()
        r = b - apply_a(x)
        p = r.clone()
        rr = torch.sum(r * r)
        bb = torch.sum(b * b)
        for k in range(100000):
            if rr <= rel * rel * bb:
                return k
            ap = apply_a(p)
            alpha = rr / torch.sum(p * ap)
            x += alpha * p
            r -= alpha * ap
            rr_new = torch.sum(r * r)
            p = r + (rr_new / rr) * p
            rr = rr_new
    raise RuntimeError("CG did not converge")


def unrolled_loss(s, e0, target, r, coefficient, steps, rel=1e-14, extra=60):
    """Loss and the per-step iteration counts of the unrolled forward; s, e0
    and target are interior arrays (torch float64), r = (rx, ry[, rz]).

    The derivative of the CG iterates converges later than the iterates: at
    128^2, stopping where the forward reaches 1e-14 leaves the autodiff
    gradient 1e-4 off the exact one, 20 more iterations 5e-9 and 60 more
    2e-12 (test_adjoint_cpu checks the last against sparse LU solves)."""
    rho = torch.exp(s)
    op = Operator2D(rho, *r, coefficient) if s.dim() == 2 else Operator3D(rho, *r, coefficient)
    u = e0 * rho
    its = []
    for _ in range(steps):
        k = iterations_to(op, u.detach(), u.detach(), rel) + extra
        u = cg(op, u, u, k)
        its.append(k)
    return torch.sum((u - target) ** 2), its


class AdjointLoss(torch.autograd.Function):
    """loss(s) for s = ln(density) over the interior, through the suite's
    forward steps and C-ABI adjoint.  `make_problem()` returns a fresh
    tealeaf_amd Problem in its initial state; energy0 is taken from it."""

    @staticmethod
    def forward(ctx, s, make_problem, energy0, target, steps, hd):
        prob = make_problem()
        try:
            shape = prob.settings.shape()
            dens = np.empty(shape)
            inner = tuple(slice(hd, n - hd) for n in shape)
            dens[inner] = np.exp(s.detach().cpu().numpy())
            dens = reflect_halo(dens, hd)
            prob.inject(dens, energy0)
            loss, grad, _ = prob.adjoint(steps, target)
        finally:
            prob.close()
        ctx.save_for_backward(torch.from_numpy(grad.reshape(shape)[inner].copy()))
        return torch.tensor(loss, dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        (grad,) = ctx.saved_tensors
        return g * grad, None, None, None, None, None


def reflect_halo(a, hd):
    """Fill the halo of `a` by reflection (the reference's halo update of
    density at depth hd on a single chunk)."""
    a = a.copy()
    for ax in range(a.ndim):
        n = a.shape[ax]
        for k in range(hd):
            src = [slice(None)] * a.ndim
            dst = [slice(None)] * a.ndim
            src[ax], dst[ax] = hd + k, hd - 1 - k
            a[tuple(dst)] = a[tuple(src)]
            src[ax], dst[ax] = n - hd - 1 - k, n - hd + k
            a[tuple(dst)] = a[tuple(src)]
    return a


def faces(shape):
    """(a, b, axis) index arrays of the interior faces of a grid, flattened C order."""
    idx = np.arange(int(np.prod(shape))).reshape(shape)
    out = []
    for ax in range(len(shape)):
        lo = [slice(None)] * len(shape)
        hi = [slice(None)] * len(shape)
        lo[ax], hi[ax] = slice(None, -1), slice(1, None)
        out.append((idx[tuple(lo)].ravel(), idx[tuple(hi)].ravel(), len(shape) - 1 - ax))
    return out          # axis numbered x=0, y=1, z=2


def sparse_operator(rho, r, coefficient):
    """A as a scipy CSC matrix on the interior cells."""
    import scipy.sparse as sp
    n = rho.size
    w = (rho if coefficient == 1 else 1.0 / rho).ravel()
    rows, cols, vals = [np.arange(n)], [np.arange(n)], [np.ones(n)]
    for a, b, ax in faces(rho.shape):
        k = 0.5 * r[ax] * (1.0 / w[a] + 1.0 / w[b])
        rows += [a, b, a, b]
        cols += [a, b, b, a]
        vals += [k, k, -k, -k]
    return sp.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(n, n))


def adjoint_exact(rho, e0, target, r, coefficient, steps):
    """The face-sum adjoint with exact (sparse LU) solves: the suite's
    formula without iteration error, for grids too large for dense solves."""
    import scipy.sparse.linalg as spl
    lu = spl.splu(sparse_operator(rho, r, coefficient))
    b0 = (e0 * rho).ravel()
    us, u = [], b0
    for _ in range(steps):
        u = lu.solve(u)
        us.append(u)
    lam = 2.0 * (us[-1] - target.ravel())
    rr = rho.ravel()
    dks = -0.5 / rr if coefficient == 1 else 0.5 * rr
    g = np.zeros(rho.size)
    mu = None
    for k in reversed(range(steps)):
        mu = lu.solve(lam)
        u1 = us[k]
        for a, b, ax in faces(rho.shape):
            d = -(mu[a] - mu[b]) * (u1[a] - u1[b]) * r[ax]
            np.add.at(g, a, d * dks[a])
            np.add.at(g, b, d * dks[b])
        lam = mu
    g += mu * b0
    return float(np.sum((us[-1] - target.ravel()) ** 2)), g.reshape(rho.shape)


def dense_operator(rho, r, coefficient):
    """A as a dense matrix (numpy) on the interior cells; r = (rx, ry[, rz])."""
    n = rho.size
    w = (rho if coefficient == 1 else 1.0 / rho).ravel()
    A = np.eye(n)
    for a, b, ax in faces(rho.shape):
        k = 0.5 * r[ax] * (1.0 / w[a] + 1.0 / w[b])
        A[a, a] += k
        A[b, b] += k
        A[a, b] -= k
        A[b, a] -= k
    return A


def adjoint_dense(rho, e0, target, r, coefficient, steps):
    """Exact forward by dense solves and the gradient by the face-sum adjoint
    the suite's kernels implement (dL/d ln rho per interior cell)."""
    A = dense_operator(rho, r, coefficient)
    b0 = (e0 * rho).ravel()
    us, u = [], b0
    for _ in range(steps):
        u = np.linalg.solve(A, u)
        us.append(u)
    lam = 2.0 * (us[-1] - target.ravel())
    rr = rho.ravel()
    dks = -0.5 / rr if coefficient == 1 else 0.5 * rr
    g = np.zeros(rho.size)
    mu = None
    for k in reversed(range(steps)):
        mu = np.linalg.solve(A, lam)
        u1 = us[k]
        for a, b, ax in faces(rho.shape):
            d = -(mu[a] - mu[b]) * (u1[a] - u1[b]) * r[ax]
            np.add.at(g, a, d * dks[a])
            np.add.at(g, b, d * dks[b])
        lam = mu
    g += mu * b0
    return float(np.sum((us[-1] - target.ravel()) ** 2)), g.reshape(rho.shape)


def exact_loss_torch(s, e0, target, r, coefficient, steps):
    """The same loss with dense torch solves (exact forward, autograd backward)."""
    rho = torch.exp(s)
    w = (rho if coefficient == 1 else 1.0 / rho).reshape(-1)
    n = w.numel()
    A = torch.eye(n, dtype=torch.float64)
    for a, b, ax in faces(tuple(s.shape)):
        a_t, b_t = torch.from_numpy(a), torch.from_numpy(b)
        k = 0.5 * r[ax] * (1.0 / w[a_t] + 1.0 / w[b_t])
        A = A.index_put((a_t, a_t), k, accumulate=True)
        A = A.index_put((b_t, b_t), k, accumulate=True)
        A = A.index_put((a_t, b_t), -k, accumulate=True)
        A = A.index_put((b_t, a_t), -k, accumulate=True)
    u = (e0 * rho).reshape(-1)
    for _ in range(steps):
        u = torch.linalg.solve(A, u)
    return torch.sum((u - target.reshape(-1)) ** 2)
