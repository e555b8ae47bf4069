"""The C-ABI adjoint (tlhip_problem_adjoint) on the GPU.

Loss = sum over the interior of (u_N - u_target)^2 after N implicit steps
(CG, FP64); u_target is the solve from a uniform density (1.0) perturbed by
seeded N(0, 0.01) noise.  The gradient with respect to the per-cell
log-density is checked
* against autodiff through an FP64 unrolled CG written independently in
  torch (CPU), and against the exact gradient from sparse LU solves, at
  128^2;
* against central finite differences (h = 1e-6 and 1e-4) of the suite's
  own forward at 32 seeded cells, at the uniform density, on the 2048^2
  layered deck and the 128^3 cuboid deck (fixed-length solves, see the
  test);
* through a torch.autograd.Function wrapper (test harness only), against a
  finite difference along a random direction.
The adjoint's formula itself is pinned on the CPU (test_adjoint_cpu.py).
"""
import numpy as np
import pytest
import torch

from tealeaf_amd import deck as tdeck
from tests import adjoint_ref as ar

pytestmark = pytest.mark.gpu

NOISE = 0.01


def _setup(text, dims, steps, rel, seed):
    """(make_problem, settings, density, energy0, target)"""
    from tealeaf_amd.problem import Problem
    s, states = tdeck.parse_deck(text, dims)

    def make():
        p = Problem(s, states)
        p.set_solve(0, rel)
        return p

    with make() as p:
        dens, e0 = p.get("density"), p.get("energy0")
    with make() as p:
        p.inject(np.ones_like(dens), e0)
        p.run(steps)
        u_uni = p.get("u").reshape(s.shape())
    hd = s.halo_depth
    inner = tuple(slice(hd, n - hd) for n in s.shape())
    target = np.zeros(s.shape())
    rng = np.random.default_rng(seed)
    target[inner] = u_uni[inner] + NOISE * rng.standard_normal(u_uni[inner].shape)
    return make, s, dens, e0, target


def _loss_forward(make, dens, e0, target, steps, inner, shape):
    with make() as p:
        p.inject(dens, e0)
        p.run(steps)
        u = p.get("u").reshape(shape)
    d = (u - target)[inner]
    return float(np.sum(d * d))


def _r(s):
    r = [s.dt_init / s.dx ** 2, s.dt_init / s.dy ** 2]
    if s.dims == 3:
        r.append(s.dt_init / s.dz ** 2)
    return tuple(r)


@pytest.mark.parametrize("coefficient", [1, 2])
def test_adjoint_matches_unrolled_autodiff(hip, coefficient):
    steps = 5
    text = tdeck.layered_2d(128, steps, solver=tdeck.CG, eps=1e-12, max_iters=5000)
    if coefficient == 2:
        text = text.replace("*endtea", "coefficient_inverse_density\n*endtea")
    make, s, dens, e0, target = _setup(text, 2, steps, 1e-13, seed=40 + coefficient)
    assert s.coefficient == coefficient
    hd = s.halo_depth
    inner = tuple(slice(hd, n - hd) for n in s.shape())
    shape = s.shape()
    with make() as p:
        loss, grad, reps = p.adjoint(steps, target)
    assert all(r["rel_residual"] <= 1e-13 for r in reps)
    g = grad.reshape(shape)
    st = torch.tensor(np.log(dens.reshape(shape)[inner]), requires_grad=True)
    lt, its = ar.unrolled_loss(st, torch.from_numpy(e0.reshape(shape)[inner]),
                               torch.from_numpy(target[inner]), _r(s), coefficient, steps)
    (gt,) = torch.autograd.grad(lt, st)
    gt = gt.numpy()
    assert abs(loss - lt.item()) <= 1e-9 * lt.item(), (loss, lt.item())
    err = np.linalg.norm(g[inner] - gt) / np.linalg.norm(gt)
    assert err < 1e-6, err
    assert np.abs(g[inner] - gt).max() <= 1e-6 * np.abs(gt).max()
    # and the exact gradient (sparse LU solves, CPU)
    loss_x, gx = ar.adjoint_exact(dens.reshape(shape)[inner], e0.reshape(shape)[inner],
                                  target[inner], _r(s), coefficient, steps)
    assert np.linalg.norm(g[inner] - gx) <= 1e-6 * np.linalg.norm(gx)
    # the halo carries no gradient
    halo = np.ones(shape, dtype=bool)
    halo[inner] = False
    assert not g[halo].any()


FD_CASES = [
    ("2d-2048", lambda eps, it: tdeck.layered_2d(2048, 5, solver=tdeck.CG, eps=eps, max_iters=it), 2),
    ("3d-128", lambda eps, it: tdeck.cuboid_3d(128, 5, solver=tdeck.CG, eps=eps, max_iters=it), 3),
]


@pytest.mark.timeout(900)
@pytest.mark.parametrize("name,deck,dims", FD_CASES, ids=[c[0] for c in FD_CASES])
def test_adjoint_matches_central_differences(hip, name, deck, dims):
    """The gradient at the uniform density (1.0) the target was made from,
    the first step of recovering the density from the noisy target; 32
    seeded cells: 16 uniformly drawn, 16 drawn from the 1% of cells with the
    largest |g|.

    Central differences resolve small errors only if the forward map is
    smooth and its derivative has converged.  A solve that stops on a
    residual test is neither (the iteration count can change between s+h
    and s-h), and the derivative of CG iterates settles later than the
    iterates.  So every solve here runs a fixed K iterations (eps 0): twice
    the count that reaches a relative residual of 1e-10, plus 200, for the
    adjoint's solves as for the forward ones.  The loss difference is summed
    per cell, (u+ - u-)(u+ + u- - 2 target), so that it does not drown in
    the rounding of the totals.  What is left is FP64 rounding in u+ - u-,
    which grows as 1/h and limits a difference quotient at h = 1e-6 to a few
    1e-6 max|g| even with exact solves (measured on the CPU with sparse LU
    at 512^2 and the same conditioning; 1.4e-5 (2D) and 3.9e-5 (3D) max|g|
    here, 100x the h = 1e-4 errors, as rounding predicts).  So h = 1e-6 is
    held to 1e-4 max|g|, and h = 1e-4 (truncation about 1e-8) to 1e-6
    max|g| at every cell and 1e-5 relative at the large-gradient cells."""
    from tealeaf_amd.problem import Problem
    steps = 5
    s0, states0 = tdeck.parse_deck(deck(1e-12, 20000), dims)
    with Problem(s0, states0) as p:
        p.set_solve(0, 1e-10)
        p.inject(np.ones(p.n), p.get("energy0"))
        k = 2 * max(r["printed_cg_iters"] for r in p.run(steps)) + 200
    make, s, _, e0, target = _setup(deck(0.0, k), dims, steps, 0.0, seed=77 + dims)
    shape = s.shape()
    hd = s.halo_depth
    inner = tuple(slice(hd, n - hd) for n in shape)
    ones = np.ones(shape)
    with make() as p:
        p.inject(ones, e0)
        loss, grad, reps = p.adjoint(steps, target)
    assert all(r["printed_cg_iters"] == k for r in reps)
    g = grad.reshape(shape)
    gi = np.abs(g[inner])
    gmax = gi.max()
    rng = np.random.default_rng(123 + dims)
    cells = [tuple(int(rng.integers(hd, n - hd)) for n in shape) for _ in range(16)]
    top = np.argsort(gi.ravel())[-max(16, gi.size // 100):]
    for f in rng.choice(top, 16, replace=False):
        cells.append(tuple(int(i) + hd for i in np.unravel_index(f, gi.shape)))
    t = target[inner]
    for h, tol in ((1e-4, 1e-6), (1e-6, 1e-4)):
        worst, worst_top = 0.0, 0.0
        for i, c in enumerate(cells):
            us = []
            for sign in (1, -1):
                d = ones.copy()
                d[c] = np.exp(sign * h)
                with make() as p:
                    p.inject(d, e0)
                    p.run(steps)
                    us.append(p.get("u").reshape(shape)[inner])
            fd = float(np.sum((us[0] - us[1]) * (us[0] + us[1] - 2.0 * t))) / (2 * h)
            err = abs(fd - g[c]) / gmax
            rel = abs(fd - g[c]) / abs(g[c])
            worst = max(worst, err)
            if i >= 16:
                worst_top = max(worst_top, rel)
            print(f"{name} K={k} h={h:g} cell {c}: adjoint {g[c]:.10e} fd {fd:.10e} "
                  f"|fd-g|/max|g| {err:.2e} rel {rel:.2e}", flush=True)
        print(f"{name} h={h:g}: worst |fd-g|/max|g| {worst:.2e}, worst relative at the "
              f"large-gradient cells {worst_top:.2e}", flush=True)
        assert worst < tol, (h, worst)
        if h == 1e-4:
            assert worst_top < 1e-5, worst_top


@pytest.mark.parametrize("coefficient", [1, 2])
def test_adjoint_matches_exact_gradient_512(hip, coefficient):
    """At 512^2 with dt 0.064 (the conditioning of 2048^2 at dt 0.004), on
    the layered deck's density: the whole gradient against the face-sum
    formula evaluated with exact sparse LU solves on the CPU."""
    steps = 5
    text = tdeck.layered_2d(512, steps, solver=tdeck.CG, max_iters=20000, dt=0.064)
    if coefficient == 2:
        text = text.replace("*endtea", "coefficient_inverse_density\n*endtea")
    make, s, dens, e0, target = _setup(text, 2, steps, 1e-12, seed=60 + coefficient)
    shape = s.shape()
    inner = tuple(slice(s.halo_depth, n - s.halo_depth) for n in shape)
    with make() as p:
        loss, grad, _ = p.adjoint(steps, target)
    g = grad.reshape(shape)[inner]
    loss_x, gx = ar.adjoint_exact(dens.reshape(shape)[inner], e0.reshape(shape)[inner],
                                  target[inner], _r(s), coefficient, steps)
    err = np.linalg.norm(g - gx) / np.linalg.norm(gx)
    print(f"512^2 coefficient {coefficient}: loss {loss:.12e} exact {loss_x:.12e}; "
          f"gradient rel err {err:.2e}, max {np.abs(g - gx).max() / np.abs(gx).max():.2e}", flush=True)
    assert abs(loss - loss_x) <= 1e-9 * loss_x
    assert err < 1e-6
    assert np.abs(g - gx).max() <= 1e-6 * np.abs(gx).max()


def test_autograd_function_wrapper(hip):
    """The torch.autograd.Function over the C ABI: its gradient is the
    adjoint's, and a central difference along a random direction of the
    loss it returns agrees (fixed-length solves, h = 1e-4, whose rounding
    and truncation errors are both below 1e-7 here)."""
    from tealeaf_amd.problem import Problem
    steps = 3
    s1, st1 = tdeck.parse_deck(tdeck.layered_2d(64, steps, solver=tdeck.CG, max_iters=5000), 2)
    with Problem(s1, st1) as p:
        p.set_solve(0, 1e-10)
        # 40 past convergence: enough for the derivative at 64^2, and short of
        # the point where the recurrence residual reaches exactly zero (CG's
        # next alpha would then be 0/0)
        k = max(r["printed_cg_iters"] for r in p.run(steps)) + 40
    text = tdeck.layered_2d(64, steps, solver=tdeck.CG, eps=0.0, max_iters=k)
    make, s, dens, e0, target = _setup(text, 2, steps, 0.0, seed=9)
    hd = s.halo_depth
    shape = s.shape()
    inner = tuple(slice(hd, n - hd) for n in shape)
    s0 = torch.tensor(np.log(dens.reshape(shape)[inner]), requires_grad=True)
    loss = ar.AdjointLoss.apply(s0, make, e0, target, steps, hd)
    (g,) = torch.autograd.grad(loss, s0)
    with make() as p:
        d2 = np.zeros(shape)
        d2[inner] = np.exp(s0.detach().numpy())
        p.inject(ar.reflect_halo(d2, hd), e0)
        loss2, grad2, _ = p.adjoint(steps, target)
    assert loss.item() == loss2
    np.testing.assert_array_equal(g.numpy(), grad2.reshape(shape)[inner])
    v = torch.from_numpy(np.random.default_rng(2).standard_normal(tuple(s0.shape)))
    h = 1e-4
    with torch.no_grad():
        lp = ar.AdjointLoss.apply(s0 + h * v, make, e0, target, steps, hd)
        lm = ar.AdjointLoss.apply(s0 - h * v, make, e0, target, steps, hd)
    fd = (lp - lm).item() / (2 * h)
    dd = torch.sum(g * v).item()
    assert abs(fd - dd) <= 1e-6 * abs(dd), (fd, dd)
