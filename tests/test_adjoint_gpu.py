"""The C-ABI adjoint (tlhip_problem_adjoint) on the GPU.

Loss = sum over the interior of (u_N - u_target)^2 after N implicit steps
(CG, FP64); u_target is the solve from a uniform density (1.0) perturbed by
seeded N(0, 0.01) noise.  The gradient with respect to the per-cell
log-density is checked
* against autodiff through an FP64 unrolled CG written independently in
  torch (CPU), and against the exact gradient from sparse LU solves, at
  128^2;
* against central finite differences (second order at h = 1e-4, fourth
  order at h = 3e-4) of the suite's own forward at 32 seeded cells, at the
  uniform density, on the 2048^2 layered deck and the 128^3 cuboid deck,
  and at the layered deck's own density at 512^2 (fixed-length solves, see
  the tests);
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


def _fixed_k(text_of, dims, steps, base):
    """Iterations for the fixed-length solves: twice the count that reaches a
    relative residual of 1e-10 at the base density, plus 200."""
    from tealeaf_amd.problem import Problem
    s0, states0 = tdeck.parse_deck(text_of(1e-12, 20000), dims)
    with Problem(s0, states0) as p:
        p.set_solve(0, 1e-10)
        p.inject(base if base is not None else np.ones(p.n), p.get("energy0"))
        return 2 * max(r["printed_cg_iters"] for r in p.run(steps)) + 200


class _Diff:
    """Loss differences along one cell's log-density, summed per cell,
    (u+ - u-)(u+ + u- - 2 target), so that they do not drown in the rounding
    of the totals."""

    def __init__(self, make, base, e0, target, steps, shape, inner):
        self.make, self.base, self.e0, self.t = make, base, e0, target[inner]
        self.steps, self.shape, self.inner = steps, shape, inner

    def _u(self, c, x):
        d = self.base.copy()
        d[c] = d[c] * np.exp(x)
        with self.make() as p:
            p.inject(d, self.e0)
            p.run(self.steps)
            return p.get("u").reshape(self.shape)[self.inner]

    def delta(self, c, h):           # L(+h) - L(-h)
        up, um = self._u(c, h), self._u(c, -h)
        return float(np.sum((up - um) * (up + um - 2.0 * self.t)))

    def central(self, c, h):         # second order
        return self.delta(c, h) / (2 * h)

    def central4(self, c, h):        # fourth order: (8 D(h) - D(2h)) / 12h
        return (8.0 * self.delta(c, h) - self.delta(c, 2 * h)) / (12 * h)


def _cells(g, inner, hd, n_rand, n_top, seed):
    gi = np.abs(g[inner])
    rng = np.random.default_rng(seed)
    rand = [tuple(int(rng.integers(hd, m + hd)) for m in gi.shape) for _ in range(n_rand)]
    top_pool = np.argsort(gi.ravel())[-max(n_top, gi.size // 100):]
    top = [tuple(int(i) + hd for i in np.unravel_index(f, gi.shape))
           for f in rng.choice(top_pool, n_top, replace=False)]
    return rand, top


@pytest.mark.timeout(900)
@pytest.mark.parametrize("name,deck,dims", FD_CASES, ids=[c[0] for c in FD_CASES])
def test_adjoint_matches_central_differences(hip, name, deck, dims):
    """The gradient at the uniform density (1.0) the target was made from,
    the first step of recovering the density from the noisy target; 16
    seeded cells drawn uniformly and 16 from the 1% with the largest |g|.

    Central differences resolve small errors only if the forward map is
    smooth and its derivative has converged.  A solve that stops on a
    residual test is neither (the iteration count can change between s+h
    and s-h), and the derivative of CG iterates settles later than the
    iterates.  So every solve here runs a fixed K iterations (eps 0).  What
    limits a difference quotient then is truncation, O(h^2) (5e-6 relative
    at h = 1e-3, 4e-7 at 3e-4, measured at 2048^2), and FP64 rounding in
    u+ - u-, which grows as 1/h (a floor of 2-5e-7 relative at h = 1e-4,
    1e-6 at 3e-5).  The second-order quotient at h = 1e-4 sits on that
    floor: held to 1e-6 max|g| everywhere and 2e-6 relative at the
    large-gradient cells.  The fourth-order quotient at h = 3e-4 (truncation
    O(h^4), rounding 1.5x the floor at that h) is held to 1e-6 relative at
    the large-gradient cells."""
    steps = 5
    k = _fixed_k(deck, dims, steps, None)
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
    gmax = np.abs(g[inner]).max()
    rand, top = _cells(g, inner, hd, 16, 16, 123 + dims)
    D = _Diff(make, ones, e0, target, steps, shape, inner)
    worst_abs = worst_rel2 = worst_rel4 = 0.0
    for c in rand + top:
        fd = D.central(c, 1e-4)
        worst_abs = max(worst_abs, abs(fd - g[c]) / gmax)
        line = f"{name} K={k} cell {c}: adjoint {g[c]:.10e} fd2 {fd:.10e}"
        if c in top:
            worst_rel2 = max(worst_rel2, abs(fd - g[c]) / abs(g[c]))
            fd4 = D.central4(c, 3e-4)
            worst_rel4 = max(worst_rel4, abs(fd4 - g[c]) / abs(g[c]))
            line += f" fd4 {fd4:.10e} rel2 {abs(fd - g[c]) / abs(g[c]):.2e} " \
                    f"rel4 {abs(fd4 - g[c]) / abs(g[c]):.2e}"
        print(line, flush=True)
    print(f"{name}: worst |fd2-g|/max|g| {worst_abs:.2e}, large-gradient cells: worst relative "
          f"fd2 {worst_rel2:.2e}, fd4 {worst_rel4:.2e}", flush=True)
    assert worst_abs < 1e-6, worst_abs
    assert worst_rel2 < 2e-6, worst_rel2
    assert worst_rel4 < 1e-6, worst_rel4


@pytest.mark.timeout(600)
@pytest.mark.parametrize("coefficient", [1, 2])
def test_adjoint_central_differences_layered_density(hip, coefficient):
    """The gradient at the layered deck's own (non-uniform) density, 512^2,
    dt 0.064 (the conditioning of 2048^2 at dt 0.004), fixed-length solves:
    fourth-order central differences at h = 3e-4 at 8 uniformly drawn cells
    (within 1e-6 max|g|) and 8 of the largest-|g| cells (within 1e-6
    relative)."""
    steps = 5

    def deck(eps, it):
        text = tdeck.layered_2d(512, steps, solver=tdeck.CG, eps=eps, max_iters=it, dt=0.064)
        if coefficient == 2:
            text = text.replace("*endtea", "coefficient_inverse_density\n*endtea")
        return text

    from tealeaf_amd.problem import Problem
    s0, states0 = tdeck.parse_deck(deck(1e-12, 20000), 2)
    with Problem(s0, states0) as p:
        dens0 = p.get("density")
    k = _fixed_k(deck, 2, steps, dens0)
    make, s, dens, e0, target = _setup(deck(0.0, k), 2, steps, 0.0, seed=90 + coefficient)
    shape = s.shape()
    hd = s.halo_depth
    inner = tuple(slice(hd, n - hd) for n in shape)
    base = dens.reshape(shape)
    assert np.ptp(base[inner]) > 0                       # non-uniform
    with make() as p:
        p.inject(base, e0)
        loss, grad, reps = p.adjoint(steps, target)
    assert all(r["printed_cg_iters"] == k for r in reps)
    g = grad.reshape(shape)
    gmax = np.abs(g[inner]).max()
    rand, top = _cells(g, inner, hd, 8, 8, 300 + coefficient)
    D = _Diff(make, base, e0, target, steps, shape, inner)
    worst_abs = worst_rel = 0.0
    for c in rand + top:
        fd4 = D.central4(c, 3e-4)
        worst_abs = max(worst_abs, abs(fd4 - g[c]) / gmax)
        if c in top:
            worst_rel = max(worst_rel, abs(fd4 - g[c]) / abs(g[c]))
        print(f"layered coefficient {coefficient} K={k} cell {c}: adjoint {g[c]:.10e} "
              f"fd4 {fd4:.10e}", flush=True)
    print(f"layered coefficient {coefficient}: worst |fd4-g|/max|g| {worst_abs:.2e}, worst "
          f"relative at the large-gradient cells {worst_rel:.2e}", flush=True)
    assert worst_abs < 1e-6, worst_abs
    assert worst_rel < 1e-6, worst_rel


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
