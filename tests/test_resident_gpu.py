"""The device-resident CG / PPCG / Jacobi loops against the host-driven loop.

With tlhip_set_resident(1) (the default) the driver forms alpha and beta on
the device, evaluates the stop and switch tests there and keeps several
iterations enqueued ahead of the host; with 0 it takes pw and rrn to the
host every iteration, as the reference's drivers do (cg_driver.c:78-99,
ppcg_driver.c:95-131).  Both run the same kernels with the same arguments in
the same order, so every report, every field and the summary must be the
same bits.  The resident Jacobi loop sweeps between u and a partner buffer
instead of copying u into r first (jacobi_driver.c:51-55): the same values,
so there too every bit must agree, r included.  The cases cover each place
the loop can end: convergence, the CG -> Chebyshev / PPCG switch by presteps
and by the error switch, max_iters (in the CG phase and in the PPCG phase),
the relative-residual stop whose true-residual check can refuse a
candidate, PPCG inner-step counts whose fused launches leave the ping-pong
buffers swapped, 3D (sqrt of the error in the stop test; PPCG's outer loop
resident around step-by-step inner sweeps), and seeded block densities.  The parity of the resident
path with the reference itself is held by the other driver tests, which run
with the default.
"""
import numpy as np
import pytest

from tealeaf_amd import blocks
from tealeaf_amd import deck as tdeck

pytestmark = pytest.mark.gpu


def _run(hip, text, resident, dims=None, inject=None, rel=0.0, extra=()):
    from tealeaf_amd.problem import Problem
    hip.tlhip_set_resident(resident)
    try:
        with Problem.from_deck(text, dims) as p:
            if inject is not None:
                p.inject(*inject)
            if rel > 0.0:
                p.set_solve(0, rel)
            reps = p.run()
            fields = [p.get(f) for f in ("u", "energy") + tuple(extra)]
            return reps, fields, p.field_summary()
    finally:
        hip.tlhip_set_resident(1)


def _same(hip, text, dims=None, inject=None, rel=0.0, extra=()):
    a = _run(hip, text, 0, dims, inject, rel, extra)
    b = _run(hip, text, 1, dims, inject, rel, extra)
    assert len(a[0]) == len(b[0])
    for ra, rb in zip(a[0], b[0]):
        assert ra == rb, (ra, rb)
    for fa, fb in zip(a[1], b[1]):
        np.testing.assert_array_equal(fa, fb)
    assert a[2] == b[2]
    return b[0]


@pytest.mark.parametrize("solver", [tdeck.CG, tdeck.CHEBY, tdeck.PPCG])
def test_2d_layered(hip, solver):
    reps = _same(hip, tdeck.layered_2d(96, 2, solver=solver, presteps=25, max_iters=3000))
    assert all(r["printed_cg_iters"] > 0 for r in reps)


@pytest.mark.parametrize("solver", [tdeck.CG, tdeck.CHEBY, tdeck.PPCG, tdeck.JACOBI])
def test_2d_odd_width(hip, solver):
    """An odd row length (77 + 4 halo cells): one cell per lane, no fused
    calc_p + calc_w and no fused PPCG or Chebyshev loop, so the resident CG
    body runs calc_w, calc_ur, calc_p and a halo update per iteration, and
    the PPCG outer loop stays host-driven."""
    eps = 1e-8 if solver == tdeck.JACOBI else 1e-15
    _same(hip, tdeck.layered_2d(77, 2, solver=solver, presteps=20, eps=eps, max_iters=3000),
          extra=("r",))


@pytest.mark.parametrize("solver", [tdeck.CHEBY, tdeck.PPCG])
def test_2d_error_switch(hip, solver):
    _same(hip, tdeck.layered_2d(80, 2, solver=solver, max_iters=3000, epslim=1e-3))


@pytest.mark.parametrize("inner", [3, 7, 12, 15])
def test_2d_ppcg_inner_counts(hip, inner):
    """Inner-step counts whose fused launches (5 steps each) end with the
    sd / r / u ping-pong swapped or not."""
    reps = _same(hip, tdeck.layered_2d(72, 1, solver=tdeck.PPCG, presteps=20, inner=inner,
                                       max_iters=3000))
    assert reps[0]["n_ppcg_outer"] > 0


@pytest.mark.parametrize("solver,max_iters", [(tdeck.CG, 50), (tdeck.PPCG, 25), (tdeck.PPCG, 34),
                                              (tdeck.CHEBY, 60)])
def test_max_iters_cutoff(hip, solver, max_iters):
    """Runs that end by max_iters.  At 128^2 (eps 1e-15) the oracle needs
    74 CG iterations, or 31 CG + 6 PPCG outer, or 31 CG + 49 Chebyshev; so
    the cut falls inside CG, inside PPCG's CG phase (25), inside its PPCG
    phase (34) and inside the Chebyshev phase."""
    reps = _same(hip, tdeck.layered_2d(128, 2, solver=solver, presteps=30, max_iters=max_iters))
    assert all(r["printed_cg_iters"] + r["printed_inner_iters"] >= max_iters - 1 for r in reps)


@pytest.mark.parametrize("solver", [tdeck.CG, tdeck.CHEBY, tdeck.PPCG])
@pytest.mark.parametrize("max_iters", [1, 2, 5])
def test_tiny_max_iters(hip, solver, max_iters):
    """Fewer iterations than the loop keeps in flight."""
    reps = _same(hip, tdeck.layered_2d(48, 2, solver=solver, presteps=1, max_iters=max_iters))
    assert all(r["printed_cg_iters"] + r["printed_inner_iters"] >= max_iters for r in reps)


@pytest.mark.parametrize("solver", [tdeck.CG, tdeck.PPCG])
def test_relative_residual_stop(hip, solver):
    """The relative-residual stop: the device flags a candidate, the host
    confirms it with the true residual (or refuses it and the loop goes on)."""
    reps = _same(hip, tdeck.layered_2d(80, 2, solver=solver, presteps=20, max_iters=3000),
                 rel=1e-9)
    assert all(r["n_true_residuals"] >= 1 for r in reps)
    assert all(0.0 < r["rel_residual"] <= 1e-9 for r in reps)


@pytest.mark.parametrize("solver", [tdeck.CG, tdeck.CHEBY, tdeck.PPCG])
def test_3d(hip, solver):
    _same(hip, tdeck.cuboid_3d(20, 2, solver=solver, presteps=15, max_iters=2000), dims=3)


@pytest.mark.parametrize("max_iters", [5, 14, 20])
def test_3d_ppcg_max_iters(hip, max_iters):
    """3D PPCG runs resident in its outer loop too (step by step inside).
    At 24^3 (presteps 10) the oracle needs 11 CG + 16 PPCG iterations per
    step, so the cut falls inside the CG phase (5) and inside the PPCG
    phase (14, 20)."""
    reps = _same(hip, tdeck.cuboid_3d(24, 2, solver=tdeck.PPCG, presteps=10, inner=10,
                                      max_iters=max_iters), dims=3)
    assert all(r["printed_cg_iters"] + r["printed_inner_iters"] >= max_iters - 1 for r in reps)


def test_3d_ppcg_relative_stop(hip):
    reps = _same(hip, tdeck.cuboid_3d(24, 2, solver=tdeck.PPCG, presteps=10, inner=10,
                                      max_iters=2000), dims=3, rel=1e-9)
    assert all(r["n_true_residuals"] >= 1 and r["n_ppcg_outer"] > 0 for r in reps)


def test_3d_blocks_recip(hip):
    text = tdeck.cuboid_3d(32, 1, solver=tdeck.PPCG, inner=10, coefficient=2, dt=0.004,
                           eps=1e-9, max_iters=3000)
    s, _ = tdeck.parse_deck(text, 3)
    inject = blocks.block_fields((32, 32, 32), 16, seed=5, rho=(0.1, 100.0), energy=(0.1, 25.0),
                                 halo_depth=s.halo_depth)
    _same(hip, text, dims=3, inject=inject)


def test_2d_blocks(hip):
    text = tdeck.layered_2d(128, 1, solver=tdeck.PPCG, max_iters=4000, epslim=1e-4)
    s, _ = tdeck.parse_deck(text, 2)
    inject = blocks.block_fields((128, 128), 32, seed=9, rho=(0.01, 1.0), energy=(0.1, 25.0),
                                 halo_depth=s.halo_depth)
    reps = _same(hip, text, dims=2, inject=inject)
    assert reps[0]["n_ppcg_outer"] > 0


# Jacobi: at 40^2 (eps 1e-8) the oracle stops after 49 and 48 iterations;
# at 12^3 (eps 1e-8, the 3D test on sqrt(error)) after 171 and 167.
@pytest.mark.parametrize("dims,text", [
    (2, tdeck.layered_2d(40, 2, solver=tdeck.JACOBI, eps=1e-8, max_iters=2000)),
    (2, tdeck.layered_2d(32, 2, solver=tdeck.JACOBI, eps=1e-10, max_iters=4000)),
    (3, tdeck.cuboid_3d(12, 2, solver=tdeck.JACOBI, eps=1e-8, max_iters=2000)),
], ids=["2d-40", "2d-32", "3d-12"])
def test_jacobi(hip, dims, text):
    reps = _same(hip, text, dims=dims, extra=("r", "u0"))
    assert all(r["printed_cg_iters"] >= 40 for r in reps)


@pytest.mark.parametrize("dims", [2, 3])
@pytest.mark.parametrize("max_iters", [1, 2, 5, 50, 51, 52, 101])
def test_jacobi_max_iters(hip, dims, max_iters):
    """Runs cut by max_iters: inside the first iterations in flight, just
    before, at and after an iteration that adds the residual's norm (every
    50th), whose r the loop must leave as the host-driven loop does."""
    text = (tdeck.layered_2d(64, 2, solver=tdeck.JACOBI, max_iters=max_iters) if dims == 2 else
            tdeck.cuboid_3d(12, 2, solver=tdeck.JACOBI, max_iters=max_iters))
    reps = _same(hip, text, dims=dims, extra=("r",))
    assert all(r["n_jacobi"] == max_iters and r["printed_cg_iters"] == max_iters for r in reps)


def test_jacobi_relative_stop(hip):
    reps = _same(hip, tdeck.layered_2d(40, 2, solver=tdeck.JACOBI, max_iters=3000), rel=1e-4,
                 extra=("r",))
    assert all(r["n_true_residuals"] >= 1 for r in reps)


def test_jacobi_blocks(hip):
    text = tdeck.layered_2d(64, 1, solver=tdeck.JACOBI, eps=1e-9, max_iters=3000)
    s, _ = tdeck.parse_deck(text, 2)
    inject = blocks.block_fields((64, 64), 16, seed=4, rho=(0.1, 10.0), energy=(0.1, 25.0),
                                 halo_depth=s.halo_depth)
    _same(hip, text, dims=2, inject=inject, extra=("r",))


@pytest.mark.parametrize("dims,text", [
    (2, tdeck.layered_2d(64, 2, solver=tdeck.CG, max_iters=3000)),
    (2, tdeck.layered_2d(64, 2, solver=tdeck.CHEBY, presteps=20, max_iters=3000)),
    (2, tdeck.layered_2d(64, 2, solver=tdeck.PPCG, presteps=20, inner=10, max_iters=3000)),
    (2, tdeck.layered_2d(64, 1, solver=tdeck.PPCG, presteps=20, inner=12, max_iters=41)),
    (3, tdeck.cuboid_3d(20, 2, solver=tdeck.CG, max_iters=2000)),
    (3, tdeck.cuboid_3d(20, 2, solver=tdeck.PPCG, presteps=15, max_iters=2000)),
], ids=["2d-cg", "2d-cheby", "2d-ppcg", "2d-ppcg-inner12-cut41", "3d-cg", "3d-ppcg"])
def test_graph_replays_match_eager(hip, dims, text):
    """Small grids can replay pairs of resident iterations from a HIP
    graph (tlhip_set_graphs; off by default); otherwise the same iterations
    are enqueued launch by launch.  Every report, u, energy, r and the summary
    must be the same bits; each solve replays the graph hundreds of times
    with the host's waits and the loop's own eager launches in between."""
    runs = []
    for on in (0, 1):
        hip.tlhip_set_graphs(on)
        try:
            runs.append(_run(hip, text, 1, dims, extra=("r", "p")))
        finally:
            hip.tlhip_set_graphs(0)
    a, b = runs
    assert len(a[0]) == len(b[0])
    for ra, rb in zip(a[0], b[0]):
        assert ra == rb, (ra, rb)
    for fa, fb in zip(a[1], b[1]):
        np.testing.assert_array_equal(fa, fb)
    assert a[2] == b[2]


def test_set_resident_returns_zero(hip):
    assert hip.dll.tlhip_set_resident(0) == 0
    assert hip.dll.tlhip_set_resident(1) == 0
