"""Solves to a relative residual, in FP64 and in mixed precision (GPU).

* FP64 with the relative-residual stop: the native driver against the
  oracle driver with the same stop -- printed iterations within +-1, field
  summary within 1e-12 -- and the true residual, recomputed here in numpy,
  at or below the target.
* Mixed precision (FP32 CG / Chebyshev / PPCG sweeps, FP64 residual
  replacement): the
  true residual, recomputed here in numpy from the returned u, at or below
  1e-10 ||u0||; the solution against the oracle's FP64 solve to the same
  target; and the work against the FP64 GPU solve (iteration counts within
  a factor).  These solves have no counterpart in the reference, so they are
  held to what the mathematics fixes -- the residual of the returned u --
  rather than to bitwise agreement.
"""
import numpy as np
import pytest

from tealeaf_amd import blocks
from tealeaf_amd import deck as tdeck
from tests import numerics
from tests.problem_helpers import printed_iters, rel

pytestmark = pytest.mark.gpu

TOL = 1e-10
FIELDS = (("u", 3), ("u0", 4), ("kx", 8), ("ky", 9), ("kz", 10))


def _inject_for(s, blk):
    if blk is None:
        return None
    cells = (s.x_cells, s.y_cells) if s.dims == 2 else (s.x_cells, s.y_cells, s.z_cells)
    return blocks.block_fields(cells, blk[0], seed=blk[2], rho=(1.0, blk[1]),
                               halo_depth=s.halo_depth)


def _gpu_solve(s, states, inject, precision, tol, steps=1):
    from tealeaf_amd.problem import Problem
    with Problem(s, states) as prob:
        if inject is not None:
            prob.inject(*inject)
        prob.set_solve(precision, tol)
        reps = prob.run(steps)
        f = {name: prob.get(name).reshape(s.shape()) for name, _ in FIELDS if name != "kz" or s.dims == 3}
        return reps, f, prob.field_summary()


def _residual(s, f):
    return numerics.rel_residual(f["u"], f["u0"], f["kx"], f["ky"], s.halo_depth, f.get("kz"))


REL_CASES = [
    ("2d-cg", tdeck.layered_2d(64, 2, solver=tdeck.CG, max_iters=5000), 2, None),
    ("2d-ppcg", tdeck.layered_2d(96, 2, solver=tdeck.PPCG, presteps=20, inner=10, max_iters=3000), 2, None),
    ("2d-ppcg-errswitch", tdeck.layered_2d(96, 1, solver=tdeck.PPCG, inner=10, max_iters=3000,
                                           dt=0.05, epslim=1e30), 2, None),
    ("2d-cheby", tdeck.layered_2d(64, 1, solver=tdeck.CHEBY, presteps=20, max_iters=3000), 2, None),
    ("3d-cg", tdeck.cuboid_3d(18, 2), 3, None),
    ("3d-ppcg", tdeck.cuboid_3d(20, 1, solver=tdeck.PPCG, inner=8, coefficient=2), 3, None),
    ("2d-ppcg-blocks", tdeck.layered_2d(80, 1, solver=tdeck.PPCG, presteps=20, inner=10,
                                        max_iters=4000, dt=0.5, epslim=1e30), 2, (16, 1e3, 9)),
]


@pytest.mark.parametrize("name,text,dims,blk", REL_CASES, ids=[c[0] for c in REL_CASES])
def test_fp64_relative_stop_matches_oracle(hip, name, text, dims, blk):
    from oracle.pyoracle import run_problem
    s, states = tdeck.parse_deck(text, dims)
    inject = _inject_for(s, blk)
    want = run_problem(s, states, inject=inject, rel_tol=TOL)
    reps, f, summary = _gpu_solve(s, states, inject, 0, TOL, steps=s.end_step)
    for a, b in zip(reps, want["steps"]):
        pa, pb = printed_iters(a, s.solver), printed_iters(b, s.solver)
        for k in pa:
            assert abs(pa[k] - pb[k]) <= 1, (k, pa, pb)
        assert a["rel_residual"] <= TOL
    for a, b in zip(summary, want["summary"]):
        assert rel(a, b) < 1e-12, (summary, want["summary"])
    assert _residual(s, f) <= TOL * (1 + 1e-6)


# Larger timesteps than the reference deck's give these small grids the
# conditioning of the 4096^2-8192^2 runs; the error switch with a large
# epslim starts PPCG after the reference's 20 CG iterations even on the
# block fields, whose absolute residuals stay far above the presteps rule's
# threshold of 1.
MIXED_CASES = [
    ("cg-layered", tdeck.layered_2d(64, 2, solver=tdeck.CG, max_iters=5000, dt=0.05), None),
    ("ppcg-layered", tdeck.layered_2d(96, 2, solver=tdeck.PPCG, presteps=20, inner=10,
                                      max_iters=3000, dt=0.05, epslim=1e30), None),
    ("ppcg-odd", tdeck.layered_2d(77, 1, solver=tdeck.PPCG, presteps=20, inner=7,
                                  max_iters=3000, dt=0.05, epslim=1e30), None),
    ("ppcg-contrast1e3", tdeck.layered_2d(128, 1, solver=tdeck.PPCG, presteps=30, inner=10,
                                          max_iters=6000, dt=1.0, epslim=1e30), (16, 1e3, 21)),
    ("ppcg-contrast1e6", tdeck.layered_2d(128, 1, solver=tdeck.PPCG, presteps=30, inner=10,
                                          max_iters=8000, dt=1.0, epslim=1e30), (16, 1e6, 22)),
    ("cg-contrast1e6", tdeck.layered_2d(128, 1, solver=tdeck.CG, max_iters=8000, dt=1.0),
     (16, 1e6, 23)),
    ("cheby-layered", tdeck.layered_2d(64, 2, solver=tdeck.CHEBY, presteps=20, max_iters=5000,
                                       dt=0.05, epslim=1e30), None),
    ("cheby-odd", tdeck.layered_2d(77, 1, solver=tdeck.CHEBY, presteps=20, max_iters=5000,
                                   dt=0.05, epslim=1e30), None),
    ("cheby-contrast1e3", tdeck.layered_2d(128, 1, solver=tdeck.CHEBY, presteps=30,
                                           max_iters=10000, dt=1.0, epslim=1e30), (16, 1e3, 24)),
    ("3d-cg", tdeck.cuboid_3d(24, 2, max_iters=5000), None),
    ("3d-ppcg", tdeck.cuboid_3d(24, 1, solver=tdeck.PPCG, inner=8, coefficient=2,
                                max_iters=5000), None),
    ("3d-cheby", tdeck.cuboid_3d(22, 1, solver=tdeck.CHEBY, max_iters=5000), None),
    ("3d-cg-blocks", tdeck.cuboid_3d(32, 1, max_iters=8000, dt=0.5), (8, 1e3, 26)),
]


def _work(reps):
    return sum(r["n_cg_steps"] + r["n_ppcg_outer"] + r["n_cheby_steps"] for r in reps)


@pytest.mark.parametrize("name,text,blk", MIXED_CASES, ids=[c[0] for c in MIXED_CASES])
def test_mixed_precision_reaches_fp64_residual(hip, name, text, blk):
    from oracle.pyoracle import run_problem
    s, states = tdeck.parse_deck(text, 3 if name.startswith("3d") else 2)
    inject = _inject_for(s, blk)
    reps_m, f_m, _ = _gpu_solve(s, states, inject, 1, TOL, steps=s.end_step)
    reps_d, f_d, _ = _gpu_solve(s, states, inject, 0, TOL, steps=s.end_step)
    want = run_problem(s, states, inject=inject, rel_tol=TOL, fields=FIELDS[:2])
    for r in reps_m:
        assert r["rel_residual"] <= TOL and r["n_true_residuals"] >= 1
    # the returned u meets the target, measured independently
    res_m = _residual(s, f_m)
    assert res_m <= TOL * (1 + 1e-6), res_m
    # the same state as the FP64 solves (u0 of the last step derives from it)
    u_ref = want["fields"]["u"].reshape(s.shape())
    inner = (slice(s.halo_depth, -s.halo_depth),) * s.dims
    err = np.linalg.norm((f_m["u"] - u_ref)[inner]) / np.linalg.norm(u_ref[inner])
    assert err < 1e-6, err
    # the same algorithm as the FP64 solve, and about as much of it
    if s.solver == tdeck.PPCG:
        assert all(r["n_ppcg_outer"] > 0 for r in reps_m + reps_d)
    if s.solver == tdeck.CHEBY:
        assert all(r["n_cheby_steps"] > 0 for r in reps_m + reps_d)
    it_m, it_d = _work(reps_m), _work(reps_d)
    assert it_m <= 1.25 * it_d + 10, (it_m, it_d)


# 1024^2: CG and PPCG on seeded block fields (64^2-cell blocks, log-uniform
# density over a 1e3 contrast, dt 1, PPCG after 20 CG iterations);
# Chebyshev on the reference's layered deck (presteps 30) -- Chebyshev needs
# the eigenvalue bounds that the reference's switch rule gives it.  The
# returned u meets 1e-10 (numpy residual) and lies within 1e-6 (relative) of
# the FP64 GPU solve to the same target.
LARGE_CASES = [("cg-blocks", tdeck.CG, True), ("cheby-layered", tdeck.CHEBY, False),
               ("ppcg-blocks", tdeck.PPCG, True)]


@pytest.mark.parametrize("name,solver,blk", LARGE_CASES, ids=[c[0] for c in LARGE_CASES])
def test_mixed_precision_1024(hip, name, solver, blk):
    if blk:
        text = tdeck.layered_2d(1024, 1, solver=solver, inner=10, max_iters=20000, dt=1.0,
                                epslim=1e30)
    else:
        text = tdeck.layered_2d(1024, 1, solver=solver, presteps=30, max_iters=20000)
    s, states = tdeck.parse_deck(text, 2)
    inject = (blocks.block_fields((1024, 1024), 64, seed=31, rho=(1.0, 1e3), energy=(0.1, 25.0),
                                  halo_depth=s.halo_depth) if blk else None)
    reps_m, f_m, _ = _gpu_solve(s, states, inject, 1, TOL)
    reps_d, f_d, _ = _gpu_solve(s, states, inject, 0, TOL)
    assert reps_m[0]["rel_residual"] <= TOL and reps_d[0]["rel_residual"] <= TOL
    res_m = _residual(s, f_m)
    assert res_m <= TOL * (1 + 1e-6), res_m
    inner = (slice(s.halo_depth, -s.halo_depth),) * 2
    err = np.linalg.norm((f_m["u"] - f_d["u"])[inner]) / np.linalg.norm(f_d["u"][inner])
    assert err < 1e-6, err
    assert _work(reps_m) <= 1.25 * _work(reps_d) + 10, (_work(reps_m), _work(reps_d))


def test_mixed_precision_rejects_unsupported(hip):
    from tealeaf_amd.abi import TealeafError
    from tealeaf_amd.problem import Problem
    with Problem.from_deck(tdeck.cuboid_3d(8, 1, solver=tdeck.JACOBI), 3) as prob:
        with pytest.raises(TealeafError):
            prob.set_solve(1, TOL)
    with Problem.from_deck(tdeck.layered_2d(16, 1, solver=tdeck.JACOBI), 2) as prob:
        with pytest.raises(TealeafError):
            prob.set_solve(1, TOL)
        with pytest.raises(TealeafError):
            prob.set_solve(0, -1.0)


# 3D at 128^3: CG on seeded 16^3-cell blocks (density log-uniform over a 1e3
# contrast), PPCG on the reference's cuboid deck (recip conductivity); the
# returned u meets 1e-10 (numpy residual) and lies within 1e-6 of the FP64
# GPU solve.
LARGE3_CASES = [
    ("cg-blocks", tdeck.cuboid_3d(128, 1, max_iters=20000, dt=0.5), True),
    ("ppcg", tdeck.cuboid_3d(128, 1, solver=tdeck.PPCG, inner=10, coefficient=2,
                             max_iters=20000), False),
]


@pytest.mark.parametrize("name,text,blk", LARGE3_CASES, ids=[c[0] for c in LARGE3_CASES])
def test_mixed_precision_3d_128(hip, name, text, blk):
    s, states = tdeck.parse_deck(text, 3)
    inject = (blocks.block_fields((128, 128, 128), 16, seed=41, rho=(1.0, 1e3),
                                  energy=(0.1, 25.0), halo_depth=s.halo_depth) if blk else None)
    reps_m, f_m, _ = _gpu_solve(s, states, inject, 1, TOL)
    reps_d, f_d, _ = _gpu_solve(s, states, inject, 0, TOL)
    assert reps_m[0]["rel_residual"] <= TOL and reps_d[0]["rel_residual"] <= TOL
    assert reps_m[0]["n_true_residuals"] >= 1
    res_m = _residual(s, f_m)
    assert res_m <= TOL * (1 + 1e-6), res_m
    inner = (slice(s.halo_depth, -s.halo_depth),) * 3
    err = np.linalg.norm((f_m["u"] - f_d["u"])[inner]) / np.linalg.norm(f_d["u"][inner])
    assert err < 1e-6, err
    assert _work(reps_m) <= 1.25 * _work(reps_d) + 10, (_work(reps_m), _work(reps_d))


@pytest.mark.parametrize("n,inner", [(20, 8), (26, 7), (24, 10)])
def test_mixed_3d_pairs_bit_identical(hip, n, inner):
    """The FP32-refined 3D PPCG with two FP32 inner steps per pass (fusion >
    1, even x) gives the same bits as one pass per step (fusion 1): u,
    summary, iteration counts and true-residual evaluations; an odd inner
    count ends on a single step."""
    text = tdeck.cuboid_3d(n, 2, solver=tdeck.PPCG, inner=inner, coefficient=2,
                           max_iters=4000)
    s, states = tdeck.parse_deck(text, 3)
    out = {}
    try:
        for fusion in (1, 10):
            hip.tlhip_set_ppcg_fusion(fusion)
            out[fusion] = _gpu_solve(s, states, None, 1, TOL, steps=2)
    finally:
        hip.tlhip_set_ppcg_fusion(10)
    (r1, f1, s1), (r10, f10, s10) = out[1], out[10]
    for key in ("n_cg_steps", "n_ppcg_outer", "n_ppcg_inner", "n_true_residuals"):
        assert [r[key] for r in r1] == [r[key] for r in r10], key
    assert sum(r["n_ppcg_outer"] for r in r10) > 0
    np.testing.assert_array_equal(f10["u"], f1["u"])
    assert s10 == s1
