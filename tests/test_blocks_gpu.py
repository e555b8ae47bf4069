"""Seeded block-density workloads on the GPU (3D 256^3 PPCG with
recip_conductivity on 16^3-cell blocks; smaller versions against the oracle).

* At 48^3 (same 16^3 blocks, same generator): the native driver against the
  oracle driver, iteration counts within +-1 and temp within 1e-11; vol, mass
  and ie (which the solve does not change) against exactly rounded sums
  (math.fsum) of the injected fields within 1e-13.  The reference sums the
  summary serially (field_summary.c:9-45), so over 110 592 cells its own
  rounding reaches a few 1e-12 (measured: vol 2.1e-12 below 1000); the
  GPU's tree reduction is closer to the exact sum.
* At the full 256^3 the oracle is too slow, so the run is held to what the
  mathematics fixes: each step's true residual (recomputed in numpy) at or
  below the 1e-10 relative stop, the volume and mass of the summary against
  numpy sums of the injected density, and conservation -- the reflective
  operator has zero column sums beyond the identity, so sum(u) over the
  interior is carried from step to step up to the residual.
"""
import numpy as np
import pytest

from tealeaf_amd import blocks
from tealeaf_amd import deck as tdeck
from tests import numerics
from tests.problem_helpers import printed_iters, rel

pytestmark = pytest.mark.gpu

RHO = (0.1, 100.0)
ENERGY = (0.1, 25.0)


def _fields(s, seed):
    return blocks.block_fields((s.x_cells, s.y_cells, s.z_cells), 16, seed=seed, rho=RHO,
                               energy=ENERGY, halo_depth=s.halo_depth)


def test_3d_blocks_match_oracle(hip):
    from oracle.pyoracle import run_problem
    from tealeaf_amd.problem import Problem
    text = tdeck.cuboid_3d(48, 2, solver=tdeck.PPCG, inner=10, coefficient=2, dt=0.004,
                           eps=1e-9, max_iters=3000)
    s, states = tdeck.parse_deck(text, 3)
    inject = _fields(s, 31)
    want = run_problem(s, states, inject=inject)
    with Problem(s, states) as p:
        p.inject(*inject)
        reps = p.run()
        summary = p.field_summary()
    assert any(r["n_ppcg_outer"] > 0 for r in reps)
    for a, b in zip(reps, want["steps"]):
        pa, pb = printed_iters(a, s.solver), printed_iters(b, s.solver)
        for k in pa:
            assert abs(pa[k] - pb[k]) <= 1, (k, pa, pb)
    assert rel(summary[3], want["summary"][3]) < 1e-11, (summary, want["summary"])
    import math
    hd = s.halo_depth
    inner = (slice(hd, -hd),) * 3
    cell = (s.xmax - s.xmin) / s.x_cells * (s.ymax - s.ymin) / s.y_cells * (s.zmax - s.zmin) / s.z_cells
    d, e = inject[0][inner].ravel(), inject[1][inner].ravel()
    exact = (math.fsum([cell] * d.size), math.fsum(cell * d), math.fsum(cell * d * e))
    for got, ex in zip(summary[:3], exact):
        assert rel(got, ex) < 1e-13, (summary, exact)


@pytest.mark.timeout(600)
def test_3d_256_blocks_ppcg(hip):
    from tealeaf_amd.problem import Problem
    steps = 5
    text = tdeck.cuboid_3d(256, steps, solver=tdeck.PPCG, inner=10, coefficient=2, dt=0.004,
                           max_iters=20000)
    s, states = tdeck.parse_deck(text, 3)
    dens, en = _fields(s, 256)
    hd = s.halo_depth
    inner = (slice(hd, -hd),) * 3
    with Problem(s, states) as p:
        p.inject(dens, en)
        p.set_solve(0, 1e-10)
        total0 = float(np.sum((en * dens)[inner]))
        for tt in range(steps):
            rep = p.step(tt)
            assert rep["n_ppcg_outer"] > 0 or tt > 0
            assert rep["rel_residual"] <= 1e-10
            f = {k: p.get(k).reshape(s.shape()) for k in ("u", "u0", "kx", "ky", "kz")}
            res = numerics.rel_residual(f["u"], f["u0"], f["kx"], f["ky"], hd, f["kz"])
            assert res <= 1e-10 * (1 + 1e-6), (tt, res)
            total = float(np.sum(f["u"][inner]))
            assert abs(total - total0) <= 1e-9 * total0, (tt, total, total0)
            print(f"256^3 step {tt}: CG {rep['n_cg_steps']} PPCG {rep['n_ppcg_outer']} "
                  f"true rel residual {res:.2e} sum(u) drift {abs(total - total0) / total0:.1e}",
                  flush=True)
        vol, mass, _, _ = p.field_summary()
    cell = (s.xmax - s.xmin) / s.x_cells * (s.ymax - s.ymin) / s.y_cells * (s.zmax - s.zmin) / s.z_cells
    assert rel(vol, cell * 256 ** 3) < 1e-12
    assert rel(mass, cell * float(np.sum(dens[inner]))) < 1e-12
