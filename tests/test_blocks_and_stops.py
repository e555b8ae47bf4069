"""CPU checks of the input generator and of the relative-residual stop.

* tealeaf_amd.blocks: SplitMix64 against its published output sequences, and
  the block layout of the generated density / energy fields.
* The oracle's relative-residual stop (an extension of the reference's eps
  test used to time solves to ||u0 - A u|| <= 1e-10 ||u0||): the solution it
  returns is checked with an independent numpy statement of A.
"""
import numpy as np
import pytest

from tealeaf_amd import blocks
from tealeaf_amd import deck as tdeck
from tests import numerics

# Published SplitMix64 outputs: seed 0 (Vigna's splitmix64.c reference
# sequence) and seed 1234567 (the sequence used by common test suites).
SEED0 = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
SEED1234567 = [6457827717110365317, 3203168211198807973, 9817491932198370423,
               4593380528125082431, 16408922859458223821]


def test_splitmix64_published_sequences():
    assert [int(v) for v in blocks.splitmix64(0, np.arange(1, 4))] == SEED0
    assert [int(v) for v in blocks.splitmix64(1234567, np.arange(1, 6))] == SEED1234567


def test_uniform_range_and_determinism():
    u = blocks.uniform(42, np.arange(1, 100001, dtype=np.uint64))
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.01
    np.testing.assert_array_equal(u, blocks.uniform(42, np.arange(1, 100001, dtype=np.uint64)))


def test_block_fields_2d_layout():
    nx, ny, b, hd = 40, 30, 16, 2
    d, e = blocks.block_fields((nx, ny), b, seed=7, rho=(1.0, 1e3), halo_depth=hd)
    assert d.shape == e.shape == (ny + 2 * hd, nx + 2 * hd)
    nbx, nby = 3, 2
    dens_b, en_b = blocks.block_values(7, nbx * nby, (1.0, 1e3), (0.1, 25.0))
    for j in range(ny):
        for k in range(nx):
            bid = (j // b) * nbx + (k // b)
            assert d[j + hd, k + hd] == dens_b[bid] and e[j + hd, k + hd] == en_b[bid]
    # halos repeat the nearest interior cell
    np.testing.assert_array_equal(d[:hd, hd:-hd], np.repeat(d[hd:hd + 1, hd:-hd], hd, axis=0))
    np.testing.assert_array_equal(d[hd:-hd, -hd:], np.repeat(d[hd:-hd, -hd - 1:-hd], hd, axis=1))
    assert d.min() >= 1.0 and d.max() <= 1e3
    assert e.min() >= 0.1 and e.max() <= 25.0


def test_block_fields_3d_ids_are_z_major():
    d, _ = blocks.block_fields((20, 18, 17), 16, seed=3, halo_depth=1)
    dens_b, _ = blocks.block_values(3, 2 * 2 * 2, (0.1, 100.0), (0.1, 25.0))
    # block (bx, by, bz) = (1, 0, 1) -> id (1 * 2 + 0) * 2 + 1 = 5
    assert d[1 + 16, 1 + 0, 1 + 16] == dens_b[5]
    assert d[1, 1, 1] == dens_b[0]


def test_log_uniform_density_spans_contrast():
    dens, _ = blocks.block_values(11, 4096, (1.0, 1e6), (0.1, 25.0))
    lg = np.log10(dens)
    assert lg.min() < 0.1 and lg.max() > 5.9
    assert abs(lg.mean() - 3.0) < 0.15


STOP_CASES = [
    ("2d-cg", tdeck.layered_2d(48, 1, solver=tdeck.CG, max_iters=5000), 2, None),
    ("2d-ppcg", tdeck.layered_2d(64, 1, solver=tdeck.PPCG, presteps=20, inner=10, max_iters=2000), 2, None),
    ("2d-cheby", tdeck.layered_2d(64, 1, solver=tdeck.CHEBY, presteps=20, max_iters=3000), 2, None),
    ("3d-cg", tdeck.cuboid_3d(14, 1), 3, None),
    ("2d-ppcg-blocks", tdeck.layered_2d(64, 1, solver=tdeck.PPCG, presteps=20, inner=10,
                                        max_iters=4000), 2, (16, 1e3)),
]


@pytest.mark.parametrize("name,text,dims,blk", STOP_CASES, ids=[c[0] for c in STOP_CASES])
def test_oracle_relative_residual_stop(name, text, dims, blk):
    from oracle.pyoracle import run_problem
    s, states = tdeck.parse_deck(text, dims)
    inject = None
    if blk is not None:
        cells = (s.x_cells, s.y_cells) if dims == 2 else (s.x_cells, s.y_cells, s.z_cells)
        inject = blocks.block_fields(cells, blk[0], seed=5, rho=(1.0, blk[1]), halo_depth=s.halo_depth)
    tol = 1e-10
    res = run_problem(s, states, steps=1, inject=inject, rel_tol=tol,
                      fields=(("u", 3), ("u0", 4), ("kx", 8), ("ky", 9), ("kz", 10)))
    rep = res["steps"][0]
    f = {k: v.reshape(s.shape()) for k, v in res["fields"].items()}
    got = numerics.rel_residual(f["u"], f["u0"], f["kx"], f["ky"], s.halo_depth,
                                f["kz"] if dims == 3 else None)
    assert rep["n_true_residuals"] >= 1
    assert got <= tol * (1 + 1e-6), got
    assert abs(got - rep["rel_residual"]) <= 1e-6 * tol + 1e-3 * got


def test_oracle_more_inner_steps_than_max_iters():
    """A PPCG deck whose 10 inner steps exceed max_iters (5): the reference
    sizes its coefficient arrays max_iters and would overrun them; the
    oracle holds both and runs the deck to the max_iters cut, PPCG iterations
    included (the switch comes after presteps = 1)."""
    from tests.problem_helpers import oracle_run
    res = oracle_run(tdeck.layered_2d(48, 2, solver=tdeck.PPCG, presteps=1, max_iters=5))
    counts = [(s["printed_cg_iters"], s["printed_inner_iters"]) for s in res["steps"]]
    assert all(cg + inner >= 5 for cg, inner in counts), counts
    assert all(inner > 0 for _, inner in counts), counts
