"""Pins the CPU oracle before anything is checked against it.

1. Kernels: bit-exact against the reference omp3 kernels compiled from
   source (oracle/_ref/libomp3_{2d,3d}.so, serial, -ffp-contract=off) and
   against the committed fixtures that build wrote (tests/golden/).
2. Closed forms that a dropped term, a wrong sign or a transposed operand
   would break: A*const = const under mirrored halos, symmetry of A,
   uniform-density coefficients, the exact solve of a cosine mode, and a
   dense numpy solve.
3. Drivers end to end: the reference's known answers (tea.problems) and the
   reference host binaries built from source.
"""
import ctypes

import numpy as np
import pytest

from tealeaf_amd import deck as tdeck
from tests import kernel_cases as kc
from oracle.pyoracle import load_oracle, load_ref
from tests.problem_helpers import (known_answer, oracle_run, printed_iters, ref_run,
                                   reference_check_passes, rel)

CASES = (kc.make_cases(2, 10, 7, seed=11) + kc.make_cases(2, 9, 6, seed=12, coefficient=2)
         + kc.make_cases(2, 11, 6, hd=1, seed=15) + kc.make_cases(2, 8, 9, hd=3, seed=16)
         + kc.make_cases(3, 6, 5, 4, seed=13) + kc.make_cases(3, 5, 6, 3, seed=14, coefficient=2)
         + kc.make_cases(3, 5, 4, 6, hd=1, seed=17)
         + [kc.state_case(2, 12, 10), kc.state_case(3, 8, 7, 6)])


def _prefix(c):
    return "or%dd_" % c.dims


def _same(a, b):
    for k in a:
        if isinstance(a[k], np.ndarray):
            assert np.array_equal(a[k], b[k]), k
        else:
            assert a[k] == b[k], k


@pytest.mark.parametrize("case", CASES, ids=[kc.case_id(c) for c in CASES])
def test_oracle_kernel_bit_exact_vs_reference(case):
    ref = load_ref(case.dims)
    if ref is None:
        pytest.skip("reference omp3 not built here (oracle/_ref)")
    got = kc.run_host(load_oracle().dll, _prefix(case), case)
    want = kc.run_host(ref, "", case, ref=True)
    _same(got, want)


def test_oracle_kernels_match_golden_fixtures():
    path = kc.GOLDEN_FILE
    data = np.load(path, allow_pickle=False)
    checked = 0
    for case in kc.golden_cases():
        kc.with_golden_inputs(case, data)
        got = kc.run_host(load_oracle().dll, _prefix(case), case)
        for k, v in got.items():
            key = f"{kc.case_id(case)}:{k}"
            want = data[key]
            assert np.array_equal(np.asarray(v), want), key
            checked += 1
    assert checked > 100


def test_set_chunk_data_closed_form():
    O = load_oracle()
    for dims in (2, 3):
        c = kc.data_case(dims, 8, 6, 5)
        got = kc.run_host(O.dll, _prefix(c), c)
        dx, dy = 10.0 / 8, 10.0 / 6
        i = np.arange(c.x + 1)
        assert np.array_equal(got["vertex_x"][: c.x + 1], 0.0 + dx * 3 + dx * (i - 2))
        assert np.array_equal(got["vertex_y"][: c.y + 1], -1.0 + dy * 5 + dy * (np.arange(c.y + 1) - 2))
        assert np.allclose(got["cell_x"][: c.x], 0.5 * (got["vertex_x"][: c.x] + got["vertex_x"][1: c.x + 1]))
        n = c.x * c.y * c.z
        vol = dx * dy * (10.0 / 5 if dims == 3 else 1.0)
        assert np.allclose(got["volume"][:n], vol, rtol=0, atol=1e-15)


# ---------------------------------------------------------------- closed forms

def _kx_ky_uniform(nx, ny, rho, rx, ry, coef, hd=2):
    """cg_init on a uniform density: kx = rx/rho (conductivity) or rx*rho (recip)."""
    O = load_oracle()
    x, y = nx + 2 * hd, ny + 2 * hd
    n = x * y
    d, e = np.full(n, rho), np.full(n, 2.0)
    u, p, r, w, kx, ky = (np.zeros(n) for _ in range(6))
    rro = ctypes.c_double(0)
    O.or2d_cg_init(x, y, hd, coef, rx, ry, ctypes.addressof(rro), d.ctypes.data, e.ctypes.data,
                   u.ctypes.data, p.ctypes.data, r.ctypes.data, w.ctypes.data, kx.ctypes.data,
                   ky.ctypes.data)
    return kx.reshape(y, x), ky.reshape(y, x), w.reshape(y, x), rro.value


def test_uniform_density_coefficients():
    for coef, want in ((1, lambda r, rho: r / rho), (2, lambda r, rho: r * rho)):
        kx, ky, _, rro = _kx_ky_uniform(9, 7, 4.0, 0.3, 0.7, coef)
        assert np.allclose(kx[2:-1, 2:-1], want(0.3, 4.0), rtol=1e-15)
        assert np.allclose(ky[2:-1, 2:-1], want(0.7, 4.0), rtol=1e-15)
        # A u = u for a constant u with mirrored halos -> r = u - Au = 0 except
        # where the (unmirrored) halo of u is zero-free: u = e*rho is constant
        # everywhere here, so r and rro vanish.
        assert abs(rro) < 1e-20


def _apply_dense(nx, ny, kx, ky, hd=2):
    """Dense matrix of the 5-point operator with mirrored (zero-flux) halos."""
    N = nx * ny
    A = np.zeros((N, N))
    for j in range(ny):
        for i in range(nx):
            row = i + j * nx
            jj, kk = j + hd, i + hd
            kc_, kr = kx[jj, kk], kx[jj, kk + 1]
            yc, yu = ky[jj, kk], ky[jj + 1, kk]
            A[row, row] += 1.0 + kr + kc_ + yu + yc
            for (di, dj, k) in ((1, 0, kr), (-1, 0, kc_), (0, 1, yu), (0, -1, yc)):
                i2, j2 = i + di, j + dj
                # mirrored halo: neighbour outside the interior is the cell itself
                i2 = min(max(i2, 0), nx - 1)
                j2 = min(max(j2, 0), ny - 1)
                A[row, i2 + j2 * nx] -= k
    return A


def _random_k(nx, ny, rng, hd=2):
    x, y = nx + 2 * hd, ny + 2 * hd
    rho = np.exp(rng.uniform(np.log(0.1), np.log(10), (y, x)))
    kx, ky, _, _ = _kx_ky_from(rho, 0.5, 0.8, hd)
    return kx, ky, rho


def _kx_ky_from(rho, rx, ry, hd):
    O = load_oracle()
    y, x = rho.shape
    n = x * y
    d = np.ascontiguousarray(rho.ravel())
    e = np.ones(n)
    u, p, r, w, kx, ky = (np.zeros(n) for _ in range(6))
    rro = ctypes.c_double(0)
    O.or2d_cg_init(x, y, hd, 1, rx, ry, ctypes.addressof(rro), d.ctypes.data, e.ctypes.data,
                   u.ctypes.data, p.ctypes.data, r.ctypes.data, w.ctypes.data, kx.ctypes.data,
                   ky.ctypes.data)
    return kx.reshape(y, x), ky.reshape(y, x), w, rro.value


def _mirror(f, hd=2):
    O = load_oracle()
    y, x = f.shape
    g = np.ascontiguousarray(f.ravel().copy())
    nb = np.full(4, -1, dtype=np.int32)
    flags = np.array([0, 0, 0, 0, 1, 0], dtype=np.bool_)
    O.or2d_local_halos(x, y, hd, hd, nb.ctypes.data, flags.ctypes.data, 0, 0, 0, 0,
                       g.ctypes.data, 0)
    return g.reshape(y, x)


def _calc_w(p, kx, ky, hd=2):
    O = load_oracle()
    y, x = p.shape
    pp = np.ascontiguousarray(p.ravel())
    w = np.zeros(x * y)
    pw = ctypes.c_double(0)
    O.or2d_cg_calc_w(x, y, hd, ctypes.addressof(pw), pp.ctypes.data, w.ctypes.data,
                     np.ascontiguousarray(kx.ravel()).ctypes.data,
                     np.ascontiguousarray(ky.ravel()).ctypes.data)
    return w.reshape(y, x), pw.value


def test_operator_closed_forms():
    rng = np.random.default_rng(5)
    nx, ny, hd = 11, 9, 2
    kx, ky, rho = _random_k(nx, ny, rng)
    # A * const = const (zero-flux boundaries)
    c = _mirror(np.full((ny + 4, nx + 4), 3.25))
    w, _ = _calc_w(c, kx, ky)
    assert np.allclose(w[2:-2, 2:-2], 3.25, rtol=1e-13)
    # against a dense assembly, and symmetry <Ax, z> = <x, Az>
    A = _apply_dense(nx, ny, kx, ky)
    assert np.allclose(A, A.T, rtol=0, atol=1e-12)
    xv = _mirror(rng.normal(size=(ny + 4, nx + 4)))
    zv = _mirror(rng.normal(size=(ny + 4, nx + 4)))
    Ax, xAx = _calc_w(xv, kx, ky)
    Az, _ = _calc_w(zv, kx, ky)
    dense = A @ xv[2:-2, 2:-2].ravel()
    assert np.allclose(Ax[2:-2, 2:-2].ravel(), dense, rtol=1e-12, atol=1e-12)
    assert np.isclose(np.sum(Ax[2:-2, 2:-2] * zv[2:-2, 2:-2]),
                      np.sum(xv[2:-2, 2:-2] * Az[2:-2, 2:-2]), rtol=1e-12)
    assert np.isclose(xAx, np.sum(Ax[2:-2, 2:-2] * xv[2:-2, 2:-2]), rtol=1e-13)


def test_cosine_mode_exact_solve():
    """Uniform rho: a cosine mode is an eigenvector, one CG solve divides it by
    lambda = 1 + 2rx(1-cos(pi k/nx))/rho + 2ry(1-cos(pi l/ny))/rho."""
    nx, ny, rho, dt = 32, 24, 2.0, 0.004
    # CG converges in two steps on two eigenvectors; stop before 0/0.
    text = tdeck.layered_2d(nx, 1, eps=1e-20, max_iters=500)
    s, states = tdeck.parse_deck(text)
    s.xmax, s.ymax, s.y_cells = 1.0, 1.0, ny
    kmode, lmode = 3, 2
    xs = (np.arange(nx) + 0.5) / nx
    ys = (np.arange(ny) + 0.5) / ny
    mode = 1.0 + 0.5 * np.outer(np.cos(np.pi * lmode * ys), np.cos(np.pi * kmode * xs))
    dens = np.full((ny + 4, nx + 4), rho)
    en = np.zeros((ny + 4, nx + 4))
    en[2:-2, 2:-2] = mode / rho
    en = _mirror(en)
    en[:2, :2] = en[:2, -2:] = en[-2:, :2] = en[-2:, -2:] = 0.0
    u = oracle_run_injected(s, states, dens, en)["u"][2:-2, 2:-2]
    rx, ry = dt / (1.0 / nx) ** 2, dt / (1.0 / ny) ** 2
    lam = 1 + 2 * rx * (1 - np.cos(np.pi * kmode / nx)) / rho + 2 * ry * (1 - np.cos(np.pi * lmode / ny)) / rho
    want = 1.0 + 0.5 * np.outer(np.cos(np.pi * lmode * ys), np.cos(np.pi * kmode * xs)) / lam
    assert np.allclose(u, want, rtol=1e-10, atol=1e-10)


def test_dense_solve_matches_cg():
    """One timestep on a 12^2 random-density grid equals numpy.linalg.solve."""
    rng = np.random.default_rng(9)
    nx = ny = 12
    text = tdeck.layered_2d(nx, 1, eps=1e-30, max_iters=400)
    s, states = tdeck.parse_deck(text)
    dens = np.exp(rng.uniform(np.log(0.1), np.log(10), (ny + 4, nx + 4)))
    dens = _mirror(dens)
    en = rng.uniform(0.5, 2.0, (ny + 4, nx + 4))
    en = _mirror(en)
    res = oracle_run_injected(s, states, dens, en)
    rx, ry = s.dt_init / s.dx ** 2, s.dt_init / s.dy ** 2
    kx, ky, _, _ = _kx_ky_from(dens, rx, ry, 2)
    A = _apply_dense(nx, ny, kx, ky)
    b = (en * dens)[2:-2, 2:-2].ravel()
    want = np.linalg.solve(A, b)
    assert np.allclose(res["u"][2:-2, 2:-2].ravel(), want, rtol=1e-11)


def oracle_run_injected(s, states, dens, en):
    O = load_oracle()
    cs, cst = tdeck.c_settings(s), tdeck.c_states(states)
    P = O.or_problem_create(ctypes.addressof(cs), ctypes.addressof(cst), len(states))
    try:
        assert O.or_problem_field_len(P) == dens.size == en.size
        O.or_problem_inject(P, np.ascontiguousarray(dens.ravel()).ctypes.data,
                            np.ascontiguousarray(en.ravel()).ctypes.data)
        rep = tdeck.CStepReport()
        assert O.or_problem_step(P, 0, ctypes.addressof(rep)) == 0
        u = np.zeros(dens.size)
        O.or_problem_get_field(P, 3, u.ctypes.data)
    finally:
        O.or_problem_destroy(P)
    return {"u": u.reshape(dens.shape), "rep": rep.as_dict()}


def test_cheby_coefs_recurrence():
    """Chebyshev coefficients against the closed form of the recurrence:
    rho_k = 1/(2 sigma - rho_{k-1}) with rho_0 = 1/sigma is T_{k}(sigma)/T_{k+1}(sigma)."""
    O = load_oracle()
    emin, emax, n = 1.5, 900.0, 25
    al, be = np.zeros(n), np.zeros(n)
    theta = ctypes.c_double(0)
    O.or_cheby_coefs(ctypes.c_double(emin), ctypes.c_double(emax), n, ctypes.addressof(theta),
                     al.ctypes.data, be.ctypes.data)
    th = (emax + emin) / 2
    de = (emax - emin) / 2
    sig = th / de
    T = np.polynomial.chebyshev.chebval
    rho = [T(sig, [0] * k + [1]) / T(sig, [0] * (k + 1) + [1]) for k in range(0, n + 2)]
    assert theta.value == th
    for i in range(n):
        assert np.isclose(al[i], rho[i + 1] * rho[i], rtol=1e-12)
        assert np.isclose(be[i], 2 * rho[i + 1] / de, rtol=1e-12)


def test_eigenvalues_of_lanczos_matrix():
    """The QL eigenvalues equal numpy's for the CG Lanczos tridiagonal."""
    O = load_oracle()
    rng = np.random.default_rng(2)
    n = 30
    al = rng.uniform(0.01, 1.0, n)
    be = rng.uniform(0.01, 0.9, n)
    d = 1 / al + np.concatenate([[0], be[:-1] / al[:-1]])
    e = np.sqrt(be[:-1]) / al[:-1]
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    ev = np.linalg.eigvalsh(T)
    lo, hi = ctypes.c_double(), ctypes.c_double()
    assert O.or_eigenvalues(al.ctypes.data, be.ctypes.data, n, ctypes.addressof(lo),
                            ctypes.addressof(hi)) == 0
    assert np.isclose(lo.value, ev.min() * 0.95, rtol=1e-12)
    assert np.isclose(hi.value, ev.max() * 1.05, rtol=1e-12)


# ---------------------------------------------------------------- end to end

@pytest.mark.parametrize("dims,n,steps", [(2, 64, 1), (3, 25, 20)])
def test_oracle_known_answers(dims, n, steps):
    text = tdeck.layered_2d(n, steps, max_iters=5000) if dims == 2 else \
        tdeck.cuboid_3d(n, steps)
    res = oracle_run(text, dims)
    want = known_answer(res["settings"])
    assert want is not None
    assert reference_check_passes(res["summary"][3], want), (res["summary"][3], want)


E2E = [
    ("2d", tdeck.layered_2d(64, 2, max_iters=5000)),
    ("2d", tdeck.layered_2d(96, 3, solver=tdeck.CHEBY, presteps=20, max_iters=2000)),
    ("2d", tdeck.layered_2d(96, 2, solver=tdeck.PPCG, presteps=20, inner=12, max_iters=2000)),
    ("2d", tdeck.layered_2d(40, 2, solver=tdeck.JACOBI, eps=1e-8, max_iters=400)),
    ("3d", tdeck.cuboid_3d(20, 3)),
    ("3d", tdeck.cuboid_3d(20, 2, solver=tdeck.PPCG, inner=8, coefficient=2)),
    ("3d", tdeck.cuboid_3d(20, 2, solver=tdeck.CHEBY)),
]


@pytest.mark.parametrize("dim,text", E2E, ids=[f"{d}-{i}" for i, (d, _) in enumerate(E2E)])
def test_oracle_matches_reference_binary(dim, text):
    try:
        ref = ref_run(f"tealeaf{dim}_omp3", text, threads=2)
    except FileNotFoundError:
        pytest.skip("reference binary not built here (oracle/_ref)")
    dims = int(dim[0])
    res = oracle_run(text, dims)
    s = res["settings"]
    assert len(ref["steps"]) == len(res["steps"])
    for r_step, o_step in zip(ref["steps"], res["steps"]):
        mine = printed_iters(o_step, s.solver)
        for k, v in mine.items():
            assert abs(v - r_step["iters"][k]) <= 1, (k, v, r_step["iters"])
    for a, b in zip(res["summary"], ref["summary"]):
        assert rel(a, b) < 1e-12, (res["summary"], ref["summary"])
