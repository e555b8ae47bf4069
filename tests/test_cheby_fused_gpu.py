"""Temporally blocked Chebyshev loop (2D) against separate reference steps.

tlhip2d_cheby_iterations runs `count` turns of the Chebyshev loop of
2d/drivers/cheby_driver.c in one pass.  Its u, p and r must equal, bit for
bit, those of the oracle doing what the reference does per turn: a depth-1
reflective halo update of u (2d/c_kernels/omp3/local_halos.c), then
cheby_iterate and cheby_calc_u (2d/c_kernels/omp3/cheby.c:8-24,57-85).  The
halo rings of u and p pass through unchanged.  The driver with the fused
loop must give the same bits as with one launch per iteration, on every
way the Chebyshev phase can end.
"""
import ctypes

import numpy as np
import pytest

from oracle.pyoracle import ptr
from tests.fp32_emulation import fma32
from tealeaf_amd import blocks
from tealeaf_amd import deck as tdeck

pytestmark = pytest.mark.gpu

# (nx, ny, hd, tile_rows): nx even (x = nx + 2 hd even); 0 = default tiles.
GRIDS = [(32, 32, 2, 0), (250, 70, 2, 0), (300, 129, 2, 8), (6, 5, 2, 3),
         (124, 3, 2, 0), (118, 41, 4, 2), (460, 17, 2, 1), (64, 20, 1, 0), (122, 33, 3, 4)]


def _oracle_steps(O, x, y, hd, alphas, betas, u, u0, p, kx, ky):
    u, p = u.copy(), p.copy()
    r = np.zeros_like(u)
    w = np.zeros_like(u)
    nb = np.full(4, -1, dtype=np.int32)
    flags = np.zeros(6, dtype=np.bool_)
    flags[3] = True                                   # u only
    dummy = np.zeros(x * y)
    for a, b in zip(alphas, betas):
        O.or2d_local_halos(x, y, 1, hd, ptr(nb), ptr(flags), ptr(dummy), ptr(dummy), ptr(dummy),
                           ptr(u), ptr(dummy), ptr(dummy))
        O.or2d_cheby_iterate(x, y, hd, a, b, ptr(u), ptr(u0), ptr(p), ptr(r), ptr(w), ptr(kx),
                             ptr(ky))
    return u, p, r


def _interior(a, x, y, hd):
    return a.reshape(y, x)[hd:y - hd, hd:x - hd]


def _ring(a, x, y, hd):
    m = np.ones((y, x), dtype=bool)
    m[hd:y - hd, hd:x - hd] = False
    return a.reshape(y, x)[m]


@pytest.mark.parametrize("grid", GRIDS, ids=[f"{g[0]}x{g[1]}h{g[2]}t{g[3]}" for g in GRIDS])
@pytest.mark.parametrize("count", [1, 2, 3, 5, 7, 10])
@pytest.mark.parametrize("pf", [1, 0], ids=["pf", "nopf"])
def test_fused_cheby_matches_separate_steps(hip, oracle, grid, count, pf):
    from tealeaf_amd.device import DeviceArray
    nx, ny, hd, ty = grid
    x, y = nx + 2 * hd, ny + 2 * hd
    rng = np.random.default_rng(3000 * nx + 10 * count + hd)
    n = x * y
    u = rng.uniform(-1, 1, n)
    u0 = rng.uniform(-1, 1, n)
    p = rng.uniform(-1, 1, n)
    kx = rng.uniform(0.05, 0.3, n)
    ky = rng.uniform(0.05, 0.3, n)
    alphas = rng.uniform(0.1, 0.9, count)
    betas = rng.uniform(0.1, 0.9, count)
    want_u, want_p, want_r = _oracle_steps(oracle, x, y, hd, alphas, betas, u, u0, p, kx, ky)

    d = {k: DeviceArray.from_host(v) for k, v in dict(u=u, u0=u0, p=p, kx=kx, ky=ky).items()}
    out = {k: DeviceArray.from_host(np.full(n, np.nan)) for k in ("u", "p")}
    r_dev = DeviceArray.from_host(np.full(n, -7.0))
    nb = np.full(4, -1, dtype=np.int32)
    hip.tlhip_set_fused_tuning(ty if ty else -1, pf)
    try:
        hip.tlhip2d_cheby_iterations(x, y, hd, count, alphas.ctypes.data, betas.ctypes.data,
                                     nb.ctypes.data, d["u"].ptr, out["u"].ptr, d["u0"].ptr,
                                     d["p"].ptr, out["p"].ptr, r_dev.ptr, d["kx"].ptr, d["ky"].ptr)
    finally:
        hip.tlhip_set_fused_tuning(-1, -1)
    hip.tlhip_synchronize()
    got_u, got_p, got_r = out["u"].download(), out["p"].download(), r_dev.download()
    np.testing.assert_array_equal(_interior(got_u, x, y, hd), _interior(want_u, x, y, hd))
    np.testing.assert_array_equal(_interior(got_p, x, y, hd), _interior(want_p, x, y, hd))
    np.testing.assert_array_equal(_interior(got_r, x, y, hd), _interior(want_r, x, y, hd))
    np.testing.assert_array_equal(_ring(got_u, x, y, hd), _ring(u, x, y, hd))
    np.testing.assert_array_equal(_ring(got_p, x, y, hd), _ring(p, x, y, hd))
    assert np.all(_ring(got_r, x, y, hd) == -7.0)          # r's halo is not written
    for v in list(d.values()) + list(out.values()) + [r_dev]:
        v.free()


def _f32_turns(x, y, hd, alphas, betas, u, u0, p, kx, ky):
    """Chebyshev turns on (rows, pitch) float32 arrays, the FP32 kernel's
    arithmetic: the reference's expressions with the products that feed an
    addition fused (A u as in the FP32 PPCG loop, p = fma(alpha, p, beta*r))."""
    f = np.float32
    u, p = u.copy(), p.copy()
    r = np.zeros_like(u)
    J, K = slice(hd, y - hd), slice(hd, x - hd)
    for a, b in zip(alphas, betas):
        a, b = f(a), f(b)
        s = u.copy()                                      # depth-1 reflective halo
        s[J, hd - 1] = s[J, hd]
        s[J, x - hd] = s[J, x - hd - 1]
        s[hd - 1, K] = s[hd, K]
        s[y - hd, K] = s[y - hd - 1, K]
        cen = s[J, K]
        lft, rgt = s[J, hd - 1:x - hd - 1], s[J, hd + 1:x - hd + 1]
        dn, up = s[hd - 1:y - hd - 1, K], s[hd + 1:y - hd + 1, K]
        kxc, kxr = kx[J, K], kx[J, hd + 1:x - hd + 1]
        kyc, kyu = ky[J, K], ky[hd + 1:y - hd + 1, K]
        d = (f(1) + (kxr + kxc)) + (kyu + kyc)
        wv = fma32(d, cen, -fma32(kxr, rgt, kxc * lft)) - fma32(kyu, up, kyc * dn)
        r[J, K] = u0[J, K] - wv
        p[J, K] = fma32(a, p[J, K], b * r[J, K])
        u[J, K] = u[J, K] + p[J, K]
    return u, p, r


F32_GRIDS = [(32, 30, 2, 0), (250, 70, 2, 0), (510, 37, 2, 4), (6, 5, 2, 0), (61, 9, 1, 0),
             (1000, 21, 3, 8)]


@pytest.mark.parametrize("grid", F32_GRIDS, ids=[f"{g[0]}x{g[1]}h{g[2]}t{g[3]}" for g in F32_GRIDS])
@pytest.mark.parametrize("count", [1, 3, 10])
@pytest.mark.parametrize("norm", [False, True], ids=["plain", "norm"])
def test_fused_cheby_f32_matches_float32_turns(hip, grid, count, norm):
    from tealeaf_amd.device import DeviceArray
    nx, ny, hd, ty = grid
    x, y = nx + 2 * hd, ny + 2 * hd
    pitch = (x + 3) // 4 * 4
    shape = (y, pitch)
    rng = np.random.default_rng(9000 + nx + 10 * count)
    f32 = lambda lo, hi: rng.uniform(lo, hi, shape).astype(np.float32)  # noqa: E731
    u, u0, p = f32(-1, 1), f32(-1, 1), f32(-1, 1)
    kx, ky = f32(0.05, 0.3), f32(0.05, 0.3)
    alphas = rng.uniform(0.1, 0.9, count)
    betas = rng.uniform(0.1, 0.9, count)
    want_u, want_p, want_r = _f32_turns(x, y, hd, alphas, betas, u, u0, p, kx, ky)
    d = {k: DeviceArray.from_host(v) for k, v in dict(u=u, u0=u0, p=p, kx=kx, ky=ky).items()}
    out = {k: DeviceArray.from_host(np.full(shape, np.nan, dtype=np.float32))
           for k in ("u", "p", "r")}
    nb = np.full(4, -1, dtype=np.int32)
    rrn = ctypes.c_double(-1.0)
    hip.tlhip_set_fused_tuning(ty if ty else -1, -1)
    try:
        hip.tlhip2d_cheby_iterations_f32(x, y, hd, pitch, count, alphas.ctypes.data,
                                         betas.ctypes.data, nb.ctypes.data, d["u"].ptr,
                                         out["u"].ptr, d["u0"].ptr, d["p"].ptr, out["p"].ptr,
                                         out["r"].ptr, d["kx"].ptr, d["ky"].ptr,
                                         ctypes.addressof(rrn) if norm else None)
    finally:
        hip.tlhip_set_fused_tuning(-1, -1)
    hip.tlhip_synchronize()
    J, K = slice(hd, y - hd), slice(hd, x - hd)
    got = {k: v.download().reshape(shape) for k, v in out.items()}
    np.testing.assert_array_equal(got["u"][J, K], want_u[J, K])
    np.testing.assert_array_equal(got["p"][J, K], want_p[J, K])
    np.testing.assert_array_equal(got["r"][J, K], want_r[J, K])
    if norm:
        want_rr = float(np.sum(want_r[J, K].astype(np.float64) ** 2))
        assert abs(rrn.value - want_rr) <= 1e-13 * want_rr, (rrn.value, want_rr)
    for v in list(d.values()) + list(out.values()):
        v.free()


def test_fused_cheby_rejects_bad_arguments(hip):
    from tealeaf_amd.device import DeviceArray
    raw = hip.dll
    nb = np.full(4, -1, dtype=np.int32)
    a = np.ones(11)
    x, y = 16, 14
    bufs = [DeviceArray.from_host(np.zeros(x * y)) for _ in range(8)]
    q = [b.ptr for b in bufs]
    args = (q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7])
    assert raw.tlhip2d_cheby_iterations(x, y, 2, 11, a.ctypes.data, a.ctypes.data, nb.ctypes.data,
                                        *args) != 0                      # count above 10
    assert raw.tlhip2d_cheby_iterations(15, y, 2, 2, a.ctypes.data, a.ctypes.data, nb.ctypes.data,
                                        *args) != 0                      # odd x
    assert raw.tlhip2d_cheby_iterations(x, y, 2, 2, a.ctypes.data, a.ctypes.data, nb.ctypes.data,
                                        q[0], q[0], *q[2:]) != 0          # u in place
    nb[2] = 5                                                            # an internal face
    assert raw.tlhip2d_cheby_iterations(x, y, 2, 2, a.ctypes.data, a.ctypes.data, nb.ctypes.data,
                                        *args) != 0
    assert "external" in hip.last_error()
    for b in bufs:
        b.free()


def _run(hip, text, fusion, inject=None):
    from tealeaf_amd.problem import Problem
    hip.tlhip_set_cheby_fusion(fusion)
    try:
        with Problem.from_deck(text, 2) as prob:
            if inject is not None:
                prob.inject(*inject)
            steps = prob.run()
            return steps, prob.get("u"), prob.get("energy"), prob.field_summary()
    finally:
        hip.tlhip_set_cheby_fusion(10)


@pytest.mark.parametrize("case", ["layered", "errswitch", "cut_cg", "cut_cheby", "blocks"])
def test_driver_cheby_fusion_is_bit_identical(hip, case):
    """The driver with fused Chebyshev launches gives the same bits as one
    launch per iteration: the presteps and error switches, max_iters inside
    the CG phase and inside the Chebyshev phase (at 128^2 the oracle needs
    31 CG + 49 Chebyshev iterations), and seeded block densities."""
    inject = None
    if case == "layered":
        text = tdeck.layered_2d(96, 2, solver=tdeck.CHEBY, presteps=25, max_iters=3000)
    elif case == "errswitch":
        text = tdeck.layered_2d(80, 2, solver=tdeck.CHEBY, max_iters=3000, epslim=1e-3)
    elif case == "cut_cg":
        text = tdeck.layered_2d(128, 2, solver=tdeck.CHEBY, presteps=30, max_iters=25)
    elif case == "cut_cheby":
        text = tdeck.layered_2d(128, 2, solver=tdeck.CHEBY, presteps=30, max_iters=63)
    else:
        text = tdeck.layered_2d(128, 1, solver=tdeck.CHEBY, max_iters=4000, epslim=1e-4)
        s, _ = tdeck.parse_deck(text, 2)
        inject = blocks.block_fields((128, 128), 32, seed=9, rho=(0.01, 1.0), energy=(0.1, 25.0),
                                     halo_depth=s.halo_depth)
    a = _run(hip, text, 1, inject)
    b = _run(hip, text, 10, inject)
    assert len(a[0]) == len(b[0])
    for ra, rb in zip(a[0], b[0]):
        assert ra == rb, (ra, rb)
    np.testing.assert_array_equal(a[1], b[1])
    np.testing.assert_array_equal(a[2], b[2])
    assert a[3] == b[3]
    if case != "cut_cg":
        assert all(r["n_cheby_steps"] > 1 for r in b[0])


def test_set_cheby_fusion_range(hip):
    assert hip.dll.tlhip_set_cheby_fusion(11) != 0
    assert hip.dll.tlhip_set_cheby_fusion(-1) != 0
    assert hip.dll.tlhip_set_cheby_fusion(10) == 0
