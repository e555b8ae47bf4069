"""The CG pair of passes with kx and ky formed from the conductivity.

tlhip2d_cg_calc_pw_cond and tlhip2d_cg_calc_ur_cond read cond (cg_init's
intermediate w: density, or 1/density for the reciprocal coefficient,
2d/c_kernels/omp3/cg.c:45-53) instead of kx and ky and form every face
coefficient with cg_init's expression (cg.c:55-64).  Against the oracle's
cg_init (kx, ky), calc_p, a depth-1 reflective halo update of p, calc_w and
calc_ur (cg.c:88-164) on the same inputs: p, w, u and r bit for bit, the
dot products within 1e-13 (summation order).  Grids cover several strips and
row blocks, partial ones, halo depths 2 and 3, and both coefficients.  The
3D pair (tlhip3d_cg_*_cond, kz too; 3d/c_kernels/omp3/cg_solver.c) is held
to the same bar against the oracle's 3D kernels, on grids with partial row
blocks and short z-tiles; the driver with either pair agrees with its
kx/ky(/kz) passes to rounding.
"""
import numpy as np
import pytest

from oracle.pyoracle import ptr

pytestmark = pytest.mark.gpu

# (nx, ny, hd, coefficient); x = nx + 2 hd even
GRIDS = [(64, 48, 2, 1), (64, 48, 2, 2), (250, 37, 2, 2), (126, 9, 3, 1), (6, 5, 2, 2),
         (300, 130, 2, 1), (1000, 70, 3, 2)]
RX, RY = 0.37, 0.21


def _halo_p(O, x, y, hd, p):
    nb = np.full(4, -1, dtype=np.int32)
    flags = np.zeros(6, dtype=np.bool_)
    flags[4] = True                                   # p only
    dummy = np.zeros(x * y)
    O.or2d_local_halos(x, y, 1, hd, ptr(nb), ptr(flags), ptr(dummy), ptr(dummy), ptr(dummy),
                       ptr(dummy), ptr(p), ptr(dummy))


def _inputs(O, nx, ny, hd, coef):
    x, y = nx + 2 * hd, ny + 2 * hd
    n = x * y
    rng = np.random.default_rng(900 + nx + 7 * ny + hd + coef)
    dens = rng.uniform(0.1, 10.0, n)
    energy = rng.uniform(0.5, 2.0, n)
    u, p, r, w, kx, ky = (np.zeros(n) for _ in range(6))
    rro = np.zeros(1)
    O.or2d_cg_init(x, y, hd, coef, RX, RY, ptr(rro), ptr(dens), ptr(energy), ptr(u), ptr(p),
                   ptr(r), ptr(w), ptr(kx), ptr(ky))
    f = dict(dens=dens, kx=kx, ky=ky, p=rng.uniform(-1, 1, n), r=rng.uniform(-1, 1, n),
             u=rng.uniform(-1, 1, n))
    return x, y, f, rng


def _interior(a, x, y, hd):
    return a.reshape(y, x)[hd:y - hd, hd:x - hd]


@pytest.mark.parametrize("grid", GRIDS, ids=[f"{g[0]}x{g[1]}h{g[2]}c{g[3]}" for g in GRIDS])
def test_calc_pw_cond_matches_reference(hip, oracle, grid):
    from tealeaf_amd.device import DeviceArray
    nx, ny, hd, coef = grid
    O = oracle
    x, y, f, rng = _inputs(O, nx, ny, hd, coef)
    n = x * y
    beta = float(rng.uniform(0.1, 0.9))
    # reference: calc_p, halo update of p, calc_w
    p1 = f["p"].copy()
    O.or2d_cg_calc_p(x, y, hd, beta, ptr(p1), ptr(f["r"]))
    _halo_p(O, x, y, hd, p1)
    w_ref, pw_ref = np.zeros(n), np.zeros(1)
    O.or2d_cg_calc_w(x, y, hd, ptr(pw_ref), ptr(p1), ptr(w_ref), ptr(f["kx"]), ptr(f["ky"]))
    # calc_w alone on p (after its halo update)
    p0 = f["p"].copy()
    _halo_p(O, x, y, hd, p0)
    w0_ref, pw0_ref = np.zeros(n), np.zeros(1)
    O.or2d_cg_calc_w(x, y, hd, ptr(pw0_ref), ptr(p0), ptr(w0_ref), ptr(f["kx"]), ptr(f["ky"]))

    d = {k: DeviceArray.from_host(f[k]) for k in ("dens", "p", "r")}
    cond = DeviceArray.from_host(np.full(n, np.nan))
    p_out = DeviceArray.from_host(np.full(n, np.nan))
    w = DeviceArray.from_host(np.full(n, np.nan))
    hip.tlhip2d_cg_cond(x, y, hd, coef, d["dens"].ptr, cond.ptr)
    pw, pw_nw, pw0 = np.zeros(1), np.zeros(1), np.zeros(1)
    hip.tlhip2d_cg_calc_pw_cond(x, y, hd, 1, beta, RX, RY, ptr(pw), d["p"].ptr, p_out.ptr,
                                d["r"].ptr, w.ptr, cond.ptr)
    got_p, got_w = p_out.download(), w.download()
    np.testing.assert_array_equal(_interior(got_p, x, y, hd), _interior(p1, x, y, hd))
    np.testing.assert_array_equal(_interior(got_w, x, y, hd), _interior(w_ref, x, y, hd))
    assert abs(pw[0] - pw_ref[0]) <= 1e-13 * abs(pw_ref[0])
    # w not stored: the same p_out and the same bits of pw
    p_out2 = DeviceArray.from_host(np.full(n, np.nan))
    hip.tlhip2d_cg_calc_pw_cond(x, y, hd, 1, beta, RX, RY, ptr(pw_nw), d["p"].ptr, p_out2.ptr,
                                d["r"].ptr, None, cond.ptr)
    np.testing.assert_array_equal(p_out2.download(), got_p)
    assert pw_nw[0] == pw[0]
    # no update: calc_w on p
    w0 = DeviceArray.from_host(np.full(n, np.nan))
    hip.tlhip2d_cg_calc_pw_cond(x, y, hd, 0, 0.0, RX, RY, ptr(pw0), d["p"].ptr, None, None,
                                w0.ptr, cond.ptr)
    np.testing.assert_array_equal(_interior(w0.download(), x, y, hd), _interior(w0_ref, x, y, hd))
    assert abs(pw0[0] - pw0_ref[0]) <= 1e-13 * abs(pw0_ref[0])
    for v in list(d.values()) + [cond, p_out, p_out2, w, w0]:
        v.free()


@pytest.mark.parametrize("grid", GRIDS, ids=[f"{g[0]}x{g[1]}h{g[2]}c{g[3]}" for g in GRIDS])
def test_calc_ur_cond_matches_reference(hip, oracle, grid):
    from tealeaf_amd.device import DeviceArray
    nx, ny, hd, coef = grid
    O = oracle
    x, y, f, rng = _inputs(O, nx, ny, hd, coef)
    n = x * y
    alpha = float(rng.uniform(0.1, 0.9))
    # reference: w = A p after p's halo update, then calc_ur
    p = f["p"].copy()
    _halo_p(O, x, y, hd, p)
    w_ref, pw_ref = np.zeros(n), np.zeros(1)
    O.or2d_cg_calc_w(x, y, hd, ptr(pw_ref), ptr(p), ptr(w_ref), ptr(f["kx"]), ptr(f["ky"]))
    u_ref, r_ref, rrn_ref = f["u"].copy(), f["r"].copy(), np.zeros(1)
    O.or2d_cg_calc_ur(x, y, hd, alpha, ptr(rrn_ref), ptr(u_ref), ptr(p), ptr(r_ref), ptr(w_ref))

    d = {k: DeviceArray.from_host(f[k]) for k in ("dens", "p", "r", "u")}
    cond = DeviceArray.from_host(np.full(n, np.nan))
    hip.tlhip2d_cg_cond(x, y, hd, coef, d["dens"].ptr, cond.ptr)
    rrn = np.zeros(1)
    hip.tlhip2d_cg_calc_ur_cond(x, y, hd, alpha, RX, RY, ptr(rrn), d["u"].ptr, d["p"].ptr,
                                d["r"].ptr, cond.ptr)
    np.testing.assert_array_equal(d["u"].download(), u_ref)     # halo untouched
    np.testing.assert_array_equal(d["r"].download(), r_ref)
    assert abs(rrn[0] - rrn_ref[0]) <= 1e-13 * abs(rrn_ref[0])
    for v in list(d.values()) + [cond]:
        v.free()


def test_cond_pair_rejects_bad_grids(hip):
    from tealeaf_amd.device import DeviceArray
    a = [DeviceArray.from_host(np.ones(20 * 20)) for _ in range(5)]
    pw = np.zeros(1)
    L = hip.dll
    # halo depth 1 (cg_init's kx there reads a stale w), odd x
    assert L.tlhip2d_cg_calc_pw_cond(20, 20, 1, 1, 0.5, RX, RY, ptr(pw), a[0].ptr, a[1].ptr,
                                     a[2].ptr, a[3].ptr, a[4].ptr) != 0
    assert L.tlhip2d_cg_calc_ur_cond(19, 20, 2, 0.5, RX, RY, ptr(pw), a[0].ptr, a[1].ptr,
                                     a[2].ptr, a[4].ptr) != 0
    assert L.tlhip2d_cg_calc_ur_cond(20, 20, 2, 0.5, RX, RY, ptr(pw), a[0].ptr, a[1].ptr,
                                     a[0].ptr, a[4].ptr) != 0
    assert L.tlhip2d_cg_cond(20, 20, 2, 3, a[0].ptr, a[1].ptr) != 0
    for v in a:
        v.free()


def _run(hip, text, cond, resident):
    from tealeaf_amd.problem import Problem
    hip.tlhip_set_cg_cond(int(cond))
    hip.tlhip_set_resident(int(resident))
    try:
        with Problem.from_deck(text, 2) as prob:
            steps = prob.run()
            return steps, prob.get("u"), prob.get("energy"), prob.field_summary()
    finally:
        hip.tlhip_set_cg_cond(1)
        hip.tlhip_set_resident(1)


@pytest.mark.parametrize("solver", ["CG", "CHEBY", "PPCG"])
def test_driver_cond_pair_agrees(hip, solver):
    """The driver with the cond pair (the default) against kx/ky passes:
    the resident and host-driven loops give the same bits with it; against
    the kx/ky passes (rrn summed in another order) the iteration counts
    agree within one and u, energy and the summary within 1e-12 relative."""
    from tealeaf_amd import deck as tdeck
    text = tdeck.layered_2d(96, 3, solver=getattr(tdeck, solver), presteps=20, inner=10,
                            max_iters=4000)
    on = _run(hip, text, True, True)
    host = _run(hip, text, True, False)
    off = _run(hip, text, False, True)
    np.testing.assert_array_equal(host[1], on[1])
    assert host[3] == on[3]
    for key in ("n_cg_steps", "n_cheby_steps", "n_ppcg_outer"):
        a = [s.get(key, 0) for s in on[0]]
        b = [s.get(key, 0) for s in off[0]]
        assert all(abs(i - j) <= 1 for i, j in zip(a, b)), (key, a, b)
    for got, want in ((on[1], off[1]), (on[2], off[2])):
        assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))
    for a, b in zip(on[3], off[3]):
        assert abs(a - b) <= 1e-12 * abs(b), (on[3], off[3])


# 3D: (nx, ny, nz, hd, coefficient, planes per z-tile; 0 = default); x even
GRIDS3 = [(24, 20, 18, 2, 1, 0), (250, 9, 7, 2, 2, 0), (126, 17, 33, 3, 1, 5),
          (6, 5, 4, 2, 2, 1), (368, 6, 21, 2, 1, 16), (60, 31, 12, 3, 2, 3)]
RZ = 0.29


def _halo_p3(O, x, y, z, hd, p):
    nb = np.full(6, -1, dtype=np.int32)
    flags = np.zeros(6, dtype=np.bool_)
    flags[4] = True                                   # p only
    dummy = np.zeros(x * y * z)
    O.or3d_local_halos(x, y, z, 1, hd, ptr(nb), ptr(flags), ptr(dummy), ptr(dummy), ptr(dummy),
                       ptr(dummy), ptr(p), ptr(dummy))


def _inputs3(O, nx, ny, nz, hd, coef):
    x, y, z = nx + 2 * hd, ny + 2 * hd, nz + 2 * hd
    n = x * y * z
    rng = np.random.default_rng(1900 + nx + 7 * ny + 13 * nz + hd + coef)
    dens = rng.uniform(0.1, 10.0, n)
    energy = rng.uniform(0.5, 2.0, n)
    u, p, r, w, kx, ky, kz = (np.zeros(n) for _ in range(7))
    rro = np.zeros(1)
    O.or3d_cg_solver_init(x, y, z, hd, coef, RX, RY, RZ, ptr(rro), ptr(dens), ptr(energy), ptr(u),
                          ptr(p), ptr(r), ptr(w), ptr(kx), ptr(ky), ptr(kz))
    f = dict(dens=dens, kx=kx, ky=ky, kz=kz, p=rng.uniform(-1, 1, n), r=rng.uniform(-1, 1, n),
             u=rng.uniform(-1, 1, n))
    return x, y, z, f, rng


def _interior3(a, x, y, z, hd):
    return a.reshape(z, y, x)[hd:z - hd, hd:y - hd, hd:x - hd]


@pytest.mark.parametrize("grid", GRIDS3, ids=[f"{g[0]}x{g[1]}x{g[2]}h{g[3]}c{g[4]}t{g[5]}"
                                              for g in GRIDS3])
def test_calc_pw_cond3_matches_reference(hip, oracle, grid):
    from tealeaf_amd.device import DeviceArray
    nx, ny, nz, hd, coef, tz = grid
    O = oracle
    x, y, z, f, rng = _inputs3(O, nx, ny, nz, hd, coef)
    n = x * y * z
    beta = float(rng.uniform(0.1, 0.9))
    p1 = f["p"].copy()
    O.or3d_cg_solver_calc_p(x, y, z, hd, beta, ptr(p1), ptr(f["r"]))
    _halo_p3(O, x, y, z, hd, p1)
    w_ref, pw_ref = np.zeros(n), np.zeros(1)
    O.or3d_cg_solver_calc_w(x, y, z, hd, ptr(pw_ref), ptr(p1), ptr(w_ref), ptr(f["kx"]),
                            ptr(f["ky"]), ptr(f["kz"]))
    p0 = f["p"].copy()
    _halo_p3(O, x, y, z, hd, p0)
    w0_ref, pw0_ref = np.zeros(n), np.zeros(1)
    O.or3d_cg_solver_calc_w(x, y, z, hd, ptr(pw0_ref), ptr(p0), ptr(w0_ref), ptr(f["kx"]),
                            ptr(f["ky"]), ptr(f["kz"]))

    d = {k: DeviceArray.from_host(f[k]) for k in ("dens", "p", "r")}
    cond = DeviceArray.from_host(np.full(n, np.nan))
    p_out, p_out2, w, w0 = (DeviceArray.from_host(np.full(n, np.nan)) for _ in range(4))
    hip.tlhip3d_cg_cond(x, y, z, hd, coef, d["dens"].ptr, cond.ptr)
    pw, pw_nw, pw0 = np.zeros(1), np.zeros(1), np.zeros(1)
    hip.tlhip_set_tile_planes(tz if tz else -1)
    try:
        hip.tlhip3d_cg_calc_pw_cond(x, y, z, hd, 1, beta, RX, RY, RZ, ptr(pw), d["p"].ptr,
                                    p_out.ptr, d["r"].ptr, w.ptr, cond.ptr)
        hip.tlhip3d_cg_calc_pw_cond(x, y, z, hd, 1, beta, RX, RY, RZ, ptr(pw_nw), d["p"].ptr,
                                    p_out2.ptr, d["r"].ptr, None, cond.ptr)
        hip.tlhip3d_cg_calc_pw_cond(x, y, z, hd, 0, 0.0, RX, RY, RZ, ptr(pw0), d["p"].ptr, None,
                                    None, w0.ptr, cond.ptr)
    finally:
        hip.tlhip_set_tile_planes(-1)
    got_p = p_out.download()
    np.testing.assert_array_equal(_interior3(got_p, x, y, z, hd), _interior3(p1, x, y, z, hd))
    np.testing.assert_array_equal(_interior3(w.download(), x, y, z, hd),
                                  _interior3(w_ref, x, y, z, hd))
    assert abs(pw[0] - pw_ref[0]) <= 1e-13 * abs(pw_ref[0])
    np.testing.assert_array_equal(p_out2.download(), got_p)
    assert pw_nw[0] == pw[0]
    np.testing.assert_array_equal(_interior3(w0.download(), x, y, z, hd),
                                  _interior3(w0_ref, x, y, z, hd))
    assert abs(pw0[0] - pw0_ref[0]) <= 1e-13 * abs(pw0_ref[0])
    for v in list(d.values()) + [cond, p_out, p_out2, w, w0]:
        v.free()


@pytest.mark.parametrize("grid", GRIDS3, ids=[f"{g[0]}x{g[1]}x{g[2]}h{g[3]}c{g[4]}t{g[5]}"
                                              for g in GRIDS3])
def test_calc_ur_cond3_matches_reference(hip, oracle, grid):
    from tealeaf_amd.device import DeviceArray
    nx, ny, nz, hd, coef, tz = grid
    O = oracle
    x, y, z, f, rng = _inputs3(O, nx, ny, nz, hd, coef)
    n = x * y * z
    alpha = float(rng.uniform(0.1, 0.9))
    p = f["p"].copy()
    _halo_p3(O, x, y, z, hd, p)
    w_ref, pw_ref = np.zeros(n), np.zeros(1)
    O.or3d_cg_solver_calc_w(x, y, z, hd, ptr(pw_ref), ptr(p), ptr(w_ref), ptr(f["kx"]),
                            ptr(f["ky"]), ptr(f["kz"]))
    u_ref, r_ref, rrn_ref = f["u"].copy(), f["r"].copy(), np.zeros(1)
    O.or3d_cg_solver_calc_ur(x, y, z, hd, alpha, ptr(rrn_ref), ptr(u_ref), ptr(p), ptr(r_ref),
                             ptr(w_ref))
    d = {k: DeviceArray.from_host(f[k]) for k in ("dens", "p", "r", "u")}
    cond = DeviceArray.from_host(np.full(n, np.nan))
    hip.tlhip3d_cg_cond(x, y, z, hd, coef, d["dens"].ptr, cond.ptr)
    rrn = np.zeros(1)
    hip.tlhip_set_tile_planes(tz if tz else -1)
    try:
        hip.tlhip3d_cg_calc_ur_cond(x, y, z, hd, alpha, RX, RY, RZ, ptr(rrn), d["u"].ptr,
                                    d["p"].ptr, d["r"].ptr, cond.ptr)
    finally:
        hip.tlhip_set_tile_planes(-1)
    np.testing.assert_array_equal(d["u"].download(), u_ref)     # halo untouched
    np.testing.assert_array_equal(d["r"].download(), r_ref)
    assert abs(rrn[0] - rrn_ref[0]) <= 1e-13 * abs(rrn_ref[0])
    for v in list(d.values()) + [cond]:
        v.free()


def _run3(hip, text, cond, resident):
    from tealeaf_amd.problem import Problem
    hip.tlhip_set_cg_cond(int(cond))
    hip.tlhip_set_resident(int(resident))
    try:
        with Problem.from_deck(text, 3) as prob:
            steps = prob.run()
            return steps, prob.get("u"), prob.get("energy"), prob.field_summary()
    finally:
        hip.tlhip_set_cg_cond(1)
        hip.tlhip_set_resident(1)


@pytest.mark.parametrize("solver", ["CG", "CHEBY", "PPCG"])
def test_driver_cond3_pair_agrees(hip, solver):
    """3D: the driver with the cond pair (the default) against the kx/ky/kz
    passes, as test_driver_cond_pair_agrees."""
    from tealeaf_amd import deck as tdeck
    text = tdeck.cuboid_3d(30, 3, solver=getattr(tdeck, solver), presteps=10, inner=10,
                           max_iters=2000)
    on = _run3(hip, text, True, True)
    host = _run3(hip, text, True, False)
    off = _run3(hip, text, False, True)
    np.testing.assert_array_equal(host[1], on[1])
    assert host[3] == on[3]
    for key in ("n_cg_steps", "n_cheby_steps", "n_ppcg_outer"):
        a = [s.get(key, 0) for s in on[0]]
        b = [s.get(key, 0) for s in off[0]]
        assert all(abs(i - j) <= 1 for i, j in zip(a, b)), (key, a, b)
    assert sum(s["n_cg_steps"] for s in on[0]) > 0
    for got, want in ((on[1], off[1]), (on[2], off[2])):
        assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))
    for a, b in zip(on[3], off[3]):
        assert abs(a - b) <= 1e-12 * abs(b), (on[3], off[3])
