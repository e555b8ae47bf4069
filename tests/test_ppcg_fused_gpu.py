"""Temporally blocked PPCG inner loop (2D) against S separate reference steps.

tlhip2d_ppcg_inner_iterations runs `count` inner steps of
2d/drivers/ppcg_driver.c:165-183 in one pass.  Its r, u and sd must equal,
bit for bit, those of the oracle doing what the reference does per step:
a depth-1 reflective halo update of sd (2d/c_kernels/omp3/local_halos.c)
then ppcg_inner_iteration (2d/c_kernels/omp3/ppcg.c:28-61), optionally
preceded by ppcg_init (ppcg.c:8-25).  The halo ring of r and u passes
through unchanged.  The cases cover partial x windows and y tiles, tiles
shorter than the step count, grids narrower than one window, odd halo
depths (lanes that straddle the interior edge) and both loop schedules.

The FP32 variant (the mixed-precision solve's inner sweeps) is held to the
same per-step arithmetic done in numpy float32 with the kernel's fused
multiply-adds emulated exactly (tests/fp32_emulation.py) -- so it too must
match bit for bit; its fused 2-norm variant to the FP64 sum of squares.
"""
import ctypes

import numpy as np
import pytest

from oracle.pyoracle import ptr
from tests.fp32_emulation import fma32
from tealeaf_amd import deck as tdeck

pytestmark = pytest.mark.gpu

# (nx, ny, hd, tile_rows): nx even; windows are 128 - 2*SP wide (SP = S
# rounded up to even), tiles `tile_rows` high (0 = default).
GRIDS = [(32, 32, 2, 0), (250, 70, 2, 0), (300, 129, 2, 8), (6, 5, 2, 3),
         (124, 3, 2, 0), (118, 41, 4, 2), (460, 17, 2, 1), (64, 20, 1, 0), (122, 33, 3, 4)]


def _oracle_steps(O, x, y, hd, alphas, betas, theta, from_r, sd, r, u, kx, ky):
    sd, r, u = sd.copy(), r.copy(), u.copy()
    if from_r:
        O.or2d_ppcg_init(x, y, hd, theta, ptr(r), ptr(sd))
    nb = np.full(4, -1, dtype=np.int32)
    flags = np.zeros(6, dtype=np.bool_)
    flags[5] = True                                   # sd only
    dummy = np.zeros(x * y)
    for a, b in zip(alphas, betas):
        O.or2d_local_halos(x, y, 1, hd, ptr(nb), ptr(flags), ptr(dummy), ptr(dummy), ptr(dummy),
                           ptr(dummy), ptr(dummy), ptr(sd))
        O.or2d_ppcg_inner_iteration(x, y, hd, a, b, ptr(u), ptr(r), ptr(kx), ptr(ky), ptr(sd))
    return sd, r, u


def _interior(a, x, y, hd):
    return a.reshape(y, x)[hd:y - hd, hd:x - hd]


def _ring(a, x, y, hd):
    m = np.ones((y, x), dtype=bool)
    m[hd:y - hd, hd:x - hd] = False
    return a.reshape(y, x)[m]


@pytest.mark.parametrize("grid", GRIDS, ids=[f"{g[0]}x{g[1]}h{g[2]}t{g[3]}" for g in GRIDS])
@pytest.mark.parametrize("count", [1, 2, 3, 4, 5, 7, 10])
@pytest.mark.parametrize("from_r", [True, False], ids=["init", "cont"])
@pytest.mark.parametrize("pf", [1, 0], ids=["pf", "nopf"])
def test_fused_matches_separate_steps(hip, oracle, grid, count, from_r, pf):
    from tealeaf_amd.device import DeviceArray
    nx, ny, hd, ty = grid
    x, y = nx + 2 * hd, ny + 2 * hd
    rng = np.random.default_rng(1000 * nx + 10 * count + from_r)
    n = x * y
    sd = rng.uniform(-1, 1, n)
    r = rng.uniform(-1, 1, n)
    u = rng.uniform(-1, 1, n)
    kx = rng.uniform(0.05, 0.3, n)
    ky = rng.uniform(0.05, 0.3, n)
    alphas = rng.uniform(0.1, 0.9, count)
    betas = rng.uniform(0.1, 0.9, count)
    theta = 1.7
    want_sd, want_r, want_u = _oracle_steps(oracle, x, y, hd, alphas, betas, theta, from_r, sd,
                                            r, u, kx, ky)

    d = {k: DeviceArray.from_host(v) for k, v in
         dict(sd=sd, r=r, u=u, kx=kx, ky=ky).items()}
    out = {k: DeviceArray.from_host(np.full(n, np.nan)) for k in ("sd", "r", "u")}
    nb = np.full(4, -1, dtype=np.int32)
    hip.tlhip_set_fused_tuning(ty if ty else -1, pf)
    try:
        hip.tlhip2d_ppcg_inner_iterations(
            x, y, hd, count, alphas.ctypes.data, betas.ctypes.data, theta, int(from_r),
            nb.ctypes.data, None if from_r else d["sd"].ptr, out["sd"].ptr, d["r"].ptr,
            out["r"].ptr, d["u"].ptr, out["u"].ptr, d["kx"].ptr, d["ky"].ptr, 0.0, None, None)
    finally:
        hip.tlhip_set_fused_tuning(-1, -1)
    hip.tlhip_synchronize()
    got = {k: v.download() for k, v in out.items()}
    for name, want, src in (("r", want_r, r), ("u", want_u, u), ("sd", want_sd, sd)):
        g = got[name]
        np.testing.assert_array_equal(_interior(g, x, y, hd), _interior(want, x, y, hd),
                                      err_msg=f"{name} interior")
        if name != "sd" or not from_r:
            np.testing.assert_array_equal(_ring(g, x, y, hd), _ring(src, x, y, hd),
                                          err_msg=f"{name} halo ring")
    for v in list(d.values()) + list(out.values()):
        v.free()


def test_fused_rejects_bad_layout(hip):
    from tealeaf_amd.device import DeviceArray
    raw = hip.dll
    nb = np.full(4, -1, dtype=np.int32)
    a = np.ones(2)
    x, y = 15, 14                        # odd row length
    bufs = [DeviceArray.from_host(np.zeros(x * y)) for _ in range(8)]
    p = [b.ptr for b in bufs]
    rc = raw.tlhip2d_ppcg_inner_iterations(x, y, 2, 2, a.ctypes.data, a.ctypes.data, 1.0, 1,
                                           nb.ctypes.data, None, *p[:7], 0.0, None, None)
    assert rc != 0 and "pitch" in hip.last_error()
    x = 16                               # in-place output
    rc = raw.tlhip2d_ppcg_inner_iterations(x, y, 2, 2, a.ctypes.data, a.ctypes.data, 1.0, 0,
                                           nb.ctypes.data, p[0], p[0], p[1], p[2], p[3], p[4],
                                           p[5], p[6], 0.0, None, None)
    assert rc != 0
    rc = raw.tlhip2d_ppcg_inner_iterations(x, y, 2, 11, a.ctypes.data, a.ctypes.data, 1.0, 0,
                                           nb.ctypes.data, p[0], p[1], p[2], p[3], p[4], p[5],
                                           p[6], p[7], 0.0, None, None)
    assert rc != 0                       # count above 10
    nb[1] = 3                            # an internal face
    rc = raw.tlhip2d_ppcg_inner_iterations(x, y, 2, 2, a.ctypes.data, a.ctypes.data, 1.0, 0,
                                           nb.ctypes.data, p[0], p[1], p[2], p[3], p[4], p[5],
                                           p[6], p[7], 0.0, None, None)
    assert rc != 0
    nb[1] = -1                           # calc_ur folded in without init_from_r
    rc = raw.tlhip2d_ppcg_inner_iterations(x, y, 2, 2, a.ctypes.data, a.ctypes.data, 1.0, 0,
                                           nb.ctypes.data, p[0], p[1], p[2], p[3], p[4], p[5],
                                           p[6], p[7], 0.5, p[0], p[1])
    assert rc != 0 and "init_from_r" in hip.last_error()
    for b in bufs:
        b.free()


def _run_fusions(hip, text, fusions, fused_norm, dims=2):
    from tealeaf_amd.problem import Problem
    res = {}
    hip.tlhip_set_ppcg_fused_norm(int(fused_norm))
    try:
        for fusion in fusions:
            hip.tlhip_set_ppcg_fusion(fusion)
            with Problem.from_deck(text, dims) as prob:
                steps = prob.run()
                res[fusion] = (steps, prob.get("u"), prob.get("energy"), prob.field_summary())
    finally:
        hip.tlhip_set_ppcg_fusion(10)
        hip.tlhip_set_ppcg_fused_norm(1)
    return res


@pytest.mark.parametrize("inner", [1, 7, 10, 12])
def test_driver_fusion_is_bit_identical(hip, inner):
    """The native driver with the fused inner loop gives the same bits as
    one launch per inner step, for inner-step counts that split unevenly
    (the 2-norm of r in its own pass, as the unfused path does); with the
    2-norm formed in the last fused launch (the default), rrn is summed in
    another order: the same iteration counts, fields within 1e-13 relative,
    and the summary within 1e-13."""
    text = tdeck.layered_2d(90, 2, solver=tdeck.PPCG, presteps=15, inner=inner, max_iters=2000)
    res = _run_fusions(hip, text, (1, 10, 5, 3), False)
    base = res[1]
    for f in (3, 5, 10):
        steps, u, e, summ = res[f]
        assert [s["n_ppcg_inner"] for s in steps] == [s["n_ppcg_inner"] for s in base[0]]
        assert [s["n_ppcg_outer"] for s in steps] == [s["n_ppcg_outer"] for s in base[0]]
        np.testing.assert_array_equal(u, base[1])
        np.testing.assert_array_equal(e, base[2])
        assert summ == base[3]
    normed = _run_fusions(hip, text, (10, 3), True)
    for f in (10, 3):
        steps, u, e, summ = normed[f]
        assert [s["n_ppcg_outer"] for s in steps] == [s["n_ppcg_outer"] for s in base[0]]
        assert np.max(np.abs(u - base[1])) <= 1e-13 * np.max(np.abs(base[1]))
        assert np.max(np.abs(e - base[2])) <= 1e-13 * np.max(np.abs(base[2]))
        for a, b in zip(summ, base[3]):
            assert abs(a - b) <= 1e-13 * abs(b), (summ, base[3])


@pytest.mark.parametrize("resident", [1, 0])
@pytest.mark.parametrize("n,inner", [(24, 1), (24, 7), (30, 10), (23, 10)])
def test_driver_3d_pairs_bit_identical(hip, resident, n, inner):
    """3D PPCG with two inner steps per launch (fusion > 1, even x) gives
    the same bits as one launch per step after a halo update (fusion 1): odd
    inner counts end on a single step, an odd x (23 + 2*2 cells) keeps the
    single steps; on the resident and the host-driven driver."""
    text = tdeck.cuboid_3d(n, 2, solver=tdeck.PPCG, presteps=10, inner=inner, max_iters=400)
    hip.tlhip_set_resident(resident)
    try:
        res = _run_fusions(hip, text, (1, 10), False, dims=3)
    finally:
        hip.tlhip_set_resident(1)
    base = res[1]
    steps, u, e, summ = res[10]
    assert [s["n_ppcg_inner"] for s in steps] == [s["n_ppcg_inner"] for s in base[0]]
    assert [s["n_ppcg_outer"] for s in steps] == [s["n_ppcg_outer"] for s in base[0]]
    assert sum(s["n_ppcg_outer"] for s in steps) > 0
    np.testing.assert_array_equal(u, base[1])
    np.testing.assert_array_equal(e, base[2])
    assert summ == base[3]


def test_set_ppcg_fusion_range(hip):
    assert hip.dll.tlhip_set_ppcg_fusion(11) != 0
    assert hip.dll.tlhip_set_ppcg_fusion(-1) != 0
    assert hip.dll.tlhip_set_ppcg_fusion(10) == 0


def _f32_steps(x, y, hd, alphas, betas, theta, from_r, sd, r, u, kx, ky, ur=None):
    """The reference's per-step expressions on (rows, pitch) float32 arrays,
    with the FP32 kernel's fused multiply-adds: tx = kxr*rgt + kxc*lft and
    ty = kyu*up + kyc*dn each as one fma after a product, A sd = fma(d, cen,
    -tx) - ty, sd' = fma(alpha, cen, beta*r'), and sd = r * (1/theta).
    ur = (alpha, p, w): cg_calc_ur first, r = fma(-alpha, w, r) and
    u = fma(alpha, p, u)."""
    f = np.float32
    sd, r, u = sd.copy(), r.copy(), u.copy()
    J, K = slice(hd, y - hd), slice(hd, x - hd)
    if ur is not None:
        al, p, w = ur
        r[J, K] = fma32(-f(al), w[J, K], r[J, K])
        u[J, K] = fma32(f(al), p[J, K], u[J, K])
    if from_r:
        sd[J, K] = r[J, K] * (f(1) / f(theta))
    for a, b in zip(alphas, betas):
        a, b = f(a), f(b)
        s = sd.copy()                                     # depth-1 reflective halo
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
        tx = fma32(kxr, rgt, kxc * lft)
        ty = fma32(kyu, up, kyc * dn)
        av = fma32(d, cen, -tx) - ty
        r[J, K] = r[J, K] - av
        u[J, K] = u[J, K] + cen
        sd[J, K] = fma32(a, cen, b * r[J, K])
    return sd, r, u


F32_GRIDS = [(32, 30, 2, 0), (250, 70, 2, 0), (510, 37, 2, 4), (6, 5, 2, 0), (61, 9, 1, 0),
             (1000, 21, 3, 8), (248, 17, 2, 2)]


@pytest.mark.parametrize("grid", F32_GRIDS, ids=[f"{g[0]}x{g[1]}h{g[2]}t{g[3]}" for g in F32_GRIDS])
@pytest.mark.parametrize("count", [1, 3, 5, 10])
@pytest.mark.parametrize("mode", [1, 0, 2], ids=["init", "cont", "init_ur"])
@pytest.mark.parametrize("pad", [0, 4])
def test_fused_f32_matches_float32_steps(hip, grid, count, mode, pad):
    from tealeaf_amd.device import DeviceArray
    nx, ny, hd, ty = grid
    x, y = nx + 2 * hd, ny + 2 * hd
    pitch = (x + 3) // 4 * 4 + pad
    rng = np.random.default_rng(7000 + nx + 10 * count + pad)
    shape = (y, pitch)
    f32 = lambda lo, hi: rng.uniform(lo, hi, shape).astype(np.float32)  # noqa: E731
    sd, r, u = f32(-1, 1), f32(-1, 1), f32(-1, 1)
    kx, ky = f32(0.05, 0.3), f32(0.05, 0.3)
    alphas = rng.uniform(0.1, 0.9, count)
    betas = rng.uniform(0.1, 0.9, count)
    theta = 1.7
    from_r = mode >= 1
    p, w = f32(-1, 1), f32(-1, 1)
    al_ur = 0.37 if mode == 2 else 0.0
    want = dict(zip(("sd", "r", "u"), _f32_steps(x, y, hd, alphas, betas, theta, from_r, sd, r,
                                                  u, kx, ky,
                                                  ur=(al_ur, p, w) if mode == 2 else None)))
    src = dict(sd=sd, r=r, u=u)
    d = {k: DeviceArray.from_host(v) for k, v in dict(sd=sd, r=r, u=u, kx=kx, ky=ky, p=p,
                                                      w=w).items()}
    out = {k: DeviceArray.from_host(np.full(shape, np.nan, dtype=np.float32)) for k in ("sd", "r", "u")}
    nb = np.full(4, -1, dtype=np.int32)
    hip.tlhip_set_fused_tuning(ty if ty else -1, -1)
    try:
        hip.tlhip2d_ppcg_inner_iterations_f32(
            x, y, hd, pitch, count, alphas.ctypes.data, betas.ctypes.data, theta, int(from_r),
            nb.ctypes.data, None if from_r else d["sd"].ptr, out["sd"].ptr, d["r"].ptr,
            out["r"].ptr, d["u"].ptr, out["u"].ptr, d["kx"].ptr, d["ky"].ptr, al_ur,
            d["p"].ptr if mode == 2 else None, d["w"].ptr if mode == 2 else None)
    finally:
        hip.tlhip_set_fused_tuning(-1, -1)
    hip.tlhip_synchronize()
    J, K = slice(hd, y - hd), slice(hd, x - hd)
    ring = np.ones(shape, dtype=bool)
    ring[J, K] = False
    got = {}
    for name in ("r", "u", "sd"):
        g = got[name] = out[name].download().reshape(shape)
        np.testing.assert_array_equal(g[J, K], want[name][J, K], err_msg=f"{name} interior")
        if name != "sd" or not from_r:
            np.testing.assert_array_equal(g[ring], src[name][ring], err_msg=f"{name} halo ring")
    # the same launch forming r.r: identical fields, and the FP64 sum of the
    # interior's squares
    out2 = {k: DeviceArray.from_host(np.full(shape, np.nan, dtype=np.float32)) for k in ("sd", "r", "u")}
    rrn = ctypes.c_double(-1.0)
    hip.tlhip2d_ppcg_inner_iterations_f32_norm(
        x, y, hd, pitch, count, alphas.ctypes.data, betas.ctypes.data, theta, int(from_r),
        nb.ctypes.data, None if from_r else d["sd"].ptr, out2["sd"].ptr, d["r"].ptr,
        out2["r"].ptr, d["u"].ptr, out2["u"].ptr, d["kx"].ptr, d["ky"].ptr, al_ur,
        d["p"].ptr if mode == 2 else None, d["w"].ptr if mode == 2 else None,
        ctypes.addressof(rrn))
    for name in ("r", "u", "sd"):
        np.testing.assert_array_equal(out2[name].download().reshape(shape)[J, K], got[name][J, K])
    want_rr = float(np.sum(got["r"][J, K].astype(np.float64) ** 2))
    assert abs(rrn.value - want_rr) <= 1e-13 * want_rr, (rrn.value, want_rr)
    for v in list(d.values()) + list(out.values()) + list(out2.values()):
        v.free()


@pytest.mark.parametrize("grid", [(250, 70, 2, 0), (118, 41, 4, 2), (64, 20, 1, 0)],
                         ids=["250x70", "118x41h4", "64x20h1"])
@pytest.mark.parametrize("count", [1, 5, 10])
def test_fused_with_calc_ur_matches_oracle(hip, oracle, grid, count):
    """calc_ur folded into the first launch: the oracle runs cg_calc_ur
    (cg.c:116-143), ppcg_init and `count` halo-update + inner steps."""
    from tealeaf_amd.device import DeviceArray
    nx, ny, hd, ty = grid
    x, y = nx + 2 * hd, ny + 2 * hd
    n = x * y
    rng = np.random.default_rng(500 + nx + count)
    r, u, p, w = (rng.uniform(-1, 1, n) for _ in range(4))
    kx, ky = rng.uniform(0.05, 0.3, n), rng.uniform(0.05, 0.3, n)
    alphas, betas = rng.uniform(0.1, 0.9, count), rng.uniform(0.1, 0.9, count)
    theta, alpha = 1.3, 0.37
    r_ur, u_ur = r.copy(), u.copy()
    rrn = np.zeros(1)
    oracle.or2d_cg_calc_ur(x, y, hd, alpha, ptr(rrn), ptr(u_ur), ptr(p), ptr(r_ur), ptr(w))
    want_sd, want_r, want_u = _oracle_steps(oracle, x, y, hd, alphas, betas, theta, True,
                                            np.zeros(n), r_ur, u_ur, kx, ky)
    d = {k: DeviceArray.from_host(v) for k, v in dict(r=r, u=u, p=p, w=w, kx=kx, ky=ky).items()}
    out = {k: DeviceArray.from_host(np.full(n, np.nan)) for k in ("sd", "r", "u")}
    nb = np.full(4, -1, dtype=np.int32)
    hip.tlhip_set_fused_tuning(ty if ty else -1, -1)
    try:
        hip.tlhip2d_ppcg_inner_iterations(
            x, y, hd, count, alphas.ctypes.data, betas.ctypes.data, theta, 1, nb.ctypes.data,
            None, out["sd"].ptr, d["r"].ptr, out["r"].ptr, d["u"].ptr, out["u"].ptr, d["kx"].ptr,
            d["ky"].ptr, alpha, d["p"].ptr, d["w"].ptr)
    finally:
        hip.tlhip_set_fused_tuning(-1, -1)
    hip.tlhip_synchronize()
    for name, want, src in (("r", want_r, r), ("u", want_u, u), ("sd", want_sd, None)):
        g = out[name].download()
        np.testing.assert_array_equal(_interior(g, x, y, hd), _interior(want, x, y, hd),
                                      err_msg=f"{name} interior")
        if src is not None:
            np.testing.assert_array_equal(_ring(g, x, y, hd), _ring(src, x, y, hd))
    for v in list(d.values()) + list(out.values()):
        v.free()


@pytest.mark.parametrize("rows", [-1, 1, 5])
@pytest.mark.parametrize("grid", [(250, 70, 2), (74, 29, 1), (300, 21, 3), (6, 5, 2)],
                         ids=["250x70", "74x29h1", "300x21h3", "6x5"])
def test_calc_pw_matches_calc_p_halo_calc_w(hip, oracle, grid, rows):
    """tlhip2d_cg_calc_pw against the oracle's cg_calc_p, depth-1 halo update
    of p and cg_calc_w: p_out and w bit-exact, pw to reduction order; with
    the default rows per wave and with 1 and 5 (tiles that end inside the
    grid, so the y window restarts at every tile)."""
    import ctypes
    from tealeaf_amd.device import DeviceArray
    nx, ny, hd = grid
    x, y = nx + 2 * hd, ny + 2 * hd
    n = x * y
    rng = np.random.default_rng(900 + nx)
    p, r = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    kx, ky = rng.uniform(0.05, 0.3, n), rng.uniform(0.05, 0.3, n)
    beta = 0.61
    pw_ref = np.zeros(1)
    p_ref, w_ref = p.copy(), np.zeros(n)
    oracle.or2d_cg_calc_p(x, y, hd, beta, ptr(p_ref), ptr(r))
    nbr = np.full(4, -1, dtype=np.int32)
    flags = np.zeros(6, dtype=np.bool_)
    flags[4] = True                                   # p
    dummy = np.zeros(n)
    oracle.or2d_local_halos(x, y, 1, hd, ptr(nbr), ptr(flags), ptr(dummy), ptr(dummy), ptr(dummy),
                            ptr(dummy), ptr(p_ref), ptr(dummy))
    oracle.or2d_cg_calc_w(x, y, hd, ptr(pw_ref), ptr(p_ref), ptr(w_ref), ptr(kx), ptr(ky))
    d = {k: DeviceArray.from_host(v) for k, v in dict(p=p, r=r, kx=kx, ky=ky).items()}
    p_out, w = DeviceArray.from_host(np.full(n, np.nan)), DeviceArray.from_host(np.full(n, np.nan))
    pw = ctypes.c_double(0.0)
    hip.tlhip_set_tile_rows(rows)
    try:
        hip.tlhip2d_cg_calc_pw(x, y, hd, beta, ctypes.addressof(pw), d["p"].ptr, p_out.ptr,
                               d["r"].ptr, w.ptr, d["kx"].ptr, d["ky"].ptr)
    finally:
        hip.tlhip_set_tile_rows(-1)
    gp, gw = p_out.download(), w.download()
    np.testing.assert_array_equal(_interior(gp, x, y, hd), _interior(p_ref, x, y, hd))
    np.testing.assert_array_equal(_interior(gw, x, y, hd), _interior(w_ref, x, y, hd))
    np.testing.assert_array_equal(_ring(gp, x, y, hd), _ring(p, x, y, hd))
    terms = np.sum(np.abs(_interior(w_ref, x, y, hd) * _interior(p_ref, x, y, hd)))
    assert abs(pw.value - pw_ref[0]) <= 1e-13 * max(terms, 1e-300)
    for v in list(d.values()) + [p_out, w]:
        v.free()


@pytest.mark.parametrize("planes", [-1, 3])
@pytest.mark.parametrize("grid", [(20, 13, 9, 2), (17, 10, 12, 1), (9, 7, 5, 3), (130, 6, 5, 2)],
                         ids=["20x13x9", "17x10x12h1", "9x7x5h3", "130x6x5"])
def test_calc_pw_3d_matches_calc_p_halo_calc_w(hip, oracle, grid, planes):
    """tlhip3d_cg_solver_calc_pw against the oracle's cg_solver_calc_p, a
    depth-1 halo update of p and cg_solver_calc_w: p_out and w bit-exact
    (the mirror on all six faces, odd widths with one cell per lane, tiles
    of 3 planes), the halo ring of p copied, pw to reduction order."""
    import ctypes
    from tealeaf_amd.device import DeviceArray
    nx, ny, nz, hd = grid
    x, y, z = nx + 2 * hd, ny + 2 * hd, nz + 2 * hd
    n = x * y * z
    rng = np.random.default_rng(1900 + nx + ny)
    p, r = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    kx, ky, kz = (rng.uniform(0.05, 0.3, n) for _ in range(3))
    beta = 0.43
    pw_ref = np.zeros(1)
    p_ref, w_ref = p.copy(), np.zeros(n)
    oracle.or3d_cg_solver_calc_p(x, y, z, hd, beta, ptr(p_ref), ptr(r))
    nbr = np.full(6, -1, dtype=np.int32)
    flags = np.zeros(6, dtype=np.bool_)
    flags[4] = True                                   # p
    dummy = np.zeros(n)
    oracle.or3d_local_halos(x, y, z, 1, hd, ptr(nbr), ptr(flags), ptr(dummy), ptr(dummy),
                            ptr(dummy), ptr(dummy), ptr(p_ref), ptr(dummy))
    oracle.or3d_cg_solver_calc_w(x, y, z, hd, ptr(pw_ref), ptr(p_ref), ptr(w_ref), ptr(kx),
                                 ptr(ky), ptr(kz))
    d = {k: DeviceArray.from_host(v) for k, v in dict(p=p, r=r, kx=kx, ky=ky, kz=kz).items()}
    p_out, w = DeviceArray.from_host(np.full(n, np.nan)), DeviceArray.from_host(np.full(n, np.nan))
    pw = ctypes.c_double(0.0)
    hip.tlhip_set_tile_planes(planes)
    try:
        hip.tlhip3d_cg_solver_calc_pw(x, y, z, hd, beta, ctypes.addressof(pw), d["p"].ptr,
                                      p_out.ptr, d["r"].ptr, w.ptr, d["kx"].ptr, d["ky"].ptr,
                                      d["kz"].ptr)
    finally:
        hip.tlhip_set_tile_planes(-1)
    inner = (slice(hd, z - hd), slice(hd, y - hd), slice(hd, x - hd))
    gp, gw = p_out.download().reshape(z, y, x), w.download().reshape(z, y, x)
    np.testing.assert_array_equal(gp[inner], p_ref.reshape(z, y, x)[inner])
    np.testing.assert_array_equal(gw[inner], w_ref.reshape(z, y, x)[inner])
    ring = np.ones((z, y, x), dtype=bool)
    ring[inner] = False
    np.testing.assert_array_equal(gp[ring], p.reshape(z, y, x)[ring])
    terms = np.sum(np.abs(w_ref.reshape(z, y, x)[inner] * p_ref.reshape(z, y, x)[inner]))
    assert abs(pw.value - pw_ref[0]) <= 1e-13 * max(terms, 1e-300)
    for v in list(d.values()) + [p_out, w]:
        v.free()
