"""Two 3D PPCG inner steps in one pass against two separate reference steps.

tlhip3d_ppcg_inner_iterations runs two turns of
3d/drivers/ppcg_solver_driver.c:170-183 (a depth-1 reflective halo update of
sd, then ppcg_inner_iteration, 3d/c_kernels/omp3/ppcg_solver.c:32-73) in one
pass.  Its r, u and sd must equal, bit for bit, those of the oracle doing
the two turns; the halo rings of r, u and sd pass through unchanged.  The
grids cover several windows along x, partial row blocks, short and partial
z-tiles, tiles of one plane, and halo depths 1, 2 and 3.
"""
import numpy as np
import pytest

from oracle.pyoracle import ptr

pytestmark = pytest.mark.gpu

# (nx, ny, nz, hd, planes per z-tile; 0 = default); x = nx + 2 hd even
GRIDS = [(24, 20, 18, 2, 0), (250, 9, 7, 2, 0), (130, 13, 33, 1, 5), (6, 5, 4, 2, 1),
         (368, 6, 21, 2, 16), (60, 17, 12, 3, 3), (128, 128, 40, 2, 0)]


def _oracle_pair(O, x, y, z, hd, alphas, betas, sd, r, u, kx, ky, kz):
    sd, r, u = sd.copy(), r.copy(), u.copy()
    nb = np.full(6, -1, dtype=np.int32)
    flags = np.zeros(6, dtype=np.bool_)
    flags[5] = True                                   # sd only
    dummy = np.zeros(x * y * z)
    for a, b in zip(alphas, betas):
        O.or3d_local_halos(x, y, z, 1, hd, ptr(nb), ptr(flags), ptr(dummy), ptr(dummy),
                           ptr(dummy), ptr(dummy), ptr(dummy), ptr(sd))
        O.or3d_ppcg_inner_iteration(x, y, z, hd, a, b, ptr(u), ptr(r), ptr(kx), ptr(ky), ptr(kz),
                                    ptr(sd))
    return sd, r, u


def _interior(a, x, y, z, hd):
    return a.reshape(z, y, x)[hd:z - hd, hd:y - hd, hd:x - hd]


def _ring(a, x, y, z, hd):
    m = np.ones((z, y, x), dtype=bool)
    m[hd:z - hd, hd:y - hd, hd:x - hd] = False
    return a.reshape(z, y, x)[m]


@pytest.mark.parametrize("grid", GRIDS, ids=[f"{g[0]}x{g[1]}x{g[2]}h{g[3]}t{g[4]}" for g in GRIDS])
@pytest.mark.parametrize("rows", [8, 4, 14, 0])
def test_pair_matches_two_steps(hip, oracle, grid, rows):
    from tealeaf_amd.device import DeviceArray
    nx, ny, nz, hd, tz = grid
    x, y, z = nx + 2 * hd, ny + 2 * hd, nz + 2 * hd
    n = x * y * z
    rng = np.random.default_rng(500 + nx + ny + nz + hd)
    sd, r, u = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    kx, ky, kz = (rng.uniform(0.02, 0.2, n) for _ in range(3))
    alphas, betas = rng.uniform(0.1, 0.9, 2), rng.uniform(0.1, 0.9, 2)
    want_sd, want_r, want_u = _oracle_pair(oracle, x, y, z, hd, alphas, betas, sd, r, u, kx, ky,
                                           kz)
    d = {k: DeviceArray.from_host(v) for k, v in dict(sd=sd, r=r, u=u, kx=kx, ky=ky, kz=kz).items()}
    out = {k: DeviceArray.from_host(np.full(n, np.nan)) for k in ("sd", "r", "u")}
    nb = np.full(6, -1, dtype=np.int32)
    hip.tlhip_set_tile_planes(tz if tz else -1)
    hip.tlhip_set_pair_rows(rows)
    try:
        hip.tlhip3d_ppcg_inner_iterations(x, y, z, hd, 2, alphas.ctypes.data, betas.ctypes.data,
                                          nb.ctypes.data, d["sd"].ptr, out["sd"].ptr, d["r"].ptr,
                                          out["r"].ptr, d["u"].ptr, out["u"].ptr, d["kx"].ptr,
                                          d["ky"].ptr, d["kz"].ptr)
    finally:
        hip.tlhip_set_tile_planes(-1)
        hip.tlhip_set_pair_rows(8)
    hip.tlhip_synchronize()
    got = {k: v.download() for k, v in out.items()}
    for name, want, src in (("sd", want_sd, sd), ("r", want_r, r), ("u", want_u, u)):
        np.testing.assert_array_equal(_interior(got[name], x, y, z, hd),
                                      _interior(want, x, y, z, hd), err_msg=name)
        np.testing.assert_array_equal(_ring(got[name], x, y, z, hd), _ring(src, x, y, z, hd),
                                      err_msg=name + " ring")
    for v in list(d.values()) + list(out.values()):
        v.free()


def test_pair_rejects_bad_arguments(hip):
    from tealeaf_amd.device import DeviceArray
    raw = hip.dll
    nb = np.full(6, -1, dtype=np.int32)
    a = np.ones(2)
    x, y, z = 12, 10, 8
    bufs = [DeviceArray.from_host(np.zeros(x * y * z)) for _ in range(9)]
    q = [b.ptr for b in bufs]
    assert raw.tlhip3d_ppcg_inner_iterations(x, y, z, 2, 3, a.ctypes.data, a.ctypes.data,
                                             nb.ctypes.data, *q) != 0          # count 3
    assert raw.tlhip3d_ppcg_inner_iterations(11, y, z, 2, 2, a.ctypes.data, a.ctypes.data,
                                             nb.ctypes.data, *q) != 0          # odd x
    assert raw.tlhip3d_ppcg_inner_iterations(x, y, z, 2, 2, a.ctypes.data, a.ctypes.data,
                                             nb.ctypes.data, q[0], q[0], *q[2:]) != 0   # sd in place
    nb[4] = 2
    assert raw.tlhip3d_ppcg_inner_iterations(x, y, z, 2, 2, a.ctypes.data, a.ctypes.data,
                                             nb.ctypes.data, *q) != 0          # internal face
    assert "external" in hip.last_error()
    for b in bufs:
        b.free()
