"""Per-kernel parity of the HIP suite against the CPU oracle (GPU).

The library is built with -ffp-contract=off, so every field a kernel writes
must equal the oracle's BIT FOR BIT (same IEEE operations, same order).  Only
reduction scalars differ, by summation order: they are held to a relative
1e-13 of the larger of the result and the sum of absolute terms.
"""
import ctypes

import numpy as np
import pytest

from tests import kernel_cases as kc
from oracle.pyoracle import load_oracle

pytestmark = pytest.mark.gpu

SCALAR_RTOL = 1e-13   # reduction order only (FP64)

CASES = (kc.make_cases(2, 10, 7, seed=21) + kc.make_cases(2, 37, 29, seed=22, coefficient=2)
         + kc.make_cases(2, 300, 21, seed=23) + kc.make_cases(2, 129, 33, hd=1, seed=24)
         + kc.make_cases(2, 64, 17, hd=3, seed=25)
         + kc.make_cases(3, 6, 5, 4, seed=31) + kc.make_cases(3, 19, 11, 7, seed=32, coefficient=2)
         + kc.make_cases(3, 130, 9, 5, seed=33) + kc.make_cases(3, 12, 10, 9, hd=1, seed=34)
         + [kc.state_case(2, 12, 10), kc.state_case(3, 8, 7, 6), kc.data_case(2, 9, 7),
            kc.data_case(3, 7, 6, 5)])


def _interior(case, a):
    h = case.hd
    if case.dims == 3:
        return a.reshape(case.z, case.y, case.x)[h:-h, h:-h, h:-h]
    return a.reshape(case.y, case.x)[h:-h, h:-h]


def _abs_terms(case, want):
    """Sum of |terms| of each reduction: the scale its rounding error follows."""
    f, k = case.fields, case.kernel
    I = lambda a: _interior(case, a)  # noqa: E731
    if k in ("cg_init", "cg_solver_init"):
        return {"rro": np.sum(I(want["r"]) ** 2)}
    if k in ("cg_calc_w", "cg_solver_calc_w"):
        return {"pw": np.sum(np.abs(I(want["w"]) * I(f["p"])))}
    if k in ("cg_calc_ur", "cg_solver_calc_ur"):
        return {"rrn": np.sum(I(want["r"]) ** 2)}
    if k == "calculate_2norm":
        return {"norm": np.sum(I(f["r"]) ** 2)}
    if k in ("jacobi_iterate", "jacobi_solver_iterate"):
        return {"error": np.sum(np.abs(I(want["u"]) - I(want["r"])))}
    if k == "field_summary":
        V = np.abs(I(f["volume"]))
        m = V * np.abs(I(f["density"]))
        return {"vol": V.sum(), "mass": m.sum(), "ie": np.sum(m * np.abs(I(f["energy0"]))),
                "temp": np.sum(m * np.abs(I(f["u"])))}
    return {}


def _check(case, got, want):
    scale = _abs_terms(case, want)
    for k, w in want.items():
        g = got[k]
        if isinstance(w, np.ndarray):
            if not np.array_equal(g, w):
                bad = np.flatnonzero(g != w)
                raise AssertionError(f"{k}: {bad.size} cells differ, first {bad[:5]}: "
                                     f"gpu {g[bad[:3]]} oracle {w[bad[:3]]}")
        else:
            tol = SCALAR_RTOL * (scale.get(k, 0.0) + abs(case.scalars.get(k, 0.0)))
            assert abs(g - w) <= tol, (k, g, w, tol)


@pytest.mark.parametrize("case", CASES, ids=[kc.case_id(c) + f"-{c.x}x{c.y}x{c.z}h{c.hd}" for c in CASES])
def test_kernel_matches_oracle(hip, case):
    want = kc.run_host(load_oracle().dll, "or%dd_" % case.dims, case)
    got = kc.run_gpu(case)
    _check(case, got, want)


def test_kernels_match_golden_fixtures(hip):
    data = np.load(kc.GOLDEN_FILE, allow_pickle=False)
    for case in kc.golden_cases():
        kc.with_golden_inputs(case, data)
        got = kc.run_gpu(case)
        cid = kc.case_id(case)
        for k, v in got.items():
            want = data[f"{cid}:{k}"]
            if isinstance(v, np.ndarray):
                assert np.array_equal(v, want), f"{cid}:{k}"
            else:
                assert abs(v - float(want)) <= SCALAR_RTOL * max(abs(float(want)), 1.0), f"{cid}:{k}"


def test_lane_shift_helpers(hip):
    from tealeaf_amd.device import DeviceArray
    src = np.arange(64, dtype=np.float64) * 1.5 + 0.25
    d_in = DeviceArray.from_host(src)
    d_out = DeviceArray(128)
    hip.tlhip_selftest_lanes(d_in.ptr, d_out.ptr)
    out = d_out.download()
    prev = np.concatenate([[-1.0], src[:-1]])
    nxt = np.concatenate([src[1:], [-1.0]])
    assert np.array_equal(out[:64], prev)
    assert np.array_equal(out[64:], nxt)


def test_reductions_are_deterministic(hip):
    case = kc.make_cases(2, 300, 210, seed=7)[1]     # cg_calc_w
    assert case.kernel == "cg_calc_w"
    a = kc.run_gpu(case)["pw"]
    b = kc.run_gpu(case)["pw"]
    assert a == b


def test_tile_rows_do_not_change_fields(hip):
    case = kc.make_cases(2, 200, 97, seed=8)[1]
    base = kc.run_gpu(case)
    for rows in (1, 3, 17, 200):
        hip.tlhip_set_tile_rows(rows)
        try:
            got = kc.run_gpu(case)
        finally:
            hip.tlhip_set_tile_rows(-1)
        assert np.array_equal(got["w"], base["w"])
        assert abs(got["pw"] - base["pw"]) <= SCALAR_RTOL * abs(base["pw"])


STENCIL_3D = ("cg_solver_calc_w", "ppcg_inner_iteration", "calculate_residual",
              "cheby_solver_iterate", "cheby_solver_init", "jacobi_solver_iterate",
              "cg_solver_init")


@pytest.mark.parametrize("planes", [1, 2, 3, 7, 64])
def test_tile_planes_match_oracle(hip, planes):
    """The 3D stencil sweeps carry planes k-1, k of the field and kz along z
    within a tile of `planes` planes: every tile height (one plane, tiles
    that split z unevenly, one tile for the whole depth) gives the oracle's
    fields bit for bit."""
    cases = [c for c in kc.make_cases(3, 19, 11, 13, seed=35) + kc.make_cases(3, 9, 8, 10, hd=1, seed=36)
             if c.kernel in STENCIL_3D]
    assert {c.kernel for c in cases} >= set(STENCIL_3D) - {"cg_solver_init"}
    hip.tlhip_set_tile_planes(planes)
    try:
        for case in cases:
            want = kc.run_host(load_oracle().dll, "or3d_", case)
            _check(case, kc.run_gpu(case), want)
    finally:
        hip.tlhip_set_tile_planes(-1)


def test_errors_are_reported_not_fatal(hip):
    from tealeaf_amd.abi import TealeafError
    from tealeaf_amd.device import DeviceArray
    d = DeviceArray(16 * 16)
    pw = ctypes.c_double(0)
    with pytest.raises(TealeafError, match="distinct"):
        hip.tlhip2d_cg_calc_w(16, 16, 2, ctypes.addressof(pw), d.ptr, d.ptr, d.ptr, d.ptr)
    with pytest.raises(TealeafError, match="Coefficient 3"):
        hip.tlhip2d_cg_init(16, 16, 2, 3, 0.1, 0.1, ctypes.addressof(pw), d.ptr, d.ptr, d.ptr,
                            d.ptr, d.ptr, d.ptr, d.ptr, d.ptr)
    with pytest.raises(TealeafError, match="Incorrect face"):
        buf = np.zeros(64)
        hip.tlhip2d_pack_or_unpack(16, 16, 1, 2, 7, True, d.ptr, buf.ctypes.data)
