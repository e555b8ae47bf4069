"""The allocation-free entry point tlhip2d_cg_calc_w_noalloc against
tlhip2d_cg_calc_w through the C ABI: w bit-exact and the dot product
bit-identical (the same kernel and the same summation order; the result
arrives in a device double instead of through the runtime's mailbox), on
two calls in a row, the second with a smaller grid, so the shared scratch
array holds the first call's partials when the second runs; an odd row
length (one cell per lane); and shapes beyond what the entry point
accepts, which are rejected with a message before any launch.
"""
import ctypes

import numpy as np
import pytest

from tealeaf_amd.device import DeviceArray

pytestmark = pytest.mark.gpu


def _fields(x, y, seed):
    rng = np.random.default_rng(seed)
    n = x * y
    return {"p": rng.uniform(-1, 1, n), "kx": rng.uniform(0.05, 0.3, n),
            "ky": rng.uniform(0.05, 0.3, n)}


def _both(hip, x, y, hd, seed, pw0=0.25):
    f = _fields(x, y, seed)
    d = {k: DeviceArray.from_host(v) for k, v in f.items()}
    w_a = DeviceArray.from_host(np.full(x * y, np.nan))
    w_b = DeviceArray.from_host(np.full(x * y, np.nan))
    # the entry point with the library's allocated slots: pw on the host
    pw = ctypes.c_double(pw0)
    hip.tlhip2d_cg_calc_w(x, y, hd, ctypes.addressof(pw), d["p"].ptr, w_a.ptr, d["kx"].ptr,
                          d["ky"].ptr)
    # the allocation-free one: pw in a device double
    pw_dev = DeviceArray.from_host(np.array([pw0]))
    hip.tlhip2d_cg_calc_w_noalloc(x, y, hd, pw_dev.ptr, d["p"].ptr, w_b.ptr, d["kx"].ptr,
                                  d["ky"].ptr)
    hip.tlhip_synchronize()
    got_pw = float(pw_dev.download()[0])
    a, b = w_a.download(), w_b.download()
    for v in list(d.values()) + [w_a, w_b, pw_dev]:
        v.free()
    return pw.value, got_pw, a, b


def _interior(a, x, y, hd):
    return a.reshape(y, x)[hd:y - hd, hd:x - hd]


def test_noalloc_matches_calc_w_twice(hip):
    print("[no_alloc] 1030x900 then 130x70", flush=True)
    for (x, y, hd), seed in (((1030, 900, 2), 1), ((130, 70, 2), 2)):
        pw, got, a, b = _both(hip, x, y, hd, seed)
        assert got == pw, (x, y, got, pw)
        np.testing.assert_array_equal(_interior(b, x, y, hd), _interior(a, x, y, hd))


def test_noalloc_odd_rows(hip):
    print("[no_alloc] 77x41 hd 1", flush=True)
    pw, got, a, b = _both(hip, 77, 41, 1, 3)
    assert got == pw
    np.testing.assert_array_equal(_interior(b, 77, 41, 1), _interior(a, 77, 41, 1))


@pytest.mark.parametrize("x,y", [(8388868, 5), (46342, 46342)],
                         ids=["rows-over-65536-strips", "over-2^31-cells"])
def test_noalloc_rejects_oversize(hip, x, y):
    """Shapes the host checks turn into an error before any launch: a row of
    8388864 interior cells, 65538 strips of 128 (or 131076 of 64) where the
    scratch holds 65536, and a grid of more than 2^31 - 1 cells.  The
    buffers are never touched."""
    dummy = DeviceArray(16)
    rc = hip.dll.tlhip2d_cg_calc_w_noalloc(x, y, 2, dummy.ptr, dummy.ptr + 32, dummy.ptr + 64,
                                           dummy.ptr + 96, dummy.ptr + 112)
    dummy.free()
    assert rc != 0
    msg = hip.last_error()
    assert "too long" in msg or "too large" in msg, msg
