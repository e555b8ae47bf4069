"""The FP32 fused multiply-add emulation against exact rational arithmetic."""
from fractions import Fraction

import numpy as np

from tests.fp32_emulation import fma32

f = np.float32


def _round32(q: Fraction) -> np.float32:
    """Nearest float32 to the rational q, ties to even (normal range)."""
    lo = np.float32(float(q))
    cands = [lo, np.nextafter(lo, f(np.inf)), np.nextafter(lo, f(-np.inf))]
    return min(cands, key=lambda v: (abs(Fraction(float(v)) - q),
                                     int(np.asarray(v).view(np.int32)) & 1))


def test_exact_cases():
    assert fma32(f(3), f(5), f(7)) == f(22)
    # (1 + 2^-12)^2 - 1 = 2^-11 + 2^-24: the fused result keeps 2^-24, a
    # rounded product loses it
    t = f(1 + 2.0 ** -12)
    assert fma32(t, t, f(-1)) == f(2.0 ** -11 + 2.0 ** -24)
    assert (t * t) - f(1) == f(2.0 ** -11)


def test_double_rounding_case():
    """a*b + 1 lies just above the float32 midpoint 1 + 2^-24 but rounds to
    it in float64: rounding the float64 sum gives 1, the fused result
    1 + 2^-23."""
    a = f(float.fromhex("0x1.b9acdep-25"))
    b = f(float.fromhex("0x1.28c2d2p+0"))
    naive = np.float32(np.float64(a) * np.float64(b) + 1.0)
    assert naive == f(1.0)
    want = _round32(Fraction(float(a)) * Fraction(float(b)) + 1)
    assert want == f(1 + 2.0 ** -23)
    assert fma32(a, b, f(1)) == want


def test_random_against_rationals():
    rng = np.random.default_rng(3)
    v = rng.uniform(-2, 2, (3, 500)).astype(np.float32)
    got = fma32(v[0], v[1], v[2])
    for i in range(v.shape[1]):
        ex = Fraction(float(v[0, i])) * Fraction(float(v[1, i])) + Fraction(float(v[2, i]))
        assert got[i] == _round32(ex), i
