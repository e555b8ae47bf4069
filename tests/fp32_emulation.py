"""Exact emulation of the FP32 fused multiply-add on the CPU, for checking
FP32 kernels that use it."""
import numpy as np


def fma32(a, b, c):
    """a * b + c rounded once to float32, as the device's fused multiply-add.
    The product of two floats is exact in float64; TwoSum gives the error of
    the float64 sum; rounding that sum to odd keeps it exact in the last bit
    a float32 rounding can see (53 >= 24 + 2), so the final conversion
    rounds the exact value."""
    a64 = np.asarray(a, dtype=np.float32).astype(np.float64)
    b64 = np.asarray(b, dtype=np.float32).astype(np.float64)
    c64 = np.asarray(c, dtype=np.float32).astype(np.float64)
    p = a64 * b64
    s = p + c64
    bb = s - p
    e = (p - (s - bb)) + (c64 - bb)
    even = (np.asarray(s).view(np.int64) & 1) == 0
    s = np.where((e != 0) & even, np.nextafter(s, np.where(e > 0, np.inf, -np.inf)), s)
    return s.astype(np.float32)
