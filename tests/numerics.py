"""Independent numpy statements of the operator, for checks that must not
reuse the code under test.

A u is the reference's SMVP (2d/shared.h:50-54, 3d/shared.h:52-58) over the
interior, with the depth-1 reflective halo a single chunk gets from
local_halos (a neighbour outside the interior is the cell itself).
"""
import numpy as np


def _mirror1(a, hd):
    """Copy of `a` whose depth-1 halo layer mirrors the interior faces."""
    b = a.copy()
    if a.ndim == 2:
        y, x = a.shape
        J, K = slice(hd, y - hd), slice(hd, x - hd)
        b[J, hd - 1], b[J, x - hd] = b[J, hd], b[J, x - hd - 1]
        b[hd - 1, K], b[y - hd, K] = b[hd, K], b[y - hd - 1, K]
        return b
    z, y, x = a.shape
    I, J, K = slice(hd, z - hd), slice(hd, y - hd), slice(hd, x - hd)
    b[I, J, hd - 1], b[I, J, x - hd] = b[I, J, hd], b[I, J, x - hd - 1]
    b[I, hd - 1, K], b[I, y - hd, K] = b[I, hd, K], b[I, y - hd - 1, K]
    b[hd - 1, J, K], b[z - hd, J, K] = b[hd, J, K], b[z - hd - 1, J, K]
    return b


def apply_a(u, kx, ky, hd, kz=None):
    """Interior of A u, same shape as the interior."""
    m = _mirror1(u, hd)
    if u.ndim == 2:
        y, x = u.shape
        J, K = slice(hd, y - hd), slice(hd, x - hd)
        Jp, Kp = slice(hd + 1, y - hd + 1), slice(hd + 1, x - hd + 1)
        Jm, Km = slice(hd - 1, y - hd - 1), slice(hd - 1, x - hd - 1)
        c = m[J, K]
        kxc, kxr, kyc, kyu = kx[J, K], kx[J, Kp], ky[J, K], ky[Jp, K]
        diag = (1.0 + (kxr + kxc)) + (kyu + kyc)
        return diag * c - (kxr * m[J, Kp] + kxc * m[J, Km]) - (kyu * m[Jp, K] + kyc * m[Jm, K])
    z, y, x = u.shape
    I, J, K = slice(hd, z - hd), slice(hd, y - hd), slice(hd, x - hd)
    Ip, Jp, Kp = slice(hd + 1, z - hd + 1), slice(hd + 1, y - hd + 1), slice(hd + 1, x - hd + 1)
    Im, Jm, Km = slice(hd - 1, z - hd - 1), slice(hd - 1, y - hd - 1), slice(hd - 1, x - hd - 1)
    c = m[I, J, K]
    kxc, kxr = kx[I, J, K], kx[I, J, Kp]
    kyc, kyu = ky[I, J, K], ky[I, Jp, K]
    kzc, kzf = kz[I, J, K], kz[Ip, J, K]
    diag = (1.0 + (kxr + kxc)) + (kyu + kyc) + (kzf + kzc)
    return (diag * c - (kxr * m[I, J, Kp] + kxc * m[I, J, Km]) - (kyu * m[I, Jp, K] + kyc * m[I, Jm, K])
            - (kzf * m[Ip, J, K] + kzc * m[Im, J, K]))


def rel_residual(u, u0, kx, ky, hd, kz=None):
    """||u0 - A u||_2 / ||u0||_2 over the interior."""
    inner = tuple(slice(hd, s - hd) for s in u.shape)
    r = u0[inner] - apply_a(u, kx, ky, hd, kz)
    return float(np.sqrt(np.sum(r * r) / np.sum(u0[inner] ** 2)))
