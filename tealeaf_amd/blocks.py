"""Seeded piecewise-constant input fields (density, energy per block).

The reference builds its fields from geometric states only
(2d/c_kernels/omp3/set_chunk_state.c:9-90).  The heterogeneous workloads
(3D 256^3 on 16^3-cell blocks with log-uniform density in [0.1, 100]; 2D
8192^2 on 64x64-cell blocks with density contrast 1e3 / 1e6) instead draw one
density and one energy per block from a counter-based generator, so the CPU
oracle and the GPU suite see identical inputs without storing them:

* generator: SplitMix64 (Steele, Lea & Flood, "Fast splittable pseudorandom
  number generators", OOPSLA 2014; the 64-bit finaliser published with
  Vigna's xorshift/xoroshiro family), i.e. output n of a stream seeded with
  `seed` is mix(seed + n * 0x9E3779B97F4A7C15);
* block b (blocks numbered x fastest, then y, then z) takes outputs 2b+1
  (density) and 2b+2 (energy); a uniform in [0, 1) is the top 53 bits
  times 2^-53;
* density = exp(ln(lo) + U * (ln(hi) - ln(lo))) (log-uniform),
  energy = e_lo + U * (e_hi - e_lo).

The arrays cover the whole chunk including halos; halo cells take the value
of the nearest interior cell.  They are handed to Problem.inject, which (as
the reference does after set_chunk_state) sets u = energy * density.
"""
from __future__ import annotations

import numpy as np

GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: np.ndarray) -> np.ndarray:
    """Output number n (1-based: n=1 is the first) of SplitMix64 seeded with `seed`."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + np.asarray(n, dtype=np.uint64) * GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, n: np.ndarray) -> np.ndarray:
    """Uniform doubles in [0, 1) from outputs n of the stream."""
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def block_values(seed: int, nblocks: int, rho: tuple[float, float], energy: tuple[float, float]):
    """(density, energy) per block, block ids 0..nblocks-1."""
    b = np.arange(nblocks, dtype=np.uint64)
    u_rho = uniform(seed, 2 * b + 1)
    u_e = uniform(seed, 2 * b + 2)
    lo, hi = np.log(rho[0]), np.log(rho[1])
    dens = np.exp(lo + u_rho * (hi - lo))
    en = energy[0] + u_e * (energy[1] - energy[0])
    return dens, en


def block_fields(cells: tuple[int, ...], block: int, seed: int, rho=(0.1, 100.0),
                 energy=(0.1, 25.0), halo_depth: int = 2):
    """Full-chunk density and energy arrays (halos included), shaped
    (y, x) in 2D or (z, y, x) in 3D, for `cells` = (nx, ny[, nz]) interior
    cells cut into cubes/squares of `block` cells (the last block along an
    axis may be partial)."""
    dims = len(cells)
    nb = [-(-c // block) for c in cells]                    # blocks per axis
    dens_b, en_b = block_values(seed, int(np.prod(nb)), rho, energy)
    hd = halo_depth
    # block coordinate of every padded index along each axis (halo clamps)
    idx = [np.clip(np.arange(c + 2 * hd) - hd, 0, c - 1) // block for c in cells]
    if dims == 2:
        bid = idx[1][:, None] * nb[0] + idx[0][None, :]
    else:
        bid = (idx[2][:, None, None] * nb[1] + idx[1][None, :, None]) * nb[0] + idx[0][None, None, :]
    return dens_b[bid], en_b[bid]
