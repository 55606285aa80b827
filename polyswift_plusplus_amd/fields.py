"""Seeded initial fields for the SCFT workloads.

The reference seeds w with U(-0.5, 0.5) or slab/sine profiles
(psphysf/PsMonomerDens.cpp:116-161); the workloads here instead ask for
i.i.d. Gaussian fields N(0, sigma) plus a lamellar or hexagonal-cylinder
seed.  The generator is counter-based so that every consumer (engine, tests,
oracle runs) gets bit-identical arrays on any host:

* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) with counter
  (cell, field, instance, 0) and key (seed low 32 bits, seed high 32 bits);
* one Box-Muller normal per cell from the first two output words:
  u1 = (x0 + 1) / 2^32, u2 = x1 / 2^32, z = sqrt(-2 ln u1) cos(2 pi u2).

Field ids: 0 = w_A, 1 = w_B.
"""
from __future__ import annotations

import numpy as np

_M0 = np.uint64(0xD2511F53)
_M1 = np.uint64(0xCD9E8D57)
_W0 = 0x9E3779B9
_W1 = 0xBB67AE85
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0: int, k1: int):
    """Vectorised Philox4x32-10 over uint32 counter arrays."""
    c0 = np.asarray(c0, dtype=np.uint64) & _MASK
    c1 = np.asarray(c1, dtype=np.uint64) & _MASK
    c2 = np.asarray(c2, dtype=np.uint64) & _MASK
    c3 = np.asarray(c3, dtype=np.uint64) & _MASK
    for _ in range(10):
        p0 = _M0 * c0
        p1 = _M1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK
        c0, c1, c2, c3 = (hi1 ^ c1 ^ np.uint64(k0)) & _MASK, lo1, (hi0 ^ c3 ^ np.uint64(k1)) & _MASK, lo0
        k0 = (k0 + _W0) & 0xFFFFFFFF
        k1 = (k1 + _W1) & 0xFFFFFFFF
    return c0, c1, c2, c3


def gaussian(shape, sigma: float, seed: int, field: int, instance: int = 0) -> np.ndarray:
    """i.i.d. N(0, sigma^2) per cell, C order."""
    n = int(np.prod(shape))
    cell = np.arange(n, dtype=np.uint64)
    x0, x1, _, _ = philox4x32_10(cell, np.full(n, field, np.uint64),
                                 np.full(n, instance, np.uint64), np.zeros(n, np.uint64),
                                 seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)
    u1 = (x0.astype(np.float64) + 1.0) / 4294967296.0
    u2 = x1.astype(np.float64) / 4294967296.0
    z = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
    return (sigma * z).reshape(shape)


def coords(shape, cell):
    """Cell-corner coordinates x_i = i dx (the reference places field value i
    at i*dx, psgrid/PsUniCartGrid.cpp:30-43)."""
    ax = [np.arange(n) * h for n, h in zip(shape, cell)]
    return np.meshgrid(*ax, indexing="ij")


def lamellar(shape, cell, amplitude: float = 1.0) -> np.ndarray:
    """amplitude * cos(2 pi x / Lx)."""
    x, _, _ = coords(shape, cell)
    Lx = shape[0] * cell[0]
    return amplitude * np.cos(2.0 * np.pi * x / Lx)


# Hexagonal-cylinder seed in a cubic box.  A cube is not hex-commensurate, so
# the three in-plane wavevectors (in units of 2 pi / L) are chosen as
# (2, 0), (-1, 2), (-1, -2): they sum to zero (the triangle condition of a
# hexagonal lattice) with magnitudes 2, sqrt5, sqrt5.  Cylinders run along z.
HEX_MODES = ((2, 0), (-1, 2), (-1, -2))


def hex_cylinders(shape, cell, amplitude: float = 1.0) -> np.ndarray:
    """amplitude * mean of cos(k_i . r) over HEX_MODES."""
    x, y, _ = coords(shape, cell)
    Lx, Ly = shape[0] * cell[0], shape[1] * cell[1]
    h = np.zeros(shape)
    for a, b in HEX_MODES:
        h += np.cos(2.0 * np.pi * (a * x / Lx + b * y / Ly))
    return amplitude * h / len(HEX_MODES)


def seeded_fields(shape, cell, sigma=0.1, seed=1, instance=0, seed_kind=None, amplitude=1.0):
    """(w_A, w_B): Gaussian noise plus an optional ordered seed s(r) added as
    w_A += s, w_B -= s (the A-rich regions sit where s < 0)."""
    wA = gaussian(shape, sigma, seed, 0, instance)
    wB = gaussian(shape, sigma, seed, 1, instance)
    if seed_kind == "lamellar":
        s = lamellar(shape, cell, amplitude)
    elif seed_kind == "hex":
        s = hex_cylinders(shape, cell, amplitude)
    elif seed_kind in (None, "none"):
        s = 0.0
    else:
        raise ValueError(f"unknown seed kind {seed_kind!r}")
    return wA + s, wB - s
