"""Deterministic BCC-seeded Pmmm fields and the full -> reduced extractions.

The reference reads its initial fields from ``BCC_fields_test.txt``
(crysSCFT_sample.cpp:196), which is not shipped.  These generators make
fields of the same kind: Gaussians at the BCC sites (0,0,0) and (1/2,1/2,1/2)
on cell-centred coordinates x_j = (j + 1/2) L / N, so that every axis has the
mirror f(N-1-j) = f(j) the 3m and Pmmm solvers assume, plus seeded noise
averaged over the 8 reflections.  The noise comes from splitmix64 and
Box-Muller, so any platform reproduces the same doubles.
"""
from __future__ import annotations

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed: int, count: int) -> np.ndarray:
    """``count`` consecutive outputs of splitmix64 started at ``seed``."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def normal_noise(seed: int, shape) -> np.ndarray:
    """Standard normals by Box-Muller from splitmix64 uniforms."""
    n = int(np.prod(shape))
    m = (n + 1) // 2
    r = splitmix64(seed, 2 * m)
    u1 = ((r[0::2] >> np.uint64(11)).astype(np.float64) + 1.0) * 2.0 ** -53  # (0, 1]
    u2 = (r[1::2] >> np.uint64(11)).astype(np.float64) * 2.0 ** -53          # [0, 1)
    rad = np.sqrt(-2.0 * np.log(u1))
    z = np.empty(2 * m)
    z[0::2] = rad * np.cos(2.0 * np.pi * u2)
    z[1::2] = rad * np.sin(2.0 * np.pi * u2)
    return z[:n].reshape(shape)


def mirror_symmetrise(f: np.ndarray) -> np.ndarray:
    """Average over the 8 reflections j -> N-1-j of the three axes."""
    g = f
    for ax in range(3):
        g = 0.5 * (g + np.flip(g, axis=ax))
    return np.ascontiguousarray(g)


def bcc_phi(grid, cell, sigma_frac=0.12, amplitude=0.6) -> np.ndarray:
    """Sum of isotropic Gaussians at the BCC sites, minimum-image distance."""
    grid = tuple(int(v) for v in grid)
    cell = (cell,) * 3 if np.isscalar(cell) else tuple(float(v) for v in cell)
    coords = [(np.arange(N) + 0.5) * L / N for N, L in zip(grid, cell)]
    sig = sigma_frac * min(cell)
    phi = np.zeros(grid)
    for site in ((0.0, 0.0, 0.0), (0.5, 0.5, 0.5)):
        d2 = np.zeros(grid)
        for ax, (x, L) in enumerate(zip(coords, cell)):
            d = x - site[ax] * L
            d = d - L * np.round(d / L)
            shape = [1, 1, 1]
            shape[ax] = -1
            d2 = d2 + (d * d).reshape(shape)
        phi += amplitude * np.exp(-d2 / (2.0 * sig * sig))
    return phi


def bcc_fields(grid, cell, fA, xN, seed=20241203, noise=1e-3, phi_ref=None):
    """Full-grid (wA, wB) with wA = xN (phi_ref - phiA0), wB = -wA.

    ``phi_ref`` defaults to fA (the 0.25 of the plain-method workload is its
    fA).  The result is invariant under the three mirrors."""
    phi0 = bcc_phi(grid, cell)
    phi0 = phi0 + noise * normal_noise(seed, phi0.shape)
    phi0 = mirror_symmetrise(phi0)
    ref = fA if phi_ref is None else phi_ref
    wA = xN * (ref - phi0)
    wB = -wA
    return np.ascontiguousarray(wA), np.ascontiguousarray(wB)


def extract(full: np.ndarray, solver: str) -> np.ndarray:
    """Reduced field a solver works on (crysSCFT_sample.cpp:202-258).

    P1: the full grid; 3m: even indices f[2x,2y,2z]; PmmmDCT: first octant."""
    if solver == "P1":
        return np.ascontiguousarray(full)
    if solver == "3m":
        return np.ascontiguousarray(full[::2, ::2, ::2])
    if solver in ("PmmmDCT", "Pmmm"):
        nx, ny, nz = (s // 2 for s in full.shape)
        return np.ascontiguousarray(full[:nx, :ny, :nz])
    raise ValueError(solver)


def m3_operations(g):
    """The 8 symmetry operations of a 3m solver with translational part g
    (generate_m3_operation, solver_utilities.h:55-66): (mirrored axes,
    translation in cell fractions) for identity, -z, -y, -y-z, -x, -x-z,
    -x-y, -x-y-z."""
    g = [float(v) for v in g]
    return [((), (0.0, 0.0, 0.0)),
            ((2,), (g[0], g[1], g[2])),
            ((1,), (g[3], g[4], g[5])),
            ((1, 2), (g[0] + g[3], -g[1] + g[4], g[2] + g[5])),
            ((0,), (g[6], g[7], g[8])),
            ((0, 2), (g[6] - g[0], g[1] + g[7], g[2] + g[8])),
            ((0, 1), (-g[3] + g[6], g[4] + g[7], g[5] + g[8])),
            ((0, 1, 2), (-g[0] - g[3] + g[6], -g[1] + g[4] + g[7], g[2] + g[5] + g[8]))]


def expand_3m(reduced: np.ndarray, g) -> np.ndarray:
    """Full-grid field of a 3m solver's reduced field (the even sub-lattice
    f[2x,2y,2z]) with translational part g: f(op(r)) = f(r) for the 8
    operations, where on an axis of length N an image index is 2j + tN, or
    -1 - 2j + tN on a mirrored axis (mod N).  Raises ValueError unless the
    operations map the sub-lattice one to one onto the grid."""
    shape = tuple(2 * s for s in reduced.shape)
    full = np.zeros(shape)
    hits = np.zeros(shape, dtype=np.int64)
    idx = np.indices(reduced.shape)
    for mirrored, t in m3_operations(g):
        im = []
        for a in range(3):
            s = t[a] * shape[a]
            if abs(s - round(s)) > 1e-9:
                raise ValueError("translational part is not a whole number of grid points")
            s = int(round(s))
            j = 2 * idx[a]
            im.append(((-1 - j + s) if a in mirrored else (j + s)) % shape[a])
        full[im[0], im[1], im[2]] = reduced
        np.add.at(hits, (im[0], im[1], im[2]), 1)
    if not (hits == 1).all():
        raise ValueError("the operations do not map the 2x2y2z sub-lattice onto the grid")
    return full


def expand(reduced: np.ndarray, solver: str) -> np.ndarray:
    """Inverse of :func:`extract` for mirror-symmetric fields."""
    if solver == "P1":
        return np.ascontiguousarray(reduced)
    f = reduced
    if solver in ("PmmmDCT", "Pmmm"):
        for ax in range(3):
            f = np.concatenate([f, np.flip(f, axis=ax)], axis=ax)
        return np.ascontiguousarray(f)
    if solver == "3m":
        # f[2j] = e[j], f[2j+1] = e[M-1-j] on every axis
        for ax in range(3):
            M = f.shape[ax]
            shape = list(f.shape)
            shape[ax] = 2 * M
            g = np.empty(shape)
            sl_even = [slice(None)] * 3
            sl_odd = [slice(None)] * 3
            sl_even[ax] = slice(0, 2 * M, 2)
            sl_odd[ax] = slice(1, 2 * M, 2)
            g[tuple(sl_even)] = f
            g[tuple(sl_odd)] = np.flip(f, axis=ax)
            f = g
        return np.ascontiguousarray(f)
    raise ValueError(solver)
