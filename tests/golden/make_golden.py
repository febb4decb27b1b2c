"""Writes the golden fixtures of tests/golden from scipy / numpy.

The reference ships no fixtures and its transforms come from FFTW3, which is
absent here.  scipy.fft.dctn(type=2) is FFTW's REDFT10 and type=3 is
REDFT01 (both unnormalised with norm=None); numpy.fft.rfftn / irfftn are
r2c / c2r (irfftn divides by N, undone below).  From those definitions this
script writes inputs and expected outputs for:

  transforms.npz  REDFT10^3, REDFT01^3, r2c, c2r on small grids
  diffusion.npz   one diffusion() of each solver (solver_p1.h:83-202,
                  solver_3m.h:143-324, solver_Pmmm.h:125-215) computed with
                  the library routines above and the reference's K
                  (solver_p1.h:142-172, solver_3m.h:184-235,
                  solver_Pmmm.h:159-186)

Run:  python tests/golden/make_golden.py   (needs scipy; only the .npz files
are read by the tests).
"""
from __future__ import annotations

import os

import numpy as np
import scipy.fft as sf

HERE = os.path.dirname(os.path.abspath(__file__))
PI = 3.1415926535897932384626


def k2_p1(N, L):
    k = np.array([i if i < (N + 1) // 2 else N - i for i in range(N)], dtype=np.float64)
    return (2 * PI * k * 1.0 / L) ** 2


def diffusion_p1(f, cell, ds):
    Nx, Ny, Nz = f.shape
    kx, ky = k2_p1(Nx, cell[0]), k2_p1(Ny, cell[1])
    kz = (2 * PI * np.arange(Nz // 2 + 1) * 1.0 / cell[2]) ** 2
    K = np.exp(-ds * (kx[:, None, None] + ky[None, :, None] + kz[None, None, :]))
    F = np.fft.rfftn(f)
    g = np.fft.irfftn(F * K, s=f.shape, axes=(0, 1, 2)) * f.size  # c2r, unnormalised
    return g * (1.0 / f.size)


def diffusion_pmmm(e, L, ds):
    X, Y, Z = e.shape
    kx = (2 * PI * np.arange(X) * 1.0 / L) ** 2
    ky = (2 * PI * np.arange(Y) * 1.0 / L) ** 2
    kz = (2 * PI * np.arange(Z) * 1.0 / L) ** 2
    K = np.exp(-(kx[:, None, None] + ky[None, :, None] + kz[None, None, :]) * ds)
    s1 = sf.dctn(e, type=2)
    out = sf.dctn(s1 * K, type=3)
    return out * (1.0 / (8 * e.size))


def expand_3m(e):
    f = e
    for ax in range(3):
        M = f.shape[ax]
        shape = list(f.shape)
        shape[ax] = 2 * M
        g = np.empty(shape)
        ev = [slice(None)] * 3
        od = [slice(None)] * 3
        ev[ax] = slice(0, 2 * M, 2)
        od[ax] = slice(1, 2 * M, 2)
        g[tuple(ev)] = f
        g[tuple(od)] = np.flip(f, axis=ax)
        f = g
    return f


def main():
    rng = np.random.default_rng(20241203)
    tr = {}
    for i, sh in enumerate([(4, 6, 8), (8, 8, 16), (3, 5, 8), (6, 12, 16)]):
        a = rng.standard_normal(sh)
        c = np.fft.rfftn(rng.standard_normal(sh))
        tr[f"in{i}"] = a
        tr[f"dct2_{i}"] = sf.dctn(a, type=2)
        tr[f"dct3_{i}"] = sf.dctn(a, type=3)
        tr[f"rfft_{i}"] = np.fft.rfftn(a)
        tr[f"cin{i}"] = c
        tr[f"irfft_{i}"] = np.fft.irfftn(c, s=sh, axes=(0, 1, 2)) * np.prod(sh)
    np.savez(os.path.join(HERE, "transforms.npz"), **tr)

    df = {}
    cases = [((8, 12, 16), (4.0, 4.1, 4.3), 0.01), ((16, 16, 16), (4.0, 4.0, 4.0), 0.01),
             ((12, 8, 16), (3.7, 4.4, 4.0), 0.005)]
    for i, (N, cell, ds) in enumerate(cases):
        full = rng.standard_normal(N)
        df[f"grid{i}"] = np.array(N)
        df[f"cell{i}"] = np.array(cell)
        df[f"ds{i}"] = np.array(ds)
        df[f"p1_in{i}"] = full
        df[f"p1_out{i}"] = diffusion_p1(full, cell, ds)
        e = rng.standard_normal(tuple(n // 2 for n in N))
        df[f"pmmm_in{i}"] = e
        df[f"pmmm_out{i}"] = diffusion_pmmm(e, cell[0], ds)  # cubic: cell_para[0]
        e3 = rng.standard_normal(tuple(n // 2 for n in N))
        df[f"m3_in{i}"] = e3
        df[f"m3_out{i}"] = diffusion_p1(expand_3m(e3), cell, ds)[::2, ::2, ::2]
    np.savez(os.path.join(HERE, "diffusion.npz"), **df)


if __name__ == "__main__":
    main()
