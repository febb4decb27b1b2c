"""Field files and full <-> reduced maps through the C ABI (host only).

``output_fields`` / ``input_fields`` follow crysSCFT_sample.cpp:271-292 (one
"phiA phiB wA wB" line per point); ``extract`` / ``expand`` follow the init
lambdas of main() (crysSCFT_sample.cpp:202-258) and their mirror inverse.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check, dptr, lib

KINDS = {"P1": _lib.P1, "3m": _lib.M3, "PmmmDCT": _lib.PMMM_DCT}


def output_fields(filename, phiA, phiB, wA, wB, digits=0):
    arrs = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (phiA, phiB, wA, wB)]
    n = arrs[0].size
    check(lib().crys_output_fields(*[dptr(a) for a in arrs], n, str(filename).encode(),
                                   int(digits)))


def input_fields(filename, size):
    arrs = [np.zeros(size) for _ in range(4)]
    check(lib().crys_input_fields(*[dptr(a) for a in arrs], int(size), str(filename).encode()))
    return tuple(arrs)  # phiA, phiB, wA, wB


def _reduced_shape(kind, grid):
    return tuple(grid) if kind == _lib.P1 else tuple(g // 2 for g in grid)


def extract(full, solver):
    kind = KINDS[solver]
    full = np.ascontiguousarray(full, dtype=np.float64)
    g = (ctypes.c_size_t * 3)(*full.shape)
    out = np.empty(_reduced_shape(kind, full.shape))
    check(lib().crys_extract_fields(kind, g, dptr(full), dptr(out)))
    return out


def expand(reduced, solver, grid_log):
    kind = KINDS[solver]
    reduced = np.ascontiguousarray(reduced, dtype=np.float64)
    g = (ctypes.c_size_t * 3)(*grid_log)
    out = np.empty(tuple(grid_log))
    check(lib().crys_expand_fields(kind, g, dptr(reduced), dptr(out)))
    return out
