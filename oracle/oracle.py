"""ctypes view of the CPU oracle (``liboracle.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` as the checker.  The product
package never imports it.

The oracle restates the reference's algorithm in C (crysSCFT_sample.cpp:294-531,
solver_p1.h, solver_3m.h, solver_Pmmm.h, solver_utilities.h); see ``oracle.h``
for what pins each function.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

P1, M3, PMMM = 0, 1, 2
KIND = {"P1": P1, "3m": M3, "M3": M3, "PmmmDCT": PMMM, "PMMM": PMMM, "Pmmm": PMMM}
TRACE_COLS = 7

_dp = ctypes.POINTER(ctypes.c_double)
_sz = ctypes.c_size_t


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc only)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", HERE, "liboracle.so"],
                       check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_dct2_3d.argtypes = [_dp, _dp, _sz, _sz, _sz]
        L.orc_dct3_3d.argtypes = [_dp, _dp, _sz, _sz, _sz]
        L.orc_rfft_3d.argtypes = [_dp, _dp, _sz, _sz, _sz]
        L.orc_irfft_3d.argtypes = [_dp, _dp, _sz, _sz, _sz]
        L.orc_solver_create.restype = ctypes.c_void_p
        L.orc_solver_create.argtypes = [ctypes.c_int, ctypes.POINTER(_sz), _dp,
                                        ctypes.c_char_p, _sz]
        L.orc_solver_destroy.argtypes = [ctypes.c_void_p]
        L.orc_set_cell_para.argtypes = [ctypes.c_void_p, _dp]
        L.orc_set_contour_step.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_diffusion.argtypes = [ctypes.c_void_p]
        L.orc_get_iomat.restype = _dp
        L.orc_get_iomat.argtypes = [ctypes.c_void_p]
        L.orc_get_space_size_phy.restype = _sz
        L.orc_get_space_size_phy.argtypes = [ctypes.c_void_p]
        L.orc_get_grid_size_phy.argtypes = [ctypes.c_void_p, ctypes.POINTER(_sz)]
        L.orc_get_cell_para.argtypes = [ctypes.c_void_p, _dp]
        L.orc_get_ds.restype = ctypes.c_double
        L.orc_get_ds.argtypes = [ctypes.c_void_p]
        L.orc_propagate.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _sz, ctypes.c_int]
        L.orc_scft_simple_iterate.restype = ctypes.c_int
        L.orc_scft_simple_iterate.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                              ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                              ctypes.c_double, ctypes.c_double, _sz, _dp, _dp,
                                              _dp, _dp, _dp]
        L.orc_set_num_threads.argtypes = [ctypes.c_int]
        L.orc_get_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def set_num_threads(n: int) -> None:
    lib().orc_set_num_threads(int(n))


def _t3(fn, a: np.ndarray, out_shape) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float64)
    out = np.zeros(out_shape, dtype=np.float64)
    fn(_p(a), _p(out), *[_sz(d) for d in a.shape[:3]])
    return out


def dct2_3d(a):
    return _t3(lib().orc_dct2_3d, a, a.shape)


def dct3_3d(a):
    return _t3(lib().orc_dct3_3d, a, a.shape)


def rfft_3d(a):
    nx, ny, nz = a.shape
    out = _t3(lib().orc_rfft_3d, a, (nx, ny, nz // 2 + 1, 2))
    return out[..., 0] + 1j * out[..., 1]


def irfft_3d(c, nz):
    nx, ny, nzc = c.shape
    assert nzc == nz // 2 + 1
    a = np.ascontiguousarray(np.stack([c.real, c.imag], axis=-1), dtype=np.float64)
    out = np.zeros((nx, ny, nz), dtype=np.float64)
    lib().orc_irfft_3d(_p(a), _p(out), _sz(nx), _sz(ny), _sz(nz))
    return out


class Solver:
    """Oracle restatement of CrysPSSolverBase (solver_base.h:8-51)."""

    def __init__(self, kind, grid_log, g=None):
        self.kind = KIND[kind] if isinstance(kind, str) else int(kind)
        grid = (_sz * 3)(*[int(v) for v in grid_log])
        gg = np.zeros(9) if g is None else np.ascontiguousarray(g, dtype=np.float64)
        err = ctypes.create_string_buffer(256)
        h = lib().orc_solver_create(self.kind, grid, _p(gg), err, 256)
        if not h:
            raise ValueError(err.value.decode())
        self._h = h
        self.grid_log = tuple(int(v) for v in grid_log)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_solver_destroy(h)
            self._h = None

    def set_cell_para(self, cell):
        c = np.zeros(6)
        c[: len(cell)] = cell
        lib().orc_set_cell_para(self._h, _p(c))

    def set_contour_step(self, ds):
        lib().orc_set_contour_step(self._h, float(ds))

    def diffusion(self):
        lib().orc_diffusion(self._h)

    @property
    def space_size_phy(self):
        return int(lib().orc_get_space_size_phy(self._h))

    @property
    def grid_size_phy(self):
        out = (_sz * 3)()
        lib().orc_get_grid_size_phy(self._h, out)
        return tuple(int(v) for v in out)

    @property
    def cell_para(self):
        c = np.zeros(6)
        lib().orc_get_cell_para(self._h, _p(c))
        return c

    @property
    def ds(self):
        return float(lib().orc_get_ds(self._h))

    def iomat(self) -> np.ndarray:
        """A numpy view of the solver-owned in/out buffer (physical layout)."""
        n = self.space_size_phy
        ptr = lib().orc_get_iomat(self._h)
        return np.ctypeslib.as_array(ptr, shape=(n,)).reshape(self.grid_size_phy)

    def apply(self, field: np.ndarray) -> np.ndarray:
        io = self.iomat()
        io[...] = field.reshape(io.shape)
        self.diffusion()
        return io.copy()

    def propagate(self, E, init, Ns, direction):
        n = self.space_size_phy
        q = np.zeros((Ns + 1, n))
        E = np.ascontiguousarray(E, dtype=np.float64).reshape(-1)
        init = np.ascontiguousarray(init, dtype=np.float64).reshape(-1)
        lib().orc_propagate(self._h, _p(q), _p(init), _p(E), _sz(Ns), int(direction))
        return q

    def scft(self, fA, ds, xN, L, acceptance, nstep, wA, wB):
        """diblock_scft_simple_iterate; returns (trace[nstep,7], wA, wB, phiA, phiB)."""
        n = self.space_size_phy
        wA = np.ascontiguousarray(wA, dtype=np.float64).reshape(-1).copy()
        wB = np.ascontiguousarray(wB, dtype=np.float64).reshape(-1).copy()
        phiA = np.zeros(n)
        phiB = np.zeros(n)
        trace = np.zeros((nstep, TRACE_COLS))
        Lx, Ly, Lz = (L, L, L) if np.isscalar(L) else tuple(L)
        rc = lib().orc_scft_simple_iterate(self._h, fA, ds, xN, Lx, Ly, Lz, acceptance,
                                           _sz(nstep), _p(wA), _p(wB), _p(phiA), _p(phiB),
                                           _p(trace))
        if rc != 0:
            raise ValueError("orc_scft_simple_iterate rejected its arguments")
        sh = self.grid_size_phy
        return trace, wA.reshape(sh), wB.reshape(sh), phiA.reshape(sh), phiB.reshape(sh)
