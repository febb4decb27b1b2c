"""Python mirror of the reference's solver plugin API over the C ABI.

``CrysPSSolverBase`` follows solver_base.h:8-51; ``SolverP1``, ``Solver3m`` and
``SolverPmmmDCT`` (aliases ``CrysPSSolverP1``, ``CrysPSSolver3m``,
``CrysPSSolverPmmm``) follow the constructors of solver_p1.h:16,
solver_3m.h:17 and solver_Pmmm.h:16.  The solver-owned ``iomat`` lives in
device memory; ``upload_iomat`` / ``download_iomat`` stage it explicitly.
Errors the reference answers with ``exit(1)`` raise :class:`CrysError`.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import CrysError, check, dptr, lib

_PRECISION = {"fp64": _lib.FP64, "fp32": _lib.FP32, "f64": _lib.FP64, "f32": _lib.FP32}


class CrysPSSolverBase:
    """Device-resident solver handle (one HIP stream per handle)."""

    kind: int = -1
    name: str = "CrysPSSolverBase"

    def __init__(self, grid_size_log, translational_part=None, precision="fp64",
                 orthorhombic=False):
        if self.kind < 0:
            raise TypeError("instantiate SolverP1, Solver3m or SolverPmmmDCT")
        grid = (ctypes.c_size_t * 3)(*[int(v) for v in grid_size_log])
        g = None
        if translational_part is not None:
            g = np.ascontiguousarray(translational_part, dtype=np.float64)
            if g.size != 9:
                raise ValueError("translational_part needs 9 values")
        prec = _PRECISION[precision] if isinstance(precision, str) else int(precision)
        flags = _lib.FLAG_PMMM_ORTHORHOMBIC if orthorhombic else 0
        h = ctypes.c_void_p()
        check(lib().crys_solver_create(self.kind, grid, None if g is None else dptr(g), prec, flags,
                                       ctypes.byref(h)))
        self._h = h
        self.precision = prec
        # PmmmDCT keeps only cell_para[0] unless orthorhombic (solver_Pmmm.h:101)
        self.cubic_only = self.kind == _lib.PMMM_DCT and not orthorhombic

    # -- lifetime --------------------------------------------------------
    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().crys_solver_destroy(h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self) -> ctypes.c_void_p:
        if self._h is None:
            raise CrysError(_lib.ERR_INVALID_ARGUMENT, "solver is closed")
        return self._h

    # -- virtual API (solver_base.h:16-20) -------------------------------
    def set_cell_para(self, cell_para):
        c = np.zeros(6)
        c[: len(cell_para)] = cell_para
        check(lib().crys_set_cell_para(self.handle, dptr(c)))

    def set_contour_step(self, ds: float):
        check(lib().crys_set_contour_step(self.handle, float(ds)))

    def diffusion(self):
        check(lib().crys_diffusion(self.handle))

    def time_diffusion(self, reps: int = 200) -> float:
        """Mean milliseconds per diffusion() call, timed on the device with
        HIP events over `reps` back-to-back calls (changes the iomat)."""
        ms = ctypes.c_double()
        check(lib().crys_diffusion_time(self.handle, int(reps), ctypes.byref(ms)))
        return ms.value

    # -- getters (solver_base.h:22-28) ----------------------------------
    def get_cell_para(self):
        c = np.zeros(6)
        check(lib().crys_get_cell_para(self.handle, dptr(c)))
        return tuple(c)

    def get_grid_size_log(self):
        out = (ctypes.c_size_t * 3)()
        check(lib().crys_get_grid_size_log(self.handle, out))
        return tuple(int(v) for v in out)

    def get_grid_size_phy(self):
        out = (ctypes.c_size_t * 3)()
        check(lib().crys_get_grid_size_phy(self.handle, out))
        return tuple(int(v) for v in out)

    def get_space_size_log(self) -> int:
        v = ctypes.c_size_t()
        check(lib().crys_get_space_size_log(self.handle, ctypes.byref(v)))
        return int(v.value)

    def get_space_size_phy(self) -> int:
        v = ctypes.c_size_t()
        check(lib().crys_get_space_size_phy(self.handle, ctypes.byref(v)))
        return int(v.value)

    def get_ds(self) -> float:
        v = ctypes.c_double()
        check(lib().crys_get_ds(self.handle, ctypes.byref(v)))
        return float(v.value)

    def effective_cell(self, cell_para):
        """The cell lengths the solver applies for `cell_para`: (L0, L0, L0)
        for a PmmmDCT solver that keeps the reference's cubic quirk."""
        c = tuple(float(v) for v in cell_para[:3])
        return (c[0],) * 3 if self.cubic_only else c

    def get_iomat_for_diffusion(self) -> int:
        """Device address of the solver-owned in/out buffer."""
        p = ctypes.c_void_p()
        check(lib().crys_get_iomat_for_diffusion(self.handle, ctypes.byref(p)))
        return int(p.value or 0)

    # -- host staging ----------------------------------------------------
    def upload_iomat(self, host: np.ndarray):
        a = np.ascontiguousarray(host, dtype=np.float64).reshape(-1)
        if a.size != self.get_space_size_phy():
            raise ValueError(f"iomat needs {self.get_space_size_phy()} values, got {a.size}")
        check(lib().crys_iomat_upload(self.handle, dptr(a)))

    def download_iomat(self) -> np.ndarray:
        out = np.empty(self.get_grid_size_phy(), dtype=np.float64)
        check(lib().crys_iomat_download(self.handle, dptr(out)))
        return out

    def synchronize(self):
        check(lib().crys_solver_synchronize(self.handle))

    def apply(self, field: np.ndarray) -> np.ndarray:
        """Upload, diffusion(), download: one triplet on the device."""
        self.upload_iomat(field)
        self.diffusion()
        return self.download_iomat()


class SolverP1(CrysPSSolverBase):
    """CrysPSSolverP1 (solver_p1.h:14): full grid, no symmetry."""

    kind = _lib.P1
    name = "SolverNormal"


class Solver3m(CrysPSSolverBase):
    """CrysPSSolver3m (solver_3m.h:14): 2x2y2z even sub-lattice, 1/8 size."""

    kind = _lib.M3
    name = "Solver3m"

    def __init__(self, grid_size_log, translational_part=(0,) * 9, precision="fp64"):
        super().__init__(grid_size_log, translational_part, precision)


class SolverPmmmDCT(CrysPSSolverBase):
    """CrysPSSolverPmmm (solver_Pmmm.h:13): DCT-II/III on the first octant.

    Cubic-only like the reference (cell_para[0] for every axis,
    solver_Pmmm.h:101) unless ``orthorhombic=True``.
    """

    kind = _lib.PMMM_DCT
    name = "SolverPmmmDCT"

    def __init__(self, grid_size_log, precision="fp64", orthorhombic=False):
        super().__init__(grid_size_log, None, precision, orthorhombic)


CrysPSSolverP1 = SolverP1
CrysPSSolver3m = Solver3m
CrysPSSolverPmmm = SolverPmmmDCT

SOLVERS = {"P1": SolverP1, "3m": Solver3m, "PmmmDCT": SolverPmmmDCT}


def axis_operator(kind: int, n_logical: int, L: float, ds: float) -> np.ndarray:
    """The per-axis factor D of a solver's diffusion() (host computed).

    Returns D with out = D @ v along that axis (shape m x m)."""
    m = n_logical if kind == _lib.P1 else n_logical // 2
    out = np.empty(m * m)
    check(lib().crys_axis_operator(int(kind), int(n_logical), float(L), float(ds), dptr(out)))
    return out.reshape(m, m).T.copy()
