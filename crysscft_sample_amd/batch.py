"""Many independent SCFT instances on one GPU (crys_batch_* of the C ABI).

Each instance is one ``diblock_scft_simple_iterate`` run
(crysSCFT_sample.cpp:426-531) with its own grid, chemistry, cell and contour;
all of them advance together, one launch per axis pass and contour step.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, dptr, lib

_KIND = {"P1": _lib.P1, "3m": _lib.M3, "PmmmDCT": _lib.PMMM_DCT}


class CInstance(ctypes.Structure):
    """Layout of crys_instance_t."""

    _fields_ = [("kind", ctypes.c_int), ("flags", ctypes.c_int),
                ("grid_log", ctypes.c_size_t * 3), ("cell", ctypes.c_double * 3),
                ("fA", ctypes.c_double), ("xN", ctypes.c_double), ("ds", ctypes.c_double),
                ("acceptance", ctypes.c_double)]


@dataclass
class Instance:
    solver: str
    grid_log: tuple
    cell: tuple
    fA: float
    xN: float
    ds: float
    acceptance: float = 0.1
    orthorhombic: bool = False

    def to_c(self) -> CInstance:
        c = CInstance()
        c.kind = _KIND[self.solver]
        c.flags = _lib.FLAG_PMMM_ORTHORHOMBIC if self.orthorhombic else 0
        for a in range(3):
            c.grid_log[a] = int(self.grid_log[a])
            c.cell[a] = float(self.cell[a])
        c.fA, c.xN, c.ds, c.acceptance = self.fA, self.xN, self.ds, self.acceptance
        return c


class Batch:
    def __init__(self, instances, precision="fp32"):
        self.instances = list(instances)
        arr = (CInstance * len(self.instances))(*[i.to_c() for i in self.instances])
        prec = {"fp64": _lib.FP64, "fp32": _lib.FP32}[precision]
        h = ctypes.c_void_p()
        check(lib().crys_batch_create(arr, len(self.instances), prec, ctypes.byref(h)))
        self._h = h
        self.info = []
        for k in range(len(self.instances)):
            n, a, b = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_size_t()
            check(lib().crys_batch_get_info(h, k, ctypes.byref(n), ctypes.byref(a),
                                            ctypes.byref(b)))
            self.info.append((int(n.value), int(a.value), int(b.value)))

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().crys_batch_destroy(h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape(self, k):
        inst = self.instances[k]
        red = 1 if inst.solver == "P1" else 2
        return tuple(g // red for g in inst.grid_log)

    def set_fields(self, k, wA, wB):
        a = np.ascontiguousarray(wA, dtype=np.float64).reshape(-1)
        b = np.ascontiguousarray(wB, dtype=np.float64).reshape(-1)
        check(lib().crys_batch_set_fields(self._h, k, dptr(a), dptr(b)))

    def run(self, nstep):
        out = np.zeros((len(self.instances), nstep, _lib.TRACE_COLS))
        check(lib().crys_batch_run(self._h, int(nstep), dptr(out)))
        return out

    PATHS = {"auto": 0, "phases": 1, "passes": 2}

    def set_path(self, path: str):
        """'auto', 'phases' (one launch per contour step) or 'passes' (three)."""
        check(lib().crys_batch_set_path(self._h, self.PATHS[path]))

    def path(self) -> str:
        p = ctypes.c_int()
        check(lib().crys_batch_get_path(self._h, ctypes.byref(p)))
        return {1: "phases", 2: "passes"}[p.value]

    def autotune(self, reps: int = 3):
        """Times the contour steps on each possible path, keeps the faster.
        Returns (path, ms of the phases or None, ms of the passes)."""
        p = ctypes.c_int()
        ms = (ctypes.c_double * 2)()
        check(lib().crys_batch_autotune(self._h, int(reps), ctypes.byref(p), ms))
        return ({1: "phases", 2: "passes"}[p.value], ms[0] if ms[0] >= 0 else None, ms[1])

    def set_phase_grid(self, workgroups: int):
        """0: one workgroup per phase item; G > 0: a persistent grid of G
        workgroups that prefetch the next item's plane into LDS."""
        check(lib().crys_batch_set_phase_grid(self._h, int(workgroups)))

    def phase_grid(self) -> int:
        g = ctypes.c_int()
        check(lib().crys_batch_get_phase_grid(self._h, ctypes.byref(g)))
        return g.value

    def fields(self, k):
        sh = self.shape(k)
        phiA, phiB, wA, wB = (np.empty(sh) for _ in range(4))
        check(lib().crys_batch_get_fields(self._h, k, dptr(phiA), dptr(phiB), dptr(wA),
                                          dptr(wB)))
        return phiA, phiB, wA, wB


def random_instances(count, seed=7, sizes=(16, 24, 32, 48, 64), ns=(100, 150, 200)):
    """Batched workload: reduced grids drawn from ``sizes`` (cubed), fA ~ U[0.15, 0.35],
    xN ~ U[15, 40], L ~ U[3.6, 4.6], Ns from ``ns``; PmmmDCT."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        m = int(rng.choice(sizes))
        Ns = int(rng.choice(ns))
        L = float(rng.uniform(3.6, 4.6))
        out.append(Instance("PmmmDCT", (2 * m,) * 3, (L,) * 3, float(rng.uniform(0.15, 0.35)),
                            float(rng.uniform(15, 40)), 1.0 / Ns, 0.1))
    return out
