"""SCFT iteration of an AB diblock melt on the device, over the C ABI.

:func:`diblock_scft_simple_iterate` mirrors the reference's function of the
same name (crysSCFT_sample.cpp:426-531): truncated NsA/NsB (:443-444), cell
and contour step set on the solver (:466-467), a warm-up diffusion (:471),
then ``Nstep`` simple-mixing iterations (:476-508).  Every iteration runs on
the GPU; the energies of every iteration come back as a trace (the reference
prints them at ``itr % 10 == 0`` and at the last iteration).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, dptr, lib
from .solver import CrysPSSolverBase


@dataclass
class SCFTResult:
    trace: np.ndarray  # (Nstep, 7): F, enthalpic, -lnQ, entropic, ksi, max_incomp, Q
    phiA: np.ndarray
    phiB: np.ndarray
    wA: np.ndarray
    wB: np.ndarray
    NsA: int
    NsB: int

    def lines(self, every: int = 10):
        """The reference's print_energy lines (crysSCFT_sample.cpp:419-423)."""
        out = []
        n = len(self.trace)
        for itr, r in enumerate(self.trace):
            if itr % every == 0 or itr == n - 1:
                out.append(f"iter {itr}: FreeEnergy = {r[0]:.10g} ( {r[1]:.10g} + {r[2]:.10g} + "
                           f"{r[3]:.10g} + {r[4]:.10g} )  , MaxIncompressibility = {r[5]:.10g}")
        return out


class SCFT:
    """Device-resident SCFT state bound to one solver (crys_scft_t)."""

    def __init__(self, solver: CrysPSSolverBase, fA, ds, xN, L, acceptance):
        Lx, Ly, Lz = (L, L, L) if np.isscalar(L) else tuple(float(v) for v in L)
        h = ctypes.c_void_p()
        check(lib().crys_scft_create(solver.handle, float(fA), float(ds), float(xN), Lx, Ly, Lz,
                                     float(acceptance), ctypes.byref(h)))
        self._h = h
        self.solver = solver
        self.fA, self.ds, self.xN = float(fA), float(ds), float(xN)
        self.shape = solver.get_grid_size_phy()
        a, b = ctypes.c_size_t(), ctypes.c_size_t()
        check(lib().crys_scft_get_contour(h, ctypes.byref(a), ctypes.byref(b)))
        self.NsA, self.NsB = int(a.value), int(b.value)

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().crys_scft_destroy(h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_checkpointing(self, every: int = 0, budget_bytes: int = 0):
        """Contour checkpointing (crys_scft_set_checkpointing): keep the
        forward chain every ``every`` slots and recompute between them; with
        every=0, full histories when they fit ``budget_bytes`` (0: free device
        memory), else every = ceil(sqrt(Ns))."""
        check(lib().crys_scft_set_checkpointing(self._h, int(every), int(budget_bytes)))

    def checkpointing(self):
        """(interval in use, 0 for full histories; contour storage in bytes)."""
        e = ctypes.c_size_t()
        b = ctypes.c_size_t()
        check(lib().crys_scft_get_checkpointing(self._h, ctypes.byref(e), ctypes.byref(b)))
        return e.value, b.value

    def set_fields(self, wA, wB):
        a = np.ascontiguousarray(wA, dtype=np.float64).reshape(-1)
        b = np.ascontiguousarray(wB, dtype=np.float64).reshape(-1)
        n = int(np.prod(self.shape))
        if a.size != n or b.size != n:
            raise ValueError(f"fields need {n} values")
        check(lib().crys_scft_set_fields(self._h, dptr(a), dptr(b)))

    def run(self, nstep: int, trace: bool = True) -> np.ndarray | None:
        out = np.zeros((nstep, _lib.TRACE_COLS)) if trace else None
        check(lib().crys_scft_run(self._h, int(nstep), dptr(out) if trace else None))
        return out

    def fields(self):
        n = self.shape
        phiA, phiB, wA, wB = (np.empty(n) for _ in range(4))
        check(lib().crys_scft_get_fields(self._h, dptr(phiA), dptr(phiB), dptr(wA), dptr(wB)))
        return phiA, phiB, wA, wB

    def evaluate(self, wA=None, wB=None, stress=True):
        """H[wA, wB], dH/dwA, dH/dwB and the cell stress dH/dL (crys_scft_evaluate).

        Returns a dict with ``H``, ``dH_dwA``, ``dH_dwB`` (physical shape),
        ``dH_dL`` (3,) or None, and ``trace_row`` (the print_energy terms)."""
        shape = self.shape
        a = b = None
        if wA is not None:
            a = np.ascontiguousarray(wA, dtype=np.float64).reshape(-1)
            b = np.ascontiguousarray(wB, dtype=np.float64).reshape(-1)
        H = ctypes.c_double()
        gA = np.empty(shape)
        gB = np.empty(shape)
        dL = np.empty(3) if stress else None
        row = np.empty(_lib.TRACE_COLS)
        check(lib().crys_scft_evaluate(self._h, None if a is None else dptr(a),
                                       None if b is None else dptr(b), ctypes.byref(H), dptr(gA),
                                       dptr(gB), None if dL is None else dptr(dL), dptr(row)))
        return {"H": H.value, "dH_dwA": gA, "dH_dwB": gB, "dH_dL": dL, "trace_row": row}

    def evaluate_device(self, wA_ptr, wB_ptr, gA_ptr, gB_ptr, stress=True):
        """crys_scft_evaluate_device: fields and gradients are device addresses
        (ints) of space_size_phy doubles; the inputs must be ready.  Returns a
        dict with ``H``, ``dH_dL`` (3,) or None and ``trace_row``."""
        H = ctypes.c_double()
        dL = np.empty(3) if stress else None
        row = np.empty(_lib.TRACE_COLS)
        check(lib().crys_scft_evaluate_device(self._h, wA_ptr, wB_ptr, ctypes.byref(H), gA_ptr,
                                              gB_ptr, None if dL is None else dptr(dL),
                                              dptr(row)))
        return {"H": H.value, "dH_dL": dL, "trace_row": row}

    def time_propagation(self, reps: int = 5):
        """(ms per contour step of both chains, launches per step) over the
        iteration's own launch sequence, captured in a hipGraph."""
        a = ctypes.c_double()
        nl = ctypes.c_int()
        check(lib().crys_scft_time_propagation(self._h, int(reps), ctypes.byref(a),
                                               ctypes.byref(nl)))
        return a.value, nl.value

    def time_step(self, reps: int = 50):
        """(ms per contour step of both chains, [ms per launch of each kernel
        of the step, in issue order]), timed with HIP events on the stream."""
        a = ctypes.c_double()
        per = (ctypes.c_double * 3)()
        nk = ctypes.c_int()
        check(lib().crys_scft_time_step(self._h, int(reps), ctypes.byref(a), per,
                                        ctypes.byref(nk)))
        return a.value, [per[k] for k in range(nk.value)]


def diblock_scft_simple_iterate(Nx, Ny, Nz, fA, ds, xN, Lx, Ly, Lz, acceptance, Nstep,
                                solvername, solver: CrysPSSolverBase, init_func,
                                verbose=False) -> SCFTResult:
    """Mirror of crysSCFT_sample.cpp:426-531.

    ``init_func(wA, wB)`` fills two arrays of the solver's physical shape,
    as the reference's init_func does (:463).  Nx, Ny, Nz are the logical
    grid and must match the solver's."""
    if tuple(solver.get_grid_size_log()) != (int(Nx), int(Ny), int(Nz)):
        raise ValueError("logical grid does not match the solver")
    shape = solver.get_grid_size_phy()
    wA = np.zeros(shape)
    wB = np.zeros(shape)
    init_func(wA, wB)
    run = SCFT(solver, fA, ds, xN, (Lx, Ly, Lz), acceptance)
    try:
        if verbose:
            print("*" * 66)
            print(f"Warm up the solver {solvername}...")
            print(f"Start solving with solver {solvername}...")
        run.set_fields(wA, wB)
        tr = run.run(int(Nstep))
        phiA, phiB, wA1, wB1 = run.fields()
        res = SCFTResult(tr, phiA, phiB, wA1, wB1, run.NsA, run.NsB)
        if verbose:
            for line in res.lines():
                print(line)
        return res
    finally:
        run.close()
