"""torch.autograd binding of the free energy, its field gradient and stress.

``free_energy(wA, wB, cell, run)`` returns H[wA, wB; cell] (see
``crys_scft_evaluate`` in include/crys_solver.h) as a differentiable torch
scalar.  The backward pass returns dH/dwA and dH/dwB from the forward and
backward propagators and dH/dL from the adjoint inner products with the L
derivative of the diffusion operator, all computed by libcrys_solver on the
GPU.

Tensors on a HIP device are passed by device address
(``crys_scft_evaluate_device``): the fields and gradients never leave HBM.
Tensors on the CPU go through ``crys_scft_evaluate``'s host staging.

A PmmmDCT solver without ``orthorhombic=True`` keeps the reference's cubic
cell (solver_Pmmm.h:101): H then depends on cell[0] alone, so the gradient of
``cell`` is (dH/dLx + dH/dLy + dH/dLz, 0, 0).
"""
from __future__ import annotations

import numpy as np

from .scft import SCFT


def _torch():
    import torch

    return torch


def _set_cell(run: SCFT, cell):
    want = run.solver.effective_cell(cell)
    if tuple(run.solver.get_cell_para()[:3]) != want or run.solver.get_ds() != run.ds:
        run.solver.set_cell_para(want)
        run.solver.set_contour_step(run.ds)


def _cell_grad(run: SCFT, dL):
    """dH/dcell from the per-axis derivatives of the applied cell."""
    if run.solver.cubic_only:
        return np.array([dL[0] + dL[1] + dL[2], 0.0, 0.0])
    return dL


def _function():
    torch = _torch()

    class FreeEnergy(torch.autograd.Function):
        @staticmethod
        def forward(ctx, wA, wB, cell, run):
            _set_cell(run, [float(v) for v in cell.detach().cpu().reshape(-1)[:3]])
            shape = run.shape
            stress = bool(cell.requires_grad)
            dev, dt = wA.device, wA.dtype
            if dev.type == "cuda" and wB.device == dev:
                # device path: contiguous float64 views, the library's stream
                # starts after the work that produced them
                a = wA.detach().to(torch.float64).contiguous()
                b = wB.detach().to(torch.float64).contiguous()
                if a.numel() != int(np.prod(shape)) or b.numel() != a.numel():
                    raise ValueError(f"fields need {int(np.prod(shape))} values")
                gA64 = torch.empty_like(a)
                gB64 = torch.empty_like(b)
                torch.cuda.current_stream(dev).synchronize()
                res = run.evaluate_device(a.data_ptr(), b.data_ptr(), gA64.data_ptr(),
                                          gB64.data_ptr(), stress=stress)
                gA = gA64.reshape(wA.shape).to(dt)
                gB = gB64.reshape(wB.shape).to(dt)
            else:
                a = wA.detach().cpu().numpy().astype(np.float64).reshape(shape)
                b = wB.detach().cpu().numpy().astype(np.float64).reshape(shape)
                res = run.evaluate(a, b, stress=stress)
                gA = torch.from_numpy(res["dH_dwA"]).reshape(wA.shape).to(dev, dt)
                gB = torch.from_numpy(res["dH_dwB"]).reshape(wB.shape).to(dev, dt)
            if res["dH_dL"] is not None:
                dL = torch.from_numpy(_cell_grad(run, res["dH_dL"])).to(cell.device, cell.dtype)
                dL = dL.reshape(cell.shape) if cell.numel() == 3 else dL
            else:
                dL = torch.zeros_like(cell)
            ctx.save_for_backward(gA, gB, dL)
            return torch.tensor(res["H"], dtype=dt, device=dev)

        @staticmethod
        def backward(ctx, gH):
            gA, gB, dL = ctx.saved_tensors
            return gH * gA, gH * gB, gH.to(dL.device) * dL, None

    return FreeEnergy


_FN = None


def free_energy(wA, wB, cell, run: SCFT):
    """Differentiable H[wA, wB; cell] of an SCFT handle (torch tensors)."""
    global _FN
    if _FN is None:
        _FN = _function()
    return _FN.apply(wA, wB, cell, run)
