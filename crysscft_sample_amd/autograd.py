"""torch.autograd binding of the free energy, its field gradient and stress.

``free_energy(wA, wB, cell, run)`` returns H[wA, wB; cell] (see
``crys_scft_evaluate`` in include/crys_solver.h) as a differentiable torch
scalar.  The backward pass returns dH/dwA and dH/dwB from the forward and
backward propagators and dH/dL from the adjoint inner products with the L
derivative of the diffusion operator, all computed by libcrys_solver on the
GPU.  Tensors cross the boundary through host memory: the library keeps its
own device buffers, so the binding works with tensors on any device.
"""
from __future__ import annotations

import numpy as np

from .scft import SCFT


def _torch():
    import torch

    return torch


def _set_cell(run: SCFT, cell):
    c = [float(v) for v in cell]
    if tuple(run.solver.get_cell_para()[:3]) != tuple(c) or run.solver.get_ds() != run.ds:
        run.solver.set_cell_para(c)
        run.solver.set_contour_step(run.ds)


def _function():
    torch = _torch()

    class FreeEnergy(torch.autograd.Function):
        @staticmethod
        def forward(ctx, wA, wB, cell, run):
            _set_cell(run, cell.detach().cpu().numpy())
            shape = run.shape
            a = wA.detach().cpu().numpy().astype(np.float64).reshape(shape)
            b = wB.detach().cpu().numpy().astype(np.float64).reshape(shape)
            res = run.evaluate(a, b, stress=bool(cell.requires_grad))
            dev, dt = wA.device, wA.dtype
            gA = torch.from_numpy(res["dH_dwA"]).reshape(wA.shape).to(dev, dt)
            gB = torch.from_numpy(res["dH_dwB"]).reshape(wB.shape).to(dev, dt)
            dL = (torch.from_numpy(res["dH_dL"]).to(cell.device, cell.dtype)
                  if res["dH_dL"] is not None else torch.zeros_like(cell))
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
