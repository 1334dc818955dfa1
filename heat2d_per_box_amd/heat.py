"""Explicit heat solve through the box FGT, differentiable (tests and bench).

One step of the reaction-diffusion update used to exercise the adjoint:

    u^{n+1} = G_delta (u^n + dt (u^n - (u^n)^3)),   G_delta f = boxfgt(delta, f) / (pi delta),

with the callers' normalisation of the Gauss transform into the heat
semigroup at time delta/4 (unsteady_stokes2d_box_per_order4.f:483-484).
Each step is wrapped in ``torch.utils.checkpoint`` so the backward pass keeps
only u^n per step and recomputes the step's nonlinearity.
"""
from __future__ import annotations

import math

import torch
from torch.utils.checkpoint import checkpoint

from .autograd import GaussTransform

__all__ = ["heat_step", "heat_solve"]


def heat_step(gt: GaussTransform, u: torch.Tensor, delta: torch.Tensor, dt) -> torch.Tensor:
    v = u + dt * (u - u * u * u)
    return gt(v, delta) / (math.pi * delta.to(u.device))


def heat_solve(gt: GaussTransform, u0: torch.Tensor, delta: torch.Tensor, nsteps: int, dt=None,
               checkpoint_steps: bool = True) -> torch.Tensor:
    """u^nsteps from u^0; dt defaults to delta (the heat solve's time step)."""
    u = u0
    for _ in range(nsteps):
        step_dt = delta.to(u.device) if dt is None else dt
        if checkpoint_steps and (u.requires_grad or delta.requires_grad):
            u = checkpoint(heat_step, gt, u, delta, step_dt, use_reentrant=False)
        else:
            u = heat_step(gt, u, delta, step_dt)
    return u
