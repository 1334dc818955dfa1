"""The callers' time step around boxfgt, on a resident tree.

``VolPot`` is the volume-potential step of the heat solver (``volpot``,
hpots/hpots2dadap.f:178-381): with a time rule (t_j, w_j) on [0, dt] whose
last node is dt,

    vpot = sum_{j < m} dt w_j / (4 pi (dt - t_j)) boxfgt(4 (dt - t_j), f(., tpre + t_j))
           + dt w_m f(., tpre + dt),

the transform at delta_j = 4 (dt - t_j) normalised into the heat semigroup at
time dt - t_j.  Here every node keeps its own resident plan (the tree, lists
and FP64 tables are built once, keyed on delta_j), the forcing is evaluated
on the device at the leaf nodes returned by ``bfgt_plan_leaf_points``, and
each node's transform accumulates straight into one device array through
``bfgt_plan_execute_scaled`` -- no temporaries, no host round trips.

The reference's rules: trapezoid (nordert = 2: nodes 0, dt; weights 1/2,
1/2) and Simpson (nordert = 4: 0, dt/2, dt; 1/6, 4/6, 1/6)
(hpots/hpots2dadap.f:253-275).
"""
from __future__ import annotations

import math
from typing import Callable, Optional

import torch

from .boxfgt import BoxFGT

__all__ = ["VolPot", "time_rule"]


def time_rule(nordert: int, dt: float):
    """Nodes and weights of the reference's time quadrature on [0, dt]."""
    if nordert == 2:
        return [0.0, dt], [0.5, 0.5]
    if nordert == 4:
        return [0.0, 0.5 * dt, dt], [1.0 / 6.0, 4.0 / 6.0, 1.0 / 6.0]
    raise ValueError("nordert must be 2 (trapezoid) or 4 (Simpson)")


class VolPot:
    """Volume potential of one time step on a resident tree.

    ``fforce(points, t)`` returns the forcing at time t on the device, shape
    (nboxes, npbox, nd), from the leaf points (nboxes, npbox, 2); its values on
    non-leaf boxes are ignored.
    """

    def __init__(self, tree, norder: int, dt: float, nordert: int = 2, eps: float = 1e-9, ipoly: int = 1,
                 nd: int = 1, ifpgh: int = 1, iperiod: Optional[int] = None, precision: str = "fp64"):
        self.tree, self.norder, self.dt, self.nd, self.ifpgh = tree, norder, float(dt), nd, ifpgh
        self.tnodes, self.twhts = time_rule(nordert, self.dt)
        self.plans = []
        for tj, wj in zip(self.tnodes[:-1], self.twhts[:-1]):
            tau = self.dt - tj
            plan = BoxFGT(tree.copy(), nd, 4.0 * tau, eps, ipoly, norder, ifpgh, precision=precision,
                          iperiod=iperiod)
            self.plans.append((tj, self.dt * wj / (4.0 * math.pi * tau), plan))
        self.points = self.plans[0][2].leaf_points() if self.plans else None
        self.leaf = torch.from_numpy(tree.nchild == 0).to("cuda")

    def __call__(self, fforce: Callable, tpre: float, vpot: Optional[torch.Tensor] = None,
                 grad: Optional[torch.Tensor] = None, hess: Optional[torch.Tensor] = None) -> torch.Tensor:
        nb, npb = self.tree.centers.shape[0], self.norder ** 2
        if vpot is None:
            vpot = torch.zeros((nb, npb, self.nd), dtype=torch.float64, device="cuda")
        for tj, scale, plan in self.plans:
            f = fforce(self.points, tpre + tj).contiguous()
            plan(f, vpot, grad, hess, scale=scale)
        # last node (t = dt): the forcing itself, on leaf boxes
        f = fforce(self.points, tpre + self.tnodes[-1])
        vpot += self.dt * self.twhts[-1] * f * self.leaf[:, None, None]
        return vpot

    def close(self):
        for _, _, p in self.plans:
            p.close()
        self.plans = []
