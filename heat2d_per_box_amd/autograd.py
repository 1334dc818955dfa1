"""Differentiable box FGT: a ``torch.autograd.Function`` over the HIP library.

The forward is ``bfgt_plan_execute`` (the drop-in transform); the backward is
``bfgt_plan_adjoint``, the hand-written transpose of every stage, through the
same C ABI.  The derivative with respect to delta uses the heat-kernel
identity for G(x) = exp(-|x|^2/delta):

    dG/ddelta = |x|^2/delta^2 G = (Laplacian G)/4 + G/delta,

so d pot/d delta = (hxx + hyy)/4 + pot/delta, with the Hessian taken from the
same transform (ifpgh = 3) on the same tree.  Only the potential output can
carry a delta cotangent (its delta derivative needs no third derivatives).

Plans are resident per (delta, ifpgh) on one tree, because the solvers call
the transform 2-3 times per time step on one tree with one delta
(hpots/hpots2dadap.f:291-381).  At most ``max_plans`` of them are kept (least
recently used closed first), so that a loop over many deltas -- an
optimisation in delta, finite-difference probes -- holds a bounded amount of
device memory.
"""
from __future__ import annotations

from collections import OrderedDict
from typing import Optional, Tuple

import torch

from .boxfgt import BoxFGT

__all__ = ["GaussTransform", "boxfgt_apply"]


class GaussTransform:
    """pot (and grad/hess) = boxfgt(delta, f) on one tree, differentiable in f and delta.

    ``tree`` is a :class:`heat2d_per_box_amd.tree.QuadTree`; leaf values are
    tensors of shape (nboxes, npbox, nd) on the GPU, float64.
    """

    def __init__(self, tree, nd: int, eps: float, ipoly: int, norder: int, ifpgh: int = 1,
                 iperiod: Optional[int] = None, max_plans: int = 4):
        if ifpgh not in (1, 2, 3):
            raise ValueError("ifpgh must be 1, 2 or 3")
        if max_plans < 2:
            raise ValueError("max_plans must be at least 2 (a forward and a backward plan of one delta)")
        self.tree, self.nd, self.eps, self.ipoly, self.norder, self.ifpgh = tree, nd, eps, ipoly, norder, ifpgh
        self.iperiod = tree.iperiod if iperiod is None else iperiod
        self.max_plans = int(max_plans)
        self._plans: "OrderedDict[Tuple[float, int], BoxFGT]" = OrderedDict()

    def plan(self, delta: float, ifpgh: int) -> BoxFGT:
        key = (float(delta), int(ifpgh))
        p = self._plans.get(key)
        if p is not None:
            self._plans.move_to_end(key)
            return p
        while len(self._plans) >= self.max_plans:
            _, old = self._plans.popitem(last=False)
            # the plan's stream may still run work on device arrays the
            # caller owns; closing frees only the plan's own buffers
            old.stream.synchronize()
            old.close()
        p = BoxFGT(self.tree.copy(), self.nd, float(delta), self.eps, self.ipoly, self.norder, ifpgh,
                   iperiod=self.iperiod)
        self._plans[key] = p
        return p

    @property
    def n_plans(self) -> int:
        return len(self._plans)

    def clear(self) -> None:
        for p in self._plans.values():
            p.close()
        self._plans.clear()

    def __call__(self, f: torch.Tensor, delta):
        """Returns pot (ifpgh=1) or the tuple (pot, grad[, hess])."""
        if not torch.is_tensor(delta):
            delta = torch.tensor(float(delta), dtype=torch.float64)
        out = _BoxFGTFn.apply(f, delta, self)
        return out[0] if self.ifpgh == 1 else out


def boxfgt_apply(gt: GaussTransform, f: torch.Tensor, delta):
    return gt(f, delta)


class _BoxFGTFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f, delta, gt: GaussTransform):
        d = float(delta)
        nb, npb, nd = f.shape
        fc = f.detach().contiguous()
        need_dd = bool(delta.requires_grad)
        ifpgh = 3 if need_dd else gt.ifpgh
        plan = gt.plan(d, ifpgh)
        pot = torch.zeros((nb, npb, nd), dtype=torch.float64, device=f.device)
        grad = torch.zeros((nb, npb, 2, nd), dtype=torch.float64, device=f.device) if ifpgh >= 2 else None
        hess = torch.zeros((nb, npb, 3, nd), dtype=torch.float64, device=f.device) if ifpgh >= 3 else None
        plan(fc, pot, grad, hess)
        dpot_dd = None
        if need_dd:
            dpot_dd = 0.25 * (hess[:, :, 0, :] + hess[:, :, 2, :]) + pot / d
        ctx.gt, ctx.delta, ctx.ddev = gt, d, delta.device
        ctx.save_for_backward(dpot_dd if dpot_dd is not None else torch.empty(0))
        ctx.need_dd = need_dd
        outs = [pot]
        if gt.ifpgh >= 2:
            outs.append(grad)
        if gt.ifpgh >= 3:
            outs.append(hess)
        return tuple(outs)

    @staticmethod
    def backward(ctx, *gouts):
        gt, d = ctx.gt, ctx.delta
        (dpot_dd,) = ctx.saved_tensors
        gpot = gouts[0]
        ref = gpot if gpot is not None else next(g for g in gouts if g is not None)
        nb, npb, nd = ref.shape[0], ref.shape[1], ref.shape[-1]
        dev = ref.device

        def dense(g, shape):
            return torch.zeros(shape, dtype=torch.float64, device=dev) if g is None else g.contiguous()

        gp = dense(gpot, (nb, npb, nd))
        gg = dense(gouts[1], (nb, npb, 2, nd)) if gt.ifpgh >= 2 else None
        gh = dense(gouts[2], (nb, npb, 3, nd)) if gt.ifpgh >= 3 else None
        fbar = None
        if ctx.needs_input_grad[0]:
            fbar = torch.zeros((nb, npb, nd), dtype=torch.float64, device=dev)
            gt.plan(d, gt.ifpgh).adjoint(gp, fbar, gg, gh)
        dbar = None
        if ctx.needs_input_grad[1]:
            if (gg is not None and gouts[1] is not None and bool(gg.abs().max() > 0)) or \
               (gh is not None and gouts[2] is not None and bool(gh.abs().max() > 0)):
                raise NotImplementedError("delta derivative of grad/hess outputs needs third derivatives")
            dbar = (gp * dpot_dd).sum().reshape(()).to(ctx.ddev)
        return fbar, dbar, None
