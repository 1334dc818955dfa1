"""bfgt_plan_execute_noalloc: the evaluation with its workspace in the
library's static device array, called through the C ABI.

It runs the same kernels as bfgt_plan_execute on the same inputs, so its
outputs must equal that entry point's bit for bit (and therefore the
oracle's, to the parity bar).  Two calls in a row on one GPU share the static
array: the second, on a smaller plan, must not see what the first left in it.
A plan whose workspace exceeds the array is refused before any device work.
"""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from heat2d_per_box_amd.densities import make_workload
from tests.fields import leaf_rel_err

pytestmark = pytest.mark.gpu


def _noalloc_fn():
    from heat2d_per_box_amd._lib import lib

    return lib().bfgt_plan_execute_noalloc


def _run(w, noalloc, precision="fp64"):
    import torch
    from heat2d_per_box_amd._lib import check
    from heat2d_per_box_amd.boxfgt import BoxFGT

    plan = BoxFGT(w.tree.copy(), w.nd, w.delta, w.eps, w.ipoly, w.norder, w.ifpgh, precision=precision)
    nb, npb, nd = w.fvals.shape
    fv = torch.from_numpy(np.ascontiguousarray(w.fvals)).cuda()
    pot = torch.zeros((nb, npb, nd), dtype=torch.float64, device="cuda")
    grad = torch.zeros((nb, npb, 2, nd), dtype=torch.float64, device="cuda") if w.ifpgh >= 2 else None
    hess = torch.zeros((nb, npb, 3, nd), dtype=torch.float64, device="cuda") if w.ifpgh >= 3 else None
    ptr = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
    if noalloc:
        check(_noalloc_fn()(plan._h, ptr(fv), ptr(pot), ptr(grad), ptr(hess), None))
    else:
        plan(fv, pot, grad, hess)
    torch.cuda.synchronize()
    out = {"pot": pot.cpu().numpy(), "grad": grad.cpu().numpy() if grad is not None else None,
           "hess": hess.cpu().numpy() if hess is not None else None}
    plan.close()
    return out


# the first workload (nd = 3, 961 leaves, 37 MB of workspace) leaves its
# expansions in the static array; the second, smaller one (nd = 1, 406
# leaves, cutoff boxes that are leaves themselves and whose mp slot must read
# as zero) reuses it
CALLS = [("c3", "tiny", {"delta": 1e-4, "eps": 1e-9}), ("c2", "small", {})]


def test_noalloc_matches_execute_twice(device):
    for name, scale, kw in CALLS:
        w = make_workload(name, scale, **kw)
        print(f"noalloc case {name}/{scale}/{kw}: leaves={w.tree.nleaves}", flush=True)
        a = _run(w, noalloc=False)
        b = _run(w, noalloc=True)
        for k in ("pot", "grad", "hess"):
            if a[k] is not None:
                assert np.array_equal(a[k], b[k]), (name, k)
        ref = O.oracle_boxfgt(w.tree, w.fvals, w.nd, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, w.ifpgh)
        assert leaf_rel_err(w.tree, b["pot"], ref.pot) < 1e-12
        if w.ifpgh >= 2:
            assert leaf_rel_err(w.tree, b["grad"], ref.grad) < 1e-11
        if w.ifpgh >= 3:
            assert leaf_rel_err(w.tree, b["hess"], ref.hess) < 1e-11


def test_noalloc_refuses_plans_above_the_limit(device):
    """The full c3 tree at delta = 1e-5 (4,096 cutoff boxes, nd = 3) needs
    more than the 64 MiB static workspace: BFGT_ERR_RESOURCE, outputs untouched."""
    import torch
    from heat2d_per_box_amd.boxfgt import BoxFGT

    w = make_workload("c3", "full", delta=1e-5, eps=1e-9)
    plan = BoxFGT(w.tree.copy(), w.nd, w.delta, w.eps, w.ipoly, w.norder, w.ifpgh)
    nb, npb, nd = w.fvals.shape
    fv = torch.from_numpy(np.ascontiguousarray(w.fvals)).cuda()
    outs = [torch.zeros((nb, npb, c, nd) if c > 1 else (nb, npb, nd), dtype=torch.float64, device="cuda")
            for c in (1, 2, 3)]
    rc = _noalloc_fn()(plan._h, ctypes.c_void_p(fv.data_ptr()), *[ctypes.c_void_p(t.data_ptr()) for t in outs],
                       None)
    assert rc == 12
    torch.cuda.synchronize()
    assert all(float(t.abs().max()) == 0.0 for t in outs)
    plan.close()


@pytest.mark.parametrize("precision", ["split32", "bf16x3"])
def test_noalloc_other_arithmetics(device, precision):
    """Split-FP32 and three-slice bf16 plans (their own partial-sum and
    folded-expansion use of the workspace): the same two calls in a row, bit
    for bit against bfgt_plan_execute, and within those arithmetics' bar of
    the oracle."""
    for name, scale, kw in CALLS:
        w = make_workload(name, scale, **kw)
        print(f"noalloc {precision} case {name}/{scale}/{kw}: leaves={w.tree.nleaves}", flush=True)
        a = _run(w, noalloc=False, precision=precision)
        b = _run(w, noalloc=True, precision=precision)
        for k in ("pot", "grad", "hess"):
            if a[k] is not None:
                assert np.array_equal(a[k], b[k]), (name, k)
        ref = O.oracle_boxfgt(w.tree, w.fvals, w.nd, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, w.ifpgh)
        assert leaf_rel_err(w.tree, b["pot"], ref.pot) < 1e-6
