"""Overwriting execution (bfgt_plan_execute_overwrite) against the
accumulating one.

pot (+grad, hess) = scale * A fvals on the leaf boxes, whatever the arrays
held, must equal scale * A fvals accumulated into zeroed arrays bit for bit
(both add the same FP64 terms in the same order onto 0.0), and entries of
non-leaf boxes must keep what they held.  Covered: PW-only, mixed and
near-field-dominated trees (leaves above the cutoff level, which the near
field writes instead of adding to), a plan with no PW leaves, periodic and
free space, n <= 8 (2 x 2 leaf blocks, leaf pairs), n = 16 (column groups),
the vector kernels (n = 20), split-FP32, one to three densities, and the
timed c2 and c5 workloads at full size.
"""
import numpy as np
import pytest

from heat2d_per_box_amd.densities import make_workload
from heat2d_per_box_amd.tree import build_uniform

pytestmark = pytest.mark.gpu


def _both(tree, fvals, nd, delta, eps, ipoly, iperiod, n, ifpgh, precision="fp64", scale=1.0):
    import torch
    from heat2d_per_box_amd.boxfgt import BoxFGT

    plan = BoxFGT(tree.copy(), nd, delta, eps, ipoly, n, ifpgh, precision=precision, iperiod=iperiod)
    nb, npb = fvals.shape[0], fvals.shape[1]
    fv = torch.from_numpy(np.ascontiguousarray(fvals)).cuda()
    shapes = {"pot": (nb, npb, nd), "grad": (nb, npb, 2, nd), "hess": (nb, npb, 3, nd)}
    keys = ["pot", "grad", "hess"][:ifpgh]
    zero = {k: torch.zeros(shapes[k], dtype=torch.float64, device="cuda") for k in keys}
    gen = torch.Generator(device="cuda").manual_seed(3)
    held = {k: torch.randn(shapes[k], dtype=torch.float64, device="cuda", generator=gen) for k in keys}
    before = {k: v.cpu().numpy() for k, v in held.items()}
    plan(fv, zero["pot"], zero.get("grad"), zero.get("hess"), scale=scale)
    # twice: the second call must not see the first one's outputs
    for _ in range(2):
        plan(fv, held["pot"], held.get("grad"), held.get("hess"), scale=scale, overwrite=True)
    torch.cuda.synchronize()
    info = plan.info
    plan.close()
    return ({k: v.cpu().numpy() for k, v in zero.items()}, {k: v.cpu().numpy() for k, v in held.items()},
            before, info)


def _check(tree, acc, ovw, before):
    leaf = tree.nchild == 0
    for k in acc:
        assert np.array_equal(ovw[k][leaf], acc[k][leaf]), k
        assert np.array_equal(ovw[k][~leaf], before[k][~leaf]), k


CASES = [
    ("c1", "full", {}, "fp64"),
    ("c2", "small", {}, "fp64"),
    ("c2", "small", {}, "split32"),
    ("c3", "tiny", {"delta": 1e-5, "eps": 1e-9}, "fp64"),  # near-field dominated
    ("c3", "tiny", {"delta": 1e-3, "eps": 1e-6}, "fp64"),
    ("c3", "tiny", {"delta": 1e-1, "eps": 1e-9}, "fp64"),  # Fourier regime
    ("c3", "tiny", {"delta": 1e-5, "eps": 1e-9}, "split32"),
    ("c4", "small", {}, "fp64"),
    ("c5", "small", {}, "fp64"),
]


@pytest.mark.parametrize("name,scale_kind,kw,precision", CASES, ids=lambda v: str(v))
def test_overwrite_equals_accumulate_on_zeros(device, name, scale_kind, kw, precision):
    w = make_workload(name, scale_kind, **kw)
    acc, ovw, before, info = _both(w.tree, w.fvals, w.nd, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, w.ifpgh,
                                   precision=precision, scale=0.75)
    print(f"{name}/{scale_kind} {kw} {precision}: pw leaves {info['n_pw_leaves']}, near pairs "
          f"{info['n_near_pairs']}", flush=True)
    _check(w.tree, acc, ovw, before)


@pytest.mark.parametrize("name,scale_kind", [("c2", "full"), ("c5", "full"), ("c2", "mixed"), ("c5", "mixed")])
def test_overwrite_full_size(device, name, scale_kind):
    """The timed workloads at full size (c2 is the bench's default), and the
    headline shapes with leaves above the cutoff level, which the near field
    writes instead of adding to."""
    w = make_workload(name, scale_kind)
    acc, ovw, before, info = _both(w.tree, w.fvals, w.nd, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, w.ifpgh)
    if scale_kind == "mixed":
        assert info["n_near_pairs"] > 0 and info["n_pw_leaves"] < w.tree.nleaves
    _check(w.tree, acc, ovw, before)


def test_overwrite_without_pw_leaves(device):
    """delta small enough that every leaf is a near-field target: the near
    field writes every leaf (nd = 2, pot + grad + hess)."""
    t = build_uniform(3, iperiod=1)
    rng = np.random.default_rng(4)
    fv = rng.standard_normal((t.centers.shape[0], 64, 2))
    acc, ovw, before, info = _both(t, fv, 2, 1e-6, 1e-9, 0, 1, 8, 3)
    assert not info["has_pw"]
    _check(t, acc, ovw, before)


@pytest.mark.parametrize("iperiod", [0, 1])
def test_overwrite_vector_kernels(device, iperiod):
    """Leaf order 20: the vector-FMA evaluation overwrites its leaves."""
    from tests.test_gpu_parity import _vector_tree

    t, fv = _vector_tree(20, iperiod, 2)
    acc, ovw, before, info = _both(t, fv, 2, 5e-4, 1e-9, 0, iperiod, 20, 3)
    assert info["has_pw"] and not info["mfma"] and info["n_pw_leaves"] > 0
    _check(t, acc, ovw, before)


@pytest.mark.parametrize("n", [4, 8])
def test_overwrite_leaf_pairs_and_quads(device, n):
    """n <= 8 on a small tree (leaf pairs) and on the uniform level-6 tree
    (2 x 2 leaf blocks), nd = 3."""
    for lev in (3, 6):
        t = build_uniform(lev, iperiod=1)
        rng = np.random.default_rng(lev)
        fv = rng.standard_normal((t.centers.shape[0], n * n, 3))
        acc, ovw, before, info = _both(t, fv, 3, 1e-3, 1e-9, 0, 1, n, 2)
        assert info["has_pw"] and info["mfma"]
        _check(t, acc, ovw, before)
