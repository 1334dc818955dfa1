"""The oracle's ifnewtree=1 path (boxfgt STEPS 8-10: refine the tree where
pot is not resolved, coarsen, 2:1 rebalance; bfgt/boxfgt.f:1172-1483)
against the reference boxfgt compiled from its sources (oracle/_ref).

The reference runs on one thread here: with several, its level-restriction
step (vol_tree_refine_boxes_flag_interp, common/tree_vol_coeffs_nd.f:3127,
an OpenMP loop whose gradient/Hessian coefficient work arrays are shared)
races, and the gradients and Hessians it interpolates onto the boxes that
step adds vary from run to run.  Cases are chosen so that no leaf of the
finest level asks to be refined (the reference would then write a level its
arrays do not hold); the oracle counts such leaves in ``ncapped``."""
import numpy as np
import pytest

from heat2d_per_box_amd.densities import gaussian_sum
from heat2d_per_box_amd.tree import build_adaptive, with_capacity
from oracle import oracle as O
from tests.treecmp import same_tree_arrays

pytestmark = pytest.mark.skipif(not O.reference_available(), reason="reference library not built")

# (seed, n, delta, iperiod, ifpgh, eps, max level, nd, ipoly)
CASES = [
    (1, 8, 1e-3, 0, 3, 1e-6, 5, 1, 0),     # free space, refinement only
    (3, 16, 1e-4, 1, 2, 1e-9, 5, 1, 0),    # direct-only transform, coarsening and rebalance
    (12, 8, 3e-3, 1, 3, 1e-9, 6, 2, 0),    # nd = 2, PW and near field, 516 new boxes
    (16, 8, 5e-4, 0, 2, 1e-9, 6, 1, 0),    # free space, 852 new boxes
    (11, 8, 1e-1, 0, 3, 1e-9, 5, 1, 0),    # free space at delta = 0.1 (cutoff at the root)
    (4, 8, 1e-1, 1, 3, 1e-9, 5, 1, 0),     # periodic Fourier regime
    (14, 12, 3e-4, 1, 2, 1e-9, 5, 1, 1),   # Chebyshev nodes
]


def _case(seed, n, delta, iperiod, ifpgh, eps, lev, nd, ipoly):
    g = gaussian_sum(np.random.default_rng(seed), 3, 0.02, 0.06)
    t = build_adaptive(g, 1, n, ipoly, 1e-6, lev, iperiod=iperiod, min_level=2)
    fv = t.sample(g, n, ipoly, 1)
    if nd == 2:
        fv = np.concatenate([fv, 0.5 * fv**2], axis=2)
    return with_capacity(t, 4 * t.nboxes + 400), fv


# random trees and parameters whose refinement stays above the finest level
RANDOM_SEEDS = [0, 3, 6, 7, 12, 14, 22, 23]


def _random_case(seed):
    """(capacity tree, fvals, nd, delta, eps, ipoly, iperiod, n, ifpgh) drawn
    from seed: leaf order 4..16, nd 1..2, every ifpgh, delta 1e-5..0.3, eps
    1e-6..1e-12, Legendre or Chebyshev, periodic or free space."""
    from heat2d_per_box_amd.densities import GaussianSum

    rng = np.random.default_rng(1000 + seed)
    iperiod = int(rng.integers(0, 2))
    ipoly = int(rng.integers(0, 2))
    n = int(rng.choice([4, 6, 8, 10, 12, 16]))
    nd = int(rng.integers(1, 3))
    ifpgh = int(rng.integers(1, 4))
    delta = float(10 ** rng.uniform(-5, -0.5))
    eps = float(rng.choice([1e-6, 1e-9, 1e-12]))
    m = int(rng.integers(1, 5))
    g = GaussianSum(rng.uniform(-0.5, 0.5, (m, 2)), rng.uniform(0.01, 0.08, m), rng.uniform(-1, 1, m))
    lev = int(rng.integers(4, 7))
    fun = lambda p: np.stack([g(p), np.cos(2 * p[..., 0]) * g(p)][:nd], axis=-1)  # noqa: E731
    t = build_adaptive(fun, nd, n, ipoly, 10 ** rng.uniform(-8, -4), lev, iperiod=iperiod,
                       min_level=int(rng.integers(1, 3)))
    fv = t.sample(fun, n, ipoly, nd)
    return with_capacity(t, 8 * t.nboxes + 200), fv, nd, delta, eps, ipoly, iperiod, n, ifpgh


@pytest.mark.parametrize("seed", RANDOM_SEEDS)
def test_newtree_random_trees_match_reference(seed):
    tc, fv, nd, delta, eps, ipoly, iperiod, n, ifpgh = _random_case(seed)
    ro = O.oracle_boxfgt_newtree(tc, fv, nd, delta, eps, ipoly, iperiod, n, ifpgh)
    rr = O.reference_boxfgt(tc, fv, nd, delta, eps, ipoly, iperiod, n, ifpgh, ifnewtree=1, nthreads=1)
    assert ro.info["ncapped"] == 0
    msgs = same_tree_arrays(ro.tree, rr.tree)
    assert not msgs, msgs
    nb = ro.tree.nboxes
    lv = ro.tree.nchild == 0
    for name in ("pot", "grad", "hess")[:ifpgh]:
        a, b = getattr(ro, name)[:nb][lv], getattr(rr, name)[:nb][lv]
        assert np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(b), name


@pytest.mark.parametrize("case", CASES, ids=[f"s{c[0]}-n{c[1]}-d{c[2]:g}-p{c[3]}-f{c[4]}-nd{c[7]}-q{c[8]}"
                                             for c in CASES])
def test_newtree_matches_reference(case):
    seed, n, delta, iperiod, ifpgh, eps, lev, nd, ipoly = case
    tc, fv = _case(*case)
    ro = O.oracle_boxfgt_newtree(tc, fv, nd, delta, eps, ipoly, iperiod, n, ifpgh)
    rr = O.reference_boxfgt(tc, fv, nd, delta, eps, ipoly, iperiod, n, ifpgh, ifnewtree=1, nthreads=1)
    assert ro.info["ncapped"] == 0
    msgs = same_tree_arrays(ro.tree, rr.tree)
    assert not msgs, msgs
    nb = ro.tree.nboxes
    lv = ro.tree.nchild == 0
    for name in ("pot", "grad", "hess")[:ifpgh]:
        a, b = getattr(ro, name)[:nb][lv], getattr(rr, name)[:nb][lv]
        err = np.linalg.norm(a - b) / np.linalg.norm(b)
        assert err < 1e-12, (name, err)


def test_newtree_without_refinement_is_the_plain_call():
    """A uniform tree (only finest-level leaves, which are never checked)
    comes back unchanged, with the plain call's values."""
    from heat2d_per_box_amd.tree import build_uniform

    g = gaussian_sum(np.random.default_rng(2), 3, 0.05, 0.1)
    t = build_uniform(3)
    fv = t.sample(g, 8, 0, 1)
    tc = with_capacity(t, 200)
    ro = O.oracle_boxfgt_newtree(tc, fv, 1, 1e-3, 1e-9, 0, 1, 8, 3)
    rp = O.oracle_boxfgt(t, fv, 1, 1e-3, 1e-9, 0, 1, 8, 3)
    assert ro.info["nnew"] == 0 and ro.tree.nboxes == t.nboxes
    for name in ("pot", "grad", "hess"):
        np.testing.assert_allclose(getattr(ro, name)[: t.nboxes], getattr(rp, name), rtol=0, atol=1e-13)
