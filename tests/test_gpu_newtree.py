"""boxfgt with ifnewtree=1 on the GPU (the host-array drop-in,
bfgt_hip_boxfgt): the tree is refined where pot is not resolved, coarsened
and 2:1-rebalanced (bfgt/boxfgt.f:1172-1483).  Checked box for box against
the oracle's restatement (same numbering, colleague lists and centres; values
to 1e-12) and against the reference compiled from its sources, run on one
thread (tests/test_oracle_newtree.py explains why)."""
import numpy as np
import pytest

from heat2d_per_box_amd.boxfgt import BoxFGTError, boxfgt
from heat2d_per_box_amd.densities import gaussian_sum
from heat2d_per_box_amd.tree import build_adaptive, build_uniform, with_capacity
from oracle import oracle as O
from tests.test_oracle_newtree import CASES, RANDOM_SEEDS, _case, _random_case
from tests.treecmp import same_tree_arrays

pytestmark = pytest.mark.gpu


def _run_gpu(tc, fv, nd, delta, eps, ipoly, iperiod, n, ifpgh, targs=None, ifpghtarg=0, precision="fp64"):
    """The drop-in on copies of the capacity-laid-out tree; returns the new
    tree, pot/grad/hess (capacity), timeinfo and the target arrays."""
    t = tc.copy()
    nbmax = t.centers.shape[0]
    npbox = n * n
    fvals = np.zeros((nbmax, npbox, nd))
    fvals[: fv.shape[0]] = fv
    pot = np.zeros((nbmax, npbox, nd))
    grad = np.zeros((nbmax, npbox, 2, nd))
    hess = np.zeros((nbmax, npbox, 3, nd))
    ti = np.zeros(20)
    nt = 0 if targs is None else targs.shape[0]
    pote = np.zeros((max(nt, 1), nd))
    grade = np.zeros((max(nt, 1), 2, nd))
    hesse = np.zeros((max(nt, 1), 3, nd))
    tg = np.zeros((1, 2)) if targs is None else np.ascontiguousarray(targs)
    nl = boxfgt(nd, 2, delta, eps, ipoly, iperiod, n, npbox, nbmax, t.nlevels, t.ltree, t.itree, t.iptr, t.centers,
                t.boxsize, fvals, ifpgh, pot, grad, hess, ifnewtree=1, ntarg=nt, targs=tg, ifpghtarg=ifpghtarg,
                pote=pote, grade=grade, hesse=hesse, timeinfo=ti, precision=precision)
    t.nlevels = nl
    return t, pot, grad, hess, ti, (pote[:nt], grade[:nt], hesse[:nt])


def _rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.mark.parametrize("case", CASES, ids=[f"s{c[0]}-n{c[1]}-d{c[2]:g}-p{c[3]}-f{c[4]}-nd{c[7]}-q{c[8]}"
                                             for c in CASES])
def test_newtree_matches_oracle_and_reference(device, case):
    seed, n, delta, iperiod, ifpgh, eps, lev, nd, ipoly = case
    print(f"[newtree] case {case}", flush=True)
    tc, fv = _case(*case)
    t, pot, grad, hess, ti, _ = _run_gpu(tc, fv, nd, delta, eps, ipoly, iperiod, n, ifpgh)
    ro = O.oracle_boxfgt_newtree(tc, fv, nd, delta, eps, ipoly, iperiod, n, ifpgh)
    assert ti[18] == ro.info["nnew"] and ti[19] == 0
    msgs = same_tree_arrays(t, ro.tree)
    assert not msgs, msgs
    nb = t.nboxes
    lv = t.nchild == 0
    for name, g in (("pot", pot), ("grad", grad), ("hess", hess))[:ifpgh]:
        err = _rel(g[:nb][lv], getattr(ro, name)[:nb][lv])
        print(f"  {name} vs oracle {err:.2e}", flush=True)
        assert err < 1e-12, (name, err)
    if O.reference_available():
        rr = O.reference_boxfgt(tc, fv, nd, delta, eps, ipoly, iperiod, n, ifpgh, ifnewtree=1, nthreads=1)
        assert not same_tree_arrays(t, rr.tree)
        for name, g in (("pot", pot), ("grad", grad), ("hess", hess))[:ifpgh]:
            assert _rel(g[:nb][lv], getattr(rr, name)[:nb][lv]) < 1e-12, name


@pytest.mark.parametrize("seed", RANDOM_SEEDS)
def test_newtree_random_trees(device, seed):
    tc, fv, nd, delta, eps, ipoly, iperiod, n, ifpgh = _random_case(seed)
    print(f"[newtree] random seed {seed}: n={n} nd={nd} ifpgh={ifpgh} delta={delta:.3g} eps={eps:g} "
          f"iperiod={iperiod} ipoly={ipoly} boxes={tc.nboxes}", flush=True)
    t, pot, grad, hess, ti, _ = _run_gpu(tc, fv, nd, delta, eps, ipoly, iperiod, n, ifpgh)
    ro = O.oracle_boxfgt_newtree(tc, fv, nd, delta, eps, ipoly, iperiod, n, ifpgh)
    assert ti[18] == ro.info["nnew"]
    msgs = same_tree_arrays(t, ro.tree)
    assert not msgs, msgs
    nb = t.nboxes
    lv = t.nchild == 0
    for name, g in (("pot", pot), ("grad", grad), ("hess", hess))[:ifpgh]:
        assert _rel(g[:nb][lv], getattr(ro, name)[:nb][lv]) < 1e-12, name


def test_newtree_c3_workload(device):
    """The nd = 3 spot workload on its 4,017-box tree at delta = 1e-3
    (pot + grad + hess): 320 boxes added, tree and values as the oracle's."""
    from heat2d_per_box_amd.densities import make_workload

    w = make_workload("c3", "small", delta=1e-3, eps=1e-9)
    tc = with_capacity(w.tree, 6 * w.tree.nboxes)
    t, pot, grad, hess, ti, _ = _run_gpu(tc, w.fvals, w.nd, w.delta, w.eps, 0, 1, w.norder, w.ifpgh)
    ro = O.oracle_boxfgt_newtree(tc, w.fvals, w.nd, w.delta, w.eps, 0, 1, w.norder, w.ifpgh)
    assert ro.info["nnew"] > 0 and ti[18] == ro.info["nnew"]
    msgs = same_tree_arrays(t, ro.tree)
    assert not msgs, msgs
    nb = t.nboxes
    lv = t.nchild == 0
    for name, g in (("pot", pot), ("grad", grad), ("hess", hess)):
        assert _rel(g[:nb][lv], getattr(ro, name)[:nb][lv]) < 1e-12, name


def test_newtree_targets_in_the_new_tree(device):
    """Extra targets after the refinement are interpolated in the new tree
    (bfgtmain STEP 11 runs after STEPS 8-10)."""
    case = CASES[2]
    seed, n, delta, iperiod, ifpgh, eps, lev, nd, ipoly = case
    tc, fv = _case(*case)
    rng = np.random.default_rng(5)
    tg = rng.uniform(-0.5, 0.5, size=(500, 2))
    t, pot, grad, hess, ti, (pe, ge, he) = _run_gpu(tc, fv, nd, delta, eps, ipoly, iperiod, n, ifpgh, targs=tg,
                                                     ifpghtarg=3)
    ro = O.oracle_boxfgt_newtree(tc, fv, nd, delta, eps, ipoly, iperiod, n, ifpgh)
    rp, rg, rh = O.oracle_eval_targets(ro.tree, ro.tree.itree, nd, ipoly, n, tg, 3, ro.pot, ro.grad, ro.hess)
    assert _rel(pe, rp) < 1e-12 and _rel(ge, rg) < 1e-12 and _rel(he, rh) < 1e-12


def test_newtree_unrefined_tree_is_the_plain_call(device):
    """A uniform tree (leaves only at the finest level, never checked) comes
    back unchanged with the plain call's values."""
    g = gaussian_sum(np.random.default_rng(2), 3, 0.05, 0.1)
    t0 = build_uniform(3)
    fv = t0.sample(g, 8, 0, 1)
    t, pot, grad, hess, ti, _ = _run_gpu(with_capacity(t0, 200), fv, 1, 1e-3, 1e-9, 0, 1, 8, 3)
    rp = O.oracle_boxfgt(t0, fv, 1, 1e-3, 1e-9, 0, 1, 8, 3)
    assert ti[18] == 0 and t.nboxes == t0.nboxes and not same_tree_arrays(t, t0)
    for g_, r in ((pot, rp.pot), (grad, rp.grad), (hess, rp.hess)):
        assert _rel(g_[: t0.nboxes], r) < 1e-12


def test_newtree_capacity_exceeded(device):
    """A tree that would outgrow its capacity is refused with
    BFGT_ERR_RESOURCE (the reference writes past its arrays)."""
    case = CASES[3]
    seed, n, delta, iperiod, ifpgh, eps, lev, nd, ipoly = case
    tc, fv = _case(*case)
    small = with_capacity(build_adaptive(gaussian_sum(np.random.default_rng(seed), 3, 0.02, 0.06), 1, n, ipoly,
                                         1e-6, lev, iperiod=iperiod, min_level=2), tc.nboxes + 8)
    with pytest.raises(BoxFGTError) as e:
        _run_gpu(small, fv, nd, delta, eps, ipoly, iperiod, n, ifpgh)
    assert e.value.code == 12


def test_newtree_at_the_finest_level_is_not_split(device):
    """Leaves of the finest level that ask to be split stay leaves (counted
    in timeinfo(20)); the oracle does the same."""
    case = (6, 8, 1e-4, 1, 3, 1e-9, 6, 1, 0)
    tc, fv = _case(*case)
    seed, n, delta, iperiod, ifpgh, eps, lev, nd, ipoly = case
    t, pot, grad, hess, ti, _ = _run_gpu(tc, fv, nd, delta, eps, ipoly, iperiod, n, ifpgh)
    ro = O.oracle_boxfgt_newtree(tc, fv, nd, delta, eps, ipoly, iperiod, n, ifpgh)
    assert ro.info["ncapped"] > 0 and ti[19] == ro.info["ncapped"]
    assert not same_tree_arrays(t, ro.tree)
    nb = t.nboxes
    lv = t.nchild == 0
    assert _rel(pot[:nb][lv], ro.pot[:nb][lv]) < 1e-12
