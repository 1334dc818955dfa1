"""Split-FP32 leaf transforms and near field: error against the FP64 oracle.

The split-FP32 path (BFGT_PREC_SPLIT32) runs the leaf-to-PW and PW-to-leaf
contractions and the near-field table contractions as hi/lo FP32 MFMA
products accumulated in FP32, with FP64 phases, tables, translation and
output accumulation.  FP32 accumulation bounds its error near the FP32 unit
roundoff (6e-8) times the conditioning of the sums, independent of eps, so
the bar here is 2e-6 relative L2 over leaf points for pot and 1e-5 for
grad/hess.  The measured figures are printed so the per-config error can be
read off the log.
"""
import numpy as np
import pytest

from oracle import oracle as O
from heat2d_per_box_amd.densities import make_workload
from tests.fields import leaf_rel_err
from tests.test_gpu_parity import run_gpu

pytestmark = pytest.mark.gpu

CASES = [
    ("c1", "full", {}),
    ("c2", "small", {}),
    ("c3", "tiny", {"delta": 1e-5, "eps": 1e-6}),
    ("c3", "tiny", {"delta": 1e-3, "eps": 1e-9}),
    ("c3", "tiny", {"delta": 1e-1, "eps": 1e-9}),
    ("c4", "small", {}),
    ("c5", "small", {}),
]


@pytest.mark.parametrize("name,scale,kw", CASES, ids=lambda v: str(v))
def test_split32_error_vs_oracle(device, name, scale, kw):
    w = make_workload(name, scale, **kw)
    ref = O.oracle_boxfgt(w.tree, w.fvals, w.nd, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, w.ifpgh)
    got = run_gpu(w.tree, w.fvals, w.nd, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, w.ifpgh,
                  precision="split32")
    assert got["plan"].info["precision"] == 1
    errs = {"pot": leaf_rel_err(w.tree, got["pot"], ref.pot)}
    if w.ifpgh >= 2:
        errs["grad"] = leaf_rel_err(w.tree, got["grad"], ref.grad)
    if w.ifpgh >= 3:
        errs["hess"] = leaf_rel_err(w.tree, got["hess"], ref.hess)
    print(f"split32 {name}/{scale}/{kw}: npw={got['plan'].info['npw']} errors {errs}", flush=True)
    assert errs["pot"] < 2e-6, errs
    for k in ("grad", "hess"):
        if k in errs:
            assert errs[k] < 1e-5, errs


@pytest.mark.parametrize("n,nd,ifpgh", [(8, 1, 2), (8, 3, 3), (4, 2, 2), (6, 4, 1), (5, 3, 3), (16, 2, 3),
                                         (12, 1, 1)])
def test_split32_near_field_only(device, n, nd, ifpgh):
    """A tree whose leaves are all near-field targets (delta = 1e-6): the
    split path's near field alone, against the oracle, with the bars above
    (n <= 8 with two or more densities packs two densities per tile, as the
    FP64 path does, odd nd leaving the last tile half used; otherwise one
    wave per target and density)."""
    from heat2d_per_box_amd.tree import build_uniform

    t = build_uniform(2)
    rng = np.random.default_rng(3 + n + nd)
    fv = rng.standard_normal((t.centers.shape[0], n * n, nd))
    fv[t.nchild != 0] = 0.0
    ref = O.oracle_boxfgt(t, fv, nd, 1e-6, 1e-9, 0, 1, n, ifpgh)
    got = run_gpu(t, fv, nd, 1e-6, 1e-9, 0, 1, n, ifpgh, precision="split32")
    assert got["plan"].info["n_pw_leaves"] == 0 and got["plan"].info["n_near_pairs"] > 0
    errs = {"pot": leaf_rel_err(t, got["pot"], ref.pot)}
    if ifpgh >= 2:
        errs["grad"] = leaf_rel_err(t, got["grad"], ref.grad)
    if ifpgh >= 3:
        errs["hess"] = leaf_rel_err(t, got["hess"], ref.hess)
    print(f"split32 near field n={n} nd={nd} ifpgh={ifpgh}: errors {errs}", flush=True)
    assert 0 < errs["pot"] < 2e-6, errs
    for k in ("grad", "hess"):
        if k in errs:
            assert errs[k] < 1e-5, errs


FULL = [
    ("c2", {}),
    ("c3", {"delta": 1e-5, "eps": 1e-6}),
    ("c3", {"delta": 1e-3, "eps": 1e-9}),
    ("c3", {"delta": 1e-1, "eps": 1e-9}),
    ("c5", {}),
]


@pytest.mark.parametrize("name,kw", FULL, ids=lambda v: str(v))
def test_split32_full_size(device, name, kw):
    """The split kernels at the workload sizes the bench runs: c2 (19,834
    leaves, n = 16), the c3 sweep legs on the 6,709-leaf tree (n = 8, nd = 3,
    pot + grad + hess; delta = 1e-1 is the Fourier regime with one cutoff box)
    and c5 (104,317 leaves, 16-leaf chunks).  Same bars as above."""
    w = make_workload(name, "full", **kw)
    ref = O.oracle_boxfgt(w.tree, w.fvals, w.nd, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, w.ifpgh)
    got = run_gpu(w.tree, w.fvals, w.nd, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, w.ifpgh,
                  precision="split32")
    assert got["plan"].info["precision"] == 1
    errs = {"pot": leaf_rel_err(w.tree, got["pot"], ref.pot)}
    if w.ifpgh >= 2:
        errs["grad"] = leaf_rel_err(w.tree, got["grad"], ref.grad)
    if w.ifpgh >= 3:
        errs["hess"] = leaf_rel_err(w.tree, got["hess"], ref.hess)
    print(f"split32 {name}/full/{kw}: leaves={w.tree.nleaves} npw={got['plan'].info['npw']} errors {errs}",
          flush=True)
    assert errs["pot"] < 2e-6, errs
    for k in ("grad", "hess"):
        if k in errs:
            assert errs[k] < 1e-5, errs


def test_split32_on_vector_kernels_runs_fp64(device):
    """Leaf orders above 16 have no split kernels: the plan runs FP64, says so
    in its precision field, and matches the FP64 plan bit for bit."""
    from tests.test_gpu_parity import _vector_tree

    t, fv = _vector_tree(20, 1, 2)
    a = run_gpu(t, fv, 2, 5e-4, 1e-9, 0, 1, 20, 2, precision="fp64")
    b = run_gpu(t, fv, 2, 5e-4, 1e-9, 0, 1, 20, 2, precision="split32")
    assert b["plan"].info["precision"] == 0 and not b["plan"].info["mfma"]
    assert np.array_equal(a["pot"], b["pot"]) and np.array_equal(a["grad"], b["grad"])
