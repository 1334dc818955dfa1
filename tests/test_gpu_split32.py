"""Split-FP32 leaf transforms: error against the FP64 oracle.

The split-FP32 path (BFGT_PREC_SPLIT32) runs the leaf-to-PW and PW-to-leaf
contractions as hi/lo FP32 MFMA products accumulated in FP32, with FP64
phases, tables, translation, near field and output accumulation.  FP32
accumulation bounds its error near the FP32 unit roundoff (6e-8) times the
conditioning of the sums, independent of eps, so the bar here is 2e-6
relative L2 over leaf points for pot and 1e-5 for grad/hess.  The measured
figures are printed so the per-config error can be read off the log.
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


def test_split32_differs_from_fp64_only_in_pw_stages(device):
    """In the Fourier-series regime with a single PW level the split path
    changes only the leaf transforms: the near field and translation stay
    FP64, so on a tree without PW leaves both precisions agree exactly."""
    from heat2d_per_box_amd.tree import build_uniform

    t = build_uniform(2)
    rng = np.random.default_rng(3)
    n = 8
    fv = rng.standard_normal((t.centers.shape[0], n * n, 1))
    fv[t.nchild != 0] = 0.0
    # delta small enough that every leaf is at or above the cutoff: near field only
    a = run_gpu(t, fv, 1, 1e-6, 1e-9, 0, 1, n, 2, precision="fp64")
    b = run_gpu(t, fv, 1, 1e-6, 1e-9, 0, 1, n, 2, precision="split32")
    assert a["plan"].info["n_pw_leaves"] == 0
    assert np.array_equal(a["pot"], b["pot"]) and np.array_equal(a["grad"], b["grad"])


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
