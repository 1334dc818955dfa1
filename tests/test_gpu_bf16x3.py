"""Three-slice bf16 leaf transforms (BFGT_PREC_BF16X3): error against the
FP64 oracle.

Every operand of the two leaf passes is cut into three bf16 slices and each
product formed from the six slice products above 2^-24, accumulated in FP32
on v_mfma_f32_16x16x16_bf16; the near field runs split-FP32.  As for the
split-FP32 path the error is set by FP32 accumulation, near the FP32 unit
roundoff times the conditioning of the sums and independent of eps: the bar
is 2e-6 relative L2 over leaf points for pot and 1e-5 for grad/hess.  The
measured figures are printed so the per-config error can be read off the log.
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
    ("c2", "full", {}),
    ("c3", "full", {"delta": 1e-5, "eps": 1e-9}),
    ("c5", "full", {}),
]


def _errs(tree, got, ref, ifpgh):
    e = {"pot": leaf_rel_err(tree, got["pot"], ref.pot)}
    if ifpgh >= 2:
        e["grad"] = leaf_rel_err(tree, got["grad"], ref.grad)
    if ifpgh >= 3:
        e["hess"] = leaf_rel_err(tree, got["hess"], ref.hess)
    return e


@pytest.mark.parametrize("name,scale,kw", CASES, ids=lambda v: str(v))
def test_bf16x3_error_vs_oracle(device, name, scale, kw):
    w = make_workload(name, scale, **kw)
    ref = O.oracle_boxfgt(w.tree, w.fvals, w.nd, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, w.ifpgh)
    got = run_gpu(w.tree, w.fvals, w.nd, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, w.ifpgh,
                  precision="bf16x3")
    info = got["plan"].info
    assert info["precision"] == 2
    if info["n_pw_leaves"] > 0:
        # the column kernels share each x contraction over a column (group)
        assert info["xc_up"] <= info["n_pw_leaves"] and info["xc_down"] <= info["n_pw_leaves"]
        if scale == "full":
            assert info["xc_down"] < info["n_pw_leaves"]
    errs = _errs(w.tree, got, ref, w.ifpgh)
    print(f"bf16x3 {name}/{scale}/{kw}: npw={got['plan'].info['npw']} errors {errs}", flush=True)
    assert 0 < errs["pot"] < 2e-6, errs
    for k in ("grad", "hess"):
        if k in errs:
            assert errs[k] < 1e-5, errs


def test_bf16x3_on_vector_kernels_runs_fp64(device):
    """Leaf order 20 (vector kernels): the plan reports and runs FP64."""
    from tests.test_gpu_parity import _vector_tree

    t, fv = _vector_tree(20, 1, 2)
    a = run_gpu(t, fv, 2, 5e-4, 1e-9, 0, 1, 20, 2, precision="fp64")
    b = run_gpu(t, fv, 2, 5e-4, 1e-9, 0, 1, 20, 2, precision="bf16x3")
    assert b["plan"].info["precision"] == 0
    assert np.array_equal(a["pot"], b["pot"]) and np.array_equal(a["grad"], b["grad"])
