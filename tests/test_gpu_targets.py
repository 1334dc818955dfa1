"""GPU extra-target evaluation (ifpghtarg = 1..3) against the oracle.

* The host-array drop-in with targets vs the oracle's boxfgt with targets
  (and vs the compiled reference where it is built), including
  ifpghtarg > ifpgh, where the extra channels go to scratch arrays and the
  caller's grad/hess stay untouched (bfgt/boxfgt.f:207-277).
* The device entry bfgt_plan_eval_targets on the same leaf arrays as the
  oracle's interpolation, and exactness on polynomial leaf data.
FP64 vs FP64 on the same interpolant: bar 1e-12 relative L2.
"""
import numpy as np
import pytest

from oracle import oracle as O
from heat2d_per_box_amd.densities import GaussianSum, make_workload
from heat2d_per_box_amd.tree import build_adaptive

pytestmark = pytest.mark.gpu


def _targets(rng, nt):
    tg = rng.uniform(-0.5, 0.5, (nt, 2))
    # box corners and edges (ties go to the plus side), the domain corner
    tg[:6] = [[0.0, 0.0], [0.25, -0.25], [-0.5, -0.5], [0.125, 0.3], [-0.25, 0.0], [0.0, 0.4375]]
    return tg


def _host(w, ifpgh, ifpt, tg, pot0):
    from heat2d_per_box_amd.boxfgt import boxfgt

    t = w.tree.copy()
    nb, npb, nd = w.fvals.shape
    pot = pot0.copy()
    # arrays the transform does not own (ifpgh < 2, < 3) are passed as sentinels that must stay untouched
    grad = np.zeros((nb, npb, 2, nd)) if ifpgh >= 2 else np.full((nb, npb, 2, nd), 7.0)
    hess = np.zeros((nb, npb, 3, nd)) if ifpgh >= 3 else np.full((nb, npb, 3, nd), 7.0)
    nt = tg.shape[0]
    pote = np.zeros((nt, nd))
    grade = np.zeros((nt, 2, nd))
    hesse = np.zeros((nt, 3, nd))
    ti = np.zeros(20)
    boxfgt(nd, 2, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, npb, nb, t.nlevels, t.ltree, t.itree, t.iptr,
           t.centers, t.boxsize, w.fvals, ifpgh, pot, grad, hess, 0, nt, np.ascontiguousarray(tg), ifpt, pote,
           grade, hesse, ti)
    return pot, grad, hess, pote, grade, hesse, ti


CASES = [
    ("c1", "full", {}, 1, 1),
    ("c3", "tiny", {"delta": 1e-3, "eps": 1e-9}, 3, 3),
    ("c3", "tiny", {"delta": 1e-5, "eps": 1e-6}, 1, 3),
    ("c3", "tiny", {"delta": 1e-1, "eps": 1e-9}, 2, 3),
    ("c4", "small", {}, 1, 2),
    ("c2", "small", {}, 2, 2),
]


@pytest.mark.parametrize("name,scale,kw,ifpgh,ifpt", CASES, ids=lambda v: str(v))
def test_host_targets_match_oracle(device, name, scale, kw, ifpgh, ifpt):
    w = make_workload(name, scale, **kw)
    rng = np.random.default_rng(8)
    tg = _targets(rng, 700)
    pot0 = 0.01 * rng.standard_normal(w.fvals.shape)
    print(f"case {name}/{scale}/{kw} ifpgh={ifpgh} ifpghtarg={ifpt}: leaves={w.tree.nleaves}", flush=True)
    pot, grad, hess, pote, grade, hesse, ti = _host(w, ifpgh, ifpt, tg, pot0)
    ref = O.oracle_boxfgt(w.tree, w.fvals, w.nd, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, ifpgh, targs=tg,
                          ifpghtarg=ifpt, pot0=pot0)
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)  # noqa: E731
    assert rel(pote, ref.pote) < 1e-12
    if ifpt >= 2:
        assert rel(grade, ref.grade) < 1e-11
    if ifpt >= 3:
        assert rel(hesse, ref.hesse) < 1e-11
    lv = w.tree.nchild == 0
    assert rel(pot[lv], ref.pot[lv]) < 1e-12
    if ifpgh < 2:
        assert np.all(grad == 7.0)
    if ifpgh < 3:
        assert np.all(hess == 7.0)
    assert ti[11] > 0.0


@pytest.mark.skipif(not O.reference_available(), reason="reference library not built")
@pytest.mark.parametrize("name,scale,kw,ifpgh,ifpt", CASES[:4], ids=lambda v: str(v))
def test_host_targets_match_reference(device, name, scale, kw, ifpgh, ifpt):
    w = make_workload(name, scale, **kw)
    rng = np.random.default_rng(9)
    tg = _targets(rng, 500)
    pot0 = 0.01 * rng.standard_normal(w.fvals.shape)
    pot, grad, hess, pote, grade, hesse, _ = _host(w, ifpgh, ifpt, tg, pot0)
    ref = O.reference_boxfgt(w.tree, w.fvals, w.nd, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, ifpgh,
                             targs=tg, ifpghtarg=ifpt, pot0=pot0)
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)  # noqa: E731
    assert rel(pote, ref.pote) < 1e-12
    if ifpt >= 2:
        assert rel(grade, ref.grade) < 1e-11
    if ifpt >= 3:
        assert rel(hesse, ref.hesse) < 1e-11


@pytest.mark.parametrize("ipoly,n", [(0, 4), (0, 16), (1, 8), (0, 24)])
def test_device_targets_polynomial_exact_and_match_oracle(device, ipoly, n):
    import torch
    from heat2d_per_box_amd.boxfgt import BoxFGT

    g = GaussianSum(np.array([[0.1, -0.2]]), np.array([0.05]), np.array([1.0]))
    t = build_adaptive(lambda p: g(p)[..., None], 1, min(n, 8), ipoly, 1e-4, 4, iperiod=1, min_level=1, eta=0.0)

    def poly(p):
        x, y = p[..., 0], p[..., 1]
        return np.stack([(1 + 0.3 * x - 0.7 * y + 0.5 * x * y) * (1 + 0.2 * x ** (n - 2)) + y ** (n - 1),
                         np.cos(3 * x) * y], axis=-1)

    nd = 2
    vals = t.sample(poly, n, ipoly, nd)
    grad = np.ascontiguousarray(np.stack([vals, -vals], axis=2))
    hess = np.ascontiguousarray(np.stack([vals, 2 * vals, 3 * vals], axis=2))
    plan = BoxFGT(t.copy(), nd, 1e-3, 1e-6, ipoly, n, 3)
    rng = np.random.default_rng(3)
    tg = _targets(rng, 5000)
    cu = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    pote = torch.zeros((tg.shape[0], nd), dtype=torch.float64, device="cuda")
    grade = torch.zeros((tg.shape[0], 2, nd), dtype=torch.float64, device="cuda")
    hesse = torch.zeros((tg.shape[0], 3, nd), dtype=torch.float64, device="cuda")
    plan.eval_targets(cu(tg), 3, cu(vals), pote, cu(grad), cu(hess), grade, hesse)
    torch.cuda.synchronize()
    pe, ge, he = pote.cpu().numpy(), grade.cpu().numpy(), hesse.cpu().numpy()
    want = poly(tg)[:, 0]
    assert np.max(np.abs(pe[:, 0] - want)) < 1e-11 * np.max(np.abs(want))
    rp, rg, rh = O.oracle_eval_targets(t, t.itree, nd, ipoly, n, tg, 3, vals, grad, hess)
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)  # noqa: E731
    assert rel(pe, rp) < 1e-13 and rel(ge, rg) < 1e-13 and rel(he, rh) < 1e-13


@pytest.mark.parametrize("case", [1, 2])
def test_host_targets_on_kept_plan(device, case):
    """The same host call with extra targets twice: the second runs on the kept
    plan (target arrays of its own; with ifpghtarg > ifpgh the scratch channels
    zeroed again) and gives the first call's outputs bit for bit."""
    from heat2d_per_box_amd.boxfgt import release_plans

    name, scale, kw, ifpgh, ifpt = CASES[case]
    w = make_workload(name, scale, **kw)
    rng = np.random.default_rng(9)
    tg = _targets(rng, 500)
    pot0 = 0.01 * rng.standard_normal(w.fvals.shape)
    release_plans()
    first = _host(w, ifpgh, ifpt, tg, pot0)
    second = _host(w, ifpgh, ifpt, tg, pot0)
    release_plans()
    for a, b in zip(first[:6], second[:6]):
        assert np.array_equal(a, b)
