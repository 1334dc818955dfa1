"""GPU parity of the HIP box FGT against the oracle (and the compiled reference).

Every case runs the HIP library through its C ABI (ctypes) on device arrays
and compares with the CPU oracle on the same tree and density.  Both compute
in FP64 and follow the same algorithm, so the bar is a relative L2 difference
over leaf points of 1e-12 (1e-11 for gradients/Hessians), far below eps.
"""
import numpy as np
import pytest

from oracle import oracle as O
from heat2d_per_box_amd.densities import make_workload, GaussianSum
from heat2d_per_box_amd.tree import build_uniform, build_adaptive
from tests.fields import FourierMode, leaf_rel_err

pytestmark = pytest.mark.gpu

TOL = {"pot": 1e-12, "grad": 1e-11, "hess": 1e-11}


def run_gpu(tree, fvals, nd, delta, eps, ipoly, iperiod, norder, ifpgh, precision="fp64", repeat=1):
    import torch
    from heat2d_per_box_amd.boxfgt import BoxFGT

    t = tree.copy()
    plan = BoxFGT(t, nd, delta, eps, ipoly, norder, ifpgh, precision=precision, iperiod=iperiod)
    fv = torch.from_numpy(np.ascontiguousarray(fvals)).cuda()
    nb, npb = fvals.shape[0], fvals.shape[1]
    pot = torch.zeros((nb, npb, nd), dtype=torch.float64, device="cuda")
    grad = torch.zeros((nb, npb, 2, nd), dtype=torch.float64, device="cuda") if ifpgh >= 2 else None
    hess = torch.zeros((nb, npb, 3, nd), dtype=torch.float64, device="cuda") if ifpgh >= 3 else None
    for _ in range(repeat):
        plan(fv, pot, grad, hess)
    torch.cuda.synchronize()
    out = {
        "pot": pot.cpu().numpy(),
        "grad": grad.cpu().numpy() if grad is not None else None,
        "hess": hess.cpu().numpy() if hess is not None else None,
        "itree": t.itree,
        "plan": plan,
    }
    return out


def compare(tree, got, ref, ifpgh, scale=1.0):
    errs = {"pot": leaf_rel_err(tree, got["pot"], scale * ref.pot)}
    if ifpgh >= 2:
        errs["grad"] = leaf_rel_err(tree, got["grad"], scale * ref.grad)
    if ifpgh >= 3:
        errs["hess"] = leaf_rel_err(tree, got["hess"], scale * ref.hess)
    return errs


def assert_close(errs, tol=TOL, mult=1.0):
    for k, v in errs.items():
        assert v < tol[k] * mult, errs


WORKLOADS = [
    ("c1", "full", {}),
    ("c2", "small", {}),
    ("c3", "tiny", {"delta": 1e-5, "eps": 1e-6}),
    ("c3", "tiny", {"delta": 1e-5, "eps": 1e-9}),
    ("c3", "tiny", {"delta": 1e-3, "eps": 1e-6}),
    ("c3", "tiny", {"delta": 1e-3, "eps": 1e-9}),
    ("c3", "tiny", {"delta": 1e-1, "eps": 1e-6}),
    ("c3", "tiny", {"delta": 1e-1, "eps": 1e-9}),
    ("c4", "small", {}),
    ("c5", "small", {}),
]


@pytest.mark.parametrize("name,scale,kw", WORKLOADS, ids=lambda v: str(v))
def test_workload_matches_oracle(device, name, scale, kw):
    w = make_workload(name, scale, **kw)
    print(f"case {name}/{scale}/{kw}: leaves={w.tree.nleaves}", flush=True)
    ref = O.oracle_boxfgt(w.tree, w.fvals, w.nd, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, w.ifpgh)
    got = run_gpu(w.tree, w.fvals, w.nd, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, w.ifpgh)
    info = got["plan"].info
    assert info["npwlevel"] == ref.info["npwlevel"] and info["npw"] == ref.info["npw"]
    assert np.array_equal(got["itree"], ref.itree)
    assert_close(compare(w.tree, got, ref, w.ifpgh))


@pytest.mark.parametrize("delta", [1e-5, 1e-3])
def test_full_size_c3_matches_oracle(device, delta):
    """The 6,709-leaf n = 8 workload at full size: the downward pass runs as
    2 x 2 leaf blocks over column pairs (fewer x contractions than leaves);
    at delta = 1e-5 the translation runs in groups of four sibling cutoff
    boxes (2,676 boxes), each distinct neighbour expansion read once."""
    w = make_workload("c3", "full", delta=delta, eps=1e-9)
    print(f"case c3/full delta={delta}: leaves={w.tree.nleaves}", flush=True)
    ref = O.oracle_boxfgt(w.tree, w.fvals, w.nd, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, w.ifpgh)
    got = run_gpu(w.tree, w.fvals, w.nd, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, w.ifpgh)
    info = got["plan"].info
    assert info["xc_down"] < info["n_pw_leaves"]
    if delta == 1e-5:
        assert info["translation_group"] == 4
    assert_close(compare(w.tree, got, ref, w.ifpgh))


@pytest.mark.parametrize("delta", [1e-5, 1e-3])
def test_full_size_c3_one_density(device, delta):
    """The c3 tree with its first density only: one wave per leaf group in the
    n <= 8 downward pass, and the near field's tables split over two waves
    per target (x table 0 | the others)."""
    w = make_workload("c3", "full", delta=delta, eps=1e-9)
    fv = np.ascontiguousarray(w.fvals[..., :1])
    ref = O.oracle_boxfgt(w.tree, fv, 1, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, w.ifpgh)
    got = run_gpu(w.tree, fv, 1, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, w.ifpgh)
    assert got["plan"].info["xc_down"] < got["plan"].info["n_pw_leaves"]
    assert_close(compare(w.tree, got, ref, w.ifpgh))


def test_full_size_c3_free_space_matches_oracle(device):
    """The c3 tree in free space (iperiod = 0) at delta = 1e-5: translation
    groups of four whose 4 x 4 source patches are cut at the domain edge
    (empty cells, partial target lists)."""
    w = make_workload("c3", "full", delta=1e-5, eps=1e-9)
    ref = O.oracle_boxfgt(w.tree, w.fvals, w.nd, w.delta, w.eps, w.ipoly, 0, w.norder, w.ifpgh)
    got = run_gpu(w.tree, w.fvals, w.nd, w.delta, w.eps, w.ipoly, 0, w.norder, w.ifpgh)
    assert got["plan"].info["translation_group"] == 4
    assert_close(compare(w.tree, got, ref, w.ifpgh))


@pytest.mark.parametrize("iperiod,n,delta,ifpgh", [(0, 12, 1e-3, 3), (1, 16, 1e-3, 2), (1, 12, 3e-4, 1),
                                                    (0, 16, 1e-2, 3)])
def test_column_passes_match_oracle(device, iperiod, n, delta, ifpgh):
    """Leaf orders > 8 share each x contraction over a column (same cutoff
    box, level and x cell) in both leaf passes; eps = 1e-12 gives 3 tiles of
    16 y frequencies (npw = 66, 72) besides the 2-tile cases."""
    g = GaussianSum(np.array([[0.1, -0.2], [-0.3, 0.25]]), np.array([0.03, 0.06]), np.array([1.0, -0.5]))
    fun = lambda p: np.stack([g(p), np.cos(2 * p[..., 0]) * g(p)], axis=-1)  # noqa: E731
    t = build_adaptive(fun, 2, n, 0, 1e-7, 6, iperiod=iperiod, min_level=2, eta=0.0)
    fv = t.sample(fun, n, 0, 2)
    ref = O.oracle_boxfgt(t, fv, 2, delta, 1e-12, 0, iperiod, n, ifpgh)
    got = run_gpu(t, fv, 2, delta, 1e-12, 0, iperiod, n, ifpgh)
    info = got["plan"].info
    print(f"iperiod={iperiod} n={n} delta={delta}: npw={info['npw']} leaves={info['n_pw_leaves']} "
          f"x contractions up/down {info['xc_up']}/{info['xc_down']}", flush=True)
    assert info["xc_up"] < info["n_pw_leaves"] and info["xc_down"] < info["n_pw_leaves"]
    assert_close(compare(t, got, ref, ifpgh))


def test_full_size_c5_matches_oracle(device):
    """The 104,317-leaf n = 16 workload at full size: 16-leaf upward chunks
    (fewer chunk sums than one per 8 leaves) and grouped translation."""
    w = make_workload("c5", "full")
    print(f"case c5/full: leaves={w.tree.nleaves}", flush=True)
    ref = O.oracle_boxfgt(w.tree, w.fvals, w.nd, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, w.ifpgh)
    got = run_gpu(w.tree, w.fvals, w.nd, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, w.ifpgh)
    info = got["plan"].info
    assert info["n_chunks"] < info["n_pw_leaves"] / 8
    assert_close(compare(w.tree, got, ref, w.ifpgh))


@pytest.mark.skipif(not O.reference_available(), reason="reference library not built")
@pytest.mark.parametrize("name,kw", [("c1", {}), ("c3", {"delta": 1e-5, "eps": 1e-9}), ("c3", {"delta": 1e-1, "eps": 1e-9})])
def test_workload_matches_reference(device, name, kw):
    w = make_workload(name, "full" if name == "c1" else "tiny", **kw)
    ref = O.reference_boxfgt(w.tree, w.fvals, w.nd, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, w.ifpgh)
    got = run_gpu(w.tree, w.fvals, w.nd, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, w.ifpgh)
    assert np.array_equal(got["itree"], ref.itree)
    assert_close(compare(w.tree, got, ref, w.ifpgh))


@pytest.mark.parametrize("seed", range(8))
def test_random_trees_match_oracle(device, seed):
    """Random regimes: periodic/free space, Legendre/Chebyshev, n, nd, ifpgh, delta, eps."""
    rng = np.random.default_rng(200 + seed)
    iperiod = int(rng.integers(0, 2))
    ipoly = int(rng.integers(0, 2))
    n = int(rng.choice([4, 6, 8, 12, 16]))
    nd = int(rng.integers(1, 4))
    ifpgh = int(rng.integers(1, 4))
    delta = float(10 ** rng.uniform(-5.5, -0.5))
    eps = float(rng.choice([1e-6, 1e-9, 1e-12]))
    g = GaussianSum(rng.uniform(-0.5, 0.5, (3, 2)), rng.uniform(0.02, 0.1, 3), rng.uniform(-1, 1, 3))
    fun = lambda p: np.stack([g(p), np.sin(3 * p[..., 0]) * g(p), g(p) ** 2][:nd], axis=-1)  # noqa: E731
    t = build_adaptive(fun, nd, n, ipoly, 1e-6, 6, iperiod=iperiod, min_level=1, eta=0.0)
    fv = t.sample(fun, n, ipoly, nd)
    print(f"seed {seed}: iperiod={iperiod} ipoly={ipoly} n={n} nd={nd} ifpgh={ifpgh} delta={delta:.3g} eps={eps:g} leaves={t.nleaves}", flush=True)
    ref = O.oracle_boxfgt(t, fv, nd, delta, eps, ipoly, iperiod, n, ifpgh)
    got = run_gpu(t, fv, nd, delta, eps, ipoly, iperiod, n, ifpgh)
    assert np.array_equal(got["itree"], ref.itree)
    assert_close(compare(t, got, ref, ifpgh))


@pytest.mark.parametrize("n,nd,ifpgh,iperiod,ipoly", [(8, 3, 3, 1, 0), (8, 2, 2, 0, 1), (6, 3, 1, 1, 0), (4, 2, 3, 0, 0),
                                                      (4, 3, 2, 1, 1), (8, 4, 3, 0, 0), (6, 2, 3, 1, 1),
                                                      (8, 1, 3, 1, 0), (8, 1, 2, 0, 1), (6, 1, 1, 1, 1), (4, 1, 3, 0, 0),
                                                      (4, 1, 2, 1, 0)])
def test_packed_near_field(device, n, nd, ifpgh, iperiod, ipoly):
    """n <= 8: with nd >= 2 the near field holds two densities and two tables
    per matrix tile (an odd nd leaves the last tile's second density empty);
    nd = 1 runs the one-density kernel.  Small delta, so that most leaves are
    near-field targets."""
    rng = np.random.default_rng(300 + 10 * n + nd)
    g = GaussianSum(rng.uniform(-0.5, 0.5, (3, 2)), rng.uniform(0.02, 0.1, 3), rng.uniform(-1, 1, 3))
    comps = lambda p: [g(p), np.sin(3 * p[..., 0]) * g(p), g(p) ** 2, np.cos(2 * p[..., 1]) + 0.1 * g(p)]  # noqa: E731
    fun = lambda p: np.stack(comps(p)[:nd], axis=-1)  # noqa: E731
    t = build_adaptive(fun, nd, n, ipoly, 1e-6, 6, iperiod=iperiod, min_level=1, eta=0.0)
    fv = t.sample(fun, n, ipoly, nd)
    delta, eps = 1e-5, 1e-9
    print(f"n={n} nd={nd} ifpgh={ifpgh} iperiod={iperiod} ipoly={ipoly} leaves={t.nleaves}", flush=True)
    ref = O.oracle_boxfgt(t, fv, nd, delta, eps, ipoly, iperiod, n, ifpgh)
    got = run_gpu(t, fv, nd, delta, eps, ipoly, iperiod, n, ifpgh)
    assert got["plan"].info["n_near_pairs"] > 0
    assert np.array_equal(got["itree"], ref.itree)
    assert_close(compare(t, got, ref, ifpgh))


@pytest.mark.parametrize("nd,ifpgh", [(1, 3), (1, 2), (2, 3)])
def test_near_field_n16(device, nd, ifpgh):
    """n = 16 on a near-field-dominated tree (the 406-leaf c2 tree at
    delta = 1e-6): with one density each target's tables are split over two
    waves (x table 0 | the others), with two one wave per (target, density)."""
    w = make_workload("c2", "small")
    fv = np.ascontiguousarray(np.concatenate([w.fvals, 0.5 * w.fvals], axis=-1)[..., :nd])
    ref = O.oracle_boxfgt(w.tree, fv, nd, 1e-6, 1e-9, w.ipoly, w.iperiod, w.norder, ifpgh)
    got = run_gpu(w.tree, fv, nd, 1e-6, 1e-9, w.ipoly, w.iperiod, w.norder, ifpgh)
    assert got["plan"].info["n_near_pairs"] > 0
    assert_close(compare(w.tree, got, ref, ifpgh))


def test_fourier_mode_closed_form(device):
    """Heat semigroup on a Fourier mode: pot = pi delta exp(-delta|k|^2/4) f."""
    t = build_uniform(4)
    f = FourierMode(2, -1, phi=0.7)
    n = 16
    fv = t.sample(f, n, 0, 1)
    for delta in (1e-1, 1e-3, 1e-4):
        got = run_gpu(t, fv, 1, delta, 1e-9, 0, 1, n, 3)
        pot, grad, hess = f.exact(t.grid_points(n, 0), delta)
        assert leaf_rel_err(t, got["pot"][..., 0], pot) < 2e-8
        assert leaf_rel_err(t, got["grad"][..., 0], grad) < 2e-8
        assert leaf_rel_err(t, got["hess"][..., 0], hess) < 5e-8


def test_plan_reuse_accumulates(device):
    w = make_workload("c1", "full")
    ref = O.oracle_boxfgt(w.tree, w.fvals, w.nd, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, w.ifpgh)
    got = run_gpu(w.tree, w.fvals, w.nd, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, w.ifpgh, repeat=3)
    assert_close(compare(w.tree, got, ref, w.ifpgh, scale=3.0))


def test_gpu_tables_match_oracle(device):
    """FP64 tables generated on the GPU against the oracle's tables."""
    w = make_workload("c2", "tiny")
    from heat2d_per_box_amd.boxfgt import BoxFGT

    t = w.tree.copy()
    plan = BoxFGT(t, 1, w.delta, w.eps, 0, w.norder, 2)
    n, npw = w.norder, plan.info["npw"]
    L0 = plan.info["nlevstart"]
    pmax, npw_o = O.fgtpwterms(2 * t.boxsize[L0], w.delta, w.eps, 0)
    assert npw_o == npw
    ws, ts = O.pwnodes(pmax, npw)
    from heat2d_per_box_amd.tree import poly_nodes

    for lev in range(L0, t.nlevels + 1):
        tp = plan.table(0, lev)  # [m][k]
        ref = O.poly2pw(n, 0, ws, ts, w.delta, t.boxsize[lev])
        assert np.max(np.abs(tp - ref)) < 1e-13 * np.max(np.abs(ref))
        e = plan.table(1, lev)
        r = O.pw2pgh(n, ts, poly_nodes(n, 0), w.delta, t.boxsize[lev])
        for k in range(3):
            assert np.max(np.abs(e[k] - r[k])) < 1e-12 * np.max(np.abs(r[k]))
    for lev in range(0, t.nlevels + 1):
        tab, tabx, tabxx, ind = O.loctab_level(w.eps, 0, n, w.delta, t.boxsize[lev])
        gi = plan.table(3, lev)
        used = [q for q in range(13) if q not in (1, 11)]  # offsets +-5 never occur in a 2:1 tree
        assert np.array_equal(gi[used], ind[used])
        g = plan.table(2, lev)  # [q][t][j][m], masked
        for q in used:
            mask = np.zeros((n, n), dtype=bool)
            j0, j1 = ind[q][n]
            for j in range(n):
                if j0 <= j + 1 <= j1:
                    a, b = ind[q][j]
                    mask[j, max(a - 1, 0):max(b, 0)] = True
            for tt, T in enumerate((tab, tabx, tabxx)):
                want = np.where(mask, T[q], 0.0)
                big = max(np.max(np.abs(T[q])), 1e-300)
                assert np.max(np.abs(g[q, tt] - want)) < 1e-12 * big, (lev, q, tt)


def test_host_entry_point_matches_oracle(device):
    """bfgt_hip_eval on host arrays: accumulation into existing pot, itree side effects."""
    from heat2d_per_box_amd.boxfgt import boxfgt

    w = make_workload("c3", "tiny", delta=1e-5, eps=1e-9)
    t = w.tree.copy()
    nb, npb, nd = w.fvals.shape
    rng = np.random.default_rng(7)
    pot0 = 1e-6 * rng.standard_normal((nb, npb, nd))
    grad0 = 1e-4 * rng.standard_normal((nb, npb, 2, nd))
    hess0 = 1e-2 * rng.standard_normal((nb, npb, 3, nd))
    pot, grad, hess = pot0.copy(), grad0.copy(), hess0.copy()
    ti = np.zeros(20)
    boxfgt(nd, 2, w.delta, w.eps, 0, 1, w.norder, npb, nb, t.nlevels, t.ltree, t.itree, t.iptr, t.centers, t.boxsize,
           w.fvals, 3, pot, grad, hess, timeinfo=ti)
    ref = O.oracle_boxfgt(w.tree, w.fvals, nd, w.delta, w.eps, 0, 1, w.norder, 3)
    lv = t.nchild == 0
    assert np.array_equal(t.itree, ref.itree)
    assert leaf_rel_err(t, pot, pot0 + ref.pot) < 1e-12
    assert leaf_rel_err(t, grad, grad0 + ref.grad) < 1e-11
    assert leaf_rel_err(t, hess, hess0 + ref.hess) < 1e-11
    assert np.array_equal(pot[~lv], pot0[~lv])  # non-leaf boxes untouched
    assert ti[0] > 0 and np.all(ti >= 0)


def test_empty_and_edge_trees(device):
    """Root-only tree (no near field, no PW leaves) and a single-level tree."""
    from heat2d_per_box_amd.tree import build_tree

    for refined in (set(), {(0, 0, 0)}):
        t = build_tree(refined)
        n = 8
        fv = t.sample(lambda p: np.cos(2 * np.pi * p[..., 0]), n, 0, 1)
        for delta in (1e-1, 1e-2):
            ref = O.oracle_boxfgt(t, fv, 1, delta, 1e-9, 0, 1, n, 2)
            got = run_gpu(t, fv, 1, delta, 1e-9, 0, 1, n, 2)
            if np.linalg.norm(ref.pot) == 0:
                assert np.linalg.norm(got["pot"]) == 0
            else:
                assert_close(compare(t, got, ref, 2))


def test_npw_limit_is_an_error(device):
    """A shallow periodic tree with small delta needs more than 200 plane waves: the
    reference overflows its ws/ts(200) arrays; here the plan refuses with BFGT_ERR_NPW."""
    from heat2d_per_box_amd.boxfgt import BoxFGT, BoxFGTError
    from heat2d_per_box_amd.tree import build_tree

    t = build_tree(set())
    with pytest.raises(BoxFGTError) as ei:
        BoxFGT(t, 1, 1e-4, 1e-9, 0, 8, 1)
    assert ei.value.code == 7


def test_stage_profiling_windows(device):
    """Profiling windows: all stages, then one stage only (its two events alone),
    then calls outside any window; none of them changes the results."""
    import torch
    from heat2d_per_box_amd.boxfgt import BoxFGTError

    w = make_workload("c3", "tiny", delta=1e-5, eps=1e-9)
    ref = O.oracle_boxfgt(w.tree, w.fvals, w.nd, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, w.ifpgh)
    got = run_gpu(w.tree, w.fvals, w.nd, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, w.ifpgh)
    plan = got["plan"]
    fv = torch.from_numpy(np.ascontiguousarray(w.fvals)).cuda()
    nb, npb, nd = w.fvals.shape
    pot = torch.zeros((nb, npb, nd), dtype=torch.float64, device="cuda")
    grad = torch.zeros((nb, npb, 2, nd), dtype=torch.float64, device="cuda")
    hess = torch.zeros((nb, npb, 3, nd), dtype=torch.float64, device="cuda")
    plan.profile_begin(2)
    for _ in range(3):  # the third call is past the window
        plan(fv, pot, grad, hess)
    s_all, n_all = plan.profile_end()
    assert n_all == 2 and np.all(s_all >= 0) and s_all[3] == 0
    assert s_all[0] > 0 and s_all[4] > 0 and s_all[5] > 0
    plan.profile_begin(4, 4)
    for _ in range(2):
        plan(fv, pot, grad, hess)
    s_one, n_one = plan.profile_end()
    assert n_one == 2 and s_one[4] > 0
    assert np.all(np.delete(s_one, 4) == 0)
    with pytest.raises(BoxFGTError):
        plan.profile_begin(1, 3)
    torch.cuda.synchronize()
    out = {"pot": pot.cpu().numpy(), "grad": grad.cpu().numpy(), "hess": hess.cpu().numpy()}
    assert_close(compare(w.tree, out, ref, w.ifpgh, scale=5.0))


def test_density_loop_with_workgroup_sums(device):
    """Two densities on the c2 tree (19,834 leaves, n = 16): with at least 2048
    (chunk, y tile) waves each upward wave runs the densities in turn, and with
    merges the chunk sums of a cutoff box are added per workgroup -- the
    density loop of the reducing upward form (its LDS slots are refilled per
    density)."""
    w = make_workload("c2", "full")
    f2 = np.concatenate([w.fvals, np.random.default_rng(3).standard_normal(w.fvals.shape)], axis=2)
    f2 = np.ascontiguousarray(f2)
    ref = O.oracle_boxfgt(w.tree, f2, 2, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, w.ifpgh)
    got = run_gpu(w.tree, f2, 2, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, w.ifpgh)
    info = got["plan"].info
    ntn = ((info["npw"] + 1) // 2 + 15) // 16
    assert info["n_chunks"] * ntn >= 2048 and info["merge_writes"] > 0
    assert_close(compare(w.tree, got, ref, w.ifpgh))


def test_host_entry_point_keeps_plans(device, monkeypatch):
    """Repeated host-array calls on one tree reuse the kept plan: a second call
    gives the same results and writes the colleague lists into itree again
    (here over lists the caller scribbled on); other deltas and another tree get
    plans of their own (more than the four kept: the least recently used is
    dropped and rebuilt); keeping can be switched off and the plans released."""
    from heat2d_per_box_amd.boxfgt import boxfgt, release_plans

    def call(w, delta, scribble=False):
        t = w.tree.copy()
        nb, npb, nd = w.fvals.shape
        if scribble:
            nused = t.itree[2 * t.nlevels + 1]
            t.itree[t.iptr[5] - 1:t.iptr[5] - 1 + nused] = -7
            t.itree[t.iptr[6] - 1:t.iptr[6] - 1 + 9 * nused] = -7
        pot = np.zeros((nb, npb, nd))
        grad = np.zeros((nb, npb, 2, nd))
        boxfgt(nd, 2, delta, w.eps, w.ipoly, w.iperiod, w.norder, npb, nb, t.nlevels, t.ltree, t.itree, t.iptr,
               t.centers, t.boxsize, w.fvals, 2, pot, grad)
        return t, pot, grad

    def check_call(w, delta, scribble=False):
        t, pot, grad = call(w, delta, scribble)
        ref = O.oracle_boxfgt(w.tree, w.fvals, w.nd, delta, w.eps, w.ipoly, w.iperiod, w.norder, 2)
        assert np.array_equal(t.itree, ref.itree)
        assert leaf_rel_err(t, pot, ref.pot) < 1e-12 and leaf_rel_err(t, grad, ref.grad) < 1e-11
        return pot, grad

    release_plans()
    w = make_workload("c3", "tiny", delta=1e-5, eps=1e-9)
    p1, g1 = check_call(w, 1e-5)
    p2, g2 = check_call(w, 1e-5, scribble=True)  # the kept plan
    assert np.array_equal(p1, p2) and np.array_equal(g1, g2)
    for delta in (2e-5, 4e-5, 8e-5, 1.6e-4):
        check_call(w, delta)
    check_call(w, 1e-5)  # dropped as least recently used, built again
    check_call(make_workload("c1", "full"), make_workload("c1", "full").delta)
    monkeypatch.setenv("BFGT_HOST_PLAN_CACHE", "0")
    check_call(w, 1e-5)
    monkeypatch.delenv("BFGT_HOST_PLAN_CACHE")
    release_plans()
    check_call(w, 1e-5)
    release_plans()


def test_host_entry_staged_copies(device):
    """Host arrays above the staging threshold (c2: 40 MB pot, 80 MB grad) go
    through the pinned buffers in 32 MB chunks (several per array, both
    buffers): the first call copies its inputs on a second host thread while
    it builds the plan, the call on the kept plan copies them itself; both add
    the same transform to nonzero starting arrays, bit for bit, and match the
    oracle."""
    from heat2d_per_box_amd.boxfgt import boxfgt, release_plans

    w = make_workload("c2", "full")
    nb, npb, nd = w.fvals.shape
    rng = np.random.default_rng(11)
    pot0 = 1e-6 * rng.standard_normal((nb, npb, nd))
    grad0 = 1e-4 * rng.standard_normal((nb, npb, 2, nd))
    assert pot0.nbytes >= 32 << 20 and grad0.nbytes >= 64 << 20

    def call():
        t = w.tree.copy()
        pot, grad = pot0.copy(), grad0.copy()
        boxfgt(nd, 2, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, npb, nb, t.nlevels, t.ltree, t.itree, t.iptr,
               t.centers, t.boxsize, w.fvals, 2, pot, grad)
        return t, pot, grad

    release_plans()
    t1, p1, g1 = call()
    t2, p2, g2 = call()
    release_plans()
    assert np.array_equal(p1, p2) and np.array_equal(g1, g2) and np.array_equal(t1.itree, t2.itree)
    ref = O.oracle_boxfgt(w.tree, w.fvals, nd, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, 2)
    assert leaf_rel_err(t1, p1, pot0 + ref.pot) < 1e-12
    assert leaf_rel_err(t1, g1, grad0 + ref.grad) < 1e-11


@pytest.mark.parametrize("delta,eps", [(1e-1, 1e-6), (1e-1, 1e-9), (1e-5, 1e-6), (1e-3, 1e-6)])
def test_full_size_c3_regimes(device, delta, eps):
    """The remaining legs of the delta x eps sweep on the full 6,709-leaf c3
    tree.  delta = 1e-1 is the Fourier-series regime at full size: one cutoff
    box (the root) whose expansion is the sum of every upward chunk (hundreds
    of chunk sums merged into one), no translation partners and no near
    field; delta = 1e-5 and 1e-3 at eps = 1e-6 give the shorter expansions
    (npw = 34, 28) with the quad passes, and at 1e-5 the grouped translation
    and the packed near field."""
    w = make_workload("c3", "full", delta=delta, eps=eps)
    print(f"case c3/full delta={delta} eps={eps}: leaves={w.tree.nleaves}", flush=True)
    ref = O.oracle_boxfgt(w.tree, w.fvals, w.nd, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, w.ifpgh)
    got = run_gpu(w.tree, w.fvals, w.nd, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, w.ifpgh)
    info = got["plan"].info
    assert info["npwlevel"] == ref.info["npwlevel"] and info["npw"] == ref.info["npw"]
    assert np.array_equal(got["itree"], ref.itree)
    if delta == 1e-1:
        assert info["n_ancestors"] == 1 and info["n_near_pairs"] == 0
        assert info["n_pw_leaves"] == w.tree.nleaves and info["merge_reads"] > 100
    else:
        assert info["xc_down"] < info["n_pw_leaves"]
        if delta == 1e-5:
            assert info["n_near_pairs"] > 0 and info["translation_group"] == 4
    assert_close(compare(w.tree, got, ref, w.ifpgh))


def _vector_tree(n, iperiod, nd):
    g = GaussianSum(np.array([[0.12, -0.2], [-0.3, 0.25]]), np.array([0.03, 0.07]), np.array([1.0, -0.6]))
    fun = lambda p: np.stack([g(p), np.cos(3 * p[..., 1]) * g(p)][:nd], axis=-1)  # noqa: E731
    t = build_adaptive(fun, nd, n, 0, 1e-8, 5, iperiod=iperiod, min_level=2, eta=0.0)
    return t, t.sample(fun, n, 0, nd)


@pytest.mark.parametrize("n", [20, 24])
@pytest.mark.parametrize("iperiod", [0, 1])
def test_vector_fallback_matches_oracle(device, n, iperiod):
    """Leaf orders 17..32 run the vector-FMA kernels (upward, downward and
    near field): nd = 2, pot + grad + hess, a tree with PW leaves and near
    pairs (delta = 5e-4 puts the cutoff at level 3 of a level 2-4 tree)."""
    t, fv = _vector_tree(n, iperiod, 2)
    ref = O.oracle_boxfgt(t, fv, 2, 5e-4, 1e-9, 0, iperiod, n, 3)
    got = run_gpu(t, fv, 2, 5e-4, 1e-9, 0, iperiod, n, 3)
    info = got["plan"].info
    print(f"n={n} iperiod={iperiod}: leaves={t.nleaves} npw={info['npw']} pw leaves={info['n_pw_leaves']} "
          f"near pairs={info['n_near_pairs']}", flush=True)
    assert not info["mfma"] and info["n_pw_leaves"] > 0 and info["n_near_pairs"] > 0
    assert np.array_equal(got["itree"], ref.itree)
    assert_close(compare(t, got, ref, 3))


def test_npw_above_128_matches_oracle(device):
    """An expansion longer than the matrix-core kernels take (npw = 142 > 128:
    eps = 1e-20 in free space with the cutoff box just under twice the cutoff
    distance) runs the vector-FMA kernels at n = 8."""
    t = build_uniform(5, iperiod=0)
    g = GaussianSum(np.array([[0.1, 0.05]]), np.array([0.08]), np.array([1.0]))
    fv = t.sample(lambda p: g(p)[..., None], 8, 0, 1)
    delta, eps = 9.399e-05, 1e-20
    ref = O.oracle_boxfgt(t, fv, 1, delta, eps, 0, 0, 8, 2)
    got = run_gpu(t, fv, 1, delta, eps, 0, 0, 8, 2)
    info = got["plan"].info
    assert info["npw"] == ref.info["npw"] and info["npw"] > 128 and not info["mfma"]
    assert_close(compare(t, got, ref, 2))


@pytest.mark.parametrize("name", ["c2", "c5"])
def test_mixed_workloads_match_oracle(device, name):
    """The headline shapes with a near field: c2/mixed (20,353 leaves, 5.2 M
    points, 231 near-field target leaves at the cutoff level 4) and c5/mixed
    (106,246 leaves, 27.2 M points, leaves at levels 5 and 6 against the
    cutoff level 6): every stage runs at full size, the near field of n = 16
    beside 16-leaf upward chunks and grouped translation."""
    w = make_workload(name, "mixed")
    print(f"case {name}/mixed: leaves={w.tree.nleaves}", flush=True)
    ref = O.oracle_boxfgt(w.tree, w.fvals, w.nd, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, w.ifpgh)
    got = run_gpu(w.tree, w.fvals, w.nd, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, w.ifpgh)
    info = got["plan"].info
    assert info["n_near_pairs"] == ref.info["npairs"] and info["n_near_pairs"] > 1000
    assert info["n_pw_leaves"] > 0.9 * w.tree.nleaves
    assert np.array_equal(got["itree"], ref.itree)
    assert_close(compare(w.tree, got, ref, w.ifpgh))
