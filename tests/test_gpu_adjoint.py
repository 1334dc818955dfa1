"""GPU adjoint (transposed operator) of the box FGT.

The adjoint must be the exact transpose of the computed forward map
A: fvals -> (pot, grad, hess) on leaf boxes.  Checked three ways:
  * against the dense FP64 transpose of the oracle's operator on tiny trees
    (columns of A from the oracle on unit leaf values, batched as densities);
  * by the dot-product identity <g, A f> = <A^T g, f> on the benchmark
    workloads (every regime: PW only, mixed, near-field dominated, Fourier
    series, free space, periodic images of level-1 sources);
  * through torch.autograd on a multi-step heat solve, against central
    finite differences (h = 1e-6) in the initial values and in delta.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from heat2d_per_box_amd.densities import GaussianSum, make_workload
from heat2d_per_box_amd.tree import build_adaptive, build_uniform

pytestmark = pytest.mark.gpu


def _leaf_mask(tree, npb, nd):
    lv = (tree.nchild == 0)
    return np.broadcast_to(lv[:, None, None], (tree.centers.shape[0], npb, nd))


def _rand_leaf(rng, tree, shape):
    a = rng.standard_normal(shape)
    a[tree.nchild != 0] = 0.0
    return a


def _gpu_forward_adjoint(tree, nd, delta, eps, ipoly, iperiod, n, ifpgh, f, gp, gg, gh):
    import torch
    from heat2d_per_box_amd.boxfgt import BoxFGT

    plan = BoxFGT(tree.copy(), nd, delta, eps, ipoly, n, ifpgh, iperiod=iperiod)
    nb, npb = f.shape[0], f.shape[1]
    cu = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda() if a is not None else None  # noqa: E731
    pot = torch.zeros((nb, npb, nd), dtype=torch.float64, device="cuda")
    grad = torch.zeros((nb, npb, 2, nd), dtype=torch.float64, device="cuda") if ifpgh >= 2 else None
    hess = torch.zeros((nb, npb, 3, nd), dtype=torch.float64, device="cuda") if ifpgh >= 3 else None
    plan(cu(f), pot, grad, hess)
    fbar = torch.zeros((nb, npb, nd), dtype=torch.float64, device="cuda")
    plan.adjoint(cu(gp), fbar, cu(gg) if ifpgh >= 2 else None, cu(gh) if ifpgh >= 3 else None)
    torch.cuda.synchronize()
    out = {"pot": pot.cpu().numpy(), "grad": grad.cpu().numpy() if grad is not None else None,
           "hess": hess.cpu().numpy() if hess is not None else None, "fbar": fbar.cpu().numpy(), "plan": plan}
    return out


def _dot_identity(tree, nd, delta, eps, ipoly, iperiod, n, ifpgh, seed=0):
    rng = np.random.default_rng(seed)
    nb, npb = tree.centers.shape[0], n * n
    f = _rand_leaf(rng, tree, (nb, npb, nd))
    gp = _rand_leaf(rng, tree, (nb, npb, nd))
    gg = _rand_leaf(rng, tree, (nb, npb, 2, nd)) if ifpgh >= 2 else None
    gh = _rand_leaf(rng, tree, (nb, npb, 3, nd)) if ifpgh >= 3 else None
    r = _gpu_forward_adjoint(tree, nd, delta, eps, ipoly, iperiod, n, ifpgh, f, gp, gg, gh)
    lhs = float(np.sum(gp * r["pot"]))
    norm_l = np.linalg.norm(gp) * np.linalg.norm(r["pot"])
    if ifpgh >= 2:
        lhs += float(np.sum(gg * r["grad"]))
        norm_l += np.linalg.norm(gg) * np.linalg.norm(r["grad"])
    if ifpgh >= 3:
        lhs += float(np.sum(gh * r["hess"]))
        norm_l += np.linalg.norm(gh) * np.linalg.norm(r["hess"])
    rhs = float(np.sum(r["fbar"] * f))
    # fbar stays zero on non-leaf boxes
    assert np.all(r["fbar"][tree.nchild != 0] == 0.0)
    return abs(lhs - rhs) / norm_l, r["plan"].info


DOT_CASES = [
    ("c1", "full", {}),
    ("c2", "small", {}),
    ("c3", "tiny", {"delta": 1e-5, "eps": 1e-9}),
    ("c3", "tiny", {"delta": 1e-3, "eps": 1e-6}),
    ("c3", "tiny", {"delta": 1e-1, "eps": 1e-9}),
    ("c4", "small", {}),
    ("c5", "small", {}),
]


@pytest.mark.parametrize("name,scale,kw", DOT_CASES, ids=lambda v: str(v))
@pytest.mark.parametrize("ifpgh", [1, 2, 3])
def test_dot_product_identity(device, name, scale, kw, ifpgh):
    w = make_workload(name, scale, **kw)
    print(f"case {name}/{scale}/{kw} ifpgh={ifpgh}: leaves={w.tree.nleaves}", flush=True)
    err, info = _dot_identity(w.tree, w.nd, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, ifpgh)
    assert err < 1e-13, (err, info)


@pytest.mark.parametrize("delta", [1e-5, 1e-3])
def test_dot_product_identity_full_size_c3(device, delta):
    """Full-size c3: the transposed translation lists in groups of four
    (2,676 cutoff boxes at delta = 1e-5) and the leaf transform's transpose
    as 2 x 2 leaf blocks over column pairs."""
    w = make_workload("c3", "full", delta=delta, eps=1e-9)
    err, info = _dot_identity(w.tree, w.nd, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, 3)
    assert info["xc_adj_transform"] < info["n_pw_leaves"]
    if delta == 1e-5:
        assert info["translation_group"] == 4
    assert err < 1e-13, (err, info)


def test_dot_product_identity_full_size_c3_free_space(device):
    """The transposed translation groups of four in free space (source patches
    cut at the domain edge)."""
    w = make_workload("c3", "full", delta=1e-5, eps=1e-9)
    err, info = _dot_identity(w.tree, w.nd, w.delta, w.eps, w.ipoly, 0, w.norder, 3)
    assert info["translation_group"] == 4
    assert err < 1e-13, (err, info)


def test_dot_product_identity_full_size_c5(device):
    """16-leaf upward chunks in the adjoint's chunk sums (104,317 leaves)."""
    w = make_workload("c5", "full")
    err, info = _dot_identity(w.tree, w.nd, w.delta, w.eps, w.ipoly, w.iperiod, w.norder, 2)
    assert info["n_chunks"] < info["n_pw_leaves"] / 8
    assert err < 1e-13, (err, info)


@pytest.mark.parametrize("seed", range(4))
def test_dot_product_random_trees(device, seed):
    rng = np.random.default_rng(700 + seed)
    iperiod = int(rng.integers(0, 2))
    ipoly = int(rng.integers(0, 2))
    n = int(rng.choice([4, 6, 8, 12, 16]))
    nd = int(rng.integers(1, 3))
    ifpgh = int(rng.integers(1, 4))
    delta = float(10 ** rng.uniform(-5.5, -0.5))
    eps = float(rng.choice([1e-6, 1e-9, 1e-12]))
    g = GaussianSum(rng.uniform(-0.5, 0.5, (3, 2)), rng.uniform(0.02, 0.1, 3), rng.uniform(-1, 1, 3))
    fun = lambda p: np.stack([g(p), np.sin(3 * p[..., 0]) * g(p)][:nd], axis=-1)  # noqa: E731
    t = build_adaptive(fun, nd, n, ipoly, 1e-6, 6, iperiod=iperiod, min_level=1, eta=0.0)
    print(f"seed {seed}: iperiod={iperiod} ipoly={ipoly} n={n} nd={nd} ifpgh={ifpgh} delta={delta:.3g} "
          f"eps={eps:g} leaves={t.nleaves}", flush=True)
    err, info = _dot_identity(t, nd, delta, eps, ipoly, iperiod, n, ifpgh, seed=seed)
    assert err < 1e-13, (err, info)


def _oracle_dense(tree, delta, eps, ipoly, iperiod, n, ifpgh):
    """Dense FP64 operator of the oracle: rows (channel, leaf point), columns leaf points."""
    nb, npb = tree.centers.shape[0], n * n
    leaves = np.flatnonzero(tree.nchild == 0)
    cols = [(b, i) for b in leaves for i in range(npb)]
    ncol = len(cols)
    nchan = {1: 1, 2: 3, 3: 6}[ifpgh]
    A = np.zeros((nchan, len(leaves), npb, ncol))
    batch = 64
    for c0 in range(0, ncol, batch):
        c1 = min(ncol, c0 + batch)
        fv = np.zeros((nb, npb, c1 - c0))
        for k, (b, i) in enumerate(cols[c0:c1]):
            fv[b, i, k] = 1.0
        r = O.oracle_boxfgt(tree, fv, c1 - c0, delta, eps, ipoly, iperiod, n, ifpgh)
        A[0, :, :, c0:c1] = r.pot[leaves]
        if ifpgh >= 2:
            A[1, :, :, c0:c1] = r.grad[leaves][:, :, 0, :]
            A[2, :, :, c0:c1] = r.grad[leaves][:, :, 1, :]
        if ifpgh >= 3:
            for j in range(3):
                A[3 + j, :, :, c0:c1] = r.hess[leaves][:, :, j, :]
    return A.reshape(nchan * len(leaves) * npb, ncol), leaves


TINY = [
    # (tree builder, delta, eps, ipoly, iperiod, n, ifpgh)
    ("uniform3", 1e-3, 1e-9, 0, 1, 4, 3),   # PW + near field
    ("adaptive", 1e-4, 1e-9, 0, 1, 4, 2),   # near field dominated
    ("adaptive", 1e-2, 1e-9, 1, 1, 4, 3),   # Fourier series regime
    ("adaptive", 2e-3, 1e-6, 0, 1, 6, 1),   # periodic images of level-1 sources
    ("adaptive", 2e-3, 1e-6, 1, 0, 4, 3),   # free space
    ("uniform2", 1e-1, 1e-9, 0, 1, 4, 3),   # Fourier series, uniform
]


@pytest.mark.parametrize("case", TINY, ids=lambda c: f"{c[0]}-d{c[1]:g}-ip{c[4]}-pgh{c[6]}")
def test_dense_transpose_matches_oracle(device, case):
    kind, delta, eps, ipoly, iperiod, n, ifpgh = case
    if kind.startswith("uniform"):
        t = build_uniform(int(kind[-1]), iperiod=iperiod)
    else:
        g = GaussianSum(np.array([[0.1, -0.2]]), np.array([0.03]), np.array([1.0]))
        t = build_adaptive(lambda p: g(p)[..., None], 1, n, ipoly, 1e-4, 4, iperiod=iperiod, min_level=1, eta=0.0)
    A, leaves = _oracle_dense(t, delta, eps, ipoly, iperiod, n, ifpgh)
    rng = np.random.default_rng(5)
    nb, npb = t.centers.shape[0], n * n
    gp = _rand_leaf(rng, t, (nb, npb, 1))
    gg = _rand_leaf(rng, t, (nb, npb, 2, 1)) if ifpgh >= 2 else None
    gh = _rand_leaf(rng, t, (nb, npb, 3, 1)) if ifpgh >= 3 else None
    f = np.zeros((nb, npb, 1))
    r = _gpu_forward_adjoint(t, 1, delta, eps, ipoly, iperiod, n, ifpgh, f, gp, gg, gh)
    gvec = [gp[leaves, :, 0]]
    if ifpgh >= 2:
        gvec += [gg[leaves, :, 0, 0], gg[leaves, :, 1, 0]]
    if ifpgh >= 3:
        gvec += [gh[leaves, :, j, 0] for j in range(3)]
    want = A.T @ np.concatenate([v.reshape(-1) for v in gvec])
    got = r["fbar"][leaves, :, 0].reshape(-1)
    print(f"case {case}: leaves={len(leaves)} npw={r['plan'].info['npw']} npwlevel={r['plan'].info['npwlevel']}",
          flush=True)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-12


PACKED = [
    # (n, nd, ifpgh, delta, iperiod): the adjoint near field for n <= 8 with
    # two or more densities holds two densities per product (k_near_adj_pack)
    (4, 2, 3, 1e-4, 1),
    (6, 3, 2, 1e-4, 0),
    (8, 2, 1, 2e-4, 1),
    (8, 3, 3, 1e-3, 1),
]


@pytest.mark.parametrize("case", PACKED, ids=lambda c: f"n{c[0]}-nd{c[1]}-pgh{c[2]}-d{c[3]:g}-ip{c[4]}")
def test_dense_transpose_packed_near_field(device, case):
    """Every density against the oracle's dense transpose (the densities are
    independent: fbar_d = A^T g_d), near-field-dominated and mixed trees."""
    n, nd, ifpgh, delta, iperiod = case
    g = GaussianSum(np.array([[0.1, -0.2]]), np.array([0.03]), np.array([1.0]))
    t = build_adaptive(lambda p: g(p)[..., None], 1, n, 0, 1e-4, 4, iperiod=iperiod, min_level=1, eta=0.0)
    A, leaves = _oracle_dense(t, delta, 1e-9, 0, iperiod, n, ifpgh)
    rng = np.random.default_rng(9)
    nb, npb = t.centers.shape[0], n * n
    gp = _rand_leaf(rng, t, (nb, npb, nd))
    gg = _rand_leaf(rng, t, (nb, npb, 2, nd)) if ifpgh >= 2 else None
    gh = _rand_leaf(rng, t, (nb, npb, 3, nd)) if ifpgh >= 3 else None
    r = _gpu_forward_adjoint(t, nd, delta, 1e-9, 0, iperiod, n, ifpgh, np.zeros((nb, npb, nd)), gp, gg, gh)
    assert r["plan"].info["n_near_pairs"] > 0
    for d in range(nd):
        gvec = [gp[leaves, :, d]]
        if ifpgh >= 2:
            gvec += [gg[leaves, :, 0, d], gg[leaves, :, 1, d]]
        if ifpgh >= 3:
            gvec += [gh[leaves, :, j, d] for j in range(3)]
        want = A.T @ np.concatenate([v.reshape(-1) for v in gvec])
        got = r["fbar"][leaves, :, d].reshape(-1)
        assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-12, d


def test_autograd_matches_adjoint_and_fd(device):
    """torch.autograd through the Function: gradient in f and delta vs central FD."""
    import torch
    from heat2d_per_box_amd.autograd import GaussTransform

    w = make_workload("c4", "small")
    gt = GaussTransform(w.tree, w.nd, w.eps, w.ipoly, w.norder, ifpgh=1)
    rng = np.random.default_rng(11)
    f = torch.from_numpy(w.fvals).cuda().requires_grad_(True)
    # a target near the transform's output keeps the loss, and so the rounding
    # error of its central differences, small
    tgt = torch.from_numpy(0.05 * _rand_leaf(rng, w.tree, w.fvals.shape)).cuda()
    delta = torch.tensor(w.delta, dtype=torch.float64, requires_grad=True)

    def loss_fn(ff, dd):
        return ((gt(ff, dd) - tgt) ** 2).sum()

    loss = loss_fn(f, delta)
    loss.backward()
    v = torch.from_numpy(_rand_leaf(rng, w.tree, w.fvals.shape)).cuda()
    h = 1e-6
    with torch.no_grad():
        fd_f = (loss_fn(f + h * v, delta) - loss_fn(f - h * v, delta)) / (2 * h)
        fd_d = (loss_fn(f, delta + h) - loss_fn(f, delta - h)) / (2 * h)
    ad_f = (f.grad * v).sum()
    print(f"f: fd={fd_f.item():.12e} ad={ad_f.item():.12e}; delta: fd={fd_d.item():.12e} ad={delta.grad.item():.12e}",
          flush=True)
    assert abs(fd_f.item() - ad_f.item()) <= 1e-6 * abs(ad_f.item())
    assert abs(fd_d.item() - delta.grad.item()) <= 1e-6 * abs(delta.grad.item())
    gt.clear()


def test_heat_solve_gradients_fd(device):
    """Ten steps of u <- G_delta(u + dt (u - u^3)) on the uniform level-6 tree (nd=2,
    delta = dt = 1e-3), loss ||u^10 - target||^2, gradients in u^0 and delta through
    checkpointed steps, against central finite differences with h = 1e-6."""
    import torch
    from heat2d_per_box_amd.autograd import GaussTransform
    from heat2d_per_box_amd.heat import heat_solve

    w = make_workload("c4", "full")
    gt = GaussTransform(w.tree, w.nd, w.eps, w.ipoly, w.norder, ifpgh=1)
    rng = np.random.default_rng(44)
    u0 = torch.from_numpy(w.fvals).cuda().requires_grad_(True)
    g2 = GaussianSum(rng.uniform(-0.5, 0.5, (3, 2)), rng.uniform(0.05, 0.1, 3), rng.uniform(-1, 1, 3))
    tgt_np = w.tree.sample(lambda p: np.stack([g2(p), 0.5 * g2(p)], axis=-1), w.norder, w.ipoly, w.nd)
    tgt = torch.from_numpy(tgt_np).cuda()
    delta = torch.tensor(1e-3, dtype=torch.float64, requires_grad=True)

    def loss_fn(u, d, ckpt=True):
        return ((heat_solve(gt, u, d, 10, checkpoint_steps=ckpt) - tgt) ** 2).sum()

    loss = loss_fn(u0, delta)
    loss.backward()
    v = torch.from_numpy(_rand_leaf(rng, w.tree, w.fvals.shape)).cuda()
    h = 1e-6
    with torch.no_grad():
        fd_u = (loss_fn(u0 + h * v, delta, False) - loss_fn(u0 - h * v, delta, False)) / (2 * h)
        fd_d = (loss_fn(u0, delta + h, False) - loss_fn(u0, delta - h, False)) / (2 * h)
    ad_u = (u0.grad * v).sum()
    print(f"loss={loss.item():.6e} u0: fd={fd_u.item():.12e} ad={ad_u.item():.12e}; "
          f"delta: fd={fd_d.item():.12e} ad={delta.grad.item():.12e}", flush=True)
    assert abs(fd_u.item() - ad_u.item()) <= 1e-6 * abs(ad_u.item())
    assert abs(fd_d.item() - delta.grad.item()) <= 1e-6 * abs(delta.grad.item())
    assert math.isfinite(loss.item())
    gt.clear()


@pytest.mark.parametrize("n,iperiod", [(20, 1), (24, 0)])
def test_dot_product_identity_vector_near_field(device, n, iperiod):
    """Leaf orders above 16: a plan whose leaves are all near-field targets
    (delta = 1e-6) has the adjoint of the vector near-field kernel (one block
    per source, tables read transposed); nd = 2, pot + grad + hess."""
    from tests.test_gpu_parity import _vector_tree

    t, _ = _vector_tree(n, iperiod, 2)
    err, info = _dot_identity(t, 2, 1e-6, 1e-9, 0, iperiod, n, 3, seed=n)
    assert not info["has_pw"] and info["n_near_pairs"] > 0
    assert err < 1e-13, (err, info)


VECTOR_DENSE = [
    # (uniform level, n, delta, iperiod, ifpgh)
    (1, 20, 2e-2, 0, 3),  # free space, translation at level 0
    (2, 17, 1e-2, 0, 1),  # free space, PW lists among level-1 boxes
    (1, 20, 5e-2, 1, 3),  # periodic: Fourier series regime
]


@pytest.mark.parametrize("case", VECTOR_DENSE, ids=lambda c: f"l{c[0]}-n{c[1]}-d{c[2]:g}-ip{c[3]}-pgh{c[4]}")
def test_dense_transpose_vector_pw_kernels(device, case):
    """Leaf orders above 16 run the vector kernels; their adjoint (evaluation
    transposed with the conjugated tables, transposed translation lists,
    leaf transform transposed) against the oracle's dense FP64 transpose."""
    lev, n, delta, iperiod, ifpgh = case
    t = build_uniform(lev, iperiod=iperiod)
    A, leaves = _oracle_dense(t, delta, 1e-9, 0, iperiod, n, ifpgh)
    rng = np.random.default_rng(11)
    nb, npb = t.centers.shape[0], n * n
    gp = _rand_leaf(rng, t, (nb, npb, 1))
    gg = _rand_leaf(rng, t, (nb, npb, 2, 1)) if ifpgh >= 2 else None
    gh = _rand_leaf(rng, t, (nb, npb, 3, 1)) if ifpgh >= 3 else None
    r = _gpu_forward_adjoint(t, 1, delta, 1e-9, 0, iperiod, n, ifpgh, np.zeros((nb, npb, 1)), gp, gg, gh)
    info = r["plan"].info
    assert info["has_pw"] and not info["mfma"] and info["n_pw_leaves"] > 0
    gvec = [gp[leaves, :, 0]]
    if ifpgh >= 2:
        gvec += [gg[leaves, :, 0, 0], gg[leaves, :, 1, 0]]
    if ifpgh >= 3:
        gvec += [gh[leaves, :, j, 0] for j in range(3)]
    want = A.T @ np.concatenate([v.reshape(-1) for v in gvec])
    got = r["fbar"][leaves, :, 0].reshape(-1)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-12


@pytest.mark.parametrize("n,iperiod", [(20, 1), (24, 0)])
def test_dot_product_identity_vector_pw_kernels(device, n, iperiod):
    """Leaf orders above 16 on a tree with PW leaves and near pairs
    (delta = 5e-4: cutoff at level 3 of a level 2-4 tree), nd = 2, pot +
    grad + hess: the vector kernels' adjoint of every stage."""
    from tests.test_gpu_parity import _vector_tree

    t, _ = _vector_tree(n, iperiod, 2)
    err, info = _dot_identity(t, 2, 5e-4, 1e-9, 0, iperiod, n, 3, seed=n)
    assert info["has_pw"] and not info["mfma"] and info["n_pw_leaves"] > 0 and info["n_near_pairs"] > 0
    assert err < 1e-13, (err, info)


def test_dot_product_identity_npw_above_128(device):
    """An expansion longer than the matrix-core kernels take (npw = 142, n = 8,
    free space) runs the vector kernels forward and adjoint."""
    t = build_uniform(5, iperiod=0)
    err, info = _dot_identity(t, 1, 9.399e-05, 1e-20, 0, 0, 8, 2, seed=3)
    assert info["npw"] > 128 and not info["mfma"]
    assert err < 1e-13, (err, info)


def test_autograd_plan_cache_is_bounded(device):
    """A loop over many deltas keeps at most max_plans resident plans (least
    recently used closed first); a delta whose plan was closed is rebuilt and
    gives the same result bit for bit."""
    import torch
    from heat2d_per_box_amd.autograd import GaussTransform

    w = make_workload("c1", "full")
    gt = GaussTransform(w.tree, w.nd, w.eps, w.ipoly, w.norder, ifpgh=1, max_plans=3)
    f = torch.from_numpy(w.fvals).cuda()
    first = gt(f, 1e-3).cpu().numpy()
    for k in range(6):
        gt(f, 1e-3 * (1.1 + 0.1 * k))
        assert gt.n_plans <= 3
    again = gt(f, 1e-3).cpu().numpy()
    assert np.array_equal(first, again)
    ref = O.oracle_boxfgt(w.tree, w.fvals, w.nd, 1e-3, w.eps, w.ipoly, w.iperiod, w.norder, 1)
    lv = w.tree.nchild == 0
    assert np.linalg.norm((again - ref.pot)[lv]) / np.linalg.norm(ref.pot[lv]) < 1e-12
    gt.clear()
    assert gt.n_plans == 0
