"""Device tree build and re-meshing against the reference's own tree code.

The reference's vol_tree_mem2 / vol_tree_build2 / vol_tree_adap_fv
(common/tree_ext.f, compiled from its sources into oracle/_ref) and the
device tree (bfgt_tree_*) run on the same test function (oracle/tree_funcs.c,
fed to the device through a host callback so both see identical values).
Trees are compared box by box by cell: the same boxes, parents, children,
colleague lists and centres, and values on the nodes to rounding; the box
numbering of the boxes that level restriction and re-meshing add differs.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.test_tree_reference import BUMPS, as_quadtree, ref_build
from tests.treecmp import compare_trees, compare_values, level_restricted

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not O.reference_available(), reason="reference library not built")]

CASES = [
    # nd, ipoly, iperiod, iptype, norder, eps
    (1, 0, 1, 2, 8, 1e-6),
    (2, 1, 0, 2, 8, 1e-6),
    (1, 0, 1, 1, 6, 1e-6),
    (1, 0, 0, 0, 10, 1e-6),
    (1, 0, 1, 2, 16, 1e-9),
    (1, 0, 1, 2, 4, 1e-6),
]


def host_fun(nd, dpars):
    return lambda xy: O.tree_fun_eval(nd, xy, dpars)


def smooth_field(points):
    x, y = points[..., 0], points[..., 1]
    return np.sin(3 * x + 1) * np.cos(2 * y) + 0.3 * x * y


@pytest.mark.parametrize("nd,ipoly,iperiod,iptype,norder,eps", CASES, ids=lambda v: str(v))
def test_build_matches_reference(device, nd, ipoly, iperiod, iptype, norder, eps):
    from heat2d_per_box_amd.devtree import DeviceTree

    T, rintl, caps, dp = ref_build(nd, ipoly, iperiod, eps, norder, iptype)
    maxlev = (caps[1] - 1) // 2  # vol_tree_build2's level cap for these capacities
    # norms computed while the tree grows (vol_tree_mem2)
    D = DeviceTree.build(host_fun(nd, dp), nd, norder, eps, ipoly, iperiod, iptype=iptype, max_levels=maxlev,
                         max_boxes=caps[0], host=True)
    q, v = D.to_host()
    print(f"case nd={nd} ipoly={ipoly} iperiod={iperiod} iptype={iptype} n={norder}: reference {T.nboxes} boxes "
          f"{T.nlevels} levels, device {q.nboxes} boxes {q.nlevels} levels", flush=True)
    nl = T.nlevels
    assert np.allclose(D.rintl[: nl + 1], rintl[: nl + 1], rtol=1e-13, atol=0)
    R = as_quadtree(T, iperiod)
    compare_trees(q, R)
    assert np.array_equal(np.sort(q.centers, axis=0), np.sort(R.centers, axis=0))
    # the node coordinates may differ in the last bit (fused or separate
    # multiply-add), hence a few units of rounding on the values
    err = compare_values(q, v, R, T.vals, leaves_only=False)
    assert err <= 1e-14, err
    assert level_restricted(q, iperiod)
    D.close()
    # norms given (vol_tree_build2)
    D2 = DeviceTree.build(host_fun(nd, dp), nd, norder, eps, ipoly, iperiod, iptype=iptype, max_levels=maxlev,
                          max_boxes=caps[0], rintl=rintl[: maxlev + 1], host=True)
    q2, v2 = D2.to_host()
    compare_trees(q2, R)
    err = compare_values(q2, v2, R, T.vals, leaves_only=False)
    assert err <= 1e-14, err
    D2.close()


@pytest.mark.parametrize("nd,ipoly,iperiod,iptype,norder,eps", [c for c in CASES if c[4] % 2 == 0],
                         ids=lambda v: str(v))
def test_adapt_matches_reference(device, nd, ipoly, iperiod, iptype, norder, eps):
    import torch

    from heat2d_per_box_amd.devtree import DeviceTree
    from heat2d_per_box_amd.tree import box_grid

    T, rintl, caps, dp = ref_build(nd, ipoly, iperiod, eps, norder, iptype)
    maxlev = (caps[1] - 1) // 2
    D = DeviceTree.build(host_fun(nd, dp), nd, norder, eps, ipoly, iperiod, iptype=iptype, max_levels=maxlev,
                         max_boxes=caps[0], rintl=rintl[: maxlev + 1], host=True)
    q, _ = D.to_host()
    # carried values: a smooth field sampled on each tree's own nodes
    u_ref = np.repeat(smooth_field(box_grid(T.centers, T.boxsize[T.ilevel], norder, ipoly))[..., None], nd, axis=2)
    u_dev = np.repeat(smooth_field(box_grid(q.centers, q.boxsize[q.ilevel], norder, ipoly))[..., None], nd, axis=2)
    D.values().copy_(torch.from_numpy(np.ascontiguousarray(u_dev)))
    dp1 = O.tree_fun_dpars(BUMPS, 0.05)
    A = O.reference_tree_adapt(T, u_ref, nd, ipoly, iperiod, eps, norder, iptype, rintl, dp1, *caps)
    mlev = caps[1]
    D.adapt(host_fun(nd, dp1), eps, rintl=rintl[: mlev + 1], max_levels=mlev, host=True)
    qa, va = D.to_host()
    print(f"adapt nd={nd} ipoly={ipoly} iperiod={iperiod} n={norder}: {T.nboxes} -> reference {A.nboxes}, "
          f"device {qa.nboxes} boxes", flush=True)
    R = as_quadtree(A, iperiod)
    compare_trees(qa, R)
    err = compare_values(qa, va, R, A.vals, leaves_only=True)
    print(f"  carried values: max rel diff {err:.2e}", flush=True)
    assert err <= 1e-13, err
    assert level_restricted(qa, iperiod)
    # a second step from the adapted trees
    dp2 = O.tree_fun_dpars(BUMPS, 0.1)
    A2 = O.reference_tree_adapt(A, A.vals, nd, ipoly, iperiod, eps, norder, iptype, rintl, dp2, *caps)
    D.adapt(host_fun(nd, dp2), eps, rintl=rintl[: mlev + 1], max_levels=mlev, host=True)
    qb, vb = D.to_host()
    R2 = as_quadtree(A2, iperiod)
    compare_trees(qb, R2)
    err = compare_values(qb, vb, R2, A2.vals, leaves_only=True)
    assert err <= 1e-13, err
    D.close()


def test_device_function_builds_the_same_tree(device):
    """A GPU-side function (torch) instead of the host callback: same tree,
    values to rounding."""
    import torch

    from heat2d_per_box_amd.devtree import DeviceTree

    nd, norder, eps = 1, 8, 1e-6
    T, rintl, caps, dp = ref_build(nd, 0, 1, eps, norder, 2)

    def fun(xy):
        s = torch.zeros(xy.shape[0], dtype=torch.float64, device=xy.device)
        for (a, cx, cy, vx, vy, w) in BUMPS:
            s += a * torch.exp(-((xy[:, 0] - cx) ** 2 + (xy[:, 1] - cy) ** 2) / w)
        return s

    maxlev = (caps[1] - 1) // 2
    D = DeviceTree.build(fun, nd, norder, eps, 0, 1, max_levels=maxlev, max_boxes=caps[0])
    q, v = D.to_host()
    R = as_quadtree(T, 1)
    compare_trees(q, R)
    err = compare_values(q, v, R, T.vals, leaves_only=False)
    assert err <= 1e-14, err
    D.close()


def test_tree_plan_matches_oracle(device):
    """The resident FGT on a device-built tree and its values, against the
    oracle on the exported tree."""
    import torch

    from heat2d_per_box_amd.devtree import DeviceTree
    from tests.fields import leaf_rel_err

    nd, norder = 2, 8
    dp = O.tree_fun_dpars(BUMPS, 0.0)
    D = DeviceTree.build(host_fun(nd, dp), nd, norder, 1e-6, 0, 1, max_levels=12, max_boxes=20000, host=True)
    q, v = D.to_host()
    plan = D.plan(1e-3, 1e-9, 2)
    nb, npb = q.nboxes, norder * norder
    pot = torch.zeros((nb, npb, nd), dtype=torch.float64, device="cuda")
    grad = torch.zeros((nb, npb, 2, nd), dtype=torch.float64, device="cuda")
    plan(D.values(), pot, grad)
    torch.cuda.synchronize()
    ref = O.oracle_boxfgt(q, v, nd, 1e-3, 1e-9, 0, 1, norder, 2)
    assert leaf_rel_err(q, pot.cpu().numpy(), ref.pot) < 1e-12
    assert leaf_rel_err(q, grad.cpu().numpy(), ref.grad) < 1e-11
    plan.close()
    D.close()


def test_capacity_and_callback_errors(device):
    from heat2d_per_box_amd._lib import BoxFGTError
    from heat2d_per_box_amd.devtree import DeviceTree

    dp = O.tree_fun_dpars(BUMPS, 0.0)
    with pytest.raises(BoxFGTError) as ei:
        DeviceTree.build(host_fun(1, dp), 1, 8, 1e-6, 0, 1, max_levels=12, max_boxes=50, host=True)
    assert ei.value.code == 12

    class Boom(RuntimeError):
        pass

    def bad(xy):
        raise Boom("user function failed")

    with pytest.raises(Boom):
        DeviceTree.build(bad, 1, 8, 1e-6, 0, 1, max_levels=12, max_boxes=20000, host=True)


def capacity_layout(q, vals, mboxes, mlevels):
    """A QuadTree in the reference builders' capacity layout (laddr for
    mlevels + 1 levels, arrays for mboxes boxes), as an oracle.RefTree."""
    nb, nl = q.nboxes, q.nlevels
    ltree = 2 * (mlevels + 1) + 17 * mboxes
    itree = np.zeros(ltree, dtype=np.int32)
    lad = 2 * (mlevels + 1)
    iptr = np.array([1, lad + 1, lad + 1 + mboxes, lad + 1 + 2 * mboxes, lad + 1 + 3 * mboxes,
                     lad + 1 + 7 * mboxes, lad + 1 + 8 * mboxes, lad + 1 + 17 * mboxes], dtype=np.int32)
    itree[: 2 * (nl + 1)] = q.laddr.reshape(-1)
    for k, arr in ((1, q.ilevel), (2, q.iparent), (3, q.nchild), (4, q.ichild.reshape(-1)), (5, q.ncoll),
                   (6, q.coll.reshape(-1))):
        o = iptr[k] - 1
        itree[o : o + arr.shape[0]] = arr
    return O.RefTree(nb, nl, itree, iptr, q.centers.copy(), q.boxsize.copy(), np.ascontiguousarray(vals))


def test_import_round_trip(device):
    """A tree built elsewhere (the 19,834-leaf benchmark tree, and a tree of
    the reference's builder in its capacity layout and box order) comes back
    from the device with the same boxes, links, colleagues and values."""
    from heat2d_per_box_amd.densities import make_workload
    from heat2d_per_box_amd.devtree import DeviceTree

    w = make_workload("c2", "full")
    D = DeviceTree.from_tree(w.tree, w.norder, w.fvals, nd=w.nd)
    q, v = D.to_host()
    assert q.nboxes == w.tree.nboxes
    compare_trees(q, w.tree)
    assert compare_values(q, v, w.tree, w.fvals, leaves_only=False) == 0.0
    D.close()
    T, rintl, caps, dp = ref_build(2, 0, 1, 1e-6, 8, 2)
    R = as_quadtree(T, 1)
    D = DeviceTree.from_tree(R, 8, T.vals, nd=2)
    q, v = D.to_host()
    compare_trees(q, R)
    assert compare_values(q, v, R, T.vals, leaves_only=False) == 0.0
    D.close()


def test_adapt_benchmark_tree_matches_reference(device):
    """Re-meshing the 19,834-leaf benchmark tree (n = 16, 26,445 boxes) for
    two moving Gaussians, the benchmark density carried along: the device and
    the reference's vol_tree_adap_fv (on the same tree in its capacity
    layout) give the same tree and the carried values to rounding."""
    from heat2d_per_box_amd.densities import make_workload
    from heat2d_per_box_amd.devtree import DeviceTree

    w = make_workload("c2", "full")
    nd, n = w.nd, w.norder
    bumps = [(1.0, 0.1, -0.05, 1.0, 0.5, 2e-4), (0.7, -0.3, 0.25, -0.5, 0.3, 5e-4)]
    dp = O.tree_fun_dpars(bumps, 0.02)
    mlev = w.tree.nlevels + 6
    rintl = np.full(201, 0.05)
    mboxes = 3 * w.tree.nboxes
    R0 = capacity_layout(w.tree, w.fvals, mboxes, mlev)
    A = O.reference_tree_adapt(R0, w.fvals, nd, 0, 1, 1e-9, n, 2, rintl, dp, mboxes, mlev, R0.itree.shape[0])
    D = DeviceTree.from_tree(w.tree, n, w.fvals, nd=nd, max_boxes=mboxes)
    D.adapt(host_fun(nd, dp), 1e-9, rintl=rintl[: mlev + 1], max_levels=mlev, host=True)
    q, v = D.to_host()
    R = as_quadtree(A, 1)
    print(f"benchmark tree re-meshed: {w.tree.nboxes} -> reference {A.nboxes}, device {q.nboxes} boxes", flush=True)
    compare_trees(q, R)
    err = compare_values(q, v, R, A.vals, leaves_only=True)
    print(f"  carried values: max rel diff {err:.2e}", flush=True)
    assert err <= 1e-13, err
    D.close()
