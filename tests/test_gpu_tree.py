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
