"""The callers' time step around boxfgt on a resident tree (VolPot).

* bfgt_plan_leaf_points reproduces the reference grid (centre + node * b/2).
* bfgt_plan_execute_scaled adds scale * (the unscaled result).
* A trapezoid and a Simpson step match the same sum built from the oracle
  (the reference's volpot loop, hpots/hpots2dadap.f:280-381) to 1e-12.
* Closed form: for f(x, t) = exp(a t) cos(k.x + phi) on the periodic unit box
  the step equals sum_j dt w_j exp(a (tpre + t_j)) exp(-|k|^2 (dt - t_j)) f,
  the heat semigroup applied node by node, to the transform's eps.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from heat2d_per_box_amd.tree import box_grid, build_adaptive, build_uniform
from heat2d_per_box_amd.densities import GaussianSum

pytestmark = pytest.mark.gpu


def _tree():
    g = GaussianSum(np.array([[0.1, -0.2], [-0.3, 0.25]]), np.array([0.03, 0.06]), np.array([1.0, -0.5]))
    return build_adaptive(lambda p: g(p)[..., None], 1, 8, 1, 1e-7, 7, iperiod=1, min_level=2, eta=0.0)


def test_leaf_points_and_scaled_execute(device):
    import torch
    from heat2d_per_box_amd.boxfgt import BoxFGT

    t = _tree()
    n = 8
    plan = BoxFGT(t.copy(), 1, 2e-3, 1e-9, 1, n, 2)
    xy = plan.leaf_points().cpu().numpy()
    lv = t.nchild == 0
    want = box_grid(t.centers[lv], t.boxsize[t.ilevel[lv]], n, 1)
    assert np.max(np.abs(xy[lv] - want)) < 1e-15
    assert np.all(xy[~lv] == 0.0)
    rng = np.random.default_rng(1)
    f = rng.standard_normal((t.centers.shape[0], n * n, 1))
    f[~lv] = 0
    fv = torch.from_numpy(f).cuda()
    p1 = torch.zeros_like(fv)
    g1 = torch.zeros((fv.shape[0], n * n, 2, 1), dtype=torch.float64, device="cuda")
    p2, g2 = torch.zeros_like(p1), torch.zeros_like(g1)
    plan(fv, p1, g1)
    plan(fv, p2, g2, scale=-0.37)
    torch.cuda.synchronize()
    rel = lambda a, b: float((a - b).norm() / b.norm())  # noqa: E731
    assert rel(p2, -0.37 * p1) < 1e-15 and rel(g2, -0.37 * g1) < 1e-15


def _oracle_step(t, n, dt, nordert, eps, fforce_np, tpre):
    from heat2d_per_box_amd.volpot import time_rule

    tn, tw = time_rule(nordert, dt)
    lv = t.nchild == 0
    pts = np.zeros((t.centers.shape[0], n * n, 2))
    pts[lv] = box_grid(t.centers[lv], t.boxsize[t.ilevel[lv]], n, 1)
    vpot = np.zeros((t.centers.shape[0], n * n, 1))
    for tj, wj in zip(tn[:-1], tw[:-1]):
        f = fforce_np(pts, tpre + tj)
        f[~lv] = 0
        r = O.oracle_boxfgt(t, f, 1, 4 * (dt - tj), eps, 1, 1, n, 1)
        vpot += r.pot * dt * wj / (4 * math.pi * (dt - tj))
    f = fforce_np(pts, tpre + tn[-1])
    vpot[lv] += dt * tw[-1] * f[lv]
    return vpot


@pytest.mark.parametrize("nordert", [2, 4])
def test_volpot_matches_oracle_loop(device, nordert):
    import torch
    from heat2d_per_box_amd.volpot import VolPot

    t = _tree()
    n, dt, eps, tpre = 8, 1e-3, 1e-9, 0.3

    def f_np(p, tt):
        x, y = p[..., 0], p[..., 1]
        return (np.exp(-((x - 0.1) ** 2 + (y + 0.2) ** 2) / 0.002) * np.cos(5 * tt) + 0.3 * np.sin(2 * np.pi * x))[
            ..., None]

    def f_t(p, tt):
        x, y = p[..., 0], p[..., 1]
        return (torch.exp(-((x - 0.1) ** 2 + (y + 0.2) ** 2) / 0.002) * math.cos(5 * tt)
                + 0.3 * torch.sin(2 * math.pi * x))[..., None]

    vp = VolPot(t, n, dt, nordert=nordert, eps=eps, ipoly=1)
    got = vp(f_t, tpre).cpu().numpy()
    torch.cuda.synchronize()
    want = _oracle_step(t, n, dt, nordert, eps, f_np, tpre)
    lv = t.nchild == 0
    err = np.linalg.norm((got - want)[lv]) / np.linalg.norm(want[lv])
    print(f"volpot nordert={nordert}: rel err vs oracle loop {err:.2e}", flush=True)
    assert err < 1e-12
    vp.close()


@pytest.mark.parametrize("nordert", [2, 4])
def test_volpot_fourier_closed_form(device, nordert):
    import torch
    from heat2d_per_box_amd.volpot import VolPot, time_rule

    t = build_uniform(4)
    n, dt, eps, tpre, a = 16, 2e-3, 1e-10, 0.1, 1.7
    k = 2 * math.pi * np.array([2.0, -1.0])
    phi = 0.4

    def f_t(p, tt):
        return (math.exp(a * tt) * torch.cos(k[0] * p[..., 0] + k[1] * p[..., 1] + phi))[..., None]

    vp = VolPot(t, n, dt, nordert=nordert, eps=eps, ipoly=1)
    got = vp(f_t, tpre)
    torch.cuda.synchronize()
    tn, tw = time_rule(nordert, dt)
    factor = sum(dt * wj * math.exp(a * (tpre + tj)) * math.exp(-float(k @ k) * (dt - tj)) for tj, wj in zip(tn, tw))
    lv = torch.from_numpy(t.nchild == 0).cuda()
    pts = vp.points
    want = factor * torch.cos(k[0] * pts[..., 0] + k[1] * pts[..., 1] + phi)[..., None]
    err = float(((got - want)[lv]).norm() / want[lv].norm())
    print(f"volpot closed form nordert={nordert}: rel err {err:.2e}", flush=True)
    assert err < 1e-8
    vp.close()
