"""Seeded synthetic densities and the trees they live on.

The workloads of the box FGT benchmarks and parity tests: sums of Gaussians
``a * exp(-|x - c|^2 / w^2)`` periodised over the 3x3 neighbouring cells of
the unit box ``[-0.5, 0.5]^2``, spot patterns for the multi-density case,
and the quadtrees they are tabulated on.  Random draws use
``numpy.random.default_rng(seed)`` in the order centres -> widths ->
amplitudes.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Callable, List, Optional, Sequence

import numpy as np

from .tree import QuadTree, build_adaptive, build_uniform

__all__ = ["GaussianSum", "Workload", "make_workload", "WORKLOADS"]


@dataclass
class GaussianSum:
    centers: np.ndarray  # (m, 2)
    widths: np.ndarray  # (m,)
    amps: np.ndarray  # (m,)
    nimage: int = 1  # images per side: 1 -> 3x3 cells
    offset: float = 0.0
    clip: Optional[tuple] = None

    def __call__(self, p: np.ndarray) -> np.ndarray:
        x = p[..., 0]
        y = p[..., 1]
        out = np.full(x.shape, self.offset, dtype=np.float64)
        r = range(-self.nimage, self.nimage + 1)
        for c, w, a in zip(self.centers, self.widths, self.amps):
            for mx in r:
                for my in r:
                    out += a * np.exp(-((x - c[0] - mx) ** 2 + (y - c[1] - my) ** 2) / w**2)
        if self.clip is not None:
            out = np.clip(out, *self.clip)
        return out

    def sample_boxes(self, centers: np.ndarray, size, norder: int, ipoly: int) -> np.ndarray:
        """Values on the grids of boxes (m, norder**2), evaluating each Gaussian image only on
        boxes within 6.5 widths of it (the neglected terms are below 1e-18 of its amplitude)."""
        from .tree import box_grid

        pts = box_grid(centers, size, norder, ipoly)
        m = centers.shape[0]
        half = 0.5 * np.broadcast_to(np.asarray(size, dtype=np.float64), (m,))
        out = np.full(pts.shape[:2], self.offset, dtype=np.float64)
        r = range(-self.nimage, self.nimage + 1)
        for c, w, a in zip(self.centers, self.widths, self.amps):
            reach = 6.5 * w
            for mx in r:
                for my in r:
                    gx, gy = c[0] + mx, c[1] + my
                    dx = np.maximum(np.abs(centers[:, 0] - gx) - half, 0.0)
                    dy = np.maximum(np.abs(centers[:, 1] - gy) - half, 0.0)
                    sel = np.nonzero(dx * dx + dy * dy < reach * reach)[0]
                    if sel.size == 0:
                        continue
                    q = pts[sel]
                    out[sel] += a * np.exp(-((q[..., 0] - gx) ** 2 + (q[..., 1] - gy) ** 2) / w**2)
        if self.clip is not None:
            out = np.clip(out, *self.clip)
        return out

    def heat_potential(self, p: np.ndarray, delta: float) -> np.ndarray:
        """Closed form of ``int exp(-|x-y|^2/delta) f(y) dy`` over R^2 for the image sum
        (unnormalised kernel): each Gaussian maps to
        ``a pi delta w^2/(delta+w^2) exp(-|x-c|^2/(delta+w^2))``.  Only exact for
        ``clip is None``."""
        x = p[..., 0]
        y = p[..., 1]
        out = np.zeros(x.shape)
        r = range(-self.nimage, self.nimage + 1)
        for c, w, a in zip(self.centers, self.widths, self.amps):
            s = delta + w * w
            for mx in r:
                for my in r:
                    out += a * np.pi * delta * w * w / s * np.exp(-((x - c[0] - mx) ** 2 + (y - c[1] - my) ** 2) / s)
        return out


def gaussian_sum(rng: np.random.Generator, m: int, wlo: float, whi: float, alo=-1.0, ahi=1.0, logw=False) -> GaussianSum:
    c = rng.uniform(-0.5, 0.5, size=(m, 2))
    if logw:
        w = np.exp(rng.uniform(np.log(wlo), np.log(whi), size=m))
    else:
        w = rng.uniform(wlo, whi, size=m)
    a = rng.uniform(alo, ahi, size=m)
    return GaussianSum(c, w, a)


class Stack:
    """nd densities evaluated together: returns (..., nd)."""

    def __init__(self, fields: Sequence[Callable[[np.ndarray], np.ndarray]]):
        self.fields = list(fields)

    def __call__(self, p):
        return np.stack([f(p) for f in self.fields], axis=-1)

    def sample_boxes(self, centers, size, norder, ipoly):
        from .tree import eval_on_boxes

        return np.stack([eval_on_boxes(f, centers, size, norder, ipoly, 1)[..., 0] for f in self.fields], axis=-1)


@dataclass
class Workload:
    name: str
    tree: QuadTree
    fvals: np.ndarray  # (nboxes, npbox, nd)
    nd: int
    norder: int
    ipoly: int
    delta: float
    eps: float
    ifpgh: int
    iperiod: int
    density: Callable
    description: str

    @property
    def npts(self) -> int:
        return self.tree.nleaves * self.norder**2


def _c1(scale: str):
    rng = np.random.default_rng(1)
    g = gaussian_sum(rng, 3, 0.05, 0.1)
    t = build_uniform(4)
    return t, g, 1, 8, 1e-3, 1e-6, 1, "uniform level-4 tree (256 leaves), n=8 Legendre, nd=1, delta=1e-3, eps=1e-6, pot"


def _c2(scale: str):
    rng = np.random.default_rng(2)
    c = rng.uniform(-0.5, 0.5, size=(5, 2))
    w = np.r_[0.005, rng.uniform(0.05, 0.1, size=4)]
    a = rng.uniform(-1, 1, size=5)
    g = GaussianSum(c, w, a)
    # "mixed": a level-4 floor (the cutoff level of delta = 1e-4 at eps = 1e-9),
    # so that the leaves away from the sharp Gaussian are near-field targets,
    # and a wider disc of level-10 leaves around it (~20k leaves in all)
    floor = {"full": 7, "mixed": 4, "small": 4, "tiny": 3}[scale]
    maxl = {"full": 10, "mixed": 10, "small": 7, "tiny": 5}[scale]
    rad = {"full": 0.03, "mixed": 0.075, "small": 0.03, "tiny": 0.02}[scale]
    # the sharp Gaussian is refined to the deepest level within 6 widths of its centre
    t = build_adaptive(g, 1, 16, 0, 1e-9, maxl, min_level=floor, discs=[(c[0, 0], c[0, 1], rad, maxl)])
    return t, g, 1, 16, 1e-4, 1e-9, 2, (
        f"adaptive periodic tree: tail rule (tol 1e-9), uniform level-{floor} floor, levels up to {maxl} "
        f"within {rad} of the sharp Gaussian; n=16 Legendre, nd=1, delta=1e-4, eps=1e-9, pot+grad")


def _c3(scale: str, delta=1e-3, eps=1e-9):
    rng = np.random.default_rng(3)
    spots = []
    for _ in range(2):
        c = rng.uniform(-0.5, 0.5, size=(20, 2))
        spots.append(GaussianSum(c, np.full(20, 0.02), np.ones(20), clip=(0.0, 1.0)))
    g3 = gaussian_sum(rng, 5, 0.05, 0.1)
    # zero-mean: subtract the mean of the 3x3 image sum over the cell
    g3.offset = -float(np.sum(g3.amps * np.pi * g3.widths**2))
    dens = Stack([spots[0], spots[1], g3])
    maxl = {"full": 9, "small": 6, "tiny": 5}[scale]
    t = build_adaptive(dens, 3, 8, 0, 1e-9, maxl, min_level=2, eta=1.0)
    return t, dens, 3, 8, delta, eps, 3, (
        f"nd=3 (two spot fields, one zero-mean Gaussian sum) on an adaptive tree (max level {maxl}), "
        f"n=8 Legendre, delta={delta:g}, eps={eps:g}, pot+grad+hess")


def _c4(scale: str):
    rng = np.random.default_rng(4)
    g1 = gaussian_sum(rng, 4, 0.05, 0.1)
    g2 = gaussian_sum(rng, 4, 0.05, 0.1)
    lev = {"full": 6, "small": 4, "tiny": 3}[scale]
    t = build_uniform(lev)
    return t, Stack([g1, g2]), 2, 8, 1e-3, 1e-9, 1, f"uniform level-{lev} tree, n=8, nd=2, delta=1e-3, eps=1e-9, pot"


def _c5(scale: str):
    rng = np.random.default_rng(5)
    g = gaussian_sum(rng, 40, 0.002, 0.1, logw=True)
    # "mixed": a level-5 floor, below the cutoff level 6 of delta = 1e-5 at
    # eps = 1e-10, so that leaves at levels 5 and 6 are near-field targets,
    # and wider discs (1.7 widths) at the deepest level (~100k leaves in all)
    floor = {"full": 8, "mixed": 5, "small": 4, "tiny": 3}[scale]
    maxl = {"full": 12, "mixed": 12, "small": 7, "tiny": 5}[scale]
    reach = 1.7 if scale == "mixed" else 1.0
    # Gaussians narrower than 0.01 are refined to the deepest level within `reach` widths
    discs = [(cx, cy, reach * wi, maxl) for (cx, cy), wi in zip(g.centers, g.widths) if wi < 0.01]
    t = build_adaptive(g, 1, 16, 0, 1e-10, maxl, min_level=floor, discs=discs)
    return t, g, 1, 16, 1e-5, 1e-10, 2, (
        f"adaptive periodic tree: tail rule (tol 1e-10), uniform level-{floor} floor, levels up to {maxl} "
        f"within {reach:g} width(s) of the Gaussians narrower than 0.01; n=16 Legendre, nd=1, delta=1e-5, "
        "eps=1e-10, pot+grad")


WORKLOADS = {"c1": _c1, "c2": _c2, "c3": _c3, "c4": _c4, "c5": _c5}


def make_workload(name: str, scale: str = "full", **kw) -> Workload:
    """Build one of the seeded workloads.  ``scale`` in {"full", "small", "tiny"} shrinks
    the tree depth for CPU-sized parity tests; densities and FGT parameters are kept.
    ``"mixed"`` (c2 and c5 only) keeps the full size but lowers the uniform floor
    below the cutoff level, so that the near field runs beside the PW stages."""
    if scale == "mixed" and name not in ("c2", "c5"):
        raise ValueError("the mixed scale exists for c2 and c5 only")
    t, dens, nd, n, delta, eps, ifpgh, desc = WORKLOADS[name](scale, **kw)
    fv = t.sample(dens, n, 0, nd)
    return Workload(name, t, fv, nd, n, 0, delta, eps, ifpgh, 1, dens, desc)
