"""Level-restricted quadtrees in the reference's flattened ``itree`` layout.

The box FGT consumes a quadtree on the square ``[-L/2, L/2]^2`` stored the
way the reference's tree builder lays it out
(common/tree_vol_coeffs_nd.f:542-549):

* ``itree[0 : 2*(nlevels+1)]`` is ``laddr(2, 0:nlevels)`` -- first and last
  (1-based) box id of each level; boxes are sorted by level;
* then ``ilevel(nboxes)``, ``iparent(nboxes)``, ``nchild(nboxes)``,
  ``ichild(4, nboxes)`` (-1 = none), ``ncoll(nboxes)``, ``coll(9, nboxes)``;
* ``iptr(1..8)`` holds the 1-based start of each of these arrays, and
  ``ltree = 2*(nlevels+1) + 17*nboxes``.

Children are ordered mm, pm, mp, pp (x sign varies fastest,
common/pts_tree_nd.f:1418-1460).  Leaf values live on an ``n x n`` tensor
grid of Legendre or Chebyshev nodes scaled to the box, x index fastest
(common/voltab2d.f:18-31).

This module is host-side set-up code: it builds trees for tests and
benchmarks (uniform trees, and adaptive trees refined by the
tail-coefficient rule of vol_tree_find_box_refine,
common/tree_vol_coeffs_nd.f:689-774 and fun_err :855-943, followed by a 2:1
level-restriction pass).  It does not evaluate the transform.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Tuple

import numpy as np

__all__ = [
    "QuadTree",
    "build_tree",
    "build_uniform",
    "build_adaptive",
    "poly_nodes",
    "values_to_coefs_matrix",
]


def poly_nodes(norder: int, ipoly: int) -> np.ndarray:
    """Ascending Gauss-Legendre (ipoly=0) or Chebyshev (ipoly=1) nodes on [-1,1]
    (common/legeexps.f:140, common/chebexps.f:12)."""
    if ipoly == 0:
        x, _ = np.polynomial.legendre.leggauss(norder)
        return np.asarray(x, dtype=np.float64)
    if ipoly == 1:
        i = np.arange(1, norder + 1)
        return np.cos((2 * i - 1) * np.pi / (2 * norder))[::-1].copy()
    raise ValueError("ipoly must be 0 (Legendre) or 1 (Chebyshev)")


def values_to_coefs_matrix(norder: int, ipoly: int) -> np.ndarray:
    """``U`` with ``coef = U @ values`` (inverse of the Vandermonde matrix at the nodes)."""
    x = poly_nodes(norder, ipoly)
    if ipoly == 0:
        V = np.polynomial.legendre.legvander(x, norder - 1)
    else:
        V = np.polynomial.chebyshev.chebvander(x, norder - 1)
    return np.linalg.inv(V)


@dataclass
class QuadTree:
    """A flattened quadtree with the reference's array layout (1-based ids inside ``itree``)."""

    nlevels: int
    itree: np.ndarray  # int32 (ltree,)
    iptr: np.ndarray  # int32 (8,), 1-based offsets
    centers: np.ndarray  # float64 (nboxes, 2)  == Fortran centers(2, nboxes)
    boxsize: np.ndarray  # float64 (nlevels+1,)
    iperiod: int = 1
    _cache: dict = field(default_factory=dict, repr=False)

    # -- sizes -------------------------------------------------------------
    @property
    def nboxes(self) -> int:
        return int(self.itree[2 * self.nlevels + 1])

    @property
    def ltree(self) -> int:
        return int(self.itree.shape[0])

    def _arr(self, k: int, width: int = 1) -> np.ndarray:
        s = int(self.iptr[k]) - 1
        return self.itree[s : s + width * self.nboxes]

    # -- views (values are 1-based box ids, -1 for none) -------------------
    @property
    def laddr(self) -> np.ndarray:
        return self.itree[: 2 * (self.nlevels + 1)].reshape(self.nlevels + 1, 2)

    @property
    def ilevel(self) -> np.ndarray:
        return self._arr(1)

    @property
    def iparent(self) -> np.ndarray:
        return self._arr(2)

    @property
    def nchild(self) -> np.ndarray:
        return self._arr(3)

    @property
    def ichild(self) -> np.ndarray:
        return self._arr(4, 4).reshape(self.nboxes, 4)

    @property
    def ncoll(self) -> np.ndarray:
        return self._arr(5)

    @property
    def coll(self) -> np.ndarray:
        return self._arr(6, 9).reshape(self.nboxes, 9)

    def level_range(self, lev: int) -> range:
        """0-based box ids of level ``lev``."""
        a, b = self.laddr[lev]
        return range(int(a) - 1, int(b))

    def leaves(self) -> np.ndarray:
        """0-based ids of the leaf boxes at levels >= 1 (and the root if it is a leaf)."""
        return np.nonzero(self.nchild == 0)[0]

    @property
    def nleaves(self) -> int:
        return int(np.count_nonzero(self.nchild == 0))

    def copy(self) -> "QuadTree":
        return QuadTree(
            self.nlevels,
            self.itree.copy(),
            self.iptr.copy(),
            self.centers.copy(),
            self.boxsize.copy(),
            self.iperiod,
        )

    def grid_points(self, norder: int, ipoly: int) -> np.ndarray:
        """Coordinates of the tensor grid of every box, shape (nboxes, norder**2, 2),
        point p = i + norder*j has x node i (x fastest)."""
        xq = poly_nodes(norder, ipoly) / 2.0
        gx = np.tile(xq, norder)
        gy = np.repeat(xq, norder)
        bs = self.boxsize[self.ilevel]
        px = self.centers[:, 0:1] + gx[None, :] * bs[:, None]
        py = self.centers[:, 1:2] + gy[None, :] * bs[:, None]
        return np.stack([px, py], axis=-1)

    def sample(self, fun: Callable[[np.ndarray], np.ndarray], norder: int, ipoly: int, nd: int) -> np.ndarray:
        """Tabulate ``fun(points (..., 2)) -> (..., nd)`` on every box grid.
        Returns fvals of shape (nboxes, norder**2, nd) == Fortran fvals(nd, npbox, nboxes).
        Values of non-leaf boxes are filled too (the transform ignores them)."""
        pts = self.grid_points(norder, ipoly)
        v = np.asarray(fun(pts), dtype=np.float64)
        if v.ndim == 2:
            v = v[..., None]
        assert v.shape == (self.nboxes, norder * norder, nd), v.shape
        return np.ascontiguousarray(v)


# ---------------------------------------------------------------------------
# construction
# ---------------------------------------------------------------------------

Key = Tuple[int, int, int]  # (level, ix, iy)


def _center(key: Key, boxlen: float) -> Tuple[float, float]:
    l, ix, iy = key
    h = boxlen / (1 << l)
    return (-boxlen / 2 + (ix + 0.5) * h, -boxlen / 2 + (iy + 0.5) * h)


def _children(key: Key) -> List[Key]:
    l, ix, iy = key
    # mm, pm, mp, pp
    return [(l + 1, 2 * ix, 2 * iy), (l + 1, 2 * ix + 1, 2 * iy), (l + 1, 2 * ix, 2 * iy + 1), (l + 1, 2 * ix + 1, 2 * iy + 1)]


def _level_restrict(refined: set, leaves: set, iperiod: int) -> None:
    """Refine leaves until neighbouring leaves differ by at most one level (2:1 balance,
    periodic wrap when ``iperiod``), mutating ``refined``/``leaves`` in place."""
    changed = True
    while changed:
        changed = False
        by_level: Dict[int, List[Key]] = {}
        for k in leaves:
            by_level.setdefault(k[0], []).append(k)
        for lev in sorted(by_level.keys(), reverse=True):
            if lev < 2:
                continue
            n = 1 << lev
            for (l, ix, iy) in by_level[lev]:
                if (l, ix, iy) not in leaves:
                    continue
                for dx in (-1, 0, 1):
                    for dy in (-1, 0, 1):
                        if dx == 0 and dy == 0:
                            continue
                        jx, jy = ix + dx, iy + dy
                        if iperiod:
                            jx %= n
                            jy %= n
                        elif not (0 <= jx < n and 0 <= jy < n):
                            continue
                        # covering leaf of the neighbour cell
                        for lc in range(l, -1, -1):
                            key = (lc, jx >> (l - lc), jy >> (l - lc))
                            if key in leaves:
                                if lc < l - 1:
                                    leaves.discard(key)
                                    refined.add(key)
                                    for c in _children(key):
                                        leaves.add(c)
                                    changed = True
                                break
                            if key in refined:
                                break


def build_tree(refined: set, boxlen: float = 1.0, iperiod: int = 1) -> QuadTree:
    """Flatten a set of refined boxes (keys ``(level, ix, iy)``; the root is ``(0,0,0)``)
    into the reference layout.  Every refined box gets four children."""
    levels: List[List[Key]] = [[(0, 0, 0)]]
    while True:
        nxt = []
        for k in levels[-1]:
            if k in refined:
                nxt.extend(_children(k))
        if not nxt:
            break
        levels.append(nxt)
    nlevels = len(levels) - 1
    nboxes = sum(len(v) for v in levels)
    ids: Dict[Key, int] = {}
    order: List[Key] = []
    for lv in levels:
        for k in lv:
            ids[k] = len(order)
            order.append(k)
    ltree = 2 * (nlevels + 1) + 17 * nboxes
    itree = np.full(ltree, -1, dtype=np.int32)
    iptr = np.zeros(8, dtype=np.int32)
    iptr[0] = 1
    iptr[1] = 2 * (nlevels + 1) + 1
    iptr[2] = iptr[1] + nboxes
    iptr[3] = iptr[2] + nboxes
    iptr[4] = iptr[3] + nboxes
    iptr[5] = iptr[4] + 4 * nboxes
    iptr[6] = iptr[5] + nboxes
    iptr[7] = iptr[6] + 9 * nboxes
    pos = 0
    for lev, lv in enumerate(levels):
        itree[2 * lev] = pos + 1
        itree[2 * lev + 1] = pos + len(lv)
        pos += len(lv)
    centers = np.zeros((nboxes, 2))
    boxsize = np.array([boxlen / (1 << l) for l in range(nlevels + 1)])
    o = lambda k: int(iptr[k]) - 1  # noqa: E731
    for b, k in enumerate(order):
        centers[b] = _center(k, boxlen)
        itree[o(1) + b] = k[0]
        if k[0] > 0:
            pk = (k[0] - 1, k[1] >> 1, k[2] >> 1)
            itree[o(2) + b] = ids[pk] + 1
        else:
            itree[o(2) + b] = -1
        if k in refined:
            itree[o(3) + b] = 4
            for c, ck in enumerate(_children(k)):
                itree[o(4) + 4 * b + c] = ids[ck] + 1
        else:
            itree[o(3) + b] = 0
    t = QuadTree(nlevels, itree, iptr, centers, boxsize, iperiod)
    compute_colleagues(t, iperiod)
    return t


def compute_colleagues(t: QuadTree, iperiod: int) -> None:
    """Same-level touching boxes, periodic wrap optional (rule of computecoll,
    common/tree_routs_nd.f:106-242): candidates are the children of the
    parent's colleagues, kept in that order."""
    ncoll = t.ncoll
    coll = t.coll
    ncoll[:] = 0
    coll[:] = -1
    ncoll[0] = 1
    coll[0, 0] = 1
    bs0 = t.boxsize[0]
    ich = t.ichild
    c = t.centers
    for lev in range(1, t.nlevels + 1):
        B = np.arange(t.level_range(lev).start, t.level_range(lev).stop)
        dads = t.iparent[B] - 1
        C = coll[dads]  # (nB, 9), 1-based or -1
        kids = np.where(C[:, :, None] > 0, ich[np.maximum(C, 1) - 1], -1).reshape(len(B), 36)
        valid = kids > 0
        kk = np.maximum(kids, 1) - 1
        d = np.abs(c[kk] - c[B][:, None, :])
        if iperiod == 1:
            d = np.minimum(d, bs0 - d)
        valid &= np.all(d <= 1.05 * t.boxsize[lev], axis=-1)
        order = np.argsort(~valid, axis=1, kind="stable")[:, :9]
        packed = np.take_along_axis(np.where(valid, kids, -1), order, axis=1)
        coll[B] = packed
        ncoll[B] = valid.sum(axis=1)


def build_uniform(nlevels: int, boxlen: float = 1.0, iperiod: int = 1) -> QuadTree:
    """Uniform tree: every box of levels 0..nlevels-1 refined (4**nlevels leaves)."""
    refined = set()
    for l in range(nlevels):
        n = 1 << l
        for ix in range(n):
            for iy in range(n):
                refined.add((l, ix, iy))
    return build_tree(refined, boxlen, iperiod)


def _tail_mask(norder: int) -> np.ndarray:
    """rmask of tens_prod_get_rmask with iptype=2 (common/polytens.f:1579-1621):
    coefficient (p1,p2) is in the tail when (p1+1)^2 + (p2+1)^2 >= norder^2."""
    p = np.arange(norder)
    m = ((p[None, :] + 1) ** 2 + (p[:, None] + 1) ** 2) >= norder**2
    return m.reshape(-1).astype(np.float64)  # index p1 + norder*p2


def build_adaptive(
    fun: Callable[[np.ndarray], np.ndarray],
    nd: int,
    norder: int,
    ipoly: int,
    eps: float,
    max_level: int,
    boxlen: float = 1.0,
    iperiod: int = 1,
    eta: float = 1.0,
    min_level: int = 0,
) -> QuadTree:
    """Adaptive tree: a box is refined when the scaled L2 norm of the tail of its
    tensor-product coefficients exceeds ``eps * rint`` (the iptype=2 rule of
    vol_tree_find_box_refine, common/tree_vol_coeffs_nd.f:689-774, with the
    function's L2 norm as ``rint``), up to ``max_level``; then 2:1 balanced."""
    U = values_to_coefs_matrix(norder, ipoly)
    U2 = np.kron(U, U)  # acts on p = i + n*j with x fastest: kron(Uy, Ux)
    mask = _tail_mask(norder)
    rsum = np.sqrt(mask.sum())
    xq = poly_nodes(norder, ipoly) / 2.0
    gx = np.tile(xq, norder)
    gy = np.repeat(xq, norder)

    # L2 norm of the density over the unit cell (fine midpoint sampling)
    m = 512
    s = (np.arange(m) + 0.5) / m * boxlen - boxlen / 2
    X, Y = np.meshgrid(s, s, indexing="xy")
    fv = np.asarray(fun(np.stack([X, Y], axis=-1)), dtype=np.float64)
    if fv.ndim == 2:
        fv = fv[..., None]
    rint = np.sqrt(np.max(np.sum(fv**2, axis=(0, 1)) * (boxlen / m) ** 2))
    rsc = np.sqrt(1.0 / boxlen**2) * max(rint, 1e-300)

    refined: set = set()
    leaves = {(0, 0, 0)}
    frontier = [(0, 0, 0)]
    for lev in range(0, max_level):
        if not frontier:
            break
        h = boxlen / (1 << lev)
        keys = frontier
        cx = np.array([_center(k, boxlen)[0] for k in keys])
        cy = np.array([_center(k, boxlen)[1] for k in keys])
        pts = np.stack([cx[:, None] + gx[None, :] * h, cy[:, None] + gy[None, :] * h], axis=-1)
        v = np.asarray(fun(pts), dtype=np.float64)
        if v.ndim == 2:
            v = v[..., None]
        coefs = np.einsum("pq,bqd->bpd", U2, v)
        err = np.sqrt(np.einsum("p,bpd->bd", mask, coefs**2)).max(axis=1) / rsum
        err = err * (h / 2.0) ** eta
        new = []
        for k, e in zip(keys, err):
            if e > eps * rsc or lev < min_level:
                refined.add(k)
                leaves.discard(k)
                for c in _children(k):
                    leaves.add(c)
                    new.append(c)
        frontier = new
    _level_restrict(refined, leaves, iperiod)
    return build_tree(refined, boxlen, iperiod)
