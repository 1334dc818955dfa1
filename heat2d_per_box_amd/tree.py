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
common/pts_tree_nd.f:1418-1460), and the children of a level are listed in
the order of their parents (vol_tree_refine_boxes,
common/tree_vol_coeffs_nd.f:780-850).  Leaf values live on an ``n x n``
tensor grid of Legendre or Chebyshev nodes scaled to the box, x index fastest
(common/voltab2d.f:18-31).

This module is host-side set-up code for tests and benchmarks: uniform
trees, and adaptive trees refined by the tail-coefficient rule of
vol_tree_find_box_refine (common/tree_vol_coeffs_nd.f:689-774, fun_err
:855-943) and/or around given discs, then 2:1 level-restricted.  All steps
are vectorised over the boxes of a level.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable, Dict, Iterable, List, Optional, Sequence, Tuple

import numpy as np

__all__ = [
    "QuadTree",
    "build_tree",
    "build_uniform",
    "build_adaptive",
    "poly_nodes",
    "values_to_coefs_matrix",
    "compute_colleagues",
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


def box_grid(centers: np.ndarray, size, norder: int, ipoly: int) -> np.ndarray:
    """Tensor-grid points of boxes with the given centres (m, 2) and side(s):
    shape (m, norder**2, 2), point p = i + norder*j has x node i."""
    xq = poly_nodes(norder, ipoly) / 2.0
    gx = np.tile(xq, norder)
    gy = np.repeat(xq, norder)
    s = np.broadcast_to(np.asarray(size, dtype=np.float64), (centers.shape[0],))
    px = centers[:, 0:1] + gx[None, :] * s[:, None]
    py = centers[:, 1:2] + gy[None, :] * s[:, None]
    return np.stack([px, py], axis=-1)


def eval_on_boxes(fun, centers: np.ndarray, size, norder: int, ipoly: int, nd: int) -> np.ndarray:
    """Values (m, norder**2, nd) of ``fun`` on the grids of the given boxes.  A density
    object with a ``sample_boxes`` method evaluates itself with culling."""
    if hasattr(fun, "sample_boxes"):
        v = fun.sample_boxes(centers, size, norder, ipoly)
    else:
        v = np.asarray(fun(box_grid(centers, size, norder, ipoly)), dtype=np.float64)
    if v.ndim == 2:
        v = v[..., None]
    assert v.shape == (centers.shape[0], norder * norder, nd), (v.shape, nd)
    return v


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

    @property
    def nboxes(self) -> int:
        return int(self.itree[2 * self.nlevels + 1])

    @property
    def ltree(self) -> int:
        return int(self.itree.shape[0])

    def _arr(self, k: int, width: int = 1) -> np.ndarray:
        s = int(self.iptr[k]) - 1
        return self.itree[s : s + width * self.nboxes]

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
        """0-based ids of the leaf boxes."""
        return np.nonzero(self.nchild == 0)[0]

    @property
    def nleaves(self) -> int:
        return int(np.count_nonzero(self.nchild == 0))

    def copy(self) -> "QuadTree":
        return QuadTree(self.nlevels, self.itree.copy(), self.iptr.copy(), self.centers.copy(), self.boxsize.copy(),
                        self.iperiod)

    def grid_points(self, norder: int, ipoly: int) -> np.ndarray:
        """Coordinates of the tensor grid of every box, shape (nboxes, norder**2, 2)."""
        return box_grid(self.centers, self.boxsize[self.ilevel], norder, ipoly)

    def sample(self, fun, norder: int, ipoly: int, nd: int, leaves_only: bool = True) -> np.ndarray:
        """fvals (nboxes, norder**2, nd) == Fortran fvals(nd, npbox, nboxes) of ``fun`` on the
        leaf grids (non-leaf boxes are zero: the transform ignores them)."""
        out = np.zeros((self.nboxes, norder * norder, nd))
        sel = self.leaves() if leaves_only else np.arange(self.nboxes)
        lev = self.ilevel[sel]
        for l in np.unique(lev):
            b = sel[lev == l]
            out[b] = eval_on_boxes(fun, self.centers[b], self.boxsize[l], norder, ipoly, nd)
        return out


# ---------------------------------------------------------------------------
# construction (boxes of one level are integer cells (ix, iy), key ix<<32 | iy)
# ---------------------------------------------------------------------------

def _key(c: np.ndarray) -> np.ndarray:
    return (c[:, 0].astype(np.int64) << 32) | c[:, 1].astype(np.int64)


def _unkey(k: np.ndarray) -> np.ndarray:
    return np.stack([(k >> 32).astype(np.int64), (k & 0xFFFFFFFF).astype(np.int64)], axis=1)


def _children(c: np.ndarray) -> np.ndarray:
    """Children of cells (m, 2) in the order mm, pm, mp, pp per parent: (4m, 2)."""
    off = np.array([[0, 0], [1, 0], [0, 1], [1, 1]], dtype=np.int64)
    return (2 * c[:, None, :] + off[None, :, :]).reshape(-1, 2)


def _centers(c: np.ndarray, lev: int, boxlen: float) -> np.ndarray:
    h = boxlen / (1 << lev)
    return -boxlen / 2 + (c.astype(np.float64) + 0.5) * h


def _level_restrict(leaves: Dict[int, np.ndarray], refined: Dict[int, np.ndarray], iperiod: int) -> None:
    """Refine leaves until touching leaves differ by at most one level (2:1 balance,
    periodic wrap when ``iperiod``).  ``leaves``/``refined`` map level -> sorted keys."""
    dirs = np.array([[dx, dy] for dx in (-1, 0, 1) for dy in (-1, 0, 1) if dx or dy], dtype=np.int64)
    changed = True
    while changed:
        changed = False
        for lev in sorted(leaves.keys(), reverse=True):
            if lev < 2 or leaves[lev].size == 0:
                continue
            n = 1 << lev
            nb = (_unkey(leaves[lev])[:, None, :] + dirs[None, :, :]).reshape(-1, 2)
            if iperiod:
                nb %= n
            else:
                nb = nb[np.all((nb >= 0) & (nb < n), axis=1)]
            for c in range(lev - 2, -1, -1):
                if c not in leaves or leaves[c].size == 0:
                    continue
                anc = np.unique(_key(nb >> (lev - c)))
                hit = np.intersect1d(anc, leaves[c], assume_unique=True)
                if hit.size:
                    leaves[c] = np.setdiff1d(leaves[c], hit, assume_unique=True)
                    refined[c] = np.union1d(refined.get(c, np.zeros(0, np.int64)), hit)
                    kids = np.sort(_key(_children(_unkey(hit))))
                    leaves[c + 1] = np.union1d(leaves.get(c + 1, np.zeros(0, np.int64)), kids)
                    changed = True


def _flatten(refined: Dict[int, np.ndarray], boxlen: float, iperiod: int) -> QuadTree:
    """Flatten refined-cell sets into the reference layout (children listed in parent order)."""
    level_cells: List[np.ndarray] = [np.zeros((1, 2), dtype=np.int64)]
    while True:
        lev = len(level_cells) - 1
        cur = level_cells[-1]
        ref = refined.get(lev)
        if ref is None or ref.size == 0:
            break
        mask = np.isin(_key(cur), ref)
        if not mask.any():
            break
        level_cells.append(_children(cur[mask]))
    nlevels = len(level_cells) - 1
    counts = [c.shape[0] for c in level_cells]
    starts = np.concatenate([[0], np.cumsum(counts)])
    nboxes = int(starts[-1])
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
    for lev in range(nlevels + 1):
        itree[2 * lev] = starts[lev] + 1
        itree[2 * lev + 1] = starts[lev + 1]
    centers = np.zeros((nboxes, 2))
    boxsize = np.array([boxlen / (1 << l) for l in range(nlevels + 1)])
    o = [int(v) - 1 for v in iptr]
    ilevel = itree[o[1] : o[1] + nboxes]
    iparent = itree[o[2] : o[2] + nboxes]
    nchild = itree[o[3] : o[3] + nboxes]
    ichild = itree[o[4] : o[4] + 4 * nboxes].reshape(nboxes, 4)
    for lev, cells in enumerate(level_cells):
        s0, s1 = starts[lev], starts[lev + 1]
        centers[s0:s1] = _centers(cells, lev, boxlen)
        ilevel[s0:s1] = lev
        nchild[s0:s1] = 0
        if lev < nlevels:
            mask = np.isin(_key(cells), refined.get(lev, np.zeros(0, np.int64)))
            par = np.nonzero(mask)[0] + s0  # 0-based parent ids, in order
            kid0 = starts[lev + 1]
            nchild[par] = 4
            ichild[par] = (kid0 + np.arange(4 * par.size).reshape(-1, 4) + 1).astype(np.int32)
            iparent[kid0 : kid0 + 4 * par.size] = np.repeat(par + 1, 4)
    iparent[0] = -1
    t = QuadTree(nlevels, itree, iptr, centers, boxsize, iperiod)
    compute_colleagues(t, iperiod)
    return t


def compute_colleagues(t: QuadTree, iperiod: int) -> None:
    """Same-level touching boxes, periodic wrap optional (rule of computecoll,
    common/tree_routs_nd.f:106-242): candidates are the children of the parent's
    colleagues, kept in that order."""
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
        r = t.level_range(lev)
        B = np.arange(r.start, r.stop)
        dads = t.iparent[B] - 1
        C = coll[dads]
        kids = np.where(C[:, :, None] > 0, ich[np.maximum(C, 1) - 1], -1).reshape(len(B), 36)
        valid = kids > 0
        kk = np.maximum(kids, 1) - 1
        d = np.abs(c[kk] - c[B][:, None, :])
        if iperiod == 1:
            d = np.minimum(d, bs0 - d)
        valid &= np.all(d <= 1.05 * t.boxsize[lev], axis=-1)
        order = np.argsort(~valid, axis=1, kind="stable")[:, :9]
        coll[B] = np.take_along_axis(np.where(valid, kids, -1), order, axis=1)
        ncoll[B] = valid.sum(axis=1)


def with_capacity(t: QuadTree, nbmax: int) -> QuadTree:
    """The same tree laid out for ``nbmax`` boxes (itree and centers sized for
    them, iptr offsets accordingly), the way callers allocate room for a tree
    that may grow (vol_tree_mem sizes, common/tree_vol_coeffs_nd.f:542-549)."""
    nb, nl = t.nboxes, t.nlevels
    if nbmax < nb:
        raise ValueError("capacity below the box count")
    ltree = 2 * (nl + 1) + 17 * nbmax
    itree = np.full(ltree, -1, dtype=np.int32)
    iptr = np.zeros(8, dtype=np.int32)
    iptr[0] = 1
    iptr[1] = 2 * (nl + 1) + 1
    widths = (1, 1, 1, 4, 1, 9)  # ilevel, iparent, nchild, ichild, ncoll, coll
    for k in range(2, 8):
        iptr[k] = iptr[k - 1] + widths[k - 2] * nbmax
    itree[: 2 * (nl + 1)] = t.itree[: 2 * (nl + 1)]
    for k, w in zip(range(1, 7), widths):
        s0, d0 = int(t.iptr[k]) - 1, int(iptr[k]) - 1
        itree[d0 : d0 + w * nb] = t.itree[s0 : s0 + w * nb]
    centers = np.zeros((nbmax, 2))
    centers[:nb] = t.centers[:nb]
    return QuadTree(nl, itree, iptr, centers, t.boxsize.copy(), t.iperiod)


def build_tree(refined: Iterable[Tuple[int, int, int]], boxlen: float = 1.0, iperiod: int = 1) -> QuadTree:
    """Tree from an explicit set of refined boxes ``(level, ix, iy)`` (root ``(0,0,0)``)."""
    by: Dict[int, list] = {}
    for (l, ix, iy) in refined:
        by.setdefault(l, []).append((ix, iy))
    ref = {l: np.sort(_key(np.array(v, dtype=np.int64))) for l, v in by.items()}
    return _flatten(ref, boxlen, iperiod)


def build_uniform(nlevels: int, boxlen: float = 1.0, iperiod: int = 1) -> QuadTree:
    """Uniform tree: every box of levels 0..nlevels-1 refined (4**nlevels leaves)."""
    ref = {}
    for l in range(nlevels):
        n = 1 << l
        ix, iy = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
        ref[l] = np.sort(_key(np.stack([ix.ravel(), iy.ravel()], axis=1)))
    return _flatten(ref, boxlen, iperiod)


def _tail_mask(norder: int) -> np.ndarray:
    """rmask of tens_prod_get_rmask with iptype=2 (common/polytens.f:1579-1621):
    coefficient (p1,p2) is in the tail when (p1+1)^2 + (p2+1)^2 >= norder^2."""
    p = np.arange(norder)
    m = ((p[None, :] + 1) ** 2 + (p[:, None] + 1) ** 2) >= norder**2
    return m.reshape(-1).astype(np.float64)


def build_adaptive(
    fun,
    nd: int,
    norder: int,
    ipoly: int,
    eps: float,
    max_level: int,
    boxlen: float = 1.0,
    iperiod: int = 1,
    eta: float = 1.0,
    min_level: int = 0,
    discs: Sequence[Tuple[float, float, float, int]] = (),
    rint: Optional[float] = None,
) -> QuadTree:
    """Adaptive tree.  A box is refined when

    * its level is below ``min_level`` (uniform floor), or
    * the scaled L2 norm of the tail of its tensor-product coefficients exceeds
      ``eps * rint`` (the iptype=2 rule of vol_tree_find_box_refine,
      common/tree_vol_coeffs_nd.f:689-774, with ``rint`` the density's L2 norm), or
    * it meets one of the ``discs`` (cx, cy, radius, level): boxes closer than
      ``radius`` to the centre (periodic distance) are refined down to ``level``,

    up to ``max_level``; the result is 2:1 level-restricted."""
    U = values_to_coefs_matrix(norder, ipoly)
    U2 = np.kron(U, U)
    mask = _tail_mask(norder)
    rsum = np.sqrt(mask.sum())
    if rint is None:
        m = 256
        s = (np.arange(m) + 0.5) / m * boxlen - boxlen / 2
        X, Y = np.meshgrid(s, s, indexing="xy")
        fv = np.asarray(fun(np.stack([X, Y], axis=-1)), dtype=np.float64)
        if fv.ndim == 2:
            fv = fv[..., None]
        rint = float(np.sqrt(np.max(np.sum(fv**2, axis=(0, 1)) * (boxlen / m) ** 2)))
    rsc = np.sqrt(1.0 / boxlen**2) * max(rint, 1e-300)

    leaves: Dict[int, np.ndarray] = {}
    refined: Dict[int, np.ndarray] = {}
    frontier = np.zeros((1, 2), dtype=np.int64)
    for lev in range(0, max_level + 1):
        if frontier.shape[0] == 0:
            break
        h = boxlen / (1 << lev)
        if lev == max_level:
            leaves[lev] = np.sort(_key(frontier))
            break
        dec = np.zeros(frontier.shape[0], dtype=bool)
        if lev < min_level:
            dec[:] = True
        else:
            cen = _centers(frontier, lev, boxlen)
            for (cx, cy, rad, dl) in discs:
                if lev < dl:
                    d = np.abs(cen - np.array([cx, cy]))
                    if iperiod:
                        d = np.minimum(d, boxlen - d)
                    d = np.maximum(d - h / 2, 0.0)
                    dec |= np.hypot(d[:, 0], d[:, 1]) < rad
            todo = np.nonzero(~dec)[0]
            for s in range(0, todo.size, 4096):
                idx = todo[s : s + 4096]
                v = eval_on_boxes(fun, cen[idx], h, norder, ipoly, nd)
                coefs = np.einsum("pq,bqd->bpd", U2, v)
                err = np.sqrt(np.einsum("p,bpd->bd", mask, coefs**2)).max(axis=1) / rsum * (h / 2.0) ** eta
                dec[idx] = err > eps * rsc
        refined[lev] = np.sort(_key(frontier[dec]))
        leaves[lev] = np.sort(_key(frontier[~dec]))
        frontier = _children(frontier[dec])
    _level_restrict(leaves, refined, iperiod)
    return _flatten(refined, boxlen, iperiod)
