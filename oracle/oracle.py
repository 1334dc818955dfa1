"""ctypes front end of the checkers.  TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg
import this module.  It wraps

* ``_build/liboracle.so`` -- the C restatement of the reference boxfgt
  (boxfgt_oracle.c), and
* ``_ref/libref_boxfgt.so`` -- the reference Fortran ``boxfgt`` itself,
  compiled from its sources by ``make -C oracle ref`` (present only where the
  reference sources were available at build time).

Array conventions follow the Fortran ones with C order reversed:
fvals/pot ``(nboxes, npbox, nd)``, grad ``(nboxes, npbox, 2, nd)``, hess
``(nboxes, npbox, 3, nd)``.  Both entry points mutate a copy of ``itree``
(the colleague lists are recomputed) and return it.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref_boxfgt.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)

_oracle = None
_ref = None


def build(with_reference: Optional[bool] = None) -> None:
    """Compile the oracle (and the reference when its sources exist)."""
    import subprocess

    targets = ["all"]
    if with_reference is None:
        with_reference = os.path.isdir(os.environ.get("REF_DIR", "/root/reference"))
    if with_reference:
        targets += ["ref", "initpot"]
    subprocess.run(["make", "-s", "-C", HERE, "-j8"] + targets, check=True)


def load_oracle() -> ctypes.CDLL:
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle library missing: {ORACLE_SO} (run `make -C oracle`)")
        lib = ctypes.CDLL(ORACLE_SO)
        lib.oracle_boxfgt.restype = ctypes.c_int
        lib.oracle_boxfgt.argtypes = [
            ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
            ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip, _ip, _dp, _dp, _dp, ctypes.c_int, _dp, _dp, _dp, _dp, _ip,
        ]
        lib.oracle_boxfgt_nt.restype = ctypes.c_int
        lib.oracle_boxfgt_nt.argtypes = [
            ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
            ctypes.c_int, ctypes.c_int, _ip, _ip, _ip, _dp, _dp, _dp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int,
            _dp, _ip,
        ]
        lib.oracle_eval_targets.restype = ctypes.c_int
        lib.oracle_eval_targets.argtypes = [
            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip, _ip, _dp, _dp, ctypes.c_int, _dp,
            ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp,
        ]
        for name in ("oracle_legeexps", "oracle_chebexps"):
            f = getattr(lib, name)
            f.restype = None
            f.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        lib.oracle_fgtpwterms.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp, _ip]
        lib.oracle_pwnodes.argtypes = [ctypes.c_double, ctypes.c_int, _dp, _dp]
        lib.oracle_periodic_pwnodes.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp, _dp]
        lib.oracle_sph_besselj.argtypes = [ctypes.c_int, ctypes.c_double, _dp]
        lib.oracle_poly2pw.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_double,
                                       ctypes.c_double, _dp]
        lib.oracle_pw2pgh.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_double, ctypes.c_double,
                                      _dp, _dp, _dp]
        lib.oracle_tree_fun_eval.restype = None
        lib.oracle_loctab_level.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                            ctypes.c_double, _dp, _dp, _dp, _ip]
        _oracle = lib
    return _oracle


def reference_available() -> bool:
    return os.path.exists(REF_SO)


def load_reference() -> ctypes.CDLL:
    global _ref
    if _ref is None:
        if not reference_available():
            raise RuntimeError(f"reference library missing: {REF_SO} (run `make -C oracle ref`)")
        _ref = ctypes.CDLL(REF_SO, mode=os.RTLD_LAZY | os.RTLD_LOCAL)
        _ref.boxfgt_.restype = None
    return _ref


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


@dataclass
class FGTResult:
    pot: np.ndarray
    grad: np.ndarray
    hess: np.ndarray
    timeinfo: np.ndarray
    itree: np.ndarray
    info: Optional[dict] = None
    tree: Optional[object] = None       # the tree after ifnewtree=1
    pote: Optional[np.ndarray] = None   # (ntarg, nd)
    grade: Optional[np.ndarray] = None  # (ntarg, 2, nd)
    hesse: Optional[np.ndarray] = None  # (ntarg, 3, nd)


def _alloc(nboxes, npbox, nd):
    return (np.zeros((nboxes, npbox, nd)), np.zeros((nboxes, npbox, 2, nd)), np.zeros((nboxes, npbox, 3, nd)))


def oracle_eval_targets(tree, itree, nd, ipoly, norder, targs, ifpghtarg, pot, grad, hess):
    """Leaf arrays interpolated at targets (ntarg, 2): (pote, grade, hesse)."""
    lib = load_oracle()
    iptr = np.ascontiguousarray(tree.iptr, dtype=np.int32)
    centers = np.ascontiguousarray(tree.centers, dtype=np.float64)
    boxsize = np.ascontiguousarray(tree.boxsize, dtype=np.float64)
    tg = np.ascontiguousarray(targs, dtype=np.float64)
    nt = tg.shape[0]
    pote, grade, hesse = np.zeros((nt, nd)), np.zeros((nt, 2, nd)), np.zeros((nt, 3, nd))
    g = np.ascontiguousarray(grad) if grad is not None else np.zeros(1)
    h = np.ascontiguousarray(hess) if hess is not None else np.zeros(1)
    rc = lib.oracle_eval_targets(nd, ipoly, norder, tree.nlevels, _p(itree, _ip), _p(iptr, _ip), _p(centers),
                                 _p(boxsize), nt, _p(tg), ifpghtarg, _p(np.ascontiguousarray(pot)), _p(g), _p(h),
                                 _p(pote), _p(grade), _p(hesse))
    if rc != 0:
        raise RuntimeError(f"oracle_eval_targets failed with status {rc}")
    return pote, (grade if ifpghtarg >= 2 else None), (hesse if ifpghtarg >= 3 else None)


def oracle_boxfgt(tree, fvals, nd, delta, eps, ipoly, iperiod, norder, ifpgh, targs=None, ifpghtarg=0,
                  pot0=None) -> FGTResult:
    """Run the C restatement on a copy of the tree arrays.

    With targets, follows boxfgt's dispatch (bfgt/boxfgt.f:207-277): when
    ifpghtarg > ifpgh the transform runs with ifpghtarg on zeroed scratch
    arrays for the extra channels, which the caller's grad/hess never see;
    the targets then interpolate pot (accumulated onto pot0) and the
    gradient/Hessian arrays the transform wrote (bfgt/boxfgt.f:1494-1563)."""
    if targs is not None and ifpghtarg > 0:
        ifpgh2 = max(ifpgh, ifpghtarg)
        r = oracle_boxfgt(tree, fvals, nd, delta, eps, ipoly, iperiod, norder, ifpgh2)
        if pot0 is not None:
            r.pot += pot0
        r.pote, r.grade, r.hesse = oracle_eval_targets(tree, r.itree, nd, ipoly, norder, targs, ifpghtarg, r.pot,
                                                       r.grad, r.hess)
        if ifpgh < 2:
            r.grad = np.zeros_like(r.grad)
        if ifpgh < 3:
            r.hess = np.zeros_like(r.hess)
        return r
    lib = load_oracle()
    itree = np.ascontiguousarray(tree.itree, dtype=np.int32).copy()
    iptr = np.ascontiguousarray(tree.iptr, dtype=np.int32)
    centers = np.ascontiguousarray(tree.centers, dtype=np.float64)
    boxsize = np.ascontiguousarray(tree.boxsize, dtype=np.float64)
    nboxes = centers.shape[0]
    npbox = norder * norder
    fv = np.ascontiguousarray(fvals, dtype=np.float64)
    assert fv.shape == (nboxes, npbox, nd)
    pot, grad, hess = _alloc(nboxes, npbox, nd)
    ti = np.zeros(20)
    info = np.zeros(8, dtype=np.int32)
    rc = lib.oracle_boxfgt(nd, delta, eps, ipoly, iperiod, norder, npbox, nboxes, tree.nlevels, _p(itree, _ip),
                           _p(iptr, _ip), _p(centers), _p(boxsize), _p(fv), ifpgh, _p(pot), _p(grad), _p(hess),
                           _p(ti), _p(info, _ip))
    if rc != 0:
        raise RuntimeError(f"oracle_boxfgt failed with status {rc}")
    keys = ["npwlevel", "npw", "nlevstart", "nlevend", "nleafbox", "npairs"]
    return FGTResult(pot, grad, hess, ti, itree, {k: int(v) for k, v in zip(keys, info)})


def _adapted_tree(tree, itree, nlevels, centers):
    from heat2d_per_box_amd.tree import QuadTree

    return QuadTree(int(nlevels), itree, np.ascontiguousarray(tree.iptr, dtype=np.int32).copy(), centers,
                    np.ascontiguousarray(tree.boxsize, dtype=np.float64).copy(), tree.iperiod)


def oracle_boxfgt_newtree(tree, fvals, nd, delta, eps, ipoly, iperiod, norder, ifpgh) -> FGTResult:
    """boxfgt with ifnewtree=1 in the C restatement, on copies of the tree
    arrays laid out for ``centers.shape[0]`` boxes (the capacity): the
    transform, then the tree refined, coarsened and rebalanced to resolve pot
    (bfgt/boxfgt.f:1172-1483).  ``r.tree`` is the new tree (nlevels may
    change), ``r.pot`` etc. hold the values on its boxes, and ``r.info``
    includes ``nnew`` (boxes created by the refinement) and ``ncapped`` (leaves
    at the finest level whose refinement was not done)."""
    lib = load_oracle()
    itree = np.ascontiguousarray(tree.itree, dtype=np.int32).copy()
    iptr = np.ascontiguousarray(tree.iptr, dtype=np.int32)
    centers = np.ascontiguousarray(tree.centers, dtype=np.float64).copy()
    boxsize = np.ascontiguousarray(tree.boxsize, dtype=np.float64)
    nbmax = centers.shape[0]
    npbox = norder * norder
    fv = np.zeros((nbmax, npbox, nd))
    fv[: fvals.shape[0]] = fvals
    pot, grad, hess = _alloc(nbmax, npbox, nd)
    ti = np.zeros(20)
    info = np.zeros(8, dtype=np.int32)
    nl = ctypes.c_int(tree.nlevels)
    rc = lib.oracle_boxfgt_nt(nd, delta, eps, ipoly, iperiod, norder, npbox, nbmax, ctypes.byref(nl),
                              _p(itree, _ip), _p(iptr, _ip), _p(centers), _p(boxsize), _p(fv), ifpgh, _p(pot),
                              _p(grad), _p(hess), 1, _p(ti), _p(info, _ip))
    if rc != 0:
        raise RuntimeError(f"oracle_boxfgt_nt failed with status {rc}")
    keys = ["npwlevel", "npw", "nlevstart", "nlevend", "nleafbox", "npairs", "nnew", "ncapped"]
    r = FGTResult(pot, grad, hess, ti, itree, {k: int(v) for k, v in zip(keys, info)})
    r.tree = _adapted_tree(tree, itree, nl.value, centers)
    return r


def reference_boxfgt(tree, fvals, nd, delta, eps, ipoly, iperiod, norder, ifpgh, nthreads: Optional[int] = None,
                     targs=None, ifpghtarg=0, pot0=None, ifnewtree=0) -> FGTResult:
    """Run the reference Fortran boxfgt on a copy of the tree, optionally
    with extra targets (ntarg, 2) and pot pre-loaded with pot0.  With
    ifnewtree=1 the tree arrays (laid out for ``centers.shape[0]`` boxes) are
    refined, coarsened and rebalanced and ``r.tree`` is the new tree."""
    lib = load_reference()
    if nthreads is not None:
        try:
            omp = ctypes.CDLL("libomp.so")
            omp.omp_set_num_threads(int(nthreads))
        except OSError:
            pass
    c_int = ctypes.c_int
    c_dbl = ctypes.c_double
    itree = np.ascontiguousarray(tree.itree, dtype=np.int32).copy()
    iptr = np.ascontiguousarray(tree.iptr, dtype=np.int32).copy()
    centers = np.ascontiguousarray(tree.centers, dtype=np.float64).copy()
    boxsize = np.ascontiguousarray(tree.boxsize, dtype=np.float64).copy()
    nboxes = centers.shape[0]
    npbox = norder * norder
    fv = np.zeros((nboxes, npbox, nd))
    fv[: fvals.shape[0]] = fvals
    pot, grad, hess = _alloc(nboxes, npbox, nd)
    if pot0 is not None:
        pot[: pot0.shape[0]] += pot0
    ti = np.zeros(20)
    if targs is not None and ifpghtarg > 0:
        tg = np.ascontiguousarray(targs, dtype=np.float64).copy()
        nt = tg.shape[0]
    else:
        tg, nt, ifpghtarg = np.zeros((1, 2)), 0, 0
    pote = np.zeros((max(nt, 1), nd))
    grade = np.zeros((max(nt, 1), 2, nd))
    hesse = np.zeros((max(nt, 1), 3, nd))
    args = [
        c_int(nd), c_int(2), c_dbl(delta), c_dbl(eps), c_int(ipoly), c_int(iperiod), c_int(norder), c_int(npbox),
        c_int(nboxes), c_int(tree.nlevels), c_int(itree.shape[0]),
    ]
    refs = [ctypes.byref(a) for a in args]
    refs += [_p(itree, _ip), _p(iptr, _ip), _p(centers), _p(boxsize), _p(fv)]
    ifpgh_c, ifnew, ntarg, ifpt = c_int(ifpgh), c_int(ifnewtree), c_int(nt), c_int(ifpghtarg)
    refs += [ctypes.byref(ifpgh_c), _p(pot), _p(grad), _p(hess), ctypes.byref(ifnew), ctypes.byref(ntarg), _p(tg),
             ctypes.byref(ifpt), _p(pote), _p(grade), _p(hesse), _p(ti)]
    lib.boxfgt_(*refs)
    if args[5].value != iperiod:
        raise RuntimeError("reference did not restore iperiod")
    r = FGTResult(pot, grad, hess, ti, itree)
    if ifnewtree:
        r.tree = _adapted_tree(tree, itree, args[9].value, centers)
    if nt > 0:
        r.pote = pote[:nt]
        r.grade = grade[:nt] if ifpghtarg >= 2 else None
        r.hesse = hesse[:nt] if ifpghtarg >= 3 else None
    return r


# -- 1-D pieces, for pinning the oracle against closed forms ----------------

def legeexps(n):
    lib = load_oracle()
    x, w, u = np.zeros(n), np.zeros(n), np.zeros((n, n))
    lib.oracle_legeexps(n, _p(x), _p(w), _p(u))
    return x, w, u.T.copy()  # u[i, j] = U(i, j): coefficient i from value j


def chebexps(n):
    lib = load_oracle()
    x, w, u = np.zeros(n), np.zeros(n), np.zeros((n, n))
    lib.oracle_chebexps(n, _p(x), _p(w), _p(u))
    return x, w, u.T.copy()


def fgtpwterms(bsize, delta, eps, iperiod):
    lib = load_oracle()
    pmax = ctypes.c_double(0)
    npw = ctypes.c_int(0)
    lib.oracle_fgtpwterms(bsize, delta, eps, iperiod, ctypes.byref(pmax), ctypes.byref(npw))
    return pmax.value, npw.value


def pwnodes(pmax, npw):
    lib = load_oracle()
    ws, ts = np.zeros(npw), np.zeros(npw)
    lib.oracle_pwnodes(pmax, npw, _p(ws), _p(ts))
    return ws, ts


def periodic_pwnodes(bs0, delta, npw):
    lib = load_oracle()
    ws, ts = np.zeros(npw), np.zeros(npw)
    lib.oracle_periodic_pwnodes(bs0, delta, npw, _p(ws), _p(ts))
    return ws, ts


def sph_besselj(n, x):
    lib = load_oracle()
    f = np.zeros(n)
    lib.oracle_sph_besselj(n, x, _p(f))
    return f


def poly2pw(n, ipoly, ws, ts, delta, boxdim):
    """tab[m, j] (m = node, j = PW index) complex."""
    lib = load_oracle()
    npw = len(ts)
    out = np.zeros((npw, n), dtype=np.complex128)
    lib.oracle_poly2pw(n, ipoly, npw, _p(np.ascontiguousarray(ws)), _p(np.ascontiguousarray(ts)), delta, boxdim,
                       out.ctypes.data_as(_dp))
    return out.T.copy()


def pw2pgh(n, ts, xs, delta, boxdim):
    """E[j, i] tables (npw x n) for pot, d/dx, d2/dx2."""
    lib = load_oracle()
    npw = len(ts)
    outs = [np.zeros((n, npw), dtype=np.complex128) for _ in range(3)]
    lib.oracle_pw2pgh(n, npw, _p(np.ascontiguousarray(ts)), _p(np.ascontiguousarray(xs)), delta, boxdim,
                      *[o.ctypes.data_as(_dp) for o in outs])
    return [o.T.copy() for o in outs]


def loctab_level(eps, ipoly, n, delta, boxdim):
    """Near-field tables of one target level: arrays (13, n_targ, n_src) for k=-6..6
    (entry [k+6][j, m] couples source node m to target node j) and ind (13, n+1, 2)
    1-based sparsity bounds (row j: source range of target j; row n: target range)."""
    lib = load_oracle()
    t = [np.zeros((13, n, n)) for _ in range(3)]
    ind = np.zeros((13, n + 1, 2), dtype=np.int32)
    lib.oracle_loctab_level(eps, ipoly, n, delta, boxdim, *[_p(a) for a in t], _p(ind, _ip))
    # stored column-major per offset: element (m, j) at m + n*j, i.e. C array [k][j][m]
    return t + [ind]


# -- the reference's tree builder and adapter (common/tree_ext.f) ------------

def tree_fun_dpars(bumps, t=0.0):
    """dpars of the test function in tree_funcs.c: bumps = [(a, cx, cy, vx, vy, s), ...]."""
    return np.array([float(t), float(len(bumps))] + [float(v) for b in bumps for v in b], dtype=np.float64)


def tree_fun_eval(nd, xy, dpars):
    """The test function of tree_funcs.c at points xy (npts, 2): (npts, nd)."""
    lib = load_oracle()
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    out = np.zeros((xy.shape[0], nd))
    lib.oracle_tree_fun_eval(ctypes.c_int(nd), ctypes.c_int(xy.shape[0]), _p(xy),
                             _p(np.ascontiguousarray(dpars, dtype=np.float64)), _p(out))
    return out


@dataclass
class RefTree:
    """A tree as the reference's builders leave it: itree laid out for
    (mlevels, mboxes) capacities (iptr gives the offsets), the first nboxes
    boxes and nlevels+1 levels in use."""
    nboxes: int
    nlevels: int
    itree: np.ndarray
    iptr: np.ndarray
    centers: np.ndarray   # (nboxes, 2)
    boxsize: np.ndarray   # (nlevels + 1,)
    vals: np.ndarray      # (nboxes, npbox, nd)

    def _a(self, k, width=1):
        o = int(self.iptr[k]) - 1
        a = self.itree[o : o + width * self.nboxes]
        return a.reshape(self.nboxes, width) if width > 1 else a

    @property
    def laddr(self):
        return self.itree[: 2 * (self.nlevels + 1)].reshape(-1, 2)

    @property
    def ilevel(self):
        return self._a(1)

    @property
    def iparent(self):
        return self._a(2)

    @property
    def nchild(self):
        return self._a(3)

    @property
    def ichild(self):
        return self._a(4, 4)

    @property
    def ncoll(self):
        return self._a(5)

    @property
    def coll(self):
        return self._a(6, 9)


def _tree_fun_ptr():
    lib = load_oracle()
    return ctypes.c_void_p(ctypes.cast(lib.oracle_tree_fun, ctypes.c_void_p).value)


def set_reference_threads(nthreads: int) -> None:
    """OpenMP threads of the compiled reference (its runtime is libomp)."""
    try:
        ctypes.CDLL("libomp.so").omp_set_num_threads(int(nthreads))
    except OSError:
        pass


def reference_tree_mem(nd, ipoly, iperiod, eps, boxlen, norder, iptype, eta, dpars, ifnewtree=1):
    """vol_tree_mem2 (common/tree_ext.f:996): (mboxes, mlevels, ltree, rintl(0:200))."""
    lib = load_reference()
    c_int, c_dbl = ctypes.c_int, ctypes.c_double
    zk = np.array([1.0, 0.0])
    zpars = np.zeros(2)
    ipars = np.zeros(1, dtype=np.int32)
    dp = np.ascontiguousarray(dpars, dtype=np.float64).copy()
    mboxes, mlevels, ltree = c_int(0), c_int(0), c_int(0)
    rintl = np.zeros(201)
    args = [c_int(2), c_int(ipoly), c_int(iperiod), c_dbl(eps)]
    refs = [ctypes.byref(a) for a in args] + [_p(zk)]
    a2 = [c_dbl(boxlen), c_int(norder), c_int(iptype), c_dbl(eta)]
    refs += [ctypes.byref(a) for a in a2] + [_tree_fun_ptr()]
    ndc, ifn = c_int(nd), c_int(ifnewtree)
    refs += [ctypes.byref(ndc), _p(dp), _p(zpars), _p(ipars, _ip), ctypes.byref(ifn), ctypes.byref(mboxes),
             ctypes.byref(mlevels), ctypes.byref(ltree), _p(rintl)]
    lib.vol_tree_mem2_(*refs)
    return mboxes.value, mlevels.value, ltree.value, rintl


def reference_tree_build(nd, ipoly, iperiod, eps, boxlen, norder, iptype, eta, dpars, rintl, mboxes, mlevels,
                         ltree) -> RefTree:
    """vol_tree_build2 (common/tree_ext.f:1363) with capacities from vol_tree_mem2."""
    lib = load_reference()
    c_int, c_dbl = ctypes.c_int, ctypes.c_double
    npbox = norder * norder
    zk = np.array([1.0, 0.0])
    zpars = np.zeros(2)
    ipars = np.zeros(1, dtype=np.int32)
    dp = np.ascontiguousarray(dpars, dtype=np.float64).copy()
    rl = np.zeros(max(mlevels + 1, 201))
    rl[: len(rintl)] = rintl[: len(rl)]
    itree = np.zeros(ltree, dtype=np.int32)
    iptr = np.zeros(8, dtype=np.int32)
    centers = np.zeros((mboxes, 2))
    boxsize = np.zeros(mlevels + 1)
    fvals = np.zeros((mboxes, npbox, nd))
    nboxes, nlevels = c_int(0), c_int(0)
    args = [c_int(2), c_int(ipoly), c_int(iperiod), c_dbl(eps)]
    refs = [ctypes.byref(a) for a in args] + [_p(zk)]
    a2 = [c_dbl(boxlen), c_int(norder), c_int(iptype), c_dbl(eta)]
    refs += [ctypes.byref(a) for a in a2] + [_tree_fun_ptr()]
    ndc, mb, ml, lt = c_int(nd), c_int(mboxes), c_int(mlevels), c_int(ltree)
    refs += [ctypes.byref(ndc), _p(dp), _p(zpars), _p(ipars, _ip), _p(rl), ctypes.byref(nboxes), ctypes.byref(mb),
             ctypes.byref(nlevels), ctypes.byref(ml), ctypes.byref(lt), _p(itree, _ip), _p(iptr, _ip), _p(centers),
             _p(boxsize), _p(fvals)]
    lib.vol_tree_build2_(*refs)
    nb, nl = nboxes.value, nlevels.value
    return RefTree(nb, nl, itree, iptr, centers[:nb].copy(), boxsize[: nl + 1].copy(), fvals[:nb].copy())


def reference_tree_adapt(tree: RefTree, uvals, nd, ipoly, iperiod, eps, norder, iptype, rintl, dpars, mboxes,
                         mlevels, ltree) -> RefTree:
    """vol_tree_adap_fv (common/tree_ext.f:2337) on copies of a tree built by
    reference_tree_build with the same capacities; returns the adapted tree
    with uvals interpolated onto it (in RefTree.vals)."""
    lib = load_reference()
    c_int, c_dbl = ctypes.c_int, ctypes.c_double
    npbox = norder * norder
    zpars = np.zeros(2)
    ipars = np.zeros(1, dtype=np.int32)
    dp = np.ascontiguousarray(dpars, dtype=np.float64).copy()
    rl = np.zeros(max(mlevels + 1, 201))
    rl[: len(rintl)] = rintl[: len(rl)]
    itree = tree.itree.copy()
    iptr = tree.iptr.copy()
    centers = np.zeros((mboxes, 2))
    centers[: tree.nboxes] = tree.centers
    boxsize = np.zeros(mlevels + 1)
    boxsize[: tree.nlevels + 1] = tree.boxsize
    uv = np.zeros((mboxes, npbox, nd))
    uv[: tree.nboxes] = uvals
    nboxes, nlevels = c_int(tree.nboxes), c_int(tree.nlevels)
    args = [c_int(nd), c_int(2), c_dbl(eps), c_int(ipoly), c_int(iptype), c_int(norder), c_int(npbox)]
    refs = [ctypes.byref(a) for a in args]
    mb, ml, lt = c_int(mboxes), c_int(mlevels), c_int(ltree)
    refs += [ctypes.byref(nboxes), ctypes.byref(mb), ctypes.byref(nlevels), ctypes.byref(ml), ctypes.byref(lt),
             _p(itree, _ip), _p(iptr, _ip), _p(centers), _p(boxsize), _p(rl), _p(dp), _p(zpars), _p(ipars, _ip)]
    ip = c_int(iperiod)
    refs += [ctypes.byref(ip), _tree_fun_ptr(), _p(uv)]
    lib.vol_tree_adap_fv_(*refs)
    nb, nl = nboxes.value, nlevels.value
    return RefTree(nb, nl, itree, iptr, centers[:nb].copy(), boxsize[: nl + 1].copy(), uv[:nb].copy())
