"""Python mirror of the reference ``boxfgt`` entry point, over the HIP C ABI.

``boxfgt(...)`` takes the reference's 28 arguments (bfgt/boxfgt.f:5-8) as
scalars and numpy arrays, with the same meaning, in-place side effects
(``itree``'s colleague lists are recomputed, ``pot/grad/hess`` are
accumulated on leaf boxes, ``timeinfo`` is zeroed and filled) and the same
restrictions the GPU library states (2-D), extra targets included
(``ifpghtarg``, ``targs``, ``pote/grade/hesse``), raising
:class:`BoxFGTError` where the reference would silently misbehave.  With
``ifnewtree=1`` the tree (``itree``, ``centers``) and pot/grad/hess are
adapted to pot in place and the new ``nlevels`` is returned.

``BoxFGT`` is the resident form: the plan (lists, FP64 tables, schedules)
lives on the GPU and each call runs on device tensors already in HBM.

Array shapes (C order, i.e. the Fortran shapes reversed):
``fvals/pot (nboxes, npbox, nd)``, ``grad (nboxes, npbox, 2, nd)``,
``hess (nboxes, npbox, 3, nd)``, ``centers (nboxes, 2)``.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from ._lib import PRECISIONS, BoxFGTError, check, lib

__all__ = ["boxfgt", "BoxFGT", "BoxFGTError", "PlanInfo"]

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _ptr(a, t=_dp):
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("arrays must be C-contiguous")
    return a.ctypes.data_as(t)


def _f64(a, name):
    if a is None:
        return None
    if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]):
        raise TypeError(f"{name} must be a C-contiguous float64 numpy array")
    return a


def _i32(a, name):
    if not (isinstance(a, np.ndarray) and a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]):
        raise TypeError(f"{name} must be a C-contiguous int32 numpy array")
    return a


def boxfgt(nd, ndim, delta, eps, ipoly, iperiod, norder, npbox, nboxes, nlevels, ltree, itree, iptr, centers,
           boxsize, fvals, ifpgh, pot, grad=None, hess=None, ifnewtree=0, ntarg=0, targs=None, ifpghtarg=0,
           pote=None, grade=None, hesse=None, timeinfo=None, precision: str = "fp64") -> int:
    """Drop-in for the reference ``boxfgt`` call on host arrays (synchronous).
    Returns nlevels, which a call with ifnewtree=1 may change."""
    itree = _i32(itree, "itree")
    iptr = _i32(iptr, "iptr")
    centers = _f64(centers, "centers")
    boxsize = _f64(boxsize, "boxsize")
    fvals = _f64(fvals, "fvals")
    pot = _f64(pot, "pot")
    grad = _f64(grad, "grad")
    hess = _f64(hess, "hess")
    nv = nboxes * npbox * nd
    if fvals.size < nv or pot.size < nv:
        raise ValueError("fvals/pot smaller than (nd, npbox, nboxes)")
    if ifpgh >= 2 and (grad is None or grad.size < 2 * nv):
        raise ValueError("grad must hold (nd, 2, npbox, nboxes)")
    if ifpgh >= 3 and (hess is None or hess.size < 3 * nv):
        raise ValueError("hess must hold (nd, 3, npbox, nboxes)")
    if ifpghtarg > 0 and ntarg > 0:
        targs = _f64(targs, "targs")
        pote = _f64(pote, "pote")
        grade = _f64(grade, "grade")
        hesse = _f64(hesse, "hesse")
        if targs is None or targs.size < 2 * ntarg or pote is None or pote.size < ntarg * nd:
            raise ValueError("targs must hold (2, ntarg) and pote (nd, ntarg)")
        if ifpghtarg >= 2 and (grade is None or grade.size < 2 * ntarg * nd):
            raise ValueError("grade must hold (nd, 2, ntarg)")
        if ifpghtarg >= 3 and (hesse is None or hesse.size < 3 * ntarg * nd):
            raise ValueError("hesse must hold (nd, 3, ntarg)")
    ti = np.zeros(20) if timeinfo is None else _f64(timeinfo, "timeinfo")
    if ifnewtree == 1 and centers.shape[0] * centers.shape[-1] < 2 * nboxes:
        raise ValueError("centers must hold (2, nboxes)")
    nl = ctypes.c_int(nlevels)
    rc = lib().bfgt_hip_boxfgt(
        nd, ndim, delta, eps, ipoly, iperiod, norder, npbox, nboxes, ctypes.byref(nl), ltree, _ptr(itree, _ip),
        _ptr(iptr, _ip), _ptr(centers), _ptr(boxsize), _ptr(fvals), ifpgh, _ptr(pot), _ptr(grad), _ptr(hess),
        ifnewtree, ntarg, _ptr(targs), ifpghtarg, _ptr(pote), _ptr(grade), _ptr(hesse), _ptr(ti),
        PRECISIONS[precision])
    check(rc)
    return nl.value


def release_plans() -> None:
    """Free the plans that :func:`boxfgt` keeps for later calls on the same tree."""
    check(lib().bfgt_hip_release_plans())


class PlanInfo(dict):
    pass


class _StreamOrder:
    def __init__(self, cur, plan_stream):
        self.cur, self.ps = cur, plan_stream

    def __enter__(self):
        self.ps.wait_stream(self.cur)

    def __exit__(self, *exc):
        self.cur.wait_stream(self.ps)
        return False


class _NoOrder:
    def __enter__(self):
        return None

    def __exit__(self, *exc):
        return False


class BoxFGT:
    """Resident box FGT plan on the current GPU.

    ``tree`` is a :class:`heat2d_per_box_amd.tree.QuadTree`; its colleague lists
    are recomputed in place, as the reference does.  Call with torch tensors on
    the GPU (float64, contiguous); outputs are accumulated.
    """

    def __init__(self, tree, nd: int, delta: float, eps: float, ipoly: int, norder: int, ifpgh: int,
                 precision: str = "fp64", iperiod: Optional[int] = None, stream=None):
        import torch

        self.tree = tree
        self.nd, self.delta, self.eps, self.ipoly, self.norder, self.ifpgh = nd, delta, eps, ipoly, norder, ifpgh
        self.npbox = norder * norder
        self.iperiod = tree.iperiod if iperiod is None else iperiod
        self.precision = precision
        self.nboxes = tree.centers.shape[0]
        if stream is None:
            stream = torch.cuda.current_stream()
        self.stream = stream
        self._h = ctypes.c_void_p()
        it = _i32(tree.itree, "itree")
        rc = lib().bfgt_plan_create(ctypes.byref(self._h), nd, 2, delta, eps, ipoly, self.iperiod, norder,
                                    self.nboxes, tree.nlevels, it.shape[0], _ptr(it, _ip),
                                    _ptr(_i32(tree.iptr, "iptr"), _ip), _ptr(_f64(tree.centers, "centers")),
                                    _ptr(_f64(tree.boxsize, "boxsize")), ifpgh, PRECISIONS[precision],
                                    ctypes.c_void_p(stream.cuda_stream))
        check(rc)
        iv = (ctypes.c_int * 16)()
        dv = (ctypes.c_double * 8)()
        check(lib().bfgt_plan_query(self._h, iv, dv))
        keys = ["npwlevel", "npw", "nexp", "nlevstart", "nlevend", "nleafbox", "n_pw_leaves", "n_chunks",
                "n_ancestors", "n_near_targets", "n_near_pairs", "has_pw", "has_near", "iperiod_eff", "nboxes",
                "precision"]
        self.info = PlanInfo({k: int(v) for k, v in zip(keys, iv)})
        self.info.update({"delta0": dv[0], "pmax": dv[1], "t_plan_host": dv[2], "t_tables": dv[3],
                          "mfma": bool(dv[4]), "n_m2l": int(dv[5]), "n_near_sources": int(dv[6]),
                          "translation_group": int(dv[7])})
        wk = (ctypes.c_longlong * 7)()
        check(lib().bfgt_plan_work(self._h, wk, 7))
        self.info.update({"xc_up": int(wk[0]), "xc_down": int(wk[1]), "xc_adj_eval": int(wk[3]),
                          "xc_adj_transform": int(wk[4]), "merge_reads": int(wk[5]),
                          "merge_writes": int(wk[6])})

    def _ordered(self):
        """Order the plan's stream after the caller's current stream before a
        launch, and the current stream after the plan's stream afterwards, so
        that tensors produced and consumed on whatever stream is current at the
        call see the library's work in order."""
        import torch

        cur = torch.cuda.current_stream()
        return _StreamOrder(cur, self.stream) if cur != self.stream else _NoOrder()

    def _check_tensor(self, t, shape, name):
        import torch

        if t is None:
            return None
        if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
            raise TypeError(f"{name} must be a contiguous float64 GPU tensor")
        if tuple(t.shape) != shape:
            raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {shape}")
        return ctypes.c_void_p(t.data_ptr())

    def __call__(self, fvals, pot, grad=None, hess=None, timeinfo: bool = False, scale: float = 1.0,
                 overwrite: bool = False):
        """pot (+grad, hess) += scale * A fvals on device tensors; with
        ``overwrite`` the leaf entries are set to scale * A fvals instead."""
        nb, npb, nd = self.nboxes, self.npbox, self.nd
        f = self._check_tensor(fvals, (nb, npb, nd), "fvals")
        p = self._check_tensor(pot, (nb, npb, nd), "pot")
        g = self._check_tensor(grad, (nb, npb, 2, nd), "grad")
        h = self._check_tensor(hess, (nb, npb, 3, nd), "hess")
        ti = np.zeros(20) if timeinfo else None
        tip = _ptr(ti) if ti is not None else None
        with self._ordered():
            if overwrite:
                check(lib().bfgt_plan_execute_overwrite(self._h, float(scale), f, p, g, h, tip))
            elif scale == 1.0:
                check(lib().bfgt_plan_execute(self._h, f, p, g, h, tip))
            else:
                check(lib().bfgt_plan_execute_scaled(self._h, float(scale), f, p, g, h, tip))
        return ti

    def leaf_points(self, xy=None):
        """Device tensor (nboxes, npbox, 2) of the leaf tensor-grid nodes (zeros on non-leaf boxes)."""
        import torch

        if xy is None:
            xy = torch.zeros((self.nboxes, self.npbox, 2), dtype=torch.float64, device="cuda")
        x = self._check_tensor(xy, (self.nboxes, self.npbox, 2), "xy")
        with self._ordered():
            check(lib().bfgt_plan_leaf_points(self._h, x))
        return xy

    def eval_targets(self, targs, ifpghtarg: int, pot, pote, grad=None, hess=None, grade=None, hesse=None):
        """Interpolate the leaf arrays at targets (ntarg, 2) into pote (ntarg, nd),
        grade (ntarg, 2, nd), hesse (ntarg, 3, nd): device tensors, overwritten."""
        nb, npb, nd = self.nboxes, self.npbox, self.nd
        nt = int(targs.shape[0])
        tg = self._check_tensor(targs, (nt, 2), "targs")
        p = self._check_tensor(pot, (nb, npb, nd), "pot")
        g = self._check_tensor(grad, (nb, npb, 2, nd), "grad")
        h = self._check_tensor(hess, (nb, npb, 3, nd), "hess")
        pe = self._check_tensor(pote, (nt, nd), "pote")
        ge = self._check_tensor(grade, (nt, 2, nd), "grade")
        he = self._check_tensor(hesse, (nt, 3, nd), "hesse")
        with self._ordered():
            check(lib().bfgt_plan_eval_targets(self._h, nt, tg, int(ifpghtarg), p, g, h, pe, ge, he))

    def adjoint(self, gpot, fbar, ggrad=None, ghess=None, timeinfo: bool = False):
        """Transposed operator: ``fbar += A^T (gpot, ggrad, ghess)`` on device
        tensors, A the map fvals -> (pot, grad, hess) of this plan's ifpgh."""
        nb, npb, nd = self.nboxes, self.npbox, self.nd
        gp = self._check_tensor(gpot, (nb, npb, nd), "gpot")
        fb = self._check_tensor(fbar, (nb, npb, nd), "fbar")
        gg = self._check_tensor(ggrad, (nb, npb, 2, nd), "ggrad")
        gh = self._check_tensor(ghess, (nb, npb, 3, nd), "ghess")
        ti = np.zeros(20) if timeinfo else None
        with self._ordered():
            check(lib().bfgt_plan_adjoint(self._h, gp, gg, gh, fb, _ptr(ti) if ti is not None else None))
        return ti

    def profile_begin(self, max_steps: int, stage: int = -1) -> None:
        """Record per-stage HIP events for the next ``max_steps`` calls: around
        every stage (``stage=-1``) or only around one (0 leaf->PW, 1 merge,
        2 translate, 4 PW->values, 5 near field)."""
        check(lib().bfgt_plan_profile_stage(self._h, int(max_steps), int(stage)))

    def profile_end(self):
        """Summed seconds per stage over the profiled calls and their count.
        Stages: leaf->PW, merge, translate, split (0), PW->values, near field."""
        out = np.zeros(6)
        cnt = ctypes.c_int(0)
        check(lib().bfgt_plan_profile_end(self._h, _ptr(out), ctypes.byref(cnt)))
        return out, cnt.value

    def table(self, which: int, level: int) -> np.ndarray:
        n, npw, nexp = self.norder, self.info["npw"], self.info["nexp"]
        shapes = {
            0: ((n, npw), np.complex128),
            1: ((3, npw, n), np.complex128),
            2: ((13, 3, n, n), np.float64),
            3: ((13, n + 1, 2), np.int32),
            4: ((9, npw, (npw + 1) // 2), np.complex128),
        }
        H = (npw + 1) // 2
        if self.info["has_pw"]:
            shapes[5] = ((1 << max(level - self.info["nlevstart"], 0), npw), np.complex128)
            shapes[6] = ((self.info["n_ancestors"], self.nd, npw, H), np.complex128)
            # local expansions as stored: mirror-folded (S, D) on the MFMA path
            mh4, h16 = -(-H // 4) * 4, -(-H // 16) * 16
            shapes[7] = (((self.info["n_ancestors"], self.nd, mh4, h16, 2), np.complex128) if self.info["mfma"]
                         else shapes[6])
        shapes[8] = ((self.info["n_near_targets"] + 1,), np.int32)
        shape, dt = shapes[which]
        out = np.zeros(shape, dtype=dt)
        check(lib().bfgt_plan_get_table(self._h, which, level, out.ctypes.data_as(ctypes.c_void_p), out.nbytes))
        return out

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().bfgt_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
