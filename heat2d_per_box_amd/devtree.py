"""Adaptive quadtrees resident in GPU memory (the ``bfgt_tree_*`` C ABI).

``DeviceTree.build`` is the reference's ``vol_tree_mem2`` + ``vol_tree_build2``
(common/tree_ext.f:996, :1363) and ``DeviceTree.adapt`` its per-step
re-meshing ``vol_tree_adap_fv`` (common/tree_ext.f:2337), both run on the GPU;
the tree arrays and the values on the leaf nodes stay in device memory, and
``DeviceTree.plan`` makes the resident box FGT on the tree as it is.

The user function is a Python callable.  With ``host=False`` (default) it
receives the points as a float64 GPU tensor ``(npts, 2)`` and returns
``(npts, nd)`` (or ``(npts,)`` for nd = 1) on the GPU; with ``host=True`` it
receives and returns numpy arrays, like the reference's ``fun``, which is
called on the host.
"""
from __future__ import annotations

import ctypes
from typing import Callable, Optional

import numpy as np

from ._lib import TREE_FUN, check, lib

__all__ = ["DeviceTree"]

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class _DeviceArray:
    """A float64 device array seen through ``__cuda_array_interface__``."""

    def __init__(self, ptr: int, shape):
        self.__cuda_array_interface__ = {"shape": tuple(shape), "typestr": "<f8", "data": (int(ptr), False),
                                         "version": 2, "strides": None}


def _device_tensor(ptr: int, shape):
    import torch

    return torch.as_tensor(_DeviceArray(ptr, shape), device="cuda")


class _Callback:
    """ctypes trampoline from the library's callback to a Python function."""

    def __init__(self, fun: Callable, host: bool):
        self.fun, self.host = fun, host
        self.error: Optional[BaseException] = None
        self.cfun = TREE_FUN(self._call)

    def _call(self, ctx, nd, npts, d_xy, d_vals, stream):
        import torch

        try:
            # torch work goes on the library's stream, ordered before its next kernels
            with torch.cuda.stream(torch.cuda.ExternalStream(stream)):
                xy = _device_tensor(d_xy, (npts, 2))
                out = _device_tensor(d_vals, (npts, nd))
                if self.host:
                    v = np.asarray(self.fun(xy.cpu().numpy()), dtype=np.float64).reshape(npts, nd)
                    out.copy_(torch.from_numpy(np.ascontiguousarray(v)))
                else:
                    out.copy_(self.fun(xy).reshape(npts, nd))
            return 0
        except BaseException as e:  # reported by the caller after the C call returns
            self.error = e
            return 1


class DeviceTree:
    """A level-restricted quadtree on [-boxlen/2, boxlen/2]^2 in GPU memory,
    with values ``(nboxes, npbox, nd)`` on its nodes (the reference's
    ``fvals(nd, npbox, nboxes)``)."""

    def __init__(self):
        raise TypeError("use DeviceTree.build(...)")

    @classmethod
    def build(cls, fun: Callable, nd: int, norder: int, eps: float, ipoly: int = 0, iperiod: int = 1,
              boxlen: float = 1.0, iptype: int = 2, eta: float = 1.0, rintl=None, max_levels: int = 20,
              max_boxes: int = 1 << 20, host: bool = False, stream=None) -> "DeviceTree":
        """Build the tree that resolves ``fun`` (its values become the tree's
        values).  ``rintl`` None computes the per-level norms as the tree grows
        (``vol_tree_mem2``); they are kept in ``self.rintl`` (length
        ``max_levels + 1``) for :meth:`adapt`."""
        import torch

        self = object.__new__(cls)
        self._h = ctypes.c_void_p()
        self.stream = torch.cuda.current_stream() if stream is None else stream
        self.nd, self.norder, self.ipoly, self.iperiod, self.boxlen = nd, norder, ipoly, iperiod, boxlen
        self.iptype, self.eta = iptype, eta
        cb = _Callback(fun, host)
        rin = None
        if rintl is not None:
            rin = np.ascontiguousarray(rintl, dtype=np.float64)
            if rin.shape[0] < max_levels + 1:
                raise ValueError(f"rintl needs {max_levels + 1} entries")
        rout = np.zeros(max_levels + 1)
        rc = lib().bfgt_tree_build(ctypes.byref(self._h), nd, ipoly, iperiod, eps, boxlen, norder, iptype, eta,
                                   cb.cfun, None, rin.ctypes.data_as(_dp) if rin is not None else None, max_levels,
                                   max_boxes, rout.ctypes.data_as(_dp), ctypes.c_void_p(self.stream.cuda_stream))
        if cb.error is not None:
            raise cb.error
        check(rc)
        self.rintl = rout
        return self

    @classmethod
    def from_tree(cls, tree, norder: int, vals=None, nd: int = 1, ipoly: int = 0, iptype: int = 2,
                  eta: float = 1.0, rintl=None, max_boxes: Optional[int] = None, stream=None) -> "DeviceTree":
        """A device tree from a :class:`heat2d_per_box_amd.tree.QuadTree` (or any
        tree in the reference layout with ``nlevels, itree, iptr, centers,
        boxsize, iperiod``) and values ``(nboxes, npbox, nd)``; ``rintl`` are
        the norms :meth:`adapt` uses by default."""
        import torch

        self = object.__new__(cls)
        self._h = ctypes.c_void_p()
        self.stream = torch.cuda.current_stream() if stream is None else stream
        nb = tree.centers.shape[0]
        self.nd, self.norder, self.ipoly, self.iperiod = nd, norder, ipoly, tree.iperiod
        self.boxlen, self.iptype, self.eta = float(tree.boxsize[0]), iptype, eta
        itree = np.ascontiguousarray(tree.itree, dtype=np.int32)
        iptr = np.ascontiguousarray(tree.iptr, dtype=np.int32)
        centers = np.ascontiguousarray(tree.centers, dtype=np.float64)
        boxsize = np.ascontiguousarray(tree.boxsize, dtype=np.float64)
        v = None
        if vals is not None:
            v = np.ascontiguousarray(vals, dtype=np.float64)
            if v.shape != (nb, norder * norder, nd):
                raise ValueError(f"vals must have shape {(nb, norder * norder, nd)}")
        rc = lib().bfgt_tree_import(ctypes.byref(self._h), nd, ipoly, tree.iperiod, norder, iptype, eta, nb,
                                    tree.nlevels, itree.ctypes.data_as(_ip), iptr.ctypes.data_as(_ip),
                                    centers.ctypes.data_as(_dp), boxsize.ctypes.data_as(_dp),
                                    v.ctypes.data_as(_dp) if v is not None else None,
                                    max_boxes if max_boxes is not None else 4 * nb,
                                    ctypes.c_void_p(self.stream.cuda_stream))
        check(rc)
        self.rintl = None if rintl is None else np.ascontiguousarray(rintl, dtype=np.float64)
        return self

    def adapt(self, fun: Callable, eps: float, rintl=None, max_levels: int = 20, host: bool = False) -> None:
        """Re-mesh for a new ``fun`` and interpolate the tree's values onto the
        new tree (``vol_tree_adap_fv``).  ``rintl`` defaults to the norms of the
        build."""
        r = self.rintl if rintl is None else np.ascontiguousarray(rintl, dtype=np.float64)
        if r is None:
            raise ValueError("no norms: pass rintl")
        if r.shape[0] < max_levels + 1:
            raise ValueError(f"rintl needs {max_levels + 1} entries")
        cb = _Callback(fun, host)
        rc = lib().bfgt_tree_adapt(self._h, eps, r.ctypes.data_as(_dp), cb.cfun, None, max_levels)
        if cb.error is not None:
            raise cb.error
        check(rc)

    @property
    def info(self) -> dict:
        iv = (ctypes.c_int * 10)()
        check(lib().bfgt_tree_query(self._h, iv, None))
        keys = ["nboxes", "nlevels", "ltree", "nleaves", "max_boxes", "nd", "norder", "ipoly", "iperiod", "iptype"]
        return {k: int(v) for k, v in zip(keys, iv)}

    def values(self):
        """The tree's values as a GPU tensor view ``(nboxes, npbox, nd)``;
        valid until the next :meth:`adapt`."""
        iv = (ctypes.c_int * 10)()
        p = ctypes.c_void_p()
        check(lib().bfgt_tree_query(self._h, iv, ctypes.byref(p)))
        return _device_tensor(p.value, (iv[0], self.norder * self.norder, self.nd))

    def to_host(self):
        """(QuadTree, values (nboxes, npbox, nd)) in the reference layout."""
        from .tree import QuadTree

        info = self.info
        nb, nl = info["nboxes"], info["nlevels"]
        itree = np.zeros(info["ltree"], dtype=np.int32)
        iptr = np.zeros(8, dtype=np.int32)
        centers = np.zeros((nb, 2))
        boxsize = np.zeros(nl + 1)
        vals = np.zeros((nb, self.norder * self.norder, self.nd))
        check(lib().bfgt_tree_export(self._h, itree.ctypes.data_as(_ip), iptr.ctypes.data_as(_ip),
                                     centers.ctypes.data_as(_dp), boxsize.ctypes.data_as(_dp),
                                     vals.ctypes.data_as(_dp)))
        return QuadTree(nl, itree, iptr, centers, boxsize, self.iperiod), vals

    def plan(self, delta: float, eps: float, ifpgh: int, precision: str = "fp64"):
        """Resident box FGT plan on the tree as it is now; run it on
        :meth:`values`."""
        from .boxfgt import BoxFGT

        tree, _ = self.to_host()
        return BoxFGT(tree, self.nd, delta, eps, self.ipoly, self.norder, ifpgh, precision=precision,
                      stream=self.stream)

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            lib().bfgt_tree_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
