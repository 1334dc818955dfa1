"""Loader of the HIP library ``libboxfgt_hip.so`` (C ABI in include/boxfgt_hip.h).

There is no fallback: if the library is missing or cannot be loaded the
import of the transform raises, naming the file and how to build it.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libboxfgt_hip.so")

BFGT_OK = 0
STATUS = {
    1: "BFGT_ERR_NDIM",
    2: "BFGT_ERR_IPOLY",
    3: "BFGT_ERR_IFPGH",
    4: "BFGT_ERR_IFNEWTREE",
    5: "BFGT_ERR_TARGETS",
    6: "BFGT_ERR_NORDER",
    7: "BFGT_ERR_NPW",
    8: "BFGT_ERR_TREE",
    9: "BFGT_ERR_ARG",
    10: "BFGT_ERR_HIP",
    11: "BFGT_ERR_NODEVICE",
    12: "BFGT_ERR_RESOURCE",
    13: "BFGT_ERR_CALLBACK",
}
PRECISIONS = {"fp64": 0, "split32": 1, "bf16x3": 2}

EXPORTS = [
    "bfgt_hip_eval",
    "bfgt_hip_boxfgt",
    "bfgt_hip_release_plans",
    "bfgt_plan_create",
    "bfgt_plan_execute",
    "bfgt_plan_execute_scaled",
    "bfgt_plan_execute_overwrite",
    "bfgt_plan_execute_noalloc",
    "bfgt_plan_leaf_points",
    "bfgt_plan_adjoint",
    "bfgt_plan_eval_targets",
    "bfgt_plan_destroy",
    "bfgt_plan_query",
    "bfgt_plan_work",
    "bfgt_plan_get_table",
    "bfgt_plan_profile_begin",
    "bfgt_plan_profile_stage",
    "bfgt_plan_profile_end",
    "bfgt_tree_build",
    "bfgt_tree_adapt",
    "bfgt_tree_import",
    "bfgt_tree_query",
    "bfgt_tree_export",
    "bfgt_tree_plan_create",
    "bfgt_tree_destroy",
    "bfgt_last_error",
    "bfgt_device_count",
]

# int fun(void* ctx, int nd, int npts, const double* d_xy, double* d_vals, void* stream)
TREE_FUN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                            ctypes.c_void_p, ctypes.c_void_p)


class BoxFGTError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{STATUS.get(code, code)}: {message}")
        self.code = code


_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"HIP library not built: {LIB_PATH} is missing (run `python -m heat2d_per_box_amd.build` "
            "or __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    L.bfgt_hip_release_plans.restype = _i
    L.bfgt_hip_release_plans.argtypes = []
    L.bfgt_hip_eval.restype = _i
    L.bfgt_hip_eval.argtypes = [_i, _i, _d, _d, _i, _i, _i, _i, _i, _i, _i, _ip, _ip, _dp, _dp, _dp, _i, _dp, _dp,
                                _dp, _i, _i, _dp, _i, _dp, _dp, _dp, _dp, _i]
    L.bfgt_hip_boxfgt.restype = _i
    L.bfgt_hip_boxfgt.argtypes = [_i, _i, _d, _d, _i, _i, _i, _i, _i, _ip, _i, _ip, _ip, _dp, _dp, _dp, _i, _dp,
                                  _dp, _dp, _i, _i, _dp, _i, _dp, _dp, _dp, _dp, _i]
    L.bfgt_plan_create.restype = _i
    L.bfgt_plan_create.argtypes = [ctypes.POINTER(_vp), _i, _i, _d, _d, _i, _i, _i, _i, _i, _i, _ip, _ip, _dp, _dp,
                                   _i, _i, _vp]
    L.bfgt_plan_execute.restype = _i
    L.bfgt_plan_execute.argtypes = [_vp, _vp, _vp, _vp, _vp, _dp]
    L.bfgt_plan_eval_targets.restype = _i
    L.bfgt_plan_eval_targets.argtypes = [_vp, _i, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp]
    L.bfgt_plan_adjoint.restype = _i
    L.bfgt_plan_adjoint.argtypes = [_vp, _vp, _vp, _vp, _vp, _dp]
    L.bfgt_plan_execute_noalloc.restype = _i
    L.bfgt_plan_execute_noalloc.argtypes = [_vp, _vp, _vp, _vp, _vp, _dp]
    L.bfgt_plan_execute_scaled.restype = _i
    L.bfgt_plan_execute_scaled.argtypes = [_vp, _d, _vp, _vp, _vp, _vp, _dp]
    L.bfgt_plan_execute_overwrite.restype = _i
    L.bfgt_plan_execute_overwrite.argtypes = [_vp, _d, _vp, _vp, _vp, _vp, _dp]
    L.bfgt_plan_leaf_points.restype = _i
    L.bfgt_plan_leaf_points.argtypes = [_vp, _vp]
    L.bfgt_plan_destroy.restype = _i
    L.bfgt_plan_destroy.argtypes = [_vp]
    L.bfgt_plan_query.restype = _i
    L.bfgt_plan_query.argtypes = [_vp, _ip, _dp]
    L.bfgt_plan_work.restype = _i
    L.bfgt_plan_work.argtypes = [_vp, ctypes.POINTER(ctypes.c_longlong), _i]
    L.bfgt_plan_get_table.restype = _i
    L.bfgt_plan_get_table.argtypes = [_vp, _i, _i, _vp, ctypes.c_size_t]
    L.bfgt_plan_profile_begin.restype = _i
    L.bfgt_plan_profile_begin.argtypes = [_vp, _i]
    L.bfgt_plan_profile_stage.restype = _i
    L.bfgt_plan_profile_stage.argtypes = [_vp, _i, _i]
    L.bfgt_plan_profile_end.restype = _i
    L.bfgt_plan_profile_end.argtypes = [_vp, _dp, _ip]
    L.bfgt_tree_build.restype = _i
    L.bfgt_tree_build.argtypes = [ctypes.POINTER(_vp), _i, _i, _i, _d, _d, _i, _i, _d, TREE_FUN, _vp, _dp, _i, _i,
                                  _dp, _vp]
    L.bfgt_tree_import.restype = _i
    L.bfgt_tree_import.argtypes = [ctypes.POINTER(_vp), _i, _i, _i, _i, _i, _d, _i, _i, _ip, _ip, _dp, _dp, _dp, _i,
                                   _vp]
    L.bfgt_tree_adapt.restype = _i
    L.bfgt_tree_adapt.argtypes = [_vp, _d, _dp, TREE_FUN, _vp, _i]
    L.bfgt_tree_query.restype = _i
    L.bfgt_tree_query.argtypes = [_vp, _ip, ctypes.POINTER(_vp)]
    L.bfgt_tree_export.restype = _i
    L.bfgt_tree_export.argtypes = [_vp, _ip, _ip, _dp, _dp, _dp]
    L.bfgt_tree_plan_create.restype = _i
    L.bfgt_tree_plan_create.argtypes = [ctypes.POINTER(_vp), _vp, _d, _d, _i, _i, _vp]
    L.bfgt_tree_destroy.restype = _i
    L.bfgt_tree_destroy.argtypes = [_vp]
    L.bfgt_last_error.restype = ctypes.c_char_p
    L.bfgt_last_error.argtypes = []
    L.bfgt_device_count.restype = _i
    L.bfgt_device_count.argtypes = []
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != BFGT_OK:
        raise BoxFGTError(rc, lib().bfgt_last_error().decode(errors="replace"))
