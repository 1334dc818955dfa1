"""The C ABI library loads and exports every symbol include/boxfgt_hip.h declares;
argument errors are reported with the documented status codes before any
device work (these run without a GPU)."""
import os
import re

import numpy as np
import pytest

from heat2d_per_box_amd import _lib
from heat2d_per_box_amd.tree import build_uniform

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            src = open(os.path.join(ROOT, "include", fn)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names += re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(bfgt_\w+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_functions()
    assert "bfgt_hip_eval" in names and "bfgt_plan_execute" in names
    for n in names:
        assert hasattr(L, n), n
    assert sorted(_lib.EXPORTS) == names


def _eval(**over):
    t = build_uniform(2)
    n = 4
    nb = t.nboxes
    fv = np.zeros((nb, n * n, 1))
    pot = np.zeros_like(fv)
    a = dict(nd=1, ndim=2, delta=1e-3, eps=1e-6, ipoly=0, iperiod=1, norder=n, npbox=n * n, nboxes=nb,
             nlevels=t.nlevels, ltree=t.ltree, ifpgh=1, ifnewtree=0, ifpghtarg=0)
    a.update(over)
    L = _lib.lib()
    import ctypes

    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    return L.bfgt_hip_eval(a["nd"], a["ndim"], a["delta"], a["eps"], a["ipoly"], a["iperiod"], a["norder"],
                           a["npbox"], a["nboxes"], a["nlevels"], a["ltree"], t.itree.ctypes.data_as(ip),
                           t.iptr.ctypes.data_as(ip), t.centers.ctypes.data_as(dp), t.boxsize.ctypes.data_as(dp),
                           fv.ctypes.data_as(dp), a["ifpgh"], pot.ctypes.data_as(dp), None, None, a["ifnewtree"], 0,
                           None, a["ifpghtarg"], None, None, None, None, 0)


@pytest.mark.parametrize("over,code", [
    ({"ifnewtree": 1}, 4),
    ({"ifpghtarg": 4}, 5),
    ({"ifpghtarg": -1}, 5),
    ({"npbox": 15}, 6),
    ({"ndim": 3}, 1),
    ({"ipoly": 2}, 2),
    ({"ifpgh": 4}, 3),
    ({"norder": 40, "npbox": 1600}, 6),
])
def test_invalid_requests_are_refused(over, code):
    assert _eval(**over) == code
    assert len(_lib.lib().bfgt_last_error()) > 0
