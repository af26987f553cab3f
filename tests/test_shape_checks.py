"""Solid shapes are checked before any kernel draws them: sg_geometry_rebuild
and sg_set_moving_solid reject the same bad shapes with SG_ERR_ARG and a
message naming the call (include/simcogpu.h).  The checks come before the
context is used, so a null context shows them without a GPU."""
import ctypes
import math

import pytest

from simcoflow_amd import _lib

SG_ERR_ARG = _lib.SG_ERR_ARG

BAD = [
    ("unknown kind", dict(kind=7, cx=0.0, cy=0.0, a=1.0, b=1.0, c=0.0), "known kind"),
    ("centre not finite", dict(kind=2, cx=math.nan, cy=0.0, a=1.0, b=0.0, c=0.0), "finite centre"),
    ("radius zero", dict(kind=2, cx=0.0, cy=0.0, a=0.0, b=0.0, c=0.0), "positive sizes"),
    ("box without height", dict(kind=1, cx=0.0, cy=0.0, a=1.0, b=0.0, c=0.0), "positive sizes"),
    ("boom bar too wide", dict(kind=4, cx=0.0, cy=0.0, a=0.1, b=0.2, c=0.5), "b < radius a"),
    ("boom bar negative", dict(kind=4, cx=0.0, cy=0.0, a=0.2, b=0.1, c=-0.5), "c >= 0"),
]


def _shape(d):
    s = _lib.SolidShape()
    for k, v in d.items():
        setattr(s, k, v)
    return s


@pytest.mark.parametrize("name,shape,msg", BAD, ids=[b[0] for b in BAD])
def test_bad_shapes_rejected_by_both_calls(name, shape, msg):
    lib = _lib.load()
    s = _shape(shape)
    assert lib.sg_geometry_rebuild(None, ctypes.byref(s)) == SG_ERR_ARG
    e = _lib.last_error()
    assert e.startswith("sg_geometry_rebuild") and msg in e, e
    assert lib.sg_set_moving_solid(None, ctypes.byref(s), 0.0, 0.0, 0.0, 0.0) == SG_ERR_ARG
    e = _lib.last_error()
    assert e.startswith("sg_set_moving_solid") and msg in e, e


def test_good_shape_reaches_the_context_check():
    lib = _lib.load()
    s = _shape(dict(kind=4, cx=0.5, cy=0.5, a=0.2, b=0.05, c=0.3))
    assert lib.sg_geometry_rebuild(None, ctypes.byref(s)) == SG_ERR_ARG
    assert "null context" in _lib.last_error()
    assert lib.sg_set_moving_solid(None, ctypes.byref(s), 0.0, 0.0, 0.0, 0.0) == SG_ERR_ARG
    assert "null context" in _lib.last_error()


@pytest.mark.parametrize("k", range(4))
def test_moving_solid_rejects_non_finite_motion(k):
    lib = _lib.load()
    s = _shape(dict(kind=2, cx=0.5, cy=0.5, a=0.2, b=0.0, c=0.0))
    motion = [0.0] * 4
    motion[k] = math.inf if k % 2 else math.nan
    assert lib.sg_set_moving_solid(None, ctypes.byref(s), *motion) == SG_ERR_ARG
    assert "must be finite" in _lib.last_error()
