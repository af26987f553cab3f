"""The solid-shape struct of the C ABI (include/simcogpu.h) is declared the
same way in the ctypes binding and the Fortran module: same members in the
same order and types, so a shape passed from Python or Fortran lands in the
fields the library reads."""
import ctypes
import os
import re

from simcoflow_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _c_members():
    text = open(os.path.join(ROOT, "include", "simcogpu.h")).read()
    body = re.search(r"typedef struct sg_solid_shape \{(.*?)\} sg_solid_shape;", text, re.S).group(1)
    out = []
    for line in body.splitlines():
        line = re.sub(r"/\*.*?\*/", "", line).strip().rstrip(";")
        if not line:
            continue
        ctype, names = line.split(None, 1)
        out += [(n.strip(), ctype) for n in names.split(",")]
    return out


def _fortran_members():
    text = open(os.path.join(ROOT, "fortran", "simcogpu_mod.f90")).read()
    body = re.search(r"type, bind\(C\), public :: sg_solid_shape\n(.*?)end type sg_solid_shape", text, re.S).group(1)
    out = []
    for line in body.splitlines():
        m = re.match(r"\s*(integer|real)\((c_int|c_double)\)\s*::\s*(.*)", line)
        if m:
            out += [(n.strip(), {"c_int": "int", "c_double": "double"}[m.group(2)]) for n in m.group(3).split(",")]
    return out


def test_solid_shape_members_agree():
    c = _c_members()
    assert [n for n, _ in c] == ["kind", "cx", "cy", "a", "b", "c"]
    py = [(n, {ctypes.c_int: "int", ctypes.c_double: "double"}[t]) for n, t in _lib.SolidShape._fields_]
    assert py == c
    assert _fortran_members() == c
    assert ctypes.sizeof(_lib.SolidShape) == 8 + 8 * (len(c) - 1)
    assert _lib.SOLID_KINDS == {"none": 0, "box": 1, "circle": 2, "ellipse": 3, "boom": 4}
    text = open(os.path.join(ROOT, "include", "simcogpu.h")).read()
    for name, k in _lib.SOLID_KINDS.items():
        assert re.search(rf"#define SG_SOLID_{name.upper()} {k}\b", text), name
