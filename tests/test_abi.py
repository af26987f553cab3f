"""The C ABI library loads and exports every entry point declared in
include/simcogpu.h (no device call is made: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "simcogpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sg_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("sg_ctx_create", "sg_poisson_solve", "sg_poisson_solve_resident", "sg_velocity_correct_resident"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    from simcoflow_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libsimcogpu.so not built (python -m simcoflow_amd.build)")
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    # the Python binding declares a prototype for each of them
    assert set(declared_symbols()) <= set(_lib.PROTOTYPES) | {"sg_time_kernel"}


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from simcoflow_amd import _lib
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(ImportError):
        _lib.load()
