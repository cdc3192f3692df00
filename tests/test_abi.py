"""The C-ABI libraries load and export every symbol their headers declare
(CPU only: no compute call is made without a GPU)."""
import ctypes
import os
import re

import pytest

from rchol_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rchol_\w+)\s*\(", text)))


def test_engine_exports_every_declared_symbol():
    lib = _lib.engine()
    names = declared("rchol_amd.h")
    assert len(names) >= 14
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_lib.ENGINE_SYMBOLS)


def test_factor_library_exports_every_declared_symbol():
    lib = _lib.factor()
    names = declared("rchol_factor.h")
    assert len(names) >= 8
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_lib.FACTOR_SYMBOLS)


def test_engine_reports_errors_without_exceptions():
    lib = _lib.engine()
    if lib.rchol_amd_device_count() > 0:
        pytest.skip("a GPU is visible; the no-device path is not reachable")
    h = ctypes.c_void_p()
    rc = lib.rchol_amd_create(0, ctypes.byref(h))
    assert rc == -2
    assert b"no HIP device" in lib.rchol_amd_last_error()
    with pytest.raises(RuntimeError):
        _lib.check(rc)


def test_null_arguments_are_rejected():
    lib = _lib.engine()
    assert lib.rchol_amd_upload_A(None, 4, None, None, None) == -1
    assert lib.rchol_amd_precond(None, None, None) == -1
    assert lib.rchol_amd_set_option(None, 1, 0.0) == -1
    assert lib.rchol_amd_update_A_values(None, 4, None) == -1
