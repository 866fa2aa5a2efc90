"""The C-ABI library loads and exports every entry point include/ygrid.h declares (no GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "ygrid.h")).read()
    return sorted(set(re.findall(r"\b(ygrid_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = _declared()
    for must in ("ygrid_initialise_grid", "ygrid_grid", "ygrid_finalise_grid", "ygrid_finalise_weights",
                 "ygrid_initialise_degrid", "ygrid_degrid", "ygrid_gridder_create", "ygrid_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import yandasoft_amd as ya
    L = ya.lib()
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing
    assert set(_declared()) <= set(ya.abi.EXPORTED) | {"ygrid_last_error"}


def test_error_path_without_device_is_reported_not_thrown():
    import yandasoft_amd as ya
    L = ya.lib()
    h = ctypes.c_void_p()
    rc = L.ygrid_ctx_create(1 << 20, ctypes.byref(h))  # no such device anywhere
    assert rc != 0
    assert len(L.ygrid_last_error()) > 0
    assert L.ygrid_version() == 1


def test_null_arguments_are_rejected():
    import yandasoft_amd as ya
    L = ya.lib()
    assert L.ygrid_grid(None, None) == 1
    assert b"null" in L.ygrid_last_error()
