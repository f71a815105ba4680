"""The C-ABI library loads on a host without a GPU and exports every entry
point include/pdekit_rd.h declares; argument errors come back as status codes."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from pdekit_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "pdekit_rd.h")).read()
    return sorted(set(re.findall(r"\b(pdrd_[A-Za-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert sorted(L.EXPORTED) == _declared()


def test_every_declared_symbol_is_exported():
    lib = L.lib()
    for name in _declared():
        assert hasattr(lib, name), name


def test_version_string():
    assert b"gfx950" in L.lib().pdrd_version()


def test_bad_arguments_return_status():
    lib = L.lib()
    d = L.FeDesc()
    rc = lib.pdrd_fe_describe(7, 1, 1, 4, 1, C.byref(d))
    assert rc == -1 and b"shape" in lib.pdrd_last_error()
    rc = lib.pdrd_fe_describe(0, 1, 1, 99, 1, C.byref(d))
    assert rc == -1 and b"quadrature" in lib.pdrd_last_error()
    h = C.c_void_p()
    coords = np.zeros((3, 2))
    # physics of the wrong dimension
    rc = lib.pdrd_create(2, 3, L.ptr(coords), 3, L.EULER3D, L.LDA, None, 0, C.byref(h))
    assert rc == -1
    # unknown scheme
    rc = lib.pdrd_create(2, 3, L.ptr(coords), 3, L.BURGERS2D, 9, None, 0, C.byref(h))
    assert rc == -1
    with pytest.raises(L.PdrdError):
        L.check(rc, "pdrd_create")


def test_cells_out_of_range_are_rejected():
    lib = L.lib()
    h = C.c_void_p()
    coords = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])
    assert lib.pdrd_create(2, 3, L.ptr(coords), 3, L.BURGERS2D, L.B, None, 0, C.byref(h)) == 0
    d = L.FeDesc()
    assert lib.pdrd_fe_describe(L.TRIAG, 1, 1, 4, L.WARPBLEND, C.byref(d)) == 0
    tab = np.zeros(d.table_size)
    assert lib.pdrd_fe_fill(C.byref(d), L.ptr(tab)) == 0
    g = np.array([[0, 1, 5]], dtype=np.uint32)
    rc = lib.pdrd_add_cells(h, C.byref(d), L.ptr(tab), 1, L.ptr(g, C.c_uint32), L.ptr(g, C.c_uint32))
    assert rc == -1 and b"range" in lib.pdrd_last_error()
    # no cells: initialize refuses before touching the device
    assert lib.pdrd_initialize(h) == -1
    assert lib.pdrd_destroy(h) == 0
