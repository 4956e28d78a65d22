"""The C-ABI library loads and exports every function include/mdx.h declares
(no device calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT
from mdacp_amd import mdx

HEADERS = [os.path.join(ROOT, "include", h) for h in sorted(os.listdir(os.path.join(ROOT, "include")))
           if h.endswith(".h")]


def declared(header):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(mdx\w+)\s*\(", text, flags=re.M)


def test_headers_declare_entry_points():
    names = [n for h in HEADERS for n in declared(h)]
    assert len(names) >= 20
    assert set(mdx.EXPORTS) <= set(names)


def test_library_exports_every_declared_symbol():
    lib = mdx.load()
    missing = [n for h in HEADERS for n in declared(h) if not hasattr(lib, n)]
    assert not missing, missing


def test_library_is_gfx950_code_object():
    # the offload bundle of our own build names its target triple
    blob = open(mdx.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_default_params_without_device():
    lib = mdx.load()
    p = mdx.MdxParams()
    lib.mdx_default_params(ctypes.byref(p))
    assert (p.dt, p.cutoff, p.buffer_length, p.check_list_length) == (0.001, 3.0, 0.3, 500)


def test_create_rejects_small_box_before_touching_device():
    lib = mdx.load()
    p = mdx.MdxParams()
    lib.mdx_default_params(ctypes.byref(p))
    for d in range(3):
        p.L[d] = 5.0
    h = ctypes.c_void_p()
    assert lib.mdx_create(ctypes.byref(p), ctypes.byref(h)) == -5
    assert b"search length" in lib.mdx_last_error(None)


def test_ensemble_create_rejects_small_box_before_touching_device():
    lib = mdx.load()
    p = mdx.MdxParams()
    lib.mdx_default_params(ctypes.byref(p))
    n = (ctypes.c_int * 2)(500, 500)
    L3 = (ctypes.c_double * 6)(10.0, 10.0, 10.0, 10.0, 5.0, 10.0)
    h = ctypes.c_void_p()
    assert lib.mdx_ensemble_create(ctypes.byref(p), 2, n, L3, ctypes.byref(h)) == -5
    assert b"replica 1" in lib.mdx_ensemble_last_error(None)
    p.check_list_length = 5000
    L3[4] = 10.0
    assert lib.mdx_ensemble_create(ctypes.byref(p), 2, n, L3, ctypes.byref(h)) == -1
