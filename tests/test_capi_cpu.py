"""CPU checks of the C ABI boundary: the built HIP library loads and exports
every symbol that include/nsfv1.h declares (no device call is made)."""
import ctypes
import os
import re

import pytest

import plugin_navierstokes_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "nsfv1.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+char\s*\*|int)\s+(nsfv1_\w+)\s*\(", txt, re.M)))


def test_header_lists_entry_points():
    syms = declared_symbols()
    assert "nsfv1_assemble" in syms and "nsfv1_create" in syms
    assert set(syms) == set(P.EXPORTS), set(syms) ^ set(P.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(P.LIB_PATH):
        pytest.fail("libnsfv1.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(P.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_upwind_and_stabilization_keys_documented():
    # the string keys of upwind_interface.cpp:46-61 and stabilization.cpp:49-56, 91-99
    txt = open(os.path.join(ROOT, "include", "nsfv1.h")).read()
    for key in ["no", "full", "skewed", "lps", "fields", "flow", "raw", "fivepoint", "cor"]:
        assert key in txt


def test_function_count_checked_before_any_device_call():
    """navier_stokes_fv1.cpp:66-68: the constructor refuses a function list
    of the wrong length; the check precedes every HIP call, so it also runs
    without a GPU."""
    if not os.path.exists(P.LIB_PATH):
        pytest.fail("libnsfv1.so not built (run __graft_entry__.build())")
    L = P.load_library()
    h = ctypes.c_void_p()
    rc = L.nsfv1_create(2, b"u,v,w,p", b"Inner", ctypes.byref(h))
    assert rc == -2 and not h.value
    assert b"needs exactly 3 symbolic function" in L.nsfv1_last_error()
    rc = L.nsfv1_create(4, b"u,v,w,p", b"Inner", ctypes.byref(h))
    assert rc == -1 and b"dim must be 2 or 3" in L.nsfv1_last_error()
