"""The C ABI library (CPU checks: no kernel is launched).

The library must load and export every function include/interax_hip.h declares, and the Python binding must declare
exactly those functions.
"""
import ctypes
import os
import re

import pytest

import interax_hip as ih

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "interax_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ixh_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("ixh_pair_create", "ixh_pair_execute", "ixh_pair_set_parameters", "ixh_threebody_create",
                     "ixh_threebody_execute", "ixh_threebody_set_parameters", "ixh_last_error"):
        assert required in names


def test_library_exports_every_declared_symbol():
    lib = ih.load_library()
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_binding_covers_the_header():
    assert sorted(ih.SIGNATURES) == declared_functions()


def test_version_without_a_device():
    lib = ih.load_library()
    major, minor = ctypes.c_int(), ctypes.c_int()
    assert lib.ixh_version(ctypes.byref(major), ctypes.byref(minor)) == 0
    assert (major.value, minor.value) == (0, 1)


def test_invalid_arguments_are_rejected_before_any_device_work():
    lib = ih.load_library()
    handle = ctypes.c_void_p()
    assert lib.ixh_pair_create(0, None, ctypes.byref(handle)) == 1
    assert b"null descriptor" in lib.ixh_last_error()
    d = ih.PairDesc()
    d.num_particles = 2
    d.terms = 0
    assert lib.ixh_pair_create(0, ctypes.byref(d), ctypes.byref(handle)) == 1
    d.terms = ih.IXH_TERM_EXCHANGE
    d.method = 5
    assert lib.ixh_pair_create(0, ctypes.byref(d), ctypes.byref(handle)) == 1
    assert lib.ixh_last_error() == b"InteraxPairForce: Illegal value for nonbonded method"
    # the state and list-check calls reject missing handles the same way
    assert lib.ixh_state_set_stream(None, None) == 1
    assert b"ixh_state_set_stream" in lib.ixh_last_error()
    valid = ctypes.c_int()
    assert lib.ixh_pair_list_was_valid(None, ctypes.byref(valid)) == 1
    assert b"null argument" in lib.ixh_last_error()
