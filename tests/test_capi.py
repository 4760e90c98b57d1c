"""The C ABI: the library loads, exports every symbol include/adrt_hip.h
declares, and its shape rules match the reference's
(adrt_cdefs_common.cpp:140-330).  CPU-only except the marked launch tests."""

import ctypes
import os
import re

import numpy as np
import pytest

from adrt_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "adrt_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(adrt_hip_\w+)\s*\(", text)))


def test_library_loads():
    lib = _lib.load()
    assert lib.adrt_hip_abi_version() >= 100


def test_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 20
    raw = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in names if not hasattr(raw, n)]
    assert not missing, missing


def test_signature_table_covers_header():
    assert set(declared_functions()) == set(_lib.SIGNATURES)


def test_status_strings():
    lib = _lib.load()
    for code in range(9):
        assert lib.adrt_hip_status_string(code)
    assert lib.adrt_hip_status_string(0) == b"ok"


@pytest.mark.parametrize("n", [0, 1, 2, 3, 4, 5, 8, 9, 16, 17, 2**32, 2**32 + 1, 2**63, 2**64 - 1])
def test_num_iters_matches_closed_form(n):
    expect = n.bit_length() - (bin(n).count("1") == 1)
    assert _lib.load().adrt_hip_num_iters(n) == expect


def test_adrt_shape_rule():
    lib = _lib.load()
    ok = lambda s: bool(lib.adrt_hip_valid_adrt_shape(_lib.shape3(s)))  # noqa: E731
    assert ok((1, 1, 1)) and ok((5, 16, 16)) and ok((1, 2**20, 2**20))
    assert not ok((1, 16, 32)) and not ok((1, 31, 31)) and not ok((0, 16, 16))
    assert not ok((1, 0, 0))


def test_output_shape_rule():
    lib = _lib.load()
    ok = lambda s: bool(lib.adrt_hip_valid_output_shape(_lib.shape4(s)))  # noqa: E731
    assert ok((1, 4, 31, 16)) and ok((3, 4, 1, 1))
    assert not ok((1, 4, 31, 15)) and not ok((1, 4, 30, 16)) and not ok((1, 3, 31, 16))
    assert not ok((1, 4, 33, 17)) and not ok((0, 4, 31, 16))


def test_step_rule():
    lib = _lib.load()
    s = _lib.shape4((1, 4, 31, 16))
    assert [bool(lib.adrt_hip_valid_step(s, k)) for k in range(-1, 5)] == [
        False, True, True, True, True, False]


def test_fmg_shape_rules():
    lib = _lib.load()
    r = lambda s: bool(lib.adrt_hip_valid_restriction_shape(_lib.shape4(s)))  # noqa: E731
    assert r((1, 4, 11, 6)) and r((2, 4, 3, 2)) and not r((1, 4, 5, 3)) and not r((1, 2, 11, 6))
    assert not r((1, 4, 10, 6)) and not r((1, 4, 1, 1))
    p = lambda s: bool(lib.adrt_hip_valid_prolongation_shape(_lib.shape3(s)))  # noqa: E731
    assert p((1, 1, 1)) and not p((1, 0, 3)) and not p((1, 2**63, 1))
    h = lambda s: bool(lib.adrt_hip_valid_highpass_shape(_lib.shape3(s)))  # noqa: E731
    assert h((1, 2, 2)) and not h((1, 1, 2)) and not h((1, 2, 1))


def test_entry_points_reject_before_launch():
    """Invalid arguments return status codes without touching a device."""
    lib = _lib.load()
    z = ctypes.c_void_p(0)
    assert lib.adrt_hip_adrt(z, z, z, 0, 0, 1, 3, z) == _lib.ERR_SHAPE
    assert lib.adrt_hip_adrt(z, z, z, 0, 7, 1, 4, z) == _lib.ERR_DTYPE
    assert lib.adrt_hip_adrt(z, z, z, 0, 0, 1, 4, z) == _lib.ERR_NULL
    assert lib.adrt_hip_adrt_step(z, z, 0, 1, 16, 4, z) == _lib.ERR_STEP
    assert lib.adrt_hip_bdrt_step(z, z, 0, 1, 16, -1, z) == _lib.ERR_STEP
    assert lib.adrt_hip_iadrt(z, z, z, 0, 0, 1, 12, z) == _lib.ERR_SHAPE
    assert lib.adrt_hip_fmg_restriction(z, z, 0, 1, 1, z) == _lib.ERR_SHAPE
    assert lib.adrt_hip_fmg_highpass(z, z, 0, 1, 1, 5, z) == _lib.ERR_SHAPE
    assert lib.adrt_hip_truncate_combine(z, z, 0, 1, 4, 1.0, 3, z) == _lib.ERR_ARG
    assert lib.adrt_hip_bdrt_rows(z, z, z, 0, 0, 1, 12, 5, z) == _lib.ERR_SHAPE
    assert lib.adrt_hip_bdrt_rows(z, z, z, 0, 0, 1, 16, 0, z) == _lib.ERR_ARG
    assert lib.adrt_hip_bdrt_rows(z, z, z, 0, 0, 1, 16, 32, z) == _lib.ERR_ARG
    assert lib.adrt_hip_bdrt_rows(z, z, z, 0, 0, 1, 16, 31, z) == _lib.ERR_NULL
    one = ctypes.c_void_p(1)
    need = lib.adrt_hip_workspace_bytes(_lib.OP_ADRT, 0, 2, 128)
    assert need > 0
    assert lib.adrt_hip_adrt(one, one, z, 0, 0, 2, 128, z) == _lib.ERR_WORKSPACE
    assert lib.adrt_hip_workspace_bytes(_lib.OP_ADRT, 0, 2, 64) == 0  # LDS kernel


def test_workspace_sizes_scale():
    lib = _lib.load()
    for op in (_lib.OP_ADRT, _lib.OP_BDRT, _lib.OP_IADRT):
        a = lib.adrt_hip_workspace_bytes(op, 0, 1, 256)
        b = lib.adrt_hip_workspace_bytes(op, 1, 2, 256)
        assert 0 < a <= b
    assert lib.adrt_hip_workspace_bytes(9, 0, 1, 16) == 0
    assert lib.adrt_hip_workspace_bytes(0, 0, 1, 15) == 0
