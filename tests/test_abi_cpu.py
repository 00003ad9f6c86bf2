"""CPU-side checks of the C-ABI library: it loads, exports every function
include/cc_abi.h declares, and reports a clean error without a GPU."""
import ctypes as C
import os
import re

import pytest

from compchan3dmpi_amd import channel

HDR = os.path.join(os.path.dirname(os.path.dirname(__file__)), "include", "cc_abi.h")


def declared():
    txt = open(HDR).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(cc_[a-z_]+)\s*\(", txt)))


def test_header_matches_binding_list():
    assert declared() == sorted(channel.ABI_SYMBOLS)


def test_library_exports_every_symbol():
    L = channel.load_library()
    for name in declared():
        assert hasattr(L, name), name
    assert L.cc_abi_version() == 1


def test_struct_layout_matches_header():
    # 6 ints then 8 doubles, no padding surprises
    assert C.sizeof(channel.cc_params) == 6 * 4 + 8 * 8
    assert channel.cc_params.Re.offset == 24


def test_create_rejects_bad_arguments():
    L = channel.load_library()
    h = C.c_void_p()
    p = channel.cc_params(4, 8, 8, 10, 1, 0, 3000.0, 1.5, 1e-4, 2.0, 2.0, 0.2, 0.0, 0.0)
    assert L.cc_create(C.byref(p), C.byref(h)) < 0
    assert b"grid too small" in L.cc_last_error()
    p = channel.cc_params(12, 8, 16, 10, 7, 0, 3000.0, 1.5, 1e-4, 2.0, 2.0, 0.2, 0.0, 0.0)
    assert L.cc_create(C.byref(p), C.byref(h)) < 0
    # a plane of 2^31 or more unique nodes: the kernels index within a plane
    # in 32-bit arithmetic
    p = channel.cc_params(50002, 50002, 16, 10, 1, 0, 3000.0, 1.5, 1e-4, 2.0, 2.0, 0.2, 0.0, 0.0)
    assert L.cc_create(C.byref(p), C.byref(h)) < 0
    assert b"grid too large" in L.cc_last_error()
    # (n1-2)(n2-2) = 46340^2 < 2^31 passes the size check: with dt = 0 the
    # next check rejects it, before any device call
    p = channel.cc_params(46342, 46342, 16, 10, 1, 0, 3000.0, 1.5, 0.0, 2.0, 2.0, 0.2, 0.0, 0.0)
    assert L.cc_create(C.byref(p), C.byref(h)) < 0
    assert b"must be positive" in L.cc_last_error()


def test_no_device_is_an_error_not_a_fallback():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is visible")
    with pytest.raises(channel.CCError):
        channel.Channel(12, 8, 16)
