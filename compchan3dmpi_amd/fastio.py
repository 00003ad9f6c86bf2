"""The reference's formatted point files through the native parallel
reader/writer of libcc_hip.so (csrc/cc_io.cpp).

Solution files 3Dnnn.dat (main_v1.9.f90:759-773), fluc.dat (fluc.f90:64-80)
and qInvariant.dat (qInvariant.f90:84-96) are written byte-identical to the
Fortran runtime's output; readers accept any whitespace-separated layout of
the same records (the reference's own fixed-column files included).
Arrays use the reference's Fortran order exposed as C-order NumPy arrays:
fields (nf, n3, n2, n1).  Host-only: no device is needed.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .channel import CCError, load_library

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ready = False


def _lib():
    global _ready
    L = load_library()
    if not _ready:
        L.cc_io_last_error.restype = C.c_char_p
        L.cc_write_points.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int, _dp, _dp,
                                      _dp, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int]
        L.cc_read_points.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_int]
        L.cc_write_solution.argtypes = [C.c_char_p, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int,
                                        _dp, _dp, _dp, _dp, C.c_void_p, C.c_int]
        L.cc_read_solution.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp,
                                       C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_int]
        L.cc_solution_dims.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                       C.POINTER(C.c_int)]
        L.cc_format_field.argtypes = [C.c_double, C.c_int, C.c_char_p]
        _ready = True
    return L


def _check(L, rc):
    if rc < 0:
        raise CCError(f"cc io error {rc}: {L.cc_io_last_error().decode()}")
    return rc


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def write_solution(path, time, nstep, x1, x2, x3, var, pattern=None, nthreads=0):
    """A solution file (the reference's 3D file, with the "Pattern" column of
    the study programs when pattern (n3, n2, n1) is given)."""
    L = _lib()
    var = _f64(var)
    _, n3, n2, n1 = var.shape
    pat = None
    if pattern is not None:
        pat = np.ascontiguousarray(pattern, dtype=np.int32)
    _check(L, L.cc_write_solution(str(path).encode(), float(time), int(nstep), n1, n2, n3,
                                  _f64(x1), _f64(x2), _f64(x3), var,
                                  pat.ctypes.data if pat is not None else None, nthreads))


def read_solution(path, nthreads=0):
    """(time, nstep, x1, x2, x3, var (7, n3, n2, n1)) of a solution file;
    the grid comes from its ZONE line."""
    L = _lib()
    n1, n2, n3 = dims(path)
    x1, x2, x3 = np.empty(n1), np.empty(n2), np.empty(n3)
    var = np.empty((7, n3, n2, n1))
    t = C.c_double()
    ns = C.c_int()
    _check(L, L.cc_read_solution(str(path).encode(), n1, n2, n3, x1, x2, x3, var, C.byref(t),
                                 C.byref(ns), nthreads))
    return t.value, ns.value, x1, x2, x3, var


def dims(path):
    """(n1, n2, n3) from a solution file's ZONE k=,j=,i= line."""
    L = _lib()
    a, b, c = C.c_int(), C.c_int(), C.c_int()
    _check(L, L.cc_solution_dims(str(path).encode(), C.byref(a), C.byref(b), C.byref(c)))
    return a.value, b.value, c.value


def write_points(path, header, x1, x2, x3, fields, blank_records=True, nthreads=0):
    """Records (3(1X,F10.6),nf(1X,E22.15)) in k, j, i order after a header
    (the caller's lines, each ending in a newline); blank_records puts an
    empty record before every k and every j as the reference's writers do."""
    L = _lib()
    f = _f64(fields)
    nf, n3, n2, n1 = f.shape
    _check(L, L.cc_write_points(str(path).encode(), header.encode(), n1, n2, n3, _f64(x1),
                                _f64(x2), _f64(x3), nf, f.ctypes.data, None,
                                1 if blank_records else 0, nthreads))


def read_points(path, n1, n2, n3, nf, nheader=3, nthreads=0):
    """(x1, x2, x3, fields (nf, n3, n2, n1)) of a point file."""
    L = _lib()
    x1, x2, x3 = np.empty(n1), np.empty(n2), np.empty(n3)
    f = np.empty((nf, n3, n2, n1))
    _check(L, L.cc_read_points(str(path).encode(), nheader, n1, n2, n3, nf, 0, x1.ctypes.data,
                               x2.ctypes.data, x3.ctypes.data, f.ctypes.data, nthreads))
    return x1, x2, x3, f


def format_field(x, kind="E22.15"):
    """One value as the writer formats it ("E22.15" or "F10.6")."""
    L = _lib()
    b = C.create_string_buffer(32)
    _check(L, L.cc_format_field(float(x), 0 if kind == "E22.15" else 1, b))
    return b.value.decode()
