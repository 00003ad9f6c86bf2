"""ctypes front end of the CPU restatement (oracle/cc_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker; never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libccoracle.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                 C.c_int, C.c_double, C.c_double, C.c_double]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_threads.restype = C.c_int
        L.orc_load.argtypes = [C.c_void_p, _dp, C.c_int, C.c_double]
        L.orc_set_x3_file.argtypes = [C.c_void_p, _dp, C.c_double]
        L.orc_set_ext.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.orc_set_wall.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_set_forcing.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_void_p]
        L.orc_stage.argtypes = [C.c_void_p, C.c_int]
        L.orc_step.argtypes = [C.c_void_p, C.c_int]
        L.orc_step.restype = C.c_int
        L.orc_get_var.argtypes = [C.c_void_p, _dp]
        L.orc_get_solv.argtypes = [C.c_void_p, C.c_int, _dp]
        L.orc_get_mu.argtypes = [C.c_void_p, _dp]
        L.orc_get_nstep.argtypes = [C.c_void_p]
        L.orc_get_nstep.restype = C.c_int
        L.orc_get_time.argtypes = [C.c_void_p]
        L.orc_get_time.restype = C.c_double
        L.orc_get_stats.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_get_mesh.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.orc_get_residual.argtypes = [C.c_void_p, _dp]
        L.orc_get_blown.argtypes = [C.c_void_p, _ip]
        L.orc_diag.argtypes = [C.c_void_p, _dp]
        L.orc_calc_tauwall.argtypes = [C.c_void_p]
        L.orc_avg_spatial.argtypes = [C.c_void_p]
        _lib = L
    return _lib


class OrcWall(C.Structure):
    """Wall condition of one wall (cc_oracle.h orc_wall); zero = reference wall."""
    _fields_ = [("vel", C.c_int), ("stripes", C.c_int), ("temp", C.c_int), ("reserved", C.c_int),
                ("amp", C.c_double), ("kx", C.c_double), ("omega", C.c_double),
                ("phase_odd", C.c_double), ("disc_omega", C.c_double), ("grad", C.c_double),
                ("Tw", C.c_double), ("dTdz", C.c_double)]


class Oracle:
    """One channel instance on the CPU, FP64, reference semantics.

    Arrays use the reference's Fortran order, exposed as C-order numpy arrays
    of shape (ncomp, n3, n2, n1)."""

    def __init__(self, n1, n2, n3, Re=3000.0, Mac=1.5, dt=1e-4, istat=10,
                 Pex=2.0, Pey=2.0, Pez=0.2):
        self.n = (n1, n2, n3)
        self.L = lib()
        self.h = self.L.orc_create(n1, n2, n3, Re, Mac, dt, istat, Pex, Pey, Pez)
        if not self.h:
            raise ValueError("orc_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    @property
    def shape(self):
        n1, n2, n3 = self.n
        return (n3, n2, n1)

    def load(self, var, nstep=0, time=0.0, x3_file=None, x2_end=None):
        var = np.ascontiguousarray(var, dtype=np.float64)
        assert var.shape == (7,) + self.shape
        self.L.orc_load(self.h, var, nstep, time)
        if x3_file is not None:
            self.L.orc_set_x3_file(self.h, np.ascontiguousarray(x3_file, dtype=np.float64),
                                   float(x2_end))

    def set_ext(self, Ri=0.0, Ls=0.0):
        self.L.orc_set_ext(self.h, Ri, Ls)

    def set_wall(self, wall, spec=None, x1=None, x2=None):
        """Wall condition of wall 0 (bottom) or 1 (top): spec is a dict of
        OrcWall fields (None = the reference wall); x1, x2 the solution
        file's coordinates (default: the analytic mesh)."""
        w = OrcWall(**(spec or {}))
        keep = []
        def ptr(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64)
            keep.append(a)
            return a.ctypes.data
        self.L.orc_set_wall(self.h, int(wall), C.byref(w), ptr(x1), ptr(x2))

    def set_forcing(self, tauL, tauU, rhobar=None):
        if rhobar is None:
            self.L.orc_set_forcing(self.h, tauL, tauU, None)
        else:
            rb = np.ascontiguousarray(rhobar, dtype=np.float64)
            self.L.orc_set_forcing(self.h, tauL, tauU, rb.ctypes.data_as(C.c_void_p))

    def stage(self, it):
        self.L.orc_stage(self.h, it)

    def step(self, n=1):
        return self.L.orc_step(self.h, n)

    def var(self):
        out = np.empty((7,) + self.shape)
        self.L.orc_get_var(self.h, out)
        return out

    def solv(self, level=0):
        out = np.empty((5,) + self.shape)
        self.L.orc_get_solv(self.h, level, out)
        return out

    def mu(self):
        out = np.empty(self.shape)
        self.L.orc_get_mu(self.h, out)
        return out

    def residual(self):
        out = np.empty((5,) + self.shape)
        self.L.orc_get_residual(self.h, out)
        return out

    @property
    def nstep(self):
        return self.L.orc_get_nstep(self.h)

    @property
    def time(self):
        return self.L.orc_get_time(self.h)

    def stats(self):
        n3 = self.n[2]
        tau = np.empty(2)
        s1 = np.empty((7, n3))
        s2 = np.empty((7, n3))
        cr = np.empty((9, n3))
        self.L.orc_get_stats(self.h, tau, s1, s2, cr)
        return tau, s1, s2, cr

    def mesh(self):
        n3 = self.n[2]
        coef = np.empty(12)
        mesh = np.empty(4 * n3)
        scal = np.empty(8)
        self.L.orc_get_mesh(self.h, coef, mesh, scal)
        return coef, mesh.reshape(4, n3), scal

    def blown(self):
        ijk = np.zeros(3, dtype=np.int32)
        self.L.orc_get_blown(self.h, ijk)
        return tuple(int(x) for x in ijk)

    def diag(self):
        out = np.empty(16)
        self.L.orc_diag(self.h, out)
        return out

    def calc_tauwall(self):
        self.L.orc_calc_tauwall(self.h)

    def avg_spatial(self):
        self.L.orc_avg_spatial(self.h)
