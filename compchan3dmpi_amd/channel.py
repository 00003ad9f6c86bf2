"""Python host side of the MI355X channel time step: a ctypes binding of the
C ABI in include/cc_abi.h (libcc_hip.so, built in-tree).

The binding mirrors the reference's step seam (v1.10/codes/main_v1.10.f90:
37-53): upload the state read from a solution file, advance it, read the
statistics that the reference appends to tauWall.dat / moment*.dat, and
download var(n1,n2,n3,7) for the solution writer.  It refuses to run when
the HIP library is missing; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcc_hip.so")
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include", "cc_abi.h")

FP64_STRICT, FP64, MIXED = 0, 1, 2
PRECISIONS = {"fp64-strict": FP64_STRICT, "fp64": FP64, "mixed": MIXED}

# Every function the C ABI exports (include/cc_abi.h).
ABI_SYMBOLS = ("cc_create", "cc_upload", "cc_step", "cc_stats", "cc_diag", "cc_diag_noalloc",
               "cc_download",
               "cc_blown", "cc_profile", "cc_tiling", "cc_set_tiling", "cc_destroy", "cc_last_error",
               "cc_abi_version",
               "cc_batch_create", "cc_batch_member", "cc_batch_step", "cc_batch_tune",
               "cc_batch_destroy", "cc_mlp_nparams", "cc_mlp_create", "cc_mlp_set_params",
               "cc_mlp_get_params", "cc_mlp_forward", "cc_mlp_grad", "cc_mlp_train",
               "cc_mlp_destroy", "cc_mlp_last_error", "cc_filter_samples", "cc_set_closure",
               "cc_set_wall", "cc_wall_preset", "cc_average", "cc_postprocess",
               "cc_pp_fluctuation", "cc_pp_qcriterion", "cc_write_solution", "cc_read_solution",
               "cc_solution_dims", "cc_write_points", "cc_read_points", "cc_format_field",
               "cc_io_last_error")

# Study names accepted by cc_wall_preset: the reference's wall-condition
# parameter studies (v1.9_STWNV, v1.9_RD, v1.9_groov, v1.9_groov_strat,
# v1.9_strat).
STUDIES = ("s1", "s2", "s3", "s4", "s5", "s6", "rd1", "rd2", "rd3", "g1", "gst1", "gst2",
           "gst3", "gst4", "gst5", "st1", "st2", "st3", "st4")


class cc_params(C.Structure):
    _fields_ = [("n1", C.c_int), ("n2", C.c_int), ("n3", C.c_int), ("istat", C.c_int),
                ("precision", C.c_int), ("reserved", C.c_int),
                ("Re", C.c_double), ("Mac", C.c_double), ("dt", C.c_double),
                ("PexVal", C.c_double), ("PeyVal", C.c_double), ("PezVal", C.c_double),
                ("Ri_b", C.c_double), ("slip_len", C.c_double)]


class cc_wall(C.Structure):
    """Wall condition (include/cc_abi.h cc_wall); zero = the reference wall."""
    _fields_ = [("vel", C.c_int), ("stripes", C.c_int), ("temp", C.c_int), ("reserved", C.c_int),
                ("amp", C.c_double), ("kx", C.c_double), ("omega", C.c_double),
                ("phase_odd", C.c_double), ("disc_omega", C.c_double), ("grad", C.c_double),
                ("Tw", C.c_double), ("dTdz", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class CCError(RuntimeError):
    pass


def build() -> str:
    """Compile libcc_hip.so for gfx950 (hipcc, in-tree)."""
    subprocess.run(["make", "-s", "-C", CSRC], check=True)
    return LIB_PATH


_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def load_library():
    """Load libcc_hip.so; raise if it is not built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise CCError(f"{LIB_PATH} is not built; run compchan3dmpi_amd.channel.build()")
    L = C.CDLL(LIB_PATH)
    L.cc_create.argtypes = [C.POINTER(cc_params), C.POINTER(C.c_void_p)]
    L.cc_upload.argtypes = [C.c_void_p, _dp, C.c_int, C.c_double]
    L.cc_step.argtypes = [C.c_void_p, C.c_int]
    L.cc_stats.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
    L.cc_diag.argtypes = [C.c_void_p, _dp, C.c_double, _dp]
    L.cc_diag_noalloc.argtypes = [C.c_void_p, _dp, C.c_double, _dp]
    L.cc_download.argtypes = [C.c_void_p, _dp, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.cc_blown.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
    L.cc_profile.argtypes = [C.c_void_p, C.c_int, _dp]
    L.cc_tiling.argtypes = [C.c_void_p, C.c_void_p]
    L.cc_set_tiling.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.cc_destroy.argtypes = [C.c_void_p]
    L.cc_destroy.restype = None
    L.cc_batch_create.argtypes = [C.POINTER(cc_params), C.c_int, C.POINTER(C.c_void_p)]
    L.cc_batch_member.argtypes = [C.c_void_p, C.c_int]
    L.cc_batch_member.restype = C.c_void_p
    L.cc_batch_step.argtypes = [C.c_void_p, C.c_int]
    L.cc_batch_tune.argtypes = [C.c_void_p, C.c_int, _dp]
    L.cc_batch_destroy.argtypes = [C.c_void_p]
    L.cc_batch_destroy.restype = None
    _fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
    L.cc_mlp_create.argtypes = [C.c_uint, C.POINTER(C.c_void_p)]
    L.cc_mlp_set_params.argtypes = [C.c_void_p, _fp]
    L.cc_mlp_get_params.argtypes = [C.c_void_p, _fp]
    L.cc_mlp_forward.argtypes = [C.c_void_p, _fp, C.c_int, _fp]
    L.cc_mlp_grad.argtypes = [C.c_void_p, _fp, _fp, C.c_int, C.POINTER(C.c_double), _fp]
    L.cc_mlp_train.argtypes = [C.c_void_p, _fp, _fp, C.c_int, C.c_int, C.c_int, C.c_float,
                               C.c_uint, _dp]
    L.cc_mlp_destroy.argtypes = [C.c_void_p]
    L.cc_mlp_destroy.restype = None
    L.cc_mlp_last_error.restype = C.c_char_p
    L.cc_filter_samples.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_int)]
    L.cc_set_closure.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.cc_set_wall.argtypes = [C.c_void_p, C.c_int, C.POINTER(cc_wall), C.c_void_p, C.c_void_p]
    L.cc_wall_preset.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(cc_wall)]
    L.cc_average.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                             C.POINTER(C.c_int)]
    L.cc_postprocess.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_void_p, C.c_void_p]
    L.cc_pp_fluctuation.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]
    L.cc_pp_qcriterion.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp]
    L.cc_last_error.restype = C.c_char_p
    for f in ("cc_create", "cc_upload", "cc_step", "cc_stats", "cc_diag", "cc_diag_noalloc",
              "cc_download",
              "cc_blown", "cc_profile", "cc_tiling", "cc_set_tiling", "cc_abi_version",
              "cc_batch_create",
              "cc_batch_step",
              "cc_batch_tune", "cc_mlp_nparams", "cc_mlp_create", "cc_mlp_set_params",
              "cc_mlp_get_params", "cc_mlp_forward", "cc_mlp_grad", "cc_mlp_train",
              "cc_filter_samples", "cc_set_closure", "cc_set_wall", "cc_wall_preset",
              "cc_average", "cc_postprocess", "cc_pp_fluctuation", "cc_pp_qcriterion"):
        getattr(L, f).restype = C.c_int
    _lib = L
    return L


def _check(L, rc):
    if rc < 0:
        raise CCError(f"cc error {rc}: {L.cc_last_error().decode()}")
    return rc


def wall_preset(name):
    """(wall, cc_wall) of a named study (host-only call, no device needed)."""
    L = load_library()
    w = cc_wall()
    wall = C.c_int()
    _check(L, L.cc_wall_preset(name.encode(), C.byref(wall), C.byref(w)))
    return wall.value, w


def _params(n1, n2, n3, Re=3000.0, Mac=1.5, dt=1e-4, istat=10, PexVal=2.0, PeyVal=2.0,
            PezVal=0.2, precision="fp64", Ri_b=0.0, slip_len=0.0):
    prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    return cc_params(n1, n2, n3, istat, prec, 0, Re, Mac, dt, PexVal, PeyVal, PezVal, Ri_b,
                     slip_len)


class Channel:
    """One channel instance on the current HIP device.

    Arrays use the reference's Fortran order exposed as C-order numpy arrays:
    var has shape (7, n3, n2, n1); stat1/stat2 (7, n3); corr (9, n3)."""

    def __init__(self, n1, n2, n3, Re=3000.0, Mac=1.5, dt=1e-4, istat=10, PexVal=2.0,
                 PeyVal=2.0, PezVal=0.2, precision="fp64", Ri_b=0.0, slip_len=0.0):
        self.L = load_library()
        self.p = _params(n1, n2, n3, Re, Mac, dt, istat, PexVal, PeyVal, PezVal, precision, Ri_b,
                         slip_len)
        self.n = (n1, n2, n3)
        self.owned = True
        h = C.c_void_p()
        _check(self.L, self.L.cc_create(C.byref(self.p), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None) and getattr(self, "owned", True):
            self.L.cc_destroy(self.h)
        self.h = None

    def __del__(self):
        self.close()

    @property
    def shape(self):
        n1, n2, n3 = self.n
        return (n3, n2, n1)

    def upload(self, var, nstep=0, time=0.0):
        var = np.ascontiguousarray(var, dtype=np.float64)
        if var.shape != (7,) + self.shape:
            raise ValueError(f"var has shape {var.shape}, expected {(7,) + self.shape}")
        _check(self.L, self.L.cc_upload(self.h, var, nstep, time))

    def step(self, nsteps=1):
        """Advance nsteps; returns 0 or the step at which the blow-up check fired."""
        return _check(self.L, self.L.cc_step(self.h, nsteps))

    def stats(self):
        n3 = self.n[2]
        tau = np.empty(2)
        s1 = np.empty((7, n3))
        s2 = np.empty((7, n3))
        cr = np.empty((9, n3))
        _check(self.L, self.L.cc_stats(self.h, tau, s1, s2, cr))
        return tau, s1, s2, cr

    def diag(self, x3_file, x2_end):
        out = np.empty(16)
        _check(self.L, self.L.cc_diag(self.h, np.ascontiguousarray(x3_file, dtype=np.float64),
                                      float(x2_end), out))
        return out

    def diag_noalloc(self, x3_file, x2_end):
        """diag() through cc_diag_noalloc: no device allocation, partials in
        the library's static per-GPU scratch (n3 <= 4096)."""
        out = np.empty(16)
        _check(self.L, self.L.cc_diag_noalloc(self.h,
                                              np.ascontiguousarray(x3_file, dtype=np.float64),
                                              float(x2_end), out))
        return out

    def download(self):
        out = np.empty((7,) + self.shape)
        ns = C.c_int()
        t = C.c_double()
        _check(self.L, self.L.cc_download(self.h, out, C.byref(ns), C.byref(t)))
        return out, ns.value, t.value

    def profile(self, mode=2):
        """mode 1: reset and enable per-launch HIP-event timing of the stage
        kernel; -k (k >= 2): the same for every k-th step only; 0: disable;
        2: query.  Returns (total_ms, launches)."""
        out = np.zeros(2)
        _check(self.L, self.L.cc_profile(self.h, mode, out))
        return float(out[0]), int(out[1])

    def tiling(self):
        """Tiling of the fused stage kernel: {"tile": (tx, ty), "kchunk": planes
        per block, "blocks": blocks per launch, "cus": the device's CUs,
        "blocks_per_cu": resident blocks per CU}."""
        out = np.zeros(6, dtype=np.int32)
        _check(self.L, self.L.cc_tiling(self.h, out.ctypes.data_as(C.c_void_p)))
        return {"tile": (int(out[0]), int(out[1])), "kchunk": int(out[2]), "blocks": int(out[3]),
                "cus": int(out[4]), "blocks_per_cu": int(out[5])}

    def set_tiling(self, kchunk, variant=-1):
        """Force kchunk planes per block (and, mixed only, a tile shape
        variant 0..3); kchunk <= 0 restores the automatic choice."""
        _check(self.L, self.L.cc_set_tiling(self.h, int(variant), int(kchunk)))
        return self.tiling()

    def filter_samples(self):
        """4x4x4 box-filtered features (N,16) and subgrid targets (N,16) of
        the current state (float32)."""
        M = C.c_int()
        _check(self.L, self.L.cc_filter_samples(self.h, None, None, C.byref(M)))
        X = np.zeros((M.value, 16), dtype=np.float32)
        T = np.zeros((M.value, 16), dtype=np.float32)
        _check(self.L, self.L.cc_filter_samples(self.h, X.ctypes.data_as(C.c_void_p),
                                                T.ctypes.data_as(C.c_void_p), C.byref(M)))
        return X, T

    def set_closure(self, mlp, norm):
        """Attach a trained closure (an Mlp) with its normalisation[40]; None detaches."""
        if mlp is None:
            _check(self.L, self.L.cc_set_closure(self.h, None, None))
            self._closure = None
            return
        nv = np.ascontiguousarray(norm, dtype=np.float32)
        _check(self.L, self.L.cc_set_closure(self.h, mlp.h, nv.ctypes.data_as(C.c_void_p)))
        self._closure = (mlp, nv)  # keep both alive while attached

    def set_wall(self, wall, spec=None, x1=None, x2=None):
        """Wall condition of wall 0 (bottom) or 1 (top): spec is a cc_wall, a
        dict of its fields, a study name (applied to the study's wall; then
        `wall` may be None) or None for the reference wall.  x1 (n1), x2 (n2):
        the solution file's coordinates (None: the analytic mesh)."""
        if isinstance(spec, str):
            w_study, spec = wall_preset(spec)
            wall = w_study if wall is None else wall
        elif isinstance(spec, dict):
            spec = cc_wall(**spec)
        elif spec is None:
            spec = cc_wall()
        keep = []

        def ptr(a, n):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64)
            if a.shape != (n,):
                raise ValueError(f"coordinate array has shape {a.shape}, expected ({n},)")
            keep.append(a)
            return a.ctypes.data

        _check(self.L, self.L.cc_set_wall(self.h, int(wall), C.byref(spec), ptr(x1, self.n[0]),
                                          ptr(x2, self.n[1])))

    def average(self, mode=0):
        """Time averages of the plane statistics (mode 1: restart the
        accumulation, 2: stop it, 0: query).  Returns (mean (7, n3), rms
        (7, n3), covar (9, n3), count)."""
        n3 = self.n[2]
        mean, rms, cov = np.zeros((7, n3)), np.zeros((7, n3)), np.zeros((9, n3))
        cnt = C.c_int()
        _check(self.L, self.L.cc_average(self.h, int(mode), mean.ctypes.data, rms.ctypes.data,
                                         cov.ctypes.data, C.byref(cnt)))
        return mean, rms, cov, cnt.value

    def postprocess(self, mean=None, coords=None, fluc=True, q=True):
        """Fluctuation field (7, n3, n2, n1) and Q-criterion (n3, n2, n1) of
        the current state on the device; mean (7, n3) or None for the
        accumulated time average; coords = (x1, x2, x3) for Q."""
        n1, n2, n3 = self.n
        keep = []

        def ptr(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64)
            keep.append(a)
            return a.ctypes.data

        F = np.empty((7, n3, n2, n1)) if fluc else None
        Q = np.empty((n3, n2, n1)) if q else None
        x1, x2, x3 = coords if coords is not None else (None, None, None)
        _check(self.L, self.L.cc_postprocess(self.h, ptr(mean), ptr(x1), ptr(x2), ptr(x3),
                                             F.ctypes.data if fluc else None,
                                             Q.ctypes.data if q else None))
        return F, Q

    def blown(self):
        ijk = (C.c_int * 3)()
        step = self.L.cc_blown(self.h, ijk)
        return step, tuple(ijk)


class Member(Channel):
    """A batch member: the batch owns the handle."""

    def __init__(self, L, h, n, p):
        self.L = L
        self.h = C.c_void_p(h)
        self.n = n
        self.p = p
        self.owned = False


class Batch:
    """Several channel instances (each a dict of Channel keyword arguments)
    advanced together by persistent batched launches."""

    def __init__(self, specs):
        self.L = load_library()
        arr = (cc_params * len(specs))()
        for i, sp in enumerate(specs):
            arr[i] = _params(**sp)
        h = C.c_void_p()
        _check(self.L, self.L.cc_batch_create(arr, len(specs), C.byref(h)))
        self.h = h
        self.members = [Member(self.L, self.L.cc_batch_member(h, i),
                               (sp["n1"], sp["n2"], sp["n3"]), arr[i])
                        for i, sp in enumerate(specs)]

    def step(self, nsteps=1):
        return _check(self.L, self.L.cc_batch_step(self.h, nsteps))

    def tune(self, trial_steps=3):
        out = np.zeros(4)
        _check(self.L, self.L.cc_batch_tune(self.h, trial_steps, out))
        return {"ms_per_step": out[0], "tile": (int(out[1]), int(out[2])), "kchunk": int(out[3])}

    def close(self):
        if getattr(self, "h", None):
            for m in self.members:
                m.h = None
            self.L.cc_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def _mcheck(L, rc):
    if rc < 0:
        raise CCError(f"cc_mlp error {rc}: {L.cc_mlp_last_error().decode()}")
    return rc


class Mlp:
    """The subgrid-closure network on the device (BF16 MFMA, FP32 master)."""

    def __init__(self, seed=0):
        self.L = load_library()
        h = C.c_void_p()
        _mcheck(self.L, self.L.cc_mlp_create(seed, C.byref(h)))
        self.h = h
        self.nparams = self.L.cc_mlp_nparams()

    @property
    def params(self):
        out = np.empty(self.nparams, dtype=np.float32)
        _mcheck(self.L, self.L.cc_mlp_get_params(self.h, out))
        return out

    @params.setter
    def params(self, p):
        _mcheck(self.L, self.L.cc_mlp_set_params(self.h, np.ascontiguousarray(p, dtype=np.float32)))

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.empty_like(X)
        _mcheck(self.L, self.L.cc_mlp_forward(self.h, X, X.shape[0], Y))
        return Y

    def grad(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.empty(self.nparams, dtype=np.float32)
        loss = C.c_double()
        _mcheck(self.L, self.L.cc_mlp_grad(self.h, X, T, X.shape[0], C.byref(loss), g))
        return loss.value, g

    def train(self, X, T, epochs=10, batch=4096, lr=1e-3, seed=0):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        hist = np.zeros(epochs)
        _mcheck(self.L, self.L.cc_mlp_train(self.h, X, T, X.shape[0], epochs, batch, lr, seed,
                                            hist))
        return hist

    def close(self):
        if getattr(self, "h", None):
            self.L.cc_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
