"""Post-processing of a run directory, as the reference's stand-alone
programs do it, with the field work on the device and the files through the
native parallel reader/writer.

  average  v1.9/codes/avg_v1.7.f90      ../output/moment1st.dat,
           moment2nd.dat, momentJoint.dat -> ../output/mean.dat, rms.dat,
           corr.dat and plotInfo.dat (in the working directory)
  fluc     v1.9/codes/fluc.f90          ../input/3D.dat + ../output/mean.dat
           -> ../output/fluc.dat          (device: cc_pp_fluctuation)
  qinv     v1.8_validation/codes/qInvariant.f90   ../output/fluc.dat
           -> ../output/qInvariant.dat    (device: cc_pp_qcriterion)

Paths are relative to the run's codes/ directory like the reference's.  The
outputs are byte-identical to those programs' on the same input files.

On a live instance, Channel.average() accumulates the time averages from the
device's full-precision statistics and Channel.postprocess() computes the
fluctuation field and Q without any text round trip.

    python -m compchan3dmpi_amd.postproc avg|fluc|qinv [codes_dir]
"""
from __future__ import annotations

import os
import sys

import numpy as np

from . import fastio, formats
from .channel import _check, load_library

PAIRS = ((0, 3), (0, 5), (1, 2), (1, 3), (2, 3), (2, 5), (5, 3), (6, 3), (6, 5))


def _moment_rows(path, ncol):
    """(time[rows], data[rows, n3, ncol]) of a moment file: (1X,F15.7) time,
    then n3 groups of ncol (1X,F15.7) values (avg_v1.7.f90:115-140)."""
    t, d = [], []
    for line in open(path):
        s = line.rstrip("\n")
        if not s.strip():
            continue
        vals = np.array([float(s[q:q + 16]) for q in range(0, len(s) - len(s) % 16, 16)])
        t.append(vals[0])
        d.append(vals[1:].reshape(-1, ncol))
    return np.array(t), np.array(d)


def average(codes_dir=".", first=1):
    """avg_v1.7 on the run's moment files, single time interval from line
    `first` (its one-argument form).  Like avg_v1.7 the last line is left out
    (the interval ends at line cnt-1, avg_v1.7.f90:59-60).  Writes mean.dat,
    rms.dat, corr.dat and plotInfo.dat; returns (mean, rms, covar)."""
    out = os.path.join(codes_dir, "..", "output")
    t, s1 = _moment_rows(os.path.join(out, "moment1st.dat"), 7)
    _, s2 = _moment_rows(os.path.join(out, "moment2nd.dat"), 7)
    _, cr = _moment_rows(os.path.join(out, "momentJoint.dat"), 9)
    if np.any(np.diff(t[first - 1:]) < 0):
        raise ValueError("the moment files hold several time intervals: average them one at a "
                         "time by choosing `first`")
    rows, n3 = s1.shape[0], s1.shape[1]
    alc = rows - first
    if alc < 1:
        raise ValueError("fewer than two lines to average")
    a1, a2, ac = np.zeros(s1.shape[1:]), np.zeros(s2.shape[1:]), np.zeros(cr.shape[1:])
    for i in range(first - 1, rows - 1):  # sum(...,3) in line order
        a1 = a1 + s1[i]
        a2 = a2 + s2[i]
        ac = ac + cr[i]
    stat1, stat2, corr = a1 / float(alc), a2 / float(alc), ac / float(alc)
    rms = np.sqrt(stat2 - stat1 * stat1)
    covar = np.stack([corr[:, q] - stat1[:, a] * stat1[:, b] for q, (a, b) in enumerate(PAIRS)],
                     axis=1)
    # x3 of the analytic mesh (setupMesh, comdata.f90:70-110); n1, n2 do not enter it
    x3 = formats.mesh(4, 4, n3)[2]
    with open(os.path.join(out, "mean.dat"), "w") as f:
        f.write("".join("".join(f" {v:22.18f}" for v in (x3[k], *stat1[k])) + "\n" for k in range(n3)))
    with open(os.path.join(out, "rms.dat"), "w") as f:
        f.write("".join("".join(f" {v:22.18f}" for v in (x3[k], *rms[k])) + "\n" for k in range(n3)))
    with open(os.path.join(out, "corr.dat"), "w") as f:
        f.write("".join("".join(f" {v:10.5f}" for v in (x3[k], *covar[k])) + "\n" for k in range(n3)))
    with open(os.path.join(codes_dir, "plotInfo.dat"), "w") as f:
        f.write(f"Averaged b/w nondimensional time{t[first - 1]:10.5f} to {t[rows - 2]:10.5f}\n")
        f.write("-" * 59 + "\n")
    return stat1, rms, covar


def read_mean(path):
    """stat1 [n3, 7] from mean.dat (23X,7(1X,F22.18); fluc.f90:39-52)."""
    rows = [line for line in open(path) if line.strip()]
    return np.array([[float(r[23 + 23 * c:23 + 23 * c + 23]) for c in range(7)] for r in rows])


def fluc(codes_dir="."):
    """fluc.f90: the solution ../input/3D.dat minus the plane means of
    ../output/mean.dat, on the device; writes ../output/fluc.dat."""
    L = load_library()
    t, ns, x1, x2, x3, var = fastio.read_solution(os.path.join(codes_dir, "..", "input", "3D.dat"))
    mean = read_mean(os.path.join(codes_dir, "..", "output", "mean.dat"))
    n3, n2, n1 = var.shape[1:]
    out = np.empty_like(var)
    _check(L, L.cc_pp_fluctuation(n1, n2, n3, var, np.ascontiguousarray(mean.T), out))
    header = (f'TITLE ="{t:20.10f}{ns:9d}"\n'
              'Variables ="x","y","z","rho_prime","U_prime","V_prime","W_prime","E_prime",'
              '"T_prime","P_prime"\n'
              f'ZONE k={n3:3d},j={n2:3d},i={n1:3d},DATAPACKING="POINT"\n')
    fastio.write_points(os.path.join(codes_dir, "..", "output", "fluc.dat"), header, x1, x2, x3,
                        out)
    return out


def qinv(codes_dir="."):
    """qInvariant.f90: Q of the velocity fluctuations in ../output/fluc.dat,
    on the device; writes ../output/qInvariant.dat."""
    L = load_library()
    path = os.path.join(codes_dir, "..", "output", "fluc.dat")
    n1, n2, n3 = fastio.dims(path)
    x, y, z, F = fastio.read_points(path, n1, n2, n3, 7)
    Q = np.empty((n3, n2, n1))
    _check(L, L.cc_pp_qcriterion(n1, n2, n3, x, y, z, np.ascontiguousarray(F[1:4]), Q))
    header = (' TITLE ="qInvariant"\n Variables = "x","y","z","Q"\n'
              f'ZONE i={n1:3d},j={n2:3d},k={n3:3d},DATAPACKING="POINT"\n')
    fastio.write_points(os.path.join(codes_dir, "..", "output", "qInvariant.dat"), header, x, y,
                        z, Q[None], blank_records=False)
    return Q


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if not argv or argv[0] not in ("avg", "fluc", "qinv"):
        print(__doc__)
        return 2
    d = argv[1] if len(argv) > 1 else "."
    {"avg": average, "fluc": fluc, "qinv": qinv}[argv[0]](d)
    print(f"{argv[0]}: done")
    return 0


if __name__ == "__main__":
    sys.exit(main())
