"""NumPy restatement of the reference's post-processing programs.

TEST INFRASTRUCTURE ONLY: the checker for compchan3dmpi_amd.postproc and the
post-processing kernels; never imported by the product package.  Each
function follows its Fortran source operation by operation (NumPy's
elementwise IEEE operations, no fused multiply-adds, sums accumulated in the
Fortran loop order), so that its results are bit-identical to the programs
built by oracle/build_post.sh.

  average     v1.9/codes/avg_v1.7.f90:18-283 (single time interval)
  fluc        v1.9/codes/fluc.f90:55-62
  q_invariant v1.8_validation/codes/qInvariant.f90:26-81
"""
from __future__ import annotations

import numpy as np


def read_moment_rows(path: str, ncol: int) -> tuple[np.ndarray, np.ndarray]:
    """Rows of a moment file: time (1X,F15.7) then n3 groups of ncol values
    (avg_v1.7.f90:115-140).  Returns (time[rows], data[rows, n3, ncol])."""
    t, d = [], []
    for line in open(path):
        if not line.strip():
            continue
        vals = [float(line[16 * q:16 * q + 16]) for q in range(len(line.rstrip("\n")) // 16)]
        t.append(vals[0])
        d.append(np.array(vals[1:]).reshape(-1, ncol))
    return np.array(t), np.array(d)


def average(s1: np.ndarray, s2: np.ndarray, cr: np.ndarray, first: int = 1):
    """Time averages of the moment histories s1, s2 [rows, n3, 7], cr [rows,
    n3, 9] over rows first..rows-1 (1-based; avg_v1.7 ends the interval at
    line cnt-1, :59-60).  Returns mean, rms [n3, 7] and covar [n3, 9]."""
    rows = s1.shape[0]
    alc = rows - 1 - first + 1
    idx = range(first - 1, rows - 1)
    a1 = np.zeros(s1.shape[1:])
    a2 = np.zeros(s2.shape[1:])
    ac = np.zeros(cr.shape[1:])
    for i in idx:  # sum(stat1Spc, 3): in line order
        a1 = a1 + s1[i]
        a2 = a2 + s2[i]
        ac = ac + cr[i]
    stat1 = a1 / float(alc)
    stat2 = a2 / float(alc)
    rms = np.sqrt(stat2 - stat1 * stat1)
    corr = ac / float(alc)
    pairs = ((0, 3), (0, 5), (1, 2), (1, 3), (2, 3), (2, 5), (5, 3), (6, 3), (6, 5))
    covar = np.stack([corr[:, q] - stat1[:, a] * stat1[:, b] for q, (a, b) in enumerate(pairs)],
                     axis=1)
    return stat1, rms, covar


def fluc(var: np.ndarray, mean: np.ndarray) -> np.ndarray:
    """var (7, n3, n2, n1) minus the plane means mean [n3, 7] (fluc.f90:55-62)."""
    return var - mean.T[:, :, None, None]


def q_invariant(x: np.ndarray, y: np.ndarray, z: np.ndarray, U: np.ndarray) -> np.ndarray:
    """Q = (P^2 - S:S + R:R)/2 of the velocity fluctuations U (3, n3, n2, n1)
    with the three-point derivatives of a non-uniform mesh
    (qInvariant.f90:26-81); 0 on the boundary nodes (the reference's static
    Q array is never written there)."""
    _, n3, n2, n1 = U.shape
    Q = np.zeros((n3, n2, n1))
    haf = 0.5
    I = slice(1, n1 - 1)
    J = slice(1, n2 - 1)
    K = slice(1, n3 - 1)
    dx1 = (x[2:] - x[1:-1])[None, None, :]
    dx2 = (x[1:-1] - x[:-2])[None, None, :]
    dy1 = (y[2:] - y[1:-1])[None, :, None]
    dy2 = (y[1:-1] - y[:-2])[None, :, None]
    dz1 = (z[2:] - z[1:-1])[:, None, None]
    dz2 = (z[1:-1] - z[:-2])[:, None, None]
    Jt = [[None] * 3 for _ in range(3)]
    for c in range(3):
        u = U[c]
        c0 = u[K, J, I]
        Jt[c][0] = ((dx2 / dx1) * (u[K, J, 2:] - c0) + (dx1 / dx2) * (c0 - u[K, J, :-2])) / (dx1 + dx2)
        Jt[c][1] = ((dy2 / dy1) * (u[K, 2:, I] - c0) + (dy1 / dy2) * (c0 - u[K, :-2, I])) / (dy1 + dy2)
        Jt[c][2] = ((dz2 / dz1) * (u[2:, J, I] - c0) + (dz1 / dz2) * (c0 - u[:-2, J, I])) / (dz1 + dz2)
    S = [[None] * 3 for _ in range(3)]
    R = [[None] * 3 for _ in range(3)]
    S[0][0] = Jt[0][0]
    S[0][1] = haf * (Jt[0][1] + Jt[1][0])
    S[0][2] = haf * (Jt[0][2] + Jt[2][0])
    S[1][0] = haf * (Jt[1][0] + Jt[0][1])
    S[1][1] = Jt[1][1]
    S[1][2] = haf * (Jt[2][1] + Jt[1][2])
    S[2][0] = haf * (Jt[2][0] + Jt[0][2])
    S[2][1] = haf * (Jt[2][1] + Jt[1][2])
    S[2][2] = Jt[2][2]
    zero = np.zeros_like(Jt[0][0])
    R[0][0] = zero
    R[0][1] = haf * (Jt[0][1] - Jt[1][0])
    R[0][2] = haf * (Jt[0][2] - Jt[2][0])
    R[1][0] = haf * (Jt[1][0] - Jt[0][1])
    R[1][1] = zero
    R[1][2] = haf * (Jt[1][2] - Jt[2][1])
    R[2][0] = haf * (Jt[2][0] - Jt[0][2])
    R[2][1] = haf * (Jt[2][1] - Jt[1][2])
    R[2][2] = zero
    # sum over the 3x3 array in Fortran (column-major) order
    a1 = np.zeros_like(zero)
    a2 = np.zeros_like(zero)
    for col in range(3):
        for row in range(3):
            a1 = a1 + S[row][col] * S[row][col]
            a2 = a2 + (-(R[row][col] * R[row][col]))
    P = -(S[0][0] + S[1][1] + S[2][2])
    Q[K, J, I] = haf * (P * P - a1 - a2)
    return Q


def mean_text(x3: np.ndarray, a: np.ndarray) -> str:
    """mean.dat / rms.dat lines: 8(1X,F22.18) (avg_v1.7.f90:240-245)."""
    return "".join("".join(f" {v:22.18f}" for v in (x3[k], *a[k])) + "\n" for k in range(len(x3)))


def corr_text(x3: np.ndarray, c: np.ndarray) -> str:
    """corr.dat lines: 10(1X,F10.5) (avg_v1.7.f90:277-280)."""
    return "".join("".join(f" {v:10.5f}" for v in (x3[k], *c[k])) + "\n" for k in range(len(x3)))
