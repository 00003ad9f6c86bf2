"""File contracts of the chanCompr solver: ``param.dat`` and the Tecplot-ASCII
solution files, plus the seeded laminar-plus-noise initial condition.

Reference formats:

* ``param.dat``: 24 ``description:value`` lines with the value from column 49,
  read by ``v1.9/codes/comdata.f90:31-68`` (``48X`` edit descriptors) and by
  ``v1.9/bin/readParam.sh:4-36``.
* solution file ``3D.dat`` / ``3Dnnn.dat``: header ``(A,F20.10,I9,A)``, a
  ``Variables`` line, a ``ZONE`` line ``(A,I3,A,I3,A,I3,A)``, then for each k a
  blank record, for each j a blank record and n1 records
  ``(3(1X,F10.6),7(1X,E22.15))`` of x y z rho u v w E T p
  (``v1.9/codes/main_v1.9.f90:210-251, 749-794``).

Arrays follow the reference's Fortran order: ``var`` has shape
``(7, n3, n2, n1)`` in C order, i.e. ``var(i,j,k,c)`` column-major.
"""
from __future__ import annotations

import dataclasses
import math
import os
import re

import numpy as np

GAMA = 1.4
PR_REF = 0.71  # comdata.f90:12 (invPr = 1/0.71)


# --------------------------------------------------------------------------
# param.dat
# --------------------------------------------------------------------------
@dataclasses.dataclass
class Params:
    """Run-time parameters of ``param.dat`` (comdata.f90:31-68)."""

    descp: str = "channel"
    resOp: int = 1
    topo: tuple = (1, 1, 1)
    cpu_list: str = "0"
    Re: float = 3000.0
    Mac: float = 1.5
    dt: float = 0.0001
    istat: int = 10
    PexVal: float = 2.0
    PeyVal: float = 2.0
    PezVal: float = 0.2
    debugOp: int = 0
    inputPath: str = "../input/3D.dat"
    flags: str = "-O3"
    mainVer: str = "v1.9"
    avgVer: str = "v1.7"

    @property
    def invRe(self) -> float:
        return 1.0 / self.Re

    @property
    def ggM(self) -> float:
        return GAMA * (GAMA - 1) * self.Mac ** 2


def _fmt_dt(dt: float) -> str:
    """dt in a 6-character field readable by ``f6.4`` (comdata.f90:47).

    Four decimals when they represent dt exactly; otherwise the leading zero
    is dropped, which gives five decimals in six characters (an explicit
    decimal point overrides the edit descriptor's d on input)."""
    s = f"{dt:6.4f}"
    if abs(float(s) - dt) <= 1e-15:
        return s
    s = f"{dt:.5f}"
    if s.startswith("0."):
        s = s[1:]
    if len(s) != 6 or abs(float(s) - dt) > 1e-15:
        raise ValueError(f"dt={dt!r} does not fit the f6.4 field of param.dat")
    return s


def _line(desc: str, value: str) -> str:
    d = (desc + " ").ljust(47, ".") + ":"
    assert len(d) == 48
    return d + value


def write_param(path: str, p: Params) -> None:
    """Write ``param.dat`` with values from column 49."""
    lines = [
        "=" * 60,
        "Case",
        _line("Description", p.descp),
        _line("Restart option [1=combined,2=per-rank]", f"{p.resOp:1d}"),
        "Process topology",
        _line("Procs in x", f"{p.topo[0]:2d}"),
        _line("Procs in y", f"{p.topo[1]:2d}"),
        _line("Procs in z", f"{p.topo[2]:2d}"),
        _line("CPU list", p.cpu_list),
        "Flow",
        _line("Reynolds number", f"{p.Re:4.0f}"),
        _line("Mach number", f"{p.Mac:3.1f}"),
        _line("Time step", _fmt_dt(p.dt)),
        _line("Statistics frequency", f"{p.istat:2d}"),
        "Hybrid scheme",
        _line("Peclet threshold x", f"{p.PexVal:3.1f}"),
        _line("Peclet threshold y", f"{p.PeyVal:3.1f}"),
        _line("Peclet threshold z", f"{p.PezVal:3.1f}"),
        "Misc",
        _line("Debug option", f"{p.debugOp:1d}"),
        _line("Input path", p.inputPath),
        _line("Compiler flags", p.flags),
        _line("Main version", p.mainVer),
        _line("Average version", p.avgVer),
    ]
    assert len(lines) == 24
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def _fortran_f(field: str, w: int, d: int) -> float:
    """Fortran F input editing of a w-wide field (blanks ignored, an explicit
    point overrides d)."""
    s = field[:w].strip()
    if not s:
        return 0.0
    if "." in s or "e" in s.lower():
        return float(s)
    return int(s) / 10 ** d


def read_param(path: str) -> Params:
    """Read ``param.dat`` the way comdata.f90:31-68 does (value at column 49)."""
    with open(path) as f:
        L = f.read().split("\n")

    def val(n):
        return L[n - 1][48:] if len(L[n - 1]) > 48 else ""

    p = Params()
    p.descp = val(3)[:10].strip()
    p.resOp = int(val(4)[:1])
    p.topo = tuple(int(val(n)[:2]) for n in (6, 7, 8))
    p.cpu_list = val(9).strip()
    p.Re = _fortran_f(val(11), 4, 0)
    p.Mac = _fortran_f(val(12), 3, 1)
    p.dt = _fortran_f(val(13), 6, 4)
    p.istat = int(val(14)[:2])
    p.PexVal = _fortran_f(val(16), 3, 1)
    p.PeyVal = _fortran_f(val(17), 3, 1)
    p.PezVal = _fortran_f(val(18), 3, 1)
    p.debugOp = int(val(20)[:1])
    p.inputPath = val(21).strip()
    if len(L) > 22:
        p.flags = val(22).strip()
    return p


# --------------------------------------------------------------------------
# mesh (comdata.f90:70-110)
# --------------------------------------------------------------------------
def mesh(n1: int, n2: int, n3: int):
    """Analytic mesh: x1, x2 uniform with n-1 intervals, x3 = 1 - cos(pi xi/2)
    with xi accumulated in steps of 2/(n3-1) (comdata.f90:74-109)."""
    Pi = 4.0 * math.atan(1.0)
    dx = 4.0 * Pi / float(n1 - 1)
    dy = 4.0 * Pi / 3.0 / float(n2 - 1)
    dXi = 2.0 / float(n3 - 1)
    x1 = np.zeros(n1)
    x2 = np.zeros(n2)
    Xi = np.zeros(n3)
    for i in range(n1 - 1):
        x1[i + 1] = x1[i] + dx
    for j in range(n2 - 1):
        x2[j + 1] = x2[j] + dy
    for k in range(n3 - 1):
        Xi[k + 1] = Xi[k] + dXi
    x3 = np.array([1.0 - math.cos(0.50 * Xi[k] * Pi) for k in range(n3)])
    return x1, x2, x3


# --------------------------------------------------------------------------
# Fortran-compatible number formatting
# --------------------------------------------------------------------------
def fortran_e(x: float, w: int = 22, d: int = 15) -> str:
    """Fortran ``Ew.d`` output: 0.ddd...E+xx with d significant digits."""
    if x == 0.0 or not math.isfinite(x):
        mant = "0." + "0" * d
        s = mant + "E+00"
        if math.copysign(1.0, x) < 0 and x == 0.0:
            s = "-" + s
        return s.rjust(w)
    s = f"{x:.{d - 1}e}"  # d.ddddde+xx with d significant digits
    mant, ex = s.split("e")
    neg = mant.startswith("-")
    digits = mant.lstrip("-").replace(".", "")
    e = int(ex) + 1
    es = f"{'+' if e >= 0 else '-'}{abs(e):02d}"
    if abs(e) > 99:
        es = f"{'+' if e >= 0 else '-'}{abs(e):03d}"
        out = ("-" if neg else "") + "0." + digits + es
    else:
        out = ("-" if neg else "") + "0." + digits + "E" + es
    return out.rjust(w)


def _fmt_e_array(a: np.ndarray, d: int = 15) -> np.ndarray:
    return np.array([fortran_e(float(x), 22, d) for x in a.ravel()]).reshape(a.shape)


# --------------------------------------------------------------------------
# solution files
# --------------------------------------------------------------------------
def write_3d(path: str, time: float, nstep: int, x1, x2, x3, var: np.ndarray,
             pattern: bool = False) -> None:
    """Write a combined solution file in the reference's POINT layout
    (main_v1.9.f90:757-773).  var has shape (7, n3, n2, n1)."""
    _, n3, n2, n1 = var.shape
    names = '"x","y","z","Rho","U","V","W","E","T","P"' + (',"Pattern"' if pattern else "")
    out = [f'TITLE ="{time:20.10f}{nstep:9d}"', f" Variables ={names}",
           f'ZONE k={n3:3d},j={n2:3d},i={n1:3d},DATAPACKING="POINT"']
    xs = [f" {v:10.6f}" for v in x1]
    ys = [f" {v:10.6f}" for v in x2]
    zs = [f" {v:10.6f}" for v in x3]
    tail = " 0" if pattern else ""
    for k in range(n3):
        out.append("")
        for j in range(n2):
            out.append("")
            cols = [_fmt_e_array(var[c, k, j, :]) for c in range(7)]
            for i in range(n1):
                out.append(xs[i] + ys[j] + zs[k] + "".join(" " + cols[c][i] for c in range(7)) + tail)
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def decompose(n, topo):
    """((bs1, bs2, bs3), (es1, es2, es3)) of every rank of an MPI topology over
    an n1 x n2 x n3 grid, 1-based and inclusive, in rank order: the reference's
    split (main_v1.9.f90:110-145; along each axis the first topo - mod(n, topo)
    ranks take n // topo nodes, the others one more) and MPI's row-major
    numbering of the Cartesian communicator it creates without reordering
    (:90)."""
    axes = []
    for d in range(3):
        sz, cut, first, b = n[d] // topo[d], topo[d] - n[d] % topo[d], 1, []
        for r in range(topo[d]):
            if r == cut:
                sz += 1
            b.append((first, first + sz - 1))
            first += sz
        axes.append(b)
    return [((axes[0][a][0], axes[1][b][0], axes[2][c][0]),
             (axes[0][a][1], axes[1][b][1], axes[2][c][1]))
            for a in range(topo[0]) for b in range(topo[1]) for c in range(topo[2])]


def dominfo_text(n, blocks) -> str:
    """domInfo.txt as the reference writes it (main_v1.9.f90:801-811): nP
    (3(i4)), the rank count (i3), each rank's bounds (6(i4,1x)); the 1X that
    ends a record writes nothing."""
    lines = [f"{n[0]:4d}{n[1]:4d}{n[2]:4d}", f"{len(blocks):3d}"]
    for bs, es in blocks:
        lines.append(" ".join(f"{v:4d}" for v in (bs[0], es[0], bs[1], es[1], bs[2], es[2])))
    return "\n".join(lines) + "\n"


def write_rank_files(outdir, time, nstep, x1, x2, x3, var, topo, suffix=""):
    """The per-rank solution files 3Dnnn<suffix>.dat and domInfo.txt that the
    reference's ranks write for this topology (main_v1.9.f90:769-811)."""
    _, n3, n2, n1 = var.shape
    blocks = decompose((n1, n2, n3), topo)
    for r, (bs, es) in enumerate(blocks):
        sl = (slice(None), slice(bs[2] - 1, es[2]), slice(bs[1] - 1, es[1]), slice(bs[0] - 1, es[0]))
        write_3d(os.path.join(outdir, f"3D{r + 1:03d}{suffix}.dat"), time, nstep,
                 x1[bs[0] - 1:es[0]], x2[bs[1] - 1:es[1]], x3[bs[2] - 1:es[2]], var[sl])
    with open(os.path.join(outdir, "domInfo.txt"), "w") as f:
        f.write(dominfo_text((n1, n2, n3), blocks))
    return blocks


def read_rank_files(outdir, suffix=""):
    """(time, nstep, x1, x2, x3, var) assembled from the per-rank files listed
    by outdir/domInfo.txt (the reference's combine program,
    v1.9/codes/combine.f90:15-58)."""
    with open(os.path.join(outdir, "domInfo.txt")) as f:
        rows = [l for l in f.read().split("\n") if l.strip()]
    n1, n2, n3 = int(rows[0][0:4]), int(rows[0][4:8]), int(rows[0][8:12])
    nproc = int(rows[1])
    var = np.zeros((7, n3, n2, n1))
    x1, x2, x3 = np.zeros(n1), np.zeros(n2), np.zeros(n3)
    t = ns = None
    for r in range(nproc):
        b = [int(v) for v in rows[2 + r].split()]
        t, ns, y1, y2, y3, v = read_3d(os.path.join(outdir, f"3D{r + 1:03d}{suffix}.dat"))
        var[:, b[4] - 1:b[5], b[2] - 1:b[3], b[0] - 1:b[1]] = v
        x1[b[0] - 1:b[1]], x2[b[2] - 1:b[3]], x3[b[4] - 1:b[5]] = y1, y2, y3
    return t, ns, x1, x2, x3, var


_HDR = re.compile(r'TITLE\s*="\s*(\S+)\s+(\d+)\s*"')


def file_coords(n1: int, n2: int, n3: int):
    """x1, x2, x3 as the reference reads them back from a solution file: the
    analytic mesh written with F10.6 (main_v1.9.f90:763-772) and read into
    the mesh arrays (:226, 254)."""
    return tuple(np.array([float(f"{v:10.6f}") for v in a]) for a in mesh(n1, n2, n3)[:3])


def read_3d(path: str, n1: int | None = None, n2: int | None = None, n3: int | None = None):
    """Read a solution file written by the reference or by write_3d.

    Returns (time, nstep, x1, x2, x3, var) with var of shape (7, n3, n2, n1).
    The header is read with (8X,F20.10,I9,1X) as main_v1.9.f90:227-231."""
    with open(path) as f:
        first = f.readline()
        f.readline()
        zone = f.readline()
        body = f.read()
    time = float(first[8:28])
    nstep = int(first[28:37])
    m = re.search(r"k=\s*(\d+)\s*,\s*j=\s*(\d+)\s*,\s*i=\s*(\d+)", zone)
    if m:
        n3, n2, n1 = int(m.group(1)), int(m.group(2)), int(m.group(3))
    vals = np.array(body.split(), dtype=np.float64)
    ncol = vals.size // (n1 * n2 * n3)
    if ncol not in (10, 11) or vals.size != ncol * n1 * n2 * n3:
        raise ValueError(f"{path}: {vals.size} numbers do not fit a {n1}x{n2}x{n3} grid")
    vals = vals.reshape(n3, n2, n1, ncol)
    x1 = vals[0, 0, :, 0].copy()
    x2 = vals[0, :, 0, 1].copy()
    x3 = vals[:, 0, 0, 2].copy()
    var = np.ascontiguousarray(np.moveaxis(vals[..., 3:10], -1, 0))
    return time, nstep, x1, x2, x3, var


# --------------------------------------------------------------------------
# seeded initial condition
# --------------------------------------------------------------------------
_M64 = (1 << 64) - 1


def splitmix64_uniform(seed: int, n: int) -> np.ndarray:
    """n draws of splitmix64 (Steele, Lea, Flood 2014) mapped to [0,1) as
    (z >> 11) * 2^-53."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed & _M64) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def laminar_ic(n1: int, n2: int, n3: int, Mac: float = 1.5, seed: int = 20261203,
               amp: float = 0.05) -> np.ndarray:
    """Laminar Poiseuille profile plus seeded uniform velocity noise.

    rho = 1, T = 1, p = 1/(gamma Ma^2), u = 1.5 x3 (2 - x3) (bulk velocity 1),
    v = w = 0, plus amp*(2r-1) on u, v, w at every updated node (i, j in
    2..n-1, k in 2..n3-1), r drawn in the order k, j, i, component (u,v,w).
    Walls carry u = v = w = 0, E = T + ggM/2 |u|^2.  Periodic duplicates are
    filled as the reference's exchange leaves them: i=1 <- n1-1, i=n1 <- 2,
    likewise in j, corners included.  Returns var of shape (7, n3, n2, n1)."""
    _, _, x3 = mesh(n1, n2, n3)
    ggM = GAMA * (GAMA - 1) * Mac ** 2
    var = np.zeros((7, n3, n2, n1))
    var[0] = 1.0
    var[1] = (1.5 * x3 * (2.0 - x3))[:, None, None]
    nu = (n1 - 2) * (n2 - 2) * (n3 - 2)
    r = splitmix64_uniform(seed, 3 * nu).reshape(n3 - 2, n2 - 2, n1 - 2, 3)
    for c in range(3):
        var[1 + c, 1:-1, 1:-1, 1:-1] += amp * (2.0 * r[..., c] - 1.0)
    var[1:4, 0] = 0.0
    var[1:4, -1] = 0.0
    var[5] = 1.0
    var[4] = var[5] + ggM * 0.5 * (var[1] ** 2 + var[2] ** 2 + var[3] ** 2)
    var[6] = var[0] * var[5] / (GAMA * Mac ** 2)
    var[:, :, :, 0] = var[:, :, :, n1 - 2]
    var[:, :, :, n1 - 1] = var[:, :, :, 1]
    var[:, :, 0, :] = var[:, :, n2 - 2, :]
    var[:, :, n2 - 1, :] = var[:, :, 1, :]
    return var


def round_e22_15(var: np.ndarray) -> np.ndarray:
    """Values as a reader sees them after an E22.15 round trip (15 significant
    digits, correctly rounded both ways)."""
    f = np.vectorize(lambda x: float(f"{x:.14e}"), otypes=[np.float64])
    return f(var)


def cfl_dt(var: np.ndarray, Mac: float, cfl: float = 0.5) -> float:
    """dt = cfl / max_nodes sum_d (|u_d| + c)/dx_d, c = sqrt(T)/Ma, with the
    wall-normal spacing the smaller neighbouring interval.  Rounded down to
    four decimals, or to five when four give zero."""
    _, n3, n2, n1 = var.shape
    x1, x2, x3 = mesh(n1, n2, n3)
    dx, dy = x1[1] - x1[0], x2[1] - x2[0]
    dz = np.minimum(np.diff(x3)[1:], np.diff(x3)[:-1])  # k = 2..n3-1
    c = np.sqrt(np.abs(var[5, 1:-1])) / Mac
    s = ((np.abs(var[1, 1:-1]) + c) / dx + (np.abs(var[2, 1:-1]) + c) / dy
         + (np.abs(var[3, 1:-1]) + c) / dz[:, None, None])
    dt = cfl / float(s.max())
    for dec in (4, 5):
        q = math.floor(dt * 10 ** dec) / 10 ** dec
        if q > 0:
            return q
    raise ValueError("CFL time step below 1e-5")


def write_run_dir(root: str, p: Params, var: np.ndarray, time: float = 0.0,
                  nstep: int = 0) -> str:
    """Lay out a run directory the reference expects: root/param.dat,
    root/input/3D.dat, root/codes, root/output, root/debug."""
    for d in ("codes", "output", "debug", "input"):
        os.makedirs(os.path.join(root, d), exist_ok=True)
    _, n3, n2, n1 = var.shape
    x1, x2, x3 = mesh(n1, n2, n3)
    p = dataclasses.replace(p, inputPath="../input/3D.dat")
    write_param(os.path.join(root, "param.dat"), p)
    write_3d(os.path.join(root, "input", "3D.dat"), time, nstep, x1, x2, x3, var)
    return root
