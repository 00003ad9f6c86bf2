"""Run the reference solver built by oracle/build_ref.sh on a given initial
state (test infrastructure: used only to write golden fixtures under
tests/golden/ in a container where /root/reference is present).

The reference reads ../param.dat and the solution file named there, and
writes ../output/3D001[_].dat every NWRITE steps (ES25.17 in this build) and
the appended diagnostic files (main_v1.9.f90:769-1037).
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from compchan3dmpi_amd import formats  # noqa: E402


def build(n1, n2, n3, maxstep, nwrite, study=""):
    args = [os.path.join(HERE, "build_ref.sh"), str(n1), str(n2), str(n3), str(maxstep),
            str(nwrite)] + ([study] if study else [])
    out = subprocess.run(args, check=True, capture_output=True, text=True)
    return out.stdout.strip().splitlines()[-1]


def run(var, params: formats.Params, maxstep: int, nwrite: int, workdir: str | None = None,
        timeout: int = 3600, study: str = ""):
    """Run the reference (or one of its wall-condition studies, see
    build_ref.sh) for maxstep steps from var (7, n3, n2, n1).

    Returns a dict with the final 3D file's var and the appended outputs."""
    _, n3, n2, n1 = var.shape
    exe = build(n1, n2, n3, maxstep, nwrite, study)
    root = workdir or tempfile.mkdtemp(prefix="ccref_")
    formats.write_run_dir(root, params, var)
    env = dict(os.environ)
    env.setdefault("OMP_NUM_THREADS", "1")
    proc = subprocess.run([exe], cwd=os.path.join(root, "codes"), capture_output=True, text=True,
                          timeout=timeout, env=env)
    res = {"stdout": proc.stdout, "returncode": proc.returncode, "root": root}
    outdir = os.path.join(root, "output")
    # the file alternates on nstep/100 (main_v1.9.f90:771-775)
    fn = "3D001_.dat" if (maxstep // 100) % 2 == 1 else "3D001.dat"
    path = os.path.join(outdir, fn)
    if os.path.exists(path):
        t, ns, x1, x2, x3, v = formats.read_3d(path, n1, n2, n3)
        res.update(time=t, nstep=ns, var=v)
    for name in ("tauWall", "moment1st", "moment2nd", "momentJoint", "KE", "rhobulk", "Ubulk",
                 "massflow"):
        p = os.path.join(outdir, name + ".dat")
        if os.path.exists(p):
            rows = [np.array(line.split(), dtype=np.float64) for line in open(p) if line.strip()]
            res[name] = np.array(rows)
    return res
