"""Writes the golden fixtures under tests/golden/ by running the reference
solver (v1.10/codes/main_v1.9.f90 + v1.9/codes/comdata.f90, compiled from
/root/reference by oracle/build_ref.sh) on seeded inputs.

Run in a container where /root/reference exists:

    python tests/golden/make_golden.py

Fixtures (all inputs are regenerated from the seed by the tests):
  tiny_12x8x16_s1.npy     var(7,16,8,12) after 1 step,  dt=1e-3, istat=100
  tiny_12x8x16_s10.npy    var after 10 steps, dt=1e-3, istat=5 (forcing on)
  tiny_12x8x16_s10_tau.npy   tauWall.dat rows of that run
  c1_64x32x48_s100.npz    config 1 after 100 steps (dt=0.0015, istat=10):
                          sha256 of the full var bytes, per-field sums, L2,
                          Linf, three full planes, tauWall/moment1st rows
  study_<name>_24x20x16_s10.npz
                          the wall-condition studies (build_ref.sh STUDY)
                          after 10 steps from the seeded 24x20x16 start,
                          dt=1e-3, istat=5: sha256 of var, per-field sums,
                          L2, Linf, the bottom wall plane (7 fields),
                          tauWall and moment1st rows

  restart_24x20x16_topo212_s20_s40.npz
                          the reference on 4 MPI ranks (topology 2x1x2), 20
                          steps from the seeded start, then restarted from its
                          per-rank files to step 40 (dt=1e-3, istat=5): plane
                          k=8 and per-field sums of the assembled state, all
                          tauWall rows, the domInfo.txt text, and sha256 and
                          first bytes of the Mesh.dat of debugOp=1

  post_g1_24x20x16_s50.npz
                          the reference's post-processing chain (oracle/
                          build_post.sh: avg_v1.7, fluc, qInvariant) on the
                          g1 study run for 50 steps (istat=5): its moment
                          files (inputs) and the text of mean/rms/corr/
                          plotInfo.dat, sha256 of the 3D.dat it read, of
                          fluc.dat and qInvariant.dat, and Q on plane k=8

    python tests/golden/make_golden.py studies [NAME ...]   # only the study fixtures
    python tests/golden/make_golden.py post       # only the post-processing one
    python tests/golden/make_golden.py restart    # only the 4-rank restart one
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from compchan3dmpi_amd import formats  # noqa: E402
import ref_run  # noqa: E402


def ic(n1, n2, n3):
    return formats.round_e22_15(formats.laminar_ic(n1, n2, n3))


STUDIES = ("s1", "s2", "s3", "s4", "s5", "s6", "rd1", "rd2", "rd3", "g1", "gst1", "gst2", "gst3",
           "gst4", "gst5", "st1", "st2", "st3", "st4")
STUDY_GRID = (24, 20, 16)


def studies(names=STUDIES):
    n1, n2, n3 = STUDY_GRID
    v = ic(n1, n2, n3)
    for st in names:
        r = ref_run.run(v, formats.Params(dt=0.001, istat=5), 10, 10, study=st)
        var = np.ascontiguousarray(r["var"])
        np.savez(
            os.path.join(HERE, f"study_{st}_{n1}x{n2}x{n3}_s10.npz"),
            sha256=np.frombuffer(hashlib.sha256(var.tobytes()).digest(), dtype=np.uint8),
            sums=var.reshape(7, -1).sum(axis=1),
            l2=np.sqrt((var.reshape(7, -1) ** 2).sum(axis=1)),
            linf=np.abs(var.reshape(7, -1)).max(axis=1),
            plane_k1=var[:, 0], tauWall=r["tauWall"], moment1st=r["moment1st"],
        )
        print("study", st, "written", flush=True)


def post():
    import subprocess
    n1, n2, n3 = STUDY_GRID
    v = ic(n1, n2, n3)
    r = ref_run.run(v, formats.Params(dt=0.001, istat=5), 50, 50, study="g1")
    root = r["root"]
    x1, x2, x3 = formats.mesh(n1, n2, n3)[:3]
    inp = os.path.join(root, "input", "3D.dat")
    formats.write_3d(inp, r["time"], r["nstep"], x1, x2, x3, r["var"])
    exe = subprocess.run([os.path.join(ROOT, "oracle", "build_post.sh"), str(n1), str(n2), str(n3)],
                         check=True, capture_output=True, text=True).stdout.strip().splitlines()[-1]
    codes = os.path.join(root, "codes")
    for prog in ("avg.x", "fluc.x", "qinv.x"):
        subprocess.run([os.path.join(exe, prog)], cwd=codes, check=True, capture_output=True,
                       input="", text=True)
    out = os.path.join(root, "output")
    rd = lambda p: np.frombuffer(open(p, "rb").read(), dtype=np.uint8)
    sha = lambda p: np.frombuffer(hashlib.sha256(open(p, "rb").read()).digest(), dtype=np.uint8)
    qv = np.array([float(l.split()[3]) for l in open(os.path.join(out, "qInvariant.dat")).readlines()[3:]])
    np.savez(
        os.path.join(HERE, f"post_g1_{n1}x{n2}x{n3}_s50.npz"),
        moment1st=rd(os.path.join(out, "moment1st.dat")),
        moment2nd=rd(os.path.join(out, "moment2nd.dat")),
        momentJoint=rd(os.path.join(out, "momentJoint.dat")),
        mean=rd(os.path.join(out, "mean.dat")), rms=rd(os.path.join(out, "rms.dat")),
        corr=rd(os.path.join(out, "corr.dat")), plotInfo=rd(os.path.join(codes, "plotInfo.dat")),
        sha_input=sha(inp), sha_fluc=sha(os.path.join(out, "fluc.dat")),
        sha_qinv=sha(os.path.join(out, "qInvariant.dat")),
        q_k8=qv.reshape(n3, n2, n1)[7],
    )
    print("post-processing fixture written", flush=True)


RESTART_GRID = (24, 20, 16)
RESTART_TOPO = (2, 1, 2)


def restart():
    """The reference on 4 MPI ranks (topology 2 x 1 x 2) restarting from its own
    per-rank files: 20 steps from the seeded start (resOp=1, debugOp=1, so
    Mesh.dat is written too), then a second build with maxstep 40 restarts
    (resOp=2) from the 3Dnnn.dat files of step 20.  Both builds keep the
    reference's E22.15 solution files (build_ref.sh REF_FMT=E22.15), which is
    what its restart reader expects."""
    import shutil
    import subprocess
    import tempfile
    n1, n2, n3 = RESTART_GRID
    v = ic(n1, n2, n3)
    env = dict(os.environ, REF_FMT="E22.15", OMP_NUM_THREADS="1")
    exes = []
    for maxstep in (20, 40):
        out = subprocess.run([os.path.join(ROOT, "oracle", "build_ref.sh"), str(n1), str(n2),
                              str(n3), str(maxstep), "10"], check=True, capture_output=True,
                             text=True, env=env).stdout.strip().splitlines()[-1]
        exes.append(out)
    root = tempfile.mkdtemp(prefix="ccrst_")
    nproc = RESTART_TOPO[0] * RESTART_TOPO[1] * RESTART_TOPO[2]
    prm = formats.Params(dt=0.001, istat=5, topo=RESTART_TOPO, debugOp=1)
    formats.write_run_dir(root, prm, v)
    codes, out = os.path.join(root, "codes"), os.path.join(root, "output")
    mpi = "/opt/conda/bin/mpiexec"
    subprocess.run([mpi, "-n", str(nproc), exes[0]], cwd=codes, check=True, capture_output=True,
                   text=True, env=env, timeout=600)
    dom20 = open(os.path.join(out, "domInfo.txt")).read()
    mesh = open(os.path.join(codes, "Mesh.dat"), "rb").read()
    formats.write_param(os.path.join(root, "param.dat"),
                        formats.Params(dt=0.001, istat=5, topo=RESTART_TOPO, resOp=2))
    p = subprocess.run([mpi, "-n", str(nproc), exes[1]], cwd=codes, check=True,
                       capture_output=True, text=True, env=env, timeout=600)
    t, ns, _, _, _, var = formats.read_rank_files(out)
    assert ns == 40, (ns, p.stdout[-500:])
    tau = np.array([np.array(l.split(), dtype=np.float64)
                    for l in open(os.path.join(out, "tauWall.dat")) if l.strip()])
    np.savez(
        os.path.join(HERE, f"restart_{n1}x{n2}x{n3}_topo212_s20_s40.npz"),
        plane_k8=var[:, 7], sums=var.reshape(7, -1).sum(axis=1), tauWall=tau,
        dominfo=np.frombuffer(dom20.encode(), dtype=np.uint8),
        mesh_sha256=np.frombuffer(hashlib.sha256(mesh).digest(), dtype=np.uint8),
        mesh_head=np.frombuffer(mesh[:400], dtype=np.uint8),
    )
    shutil.rmtree(root, ignore_errors=True)
    print("restart fixture written", flush=True)


def main():
    if sys.argv[1:] == ["restart"]:
        restart()
        return
    if sys.argv[1:2] == ["studies"]:
        studies(tuple(sys.argv[2:]) or STUDIES)
        return
    if sys.argv[1:] == ["post"]:
        post()
        return
    # tiny grid, 1 step, forcing off
    v = ic(12, 8, 16)
    r = ref_run.run(v, formats.Params(dt=0.001, istat=100), 1, 1)
    np.save(os.path.join(HERE, "tiny_12x8x16_s1.npy"), r["var"])
    # tiny grid, 10 steps, statistics every 5 steps (forcing switches on)
    r = ref_run.run(v, formats.Params(dt=0.001, istat=5), 10, 10)
    np.save(os.path.join(HERE, "tiny_12x8x16_s10.npy"), r["var"])
    np.save(os.path.join(HERE, "tiny_12x8x16_s10_tau.npy"), r["tauWall"])
    # config 1
    v = ic(64, 32, 48)
    r = ref_run.run(v, formats.Params(dt=0.0015, istat=10), 100, 100)
    var = np.ascontiguousarray(r["var"])
    np.savez(
        os.path.join(HERE, "c1_64x32x48_s100.npz"),
        sha256=np.frombuffer(hashlib.sha256(var.tobytes()).digest(), dtype=np.uint8),
        sums=var.reshape(7, -1).sum(axis=1),
        l2=np.sqrt((var.reshape(7, -1) ** 2).sum(axis=1)),
        linf=np.abs(var.reshape(7, -1)).max(axis=1),
        plane_k2=var[:, 1], plane_k24=var[:, 23], plane_j16=var[:, :, 15],
        tauWall=r["tauWall"], moment1st=r["moment1st"],
        KE=r["KE"], rhobulk=r["rhobulk"], Ubulk=r["Ubulk"], massflow=r["massflow"],
    )
    studies()
    post()
    restart()
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
