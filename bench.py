#!/usr/bin/env python3
"""Throughput of the channel-DNS time step on one MI355X.

Workload: the reference's physics (Ma 1.5, Re 3000, gamma 1.4, Pr 0.71,
mu = T^0.7, isothermal no-slip walls, periodic x/z) on the
256x192x128 grid = reference nP = [256, 128, 192] (6,291,456 nodes), from a
seeded perturbed laminar start, dt from CFL 0.5.  A "step" is one full time
step: two RK stages plus the istat-cadence wall-shear and plane statistics
that feed the forcing, all on the device.

Prints one JSON line.  value = cell-updates/s per RK stage in FP64
(= n1*n2*n3 * 2 * steps / t: the metric's node count of 6.3M for this grid,
periodic duplicates and wall planes included; config.cells also gives the
(n1-2)(n2-2)(n3-2) nodes a stage actually updates); the mixed-precision
figure is reported in config.mixed, with its own roofline and the per-field
error against FP64 after 1000 steps (config 3), whatever --steps is.
roofline: achieved algorithmic HBM bandwidth of the stage kernel over its
average launch time measured with HIP events on the library's stream, around
the stage launches of every tenth timed step (every step of a short run).  FP64
moves 100 B per node and stage (stage 1 reads U^n and writes U*, stage 2
reads U*, U^n and writes U^{n+1}; 5 FP64 words each); mixed keeps U* in
FP32, 80 B.  cpu_baseline: the reference itself (oracle/_ref) on one MPI rank
per host core the process may use, or the CPU restatement (oracle/) where
that build is absent, on a bounded sample of the same grid.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from compchan3dmpi_amd import ensemble, formats  # noqa: E402
from compchan3dmpi_amd.channel import Channel  # noqa: E402

METRIC = "cell-updates/s per RK stage (FP64 vs mixed) + % HBM roofline, 1 MI355X"
NP = (256, 128, 192)
DT = 9e-5           # CFL 0.5 on the seeded start, five decimals (formats.cfl_dt)
ISTAT = 10
HBM_PEAK_GBS = 8000.0
ALG_BYTES_PER_NODE_STAGE = 100.0        # FP64: 40 B in, 40 B out, + 40 B base state in stage 2
ALG_BYTES_PER_NODE_STAGE_MIXED = 80.0   # U* as FP32: stage 1 40 + 20, stage 2 20 + 40 + 40
ERR_STEPS = 1000                        # config 3: per-field error after 1000 steps


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def profile_every(steps):
    """Stage launches timed by HIP events: those of every tenth step of a
    long run (the events of the other steps would only add their own cost to
    the throughput), every step of a short one."""
    return 10 if steps >= 200 else 1


def run_gpu(var, precision, steps, warmup, profile, keep=False):
    n1, n2, n3 = NP
    ch = Channel(n1, n2, n3, dt=DT, istat=ISTAT, precision=precision)
    ch.upload(var)
    rc = ch.step(warmup)
    if rc:
        raise RuntimeError(f"{precision}: blow-up at warm-up step {rc}")
    if profile:
        ch.profile(-profile_every(steps))
    t0 = time.perf_counter()
    rc = ch.step(steps)          # returns after a device synchronisation
    t1 = time.perf_counter()
    if rc:
        raise RuntimeError(f"{precision}: blow-up at step {rc}")
    kms, kn = ch.profile(0) if profile else (0.0, 0)
    tau = ch.stats()[0]
    final = ch.download()[0] if keep else None
    tiling = ch.tiling()
    ch.close()
    return (t1 - t0), kms, kn, tau, final, tiling


def run_ensemble(steps, warmup):
    """The 16-instance seeded ensemble (mixed precision, persistent batched
    launches, autotuned tiles): cell-updates/s per RK stage summed over the
    members, ms per step and the tuned configuration."""
    return ensemble.bench(16, "mixed", steps, warmup)


def cpu_baseline(var, nsteps):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc  # the CPU restatement, timed as the baseline only

    n1, n2, n3 = NP
    o = orc.Oracle(n1, n2, n3, dt=DT, istat=ISTAT)
    o.load(var)
    t0 = time.perf_counter()
    o.step(nsteps)
    t1 = time.perf_counter()
    rate = n1 * n2 * n3 * 2 * nsteps / (t1 - t0)
    return rate, t1 - t0, int(o.L.orc_threads())


REF_STEPS = 20  # compile-time maxstep of the reference build timed as the CPU baseline


def host_cores():
    """Host cores this process may use: the CPU affinity set, capped by a
    cgroup CPU quota where one is set (a container's CPU share), with
    os.cpu_count() reported beside it."""
    aff = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
        if q != "max":
            quota = max(1, int(int(q) // int(per)))
    except (OSError, ValueError):
        pass
    usable = min(aff, quota) if quota else aff
    return {"usable": usable, "nproc": os.cpu_count(), "affinity": aff, "cgroup_quota": quota}


def topology(nr, n):
    """A 3-D process grid of nr ranks for nP = n: the factorisation whose
    sub-blocks are closest to cubes (smallest halo surface per rank)."""
    best = None
    for a in range(1, nr + 1):
        if nr % a:
            continue
        for b in range(1, nr // a + 1):
            if (nr // a) % b:
                continue
            c = nr // a // b
            if n[0] // a < 4 or n[1] // b < 4 or n[2] // c < 4:
                continue
            bx, by, bz = n[0] / a, n[1] / b, n[2] / c
            surf = bx * by + by * bz + bx * bz
            if best is None or surf < best[0]:
                best = (surf, (a, b, c))
    return best[1] if best else (1, 1, 1)


def cpu_reference(var):
    """The reference itself (oracle/_ref, built from its sources by
    oracle/build_ref.sh, one MPI rank per usable host core) on the same
    grid and start for REF_STEPS steps.  Its rate comes from its own
    10-step CPU-time line (main_v1.9.f90:1066-1067) of the last, warm block.
    Returns None where that build or MPI is not present."""
    import re
    import subprocess
    import tempfile

    from compchan3dmpi_amd import fastio

    n1, n2, n3 = NP
    exe = os.path.join(ROOT, "oracle", "_ref", f"n{n1}x{n2}x{n3}_s{REF_STEPS}_w1000", "main.x")
    mpiexec = "/opt/conda/bin/mpiexec"
    if not (os.path.exists(exe) and os.path.exists(mpiexec)):
        return None
    cores = host_cores()
    nr = min(cores["usable"], 128)  # each rank holds the whole var array (resOp 1)
    tp = topology(nr, NP)
    root = tempfile.mkdtemp(prefix="ccref_")
    try:
        x1, x2, x3 = formats.mesh(n1, n2, n3)[:3]
        for d in ("codes", "output", "input", "debug"):
            os.makedirs(os.path.join(root, d), exist_ok=True)
        formats.write_param(os.path.join(root, "param.dat"),
                            formats.Params(dt=DT, istat=ISTAT, topo=tp))
        fastio.write_solution(os.path.join(root, "input", "3D.dat"), 0.0, 0, x1, x2, x3, var)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        p = subprocess.run([mpiexec, "-n", str(nr), exe], cwd=os.path.join(root, "codes"),
                           capture_output=True, text=True, timeout=600, env=env)
        blocks = re.findall(r"CPU Time:\s*([0-9.]+)\s*sec for 10 iters", p.stdout)
        if p.returncode != 0 or not blocks:
            log(f"[bench] reference run failed (rc {p.returncode}): {p.stdout[-300:]} {p.stderr[-300:]}")
            return None
        sec = float(blocks[-1])
        return {"value": n1 * n2 * n3 * 2 * 10 / sec, "unit": "cell-updates/s", "cores": nr,
                "kind": "reference", "host": cores,
                "sample": f"{REF_STEPS} steps of nP={list(NP)} by the reference (flang build of "
                          f"v1.10/codes/main_v1.9.f90, {nr} MPI ranks, topology "
                          f"{tp[0]}x{tp[1]}x{tp[2]}); rate from its own CPU-time line for steps "
                          f"{REF_STEPS - 9}-{REF_STEPS}: {sec:.3f} s per 10 steps"}
    finally:
        import shutil

        shutil.rmtree(root, ignore_errors=True)


def traffic_from_profiles(precision):
    """Calibrated HBM bytes per stage launch measured with rocprofv3 for this
    grid and precision (profiles/), or None when none was measured."""
    n1, n2, n3 = NP
    path = os.path.join(ROOT, "profiles", f"traffic_stage_{precision}_{n1}x{n2}x{n3}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get("hbm_bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp64-strict", "mixed"])
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--grid", default=None,
                    help="n1,n2,n3 to time instead of the default grid, e.g. one GPU's slab "
                         "of a sharded run (34,128,192 = 1/8 of the default along x)")
    ap.add_argument("--cpu-steps", type=int, default=10,
                    help="steps of the CPU restatement timed for cpu_baseline (0 = skip)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(one channel instance per GPU); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    global NP
    if args.grid:
        NP = tuple(int(x) for x in args.grid.split(","))
    n1, n2, n3 = NP
    nodes = n1 * n2 * n3
    log(f"[bench] seeded IC nP={NP}")
    var = formats.laminar_ic(n1, n2, n3)

    log(f"[bench] {args.precision}: warmup {args.warmup}, timed {args.steps} steps")
    t, kms, kn, tau, _, tiling = run_gpu(var, args.precision, args.steps, args.warmup,
                                         profile=True)
    value = nodes * 2 * args.steps / t
    ms_step = 1000.0 * t / args.steps
    avg_launch_s = (kms / kn) / 1000.0 if kn else float("nan")
    achieved = ALG_BYTES_PER_NODE_STAGE * nodes / avg_launch_s / 1e9
    log(f"[bench] {args.precision}: {ms_step:.4f} ms/step, {value:.4e} cell-updates/s/stage, "
        f"stage kernel {1000 * avg_launch_s:.4f} ms avg over {kn} launches, tau={tau}")

    mixed = None
    if not args.no_mixed and args.precision != "mixed":
        tm, kmm, knm, taum, _, tilm = run_gpu(var, "mixed", args.steps, args.warmup,
                                              profile=True)
        vm = nodes * 2 * args.steps / tm
        mix_launch_s = (kmm / knm) / 1000.0 if knm else float("nan")
        ach_m = ALG_BYTES_PER_NODE_STAGE_MIXED * nodes / mix_launch_s / 1e9
        log(f"[bench] mixed: {1000 * tm / args.steps:.4f} ms/step, ratio {vm / value:.3f}, "
            f"stage kernel {1000 * mix_launch_s:.4f} ms")
        log(f"[bench] mixed vs fp64: {ERR_STEPS} steps each from the same start")
        _, _, _, tau64, v64, _ = run_gpu(var, "fp64", ERR_STEPS, 0, profile=False, keep=True)
        _, _, _, taum1k, vmix, _ = run_gpu(var, "mixed", ERR_STEPS, 0, profile=False, keep=True)
        names = ("rho", "u", "v", "w", "E", "T", "p")
        err = {nm: float(np.linalg.norm(vmix[c] - v64[c]) / np.linalg.norm(v64[c]))
               for c, nm in enumerate(names)}
        ub64 = v64[1, :, 1:-1, 1:-1].mean(axis=(1, 2))
        ubm = vmix[1, :, 1:-1, 1:-1].mean(axis=(1, 2))
        mixed = {"value": vm, "ms_per_step": 1000.0 * tm / args.steps,
                 "ratio_vs_fp64": vm / value,
                 "stage_kernel_ms": 1000 * mix_launch_s,
                 "tiling": tilm,
                 "roofline": {"bound": "hbm", "achieved": ach_m, "peak": HBM_PEAK_GBS,
                              "unit": "GB/s", "frac": ach_m / HBM_PEAK_GBS,
                              "traffic": traffic_from_profiles("mixed"),
                              "alg_bytes_per_node_stage": ALG_BYTES_PER_NODE_STAGE_MIXED,
                              "kernel": "k_stage_fused<float>",
                              "avg_launch_ms": 1000 * mix_launch_s,
                              "launches_timed": knm},
                 "error_steps": ERR_STEPS,
                 "rel_l2_vs_fp64": err,
                 "mean_u_profile_rel_err": float(np.abs(ubm - ub64).max() / np.abs(ub64).max()),
                 "tau_wall": [float(x) for x in taum1k],
                 "tau_wall_fp64": [float(x) for x in tau64]}
        del vmix, v64
        log(f"[bench] mixed: mean-u error {mixed['mean_u_profile_rel_err']:.3e} after {ERR_STEPS} steps")

    ens = None
    if not args.no_ensemble:
        ens = run_ensemble(max(50, args.steps // 5), 5)
        log(f"[bench] ensemble: {ens['ms_per_step']:.4f} ms/step, {ens['value']:.4e} "
            f"cell-updates/s/stage over {ens['instances']} instances, tuned {ens['tuned']}")

    cpu = None
    if args.cpu_steps > 0:
        log("[bench] cpu baseline: the reference build on the host cores")
        cpu = cpu_reference(var)
        if cpu is None:
            log(f"[bench] cpu baseline: no reference build here; oracle, {args.cpu_steps} steps")
            rate, tc, nth = cpu_baseline(var, args.cpu_steps)
            cpu = {"value": rate, "unit": "cell-updates/s", "cores": nth, "kind": "port",
                   "sample": f"{args.cpu_steps} steps of nP={list(NP)} FP64 (C restatement, "
                             f"OpenMP node loops on {nth} threads, {tc:.1f} s)"}
        log(f"[bench] cpu baseline: {cpu['value']:.4e} {cpu['unit']} ({cpu['kind']}, {cpu['cores']} cores)")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision != "mixed" else "f32/f64",
        "data": "synthetic",
        "config": {
            "workload": f"channel DNS time step, {n1}x{n3}x{n2} (reference nP=[{n1},{n2},{n3}]), "
                        "Ma 1.5, Re 3000, RK2, hybrid central/QUICK, seeded perturbed laminar start",
            "nP": list(NP), "dt": DT, "istat": ISTAT, "precision": args.precision,
            "cells": {"counted": nodes, "updated": (n1 - 2) * (n2 - 2) * (n3 - 2),
                      "rule": "value and roofline count n1*n2*n3 nodes per stage"},
            "tiling": tiling,
            "mixed": mixed,
            "ensemble": ens,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_from_profiles(args.precision),
            "kernel": "k_stage_fused" if args.precision != "fp64-strict" else "k_rhs_strict",
            "alg_bytes_per_node_stage": ALG_BYTES_PER_NODE_STAGE,
            "avg_launch_ms": 1000 * avg_launch_s,
            "launches_timed": kn,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
