"""Seeded ensembles of uneven channel instances for batched runs.

The ensemble workload: 16 channels whose grids are drawn from the reference
nP = [96,48,64], [128,64,96], [160,80,96] (streamwise, spanwise,
wall-normal; i.e. 96x64x48, 128x96x64, 160x96x80 as x-by-wall-normal-by-span),
bulk Richardson number uniform in [0, 0.5] and Navier slip length uniform in
[0, 0.02] (channel half-height units), mixed precision.  Each instance gets
its own CFL-0.5 time step and its own seeded perturbed laminar start.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import formats

GRIDS = ((96, 48, 64), (128, 64, 96), (160, 80, 96))


def spec(n=16, seed=20261203, precision="mixed", istat=10):
    """Channel keyword arguments and initial states of n seeded instances."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        n1, n2, n3 = GRIDS[int(rng.integers(0, len(GRIDS)))]
        Ri = float(rng.uniform(0.0, 0.5))
        Ls = float(rng.uniform(0.0, 0.02))
        var = formats.laminar_ic(n1, n2, n3, seed=seed + 1 + i)
        dt = formats.cfl_dt(var, 1.5)
        out.append((dict(n1=n1, n2=n2, n3=n3, dt=dt, istat=istat, precision=precision,
                         Ri_b=Ri, slip_len=Ls), var))
    return out


def bench(n=16, precision="mixed", steps=300, warmup=20):
    """The n-instance seeded ensemble in persistent batched launches with
    autotuned tiles: returns cell-updates/s per RK stage summed over the
    members, ms per step and the tuned configuration."""
    from .channel import Batch

    sp = spec(n, precision=precision)
    b = Batch([s for s, _ in sp])
    try:
        for m, (_, v) in zip(b.members, sp):
            m.upload(v)
        tuned = b.tune(trial_steps=3)
        rc = b.step(warmup)
        if rc:
            raise RuntimeError(f"ensemble: blow-up at warm-up step {rc}")
        t0 = time.perf_counter()
        rc = b.step(steps)
        t1 = time.perf_counter()
        if rc:
            raise RuntimeError(f"ensemble: blow-up at step {rc}")
    finally:
        b.close()
    nodes = sum(s["n1"] * s["n2"] * s["n3"] for s, _ in sp)
    return {"value": nodes * 2 * steps / (t1 - t0), "ms_per_step": 1000 * (t1 - t0) / steps,
            "instances": len(sp), "nodes": nodes, "precision": precision,
            "tuned": {k: (list(v) if isinstance(v, tuple) else float(v) if k == "ms_per_step" else v)
                      for k, v in tuned.items()}}


def main():
    ap = argparse.ArgumentParser(description="time the seeded ensemble on the GPU")
    ap.add_argument("--instances", type=int, default=16)
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    a = ap.parse_args()
    print(json.dumps(bench(a.instances, a.precision, a.steps, a.warmup)), flush=True)


if __name__ == "__main__":
    main()
