"""Seeded ensembles of uneven channel instances for batched runs.

The ensemble workload: 16 channels whose grids are drawn from the reference
nP = [96,48,64], [128,64,96], [160,80,96] (streamwise, spanwise,
wall-normal; i.e. 96x64x48, 128x96x64, 160x96x80 as x-by-wall-normal-by-span),
bulk Richardson number uniform in [0, 0.5] and Navier slip length uniform in
[0, 0.02] (channel half-height units), mixed precision.  Each instance gets
its own CFL-0.5 time step and its own seeded perturbed laminar start.
"""
from __future__ import annotations

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
