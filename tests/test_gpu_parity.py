"""HIP time step against the oracle (CPU restatement, itself bit-exact with the
reference build) on the same seeded inputs.  Tolerances: the north-star bar is
relative L2 <= 1e-12 in the conserved variables after N steps in FP64."""
import os

import numpy as np
import pytest

from compchan3dmpi_amd import formats
from compchan3dmpi_amd.channel import Channel
import oracle as orc

GOLD = os.path.join(os.path.dirname(__file__), "golden")

pytestmark = pytest.mark.gpu


def ic(n1, n2, n3):
    return formats.round_e22_15(formats.laminar_ic(n1, n2, n3))


def relL2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def conserved(var):
    rho = var[0]
    return np.stack([rho, rho * var[1], rho * var[2], rho * var[3], rho * var[4]])


def banner(msg):
    print(f"\n[case] {msg}", flush=True)


def run_pair(n, steps, dt, istat, precision):
    n1, n2, n3 = n
    v0 = ic(n1, n2, n3)
    o = orc.Oracle(n1, n2, n3, dt=dt, istat=istat)
    o.load(v0)
    assert o.step(steps) == 0
    ch = Channel(n1, n2, n3, dt=dt, istat=istat, precision=precision)
    ch.upload(v0)
    assert ch.step(steps) == 0
    vg, ns, t = ch.download()
    assert ns == steps
    return o, ch, vg


@pytest.mark.parametrize("precision", ["fp64-strict", "fp64"])
def test_tiny_one_step(precision):
    banner(f"tiny 12x8x16 1 step {precision}")
    o, ch, vg = run_pair((12, 8, 16), 1, 1e-3, 100, precision)
    ref = np.load(os.path.join(GOLD, "tiny_12x8x16_s1.npy"))
    for c in range(7):
        assert relL2(vg[c], ref[c]) <= 1e-13, c
    assert relL2(conserved(vg), conserved(ref)) <= 1e-13


@pytest.mark.parametrize("precision", ["fp64-strict", "fp64"])
def test_tiny_ten_steps_forcing(precision):
    banner(f"tiny 12x8x16 10 steps istat=5 {precision}")
    o, ch, vg = run_pair((12, 8, 16), 10, 1e-3, 5, precision)
    ref = np.load(os.path.join(GOLD, "tiny_12x8x16_s10.npy"))
    assert relL2(conserved(vg), conserved(ref)) <= 1e-12
    tau_g, s1g, s2g, crg = ch.stats()
    tau_o, s1o, s2o, cro = o.stats()
    assert np.allclose(tau_g, tau_o, rtol=1e-11, atol=1e-15)
    assert np.allclose(s1g, s1o, rtol=1e-12, atol=1e-15)
    assert np.allclose(s2g, s2o, rtol=1e-12, atol=1e-15)
    assert np.allclose(crg, cro, rtol=1e-11, atol=1e-14)


@pytest.mark.parametrize("precision", ["fp64-strict", "fp64"])
def test_c1_hundred_steps(precision):
    """Config 1: 64x32x48, 100 steps, dt from CFL 0.5, istat 10."""
    banner(f"c1 64x32x48 100 steps {precision}")
    o, ch, vg = run_pair((64, 32, 48), 100, 0.0015, 10, precision)
    vo = o.var()
    e = relL2(conserved(vg), conserved(vo))
    print(f"rel L2 conserved = {e:.3e}", flush=True)
    assert e <= 1e-12
    for c in range(7):
        assert relL2(vg[c], vo[c]) <= 1e-12, c


@pytest.mark.parametrize("precision", ["fp64-strict", "fp64", "mixed"])
def test_blowup_detected_at_same_step(precision):
    """A state that goes negative: the HIP check stops at the step the
    reference would stop at (main_v1.9.f90:728-766)."""
    banner(f"blow-up detection {precision}")
    n1, n2, n3 = 12, 8, 16
    v0 = ic(n1, n2, n3)
    v0[1, 5:8, 3:5, 4:7] += 5.0  # velocity spike: T < 0 at step 5, no NaN
    v0[4] = v0[5] + 1.26 * 0.5 * (v0[1] ** 2 + v0[2] ** 2 + v0[3] ** 2)
    o = orc.Oracle(n1, n2, n3, dt=0.01, istat=100)
    o.load(v0)
    rc_o = 0
    for s in range(1, 21):
        if o.step(1):
            rc_o = s
            break
    assert rc_o > 0
    ch = Channel(n1, n2, n3, dt=0.01, istat=100, precision=precision)
    ch.upload(v0)
    rc = ch.step(20)
    assert rc == rc_o
    step, ijk = ch.blown()
    assert step == rc_o and ijk == o.blown()


def test_c1_mixed_precision_error():
    """Mixed precision (FP32 flux/stencil arithmetic, FP64 state, RK and
    reductions) against the FP64 oracle at config 1 after 100 steps: the
    per-field relative L2 error is reported and the mean streamwise velocity
    profile must agree to 1e-5."""
    banner("c1 mixed precision error vs FP64 oracle")
    o, ch, vg = run_pair((64, 32, 48), 100, 0.0015, 10, "mixed")
    vo = o.var()
    names = ["rho", "u", "v", "w", "E", "T", "p"]
    for c in range(7):
        print(f"  {names[c]:>3s} rel L2 = {relL2(vg[c], vo[c]):.3e}", flush=True)
    ubar_g = vg[1, :, 1:-1, 1:-1].mean(axis=(1, 2))
    ubar_o = vo[1, :, 1:-1, 1:-1].mean(axis=(1, 2))
    e = np.abs(ubar_g - ubar_o).max() / np.abs(ubar_o).max()
    print(f"  mean-u profile rel error = {e:.3e}", flush=True)
    assert e <= 1e-5
    assert relL2(conserved(vg), conserved(vo)) <= 1e-4


def test_bulk_diagnostics_match_oracle():
    """KE, bulk density, velocity and mass flow (main_v1.9.f90:799-906) on
    the device against the oracle, with the file-rounded mesh."""
    banner("bulk diagnostics 64x32x48 after 10 steps")
    n1, n2, n3 = 64, 32, 48
    v0 = ic(n1, n2, n3)
    x1, x2, x3 = formats.mesh(n1, n2, n3)
    x3f = np.round(x3, 6)
    x2e = round(x2[-1], 6)
    o = orc.Oracle(n1, n2, n3, dt=0.0015, istat=10)
    o.load(v0, x3_file=x3f, x2_end=x2e)
    o.step(10)
    for prec in ("fp64-strict", "fp64"):
        ch = Channel(n1, n2, n3, dt=0.0015, istat=10, precision=prec)
        ch.upload(v0)
        ch.step(10)
        d = ch.diag(x3f, x2e)
        assert np.allclose(d, o.diag(), rtol=1e-12, atol=1e-15), prec


@pytest.mark.parametrize("grid", [(7, 9, 6), (37, 21, 11), (70, 19, 23)])
@pytest.mark.parametrize("precision", ["fp64-strict", "fp64", "mixed"])
def test_awkward_grids(grid, precision):
    """Grids that fill no tile evenly (x and y partial tiles, fewer unique
    nodes than a tile, the thinnest channel n3 = 6) with statistics every
    step (istat = 1, forcing on from the first step)."""
    banner(f"grid {grid} 6 steps istat=1 {precision}")
    o, ch, vg = run_pair(grid, 6, 1e-3, 1, precision)
    vo = o.var()
    tol = 1e-5 if precision == "mixed" else 1e-12
    assert relL2(conserved(vg), conserved(vo)) <= tol
    tau_g = ch.stats()[0]
    tau_o = o.stats()[0]
    assert np.allclose(tau_g, tau_o, rtol=100 * tol, atol=1e-15)
