"""Buoyancy / Navier slip on the GPU against the oracle, and batched
ensembles: every member of a batch computes bit for bit what it computes
alone, before and after the tile autotuner."""
import numpy as np
import pytest

from compchan3dmpi_amd import ensemble, formats
from compchan3dmpi_amd.channel import Batch, Channel
import oracle as orc

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"\n[case] {msg}", flush=True)


def ic(n1, n2, n3, seed=20261203):
    return formats.round_e22_15(formats.laminar_ic(n1, n2, n3, seed=seed))


def relL2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_slip_and_buoyancy_match_oracle():
    banner("slip + buoyancy 16x12x20, 20 steps, fp64 vs oracle")
    n1, n2, n3 = 16, 12, 20
    v0 = ic(n1, n2, n3)
    o = orc.Oracle(n1, n2, n3, dt=1e-3, istat=5)
    o.load(v0)
    o.set_ext(0.3, 0.01)
    assert o.step(20) == 0
    ch = Channel(n1, n2, n3, dt=1e-3, istat=5, precision="fp64", Ri_b=0.3, slip_len=0.01)
    ch.upload(v0)
    assert ch.step(20) == 0
    vg, _, _ = ch.download()
    vo = o.var()
    for c in range(7):
        assert relL2(vg[c], vo[c]) <= 1e-12, c
    tg, s1g, _, crg = ch.stats()
    to, s1o, _, cro = o.stats()
    assert np.allclose(tg, to, rtol=1e-10, atol=1e-15)
    assert np.allclose(s1g, s1o, rtol=1e-12, atol=1e-15)
    assert np.allclose(crg, cro, rtol=1e-10, atol=1e-14)


SPECS = [dict(n1=20, n2=12, n3=16, dt=1e-3, istat=4),
         dict(n1=24, n2=16, n3=20, dt=8e-4, istat=4, Ri_b=0.2, slip_len=0.01),
         dict(n1=16, n2=10, n3=12, dt=1e-3, istat=4, Ri_b=0.5, slip_len=0.02)]


def run_single(sp, var, steps):
    ch = Channel(**sp)
    ch.upload(var)
    assert ch.step(steps) == 0
    v, ns, t = ch.download()
    st = ch.stats()
    ch.close()
    return v, ns, t, st


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("istats", [(4, 4, 4), (4, 5, 4)])
def test_batch_members_equal_single_runs(precision, istats):
    """Members on one istat take their statistics in one batched pair of
    launches, members on different istats one pair each; either way every
    member's state and statistics equal its single run's, bit for bit."""
    banner(f"batch of 3 uneven instances vs single runs, {precision}, istat {istats}")
    specs = [dict(sp, precision=precision, istat=k) for sp, k in zip(SPECS, istats)]
    vars0 = [ic(sp["n1"], sp["n2"], sp["n3"], seed=7 + i) for i, sp in enumerate(specs)]
    b = Batch(specs)
    for m, v in zip(b.members, vars0):
        m.upload(v)
    assert b.step(12) == 0
    for i, (sp, m) in enumerate(zip(specs, b.members)):
        vb, nsb, tb = m.download()
        sb = m.stats()
        vs, nss, ts, ss = run_single(sp, vars0[i], 12)
        assert nsb == nss == 12 and tb == ts
        assert np.array_equal(vb, vs), i
        for a, c in zip(sb, ss):
            assert np.array_equal(a, c), i
    b.close()


def test_autotuner_restores_state_and_keeps_results():
    banner("batch autotuner, mixed")
    specs = [dict(sp, precision="mixed") for sp in SPECS]
    vars0 = [ic(sp["n1"], sp["n2"], sp["n3"], seed=11 + i) for i, sp in enumerate(specs)]
    b = Batch(specs)
    for m, v in zip(b.members, vars0):
        m.upload(v)
    best = b.tune(trial_steps=2)
    print("  tuned:", best, flush=True)
    assert best["ms_per_step"] > 0 and best["kchunk"] >= 1
    assert b.step(6) == 0
    for i, (sp, m) in enumerate(zip(specs, b.members)):
        vb, nsb, _ = m.download()
        vs, nss, _, _ = run_single(sp, vars0[i], 6)
        assert nsb == 6 and np.array_equal(vb, vs), i
    b.close()


def test_ensemble_workload_runs_tuned():
    """The 16-instance seeded ensemble (three grid sizes, Ri in [0,0.5],
    slip length in [0,0.02], mixed precision) in tuned persistent launches;
    two members are checked against their single runs."""
    banner("16-instance ensemble, mixed, tuned")
    sp = ensemble.spec(16)
    b = Batch([s for s, _ in sp])
    for m, (_, v) in zip(b.members, sp):
        m.upload(v)
    print("  tuned:", b.tune(trial_steps=2), flush=True)
    assert b.step(20) == 0
    for i in (0, 7):
        vb, _, _ = b.members[i].download()
        vs, _, _, _ = run_single(sp[i][0], sp[i][1], 20)
        assert np.array_equal(vb, vs), i
        assert np.isfinite(vb).all()
    b.close()
