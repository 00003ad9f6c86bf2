"""Study wall conditions on the device against the oracle (itself bit-exact
with each study's reference build, tests/test_oracle_walls.py): FP64 to the
1e-12 bar on the whole var (wall planes and their periodic duplicates
included), the wall shear and the plane statistics; the mixed mode to the
FP32 tolerance."""
import numpy as np
import pytest

from compchan3dmpi_amd import formats
from compchan3dmpi_amd.channel import Batch, CCError, Channel, wall_preset
import oracle as orc

pytestmark = pytest.mark.gpu

STUDIES = ("s1", "s2", "s3", "s4", "s5", "s6", "rd1", "rd2", "rd3", "g1", "gst1", "gst2", "gst3",
           "gst4", "gst5", "st1", "st2", "st3", "st4")


def banner(msg):
    print(f"\n[case] {msg}", flush=True)


def pair(name, n, steps, precision="fp64", dt=0.001, istat=5):
    n1, n2, n3 = n
    v0 = formats.round_e22_15(formats.laminar_ic(n1, n2, n3))
    x1, x2, x3 = formats.file_coords(n1, n2, n3)
    wall, spec = wall_preset(name)
    o = orc.Oracle(n1, n2, n3, dt=dt, istat=istat)
    o.load(v0, x3_file=x3, x2_end=x2[-1])
    o.set_wall(wall, spec.as_dict(), x1, x2)
    assert o.step(steps) == 0
    ch = Channel(n1, n2, n3, dt=dt, istat=istat, precision=precision)
    ch.set_wall(wall, spec, x1, x2)
    ch.upload(v0)
    assert ch.step(steps) == 0
    vg, ns, t = ch.download()
    assert ns == steps and t == o.time
    out = (vg, o.var(), ch.stats(), o.stats())
    ch.close()
    return out


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("name", STUDIES)
def test_study_fp64_matches_oracle(name):
    banner(f"study {name} 24x20x16, 10 steps, fp64")
    vg, vo, sg, so = pair(name, (24, 20, 16), 10)
    for c in range(7):
        e = rel(vg[c], vo[c]) if np.linalg.norm(vo[c]) > 0 else np.abs(vg[c]).max()
        assert e <= 1e-12, (name, c, e)
    assert np.allclose(sg[0], so[0], rtol=1e-12, atol=1e-17)          # tau
    assert np.allclose(sg[1], so[1], rtol=1e-12, atol=1e-15)          # plane means
    assert np.allclose(sg[3], so[3], rtol=1e-11, atol=1e-15)          # products


@pytest.mark.parametrize("name", ("s3", "g1", "gst5", "st3"))
def test_study_fp64_multi_tile(name):
    """64x32x48 (two tiles across x, the second partial) for 20 steps."""
    banner(f"study {name} 64x32x48, 20 steps, fp64")
    vg, vo, sg, so = pair(name, (64, 32, 48), 20, dt=0.0015, istat=10)
    for c in range(7):
        e = rel(vg[c], vo[c]) if np.linalg.norm(vo[c]) > 0 else np.abs(vg[c]).max()
        assert e <= 1e-12, (name, c, e)
    assert np.allclose(sg[0], so[0], rtol=1e-12, atol=1e-17)


@pytest.mark.parametrize("name", ("rd1", "gst2"))
def test_study_mixed(name):
    banner(f"study {name} 24x20x16, 10 steps, mixed")
    vg, vo, sg, so = pair(name, (24, 20, 16), 10, precision="mixed")
    for c in (0, 1, 4, 5, 6):
        assert rel(vg[c], vo[c]) <= 1e-5, (name, c)


def test_study_conditions_rejected_where_unsupported():
    banner("set_wall argument checks")
    _, spec = wall_preset("s1")
    ch = Channel(24, 20, 16, precision="fp64-strict")
    with pytest.raises(CCError):
        ch.set_wall(0, spec)
    ch.set_wall(0, None)  # the reference wall is fine everywhere
    ch.close()
    ch = Channel(24, 20, 16, precision="fp64", slip_len=0.01)
    with pytest.raises(CCError):
        ch.set_wall(0, spec)
    ch.close()
    b = Batch([dict(n1=24, n2=20, n3=16, dt=0.001)])
    with pytest.raises(CCError):
        b.members[0].set_wall(0, spec)
    b.close()
    ch = Channel(24, 20, 16, dt=0.001, precision="fp64")
    ch.upload(formats.laminar_ic(24, 20, 16))
    ch.step(1)
    with pytest.raises(CCError):  # conditions are set before the first step
        ch.set_wall(0, spec)
    ch.close()
