"""Post-processing on the device against the restatement of the reference's
programs (oracle/post_ref.py, itself pinned to them in tests/test_post.py):
the fluctuation field and Q-criterion kernels bit for bit on the same
inputs, the whole avg -> fluc -> qInvariant chain of a run directory
byte-identical to the reference programs' files, and the on-device time
averages and post-processing of a live instance."""
import hashlib
import os
import shutil
import tempfile

import numpy as np
import pytest

from compchan3dmpi_amd import fastio, formats, postproc
from compchan3dmpi_amd.channel import Channel, wall_preset
import oracle as orc
import post_ref

pytestmark = pytest.mark.gpu

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "post_g1_24x20x16_s50.npz"))
N1, N2, N3 = 24, 20, 16


def sha(path):
    return hashlib.sha256(open(path, "rb").read()).digest()


def g1_oracle(steps):
    v0 = formats.round_e22_15(formats.laminar_ic(N1, N2, N3))
    x1f, x2f, x3f = formats.file_coords(N1, N2, N3)
    o = orc.Oracle(N1, N2, N3, dt=0.001, istat=5)
    o.load(v0, x3_file=x3f, x2_end=x2f[-1])
    w, spec = wall_preset("g1")
    o.set_wall(w, spec.as_dict(), x1f, x2f)
    return o, v0, (x1f, x2f, x3f), (w, spec)


def test_reference_chain_on_the_device():
    """avg -> fluc -> qinv through compchan3dmpi_amd.postproc (device kernels,
    native I/O): every output file byte-identical to the reference's."""
    print("\n[case] post-processing chain g1 24x20x16", flush=True)
    root = tempfile.mkdtemp(prefix="ccpostg_")
    try:
        for d in ("codes", "output", "input"):
            os.makedirs(os.path.join(root, d))
        for name in ("moment1st", "moment2nd", "momentJoint"):
            open(os.path.join(root, "output", name + ".dat"), "wb").write(GOLD[name].tobytes())
        o, *_ = g1_oracle(50)
        assert o.step(50) == 0
        x1, x2, x3 = formats.mesh(N1, N2, N3)[:3]
        fastio.write_solution(os.path.join(root, "input", "3D.dat"), o.time, o.nstep, x1, x2, x3,
                              o.var())
        assert sha(os.path.join(root, "input", "3D.dat")) == GOLD["sha_input"].tobytes()
        codes = os.path.join(root, "codes")
        postproc.average(codes)
        assert open(os.path.join(root, "output", "mean.dat"), "rb").read() == GOLD["mean"].tobytes()
        postproc.fluc(codes)
        assert sha(os.path.join(root, "output", "fluc.dat")) == GOLD["sha_fluc"].tobytes()
        postproc.qinv(codes)
        assert sha(os.path.join(root, "output", "qInvariant.dat")) == GOLD["sha_qinv"].tobytes()
    finally:
        shutil.rmtree(root)


def test_kernels_bit_exact_with_restatement():
    print("\n[case] fluctuation and Q kernels 24x20x16", flush=True)
    from compchan3dmpi_amd.channel import load_library, _check
    L = load_library()
    o, *_ = g1_oracle(50)
    o.step(50)
    var = o.var()
    rng = np.random.default_rng(3)
    mean = rng.standard_normal((N3, 7))
    out = np.empty_like(var)
    _check(L, L.cc_pp_fluctuation(N1, N2, N3, var, np.ascontiguousarray(mean.T), out))
    F = post_ref.fluc(var, mean)
    assert np.array_equal(out, F)
    # a stretched, non-uniform mesh exercises every three-point weight
    x = np.cumsum(rng.uniform(0.5, 1.5, N1))
    y = np.cumsum(rng.uniform(0.5, 1.5, N2))
    z = np.cumsum(rng.uniform(0.5, 1.5, N3))
    U = np.ascontiguousarray(F[1:4])
    Q = np.empty((N3, N2, N1))
    _check(L, L.cc_pp_qcriterion(N1, N2, N3, x, y, z, U, Q))
    Qr = post_ref.q_invariant(x, y, z, U)
    assert np.array_equal(Q, Qr), np.abs(Q - Qr).max()


def test_live_instance_average_and_postprocess():
    """Time averages accumulated on the device over steps 10..50 against the
    oracle's statistics of the same steps; fluctuations and Q of the device
    state against the restatement applied to the oracle's state."""
    print("\n[case] live average + postprocess g1 24x20x16 fp64", flush=True)
    o, v0, (x1f, x2f, x3f), (w, spec) = g1_oracle(50)
    ch = Channel(N1, N2, N3, dt=0.001, istat=5, precision="fp64")
    ch.set_wall(w, spec, x1f, x2f)
    ch.upload(v0)
    assert ch.step(5) == 0
    ch.average(1)
    assert o.step(5) == 0
    acc1, acc2, accc, n = np.zeros((7, N3)), np.zeros((7, N3)), np.zeros((9, N3)), 0
    for _ in range(9):
        assert ch.step(5) == 0 and o.step(5) == 0
        _, s1, s2, cr = o.stats()
        acc1, acc2, accc, n = acc1 + s1, acc2 + s2, accc + cr, n + 1
    mean, rms, cov, cnt = ch.average(2)
    assert cnt == n == 9
    m_o = acc1 / n
    assert np.allclose(mean, m_o, rtol=1e-12, atol=1e-15)
    assert np.allclose(rms, np.sqrt(np.maximum(acc2 / n - m_o * m_o, 0)), rtol=1e-6, atol=1e-9)
    F, Q = ch.postprocess(None, coords=(x1f, x2f, x3f))
    Fo = post_ref.fluc(o.var(), mean.T)
    assert np.abs(F - Fo).max() <= 1e-12 * np.abs(o.var()).max()
    Qo = post_ref.q_invariant(x1f, x2f, x3f, Fo[1:4])
    assert np.abs(Q - Qo).max() <= 1e-9 * np.abs(Qo).max()
    ch.close()
