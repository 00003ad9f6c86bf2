"""The FP64 MLP reference: gradient against central finite differences."""
import numpy as np

import mlp_ref


def test_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    P = rng.normal(0, 0.3, mlp_ref.NP)
    w = mlp_ref.unpack(P)
    w["W1"][11:, :] = 0
    w["W4"][:, 9:] = 0
    w["b4"][9:] = 0
    P = mlp_ref.pack(w)
    X = np.zeros((7, 16))
    X[:, :11] = rng.normal(size=(7, 11))
    T = np.zeros((7, 16))
    T[:, :9] = rng.normal(size=(7, 9))
    _, g = mlp_ref.loss_grad(P, X, T)
    idx = np.concatenate([rng.choice(np.arange(0, 11 * 64), 12, replace=False),
                          rng.choice(np.arange(1088, 9408), 20, replace=False),
                          rng.choice(np.arange(9408, 10448), 12, replace=False)])
    h = 1e-6
    for i in idx:
        if g[i] == 0 and (9408 <= i < 10432 and (i - 9408) % 16 >= 9):
            continue
        Pp = P.copy(); Pp[i] += h
        Pm = P.copy(); Pm[i] -= h
        fd = (mlp_ref.loss_grad(Pp, X, T)[0] - mlp_ref.loss_grad(Pm, X, T)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_bf16_rounding_is_round_to_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5])
    y = mlp_ref.bf16(x)
    assert np.array_equal(y, np.array([1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -2.5]))


def test_closure_fallback_decision():
    """The DNS replaces a coarse run that blew up (non-finite shear) or whose
    mean skin friction misses the DNS's by more than the tolerance."""
    from compchan3dmpi_amd import closure
    cf = 1e-3
    good = np.array([[1.01e-3, -0.99e-3], [1.0e-3, -1.0e-3]])
    assert closure.accepted(good, cf)
    assert not closure.accepted(good * 1.05, cf)
    assert not closure.accepted(np.array([[np.nan, -1e-3]]), cf)
    assert not closure.accepted(np.zeros((0, 2)), cf)
    assert closure.accepted(good * 1.02, cf) and not closure.accepted(good * 1.02, cf, tol=0.01)
