"""Closed-form checks of the physics the reference does not have (Navier wall
slip, buoyancy), on the oracle.  The reference pins only the limits: slip
length 0 is its no-slip wall (v1.9), slip length -> infinity is the
zero-gradient wall of v1.9_groov/v1.9_g1/codes/main_v1.9_g1.f90:611-612."""
import numpy as np

from compchan3dmpi_amd import formats
import oracle as orc

GAMA, MAC = 1.4, 1.5
GGM = GAMA * (GAMA - 1) * MAC ** 2


def ic(n1, n2, n3):
    return formats.round_e22_15(formats.laminar_ic(n1, n2, n3))


def plug(n1, n2, n3, U0=0.8):
    """Uniform flow u = U0 everywhere, walls included; rho = T = 1."""
    var = np.zeros((7, n3, n2, n1))
    var[0] = 1.0
    var[1] = U0
    var[5] = 1.0
    var[4] = 1.0 + GGM * 0.5 * U0 ** 2
    var[6] = 1.0 / (GAMA * MAC ** 2)
    return var


def test_zero_slip_and_zero_buoyancy_are_the_reference():
    v0 = ic(12, 8, 16)
    a = orc.Oracle(12, 8, 16, dt=1e-3, istat=5)
    a.load(v0)
    a.step(10)
    b = orc.Oracle(12, 8, 16, dt=1e-3, istat=5)
    b.load(v0)
    b.set_ext(0.0, 0.0)
    b.step(10)
    assert np.array_equal(a.var(), b.var())


def test_free_slip_keeps_plug_flow_steady():
    """Free-slip walls (Ls -> inf) exert no shear on a uniform flow: it is an
    exact steady state of the discrete step.  No-slip walls shear it."""
    n1, n2, n3 = 10, 8, 14
    v0 = plug(n1, n2, n3)
    fs = orc.Oracle(n1, n2, n3, dt=1e-3, istat=1000)
    fs.load(v0)
    fs.set_ext(0.0, 1e30)
    fs.step(5)
    v = fs.var()
    assert np.abs(v[1, 1:-1, 1:-1, 1:-1] - 0.8).max() < 1e-12
    assert np.abs(v[1, 0, 1:-1, 1:-1] - 0.8).max() < 1e-12   # wall slides with the flow
    ns = orc.Oracle(n1, n2, n3, dt=1e-3, istat=1000)
    ns.load(v0)
    ns.step(5)
    assert np.abs(ns.var()[1, 1, 1:-1, 1:-1] - 0.8).max() > 1e-4


def test_infinite_slip_is_the_zero_gradient_wall():
    """Ls -> inf gives u_1 = -(f1 u_2 + f2 u_3)/f0 (main_v1.9_g1.f90:611),
    and the top wall the b-coefficient mirror of it."""
    n1, n2, n3 = 12, 8, 16
    o = orc.Oracle(n1, n2, n3, dt=1e-3, istat=1000)
    o.load(ic(n1, n2, n3))
    o.set_ext(0.0, 1e30)
    o.step(2)
    coef, _, _ = o.mesh()
    v = o.var()
    for c in (1, 2):
        u = v[c, :, 1:-1, 1:-1]
        g1 = -(coef[1] * u[1] + coef[2] * u[2]) / coef[0]
        assert np.allclose(u[0], g1, rtol=1e-12, atol=1e-14)
        gN = -(coef[7] * u[-2] + coef[8] * u[-3]) / coef[6]
        assert np.allclose(u[-1], gN, rtol=1e-12, atol=1e-14)


def test_small_slip_velocity_is_slip_length_times_wall_shear():
    """u_w = Ls du/dn: for a tiny slip length the wall velocity over Ls is the
    wall shear rate of the no-slip solution (one-sided derivative with the
    reference's coefficients, main_v1.9.f90:59-70)."""
    n1, n2, n3 = 12, 8, 16
    Ls = 1e-7
    o = orc.Oracle(n1, n2, n3, dt=1e-3, istat=1000)
    o.load(ic(n1, n2, n3))
    o.set_ext(0.0, Ls)
    o.step(1)
    ref = orc.Oracle(n1, n2, n3, dt=1e-3, istat=1000)
    ref.load(ic(n1, n2, n3))
    ref.step(1)
    coef, _, _ = ref.mesh()
    u = ref.var()[1, :, 1:-1, 1:-1]
    shear = coef[0] * u[0] + coef[1] * u[1] + coef[2] * u[2]
    uw = o.var()[1, 0, 1:-1, 1:-1]
    assert np.allclose(uw / Ls, shear, rtol=1e-4, atol=1e-6 * np.abs(shear).max())


def hydro(n1, n2, n3, Ri):
    """Isothermal hydrostatic state: dp/dz = -Ri rho, rho = gama Ma^2 p."""
    _, _, x3 = formats.mesh(n1, n2, n3)
    a = Ri * GAMA * MAC ** 2
    p = 1.0 / (GAMA * MAC ** 2) * np.exp(-a * x3)
    var = np.zeros((7, n3, n2, n1))
    var[6] = p[:, None, None]
    var[0] = GAMA * MAC ** 2 * var[6]
    var[5] = 1.0
    var[4] = 1.0
    return var


def max_w(n3, Ri_state, Ri_run, steps=20):
    n1 = n2 = 8
    o = orc.Oracle(n1, n2, n3, dt=2e-4, istat=100000)
    o.load(hydro(n1, n2, n3, Ri_state))
    o.set_ext(Ri_run, 0.0)
    o.step(steps)
    return np.abs(o.var()[3]).max()


def test_buoyancy_balances_the_hydrostatic_state():
    """With the body force on, the hydrostatic profile is steady up to the
    truncation error of the pressure gradient, which falls at second order
    with the wall-normal grid; without it the same profile accelerates."""
    e1 = max_w(24, 0.3, 0.3)
    e2 = max_w(48, 0.3, 0.3)
    off = max_w(24, 0.3, 0.0)
    assert e1 < 0.05 * off
    assert e2 < e1 / 3.0
