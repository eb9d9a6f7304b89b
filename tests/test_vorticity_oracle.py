"""The vorticity oracle (orc_vorticity, a restatement of pyvisc/kernel/kernel_vorticity.cl)
against closed forms: a fluid at rest has only the Bjorken-expansion Christoffel term; a
slow rigid rotation about the beam axis has thermal vorticity -hbar c omega / T in the
(x, y) component and kinetic gradients -+omega; a uniform acceleration in time appears in
the (t, x) component; boost-invariant Bjorken flow has none."""
import numpy as np
import pytest

from clvisc_amd.eos import Eos
from oracle import oracle as O

HBARC = 0.19733


@pytest.fixture(scope='module')
def tab():
    return Eos('ideal_gas').create_table()


def _grid(nx, ny, nz):
    return nx * ny * nz


def test_rest_has_only_the_expansion_term(tab):
    nx, ny, nz, tau = 7, 6, 5, 0.8
    p = O.make_params(nx, ny, nz, 0.02, 0.1, 0.1, 0.2, 0.6, tab)
    ev = np.zeros((_grid(nx, ny, nz), 4))
    ev[:, 0] = 3.0
    om = O.vorticity(p, ev, ev, tau, 'f64')
    assert np.all(om[:, 0:6] == 0.0)
    # shear1 row eta (d_eta u_nu + Christoffel): nu = 3 gets -u_tau / tau
    assert np.allclose(om[:, 6 + 12 + 3], 0.5 * (-1.0 / tau), rtol=1e-12)
    # u^eta = 0: u^mu d_mu u_eta and the acceleration terms vanish; nothing else is left
    other = [c for c in range(6, 32) if c != 21]
    assert np.abs(om[:, other]).max() < 1e-14


def test_rigid_rotation(tab):
    """v = (-w y, w x, 0), e uniform: omega^{tz} = hbar c / 2 (d_x beta_y - d_y beta_x)
    = -hbar c w / T (gamma ~ 1 for w r << 1), central differences exact for a linear u."""
    nx = ny = 9
    nz, w, h = 3, 1e-3, 0.1
    p = O.make_params(nx, ny, nz, 0.02, h, h, 0.2, 0.6, tab)
    x = (np.arange(nx) - nx // 2) * h
    X, Y, _ = np.meshgrid(x, x, np.zeros(nz), indexing='ij')
    ev = np.zeros((_grid(nx, ny, nz), 4))
    ev[:, 0] = 2.0
    ev[:, 1] = (-w * Y).reshape(-1)
    ev[:, 2] = (w * X).reshape(-1)
    om = O.vorticity(p, ev, ev, 1.0, 'f64')
    T = O.eos_funcs(p, np.array([2.0]), 'f64')[0, 2]
    inner = ((np.abs(X) < 0.35) & (np.abs(Y) < 0.35)).reshape(-1)
    assert np.allclose(om[inner, 2], -HBARC * w / T, rtol=1e-5)
    # the rest is O(w^2) (gamma varies as w^2 r^2)
    assert np.abs(om[inner][:, [0, 1, 3, 4, 5]]).max() < 1e-3 * HBARC * w / T
    # kinetic: 0.5 d_x u_y = -0.5 w, 0.5 d_y u_x = +0.5 w (u_i = -gamma v^i)
    assert np.allclose(om[inner, 6 + 4 * 1 + 2], -0.5 * w, rtol=1e-5)
    assert np.allclose(om[inner, 6 + 4 * 2 + 1], 0.5 * w, rtol=1e-5)


def test_acceleration_in_time(tab):
    """Uniform v^x growing from 0 to a in DT: omega^{yz} = hbar c / 2 d_tau beta_x = -hbar c a / (2 T DT)."""
    nx, ny, nz, a = 5, 5, 3, 1e-3
    p = O.make_params(nx, ny, nz, 0.02, 0.1, 0.1, 0.2, 0.6, tab)
    ev0 = np.zeros((_grid(nx, ny, nz), 4))
    ev0[:, 0] = 2.0
    ev1 = ev0.copy()
    ev1[:, 1] = a
    om = O.vorticity(p, ev0, ev1, 1.0, 'f64')
    T = O.eos_funcs(p, np.array([2.0]), 'f64')[0, 2]
    g = 1 / np.sqrt(1 - a * a)
    assert np.allclose(om[:, 5], 0.5 * HBARC * (-g * a / T) / 0.02, rtol=1e-9)


def test_bjorken_has_no_vorticity(tab):
    """v = 0 in Milne coordinates at any e(tau): the thermal vorticity vanishes."""
    nx, ny, nz = 5, 4, 6
    p = O.make_params(nx, ny, nz, 0.02, 0.1, 0.1, 0.2, 0.6, tab)
    ev0 = np.zeros((_grid(nx, ny, nz), 4)); ev0[:, 0] = 5.0
    ev1 = ev0.copy(); ev1[:, 0] = 4.8
    om = O.vorticity(p, ev0, ev1, 1.0, 'f64')
    assert np.all(om[:, 0:6] == 0.0)


def test_omega_mu_closed_form(tab):
    """omega^mu = Omega^{mu nu} u_nu for a fluid at rest: (T, -omega^{tx}, -omega^{ty},
    -omega^{tz}); cells below 0.1 efrz are zeroed."""
    nx, ny, nz = 3, 3, 3
    p = O.make_params(nx, ny, nz, 0.02, 0.1, 0.1, 0.2, 0.6, tab)
    ev = np.zeros((27, 4)); ev[:, 0] = 2.0; ev[0, 0] = 0.001
    rng = np.random.default_rng(1)
    om = rng.standard_normal((27, 32))
    out = O.vorticity_mu(p, ev, om, 0.5, 'f64')
    T = O.eos_funcs(p, np.array([2.0]), 'f64')[0, 2]
    np.testing.assert_allclose(out[1:, 0], T)
    np.testing.assert_allclose(out[1:, 1:], -om[1:, 0:3])
    assert np.all(out[0] == 0.0)
