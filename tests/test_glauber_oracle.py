"""The optical Glauber oracle (oracle/glauber_oracle.c, a restatement of
pyvisc/kernel/kernel_glauber.cl) pinned by what the mathematics fixes: scipy quadrature of
the Woods-Saxon thickness, the nucleus normalisation, the limits of the wounded-nucleon
formula, symmetries and the closed-form rapidity profile."""
import math

import numpy as np
import pytest
from scipy.integrate import quad

from clvisc_amd.config import read_config
from clvisc_amd.glauber import weight_mean_b
from oracle import oracle as O


@pytest.fixture(scope='module')
def g():
    return O.glauber_params(read_config())


def _ws(g, r):
    return g.ro0 / (1.0 + math.exp((r - g.ra) / g.eta))


@pytest.mark.parametrize('x,y', [(0.0, 0.0), (3.0, 1.0), (6.0, 2.0), (9.0, 0.0)])
def test_thickness_against_scipy(g, x, y):
    """T(x, y) = 2 int_0^{5R} rho(sqrt(x^2 + y^2 + z^2)) dz: the 50-point Gauss-Legendre rule
    agrees with adaptive quadrature (FP64 build to the rule's own error, FP32 to rounding)."""
    ref = 2.0 * quad(lambda z: _ws(g, math.sqrt(x * x + y * y + z * z)), 0.0, 5.0 * g.ra, epsabs=0, epsrel=1e-12)[0]
    t64 = O.glauber_thickness(g, x, y, 'f64')
    t32 = O.glauber_thickness(g, x, y, 'f32')
    assert abs(t64 - ref) <= 2e-5 * ref
    assert abs(t32 - t64) <= 2e-6 * t64


def test_thickness_integrates_to_the_nucleus(g):
    """int T dx dy = int rho d^3r (both cut at 5R along z; the plane to 5R as well)."""
    h = 0.4
    xs = np.arange(-5 * g.ra, 5 * g.ra + h / 2, h)
    t = np.array([[O.glauber_thickness(g, float(a), float(b), 'f64') for b in xs] for a in xs])
    total = t.sum() * h * h
    ref = quad(lambda r: 4 * math.pi * r * r * _ws(g, r), 0.0, 5.0 * g.ra)[0]
    assert abs(total - ref) <= 2e-3 * ref
    assert abs(ref - g.num_nucleons) <= 0.01 * g.num_nucleons      # NucleonDensity's purpose


def test_wounded_nucleon_limits(g):
    """Nw(Ta, 0) = 0; small cross section: Nw -> 2 Si0 Ta Tb; A -> infinity:
    Nw -> Ta (1 - exp(-Si0 Tb)) + Tb (1 - exp(-Si0 Ta))."""
    assert O.glauber_nw(g, 1.3, 0.0, 'f64') == 0.0
    gs = O.glauber_params(read_config())
    gs.si0 = 1e-4
    ta, tb = 1.7, 0.9
    assert O.glauber_nw(gs, ta, tb, 'f64') == pytest.approx(2 * gs.si0 * ta * tb, rel=1e-3)
    ga = O.glauber_params(read_config())
    ga.num_nucleons = 1e7
    exact = ta * (1 - math.exp(-ga.si0 * tb)) + tb * (1 - math.exp(-ga.si0 * ta))
    assert O.glauber_nw(ga, ta, tb, 'f64') == pytest.approx(exact, rel=1e-5)


def test_glauber_ini_profile_and_symmetry(g):
    """Peak Edmax at the centre for b = 0; e(x, y) = e(-x, y) = e(x, -y) for any b (the
    nuclei swap); the eta profile is the plateau with Gaussian tails."""
    nx, ny, nz, dx, dz = 41, 41, 31, 0.3, 0.3
    g0 = O.glauber_params(read_config())
    g0.b = 0.0
    ev = O.glauber_ini(g0, nx, ny, nz, dx, dx, dz, 'f64').reshape(nx, ny, nz, 4)
    assert ev[nx // 2, ny // 2, nz // 2, 0] == pytest.approx(g0.edmax, rel=1e-12)
    assert np.all(ev[..., 1:] == 0.0)
    ev = O.glauber_ini(g, nx, ny, nz, dx, dx, dz, 'f64').reshape(nx, ny, nz, 4)
    e = ev[..., 0]
    np.testing.assert_allclose(e, e[::-1, :, :], rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(e, e[:, ::-1, :], rtol=1e-12, atol=1e-300)
    etas = (np.arange(nz) - nz // 2) * dz
    d = np.abs(etas)
    h = np.where(d > g.eta_flat, np.exp(-(d - g.eta_flat) ** 2 / (2 * g.eta_gw ** 2)), 1.0)
    prof = e[nx // 2, ny // 2, :] / e[nx // 2, ny // 2, nz // 2]
    np.testing.assert_allclose(prof, h, rtol=1e-12)
    # FP32 build against FP64 (one rounding chain, libm exp/pow)
    e32 = O.glauber_ini(g, nx, ny, nz, dx, dx, dz, 'f32').reshape(nx, ny, nz, 4)[..., 0]
    big = e > 1e-3 * e.max()
    assert np.abs(e32[big] - e[big]).max() <= 2e-5 * e.max()


def test_participants_at_zero_impact_parameter(g):
    """N_part(b = 0) = int Nw(T, T) dx dy for Au+Au at 40 mb: about 380, just below both
    2 A = 394 and the 393 participants of the reference's Au+Au 0-5 % class
    (pyvisc/glauber.py:26); a missing factor 2 in T or a dropped Nw term lands far off."""
    h = 0.25
    n = int(2 * 12.0 / h) + 1
    xs = -12.0 + h * np.arange(n)
    g0 = O.glauber_params(read_config())
    g0.b = 0.0
    npart = 0.0
    for a in xs:
        for b in xs:
            t = O.glauber_thickness(g0, float(a), float(b), 'f64')
            npart += O.glauber_nw(g0, t, t, 'f64')
    npart *= h * h
    assert 370.0 < npart < 393.0


def test_weight_mean_b_inside_its_class():
    b05 = weight_mean_b(0, 5)
    b2030 = weight_mean_b(20, 30)
    assert 0.0 < b05 < 3.3 and 6.7 < b2030 < 8.2
    assert weight_mean_b(0, 5, 'Pb+Pb') > b05
