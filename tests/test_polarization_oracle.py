"""The polarization oracle (orc_polarization, a restatement of
pyvisc/kernel/kernel_polarization.cl) pinned by what the physics fixes: the density of a
static surface is the closed-form sum of dV m_T cosh(Y - eta_s) feq; without vorticity there
is no polarization; the polarization is linear in the vorticity; and, since every term is
an antisymmetric tensor contracted with p, the rest-frame time component vanishes for all
three kinds (a missing eta_s boost of the terms breaks it)."""
import numpy as np
import pytest

from oracle import oracle as O

T, M = 0.15, 1.116


def _surface(n, seed, static=False):
    rng = np.random.default_rng(seed)
    sf = np.zeros((n, 8), np.float32)
    sf[:, 0] = rng.uniform(0.01, 0.1, n)
    if not static:
        sf[:, 1:4] = rng.normal(0, 0.02, (n, 3))
        sf[:, 4:7] = rng.normal(0, 0.3, (n, 3)) * np.array([1, 1, 0.2])
    sf[:, 7] = rng.uniform(-2, 2, n)
    return sf


def _mom(m, seed):
    rng = np.random.default_rng(seed)
    pt = rng.uniform(0.2, 2.0, m)
    phi = rng.uniform(0, 2 * np.pi, m)
    mt = np.sqrt(pt ** 2 + M ** 2)
    return np.column_stack([mt, rng.uniform(-1, 1, m), pt * np.cos(phi), pt * np.sin(phi)]).astype(np.float32)


def test_density_closed_form():
    sf = _surface(700, 1, static=True)
    mom = _mom(5, 2)
    _, rho, _ = O.polarization(0, sf, np.zeros((700, 6), np.float32), None, mom, T, M, 'f64')
    for g in range(5):
        mt, Y = float(mom[g, 0]), float(mom[g, 1])
        e = mt * np.cosh(Y - sf[:, 7].astype(np.float64))
        x = np.exp(-e / T)
        ref = (sf[:, 0] * e * x / (1 + x)).sum()
        assert rho[g] == pytest.approx(ref, rel=1e-9)


@pytest.mark.parametrize('kind', [0, 1, 2])
def test_no_vorticity_no_polarization_and_linearity(kind):
    n = 500
    sf = _surface(n, 3)
    mom = _mom(6, 4)
    w = 16 if kind == 1 else 6
    z1, z2 = np.zeros((n, w), np.float32), np.zeros((n, 4), np.float32)
    pol, _, _ = O.polarization(kind, sf, z1, z2, mom, T, M, 'f64')
    assert np.all(pol == 0.0)
    rng = np.random.default_rng(5)
    o1 = rng.normal(0, 0.01, (n, w)).astype(np.float32)
    o2 = rng.normal(0, 0.01, (n, 4)).astype(np.float32)
    p1, _, _ = O.polarization(kind, sf, o1, o2, mom, T, M, 'f64')
    p2, _, _ = O.polarization(kind, sf, 2 * o1, 2 * o2, mom, T, M, 'f64')
    np.testing.assert_allclose(p2, 2 * p1, rtol=1e-9, atol=1e-20)


@pytest.mark.parametrize('kind', [0, 1, 2])
def test_rest_frame_time_component_vanishes(kind):
    n = 800
    sf = _surface(n, 6)
    mom = _mom(8, 7)
    rng = np.random.default_rng(8)
    w = 16 if kind == 1 else 6
    o1 = rng.normal(0, 0.05, (n, w)).astype(np.float32)
    o2 = rng.normal(0, 0.05, (n, 4)).astype(np.float32)
    _, _, lrf = O.polarization(kind, sf, o1, o2, mom, T, M, 'f64')
    scale = np.abs(lrf[:, 1:]).max()
    assert scale > 0
    assert np.abs(lrf[:, 0]).max() <= 1e-9 * scale
    # the Lambda at rest: the rest-frame vector is the lab one
    mom0 = np.array([[M, 0.0, 0.0, 0.0]], np.float32)
    pl, rho, lr = O.polarization(kind, sf, o1, o2, mom0, T, M, 'f64')
    np.testing.assert_allclose(lr[0, 1:], pl[0, 1:] / rho[0], rtol=1e-9)
