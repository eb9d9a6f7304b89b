"""The parton-smearing oracle (orc_smear, a restatement of the smearing kernel of
pyvisc/kernel/kernel_gaussian_smearing_new.cl) against closed forms: one parton at rest
gives a Gaussian of total energy E K exp(SIGR^2/2) (the cosh of the eta_s offset), moving
partons give cells that satisfy the fixed-point equation e = K0 - K^2/(K0 + P(e))."""
import numpy as np
import pytest

from clvisc_amd.eos import Eos
from oracle import oracle as O


def _params(tab, nx=41, ny=41, nz=41, dx=0.1, dz=0.1, tau0=0.6):
    return O.make_params(nx, ny, nz, 0.02, dx, dx, dz, tau0, tab)


@pytest.mark.parametrize('prec', ['f32', 'f64'])
def test_parton_at_rest_closed_form(prec):
    """A parton at rest at the origin: T^tautau = K w E exp(-r^2/(2 sigma^2)) cosh(eta_s) and
    T^taueta = -K w E exp(...) sinh(eta_s) (its rapidity seen from a cell at eta_s), whose
    integral tau0 int T^tautau = E K exp(sigma^2/2); for P = e/3 the fixed point is
    e = -K0 + sqrt(4 K0^2 - 3 K^2)."""
    tab = Eos('ideal_gas').create_table()
    h = 0.125                                  # exact in binary: the grids agree exactly
    p = _params(tab, 33, 33, 33, h, h)
    sig, K, E = 0.5, 1.3, 2.0
    p4x4 = np.array([[E, 0.0, 0.0, 0.0, 0.6, 0.0, 0.0, 0.0]])
    ev = O.smear(p, p4x4, sig, sig, K, prec).reshape(33, 33, 33, 4)
    x = (np.arange(33) - 16) * h
    X, Y, Z = np.meshgrid(x, x, x, indexing='ij')
    w = (1 / (2 * np.pi * sig ** 2)) * np.sqrt(1 / (2 * np.pi * sig ** 2)) / 0.6
    g = K * w * E * np.exp(-(X ** 2 + Y ** 2 + Z ** 2) / (2 * sig ** 2))
    K0, K2 = g * np.cosh(Z), (g * np.sinh(Z)) ** 2
    tot = 0.6 * K0.sum() * h ** 3
    assert tot == pytest.approx(E * K * np.exp(sig ** 2 / 2), rel=1e-3)
    expect = -K0 + np.sqrt(4 * K0 ** 2 - 3 * K2)
    big = K0 > 1e-4 * K0.max()
    err = np.abs(ev[..., 0] - expect)[big]
    # FP64: the table's FP32 node values of P = e/3 leave ~3e-8
    assert err.max() <= (2e-6 if prec == 'f32' else 2e-7) * expect.max()
    assert np.all(ev[..., 1:] == 0)


def test_moving_partons_solve_the_fixed_point():
    e = Eos('crossover')
    tab = e.create_table()
    p = _params(tab, 21, 19, 15, 0.25, 0.25)
    rng = np.random.default_rng(4)
    n = 200
    mom = rng.normal(0, 1.5, (n, 3))
    E = np.sqrt((mom ** 2).sum(1) + 0.3 ** 2)
    pos = np.column_stack([np.full(n, 0.6), rng.normal(0, 1.0, (n, 2)), rng.normal(0, 0.3, n)])
    p4x4 = np.column_stack([E, mom, pos])
    ev = O.smear(p, p4x4, 0.6, 0.6, 1.0, 'f64')
    # recompute the smeared T^{tau mu} in NumPy and check e against the inversion
    nx, ny, nz = 21, 19, 15
    gx = (np.arange(nx) - nx // 2) * 0.25
    gy = (np.arange(ny) - ny // 2) * 0.25
    ge = (np.arange(nz) - nz // 2) * 0.25
    Gx, Gy, Ge = np.meshgrid(gx, gy, ge, indexing='ij')
    etas = 0.5 * np.log(np.maximum(pos[:, 0] + pos[:, 3], 1e-14)) - 0.5 * np.log(np.maximum(pos[:, 0] - pos[:, 3], 1e-14))
    mt = np.sqrt(np.maximum(E ** 2 - mom[:, 2] ** 2, mom[:, 0] ** 2 + mom[:, 1] ** 2))
    Yr = 0.5 * (np.log(np.maximum(E + mom[:, 2], 1e-14)) - np.log(np.maximum(E - mom[:, 2], 1e-14)))
    c = 1 / (2 * 0.36)
    w = (c / np.pi) * np.sqrt(c / np.pi) / 0.6
    T = np.zeros((4,) + Gx.shape)
    for m in range(n):
        d = c * ((Gx - pos[m, 1]) ** 2 + (Gy - pos[m, 2]) ** 2 + (Ge - etas[m]) ** 2)
        dl = np.where(d < 50, w * np.exp(-d), 0.0)
        T[0] += dl * mt[m] * np.cosh(Yr[m] - Ge)
        T[1] += dl * mom[m, 0]
        T[2] += dl * mom[m, 1]
        T[3] += dl * mt[m] * np.sinh(Yr[m] - Ge)
    K0 = np.maximum(T[0], 1e-14).reshape(-1)
    K2 = (T[1] ** 2 + T[2] ** 2 + T[3] ** 2).reshape(-1)
    ed = ev[:, 0]
    P = O.eos_funcs(p, ed, 'f64')[:, 0]
    big = K0 > 1e-3 * K0.max()
    resid = np.abs(ed - (K0 - K2 / (K0 + P)))[big] / K0[big]
    assert resid.max() < 1e-9
    assert np.all(ed[big] <= K0[big]) and np.all(ed[big] > 0)
