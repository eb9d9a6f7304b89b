"""Pins of the oracle's eta-direction (Milne) sweeps, CPU only.

Every other pin of the hydro oracle runs with NZ = 1, where the five eta stencil cells are
equal and the eta flux vanishes whatever its formula.  Here the eta branches are checked at
tau = 2.5 fm (so that a factor tau is visible) with NZ > 1:

* one eta sweep of smooth profiles e(eta), v~(eta) = tau deta/dtau, pi~(eta), Pi(eta) against
  the analytic eta divergence of the covariant conservation law in the reference's tilde
  variables (kernel_ideal.cl:151-157, kernel_visc.cl:257-318, kernel_IS.cl:240-318):
    d(tau Q^mu)/dtau  gets  -d/deta F^mu,
      ideal / viscous KT:  F^mu = (T0^{tau mu} + P g0^mu) v~ + P delta^mu_eta + pi~^{eta mu}
      (P -> P + Pi with the bulk extension; T^{eta eta} carries tau^2, T^{tau eta} tau),
    d(tau u^tau pi^{mn})/dtau  gets  -(1/tau) d/deta (u^tau pi^{mn} v~)   (IS advection),
    udz^mu = (1/tau) du^mu/deta + (v~ u^tau, 0, 0, u^tau)/tau   (covariant derivative:
      Gamma^tau_{eta eta} = tau, Gamma^eta_{eta tau} = 1/tau);
  the L1 error falls at second order as DZ halves (the minmod limiter is first order at
  extrema, so the max error falls at first order), and the mistakes planted in the oracle
  (DZ instead of tau DZ; udz without its Christoffel term) miss by O(1);
* a boost-invariant field on NZ = 7 gives, slice by slice, the NZ = 1 run bit for bit;
* an eta-even field (v~ and pi~^{i eta} odd) stays even, bit for bit;
* the one-step energy balance with eta-dependent hotspots that are vacuum at the eta edges
  (the eta flux has to telescope).
"""
import ctypes as C

import numpy as np
import pytest

from clvisc_amd import ic
from clvisc_amd.eos import Eos
from oracle import oracle as O

TAU = 2.5


@pytest.fixture(scope='module')
def ideal_tab():
    return Eos('ideal_gas').create_table()


@pytest.fixture(scope='module')
def cross_tab():
    return Eos('crossover').create_table()


# --- one eta sweep against the analytic divergence --------------------------------------
def _profiles(eta):
    e = 0.5 + 10.0 * np.exp(-eta ** 2 / 2.0)
    vt = 0.3 * np.tanh(eta / 1.5)
    pi = np.zeros(eta.shape + (10,))
    for m in range(10):
        pi[..., m] = (0.05 + 0.01 * m) * np.exp(-(eta - 0.3 * (m - 4.5) / 4.5) ** 2 / (1.5 + 0.1 * m))
    bk = -0.1 * np.exp(-eta ** 2 / 2.0)
    return e, vt, pi, bk


def _fluxes(eta):
    """Analytic fluxes (ideal-gas EOS, P = e/3) whose eta derivatives the sweeps approximate."""
    e, vt, pi, bk = _profiles(eta)
    P = e / 3.0
    g2 = 1.0 / (1.0 - vt * vt)
    z = np.zeros_like(e)
    f_ideal = np.stack([(e + P) * g2 * vt, z, z, (e + P) * g2 * vt * vt + P], -1)
    Pe = P + bk
    f_visc = np.stack([(e + Pe) * g2 * vt + pi[..., 3], pi[..., 6], pi[..., 8],
                       (e + Pe) * g2 * vt * vt + Pe + pi[..., 9]], -1)
    u0 = np.sqrt(g2)
    f_is = np.concatenate([u0[..., None] * pi * vt[..., None], (u0 * bk * vt)[..., None]], -1) / TAU
    u_over_tau = np.stack([u0, z, z, u0 * vt], -1) / TAU
    return f_ideal, f_visc, f_is, u_over_tau


def _ddeta(which, eta, h=1e-3):
    """d/deta of an analytic flux (6th-order central difference of the closed form in
    float64: truncation ~1e-18, rounding ~1e-13)."""
    w = [(-1.0 / 60, -3), (3.0 / 20, -2), (-3.0 / 4, -1), (3.0 / 4, 1), (-3.0 / 20, 2), (1.0 / 60, 3)]
    return sum(c * _fluxes(eta + s * h)[which] for c, s in w) / h


def _sweep(tab, nz, plant=0, prec='f64', L=12.0):
    """One eta sweep of the ideal KT, viscous KT and IS kernels from zero sources.
    Returns eta and {name: (oracle, analytic)} per cell."""
    dz = L / nz
    eta = (np.arange(nz) - nz // 2) * dz
    e, vt, pi, bk = _profiles(eta)
    p = O.make_params(1, 1, nz, 0.02, 0.08, 0.08, dz, TAU, tab, bulk=(0.05, 0.155, 0.02))
    p.plant = plant
    L_ = O.lib(prec)
    dt = np.float64 if prec == 'f64' else np.float32
    rt = C.c_double if prec == 'f64' else C.c_float
    ev = np.zeros((nz, 4), dt)
    ev[:, 0], ev[:, 3] = e, vt
    pi_a, bk_a = np.ascontiguousarray(pi, dt), np.ascontiguousarray(bk, dt)
    s_id, s_v, ud = np.zeros((nz, 4), dt), np.zeros((nz, 4), dt), np.zeros((nz, 4), dt)
    s_is, s_bk = np.zeros((nz, 10), dt), np.zeros(nz, dt)
    P_ = O._ptr
    L_.orc_ideal_along(C.byref(p), 2, P_(s_id), P_(ev), rt(TAU))
    L_.orc_visc_along(C.byref(p), 2, P_(s_v), P_(ev), P_(pi_a), P_(bk_a), rt(TAU))
    L_.orc_is_along(C.byref(p), 2, P_(s_is), P_(ud), P_(pi_a), P_(ev), P_(s_bk), P_(bk_a), rt(TAU))
    u0 = 1.0 / np.sqrt(1.0 - vt * vt)
    chr_ = np.stack([vt * u0, 0 * u0, 0 * u0, u0], -1) / TAU
    return eta, {
        'ideal KT': (s_id, -_ddeta(0, eta)),
        'viscous KT': (s_v, -_ddeta(1, eta)),
        'IS (pi, Pi)': (np.concatenate([s_is, s_bk[:, None]], -1), -_ddeta(2, eta)),
        'udz': (ud, _ddeta(3, eta) + chr_),
    }


def _errors(eta, res, inner=4.0):
    """(relative L1 error, max error / max |analytic|) over |eta| < inner, worst component."""
    m = np.abs(eta) < inner
    out = {}
    for k, (got, want) in res.items():
        d = np.abs(np.asarray(got, np.float64)[m] - want[m])
        l1 = (d.sum(0) / np.maximum(np.abs(want[m]).sum(0), 1e-30)).max()
        mx = (d.max(0) / np.maximum(np.abs(want[m]).max(0), 1e-30)).max()
        out[k] = (l1, mx)
    return out


def test_eta_sweeps_converge_to_the_analytic_divergence(ideal_tab):
    errs = {nz: _errors(*_sweep(ideal_tab, nz)) for nz in (120, 240, 480)}
    for k in errs[120]:
        l1 = [errs[nz][k][0] for nz in (120, 240, 480)]
        mx = [errs[nz][k][1] for nz in (120, 240, 480)]
        order_l1 = np.log2(np.array(l1[:-1]) / np.array(l1[1:]))
        order_mx = np.log2(np.array(mx[:-1]) / np.array(mx[1:]))
        assert np.all(order_l1 > 1.8), (k, l1, order_l1)
        assert np.all(order_mx > 0.9), (k, mx, order_mx)
        assert l1[-1] < 2e-3, (k, l1)


def test_eta_sweep_fp32_matches_fp64(ideal_tab):
    """The FP32 build evaluates the same eta branches: within FP32 rounding of FP64."""
    eta, r32 = _sweep(ideal_tab, 120, prec='f32')
    _, r64 = _sweep(ideal_tab, 120, prec='f64')
    m = np.abs(eta) < 4.0
    for k in r32:
        a, b = np.asarray(r32[k][0], np.float64)[m], r64[k][0][m]
        assert np.abs(a - b).max() <= 2e-5 * np.abs(b).max(), k


@pytest.mark.parametrize('plant,hit', [(O.PLANT_ETA_DZ, ('ideal KT', 'viscous KT', 'IS (pi, Pi)', 'udz')),
                                       (O.PLANT_UDZ_NO_CHRISTOFFEL, ('udz',))])
def test_eta_sweep_planted_mistakes_fail(ideal_tab, plant, hit):
    """DZ instead of tau*DZ (a factor tau = 2.5 in every eta sweep) and udz without its
    Christoffel term, planted in the oracle, miss the analytic divergence by O(1), far above
    the scheme's own error at the same resolution."""
    good = _errors(*_sweep(ideal_tab, 240))
    bad = _errors(*_sweep(ideal_tab, 240, plant=plant))
    for k in good:
        if k in hit:
            assert bad[k][0] > 0.05 and bad[k][0] > 20 * good[k][0], (k, bad[k], good[k])
        else:
            assert bad[k] == good[k], k


# --- whole steps with NZ > 1 ----------------------------------------------------------------
def _run(tab, nx, ny, nz, ev, nsteps, viscous, bulk=True, dz=0.15, tau0=TAU, prec='f32'):
    p = O.make_params(nx, ny, nz, 0.02, 0.08, 0.08, dz, tau0, tab, etaos_xmin=0.0, etaos_ymin=0.08,
                      bulk=(0.05, 0.155, 0.02) if (bulk and viscous) else None)
    o = O.Oracle(p, prec, viscous=viscous)
    o.load_ini(ev)
    if viscous:
        o.visc_steps(nsteps)
    else:
        o.ideal_steps(nsteps)
    return o


@pytest.mark.parametrize('viscous', [False, True])
def test_boost_invariant_field_on_nz7_equals_nz1(cross_tab, viscous):
    """A field without eta dependence (v~ = 0) on 7 eta cells evolves, slice by slice, as the
    NZ = 1 run: bit for bit, including pi~^{eta eta} and the Christoffel sources at tau = 2.5."""
    nx, ny = 23, 21
    pos = ic.woods_saxon_positions(6, 7, R=0.5, a=0.1)
    ev1 = ic.hotspots(nx, ny, 1, 0.08, 0.08, 0.15, positions=pos, emax=30.0)
    ev7 = np.repeat(ev1.reshape(nx, ny, 1, 4), 7, axis=2).reshape(-1, 4)
    a = _run(cross_tab, nx, ny, 1, ev1, 8, viscous)
    b = _run(cross_tab, nx, ny, 7, ev7, 8, viscous)
    for name, x, y, w in [('ev', a.ev[1], b.ev[1], 4), ('pi', a.pi[1], b.pi[1], 10)] + (
            [('bulk', a.bulk[1][:, None], b.bulk[1][:, None], 1)] if viscous else []):
        if not viscous and name == 'pi':
            continue
        y = y.reshape(nx, ny, 7, w)
        for k in range(7):
            np.testing.assert_array_equal(y[:, :, k, :].reshape(-1, w), x, err_msg='%s slice %d' % (name, k))


@pytest.mark.parametrize('viscous', [False, True])
def test_eta_reflection_symmetry(cross_tab, viscous):
    """eta -> -eta: e, v^x, v^y, Pi and pi^{00,01,02,11,12,22,33} even, v~ and pi^{03,13,23} odd
    after several steps from an eta-even start with an odd v~ (exact: the stencil arithmetic
    is antisymmetric under the reflection)."""
    nx, ny, nz = 15, 13, 11
    x, y, z = ic.grid_coords(nx, ny, nz, 0.08, 0.08, 0.15)
    pos = ic.woods_saxon_positions(5, 9, R=0.4, a=0.1)
    ev = ic.hotspots(nx, ny, nz, 0.08, 0.08, 0.15, positions=pos, emax=20.0, flat=0.2, sigma_eta=0.4)
    ev = ev.reshape(nx, ny, nz, 4).astype(np.float64)
    X, Y, Z = np.meshgrid(x, y, z, indexing='ij')
    ev[..., 1] = 0.1 * X / (1.0 + X * X)
    ev[..., 2] = -0.05 * Y / (1.0 + Y * Y)
    ev[..., 3] = 0.2 * np.tanh(2.0 * Z)
    ev = ev.reshape(-1, 4).astype(np.float32)
    o = _run(cross_tab, nx, ny, nz, ev, 6, viscous, tau0=0.6)
    odd_ev = np.array([1, 1, 1, -1], np.float32)
    odd_pi = np.array([1, 1, 1, -1, 1, 1, -1, 1, -1, 1], np.float32)
    e = o.ev[1].reshape(nx, ny, nz, 4)
    np.testing.assert_array_equal(e[:, :, ::-1, :] * odd_ev, e)
    if viscous:
        pi = o.pi[1].reshape(nx, ny, nz, 10)
        np.testing.assert_array_equal(pi[:, :, ::-1, :] * odd_pi, pi)
        assert np.abs(pi[..., 3]).max() > 0 and np.abs(e[..., 3]).max() > 0
        bk = o.bulk[1].reshape(nx, ny, nz)
        np.testing.assert_array_equal(bk[:, :, ::-1], bk)


def test_eta_reflection_detects_a_planted_mistake(cross_tab):
    """The reflection check is sensitive: the same run with the eta sweeps of the viscous
    stages divided by DZ (planted in the oracle) is no longer what the correct run gives."""
    nx, ny, nz = 15, 13, 11
    pos = ic.woods_saxon_positions(5, 9, R=0.4, a=0.1)
    ev = ic.hotspots(nx, ny, nz, 0.08, 0.08, 0.15, positions=pos, emax=20.0, flat=0.2, sigma_eta=0.4)
    res = []
    for plant in (0, O.PLANT_ETA_DZ):
        p = O.make_params(nx, ny, nz, 0.02, 0.08, 0.08, 0.15, TAU, cross_tab, etaos_xmin=0.0, etaos_ymin=0.08)
        p.plant = plant
        o = O.Oracle(p, 'f32', viscous=True)
        o.load_ini(ev)
        o.visc_steps(4)
        res.append(o.ev[1].copy())
    assert np.abs(res[0][:, 0] - res[1][:, 0]).max() > 1e-3 * np.abs(res[0][:, 0]).max()


@pytest.mark.parametrize('bulk', [False, True])
def test_energy_balance_with_eta_flux(cross_tab, bulk):
    """The one-step FP64 energy balance E_{n+1} - E_n + dt/2 (W_n + W*) = 0 (the NZ = 1 form
    is in test_oracle.py) with eta-dependent hotspots on 31 eta cells that are vacuum at the eta
    and x, y edges: the eta flux telescopes, so the balance still closes to FP32 rounding."""
    from helpers import energy_sums
    nx, ny, nz = 61, 61, 31
    pos = ic.woods_saxon_positions(6, 5, R=0.4, a=0.1)
    ev = ic.hotspots(nx, ny, nz, 0.08, 0.08, 0.15, positions=pos, emax=30.0, width=0.3, flat=0.3,
                     sigma_eta=0.3)
    e3 = ev.reshape(nx, ny, nz, 4)
    assert e3[:, :, 0, 0].max() < 1e-6 * e3[..., 0].max() and e3[0, :, :, 0].max() < 1e-6 * e3[..., 0].max()
    p = O.make_params(nx, ny, nz, 0.02, 0.08, 0.08, 0.15, 0.6, cross_tab, etaos_xmin=0.0, etaos_ymin=0.08,
                      bulk=(0.05, 0.155, 0.02) if bulk else None)
    o = O.Oracle(p, 'f32', viscous=True)
    o.load_ini(ev)
    dV = 0.08 * 0.08 * 0.15
    bk = lambda k: o.bulk[k] if bulk else None
    res = []
    for n in range(8):
        E0, W0 = energy_sums(p, o.ev[1], o.pi[1], bk(1), float(o.tau), dV)
        o.visc_steps(1)
        E1, _ = energy_sums(p, o.ev[1], o.pi[1], bk(1), float(o.tau), dV)
        _, Ws = energy_sums(p, o.ev[2], o.pi[2], bk(2), 1.0, dV)
        res.append((E1 - E0 + 0.5 * p.dt * (W0 + Ws)) / E0)
    # the eta flux is active: the field moves along eta
    assert np.abs(o.ev[1][:, 3]).max() > 1e-2
    assert np.abs(res).max() < 1e-6, res
