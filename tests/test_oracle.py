"""Pins of the CPU oracle (oracle/clvisc_oracle.c) against the reference's own
known answers and against what the mathematics fixes.  CPU only.

Reference known-answer tests restated here:
  minmod4 golden vector       pyvisc/tests/test_kernel_helper.py:20-54
  root-solve round trip       pyvisc/tests/test_kernel_helper.py:57-117 (tightened)
  max reduction               pyvisc/tests/test_reduction.py:10-53
  Bjorken decay (ideal)       pyvisc/tests/test_bjorken.py:31-62
  Gubser ideal flow           pyvisc/tests/test_gubser_ideal.py:22-26
  1st-order viscous Bjorken   pyvisc/tests/test_bjorken_visc.py:74-80
  2nd-order viscous Gubser    pyvisc/tests/test_gubser_visc.py:11-27 (GUBSER_VISC_TEST)
"""
import dataclasses

import numpy as np
import pytest

from clvisc_amd import ic
from clvisc_amd.eos import Eos, octave_nodes
from oracle import oracle as O


@pytest.fixture(scope='module')
def ideal_tab():
    return Eos('ideal_gas').create_table()


@pytest.fixture(scope='module')
def cross():
    e = Eos('crossover')
    return e, e.create_table()


def params(tab, nx=1, ny=1, nz=1, dt=0.01, dx=0.08, dy=0.08, dz=0.3, tau0=0.6, **kw):
    return O.make_params(nx, ny, nz, dt, dx, dy, dz, tau0, tab, **kw)


# --- helpers -------------------------------------------------------------------------
@pytest.mark.parametrize('prec', ['f32', 'f64'])
def test_minmod4_golden(prec):
    a = np.array([1, -2, -1, 4.0])
    b = np.array([2, 0, -2, 4.0])
    c = np.array([3, 2, -3, 4.0])
    r = O.minmod4(O.minmod4(a, b, prec), c, prec)
    np.testing.assert_array_equal(r, [1, 0, -1, 4])


def test_minmod_plants_fail():
    """A minmod that took max instead of min, or dropped the sign test, fails the vector."""
    a = np.array([1, -2, -1, 4.0]); b = np.array([2, 0, -2, 4.0]); c = np.array([3, 2, -3, 4.0])
    wrong_max = lambda x, y: np.maximum(np.abs(x), np.abs(y)) * (np.sign(x) + np.sign(y)) * 0.5
    assert not np.array_equal(wrong_max(wrong_max(a, b), c), [1, 0, -1, 4])


# --- EOS -----------------------------------------------------------------------------
def test_ideal_gas_table_closed_form(ideal_tab):
    p = params(ideal_tab)
    e = np.array([0.0, 0.005, 0.01, 1.0, 9.995, 10.0, 10.005, 123.456, 1999.0], np.float32)
    out = O.eos_lookup(p, e)
    # P = e/3 is linear, so the lerp reproduces it up to FP32 rounding
    np.testing.assert_allclose(out[:, 1], e / 3.0, rtol=3e-6, atol=1e-7)
    np.testing.assert_allclose(out[:, 0], 1.0 / 3.0, rtol=1e-6)
    # T = hbarc (e/(dof pi^2/30)/hbarc)^(1/4): lerp between nodes 0.01 apart
    eos = Eos('ideal_gas')
    m = e > 1.0
    np.testing.assert_allclose(out[m, 2], eos.f_T(e[m].astype(np.float64)), rtol=2e-5)


def test_ideal_gas_row_boundary(ideal_tab):
    """e just below the end of a 1000-texel row interpolates into column 0 of the
    next row (eos_table.h:49-54) -- here 9.995 lies between texels 999 and 1000."""
    p = params(ideal_tab)
    e = np.array([9.995], np.float32)
    P = O.eos_lookup(p, e)[0, 1]
    assert abs(P - 9.995 / 3.0) < 1e-5
    # a lookup that stayed in the row (clamping col+1 to 999) would return P(9.99)
    assert abs(P - 9.99 / 3.0) > 1e-3


def test_above_table_fit(ideal_tab):
    p = params(ideal_tab)
    e = np.array([3000.0], np.float32)
    f = O.eos_funcs(p, e)[0]
    P = 0.3327 * 3000 - 0.3223 * 3000 ** 0.4585 - 0.003906 * 3000 * np.exp(-0.05697 * 3000)
    assert abs(f[0] - P) / P < 1e-5
    assert abs(f[3] - 0.333333333) < 1e-7


def test_octave_nodes_and_weights(cross):
    eos, tab = cross
    p = params(tab)
    nodes = octave_nodes(tab.kmin, tab.kmax, tab.log2j).astype(np.float32)
    out = O.eos_lookup(p, nodes)
    np.testing.assert_array_equal(out, tab.rgba)                # exactly the node values
    mids = (0.5 * (nodes[:-1].astype(np.float64) + nodes[1:])).astype(np.float32)
    out = O.eos_lookup(p, mids)
    # brute-force restatement: weight from the position inside the node interval
    lo, hi = nodes[:-1].astype(np.float64), nodes[1:].astype(np.float64)
    w = (mids - lo) / (hi - lo)
    ref = (1 - w)[:, None] * tab.rgba[:-1] + w[:, None] * tab.rgba[1:]
    np.testing.assert_allclose(out, ref, rtol=2e-6, atol=1e-12)


def test_octave_against_closed_form(cross):
    eos, tab = cross
    p = params(tab)
    e = np.geomspace(1e-4, 900.0, 4001).astype(np.float32)
    out = O.eos_lookup(p, e)
    ref = np.c_[eos.f_cs2(e), eos.f_P(e), eos.f_T(e), eos.f_S(e)]
    err = np.abs(out - ref) / np.abs(ref)
    assert err[:, 1].max() < 5e-5      # P
    assert err[:, 2].max() < 5e-5      # T
    assert err[:, 3].max() < 1e-4      # s
    # below 2^kmin P, T, s fall linearly to 0
    e_lo = np.array([0.0, 2.0 ** (tab.kmin - 1)], np.float32)
    o = O.eos_lookup(p, e_lo)
    np.testing.assert_array_equal(o[0, 1:], 0.0)
    np.testing.assert_allclose(o[1, 1:], 0.5 * tab.rgba[0, 1:], rtol=1e-7)


def test_crossover_thermodynamics(cross):
    """s = (e + P)/T and dP/de = cs2 hold for the table columns."""
    eos, tab = cross
    e = octave_nodes(tab.kmin, tab.kmax, tab.log2j)
    cs2, P, T, s = [tab.rgba[:, c].astype(np.float64) for c in range(4)]
    np.testing.assert_allclose(s, (e + P) / T, rtol=1e-6)
    dP = np.gradient(P, e)
    m = (e > 1e-3) & (e < 500)
    np.testing.assert_allclose(dP[m], cs2[m], rtol=2e-3, atol=2e-4)
    # efrz at TFRZ=0.137 lies in the crossover region
    assert 0.2 < eos.f_ed(0.137) < 0.5


# --- root solve ----------------------------------------------------------------------
@pytest.mark.parametrize('prec', ['f32', 'f64'])
def test_root_round_trip(ideal_tab, prec):
    rng = np.random.default_rng(7)
    n = 20000
    e = rng.uniform(1.0, 10.0, n)
    v = rng.uniform(0.0, 0.999, n)
    P = e / 3.0
    g2 = 1.0 / (1.0 - v * v)
    T00 = (e + P) * g2 - P
    M = (e + P) * g2 * v
    p = params(ideal_tab)
    ed, it = O.root_solve(p, T00, M, prec)
    rel = np.abs(ed - e) / e
    # reference test checks 2 decimals; Newton's quadratic convergence gives far more
    assert rel.max() < (2e-3 if prec == 'f32' else 1e-7)
    # FP64 is limited by the FP32 table values (P at the nodes rounded to float)
    assert np.median(rel) < (1e-6 if prec == 'f32' else 1e-8)
    assert it.max() <= 101


def test_root_fp64_cells_extension(ideal_tab):
    """The FP64 solve for fast cells is closer to the exact inverse than FP32."""
    rng = np.random.default_rng(3)
    e = rng.uniform(1.0, 10.0, 4000)
    v = rng.uniform(0.99, 0.9995, 4000)
    P = e / 3.0
    g2 = 1.0 / (1.0 - v * v)
    T00 = ((e + P) * g2 - P).astype(np.float32)
    M = ((e + P) * g2 * v).astype(np.float32)
    p32 = params(ideal_tab)
    p64 = params(ideal_tab, fp64_root_vmin=0.99)
    e32, _ = O.root_solve(p32, T00, M)
    e64, _ = O.root_solve(p64, T00, M)
    ed_exact, _ = O.root_solve(p32, T00.astype(np.float64), M.astype(np.float64), 'f64')
    err32 = np.abs(e32 - ed_exact) / ed_exact
    err64 = np.abs(e64 - ed_exact) / ed_exact
    assert err64.max() <= err32.max()
    assert err64.max() < 1e-6


def test_max_prop_speed_is_one_for_subluminal(ideal_tab):
    rng = np.random.default_rng(5)
    n = 5000
    ev = np.c_[rng.uniform(0, 50, n), rng.uniform(-0.57, 0.57, (n, 3))]
    ev = ev[(ev[:, 1:] ** 2).sum(1) < 0.999]
    lam = O.max_prop_speed(params(ideal_tab), ev, ev[:, 1])
    np.testing.assert_array_equal(lam, 1.0)


@pytest.mark.parametrize('cs2', ['crossover', 0.5, 0.0])
def test_max_prop_speed_is_one_below_the_face_velocity_bound(cs2, cross):
    """The device returns max(lambda, 1) = 1.0f without evaluating lambda when the face
    component vk^2 <= 0.98 and the table's cs2 lie in [0, 0.5]: lambda <= 1 - 5e-5 in exact
    arithmetic there, whatever the other components (gamma is clamped at 1 - v^2 = 1e-7).
    The reference's FP32 formula (the oracle's max_prop_speed) gives exactly 1.0 on such
    states: near the bound, at |v| -> 1 and beyond (the clamp), at NaN transverse
    components, over the table's whole energy range and beyond it, and on tables whose cs2 is
    0.5 or 0 at every node.  Beyond the bound (vk^2 -> 1) the formula does exceed 1."""
    _, tab = cross
    if cs2 != 'crossover':
        rgba = tab.rgba.copy()
        rgba[:, 0] = cs2
        tab = dataclasses.replace(tab, rgba=rgba, cs2_max=float(cs2), cs2_min=float(cs2))
    p = params(tab)
    rng = np.random.default_rng(17)
    n = 400_000
    vk = np.sqrt(0.98) * (1.0 - rng.uniform(0, 1, n) ** 3 * 0.5) * rng.choice([-1.0, 1.0], n)
    vk = vk.astype(np.float32)
    vk = vk[vk * vk <= np.float32(0.98)]
    n = vk.size
    # transverse part: from none up to |v| = 1 - 1e-9, |v| > 1 (gamma clamped), NaN
    room = np.sqrt(np.maximum(1.0 - vk.astype(np.float64) ** 2, 0.0))
    frac = np.where(rng.uniform(0, 1, n) < 0.5, 1.0 - 10.0 ** rng.uniform(-9, 0, n), rng.uniform(0, 1.5, n))
    phi = rng.uniform(0, 2 * np.pi, n)
    vt1 = (room * frac * np.cos(phi)).astype(np.float32)
    vt2 = (room * frac * np.sin(phi)).astype(np.float32)
    vt1[rng.uniform(0, 1, n) < 0.01] = np.nan
    e = (10.0 ** rng.uniform(-9, 4, n)).astype(np.float32)
    ev = np.zeros((n, 4), np.float32)
    ev[:, 0] = e
    for axis in range(3):
        ev[:, 1 + axis] = vk
        ev[:, 1 + (axis + 1) % 3] = vt1
        ev[:, 1 + (axis + 2) % 3] = vt2
        lam = O.max_prop_speed(p, ev, vk)
        np.testing.assert_array_equal(lam, np.float32(1.0))
    if cs2 != 'crossover':
        return
    # the bound is not vacuous: along the face direction close to the light cone the FP32
    # formula leaves 1 on the crossover table (the maximum is not 1.0f there)
    vk_hi = np.float32(1.0) - rng.uniform(0, 3e-7, 20000).astype(np.float32)
    ev_hi = np.zeros((vk_hi.size, 4), np.float32)
    ev_hi[:, 0] = (10.0 ** rng.uniform(-3, 2, vk_hi.size)).astype(np.float32)
    ev_hi[:, 1] = vk_hi
    assert (O.max_prop_speed(p, ev_hi, vk_hi) > 1.0).any()


def test_max_prop_speed_margin_exceeds_fp32_error():
    """The margin behind the device's lambda shortcut, with u^tau as a free input X >= 1:
    1 - lambda_KT in exact arithmetic stays >= 5.0e-5 for vk^2 <= 0.98 and cs2 in [0, 0.5],
    while helper.h:94-105's FP32 operation sequence (every operation rounded once, as the
    oracle and the device evaluate it) stays within a few ulps of the exact value for X from
    1 to 1e7, i.e. far below the margin.  The closed form of the margin is checked against
    the direct evaluation in float64."""
    f = np.float32
    rng = np.random.default_rng(23)
    n = 2_000_000
    ut = np.where(rng.uniform(0, 1, n) < 0.5, np.exp(rng.uniform(0, np.log(np.sqrt(1e7)), n)),
                  1.0 + rng.uniform(0, 1e-2, n)).astype(f)
    vk = (np.sqrt(0.98) * (1 - rng.uniform(0, 1, n) ** 4 * 0.3) * rng.choice([-1, 1], n)).astype(f)
    cs2 = (0.5 * (1 - rng.uniform(0, 1, n) ** 3)).astype(f)
    keep = vk * vk <= f(0.98)
    ut, vk, cs2 = ut[keep], vk[keep], cs2[keep]
    with np.errstate(all='ignore'):
        uk = ut * vk; ut2 = ut * ut; uk2 = uk * uk
        d = ut2 - uk2
        lam32 = (np.abs(ut * uk * (f(1) - cs2)) + np.sqrt((d - (d - f(1)) * cs2) * cs2)) \
            * (f(1) / (ut2 - (ut2 - f(1)) * cs2))
    U, V, C2 = ut.astype(np.float64), vk.astype(np.float64), cs2.astype(np.float64)
    X, a, c = U * U, 1 - C2, np.sqrt(C2)
    lam = (X * np.abs(V) * a + c * np.sqrt(X * (1 - V * V) * a + C2)) / (X * a + C2)
    margin = X * a * (1 - np.abs(V)) ** 2 / (X * a * (1 - np.abs(V)) + C2 + c * np.sqrt(X * (1 - V * V) * a + C2))
    np.testing.assert_allclose(1 - lam, margin, rtol=1e-6, atol=1e-12)
    assert margin.min() >= 5.0e-5
    err_ulps = np.abs(lam32.astype(np.float64) - lam) / 2.0 ** -24
    assert err_ulps.max() < 16
    assert lam32.max() < 1.0


# --- reduction -----------------------------------------------------------------------
def test_max_ed(ideal_tab):
    rng = np.random.default_rng(11)
    n = 50000
    p = params(ideal_tab, nx=n)
    o = O.Oracle(p, 'f32', viscous=False)
    ev = np.zeros((n, 4), np.float32)
    ev[:, 0] = rng.uniform(0, 1, n)
    o.load_ini(ev)
    assert o.max_ed() == np.float32(ev[:, 0].max())


# --- ideal hydro ---------------------------------------------------------------------
@pytest.mark.parametrize('prec', ['f32', 'f64'])
def test_bjorken_ideal(ideal_tab, prec):
    """1-cell boost-invariant expansion: e/e0 = (tau/tau0)^(-4/3) to 3 decimals."""
    p = params(ideal_tab, dt=0.01, tau0=0.6)
    o = O.Oracle(p, prec, viscous=False)
    o.load_ini(ic.bjorken(1, 30.0))
    hist = []
    for n in range(200):
        hist.append([float(o.tau), o.max_ed()])
        o.ideal_steps(1)
    h = np.array(hist)
    np.testing.assert_almost_equal((h[:, 0] / h[0, 0]) ** (-4.0 / 3.0), h[:, 1] / h[0, 1], 3)


def test_bjorken_plant_fails(ideal_tab):
    """A Christoffel term with the wrong factor (Tzz missing P) breaks the decay law."""
    p = params(ideal_tab, dt=0.01, tau0=0.6)
    o = O.Oracle(p, 'f64', viscous=False)
    o.load_ini(ic.bjorken(1, 30.0))
    o.ideal_steps(200)
    tau = float(o.tau)
    good = (tau / 0.6) ** (-4.0 / 3.0)
    wrong = (tau / 0.6) ** (-1.0)         # what d(tau e)/dtau = 0 would give
    assert abs(o.max_ed() / 30.0 - good) < 1e-3 < abs(o.max_ed() / 30.0 - wrong)


def test_gubser_ideal_analytic(ideal_tab):
    """Gubser flow, q = 1/fm, e(tau0=1, r=0) = e(T=0.3 GeV): L1 error of e and max error
    of v_x in the region r < 4 fm after tau = 1 -> 2 fm (coarser grid than C1 for speed)."""
    eos = Eos('ideal_gas')
    e0 = float(eos.f_ed(0.3))
    nx = 121
    p = params(ideal_tab, nx=nx, ny=nx, nz=1, dt=0.02, dx=0.08, dy=0.08, dz=0.15, tau0=1.0)
    o = O.Oracle(p, 'f64', viscous=False)
    ev, norm = ic.gubser_ideal(nx, nx, 1, 0.08, 0.08, 0.15, 1.0, q=1.0, e0=e0)
    o.load_ini(ev)
    o.ideal_steps(50)
    ea, _ = ic.gubser_ideal(nx, nx, 1, 0.08, 0.08, 0.15, float(o.tau), q=1.0)
    ea[:, 0] *= norm
    x = (np.arange(nx) - nx // 2) * 0.08
    X, Y = np.meshgrid(x, x, indexing='ij')
    m = np.sqrt(X * X + Y * Y).ravel() < 4.0
    e = o.ev[1][:, 0]
    l1 = np.abs(e[m] - ea[m, 0]).sum() / ea[m, 0].sum()
    assert l1 < 0.03
    assert np.abs(o.ev[1][m, 1] - ea[m, 1]).max() < 0.02


def test_riemann_flux_telescopes(ideal_tab):
    """In Cartesian (RIEMANN_TEST) mode the KT sweeps only move T^{tau mu} between cells:
    the grid sum of the flux source vanishes to rounding for a blob away from the edges."""
    nx = 24
    p = params(ideal_tab, nx=nx, ny=nx, nz=nx, dt=0.01, dx=0.1, dy=0.1, dz=0.1, tau0=1.0,
               riemann_test=True)
    o = O.Oracle(p, 'f64', viscous=False)
    x = (np.arange(nx) - nx // 2) * 0.1
    X, Y, Z = np.meshgrid(x, x, x, indexing='ij')
    # compact support: cells near the edges are exact vacuum, so no flux leaves the grid
    bump = np.maximum(0.0, 1.0 - (X ** 2 + Y ** 2 + Z ** 2).ravel() / 0.64) ** 2
    ev = np.zeros((nx ** 3, 4))
    ev[:, 0] = 10.0 * bump
    ev[:, 1] = 0.3 * bump
    o.load_ini(ev)
    o.ideal_stage(1, 1.0)
    s = o.Src.sum(axis=0)
    scale = np.abs(o.Src).sum(axis=0)
    assert np.all(np.abs(s) < 1e-12 * scale + 1e-14)


# --- viscous -------------------------------------------------------------------------
def test_bjorken_viscous_first_order(ideal_tab):
    """8x8x1 uniform cell, eta/s = 0.08: T/T0 follows the 1st-order Bjorken solution
    T/T0 = a^(1/3) (1 + 2/3 b), a = tau0/tau, b = eta/s / (tau0 T0) (1 - a^(2/3))
    (pyvisc/tests/test_bjorken_visc.py:74-80) within 1% over tau0 -> tau0 + 3 fm."""
    p = params(ideal_tab, nx=8, ny=8, nz=1, dt=0.01, dx=0.08, dy=0.08, dz=0.3, tau0=0.6,
               etaos_xmin=0.0, etaos_ymin=0.08)
    o = O.Oracle(p, 'f64', viscous=True)
    o.load_ini(ic.bjorken(64, 30.0))
    eos = Eos('ideal_gas')
    T0 = float(eos.f_T(30.0))
    taus, lhs = [], []
    for n in range(300):
        o.visc_steps(1)
        taus.append(float(o.tau))
        lhs.append(float(eos.f_T(o.max_ed())) / T0)
    tau = np.array(taus)
    a = 0.6 / tau
    b = 0.08 / 0.6 / T0 * 0.19733 * (1.0 - a ** (2.0 / 3.0))
    rhs = a ** (1.0 / 3.0) * (1 + 2.0 / 3.0 * b)
    assert np.abs(np.array(lhs) / rhs - 1).max() < 0.01


def test_fp32_vs_fp64_short_time(cross):
    """FP32 and FP64 oracles agree in the dense region after 5 viscous steps."""
    eos, tab = cross
    pos = ic.woods_saxon_positions(8, 3, R=0.6, a=0.1)
    nx, ny, nz = 25, 23, 7
    ev = ic.hotspots(nx, ny, nz, 0.08, 0.08, 0.15, positions=pos)
    res = []
    for prec in ('f32', 'f64'):
        p = params(tab, nx=nx, ny=ny, nz=nz, dt=0.02, dx=0.08, dy=0.08, dz=0.15, tau0=0.6,
                   etaos_xmin=0.0, etaos_ymin=0.08)
        o = O.Oracle(p, prec, viscous=True)
        o.load_ini(ev)
        o.visc_steps(5)
        res.append(o.ev[1].astype(np.float64))
    m = res[1][:, 0] > 0.05 * res[1][:, 0].max()
    assert np.abs(res[0][m, 0] - res[1][m, 0]).max() / res[1][:, 0].max() < 1e-5


def test_bulk_zero_reduces_to_shear(cross):
    """The bulk extension with zeta/s = 0 reproduces the shear-only oracle bit for bit."""
    eos, tab = cross
    pos = ic.woods_saxon_positions(6, 5, R=0.5, a=0.1)
    nx, ny, nz = 19, 17, 5
    ev = ic.hotspots(nx, ny, nz, 0.08, 0.08, 0.15, positions=pos)
    out = []
    for bulk in (None, (0.0, 0.155, 0.02)):
        p = params(tab, nx=nx, ny=ny, nz=nz, dt=0.02, dx=0.08, dy=0.08, dz=0.15, tau0=0.6,
                   etaos_xmin=0.0, etaos_ymin=0.08, bulk=bulk)
        o = O.Oracle(p, 'f32', viscous=True)
        o.load_ini(ev)
        o.visc_steps(6)
        out.append((o.ev[1].copy(), o.pi[1].copy()))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])


def test_gubser_viscous_analytic(ideal_tab):
    """Second-order viscous Gubser flow (L = 2 fm, lambda_1 = -10, eta/s = 0.2) with the
    GUBSER_VISC_TEST relaxation branch, started from the analytic e, u, pi at tau0 = 1
    and tau0 + DT (udiff from the two), as pyvisc/tests/test_gubser_visc.py sets it up;
    after tau = 1 -> 1.5 fm the fields agree with the analytic solution within r < 2 fm
    (grid 151^2, DX = 0.04, DT = 0.005; the reference's own test uses 501^2)."""
    nx, dx, dt, tau0, L, lam1 = 151, 0.04, 0.005, 1.0, 2.0, -10.0
    p = O.make_params(nx, nx, 1, dt, dx, dx, 0.3, tau0, ideal_tab, etaos_xmin=0.0, etaos_ymin=0.2,
                      lam1=lam1, gubser_visc_test=True, nthreads=8)
    o = O.Oracle(p, 'f32', viscous=True)
    ev1, pi1 = ic.gubser_visc(nx, nx, 1, dx, dx, 0.3, tau0, L, lam1)
    ev2, pi2 = ic.gubser_visc(nx, nx, 1, dx, dx, 0.3, tau0 + dt, L, lam1)
    o.load_ini(ev1, pi1)
    o.ev[2][...] = ev2
    o.pi[2][...] = pi2
    o.update_udiff(o.ev[1], o.ev[2])
    o.visc_steps(100)
    ea, pa = ic.gubser_visc(nx, nx, 1, dx, dx, 0.3, tau0 + 100 * dt, L, lam1)
    x = (np.arange(nx) - nx // 2) * dx
    X, Y = np.meshgrid(x, x, indexing='ij')
    m = np.sqrt(X * X + Y * Y).ravel() < 2.0
    assert np.abs(o.ev[1][m, 0] / ea[m, 0] - 1).max() < 0.01
    assert np.abs(o.ev[1][m, 1] - ea[m, 1]).max() < 0.002
    assert np.abs(o.pi[1][m, 9] / pa[m, 9] - 1).max() < 0.03
    assert np.abs(o.pi[1][m, 4] - pa[m, 4]).max() < 0.02 * np.abs(pa[m, 4]).max()
    # a plausible mistake -- the second-order coefficient lambda_1 off by a factor of 2 in
    # the evolution -- fails the same bounds (this special solution does not depend on eta/s)
    pw = O.make_params(nx, nx, 1, dt, dx, dx, 0.3, tau0, ideal_tab, etaos_xmin=0.0, etaos_ymin=0.2,
                       lam1=lam1 / 2, gubser_visc_test=True, nthreads=8)
    w = O.Oracle(pw, 'f32', viscous=True)
    w.load_ini(ev1, pi1)
    w.ev[2][...] = ev2
    w.pi[2][...] = pi2
    w.update_udiff(w.ev[1], w.ev[2])
    w.visc_steps(100)
    assert np.abs(w.pi[1][m, 9] / pa[m, 9] - 1).max() > 0.03 or \
        np.abs(w.ev[1][m, 0] / ea[m, 0] - 1).max() > 0.01


def test_bulk_dilute_edge_stays_subluminal(cross):
    """Regression: the benchmark's hotspot IC (boost-invariant slice, 385^2, bulk on)
    once drove P + Pi < 0 in the dilute periphery, so v = M/(T00 + P + Pi) exceeded 1
    by step ~66 and e blew up to 1e9.  With Pi >= -P(e) and P + Pi >= 0 in the root
    solve every cell stays below the 0.999 clamp and max e decays."""
    eos, tab = cross
    nx = 385
    pos = ic.woods_saxon_positions(400, 20240601)
    ev = ic.hotspots(nx, nx, 1, 0.08, 0.08, 0.15, positions=pos, emax=50.0)
    p = params(tab, nx=nx, ny=nx, nz=1, dt=0.02, dx=0.08, dy=0.08, dz=0.15, tau0=0.6,
               etaos_xmin=0.0, etaos_ymin=0.08, bulk=(0.05, 0.155, 0.02), nthreads=8)
    o = O.Oracle(p, 'f32', viscous=True)
    o.load_ini(ev)
    for _ in range(15):
        o.visc_steps(5)
        v = np.sqrt((o.ev[1][:, 1:4] ** 2).sum(1)).max()
        assert v <= 0.9991, v
    assert o.ev[1][:, 0].max() < 10.0


@pytest.mark.parametrize('prec', ['f32', 'f64'])
def test_slope_closed_form_equals_nested_minmod(prec):
    """The closed-form slope equals the reference's nested expression
    minmod(0.5 (c-a), minmod(THETA (c-b), THETA (b-a))) (helper.h:94-99, 213-217) in
    value on random, tied, zero and tiny inputs (only the sign of a zero may differ)."""
    dt = np.float32 if prec == 'f32' else np.float64
    rng = np.random.default_rng(7)
    n = 200000
    base = rng.standard_normal((3, n)).astype(dt)
    base[:, : n // 4] = np.round(base[:, : n // 4])            # ties and exact zeros
    base[:, n // 4: n // 2] *= dt(1e-38) if prec == 'f32' else 1e-300   # tiny / subnormal
    a, b, c = base
    th = dt(np.float32(1.1))        # THETA is the float literal 1.1f (helper.h:7)
    half = dt(0.5)

    def minmod(x, y):
        res = np.minimum(np.abs(x), np.abs(y))
        return res * (np.sign(x) + np.sign(y)) * half

    ref = minmod(half * (c - a), minmod(th * (c - b), th * (b - a)))
    got = O.slope(a, b, c, prec)
    np.testing.assert_array_equal(got, ref)     # +0 == -0 for assert_array_equal
    # a mistake of the kind the closed form invites -- dropping the 0.5 on |c-a| -- fails
    wrong = np.where(np.sign(c - b) == np.sign(b - a),
                     np.sign(c - b) * np.minimum(np.abs(c - a), np.minimum(th * np.abs(c - b), th * np.abs(b - a))),
                     0)
    assert not np.array_equal(wrong, ref)


# --- split form of the viscous update_ev (root solve in its own pass) -------------------
def _cons_of(e, v, dirn, P, Pi=0.0):
    """T^tt, T^ti of a perfect fluid with pressure P + Pi (the map the solve inverts)."""
    w = (e + P + Pi) / (1.0 - v * v)
    T00 = w - (P + Pi)
    Ti = (w * v)[:, None] * dirn
    return np.column_stack([T00, Ti])


@pytest.mark.parametrize('seeded', [False, True])
def test_visc_root_from_inverts_the_conserved_map(ideal_tab, seeded):
    """(e, v) -> T^{tau mu} -> visc_root_from recovers e and v^i, plain and seeded (a seed
    at the exact |v| must not move the converged answer); with bulk Pi in the pressure too."""
    rng = np.random.default_rng(11)
    n = 5000
    e = rng.uniform(0.5, 20.0, n)
    v = rng.uniform(0.0, 0.95, n)
    dirn = rng.standard_normal((n, 3))
    dirn /= np.linalg.norm(dirn, axis=1, keepdims=True)
    for Pi, bulk in ((0.0, None), (-0.05, (0.05, 0.155, 0.015))):
        p = params(ideal_tab, bulk=bulk)
        cons = _cons_of(e, v, dirn, e / 3.0, Pi * e / 3.0)
        bk = None if bulk is None else (Pi * e / 3.0).astype(np.float32)
        ev, it = O.visc_root_from(p, cons.astype(np.float32), bk, v if seeded else None)
        assert np.abs(ev[:, 0] - e).max() / e.min() < 2e-3
        assert np.median(np.abs(ev[:, 0] - e) / e) < 1e-5
        assert np.abs(ev[:, 1:] - v[:, None] * dirn).max() < 1e-3
        if seeded:
            assert it.mean() < 2.5


def test_visc_root_from_clamp_and_dilute_cells(ideal_tab):
    """|T^ti| > T^tt is scaled to 0.999 T^tt (subluminal result), and e < acu zeroes v."""
    p = params(ideal_tab)
    cons = np.array([[1.0, 2.0, 0.0, 0.0],       # M > T00: clamped
                     [1e-9, 0.0, 0.0, 0.0],      # vacuum
                     [2.0, 0.0, 0.0, 0.0]], np.float32)
    ev, _ = O.visc_root_from(p, cons)
    assert np.linalg.norm(ev[0, 1:]) < 1.0 and ev[0, 0] > 0
    assert np.all(ev[1, 1:] == 0.0)
    assert ev[2, 0] == pytest.approx(2.0, rel=1e-5) and np.all(ev[2, 1:] == 0.0)


def test_visc_root_from_plant_fails(ideal_tab):
    """A mistake planted on the CPU (pressure without Pi, i.e. bulk ignored) fails the
    inversion check above."""
    rng = np.random.default_rng(12)
    n = 500
    e = rng.uniform(0.5, 20.0, n)
    v = rng.uniform(0.0, 0.9, n)
    dirn = np.tile([1.0, 0.0, 0.0], (n, 1))
    p = params(ideal_tab, bulk=(0.05, 0.155, 0.015))
    Pi = -0.2 * e / 3.0
    cons = _cons_of(e, v, dirn, e / 3.0, Pi).astype(np.float32)
    ev_bad, _ = O.visc_root_from(p, cons, np.zeros(n, np.float32))
    assert np.abs(ev_bad[:, 0] - e).max() / e.min() > 1e-2


@pytest.mark.parametrize('bulk', [None, (0.05, 0.155, 0.02)])
def test_energy_balance_of_the_viscous_step(cross, bulk):
    """Q^tau = tau (T^tautau + pi^tautau) changes in one RK2 step by dt/2 (S(U_n) + S(U*)), where
    the sources sum to -W (-tau^2 T^etaeta, kt_src_christoffel) when the boundary is vacuum
    and U* is the stage-1 state (d_ev[2], d_pi[2]; the reference evaluates stage 1 at the old
    tau, so U* is not U_{n+1}).  With boost-invariant hotspots well inside the grid the FP64
    balance E_{n+1} - E_n + dt/2 (W_n + W*) vanishes to FP32 rounding.  Pins the FP64 energy
    sums (helpers.energy_sums; cv_energy_sum evaluates the same sums) and the source sign;
    E without pi^tautau or W without P fails it."""
    from helpers import energy_sums
    _, tab = cross
    nx = 101
    p = params(tab, nx, nx, 1, dt=0.02, tau0=0.6, etaos_xmin=0.0, etaos_ymin=0.08, bulk=bulk)
    pos = ic.woods_saxon_positions(8, 5, R=0.6, a=0.1)
    ev = ic.hotspots(nx, nx, 1, 0.08, 0.08, 0.3, positions=pos, emax=30.0)
    o = O.Oracle(p, 'f32', viscous=True)
    o.load_ini(ev)
    dV = 0.08 * 0.08 * 0.3
    bk = lambda k: None if bulk is None else o.bulk[k]
    res, res_nopi, res_noP = [], [], []
    for n in range(12):
        E0, W0 = energy_sums(p, o.ev[1], o.pi[1], bk(1), float(o.tau), dV)
        E0x, _ = energy_sums(p, o.ev[1], None, bk(1), float(o.tau), dV)
        o.visc_steps(1)
        E1, _ = energy_sums(p, o.ev[1], o.pi[1], bk(1), float(o.tau), dV)
        E1x, _ = energy_sums(p, o.ev[1], None, bk(1), float(o.tau), dV)
        _, Ws = energy_sums(p, o.ev[2], o.pi[2], bk(2), 1.0, dV)
        # W without P (planted on the CPU): the pressure part of W is P summed over cells
        Pn = O.eos_funcs(p, o.ev[1][:, 0])[:, 0].astype(np.float64).sum() * dV
        res.append((E1 - E0 + 0.5 * p.dt * (W0 + Ws)) / E0)
        res_nopi.append((E1x - E0x + 0.5 * p.dt * (W0 + Ws)) / E0)
        res_noP.append((E1 - E0 + 0.5 * p.dt * (W0 + Ws - 2 * Pn)) / E0)
    res = np.array(res)
    assert np.abs(res).max() < 1e-6, res
    assert np.abs(res_noP).max() > 1e-3
    assert np.abs(res_nopi).max() > 1e-5
