"""The oracle's KT step in Cartesian (RIEMANN_TEST) mode against the analytic Riemann problem
of pyvisc/tests/test_riemann.py: a P = e/3 fluid at rest with e = 3 GeV/fm^3 (pressure 1) for
z < 0 and vacuum for z > 0 on the reference's grid (NX = NY = 1, NZ = 401, DZ = 0.05,
DT = 0.01).  The exact solution is the simple rarefaction wave into vacuum: with zeta = z/t,
e = e_L and v = 0 for zeta < -c_s, v = (zeta + c_s)/(1 + zeta c_s) and
e = e_L ((1 - v)/(1 + v))^((1 + c_s^2)/(2 c_s)) for -c_s <= zeta < 1 (the C- Riemann
invariant atanh(v) + c_s/(1 + c_s^2) ln e of the undisturbed fluid), vacuum beyond z = t.
The reference's analytic module (gubser.Riemann) is not in its checkout; this is the textbook
solution (e.g. Rischke, Bernard and Maruhn, Nucl. Phys. A 595 (1995) 346)."""
import numpy as np
import pytest

from clvisc_amd.eos import Eos
from oracle import oracle as O

CS = 1.0 / np.sqrt(3.0)
EL = 3.0


def analytic(z, t):
    zeta = z / t
    v = np.clip((zeta + CS) / (1.0 + zeta * CS), 0.0, 1.0)
    e = EL * ((1.0 - v) / (1.0 + v)) ** ((1.0 + CS * CS) / (2.0 * CS))
    e = np.where(zeta < -CS, EL, e)
    e = np.where(zeta >= 1.0, 0.0, e)
    v = np.where(zeta < -CS, 0.0, v)
    return e, v


@pytest.fixture(scope='module')
def tab():
    return Eos('ideal_gas').create_table()


def run(tab, nz, dz, dt, nsteps, prec='f32', mirror=False):
    p = O.make_params(1, 1, nz, dt, 0.08, 0.08, dz, 1.0, tab, riemann_test=True)
    o = O.Oracle(p, prec, viscous=False)
    z = (np.arange(nz) - nz // 2) * dz
    ev = np.zeros((nz, 4))
    ev[:, 0] = np.where((z > 0) if mirror else (z < 0), EL, 0.0)
    o.load_ini(ev)
    o.ideal_steps(nsteps)
    return z, o.ev[1].copy()


def errors(z, ev, t):
    ea, va = analytic(z, t)
    m = np.abs(z) < 8.0
    l1e = np.abs(ev[m, 0] - ea[m]).sum() / np.abs(ea[m]).sum()
    fan = (z > -CS * t + 0.3) & (z < 0.8 * t)
    dv = np.abs(ev[fan, 3] - va[fan]).mean()
    return l1e, dv


def test_rarefaction_into_vacuum(tab):
    """t = 4 fm on the reference's grid and on grids refined twice: L1(e) below 1 % and
    halving as DZ and DT halve (first order: the solution has kinks at the fan's ends), the
    fan's velocity likewise; the head of the fan moves at -c_s."""
    res = []
    for k in range(3):
        nz, dz, dt = 400 * 2 ** k + 1, 0.05 / 2 ** k, 0.01 / 2 ** k
        z, ev = run(tab, nz, dz, dt, 400 * 2 ** k)
        res.append(errors(z, ev, 4.0))
        if k == 0:
            # undisturbed fluid ahead of the head at z = -c_s t = -2.31 fm (the scheme's
            # smearing of the head's kink reaches about 14 cells ahead of it)
            assert np.all(np.abs(ev[z < -3.0, 0] - EL) < 1e-3 * EL)
            assert np.all(np.abs(ev[z < -3.5, 0] - EL) < 1e-5 * EL)
            assert ev[:, 3].max() < 1.0
    (e0, v0), (e1, v1), (e2, v2) = res
    assert e0 < 0.01 and v0 < 0.01
    for a, b in ((e0, e1), (e1, e2), (v0, v1), (v1, v2)):
        assert 1.6 < a / b < 2.4


def test_mirror_symmetry(tab):
    """The step mirrored (vacuum on the left) gives the mirrored field with v -> -v."""
    z, ev = run(tab, 401, 0.05, 0.01, 300)
    _, em = run(tab, 401, 0.05, 0.01, 300, mirror=True)
    np.testing.assert_allclose(em[::-1, 0], ev[:, 0], rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(-em[::-1, 3], ev[:, 3], rtol=1e-5, atol=1e-6)
