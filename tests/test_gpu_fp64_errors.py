"""Per-output error of the FP32 device step against the FP64 oracle (the same
reference-ordered operations evaluated in double) on a viscous + bulk hotspot event after
100 steps.  The device equals the FP32 oracle bit for bit (asserted here again); the
numbers printed say how far FP32 itself is from FP64 for each output: e, v^x, v^y,
tau v^eta, the ten pi^{mu nu} and Pi.  The bounds are loose ceilings of the measured
values (the FP64 run is a different trajectory of a chaotic system), not parity bars."""
import sys

import numpy as np
import pytest

from clvisc_amd import _lib, ic
from clvisc_amd.eos import Eos
from oracle import oracle as O
from helpers import cv_params, orc_params, max_rel

pytestmark = pytest.mark.gpu

PI_NAMES = ('pi00', 'pi01', 'pi02', 'pi03', 'pi11', 'pi12', 'pi13', 'pi22', 'pi23', 'pi33')


def _say(msg):
    print(msg)
    sys.stdout.flush()


def _rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.sqrt((b * b).sum())
    return float(np.sqrt(((a - b) ** 2).sum()) / nb) if nb > 0 else 0.0


def test_fp32_device_vs_fp64_oracle_per_output():
    nx, ny, nz, nsteps = 45, 41, 17, 100
    eos = Eos('crossover')
    tab = eos.create_table()
    p = cv_params(nx, ny, nz, 0.02, 0.08, 0.08, 0.15, 0.6, etaos_xmin=0.0, etaos_ymin=0.08,
                  bulk=(0.05, 0.155, 0.02))
    pos = ic.woods_saxon_positions(24, 11, R=0.9, a=0.15)
    ev0 = ic.hotspots(nx, ny, nz, 0.08, 0.08, 0.15, positions=pos, emax=50.0)
    _say('case: FP32 device vs FP64 oracle, viscous+bulk %dx%dx%d, %d steps' % (nx, ny, nz, nsteps))
    h = _lib.Handle(p, tab, viscous=True)
    h.visc_initialize(p.tau0)
    h.ideal_stage(1, np.float32(p.tau0))
    h.set('ev1', ev0)
    h.visc_steps(0, nsteps)
    ev, pi, bk = h.get('ev1'), h.get('pi1'), h.get('bulk1')
    h.close()
    o32 = O.Oracle(orc_params(p, tab, nthreads=16), 'f32', viscous=True)
    o32.load_ini(ev0)
    o32.visc_steps(nsteps)
    np.testing.assert_array_equal(ev, o32.ev[1])
    np.testing.assert_array_equal(pi, o32.pi[1])
    np.testing.assert_array_equal(bk, o32.bulk[1])
    o64 = O.Oracle(orc_params(p, tab, nthreads=16), 'f64', viscous=True)
    o64.load_ini(ev0)
    o64.visc_steps(nsteps)
    rows = []
    for c, name in enumerate(('e', 'vx', 'vy', 'tau_veta')):
        rows.append((name, max_rel(ev[:, c], o64.ev[1][:, c])[0], _rel_l2(ev[:, c], o64.ev[1][:, c])))
    for c, name in enumerate(PI_NAMES):
        rows.append((name, max_rel(pi[:, c], o64.pi[1][:, c])[0], _rel_l2(pi[:, c], o64.pi[1][:, c])))
    rows.append(('Pi', max_rel(bk, o64.bulk[1])[0], _rel_l2(bk, o64.bulk[1])))
    _say('  output    max rel (floored)   rel L2')
    for name, m, l2 in rows:
        _say('  %-9s %12.3g %12.3g' % (name, m, l2))
    res = {name: (m, l2) for name, m, l2 in rows}
    # measured: e 2.5e-6 / 1.2e-6, v 2e-4 / 3e-7, pi <= 2.1e-3 / 3.6e-6, Pi 4.1e-6 / 9e-7
    assert res['e'][0] < 1e-3 and res['e'][1] < 1e-4
    for name in ('vx', 'vy', 'tau_veta'):
        assert res[name][1] < 1e-3, name
    for name in PI_NAMES + ('Pi',):
        assert res[name][1] < 1e-2, name


def test_c1_ideal_gubser_errors():
    """The ideal Gubser workload (201x201x1, q = 1/fm, e(tau0 = 1 fm, r = 0) = e(T = 0.3 GeV),
    ideal-gas EOS, dtau 0.02 fm, 200 steps to tau = 5 fm): the FP32 device run against the
    analytic e, u^x, u^y and against the FP64 oracle, per output, in r < 4 fm.  The device
    equals the FP32 oracle bit for bit."""
    nx, nsteps = 201, 200
    eos = Eos('ideal_gas')
    tab = eos.create_table()
    e0 = float(eos.f_ed(0.3))
    p = cv_params(nx, nx, 1, 0.02, 0.08, 0.08, 0.15, 1.0)
    ev0, norm = ic.gubser_ideal(nx, nx, 1, 0.08, 0.08, 0.15, 1.0, q=1.0, e0=e0)
    ev0 = ev0.astype(np.float32)
    _say('case: ideal gubser %dx%dx1, %d steps: FP32 device vs analytic and vs FP64 oracle' % (nx, nx, nsteps))
    h = _lib.Handle(p, tab, viscous=False)
    h.set('ev1', ev0)
    h.ideal_steps(0, nsteps)
    ev = h.get('ev1').astype(np.float64)
    h.close()
    o32 = O.Oracle(orc_params(p, tab, nthreads=16), 'f32', viscous=False)
    o32.load_ini(ev0)
    o32.ideal_steps(nsteps)
    np.testing.assert_array_equal(ev.astype(np.float32), o32.ev[1])
    o64 = O.Oracle(orc_params(p, tab, nthreads=16), 'f64', viscous=False)
    o64.load_ini(ev0)
    o64.ideal_steps(nsteps)
    tau = 1.0 + nsteps * 0.02
    ea, _ = ic.gubser_ideal(nx, nx, 1, 0.08, 0.08, 0.15, tau, q=1.0)
    ea[:, 0] *= norm
    x = (np.arange(nx) - nx // 2) * 0.08
    X, Y = np.meshgrid(x, x, indexing='ij')
    m = np.sqrt(X * X + Y * Y).ravel() < 4.0

    def umu(a):
        g = 1.0 / np.sqrt(1.0 - a[:, 1] ** 2 - a[:, 2] ** 2)
        return g, g * a[:, 1], g * a[:, 2]

    ud, uo, ua = umu(ev[m]), umu(o64.ev[1][m]), umu(ea[m])
    rows = [('e', np.abs(ev[m, 0] / ea[m, 0] - 1).max(), np.abs(ev[m, 0] - ea[m, 0]).sum() / ea[m, 0].sum(),
             np.abs(ev[m, 0] / o64.ev[1][m, 0] - 1).max())]
    for c, name in enumerate(('u^tau', 'u^x', 'u^y')):
        rows.append((name, np.abs(ud[c] - ua[c]).max(), np.abs(ud[c] - ua[c]).sum() / np.abs(ua[c]).sum(),
                     np.abs(ud[c] - uo[c]).max()))
    _say('  output   vs analytic: max (rel for e, abs for u)   L1 rel   | vs FP64 oracle: max')
    for name, a, l1, f in rows:
        _say('  %-7s %14.3g %14.3g %14.3g' % (name, a, l1, f))
    res = {r[0]: r[1:] for r in rows}
    # measured: e 0.073 / 0.036 / 3.3e-5; u^mu 0.051 / 0.005-0.008 / 4e-5
    assert res['e'][0] < 0.1 and res['e'][1] < 0.05 and res['e'][2] < 1e-4
    for name in ('u^tau', 'u^x', 'u^y'):
        assert res[name][0] < 0.08 and res[name][1] < 0.02 and res[name][2] < 1e-4, name
