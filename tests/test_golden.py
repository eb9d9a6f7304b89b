"""Committed golden fixtures (tests/golden/, written by tests/golden/make_golden.py).
CPU only."""
import os

import numpy as np
import pytest

from clvisc_amd import ic
from clvisc_amd.eos import Eos
from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def test_benchmark_hotspot_positions():
    """The 385x385x115 benchmark's 400 Woods-Saxon positions are the committed ones."""
    ref = np.loadtxt(os.path.join(GOLDEN, 'hotspots_400_seed20240601.txt'))
    pos = ic.woods_saxon_positions(400, 20240601)
    np.testing.assert_array_equal(pos, ref)
    r = np.hypot(ref[:, 0], ref[:, 1])
    assert r.max() < 6.62 + 10 * 0.546 and 3.0 < np.median(r) < 6.0


@pytest.fixture(scope='module')
def bulk_ode():
    return np.loadtxt(os.path.join(GOLDEN, 'bjorken_bulk_ode.txt'))


def _bjorken_bulk(dt, prec='f64', zmax=0.1):
    tab = Eos('crossover').create_table()
    p = O.make_params(8, 8, 1, dt, 0.08, 0.08, 0.3, 0.6, tab, etaos_xmin=0.0, etaos_ymin=0.0,
                      bulk=(zmax, 0.18, 0.03), nthreads=1)
    o = O.Oracle(p, prec, viscous=True)
    o.load_ini(ic.bjorken(64, 30.0))
    return o


@pytest.mark.parametrize('prec', ['f64', 'f32'])
def test_bjorken_bulk_against_ode(bulk_ode, prec):
    """Uniform 8x8x1 Bjorken cell with bulk pressure only: e(tau) and Pi(tau) follow the
    FP64 ODE within 0.5% over tau 0.6 -> 3.6 fm, and the error halves with DT (the RK2
    stage sequencing is first order in the sources), i.e. the scheme converges to it."""
    err = {}
    for dt in (0.01, 0.005):
        o = _bjorken_bulk(dt, prec)
        k = int(round(0.1 / dt))
        e_err, p_err = [], []
        for row in bulk_ode[1:]:
            o.visc_steps(k)
            assert abs(float(o.tau) - row[0]) < 1e-5
            e_err.append(o.ev[1][0, 0] / row[1] - 1)
            p_err.append(o.bulk[1][0] / row[2] - 1)
            assert np.allclose(o.ev[1][:, 0], o.ev[1][0, 0], rtol=1e-6)
        err[dt] = (np.abs(e_err).max(), np.abs(p_err).max())
        if prec == 'f32':
            break
    assert err[0.01][0] < 0.005 and err[0.01][1] < 0.005, err
    if prec == 'f64':
        for q in (0, 1):
            assert 0.4 < err[0.005][q] / err[0.01][q] < 0.6, err


def test_bjorken_bulk_plant_fails(bulk_ode):
    """zeta/s off by a factor of two misses the ODE's Pi by far more than the bound."""
    o = _bjorken_bulk(0.01, 'f64', zmax=0.05)
    o.visc_steps(100)
    row = bulk_ode[np.argmin(np.abs(bulk_ode[:, 0] - float(o.tau)))]
    assert abs(o.bulk[1][0] / row[2] - 1) > 0.1
