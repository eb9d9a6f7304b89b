"""FP64 energy sums on the device (cv_energy_sum) against the same sums in NumPy float64
on the device's own fields, and the one-step energy balance of the device run."""
import sys

import numpy as np
import pytest

from clvisc_amd import _lib, ic
from clvisc_amd.eos import Eos
from helpers import cv_params, orc_params, energy_sums

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg)
    sys.stdout.flush()


@pytest.mark.parametrize('nz,bulk', [(1, None), (1, (0.05, 0.155, 0.02)), (13, (0.05, 0.155, 0.02))])
def test_energy_sum_and_balance(nz, bulk):
    nx = 101
    dz = 0.3 if nz == 1 else 0.15
    tab = Eos('crossover').create_table()
    p = cv_params(nx, nx, nz, 0.02, 0.08, 0.08, dz, 0.6, etaos_xmin=0.0, etaos_ymin=0.08, bulk=bulk)
    pos = ic.woods_saxon_positions(8, 5, R=0.6, a=0.1)
    ev = ic.hotspots(nx, nx, nz, 0.08, 0.08, dz, positions=pos, emax=30.0)
    _say('case: energy sums %dx%dx%d bulk=%s, 12 steps' % (nx, nx, nz, bulk is not None))
    h = _lib.Handle(p, tab, viscous=True)
    h.visc_initialize(p.tau0)
    h.ideal_stage(1, np.float32(p.tau0))
    h.set('ev1', ev)
    op = orc_params(p, tab)
    dV = float(p.dx) * float(p.dy) * float(p.dz)     # the device's FP32 spacings, multiplied in double
    bk = lambda k: h.get('bulk%d' % k) if bulk else None
    tau = lambda n: float(np.float32(p.tau0 + n * p.dt))
    res = []
    for n in range(12):
        E0, W0 = h.energy_sum(tau(n), 1)
        if n % 4 == 0:
            Er, Wr = energy_sums(op, h.get('ev1'), h.get('pi1'), bk(1), tau(n), dV)
            _say('  step %d: E %.12g (numpy %.12g) W %.12g (numpy %.12g)' % (n, E0, Er, W0, Wr))
            assert abs(E0 - Er) <= 1e-12 * abs(Er) and abs(W0 - Wr) <= 1e-12 * abs(Wr)
        h.visc_steps(n, 1)
        E1, _ = h.energy_sum(tau(n + 1), 1)
        _, Ws = h.energy_sum(1.0, 2)
        res.append((E1 - E0 + 0.5 * p.dt * (W0 + Ws)) / E0)
    res = np.array(res)
    _say('  one-step balance residuals: max %.3g' % np.abs(res).max())
    if nz == 1:
        assert np.abs(res).max() < 1e-6
    # twice the same call: the same bits (fixed reduction order)
    assert h.energy_sum(tau(12), 1) == h.energy_sum(tau(12), 1)
    h.close()
