"""The jet energy-loss source of the oracle (orc_jet_src, kernel_jet_eloss.cl:8-69) pinned by
closed forms: on a uniform ideal-gas field (T(e) known) the deposit integrates over the grid to
(tau, tau, tau, tau^2) (1/tau) dE/dx (1, cos phi, sin phi, 0) dV^-1 with dE/dx = 14 (T/0.3)^3 (the
normalised Gaussian, sigma = 0.3 fm, resolved by dx = 0.06 fm); its peak sits at the start cell
moved by (tau - tau0)(cos phi, sin phi, 0); a jet that has left the grid deposits nothing."""
import numpy as np
import pytest

from clvisc_amd.eos import Eos
from oracle import oracle as O

NX, NY, NZ, DX = 81, 81, 41, 0.06
TAU0 = 0.6


def _params(direction, start, precision='f64'):
    tab = Eos('ideal_gas').create_table()
    return O.make_params(NX, NY, NZ, 0.02, DX, DX, DX, TAU0, tab, jet=(direction, start)), tab


def _start(i, j, k):
    return (i * NY + j) * NZ + k


@pytest.mark.parametrize('phi', [0.0, 0.7, 2.5])
def test_jet_deposit_integral_and_peak(phi):
    e0 = 30.0
    eos = Eos('ideal_gas')
    T = float(eos.f_T(e0))
    p, _ = _params(np.float32(phi), _start(40, 40, 20))
    ev = np.zeros((NX * NY * NZ, 4))
    ev[:, 0] = e0
    tau = 1.1
    s = O.jet_src(p, np.zeros_like(ev), ev, tau, 'f64')
    dEdx = 14.0 * (T / 0.3) ** 3
    tot = s.sum(axis=0) * DX ** 3
    c, sn = np.cos(np.float32(phi)), np.sin(np.float32(phi))
    # 1e-4: T comes from the FP32 EOS table (1.3e-5 from the closed form, cubed)
    np.testing.assert_allclose(tot[0], dEdx, rtol=1e-4)
    np.testing.assert_allclose(tot[1], dEdx * c, rtol=1e-4, atol=1e-4 * dEdx)
    np.testing.assert_allclose(tot[2], dEdx * sn, rtol=1e-4, atol=1e-4 * dEdx)
    assert tot[3] == 0.0
    pk = np.unravel_index(np.argmax(s[:, 0]), (NX, NY, NZ))
    want = (40 + (tau - TAU0) * c / DX, 40 + (tau - TAU0) * sn / DX, 20)
    assert max(abs(pk[0] - want[0]), abs(pk[1] - want[1])) <= 0.5 + 1e-6 and pk[2] == 20


def test_jet_off_grid_deposits_nothing():
    p, _ = _params(np.float32(0.0), _start(80, 40, 20))   # at the +x edge, moving outwards
    ev = np.zeros((NX * NY * NZ, 4))
    ev[:, 0] = 10.0
    s = O.jet_src(p, np.zeros_like(ev), ev, TAU0 + 0.5, 'f64')
    assert not s.any()


def test_jet_f32_close_to_f64():
    p, _ = _params(np.float32(0.7), _start(40, 40, 20))
    ev = np.zeros((NX * NY * NZ, 4))
    ev[:, 0] = 30.0
    s64 = O.jet_src(p, np.zeros_like(ev), ev, 1.1, 'f64')
    s32 = O.jet_src(p, np.zeros_like(ev, np.float32), ev, 1.1, 'f32')
    assert np.abs(s32 - s64).max() < 1e-5 * np.abs(s64).max()
