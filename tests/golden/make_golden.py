#!/usr/bin/env python
"""Writes the committed golden fixtures of this directory.

hotspots_400_seed20240601.txt
    The 400 transverse hotspot positions (x, y in fm) of the 385x385x115 benchmark
    initial condition: Woods-Saxon R = 6.62 fm, a = 0.546 fm, SplitMix64 seed 20240601
    (clvisc_amd.ic.woods_saxon_positions).  Pins the generator: a change of the stream
    or of the rejection rule changes the positions.

bjorken_bulk_ode.txt
    0+1D Bjorken flow with bulk viscous pressure only (eta/s = 0), integrated in FP64
    with scipy's DOP853 (rtol 1e-11):
        de/dtau  = -(e + P + Pi) / tau
        dPi/dtau = -(Pi + zeta / tau) / tau_Pi - (2/3) Pi / tau
    zeta = (zeta/s)(T) s hbarc, 1/tau_Pi = T / (3 (zeta/s)(T) hbarc),
    (zeta/s)(T) = zmax / (1 + ((T - Tpeak)/width)^2), crossover EOS (clvisc_amd.eos).
    This is the continuum form of the bulk relaxation the hydro step integrates
    (D Pi = -(Pi + zeta theta)/tau_Pi - (2/3) Pi theta with theta = 1/tau); the
    reference has no bulk pressure, so the ODE is the only pin for it.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

HBARC = 0.19733
BULK_PARAMS = dict(tau0=0.6, e0=30.0, zmax=0.1, tpeak=0.18, width=0.03, tau_end=3.6)


def hotspots():
    from clvisc_amd import ic
    pos = ic.woods_saxon_positions(400, 20240601)
    np.savetxt(os.path.join(HERE, 'hotspots_400_seed20240601.txt'), pos, fmt='%.17g',
               header='x y [fm]: 400 Woods-Saxon (R=6.62, a=0.546) nucleon positions, SplitMix64 seed 20240601')


def zetaos(T, zmax, tpeak, width):
    d = (T - tpeak) / width
    return zmax / (1.0 + d * d)


def bjorken_bulk():
    from scipy.integrate import solve_ivp
    from clvisc_amd.eos import Eos
    eos = Eos('crossover')
    b = BULK_PARAMS

    def rhs(tau, y):
        e, Pi = y
        T = float(eos.f_T(e))
        s = float(eos.f_S(e))
        P = float(eos.f_P(e))
        zs = zetaos(T, b['zmax'], b['tpeak'], b['width'])
        zeta = zs * s * HBARC
        inv_tau_pi = T / (3.0 * max(zs, 1e-7) * HBARC)
        return [-(e + P + Pi) / tau, -(Pi + zeta / tau) * inv_tau_pi - 2.0 / 3.0 * Pi / tau]

    taus = np.round(np.arange(b['tau0'], b['tau_end'] + 1e-9, 0.1), 10)
    sol = solve_ivp(rhs, (b['tau0'], b['tau_end']), [b['e0'], 0.0], method='DOP853', rtol=1e-11,
                    atol=1e-14, t_eval=taus)
    assert sol.success
    hdr = ('tau e Pi: 0+1D Bjorken with bulk only, FP64 DOP853; ' +
           ' '.join('%s=%r' % kv for kv in sorted(b.items())) + ' ; crossover EOS, Pi(tau0)=0')
    np.savetxt(os.path.join(HERE, 'bjorken_bulk_ode.txt'), np.column_stack([sol.t, sol.y[0], sol.y[1]]),
               fmt='%.15e', header=hdr)


if __name__ == '__main__':
    hotspots()
    bjorken_bulk()
