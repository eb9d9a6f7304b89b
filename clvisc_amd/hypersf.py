"""Freeze-out hypersurface on the device and its writers.

The search itself is the HIP kernel behind ``cv_hypersf_find`` (the reference's
``get_hypersf`` / ``visc_hypersf``, pyvisc/kernel/kernel_hypersf.cl:405-660); this
module keeps the driver-side bookkeeping of pyvisc/ideal.py:269-315 and
pyvisc/visc.py:359-419 and writes the reference's text files:

* ``hypersf.dat``  rows dS0 dS1 dS2 dS3 vx vy veta etas (pyvisc/ideal.py:301-306)
* ``sf_txyz.dat``  rows t x y z (pyvisc/ideal.py:308-312)
* ``pimnsf.dat``   rows pi00 01 02 03 11 12 13 22 23 33 (pyvisc/visc.py:400-419)
* ``bulksf.dat``   bulk pressure at each element (only when it is evolved; no
  counterpart in the reference)

The element order is not defined (atomic append, as the reference's atomic_inc).
"""
from __future__ import annotations

import os

import numpy as np


class HyperSurface(object):
    """Surface bookkeeping of one hydro handle."""

    def __init__(self, handle, cfg, efrz, viscous=False, capacity=None):
        self.handle = handle
        self.cfg = cfg
        self.efrz = float(efrz)
        self.viscous = viscous
        self.tau_old = None
        if capacity is not None:
            handle.hypersf_reserve(capacity)

    def snapshot(self, tau):
        """d_ev_old <- d_ev[1] (and d_pi_old <- d_pi[1]); tau_old <- tau
        (the n == 0 branch of get_hypersf, pyvisc/ideal.py:275-278)."""
        self.handle.hypersf_snapshot(np.float32(tau))
        self.tau_old = np.float32(tau)

    def find(self, tau_new):
        """Elements between the snapshot and d_ev[1] at tau_new, then a new snapshot
        (pyvisc/ideal.py:279-291)."""
        if self.tau_old is None:
            raise RuntimeError('HyperSurface.find before snapshot')
        self.handle.hypersf_find(np.float32(tau_new))
        self.tau_old = np.float32(tau_new)

    def count(self):
        """Number of elements found (d_num_of_sf)."""
        return self.handle.hypersf_count()[0]

    def get(self):
        """(sf (n,8), txyz (n,4), pi (n,10) or None, bulk (n,) or None)."""
        return self.handle.hypersf_get()

    def save(self, path, tfrz):
        """hypersf.dat and sf_txyz.dat (pyvisc/ideal.py:301-312)."""
        sf, txyz, _, _ = self.get()
        write_hypersf(path, sf, txyz, tfrz)

    def save_pimn(self, path, tfrz, pr, set_to_zero=False):
        """pimnsf.dat (pyvisc/visc.py:400-419) and, when the bulk pressure is
        evolved, bulksf.dat."""
        _, _, pisf, bksf = self.get()
        if pisf is None:
            raise RuntimeError('pi on the surface needs a viscous handle')
        write_pimnsf(path, pisf, self.efrz, tfrz, pr, set_to_zero)
        if bksf is not None:
            np.savetxt(os.path.join(path, 'bulksf.dat'), np.zeros_like(bksf) if set_to_zero else bksf,
                       fmt='%.6e', header='bulk pressure Pi (GeV/fm^3) at the centroid of each element')


def write_hypersf(path, sf, txyz, tfrz):
    """hypersf.dat / sf_txyz.dat as pyvisc/ideal.py:301-312 writes them."""
    np.savetxt(os.path.join(path, 'hypersf.dat'), np.asarray(sf, np.float32).reshape(-1, 8), fmt='%.6e',
               header='Tfrz=%.6e ; other rows: dS0, dS1, dS2, dS3, vx, vy, veta, etas' % tfrz)
    np.savetxt(os.path.join(path, 'sf_txyz.dat'), np.asarray(txyz, np.float32).reshape(-1, 4), fmt='%.6e',
               header='(t, x, y, z) the time-space coordinates of hypersf elements')


def write_pimnsf(path, pisf, efrz, tfrz, pr, set_to_zero=False):
    """pimnsf.dat as pyvisc/visc.py:400-419 writes it: the header carries
    1/(2 T^2 (e + P)) at freeze-out, used by the sampler's delta-f."""
    const_for_deltaf = 1.0 / (2.0 * tfrz ** 2 * (efrz + pr))
    pis = np.asarray(pisf, np.float32).reshape(-1, 10)
    if set_to_zero:
        pis = np.zeros_like(pis)
    comment = 'one_o_2TsqrEplusP=%.6e ' % const_for_deltaf + 'pi00 01 02 03 11 12 13 22 23 33'
    np.savetxt(os.path.join(path, 'pimnsf.dat'), pis, fmt='%.6e', header=comment)
