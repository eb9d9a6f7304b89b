"""Optical Glauber initial condition on the device (pyvisc/glauber.py,
pyvisc/kernel/kernel_glauber.cl).

``Glauber(cfg, d_ev1)`` fills the device field with Edmax e_T(x, y)/e_T(0, 0) H(eta_s) and
zero velocity, e_T = Hwn Nw + (1 - Hwn) Nb from the Woods-Saxon thicknesses of the two
nuclei at x -/+ b/2 (HIP kernel ``k_glauber_ini`` behind ``cv_glauber_ini``).  The
reference's positional form ``Glauber(cfg, ctx, queue, compile_options, d_ev1, save_nbc)``
is accepted; there is no OpenCL context, and ``-D BJORKEN_SCALING`` in compile_options sets
H(eta_s) = 1 as the reference's define does.
"""
from __future__ import annotations

import os

import numpy as np

from . import _lib

# centrality classes, participants and impact parameters (pyvisc/glauber.py:21-38)
_TABLES = {
    'Au+Au': (np.array([0, 5, 10, 15, 20, 30, 40, 50, 60, 70, 80, 90], float),
              np.array([393.0, 327.0, 278.0, 237.0, 202.0, 144.0, 99.0, 65.0, 44.0, 22.0, 11.0, 4.0]),
              np.array([0.0, 3.3, 4.7, 5.8, 6.7, 8.2, 9.4, 10.6, 11.6, 12.5, 13.4, 14.3])),
    'Pb+Pb': (np.array([0, 5, 10, 15, 20, 30, 40, 50, 60, 70, 80, 90], float),
              np.array([416.0, 357.0, 305.0, 261.0, 216.0, 159.0, 110.0, 70.0, 42.0, 23.0, 11.0, 4.0]),
              np.array([0.0, 3.54, 5.00, 6.13, 7.08, 8.67, 10.01, 11.19, 12.26, 13.24, 14.16, 15.07])),
}


def weight_mean_b(cent_low, cent_high, system='Au+Au'):
    """<b> of a centrality class weighted by b dN/db ~ b^2 and N_wounded(b)
    (pyvisc/glauber.py:14-58): int b b^2 Nw db / int b^2 Nw db over [b(c_low), b(c_high)]."""
    from scipy.integrate import quad
    from scipy.interpolate import interp1d
    if system not in _TABLES:
        raise ValueError('collision system must be Au+Au or Pb+Pb')
    cent, nwn, bimp = _TABLES[system]
    b_of_c = interp1d(cent, bimp, kind='cubic')
    nw_of_b = interp1d(bimp, nwn, kind='cubic')
    num = quad(lambda b: b * b * b * nw_of_b(b), b_of_c(cent_low), b_of_c(cent_high))[0]
    den = quad(lambda b: b * b * nw_of_b(b), b_of_c(cent_low), b_of_c(cent_high))[0]
    return num / den


class Glauber(object):
    """Optical Glauber IC into a device field (pyvisc/glauber.py:61-105)."""

    def __init__(self, cfg, *args, save_nbc=False, bjorken=None):
        if len(args) == 1:
            d_ev1 = args[0]
            opts = []
        elif len(args) >= 4:               # (ctx, queue, compile_options, d_ev1[, save_nbc])
            opts = list(args[2] or [])
            d_ev1 = args[3]
            if len(args) > 4:
                save_nbc = bool(args[4])
        else:
            raise TypeError('Glauber(cfg, d_ev1) or Glauber(cfg, ctx, queue, compile_options, d_ev1)')
        if bjorken is None:
            bjorken = any('BJORKEN_SCALING' in str(o) for o in opts)
        self.cfg = cfg
        self.handle = d_ev1.handle
        self.params = _lib.CvGlauber.from_cfg(cfg, bjorken)
        self.handle.glauber_ini(self.params, d_ev1.field_id)
        if save_nbc:
            self.save_nbinary(cfg)

    def nbinary(self, dx=0.3, dy=0.3, xlow=-9.75, ylow=-9.75, nx=66, ny=66):
        """Binary-collision density on a grid (num_of_binary_collisions)."""
        return self.handle.glauber_nbinary(self.params, xlow, ylow, dx, dy, nx, ny)

    def save_nbinary(self, cfg, dx=0.3, dy=0.3, xlow=-9.75, ylow=-9.75, nx=66, ny=66):
        """nbin.dat in cfg.fPathOut (pyvisc/glauber.py:95-105)."""
        np.savetxt(os.path.join(cfg.fPathOut, 'nbin.dat'), self.nbinary(dx, dy, xlow, ylow, nx, ny))
