"""PimnInfo: pi^{mu nu} along x through the grid centre, with the attributes and
files of the reference's pyvisc/pimninfo.py:15-95."""
from __future__ import annotations

import numpy as np


class PimnInfo(object):
    """pi^xx, pi^yy, pi^zz (tilde-scaled) along x at (NY/2, NZ/2) (pimninfo.py:15)."""

    def __init__(self, cfg):
        self.cfg = cfg
        NX, NY, NZ = cfg.NX, cfg.NY, cfg.NZ
        self.pixx_x, self.piyy_x, self.pizz_x = [], [], []
        self.time = []
        self.x = np.linspace(-NX / 2 * cfg.DX, NX / 2 * cfg.DX, NX)
        self.y = np.linspace(-NY / 2 * cfg.DY, NY / 2 * cfg.DY, NY)
        self.z = np.linspace(-NZ / 2 * cfg.DZ, NZ / 2 * cfg.DZ, NZ)

    def get(self, tau, d_pi):
        """Record the x line of d_pi at tau (pimninfo.py:69-88)."""
        self.time.append(tau)
        NX, NY, NZ = self.cfg.NX, self.cfg.NY, self.cfg.NZ
        cells = np.arange(NX, dtype=np.int64) * NY * NZ + (NY // 2) * NZ + NZ // 2
        pimn_x = d_pi.gather(cells).reshape(NX, 10)
        self.pixx_x.append(pimn_x[:, 4])
        self.piyy_x.append(pimn_x[:, 7])
        self.pizz_x.append(pimn_x[:, 9])

    def save(self):
        np.savetxt(self.cfg.fPathOut + '/pizz_x.dat', np.array(self.pizz_x).T)
        np.savetxt(self.cfg.fPathOut + '/pixx_x.dat', np.array(self.pixx_x).T)
        np.savetxt(self.cfg.fPathOut + '/piyy_x.dat', np.array(self.piyy_x).T)
