"""Lambda polarization from the freeze-out surface on the device
(pyvisc/polarization/polarization.py, pyvisc/kernel/kernel_polarization.cl).

``Polarization(fpath, T, Mu, mass, ...)`` reads hypersf.dat and the omegamu_*_sf.dat files
that CLVisc.evolve(save_vorticity=True) writes; ``set_momentum(list of (m_T, Y, p_x, p_y))``;
``get_themal()`` / ``get_shear()`` / ``get_accT()`` return (polarization in the lab frame
(n, 4), density (n,), polarization in the Lambda rest frame (n, 4)) as the reference's
methods do.  The sums over the surface run in the HIP kernel behind ``cv_polarization``.
"""
from __future__ import annotations

import os

import numpy as np

from . import _lib


class Polarization(object):
    """Lambda polarization on the freeze-out surface (polarization.py:25-200)."""

    def __init__(self, fpath, T=0.165, Mu=0.0, mass=1.116, path='./', Baryon_on=False, gpu_id=0, themal=True,
                 shear=False, accT=False):
        if Baryon_on:
            raise NotImplementedError('the baryon chemical potential branch is empty in the reference too')
        self.Tfrz, self.Mu, self.mass = float(T), float(Mu), float(mass)
        self.gpu_id = int(gpu_id)
        self.themal, self.shear, self.accT = themal, shear, accT
        load = lambda name: np.loadtxt(os.path.join(fpath, name), dtype=np.float32, ndmin=2)
        self.sf = load('hypersf.dat')
        self.size_sf = self.sf.shape[0]
        self.omega = load('omegamu_sf.dat') if themal else None
        self.omega_shear1 = load('omegamu_shear1_sf.dat') if shear else None
        self.omega_shear2 = load('omegamu_shear2_sf.dat') if shear else None
        self.omega_accT = load('omegamu_accT_sf.dat') if accT else None
        self.momentum = None

    @classmethod
    def from_arrays(cls, sf, omega=None, omega_shear1=None, omega_shear2=None, omega_accT=None, T=0.165,
                    mass=1.116, gpu_id=0):
        """The same object from arrays already in memory (the files' rows)."""
        self = cls.__new__(cls)
        self.Tfrz, self.Mu, self.mass, self.gpu_id = float(T), 0.0, float(mass), int(gpu_id)
        self.sf = np.ascontiguousarray(sf, np.float32).reshape(-1, 8)
        self.size_sf = self.sf.shape[0]
        self.omega, self.omega_shear1, self.omega_shear2, self.omega_accT = omega, omega_shear1, omega_shear2, omega_accT
        self.themal, self.shear, self.accT = omega is not None, omega_shear1 is not None, omega_accT is not None
        self.momentum = None
        return self

    def set_momentum(self, momentum_list):
        self.momentum = np.ascontiguousarray(momentum_list, np.float32).reshape(-1, 4)
        self.size = self.momentum.shape[0]

    def _run(self, kind, om1, om2=None):
        if self.momentum is None:
            raise RuntimeError('set_momentum first')
        L = _lib.lib()
        m = self.size
        pol, rho, lrf = np.zeros((m, 4), np.float32), np.zeros(m, np.float32), np.zeros((m, 4), np.float32)
        a = np.ascontiguousarray(om1, np.float32)
        b = None if om2 is None else np.ascontiguousarray(om2, np.float32)
        _lib.check(L.cv_polarization(self.gpu_id, kind, self.sf.ctypes.data, a.ctypes.data,
                                     None if b is None else b.ctypes.data, self.size_sf, self.momentum.ctypes.data,
                                     m, self.Tfrz, self.mass, pol.ctypes.data, rho.ctypes.data, lrf.ctypes.data))
        return pol, rho, lrf

    def get_themal(self):
        return self._run(0, self.omega)

    def get_shear(self):
        return self._run(1, self.omega_shear1, self.omega_shear2)

    def get_accT(self):
        return self._run(2, self.omega_accT)
