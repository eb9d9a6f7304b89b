"""CLIdeal: ideal (3+1)D hydro driver with the interface of the reference's
pyvisc/ideal.py:28-349 (same constructor, attributes and methods), running on the
HIP library through the C-ABI of include/clvisc_amd.h.

Device buffers are exposed as :class:`DeviceField` proxies (``d_ev[k]``, ``d_Src``)
with ``get()`` / ``set()`` in the reference's AoS layouts; the OpenCL ``ctx`` /
``queue`` of the reference do not exist here.
"""
from __future__ import annotations

import os
import time

import numpy as np

from . import _lib, chemical_potential
from .eos import Eos
from .hypersf import HyperSurface


class DeviceField(object):
    """Proxy of one device field of a handle (the reference's cl.Buffer d_ev[k] etc.)."""

    def __init__(self, handle, name):
        self.handle = handle
        self.name = name
        self.field_id = _lib.FIELDS[name]

    def get(self):
        """Copy to the host in the reference's AoS layout."""
        return self.handle.get(self.name)

    def set(self, arr):
        self.handle.set(self.name, arr)

    def gather(self, cells):
        """Rows at the given flat cell indices, without copying the whole field."""
        return self.handle.gather(self.name, cells)

    def __repr__(self):
        return 'DeviceField(%s)' % self.name


def _params_from_cfg(cfg, efrz, viscous):
    p = _lib.CvParams()
    p.nx, p.ny, p.nz = int(cfg.NX), int(cfg.NY), int(cfg.NZ)
    p.tau0, p.dt = float(cfg.TAU0), float(cfg.DT)
    p.dx, p.dy, p.dz = float(cfg.DX), float(cfg.DY), float(cfg.DZ)
    p.etaos_xmin, p.etaos_ymin = float(cfg.ETAOS_XMIN), float(cfg.ETAOS_YMIN)
    p.etaos_left_slop, p.etaos_right_slop = float(cfg.ETAOS_LEFT_SLOP), float(cfg.ETAOS_RIGHT_SLOP)
    p.lam1 = float(cfg.LAM1)
    flags = 0
    if getattr(cfg, 'riemann_test', False):
        flags |= _lib.CV_FLAG_RIEMANN_TEST
    if getattr(cfg, 'gubser_visc_test', False):
        flags |= _lib.CV_FLAG_GUBSER_VISC_TEST
    zmax = float(getattr(cfg, 'ZETAOS_MAX', 0.0))
    if viscous and zmax > 0.0:
        flags |= _lib.CV_FLAG_BULK
        p.zetaos_max, p.zetaos_tpeak, p.zetaos_width = (zmax, float(cfg.ZETAOS_TPEAK),
                                                        float(cfg.ZETAOS_WIDTH))
    vmin = float(getattr(cfg, 'FP64_ROOT_VMIN', 0.0))
    if vmin > 0.0:
        flags |= _lib.CV_FLAG_FP64_ROOT
        p.fp64_root_vmin = vmin
    if getattr(cfg, 'hypersf_fp64', False):
        flags |= _lib.CV_FLAG_SF_FP64
    # the IS vorticity couplings of update_pimn (pyvisc/visc.py:116-121, kernel_IS.cl:636-683)
    if viscous and getattr(cfg, 'pimn_omega_coupling', False):
        flags |= _lib.CV_FLAG_PIMN_OMEGA
    if viscous and getattr(cfg, 'omega_omega_coupling', False):
        flags |= _lib.CV_FLAG_OMEGA_OMEGA
    p.flags = flags
    p.efrz = float(efrz)
    p.nxskip, p.nyskip, p.nzskip = int(cfg.nxskip), int(cfg.nyskip), int(cfg.nzskip)
    return p


class CLIdeal(object):
    """Ideal hydro on the GPU (pyvisc/ideal.py:28)."""

    def __init__(self, configs, handcrafted_eos=None, gpu_id=0, _viscous=False):
        self.cfg = configs
        path = self.cfg.fPathOut
        if not os.path.exists(path):
            os.makedirs(path)
        if not getattr(self.cfg, 'use_float32', True):
            raise ValueError('use_float32=False: the device path computes in FP32 only')
        self.determine_float_size(self.cfg)
        self.gpu_id = gpu_id
        self.size = self.cfg.NX * self.cfg.NY * self.cfg.NZ
        self.tau = self.cfg.real(self.cfg.TAU0)
        self.compile_options = self.__compile_options()
        self.eos = Eos(self.cfg.eos_type) if handcrafted_eos is None else handcrafted_eos
        self.__chemical_potential(path, handcrafted_eos)
        self.eos_table = self.eos.create_table(None, self.compile_options)
        self.efrz = float(self.eos.f_ed(self.cfg.TFRZ))
        from .bulkinfo import BulkInfo
        self.bulkinfo = BulkInfo(self.cfg, self.eos)
        params = _params_from_cfg(self.cfg, self.efrz, _viscous)
        self.handle = _lib.Handle(params, self.eos_table, device=gpu_id, viscous=_viscous)
        kind = {'percell': 0, 'tiled': 1, 'march': 2}[getattr(self.cfg, 'stage_kernel', 'march')]
        self.handle.set_stage_kernel(kind, int(getattr(self.cfg, 'tile_ty', 0)), int(getattr(self.cfg, 'tile_tz', 0)))
        self.d_ev = [DeviceField(self.handle, 'ev%d' % k) for k in range(3)]
        self.d_Src = DeviceField(self.handle, 'src')
        self.h_ev1 = np.zeros((self.size, 4), self.cfg.real)
        self.hypersf = HyperSurface(self.handle, self.cfg, self.efrz, viscous=_viscous)
        self.history = []
        self.edmax = 0.0
        self.num_of_sf = 0

    def __chemical_potential(self, path, handcrafted_eos):
        """chemical_potential.dat in the output directory (pyvisc/ideal.py:65), when the s95p
        tables are at hand (cfg.eos_table_dir or $CLVISC_EOS_TABLE_DIR).  e_frz comes from the
        EOS named by cfg.eos_type, as in the reference, or from the handcrafted EOS when that
        type has no table here."""
        table_dir = getattr(self.cfg, 'eos_table_dir', None) or os.environ.get(chemical_potential.TABLE_DIR_ENV)
        if not table_dir:
            return None
        eos = self.eos
        if handcrafted_eos is not None:
            try:
                eos = Eos(self.cfg.eos_type)
            except ValueError:
                pass
        return chemical_potential.create_table(self.cfg.TFRZ, path, eos_type=self.cfg.eos_type, table_dir=table_dir,
                                               eos=eos, copy_to=getattr(self.cfg, 'chem_copy_to', ()))

    @staticmethod
    def determine_float_size(cfg):
        from .config import determine_float_size
        determine_float_size(cfg)

    def __compile_options(self):
        """The reference's -D option list (pyvisc/ideal.py:146-171), kept for callers
        that inspect it; the device reads the same values from cv_params."""
        opts = ['-D %s=%sf' % (k, getattr(self.cfg, k)) for k in
                ('TAU0', 'DT', 'DX', 'DY', 'DZ', 'ETAOS_XMIN', 'ETAOS_YMIN', 'ETAOS_LEFT_SLOP',
                 'ETAOS_RIGHT_SLOP', 'LAM1')]
        opts += ['-D NX=%d' % self.cfg.NX, '-D NY=%d' % self.cfg.NY, '-D NZ=%d' % self.cfg.NZ,
                 '-D SIZE=%d' % self.size, '-D BSZ=%d' % self.cfg.BSZ, '-D USE_SINGLE_PRECISION']
        if getattr(self.cfg, 'riemann_test', False):
            opts.append('-D RIEMANN_TEST')
        opts.append('-D EOS_TABLE')
        return opts

    def load_ini(self, dat):
        """Load (e, vx, vy, veta) of all cells into d_ev[1] (pyvisc/ideal.py:138-144)."""
        self.h_ev1 = np.asarray(dat).astype(self.cfg.real).reshape(self.size, 4)
        self.d_ev[1].set(self.h_ev1)

    def stepUpdate(self, step, jet_eloss_src=None):
        """One RK2 stage at self.tau (pyvisc/ideal.py:211-255)."""
        if not jet_eloss_src or not jet_eloss_src.get('switch_on', False):
            self.handle.ideal_stage(step, self.cfg.real(self.tau))
            return
        self.handle.set_jet_source(jet_eloss_src, stages=3)
        try:
            self.handle.ideal_stage(step, self.cfg.real(self.tau))
        finally:
            self.handle.set_jet_source(None)

    def max_energy_density(self):
        """max e over d_ev[1] (pyvisc/ideal.py:257-262)."""
        return self.handle.max_ed(1)

    def ev_to_host(self):
        self.h_ev1 = self.d_ev[1].get()

    def get_hypersf(self, n, ntskip):
        """Freeze-out surface between d_ev_old and d_ev[1] every ntskip steps
        (pyvisc/ideal.py:269-293)."""
        is_finished = self.edmax < self.efrz
        if n == 0:
            self.hypersf.snapshot(self.cfg.TAU0)
        elif (n % ntskip == 0) or is_finished:
            self.hypersf.find(self.tau)
        return is_finished

    def save(self, save_hypersf=True, save_bulk=False, viscous_on=False):
        """Write hypersf.dat / sf_txyz.dat and the bulk info (pyvisc/ideal.py:295-315)."""
        self.num_of_sf = self.hypersf.count()
        print('num of sf=', self.num_of_sf)
        if save_hypersf:
            self.hypersf.save(self.cfg.fPathOut, self.cfg.TFRZ)
        if save_bulk:
            self.bulkinfo.save(viscous_on=viscous_on)

    def update_time(self, loop):
        """tau = TAU0 + (loop+1) DT as float (pyvisc/ideal.py:317-320)."""
        self.tau = self.cfg.real(self.cfg.TAU0 + (loop + 1) * self.cfg.DT)

    def evolve(self, max_loops=2000, save_hypersf=True, save_bulk=True, plot_bulk=True, to_maxloop=False,
               jet_eloss_src=None):
        """Main loop (pyvisc/ideal.py:322-349): max e every step, surface every ntskip,
        bulk info every ntskip, then both RK2 stages."""
        # the reference passes jet_eloss_src per call: the source is on for this loop only
        self.handle.set_jet_source(jet_eloss_src)
        try:
            self._evolve_loop(max_loops, save_hypersf, save_bulk, plot_bulk, to_maxloop)
        finally:
            self.handle.set_jet_source(None)

    def _evolve_loop(self, max_loops, save_hypersf, save_bulk, plot_bulk, to_maxloop):
        for n in range(max_loops):
            t0 = time.time()
            self.edmax = self.max_energy_density()
            self.history.append([self.tau, self.edmax])
            print('tau=', self.tau, ' EdMax= ', self.edmax)
            is_finished = False
            if save_hypersf:
                is_finished = self.get_hypersf(n, self.cfg.ntskip)
            if is_finished and not to_maxloop:
                break
            if (save_bulk or plot_bulk) and n % self.cfg.ntskip == 0:
                self.bulkinfo.get(self.tau, self.d_ev[1], self.edmax)
            self.handle.ideal_steps(n, 1)
            self.update_time(loop=n)
            print('one step: {dtime}'.format(dtime=time.time() - t0))
        self.save(save_hypersf=save_hypersf, save_bulk=save_bulk)
