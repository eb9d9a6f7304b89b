"""CLVisc: (3+1)D viscous hydro driver (KT + Israel-Stewart, optional bulk pressure)
with the interface of the reference's pyvisc/visc.py:22-560, running on the HIP
library through the C-ABI of include/clvisc_amd.h.

The stage-level methods (IS_stepUpdate, visc_stepUpdate, update_udiff) call one
stage entry point each, in the reference's order.  ``evolve`` runs the same loop as
pyvisc/visc.py:473-560 but hands every run of steps between two host events (max e,
surface, bulk info: all on loops divisible by ntskip) to the fused step path
(cv_visc_steps_async), which gives the same fields as the stage-level sequence.
"""
from __future__ import annotations

import os
from time import time

import numpy as np

from .ideal import CLIdeal, DeviceField


class CLVisc(object):
    """Viscous hydro on the GPU (pyvisc/visc.py:22)."""

    def __init__(self, configs, handcrafted_eos=None, gpu_id=0):
        self.ideal = CLIdeal(configs, handcrafted_eos, gpu_id, _viscous=True)
        self.cfg = configs
        self.handle = self.ideal.handle
        self.compile_options = list(self.ideal.compile_options)
        if getattr(self.cfg, 'gubser_visc_test', False):
            self.compile_options.append('-D GUBSER_VISC_TEST')
        self.eos_table = self.ideal.eos_table
        self.size = self.ideal.size
        self.d_pi = [DeviceField(self.handle, 'pi%d' % k) for k in range(3)]
        self.d_IS_src = DeviceField(self.handle, 'is_src')
        self.d_udiff = DeviceField(self.handle, 'udiff')
        self.d_goodcell = DeviceField(self.handle, 'goodcell')
        self.bulk_on = bool(self.handle.params.flags & 0x4)
        self.d_bulk = [DeviceField(self.handle, 'bulk%d' % k) for k in range(3)] if self.bulk_on else None
        self.h_goodcell = np.ones(self.size, self.cfg.real)
        # pi(tau0) and the predicted u for the first step are set before any IC is
        # loaded (pyvisc/visc.py:101-102)
        self.IS_initialize()
        self.ideal_predict_for_first_step()

    # -- stage-level operations (reference order) --------------------------------
    def IS_initialize(self):
        """visc_initialize on d_ev[1] at tau (pyvisc/visc.py:143-149)."""
        self.handle.visc_initialize(self.ideal.tau)

    def ideal_predict_for_first_step(self):
        """Ideal stage 1 for the first u prediction (pyvisc/visc.py:468-470)."""
        self.ideal.stepUpdate(step=1)

    def IS_stepUpdate(self, step):
        """IS source, KT flux of u pi, update of pi^{mu nu} into d_pi[3-step]
        (pyvisc/visc.py:232-273)."""
        self.handle.is_stage(step, self.cfg.real(self.ideal.tau))

    def visc_stepUpdate(self, step, jet_eloss_src=None):
        """KT source with pi and the primitive update into d_ev[3-step]
        (pyvisc/visc.py:276-314)."""
        if not jet_eloss_src or not jet_eloss_src.get('switch_on', False):
            self.handle.visc_stage(step, self.cfg.real(self.ideal.tau))
            return
        self.handle.set_jet_source(jet_eloss_src, stages=3)
        try:
            self.handle.visc_stage(step, self.cfg.real(self.ideal.tau))
        finally:
            self.handle.set_jet_source(None)

    def update_udiff(self, d_ev0, d_ev1):
        """udiff = u(d_ev1) - u(d_ev0) (pyvisc/visc.py:352-357)."""
        self.handle.update_udiff(d_ev0.name, d_ev1.name)

    def update_time(self, loop):
        self.ideal.update_time(loop)

    # -- freeze-out surface --------------------------------------------------------
    def get_hypersf(self, n, ntskip, is_finished):
        """visc_hypersf between the previous surface time and now (pyvisc/visc.py:359-397)."""
        if n == 0:
            self.ideal.hypersf.snapshot(self.cfg.TAU0)
        elif (n % ntskip == 0) or is_finished:
            self.ideal.hypersf.find(self.ideal.tau)

    def save_pimn_sf(self, set_to_zero=False):
        """pimnsf.dat (pyvisc/visc.py:400-419)."""
        ed = self.ideal.efrz
        pr = float(self.ideal.eos.f_P(ed))
        print('pimn on frzsf is saved to ', os.path.join(self.cfg.fPathOut, 'pimnsf.dat'))
        self.ideal.hypersf.save_pimn(self.cfg.fPathOut, self.cfg.TFRZ, pr, set_to_zero=set_to_zero)

    def save(self, save_hypersf=True, save_bulk=False, save_pi=False, save_vorticity=False):
        """pyvisc/visc.py:423-462: surface, pi on the surface, the vorticity files."""
        self.ideal.save(save_hypersf, save_bulk, viscous_on=True)
        if save_hypersf:
            self.save_pimn_sf()
        if save_vorticity:
            om = self.handle.hypersf_get_vorticity()
            for name, a, b in (('omegamu_sf.dat', 0, 6), ('omegamu_shear1_sf.dat', 6, 22),
                               ('omegamu_shear2_sf.dat', 22, 26), ('omegamu_accT_sf.dat', 26, 32)):
                out_path = os.path.join(self.cfg.fPathOut, name)
                print('vorticity omega_{mu, nu} on surface is saved to', out_path)
                np.savetxt(out_path, om[:, a:b], fmt='%.6e', header='omega^{01, 02, 03, 12, 13, 23}')
        if save_pi and hasattr(self, 'pimn_info'):
            self.pimn_info.save()

    def check_pizz(self):
        """Print <pi^{eta eta}/(e + P)> of the mid-rapidity slice over the cells with
        e + P > 0.133 (pyvisc/visc.py:200-208, kernel_checkpi.cl) and return the per-slice
        values."""
        pizz_o_ep = self.handle.pizz_o_ep(1)
        print('<pizz/(e+P)> =', pizz_o_ep[self.cfg.NZ // 2])
        return pizz_o_ep

    def check_bad_cell(self):
        """Mid-rapidity masks of e > efrz and good cells (pyvisc/visc.py:282-300)."""
        NX, NY, NZ = self.cfg.NX, self.cfg.NY, self.cfg.NZ
        self.h_goodcell = self.d_goodcell.get()
        efrz = self.ideal.efrz
        self.ideal.ev_to_host()
        ed_mid = self.ideal.h_ev1[:, 0].reshape(NX, NY, NZ)[:, :, NZ // 2]
        frz_sf = (ed_mid > efrz).astype(float) * 4.0
        good_cells = 3 * (self.h_goodcell.reshape(NX, NY, NZ)[:, :, NZ // 2] - 1)
        comments = '''4 for e>efrz and good cell;
                    1 for e > efrz and bad cell;
                    0 for e < efrz and good cell;
                    -3 for e<efrz and bad cell'''
        outpath = os.path.join(self.cfg.fPathOut, 'bad_cell_masks_tau%s.txt' % self.ideal.tau)
        np.savetxt(outpath, frz_sf + good_cells, header=comments)

    # -- initial conditions ------------------------------------------------------------
    def optical_glauber_ini(self, system='Au+Au', cent_min=None, cent_max=None, save_binary_collisions=False):
        """Optical Glauber IC on the device into d_ev[1] (pyvisc/visc.py:151-171); with a
        centrality range the impact parameter becomes weight_mean_b of that class and is
        written into the configuration, as there."""
        from .glauber import Glauber, weight_mean_b
        if cent_min is not None and cent_max is not None:
            self.cfg.ImpactParameter = weight_mean_b(cent_min, cent_max, system)
        glauber = Glauber(self.cfg, None, None, self.compile_options, self.ideal.d_ev[1])
        if save_binary_collisions:
            glauber.save_nbinary(self.cfg)
        return glauber

    def create_ini_from_partons(self, fname, SIGR=0.6, SIGZ=0.6, KFACTOR=1.0):
        """Gaussian-smeared partons read from a text file (one header line, rows E px py pz
        t x y z) into d_ev[1] (pyvisc/visc.py:176-183, Smearing pyvisc/smearing.py:23-47)."""
        dat = np.loadtxt(fname, skiprows=1, dtype=np.float32, ndmin=2)
        self.handle.smear_partons(dat, SIGR, SIGZ, KFACTOR, which=1)

    def smear_from_p4x4(self, p4x4, SIGR=0.6, SIGZ=0.6, KFACTOR=1.0, force_bjorken=False,
                        longitudinal_profile=None):
        """Gaussian-smeared partons p4x4 (n x 8: E px py pz t x y z) into d_ev[1], then the
        optional longitudinal profile (pyvisc/visc.py:185-195, SmearingP4X4
        pyvisc/smearing.py:63-107): force_bjorken copies mid-rapidity to every eta_s."""
        self.handle.smear_partons(p4x4, SIGR, SIGZ, KFACTOR, which=1)
        if force_bjorken:
            self.handle.longitudinal_profile(np.ones(self.cfg.NZ, np.float32), which=1)
        elif longitudinal_profile is not None:
            self.handle.longitudinal_profile(longitudinal_profile, which=1)

    def get_vorticity(self, save_data=False):
        """Thermal, shear and acceleration-temperature vorticity tensors of d_ev[0] -> d_ev[1]
        into d_omega[1] on the device (pyvisc/visc.py:316-350); with save_data the vector
        omega^mu = Omega^{mu nu} u_nu (first column T) goes to omega_mu_tau<tau>.dat."""
        self.handle.vorticity_enable()
        self.handle.vorticity(self.ideal.tau)
        if save_data:
            omega_mu = self.handle.vorticity_mu(self.ideal.efrz)
            fname = ('omega_mu_tau%s' % self.ideal.tau).replace('.', 'p') + '.dat'
            np.savetxt(os.path.join(os.path.abspath(self.cfg.fPathOut), fname), omega_mu,
                       header='omega^{t, x, y, z} = Omega^{mu nu} u_{nu}')

    # -- main loop -----------------------------------------------------------------
    def evolve(self, max_loops=1000, save_hypersf=True, save_bulk=False, plot_bulk=False, save_pi=True,
               force_run_to_maxloop=False, save_vorticity=False, jet_eloss_src=None, debug=False):
        """The loop of pyvisc/visc.py:473-560.  Without force_run_to_maxloop it runs
        until max e < e_frz, as the reference does (max_loops is not consulted)."""
        if save_vorticity:
            self.handle.vorticity_enable()
        cfg = self.cfg
        # the ideal shortcut of pyvisc/visc.py:483-489 (not taken with save_vorticity, whose
        # tensors the viscous loop computes; nor with bulk viscosity, which is viscous)
        if cfg.ETAOS_YMIN < 1.0E-6 and cfg.ETAOS_LEFT_SLOP < 1.0E-6 and cfg.ETAOS_RIGHT_SLOP < 1.0E-6 \
                and not save_vorticity and not self.bulk_on:
            self.ideal.evolve(max_loops, save_hypersf, save_bulk, plot_bulk, force_run_to_maxloop,
                              jet_eloss_src)
            self.save_pimn_sf(set_to_zero=True)
            return
        # the reference passes jet_eloss_src per call: the source is on for this loop only
        self.handle.set_jet_source(jet_eloss_src)
        try:
            self._evolve_loop(max_loops, save_hypersf, save_bulk, plot_bulk, save_pi, force_run_to_maxloop,
                              save_vorticity, debug)
        finally:
            self.handle.set_jet_source(None)

    def _evolve_loop(self, max_loops, save_hypersf, save_bulk, plot_bulk, save_pi, force_run_to_maxloop,
                     save_vorticity, debug):
        cfg = self.cfg
        if save_pi:
            from .pimninfo import PimnInfo
            self.pimn_info = PimnInfo(cfg)
        ntskip = int(cfg.ntskip)
        loop = 0
        while True:
            t0 = time()
            if force_run_to_maxloop and loop == max_loops:
                break
            if loop % ntskip == 0:
                self.ideal.edmax = self.ideal.max_energy_density()
                self.ideal.history.append([self.ideal.tau, self.ideal.edmax])
                print('tau=', self.ideal.tau, ' EdMax= ', self.ideal.edmax)
            is_finished = self.ideal.edmax < self.ideal.efrz
            if save_hypersf:
                self.get_hypersf(loop, ntskip, is_finished)
            if (plot_bulk or save_bulk) and loop % ntskip == 0:
                self.ideal.bulkinfo.get(self.ideal.tau, self.ideal.d_ev[1], self.ideal.edmax, self.d_pi[1])
            if save_pi and loop % ntskip == 0:
                self.pimn_info.get(self.ideal.tau, self.d_pi[1])
            if is_finished and not force_run_to_maxloop:
                break
            # every loop up to the next multiple of ntskip has no host event (edmax is
            # unchanged, so is is_finished); after is_finished the surface is taken
            # every loop
            nrun = 1 if (is_finished or (debug and loop % ntskip == 0)) else ntskip - loop % ntskip
            if save_vorticity and loop % ntskip == 0:
                nrun = 1            # the vorticity is taken after this step (visc.py:545-547)
            if force_run_to_maxloop:
                nrun = min(nrun, max_loops - loop)
            self.handle.visc_steps(loop, nrun)
            self.update_time(loop + nrun - 1)
            if save_vorticity and loop % ntskip == 0:
                self.get_vorticity(save_data=False)
            if debug and loop % ntskip == 0:
                self.check_bad_cell()
            loop += nrun
            print('one step: {dtime}'.format(dtime=(time() - t0) / nrun))
        self.save(save_hypersf=save_hypersf, save_bulk=save_bulk, save_pi=save_pi, save_vorticity=save_vorticity)
