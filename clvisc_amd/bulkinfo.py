"""BulkInfo: slices and bulk observables along the evolution, with the attributes and
output files of the reference's pyvisc/bulkinfo.py:26-340.

The reference copies the whole field to the host at every call and evaluates the
eccentricities of every eta slice in NumPy (pyvisc/bulkinfo.py:129-260).  Here the slices
are gathered on the device (``DeviceField.gather``: the three lines and two planes, not the
field) and the per-slice sums behind eccp, <v_r> and the entropy come from one device
reduction (``cv_slice_sums``); ``device_reductions=False`` keeps the reference's host path
(whole-field copy, NumPy formulas) for comparison.
"""
from __future__ import annotations

import os
from math import floor

import numpy as np


class BulkInfo(object):
    """ed(x), ed(y), ed(eta), T, v slices, eccentricities vs rapidity (bulkinfo.py:26)."""

    def __init__(self, cfg, eos, device_reductions=True):
        self.cfg = cfg
        self.eos = eos
        self.device_reductions = device_reductions
        self.slice_sums = []          # per call: (NZ, 10) device sums (cv_slice_sums)
        self.energy, self.entropy = [], []   # get_total_energy_and_entropy_on_gpu
        NX, NY, NZ = cfg.NX, cfg.NY, cfg.NZ
        self.ex, self.ey, self.ez = [], [], []
        self.vx, self.vy, self.vz = [], [], []
        self.exy, self.vx_xy, self.vy_xy, self.vz_xy = [], [], [], []
        self.pixx_xy, self.piyy_xy, self.pitx_xy = [], [], []
        self.exz, self.vx_xz, self.vy_xz, self.vz_xz = [], [], [], []
        self.ecc2_vs_rapidity = []
        self.ecc1_vs_rapidity = []
        self.time = []
        self.edmax = []
        self.x = np.linspace(-floor(NX / 2) * cfg.DX, floor(NX / 2) * cfg.DX, NX, endpoint=True)
        self.y = np.linspace(-floor(NY / 2) * cfg.DY, floor(NY / 2) * cfg.DY, NY, endpoint=True)
        self.z = np.linspace(-floor(NZ / 2) * cfg.DZ, floor(NZ / 2) * cfg.DZ, NZ, endpoint=True)

    def get(self, tau, d_ev1, edmax, d_pi=None):
        """Record the slices of d_ev1 (and pi^xx, pi^yy, pi^tx in the mid-eta plane)
        at tau (bulkinfo.py:129-200)."""
        self.time.append(tau)
        self.edmax.append(edmax)
        NX, NY, NZ = self.cfg.NX, self.cfg.NY, self.cfg.NZ
        i0, j0, k0 = NX // 2, NY // 2, NZ // 2
        if self.device_reductions and getattr(d_ev1, 'field_id', None) in (0, 1, 2):
            self._get_device(d_ev1, i0, j0, k0)
        else:
            bulk = d_ev1.get().reshape(NX, NY, NZ, 4)
            self._ecc_vs_rapidity(bulk)
            self._slices(bulk[:, j0, k0], bulk[i0, :, k0], bulk[i0, j0, :], bulk[:, :, k0], bulk[:, j0, :])
        if d_pi is not None:
            self._pimn(d_pi)

    def _get_device(self, d_ev1, i0, j0, k0):
        """Lines and planes gathered on the device; eccentricities from cv_slice_sums."""
        NX, NY, NZ = self.cfg.NX, self.cfg.NY, self.cfg.NZ
        ii, jj, kk = np.arange(NX), np.arange(NY), np.arange(NZ)
        idx = lambda i, j, k: (np.asarray(i) * NY + np.asarray(j)) * NZ + np.asarray(k)
        gx, gy = np.meshgrid(ii, jj, indexing='ij')
        hx, hz = np.meshgrid(ii, kk, indexing='ij')
        cells = [idx(ii, j0, k0), idx(i0, jj, k0), idx(i0, j0, kk), idx(gx, gy, k0).reshape(-1),
                 idx(hx, j0, hz).reshape(-1)]
        rows = d_ev1.gather(np.concatenate(cells))
        parts = np.split(rows, np.cumsum([c.size for c in cells])[:-1])
        self._slices(parts[0], parts[1], parts[2], parts[3].reshape(NX, NY, 4), parts[4].reshape(NX, NZ, 4))
        s = d_ev1.handle.slice_sums(d_ev1.field_id)
        self.slice_sums.append(s)
        self.ecc1_vs_rapidity.append(s[:, 2] / s[:, 1])
        self.ecc2_vs_rapidity.append(s[:, 0] / s[:, 1])

    def _slices(self, evx0, evy0, evz0, evxy, evxz):
        self.ex.append(evx0[:, 0].copy())
        self.ey.append(evy0[:, 0].copy())
        self.ez.append(evz0[:, 0].copy())
        self.vx.append(evx0[:, 1].copy())
        self.vy.append(evy0[:, 2].copy())
        self.vz.append(evz0[:, 3].copy())
        self.exy.append(evxy[..., 0].copy())
        self.vx_xy.append(evxy[..., 1].copy())
        self.vy_xy.append(evxy[..., 2].copy())
        self.exz.append(evxz[..., 0].copy())
        self.vx_xz.append(evxz[..., 1].copy())
        self.vy_xz.append(evxz[..., 2].copy())
        self.vz_xz.append(evxz[..., 3].copy())

    def _pimn(self, d_pi):
        """pi^xx, pi^yy, pi^tx of the get_pimn kernel's cells (bulkinfo.py:105-121)."""
        NX, NY, NZ = self.cfg.NX, self.cfg.NY, self.cfg.NZ
        # the reference reads cell gid_x*NY*(NZ/2) + gid_y*(NZ/2) + NZ/2 (bulkinfo.py:112),
        # which is the mid-eta plane only for NZ == 1; kept as written
        gx, gy = np.meshgrid(np.arange(NX), np.arange(NY), indexing='ij')
        oid = (gx * NY * (NZ // 2) + gy * (NZ // 2) + NZ // 2).reshape(-1)
        pi = d_pi.gather(oid)
        self.pixx_xy.append(pi[:, 4].reshape(NX, NY))
        self.piyy_xy.append(pi[:, 7].reshape(NX, NY))
        self.pitx_xy.append(pi[:, 1].reshape(NX, NY))

    def eccp(self, ed, vx, vy, vz=0.0, pixx=None, piyy=None, pitx=None):
        """(<T^0x>/<Txx+Tyy>, <Txx-Tyy>/<Txx+Tyy>) (bulkinfo.py:203-229)."""
        ed = np.array(ed, copy=True)
        ed[ed < 1.0E-10] = 1.0E-10
        pre = self.eos.f_P(ed)
        vr2 = vx * vx + vy * vy + vz * vz
        vr2 = np.where(vr2 > 1.0, 0.999999, vr2)
        u0 = 1.0 / np.sqrt(1.0 - vr2)
        Tyy = (ed + pre) * u0 * u0 * vy * vy + pre
        Txx = (ed + pre) * u0 * u0 * vx * vx + pre
        T0x = (ed + pre) * u0 * u0 * vx
        if pixx is not None:
            v2 = ((Txx - Tyy).sum() + (pixx - piyy).sum()) / ((Txx + Tyy).sum() + (pixx + piyy).sum())
        else:
            v2 = (Txx - Tyy).sum() / (Txx + Tyy).sum()
        v1 = T0x.sum() / (Txx + Tyy).sum()
        return v1, v2

    def mean_vr(self, ed, vx, vy, vz=0.0):
        """<gamma e |v_T|>/<gamma e> over the plane (bulkinfo.py:231-240)."""
        ed = np.array(ed, copy=True)
        ed[ed < 1.0E-10] = 1.0E-10
        vr2 = vx * vx + vy * vy + vz * vz
        vr2 = np.where(vr2 > 1.0, 0.999999, vr2)
        u0 = 1.0 / np.sqrt(1.0 - vr2)
        return (u0 * ed * np.sqrt(vx * vx + vy * vy)).sum() / (u0 * ed).sum()

    def total_entropy(self, tau, ed, vx, vy, vz=0.0):
        """tau DX DY sum(u0 s) over the plane (bulkinfo.py:242-249)."""
        ed = np.array(ed, copy=True)
        ed[ed < 1.0E-10] = 1.0E-10
        vr2 = vx * vx + vy * vy + vz * vz
        vr2 = np.where(vr2 > 1.0, 0.999999, vr2)
        u0 = 1.0 / np.sqrt(1.0 - vr2)
        return (u0 * self.eos.f_S(ed)).sum() * tau * self.cfg.DX * self.cfg.DY

    def _etas(self):
        # eta of slice k as kernel_bulkinfo.cl:38 takes it: (K - NZ/2) DZ, integer NZ/2
        return (np.arange(self.cfg.NZ) - self.cfg.NZ // 2) * float(self.cfg.DZ)

    def _energy_entropy(self, tau, s):
        vol = tau * self.cfg.DX * self.cfg.DY * self.cfg.DZ
        eta = self._etas()
        e_lab = np.cosh(eta) * (s[:, 6] - s[:, 8]) + np.sinh(eta) * s[:, 7]
        return float(e_lab.sum() * vol), float(s[:, 9].sum() * vol)

    @staticmethod
    def _device_field(d_ev):
        if getattr(d_ev, 'field_id', None) not in (0, 1, 2):
            raise TypeError('the device reductions need a DeviceField of d_ev[0], d_ev[1] or d_ev[2]')
        return d_ev.handle, d_ev.field_id

    def get_total_energy_and_entropy_on_gpu(self, tau, d_ev):
        """Total lab-frame energy sum((e + P) u0^2 (cosh eta + v^eta sinh eta) - P cosh eta)
        and total entropy sum(u0 s), times tau DX DY DZ, of d_ev at tau (bulkinfo_h5.py:264-277,
        kernel_bulkinfo.cl:18-46); appended to energy / entropy and returned.  One device
        reduction (cv_slice_sums) in double instead of two per-cell arrays summed on the device."""
        h, fid = self._device_field(d_ev)
        e_total, s_total = self._energy_entropy(tau, h.slice_sums(fid))
        self.energy.append(e_total)
        self.entropy.append(s_total)
        return e_total, s_total

    def ecc_vs_rapidity_on_gpu(self, tau, d_ev):
        """(eccp1, eccp2) per eta slice: T^tx / (Txx + Tyy) and (Txx - Tyy) / (Txx + Tyy) summed
        over the plane (kernel_bulkinfo.cl:51-113, bulkinfo_h5.py:292-297).  The sums floor e at
        1e-10 and cap v^2 > 1 as bulkinfo.py does; kernel_bulkinfo.cl's gamma differs from that
        only for 1 - v^2 < 1e-7."""
        h, fid = self._device_field(d_ev)
        s = h.slice_sums(fid)
        return s[:, 2] / s[:, 1], s[:, 0] / s[:, 1]

    def total_energy_and_entropy(self, tau, bulk):
        """Host form of get_total_energy_and_entropy_on_gpu on a (NX, NY, NZ, 4) field."""
        ed = np.asarray(bulk[..., 0], np.float64)
        vx, vy, vz = (np.asarray(bulk[..., c], np.float64) for c in (1, 2, 3))
        pr = np.asarray(self.eos.f_P(ed), np.float64)
        u0 = 1.0 / np.sqrt(np.maximum(1.0 - vx * vx - vy * vy - vz * vz, 1.0e-7))
        eta = self._etas()[None, None, :]
        e_lab = (ed + pr) * u0 * u0 * (np.cosh(eta) + vz * np.sinh(eta)) - pr * np.cosh(eta)
        vol = tau * self.cfg.DX * self.cfg.DY * self.cfg.DZ
        return float(e_lab.sum() * vol), float((u0 * np.asarray(self.eos.f_S(ed), np.float64)).sum() * vol)

    def ecc_vs_rapidity(self, d_ev):
        """Append eccp1, eccp2 of every eta slice of d_ev (a device field proxy or an
        (N, 4) array) to ecc1_vs_rapidity / ecc2_vs_rapidity (bulkinfo.py:247-260)."""
        NX, NY, NZ = self.cfg.NX, self.cfg.NY, self.cfg.NZ
        arr = d_ev.get() if hasattr(d_ev, 'get') else np.asarray(d_ev)
        self._ecc_vs_rapidity(arr.reshape(NX, NY, NZ, 4))

    def _ecc_vs_rapidity(self, bulk):
        NZ = self.cfg.NZ
        ecc1 = np.empty(NZ)
        ecc2 = np.empty(NZ)
        for k in range(NZ):
            ecc1[k], ecc2[k] = self.eccp(bulk[:, :, k, 0], bulk[:, :, k, 1], bulk[:, :, k, 2], bulk[:, :, k, 3])
        self.ecc1_vs_rapidity.append(ecc1)
        self.ecc2_vs_rapidity.append(ecc2)

    def save(self, viscous_on=False):
        """The reference's text outputs (bulkinfo.py:268-340)."""
        path_out = os.path.abspath(self.cfg.fPathOut)
        f_T = self.eos.f_T
        np.savetxt(path_out + '/ex.dat', np.array(self.ex).T)
        np.savetxt(path_out + '/ey.dat', np.array(self.ey).T)
        np.savetxt(path_out + '/ez.dat', np.array(self.ez).T)
        np.savetxt(path_out + '/Tx.dat', f_T(np.array(self.ex)).T)
        np.savetxt(path_out + '/Ty.dat', f_T(np.array(self.ey)).T)
        np.savetxt(path_out + '/Tz.dat', f_T(np.array(self.ez)).T)
        np.savetxt(path_out + '/vx.dat', np.array(self.vx).T)
        np.savetxt(path_out + '/vy.dat', np.array(self.vy).T)
        np.savetxt(path_out + '/vz.dat', np.array(self.vz).T)
        if len(self.ecc2_vs_rapidity) != 0:
            np.savetxt(path_out + '/ecc2.dat', np.array(self.ecc2_vs_rapidity).T)
            np.savetxt(path_out + '/ecc1.dat', np.array(self.ecc1_vs_rapidity).T)
        entropy, vr, ecc2, ecc2_visc = [], [], [], []
        for idx, exy in enumerate(self.exy):
            vx, vy = self.vx_xy[idx], self.vy_xy[idx]
            np.savetxt(path_out + '/edxy%d.dat' % idx, exy)
            np.savetxt(path_out + '/Txy%d.dat' % idx, f_T(exy))
            np.savetxt(path_out + '/vx_xy%d.dat' % idx, vx)
            np.savetxt(path_out + '/vy_xy%d.dat' % idx, vy)
            ecc2.append(self.eccp(exy, vx, vy)[1])
            vr.append(self.mean_vr(exy, vx, vy))
            entropy.append(self.total_entropy(self.time[idx], exy, vx, vy))
            if viscous_on and idx < len(self.pixx_xy):
                ecc2_visc.append(self.eccp(exy, vx, vy, pixx=self.pixx_xy[idx], piyy=self.piyy_xy[idx],
                                           pitx=self.pitx_xy[idx])[1])
        for idx, exz in enumerate(self.exz):
            np.savetxt(path_out + '/ed_xz%d.dat' % idx, exz)
            np.savetxt(path_out + '/vx_xz%d.dat' % idx, self.vx_xz[idx])
            np.savetxt(path_out + '/vy_xz%d.dat' % idx, self.vy_xz[idx])
            np.savetxt(path_out + '/vz_xz%d.dat' % idx, self.vz_xz[idx])
            np.savetxt(path_out + '/T_xz%d.dat' % idx, f_T(exz))
        np.savetxt(path_out + '/eccp.dat', np.array(list(zip(self.time, ecc2))), header='tau  eccp')
        if viscous_on:
            np.savetxt(path_out + '/eccp_visc.dat', np.array(list(zip(self.time, ecc2_visc))),
                       header='tau  eccp_visc')
        np.savetxt(path_out + '/Tmax.dat', np.array(list(zip(self.time, f_T(np.array(self.edmax))))),
                   header='tau, Tmax')
        np.savetxt(path_out + '/edmax.dat', np.array(list(zip(self.time, self.edmax))), header='tau, edmax')
        np.savetxt(path_out + '/vr.dat', np.array(list(zip(self.time, vr))), header='tau <vr>')
        np.savetxt(path_out + '/entropy.dat', np.array(list(zip(self.time, entropy))), header='tau  entropy')
        if self.energy and len(self.energy) == len(self.time):
            # avg.dat of the save_to_hdf5 variant (bulkinfo_h5.py:300-306)
            k0 = self.cfg.NZ // 2
            edcent = [ez[k0] for ez in self.ez]
            np.savetxt(path_out + '/avg.dat',
                       np.array(list(zip(self.time, ecc2, edcent, self.entropy, self.energy, vr))),
                       header='tau, eccp, ed(0,0,0), stotal, Etotal, <vr>')
