"""BulkInfo on the device (cv_slice_sums + gathered slices) against the reference's host
path (whole-field copy, NumPy formulas) on an evolved viscous field."""
import sys
import time

import numpy as np
import pytest

from clvisc_amd import _lib, ic
from clvisc_amd.bulkinfo import BulkInfo
from clvisc_amd.config import read_config
from clvisc_amd.eos import Eos
from clvisc_amd.ideal import DeviceField
from helpers import cv_params, orc_params
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg)
    sys.stdout.flush()


def test_device_bulkinfo_matches_host_path():
    nx, ny, nz = 61, 57, 21
    eos = Eos('crossover')
    tab = eos.create_table()
    p = cv_params(nx, ny, nz, 0.02, 0.08, 0.08, 0.15, 0.6, etaos_xmin=0.0, etaos_ymin=0.08,
                  bulk=(0.05, 0.155, 0.02))
    pos = ic.woods_saxon_positions(30, 9, R=1.2, a=0.2)
    ev0 = ic.hotspots(nx, ny, nz, 0.08, 0.08, 0.15, positions=pos, emax=40.0)
    _say('case: bulkinfo device vs host, %dx%dx%d after 30 steps' % (nx, ny, nz))
    h = _lib.Handle(p, tab, viscous=True)
    h.visc_initialize(p.tau0)
    h.ideal_stage(1, np.float32(p.tau0))
    h.set('ev1', ev0)
    h.visc_steps(0, 30)
    cfg = read_config()
    cfg.NX, cfg.NY, cfg.NZ = nx, ny, nz
    cfg.DX, cfg.DY, cfg.DZ = 0.08, 0.08, 0.15
    d_ev1 = DeviceField(h, 'ev1')
    bd = BulkInfo(cfg, eos, device_reductions=True)
    bh = BulkInfo(cfg, eos, device_reductions=False)
    t0 = time.perf_counter()
    bd.get(1.2, d_ev1, 1.0)
    t1 = time.perf_counter()
    bh.get(1.2, d_ev1, 1.0)
    t2 = time.perf_counter()
    _say('  device %.2f ms, host %.2f ms' % (1e3 * (t1 - t0), 1e3 * (t2 - t1)))
    for name in ('ex', 'ey', 'ez', 'vx', 'vy', 'vz', 'exy', 'vx_xy', 'vy_xy', 'exz', 'vx_xz', 'vy_xz', 'vz_xz'):
        np.testing.assert_array_equal(getattr(bd, name)[0], getattr(bh, name)[0], err_msg=name)
    for name in ('ecc1_vs_rapidity', 'ecc2_vs_rapidity'):
        a, b = getattr(bd, name)[0], getattr(bh, name)[0]
        _say('  %s max |device - host| %.3g (max |host| %.3g)' % (name, np.abs(a - b).max(), np.abs(b).max()))
        assert np.abs(a - b).max() <= 1e-6 * max(1.0, np.abs(b).max())
    # the six sums against float64 NumPy with the device's FP32 table values
    ev = h.get('ev1').reshape(nx, ny, nz, 4)
    op = orc_params(p, tab)
    s = bd.slice_sums[0]
    for k in (0, nz // 2, nz - 1):
        e = np.maximum(ev[:, :, k, 0], np.float32(1e-10)).reshape(-1)
        f = O.eos_funcs(op, e).astype(np.float64)
        vx, vy, vz = (ev[:, :, k, c].reshape(-1).astype(np.float64) for c in (1, 2, 3))
        vr2 = vx * vx + vy * vy + vz * vz
        vr2 = np.where(vr2 > 1.0, 0.999999, vr2)
        u0 = 1.0 / np.sqrt(1.0 - vr2)
        w = (e + f[:, 0]) * u0 * u0
        txx, tyy = w * vx * vx + f[:, 0], w * vy * vy + f[:, 0]
        ref = [(txx - tyy).sum(), (txx + tyy).sum(), (w * vx).sum(), (u0 * f[:, 1]).sum(),
               (u0 * e * np.sqrt(vx * vx + vy * vy)).sum(), (u0 * e).sum()]
        for q in range(6):
            assert abs(s[k, q] - ref[q]) <= 1e-10 * max(abs(ref[q]), 1e-12) + 1e-12 * abs(ref[1]), (k, q)
        # kernel_bulkinfo.cl's columns: unfloored e, gamma with 1 - v^2 floored at 1e-7
        er = ev[:, :, k, 0].reshape(-1)
        fr = O.eos_funcs(op, er).astype(np.float64)
        ug = 1.0 / np.sqrt(np.maximum(1.0 - (vx * vx + vy * vy + vz * vz), 1e-7))
        wr = (er + fr[:, 0]) * ug * ug
        ref2 = [wr.sum(), (wr * vz).sum(), fr[:, 0].sum(), (ug * fr[:, 1]).sum()]
        for q in range(4):
            assert abs(s[k, 6 + q] - ref2[q]) <= 1e-10 * abs(ref2[q]) + 1e-12 * abs(ref2[0]), (k, 6 + q)
    # total energy / entropy and eccp vs rapidity through the reference's method names, against
    # the per-cell formula in float64 on the same FP32 table values
    tau = 1.2
    E, S = bd.get_total_energy_and_entropy_on_gpu(tau, d_ev1)
    e = ev[..., 0].reshape(-1)
    f = O.eos_funcs(op, e).astype(np.float64)
    vx, vy, vz = (ev[..., c].reshape(-1).astype(np.float64) for c in (1, 2, 3))
    ug = 1.0 / np.sqrt(np.maximum(1.0 - (vx * vx + vy * vy + vz * vz), 1e-7))
    eta = np.tile((np.arange(nz) - nz // 2) * float(cfg.DZ), nx * ny)
    elab = (e + f[:, 0]) * ug * ug * (np.cosh(eta) + vz * np.sinh(eta)) - f[:, 0] * np.cosh(eta)
    vol = tau * cfg.DX * cfg.DY * cfg.DZ
    _say('  E_lab %.8g (numpy %.8g), S %.8g (numpy %.8g)'
         % (E, elab.sum() * vol, S, (ug * f[:, 1]).sum() * vol))
    assert E == pytest.approx(elab.sum() * vol, rel=1e-10)
    assert S == pytest.approx((ug * f[:, 1]).sum() * vol, rel=1e-10)
    assert bd.energy == [E] and bd.entropy == [S]
    # the host form (Python EOS interpolation instead of the FP32 table): close, not equal
    Eh, Sh = bh.total_energy_and_entropy(tau, ev)
    _say('  host form: E %.8g S %.8g' % (Eh, Sh))
    assert Eh == pytest.approx(E, rel=1e-3) and Sh == pytest.approx(S, rel=1e-3)
    e1, e2 = bd.ecc_vs_rapidity_on_gpu(tau, d_ev1)
    np.testing.assert_array_equal(e1, bd.ecc1_vs_rapidity[0])
    np.testing.assert_array_equal(e2, bd.ecc2_vs_rapidity[0])
    h.close()


def test_device_bulkinfo_full_size_timing():
    """385x385x115 benchmark event: one BulkInfo.get on the device vs the reference's host
    path (272 MB field copy + NumPy over 115 slices); same eccentricities."""
    import bench
    p, tab, ev = bench.make_workload(True)
    _say('case: bulkinfo timing 385x385x115')
    h = _lib.Handle(p, tab, viscous=True)
    h.visc_initialize(bench.TAU0)
    h.ideal_stage(1, np.float32(bench.TAU0))
    h.set('ev1', ev)
    h.visc_steps(0, 20)
    cfg = read_config()
    cfg.NX, cfg.NY, cfg.NZ = bench.NX, bench.NY, bench.NZ
    eos = Eos('crossover')
    d_ev1 = DeviceField(h, 'ev1')
    bd = BulkInfo(cfg, eos, device_reductions=True)
    bh = BulkInfo(cfg, eos, device_reductions=False)
    bd.get(1.0, d_ev1, 1.0)                      # warm-up (kernels loaded)
    t0 = time.perf_counter()
    bd.get(1.0, d_ev1, 1.0)
    t1 = time.perf_counter()
    bh.get(1.0, d_ev1, 1.0)
    t2 = time.perf_counter()
    _say('  BulkInfo.get: device %.1f ms, host path %.1f ms (%.0fx)'
         % (1e3 * (t1 - t0), 1e3 * (t2 - t1), (t2 - t1) / (t1 - t0)))
    a, b = bd.ecc2_vs_rapidity[-1], bh.ecc2_vs_rapidity[-1]
    assert np.abs(a - b).max() <= 1e-6 * max(1.0, np.abs(b).max())
    assert t1 - t0 < t2 - t1
    h.close()
