"""Optical Glauber initial condition on the device (k_glauber_ini, k_glauber_nbinary)
against the FP32 oracle restatement of kernel_glauber.cl, and through the CLVisc driver."""
import sys

import numpy as np
import pytest

from clvisc_amd import _lib
from clvisc_amd.config import read_config
from clvisc_amd.eos import Eos
from helpers import cv_params
from oracle import oracle as O

pytestmark = pytest.mark.gpu
# device expf / powf (ROCm ocml) and the host's libm may differ by an ulp or two per call;
# 50 exp per thickness, so the fields agree to FP32 accumulation, not bit for bit
REL = 2e-6


def _say(msg):
    print(msg)
    sys.stdout.flush()


@pytest.mark.parametrize('b,nz,bjorken', [(0.0, 1, False), (7.0, 41, False), (7.0, 9, True)])
def test_glauber_ini_matches_oracle(b, nz, bjorken):
    nx, ny, dx, dz = 101, 97, 0.16, 0.3
    cfg = read_config(argv=[])
    cfg.ImpactParameter = b
    g = _lib.CvGlauber.from_cfg(cfg, bjorken)
    og = O.glauber_params(cfg, bjorken)
    _say('case: glauber %dx%dx%d b=%.1f bjorken=%s' % (nx, ny, nz, b, bjorken))
    p = cv_params(nx, ny, nz, 0.02, dx, dx, dz, 0.6)
    h = _lib.Handle(p, Eos('ideal_gas').create_table(), viscous=False)
    h.glauber_ini(g, 1)
    dev = h.get('ev1')
    ref = O.glauber_ini(og, nx, ny, nz, dx, dx, dz, 'f32')
    err = np.abs(dev[:, 0] - ref[:, 0]).max() / ref[:, 0].max()
    _say('  max |e - e_oracle| / max e = %.3g, bit-equal %.4f' % (err, np.mean(dev[:, 0] == ref[:, 0])))
    assert err <= REL
    assert np.all(dev[:, 1:] == 0.0)
    nb = h.glauber_nbinary(g, -9.75, -9.75, 0.3, 0.3, 66, 66)
    nbr = O.glauber_nbinary(og, -9.75, -9.75, 0.3, 0.3, 66, 66, 'f32')
    assert np.abs(nb - nbr).max() <= REL * nbr.max()
    h.close()


def test_clvisc_optical_glauber_ini(tmp_path):
    """CLVisc.optical_glauber_ini with a centrality range: the impact parameter becomes the
    class's weighted <b> and the device field is that Glauber profile; nbin.dat written."""
    from clvisc_amd import CLVisc
    from clvisc_amd.glauber import weight_mean_b
    cfg = read_config(argv=[])
    cfg.NX, cfg.NY, cfg.NZ = 65, 63, 5
    cfg.DX, cfg.DY, cfg.DZ = 0.3, 0.3, 0.3
    cfg.fPathOut = str(tmp_path)
    _say('case: CLVisc.optical_glauber_ini 20-30 %')
    visc = CLVisc(cfg)
    visc.optical_glauber_ini(system='Au+Au', cent_min=20, cent_max=30, save_binary_collisions=True)
    assert cfg.ImpactParameter == pytest.approx(weight_mean_b(20, 30))
    ev = visc.ideal.d_ev[1].get()
    ref = O.glauber_ini(O.glauber_params(cfg), 65, 63, 5, 0.3, 0.3, 0.3, 'f32')
    assert np.abs(ev[:, 0] - ref[:, 0]).max() <= REL * ref[:, 0].max()
    assert (tmp_path / 'nbin.dat').exists()
    assert np.loadtxt(tmp_path / 'nbin.dat').shape == (66, 66)


def _partons(n, seed):
    rng = np.random.default_rng(seed)
    mom = rng.normal(0, 1.2, (n, 3))
    E = np.sqrt((mom ** 2).sum(1) + 0.3 ** 2)
    pos = np.column_stack([np.full(n, 0.6), rng.normal(0, 1.5, (n, 2)), rng.normal(0, 0.4, n)])
    return np.column_stack([E, mom, pos]).astype(np.float32)


@pytest.mark.parametrize('npart', [1, 700, 1500])
def test_smear_partons_matches_oracle(npart):
    """Device smearing (blocks of 4x8x8 cells, partons through LDS in chunks of 256 with their
    per-block distances and mt cosh, sinh; list order kept) against the FP32 oracle."""
    nx, ny, nz, h = 41, 37, 21, 0.2
    tab = Eos('crossover').create_table()
    p = cv_params(nx, ny, nz, 0.02, h, h, h, 0.6)
    pp = _partons(npart, npart)
    _say('case: smearing %d partons on %dx%dx%d' % (npart, nx, ny, nz))
    hd = _lib.Handle(p, tab, viscous=False)
    hd.smear_partons(pp, 0.6, 0.6, 1.45, 1)
    dev = hd.get('ev1')
    from helpers import orc_params
    ref = O.smear(orc_params(p, tab), pp, 0.6, 0.6, 1.45, 'f32')
    err = np.abs(dev[:, 0] - ref[:, 0]).max() / ref[:, 0].max()
    _say('  max |e - e_oracle| / max e = %.3g, bit-equal %.4f' % (err, np.mean(dev[:, 0] == ref[:, 0])))
    assert err <= 1e-5
    assert np.all(dev[:, 1:] == 0.0)
    # change_longitudinal_profile: mid-rapidity copied with a profile
    prof = np.exp(-((np.arange(nz) - nz // 2) * h) ** 2 / 2).astype(np.float32)
    hd.longitudinal_profile(prof, 1)
    e2 = hd.get('ev1')[:, 0].reshape(nx, ny, nz)
    mid = dev[:, 0].reshape(nx, ny, nz)[:, :, nz // 2]
    np.testing.assert_array_equal(e2, prof[None, None, :] * mid[:, :, None])
    hd.close()


def test_clvisc_smear_from_p4x4_and_file(tmp_path):
    from clvisc_amd import CLVisc
    cfg = read_config(argv=[])
    cfg.NX, cfg.NY, cfg.NZ = 33, 31, 9
    cfg.DX, cfg.DY, cfg.DZ = 0.2, 0.2, 0.2
    cfg.fPathOut = str(tmp_path)
    _say('case: CLVisc.smear_from_p4x4 / create_ini_from_partons')
    visc = CLVisc(cfg)
    pp = _partons(300, 11)
    visc.smear_from_p4x4(pp, SIGR=0.5, SIGZ=0.5, KFACTOR=1.2, force_bjorken=True)
    e = visc.ideal.d_ev[1].get()[:, 0].reshape(33, 31, 9)
    assert np.all(e == e[:, :, 4:5])
    fname = tmp_path / 'partons.txt'
    np.savetxt(fname, pp, header='E px py pz t x y z')
    visc.create_ini_from_partons(str(fname), SIGR=0.5, SIGZ=0.5, KFACTOR=1.2)
    from helpers import orc_params
    ref = O.smear(orc_params(visc.handle.params, visc.eos_table), np.loadtxt(fname, dtype=np.float32), 0.5, 0.5,
                  1.2, 'f32')
    got = visc.ideal.d_ev[1].get()
    assert np.abs(got[:, 0] - ref[:, 0]).max() <= 1e-5 * ref[:, 0].max()


@pytest.mark.parametrize('eos_name', ['ideal_gas', 'crossover'])
def test_sd_to_ed_on_device(eos_name):
    """TRENTo entropy profile -> energy density on the device (cv_sd_to_ed) against the host
    restatement of from_sd_to_ed, including values beyond both ends of the table."""
    from clvisc_amd.ic import from_sd_to_ed, sd_to_ed_on_device
    from clvisc_amd.ideal import DeviceField
    eos = Eos(eos_name)
    nx, ny, nz = 31, 29, 3
    p = cv_params(nx, ny, nz, 0.02, 0.2, 0.2, 0.3, 0.6)
    h = _lib.Handle(p, eos.create_table(), viscous=False)
    rng = np.random.default_rng(5)
    s = np.exp(rng.uniform(np.log(1e-6), np.log(2.0 * eos.s[-1]), nx * ny * nz)).astype(np.float32)
    s[:5] = [0.0, eos.s[0], eos.s[-1], 1.5 * eos.s[-1], eos.s[1] * 0.5]
    ev = np.zeros((s.size, 4), np.float32)
    ev[:, 0] = s
    h.set('ev1', ev)
    _say('case: sd_to_ed %s, %d cells' % (eos_name, s.size))
    sd_to_ed_on_device(DeviceField(h, 'ev1'), eos)
    got = h.get('ev1')[:, 0]
    ref = from_sd_to_ed(s.astype(np.float64), eos).astype(np.float32)
    rel = np.abs(got.astype(np.float64) - ref) / np.maximum(np.abs(ref), 1e-30)
    _say('  max rel %.3g, bit-equal %.4f' % (rel.max(), np.mean(got == ref)))
    assert rel.max() <= 2.4e-7
    h.close()
