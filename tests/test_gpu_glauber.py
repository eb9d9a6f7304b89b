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
