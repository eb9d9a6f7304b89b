"""pizz_o_ep of CLVisc.check_pizz on the device (cv_pizz_o_ep) equals the oracle's bit for bit
after viscous steps, rowcol and octave tables; ideal handles refuse it."""
import sys

import numpy as np
import pytest

from clvisc_amd import _lib, ic
from clvisc_amd.eos import Eos
from oracle import oracle as O
from helpers import cv_params, orc_params

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('eos_type', ['ideal_gas', 'crossover'])
def test_pizz_o_ep_matches_oracle(eos_type):
    nx, ny, nz = 37, 33, 9
    tab = Eos(eos_type).create_table()
    p = cv_params(nx, ny, nz, 0.02, 0.08, 0.08, 0.15, 0.6, etaos_xmin=0.0, etaos_ymin=0.08,
                  bulk=(0.05, 0.155, 0.02))
    pos = ic.woods_saxon_positions(12, 9, R=0.6, a=0.1)
    ev = ic.hotspots(nx, ny, nz, 0.08, 0.08, 0.15, positions=pos, emax=30.0)
    print('case: pizz_o_ep, %s, %dx%dx%d after 6 steps' % (eos_type, nx, ny, nz))
    sys.stdout.flush()
    h = _lib.Handle(p, tab, viscous=True)
    h.visc_initialize(p.tau0)
    h.ideal_stage(1, np.float32(p.tau0))
    h.set('ev1', ev)
    h.visc_steps(0, 6)
    o = O.Oracle(orc_params(p, tab), 'f32', viscous=True)
    o.load_ini(h.get('ev1'), pi=h.get('pi1'))
    want = o.pizz_o_ep()
    got = h.pizz_o_ep(1)
    print('  mid slice %.6g, slices with cells above threshold %d of %d' % (got[nz // 2], np.isfinite(got).sum(), nz))
    assert np.isfinite(got).all() and np.any(got != 0.0)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    h.close()
    hi = _lib.Handle(cv_params(nx, ny, nz, 0.02, 0.08, 0.08, 0.15, 0.6), tab, viscous=False)
    with pytest.raises(_lib.CvError) as err:
        hi.pizz_o_ep(1)
    assert err.value.code == _lib.CV_ERR_UNSUPPORTED
    hi.close()


def test_clvisc_check_pizz(tmp_path, capsys):
    from clvisc_amd import CLVisc
    from clvisc_amd.config import read_config
    cfg = read_config(argv=[])
    cfg.NX, cfg.NY, cfg.NZ = 21, 19, 7
    cfg.eos_type = 'ideal_gas'
    cfg.fPathOut = str(tmp_path)
    cfg.eos_table_dir = ''
    visc = CLVisc(cfg)
    ev = ic.hotspots(21, 19, 7, cfg.DX, cfg.DY, cfg.DZ, positions=ic.woods_saxon_positions(5, 2, R=0.5, a=0.1),
                     emax=20.0)
    visc.ideal.load_ini(ev)
    visc.IS_initialize()
    vals = visc.check_pizz()
    assert vals.shape == (7,)
    assert '<pizz/(e+P)> =' in capsys.readouterr().out
    np.testing.assert_array_equal(vals, visc.handle.pizz_o_ep(1))
