"""Configuration interface of pyvisc/config.py:13-230: hydro.info defaults,
--KEY overrides, write_config round trip.  CPU only."""
import configparser

import numpy as np

from clvisc_amd.config import read_config, write_config


def test_defaults_and_overrides():
    cfg = read_config(argv=[])
    assert cfg.use_float32 and cfg.real is np.float32
    assert cfg.BSZ == 64
    cfg2 = read_config(argv=['--NX', '11', '--DT', '0.01', '--riemann_test', '--eos_type', 'ideal_gas'])
    assert cfg2.NX == 11 and cfg2.DT == 0.01 and cfg2.riemann_test and cfg2.eos_type == 'ideal_gas'
    assert not cfg2.gubser_visc_test


def test_write_config_round_trip(tmp_path):
    cfg = read_config(argv=['--NX', '17', '--ZETAOS_MAX', '0.05'])
    cfg.fPathOut = str(tmp_path)
    write_config(cfg, comments='test')
    again = read_config(path=str(tmp_path / 'hydro.info'), argv=[])
    assert again.NX == 17 and again.ZETAOS_MAX == 0.05
    p = configparser.ConfigParser()
    p.optionxform = str
    p.read(str(tmp_path / 'hydro.info'))
    assert set(p.sections()) >= {'path', 'glauber', 'geometry', 'intrinsic'}
