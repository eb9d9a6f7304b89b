"""CLIdeal writes chemical_potential.dat into its output directory when the s95p tables are
configured (pyvisc/ideal.py:65), the same file the host module writes, and nothing when
they are not."""
import os
import sys

import pytest

from clvisc_amd import chemical_potential as CP
from clvisc_amd.config import read_config

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'chem')


def _cfg(out, table_dir, eos_type, tfrz):
    cfg = read_config(argv=[])
    cfg.NX, cfg.NY, cfg.NZ = 9, 9, 1
    cfg.eos_type = eos_type
    cfg.TFRZ = tfrz
    cfg.fPathOut = str(out)
    cfg.eos_table_dir = table_dir
    return cfg


# freeze-out temperatures at which e_frz lies inside the fixture's table rows (0.121-0.181)
@pytest.mark.parametrize('eos_type,tfrz', [('ideal_gas', 0.095), ('crossover', 0.1251)])
def test_driver_writes_the_chemical_potential_file(tmp_path, monkeypatch, eos_type, tfrz):
    from clvisc_amd import CLIdeal
    monkeypatch.delenv(CP.TABLE_DIR_ENV, raising=False)
    print('case: CLIdeal 9x9x1, %s, TFRZ %g, chemical_potential.dat' % (eos_type, tfrz))
    sys.stdout.flush()
    on, off, host = tmp_path / 'on', tmp_path / 'off', tmp_path / 'host'
    host.mkdir()
    ideal = CLIdeal(_cfg(on, GOLD, eos_type, tfrz))
    assert 0.121 < ideal.efrz < 0.181
    fname = os.path.join(str(on), 'chemical_potential.dat')
    ref = CP.create_table(tfrz, str(host), eos_type=eos_type, table_dir=GOLD)
    assert open(fname).read() == open(ref).read()
    lines = open(fname).read().splitlines()
    assert len(lines) == 217 and all(line.split()[1] == '0.0' for line in lines)
    CLIdeal(_cfg(off, '', eos_type, tfrz))
    assert not os.path.exists(os.path.join(str(off), 'chemical_potential.dat'))
