"""Chemical potentials at freeze-out (clvisc_amd/chemical_potential.py, after
pyvisc/eos/chemical_potential.py) against the fixtures under tests/golden/chem/ (cut from
the reference's s95p tables by tests/golden/make_chem_fixture.py):

* ChemForReso.dat, the resonance potentials the PCE165 directory's own script wrote at
  e_frz = 0.144 GeV/fm^3; that script selects the stable species by the "Stable" mark alone
  (26 species against the table's 28 columns), so the fixture is compared with that rule;
* the stable list of the current rule ("Stable" or "********") is the list of numbered
  species of particles.dat, as many as the table's header names (the table README's rule);
* the decay-chain recursion equals the solution of the linear system mu = A mu + t;
* the interpolation reproduces table rows at the nodes and is linear between them.
"""
import os

import numpy as np
import pytest

from clvisc_amd import chemical_potential as CP

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'chem')
PCE165 = os.path.join(GOLD, 's95p-PCE165-v0')
PCE150 = os.path.join(GOLD, 's95p-PCE-v1')


def _gold():
    pids, mu = [], []
    with open(os.path.join(PCE165, 'ChemForReso.dat')) as f:
        for line in f:
            a, b = line.split()
            pids.append(a)
            mu.append(float(b))
    return pids, np.array(mu)


def test_resonance_potentials_match_the_reference_file():
    chem = CP.ChemicalPotential(0.144, eos_type='lattice_pce165', table_dir=GOLD, stable_marks=('Stable',))
    pids, mu = chem.chemical_potential_for_all()
    gpids, gmu = _gold()
    assert pids == gpids
    got = np.array([mu[p] for p in pids], dtype=np.float64)
    # the file holds 12 significant digits of values below 1 GeV
    np.testing.assert_allclose(got, gmu, rtol=0, atol=1e-12)
    assert np.count_nonzero(gmu) > 150


def test_written_file_matches_the_reference_file(tmp_path):
    chem = CP.ChemicalPotential(0.144, eos_type='lattice_pce165', table_dir=GOLD, stable_marks=('Stable',))
    fname = chem.get_chemical_potential_for_resonance(str(tmp_path))
    assert os.path.basename(fname) == 'chemical_potential.dat'
    rows = [line.split() for line in open(fname)]
    gpids, gmu = _gold()
    assert [r[0] for r in rows] == gpids
    np.testing.assert_allclose([float(r[1]) for r in rows], gmu, rtol=0, atol=1e-12)


def _numbered(path):
    out = []
    for line in open(os.path.join(path, 'particles.dat')):
        t = line.split()
        # numbered lines: "<number> <pid> <name> <lifetime> <mark>"; the others start at the pid
        if len(t) > 1 and t[0].isdigit() and t[1].lstrip('-').isdigit():
            out.append((int(t[0]), t[1]))
    return out


@pytest.mark.parametrize('eos_type,path', [('lattice_pce165', PCE165), ('lattice_pce150', PCE150)])
def test_stable_list_is_the_numbered_species(eos_type, path):
    chem = CP.ChemicalPotential(0.15, eos_type=eos_type, table_dir=GOLD)
    stables = chem.get_pid_for_stable()
    numbered = _numbered(path)
    assert [n for n, _ in numbered] == list(range(1, len(numbered) + 1))
    assert stables == [pid for _, pid in numbered]
    _, _, _, nstable, _ = CP.read_pichem(os.path.join(path, os.path.basename(path) + '_pichem1.dat'))
    assert len(stables) == nstable
    # the "Stable"-only rule misses the species marked ******** (short lifetimes kept stable)
    legacy = CP.ChemicalPotential(0.15, eos_type=eos_type, table_dir=GOLD, stable_marks=('Stable',))
    assert len(legacy.get_pid_for_stable()) < nstable


def _decays(path):
    """pdg05.dat parsed on its own: [(pid, [(ratio, [daughters])])]."""
    lines = open(os.path.join(path, 'pdg05.dat')).read().splitlines()
    out, i = [], 0
    while i < len(lines):
        head = lines[i].split()
        n = int(head[-1])
        ch = []
        for line in lines[i + 1:i + 1 + n]:
            t = line.split()
            ch.append((float(t[2]), t[3:3 + int(t[1])]))
        out.append((head[0], ch))
        i += n + 1
    return out


@pytest.mark.parametrize('eos_type,path,efrz', [('lattice_pce165', PCE165, 0.1377),
                                                 ('lattice_pce150', PCE150, 0.16123)])
def test_recursion_equals_the_linear_solve(eos_type, path, efrz):
    chem = CP.ChemicalPotential(efrz, eos_type=eos_type, table_dir=GOLD)
    pids, mu = chem.chemical_potential_for_all()
    table = dict(zip(chem.get_pid_for_stable(), chem.get_chemical_potential_for_stable(efrz)))
    table['22'] = 0.0
    decays = _decays(path)
    index = {pid: k for k, (pid, _) in enumerate(decays)}
    n = len(decays)
    A = np.zeros((n, n))
    t = np.zeros(n)
    for k, (pid, ch) in enumerate(decays):
        if ch == [(1.0, [pid])]:          # decays into itself: a stable species, table value
            t[k] = table[pid]
            continue
        for ratio, daughters in ch:
            for d in daughters:
                assert index[d] < k           # lighter daughters come first
                A[k, index[d]] += ratio
    sol = np.linalg.solve(np.eye(n) - A, t)
    got = np.array([mu[p] for p in pids])
    assert pids == [p for p, _ in decays]
    np.testing.assert_allclose(got, sol, rtol=1e-12, atol=1e-15)
    # a stable species keeps its table value bit for bit
    for k, (pid, ch) in enumerate(decays):
        if ch == [(1.0, [pid])]:
            assert mu[pid] == table[pid]
    assert np.count_nonzero(got) > 150


def test_interpolation_nodes_and_linearity():
    chem = CP.ChemicalPotential(0.15, eos_type='lattice_pce150', table_dir=GOLD)
    e0, de, ne, _, rows = CP.read_pichem(os.path.join(PCE150, 's95p-PCE-v1_pichem1.dat'))
    for k in (1, 7, 20):
        np.testing.assert_allclose(chem.get_chemical_potential_for_stable(e0 + k * de), rows[k], rtol=1e-13)
    a = e0 + 11.25 * de
    b = e0 + 11.75 * de
    m = chem.get_chemical_potential_for_stable(0.5 * (a + b))
    np.testing.assert_allclose(m, 0.5 * (rows[11] + rows[12]), rtol=1e-12)
    np.testing.assert_allclose(chem.get_chemical_potential_for_stable(a) + chem.get_chemical_potential_for_stable(b),
                               2.0 * m, rtol=1e-12)


def test_table_edges():
    chem = CP.ChemicalPotential(0.15, eos_type='lattice_pce150', table_dir=GOLD)
    e0, de, ne, nstable, _ = CP.read_pichem(os.path.join(PCE150, 's95p-PCE-v1_pichem1.dat'))
    # beyond row 500 the reference returns zero potentials
    z = chem.get_chemical_potential_for_stable(e0 + 600 * de)
    assert z.shape == (nstable,) and not z.any()
    for efrz in (e0 - 0.5 * de, e0 + (ne - 1) * de + 0.25 * de):
        with pytest.raises(ValueError):
            chem.get_chemical_potential_for_stable(efrz)


def test_other_eos_types_write_zeros(tmp_path):
    chem = CP.ChemicalPotential(0.15, eos_type='ideal_gas', table_dir=GOLD)
    assert chem.name == 's95p-PCE165-v0'
    fname = chem.get_chemical_potential_for_resonance(str(tmp_path))
    rows = [line.split() for line in open(fname)]
    assert [r[0] for r in rows] == [p for p, _ in _decays(PCE165)]
    assert all(r[1] == '0.0' for r in rows)


def test_create_table_takes_efrz_from_the_eos(tmp_path, monkeypatch):
    from clvisc_amd.eos import Eos
    out, extra = tmp_path / 'out', tmp_path / 'spec'
    out.mkdir()
    extra.mkdir()
    monkeypatch.setenv(CP.TABLE_DIR_ENV, GOLD)
    # the ideal gas at T = 0.095 GeV: e_frz = 0.1497 GeV/fm^3, inside the fixture's rows
    assert 0.121 < float(Eos('ideal_gas').f_ed(0.095)) < 0.181
    fname = CP.create_table(0.095, str(out), eos_type='ideal_gas', copy_to=[str(extra)])
    assert open(fname).read() == open(str(extra / 'chemical_potential.dat')).read()
    # a lattice PCE potential file at the same e_frz through an EOS object
    eos = Eos('ideal_gas')
    fname = CP.create_table(0.095, str(out), eos_type='lattice_pce150', eos=eos)
    chem = CP.ChemicalPotential(float(eos.f_ed(0.095)), eos_type='lattice_pce150')
    pids, mu = chem.chemical_potential_for_all()
    rows = [line.split() for line in open(fname)]
    assert [r[0] for r in rows] == pids
    assert [float(r[1]) for r in rows] == [float(mu[p]) for p in pids]
    monkeypatch.delenv(CP.TABLE_DIR_ENV)
    with pytest.raises(FileNotFoundError):
        CP.ChemicalPotential(0.15)
