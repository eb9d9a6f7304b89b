"""Effective chemical potentials of every hadron species at freeze-out, written as
``chemical_potential.dat`` for the particlization codes.

Follows pyvisc/eos/chemical_potential.py:24-169, which ``CLIdeal.__init__`` runs
(pyvisc/ideal.py:65).  The s95p partial-chemical-equilibrium EOS tables (Huovinen and
Petreczky) fix the particle ratios below the chemical freeze-out temperature with one
chemical potential per stable species, tabulated against energy density in
``<dir>/<dir>_pichem1.dat``.  At the freeze-out energy density e_frz:

* the stable species' potentials are interpolated linearly between the two table rows
  around e_frz (pyvisc/eos/chemical_potential.py:60-96);
* every species of ``pdg05.dat``, in file order, gets mu_i = sum over its decay channels
  of branching ratio x sum of its daughters' mu (:98-132); a stable species decays into
  itself with ratio 1, so it keeps its table value, and a resonance sees the final values
  of its lighter daughters;
* ``chemical_potential.dat`` lists ``pid mu`` in pdg05 order; EOS types other than the
  two lattice PCE ones write 0.0 for every species (:105-107, :134-140).

The tables live in the reference's ``pyvisc/eos/eos_table`` directory; pass it as
``table_dir`` or set the environment variable ``CLVISC_EOS_TABLE_DIR``.

Deviations, all on inputs where the reference fails or reads garbage: e_frz below the
table's first row (the reference indexes from the end of the table) and e_frz in the
last row (the reference raises IndexError) raise ValueError; a missing table directory
raises FileNotFoundError naming what is needed.  The reference copies the file into its
CLSmoothSpec/Resource and sampler directories (:164-169); here the destinations are the
``copy_to`` argument.
"""
from __future__ import annotations

import math
import os
import shutil

import numpy as np

TABLE_DIR_ENV = 'CLVISC_EOS_TABLE_DIR'

# table sub-directory per EOS type (pyvisc/eos/chemical_potential.py:32-35)
_PCE150 = 's95p-PCE-v1'
_PCE165 = 's95p-PCE165-v0'

# marks of the stable species in the last column of particles.dat (:48-50)
STABLE_MARKS = ('Stable', '********')

# beyond this row index the reference returns zero potentials (:80-83)
_ZERO_ABOVE_ROW = 500


def table_root(table_dir=None):
    """The directory holding the s95p table directories: ``table_dir``, else
    $CLVISC_EOS_TABLE_DIR."""
    root = table_dir if table_dir is not None else os.environ.get(TABLE_DIR_ENV)
    if not root:
        raise FileNotFoundError(
            'the chemical potentials need the s95p EOS tables: pass table_dir= (the reference '
            "checkout's pyvisc/eos/eos_table) or set %s" % TABLE_DIR_ENV)
    if not os.path.isdir(root):
        raise FileNotFoundError('EOS table directory %r does not exist' % (root,))
    return root


def table_name(eos_type):
    """s95p table directory the reference reads for an EOS type (:32-35, :62-64)."""
    return _PCE150 if eos_type == 'lattice_pce150' else _PCE165


def read_pichem(fname):
    """Header (e0, de, ne, nstable) and the rows of a ``*_pichem1.dat`` table, in order of
    increasing energy density (the file runs from the highest e down)."""
    with open(fname, 'r') as f:
        e0 = float(f.readline())
        de, ne = f.readline().split()
        de, ne = float(de), int(ne)
        nstable = int(f.readline())
    rows = np.loadtxt(fname, skiprows=3, ndmin=2)[::-1]
    if rows.shape != (ne, nstable):
        raise ValueError('%s: header says %d rows x %d columns, the table holds %s'
                         % (fname, ne, nstable, rows.shape))
    return e0, de, ne, nstable, rows


def read_decays(fname):
    """The species of ``pdg05.dat`` in file order, each with its decay channels as
    (branching ratio, daughter pids) (:112-131)."""
    with open(fname, 'r') as f:
        lines = f.readlines()
    species = []
    i = 0
    while i < len(lines):
        head = lines[i].split()
        pid, ndecays = head[0], int(head[-1])
        channels = []
        for j in range(ndecays):
            decay = lines[i + j + 1].split()
            ndaughter = int(decay[1])
            channels.append((float(decay[2]), decay[3:3 + ndaughter]))
        species.append((pid, channels))
        i += ndecays + 1
    return species


class ChemicalPotential(object):
    """Chemical potentials at the freeze-out energy density ``efrz`` (GeV/fm^3)
    (pyvisc/eos/chemical_potential.py:24-142)."""

    def __init__(self, efrz, eos_type='lattice_pce150', table_dir=None, stable_marks=STABLE_MARKS):
        self.efrz = efrz
        self.eos_type = eos_type
        self.stable_marks = tuple(stable_marks)
        self.name = table_name(eos_type)
        self.path = os.path.join(table_root(table_dir), self.name)
        if not os.path.isdir(self.path):
            raise FileNotFoundError('%s: no %s table directory' % (self.path, self.name))
        # read both inputs once, as the reference's constructor does (:37-39)
        self.mu_for_stable = self.get_chemical_potential_for_stable(efrz)
        self.pids = self.get_pid_for_stable()

    def get_pid_for_stable(self):
        """Pids of the stable species (particles.dat lines marked stable), photon excluded;
        column j of the pichem table belongs to entry j (:44-54)."""
        stables = []
        with open(os.path.join(self.path, 'particles.dat'), 'r') as f:
            for line in f:
                info = line.split()
                if info and info[-1] in self.stable_marks and info[1] != 'Gamma':
                    stables.append(info[1])
        return stables

    def get_chemical_potential_for_stable(self, efrz):
        """Stable species' potentials at ``efrz``, linear in e between the table rows
        around it (:57-96)."""
        e0, de, ne, nstable, rows = read_pichem(os.path.join(self.path, self.name + '_pichem1.dat'))
        idx = int(math.floor((efrz - e0) / de))
        if idx > _ZERO_ABOVE_ROW:
            return np.zeros_like(rows[0])
        if idx < 0 or idx + 1 >= ne:
            raise ValueError('freeze-out energy density %r GeV/fm^3 is outside the chemical '
                             'potential table [%r, %r)' % (efrz, e0, e0 + (ne - 1) * de))
        ed0, ed1 = e0 + idx * de, e0 + (idx + 1) * de
        w0 = (efrz - ed0) / de
        w1 = (ed1 - efrz) / de
        return w0 * rows[idx + 1] + w1 * rows[idx]

    def chemical_potential_for_all(self):
        """(pids in pdg05 order, {pid: mu}) for every species (:98-132)."""
        pid_for_stable = self.get_pid_for_stable()
        mu_for_stable = self.get_chemical_potential_for_stable(self.efrz)
        if len(pid_for_stable) > len(mu_for_stable):
            raise ValueError('%d stable species but %d table columns'
                             % (len(pid_for_stable), len(mu_for_stable)))
        mu_for_all = {pid: mu_for_stable[i] for i, pid in enumerate(pid_for_stable)}
        mu_for_all['22'] = 0.0
        pids = []
        for pid, channels in read_decays(os.path.join(self.path, 'pdg05.dat')):
            mu_reso = 0.0
            for branch_ratio, daughters in channels:
                for d in daughters:
                    if d not in mu_for_all:
                        raise KeyError('pdg05.dat: species %s decays into %s before %s is listed'
                                       % (pid, d, d))
                    mu_reso += branch_ratio * mu_for_all[d]
            mu_for_all[pid] = mu_reso
            pids.append(pid)
        return pids, mu_for_all

    @property
    def set_to_zero(self):
        """Only the lattice PCE EOS carry effective potentials (:105-107)."""
        return self.eos_type not in ('lattice_pce165', 'lattice_pce150')

    def get_chemical_potential_for_resonance(self, save_path):
        """Write ``save_path/chemical_potential.dat``: one ``pid mu`` line per species of
        pdg05.dat (:134-140).  Returns the file name."""
        pids, mu_for_all = self.chemical_potential_for_all()
        save_fname = os.path.join(save_path, 'chemical_potential.dat')
        with open(save_fname, 'w') as fout:
            for pid in pids:
                print(pid, 0.0 if self.set_to_zero else mu_for_all[pid], file=fout)
        return save_fname


def create_table(Tfrz=0.137, output_path='.', eos_type='lattice_pce150', table_dir=None, eos=None,
                 copy_to=()):
    """Write ``chemical_potential.dat`` for freeze-out at ``Tfrz`` (GeV) into
    ``output_path`` and copy it into each directory of ``copy_to``
    (pyvisc/eos/chemical_potential.py:145-169).  e_frz is ``f_ed(Tfrz)`` of ``eos``, by
    default ``Eos(eos_type)``.  Returns the file name."""
    if eos is None:
        from .eos import Eos
        eos = Eos(eos_type)
    efrz = float(eos.f_ed(Tfrz))
    chem = ChemicalPotential(efrz, eos_type=eos_type, table_dir=table_dir)
    save_fname = chem.get_chemical_potential_for_resonance(output_path)
    for dest in copy_to:
        shutil.copy(save_fname, dest)
    return save_fname
