"""Writes the chemical-potential fixtures under tests/golden/chem/ from the reference's
s95p EOS table directories (pyvisc/eos/eos_table/s95p-PCE165-v0 and s95p-PCE-v1).

For each directory: particles.dat and pdg05.dat as they are (pdg05.dat is the same file in
both, so the second is a link to the first), and the <dir>_pichem1.dat table cut to the
rows of energy densities ROWS (same header format: e0, "de ne", nstable; rows from the
highest e down, each row's text unchanged).  From s95p-PCE165-v0 also ChemForReso.dat, the
resonance potentials that directory's own script wrote at e_frz = 0.144 GeV/fm^3.

Run from the repository root in a checkout that has the reference at /root/reference:
    python tests/golden/make_chem_fixture.py
"""
import os

SRC = '/root/reference/pyvisc/eos/eos_table'
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'chem')
ROWS = range(60, 91)          # e = 0.121 .. 0.181 GeV/fm^3 (freeze-out densities)


def cut_pichem(src, dst):
    with open(src) as f:
        lines = f.readlines()
    e0 = float(lines[0])
    de, ne = lines[1].split()
    de, ne = float(de), int(ne)
    data = lines[3:3 + ne]              # data[k] is row ne-1-k (highest e first)
    lo, hi = ROWS[0], ROWS[-1]
    with open(dst, 'w') as f:
        f.write('%r\n' % (e0 + lo * de))
        f.write('%r %d\n' % (de, hi - lo + 1))
        f.write(lines[2])
        f.writelines(data[ne - 1 - hi:ne - lo])


def main():
    for name in ('s95p-PCE165-v0', 's95p-PCE-v1'):
        src, dst = os.path.join(SRC, name), os.path.join(DST, name)
        os.makedirs(dst, exist_ok=True)
        with open(os.path.join(src, 'particles.dat')) as f, open(os.path.join(dst, 'particles.dat'), 'w') as g:
            g.write(f.read())
        cut_pichem(os.path.join(src, name + '_pichem1.dat'), os.path.join(dst, name + '_pichem1.dat'))
        pdg = os.path.join(dst, 'pdg05.dat')
        if name == 's95p-PCE165-v0':
            with open(os.path.join(src, 'pdg05.dat')) as f, open(pdg, 'w') as g:
                g.write(f.read())
            with open(os.path.join(src, 'ChemForReso.dat')) as f, open(os.path.join(dst, 'ChemForReso.dat'), 'w') as g:
                g.write(f.read())
        else:
            with open(os.path.join(src, 'pdg05.dat')) as f, open(os.path.join(DST, 's95p-PCE165-v0', 'pdg05.dat')) as g:
                assert f.read() == g.read()
            if os.path.lexists(pdg):
                os.remove(pdg)
            os.symlink(os.path.join('..', 's95p-PCE165-v0', 'pdg05.dat'), pdg)


if __name__ == '__main__':
    main()
