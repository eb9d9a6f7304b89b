"""The C-ABI library loads and exports every entry point include/clvisc_amd.h
declares, and the ctypes structs match the header's layout.  CPU only: no call
reaches the device."""
import ctypes as C
import os
import re

from clvisc_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, 'include', 'clvisc_amd.h')).read()
    src = re.sub(r'/\*.*?\*/', '', src, flags=re.S)
    return sorted(set(re.findall(r'^\s*(?:const\s+)?\w+\s*\*?\s*(cv_\w+)\s*\(', src, flags=re.M)))


def test_header_declarations_are_exported():
    names = declared()
    assert len(names) >= 25
    L = _lib.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(_lib.exported_symbols()) == names


def test_struct_layouts(tmp_path):
    """Offsets and sizes of cv_params / cv_eos as the C compiler lays them out from
    the header, against the ctypes mirrors."""
    import shutil
    import subprocess
    import pytest
    cc = shutil.which('gcc') or shutil.which('cc')
    if cc is None:
        pytest.skip('no C compiler')
    lines = []
    for cname, py in (('cv_params', _lib.CvParams), ('cv_eos', _lib.CvEos)):
        lines.append('printf("%s sizeof %%zu\\n", sizeof(%s));' % (cname, cname))
        for f, _ in py._fields_:
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, f))
    src = ('#include <stdio.h>\n#include <stddef.h>\n#include "clvisc_amd.h"\nint main(void){\n%s\nreturn 0;}\n'
           % '\n'.join(lines))
    (tmp_path / 'layout.c').write_text(src)
    exe = str(tmp_path / 'layout')
    subprocess.check_call([cc, '-I', os.path.join(ROOT, 'include'), str(tmp_path / 'layout.c'), '-o', exe])
    out = subprocess.check_output([exe]).decode().split('\n')
    got = {}
    for line in out:
        if line:
            a, b, v = line.split()
            got[(a, b)] = int(v)
    for cname, py in (('cv_params', _lib.CvParams), ('cv_eos', _lib.CvEos)):
        assert got[(cname, 'sizeof')] == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)


def test_errors_are_reported_without_a_device():
    """cv_create on a host without a GPU fails with a message, never silently."""
    import numpy as np
    import pytest
    from clvisc_amd.eos import Eos
    p = _lib.CvParams()
    p.nx = p.ny = p.nz = 4
    p.tau0, p.dt, p.dx, p.dy, p.dz = 0.6, 0.02, 0.1, 0.1, 0.1
    p.nxskip = p.nyskip = p.nzskip = 1
    bad = _lib.CvParams.from_buffer_copy(p)
    bad.dx = -1.0
    with pytest.raises(_lib.CvError, match='positive'):
        _lib.Handle(bad, Eos('ideal_gas').create_table())
