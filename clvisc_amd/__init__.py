"""clvisc_amd: the CLVisc (3+1)D KT + Israel-Stewart hydro step on MI355X (gfx950).

Drivers with the reference's interface (pyvisc/ideal.py CLIdeal, pyvisc/visc.py
CLVisc) over the HIP library libclvisc_amd.so (C-ABI: include/clvisc_amd.h).
The configuration module (``clvisc_amd.config``) parses the command line at import,
as the reference's ``config`` does, so it is imported on demand here.
"""
from .eos import Eos, EosTable
from .ideal import CLIdeal, DeviceField
from .visc import CLVisc
from .bulkinfo import BulkInfo
from .pimninfo import PimnInfo
from .hypersf import HyperSurface

__all__ = ['Eos', 'EosTable', 'CLIdeal', 'CLVisc', 'DeviceField', 'BulkInfo', 'PimnInfo', 'HyperSurface']
