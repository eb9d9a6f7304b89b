"""Hydro configuration: ``hydro.info`` (ConfigParser) plus ``--KEY`` command-line
overrides, held in the module-level ``cfg`` -- the interface of the reference's
pyvisc/config.py:13-230 (``cfg``, ``read_config``, ``write_config``).

Keys and sections are the reference's.  Added keys: ZETAOS_MAX / ZETAOS_TPEAK /
ZETAOS_WIDTH (bulk viscosity), FP64_ROOT_VMIN (double-precision root solve in fast
cells), stage_kernel / tile_ty / tile_tz (device kernel selection), hypersf_fp64
(surface geometry in FP64).  ``use_float32``
must stay True: the device path is FP32.
"""
from __future__ import annotations

import argparse
import configparser
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def write_config(configs, comments=''):
    """Write the current setting to hydro.info in the output directory
    (pyvisc/config.py:13-66)."""
    fPathOut = configs.fPathOut
    if not os.path.exists(fPathOut):
        os.makedirs(fPathOut)
    configfile_name = os.path.join(fPathOut, 'hydro.info')
    with open(configfile_name, 'w') as cfgfile:
        C = configparser.ConfigParser()
        C.optionxform = str
        C.add_section('path')
        C.set('path', 'fPathIni', str(configs.fPathIni))
        C.set('path', 'fPathOut', str(configs.fPathOut))
        C.add_section('glauber')
        for k in ('Edmax', 'NumOfNucleons', 'SQRTS', 'NucleonDensity', 'Ra', 'Eta', 'Si0',
                  'ImpactParameter', 'Hwn', 'Eta_flat', 'Eta_gw'):
            C.set('glauber', k, str(getattr(configs, k)))
        C.add_section('geometry')
        for k in ('NX', 'NY', 'NZ', 'ntskip', 'nxskip', 'nyskip', 'nzskip', 'DT', 'DX', 'DY', 'DZ'):
            C.set('geometry', k, str(getattr(configs, k)))
        C.add_section('intrinsic')
        for k in ('TAU0', 'eos_type', 'TFRZ', 'ETAOS_XMIN', 'ETAOS_YMIN', 'ETAOS_LEFT_SLOP',
                  'ETAOS_RIGHT_SLOP', 'LAM1', 'BSZ', 'ZETAOS_MAX', 'ZETAOS_TPEAK', 'ZETAOS_WIDTH',
                  'FP64_ROOT_VMIN'):
            C.set('intrinsic', k, str(getattr(configs, k)))
        C.write(cfgfile)
        cfgfile.write('#comments: ' + comments)


def determine_float_size(args):
    args.sz_int = np.dtype('int32').itemsize
    if args.use_float32:
        args.real = np.float32
        args.sz_real = 4
        args.sz_real4 = 16
        args.sz_real8 = 32
    else:
        args.real = np.float64
        args.sz_real = 8
        args.sz_real4 = 32
        args.sz_real8 = 64
    return args


def read_config(path=None, argv=None):
    """Read the configuration from ``path`` (default: the package's hydro.info) and
    update it with ``--KEY value`` options (pyvisc/config.py:69-227)."""
    p = configparser.ConfigParser()
    p.optionxform = str
    p.read(path or os.path.join(HERE, 'hydro.info'))
    config = {}

    def g(sec, key, typ, help_, default=None):
        getter = {str: p.get, float: p.getfloat, int: p.getint}[typ]
        if default is not None and not p.has_option(sec, key):
            config[key] = (default, help_)      # keys the reference's hydro.info lacks
        else:
            config[key] = (getter(sec, key), help_)

    g('path', 'fPathIni', str, 'The absolute path for initial conditions')
    g('path', 'fPathOut', str, 'The absolute path for output directory')
    g('path', 'eos_table_dir', str, 'directory of the s95p EOS tables (the reference checkout\'s '
      'pyvisc/eos/eos_table); chemical_potential.dat is written when it is set', '')
    g('glauber', 'Edmax', float, 'maximum energy density for most central collisions')
    g('glauber', 'NumOfNucleons', float, 'Number of nucleons, A=197 for Au; A=208 for Pb')
    g('glauber', 'SQRTS', float, 'Beam energy in units of GeV/n')
    g('glauber', 'NucleonDensity', float, 'With which the woods-saxon integration = 197 for A=197')
    g('glauber', 'Ra', float, 'Radius of the nucleus')
    g('glauber', 'Eta', float, 'woods-saxon diffusiveness parameter')
    g('glauber', 'Si0', float, 'inelastic scattering cross section')
    g('glauber', 'ImpactParameter', float, 'average impact parameter')
    g('glauber', 'Hwn', float, 'in range [0,1), energy density contribution from number of wounded nucleons')
    g('glauber', 'Eta_flat', float, 'The width of the plateau along etas at mid rapidity')
    g('glauber', 'Eta_gw', float, 'the gaussian fall off at large etas where fabs(etas)>Eta_flat/2')
    g('geometry', 'NX', int, 'Grid size along x direction')
    g('geometry', 'NY', int, 'Grid size along y direction')
    g('geometry', 'NZ', int, 'Grid size along z direction')
    g('geometry', 'ntskip', int, 'Skip time steps for bulk information output')
    g('geometry', 'nxskip', int, 'Skip steps along x for bulk information output')
    g('geometry', 'nyskip', int, 'Skip steps along y for bulk information output')
    g('geometry', 'nzskip', int, 'Skip steps along z for bulk information output')
    g('geometry', 'DT', float, 'time step for hydro evolution')
    g('geometry', 'DX', float, 'x step for hydro evolution')
    g('geometry', 'DY', float, 'y step for hydro evolution')
    g('geometry', 'DZ', float, 'z step for hydro evolution')
    g('intrinsic', 'TAU0', float, 'time when hydro starts')
    g('intrinsic', 'eos_type', str, 'choices = ["crossover", "crossover_rowcol", "ideal_gas"]')
    g('intrinsic', 'TFRZ', float, 'Freeze out temperature, default=0.137')
    g('intrinsic', 'ETAOS_XMIN', float, 'temperature for minimum eta/s(T)')
    g('intrinsic', 'ETAOS_YMIN', float, 'minimum eta/s(T)')
    g('intrinsic', 'ETAOS_LEFT_SLOP', float, 'slop of eta/s(T) when T < ETAOS_XMIN')
    g('intrinsic', 'ETAOS_RIGHT_SLOP', float, 'slop of eta/s(T) when T > ETAOS_XMIN')
    g('intrinsic', 'LAM1', float, 'coefficient for pimn^2 term')
    g('intrinsic', 'ZETAOS_MAX', float, 'peak bulk zeta/s (0 disables bulk viscosity)', 0.0)
    g('intrinsic', 'ZETAOS_TPEAK', float, 'temperature of the zeta/s peak [GeV]', 0.155)
    g('intrinsic', 'ZETAOS_WIDTH', float, 'width of the zeta/s peak [GeV]', 0.02)
    g('intrinsic', 'FP64_ROOT_VMIN', float, 'FP64 root solve where M > FP64_ROOT_VMIN T00 (0 disables)', 0.0)
    # the reference reads BSZ from [opencl] and writes it to [intrinsic] (config.py:51)
    if p.has_option('opencl', 'local_workgroup_size'):
        bsz = p.getint('opencl', 'local_workgroup_size')
    elif p.has_option('intrinsic', 'BSZ'):
        bsz = p.getint('intrinsic', 'BSZ')
    else:
        bsz = 64
    config['BSZ'] = (bsz, 'Local workgroup size in one dimension (kept for compatibility)')

    parser = argparse.ArgumentParser(description='Input parameters for hydrodynamic simulations',
                                     allow_abbrev=False)
    for key, value in list(config.items()):
        parser.add_argument('--{key}'.format(key=key), nargs='?', const=1,
                            type=type(value[0]), default=value[0], help=value[1])
    for flag, help_ in (('riemann_test', 'switch on riemann test (Christoffel terms removed)'),
                        ('gubser_visc_test', '2nd order viscous gubser test'),
                        ('pimn_omega_coupling', 'pi^{mu nu} - vorticity coupling in the IS source'),
                        ('omega_omega_coupling', 'vorticity - vorticity coupling in the IS source'),
                        ('save_to_hdf5', 'accepted for compatibility; bulk info is written as text'),
                        ('opencl_interactive', 'accepted for compatibility (no effect)'),
                        ('hypersf_fp64', 'freeze-out surface geometry in FP64 instead of the FP32 of the reference')):
        parser.add_argument('--' + flag, nargs='?', const=1, type=bool, default=False, help=help_)
    parser.add_argument('--use_float32', nargs='?', const=1, type=bool, default=True,
                        help='must be True: the device path is FP32')
    parser.add_argument('--stage_kernel', type=str, default='march', choices=['march', 'tiled', 'percell'],
                        help="step kernel: 'march' (tiled, x-marching), 'tiled' or 'percell'")
    parser.add_argument('--tile_ty', type=int, default=0, help='tile rows (0 = automatic)')
    parser.add_argument('--tile_tz', type=int, default=0, help='tile columns (0 = automatic)')
    args, unknown = parser.parse_known_args(argv)
    return determine_float_size(args)


cfg = read_config()
