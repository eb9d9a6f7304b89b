"""HIP path vs the FP32 oracle on the same seeded inputs.

The HIP kernels and the oracle evaluate the same FP32 operation sequence, so the
fields are expected to be identical bit for bit; the assertion is the parity bar
(max relative difference <= 1e-4, floored at 1e-3 of each field's maximum) and the
bit-identical fraction is reported.
"""
import sys

import numpy as np
import pytest

from clvisc_amd import _lib, ic
from clvisc_amd.eos import Eos
from oracle import oracle as O
from helpers import cv_params, orc_params, max_rel, bit_equal_frac

pytestmark = pytest.mark.gpu
TOL = 1e-4


def _say(msg):
    print(msg)
    sys.stdout.flush()


def _ideal_gubser(nx=33, ny=35, q=1.0):
    eos = Eos('ideal_gas')
    tab = eos.create_table()
    p = cv_params(nx, ny, 1, 0.02, 0.08, 0.08, 0.15, 1.0)
    ev, _ = ic.gubser_ideal(nx, ny, 1, 0.08, 0.08, 0.15, 1.0, q=q, e0=float(eos.f_ed(0.3)))
    return p, tab, ev.astype(np.float32)


def test_ideal_stage_bitwise():
    p, tab, ev = _ideal_gubser()
    _say('case: ideal stage gubser %dx%dx1' % (p.nx, p.ny))
    h = _lib.Handle(p, tab, viscous=False)
    h.set('ev1', ev)
    o = O.Oracle(orc_params(p, tab), 'f32', viscous=False)
    o.load_ini(ev)
    for step, tau in ((1, np.float32(1.0)), (2, np.float32(1.02))):
        h.ideal_stage(step, tau)
        o.ideal_stage(step, tau)
        out = 'ev2' if step == 1 else 'ev1'
        g, r = h.get(out), o.ev[3 - step]
        assert np.all(np.isfinite(g))
        m = max_rel(g, r)
        _say('  step %d max rel %s bit-equal %.6f' % (step, m, bit_equal_frac(g, r)))
        assert m.max() <= TOL
        np.testing.assert_array_equal(h.get('src'), o.Src)


def test_ideal_steps_gubser_201():
    eos = Eos('ideal_gas')
    tab = eos.create_table()
    p = cv_params(201, 201, 1, 0.02, 0.08, 0.08, 0.15, 1.0)
    ev, _ = ic.gubser_ideal(201, 201, 1, 0.08, 0.08, 0.15, 1.0, q=1.0, e0=float(eos.f_ed(0.3)))
    ev = ev.astype(np.float32)
    _say('case: ideal gubser 201x201x1, 50 steps')
    h = _lib.Handle(p, tab, viscous=False)
    h.set('ev1', ev)
    h.ideal_steps(0, 50)
    o = O.Oracle(orc_params(p, tab, nthreads=16), 'f32', viscous=False)
    o.load_ini(ev)
    o.ideal_steps(50)
    g = h.get('ev1')
    m = max_rel(g, o.ev[1])
    _say('  max rel %s bit-equal %.6f' % (m, bit_equal_frac(g, o.ev[1])))
    assert m.max() <= TOL
    assert abs(h.max_ed() - o.max_ed()) == 0.0


@pytest.mark.parametrize('eos_kind,dims,kernel,chunk', [
    ('crossover', (29, 31, 11), (2, 0, 0), None), ('crossover', (29, 31, 11), (2, 7, 30), 3),
    ('crossover', (29, 31, 11), (2, 11, 20), 5), ('ideal_gas', (27, 25, 9), (2, 0, 0), None),
    ('ideal_gas', (27, 25, 9), (0, 0, 0), None), ('crossover', (45, 41, 1), (2, 64, 1), 4),
    ('crossover', (29, 31, 11), (2, 0, 0), 1)])
def test_ideal_steps_tiled_3d(eos_kind, dims, kernel, chunk):
    """The ideal stage on x-marching tiles (kind 2; kind 0 is the per-cell kernel) against
    the oracle's ideal steps: identical fields."""
    nx, ny, nz = dims
    eos = Eos(eos_kind)
    tab = eos.create_table()
    p = cv_params(nx, ny, nz, 0.02, 0.08, 0.08, 0.15, 0.6)
    pos = ic.woods_saxon_positions(12, 5, R=0.6, a=0.1)
    ev = ic.hotspots(nx, ny, nz, 0.08, 0.08, 0.15, positions=pos, emax=30.0)
    _say('case: ideal hotspots %s %dx%dx%d, 25 steps, kernel %s chunk %s' % (eos_kind, nx, ny, nz, kernel, chunk))
    h = _lib.Handle(p, tab, viscous=False)
    h.set_stage_kernel(*kernel)
    if chunk is not None:
        h.set_march_chunk(chunk)
    h.set('ev1', ev)
    h.ideal_steps(0, 25)
    o = O.Oracle(orc_params(p, tab, nthreads=16), 'f32', viscous=False)
    o.load_ini(ev)
    o.ideal_steps(25)
    for name, g, r in (('ev1', h.get('ev1'), o.ev[1]), ('ev2', h.get('ev2'), o.ev[2]), ('src', h.get('src'), o.Src)):
        _say('  %-4s bit-equal %.8f' % (name, bit_equal_frac(g, r)))
        np.testing.assert_array_equal(g, r, err_msg=name)
    h.close()


@pytest.mark.parametrize('viscous,kernel,cs2_map', [(False, (0, 0, 0), 'scale'), (False, (2, 0, 0), 'scale'),
                                                    (True, (0, 0, 0), 'scale'), (True, (1, 0, 0), 'scale'),
                                                    (True, (2, 0, 0), 'scale'), (True, (2, 0, 0), 'shift'),
                                                    (False, (2, 0, 0), 'shift')])
def test_steps_with_the_full_lambda_formula(viscous, kernel, cs2_map):
    """A crossover table whose cs2 column is scaled past 0.5 (1.6x) or shifted below 0 (-0.2;
    only lambda reads it) turns off the device's max(lambda_KT, 1) = 1 shortcut, so every face
    evaluates the reference's formula (gamma, cs2 lookup, square root, division), as the
    oracle always does, negative cs2 (a NaN square root, or lambda_KT above 1) included:
    identical fields."""
    import dataclasses
    tab0 = Eos('crossover').create_table()
    rgba = tab0.rgba.copy()
    if cs2_map == 'scale':
        rgba[:, 0] *= np.float32(1.6)
        assert rgba[:, 0].max() > 0.5
    else:
        rgba[:, 0] -= np.float32(0.2)
        assert rgba[:, 0].min() < 0.0
    tab = dataclasses.replace(tab0, rgba=rgba, cs2_max=float(rgba[:, 0].max()), cs2_min=float(rgba[:, 0].min()))
    nx, ny, nz = 29, 31, 11
    p = cv_params(nx, ny, nz, 0.02, 0.08, 0.08, 0.15, 0.6, etaos_xmin=0.0, etaos_ymin=0.08,
                  bulk=(0.05, 0.155, 0.02) if viscous else None)
    pos = ic.woods_saxon_positions(12, 3, R=0.6, a=0.1)
    ev = ic.hotspots(nx, ny, nz, 0.08, 0.08, 0.15, positions=pos, emax=50.0)
    if viscous:
        _visc_compare(p, tab, ev, 20, 'viscous + bulk, full lambda formula', kernel)
        return
    _say('case: ideal hotspots %dx%dx%d, full lambda formula, 25 steps, kernel %s' % (nx, ny, nz, kernel))
    h = _lib.Handle(p, tab, viscous=False)
    h.set_stage_kernel(*kernel)
    h.set('ev1', ev)
    h.ideal_steps(0, 25)
    o = O.Oracle(orc_params(p, tab, nthreads=16), 'f32', viscous=False)
    o.load_ini(ev)
    o.ideal_steps(25)
    for name, g, r in (('ev1', h.get('ev1'), o.ev[1]), ('ev2', h.get('ev2'), o.ev[2])):
        _say('  %-4s bit-equal %.8f' % (name, bit_equal_frac(g, r)))
        np.testing.assert_array_equal(g, r, err_msg=name)
    h.close()


@pytest.mark.parametrize('kernel', [(0, 0, 0), (2, 0, 0)])
def test_fast_periphery_matches_oracle(kernel):
    """A small hot blob expanding into a wide vacuum margin for 100 steps (viscous + bulk):
    its periphery moves at |v| > 0.9, where the device's max(lambda_KT, 1) = 1 shortcut
    applies on the face component alone (0.81 < vk^2 <= 0.98, outside the |v|^2 <= 0.81 of
    the earlier criterion).  Identical fields to the oracle, which evaluates the formula at
    every face.  (Faces with vk^2 > 0.98 are rare in such flows; the scaled-cs2 case above
    runs the formula at every face.)"""
    nx, ny, nz = 101, 101, 9
    p = cv_params(nx, ny, nz, 0.02, 0.08, 0.08, 0.15, 0.6, etaos_xmin=0.0, etaos_ymin=0.08,
                  bulk=(0.05, 0.155, 0.02))
    tab = Eos('crossover').create_table()
    pos = ic.woods_saxon_positions(5, 7, R=0.3, a=0.05)
    ev = ic.hotspots(nx, ny, nz, 0.08, 0.08, 0.15, positions=pos, emax=50.0)
    h = _lib.Handle(p, tab, viscous=True)
    h.set_stage_kernel(*kernel)
    h.visc_initialize(p.tau0)
    h.ideal_stage(1, np.float32(p.tau0))
    h.set('ev1', ev)
    h.visc_steps(0, 100)
    vk2 = h.get('ev1')[:, 1:] ** 2
    h.close()
    mid = ((vk2 > 0.81) & (vk2 <= 0.98)).any(1).mean()
    _say('case: fast periphery, cells with a component 0.81 < vk^2 <= 0.98 after 100 steps: %.4f' % mid)
    assert mid > 0.05
    _visc_compare(p, tab, ev, 100, 'fast periphery', kernel)


@pytest.mark.parametrize('dims', [(33, 1, 17), (1, 29, 13), (1, 1, 41), (27, 25, 2), (2, 3, 9)])
@pytest.mark.parametrize('viscous', [True, False])
def test_steps_thin_grids(dims, viscous):
    """Grids with one (or two) cells along an axis, and the 1x1xN longitudinal expansion of
    pyvisc/tests/test_longitudinal_expansion.py: the default kernels against the oracle."""
    nx, ny, nz = dims
    eos = Eos('crossover')
    tab = eos.create_table()
    p = cv_params(nx, ny, nz, 0.02, 0.08, 0.08, 0.15, 0.6, etaos_xmin=0.0, etaos_ymin=0.08,
                  bulk=(0.05, 0.155, 0.02) if viscous else None)
    x, y, z = ic.grid_coords(nx, ny, nz, 0.08, 0.08, 0.15)
    X, Y, Z = np.meshgrid(x, y, z, indexing='ij')
    ev = np.zeros((nx, ny, nz, 4), np.float32)
    ev[..., 0] = 20.0 * np.exp(-(X ** 2 + Y ** 2) / 0.5 - Z ** 2 / 2.0) + 0.01
    ev = ev.reshape(-1, 4)
    _say('case: thin grid %dx%dx%d %s, 15 steps' % (nx, ny, nz, 'viscous+bulk' if viscous else 'ideal'))
    h = _lib.Handle(p, tab, viscous=viscous)
    _say('  stage kernel %s march chunk %s' % (h.stage_kernel(), h.march_chunk()))
    o = O.Oracle(orc_params(p, tab, nthreads=8), 'f32', viscous=viscous)
    o.load_ini(ev)
    if viscous:
        h.visc_initialize(p.tau0)
        h.ideal_stage(1, np.float32(p.tau0))
        h.set('ev1', ev)
        h.visc_steps(0, 15)
        o.visc_steps(15)
        pairs = (('ev1', o.ev[1]), ('pi1', o.pi[1]), ('bulk1', o.bulk[1]), ('udiff', o.udiff))
    else:
        h.set('ev1', ev)
        h.ideal_steps(0, 15)
        o.ideal_steps(15)
        pairs = (('ev1', o.ev[1]), ('src', o.Src))
    for name, r in pairs:
        g = h.get(name)
        _say('  %-6s bit-equal %.8f' % (name, bit_equal_frac(g, r)))
        np.testing.assert_array_equal(g, r, err_msg=name)
    h.close()


def test_large_grid_tiled_equals_per_cell():
    """481x481x301 = 69.6M cells, near the 32-bit plane-offset limit of the tiled kernels
    (40 n < 2^32): the default marching kernel and the per-cell kernel (64-bit indexing),
    both checked against the oracle on small grids, give identical fields after 2 viscous +
    bulk steps and after 2 ideal steps."""
    nx, ny, nz = 481, 481, 301
    eos = Eos('crossover')
    tab = eos.create_table()
    p = cv_params(nx, ny, nz, 0.02, 0.08, 0.08, 0.1, 0.6, etaos_xmin=0.0, etaos_ymin=0.08, bulk=(0.05, 0.155, 0.02))
    pos = ic.woods_saxon_positions(200, 17)
    ev = ic.hotspots(nx, ny, nz, 0.08, 0.08, 0.1, positions=pos, emax=40.0)
    _say('case: large grid %dx%dx%d (%d cells), tiled vs per-cell, 2 steps' % (nx, ny, nz, nx * ny * nz))
    pi_ = _lib.CvParams.from_buffer_copy(p)
    pi_.flags &= ~_lib.CV_FLAG_BULK                       # an ideal handle takes no bulk
    for viscous in (True, False):
        out = []
        for kind in (2, 0):
            h = _lib.Handle(p if viscous else pi_, tab, viscous=viscous)
            h.set_stage_kernel(kind, 0, 0)
            if viscous:
                h.visc_initialize(p.tau0)
                h.ideal_stage(1, np.float32(p.tau0))
                h.set('ev1', ev)
                h.visc_steps(0, 2)
                out.append((h.get('ev1'), h.get('pi1'), h.get('bulk1')))
            else:
                h.set('ev1', ev)
                h.ideal_steps(0, 2)
                out.append((h.get('ev1'),))
            h.close()
        for a, b in zip(*out):
            _say('  %s bit-equal %.8f' % ('viscous' if viscous else 'ideal', bit_equal_frac(a, b)))
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize('viscous', [False, True])
def test_riemann_mode_matches_oracle(viscous):
    """CV_FLAG_RIEMANN_TEST (Cartesian mode without the Christoffel terms, pyvisc/ideal.py:164-165;
    the per-cell ideal kernel and the tiled viscous kernel's test-branch instance): a Riemann
    step along x against the oracle in the same mode, identical fields."""
    nx, ny, nz = 41, 9, 7
    eos = Eos('ideal_gas')
    tab = eos.create_table()
    p = cv_params(nx, ny, nz, 0.01, 0.08, 0.08, 0.08, 1.0, etaos_xmin=0.0, etaos_ymin=0.08,
                  flags=_lib.CV_FLAG_RIEMANN_TEST)
    ev = np.zeros((nx, ny, nz, 4), np.float32)
    ev[: nx // 2, ..., 0] = 10.0
    ev[nx // 2:, ..., 0] = 1.0
    ev = ev.reshape(-1, 4)
    _say('case: riemann mode %dx%dx%d %s, 20 steps' % (nx, ny, nz, 'viscous' if viscous else 'ideal'))
    h = _lib.Handle(p, tab, viscous=viscous)
    o = O.Oracle(orc_params(p, tab, nthreads=8), 'f32', viscous=viscous)
    o.load_ini(ev)
    if viscous:
        h.visc_initialize(p.tau0)
        h.ideal_stage(1, np.float32(p.tau0))
        h.set('ev1', ev)
        h.visc_steps(0, 20)
        o.visc_steps(20)
        pairs = (('ev1', o.ev[1]), ('pi1', o.pi[1]))
    else:
        h.set('ev1', ev)
        h.ideal_steps(0, 20)
        o.ideal_steps(20)
        pairs = (('ev1', o.ev[1]), ('src', o.Src))
    for name, r in pairs:
        g = h.get(name)
        _say('  %-4s bit-equal %.8f' % (name, bit_equal_frac(g, r)))
        np.testing.assert_array_equal(g, r, err_msg=name)
    h.close()


@pytest.mark.parametrize('viscous', [False, True])
def test_riemann_rarefaction_along_eta_matches_oracle(viscous):
    """The Riemann problem of pyvisc/tests/test_riemann.py on its grid (NX = NY = 1, NZ = 401,
    DZ = 0.05, DT = 0.01, the step along the third axis, vacuum on the right), 400 steps:
    the device equals the oracle, which follows the analytic rarefaction
    (tests/test_riemann_oracle.py).  Viscous: eta/s = 0, through the viscous kernels."""
    nz = 401
    tab = Eos('ideal_gas').create_table()
    p = cv_params(1, 1, nz, 0.01, 0.08, 0.08, 0.05, 1.0, etaos_xmin=0.0, etaos_ymin=0.0,
                  flags=_lib.CV_FLAG_RIEMANN_TEST)
    z = (np.arange(nz) - nz // 2) * 0.05
    ev = np.zeros((nz, 4), np.float32)
    ev[:, 0] = np.where(z < 0, 3.0, 0.0)
    _say('case: riemann rarefaction 1x1x%d %s, 400 steps' % (nz, 'viscous' if viscous else 'ideal'))
    h = _lib.Handle(p, tab, viscous=viscous)
    o = O.Oracle(orc_params(p, tab), 'f32', viscous=viscous)
    o.load_ini(ev)
    if viscous:
        h.visc_initialize(p.tau0)
        h.ideal_stage(1, np.float32(p.tau0))
        h.set('ev1', ev)
        h.visc_steps(0, 400)
        o.visc_steps(400)
        pairs = (('ev1', o.ev[1]), ('pi1', o.pi[1]))
    else:
        h.set('ev1', ev)
        h.ideal_steps(0, 400)
        o.ideal_steps(400)
        pairs = (('ev1', o.ev[1]),)
    for name, r in pairs:
        g = h.get(name)
        _say('  %-4s bit-equal %.8f' % (name, bit_equal_frac(g, r)))
        np.testing.assert_array_equal(g, r, err_msg=name)
    h.close()


def test_longitudinal_expansion_matches_oracle():
    """The workload of pyvisc/tests/test_longitudinal_expansion.py: NX = NY = 1, NZ = 401,
    DZ = 0.075, DT = 0.005, e = 30 exp(-eta^2 / (2 * 2^2)) at rest, ideal, 1400 steps (the
    reference's test only saves the profiles): device equal to the oracle, and the profile
    stays even in eta with tau v^eta odd."""
    nz, nsteps = 401, 1400
    tab = Eos('ideal_gas').create_table()
    p = cv_params(1, 1, nz, 0.005, 0.08, 0.08, 0.075, 0.6)
    eta = (np.arange(nz) - nz // 2) * np.float32(0.075)
    ev = np.zeros((nz, 4), np.float32)
    ev[:, 0] = 30.0 * np.exp(-eta.astype(np.float64) ** 2 / (2.0 * 2.0 * 2.0))
    _say('case: longitudinal expansion 1x1x%d, %d steps' % (nz, nsteps))
    h = _lib.Handle(p, tab, viscous=False)
    h.set('ev1', ev)
    h.ideal_steps(0, nsteps)
    o = O.Oracle(orc_params(p, tab), 'f32', viscous=False)
    o.load_ini(ev)
    o.ideal_steps(nsteps)
    g = h.get('ev1')
    _say('  ev1 bit-equal %.8f; e(0) %.4g -> %.4g' % (bit_equal_frac(g, o.ev[1]), ev[nz // 2, 0], g[nz // 2, 0]))
    np.testing.assert_array_equal(g, o.ev[1])
    np.testing.assert_array_equal(g[::-1, 0], g[:, 0])
    np.testing.assert_array_equal(-g[::-1, 3], g[:, 3])
    h.close()


def _visc_case(nx, ny, nz, bulk=None, nhot=12, seed=3, eos_kind='crossover'):
    eos = Eos(eos_kind)
    tab = eos.create_table()
    p = cv_params(nx, ny, nz, 0.02, 0.08, 0.08, 0.15, 0.6, etaos_xmin=0.0, etaos_ymin=0.08, bulk=bulk)
    pos = ic.woods_saxon_positions(nhot, seed, R=0.6, a=0.1)
    ev = ic.hotspots(nx, ny, nz, 0.08, 0.08, 0.15, positions=pos, emax=50.0)
    return p, tab, ev


def _visc_compare(p, tab, ev, nsteps, label, kernel=None, chunk=None, xcd=None):
    _say('case: %s %dx%dx%d, %d steps, kernel %s chunk %s xcd map %s'
         % (label, p.nx, p.ny, p.nz, nsteps, kernel, chunk, xcd))
    h = _lib.Handle(p, tab, viscous=True)
    if xcd is not None:
        h.set_xcd_map(xcd)
    if kernel is not None:
        h.set_stage_kernel(*kernel)
    if chunk is not None:
        h.set_march_chunk(chunk)
    _say('  stage kernel %s march chunk %s' % (h.stage_kernel(), h.march_chunk()))
    h.visc_initialize(p.tau0)
    h.ideal_stage(1, np.float32(p.tau0))
    h.set('ev1', ev)
    h.visc_steps(0, nsteps)
    o = O.Oracle(orc_params(p, tab, nthreads=16), 'f32', viscous=True)
    o.load_ini(ev)
    o.visc_steps(nsteps)
    res = {}
    for name, g, r in (('ev1', h.get('ev1'), o.ev[1]), ('pi1', h.get('pi1'), o.pi[1]),
                       ('ev2', h.get('ev2'), o.ev[2]), ('pi2', h.get('pi2'), o.pi[2]),
                       ('ev0', h.get('ev0'), o.ev[0]), ('udiff', h.get('udiff'), o.udiff),
                       ('goodcell', h.get('goodcell'), o.goodcell)):
        assert np.all(np.isfinite(g)), name
        m = max_rel(g, r)
        res[name] = m.max()
        _say('  %-8s max rel %.3g bit-equal %.6f' % (name, m.max(), bit_equal_frac(g, r)))
    if p.flags & _lib.CV_FLAG_BULK:
        g, r = h.get('bulk1'), o.bulk[1]
        m = max_rel(g, r)
        res['bulk1'] = m.max()
        _say('  %-8s max rel %.3g bit-equal %.6f' % ('bulk1', m.max(), bit_equal_frac(g, r)))
    for k, v in res.items():
        assert v <= TOL, (k, v)


@pytest.mark.parametrize('kernel,chunk', [((0, 0, 0), None), ((1, 0, 0), None), ((1, 8, 27), None),
                                          ((1, 11, 20), None), ((1, 4, 16), None), ((1, 14, 16), None),
                                          ((2, 0, 0), None), ((2, 7, 30), 1), ((2, 11, 20), 3),
                                          ((2, 4, 16), 29), ((2, 14, 16), 7)])
def test_visc_steps_small_3d(kernel, chunk):
    p, tab, ev = _visc_case(29, 31, 11)
    _visc_compare(p, tab, ev, 20, 'viscous hotspots', kernel, chunk)


@pytest.mark.parametrize('kernel,chunk', [((0, 0, 0), None), ((1, 0, 0), None), ((1, 64, 1), None),
                                          ((2, 0, 0), None), ((2, 64, 1), 4)])
def test_visc_steps_2d_boost_invariant(kernel, chunk):
    p, tab, ev = _visc_case(45, 41, 1, bulk=(0.05, 0.155, 0.02))
    _visc_compare(p, tab, ev, 15, 'viscous+bulk 2d', kernel, chunk)


@pytest.mark.parametrize('kernel,chunk,xcd', [((1, 0, 0), None, 0), ((1, 8, 27), None, 3),
                                              ((2, 0, 0), None, 0), ((2, 7, 30), 3, 3), ((2, 11, 20), 5, 16),
                                              ((2, 0, 0), 1, 8), ((2, 0, 0), 4, 8), ((2, 0, 0), 2, 5)])
def test_visc_steps_xcd_map(kernel, chunk, xcd):
    """The tiled kernels' workgroup -> work-item mapping (off, and XCD counts that do not
    divide the work) changes only the schedule: the oracle's fields bit for bit."""
    p, tab, ev = _visc_case(29, 31, 11, bulk=(0.05, 0.155, 0.02))
    _visc_compare(p, tab, ev, 20, 'viscous+bulk hotspots', kernel, chunk, xcd)


@pytest.mark.parametrize('viscous', [False, True])
def test_automatic_slab_chunks_with_a_tail(viscous):
    """Automatic march chunks on a grid with many tile columns (192x101x151, 4x16 tiles: 260
    columns): each XCD's slab of 24 planes is cut into two 11-plane chunks and a 2-plane tail
    chunk last.  Identical fields to the oracle, ideal and viscous + bulk."""
    nx, ny, nz = 192, 101, 151
    p = cv_params(nx, ny, nz, 0.02, 0.08, 0.08, 0.15, 0.6, etaos_xmin=0.0, etaos_ymin=0.08,
                  bulk=(0.05, 0.155, 0.02) if viscous else None)
    tab = Eos('crossover').create_table()
    pos = ic.woods_saxon_positions(40, 11, R=4.0, a=0.5)
    ev = ic.hotspots(nx, ny, nz, 0.08, 0.08, 0.15, positions=pos, emax=50.0)
    if viscous:
        _visc_compare(p, tab, ev, 4, 'viscous + bulk, automatic slab chunks', (2, 4, 16))
        return
    h = _lib.Handle(p, tab, viscous=False)
    h.set_stage_kernel(2, 4, 16)
    _say('case: ideal %dx%dx%d, automatic slab chunks %s, 8 steps' % (nx, ny, nz, h.march_chunk()))
    assert h.march_chunk() == (11, 24)
    h.set('ev1', ev)
    h.ideal_steps(0, 8)
    o = O.Oracle(orc_params(p, tab, nthreads=16), 'f32', viscous=False)
    o.load_ini(ev)
    o.ideal_steps(8)
    for name, g, r in (('ev1', h.get('ev1'), o.ev[1]), ('ev2', h.get('ev2'), o.ev[2])):
        _say('  %-4s bit-equal %.8f' % (name, bit_equal_frac(g, r)))
        np.testing.assert_array_equal(g, r, err_msg=name)
    h.close()


@pytest.mark.parametrize('kernel', [(0, 0, 0), (2, 0, 0)])
def test_visc_steps_fp64_root(kernel):
    """CV_FLAG_FP64_ROOT with a low threshold (M > 0.3 T^tt) so that many cells take the
    double-precision root solve: identical to the oracle's FP64 branch."""
    p, tab, ev = _visc_case(27, 25, 9, bulk=(0.05, 0.155, 0.02))
    p.flags |= _lib.CV_FLAG_FP64_ROOT
    p.fp64_root_vmin = 0.3
    _visc_compare(p, tab, ev, 20, 'viscous+bulk, FP64 root for M > 0.3 T00', kernel)


def test_visc_steps_bulk_3d():
    p, tab, ev = _visc_case(27, 25, 9, bulk=(0.05, 0.155, 0.02))
    _visc_compare(p, tab, ev, 20, 'viscous+bulk hotspots')


@pytest.mark.parametrize('kernel', [(0, 0, 0), (1, 0, 0), (2, 0, 0)])
@pytest.mark.parametrize('flags', [_lib.CV_FLAG_PIMN_OMEGA, _lib.CV_FLAG_OMEGA_OMEGA,
                                   _lib.CV_FLAG_PIMN_OMEGA | _lib.CV_FLAG_OMEGA_OMEGA])
def test_visc_steps_vorticity_couplings(kernel, flags):
    """The IS vorticity couplings (pimn_omega_coupling / omega_omega_coupling, kernel_IS.cl:636-683)
    on a viscous + bulk hotspot event: the per-cell kernel and the tiled kernels' test-branch
    instance equal the oracle with the same terms bit for bit, and the terms move pi."""
    p, tab, ev = _visc_case(29, 31, 11, bulk=(0.05, 0.155, 0.02))
    p.flags |= flags
    _say('case: vorticity couplings 0x%x, 29x31x11 viscous+bulk, 20 steps, kernel %s' % (flags, kernel))
    h = _lib.Handle(p, tab, viscous=True)
    h.set_stage_kernel(*kernel)
    h.visc_initialize(p.tau0)
    h.ideal_stage(1, np.float32(p.tau0))
    h.set('ev1', ev)
    h.visc_steps(0, 20)
    o = O.Oracle(orc_params(p, tab, nthreads=16), 'f32', viscous=True)
    o.load_ini(ev)
    o.visc_steps(20)
    for name, r in (('ev1', o.ev[1]), ('pi1', o.pi[1]), ('bulk1', o.bulk[1]), ('udiff', o.udiff)):
        g = h.get(name)
        _say('  %-6s bit-equal %.8f' % (name, bit_equal_frac(g, r)))
        np.testing.assert_array_equal(g, r, err_msg=name)
    h.close()
    p0 = _lib.CvParams.from_buffer_copy(p)
    p0.flags &= ~(_lib.CV_FLAG_PIMN_OMEGA | _lib.CV_FLAG_OMEGA_OMEGA)
    o0 = O.Oracle(orc_params(p0, tab, nthreads=16), 'f32', viscous=True)
    o0.load_ini(ev)
    o0.visc_steps(20)
    d = np.abs(o.pi[1].astype(np.float64) - o0.pi[1]).max() / np.abs(o0.pi[1]).max()
    _say('  max |pi(couplings) - pi(none)| / max|pi| = %.3g' % d)
    assert d > 1e-6


@pytest.mark.parametrize('viscous', [False, True])
@pytest.mark.parametrize('kind,stages', [(0, 2), (2, 2), (0, 3)])
def test_jet_source_matches_oracle(viscous, kind, stages):
    """jet_eloss_src (kernel_jet_eloss.cl:8-69) in the RK stages of the mask (2: stage 2 only, as
    evolve passes it, ideal.py:344, visc.py:543; 3: both): ideal and viscous + bulk steps (the
    x-marching handle runs jet stages per cell) equal the oracle with the same deposit bit for
    bit, and the deposit moves the fields."""
    nx, ny, nz, nsteps = 29, 31, 11, 15
    p, tab, ev = _visc_case(nx, ny, nz, bulk=(0.05, 0.155, 0.02) if viscous else None)
    direction, start = np.float32(0.7), (14 * ny + 15) * nz + 5
    _say('case: jet source %s, kind %d, stages %d, %dx%dx%d, %d steps, start %d direction %.2f'
         % ('viscous+bulk' if viscous else 'ideal', kind, stages, nx, ny, nz, nsteps, start, direction))
    if not viscous:
        p.flags &= ~_lib.CV_FLAG_BULK
    h = _lib.Handle(p, tab, viscous=viscous)
    h.set_stage_kernel(kind, 0, 0)
    jet = {'switch_on': True, 'start_pos_index': start, 'direction': direction}
    if viscous:
        h.visc_initialize(p.tau0)
        h.ideal_stage(1, np.float32(p.tau0))
    h.set('ev1', ev)
    h.set_jet_source(jet, stages=stages)
    (h.visc_steps if viscous else h.ideal_steps)(0, nsteps)
    prm = orc_params(p, tab, nthreads=16)
    prm.jet_on, prm.jet_direction, prm.jet_start = stages, float(direction), start
    o = O.Oracle(prm, 'f32', viscous=viscous)
    o.load_ini(ev)
    (o.visc_steps if viscous else o.ideal_steps)(nsteps)
    pairs = [('ev1', o.ev[1])] + ([('pi1', o.pi[1]), ('bulk1', o.bulk[1])] if viscous else [])
    for name, r in pairs:
        g = h.get(name)
        _say('  %-6s bit-equal %.8f' % (name, bit_equal_frac(g, r)))
        np.testing.assert_array_equal(g, r, err_msg=name)
    h.close()
    o0 = O.Oracle(orc_params(p, tab, nthreads=16), 'f32', viscous=viscous)
    o0.load_ini(ev)
    (o0.visc_steps if viscous else o0.ideal_steps)(nsteps)
    d = np.abs(o.ev[1][:, 0].astype(np.float64) - o0.ev[1][:, 0]).max() / np.abs(o0.ev[1][:, 0]).max()
    _say('  max |e(jet) - e(none)| / max e = %.3g' % d)
    assert d > 1e-4


def test_batch_rejects_vorticity_couplings():
    p, tab, _ = _visc_case(13, 11, 5)
    p.flags |= _lib.CV_FLAG_PIMN_OMEGA
    with pytest.raises(RuntimeError):
        _lib.Batch(p, tab, [(13, 11, 5), (9, 9, 3)])


def test_visc_stage_api_matches_reference_sequence():
    """cv_is_stage / cv_visc_stage / cv_update_udiff called in the order of
    CLVisc.evolve (pyvisc/visc.py:527-544) give the fused step's result."""
    p, tab, ev = _visc_case(21, 19, 7)
    _say('case: stage api 21x19x7, 3 steps')
    h = _lib.Handle(p, tab, viscous=True)
    h.visc_initialize(p.tau0)
    h.ideal_stage(1, np.float32(p.tau0))
    h.set('ev1', ev)
    o = O.Oracle(orc_params(p, tab), 'f32', viscous=True)
    o.load_ini(ev)
    for n in range(3):
        h.copy('pi0', 'pi1')
        h.copy('ev0', 'ev1')
        t1 = np.float32(p.tau0 + n * p.dt)
        t2 = np.float32(p.tau0 + (n + 1) * p.dt)
        h.is_stage(1, t1)
        h.visc_stage(1, t1)
        h.is_stage(2, t2)
        h.visc_stage(2, t2)
        h.update_udiff('ev0', 'ev1')
    o.visc_steps(3)
    for name, r in (('ev1', o.ev[1]), ('pi1', o.pi[1]), ('ev2', o.ev[2]), ('udiff', o.udiff)):
        g = h.get(name)
        m = max_rel(g, r)
        _say('  %-6s max rel %.3g bit-equal %.6f' % (name, m.max(), bit_equal_frac(g, r)))
        assert m.max() <= TOL


def _gubser_visc_handle(nx, dx, dt, tau0, L, lam1):
    tab = Eos('ideal_gas').create_table()
    p = cv_params(nx, nx, 1, dt, dx, dx, 0.3, tau0, etaos_xmin=0.0, etaos_ymin=0.2, lam1=lam1,
                  flags=_lib.CV_FLAG_GUBSER_VISC_TEST)
    h = _lib.Handle(p, tab, viscous=True)
    h.visc_initialize(tau0)
    h.ideal_stage(1, np.float32(tau0))
    ev1, pi1 = ic.gubser_visc(nx, nx, 1, dx, dx, 0.3, tau0, L, lam1)
    ev2, pi2 = ic.gubser_visc(nx, nx, 1, dx, dx, 0.3, tau0 + dt, L, lam1)
    for name, a in (('ev1', ev1), ('pi1', pi1), ('ev2', ev2), ('pi2', pi2)):
        h.set(name, a.astype(np.float32))
    h.update_udiff('ev1', 'ev2')
    return p, tab, h, (ev1, pi1, ev2, pi2)


def test_gubser_visc_bitwise_151():
    """GUBSER_VISC_TEST branch on the device vs the oracle, same start as
    pyvisc/tests/test_gubser_visc.py (analytic fields at tau0 and tau0 + DT)."""
    nx, dx, dt, tau0, L, lam1 = 151, 0.04, 0.005, 1.0, 2.0, -10.0
    _say('case: viscous gubser %dx%dx1, 100 steps vs oracle' % (nx, nx))
    p, tab, h, (ev1, pi1, ev2, pi2) = _gubser_visc_handle(nx, dx, dt, tau0, L, lam1)
    h.visc_steps(0, 100)
    o = O.Oracle(orc_params(p, tab, nthreads=16), 'f32', viscous=True)
    o.load_ini(ev1, pi1)
    o.ev[2][...] = ev2
    o.pi[2][...] = pi2
    o.update_udiff(o.ev[1], o.ev[2])
    o.visc_steps(100)
    for name, r in (('ev1', o.ev[1]), ('pi1', o.pi[1])):
        g = h.get(name)
        m = max_rel(g, r)
        _say('  %-4s max rel %.3g bit-equal %.6f' % (name, m.max(), bit_equal_frac(g, r)))
        assert m.max() <= TOL


def test_gubser_visc_analytic_501():
    """The reference's viscous Gubser setup (501^2, DX 0.04, DT 0.005, L 2, lambda_1 -10,
    pyvisc/tests/test_gubser_visc.py:53-80) to tau = 2 fm on the device: e, v_x, pi^zz,
    pi^xx within r < 2 fm against the analytic solution."""
    nx, dx, dt, tau0, L, lam1 = 501, 0.04, 0.005, 1.0, 2.0, -10.0
    _say('case: viscous gubser %dx%dx1, 200 steps vs analytic' % (nx, nx))
    p, tab, h, _ = _gubser_visc_handle(nx, dx, dt, tau0, L, lam1)
    h.visc_steps(0, 200)
    ea, pa = ic.gubser_visc(nx, nx, 1, dx, dx, 0.3, tau0 + 200 * dt, L, lam1)
    ev, pi = h.get('ev1'), h.get('pi1')
    x = (np.arange(nx) - nx // 2) * dx
    X, Y = np.meshgrid(x, x, indexing='ij')
    m = np.sqrt(X * X + Y * Y).ravel() < 2.0
    err = (np.abs(ev[m, 0] / ea[m, 0] - 1).max(), np.abs(ev[m, 1] - ea[m, 1]).max(),
           np.abs(pi[m, 9] / pa[m, 9] - 1).max(), np.abs(pi[m, 4] - pa[m, 4]).max() / np.abs(pa[m, 4]).max())
    _say('  r<2: e max rel %.3g, vx max abs %.3g, pizz max rel %.3g, pixx max abs/scale %.3g' % err)
    assert err[0] < 0.01 and err[1] < 0.003 and err[2] < 0.04 and err[3] < 0.03


def test_full_size_benchmark_grid_bitwise():
    """The 385x385x115 benchmark workload (viscous + bulk, default marching kernel) for two
    steps against the oracle on the same grid: identical fields."""
    import bench
    _say('case: full-size 385x385x115 viscous+bulk, 2 steps vs oracle')
    p, tab, ev = bench.make_workload(True)
    h = _lib.Handle(p, tab, viscous=True)
    _say('  stage kernel %s march chunk %s' % (h.stage_kernel(), h.march_chunk()))
    h.visc_initialize(p.tau0)
    h.ideal_stage(1, np.float32(p.tau0))
    h.set('ev1', ev)
    h.visc_steps(0, 2)
    o = O.Oracle(orc_params(p, tab, nthreads=16), 'f32', viscous=True)
    o.load_ini(ev)
    o.visc_steps(2)
    for name, r in (('ev1', o.ev[1]), ('pi1', o.pi[1]), ('bulk1', o.bulk[1]), ('ev0', o.ev[0]), ('udiff', o.udiff)):
        g = h.get(name)
        frac = bit_equal_frac(g, r)
        _say('  %-6s bit-equal %.8f' % (name, frac))
        assert max_rel(g, r).max() <= TOL
    h.close()


def test_clvisc_evolve_with_jet_matches_oracle(tmp_path):
    """CLVisc.evolve(jet_eloss_src=...) (pyvisc/visc.py:473-544: the deposit in stage 2 of each
    step) against the oracle's viscous steps with the jet in stage 2: identical fields."""
    from clvisc_amd.config import read_config
    from clvisc_amd import CLVisc
    nx, ny, nz, nsteps = 29, 31, 11, 12
    cfg = read_config(argv=[])
    cfg.NX, cfg.NY, cfg.NZ = nx, ny, nz
    cfg.DX, cfg.DY, cfg.DZ, cfg.DT, cfg.TAU0 = 0.08, 0.08, 0.15, 0.02, 0.6
    cfg.ETAOS_XMIN, cfg.ETAOS_YMIN, cfg.ETAOS_LEFT_SLOP, cfg.ETAOS_RIGHT_SLOP = 0.0, 0.08, 0.0, 0.0
    cfg.eos_type = 'crossover'
    cfg.fPathOut = str(tmp_path)
    start, direction = (14 * ny + 15) * nz + 5, 0.7
    _say('case: CLVisc.evolve with jet_eloss_src, %dx%dx%d, %d steps' % (nx, ny, nz, nsteps))
    visc = CLVisc(cfg)
    pos = ic.woods_saxon_positions(12, 3, R=0.6, a=0.1)
    ev = ic.hotspots(nx, ny, nz, 0.08, 0.08, 0.15, positions=pos, emax=50.0)
    visc.ideal.load_ini(ev)
    visc.evolve(max_loops=nsteps, save_hypersf=False, save_bulk=False, save_pi=False, force_run_to_maxloop=True,
                jet_eloss_src={'switch_on': True, 'start_pos_index': start, 'direction': direction})
    prm = orc_params(visc.handle.params, visc.eos_table, nthreads=16)
    prm.jet_on, prm.jet_direction, prm.jet_start = 2, float(np.float32(direction)), start
    o = O.Oracle(prm, 'f32', viscous=True)
    o.load_ini(ev)
    o.visc_steps(nsteps)
    for name, r in (('ev1', o.ev[1]), ('pi1', o.pi[1])):
        g = visc.handle.get(name)
        m = max_rel(g, r).max()
        _say('  %-4s max rel %.3g bit-equal %.8f' % (name, m, bit_equal_frac(g, r)))
        assert m <= TOL, name
    # the reference passes jet_eloss_src per call: after evolve returns the handle steps
    # without the jet (two more steps equal the oracle's steps with jet_on = 0)
    visc.handle.visc_steps(nsteps, 2)
    o.p.jet_on = 0
    o.visc_steps(2)
    np.testing.assert_array_equal(visc.handle.get('ev1'), o.ev[1])


def test_large_octave_table_runs_the_plane_tiled_kernel():
    """The x-marching kernels hold the pressure column of an octave table in LDS (at most 1540
    nodes): a finer table (3073 nodes) gets the plane-tiled kernel by default, kind 2 and
    batches refuse it, and its steps equal the oracle's."""
    tab = Eos('crossover', log2j=7).create_table()
    nx, ny, nz = 29, 31, 11
    p = cv_params(nx, ny, nz, 0.02, 0.08, 0.08, 0.15, 0.6, etaos_xmin=0.0, etaos_ymin=0.08,
                  bulk=(0.05, 0.155, 0.02))
    _say('case: octave table of %d nodes, %dx%dx%d, 8 steps' % (tab.num_ed, nx, ny, nz))
    h = _lib.Handle(p, tab, viscous=True)
    assert h.stage_kernel()[0] == 1
    with pytest.raises(_lib.CvError) as err:
        h.set_stage_kernel(2, 0, 0)
    assert err.value.code == _lib.CV_ERR_UNSUPPORTED
    with pytest.raises(_lib.CvError) as err:
        _lib.Batch(p, tab, [(nx, ny, nz)])
    assert err.value.code == _lib.CV_ERR_UNSUPPORTED
    pos = ic.woods_saxon_positions(12, 3, R=0.6, a=0.1)
    ev = ic.hotspots(nx, ny, nz, 0.08, 0.08, 0.15, positions=pos, emax=50.0)
    h.visc_initialize(p.tau0)
    h.ideal_stage(1, np.float32(p.tau0))
    h.set('ev1', ev)
    h.visc_steps(0, 8)
    o = O.Oracle(orc_params(p, tab), 'f32', viscous=True)
    o.load_ini(ev)
    o.visc_steps(8)
    for name, r in (('ev1', o.ev[1]), ('pi1', o.pi[1])):
        g = h.get(name)
        _say('  %-4s bit-equal %.8f' % (name, bit_equal_frac(g, r)))
        np.testing.assert_array_equal(g, r)
    h.close()
