"""Vorticity on the device (k_vorticity, k_vorticity_mu, the surface interpolation) against
the oracle's restatement of kernel_vorticity.cl and its centroid interpolation."""
import sys

import numpy as np
import pytest

from clvisc_amd import _lib, ic
from clvisc_amd.eos import Eos
from helpers import cv_params, orc_params, bit_equal_frac
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg)
    sys.stdout.flush()


@pytest.mark.parametrize('nz,eos_name', [(13, 'crossover'), (1, 'crossover'), (9, 'ideal_gas')])
def test_vorticity_matches_oracle(nz, eos_name):
    nx, ny = 33, 29
    tab = Eos(eos_name).create_table()
    p = cv_params(nx, ny, nz, 0.02, 0.08, 0.08, 0.15, 0.6, etaos_xmin=0.0, etaos_ymin=0.08)
    pos = ic.woods_saxon_positions(15, 4, R=0.8, a=0.1)
    ev = ic.hotspots(nx, ny, nz, 0.08, 0.08, 0.15, positions=pos, emax=40.0)
    _say('case: vorticity %dx%dx%d %s after 12 steps' % (nx, ny, nz, eos_name))
    h = _lib.Handle(p, tab, viscous=True)
    h.visc_initialize(p.tau0)
    h.ideal_stage(1, np.float32(p.tau0))
    h.set('ev1', ev)
    h.visc_steps(0, 12)
    h.vorticity_enable()
    tau = np.float32(p.tau0 + 12 * p.dt)
    h.vorticity(tau)
    g = h.get('omega1')
    op = orc_params(p, tab)
    r = O.vorticity(op, h.get('ev0'), h.get('ev1'), tau)
    _say('  bit-equal %.6f, max |omega| %.3g' % (bit_equal_frac(g, r), np.abs(r).max()))
    np.testing.assert_array_equal(g, r)
    efrz = 0.3
    np.testing.assert_array_equal(h.vorticity_mu(efrz), O.vorticity_mu(op, h.get('ev1'), r, np.float32(efrz)))
    h.close()


def test_vorticity_on_the_surface():
    """The 32 tensors at the element centroids: the same 16-corner interpolation as pi
    (checked with the oracle's pi interpolation, ten components at a time)."""
    from test_gpu_hypersf import _blob, _sorted, DX, DY, DZ
    nx, ny, nz = 25, 23, 11
    ev0, ev1, _, _, _, _ = _blob(nx, ny, nz)
    rng = np.random.default_rng(8)
    om0 = (rng.standard_normal((ev0.shape[0], 32)) * 0.1).astype(np.float32)
    om1 = (rng.standard_normal((ev0.shape[0], 32)) * 0.1).astype(np.float32)
    tab = Eos('ideal_gas').create_table()
    p = cv_params(nx, ny, nz, 0.02, DX, DY, DZ, 1.0)
    p.efrz = 0.5
    p.nxskip = p.nyskip = p.nzskip = 1
    _say('case: vorticity on the surface %dx%dx%d' % (nx, ny, nz))
    h = _lib.Handle(p, tab, viscous=True)
    h.hypersf_reserve(100000)
    h.vorticity_enable()
    t0, t1 = np.float32(1.0), np.float32(1.1)
    h.set('ev1', ev0)
    h.hypersf_snapshot(t0)
    h.set('ev1', ev1)
    h.set('omega0', om0)
    h.set('omega1', om1)
    h.hypersf_find(t1)
    sf, _, _, _ = h.hypersf_get()
    om = h.hypersf_get_vorticity()
    np.testing.assert_array_equal(h.get('omega0'), om1)       # d_omega[0] <- d_omega[1]
    chunks = []
    for c0 in (0, 10, 20, 30):
        a = np.zeros((ev0.shape[0], 10), np.float32)
        b = np.zeros_like(a)
        w = min(10, 32 - c0)
        a[:, :w], b[:, :w] = om0[:, c0:c0 + w], om1[:, c0:c0 + w]
        rc, rsf, rtx, rps, _ = O.hypersf(nx, ny, nz, DX, DY, DZ, 1, 1, 1, 0.5, ev0, ev1, t0, t1, pi_old=a, pi_new=b)
        chunks.append((rsf, rps[:, :w]))
    assert rc == sf.shape[0] > 20
    rom = np.concatenate([c[1] for c in chunks], axis=1).astype(np.float32)
    g = _sorted(sf, om)
    r = _sorted(chunks[0][0].astype(np.float32), rom)
    _say('  %d elements, bit-equal %.6f' % (rc, bit_equal_frac(g[1], r[1])))
    np.testing.assert_array_equal(g[1], r[1])
    h.close()


def test_clvisc_evolve_zero_viscosity_with_vorticity_runs_viscous_loop(tmp_path):
    """With eta/s = 0 the reference takes the ideal shortcut only without save_vorticity
    (pyvisc/visc.py:483-484): with save_vorticity=True it runs the viscous loop, which writes
    the vorticity on the surface."""
    from clvisc_amd.config import read_config
    from clvisc_amd import CLVisc
    nx, ny, nz = 29, 31, 11
    cfg = read_config(argv=[])
    cfg.NX, cfg.NY, cfg.NZ = nx, ny, nz
    cfg.DX, cfg.DY, cfg.DZ, cfg.DT, cfg.TAU0 = 0.08, 0.08, 0.15, 0.02, 0.6
    cfg.ETAOS_XMIN, cfg.ETAOS_YMIN, cfg.ETAOS_LEFT_SLOP, cfg.ETAOS_RIGHT_SLOP = 0.0, 0.0, 0.0, 0.0
    cfg.ntskip, cfg.nxskip, cfg.nyskip, cfg.nzskip = 5, 1, 1, 1
    cfg.eos_type = 'crossover'
    cfg.TFRZ = 0.30
    cfg.fPathOut = str(tmp_path)
    _say('case: CLVisc.evolve(save_vorticity=True) with eta/s = 0, %dx%dx%d, 20 steps' % (nx, ny, nz))
    visc = CLVisc(cfg)
    pos = ic.woods_saxon_positions(12, 3, R=0.6, a=0.1)
    visc.ideal.load_ini(ic.hotspots(nx, ny, nz, 0.08, 0.08, 0.15, positions=pos, emax=50.0))
    visc.evolve(max_loops=20, save_hypersf=True, save_bulk=False, save_pi=False, force_run_to_maxloop=True,
                save_vorticity=True)
    n = visc.handle.hypersf_count()[0]
    assert n > 0
    for name in ('omegamu_sf.dat', 'omegamu_shear1_sf.dat', 'omegamu_shear2_sf.dat', 'omegamu_accT_sf.dat'):
        assert np.loadtxt(tmp_path / name, ndmin=2).shape[0] == n, name
    assert np.abs(np.loadtxt(tmp_path / 'omegamu_sf.dat')).max() > 0


def test_clvisc_evolve_saves_vorticity(tmp_path):
    from clvisc_amd.config import read_config
    from clvisc_amd import CLVisc
    nx, ny, nz = 29, 31, 11
    cfg = read_config(argv=[])
    cfg.NX, cfg.NY, cfg.NZ = nx, ny, nz
    cfg.DX, cfg.DY, cfg.DZ, cfg.DT, cfg.TAU0 = 0.08, 0.08, 0.15, 0.02, 0.6
    cfg.ETAOS_XMIN, cfg.ETAOS_YMIN, cfg.ETAOS_LEFT_SLOP, cfg.ETAOS_RIGHT_SLOP = 0.0, 0.08, 0.0, 0.0
    cfg.ntskip, cfg.nxskip, cfg.nyskip, cfg.nzskip = 5, 1, 1, 1
    cfg.eos_type = 'crossover'
    cfg.TFRZ = 0.30
    cfg.fPathOut = str(tmp_path)
    _say('case: CLVisc.evolve(save_vorticity=True) %dx%dx%d, 30 steps' % (nx, ny, nz))
    visc = CLVisc(cfg)
    pos = ic.woods_saxon_positions(12, 3, R=0.6, a=0.1)
    visc.ideal.load_ini(ic.hotspots(nx, ny, nz, 0.08, 0.08, 0.15, positions=pos, emax=50.0))
    visc.evolve(max_loops=30, save_hypersf=True, save_bulk=False, save_pi=False, force_run_to_maxloop=True,
                save_vorticity=True)
    n = visc.handle.hypersf_count()[0]
    assert n > 0
    for name, ncol in (('omegamu_sf.dat', 6), ('omegamu_shear1_sf.dat', 16), ('omegamu_shear2_sf.dat', 4),
                       ('omegamu_accT_sf.dat', 6)):
        a = np.loadtxt(tmp_path / name, ndmin=2)
        assert a.shape == (n, ncol), name
    assert np.abs(np.loadtxt(tmp_path / 'omegamu_sf.dat')).max() > 0
    visc.get_vorticity(save_data=True)
    files = list(tmp_path.glob('omega_mu_tau*.dat'))
    assert files and np.loadtxt(files[0]).shape == (nx * ny * nz, 4)
    # the polarization driver reads what evolve wrote
    from clvisc_amd.polarization import Polarization
    pz = Polarization(str(tmp_path), T=0.30, shear=True, accT=True)
    mt = np.sqrt(1.0 + 1.116 ** 2)
    pz.set_momentum(np.array([[mt, 0.0, 1.0, 0.0], [mt, 0.3, 0.0, 1.0]], np.float32))
    for get in (pz.get_themal, pz.get_shear, pz.get_accT):
        pol, rho, lrf = get()
        assert pol.shape == (2, 4) and np.all(rho > 0) and np.all(np.isfinite(lrf))


@pytest.mark.parametrize('kind', [0, 1, 2])
def test_polarization_matches_oracle(kind):
    """cv_polarization (one workgroup per momentum, register sums) against the oracle's
    restatement in the reference's summation order (tree per 256 elements)."""
    from test_polarization_oracle import _surface, _mom, T, M
    from clvisc_amd.polarization import Polarization
    n = 3000
    sf = _surface(n, 11)
    mom = _mom(64, 12)
    rng = np.random.default_rng(13)
    w = 16 if kind == 1 else 6
    o1 = rng.normal(0, 0.05, (n, w)).astype(np.float32)
    o2 = rng.normal(0, 0.05, (n, 4)).astype(np.float32)
    _say('case: polarization kind %d, %d elements, %d momenta' % (kind, n, mom.shape[0]))
    pz = Polarization.from_arrays(sf, omega=o1 if kind == 0 else None, omega_shear1=o1 if kind == 1 else None,
                                  omega_shear2=o2 if kind == 1 else None, omega_accT=o1 if kind == 2 else None,
                                  T=T, mass=M)
    pz.set_momentum(mom)
    pol, rho, lrf = (pz.get_themal, pz.get_shear, pz.get_accT)[kind]()
    rp, rr, rl = O.polarization(kind, sf, o1, o2, mom, T, M, 'f32')
    e_pol = np.abs(pol - rp).max() / np.abs(rp).max()
    e_rho = np.abs(rho - rr).max() / np.abs(rr).max()
    e_lrf = np.abs(lrf[:, 1:] - rl[:, 1:]).max() / np.abs(rl[:, 1:]).max()
    _say('  rel diff: pol_lab %.3g density %.3g pol_lrf %.3g' % (e_pol, e_rho, e_lrf))
    assert e_pol < 1e-4 and e_rho < 1e-5 and e_lrf < 1e-4
