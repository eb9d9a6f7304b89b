"""Freeze-out surface and drivers on the GPU vs the oracle.

The device appends elements in no defined order (atomic append, as the reference);
both sides are sorted by their rows and compared as multisets.  FP32 geometry
follows the same operation sequence as the FP32 oracle, so the element rows are
expected bit for bit except t, z (cosh, sinh); the FP64 option is compared with the
FP64 oracle after rounding to float.
"""
import sys

import numpy as np
import pytest

from clvisc_amd import _lib, ic
from clvisc_amd.eos import Eos
from oracle import oracle as O
from helpers import cv_params, orc_params, max_rel

pytestmark = pytest.mark.gpu
DX, DY, DZ = 0.3, 0.25, 0.2


def _say(msg):
    print(msg)
    sys.stdout.flush()


def _blob(nx, ny, nz, seed=1):
    x = (np.arange(nx) - nx // 2) * DX
    y = (np.arange(ny) - ny // 2) * DY
    z = (np.arange(nz) - nz // 2) * DZ
    X, Y, Z = np.meshgrid(x, y, z, indexing='ij')
    R2 = X ** 2 + 1.3 * Y ** 2 + 0.7 * Z ** 2
    rng = np.random.default_rng(seed)
    ev0 = np.zeros((X.size, 4), np.float32)
    ev1 = np.zeros_like(ev0)
    ev0[:, 0] = (np.exp(-R2 / 4) * (1 + 0.05 * rng.random(X.shape))).reshape(-1)
    ev1[:, 0] = (0.8 * np.exp(-R2 / 4) * (1 + 0.05 * rng.random(X.shape))).reshape(-1)
    for ev, s in ((ev0, 1.0), (ev1, 1.2)):
        ev[:, 1] = (0.1 * s * X).reshape(-1)
        ev[:, 2] = (0.08 * s * Y).reshape(-1)
        ev[:, 3] = (0.05 * s * Z).reshape(-1)
    pi0 = (rng.standard_normal((X.size, 10)) * 0.01).astype(np.float32)
    pi1 = (rng.standard_normal((X.size, 10)) * 0.01).astype(np.float32)
    bk0 = (rng.standard_normal(X.size) * 0.01).astype(np.float32)
    bk1 = (rng.standard_normal(X.size) * 0.01).astype(np.float32)
    return ev0, ev1, pi0, pi1, bk0, bk1


def _sorted(sf, *others):
    order = np.lexsort(sf.T[::-1])
    return [sf[order]] + [None if o is None else o[order] for o in others]


@pytest.mark.parametrize('viscous,bulk,fp64,skip', [(False, False, False, 1), (True, False, False, 1),
                                                    (True, True, False, 2), (True, True, True, 1),
                                                    (False, False, True, 1)])
def test_hypersf_vs_oracle(viscous, bulk, fp64, skip):
    nx, ny, nz = 25, 23, 11
    _say('case: hypersf %dx%dx%d viscous=%s bulk=%s fp64=%s skip=%d' % (nx, ny, nz, viscous, bulk, fp64, skip))
    ev0, ev1, pi0, pi1, bk0, bk1 = _blob(nx, ny, nz)
    tab = Eos('ideal_gas').create_table()
    p = cv_params(nx, ny, nz, 0.02, DX, DY, DZ, 1.0, bulk=(0.05, 0.155, 0.02) if bulk else None)
    p.efrz = 0.5
    p.nxskip = p.nyskip = p.nzskip = skip
    if fp64:
        p.flags |= _lib.CV_FLAG_SF_FP64
    h = _lib.Handle(p, tab, viscous=viscous)
    h.hypersf_reserve(100000)
    t0, t1 = np.float32(1.0), np.float32(1.1)
    h.set('ev1', ev0)
    if viscous:
        h.set('pi1', pi0)
    if bulk:
        h.set('bulk1', bk0)
    h.hypersf_snapshot(t0)
    h.set('ev1', ev1)
    if viscous:
        h.set('pi1', pi1)
    if bulk:
        h.set('bulk1', bk1)
    h.hypersf_find(t1)
    cnt, cap = h.hypersf_count()
    sf, tx, ps, bs = h.hypersf_get()
    prec = 'f64' if fp64 else 'f32'
    kw = dict(pi_old=pi0, pi_new=pi1) if viscous else {}
    if bulk:
        kw.update(bk_old=bk0, bk_new=bk1)
    rc, rsf, rtx, rps, rbs = O.hypersf(nx, ny, nz, DX, DY, DZ, skip, skip, skip, 0.5, ev0, ev1, t0, t1,
                                       precision=prec, **kw)
    _say('  elements gpu %d oracle %d' % (cnt, rc))
    assert cnt == rc > 20
    g = _sorted(sf, tx, ps, bs)
    r = _sorted(rsf.astype(np.float32), rtx.astype(np.float32), None if rps is None else rps.astype(np.float32),
                None if rbs is None else rbs.astype(np.float32))
    if not fp64:
        np.testing.assert_array_equal(g[0], r[0])
        if viscous:
            np.testing.assert_array_equal(g[2], r[2])
        if bulk:
            np.testing.assert_array_equal(g[3], r[3])
    else:
        np.testing.assert_allclose(g[0], r[0], rtol=1e-6, atol=1e-9)
        if viscous:
            np.testing.assert_allclose(g[2], r[2], rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(g[1], r[1], rtol=1e-6, atol=1e-7)
    _say('  max rel txyz %.3g' % max_rel(g[1], r[1]).max())


def test_hypersf_capacity_overflow_reported():
    nx, ny, nz = 25, 23, 11
    ev0, ev1 = _blob(nx, ny, nz)[:2]
    tab = Eos('ideal_gas').create_table()
    p = cv_params(nx, ny, nz, 0.02, DX, DY, DZ, 1.0)
    p.efrz = 0.5
    p.nxskip = p.nyskip = p.nzskip = 1
    h = _lib.Handle(p, tab, viscous=False)
    h.hypersf_reserve(10)
    h.set('ev1', ev0)
    h.hypersf_snapshot(1.0)
    h.set('ev1', ev1)
    h.hypersf_find(1.1)
    cnt, cap = h.hypersf_count()
    assert cap == 10 and cnt > 10
    with pytest.raises(_lib.CvError, match='capacity'):
        h.hypersf_get()
    h.hypersf_reset()
    assert h.hypersf_count()[0] == 0


def test_gather_slices():
    nx, ny, nz = 9, 7, 5
    tab = Eos('ideal_gas').create_table()
    p = cv_params(nx, ny, nz, 0.02, DX, DY, DZ, 1.0)
    h = _lib.Handle(p, tab, viscous=True)
    rng = np.random.default_rng(5)
    pi = rng.standard_normal((nx * ny * nz, 10)).astype(np.float32)
    h.set('pi1', pi)
    cells = np.array([0, 17, nx * ny * nz - 1, 42, 42])
    np.testing.assert_array_equal(h.gather('pi1', cells), pi[cells])
    with pytest.raises(_lib.CvError):
        h.gather('pi1', [nx * ny * nz])


def test_clvisc_driver_surface_matches_oracle(tmp_path):
    """CLVisc.evolve (batched fused steps + surface every ntskip) gives the surface the
    oracle gives for the same step sequence."""
    from clvisc_amd.config import read_config
    from clvisc_amd import CLVisc
    nx, ny, nz = 29, 31, 11
    cfg = read_config(argv=[])
    cfg.NX, cfg.NY, cfg.NZ = nx, ny, nz
    cfg.DX, cfg.DY, cfg.DZ, cfg.DT, cfg.TAU0 = 0.08, 0.08, 0.15, 0.02, 0.6
    cfg.ETAOS_XMIN, cfg.ETAOS_YMIN, cfg.ETAOS_LEFT_SLOP, cfg.ETAOS_RIGHT_SLOP = 0.0, 0.08, 0.0, 0.0
    cfg.ZETAOS_MAX = 0.0
    cfg.ntskip, cfg.nxskip, cfg.nyskip, cfg.nzskip = 5, 1, 1, 1
    cfg.eos_type = 'crossover'
    cfg.TFRZ = 0.30
    cfg.fPathOut = str(tmp_path)
    _say('case: CLVisc driver %dx%dx%d, 30 steps, surface every 5' % (nx, ny, nz))
    visc = CLVisc(cfg)
    pos = ic.woods_saxon_positions(12, 3, R=0.6, a=0.1)
    ev = ic.hotspots(nx, ny, nz, 0.08, 0.08, 0.15, positions=pos, emax=50.0)
    visc.ideal.load_ini(ev)
    visc.evolve(max_loops=30, save_hypersf=True, save_bulk=False, save_pi=True, force_run_to_maxloop=True)
    sf, tx, ps, _ = visc.ideal.hypersf.get()
    # oracle: same steps, surface between snapshots every ntskip
    tab = visc.eos_table
    from helpers import orc_params
    o = O.Oracle(orc_params(visc.handle.params, tab, nthreads=16), 'f32', viscous=True)
    o.load_ini(ev)
    efrz = np.float32(visc.ideal.efrz)
    rows, pis = [], []
    prev = None
    edmax = None
    for loop in range(30):      # the loop of pyvisc/visc.py:490-555 with force_run_to_maxloop
        tau = np.float32(cfg.TAU0 + loop * cfg.DT)
        if loop % cfg.ntskip == 0:
            edmax = o.max_ed()
        finished = edmax < efrz
        if loop == 0:
            prev = (o.ev[1].copy(), o.pi[1].copy(), np.float32(cfg.TAU0))
        elif loop % cfg.ntskip == 0 or finished:
            c, rsf, rtx, rps, _ = O.hypersf(nx, ny, nz, cfg.DX, cfg.DY, cfg.DZ, 1, 1, 1, efrz, prev[0], o.ev[1],
                                            prev[2], tau, pi_old=prev[1], pi_new=o.pi[1])
            rows.append(rsf); pis.append(rps)
            prev = (o.ev[1].copy(), o.pi[1].copy(), tau)
        o.visc_steps(1)
    rsf = np.concatenate(rows).astype(np.float32)
    rps = np.concatenate(pis).astype(np.float32)
    _say('  elements gpu %d oracle %d' % (sf.shape[0], rsf.shape[0]))
    assert sf.shape[0] == rsf.shape[0] > 0
    g = _sorted(sf, ps)
    r = _sorted(rsf, rps)
    np.testing.assert_array_equal(g[0], r[0])
    np.testing.assert_array_equal(g[1], r[1])
    # the reference's files were written
    for f in ('hypersf.dat', 'sf_txyz.dat', 'pimnsf.dat', 'pixx_x.dat'):
        assert (tmp_path / f).exists(), f
    assert np.loadtxt(tmp_path / 'hypersf.dat').shape == (sf.shape[0], 8)
