"""Split viscous stage: the stage kernel writes the root solve's conserved input and
k_visc_root finishes the cells (update_ev from the clamp on), optionally seeded by the
learned network.  Without a seed it is bit-identical to the fused stage (and so to the
oracle); with a seed the device solve equals the oracle's solve from the same seeds."""
import sys

import numpy as np
import pytest

from clvisc_amd import _lib, ic, surrogate
from clvisc_amd.eos import Eos
from oracle import mlp_oracle as MO
from oracle import oracle as O
from helpers import cv_params, orc_params, bit_equal_frac

pytestmark = pytest.mark.gpu

FIELDS = ('ev0', 'ev1', 'ev2', 'pi0', 'pi1', 'pi2', 'udiff', 'goodcell')


def _say(msg):
    print(msg)
    sys.stdout.flush()


def _case(nx, ny, nz, bulk=None, nhot=12, seed=3):
    eos = Eos('crossover')
    tab = eos.create_table()
    p = cv_params(nx, ny, nz, 0.02, 0.08, 0.08, 0.15, 0.6, etaos_xmin=0.0, etaos_ymin=0.08, bulk=bulk)
    pos = ic.woods_saxon_positions(nhot, seed, R=0.6, a=0.1)
    ev = ic.hotspots(nx, ny, nz, 0.08, 0.08, 0.15, positions=pos, emax=50.0)
    return p, tab, ev


def _start(p, tab, ev, kernel=None):
    h = _lib.Handle(p, tab, viscous=True)
    if kernel is not None:
        h.set_stage_kernel(*kernel)
    h.visc_initialize(p.tau0)
    h.ideal_stage(1, np.float32(p.tau0))
    h.set('ev1', ev)
    return h


@pytest.mark.parametrize('kernel,bulk', [((2, 0, 0), None), ((2, 0, 0), (0.05, 0.155, 0.02)),
                                         ((1, 0, 0), (0.05, 0.155, 0.02)), ((2, 7, 30), None)])
def test_split_without_seed_equals_fused(kernel, bulk):
    p, tab, ev = _case(29, 31, 11, bulk=bulk)
    _say('case: split vs fused stage, 29x31x11, kernel %s, bulk %s, 20 steps' % (kernel, bulk is not None))
    hf = _start(p, tab, ev, kernel)
    hs = _start(p, tab, ev, kernel)
    hs.set_root_split(True)
    hf.visc_steps(0, 20)
    hs.visc_steps(0, 20)
    names = FIELDS + (('bulk0', 'bulk1', 'bulk2') if bulk else ())
    for name in names:
        a, b = hs.get(name), hf.get(name)
        _say('  %-8s bit-equal %.8f' % (name, bit_equal_frac(a, b)))
        np.testing.assert_array_equal(a, b, err_msg=name)
    hf.close()
    hs.close()


def test_split_seeded_root_matches_oracle():
    """A seeded step with a random network (arbitrary seeds in [0, 0.9999]): the seeds agree
    with the network oracle's on the same features (the seed kernel uses the hardware exp
    and BF16 conversion: within 5e-3) for cells with T^tt >= 1e-3 GeV/fm^3, and are the
    reference's start 0.5 for the more dilute cells; the solve from those seeds equals the
    oracle's bit for bit, iteration counts included."""
    p, tab, ev = _case(33, 29, 13, bulk=(0.05, 0.155, 0.02))
    # a dilute third of the grid (e scaled by 1e-6), so that cells on both sides of the
    # network's T^tt floor are solved
    ev = ev.copy()
    ev.reshape(33, 29, 13, 4)[:11, ..., 0] *= np.float32(1e-6)
    _say('case: seeded split stage vs oracle, 33x29x13 viscous+bulk, dilute third')
    h = _start(p, tab, ev)
    h.visc_steps(0, 5)
    rng = np.random.default_rng(21)
    P = (rng.standard_normal(MO.NPARAM) * 0.3).astype(np.float32)
    m = _lib.Mlp()
    m.set_params(P)
    h.set_root_seed(m, track_iters=True)
    h.visc_steps(5, 1)
    cons = h.get('root_cons')
    seed = h.get('root_seed')
    bk = h.get('bulk1')
    g = h.get('ev1')
    it = h.root_iters()
    T00, M, T03 = surrogate.clamp_cons(cons)
    _, _, Y = MO.forward(P, MO.features(T00, M, T03), bf16=True)
    net = cons[:, 0] >= np.float32(1e-3)
    _say('  cells above the seed floor: %d of %d' % (net.sum(), net.size))
    assert 0 < net.sum() < net.size
    np.testing.assert_array_equal(seed[~net], np.float32(0.5))
    err = np.abs(seed[net] - Y[net, 1]).max()
    _say('  seed vs network oracle: max abs %.3g; seeds in [%.3f, %.3f]' % (err, seed.min(), seed.max()))
    assert err < 5e-3 * max(1.0, np.abs(Y[:, 1]).max())
    op = orc_params(p, tab)
    r, itr = O.visc_root_from(op, cons, bk, seed)
    _say('  ev1 bit-equal %.8f; iterations mean %.3f (oracle %.3f)' % (bit_equal_frac(g, r), it.mean(), itr.mean()))
    np.testing.assert_array_equal(g, r)
    np.testing.assert_array_equal(it, itr)
    # the same stage unseeded (split, reference start): oracle's plain solve
    h2 = _start(p, tab, ev)
    h2.visc_steps(0, 5)
    h2.set_root_split(True)
    h2.set_root_seed(None, track_iters=True)
    h2.visc_steps(5, 1)
    r2, itr2 = O.visc_root_from(op, h2.get('root_cons'), h2.get('bulk1'), None)
    np.testing.assert_array_equal(h2.get('ev1'), r2)
    np.testing.assert_array_equal(h2.root_iters(), itr2)
    h.close()
    h2.close()
    m.close()


def test_split_seeded_fp64_root_matches_oracle():
    """Seed and CV_FLAG_FP64_ROOT together (vmin 0.02, so that many cells are fast): the fast
    cells take the FP64 root, the others start from the seed; the stage equals the oracle's
    split root with the same seeds and the same precedence, bit for bit."""
    p, tab, ev = _case(33, 29, 13, bulk=(0.05, 0.155, 0.02))
    p.flags |= _lib.CV_FLAG_FP64_ROOT
    p.fp64_root_vmin = 0.02
    _say('case: seeded split stage with the FP64 root (vmin 0.02) vs oracle, 33x29x13 viscous+bulk')
    h = _start(p, tab, ev)
    h.visc_steps(0, 20)
    rng = np.random.default_rng(22)
    m = _lib.Mlp()
    m.set_params((rng.standard_normal(MO.NPARAM) * 0.3).astype(np.float32))
    h.set_root_seed(m, track_iters=True)
    h.visc_steps(20, 1)
    cons, seed, bk, g, it = h.get('root_cons'), h.get('root_seed'), h.get('bulk1'), h.get('ev1'), h.root_iters()
    T00, M, _ = surrogate.clamp_cons(cons)
    fast = M > np.float32(0.02) * T00
    op = orc_params(p, tab)
    r, itr = O.visc_root_from(op, cons, bk, seed)
    _say('  %d fast cells of %d; ev1 bit-equal %.8f' % (fast.sum(), fast.size, bit_equal_frac(g, r)))
    assert fast.sum() > 100
    np.testing.assert_array_equal(g, r)
    np.testing.assert_array_equal(it[~fast], itr[~fast])
    h.close()
    m.close()


def test_trained_seed_viscous_run_close_to_plain():
    """Seed trained on the split stage's own states; the seeded run needs fewer iterations
    and stays within the root solve's tolerance of the plain run."""
    p, tab, ev = _case(41, 41, 15, bulk=(0.05, 0.155, 0.02), nhot=20)
    _say('case: trained seed, viscous+bulk 41x41x15, 30 steps seeded vs plain')
    h = _start(p, tab, ev)
    states = surrogate.visc_run_states(h, 0, 30, every=2)
    h.close()
    mlp, hist = surrogate.train_seed(states, steps=900, batch=4096, lr=3e-3)
    _say('  loss %.4g -> %.4g' % (hist[:20].mean(), hist[-20:].mean()))
    hp = _start(p, tab, ev)
    hp.set_root_split(True)
    hp.set_root_seed(None, track_iters=True)
    hs = _start(p, tab, ev)
    hs.set_root_seed(mlp, track_iters=True)
    itp, its = [], []
    for n in range(30):
        hp.visc_steps(n, 1)
        hs.visc_steps(n, 1)
        itp.append(hp.root_iters().mean())
        its.append(hs.root_iters().mean())
    ep, es = hp.get('ev1')[:, 0], hs.get('ev1')[:, 0]
    msk = ep > 1e-3 * ep.max()
    rel = np.abs(es[msk] - ep[msk]) / ep[msk]
    _say('  iterations plain %.3f seeded %.3f; max rel e diff %.3g' % (np.mean(itp), np.mean(its), rel.max()))
    assert np.mean(its) < np.mean(itp)
    assert rel.max() < 1e-3
    for x in (hp, hs, mlp):
        x.close()


@pytest.mark.parametrize('kernel,bulk,fp64', [((2, 0, 0), (0.05, 0.155, 0.02), False),
                                              ((2, 0, 0), (0.05, 0.155, 0.02), True),
                                              ((2, 0, 0), None, False),
                                              ((2, 7, 30), None, True)])
def test_fused_seed_equals_split_seed(kernel, bulk, fp64):
    """The learned seed inside the stage kernel (cv_set_root_split(0) with a network set: the
    whole wave runs the network before the root solve) gives the split stage's seeds,
    iteration counts and fields bit for bit, with and without the FP64 root for fast cells;
    the split stage is the one checked against the oracle above."""
    p, tab, ev = _case(37, 31, 13, bulk=bulk, nhot=14, seed=5)
    # dilute cells (e scaled by 1e-6) for x < 12 and for y < 10: wavefronts wholly below the
    # network's T^tt floor (which skip the network) and wavefronts on both sides of it
    ev = ev.copy()
    e4 = ev.reshape(37, 31, 13, 4)
    e4[:12, ..., 0] *= np.float32(1e-6)
    e4[12:, :10, :, 0] *= np.float32(1e-6)
    if fp64:
        p.flags |= _lib.CV_FLAG_FP64_ROOT
        p.fp64_root_vmin = 0.02
    _say('case: fused vs split seeded stage, 37x31x13, kernel %s, bulk %s, fp64 %s, 5 + 3 steps'
         % (kernel, bulk is not None, fp64))
    rng = np.random.default_rng(23)
    m = _lib.Mlp()
    m.set_params((rng.standard_normal(MO.NPARAM) * 0.3).astype(np.float32))
    hs = _start(p, tab, ev, kernel)
    hf = _start(p, tab, ev, kernel)
    hs.visc_steps(0, 5)
    hf.visc_steps(0, 5)
    hs.set_root_seed(m, track_iters=True)
    hf.set_root_seed(m, track_iters=True)
    hf.set_root_split(False)
    for n in range(5, 8):
        hs.visc_steps(n, 1)
        hf.visc_steps(n, 1)
        for name in ('root_seed',):
            a, b = hf.get(name), hs.get(name)
            _say('  step %d %-9s bit-equal %.8f' % (n, name, bit_equal_frac(a, b)))
            np.testing.assert_array_equal(a, b, err_msg='%s step %d' % (name, n))
        np.testing.assert_array_equal(hf.root_iters(), hs.root_iters(), err_msg='iterations step %d' % n)
    names = FIELDS + (('bulk0', 'bulk1', 'bulk2') if bulk else ())
    for name in names:
        a, b = hf.get(name), hs.get(name)
        _say('  %-8s bit-equal %.8f' % (name, bit_equal_frac(a, b)))
        np.testing.assert_array_equal(a, b, err_msg=name)
    it = hf.root_iters()
    _say('  iterations mean %.3f' % it.mean())
    hs.close()
    hf.close()
    m.close()


def test_fused_seed_request_on_plane_tiled_kernel_runs_split():
    """The in-stage seed needs the x-marching kernel; asked for on the plane-tiled kernel
    (kind 1) the stage runs split, with the split form's fields and seeds."""
    p, tab, ev = _case(29, 27, 11, bulk=(0.05, 0.155, 0.02), nhot=10, seed=7)
    _say('case: fused seed requested on kind 1, 29x27x11, 3 + 2 steps')
    rng = np.random.default_rng(24)
    m = _lib.Mlp()
    m.set_params((rng.standard_normal(MO.NPARAM) * 0.3).astype(np.float32))
    hs = _start(p, tab, ev, (1, 0, 0))
    hf = _start(p, tab, ev, (1, 0, 0))
    hs.visc_steps(0, 3)
    hf.visc_steps(0, 3)
    hs.set_root_seed(m, track_iters=True)
    hf.set_root_seed(m, track_iters=True)
    hf.set_root_split(False)
    hs.visc_steps(3, 2)
    hf.visc_steps(3, 2)
    for name in ('ev1', 'pi1', 'bulk1', 'root_seed'):
        np.testing.assert_array_equal(hf.get(name), hs.get(name), err_msg=name)
    np.testing.assert_array_equal(hf.root_iters(), hs.root_iters())
    hs.close()
    hf.close()
    m.close()
