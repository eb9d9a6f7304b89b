"""Event batches (one persistent launch per RK stage over events with their own grids,
graph-replayed steps) vs the same events stepped one by one.  The batched kernel runs the
same per-work-item code with the stage constants rebuilt on the device, so the fields are
expected to be identical bit for bit."""
import sys

import numpy as np
import pytest

from clvisc_amd import _lib, ic
from clvisc_amd.eos import Eos
from helpers import cv_params, bit_equal_frac

pytestmark = pytest.mark.gpu

DIMS = [(29, 31, 11), (41, 37, 1), (23, 25, 9), (33, 29, 1), (19, 21, 13)]


def _say(msg):
    print(msg)
    sys.stdout.flush()


def _ic(d, seed):
    pos = ic.woods_saxon_positions(10 + seed, seed, R=0.6, a=0.1)
    return ic.hotspots(d[0], d[1], d[2], 0.08, 0.08, 0.15, positions=pos, emax=40.0 + seed)


def _init(h, ev, tau0):
    h.visc_initialize(tau0)
    h.ideal_stage(1, np.float32(tau0))
    h.set('ev1', ev)


def _compare(hb, hs, label):
    for name in ('ev1', 'pi1', 'ev0', 'pi0', 'ev2', 'udiff', 'goodcell') + (
            ('bulk1',) if hs.params.flags & _lib.CV_FLAG_BULK else ()):
        g, r = hb.get(name), hs.get(name)
        frac = bit_equal_frac(g, r)
        if frac != 1.0:
            _say('  %s %s bit-equal %.6f' % (label, name, frac))
        np.testing.assert_array_equal(g, r, err_msg='%s %s' % (label, name))


@pytest.mark.parametrize('bulk,nxcd', [(False, None), (True, None), (True, 0), (False, 5), (True, 16)])
def test_batch_matches_single_events(bulk, nxcd):
    """Default work ranges (8), one shared counter (0) and range counts that do not divide
    the work list: the schedule changes, the fields do not."""
    tab = Eos('crossover').create_table()
    p = cv_params(1, 1, 1, 0.02, 0.08, 0.08, 0.15, 0.6, etaos_xmin=0.0, etaos_ymin=0.08,
                  bulk=(0.05, 0.155, 0.02) if bulk else None)
    _say('case: batch of %d events %s, bulk=%s, work ranges %s' % (len(DIMS), DIMS, bulk, nxcd))
    batch = _lib.Batch(p, tab, DIMS)
    if nxcd is not None:
        batch.set_xcd_map(nxcd)
    singles = []
    for e, d in enumerate(DIMS):
        ev = _ic(d, e)
        _init(batch.handle(e), ev, p.tau0)
        q = _lib.CvParams.from_buffer_copy(p)
        q.nx, q.ny, q.nz = d
        hs = _lib.Handle(q, tab, viscous=True)
        _init(hs, ev, p.tau0)
        singles.append(hs)
    # per-event tiles and chunks differ from the defaults for two events
    batch.handle(0).set_stage_kernel(2, 7, 16)
    batch.handle(0).set_march_chunk(4)
    batch.handle(2).set_march_chunk(5)
    loop = 0
    for n in (3, 4, 1):
        batch.steps(n)
        for hs in singles:
            hs.visc_steps(loop, n)
        loop += n
        _say('  after %d steps (batch info %s)' % (loop, batch.info()))
        for e, hs in enumerate(singles):
            _compare(batch.handle(e), hs, 'event %d %s' % (e, DIMS[e]))
    # event 1 leaves the batch; the others go on
    mask = [1, 0, 1, 1, 1]
    batch.set_active(mask)
    batch.steps(3)
    for e, hs in enumerate(singles):
        if mask[e]:
            hs.visc_steps(loop, 3)
    loop += 3
    _say('  after %d steps with event 1 inactive' % loop)
    for e, hs in enumerate(singles):
        _compare(batch.handle(e), hs, 'event %d (mask) %s' % (e, DIMS[e]))
    batch.close()


def test_batch_fp64_root_matches_single_events():
    """CV_FLAG_FP64_ROOT (FP64 root solve where M > fp64_root_vmin T^tt): the batch selects
    the stage kernels with that branch and stays bit-identical to the single events."""
    tab = Eos('crossover').create_table()
    p = cv_params(1, 1, 1, 0.02, 0.08, 0.08, 0.15, 0.6, etaos_xmin=0.0, etaos_ymin=0.08,
                  bulk=(0.05, 0.155, 0.02), flags=_lib.CV_FLAG_FP64_ROOT, fp64_root_vmin=0.3)
    dims = DIMS[:3]
    _say('case: batch of %d events with the FP64 root (vmin 0.3), bulk' % len(dims))
    batch = _lib.Batch(p, tab, dims)
    singles = []
    for e, d in enumerate(dims):
        ev = _ic(d, e)
        _init(batch.handle(e), ev, p.tau0)
        q = _lib.CvParams.from_buffer_copy(p)
        q.nx, q.ny, q.nz = d
        hs = _lib.Handle(q, tab, viscous=True)
        _init(hs, ev, p.tau0)
        singles.append(hs)
    batch.steps(9)
    for hs in singles:
        hs.visc_steps(0, 9)
    for e, hs in enumerate(singles):
        _compare(batch.handle(e), hs, 'event %d %s' % (e, dims[e]))
    batch.close()


def test_batch_refuses_a_jet_source():
    """The batched stage kernel has no jet source: a jet set on a member handle makes
    cv_batch_steps fail with CV_ERR_UNSUPPORTED instead of being dropped; cleared, it runs."""
    tab = Eos('crossover').create_table()
    p = cv_params(1, 1, 1, 0.02, 0.08, 0.08, 0.15, 0.6, etaos_xmin=0.0, etaos_ymin=0.08)
    dims = DIMS[:2]
    _say('case: batch of %d events, jet on event 1' % len(dims))
    batch = _lib.Batch(p, tab, dims)
    for e, d in enumerate(dims):
        _init(batch.handle(e), _ic(d, e), p.tau0)
    batch.handle(1).set_jet_source({'switch_on': True, 'start_pos_index': 0, 'direction': 0.0})
    with pytest.raises(_lib.CvError) as err:
        batch.steps(1)
    assert err.value.code == _lib.CV_ERR_UNSUPPORTED
    batch.handle(1).set_jet_source(None)
    batch.steps(1)
    batch.close()


@pytest.mark.parametrize('fp64,eos_name', [(False, 'crossover'), (True, 'crossover'), (False, 'ideal_gas'),
                                            (True, 'ideal_gas')])
def test_seeded_batch_matches_seeded_single_events(fp64, eos_name):
    """cv_batch_set_root_seed gives each event's fields exactly as its own seeded handle does
    (FP64 root for fast cells when CV_FLAG_FP64_ROOT is set): with the octave table (crossover)
    the network runs inside the persistent stage launch, with the row/col table (ideal gas)
    the stages run split, then the seed network and the root solve over all events, in one
    graph.  Deactivating an event keeps it so."""
    tab = Eos(eos_name).create_table()
    flags = _lib.CV_FLAG_FP64_ROOT if fp64 else 0
    p = cv_params(1, 1, 1, 0.02, 0.08, 0.08, 0.15, 0.6, etaos_xmin=0.0, etaos_ymin=0.08,
                  bulk=(0.05, 0.155, 0.02), flags=flags, fp64_root_vmin=0.3)
    dims = DIMS[:4]
    _say('case: seeded batch of %d events, FP64 root %s, %s EOS' % (len(dims), fp64, eos_name))
    rng = np.random.default_rng(31)
    mlp = _lib.Mlp()
    mlp.set_params((rng.standard_normal(mlp.nparam) * 0.3).astype(np.float32))
    batch = _lib.Batch(p, tab, dims)
    singles = []
    for e, d in enumerate(dims):
        ev = _ic(d, e)
        _init(batch.handle(e), ev, p.tau0)
        q = _lib.CvParams.from_buffer_copy(p)
        q.nx, q.ny, q.nz = d
        hs = _lib.Handle(q, tab, viscous=True)
        _init(hs, ev, p.tau0)
        hs.set_root_seed(mlp)
        singles.append(hs)
    batch.set_root_seed(mlp)
    batch.steps(3)
    for hs in singles:
        hs.visc_steps(0, 3)
    for e, hs in enumerate(singles):
        _compare(batch.handle(e), hs, 'event %d %s' % (e, dims[e]))
    mask = [1, 1, 0, 1]
    batch.set_active(mask)
    batch.steps(2)
    for e, hs in enumerate(singles):
        if mask[e]:
            hs.visc_steps(3, 2)
    for e, hs in enumerate(singles):
        _compare(batch.handle(e), hs, 'event %d (mask) %s' % (e, dims[e]))
    # back to the fused stages: equal to unseeded single events from here on
    batch.set_root_seed(None)
    for hs in singles:
        hs.set_root_seed(None)
        hs.set_root_split(False)
    batch.set_active([1, 1, 1, 1])
    batch.steps(1)
    for e, hs in enumerate(singles):
        hs.visc_steps(5 if mask[e] else 3, 1)
    for e, hs in enumerate(singles):
        if mask[e]:
            _compare(batch.handle(e), hs, 'event %d (unseeded) %s' % (e, dims[e]))
    batch.close()
    mlp.close()
