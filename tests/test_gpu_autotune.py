"""The autotuner on the device: it times every candidate of a grid shape on a batch of
copies, keeps the fastest, writes it to its JSON cache and reads it back without timing
again; the tuned tile and chunk give the same fields as the library's default tile."""
import sys

import numpy as np
import pytest

from clvisc_amd import _lib, ic
from clvisc_amd.autotune import Autotuner, candidates
from clvisc_amd.eos import Eos
from helpers import cv_params, bit_equal_frac

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg)
    sys.stdout.flush()


def test_autotuner_picks_a_candidate_and_keeps_the_fields(tmp_path):
    nx, ny, nz = 37, 41, 13
    tab = Eos('crossover').create_table()
    p = cv_params(nx, ny, nz, 0.02, 0.08, 0.08, 0.15, 0.6, etaos_xmin=0.0, etaos_ymin=0.08,
                  bulk=(0.05, 0.155, 0.02))
    pos = ic.woods_saxon_positions(12, 5, R=0.6, a=0.1)
    ev = ic.hotspots(nx, ny, nz, 0.08, 0.08, 0.15, positions=pos, emax=50.0)
    path = str(tmp_path / 'tiles.json')
    tuner = Autotuner(p, tab, path=path, target_cells=2e5, steps=2)
    best = tuner.tune((nx, ny, nz), ev)
    _say('case: autotune %dx%dx%d viscous+bulk -> tile %dx%d chunk %d, %.3g ms per cell-update'
         % (nx, ny, nz, best[0], best[1], best[2], best[3]))
    h0 = _lib.Handle(p, tab, viscous=True)
    _, dty, dtz = h0.stage_kernel()
    assert tuple(best[:3]) in candidates(nx, ny, nz, default=(dty, dtz))
    assert best[3] > 0
    # the cache file holds the result and a new tuner takes it without timing
    again = Autotuner(p, tab, path=path)
    assert tuple(again.cache[again.key((nx, ny, nz))]) == tuple(best)
    assert again.tune((nx, ny, nz), None) == tuple(best)
    # tuned tile vs default tile: the same 12 steps, bit for bit
    h1 = _lib.Handle(p, tab, viscous=True)
    h1.set_stage_kernel(2, best[0], best[1])
    h1.set_march_chunk(best[2])
    for h in (h0, h1):
        h.visc_initialize(p.tau0)
        h.ideal_stage(1, np.float32(p.tau0))
        h.set('ev1', ev)
        h.visc_steps(0, 12)
    for name in ('ev1', 'pi1', 'bulk1'):
        a, b = h1.get(name), h0.get(name)
        _say('  %-5s bit-equal %.8f' % (name, bit_equal_frac(a, b)))
        np.testing.assert_array_equal(a, b, err_msg=name)
    h0.close()
    h1.close()
