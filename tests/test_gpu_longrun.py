"""Long evolutions of the 385x385x115 benchmark event to freeze-out on the device (the
oracle is bit-identical over the steps it can afford; these check the evolution itself):
max e never exceeds 4x the lowest max e before it (the reference's algorithm has short
transient spikes in the dilute periphery, up to ~2x; the unregulated bulk extension ran away
by orders of magnitude) and drops below e(T = 0.137 GeV) in finite time, for the
reference's shear-only physics and with the regulated bulk pressure."""
import sys

import numpy as np
import pytest

from clvisc_amd import _lib

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg)
    sys.stdout.flush()


@pytest.mark.parametrize('bulk', [False, True])
def test_benchmark_event_reaches_freeze_out(bulk):
    import bench
    p, tab, ev = bench.make_workload(bulk)
    _say('case: 385x385x115 %s to freeze-out (e_frz %.4g)' % ('viscous+bulk' if bulk else 'shear', p.efrz))
    h = _lib.Handle(p, tab, viscous=True)
    h.visc_initialize(bench.TAU0)
    h.ideal_stage(1, np.float32(bench.TAU0))
    h.set('ev1', ev)
    loop, hist = 0, [h.max_ed(1)]
    while hist[-1] >= p.efrz and loop < 1000:
        h.visc_steps(loop, 20)
        loop += 20
        hist.append(h.max_ed(1))
    tau = bench.TAU0 + loop * bench.DT
    _say('  max e every 20 steps: %s' % ' '.join('%.3g' % m for m in hist))
    growth = max(hist[k] / min(hist[:k]) for k in range(1, len(hist)))
    _say('  frozen out at tau %.2f fm after %d steps' % (tau, loop))
    h.close()
    assert hist[-1] < p.efrz
    _say('  largest max e over the lowest before it: %.3f' % growth)
    assert growth <= 4.0, hist
