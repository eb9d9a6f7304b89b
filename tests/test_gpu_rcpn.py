"""The reciprocal of the stage kernels (hardware reciprocal + one FMA Newton step, used where
the operand lies in [2^-126, 2^126) by construction) equals the correctly rounded 1/b for
every float of that range, so the stage kernels keep the oracle's FP32 results bit for bit."""
import sys

import pytest

from clvisc_amd import _lib

pytestmark = pytest.mark.gpu


def test_rcpn_equals_correctly_rounded_reciprocal_on_its_range():
    print('case: rcpn vs 1/b, every float with 2^-126 <= |b| < 2^126')
    sys.stdout.flush()
    bad, n = _lib.selftest_rcpn(0)
    print('  %d floats checked, %d differ' % (n, bad))
    assert n == 2 * 252 * 2 ** 23
    assert bad == 0
