"""The short reciprocal and square root of the stage kernels (hardware reciprocal / reciprocal
square root + one FMA Newton step, used where the operand lies in their range by construction)
equal the correctly rounded 1/b and sqrtf for every float of those ranges, so the stage
kernels keep the oracle's FP32 results bit for bit."""
import sys

import pytest

from clvisc_amd import _lib

pytestmark = pytest.mark.gpu


def test_short_forms_equal_correctly_rounded_ones_on_their_ranges():
    print('case: rcpn vs 1/b on 2^-126 <= |b| < 2^126, sqrtn vs sqrtf on [2^-24, 1], every float')
    sys.stdout.flush()
    (rbad, rn), (sbad, sn) = _lib.selftest_fastmath(0)
    print('  reciprocal: %d floats checked, %d differ; square root: %d checked, %d differ' % (rn, rbad, sn, sbad))
    assert rn == 2 * 252 * 2 ** 23
    assert sn == 0x3f800000 - 0x33800000 + 1
    assert rbad == 0
    assert sbad == 0
