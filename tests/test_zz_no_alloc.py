"""cv_slice_sums_noalloc (the slice sums with the library's static scratch array, no device
allocation) against cv_slice_sums through the C-ABI: two calls in a row, the second on a
smaller grid (fewer slices than the first left in the shared array), and the capacity error
for grids beyond the array.  The device output buffers are allocated by the test through
the HIP runtime; the entry point itself allocates nothing."""
import ctypes as C
import sys

import numpy as np
import pytest

from clvisc_amd import _lib, ic
from clvisc_amd.eos import Eos
from helpers import cv_params

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg)
    sys.stdout.flush()


def _hip():
    for name in ('libamdhip64.so', '/opt/rocm/lib/libamdhip64.so'):
        try:
            return C.CDLL(name)
        except OSError:
            pass
    raise RuntimeError('libamdhip64.so not found')


class DevBuf(object):
    """n doubles of device memory, filled with NaN (0xFF bytes) so that unwritten entries show."""

    def __init__(self, hip, n):
        self.hip, self.n = hip, n
        self.ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), C.c_size_t(8 * n)) == 0
        assert hip.hipMemset(self.ptr, C.c_int(0xFF), C.c_size_t(8 * n)) == 0
        assert hip.hipDeviceSynchronize() == 0

    def to_host(self):
        out = np.empty(self.n, np.float64)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(8 * self.n), 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def _handle(dims, seed):
    tab = Eos('crossover').create_table()
    p = cv_params(*dims, 0.02, 0.08, 0.08, 0.15, 0.6, etaos_xmin=0.0, etaos_ymin=0.08, bulk=(0.05, 0.155, 0.02))
    h = _lib.Handle(p, tab, viscous=True)
    pos = ic.woods_saxon_positions(12, seed, R=0.8, a=0.1)
    ev = ic.hotspots(*dims, 0.08, 0.08, 0.15, positions=pos, emax=30.0)
    h.visc_initialize(0.6)
    h.ideal_stage(1, np.float32(0.6))
    h.set('ev1', ev)
    h.visc_steps(0, 3)
    return h


def test_slice_sums_noalloc_matches_slice_sums():
    hip = _hip()
    _lib.lib()
    for dims, seed in (((61, 57, 41), 3), ((33, 29, 9), 4)):
        _say('case: cv_slice_sums_noalloc vs cv_slice_sums on %dx%dx%d' % dims)
        h = _handle(dims, seed)
        for which in (1, 2):
            ref = h.slice_sums(which)
            buf = DevBuf(hip, dims[2] * 10)
            h.slice_sums_noalloc(buf.ptr.value, which)
            h.sync()
            got = buf.to_host().reshape(dims[2], 10)
            buf.free()
            assert np.all(np.isfinite(got))
            assert np.abs(ref).max() > 0
            np.testing.assert_array_equal(got, ref)
        h.close()


def test_slice_sums_noalloc_rejects_grids_beyond_the_array():
    hip = _hip()
    _say('case: cv_slice_sums_noalloc on 3x3x129 (beyond 128 slices)')
    h = _handle((3, 3, 129), 5)
    buf = DevBuf(hip, 129 * 10)
    with pytest.raises(_lib.CvError) as err:
        h.slice_sums_noalloc(buf.ptr.value, 1)
    assert err.value.code == _lib.CV_ERR_CAPACITY
    buf.free()
    h.close()
