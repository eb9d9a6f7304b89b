"""Entropy -> energy density conversion of TRENTo profiles (pyvisc/ebe_trento.py:26-33):
the host restatement against closed forms of the ideal gas (s proportional to e^{3/4})."""
import numpy as np

from clvisc_amd.eos import Eos
from clvisc_amd.ic import from_sd_to_ed


def test_ideal_gas_inverse():
    """s(e) = (e + P)/T with P = e/3 and T proportional to e^{1/4}: the conversion returns
    the e whose s is given, to the node spacing's interpolation error."""
    eos = Eos('ideal_gas')
    e = np.array([0.5, 1.0, 7.3, 42.0, 300.0])
    s = eos.f_S(e)
    back = from_sd_to_ed(s, eos)
    assert np.abs(back - e).max() <= 1e-4 * e.max()


def test_linear_extrapolation_beyond_the_table():
    eos = Eos('ideal_gas')
    s_top = eos.s[-1]
    slope = (eos.ed[-1] - eos.ed[-2]) / (eos.s[-1] - eos.s[-2])
    out = from_sd_to_ed(np.array([s_top * 1.1]), eos)
    assert abs(out[0] - (eos.ed[-1] + slope * 0.1 * s_top)) <= 1e-9 * out[0]
    # the first node: e(s_0) = ed_0
    assert from_sd_to_ed(np.array([eos.s[0]]), eos)[0] == eos.ed[0]
