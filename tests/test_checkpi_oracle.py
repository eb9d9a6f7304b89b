"""The oracle's pizz_o_ep (kernel_checkpi.cl:5-27, CLVisc.check_pizz) pinned by closed forms:
a uniform slice gives pi^{eta eta}/(e + P) itself, the cells with e + P <= 0.133 are left out
whatever their pi, each slice sees only its own cells, and a slice with no cell above the
threshold gives 0/0 = NaN, as in the reference."""
import numpy as np

from clvisc_amd.eos import Eos
from oracle import oracle as O
from helpers import cv_params, orc_params


def _oracle(nx, ny, nz, ev, pi):
    tab = Eos('ideal_gas').create_table()
    p = cv_params(nx, ny, nz, 0.02, 0.08, 0.08, 0.15, 0.6)
    o = O.Oracle(orc_params(p, tab), 'f32', viscous=True)
    o.load_ini(ev, pi=pi)
    return o


def _fields(nx, ny, nz, e, pzz, v=0.0):
    ev = np.zeros((nx, ny, nz, 4), np.float32)
    ev[..., 0] = e
    ev[..., 1] = v
    pi = np.zeros((nx, ny, nz, 10), np.float32)
    pi[..., 9] = pzz
    return ev.reshape(-1, 4), pi.reshape(-1, 10)


def test_uniform_slices_give_the_ratio():
    nx, ny, nz = 9, 7, 5
    e = np.array([0.5, 1.0, 3.0, 10.0, 40.0], np.float32)
    c = np.array([0.01, -0.02, 0.05, 0.1, -0.3], np.float32)
    ev, pi = _fields(nx, ny, nz, e[None, None, :], c[None, None, :] * (e[None, None, :] * (4.0 / 3.0)))
    out = _oracle(nx, ny, nz, ev, pi).pizz_o_ep()
    # ideal gas: e + P = 4e/3; the slice mean of 63 equal terms
    np.testing.assert_allclose(out, c, rtol=2e-6)


def test_cells_below_threshold_are_left_out():
    nx, ny, nz = 8, 6, 3
    rng = np.random.default_rng(5)
    # e + P = 4e/3: 0.067 (out), 0.147 (in, although e itself is below 0.133) or 2.67 (in)
    e = rng.choice(np.array([0.05, 0.11, 2.0], np.float32), (nx, ny, nz))
    pzz = np.where(e > 0.1, 0.2 * (4.0 / 3.0) * e, 1.0e6).astype(np.float32)     # huge below threshold
    ev, pi = _fields(nx, ny, nz, e, pzz)
    out = _oracle(nx, ny, nz, ev, pi).pizz_o_ep()
    np.testing.assert_allclose(out, 0.2, rtol=2e-6)


def test_slices_are_independent_and_empty_slice_is_nan():
    nx, ny, nz = 6, 5, 4
    rng = np.random.default_rng(6)
    e = rng.uniform(0.5, 5.0, (nx, ny, nz)).astype(np.float32)
    pzz = rng.uniform(-0.1, 0.1, (nx, ny, nz)).astype(np.float32)
    e[:, :, 3] = 0.01                                                            # below threshold
    ev, pi = _fields(nx, ny, nz, e, pzz)
    base = _oracle(nx, ny, nz, ev, pi).pizz_o_ep()
    assert np.isnan(base[3]) and np.all(np.isfinite(base[:3]))
    pzz2 = pzz.copy()
    pzz2[:, :, 1] *= 3.0
    ev2, pi2 = _fields(nx, ny, nz, e, pzz2)
    out = _oracle(nx, ny, nz, ev2, pi2).pizz_o_ep()
    np.testing.assert_array_equal(out[[0, 2]], base[[0, 2]])
    # float64 mean of the slice's terms (the sum order changes only the last bits)
    w = e[:, :, 1].astype(np.float64) * (4.0 / 3.0)
    np.testing.assert_allclose(out[1], np.mean(3.0 * pzz[:, :, 1] / w), rtol=1e-5)
