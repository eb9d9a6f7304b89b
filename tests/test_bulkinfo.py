"""BulkInfo observables (pyvisc/bulkinfo.py:203-260) on fields with closed-form answers: the
host formulas that the device reduction (cv_slice_sums) is compared against on the GPU."""
import numpy as np
import pytest

from clvisc_amd.bulkinfo import BulkInfo
from clvisc_amd.config import read_config
from clvisc_amd.eos import Eos


class _Field(object):
    """Host stand-in of a device field proxy (get / gather in AoS)."""

    def __init__(self, ev):
        self.ev = np.ascontiguousarray(ev, np.float32)

    def get(self):
        return self.ev

    def gather(self, cells):
        return self.ev[np.asarray(cells)]


def _cfg(nx, ny, nz):
    cfg = read_config()
    cfg.NX, cfg.NY, cfg.NZ = nx, ny, nz
    return cfg


def test_eccp_closed_forms():
    """Uniform e and velocity v_x: <Txx - Tyy>/<Txx + Tyy> = w vx^2 / (w vx^2 + 2 P) and
    <T^tx>/<Txx + Tyy> = w vx / (w vx^2 + 2 P), w = (e + P) u0^2; at rest both vanish."""
    eos = Eos('ideal_gas')
    bi = BulkInfo(_cfg(9, 7, 1), eos)
    e = np.full((9, 7), 3.0, np.float32)
    for vx in (0.0, 0.3, 0.8):
        v = np.full((9, 7), vx, np.float32)
        ecc1, ecc2 = bi.eccp(e, v, np.zeros_like(v))
        P = float(eos.f_P(np.float32(3.0)))
        w = (3.0 + P) / (1.0 - float(np.float32(vx)) ** 2)
        vx32 = float(np.float32(vx))
        assert ecc2 == pytest.approx(w * vx32 ** 2 / (w * vx32 ** 2 + 2 * P), rel=1e-5, abs=1e-7)
        assert ecc1 == pytest.approx(w * vx32 / (w * vx32 ** 2 + 2 * P), rel=1e-5, abs=1e-7)
    # a field elongated along y with radial flow: positive momentum anisotropy along x
    x = np.linspace(-2, 2, 41)
    X, Y = np.meshgrid(x, x, indexing='ij')
    ed = (10 * np.exp(-X ** 2 / 0.5 - Y ** 2 / 2.0)).astype(np.float32)
    vx, vy = (0.2 * X).astype(np.float32), (0.2 * Y).astype(np.float32)
    bi2 = BulkInfo(_cfg(41, 41, 1), eos)
    # pressure gradients are larger along x for this shape; the flow here is isotropic, so
    # the sign follows the e-weighted <vx^2 - vy^2> < 0 (more matter at large |y|)
    _, e2 = bi2.eccp(ed, vx, vy)
    w = ed * (1 + eos.f_P(ed) / np.maximum(ed, 1e-10))
    assert np.sign(e2) == np.sign((w * (vx ** 2 - vy ** 2)).sum())


def test_get_records_slices_and_ecc_vs_rapidity():
    """get() on a host field: the three lines, two planes and one eccentricity per eta slice;
    a field independent of eta gives the same eccentricity in every slice."""
    eos = Eos('ideal_gas')
    nx, ny, nz = 11, 9, 5
    rng = np.random.default_rng(3)
    e2d = rng.uniform(0.5, 5.0, (nx, ny))
    v2d = rng.uniform(-0.3, 0.3, (nx, ny, 2))
    ev = np.zeros((nx, ny, nz, 4), np.float32)
    ev[..., 0] = e2d[:, :, None]
    ev[..., 1] = v2d[:, :, None, 0]
    ev[..., 2] = v2d[:, :, None, 1]
    bi = BulkInfo(_cfg(nx, ny, nz), eos)
    bi.get(0.6, _Field(ev.reshape(-1, 4)), float(e2d.max()))
    np.testing.assert_array_equal(bi.ex[0], ev[:, ny // 2, nz // 2, 0])
    np.testing.assert_array_equal(bi.exz[0], ev[:, ny // 2, :, 0])
    e1, e2 = bi.ecc1_vs_rapidity[0], bi.ecc2_vs_rapidity[0]
    assert e1.shape == (nz,) and np.ptp(e2) < 1e-12 and np.ptp(e1) < 1e-12
    ref = bi.eccp(ev[:, :, 0, 0], ev[:, :, 0, 1], ev[:, :, 0, 2], ev[:, :, 0, 3])
    assert e2[0] == ref[1] and e1[0] == ref[0]
    # the public ecc_vs_rapidity(d_ev) of the reference appends the same values
    bi.ecc_vs_rapidity(_Field(ev.reshape(-1, 4)))
    assert len(bi.ecc2_vs_rapidity) == 2
    np.testing.assert_array_equal(bi.ecc2_vs_rapidity[1], e2)
    np.testing.assert_array_equal(bi.ecc1_vs_rapidity[1], e1)


def test_total_energy_and_entropy_closed_forms():
    """The lab-frame energy of total_energy_and_entropy (kernel_bulkinfo.cl:18-46) is
    T^{t mu} dSigma_mu on the tau = const surface, dSigma_mu = tau (cosh eta, 0, 0,
    -sinh eta) dx dy deta: checked against that contraction with T^{t mu} built from the
    Cartesian four-velocity u^t = u0 (cosh eta + v sinh eta), u^z = u0 (sinh eta + v cosh eta)
    on a random field; at rest and uniform it is e sum(cosh eta_k), and the entropy u0 s."""
    eos = Eos('ideal_gas')
    nx, ny, nz = 5, 4, 9
    cfg = _cfg(nx, ny, nz)
    cfg.DX, cfg.DY, cfg.DZ = 0.2, 0.3, 0.25
    bi = BulkInfo(cfg, eos)
    tau = 1.7
    vol = tau * 0.2 * 0.3 * 0.25
    eta = (np.arange(nz) - nz // 2) * 0.25
    # at rest, uniform e
    ev = np.zeros((nx, ny, nz, 4))
    ev[..., 0] = 2.5
    E, S = bi.total_energy_and_entropy(tau, ev)
    assert E == pytest.approx(2.5 * nx * ny * np.cosh(eta).sum() * vol, rel=1e-12)
    assert S == pytest.approx(float(eos.f_S(2.5)) * nx * ny * nz * vol, rel=1e-12)
    # random field: the contraction
    rng = np.random.default_rng(5)
    ev[..., 0] = rng.uniform(0.5, 5.0, (nx, ny, nz))
    v = rng.uniform(-0.5, 0.5, (nx, ny, nz, 3))
    ev[..., 1:] = v
    E, S = bi.total_energy_and_entropy(tau, ev)
    e, pr = ev[..., 0], eos.f_P(ev[..., 0])
    u0 = 1.0 / np.sqrt(1.0 - (v * v).sum(-1))
    ch, sh = np.cosh(eta)[None, None, :], np.sinh(eta)[None, None, :]
    ut = u0 * (ch + v[..., 2] * sh)
    uz = u0 * (sh + v[..., 2] * ch)
    Ttt = (e + pr) * ut * ut - pr
    Ttz = (e + pr) * ut * uz
    ref = (tau * (ch * Ttt - sh * Ttz)).sum() * 0.2 * 0.3 * 0.25
    assert E == pytest.approx(ref, rel=1e-12)
    assert S == pytest.approx((u0 * eos.f_S(e)).sum() * vol, rel=1e-12)
    # the same from per-slice sums (the device reduction's ten columns)
    s = np.zeros((nz, 10))
    w = (e + pr) * u0 * u0
    s[:, 6] = w.sum((0, 1))
    s[:, 7] = (w * v[..., 2]).sum((0, 1))
    s[:, 8] = pr.sum((0, 1))
    s[:, 9] = (u0 * eos.f_S(e)).sum((0, 1))
    E2, S2 = bi._energy_entropy(tau, s)
    assert E2 == pytest.approx(ref, rel=1e-12)
    assert S2 == pytest.approx(S, rel=1e-12)
