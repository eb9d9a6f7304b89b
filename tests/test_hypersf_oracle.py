"""Pins of the hypersurface oracle (oracle/hypersf_oracle.c) and the surface writers.
CPU only.

  unit test cube dSigma = (1,0,0,0)     pyvisc/kernel/kernel_hypersf.cl:665-690,
                                        pyvisc/tests/test_hypersf.py:26-59
  uniform cooling field: every cube is one element with dSigma_0 = tau dx dy dz and
  tau = tau_old + r dt, r = (e_old - efrz)/(e_old - e_new)
  static planar front e(x): dSigma_x = -tau dt dy dz per cube, zero time component
  multilinear centroid interpolation reproduces fields linear in (t, x, y, eta)
"""
import os

import numpy as np
import pytest

from clvisc_amd.hypersf import write_hypersf, write_pimnsf
from oracle import oracle as O

DX, DY, DZ = 0.3, 0.25, 0.2


def coords(nx, ny, nz):
    x = (np.arange(nx) - nx // 2) * DX
    y = (np.arange(ny) - ny // 2) * DY
    z = (np.arange(nz) - nz // 2) * DZ
    return np.meshgrid(x, y, z, indexing='ij')


@pytest.mark.parametrize('prec', ['f32', 'f64'])
def test_unit_cube_known_answer(prec):
    ds = O.test_cube(2.5, prec)
    np.testing.assert_allclose(ds, [1.0, 0.0, 0.0, 0.0], atol=1e-6)


def test_heating_cube_flips_orientation():
    """Corners hotter at the new time: energy flows towards the past, so dSigma_0 is
    -tau dx dy dz -- an orientation taken from the cube's mass centre alone, or a
    dropped energy-flow test, would give +."""
    n = 8
    ev0 = np.zeros((n, 4), np.float32); ev0[:, 0] = 2.0
    ev1 = np.zeros((n, 4), np.float32); ev1[:, 0] = 3.0
    cnt, sf, _, _, _ = O.hypersf(2, 2, 2, DX, DY, DZ, 1, 1, 1, 2.5, ev0, ev1, 1.0, 1.1, precision='f64')
    assert cnt == 1
    np.testing.assert_allclose(sf[0, 0], -1.05 * DX * DY * DZ, rtol=1e-6)


def test_uniform_cooling_field():
    """FP64 restatement (same source as the FP32 one, real = double).  In FP32 the
    reference's coplanar tie-break shift (1e-7) is below the rounding of coordinates
    near 1, and the hull sum of such cubes is off by up to a few times; see
    test_fp32_tie_break_loss."""
    nx, ny, nz = 6, 5, 4
    n = nx * ny * nz
    e_old, e_new, efrz = 0.5, 0.2, 0.3
    ev0 = np.zeros((n, 4), np.float32); ev0[:, 0] = e_old; ev0[:, 1] = 0.1; ev0[:, 2] = -0.2
    ev1 = np.zeros((n, 4), np.float32); ev1[:, 0] = e_new; ev1[:, 1] = 0.1; ev1[:, 2] = -0.2
    t0, t1 = 1.0, 1.1
    cnt, sf, tx, _, _ = O.hypersf(nx, ny, nz, DX, DY, DZ, 1, 1, 1, efrz, ev0, ev1, t0, t1, precision='f64')
    assert cnt == (nx - 1) * (ny - 1) * (nz - 1)
    r = (e_old - efrz) / (e_old - e_new)
    tau = t0 + r * (t1 - t0)
    np.testing.assert_allclose(sf[:, 0], tau * DX * DY * DZ, rtol=1e-5)
    np.testing.assert_allclose(sf[:, 1:4], 0.0, atol=1e-6 * DX * DY * DZ)
    np.testing.assert_allclose(sf[:, 4], 0.1, rtol=1e-6)
    np.testing.assert_allclose(sf[:, 5], -0.2, rtol=1e-6)
    # t = tau cosh(eta), z = tau sinh(eta)
    np.testing.assert_allclose(tx[:, 0], tau * np.cosh(sf[:, 7]), rtol=1e-5)
    np.testing.assert_allclose(tx[:, 3], tau * np.sinh(sf[:, 7]), rtol=1e-5, atol=1e-6)
    # element centroids sit at the cube centres in x, y, eta
    xs = np.sort(np.unique(np.round(tx[:, 1], 5)))
    np.testing.assert_allclose(xs, (np.arange(nx - 1) - nx // 2 + 0.5) * DX, atol=1e-5)


def test_static_planar_front():
    """e = e0 - g x, the same at both times: the surface is the plane x = x0 and each
    crossing cube contributes dSigma_x = -tau dt dy dz (flow of energy towards +x)."""
    nx, ny, nz = 8, 4, 3
    X, Y, Z = coords(nx, ny, nz)
    e = (1.0 - 0.9 * (X + 0.05)).astype(np.float32).reshape(-1)
    efrz = 1.0
    ev = np.zeros((e.size, 4), np.float32); ev[:, 0] = e
    t0, t1 = 2.0, 2.05
    cnt, sf, tx, _, _ = O.hypersf(nx, ny, nz, DX, DY, DZ, 1, 1, 1, efrz, ev, ev, t0, t1, precision='f64')
    assert cnt == (ny - 1) * (nz - 1)
    tau = 0.5 * (t0 + t1)
    np.testing.assert_allclose(sf[:, 1], -tau * (t1 - t0) * DY * DZ, rtol=1e-4)
    np.testing.assert_allclose(sf[:, [0, 2, 3]], 0.0, atol=1e-4 * tau * (t1 - t0) * DY * DZ)
    np.testing.assert_allclose(tx[:, 1], -0.05, atol=1e-5)


def test_centroid_interpolation_of_linear_fields():
    """v_x, v_y, veta and pi^{mu nu} linear in (x, y, eta, t) are reproduced at each
    element's centroid; a transposed corner weight would not."""
    nx, ny, nz = 7, 6, 5
    X, Y, Z = coords(nx, ny, nz)
    t0, t1 = 1.0, 1.2
    R2 = X ** 2 + 1.3 * Y ** 2 + 0.7 * Z ** 2
    ev0 = np.zeros((nx * ny * nz, 4), np.float32)
    ev1 = np.zeros_like(ev0)
    lin = lambda t, a, b, c, d: 0.01 + a * X + b * Y + c * Z + d * t
    for ev, t, scale in ((ev0, t0, 1.0), (ev1, t1, 0.8)):
        ev[:, 0] = (scale * np.exp(-R2)).reshape(-1)
        ev[:, 1] = lin(t, 0.02, 0.0, 0.01, 0.3).reshape(-1)
        ev[:, 2] = lin(t, 0.0, -0.03, 0.02, -0.1).reshape(-1)
        ev[:, 3] = lin(t, 0.01, 0.01, 0.05, 0.2).reshape(-1)
    pis = []
    for t in (t0, t1):
        pi = np.zeros((nx * ny * nz, 10), np.float32)
        for mn in range(10):
            pi[:, mn] = lin(t, 0.001 * (mn + 1), -0.002 * mn, 0.0005 * (mn - 4), 0.003).reshape(-1)
        pis.append(pi)
    bk0 = lin(t0, 0.004, 0.001, -0.002, 0.01).astype(np.float32).reshape(-1)
    bk1 = lin(t1, 0.004, 0.001, -0.002, 0.01).astype(np.float32).reshape(-1)
    cnt, sf, tx, ps, bs = O.hypersf(nx, ny, nz, DX, DY, DZ, 1, 1, 1, 0.5, ev0, ev1, t0, t1,
                                    pi_old=pis[0], pi_new=pis[1], bk_old=bk0, bk_new=bk1, precision='f64')
    assert cnt > 20
    eta = sf[:, 7]
    tau = np.sqrt(tx[:, 0] ** 2 - tx[:, 3] ** 2)
    x, y = tx[:, 1], tx[:, 2]
    at = lambda a, b, c, d: 0.01 + a * x + b * y + c * eta + d * tau
    np.testing.assert_allclose(sf[:, 4], at(0.02, 0.0, 0.01, 0.3), atol=1e-9)
    np.testing.assert_allclose(sf[:, 5], at(0.0, -0.03, 0.02, -0.1), atol=1e-9)
    np.testing.assert_allclose(sf[:, 6], at(0.01, 0.01, 0.05, 0.2), atol=1e-9)
    for mn in range(10):
        np.testing.assert_allclose(ps[:, mn], at(0.001 * (mn + 1), -0.002 * mn, 0.0005 * (mn - 4), 0.003),
                                   atol=1e-9)
    np.testing.assert_allclose(bs, at(0.004, 0.001, -0.002, 0.01), atol=1e-9)
    # the surface of a Gaussian blob is closed: the spatial normals cancel pairwise
    assert abs(sf[:, 1].sum()) < 1e-6 * np.abs(sf[:, 1]).sum() + 1e-9


def test_skip_and_single_plane_boundary():
    """nzskip > 1 thins the cubes; NZ == 1 emits nothing, as the reference's
    (k+1)*nzskip < NZ test does."""
    nx, ny = 9, 9
    e_old, e_new = 0.5, 0.2
    for nz, skip, expect in ((1, 1, 0), (9, 2, 4 * 4 * 4), (9, 4, 2 * 2 * 2)):
        n = nx * ny * nz
        ev0 = np.zeros((n, 4), np.float32); ev0[:, 0] = e_old
        ev1 = np.zeros((n, 4), np.float32); ev1[:, 0] = e_new
        cnt, sf, _, _, _ = O.hypersf(nx, ny, nz, DX, DY, DZ, skip, skip, skip, 0.3, ev0, ev1, 1.0, 1.1,
                                     precision='f64')
        assert cnt == expect, (nz, skip, cnt)
        if cnt:
            np.testing.assert_allclose(sf[:, 0], (1.0 + 0.1 * 2.0 / 3.0) * skip ** 3 * DX * DY * DZ, rtol=1e-5)


def noisy_blob(nx=21, ny=19, nz=9, seed=1):
    X, Y, Z = coords(nx, ny, nz)
    R2 = X ** 2 + 1.3 * Y ** 2 + 0.7 * Z ** 2
    rng = np.random.default_rng(seed)
    ev0 = np.zeros((X.size, 4), np.float32)
    ev1 = np.zeros_like(ev0)
    ev0[:, 0] = (np.exp(-R2 / 4) * (1 + 0.05 * rng.random(X.shape))).reshape(-1)
    ev1[:, 0] = (0.8 * np.exp(-R2 / 4) * (1 + 0.05 * rng.random(X.shape))).reshape(-1)
    ev0[:, 1] = (0.1 * X).reshape(-1)
    ev1[:, 1] = (0.12 * X).reshape(-1)
    return ev0, ev1


def test_fp32_tie_break_loss():
    """FP32 and FP64 find the same cubes; the surface totals agree to 1e-3 while single
    elements with coplanar intersections can differ by O(1) in FP32 (the reference's
    precision) -- the reason for the FP64 surface option."""
    ev0, ev1 = noisy_blob()
    out = {p: O.hypersf(21, 19, 9, DX, DY, DZ, 1, 1, 1, 0.5, ev0, ev1, np.float32(1.0), np.float32(1.1),
                        precision=p) for p in ('f32', 'f64')}
    assert out['f32'][0] == out['f64'][0] > 300
    a, b = out['f32'][1], out['f64'][1]
    np.testing.assert_allclose(a[:, 7], b[:, 7], rtol=1e-5, atol=1e-6)
    assert abs(a[:, 0].sum() - b[:, 0].sum()) < 1e-3 * b[:, 0].sum()
    # the spatial normals of a closed surface nearly cancel
    assert abs(b[:, 1].sum()) < 5e-3 * np.abs(b[:, 1]).sum()


def test_capacity_counts_overflow():
    nx, ny, nz = 6, 5, 4
    n = nx * ny * nz
    ev0 = np.zeros((n, 4), np.float32); ev0[:, 0] = 0.5
    ev1 = np.zeros((n, 4), np.float32); ev1[:, 0] = 0.2
    cnt, sf, _, _, _ = O.hypersf(nx, ny, nz, DX, DY, DZ, 1, 1, 1, 0.3, ev0, ev1, 1.0, 1.1, cap=7)
    assert cnt == 5 * 4 * 3 and sf.shape[0] == 7


def test_writers_format(tmp_path):
    """Headers the downstream sampler parses: '# Tfrz=' (substr(7)) and
    '# one_o_2TsqrEplusP=' (substr(20)) (sampler/src/sampler.cc:126-129, 182-184)."""
    sf = np.arange(16, dtype=np.float32).reshape(2, 8) * 0.1
    tx = np.arange(8, dtype=np.float32).reshape(2, 4)
    pi = np.arange(20, dtype=np.float32).reshape(2, 10) * 1e-3
    write_hypersf(str(tmp_path), sf, tx, 0.137)
    write_pimnsf(str(tmp_path), pi, 0.3135, 0.137, 0.06)
    lines = open(os.path.join(tmp_path, 'hypersf.dat')).read().splitlines()
    assert lines[0].startswith('# Tfrz=')
    assert float(lines[0][7:].split()[0]) == pytest.approx(0.137)
    assert lines[1].split() == ['%.6e' % v for v in sf[0]]
    lines = open(os.path.join(tmp_path, 'pimnsf.dat')).read().splitlines()
    c = float(lines[0][20:].split()[0])
    assert c == pytest.approx(1.0 / (2 * 0.137 ** 2 * (0.3135 + 0.06)), rel=1e-6)
    np.testing.assert_allclose(np.loadtxt(os.path.join(tmp_path, 'pimnsf.dat')), pi, rtol=1e-6)
    np.testing.assert_allclose(np.loadtxt(os.path.join(tmp_path, 'sf_txyz.dat')), tx)
