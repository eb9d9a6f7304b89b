"""Initial conditions for the hydro workloads (host side, NumPy).

These produce the (N,4) ``(e, vx, vy, veta)`` arrays that ``CLIdeal.load_ini``
takes (pyvisc/ideal.py:138-144), plus pi^{mu nu} where an analytic flow fixes it.

* ``gubser_ideal``  -- ideal conformal Gubser flow (pyvisc/tests/test_gubser_ideal.py:22-26)
* ``gubser_visc``   -- 2nd-order viscous Gubser flow of pyvisc/kernel/kernel_gubser_visc.cl
                       (used with the ``gubser_visc_test`` switch)
* ``bjorken``       -- uniform energy density (pyvisc/tests/test_bjorken.py:37-45)
* ``hotspots``      -- seeded sum of 2D Gaussian hotspots at Woods-Saxon nucleon positions
                       times a longitudinal plateau with Gaussian tails (the 385x385x115
                       viscous workload); the eta profile follows pyvisc/kernel/kernel_glauber.cl:35-47.
"""
from __future__ import annotations

import math

import numpy as np

HBARC = 0.1973269


def grid_coords(nx, ny, nz, dx, dy, dz):
    """x_i = (i - NX//2) dx (pyvisc/kernel/kernel_hypersf.cl:475-477)."""
    x = (np.arange(nx) - nx // 2) * dx
    y = (np.arange(ny) - ny // 2) * dy
    z = (np.arange(nz) - nz // 2) * dz
    return x, y, z


def gubser_ed(tau, r, q):
    return (2.0 * q) ** (8.0 / 3.0) / (tau * (1 + 2 * q * q * (tau * tau + r * r)
                                           + q ** 4 * (tau * tau - r * r) ** 2)) ** (4.0 / 3.0)


def gubser_vr(tau, r, q):
    return 2.0 * q * q * tau * r / (1.0 + q * q * tau * tau + q * q * r * r)


def gubser_ideal(nx, ny, nz, dx, dy, dz, tau, q=1.0, e0=None, offset=0.0):
    """Gubser ideal flow at proper time tau on the grid.

    ``e0``: if given, normalise so that e(tau, r=0) = e0 at this tau.
    ``offset``: 0 puts x on (i - NX//2) dx; 0.5 reproduces the reference test's
    (i - 0.5 NX) dx convention when NX is odd.
    Returns (ev (N,4) float64, norm) where e = norm * gubser_ed.
    """
    x, y, _ = grid_coords(nx, ny, nz, dx, dy, dz)
    x = x - offset * dx
    y = y - offset * dy
    X, Y = np.meshgrid(x, y, indexing='ij')
    R = np.sqrt(X * X + Y * Y)
    ed = gubser_ed(tau, R, q)
    norm = 1.0 if e0 is None else e0 / gubser_ed(tau, 0.0, q)
    ed = ed * norm
    vr = gubser_vr(tau, R, q)
    with np.errstate(invalid='ignore', divide='ignore'):
        vx = np.where(R > 0, vr * X / R, 0.0)
        vy = np.where(R > 0, vr * Y / R, 0.0)
    ev = np.zeros((nx, ny, nz, 4))
    ev[..., 0] = ed[:, :, None]
    ev[..., 1] = vx[:, :, None]
    ev[..., 2] = vy[:, :, None]
    return ev.reshape(-1, 4), norm


def bjorken(n, e0=30.0):
    ev = np.zeros((n, 4))
    ev[:, 0] = e0
    return ev


# ---------------------------------------------------------------------------------
# 2nd-order viscous Gubser flow (closed forms of pyvisc/kernel/kernel_gubser_visc.cl,
# re-derived in NumPy double precision).
# ---------------------------------------------------------------------------------
def _gv_common(tau, x, y, L, lam1):
    r2 = x * x + y * y
    A = 0.25 * (4 * L * L * tau * tau + (L * L - tau * tau + r2) ** 2) / (L * L * tau * tau)
    pw = A ** ((-1.33333333333333 * lam1 + 1) / lam1)
    B = 4 * L * L * tau * tau + (L * L - tau * tau + r2) ** 2
    Cc = 4 * L * L * r2 + (L * L + tau * tau - r2) ** 2
    S = L * L + tau * tau + r2
    return r2, pw, B, Cc, S


def gubser_visc(nx, ny, nz, dx, dy, dz, tau, L=2.0, lam1=-10.0):
    """Energy density, velocity and pi^{mu nu} (tilde-scaled eta components) of the
    viscous Gubser solution used by the reference's GUBSER_VISC_TEST.
    Returns (ev (N,4), pi (N,10)) in float64."""
    xs, ys, _ = grid_coords(nx, ny, nz, dx, dy, dz)
    X, Y = np.meshgrid(xs, ys, indexing='ij')
    r2, pw, B, Cc, S = _gv_common(tau, X, Y, L, lam1)
    eps = (1 + 0.25 * (-L * L + tau * tau - r2) ** 2 / (L * L * tau * tau)) ** (-1.33333333333333 + 1.0 / lam1) / tau ** 4
    ut = (-4 * tau * tau * r2 / S ** 2 + 1) ** -0.5
    ux = 2 * tau * X / (np.sqrt(-4 * tau * tau * r2 / S ** 2 + 1) * S)
    uy = 2 * tau * Y / (np.sqrt(-4 * tau * tau * r2 / S ** 2 + 1) * S)
    pitt = -4 * pw * r2 * B / (lam1 * tau ** 2 * Cc ** 2)
    pitx = -2 * X * pw * B * S / (lam1 * tau ** 3 * Cc ** 2)
    pity = -2 * Y * pw * B * S / (lam1 * tau ** 3 * Cc ** 2)
    Xs = np.where(np.abs(X) < 1e-6, 1e-6, X)
    Ys = np.where(np.abs(Y) < 1e-6, 1e-6, Y)
    r2s, pws, Bs, Ccs, Ss = _gv_common(tau, Xs, Ys, L, lam1)
    pixx = -pws * Bs * (Xs ** 2 * Ss ** 2 + Ys ** 2 * Ccs) / (lam1 * tau ** 4 * r2s * Ccs ** 2)
    pixy = Xs * Ys * pws * Bs * (Ccs - Ss ** 2) / (lam1 * tau ** 4 * r2s * Ccs ** 2)
    piyy = -pws * Bs * (Xs ** 2 * Ccs + Ys ** 2 * Ss ** 2) / (lam1 * tau ** 4 * r2s * Ccs ** 2)
    pizz = 2 * pw / (lam1 * tau ** 6)
    ev = np.zeros((nx, ny, nz, 4))
    ev[..., 0] = eps[:, :, None]
    ev[..., 1] = (ux / ut)[:, :, None]
    ev[..., 2] = (uy / ut)[:, :, None]
    pi = np.zeros((nx, ny, nz, 10))
    for c, v in ((0, pitt), (1, pitx), (2, pity), (4, pixx), (5, pixy), (7, piyy)):
        pi[..., c] = v[:, :, None]
    pi[..., 9] = (tau * tau * pizz)[:, :, None]
    return ev.reshape(-1, 4), pi.reshape(-1, 10)


# ---------------------------------------------------------------------------------
# Seeded hotspot initial condition
# ---------------------------------------------------------------------------------
def _splitmix64(seed, n):
    """SplitMix64 stream -> n uniform doubles in [0,1)."""
    with np.errstate(over='ignore'):
        i = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def woods_saxon_positions(nhot, seed, R=6.62, a=0.546):
    """Transverse (x, y) of ``nhot`` nucleons sampled from a 3D Woods-Saxon density
    by rejection sampling inside a sphere of radius R + 10 a (SplitMix64 stream)."""
    rmax = R + 10.0 * a
    pos = []
    batch = 4096
    k = 0
    while len(pos) < nhot:
        u = _splitmix64(seed * 1000003 + k, 4 * batch).reshape(batch, 4)
        k += 1
        xyz = (2.0 * u[:, :3] - 1.0) * rmax
        r = np.sqrt((xyz ** 2).sum(1))
        acc = (r <= rmax) & (u[:, 3] < 1.0 / (1.0 + np.exp((r - R) / a)))
        for p in xyz[acc]:
            pos.append(p[:2])
            if len(pos) == nhot:
                break
    return np.array(pos)


def eta_profile(z, flat=1.3, sigma=1.5):
    """H(eta) = 1 for |eta| <= flat, else exp(-(|eta|-flat)^2 / (2 sigma^2))."""
    az = np.abs(z)
    return np.where(az <= flat, 1.0, np.exp(-(az - flat) ** 2 / (2.0 * sigma * sigma)))


def hotspots(nx, ny, nz, dx, dy, dz, nhot=400, width=0.5, emax=50.0, seed=20240601,
             positions=None, flat=1.3, sigma_eta=1.5, dtype=np.float32):
    """Sum of ``nhot`` 2D Gaussian hotspots (width ``width`` fm) at Woods-Saxon
    positions, times the eta plateau, normalised to a peak of ``emax`` GeV/fm^3.
    v = 0.  Returns (N,4) ``dtype``."""
    x, y, z = grid_coords(nx, ny, nz, dx, dy, dz)
    if positions is None:
        positions = woods_saxon_positions(nhot, seed)
    e2 = np.zeros((nx, ny))
    inv = 1.0 / (2.0 * width * width)
    for (px, py) in positions:
        gx = np.exp(-(x - px) ** 2 * inv)
        gy = np.exp(-(y - py) ** 2 * inv)
        e2 += np.outer(gx, gy)
    h = eta_profile(z, flat, sigma_eta)
    peak = e2.max() * h.max()
    e2 *= emax / peak
    ev = np.zeros((nx, ny, nz, 4), dtype=dtype)
    ev[..., 0] = (e2[:, :, None] * h[None, None, :]).astype(dtype)
    return ev.reshape(-1, 4)


def from_sd_to_ed(entropy, eos):
    """Entropy density -> energy density through the EOS nodes, piecewise linear in s with
    linear extrapolation at both ends (pyvisc/ebe_trento.py:26-33, where
    InterpolatedUnivariateSpline(eos.s, eos.ed, k=1) does this).  Host version; the device
    version is ``sd_to_ed_on_device``."""
    s = np.asarray(eos.s, np.float64)
    ed = np.asarray(eos.ed, np.float64)
    x = np.asarray(entropy, np.float64)
    out = np.interp(x, s, ed)
    lo, hi = x < s[0], x > s[-1]
    out[lo] = ed[0] + (x[lo] - s[0]) / (s[1] - s[0]) * (ed[1] - ed[0])
    out[hi] = ed[-2] + (x[hi] - s[-2]) / (s[-1] - s[-2]) * (ed[-1] - ed[-2])
    return out


def sd_to_ed_on_device(d_ev, eos):
    """Convert the e component of a device field that holds an entropy density (a TRENTo
    profile loaded with ``set``) to energy density in place (cv_sd_to_ed)."""
    d_ev.handle.sd_to_ed(eos.s, eos.ed, d_ev.field_id)
