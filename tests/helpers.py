"""Shared test helpers: parameter builders for the HIP library and the oracle,
and the pre-registered parity metric."""
import numpy as np

from clvisc_amd import _lib
from oracle import oracle as O


def cv_params(nx, ny, nz, dt, dx, dy, dz, tau0, etaos_xmin=0.18, etaos_ymin=0.08,
              etaos_left_slop=0.0, etaos_right_slop=0.0, lam1=-10.0, flags=0,
              bulk=None, fp64_root_vmin=0.99):
    p = _lib.CvParams()
    p.nx, p.ny, p.nz = nx, ny, nz
    p.tau0, p.dt = tau0, dt
    p.dx, p.dy, p.dz = dx, dy, dz
    p.etaos_xmin, p.etaos_ymin = etaos_xmin, etaos_ymin
    p.etaos_left_slop, p.etaos_right_slop, p.lam1 = etaos_left_slop, etaos_right_slop, lam1
    if bulk is not None:
        flags |= _lib.CV_FLAG_BULK
        p.zetaos_max, p.zetaos_tpeak, p.zetaos_width = bulk
    p.fp64_root_vmin = fp64_root_vmin
    p.flags = flags
    p.nxskip = p.nyskip = p.nzskip = 1
    return p


def orc_params(cvp, table, nthreads=0):
    """Oracle parameters matching a CvParams."""
    bulk = None
    if cvp.flags & _lib.CV_FLAG_BULK:
        bulk = (cvp.zetaos_max, cvp.zetaos_tpeak, cvp.zetaos_width)
    return O.make_params(cvp.nx, cvp.ny, cvp.nz, cvp.dt, cvp.dx, cvp.dy, cvp.dz, cvp.tau0, table,
                         etaos_xmin=cvp.etaos_xmin, etaos_ymin=cvp.etaos_ymin,
                         etaos_left_slop=cvp.etaos_left_slop, etaos_right_slop=cvp.etaos_right_slop,
                         lam1=cvp.lam1, riemann_test=bool(cvp.flags & _lib.CV_FLAG_RIEMANN_TEST),
                         gubser_visc_test=bool(cvp.flags & _lib.CV_FLAG_GUBSER_VISC_TEST),
                         bulk=bulk,
                         fp64_root_vmin=cvp.fp64_root_vmin if cvp.flags & _lib.CV_FLAG_FP64_ROOT else None,
                         nthreads=nthreads,
                         couplings=(1 if cvp.flags & _lib.CV_FLAG_PIMN_OMEGA else 0)
                         | (2 if cvp.flags & _lib.CV_FLAG_OMEGA_OMEGA else 0))


def max_rel(a, b, floor_frac=1e-3):
    """Pre-registered parity metric: max over cells and components of
    |a - b| / max(|b|, floor_frac * max|b_component|), per component column."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if a.ndim == 1:
        a, b = a[:, None], b[:, None]
    out = []
    for c in range(b.shape[1]):
        scale = np.abs(b[:, c]).max()
        fl = floor_frac * scale if scale > 0 else 1e-30
        out.append(float((np.abs(a[:, c] - b[:, c]) / np.maximum(np.abs(b[:, c]), fl)).max()))
    return np.array(out)


def bit_equal_frac(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    return float(np.mean(a.view(np.uint32) == b.view(np.uint32)))


def energy_sums(oparams, ev, pi, bulk, tau, dV):
    """(E, W) in float64 from AoS fields: E = sum tau (T^tautau + pi^tautau) dV and
    W = sum (tau^2 T^etaeta + pi~^etaeta) dV, with P from the oracle's FP32 table lookup
    (+ bulk Pi).  The balance dE/dtau = -W holds up to the boundary flux."""
    ev = np.asarray(ev, np.float32).reshape(-1, 4)
    P = O.eos_funcs(oparams, ev[:, 0])[:, 0].astype(np.float64)
    if bulk is not None:
        P = P + np.asarray(bulk, np.float32).reshape(-1).astype(np.float64)
    e = ev[:, 0].astype(np.float64)
    v = ev[:, 1:].astype(np.float64)
    g2 = 1.0 / (1.0 - (v * v).sum(1))
    w = (e + P) * g2
    p00 = np.zeros_like(e) if pi is None else np.asarray(pi, np.float32).reshape(-1, 10)[:, 0].astype(np.float64)
    p33 = np.zeros_like(e) if pi is None else np.asarray(pi, np.float32).reshape(-1, 10)[:, 9].astype(np.float64)
    E = (tau * (w - P + p00)).sum() * dV
    W = (w * v[:, 2] * v[:, 2] + P + p33).sum() * dV
    return float(E), float(W)
