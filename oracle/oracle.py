"""ctypes wrapper of the CPU oracle (oracle/clvisc_oracle.c).

TEST INFRASTRUCTURE ONLY: the oracle restates the reference kernels on the CPU
so that the HIP path can be checked against it.  Only the test suite, the smoke
check and bench.py's CPU-baseline leg import this module.

State arrays use the reference's AoS layouts: ev (N,4) = (e, vx, vy, veta),
pi (N,10) in the order 00,01,02,03,11,12,13,22,23,33, cell I = (i*NY+j)*NZ+k.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, 'lib')


class OrcParams(C.Structure):
    _fields_ = [
        ('nx', C.c_int32), ('ny', C.c_int32), ('nz', C.c_int32),
        ('riemann_test', C.c_int32), ('gubser_visc_test', C.c_int32),
        ('eos_kind', C.c_int32),
        ('eos_nrow', C.c_int32), ('eos_ncol', C.c_int32), ('eos_num_ed', C.c_int32),
        ('eos_kmin', C.c_int32), ('eos_kmax', C.c_int32), ('eos_log2j', C.c_int32),
        ('dt', C.c_float), ('dx', C.c_float), ('dy', C.c_float), ('dz', C.c_float),
        ('tau0', C.c_float),
        ('etaos_xmin', C.c_float), ('etaos_ymin', C.c_float),
        ('etaos_left_slop', C.c_float), ('etaos_right_slop', C.c_float), ('lam1', C.c_float),
        ('eos_ed_start', C.c_float), ('eos_ed_step', C.c_float),
        ('eos_table', C.c_void_p),
        ('bulk_on', C.c_int32),
        ('zetaos_max', C.c_float), ('zetaos_tpeak', C.c_float), ('zetaos_width', C.c_float),
        ('fp64_root', C.c_int32), ('fp64_root_vmin', C.c_float),
        ('nthreads', C.c_int32),
        ('tau0_cfg', C.c_double), ('dt_cfg', C.c_double),
        ('couplings', C.c_int32),
        ('jet_on', C.c_int32), ('jet_direction', C.c_float), ('jet_start', C.c_int64),
        ('plant', C.c_int32),
    ]


# mistakes the tests plant in the oracle's eta branches (clvisc_oracle.c ORC_PLANT_*)
PLANT_ETA_DZ = 1
PLANT_UDZ_NO_CHRISTOFFEL = 2


class OrcState(C.Structure):
    _fields_ = [
        ('ev', C.c_void_p * 3), ('pi', C.c_void_p * 3), ('bulk', C.c_void_p * 3),
        ('Src', C.c_void_p), ('IS_src', C.c_void_p), ('bulk_src', C.c_void_p),
        ('udx', C.c_void_p), ('udy', C.c_void_p), ('udz', C.c_void_p),
        ('udiff', C.c_void_p), ('goodcell', C.c_void_p), ('iters', C.c_void_p),
    ]


_LIBS = {}


def build():
    """Compile the oracle libraries (make in oracle/)."""
    subprocess.run(['make', '-s', '-C', HERE], check=True)


def lib(precision='f32'):
    if precision not in _LIBS:
        path = os.path.join(LIBDIR, 'liborc_%s.so' % precision)
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_max_ed.restype = C.c_float if precision == 'f32' else C.c_double
        _LIBS[precision] = L
    return _LIBS[precision]


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def make_params(nx, ny, nz, dt, dx, dy, dz, tau0, eos_table, etaos_xmin=0.18, etaos_ymin=0.08,
                etaos_left_slop=0.0, etaos_right_slop=0.0, lam1=-10.0, riemann_test=False,
                gubser_visc_test=False, bulk=None, fp64_root_vmin=None, nthreads=0, couplings=0,
                jet=None):
    """Build an OrcParams.  ``eos_table`` is a clvisc_amd.eos.EosTable; ``bulk`` is
    None or (zetaos_max, zetaos_tpeak, zetaos_width)."""
    p = OrcParams()
    p.nx, p.ny, p.nz = nx, ny, nz
    p.dt, p.dx, p.dy, p.dz, p.tau0 = dt, dx, dy, dz, tau0
    p.tau0_cfg, p.dt_cfg = float(tau0), float(dt)
    p.etaos_xmin, p.etaos_ymin = etaos_xmin, etaos_ymin
    p.etaos_left_slop, p.etaos_right_slop, p.lam1 = etaos_left_slop, etaos_right_slop, lam1
    p.riemann_test, p.gubser_visc_test = int(bool(riemann_test)), int(bool(gubser_visc_test))
    p.couplings = int(couplings)     # 1: PIMUNU_OMEGA_COUPLING, 2: OMEGA_OMEGA_COUPLING
    if jet is not None:              # (direction, start_pos_index[, stage mask, default 2]) of jet_eloss_src
        p.jet_on, p.jet_direction, p.jet_start = (int(jet[2]) if len(jet) > 2 else 2), float(jet[0]), int(jet[1])
    t = eos_table
    p.eos_kind = t.kind_id
    p.eos_nrow, p.eos_ncol, p.eos_num_ed = t.nrow, t.ncol, t.num_ed
    p.eos_kmin, p.eos_kmax, p.eos_log2j = t.kmin, t.kmax, t.log2j
    p.eos_ed_start, p.eos_ed_step = t.ed_start, t.ed_step
    p._table = np.ascontiguousarray(t.rgba, dtype=np.float32)   # keep alive
    p.eos_table = p._table.ctypes.data
    if bulk is not None:
        p.bulk_on = 1
        p.zetaos_max, p.zetaos_tpeak, p.zetaos_width = bulk
    if fp64_root_vmin is not None:
        p.fp64_root, p.fp64_root_vmin = 1, fp64_root_vmin
    p.nthreads = nthreads
    return p


class Oracle(object):
    """Reference-sequenced CPU hydro (buffers named after the reference drivers)."""

    def __init__(self, params, precision='f32', viscous=True, keep_iters=False):
        self.p = params
        self.precision = precision
        self.L = lib(precision)
        self.dtype = np.float32 if precision == 'f32' else np.float64
        self.real = self.dtype
        n = params.nx * params.ny * params.nz
        self.size = n
        z = lambda *s: np.zeros(s, self.dtype)
        self.ev = [z(n, 4) for _ in range(3)]
        self.pi = [z(n, 10) for _ in range(3)]
        self.bulk = [z(n) for _ in range(3)] if params.bulk_on else [None] * 3
        self.Src = z(n, 4)
        self.IS_src = z(n, 10)
        self.bulk_src = z(n) if params.bulk_on else None
        self.udx, self.udy, self.udz, self.udiff = z(n, 4), z(n, 4), z(n, 4), z(n, 4)
        self.goodcell = np.ones(n, self.dtype)
        self.iters = np.zeros(n, np.int32) if keep_iters else None
        self.tau = self.dtype(params.tau0_cfg)
        self.loop = 0
        s = OrcState()
        for k in range(3):
            s.ev[k] = self.ev[k].ctypes.data
            s.pi[k] = self.pi[k].ctypes.data
            s.bulk[k] = self.bulk[k].ctypes.data if self.bulk[k] is not None else None
        s.Src, s.IS_src = self.Src.ctypes.data, self.IS_src.ctypes.data
        s.bulk_src = self.bulk_src.ctypes.data if self.bulk_src is not None else None
        s.udx, s.udy, s.udz = self.udx.ctypes.data, self.udy.ctypes.data, self.udz.ctypes.data
        s.udiff, s.goodcell = self.udiff.ctypes.data, self.goodcell.ctypes.data
        s.iters = self.iters.ctypes.data if self.iters is not None else None
        self.state = s
        self._rt = C.c_float if precision == 'f32' else C.c_double
        if viscous:
            # CLVisc.__init__: IS_initialize on the zero field, then the ideal prediction
            # of the zero field into d_ev[2] (pyvisc/visc.py:101-102, 143-149, 468-470)
            self.L.orc_visc_initialize(C.byref(self.p), _ptr(self.pi[1]), _ptr(self.goodcell),
                                       _ptr(self.udiff), _ptr(self.ev[1]), self._rt(self.tau))
            self.ideal_stage(1)

    # -- stage-level (reference names) -------------------------------------------------
    def ideal_stage(self, step, tau=None):
        tau = self.tau if tau is None else tau
        self.L.orc_ideal_stage(C.byref(self.p), C.byref(self.state), C.c_int(step), self._rt(tau))

    def is_stage(self, step, tau=None):
        tau = self.tau if tau is None else tau
        self.L.orc_is_stage(C.byref(self.p), C.byref(self.state), C.c_int(step), self._rt(tau))

    def visc_stage(self, step, tau=None):
        tau = self.tau if tau is None else tau
        self.L.orc_visc_stage(C.byref(self.p), C.byref(self.state), C.c_int(step), self._rt(tau))

    def update_udiff(self, a, b):
        self.L.orc_get_udiff(C.byref(self.p), _ptr(self.udiff), _ptr(a), _ptr(b))

    def update_time(self, loop):
        self.tau = self.dtype(self.p.tau0_cfg + (loop + 1) * self.p.dt_cfg)

    # -- step-level ------------------------------------------------------------------
    def ideal_steps(self, nsteps):
        self.L.orc_ideal_steps(C.byref(self.p), C.byref(self.state), C.c_int64(self.loop),
                               C.c_int64(nsteps))
        self.loop += nsteps
        self.update_time(self.loop - 1)

    def visc_steps(self, nsteps):
        self.L.orc_visc_steps(C.byref(self.p), C.byref(self.state), C.c_int64(self.loop),
                              C.c_int64(nsteps))
        self.loop += nsteps
        self.update_time(self.loop - 1)

    def max_ed(self, which=1):
        return float(self.L.orc_max_ed(C.byref(self.p), _ptr(self.ev[which])))

    def pizz_o_ep(self, which=1):
        """Per-eta-slice mean of pi^{eta eta}/(e + P) over cells with e + P > 0.133
        (CLVisc.check_pizz, kernel_checkpi.cl:5-27)."""
        out = np.empty(self.p.nz, self.dtype)
        self.L.orc_pizz_o_ep(C.byref(self.p), _ptr(self.ev[which]), _ptr(self.pi[which]), _ptr(out))
        return out

    def load_ini(self, ev, pi=None, bulk=None):
        self.ev[1][...] = np.asarray(ev, self.dtype).reshape(self.size, 4)
        if pi is not None:
            self.pi[1][...] = np.asarray(pi, self.dtype).reshape(self.size, 10)
        if bulk is not None and self.bulk[1] is not None:
            self.bulk[1][...] = np.asarray(bulk, self.dtype).reshape(self.size)


# -- unit-level helpers ------------------------------------------------------------
def eos_lookup(params, eps, precision='f32'):
    L = lib(precision)
    dt = np.float32 if precision == 'f32' else np.float64
    eps = np.ascontiguousarray(eps, dt)
    out = np.zeros((eps.size, 4), dt)
    L.orc_eos(C.byref(params), _ptr(eps), C.c_int64(eps.size), _ptr(out))
    return out


def eos_funcs(params, eps, precision='f32'):
    """(P, S, T, CS2) with the above-table fallback."""
    L = lib(precision)
    dt = np.float32 if precision == 'f32' else np.float64
    eps = np.ascontiguousarray(eps, dt)
    out = np.zeros((eps.size, 4), dt)
    L.orc_eos_funcs(C.byref(params), _ptr(eps), C.c_int64(eps.size), _ptr(out))
    return out


def minmod4(a, b, precision='f32'):
    L = lib(precision)
    dt = np.float32 if precision == 'f32' else np.float64
    a = np.ascontiguousarray(a, dt)
    b = np.ascontiguousarray(b, dt)
    out = np.zeros(4, dt)
    L.orc_minmod4(_ptr(a), _ptr(b), _ptr(out))
    return out


def root_solve(params, T00, M, precision='f32'):
    L = lib(precision)
    dt = np.float32 if precision == 'f32' else np.float64
    T00 = np.ascontiguousarray(T00, dt)
    M = np.ascontiguousarray(M, dt)
    ed = np.zeros(T00.size, dt)
    it = np.zeros(T00.size, np.int32)
    L.orc_root(C.byref(params), _ptr(T00), _ptr(M), C.c_int64(T00.size), _ptr(ed), _ptr(it))
    return ed, it


def root_solve_from(params, T00, M, v0, precision='f32'):
    """The root solve started at v0 (clamped to [0, 0.9999]), as the seeded device solve."""
    L = lib(precision)
    dt = np.float32 if precision == 'f32' else np.float64
    T00 = np.ascontiguousarray(T00, dt)
    M = np.ascontiguousarray(M, dt)
    v0 = np.ascontiguousarray(v0, dt)
    ed = np.zeros(T00.size, dt)
    it = np.zeros(T00.size, np.int32)
    L.orc_root_from(C.byref(params), _ptr(T00), _ptr(M), _ptr(v0), C.c_int64(T00.size), _ptr(ed), _ptr(it))
    return ed, it


def visc_root_from(params, cons, bulk=None, v0=None, precision='f32'):
    """The end of the viscous update_ev (kernel_visc.cl:365-392) for conserved inputs cons
    (n x 4: T^tt, T^tx, T^ty, T^t eta before the 0.999 clamp) and Pi_new (bulk, or None):
    the split form of the stage.  v0 (None: the reference's start 0.5) seeds the root solve.
    Returns (ev n x 4, iterations)."""
    L = lib(precision)
    dt = np.float32 if precision == 'f32' else np.float64
    cons = np.ascontiguousarray(np.asarray(cons, dt).reshape(-1, 4))
    n = cons.shape[0]
    b = None if bulk is None else np.ascontiguousarray(bulk, dt).reshape(-1)
    s = None if v0 is None else np.ascontiguousarray(v0, dt).reshape(-1)
    ev = np.zeros((n, 4), dt)
    it = np.zeros(n, np.int32)
    L.orc_visc_root_from(C.byref(params), _ptr(cons), None if b is None else _ptr(b), None if s is None else _ptr(s),
                         C.c_int64(n), _ptr(ev), _ptr(it))
    return ev, it


def max_prop_speed(params, ev, vk, precision='f32'):
    L = lib(precision)
    dt = np.float32 if precision == 'f32' else np.float64
    ev = np.ascontiguousarray(ev, dt)
    vk = np.ascontiguousarray(vk, dt)
    out = np.zeros(vk.size, dt)
    L.orc_max_prop_speed(C.byref(params), _ptr(ev), _ptr(vk), C.c_int64(vk.size), _ptr(out))
    return out


def slope(a, b, c, precision='f32'):
    """The generalised-minmod slope of clvisc_oracle.c at b from (a, b, c)."""
    L = lib(precision)
    dt = np.float32 if precision == 'f32' else np.float64
    a, b, c = (np.ascontiguousarray(x, dt).reshape(-1) for x in (a, b, c))
    out = np.zeros(a.size, dt)
    L.orc_slope(_ptr(a), _ptr(b), _ptr(c), C.c_int64(a.size), _ptr(out))
    return out


# -- freeze-out hypersurface ----------------------------------------------------------
class OrcSfParams(C.Structure):
    _fields_ = [('nx', C.c_int32), ('ny', C.c_int32), ('nz', C.c_int32),
                ('nxskip', C.c_int32), ('nyskip', C.c_int32), ('nzskip', C.c_int32),
                ('dx', C.c_float), ('dy', C.c_float), ('dz', C.c_float), ('efrz', C.c_float)]


def test_cube(efrz, precision='f32'):
    """dSigma of the reference's test cube (old corners e = 3, new corners e = 2)."""
    L = lib(precision)
    dt = np.float32 if precision == 'f32' else np.float64
    out = np.zeros(4, dt)
    rt = C.c_float if precision == 'f32' else C.c_double
    L.orc_test_hypersf(rt(efrz), _ptr(out))
    return out


def hypersf(nx, ny, nz, dx, dy, dz, nxskip, nyskip, nzskip, efrz, ev_old, ev_new, tau_old, tau_new,
            pi_old=None, pi_new=None, bk_old=None, bk_new=None, cap=200000, precision='f32'):
    """Surface elements between two AoS ev fields (hypersf_oracle.c: orc_hypersf).

    Returns (count, sf (n,8), txyz (n,4), pi (n,10) or None, bulk (n,) or None) with
    n = min(count, cap); count is the number of elements found."""
    L = lib(precision)
    dt = np.float32 if precision == 'f32' else np.float64
    rt = C.c_float if precision == 'f32' else C.c_double
    p = OrcSfParams(nx, ny, nz, nxskip, nyskip, nzskip, dx, dy, dz, efrz)
    viscous = pi_old is not None
    with_bulk = viscous and bk_old is not None
    sf = np.zeros((cap, 8), dt)
    tx = np.zeros((cap, 4), dt)
    ps = np.zeros((cap, 10), dt) if viscous else None
    bs = np.zeros(cap, dt) if with_bulk else None
    cnt = C.c_int64(0)
    a = lambda x: None if x is None else np.ascontiguousarray(x, dt)
    evo, evn, pio, pin, bko, bkn = a(ev_old), a(ev_new), a(pi_old), a(pi_new), a(bk_old), a(bk_new)
    L.orc_hypersf.restype = C.c_int64
    L.orc_hypersf(C.byref(p), _ptr(evo), _ptr(evn), _ptr(pio), _ptr(pin), _ptr(bko), _ptr(bkn),
                  rt(tau_old), rt(tau_new), _ptr(sf), _ptr(tx), _ptr(ps), _ptr(bs), C.byref(cnt),
                  C.c_int64(cap), C.c_int(int(viscous)))
    n = min(cnt.value, cap)
    return (cnt.value, sf[:n], tx[:n], (ps[:n] if viscous else None), (bs[:n] if with_bulk else None))


# -- optical Glauber (oracle/glauber_oracle.c) ------------------------------------------
class OrcGlauber(C.Structure):
    _fields_ = [(k, C.c_float) for k in ('num_nucleons', 'sqrts', 'ro0', 'ra', 'eta', 'si0', 'b', 'edmax', 'hwn',
                                         'eta_flat', 'eta_gw')] + [('bjorken', C.c_int32)]


def glauber_params(cfg, bjorken=False):
    """OrcGlauber from the [glauber] keys of a configuration."""
    g = OrcGlauber()
    g.num_nucleons, g.sqrts, g.ro0 = cfg.NumOfNucleons, cfg.SQRTS, cfg.NucleonDensity
    g.ra, g.eta, g.si0, g.b = cfg.Ra, cfg.Eta, cfg.Si0, cfg.ImpactParameter
    g.edmax, g.hwn, g.eta_flat, g.eta_gw = cfg.Edmax, cfg.Hwn, cfg.Eta_flat, cfg.Eta_gw
    g.bjorken = int(bjorken)
    return g


def glauber_thickness(g, x, y, precision='f32'):
    L = lib(precision)
    rt = C.c_float if precision == 'f32' else C.c_double
    L.orc_glauber_thickness.restype = rt
    L.orc_glauber_thickness.argtypes = [C.c_void_p, rt, rt]
    return L.orc_glauber_thickness(C.byref(g), x, y)


def glauber_nw(g, ta, tb, precision='f32'):
    L = lib(precision)
    rt = C.c_float if precision == 'f32' else C.c_double
    L.orc_glauber_nw.restype = rt
    L.orc_glauber_nw.argtypes = [C.c_void_p, rt, rt]
    return L.orc_glauber_nw(C.byref(g), ta, tb)


def glauber_ini(g, nx, ny, nz, dx, dy, dz, precision='f32'):
    """(nx*ny*nz, 4) AoS field of glauber_ini."""
    L = lib(precision)
    dt = np.float32 if precision == 'f32' else np.float64
    rt = C.c_float if precision == 'f32' else C.c_double
    ev = np.zeros((nx * ny * nz, 4), dt)
    L.orc_glauber_ini.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, rt, rt, rt, C.c_void_p]
    L.orc_glauber_ini(C.byref(g), nx, ny, nz, dx, dy, dz, _ptr(ev))
    return ev


def glauber_nbinary(g, xlow, ylow, dx, dy, nx, ny, precision='f32'):
    L = lib(precision)
    dt = np.float32 if precision == 'f32' else np.float64
    rt = C.c_float if precision == 'f32' else C.c_double
    out = np.zeros(nx * ny, dt)
    L.orc_glauber_nbinary.argtypes = [C.c_void_p, rt, rt, rt, rt, C.c_int, C.c_int, C.c_void_p]
    L.orc_glauber_nbinary(C.byref(g), xlow, ylow, dx, dy, nx, ny, _ptr(out))
    return out.reshape(nx, ny)


def smear_constants(sigr, sigz, tau0):
    """(1/(2 SIGR^2), w1 w2) as the device rounds them: computed in double with the float pi
    of M_PI_F (kernel_gaussian_smearing_new.cl:3-7), then rounded to float."""
    pif = float(np.float32(np.pi))
    c = 1.0 / (2.0 * float(np.float32(sigr)) ** 2)
    cz = 1.0 / (2.0 * float(np.float32(sigz)) ** 2)
    w = (c / pif) * (np.sqrt(cz / pif) / float(np.float32(tau0)))
    return np.float32(c), np.float32(w)


def smear(params, p4x4, sigr, sigz, kfactor, precision='f32'):
    """(n, 4) AoS field of the smearing kernel for a parton list (rows E px py pz t x y z)."""
    L = lib(precision)
    dt = np.float32 if precision == 'f32' else np.float64
    rt = C.c_float if precision == 'f32' else C.c_double
    a = np.ascontiguousarray(np.asarray(p4x4, dt).reshape(-1, 8))
    c, w = smear_constants(sigr, sigz, params.tau0_cfg)
    if precision != 'f32':
        cd = 1.0 / (2.0 * sigr * sigr)
        c, w = cd, (cd / np.pi) * (np.sqrt(1.0 / (2.0 * sigz * sigz) / np.pi) / params.tau0_cfg)
    n = params.nx * params.ny * params.nz
    ev = np.zeros((n, 4), dt)
    L.orc_smear.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, rt, rt, rt, C.c_void_p]
    L.orc_smear(C.byref(params), _ptr(a) if a.size else None, a.shape[0], c, w, kfactor, _ptr(ev))
    return ev


def vorticity(params, ev_old, ev_new, tau, precision='f32'):
    """(n, 32) vorticity tensors of kernel_vorticity.cl (orc_vorticity) from AoS fields."""
    L = lib(precision)
    dt = np.float32 if precision == 'f32' else np.float64
    rt = C.c_float if precision == 'f32' else C.c_double
    a, b = (np.ascontiguousarray(np.asarray(x, dt).reshape(-1, 4)) for x in (ev_old, ev_new))
    out = np.zeros((a.shape[0], 32), dt)
    L.orc_vorticity.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, rt, C.c_void_p]
    L.orc_vorticity(C.byref(params), _ptr(a), _ptr(b), tau, _ptr(out))
    return out


def jet_src(params, Src, ev, tau, precision='f32'):
    """Src (n, 4) plus the jet_eloss_src deposit at tau (orc_jet_src, kernel_jet_eloss.cl:8-69)."""
    L = lib(precision)
    dt = np.float32 if precision == 'f32' else np.float64
    rt = C.c_float if precision == 'f32' else C.c_double
    s = np.ascontiguousarray(np.array(Src, dt).reshape(-1, 4))
    e = np.ascontiguousarray(np.asarray(ev, dt).reshape(-1, 4))
    L.orc_jet_src.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, rt]
    L.orc_jet_src(C.byref(params), _ptr(s), _ptr(e), tau)
    return s


def couplings(pim, udt, ux, uy, uz, u, precision='f32'):
    """The IS vorticity couplings of one state (orc_couplings, kernel_IS.cl:354-470):
    (out[0:10] = sum_lam g_lam,lam pi_omega(lam, mu, nu), out[10:20] the same of omega_omega,
    out[20:26] = Omega^{01, 02, 03, 12, 13, 23}); pim in the 10-component upper-triangle
    order, the derivative rows and u as 4-vectors."""
    L = lib(precision)
    dt = np.float32 if precision == 'f32' else np.float64
    a = [np.ascontiguousarray(np.asarray(x, dt)) for x in (pim, udt, ux, uy, uz, u)]
    out = np.zeros(26, dt)
    L.orc_couplings.argtypes = [C.c_void_p] * 7
    L.orc_couplings(*[_ptr(x) for x in a], _ptr(out))
    return out


def vorticity_mu(params, ev, om32, efrz, precision='f32'):
    L = lib(precision)
    dt = np.float32 if precision == 'f32' else np.float64
    rt = C.c_float if precision == 'f32' else C.c_double
    e = np.ascontiguousarray(np.asarray(ev, dt).reshape(-1, 4))
    w = np.ascontiguousarray(np.asarray(om32, dt).reshape(-1, 32))
    out = np.zeros((e.shape[0], 4), dt)
    L.orc_vorticity_mu.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, rt, C.c_void_p]
    L.orc_vorticity_mu(C.byref(params), _ptr(e), _ptr(w), efrz, _ptr(out))
    return out


def polarization(kind, sf, om1, om2, mom, tfrz, mass, precision='f32'):
    """(pol_lab (m,4), density (m,), pol_lrf (m,4)) of orc_polarization (kernel_polarization.cl)."""
    L = lib(precision)
    dt = np.float32 if precision == 'f32' else np.float64
    rt = C.c_float if precision == 'f32' else C.c_double
    f = lambda a: None if a is None else np.ascontiguousarray(a, np.float32)
    sf, om1, om2, mom = f(sf), f(om1), f(om2), f(mom).reshape(-1, 4)
    m = mom.shape[0]
    pl, de, lr = np.zeros((m, 4), dt), np.zeros(m, dt), np.zeros((m, 4), dt)
    L.orc_polarization.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64,
                                   rt, rt, C.c_void_p, C.c_void_p, C.c_void_p]
    L.orc_polarization(kind, _ptr(sf), _ptr(om1), _ptr(om2), sf.shape[0], _ptr(mom), m, tfrz, mass, _ptr(pl),
                       _ptr(de), _ptr(lr))
    return pl, de, lr
