"""ctypes binding of libclvisc_amd.so (declarations in include/clvisc_amd.h).

The library is the only compute path: if it is missing or fails to load the
import raises, and nothing falls back to a CPU implementation.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'lib', 'libclvisc_amd.so')

CV_FLAG_RIEMANN_TEST = 0x1
CV_FLAG_GUBSER_VISC_TEST = 0x2
CV_FLAG_BULK = 0x4
CV_FLAG_FP64_ROOT = 0x8
CV_FLAG_SF_FP64 = 0x10
CV_FLAG_PIMN_OMEGA = 0x20
CV_FLAG_OMEGA_OMEGA = 0x40

FIELDS = {'ev0': 0, 'ev1': 1, 'ev2': 2, 'pi0': 3, 'pi1': 4, 'pi2': 5,
          'bulk0': 6, 'bulk1': 7, 'bulk2': 8, 'src': 9, 'is_src': 10,
          'udiff': 11, 'goodcell': 12, 'root_cons': 13, 'root_seed': 14, 'omega0': 15, 'omega1': 16}
NCOMP = {0: 4, 1: 4, 2: 4, 3: 10, 4: 10, 5: 10, 6: 1, 7: 1, 8: 1, 9: 4, 10: 10, 11: 4, 12: 1, 13: 4, 14: 1,
         15: 32, 16: 32}


class CvGlauber(C.Structure):
    """cv_glauber: the [glauber] keys of hydro.info (include/clvisc_amd.h)."""
    _fields_ = [(k, C.c_float) for k in ('num_nucleons', 'sqrts', 'ro0', 'ra', 'eta', 'si0', 'b', 'edmax', 'hwn',
                                         'eta_flat', 'eta_gw')] + [('bjorken', C.c_int32)]

    @classmethod
    def from_cfg(cls, cfg, bjorken=False):
        g = cls()
        g.num_nucleons, g.sqrts, g.ro0 = cfg.NumOfNucleons, cfg.SQRTS, cfg.NucleonDensity
        g.ra, g.eta, g.si0, g.b = cfg.Ra, cfg.Eta, cfg.Si0, cfg.ImpactParameter
        g.edmax, g.hwn, g.eta_flat, g.eta_gw = cfg.Edmax, cfg.Hwn, cfg.Eta_flat, cfg.Eta_gw
        g.bjorken = int(bjorken)
        return g


class CvParams(C.Structure):
    _fields_ = [
        ('nx', C.c_int32), ('ny', C.c_int32), ('nz', C.c_int32), ('flags', C.c_int32),
        ('tau0', C.c_double), ('dt', C.c_double),
        ('dx', C.c_float), ('dy', C.c_float), ('dz', C.c_float),
        ('etaos_xmin', C.c_float), ('etaos_ymin', C.c_float),
        ('etaos_left_slop', C.c_float), ('etaos_right_slop', C.c_float),
        ('lam1', C.c_float),
        ('zetaos_max', C.c_float), ('zetaos_tpeak', C.c_float), ('zetaos_width', C.c_float),
        ('fp64_root_vmin', C.c_float),
        ('efrz', C.c_float),
        ('nxskip', C.c_int32), ('nyskip', C.c_int32), ('nzskip', C.c_int32),
        ('reserved', C.c_int32),
    ]


class CvEos(C.Structure):
    _fields_ = [
        ('kind', C.c_int32),
        ('nrow', C.c_int32), ('ncol', C.c_int32), ('num_ed', C.c_int32),
        ('ed_start', C.c_float), ('ed_step', C.c_float),
        ('kmin', C.c_int32), ('kmax', C.c_int32), ('log2j', C.c_int32),
        ('n_entries', C.c_int32),
        ('rgba', C.c_void_p),
    ]


# status codes of the C-ABI (include/clvisc_amd.h)
CV_OK = 0
CV_ERR_ARG = -1
CV_ERR_HIP = -2
CV_ERR_STATE = -3
CV_ERR_CAPACITY = -4
CV_ERR_UNSUPPORTED = -5


class CvError(RuntimeError):
    """A failed C-ABI call (or a missing library); ``code`` is the CV_ERR_* status, or None."""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


_lib = None


def lib():
    """Load libclvisc_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CvError('%s is missing: build it with `python -c "import __graft_entry__ as g; '
                          'g.build()"` or `make -C clvisc_amd/csrc`' % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        L.cv_last_error.restype = C.c_char_p
        vp, i32, i64, f32 = C.c_void_p, C.c_int32, C.c_int64, C.c_float
        sig = {
            'cv_device_info': [C.c_int, C.c_char_p, C.c_int, C.POINTER(C.c_int32)],
            'cv_selftest_fastmath': [C.c_int, C.POINTER(C.c_uint64)],
            'cv_create': [C.POINTER(CvParams), C.POINTER(CvEos), C.c_int, C.c_int, C.POINTER(vp)],
            'cv_destroy': [vp],
            'cv_set_field': [vp, C.c_int, vp, i64],
            'cv_get_field': [vp, C.c_int, vp, i64],
            'cv_copy_field': [vp, C.c_int, C.c_int],
            'cv_gather': [vp, C.c_int, vp, i64, vp],
            'cv_visc_initialize': [vp, f32],
            'cv_ideal_stage': [vp, C.c_int, f32],
            'cv_is_stage': [vp, C.c_int, f32],
            'cv_visc_stage': [vp, C.c_int, f32],
            'cv_update_udiff': [vp, C.c_int, C.c_int],
            'cv_ideal_steps_async': [vp, i64, i64],
            'cv_visc_steps_async': [vp, i64, i64],
            'cv_max_ed': [vp, C.c_int, C.POINTER(f32)],
            'cv_pizz_o_ep': [vp, C.c_int, vp],
            'cv_energy_sum': [vp, C.c_int, C.c_double, vp],
            'cv_slice_sums': [vp, C.c_int, vp],
            'cv_slice_sums_noalloc': [vp, C.c_int, vp],
            'cv_glauber_ini': [vp, vp, C.c_int],
            'cv_glauber_nbinary': [vp, vp, f32, f32, f32, f32, C.c_int, C.c_int, vp],
            'cv_smear_partons': [vp, vp, i64, f32, f32, f32, C.c_int],
            'cv_longitudinal_profile': [vp, C.c_int, vp],
            'cv_sd_to_ed': [vp, C.c_int, vp, vp, i64],
            'cv_vorticity_enable': [vp],
            'cv_vorticity': [vp, f32],
            'cv_vorticity_mu': [vp, f32, vp, i64],
            'cv_hypersf_get_vorticity': [vp, vp, i64],
            'cv_polarization': [C.c_int, C.c_int, vp, vp, vp, i64, vp, i64, f32, f32, vp, vp, vp],
            'cv_set_xcd_map': [vp, C.c_int],
            'cv_sync': [vp],
            'cv_last_elapsed_ms': [vp, C.POINTER(f32)],
            'cv_profile_visc_step': [vp, i64, C.POINTER(f32), C.c_int, C.POINTER(C.c_int32)],
            'cv_set_stage_kernel': [vp, C.c_int, C.c_int, C.c_int],
            'cv_set_jet_source': [vp, C.c_int, C.c_float, C.c_int64],
            'cv_get_stage_kernel': [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32)],
            'cv_set_march_chunk': [vp, C.c_int],
            'cv_get_march_chunk': [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32)],
            'cv_hypersf_reserve': [vp, i64],
            'cv_hypersf_snapshot': [vp, f32],
            'cv_hypersf_find': [vp, f32],
            'cv_hypersf_count': [vp, C.POINTER(i64), C.POINTER(i64)],
            'cv_hypersf_get': [vp, vp, vp, vp, vp, i64],
            'cv_hypersf_reset': [vp],
            'cv_batch_create': [C.POINTER(CvParams), C.POINTER(CvEos), C.c_int, vp, C.c_int, C.POINTER(vp)],
            'cv_batch_destroy': [vp],
            'cv_batch_handle': [vp, C.c_int, C.POINTER(vp)],
            'cv_batch_set_active': [vp, vp],
            'cv_batch_steps_async': [vp, i64],
            'cv_batch_set_loop': [vp, i64],
            'cv_batch_set_xcd_map': [vp, C.c_int],
            'cv_batch_sync': [vp],
            'cv_batch_last_elapsed_ms': [vp, C.POINTER(f32)],
            'cv_batch_info': [vp, C.POINTER(i32), C.POINTER(i64), C.POINTER(i64)],
            'cv_mlp_create': [C.c_int, C.c_uint64, C.POINTER(vp)],
            'cv_mlp_destroy': [vp],
            'cv_mlp_get_params': [vp, vp, C.c_int],
            'cv_mlp_set_params': [vp, vp, C.c_int],
            'cv_mlp_grad': [vp, vp, vp, i64, vp, C.POINTER(f32)],
            'cv_mlp_set_data': [vp, vp, vp, i64],
            'cv_mlp_train': [vp, C.c_int, C.c_int, f32, vp],
            'cv_mlp_forward': [vp, vp, i64, vp],
            'cv_mlp_features': [vp, vp, vp, vp, i64, vp],
            'cv_batch_set_root_seed': [vp, vp],
            'cv_root_eval': [vp, vp, vp, vp, vp, i64, vp, vp, vp],
            'cv_set_root_seed': [vp, vp, C.c_int],
            'cv_set_root_split': [vp, C.c_int],
            'cv_get_root_iters': [vp, vp, i64],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        L.cv_mlp_nparam.argtypes = []
        L.cv_mlp_nparam.restype = C.c_int
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise CvError('clvisc_amd error %d: %s' % (rc, lib().cv_last_error().decode()), rc)


def exported_symbols():
    """Names of the C-ABI entry points this binding expects."""
    return ['cv_last_error', 'cv_device_info', 'cv_selftest_fastmath', 'cv_create', 'cv_destroy', 'cv_set_field',
            'cv_get_field', 'cv_copy_field', 'cv_gather', 'cv_visc_initialize', 'cv_ideal_stage', 'cv_is_stage',
            'cv_visc_stage', 'cv_update_udiff', 'cv_ideal_steps_async', 'cv_visc_steps_async',
            'cv_max_ed', 'cv_pizz_o_ep', 'cv_sync', 'cv_last_elapsed_ms', 'cv_profile_visc_step',
            'cv_set_stage_kernel', 'cv_get_stage_kernel', 'cv_set_jet_source', 'cv_set_march_chunk', 'cv_get_march_chunk',
            'cv_hypersf_reserve', 'cv_hypersf_snapshot',
            'cv_hypersf_find', 'cv_hypersf_count', 'cv_hypersf_get', 'cv_hypersf_reset',
            'cv_batch_create', 'cv_batch_destroy', 'cv_batch_handle', 'cv_batch_set_active',
            'cv_batch_steps_async', 'cv_batch_set_loop', 'cv_batch_set_xcd_map', 'cv_batch_sync', 'cv_batch_last_elapsed_ms',
            'cv_batch_info', 'cv_mlp_nparam', 'cv_mlp_create', 'cv_mlp_destroy', 'cv_mlp_get_params',
            'cv_mlp_set_params', 'cv_mlp_grad', 'cv_mlp_set_data', 'cv_mlp_train', 'cv_mlp_forward', 'cv_mlp_features', 'cv_batch_set_root_seed',
            'cv_root_eval', 'cv_set_root_seed', 'cv_get_root_iters', 'cv_set_root_split', 'cv_energy_sum', 'cv_slice_sums', 'cv_slice_sums_noalloc', 'cv_glauber_ini',
            'cv_glauber_nbinary', 'cv_smear_partons', 'cv_longitudinal_profile',
            'cv_sd_to_ed', 'cv_vorticity_enable', 'cv_vorticity', 'cv_vorticity_mu', 'cv_hypersf_get_vorticity',
            'cv_polarization', 'cv_set_xcd_map']


def make_eos(table):
    """CvEos from a clvisc_amd.eos.EosTable (keeps the host array alive on the struct)."""
    e = CvEos()
    e.kind = table.kind_id
    e.nrow, e.ncol, e.num_ed = table.nrow, table.ncol, table.num_ed
    e.ed_start, e.ed_step = table.ed_start, table.ed_step
    e.kmin, e.kmax, e.log2j = table.kmin, table.kmax, table.log2j
    rgba = np.ascontiguousarray(table.rgba, dtype=np.float32).reshape(-1, 4)
    e.n_entries = rgba.shape[0]
    e._keep = rgba
    e.rgba = rgba.ctypes.data
    return e


class Handle(object):
    """Owner of one device-side hydro state (cv_handle)."""

    def __init__(self, params, eos_table, device=0, viscous=True, _borrow=None):
        self.L = lib()
        self.params = params
        self.n = params.nx * params.ny * params.nz
        self.viscous = viscous
        if _borrow is not None:          # a handle owned by a Batch
            self.h, self._owner = _borrow
            return
        self._owner = None
        self.eos = make_eos(eos_table)
        h = C.c_void_p()
        check(self.L.cv_create(C.byref(params), C.byref(self.eos), device, int(viscous), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, 'h', None) and getattr(self, '_owner', None) is None:
            self.L.cv_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get(self, field):
        f = FIELDS[field] if isinstance(field, str) else field
        nc = NCOMP[f]
        out = np.empty((self.n, nc) if nc > 1 else (self.n,), np.float32)
        check(self.L.cv_get_field(self.h, f, out.ctypes.data, out.size))
        return out

    def set(self, field, arr):
        f = FIELDS[field] if isinstance(field, str) else field
        a = np.ascontiguousarray(arr, dtype=np.float32)
        check(self.L.cv_set_field(self.h, f, a.ctypes.data, a.size))

    def gather(self, field, cells):
        """Rows of a field at the given flat cell indices (AoS, float32)."""
        f = FIELDS[field] if isinstance(field, str) else field
        nc = NCOMP[f]
        idx = np.ascontiguousarray(cells, dtype=np.int64).reshape(-1)
        out = np.empty((idx.size, nc) if nc > 1 else (idx.size,), np.float32)
        check(self.L.cv_gather(self.h, f, idx.ctypes.data, idx.size, out.ctypes.data))
        return out

    def copy(self, dst, src):
        d = FIELDS[dst] if isinstance(dst, str) else dst
        s = FIELDS[src] if isinstance(src, str) else src
        check(self.L.cv_copy_field(self.h, d, s))

    def visc_initialize(self, tau):
        check(self.L.cv_visc_initialize(self.h, float(tau)))

    def ideal_stage(self, step, tau):
        check(self.L.cv_ideal_stage(self.h, step, float(tau)))

    def is_stage(self, step, tau):
        check(self.L.cv_is_stage(self.h, step, float(tau)))

    def visc_stage(self, step, tau):
        check(self.L.cv_visc_stage(self.h, step, float(tau)))

    def update_udiff(self, a, b):
        check(self.L.cv_update_udiff(self.h, FIELDS.get(a, a), FIELDS.get(b, b)))

    def ideal_steps(self, loop0, nsteps, sync=True):
        check(self.L.cv_ideal_steps_async(self.h, loop0, nsteps))
        if sync:
            self.sync()

    def visc_steps(self, loop0, nsteps, sync=True):
        check(self.L.cv_visc_steps_async(self.h, loop0, nsteps))
        if sync:
            self.sync()

    def energy_sum(self, tau, which=1):
        """(E, W) in double: E = sum tau (T^tautau + pi^tautau) dV, W = sum tau^2 T^etaeta dV
        (with pi); dE/dtau = -W plus the boundary flux."""
        out = np.zeros(2, np.float64)
        check(self.L.cv_energy_sum(self.h, which, float(tau), out.ctypes.data))
        return float(out[0]), float(out[1])

    def slice_sums(self, which=1):
        """(NZ, 10) float64 per-eta-slice sums over the (x, y) plane (cv_slice_sums): Txx - Tyy,
        Txx + Tyy, T^tx, u0 s, u0 e |v_T|, u0 e (bulkinfo.py's floors), then (e + P) u0^2,
        (e + P) u0^2 v^eta, P, u0 s (kernel_bulkinfo.cl's unfloored e and gamma)."""
        out = np.zeros((self.params.nz, 10), np.float64)
        check(self.L.cv_slice_sums(self.h, which, out.ctypes.data))
        return out

    def slice_sums_noalloc(self, d_out, which=1):
        """cv_slice_sums into caller-owned device memory (d_out: device address of NZ * 10
        doubles) with the library's static scratch array: allocates nothing, returns without
        waiting (sync() before reading d_out)."""
        check(self.L.cv_slice_sums_noalloc(self.h, which, C.c_void_p(int(d_out))))

    def glauber_ini(self, glauber, which=1):
        """Optical Glauber initial condition into slot `which` (cv_glauber_ini)."""
        check(self.L.cv_glauber_ini(self.h, C.byref(glauber), which))

    def glauber_nbinary(self, glauber, xlow, ylow, dx, dy, nx, ny):
        out = np.empty(nx * ny, np.float32)
        check(self.L.cv_glauber_nbinary(self.h, C.byref(glauber), xlow, ylow, dx, dy, nx, ny, out.ctypes.data))
        return out.reshape(nx, ny)

    def smear_partons(self, p4x4, sigr=0.6, sigz=0.6, kfactor=1.0, which=1):
        """Gaussian-smeared parton list (rows E px py pz t x y z) into slot `which`."""
        a = np.ascontiguousarray(np.asarray(p4x4, np.float32).reshape(-1, 8))
        check(self.L.cv_smear_partons(self.h, a.ctypes.data if a.size else None, a.shape[0], sigr, sigz,
                                      kfactor, which))

    def longitudinal_profile(self, profile, which=1):
        a = np.ascontiguousarray(profile, np.float32).reshape(-1)
        if a.size != self.params.nz:
            raise ValueError('profile needs NZ = %d values' % self.params.nz)
        check(self.L.cv_longitudinal_profile(self.h, which, a.ctypes.data))

    def sd_to_ed(self, s_nodes, e_nodes, which=1):
        """Component e of slot `which` from entropy to energy density (cv_sd_to_ed)."""
        s = np.ascontiguousarray(s_nodes, np.float64)
        e = np.ascontiguousarray(e_nodes, np.float64)
        check(self.L.cv_sd_to_ed(self.h, which, s.ctypes.data, e.ctypes.data, s.size))

    def vorticity_enable(self):
        check(self.L.cv_vorticity_enable(self.h))

    def vorticity(self, tau):
        """d_omega[1] from d_ev[0] and d_ev[1] at tau (cv_vorticity)."""
        check(self.L.cv_vorticity(self.h, float(tau)))

    def vorticity_mu(self, efrz):
        out = np.empty((self.n, 4), np.float32)
        check(self.L.cv_vorticity_mu(self.h, float(efrz), out.ctypes.data, out.size))
        return out

    def hypersf_get_vorticity(self):
        """(n, 32) vorticity tensors of the stored surface elements."""
        n, _ = self.hypersf_count()
        out = np.empty((n, 32), np.float32)
        check(self.L.cv_hypersf_get_vorticity(self.h, out.ctypes.data if n else out.ctypes.data, n))
        return out

    def set_xcd_map(self, nxcd):
        check(self.L.cv_set_xcd_map(self.h, int(nxcd)))

    def max_ed(self, which=1):
        out = C.c_float()
        check(self.L.cv_max_ed(self.h, which, C.byref(out)))
        return out.value

    def pizz_o_ep(self, which=1):
        """Per-eta-slice mean of pi^{eta eta}/(e + P) over the cells with e + P > 0.133
        (CLVisc.check_pizz)."""
        out = np.empty(int(self.params.nz), np.float32)
        check(self.L.cv_pizz_o_ep(self.h, which, out.ctypes.data))
        return out

    def sync(self):
        check(self.L.cv_sync(self.h))

    def profile_visc_step(self, loop0):
        ms = (C.c_float * 16)()
        nl = C.c_int32()
        check(self.L.cv_profile_visc_step(self.h, loop0, ms, 16, C.byref(nl)))
        return [ms[i] for i in range(nl.value)]

    def set_jet_source(self, jet_eloss_src=None, stages=2):
        """jet_eloss_src = {'switch_on', 'start_pos_index', 'direction'} (pyvisc/ideal.py:211) or
        None: the deposit of kernel_jet_eloss.cl added to the later RK stages in the mask
        ``stages`` (1: stage 1, 2: stage 2; evolve passes it to stage 2 only) (cv_set_jet_source)."""
        if not jet_eloss_src or not jet_eloss_src.get('switch_on', False):
            check(self.L.cv_set_jet_source(self.h, 0, 0.0, 0))
        else:
            check(self.L.cv_set_jet_source(self.h, int(stages), float(np.float32(jet_eloss_src.get('direction', 0.0))),
                                           int(jet_eloss_src.get('start_pos_index', 0))))

    def set_stage_kernel(self, kind, ty=0, tz=0):
        check(self.L.cv_set_stage_kernel(self.h, kind, ty, tz))

    def stage_kernel(self):
        k, ty, tz = C.c_int32(), C.c_int32(), C.c_int32()
        check(self.L.cv_get_stage_kernel(self.h, C.byref(k), C.byref(ty), C.byref(tz)))
        return k.value, ty.value, tz.value

    def set_march_chunk(self, xc):
        check(self.L.cv_set_march_chunk(self.h, int(xc)))

    def march_chunk(self):
        """(x planes per chunk, chunks) of the marching kernel."""
        xc, nch = C.c_int32(), C.c_int32()
        check(self.L.cv_get_march_chunk(self.h, C.byref(xc), C.byref(nch)))
        return xc.value, nch.value

    def last_elapsed_ms(self):
        out = C.c_float()
        check(self.L.cv_last_elapsed_ms(self.h, C.byref(out)))
        return out.value

    # -- learned root-solve seed
    def set_root_seed(self, mlp, track_iters=True):
        check(self.L.cv_set_root_seed(self.h, mlp.m if mlp is not None else None, int(track_iters)))

    def set_root_split(self, on=True):
        """Viscous stages in split form (stage, then seed and root solve in their own
        kernels); the fields root_cons / root_seed hold the last stage's solve input.
        on=False with a seed network set: the network runs inside the stage kernel (the same
        fields bit for bit; root_seed and the iteration counts are kept when iterations are
        tracked)."""
        check(self.L.cv_set_root_split(self.h, int(on)))

    def root_iters(self):
        out = np.empty(self.n, np.int32)
        check(self.L.cv_get_root_iters(self.h, out.ctypes.data, self.n))
        return out

    def root_eval(self, T00, M, T03, mlp=None):
        """(e, |v|, iterations) of the root solve of the given states, plain or seeded."""
        a = [np.ascontiguousarray(x, np.float32).reshape(-1) for x in (T00, M, T03)]
        n = a[0].size
        e, v, it = np.empty(n, np.float32), np.empty(n, np.float32), np.empty(n, np.int32)
        check(self.L.cv_root_eval(self.h, mlp.m if mlp is not None else None, a[0].ctypes.data, a[1].ctypes.data,
                                  a[2].ctypes.data, n, e.ctypes.data, v.ctypes.data, it.ctypes.data))
        return e, v, it

    # -- freeze-out hypersurface
    def hypersf_reserve(self, capacity):
        check(self.L.cv_hypersf_reserve(self.h, int(capacity)))

    def hypersf_snapshot(self, tau):
        check(self.L.cv_hypersf_snapshot(self.h, float(tau)))

    def hypersf_find(self, tau_new):
        check(self.L.cv_hypersf_find(self.h, float(tau_new)))

    def hypersf_count(self):
        """(elements found, capacity)."""
        c, cap = C.c_int64(), C.c_int64()
        check(self.L.cv_hypersf_count(self.h, C.byref(c), C.byref(cap)))
        return c.value, cap.value

    def hypersf_get(self):
        """(sf (n,8), txyz (n,4), pi (n,10) or None, bulk (n,) or None) of the stored elements."""
        n, _ = self.hypersf_count()
        sf = np.empty((n, 8), np.float32)
        tx = np.empty((n, 4), np.float32)
        ps = np.empty((n, 10), np.float32) if self.viscous else None
        bk = np.empty(n, np.float32) if (self.viscous and self.params.flags & CV_FLAG_BULK) else None
        ptr = lambda a: None if a is None else a.ctypes.data
        check(self.L.cv_hypersf_get(self.h, ptr(sf), ptr(tx), ptr(ps), ptr(bk), n))
        return sf, tx, ps, bk

    def hypersf_reset(self):
        check(self.L.cv_hypersf_reset(self.h))


class Batch(object):
    """Owner of an event batch (cv_batch): events with their own grids stepped together."""

    def __init__(self, params, eos_table, dims, device=0):
        self.L = lib()
        self.params = params
        self.eos = make_eos(eos_table)
        self.dims = [tuple(int(x) for x in d) for d in dims]
        d = np.ascontiguousarray(np.array(self.dims, dtype=np.int32).reshape(-1))
        b = C.c_void_p()
        check(self.L.cv_batch_create(C.byref(params), C.byref(self.eos), len(self.dims), d.ctypes.data, device,
                                     C.byref(b)))
        self.b = b
        self._handles = {}

    def handle(self, e):
        """Handle of event e (borrowed: valid while the batch lives)."""
        h = C.c_void_p()
        check(self.L.cv_batch_handle(self.b, int(e), C.byref(h)))
        if e not in self._handles:
            p = CvParams.from_buffer_copy(self.params)
            p.nx, p.ny, p.nz = self.dims[e]
            self._handles[e] = Handle(p, None, viscous=True, _borrow=(h, self))
        return self._handles[e]

    def set_active(self, mask):
        m = np.ascontiguousarray(np.asarray(mask, dtype=np.int32).reshape(-1))
        if m.size != len(self.dims):
            raise ValueError('mask needs %d entries' % len(self.dims))
        check(self.L.cv_batch_set_active(self.b, m.ctypes.data))

    def steps(self, nsteps, sync=True):
        check(self.L.cv_batch_steps_async(self.b, int(nsteps)))
        if sync:
            self.sync()

    def set_loop(self, loop):
        check(self.L.cv_batch_set_loop(self.b, int(loop)))

    def set_root_seed(self, mlp):
        """Seeded batch stages (cv_batch_set_root_seed): the split form with the network's
        seeds (FP64 root for fast cells under CV_FLAG_FP64_ROOT); None: the fused stages."""
        check(self.L.cv_batch_set_root_seed(self.b, mlp.m if mlp is not None else None))
        self._seed = mlp            # keep the network alive while the batch uses it

    def set_xcd_map(self, nxcd):
        check(self.L.cv_batch_set_xcd_map(self.b, int(nxcd)))

    def sync(self):
        check(self.L.cv_batch_sync(self.b))

    def last_elapsed_ms(self):
        out = C.c_float()
        check(self.L.cv_batch_last_elapsed_ms(self.b, C.byref(out)))
        return out.value

    def info(self):
        """(events, loop index of the next step, work items per stage)."""
        n, loop, w = C.c_int32(), C.c_int64(), C.c_int64()
        check(self.L.cv_batch_info(self.b, C.byref(n), C.byref(loop), C.byref(w)))
        return n.value, loop.value, w.value

    def close(self):
        if getattr(self, 'b', None):
            for h in self._handles.values():
                h.h = None
            self.L.cv_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Mlp(object):
    """Owner of the learned root-solve seed network (cv_mlp)."""

    def __init__(self, device=0, seed=1234):
        self.L = lib()
        self.nparam = self.L.cv_mlp_nparam()
        m = C.c_void_p()
        check(self.L.cv_mlp_create(device, C.c_uint64(seed), C.byref(m)))
        self.m = m

    def params(self):
        out = np.empty(self.nparam, np.float32)
        check(self.L.cv_mlp_get_params(self.m, out.ctypes.data, self.nparam))
        return out

    def set_params(self, p):
        a = np.ascontiguousarray(p, np.float32).reshape(-1)
        check(self.L.cv_mlp_set_params(self.m, a.ctypes.data, a.size))

    def grad(self, X, T):
        X = np.ascontiguousarray(X, np.float32).reshape(-1, 4)
        T = np.ascontiguousarray(T, np.float32).reshape(-1, 2)
        g = np.empty(self.nparam, np.float32)
        loss = C.c_float()
        check(self.L.cv_mlp_grad(self.m, X.ctypes.data, T.ctypes.data, X.shape[0], g.ctypes.data, C.byref(loss)))
        return g, loss.value

    def set_data(self, X, T):
        X = np.ascontiguousarray(X, np.float32).reshape(-1, 4)
        T = np.ascontiguousarray(T, np.float32).reshape(-1, 2)
        check(self.L.cv_mlp_set_data(self.m, X.ctypes.data, T.ctypes.data, X.shape[0]))

    def train(self, steps, batch=4096, lr=1e-3):
        hist = np.empty(max(1, steps), np.float32)
        check(self.L.cv_mlp_train(self.m, int(steps), int(batch), float(lr), hist.ctypes.data))
        return hist[:steps]

    def forward(self, X):
        X = np.ascontiguousarray(X, np.float32).reshape(-1, 4)
        Y = np.empty((X.shape[0], 2), np.float32)
        check(self.L.cv_mlp_forward(self.m, X.ctypes.data, X.shape[0], Y.ctypes.data))
        return Y

    def features(self, T00, M, T03):
        """The network input of (T^tt, |T^ti|, T^t eta) as the device forms it for the seeded
        root solve (n x 4): forward() of it gives the seeds cv_root_eval starts from."""
        a = [np.ascontiguousarray(x, np.float32).reshape(-1) for x in (T00, M, T03)]
        X = np.empty((a[0].size, 4), np.float32)
        check(self.L.cv_mlp_features(self.m, a[0].ctypes.data, a[1].ctypes.data, a[2].ctypes.data, a[0].size,
                                     X.ctypes.data))
        return X

    def close(self):
        if getattr(self, 'm', None):
            self.L.cv_mlp_destroy(self.m)
            self.m = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def selftest_fastmath(device=0):
    """cv_selftest_fastmath: ((mismatches, floats checked) of the stage kernels' reciprocal
    against 1/b on 2^-126 <= |b| < 2^126, the same of their square root against sqrtf on
    [2^-24, 1])."""
    out = (C.c_uint64 * 4)()
    check(lib().cv_selftest_fastmath(device, out))
    return (out[0], out[1]), (out[2], out[3])


def device_info(device=0):
    name = C.create_string_buffer(128)
    cus = C.c_int32()
    check(lib().cv_device_info(device, name, 128, C.byref(cus)))
    return name.value.decode(), cus.value
