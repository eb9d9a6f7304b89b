"""Equation-of-state tables for the hydro step.

Mirrors the reference ``Eos`` object (pyvisc/eos/eos.py:20-203): after
construction it carries ``ed, pr, T, s, cs2, ed_start, ed_step, num_of_ed`` and
the callables ``f_P, f_T, f_S, f_ed, f_cs2``; ``create_table`` returns the
float32 RGBA = (cs2, P, T, s) table the device looks values up in.

Two table kinds exist:

* ``'rowcol'`` -- the reference layout: a uniform-step table starting at
  ``ed_start`` stored as ``nrow x ncol`` texels, looked up with the reference's
  row/col arithmetic and an analytic fit above the table
  (pyvisc/kernel/eos_table.h:39-100).  ``Eos('ideal_gas')`` builds this one
  exactly as pyvisc/eos/eos.py:42-59 does.
* ``'octave'`` -- a logarithmically spaced table whose nodes sit uniformly
  inside each binary octave of e: 2^k (1 + j/J), k in [kmin, kmax), J = 2^log2j
  (default 2^-14..2^10 GeV/fm^3, J = 64: 1537 nodes, 24.6 KB, relative spacing
  1/128..1/64), small enough to sit in LDS on the device.  The node and the
  interpolation weight are read off the binary exponent and mantissa of e, so the
  lookup needs no log and no division and is reproducible bit for bit on any
  IEEE device.  Values are interpolated linearly in e; below 2^kmin P, T and s
  fall linearly to zero (cs2 held); at or above 2^kmax they scale linearly with e.
  ``Eos('crossover')`` builds the smooth synthetic crossover EOS used by the
  385x385x115 viscous workload.

The reference's lattice tables (s95p-PCE, HotQCD, Wuppertal-Budapest) are not
shipped with the reference repository, so only the two closed-form EOS above
are provided; a user table can be wrapped with :class:`EosCraft`.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

HBARC = 0.1973269  # pyvisc/eos/eos.py:44


@dataclass
class EosTable:
    """Device-ready EOS description (host arrays; uploaded by the driver)."""
    kind: str                 # 'rowcol' or 'octave'
    rgba: np.ndarray          # float32, shape (nrow, ncol, 4) or (n, 4)
    nrow: int = 0
    ncol: int = 0
    num_ed: int = 0
    ed_start: float = 0.0
    ed_step: float = 0.0
    kmin: int = 0
    kmax: int = 0
    log2j: int = 0
    cs2_max: float = 1.0 / 3.0
    cs2_min: float = 1.0 / 3.0

    @property
    def kind_id(self) -> int:
        return 0 if self.kind == 'rowcol' else 1


def octave_nodes(kmin, kmax, log2j):
    """Energy densities of the octave-table nodes: 2^k (1 + j/J) for k in [kmin, kmax),
    j in [0, J) with J = 2^log2j, then 2^kmax (node n = (k - kmin) J + j)."""
    J = 1 << log2j
    k = np.arange(kmin, kmax, dtype=np.float64)
    j = np.arange(J, dtype=np.float64)
    e = (np.exp2(k)[:, None] * (1.0 + j[None, :] / J)).ravel()
    return np.append(e, 2.0 ** kmax)


class Eos(object):
    """EOS with table and interpolating functions.

    ``Eos('ideal_gas')`` follows pyvisc/eos/eos.py:42-59;
    ``Eos('crossover')`` is the synthetic smooth crossover (project-defined).
    """

    def __init__(self, eos_type='ideal_gas', **kw):
        self.eos_type = eos_type
        if eos_type == 'ideal_gas':
            self.ideal_gas()
        elif eos_type == 'crossover':
            self.crossover(**kw)
        elif eos_type == 'crossover_rowcol':
            self.crossover(kind='rowcol', **kw)
        else:
            raise ValueError(
                "eos_type %r is not available: the reference's lattice tables are not "
                "shipped with it; use 'ideal_gas', 'crossover' or EosCraft" % (eos_type,))

    # -- reference ideal gas (pyvisc/eos/eos.py:42-59) -------------------------------
    def ideal_gas(self):
        hbarc = HBARC
        dof = 169.0 / 4.0
        coef = np.pi * np.pi / 30.0
        self.f_P = lambda ed: np.array(ed) / 3.0
        self.f_T = lambda ed: hbarc * (1.0 / (dof * coef) * np.array(ed) / hbarc) ** 0.25 + 1.0E-10
        self.f_S = lambda ed: (np.array(ed) + self.f_P(ed)) / self.f_T(ed)
        self.f_ed = lambda T: dof * coef * hbarc * (np.array(T) / hbarc) ** 4.0
        self.f_cs2 = lambda ed: 1.0 / 3.0 * np.ones_like(ed)
        self.ed = np.linspace(0, 1999.99, 200000)
        self.pr = self.f_P(self.ed)
        self.T = self.f_T(self.ed)
        self.s = self.f_S(self.ed)
        self.cs2 = self.f_cs2(self.ed)
        self.ed_start = 0.0
        self.ed_step = 0.01
        self.num_of_ed = 200000
        self.kind = 'rowcol'
        self.default_shape = (200, 1000)

    # -- synthetic crossover ----------------------------------------------------------
    def crossover(self, kind='octave', kmin=-14, kmax=10, log2j=6,
                  cs2_dip=0.2, e_c=0.35, width=1.2, dof=169.0 / 4.0,
                  ed_step=0.002, nrow=100, ncol=1555, emax=1.0e3):
        """Smooth crossover: cs2(e) = 1/3 - cs2_dip * exp(-ln(e/e_c)^2 / (2 width^2)).

        P(e) = int_0^e cs2 de'; d ln T = cs2 de / (e + P) with T matched to a
        massless gas of ``dof`` degrees of freedom at the top of the range; s = (e+P)/T.
        Integrated in FP64 on a fine grid, then sampled into the table.
        """
        lo, hi = math.log(1.0e-10), math.log(max(emax, 2.0e3) * 2.0)
        x = np.linspace(lo, hi, 400001)
        e = np.exp(x)
        cs2 = 1.0 / 3.0 - cs2_dip * np.exp(-(x - math.log(e_c)) ** 2 / (2.0 * width * width))
        # P(e) = e0*cs2(e0) + int cs2 de ; trapezoid in e
        de = np.diff(e)
        P = np.empty_like(e)
        P[0] = cs2[0] * e[0]
        P[1:] = P[0] + np.cumsum(0.5 * (cs2[1:] + cs2[:-1]) * de)
        # ln T from the top: T_top from a massless gas e = dof pi^2/30 T^4 / hbarc^3
        coef = np.pi * np.pi / 30.0
        T_top = HBARC * (e[-1] / (dof * coef) / HBARC) ** 0.25
        g = cs2 / (e + P)
        lnT = np.empty_like(e)
        seg = 0.5 * (g[1:] + g[:-1]) * de
        lnT[-1] = math.log(T_top)
        lnT[:-1] = lnT[-1] - np.cumsum(seg[::-1])[::-1]
        T = np.exp(lnT)
        S = (e + P) / T
        self._fine = (x, e, P, T, S, cs2)

        def interp_lnE(col):
            def f(ed):
                ed = np.asarray(ed, dtype=np.float64)
                xe = np.log(np.maximum(ed, 1e-300))
                out = np.interp(xe, x, col)
                return np.where(ed > 0, out, 0.0 * out + (col[0] if col is cs2 else 0.0))
            return f

        self.f_P = interp_lnE(P)
        self.f_T = interp_lnE(T)
        self.f_S = interp_lnE(S)
        self.f_cs2 = interp_lnE(cs2)
        self.f_ed = lambda Tq: np.exp(np.interp(np.log(np.asarray(Tq, dtype=np.float64)), lnT, x))
        self.kind = kind
        if kind == 'octave':
            self.kmin, self.kmax, self.log2j = kmin, kmax, log2j
            self.ed = octave_nodes(kmin, kmax, log2j)
            self.num_of_ed = len(self.ed)
            self.ed_start = 2.0 ** kmin
            self.ed_step = 0.0
            self.default_shape = (len(self.ed),)
        else:
            self.ed = ed_step * np.arange(nrow * ncol, dtype=np.float64)
            self.num_of_ed = nrow * ncol
            self.ed_start = 0.0
            self.ed_step = ed_step
            self.default_shape = (nrow, ncol)
        self.pr = self.f_P(self.ed)
        self.T = self.f_T(self.ed)
        self.s = self.f_S(self.ed)
        self.cs2 = self.f_cs2(self.ed)

    # -- table ------------------------------------------------------------------------
    def create_table(self, ctx=None, compile_options=None, nrow=None, ncol=None):
        """Return the float32 (cs2, P, T, s) table as an :class:`EosTable`.

        Signature kept from pyvisc/eos/eos.py:184: ``ctx`` is ignored; when a
        ``compile_options`` list is given, the reference's ``-D EOS_*`` entries are
        appended to it for callers that inspect them.
        """
        src = np.stack([self.cs2, self.pr, self.T, self.s], axis=-1).astype(np.float32)
        if self.kind == 'octave':
            tab = EosTable(kind='octave', rgba=np.ascontiguousarray(src.reshape(-1, 4)),
                           kmin=int(self.kmin), kmax=int(self.kmax), log2j=int(self.log2j),
                           num_ed=int(self.num_of_ed),
                           cs2_max=float(np.max(src[:, 0])), cs2_min=float(np.min(src[:, 0])))
        else:
            if nrow is None or ncol is None:
                nrow, ncol = self.default_shape
            if nrow * ncol != src.shape[0]:
                raise ValueError('nrow*ncol=%d does not match %d table entries'
                                 % (nrow * ncol, src.shape[0]))
            tab = EosTable(kind='rowcol', rgba=np.ascontiguousarray(src.reshape(nrow, ncol, 4)),
                           nrow=int(nrow), ncol=int(ncol), num_ed=int(self.num_of_ed),
                           ed_start=float(self.ed_start), ed_step=float(self.ed_step),
                           cs2_max=float(np.max(src[..., 0])), cs2_min=float(np.min(src[..., 0])))
        if compile_options is not None:
            compile_options.append('-D EOS_ED_START={}f'.format(self.ed_start))
            compile_options.append('-D EOS_ED_STEP={}f'.format(self.ed_step))
            compile_options.append('-D EOS_NUM_ED={}'.format(self.num_of_ed))
            if tab.kind == 'rowcol':
                compile_options.append('-D EOS_NUM_OF_ROWS=%s' % tab.nrow)
                compile_options.append('-D EOS_NUM_OF_COLS=%s' % tab.ncol)
            self.compile_options = compile_options
        self.table = tab
        return tab


class EosCraft(Eos):
    """User-supplied EOS (the reference's ``handcrafted_eos=`` plugin point,
    pyvisc/ideal.py:30,60-69): give uniform-step arrays ``ed, pr, T, s`` (and
    optionally ``cs2``; default dP/de) starting at ``ed_start`` with ``ed_step``."""

    def __init__(self, ed, pr, T, s, ed_step, ed_start=0.0, cs2=None, shape=None):
        self.eos_type = 'craft'
        self.ed = np.asarray(ed, dtype=np.float64)
        self.pr = np.asarray(pr, dtype=np.float64)
        self.T = np.asarray(T, dtype=np.float64)
        self.s = np.asarray(s, dtype=np.float64)
        self.cs2 = np.gradient(self.pr, ed_step) if cs2 is None else np.asarray(cs2, np.float64)
        self.ed_start = float(ed_start)
        self.ed_step = float(ed_step)
        self.num_of_ed = len(self.ed)
        self.kind = 'rowcol'
        self.default_shape = shape if shape is not None else (1, len(self.ed))
        self.f_P = lambda e: np.interp(e, self.ed, self.pr)
        self.f_T = lambda e: np.interp(e, self.ed, self.T)
        self.f_S = lambda e: np.interp(e, self.ed, self.s)
        self.f_cs2 = lambda e: np.interp(e, self.ed, self.cs2)
        self.f_ed = lambda t: np.interp(t, self.T, self.ed)
