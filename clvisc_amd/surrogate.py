"""Learned seed of the conserved-to-primitive root solve.

Data: the converged states (e, v) of the plain method's own runs, turned back into the
conserved input of the root solve, T^tt = (e + P) u0^2 - P, |T^ti| = (e + P) u0^2 |v|,
T^t eta = (e + P) u0^2 v_eta (the map the solve inverts).  Features and targets are the
ones of the device (include/clvisc_amd.h, cv_mlp.h): x = (log2 T^tt / 16, |T^ti| / T^tt,
T^t eta / T^tt, octave of T^tt / 24), y = (e / T^tt, |v|).  Training and inference run on
the device (_lib.Mlp); the seeded solve keeps the reference's bracket and stopping rule,
so the seed changes only where the iteration starts.
"""
from __future__ import annotations

import numpy as np

from . import _lib, ic
from .eos import Eos


def features(T00, M, T03):
    """Seed features of (T^tt, |T^ti|, T^t eta), float32 (as seed_features on the device)."""
    T00 = np.maximum(np.asarray(T00, np.float32), np.float32(1e-30))
    M = np.asarray(M, np.float32)
    T03 = np.asarray(T03, np.float32)
    _, ex = np.frexp(T00)
    return np.stack([np.log2(T00) * np.float32(1.0 / 16.0), M / T00, T03 / T00,
                     (ex - 1).astype(np.float32) * np.float32(1.0 / 24.0)], axis=1).astype(np.float32)


def states_from_fields(ev, eos):
    """(T00, M, T03, e, |v|) of every cell of an (N, 4) field (e, vx, vy, veta)."""
    ev = np.asarray(ev, np.float64).reshape(-1, 4)
    e = np.maximum(ev[:, 0], 0.0)
    v2 = np.minimum((ev[:, 1:] ** 2).sum(1), 0.999 ** 2)
    P = np.asarray(eos.f_P(e), np.float64)
    w = (e + P) / (1.0 - v2)
    T00 = w - P
    M = w * np.sqrt(v2)
    T03 = w * ev[:, 3]
    return (T00.astype(np.float32), M.astype(np.float32), T03.astype(np.float32), e.astype(np.float32),
            np.sqrt(v2).astype(np.float32))


def gubser_run_states(nsteps=200, nx=201, every=1, device=0, seed_mlp=None, return_handle=False):
    """The ideal Gubser-flow workload (nx^2 x 1, DX 0.08, DT 0.02, tau0 1, q = 1/fm, e0 from
    T = 0.3 GeV, ideal-gas EOS): run nsteps plain steps, collect the states of every
    `every`-th step.  Returns (T00, M, T03, e, |v|) arrays and the iteration counts of the
    last stage of each collected step (plain or seeded)."""
    eos = Eos('ideal_gas')
    tab = eos.create_table()
    p = _lib.CvParams()
    p.nx, p.ny, p.nz = nx, nx, 1
    p.tau0, p.dt = 1.0, 0.02
    p.dx, p.dy, p.dz = 0.08, 0.08, 0.15
    p.etaos_xmin, p.etaos_ymin = 0.18, 0.08
    p.lam1 = -10.0
    p.nxskip = p.nyskip = p.nzskip = 1
    h = _lib.Handle(p, tab, device=device, viscous=False)
    h.set_root_seed(seed_mlp, track_iters=True)
    ev, _ = ic.gubser_ideal(nx, nx, 1, 0.08, 0.08, 0.15, 1.0, q=1.0, e0=float(eos.f_ed(0.3)))
    h.set('ev1', ev.astype(np.float32))
    cols = [[], [], [], [], []]
    iters = []
    for n in range(nsteps):
        h.ideal_steps(n, 1)
        if (n + 1) % every == 0:
            for c, a in zip(cols, states_from_fields(h.get('ev1'), eos)):
                c.append(a)
            iters.append(h.root_iters())
    out = tuple(np.concatenate(c) for c in cols), np.concatenate(iters)
    if return_handle:
        return out + (h,)
    h.close()
    return out


def clamp_cons(cons):
    """(T00, M, T03) of the split stage's conserved input (n x 4, before the clamp) after the
    |T^ti| <= 0.999 T^tt clamp of the root solve (k_visc_root / k_mlp_seed_cons)."""
    c = np.asarray(cons, np.float32).reshape(-1, 4)
    T00, T03 = c[:, 0].copy(), c[:, 3].copy()
    M = np.sqrt(c[:, 1] * c[:, 1] + c[:, 2] * c[:, 2] + c[:, 3] * c[:, 3]).astype(np.float32)
    big = M > T00
    sc = (np.float32(0.999) * T00[big]) * (np.float32(1.0) / M[big])
    T03[big] = T03[big] * sc
    M[big] = np.float32(0.999) * T00[big]
    return T00, M, T03


def split_states(h):
    """(T00, M, T03, e, |v|) of every cell of the last split viscous stage of handle h:
    its conserved input (field root_cons) and the solution it stepped to (ev1 after a full
    step).  The training data of the viscous seed."""
    T00, M, T03 = clamp_cons(h.get('root_cons'))
    ev = h.get('ev1')
    v = np.sqrt((ev[:, 1:].astype(np.float64) ** 2).sum(1)).astype(np.float32)
    return T00, M, T03, ev[:, 0].copy(), v


def visc_run_states(h, n0, nsteps, every=1):
    """Run nsteps viscous steps of handle h (split form, no seed) from loop index n0,
    collecting split_states every `every` steps."""
    h.set_root_split(True)
    cols = [[], [], [], [], []]
    for k in range(nsteps):
        h.visc_steps(n0 + k, 1)
        if (k + 1) % every == 0:
            for c, a in zip(cols, split_states(h)):
                c.append(a)
    return tuple(np.concatenate(c) for c in cols)


def train_seed(states, steps=2000, batch=4096, lr=2e-3, nmax=5_000_000, seed=1234, device=0, lr_decay=True):
    """Train the seed network on (a random subset of at most nmax) states.  Returns
    (mlp, loss history)."""
    T00, M, T03, e, v = states
    rng = np.random.default_rng(seed)
    keep = np.nonzero(T00 > 1e-6)[0]
    if keep.size > nmax:
        keep = rng.choice(keep, nmax, replace=False)
    rng.shuffle(keep)
    X = features(T00[keep], M[keep], T03[keep])
    Y = np.stack([e[keep] / np.maximum(T00[keep], 1e-30), v[keep]], axis=1).astype(np.float32)
    mlp = _lib.Mlp(device=device, seed=seed)
    mlp.set_data(X, Y)
    hist = []
    if lr_decay:
        # three phases: lr, lr / 4, lr / 16
        for k, f in enumerate((1.0, 0.25, 0.0625)):
            n = steps // 3 if k < 2 else steps - 2 * (steps // 3)
            hist.append(mlp.train(n, batch, lr * f))
    else:
        hist.append(mlp.train(steps, batch, lr))
    return mlp, np.concatenate(hist)
