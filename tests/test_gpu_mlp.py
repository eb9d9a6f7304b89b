"""The learned root-solve seed on the device: forward and gradient against the CPU oracle
(oracle/mlp_oracle.py, BF16 operand rounding emulated, float64 sums), training, and the
seeded root solve against the plain one and the oracle."""
import sys

import numpy as np
import pytest

from clvisc_amd import _lib, surrogate
from clvisc_amd.eos import Eos
from oracle import mlp_oracle as MO
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg)
    sys.stdout.flush()


def _random_net(seed, n):
    rng = np.random.default_rng(seed)
    P = (rng.standard_normal(MO.NPARAM) * 0.25).astype(np.float32)
    X = rng.standard_normal((n, 4)).astype(np.float32)
    T = rng.standard_normal((n, 2)).astype(np.float32)
    return P, X, T


def test_forward_matches_oracle():
    """Also pins the MFMA operand layout: a transposed or mis-indexed tile fails."""
    _say('case: mlp forward, 1000 samples')
    P, X, _ = _random_net(1, 1000)
    m = _lib.Mlp()
    m.set_params(P)
    np.testing.assert_array_equal(m.params(), P)
    Y = m.forward(X)
    _, _, Yr = MO.forward(P, X, bf16=True)
    err = np.abs(Y - Yr).max()
    _say('  max |y - y_oracle| = %.3g (|y| max %.3g)' % (err, np.abs(Yr).max()))
    assert err < 2e-5 * max(1.0, np.abs(Yr).max())


@pytest.mark.parametrize('n', [16, 777, 4096])
def test_gradient_matches_oracle(n):
    _say('case: mlp gradient, %d samples' % n)
    P, X, T = _random_net(2 + n, n)
    m = _lib.Mlp()
    m.set_params(P)
    g, loss = m.grad(X, T)
    lr, gr = MO.loss_and_grad(P, X, T, bf16=True)
    rel = np.linalg.norm(g - gr) / np.linalg.norm(gr)
    _say('  loss %.6g vs %.6g, gradient rel err %.3g' % (loss, lr, rel))
    assert abs(loss - lr) <= 1e-5 * abs(lr)
    assert rel < 2e-3
    # per block of parameters (W1, b1, W2, b2, W3, b3)
    for a, b in ((MO.OFF_W1, MO.OFF_B1), (MO.OFF_B1, MO.OFF_W2), (MO.OFF_W2, MO.OFF_B2), (MO.OFF_B2, MO.OFF_W3),
                 (MO.OFF_W3, MO.OFF_B3), (MO.OFF_B3, MO.NPARAM)):
        assert np.linalg.norm(g[a:b] - gr[a:b]) <= 5e-3 * np.linalg.norm(gr[a:b]) + 1e-7, (a, b)


def _ideal_states(n, seed):
    """Conserved states of the ideal-gas EOS from random (e, v)."""
    rng = np.random.default_rng(seed)
    e = np.exp(rng.uniform(np.log(1e-3), np.log(200.0), n))
    v = np.minimum(rng.beta(1.2, 2.0, n), 0.995)
    dirn = rng.standard_normal((n, 3))
    dirn /= np.linalg.norm(dirn, axis=1, keepdims=True)
    ev = np.column_stack([e, v[:, None] * dirn]).astype(np.float32)
    return surrogate.states_from_fields(ev, Eos('ideal_gas'))


def test_training_and_seeded_root_solve():
    """Train on states of the ideal-gas map; the seeded solve converges to the plain
    solution (same bracket and stopping rule) in fewer iterations, and the device solve
    (plain and seeded) equals the FP32 oracle bit for bit."""
    states = _ideal_states(200000, 5)
    _say('case: train seed on 200000 ideal-gas states, 1200 steps of 4096')
    mlp, hist = surrogate.train_seed(states, steps=1200, batch=4096, lr=3e-3)
    _say('  loss %.4g -> %.4g' % (hist[:20].mean(), hist[-20:].mean()))
    assert hist[-20:].mean() < 0.1 * hist[:20].mean()
    T00, M, T03, e, v = _ideal_states(50000, 6)
    tab = Eos('ideal_gas').create_table()
    p = _lib.CvParams()
    p.nx = p.ny = p.nz = 1
    p.tau0, p.dt, p.dx, p.dy, p.dz = 1.0, 0.02, 0.08, 0.08, 0.15
    p.nxskip = p.nyskip = p.nzskip = 1
    h = _lib.Handle(p, tab, viscous=False)
    e0, v0, it0 = h.root_eval(T00, M, T03)
    e1, v1, it1 = h.root_eval(T00, M, T03, mlp)
    _say('  iterations plain mean %.3f max %d | seeded mean %.3f max %d' % (it0.mean(), it0.max(), it1.mean(),
                                                                          it1.max()))
    rel = np.abs(e1 - e0) / np.maximum(e0, 1e-6)
    _say('  seeded vs plain: max rel e %.3g, max |dv| %.3g' % (rel.max(), np.abs(v1 - v0).max()))
    assert it1.mean() < it0.mean()
    assert rel.max() < 1e-4 and np.abs(v1 - v0).max() < 2e-5
    # oracle: plain and from the same seeds
    op = O.make_params(1, 1, 1, 0.02, 0.08, 0.08, 0.15, 1.0, tab)
    er, itr = O.root_solve(op, T00, M)
    np.testing.assert_array_equal(e0, np.maximum(er, 0))
    np.testing.assert_array_equal(it0, itr)
    # the device's features equal the host's restatement (log2 within 2 ulp, the rest exactly)
    Xd, Xh = mlp.features(T00, M, T03), surrogate.features(T00, M, T03)
    np.testing.assert_array_equal(Xd[:, 1:], Xh[:, 1:])
    np.testing.assert_allclose(Xd[:, 0], Xh[:, 0], rtol=3e-7, atol=1e-9)
    # the seeds the device started from: its own features (log2f may differ from the host's
    # log2 by an ulp) through the network
    Y = mlp.forward(mlp.features(T00, M, T03))
    er1, itr1 = O.root_solve_from(op, T00, M, Y[:, 1])
    np.testing.assert_array_equal(e1, np.maximum(er1, 0))
    np.testing.assert_array_equal(it1, itr1)


def test_seeded_gubser_run_close_to_plain():
    """Ideal Gubser flow (101^2, 40 steps) with the seeded root solve stays within the
    root solve's own tolerance of the plain run."""
    _say('case: seeded vs plain gubser 101x101, 40 steps')
    (st, it_plain) = surrogate.gubser_run_states(nsteps=40, nx=101, every=1)
    mlp, _ = surrogate.train_seed(st, steps=600, batch=4096, lr=3e-3)
    st2, it_seed = surrogate.gubser_run_states(nsteps=40, nx=101, every=1, seed_mlp=mlp)
    e_p, e_s = st[3], st2[3]
    m = e_p > 1e-3 * e_p.max()
    rel = np.abs(e_s[m] - e_p[m]) / e_p[m]
    _say('  iterations plain %.3f seeded %.3f; max rel e diff %.3g' % (it_plain.mean(), it_seed.mean(), rel.max()))
    assert rel.max() < 1e-3
