"""Pins of the CPU oracle of the root-solve seed network (oracle/mlp_oracle.py).  CPU only."""
import numpy as np

from oracle import mlp_oracle as M


def _net(seed=3, n=37):
    rng = np.random.default_rng(seed)
    P = rng.standard_normal(M.NPARAM) * 0.3
    X = rng.standard_normal((n, 4))
    T = rng.standard_normal((n, 2))
    return P, X, T


def test_gradient_matches_finite_differences():
    """float64 network: analytic gradient vs central differences on every parameter."""
    P, X, T = _net()
    _, g = M.loss_and_grad(P, X, T, bf16=False)
    eps = 1e-6
    fd = np.empty(M.NPARAM)
    for q in range(M.NPARAM):
        Pp, Pm = P.copy(), P.copy()
        Pp[q] += eps
        Pm[q] -= eps
        fd[q] = (M.loss_and_grad(Pp, X, T, bf16=False)[0] - M.loss_and_grad(Pm, X, T, bf16=False)[0]) / (2 * eps)
    err = np.abs(g - fd).max() / np.abs(fd).max()
    assert err < 1e-6, err


def test_planted_derivative_fails():
    """A tanh derivative written as (1 - h) instead of (1 - h^2) is caught by the same check."""
    P, X, T = _net(5, 11)
    W1, b1, W2, b2, W3, b3 = M.unpack(P)
    h1, h2, y = M.forward(P, X, bf16=False)
    dy = 2.0 * (y - T) / X.shape[0]
    wrong_db2 = ((dy @ W3) * (1.0 - h2)).sum(0)
    _, g = M.loss_and_grad(P, X, T, bf16=False)
    assert not np.allclose(wrong_db2, g[M.OFF_B2:M.OFF_W3], rtol=1e-3)


def test_bf16_rounding_known_answers():
    one = np.float32(1.0)
    vals = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -7, -2.5, 3.0e-39, 65504.0],
                    np.float32)
    got = M.bf16r(vals)
    want = np.array([1.0, 1.0, 1.0 + 2.0 ** -6, 1.0 + 2.0 ** -7, -2.5, np.float32(3.0e-39), 65536.0], np.float32)
    # 1 + 2^-8 is a tie -> even (1.0); 1 + 3 2^-8 is a tie -> even (1 + 2^-6)
    np.testing.assert_array_equal(got[:5], want[:5])
    assert got[6] == np.float32(65536.0)
    assert abs(float(got[5]) - float(vals[5])) <= 2.0 ** -134     # half a BF16 subnormal step
    assert M.bf16r(np.array([one]))[0] == one


def test_bf16_network_close_to_float64():
    """Rounding the product operands to BF16 changes the gradient at the 1e-2 level only."""
    P, X, T = _net(9, 200)
    _, g64 = M.loss_and_grad(P, X, T, bf16=False)
    _, g16 = M.loss_and_grad(P, X, T, bf16=True)
    assert np.linalg.norm(g16 - g64) / np.linalg.norm(g64) < 3e-2


def test_features():
    T00 = np.array([1.0, 3.0, 0.25], np.float32)
    Mv = np.array([0.5, 1.5, 0.0], np.float32)
    T3 = np.array([0.1, -0.3, 0.0], np.float32)
    x = M.features(T00, Mv, T3)
    np.testing.assert_allclose(x[:, 0], np.log2(T00) / 16, rtol=1e-6)
    np.testing.assert_allclose(x[:, 1], Mv / T00, rtol=1e-6)
    np.testing.assert_allclose(x[:, 3], np.array([0, 1, -2]) / 24.0, rtol=1e-6)
