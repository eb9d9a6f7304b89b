"""The IS vorticity couplings of the oracle (orc_couplings; kernel_IS.cl:354-470, applied at
kernel_IS.cl:636-683) pinned by what the mathematics fixes, independently of the formula:

- the rest frame (u = (1,0,0,0), pi^{0 mu} = 0): Omega^{0i} = 0 and Omega^{ij} = omega^{ij} / 2;
  sum_lam g pi_omega = [P, W] / 2 (P the spatial pi, W the spatial Omega) and
  sum_lam g omega_omega = W^2 - tr(W^2) / 3 * 1 (spatial), all time components zero;
- Lorentz covariance: boosting u, pi and omega^{mu nu} (time components included, which the
  projector must remove) by one boost L turns every output X into L X L^T.
A dropped projector term, a wrong lowering sign or a transposed row breaks one of these."""
import numpy as np
import pytest

from oracle import oracle as O

G = np.diag([1.0, -1.0, -1.0, -1.0])
UPPER = [(0, 0), (0, 1), (0, 2), (0, 3), (1, 1), (1, 2), (1, 3), (2, 2), (2, 3), (3, 3)]
PAIRS = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]


def _sym10(m):
    return np.array([m[a, b] for a, b in UPPER])


def _from10(v):
    m = np.zeros((4, 4))
    for k, (a, b) in enumerate(UPPER):
        m[a, b] = m[b, a] = v[k]
    return m


def _antisym(om6):
    w = np.zeros((4, 4))
    for k, (a, b) in enumerate(PAIRS):
        w[a, b], w[b, a] = om6[k], -om6[k]
    return w


def _derivs_for(om6):
    """Derivative rows whose antisymmetric combination (kernel_IS.cl:640-645) is om6."""
    udt, ux, uy, uz = np.zeros(4), np.zeros(4), np.zeros(4), np.zeros(4)
    udt[1:4] = om6[0:3]
    ux[2], ux[3], uy[3] = om6[3], om6[4], om6[5]
    return udt, ux, uy, uz


def _boost(rng):
    """A general Lorentz boost (velocity |beta| < 0.8 in a random direction)."""
    n = rng.normal(size=3)
    n /= np.linalg.norm(n)
    beta = rng.uniform(0.2, 0.8)
    g = 1.0 / np.sqrt(1.0 - beta * beta)
    L = np.eye(4)
    L[0, 0] = g
    L[0, 1:] = L[1:, 0] = g * beta * n
    L[1:, 1:] += (g - 1.0) * np.outer(n, n)
    return L


def _rest_case(rng):
    P = rng.normal(size=(3, 3))
    P = P + P.T
    om = rng.normal(size=6)
    return P, om


def _expected_rest(P, om):
    W = _antisym(om)[1:, 1:] * 0.5
    X = np.zeros((4, 4))
    X[1:, 1:] = 0.5 * (P @ W - W @ P)
    Y = np.zeros((4, 4))
    W2 = W @ W
    Y[1:, 1:] = W2 - np.trace(W2) / 3.0 * np.eye(3)
    Om = np.zeros((4, 4))
    Om[1:, 1:] = W
    return X, Y, Om


@pytest.mark.parametrize('precision,tol', [('f64', 1e-12), ('f32', 2e-5)])
def test_rest_frame_closed_forms(precision, tol):
    rng = np.random.default_rng(7)
    for _ in range(20):
        P, om = _rest_case(rng)
        pim = np.zeros((4, 4))
        pim[1:, 1:] = P
        udt, ux, uy, uz = _derivs_for(om)
        out = O.couplings(_sym10(pim), udt, ux, uy, uz, np.array([1.0, 0, 0, 0]), precision)
        X, Y, Om = _expected_rest(P, om)
        scale = 1.0 + np.abs(out).max()
        np.testing.assert_allclose(out[20:26], [Om[a, b] for a, b in PAIRS], atol=tol * scale)
        np.testing.assert_allclose(out[0:10], _sym10(X), atol=tol * scale)
        np.testing.assert_allclose(out[10:20], _sym10(Y), atol=tol * scale)


@pytest.mark.parametrize('precision,tol', [('f64', 1e-10), ('f32', 5e-4)])
def test_boost_covariance(precision, tol):
    """Rest-frame pi (spatial) and a full omega^{mu nu} (its 0i part projected away), boosted
    by L: the outputs are L X L^T of the rest-frame closed forms."""
    rng = np.random.default_rng(11)
    for _ in range(20):
        P, om = _rest_case(rng)
        om_t = rng.normal(size=3)             # time components of omega: must drop out
        L = _boost(rng)
        w_rest = _antisym(np.concatenate([om_t, om[3:]]))
        pim = np.zeros((4, 4))
        pim[1:, 1:] = P
        u = L @ np.array([1.0, 0, 0, 0])
        pim_b = L @ pim @ L.T
        w_b = L @ w_rest @ L.T
        om_b = np.array([w_b[a, b] for a, b in PAIRS])
        udt, ux, uy, uz = _derivs_for(om_b)
        out = O.couplings(_sym10(pim_b), udt, ux, uy, uz, u, precision)
        X, Y, Om = _expected_rest(P, om)
        Xb, Yb, Omb = L @ X @ L.T, L @ Y @ L.T, L @ Om @ L.T
        scale = 1.0 + max(np.abs(Xb).max(), np.abs(Yb).max(), np.abs(Omb).max())
        np.testing.assert_allclose(out[20:26], [Omb[a, b] for a, b in PAIRS], atol=tol * scale)
        np.testing.assert_allclose(out[0:10], _sym10(Xb), atol=tol * scale)
        np.testing.assert_allclose(out[10:20], _sym10(Yb), atol=tol * scale)
        # the pi-omega tensor is transverse and traceless
        Xo = _from10(out[0:10])
        np.testing.assert_allclose(Xo @ G @ u, 0.0, atol=tol * scale)
        assert abs(np.trace(G @ Xo)) < tol * scale


def test_zero_vorticity_gives_zero():
    rng = np.random.default_rng(3)
    P, _ = _rest_case(rng)
    L = _boost(rng)
    pim = np.zeros((4, 4))
    pim[1:, 1:] = P
    z = np.zeros(4)
    out = O.couplings(_sym10(L @ pim @ L.T), z, z, z, z, L @ np.array([1.0, 0, 0, 0]), 'f64')
    assert np.abs(out).max() < 1e-12
