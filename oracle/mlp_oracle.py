"""CPU oracle of the learned root-solve seed network (TEST INFRASTRUCTURE ONLY).

Restates the arithmetic of clvisc_amd/csrc/cv_mlp.h in NumPy: a 4-64-64-2 MLP with
tanh, flat parameter vector W1[64][4], b1[64], W2[64][64], b2[64], W3[2][64], b3[2];
loss = mean over samples of sum_o (y_o - t_o)^2.  With ``bf16=True`` every operand of a
matrix product is rounded to BF16 (round to nearest even) as the device does before its
MFMA / VALU products; the products and sums themselves are evaluated in float64, so the
device (FP32 accumulation in its own order) agrees to accumulation rounding.  With
``bf16=False`` it is the plain float64 network, pinned by central finite differences in
the tests.  The reference has no counterpart of this network.

Also restates the seed features of the device (seed_features in cv_mlp.h).
"""
import numpy as np

NIN, NH, NOUT = 4, 64, 2
OFF_W1 = 0
OFF_B1 = OFF_W1 + NH * NIN
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NOUT * NH
NPARAM = OFF_B3 + NOUT


def bf16r(x):
    """Round float32 values to BF16 (round to nearest even) and back to float32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def unpack(P):
    P = np.asarray(P, np.float64)
    return (P[OFF_W1:OFF_B1].reshape(NH, NIN), P[OFF_B1:OFF_W2], P[OFF_W2:OFF_B2].reshape(NH, NH),
            P[OFF_B2:OFF_W3], P[OFF_W3:OFF_B3].reshape(NOUT, NH), P[OFF_B3:NPARAM])


def _q(x, bf16):
    return bf16r(np.asarray(x, np.float32)).astype(np.float64) if bf16 else np.asarray(x, np.float64)


def forward(P, X, bf16=True):
    """(h1, h2, y) for features X (n x 4)."""
    W1, b1, W2, b2, W3, b3 = unpack(np.asarray(P, np.float32) if bf16 else P)
    X = np.asarray(X, np.float32 if bf16 else np.float64)
    h1 = np.tanh(_q(X, bf16) @ _q(W1, bf16).T + b1)
    h2 = np.tanh(_q(h1, bf16) @ _q(W2, bf16).T + b2)
    y = _q(h2, bf16) @ _q(W3, bf16).T + b3
    return h1, h2, y


def loss_and_grad(P, X, T, bf16=True):
    """(loss, gradient vector) of mean_samples sum_o (y_o - t_o)^2."""
    W1, b1, W2, b2, W3, b3 = unpack(np.asarray(P, np.float32) if bf16 else P)
    X = np.asarray(X, np.float32 if bf16 else np.float64)
    T = np.asarray(T, np.float64)
    n = X.shape[0]
    h1, h2, y = forward(P, X, bf16)
    d = y - T
    loss = float((d * d).sum() / n)
    dy = 2.0 * d / n
    g = np.zeros(NPARAM)
    g[OFF_W3:OFF_B3] = (dy.T @ _q(h2, bf16)).reshape(-1)
    g[OFF_B3:NPARAM] = dy.sum(0)
    dz2 = (dy @ _q(W3, bf16)) * (1.0 - h2 * h2)
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    g[OFF_W2:OFF_B2] = (_q(dz2, bf16).T @ _q(h1, bf16)).reshape(-1)
    dz1 = (_q(dz2, bf16) @ _q(W2, bf16)) * (1.0 - h1 * h1)
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    g[OFF_W1:OFF_B1] = (dz1.T @ _q(X, bf16)).reshape(-1)
    return loss, g


def features(T00, M, T03):
    """Seed features (log2 T00 / 16, M / T00, T03 / T00, octave of T00 / 24), float32."""
    T00 = np.maximum(np.asarray(T00, np.float32), np.float32(1e-30))
    M = np.asarray(M, np.float32)
    T03 = np.asarray(T03, np.float32)
    _, ex = np.frexp(T00)
    return np.stack([np.log2(T00) * np.float32(1.0 / 16.0), M / T00, T03 / T00,
                     (ex - 1).astype(np.float32) * np.float32(1.0 / 24.0)], axis=1).astype(np.float32)
