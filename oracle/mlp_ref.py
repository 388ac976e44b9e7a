"""NumPy reference of the surrogate MLP -- TEST INFRASTRUCTURE ONLY.

Same architecture and parameter layout as fd_adi_amd/csrc/mlp.hip:
8 features (zero-padded to 128) -> 4 x [512, SiLU] -> 1, loss = mean squared
error of the standardised target.  ``bf16=True`` rounds to bfloat16 at the
points where the device kernels store bf16 (inputs, weights, activations,
SiLU derivatives at the pre-activations, back-propagated pre-activation
gradients); accumulation is float64 here
(float32 on the device).  With ``bf16=False`` the gradients are those of the
float64 network, which tests/test_oracle_mlp.py pins with finite differences.
"""

import numpy as np

N_IN, N_PAD, N_H, N_LAYERS = 8, 128, 512, 4


def bf16_round(x):
    """Round float64/float32 values to the nearest bfloat16 (ties to even)."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    u = (u + 0x7FFF + lsb) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def bf16_bits(x):
    f = bf16_round(x).astype(np.float32)
    return (f.view(np.uint32) >> 16).astype(np.uint16)


def layout():
    """Offsets (W_l, b_l) of the flat parameter array, then (w4, b4)."""
    offs, o = [], 0
    for l in range(N_LAYERS):
        fan_in = N_PAD if l == 0 else N_H
        offs.append((o, o + N_H * fan_in))
        o += N_H * fan_in + N_H
    return offs, o, o + N_H, o + N_H + 1


def unpack(p):
    offs, o4, ob4, n = layout()
    Ws, bs = [], []
    for l, (ow, ob) in enumerate(offs):
        fan_in = N_PAD if l == 0 else N_H
        Ws.append(p[ow:ow + N_H * fan_in].reshape(N_H, fan_in).astype(np.float64))
        bs.append(p[ob:ob + N_H].astype(np.float64))
    return Ws, bs, p[o4:o4 + N_H].astype(np.float64), float(p[ob4])


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def features(P, lo, hi):
    """Normalised input rows, zero-padded to 128 columns (P has up to 16
    feature columns; 8 for the Heston surrogate)."""
    X = np.zeros((P.shape[0], N_PAD))
    X[:, :P.shape[1]] = 2.0 * (P - lo) / (hi - lo) - 1.0
    return X


def forward(p, X, bf16=False):
    r = bf16_round if bf16 else (lambda a: a)
    Ws, bs, w4, b4 = unpack(p)
    A = r(X)
    Zs, As = [], [A]
    for W, b in zip(Ws, bs):
        Z = A.dot(r(W).T) + b
        A = r(silu(Z))
        Zs.append(Z)
        As.append(A)
    y = A.dot(r(w4)) + b4
    return y, Zs, As


def loss_and_grad(p, X, t, bf16=False):
    """MSE on standardised targets t and its gradient (flat, like p)."""
    r = bf16_round if bf16 else (lambda a: a)
    Ws, bs, w4, b4 = unpack(p)
    y, Zs, As = forward(p, X, bf16)
    B = X.shape[0]
    e = y - t
    loss = float(np.mean(e * e))
    dy = 2.0 * e / B
    g = np.zeros_like(p, dtype=np.float64)
    offs, o4, ob4, _ = layout()
    g[o4:o4 + N_H] = As[-1].T.dot(dy)
    g[ob4] = dy.sum()
    dZ = r(dy[:, None] * r(w4)[None, :] * r(dsilu(Zs[-1])))
    for l in range(N_LAYERS - 1, -1, -1):
        ow, ob = offs[l]
        fan_in = N_PAD if l == 0 else N_H
        g[ow:ow + N_H * fan_in] = dZ.T.dot(As[l]).ravel()
        g[ob:ob + N_H] = dZ.sum(axis=0)
        if l > 0:
            dZ = r(dZ.dot(r(Ws[l])) * r(dsilu(Zs[l - 1])))
    return loss, g
