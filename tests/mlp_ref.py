"""NumPy reference of the learned p* guess network (test infrastructure only).

Restates, in FP64, the network the GPU kernels run (csrc/mlp_guess.hpp,
csrc/mlp_train.hip): inputs as a BF16 hi + lo pair, BF16 W1, W2 and w3, hidden
layer 1 rounded to BF16, FP32 biases, GELU in its tanh form; the backward pass
lets gradients through the roundings unchanged.  quantize=False gives the plain
FP64 network, whose gradient a finite-difference check pins.
"""
import numpy as np

N_IN, N_H, NP = 6, 32, 1313


def bf16(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = a.view(np.uint32).astype(np.uint64)
    r = (((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16).astype(np.uint32)
    return r.view(np.float32)


def unpack(P):
    P = np.asarray(P, np.float64)
    return (P[0:192].reshape(N_H, N_IN), P[192:224], P[224:1248].reshape(N_H, N_H), P[1248:1280],
            P[1280:1312], P[1312])


def gelu(x):
    a = 0.7978845608 * (x + 0.044715 * x ** 3)
    return x / (1.0 + np.exp(-2.0 * a))


def gelu_grad(x):
    a = 0.7978845608 * (x + 0.044715 * x ** 3)
    s = 1.0 / (1.0 + np.exp(-2.0 * a))
    return s + x * s * (1.0 - s) * 2.0 * 0.7978845608 * (1.0 + 3.0 * 0.044715 * x * x)


def _q(a, quantize):
    return bf16(a).astype(np.float64) if quantize else np.asarray(a, np.float64)


def forward(P, X, quantize=True):
    W1, b1, W2, b2, w3, b3 = unpack(P)
    X = np.asarray(X, np.float32)
    if quantize:
        hi = bf16(X)
        xq = hi.astype(np.float64) + bf16(X - hi).astype(np.float64)
        W1, W2, w3 = _q(W1.astype(np.float32), 1), _q(W2.astype(np.float32), 1), _q(w3.astype(np.float32), 1)
    else:
        xq = X.astype(np.float64)
    a1 = xq @ W1.T + b1
    h1 = gelu(a1)
    if quantize:
        h1 = bf16(h1.astype(np.float32)).astype(np.float64)
    a2 = h1 @ W2.T + b2
    h2 = gelu(a2)
    y = h2 @ w3 + b3
    return y, (xq, a1, h1, a2, h2, W1, W2, w3)


def loss_grad(P, X, T, quantize=True):
    """Mean squared error over the rows and its gradient (NP values)."""
    y, (xq, a1, h1, a2, h2, W1, W2, w3) = forward(P, X, quantize)
    B = len(y)
    r = y - np.asarray(T, np.float64)
    dy = 2.0 * r / B
    g2 = (dy[:, None] * w3[None, :]) * gelu_grad(a2)
    g1 = (g2 @ W2) * gelu_grad(a1)
    grad = np.concatenate([(g1.T @ xq).ravel(), g1.sum(0), (g2.T @ h1).ravel(), g2.sum(0), h2.T @ dy, [dy.sum()]])
    return float(np.mean(r * r)), grad


def features(rho_L, u_L, p_L, rho_R, u_R, p_R, g=1.4, gR=None):
    """Network inputs and the two-rarefaction estimate p_TR, in FP64: the closed
    form with the mean exponent, two Newton steps in log p where the gammas differ."""
    gR = np.asarray(g if gR is None else gR, dtype=np.float64)
    g = np.asarray(g, dtype=np.float64)
    c_L, c_R = np.sqrt(g * p_L / rho_L), np.sqrt(gR * p_R / rho_R)
    x = np.stack(np.broadcast_arrays(np.log(rho_L / rho_R) / np.log(1e3), np.log(p_L / p_R) / np.log(1e4),
                                     (u_R - u_L) / c_L / 5, np.log(c_R / c_L) / 5, g - 1.4, gR - 1.4), axis=-1)
    zl, zr = (g - 1) / (2 * g), (gR - 1) / (2 * gR)
    z = 0.5 * (zl + zr)
    al, ar = 2 * c_L / (g - 1), 2 * c_R / (gR - 1)
    du = u_R - u_L
    y = np.log((al + ar - du) / (al * p_L ** -z + ar * p_R ** -z)) / z
    for _ in range(2):
        ql, qr = np.exp(zl * (y - np.log(p_L))), np.exp(zr * (y - np.log(p_R)))
        F = al * (ql - 1) + ar * (qr - 1) + du
        y = np.where(g != gR, y - F / (al * zl * ql + ar * zr * qr), y)
    return x, np.exp(y)
