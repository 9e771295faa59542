"""FP64 CPU oracle of the (a, b) proposal MLP (test infrastructure only).

Network: 10 residual-autocorrelation lags -> Linear(10,64) -> GELU ->
Linear(64,64) -> GELU -> Linear(64,3) -> (a_hat, b_hat, c), c = log-confidence.
Loss per sample (Gaussian likelihood with a learned shared precision e^c,
clipped to [C_MIN, C_MAX]):

    L = 0.5 e^c ((a_hat - a)^2 + (b_hat - b)^2) - c,   averaged over the batch.

Forward and backward are written out by hand (no autograd) so the HIP kernels
can be checked against them; the backward itself is pinned against central
finite differences in tests/test_mlp_oracle.py.

Parameter layout (flat, float): W1[10][64], b1[64], W2[64][64], b2[64],
W3[64][3], b3[3]  (row-major [in][out]).
"""

from __future__ import annotations

import math

import numpy as np

N_IN, N_H, N_OUT = 10, 64, 3
SIZES = [("W1", (N_IN, N_H)), ("b1", (N_H,)), ("W2", (N_H, N_H)), ("b2", (N_H,)), ("W3", (N_H, N_OUT)),
         ("b3", (N_OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)
C_MIN, C_MAX = -10.0, 15.0


def unpack(theta: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = theta[o:o + n].reshape(shape)
        o += n
    return out


def pack(parts: dict) -> np.ndarray:
    return np.concatenate([np.asarray(parts[name], dtype=np.float64).reshape(-1) for name, _ in SIZES])


_SQRT1_2 = 1.0 / math.sqrt(2.0)
_INV_SQRT_2PI = 1.0 / math.sqrt(2.0 * math.pi)
_erf = np.vectorize(math.erf)


def gelu(x):
    return 0.5 * x * (1.0 + _erf(x * _SQRT1_2))


def gelu_grad(x):
    return 0.5 * (1.0 + _erf(x * _SQRT1_2)) + x * np.exp(-0.5 * x * x) * _INV_SQRT_2PI


def forward(theta: np.ndarray, X: np.ndarray):
    p = unpack(theta)
    z1 = X @ p["W1"] + p["b1"]
    h1 = gelu(z1)
    z2 = h1 @ p["W2"] + p["b2"]
    h2 = gelu(z2)
    out = h2 @ p["W3"] + p["b3"]
    return out, (z1, h1, z2, h2)


def loss_and_grad(theta: np.ndarray, X: np.ndarray, target: np.ndarray):
    """Mean loss over the batch and its gradient w.r.t. theta (backprop by hand)."""
    p = unpack(theta)
    B = X.shape[0]
    out, (z1, h1, z2, h2) = forward(theta, X)
    c_raw = out[:, 2]
    c = np.clip(c_raw, C_MIN, C_MAX)
    ec = np.exp(c)
    da = out[:, 0] - target[:, 0]
    db = out[:, 1] - target[:, 1]
    sse = da * da + db * db
    loss = float(np.mean(0.5 * ec * sse - c))
    d3 = np.zeros_like(out)
    d3[:, 0] = ec * da / B
    d3[:, 1] = ec * db / B
    inside = (c_raw > C_MIN) & (c_raw < C_MAX)
    d3[:, 2] = np.where(inside, (0.5 * ec * sse - 1.0) / B, 0.0)
    g = {"W3": h2.T @ d3, "b3": d3.sum(0)}
    d2 = (d3 @ p["W3"].T) * gelu_grad(z2)
    g["W2"] = h1.T @ d2
    g["b2"] = d2.sum(0)
    d1 = (d2 @ p["W2"].T) * gelu_grad(z1)
    g["W1"] = X.T @ d1
    g["b1"] = d1.sum(0)
    return loss, pack(g)


def init_params(seed: int = 0) -> np.ndarray:
    """He-style init, zero biases (b3[2] = 0: unit precision at start)."""
    rng = np.random.default_rng(seed)
    parts = {"W1": rng.standard_normal((N_IN, N_H)) * math.sqrt(2.0 / N_IN), "b1": np.zeros(N_H),
             "W2": rng.standard_normal((N_H, N_H)) * math.sqrt(2.0 / N_H), "b2": np.zeros(N_H),
             "W3": rng.standard_normal((N_H, N_OUT)) * math.sqrt(1.0 / N_H), "b3": np.zeros(N_OUT)}
    return pack(parts)


def adam_step(theta, grad, m, v, step, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, bias-corrected), returns new (theta, m, v)."""
    m = beta1 * m + (1 - beta1) * grad
    v = beta2 * v + (1 - beta2) * grad * grad
    mh = m / (1 - beta1**step)
    vh = v / (1 - beta2**step)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v


def residual_acf(Y: np.ndarray, X: np.ndarray, nlags: int = N_IN) -> np.ndarray:
    """Network inputs: lag-1..nlags autocorrelations of each voxel's OLS residual
    r = y - X pinv(X) y (rows of Y are voxels), acf_k = sum_t r_t r_{t-k} /
    sum_t r_t^2 (0 where the residual is identically zero).  The per-voxel std
    normalisation that the device applies first cancels in the ratio."""
    Y = np.asarray(Y, dtype=np.float64)
    X = np.asarray(X, dtype=np.float64)
    R = Y - (Y @ np.linalg.pinv(X).T) @ X.T
    den = (R * R).sum(axis=1)
    out = np.zeros((Y.shape[0], nlags))
    for k in range(1, nlags + 1):
        num = (R[:, k:] * R[:, :-k]).sum(axis=1)
        out[:, k - 1] = np.where(den > 0, num / np.where(den > 0, den, 1.0), 0.0)
    return out
