"""Pin the FP64 MLP oracle: hand-written backprop vs central finite
differences, GELU against its defining integral, Adam against its closed form
on a quadratic."""

import math

import numpy as np
from scipy import integrate, stats

from oracle import mlp_oracle as M


def test_gelu_matches_definition():
    for x in (-3.0, -0.7, 0.0, 0.4, 2.5):
        ref = x * stats.norm.cdf(x)
        assert abs(M.gelu(np.array(x)) - ref) < 1e-12
        d = integrate.quad(lambda t: stats.norm.pdf(t), -np.inf, x)[0] + x * stats.norm.pdf(x)
        assert abs(M.gelu_grad(np.array(x)) - d) < 1e-9


def test_backprop_matches_finite_differences():
    rng = np.random.default_rng(0)
    theta = M.init_params(1) + 0.05 * rng.standard_normal(M.N_PARAMS)
    X = rng.uniform(-0.5, 0.9, (17, M.N_IN))
    tgt = rng.uniform(-0.8, 0.8, (17, 2))
    _, g = M.loss_and_grad(theta, X, tgt)
    idx = rng.choice(M.N_PARAMS, 200, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros(M.N_PARAMS)
        e[i] = h
        fd = (M.loss_and_grad(theta + e, X, tgt)[0] - M.loss_and_grad(theta - e, X, tgt)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(g[i]), (i, fd, g[i])


def test_loss_minimiser_of_confidence():
    # for fixed errors, dL/dc = 0 at e^c = 2 / sse
    theta = M.init_params(2)
    X = np.zeros((1, M.N_IN))
    out, _ = M.forward(theta, X)
    tgt = np.array([[out[0, 0] - 0.1, out[0, 1] + 0.2]])
    sse = 0.1**2 + 0.2**2
    p = M.unpack(theta.copy())
    p["b3"][2] = math.log(2.0 / sse) - (out[0, 2] - p["b3"][2])
    _, g = M.loss_and_grad(M.pack(p), X, tgt)
    assert abs(g[-1]) < 1e-12


def test_adam_closed_form_first_step():
    theta = np.array([1.0, -2.0])
    g = np.array([0.5, -3.0])
    th, m, v = M.adam_step(theta, g, np.zeros(2), np.zeros(2), 1, lr=0.1)
    # first bias-corrected step moves by lr * sign(g) (up to eps)
    np.testing.assert_allclose(th, theta - 0.1 * np.sign(g), rtol=0, atol=1e-7)


def test_residual_acf_closed_forms():
    # intercept-only design, y_t = (-1)^t with T even: the residual is y itself
    # and acf_k = (-1)^k (T - k) / T
    T = 40
    X = np.ones((T, 1))
    y = (-1.0) ** np.arange(T)
    got = M.residual_acf(y[None, :], X, 10)[0]
    want = np.array([(-1.0) ** k * (T - k) / T for k in range(1, 11)])
    assert np.abs(got - want).max() < 1e-12
    t = np.arange(T, dtype=np.float64)
    # a step series orthogonal to the intercept: r_t = +1 then -1 (T/2 each),
    # acf_k = (T - 3k) / T for k <= T/2
    z = np.where(t < T // 2, 1.0, -1.0)
    want = np.array([(T - 3 * k) / T for k in range(1, 11)])
    assert np.abs(M.residual_acf(z[None, :], X, 10)[0] - want).max() < 1e-12
