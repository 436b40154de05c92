"""The FP64 MLP oracle (oracle/mlp_oracle.py) pinned by what the mathematics fixes: central finite
differences of its loss for every parameter of a small network, and bfloat16 rounding on values
whose nearest-even rounding is known."""
import numpy as np

import mlp_oracle as MO


def test_mlp_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    dims = [5, 7, 6, 3]
    npar = sum(a * b + b for a, b in zip(dims[:-1], dims[1:]))
    p = rng.normal(0, 0.5, npar)
    X = rng.normal(0, 1, (11, dims[0]))
    T = rng.normal(0, 1, (11, dims[-1]))
    loss, g = MO.loss_and_grad(p, dims, X, T)
    h = 1e-6
    fd = np.empty(npar)
    for i in range(npar):
        e = np.zeros(npar)
        e[i] = h
        fd[i] = (MO.loss_and_grad(p + e, dims, X, T)[0] - MO.loss_and_grad(p - e, dims, X, T)[0]) / (2 * h)
    assert np.max(np.abs(g - fd)) <= 1e-7 * max(1.0, np.abs(fd).max())
    # a planted mistake (GELU derivative dropped) is caught by the same check
    good = MO.dgelu
    try:
        MO.dgelu = lambda z: np.ones_like(z)
        _, gbad = MO.loss_and_grad(p, dims, X, T)
    finally:
        MO.dgelu = good
    assert np.max(np.abs(gbad - fd)) > 1e-3


def test_bf16_rounding_nearest_even():
    ulp = 2.0 ** -7
    vals = np.array([1.0, 1 + ulp / 2, 1 + 1.5 * ulp, 1 + 0.6 * ulp, -3.0, 1e-40])
    want = np.array([1.0, 1.0, 1 + 2 * ulp, 1 + ulp, -3.0])
    got = MO.bf16r(vals)
    assert np.array_equal(got[:5], want)
    assert got[5] == 2.0 ** -133  # bfloat16 subnormals are multiples of 2^-133


def test_adam_first_step_closed_form():
    """From zero moments the bias corrections give m_hat = g and v_hat = g^2, so the first Adam
    step moves every parameter by lr g / (|g| + eps): by lr in the direction of -g wherever
    |g| >> eps, and by lr |g| / eps where |g| << eps."""
    rng = np.random.default_rng(8)
    p0 = rng.normal(0, 1, 1000)
    g = rng.normal(0, 1, 1000) * 10.0 ** rng.uniform(-4, 2, 1000)
    g[:3] = [1e-12, -1e-13, 0.0]
    lr = 3e-3
    p1, m, v = MO.adam(p0, g, np.zeros_like(g), np.zeros_like(g), 1, lr)
    step = p1 - p0
    big = np.abs(g) > 1e-6
    # |step + lr sign(g)| = lr eps / (|g| + eps) <= lr eps / |g|
    assert np.all(np.abs(step[big] + lr * np.sign(g[big])) <= lr * (1e-8 / np.abs(g[big]) + 1e-12))
    assert np.allclose(step[:3], [-lr * 1e-12 / 1e-8, lr * 1e-13 / 1e-8, 0.0], rtol=1e-3, atol=1e-20)
    assert np.allclose(m, 0.1 * g, rtol=1e-15) and np.allclose(v, 0.001 * g * g, rtol=1e-15)
    # the step does not depend on the scale of the gradient (eps aside)
    p1b, _, _ = MO.adam(p0, 1e3 * g, np.zeros_like(g), np.zeros_like(g), 1, lr)
    assert np.all(np.abs(p1b[big] - p1[big]) <= lr * 1e-8 / np.abs(g[big]) + 1e-15)


def test_adam_second_step_by_hand():
    """Two steps on one parameter with gradients 1 then -2, lr 0.1, betas (0.9, 0.999), computed
    by hand: m_2 = 0.9 * 0.1 * 1 + 0.1 * (-2) = -0.11, m_hat = -0.11 / (1 - 0.81) = -11/19;
    v_2 = 0.999 * 0.001 * 1 + 0.001 * 4 = 0.004999, v_hat = 0.004999 / (1 - 0.998001) = 4999/1999;
    p_2 = p_1 - 0.1 * m_hat / (sqrt(v_hat) + 1e-8), with p_1 = p_0 - 0.1 * 1 / (1 + 1e-8)."""
    p, m, v = MO.adam(np.array([0.5]), np.array([1.0]), np.zeros(1), np.zeros(1), 1, 0.1)
    assert abs(p[0] - (0.5 - 0.1 / (1 + 1e-8))) < 1e-15
    p, m, v = MO.adam(p, np.array([-2.0]), m, v, 2, 0.1)
    assert abs(m[0] - (-0.11)) < 1e-16 and abs(v[0] - 0.004999) < 1e-17
    want = 0.5 - 0.1 / (1 + 1e-8) - 0.1 * (-11.0 / 19.0) / (np.sqrt(4999.0 / 1999.0) + 1e-8)
    assert abs(p[0] - want) < 1e-14
    assert abs(p[0] - 0.4366103) < 1e-6  # 0.5 - 0.1 + 0.0366103...
