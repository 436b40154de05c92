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
