"""MFMA MLP kernels (flint_mlp_*) against the FP64 oracle (oracle/mlp_oracle.py, pinned by finite
differences).  The oracle rounds to bfloat16 where the kernels store BF16, so the remaining
difference is FP32 accumulation order (and an occasional activation rounded the other way)."""
import numpy as np
import pytest

import mlp_oracle as MO

pytestmark = pytest.mark.gpu
DIMS = [33, 256, 256, 256, 32]


@pytest.fixture(scope="module")
def net(cuda):
    from flint_amd.mlp import MLP
    return MLP(DIMS, seed=5)


def _data(n, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(0, 1, (n, DIMS[0])).astype(np.float32), rng.normal(0, 1, (n, DIMS[-1])).astype(np.float32)


def test_mlp_forward_matches_oracle(net, cuda):
    import torch
    X, _ = _data(777, 1)
    print("[gpu] mlp forward", flush=True)
    Y = net.forward(torch.from_numpy(X).to(cuda)).cpu().numpy()
    ref, _, _ = MO.forward(net.get_params(), DIMS, X, bf16=True)
    assert np.linalg.norm(Y - ref) / np.linalg.norm(ref) < 2e-3
    assert np.max(np.abs(Y - ref)) < 1e-2 * np.abs(ref).max()


def test_mlp_gradients_match_oracle(net, cuda):
    import torch
    X, T = _data(512, 2)
    print("[gpu] mlp gradients", flush=True)
    g, loss = net.gradients(torch.from_numpy(X).to(cuda), torch.from_numpy(T).to(cuda))
    g = g.cpu().numpy()
    lref, gref = MO.loss_and_grad(net.get_params(), DIMS, X, T, bf16=True)
    assert abs(loss.item() - lref) < 1e-3 * lref
    o = 0
    for a, b in zip(DIMS[:-1], DIMS[1:]):
        for sz in (a * b, b):
            gl, rl = g[o:o + sz], gref[o:o + sz]
            assert np.linalg.norm(gl - rl) <= 2e-2 * np.linalg.norm(rl), (o, sz)
            o += sz


def test_mlp_adam_step_matches_numpy(cuda):
    import torch
    from flint_amd.mlp import MLP
    m = MLP([8, 32, 4], seed=9)
    X, T = _data(256, 3)
    X, T = X[:, :8].copy(), T[:, :4].copy()
    Xd, Td = torch.from_numpy(X).to(cuda), torch.from_numpy(T).to(cuda)
    p0 = m.get_params().astype(np.float64)
    g, _ = m.gradients(Xd, Td)
    m.train_step(Xd, Td, 1e-2)
    torch.cuda.synchronize()
    p1 = m.get_params()
    ref, _, _ = MO.adam(p0, g.cpu().numpy().astype(np.float64), np.zeros_like(p0), np.zeros_like(p0), 1, 1e-2)
    assert np.max(np.abs(p1 - ref)) < 1e-6


def test_mlp_training_reduces_loss(cuda):
    import torch
    from flint_amd.mlp import MLP
    rng = np.random.default_rng(4)
    X = rng.uniform(-2, 2, (16384, 6)).astype(np.float32)
    T = np.stack([np.sin(X[:, 0]) * X[:, 1], np.tanh(X[:, 2] - X[:, 3]), X[:, 4] * X[:, 5]], 1).astype(np.float32)
    m = MLP([6, 64, 64, 3], seed=1)
    Xd, Td = torch.from_numpy(X).to(cuda), torch.from_numpy(T).to(cuda)
    print("[gpu] mlp training", flush=True)
    hist = m.fit(Xd, Td, epochs=30, batch=512, lr=3e-3)
    assert hist[-1] < 0.05 * hist[0]
