"""FP64 restatement of the library's MLP (flint_amd/csrc/mlp.hip) -- TEST INFRASTRUCTURE ONLY.

Forward pass, mean-squared-error loss and its gradient by backpropagation, for the flat parameter
layout of flint_mlp (per layer W [d_l][d_{l+1}] row-major, then b).  With ``bf16=True`` it rounds
the input, the weights, the hidden activations and the back-propagated gradients to bfloat16
(round to nearest even), the points where the GPU kernels store BF16, so that the remaining
difference to the GPU is FP32 accumulation order."""
import numpy as np
from scipy.special import erf


def bf16r(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def dgelu(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def unpack(params, dims):
    p = np.asarray(params, dtype=np.float64)
    out, o = [], 0
    for a, b in zip(dims[:-1], dims[1:]):
        W = p[o:o + a * b].reshape(a, b)
        o += a * b
        bb = p[o:o + b]
        o += b
        out.append((W, bb))
    assert o == p.size
    return out


def forward(params, dims, X, bf16=False):
    """Returns (output, [inputs of each layer], [pre-activations])."""
    r = bf16r if bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    layers = unpack(params, dims)
    a = r(X)
    acts, zs = [a], []
    for l, (W, b) in enumerate(layers):
        z = a @ r(W) + b
        zs.append(z)
        if l + 1 < len(layers):
            a = r(gelu(z))
            acts.append(a)
        else:
            a = z
    return a, acts, zs


def loss_and_grad(params, dims, X, T, bf16=False):
    r = bf16r if bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    out, acts, zs = forward(params, dims, X, bf16)
    n, d = out.shape
    loss = np.mean((out - T) ** 2)
    dz = 2.0 * (out - T) / (n * d)
    layers = unpack(params, dims)
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        dzr = r(dz)
        gW = acts[l].T @ dzr
        gb = dz.sum(0)
        grads[l] = (gW, gb)
        if l > 0:
            dz = (dzr @ r(W).T) * dgelu(zs[l - 1])
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, flat


def adam(params, grad, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    p = params - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return p, m, v
