"""The MFMA multilayer perceptron of libflint_amd (flint_mlp_* in include/flint_amd.h).

Device arrays are torch tensors (FP32, row-major [n][features]); training and inference run in the
library's HIP kernels.  Feature scaling lives here: a network maps standardised inputs to
standardised outputs, and ``Scaler`` keeps the per-feature mean and spread."""
import ctypes as C

import numpy as np

from .native import check, lib, _p, _stream


def _bind():
    L = lib()
    if getattr(L, "_mlp_bound", False):
        return L
    vp, l, i = C.c_void_p, C.c_long, C.c_int
    L.flint_mlp_create.argtypes = [i, C.POINTER(C.c_int), C.c_ulonglong, C.POINTER(vp)]
    L.flint_mlp_destroy.argtypes = [vp]
    L.flint_mlp_num_params.argtypes = [vp, C.POINTER(C.c_long)]
    L.flint_mlp_get_params.argtypes = [vp, vp]
    L.flint_mlp_set_params.argtypes = [vp, vp]
    L.flint_mlp_forward.argtypes = [vp, l, vp, vp, vp]
    L.flint_mlp_gradients.argtypes = [vp, l, vp, vp, vp, vp, vp]
    L.flint_mlp_train_step.argtypes = [vp, l, vp, vp, C.c_float, vp, vp]
    L._mlp_bound = True
    return L


class Scaler:
    """Per-feature standardisation y = (x - mean) / scale."""

    def __init__(self, mean, scale):
        self.mean = np.asarray(mean, dtype=np.float64)
        self.scale = np.asarray(scale, dtype=np.float64)

    @classmethod
    def fit(cls, x, floor=1e-12):
        x = np.asarray(x, dtype=np.float64)
        return cls(x.mean(0), np.maximum(x.std(0), floor))

    def to_dict(self):
        return {"mean": self.mean.tolist(), "scale": self.scale.tolist()}

    @classmethod
    def from_dict(cls, d):
        return cls(d["mean"], d["scale"])


class MLP:
    def __init__(self, dims, seed=0):
        L = _bind()
        self.dims = [int(d) for d in dims]
        arr = (C.c_int * len(self.dims))(*self.dims)
        h = C.c_void_p()
        check(L.flint_mlp_create(len(self.dims) - 1, arr, C.c_ulonglong(seed), C.byref(h)))
        self._h = h
        n = C.c_long()
        check(L.flint_mlp_num_params(h, C.byref(n)))
        self.nparams = n.value

    def close(self):
        if self._h:
            lib().flint_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        w = np.empty(self.nparams, dtype=np.float32)
        check(lib().flint_mlp_get_params(self._h, _p(w)))
        return w

    def set_params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == self.nparams
        check(lib().flint_mlp_set_params(self._h, _p(w)))

    def forward(self, X, Y=None, stream=None):
        import torch
        n = X.shape[0]
        if Y is None:
            Y = torch.empty((n, self.dims[-1]), dtype=torch.float32, device=X.device)
        check(lib().flint_mlp_forward(self._h, n, _p(X), _p(Y), _stream(stream)))
        return Y

    def gradients(self, X, T, stream=None):
        import torch
        g = torch.empty(self.nparams, dtype=torch.float32, device=X.device)
        loss = torch.empty(1, dtype=torch.float32, device=X.device)
        check(lib().flint_mlp_gradients(self._h, X.shape[0], _p(X), _p(T), _p(g), _p(loss), _stream(stream)))
        return g, loss

    def train_step(self, X, T, lr, loss=None, stream=None):
        check(lib().flint_mlp_train_step(self._h, X.shape[0], _p(X), _p(T), C.c_float(lr), _p(loss),
                                         _stream(stream)))

    def fit(self, X, T, epochs=10, batch=4096, lr=1e-3, lr_final=None, seed=0, log=None):
        """Adam over shuffled mini-batches of device tensors X [n][d0], T [n][dL]; cosine decay of
        the learning rate to lr_final.  Returns the mean loss of every epoch."""
        import torch
        n = X.shape[0]
        gen = torch.Generator(device=X.device)
        gen.manual_seed(seed)
        steps = epochs * (n // batch)
        lr_final = lr * 0.05 if lr_final is None else lr_final
        loss = torch.zeros(1, dtype=torch.float32, device=X.device)
        history, step = [], 0
        for ep in range(epochs):
            perm = torch.randperm(n, device=X.device, generator=gen)
            acc = torch.zeros(1, dtype=torch.float32, device=X.device)
            nb = n // batch
            for b in range(nb):
                idx = perm[b * batch:(b + 1) * batch]
                xb, tb = X[idx].contiguous(), T[idx].contiguous()
                lr_t = lr_final + 0.5 * (lr - lr_final) * (1.0 + np.cos(np.pi * step / max(steps, 1)))
                self.train_step(xb, tb, lr_t, loss)
                acc += loss
                step += 1
            history.append(float(acc.item()) / max(nb, 1))
            if log:
                log(f"epoch {ep + 1}/{epochs} loss {history[-1]:.4e}")
        return history
