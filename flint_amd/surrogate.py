"""Learned reactor step (no counterpart in the FLINT reference).

The plain batched BDF (flint_reactor_advance, constant pressure) generates (state, increment)
pairs over a fixed step dt from seeded states; an MFMA MLP (flint_amd.mlp) learns the standardised
map (log max(Y_k, ymin), T, log p, and the rate features asinh(dt dY_k/dt / yfloor),
asinh(dt dT/dt / tfloor)) -> o, the increments in units of the local rate,
o_k = dY_k / (dt |dY_k/dt| + yfloor) and o_T = dT / (dt |dT/dt| + tfloor) with dX/dt at the old state;
flint_reactor_advance_mlp applies it with the BDF as per-cell fallback and exact
element/mass/enthalpy conservation for the accepted cells.  Each prediction is judged in FP64
before it is accepted: the Newton corrections from it towards the backward-Euler and the
trapezoidal step over dt (F and J at the prediction) must stay within max_newton_dY and
max_newton_dT, half the step's accuracy bars (1e-4 in Y, 1 K); the raw element residual only
screens out implausible outputs.

    python -m flint_amd.surrogate --cells 262144 --steps 16 --epochs 60 --out flint_amd/data/surrogate_rich31.npz
"""
import argparse
import ctypes as C
import json
import os
import time

import numpy as np

from . import states as S
from .elements import element_weights
from .mlp import MLP, Scaler
from .native import check, lib, _p, _stream

YMIN = 1e-12      # floor of Y_k in the log features
YFLOOR = 1e-9     # increment scale floors: mass fraction, K
TFLOOR = 1e-3
OCLIP = 50.0      # |o| clip of the training targets
HIDDEN = (256, 256, 256)
MAX_RESIDUAL = 1e-4     # plausibility of the raw prediction (element and mass residual)
MAX_NEWTON_DY = 5e-5    # the FP64 judge: implicit-step Newton correction bounds (half the bars)
MAX_NEWTON_DT = 0.5
HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_MODEL = os.path.join(HERE, "data", "surrogate_rich31.npz")


def features(X, p, ymin=YMIN, F=None, dt=None, yfloor=YFLOOR, tfloor=TFLOOR):
    """X [ns+1, n] (Y_k, T), p [n] (torch, FP64) -> [n, ns+2]; with the right-hand side F [ns+1, n]
    at X and the step dt also the rate features asinh(dt F_k / yfloor), asinh(dt F_T / tfloor)
    -> [n, 2 ns + 3] (the device features kernel computes the same)."""
    import torch
    ns = X.shape[0] - 1
    cols = [torch.log(torch.clamp(X[:ns], min=ymin)).T, X[ns:].T, torch.log(p)[:, None]]
    if F is not None:
        fl = torch.full((ns + 1, 1), yfloor, dtype=F.dtype, device=F.device)
        fl[ns] = tfloor
        cols.append(torch.asinh(dt * F / fl).T)
    return torch.cat(cols, 1)


def rate_scale(F, dt, yfloor=YFLOOR, tfloor=TFLOOR):
    """[n, ns+1] increment units dt |dX/dt| + floor from the right-hand side F [ns+1, n]."""
    import torch
    ns = F.shape[0] - 1
    fl = torch.full((ns + 1, 1), yfloor, dtype=F.dtype, device=F.device)
    fl[ns] = tfloor
    return (dt * F.abs() + fl).T


def increments(X0, X1, F, dt, clip=OCLIP):
    """Rate-scaled increments o [n, ns+1] of X0 -> X1 (F = dX/dt at X0)."""
    return torch_clamp((X1 - X0).T / rate_scale(F, dt), clip)


def torch_clamp(x, c):
    import torch
    return torch.clamp(x, min=-c, max=c)


def generate_pairs(mech, md, n_cells, n_steps, dt, seed, rtol=1e-6, atol=1e-12, log=None, rates=False):
    """(features, increments) [N, ns+2 (2 ns + 3 with rates)], [N, ns+1] (FP64, device) along BDF
    trajectories of n_cells seeded states (half near ignition) over n_steps steps of dt."""
    import torch
    dev = torch.device("cuda:0")
    st = S.radical_states(md.species, md.W, seed, 0, n_cells, ignition_half=True)
    X = torch.from_numpy(np.vstack([S.soa(st["Y"]), st["T"][None, :]])).to(dev)
    p = torch.from_numpy(st["p"]).to(dev)
    status = torch.zeros(n_cells, dtype=torch.int32, device=dev)
    alive = torch.ones(n_cells, dtype=torch.bool, device=dev)
    F = torch.empty_like(X)
    Fs, Ds = [], []
    for s in range(n_steps):
        X0 = X.clone()
        mech.reactor_jacobian(X0, F, None, p=p, mode=1)
        t = time.time()
        mech.advance(X, p, mode=1, t_end=dt, rtol=rtol, atol=atol, status=status)
        torch.cuda.synchronize()
        alive &= status == 0
        Fs.append(features(X0[:, alive], p[alive], F=F[:, alive] if rates else None, dt=dt))
        Ds.append(increments(X0[:, alive], X[:, alive], F[:, alive], dt))
        if log:
            log(f"step {s + 1}/{n_steps}: {int(alive.sum())} cells ok, {time.time() - t:.1f} s")
    return torch.cat(Fs), torch.cat(Ds)


class Surrogate:
    """A trained network with its feature scalings and acceptance thresholds."""

    def __init__(self, net, xs, ys, el_weight, dt, ymin=YMIN, yfloor=YFLOOR, tfloor=TFLOOR, max_dT=50.0,
                 max_residual=MAX_RESIDUAL, max_T_mismatch=1.0, max_newton_dY=MAX_NEWTON_DY,
                 max_newton_dT=MAX_NEWTON_DT, meta=None):
        self.net, self.xs, self.ys = net, xs, ys
        self.el_weight = np.ascontiguousarray(el_weight, dtype=np.float64)
        self.dt, self.ymin, self.yfloor, self.tfloor = dt, ymin, yfloor, tfloor
        self.max_dT, self.max_residual, self.max_T_mismatch = max_dT, max_residual, max_T_mismatch
        self.max_newton_dY, self.max_newton_dT = max_newton_dY, max_newton_dT
        self.meta = meta or {}

    @property
    def scaling(self):
        return np.ascontiguousarray(np.concatenate([self.xs.mean, self.xs.scale, self.ys.mean, self.ys.scale]))

    def save(self, path):
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        np.savez_compressed(path, params=self.net.get_params(), dims=np.array(self.net.dims), x_mean=self.xs.mean,
                            x_scale=self.xs.scale, y_mean=self.ys.mean, y_scale=self.ys.scale,
                            el_weight=self.el_weight, dt=self.dt, ymin=self.ymin, yfloor=self.yfloor,
                            tfloor=self.tfloor, max_dT=self.max_dT, max_residual=self.max_residual,
                            max_T_mismatch=self.max_T_mismatch, meta=json.dumps(self.meta))

    @classmethod
    def load(cls, path):
        z = np.load(path, allow_pickle=False)
        net = MLP(z["dims"].tolist())
        net.set_params(z["params"])
        # the acceptance thresholds are those of this module (the judge replaced the tight residual
        # gate the file may carry)
        return cls(net, Scaler(z["x_mean"], z["x_scale"]), Scaler(z["y_mean"], z["y_scale"]), z["el_weight"],
                   float(z["dt"]), float(z["ymin"]), float(z["yfloor"]), float(z["tfloor"]), float(z["max_dT"]),
                   meta=json.loads(str(z["meta"])))

    def advance(self, mech, X, p, t_end=None, rtol=1e-6, atol=1e-12, fallback=None, check_out=None, stats=None,
                status=None, n=None, stream=None):
        """One learned step of X [ns+1, ld] (torch, device, in place); returns the fallback count.
        check_out [2, ld] (FP64, optional) receives the judge's max |delta Y_k| and |delta T|."""
        L = lib()
        if not getattr(L, "_sur_bound", False):
            vp, l = C.c_void_p, C.c_long
            d = C.c_double
            L.flint_reactor_advance_mlp.argtypes = [vp, vp, l, l, vp, vp, d, d, d, vp, vp, d, d, d, d, d, d, d, d, vp,
                                                    C.POINTER(C.c_long), vp, vp, vp, vp]
            L._sur_bound = True
        ld = X.shape[-1]
        n = ld if n is None else n
        nf = C.c_long()
        sc = self.scaling
        check(L.flint_reactor_advance_mlp(mech._h, self.net._h, n, ld, _p(X), _p(p), self.dt if t_end is None else t_end,
                                          rtol, atol, _p(sc), _p(self.el_weight), self.ymin, self.yfloor,
                                          self.tfloor, self.max_dT, self.max_residual, self.max_T_mismatch,
                                          self.max_newton_dY, self.max_newton_dT, _p(fallback), C.byref(nf),
                                          _p(check_out), _p(stats), _p(status), _stream(stream)))
        return nf.value


def train(mech, md, n_cells=1 << 16, n_steps=16, dt=1e-5, seed=7, epochs=30, batch=4096, lr=1e-3, log=print,
          rates=False):
    import torch
    t0 = time.time()
    Fx, Dy = generate_pairs(mech, md, n_cells, n_steps, dt, seed, log=log, rates=rates)
    t_gen = time.time() - t0
    xs = Scaler.fit(Fx.cpu().numpy())
    ys = Scaler.fit(Dy.cpu().numpy())
    dev = Fx.device
    Xn = ((Fx - torch.from_numpy(xs.mean).to(dev)) / torch.from_numpy(xs.scale).to(dev)).float().contiguous()
    Yn = ((Dy - torch.from_numpy(ys.mean).to(dev)) / torch.from_numpy(ys.scale).to(dev)).float().contiguous()
    ns = md.ns
    net = MLP([Fx.shape[1], *HIDDEN, ns + 1], seed=seed)
    t1 = time.time()
    hist = net.fit(Xn, Yn, epochs=epochs, batch=batch, lr=lr, seed=seed, log=log)
    torch.cuda.synchronize()
    t_train = time.time() - t1
    meta = dict(pairs=int(Fx.shape[0]), cells=n_cells, steps=n_steps, seed=seed, epochs=epochs, batch=batch, lr=lr,
                rates=rates, loss=hist, t_generate=t_gen, t_train=t_train)
    log(f"{Fx.shape[0]} pairs in {t_gen:.1f} s, trained in {t_train:.1f} s, final loss {hist[-1]:.3e}")
    return Surrogate(net, xs, ys, element_weights(md.elements), dt, meta=meta)


def main():
    import torch  # noqa: F401
    from .mechdata import mech_data
    from .native import Mechanism
    ap = argparse.ArgumentParser()
    ap.add_argument("--mech", default="TSR-Rich-31")
    ap.add_argument("--cells", type=int, default=1 << 16)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--dt", type=float, default=1e-5)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--out", default=DEFAULT_MODEL)
    ap.add_argument("--state-only", action="store_true", help="inputs (log Y, T, log p) without the rate features")
    a = ap.parse_args()
    md = mech_data(a.mech)
    mech = Mechanism.from_data(md)
    log = lambda s: print(s, flush=True)
    sur = train(mech, md, a.cells, a.steps, a.dt, a.seed, a.epochs, lr=a.lr, log=log, rates=not a.state_only)
    sur.save(a.out)
    log(f"saved {a.out}")


if __name__ == "__main__":
    main()
