"""Learned warm start of the batched equilibrium solver (no counterpart in the FLINT reference).

An MFMA MLP (flint_amd.mlp, 26 -> 128 -> 128 -> 25 for TSR-GP-24) maps the unburnt state (mass
fractions Y0_k, mixture enthalpy h0 = sum_k Y0_k h_k(T0), log p) to the equilibrium (T, ln n_k);
flint_cea_solve starts its Newton iteration there.  States that do not converge from the predicted
guess are solved again from the reference's default guess, so the warm-started batch returns
exactly the states the default path returns a solution for.  `solve` runs all of it in one native
call (flint_cea_solve_warm); `guess` is the same network evaluated through torch, for inspection.

    python -m flint_amd.warmstart --states 262144 --out flint_amd/data/cea_warmstart_gp24.npz
"""
import argparse
import json
import os
import time

import numpy as np

from . import states as S
from .mlp import MLP, Scaler

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_MODEL = os.path.join(HERE, "data", "cea_warmstart_gp24.npz")
NMIN = 1e-30
HIDDEN = (128, 128)


def _tab(md, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(md.h)).to(dev)


def features(md, Y0, T0, p_bar, htab=None):
    """Y0 [ns, n], T0 [n], p [n] (torch FP64, device) -> [n, ns+2] (Y0, h0, log p)."""
    import torch
    htab = _tab(md, Y0.device) if htab is None else htab
    Ti = torch.floor(T0).long()
    d = (T0 - Ti)[:, None]
    a, b = htab[Ti - md.Th_min], htab[Ti - md.Th_min + 1]
    h0 = ((a + (b - a) * d) * Y0.T).sum(1)
    return torch.cat([Y0.T, h0[:, None], torch.log(p_bar)[:, None]], 1)


def targets(md, T, Y):
    """equilibrium T [n], Y [ns, n] -> [n, ns+1] (T, ln n_k)."""
    import torch
    W = torch.from_numpy(np.asarray(md.W)).to(Y.device)
    return torch.cat([T[:, None], torch.log(torch.clamp(Y / W[:, None], min=NMIN)).T], 1)


class WarmStart:
    def __init__(self, net, xs, ys, meta=None):
        self.net, self.xs, self.ys, self.meta = net, xs, ys, meta or {}

    def save(self, path):
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        np.savez_compressed(path, params=self.net.get_params(), dims=np.array(self.net.dims), x_mean=self.xs.mean,
                            x_scale=self.xs.scale, y_mean=self.ys.mean, y_scale=self.ys.scale,
                            meta=json.dumps(self.meta))

    @classmethod
    def load(cls, path):
        z = np.load(path, allow_pickle=False)
        net = MLP(z["dims"].tolist())
        net.set_params(z["params"])
        return cls(net, Scaler(z["x_mean"], z["x_scale"]), Scaler(z["y_mean"], z["y_scale"]),
                   json.loads(str(z["meta"])))

    def guess(self, md, Y0, T0, p_bar):
        """[ns+1, n] (T, ln n_k) predicted for the states (torch, device)."""
        import torch
        dev = Y0.device
        F = features(md, Y0, T0, p_bar)
        Fn = ((F - torch.from_numpy(self.xs.mean).to(dev)) / torch.from_numpy(self.xs.scale).to(dev)).float()
        O = self.net.forward(Fn.contiguous()).double()
        O = O * torch.from_numpy(self.ys.scale).to(dev) + torch.from_numpy(self.ys.mean).to(dev)
        return O.T.contiguous()

    def solve(self, mech, md, Y0, T0, p_bar, T_out, Y_out, iters=None, status=None, stream=None):
        """HP equilibrium of the states warm-started by the network, in one native call
        (flint_cea_solve_warm: features, MFMA network, warm solve, and a second pass from the
        default guess for the states that did not converge from theirs)."""
        mech.equilibrium_warm(self.net, (self.xs.mean, self.xs.scale, self.ys.mean, self.ys.scale), T0, Y0, T_out,
                              Y_out, p_bar, iters=iters, status=status, stream=stream)


def train(mech, md, n=1 << 18, seed=41, epochs=60, batch=4096, lr=2e-3, log=print):
    import torch
    dev = torch.device("cuda:0")
    st = S.cea_states(md.species, md.W, seed, 0, n)
    Y0 = torch.from_numpy(S.soa(st["Y"])).to(dev)
    T0 = torch.from_numpy(st["T0"]).to(dev)
    p = torch.from_numpy(st["p_bar"]).to(dev)
    T, Y = torch.empty_like(T0), torch.empty_like(Y0)
    status = torch.empty(n, dtype=torch.int32, device=dev)
    t0 = time.time()
    mech.equilibrium(T0, Y0, T, Y, p=p, hp=True, status=status)
    torch.cuda.synchronize()
    ok = status == 0
    F = features(md, Y0, T0, p)[ok]
    Tg = targets(md, T, Y)[ok]
    xs, ys = Scaler.fit(F.cpu().numpy()), Scaler.fit(Tg.cpu().numpy())
    Xn = ((F - torch.from_numpy(xs.mean).to(dev)) / torch.from_numpy(xs.scale).to(dev)).float().contiguous()
    Yn = ((Tg - torch.from_numpy(ys.mean).to(dev)) / torch.from_numpy(ys.scale).to(dev)).float().contiguous()
    net = MLP([md.ns + 2, *HIDDEN, md.ns + 1], seed=seed)
    hist = net.fit(Xn, Yn, epochs=epochs, batch=batch, lr=lr, seed=seed, log=log)
    meta = dict(states=int(ok.sum()), seed=seed, epochs=epochs, batch=batch, lr=lr, loss=hist, t=time.time() - t0)
    log(f"{int(ok.sum())} equilibria, trained in {time.time() - t0:.1f} s, final loss {hist[-1]:.3e}")
    return WarmStart(net, xs, ys, meta)


def main():
    from .mechdata import mech_data
    from .native import Mechanism
    ap = argparse.ArgumentParser()
    ap.add_argument("--states", type=int, default=1 << 18)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--seed", type=int, default=41)
    ap.add_argument("--out", default=DEFAULT_MODEL)
    a = ap.parse_args()
    md = mech_data("TSR-GP-24")
    mech = Mechanism.from_data(md)
    ws = train(mech, md, a.states, a.seed, a.epochs, log=lambda s: print(s, flush=True))
    ws.save(a.out)
    print("saved", a.out, flush=True)


if __name__ == "__main__":
    main()
