"""Learned reactor step (flint_reactor_advance_mlp) with the trained network in
flint_amd/data/surrogate_rich31.npz.

* cells that fall back are advanced by the plain BDF kernel: bit-identical to flint_reactor_advance
  on the same cells;
* accepted cells keep element amounts, total mass and mixture enthalpy of the old state to
  round-off (the projection and the enthalpy solve);
* accepted cells stay within max |dY| <= 1e-4 and max |dT| <= 1 K of the plain FP64 BDF step, and
  each passed the FP64 judge (the backward-Euler and trapezoidal Newton corrections from the
  prediction within the thresholds);
* mid-trajectory, on 2^16 cells, the learned step is faster than the plain step (measured
  1.79x at 2^18 cells);
* flint_reactor_check_step, the judge on its own, passes the plain step and rejects a perturbed
  one without changing either state."""
import numpy as np
import pytest

from flint_amd import states as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(mechs, cuda):
    import torch
    from flint_amd.native import Mechanism
    from flint_amd.surrogate import Surrogate, DEFAULT_MODEL
    md = mechs["TSR-Rich-31"]
    mech = Mechanism.from_data(md)
    sur = Surrogate.load(DEFAULT_MODEL)
    n = 4096
    st = S.radical_states(md.species, md.W, 31337, 0, n, ignition_half=True)
    X = torch.from_numpy(np.vstack([S.soa(st["Y"]), st["T"][None, :]])).to(cuda)
    p = torch.from_numpy(st["p"]).to(cuda)
    print("[gpu] surrogate setup: plain BDF to t = 5e-5", flush=True)
    mech.advance(X, p, mode=1, t_end=5e-5)
    torch.cuda.synchronize()
    return md, mech, sur, X, p


def _mix_h(md, Y, T):
    Ti = np.floor(T).astype(int)
    d = T - Ti
    a, b = md.h[Ti - md.Th_min], md.h[Ti - md.Th_min + 1]
    return ((a + (b - a) * d[:, None]) * Y.T).sum(1)


def test_surrogate_step_conservation_and_fallback(setup, cuda):
    import torch
    md, mech, sur, X, p = setup
    ns, n = md.ns, X.shape[1]
    Xp, Xm = X.clone(), X.clone()
    sp = torch.zeros(n, dtype=torch.int32, device=cuda)
    mech.advance(Xp, p, mode=1, t_end=sur.dt, status=sp)
    fb = torch.full((n,), -1, dtype=torch.int32, device=cuda)
    sm = torch.full((n,), -1, dtype=torch.int32, device=cuda)
    print("[gpu] surrogate step", flush=True)
    chk = torch.full((2, n), -1.0, dtype=torch.float64, device=cuda)
    nf = sur.advance(mech, Xm, p, fallback=fb, check_out=chk, status=sm)
    torch.cuda.synchronize()
    X0, Xp, Xm = X.cpu().numpy(), Xp.cpu().numpy(), Xm.cpu().numpy()
    fb, sm, sp, chk = fb.cpu().numpy(), sm.cpu().numpy(), sp.cpu().numpy(), chk.cpu().numpy()
    assert set(np.unique(fb)) <= {0, 1} and nf == int(fb.sum())
    f = fb == 1
    # fallback cells: the plain BDF kernel on the same cells
    assert np.array_equal(Xm[:, f], Xp[:, f]) and np.array_equal(sm[f], sp[f])
    a = ~f
    assert a.sum() > 0.5 * n
    assert np.all(sm[a] == 0)
    assert np.all(chk[0, a] <= sur.max_newton_dY) and np.all(chk[1, a] <= sur.max_newton_dT)
    comp = np.asarray(md.composition, dtype=np.float64)
    W = np.asarray(md.W)
    b0, b1 = comp @ (X0[:ns, a] / W[:, None]), comp @ (Xm[:ns, a] / W[:, None])
    assert np.max(np.abs(b1 - b0) / np.abs(b0).max(0)) < 1e-11
    assert np.max(np.abs(Xm[:ns, a].sum(0) - X0[:ns, a].sum(0))) < 1e-12
    h0, h1 = _mix_h(md, X0[:ns, a], X0[ns, a]), _mix_h(md, Xm[:ns, a], Xm[ns, a])
    assert np.max(np.abs(h1 - h0) / np.maximum(np.abs(h0), 1e5)) < 1e-9
    # accuracy against the plain FP64 step: the learned step's bars, max |dY| <= 1e-4 and
    # max |dT| <= 1 K over the accepted cells (the judge's thresholds, half the bars, keep them;
    # at 2^18 cells the worst accepted cells measured 5.0e-5 and 0.49 K)
    assert np.max(np.abs(Xm[ns, a] - Xp[ns, a])) <= 1.0
    assert np.max(np.abs(Xm[:ns, a] - Xp[:ns, a])) <= 1e-4


def test_surrogate_with_zero_tolerance_is_the_plain_step(setup, cuda):
    import torch
    md, mech, sur, X, p = setup
    Xp, Xm = X.clone(), X.clone()
    mech.advance(Xp, p, mode=1, t_end=sur.dt)
    keep = sur.max_residual
    try:
        sur.max_residual = 0.0
        nf = sur.advance(mech, Xm, p)
    finally:
        sur.max_residual = keep
    torch.cuda.synchronize()
    assert nf == X.shape[1]
    assert torch.equal(Xm, Xp)


def test_surrogate_rejects_mismatched_network(setup, cuda):
    from flint_amd.mlp import MLP
    from flint_amd.native import FlintError
    from flint_amd.surrogate import Surrogate
    md, mech, sur, X, p = setup
    bad = Surrogate(MLP([10, 16, 5]), sur.xs, sur.ys, sur.el_weight, sur.dt)
    with pytest.raises(FlintError):
        bad.advance(mech, X.clone(), p)


def test_surrogate_gain_mid_trajectory(mechs, cuda):
    """2^16 cells at t = 5e-5 s: the learned step (network, judge and BDF fallback, timed as one)
    against the plain BDF step over the same dt, with every accepted cell within the bars."""
    import time
    import torch
    from flint_amd.native import Mechanism
    from flint_amd.surrogate import Surrogate, DEFAULT_MODEL
    md = mechs["TSR-Rich-31"]
    mech = Mechanism.from_data(md)
    sur = Surrogate.load(DEFAULT_MODEL)
    ns, n = md.ns, 1 << 16
    st = S.radical_states(md.species, md.W, 4048, 0, n, ignition_half=True)
    X = torch.from_numpy(np.vstack([S.soa(st["Y"]), st["T"][None, :]])).to(cuda)
    p = torch.from_numpy(st["p"]).to(cuda)
    print(f"[gpu] surrogate gain n={n}: plain BDF to t = 5e-5", flush=True)
    mech.advance(X, p, mode=1, t_end=5e-5)
    w = X[:, :1024].clone()
    sur.advance(mech, w, p[:1024].clone())  # first-call set-up outside the timing

    def timed(f):
        torch.cuda.synchronize()
        t = time.perf_counter()
        f()
        torch.cuda.synchronize()
        return time.perf_counter() - t

    Xp, Xm = X.clone(), X.clone()
    fb = torch.zeros(n, dtype=torch.int32, device=cuda)
    t_plain = timed(lambda: mech.advance(Xp, p, mode=1, t_end=sur.dt))
    t_mlp = timed(lambda: sur.advance(mech, Xm, p, fallback=fb))
    acc = fb == 0
    dY = float((Xm[:ns, acc] - Xp[:ns, acc]).abs().max())
    dT = float((Xm[ns, acc] - Xp[ns, acc]).abs().max())
    print(f"[gpu] plain {t_plain * 1e3:.1f} ms, learned {t_mlp * 1e3:.1f} ms ({t_plain / t_mlp:.2f}x), "
          f"accepted {float(acc.float().mean()):.3f}, max |dY| {dY:.2e}, max |dT| {dT:.3f} K", flush=True)
    assert dY <= 1e-4 and dT <= 1.0
    assert t_plain / t_mlp >= 1.25


def test_check_step_judges_candidate_states(setup, cuda):
    """flint_reactor_check_step, the learned step's FP64 judge for any predictor: mid-trajectory
    the plain BDF's own step passes for most cells, the same step with one species moved by 1e-3
    fails for every cell, and neither state is changed."""
    import torch
    md, mech, sur, X, p = setup
    ns, n = md.ns, X.shape[1]
    Xt = X.clone()
    mech.advance(Xt, p, mode=1, t_end=sur.dt)
    X0, Xt0 = X.clone(), Xt.clone()
    flag = torch.full((n,), -1, dtype=torch.int32, device=cuda)
    chk = torch.full((2, n), -1.0, dtype=torch.float64, device=cuda)
    print("[gpu] check_step on the plain step", flush=True)
    mech.check_step(X, Xt, p, sur.dt, flag, 5e-5, 0.5, check_out=chk)
    torch.cuda.synchronize()
    f, c = flag.cpu().numpy(), chk.cpu().numpy()
    assert set(np.unique(f)) <= {0, 1}
    assert (f == 0).mean() > 0.8
    assert np.all((c[0] <= 5e-5) & (c[1] <= 0.5) == (f == 0))
    assert torch.equal(X, X0) and torch.equal(Xt, Xt0)
    # the corrections themselves, against numpy's solve with F and J of the same reactor model
    # (flint_reactor_jacobian) on 256 cells
    m = 256
    nz = ns + 1
    F0 = torch.empty((nz, n), dtype=torch.float64, device=cuda)
    Fp = torch.empty_like(F0)
    Jp = torch.empty((nz * nz, n), dtype=torch.float64, device=cuda)
    mech.reactor_jacobian(X, F0, None, p=p, mode=1)
    mech.reactor_jacobian(Xt, Fp, Jp, p=p, mode=1)
    torch.cuda.synchronize()
    x0, xt = X0.cpu().numpy()[:, :m], Xt0.cpu().numpy()[:, :m]
    f0, fp, jp = F0.cpu().numpy()[:, :m], Fp.cpu().numpy()[:, :m], Jp.cpu().numpy()[:, :m]
    dt = sur.dt
    for i in range(m):
        Ji = jp[:, i].reshape(nz, nz)
        dbe = -np.linalg.solve(np.eye(nz) - dt * Ji, xt[:, i] - x0[:, i] - dt * fp[:, i])
        dtr = -np.linalg.solve(np.eye(nz) - 0.5 * dt * Ji, xt[:, i] - x0[:, i] - 0.5 * dt * (f0[:, i] + fp[:, i]))
        ey = max(np.abs(dbe[:ns]).max(), np.abs(dtr[:ns]).max())
        et = max(abs(dbe[ns]), abs(dtr[ns]))
        assert abs(c[0, i] - ey) <= 1e-9 * ey + 1e-16
        assert abs(c[1, i] - et) <= 1e-9 * et + 1e-12
    bad = Xt.clone()
    bad[3] += 1e-3
    mech.check_step(X, bad, p, sur.dt, flag, 5e-5, 0.5, check_out=chk)
    torch.cuda.synchronize()
    assert int(flag.sum()) == n
    assert float(chk[0].min()) > 5e-5

