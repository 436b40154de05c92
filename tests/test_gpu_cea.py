"""Batched equilibrium kernel (flint_cea_solve) against the CEA oracle, which equals the reference
solver bit for bit (tests/test_cea_oracle.py).  Pass bar: |dT| <= 1e-6 K and composition within
rel. 1e-10 (of the largest mass fraction) per state, the same iteration count, for HP and UV."""
import os

import numpy as np
import pytest

from flint_amd import states as S

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cea_TSR-GP-24.npz")


@pytest.fixture(scope="module")
def gp24(mechs, cuda):
    from flint_amd.native import Mechanism
    md = mechs["TSR-GP-24"]
    return md, Mechanism.from_data(md)


def _solve(mech, T0, Y, p, hp, cuda, guess=None):
    import torch
    ns, n = Y.shape
    ld = n + 3
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda)
    Td = torch.zeros(ld, dtype=torch.float64, device=cuda)
    Td[:n] = dev(T0)
    Yd = torch.zeros((ns, ld), dtype=torch.float64, device=cuda)
    Yd[:, :n] = dev(Y)
    pd = torch.ones(ld, dtype=torch.float64, device=cuda)
    pd[:n] = dev(p)
    gd = None
    if guess is not None:
        gd = torch.zeros((ns + 1, ld), dtype=torch.float64, device=cuda)
        gd[:, :n] = dev(guess)
    To = torch.full((ld,), 7.0, dtype=torch.float64, device=cuda)
    Yo = torch.full((ns, ld), 7.0, dtype=torch.float64, device=cuda)
    it = torch.full((ld,), -1, dtype=torch.int32, device=cuda)
    st = torch.full((ld,), -1, dtype=torch.int32, device=cuda)
    mech.equilibrium(Td, Yd, To, Yo, p=pd, hp=hp, guess=gd, iters=it, status=st, n=n)
    torch.cuda.synchronize()
    To, Yo, it, st = To.cpu().numpy(), Yo.cpu().numpy(), it.cpu().numpy(), st.cpu().numpy()
    assert np.all(To[n:] == 7.0) and np.all(Yo[:, n:] == 7.0) and np.all(st[n:] == -1)
    return To[:n], Yo[:, :n], it[:n], st[:n]


def _close(T, Y, To, Yo):
    assert np.max(np.abs(T - To)) <= 1e-6
    assert np.max(np.abs(Y - Yo) / np.abs(Yo).max(0)) <= 1e-10


@pytest.mark.parametrize("mode", ["hp", "uv"])
def test_cea_gpu_matches_reference_outputs(gp24, oracle_mod, cuda, mode):
    md, mech = gp24
    z = np.load(GOLD)
    st = S.cea_states(md.species, md.W, int(z["seed"]), int(z["start"]), int(z["count"]))
    print(f"[gpu] cea golden {mode}", flush=True)
    T, Y, it, status = _solve(mech, st["T0"], S.soa(st["Y"]), st["p_bar"], mode == "hp", cuda)
    assert np.all(status == 0)
    _close(T, Y, z[f"{mode}_T"], z[f"{mode}_Y"])


@pytest.mark.parametrize("mode", ["hp", "uv"])
def test_cea_gpu_matches_oracle(gp24, oracle_mod, cuda, mode):
    md, mech = gp24
    n = 8192
    st = S.cea_states(md.species, md.W, 21, 0, n)
    Y0 = S.soa(st["Y"])
    print(f"[gpu] cea {mode} n={n}", flush=True)
    T, Y, it, status = _solve(mech, st["T0"], Y0, st["p_bar"], mode == "hp", cuda)
    To, Yo, ito, sto = oracle_mod.Oracle(md).cea(st["p_bar"], st["T0"], Y0, hp=(mode == "hp"))
    assert np.array_equal(status, sto) and np.all(status == 0)
    _close(T, Y, To, Yo)
    assert np.mean(it == ito) > 0.99


def test_cea_gpu_known_answer(gp24, cuda):
    md, mech = gp24
    Y = np.full((md.ns, 1), 1e-20)
    Y[md.species.index("O2")] = 0.8
    Y[md.species.index("CH4")] = 0.2
    T, Yo, it, st = _solve(mech, np.array([1000.0]), Y, np.array([10.0]), False, cuda)
    assert st[0] == 0 and abs(T[0] - 3616.638618717671) / 3616.638618717671 < 0.01
    assert abs(Yo[3, 0] - 1.00013716e-01) / 1.00013716e-01 < 0.01


def test_cea_gpu_warm_start(gp24, oracle_mod, cuda):
    md, mech = gp24
    n = 1024
    st = S.cea_states(md.species, md.W, 22, 0, n)
    Y0 = S.soa(st["Y"])
    To, Yo, ito, sto = oracle_mod.Oracle(md).cea(st["p_bar"], st["T0"], Y0, hp=True)
    guess = np.vstack([To[None, :], np.log(np.maximum(Yo / np.asarray(md.W)[:, None], 1e-300))])
    print("[gpu] cea warm start", flush=True)
    T, Y, it, status = _solve(mech, st["T0"], Y0, st["p_bar"], True, cuda, guess=guess)
    Tg, Yg, itg, stg = oracle_mod.Oracle(md).cea(st["p_bar"], st["T0"], Y0, hp=True, guess=guess)
    assert np.all(status == 0)
    _close(T, Y, Tg, Yg)
    assert it.mean() < 0.5 * ito.mean()


def test_cea_gpu_rejects_bad_arguments(gp24, cuda):
    import torch
    from flint_amd.native import FlintError
    md, mech = gp24
    T = torch.zeros(4, dtype=torch.float64, device=cuda)
    Y = torch.zeros((md.ns, 4), dtype=torch.float64, device=cuda)
    with pytest.raises(FlintError):
        mech.equilibrium(T, Y, T.clone(), Y.clone(), p=None, hp=True)  # HP needs p
    with pytest.raises(FlintError):
        mech.equilibrium(T, Y, T.clone(), Y.clone(), p=T, hp=2)


def test_cea_learned_warm_start(gp24, cuda):
    """The trained warm-start network (flint_amd/data/cea_warmstart_gp24.npz) reaches the same
    equilibria as the default guess, within the solver's own convergence tolerances, in fewer
    Newton iterations."""
    import torch
    from flint_amd.warmstart import WarmStart, DEFAULT_MODEL
    md, mech = gp24
    ws = WarmStart.load(DEFAULT_MODEL)
    n = 4096
    st = S.cea_states(md.species, md.W, 23, 0, n)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda)
    Y0, T0, p = dev(S.soa(st["Y"])), dev(st["T0"]), dev(st["p_bar"])
    T1, Y1, T2, Y2 = torch.empty_like(T0), torch.empty_like(Y0), torch.empty_like(T0), torch.empty_like(Y0)
    it1, s1 = torch.empty(n, dtype=torch.int32, device=cuda), torch.empty(n, dtype=torch.int32, device=cuda)
    it2, s2 = torch.empty_like(it1), torch.empty_like(s1)
    print("[gpu] cea learned warm start", flush=True)
    mech.equilibrium(T0, Y0, T1, Y1, p=p, hp=True, iters=it1, status=s1)
    ws.solve(mech, md, Y0, T0, p, T2, Y2, iters=it2, status=s2)
    torch.cuda.synchronize()
    assert int((s1 != 0).sum()) == 0 and int((s2 != 0).sum()) == 0
    dT = (T2 - T1).abs().cpu().numpy()
    dY = ((Y2 - Y1).abs() / Y1.abs().max(0).values).max(0).values.cpu().numpy()
    # with the reference's convergence tests (corrections <= 0.5e-5, |dln T| <= 1e-4) the converged
    # state depends on the start: a regression bound near the measured spread (3e-4 K), not the
    # workload's 1e-6 K bar, which the tightened mode below meets
    assert dT.max() < 1e-3 and np.quantile(dT, 0.99) < 1e-4
    assert dY.max() < 1e-5
    assert it2.float().mean().item() < 0.5 * it1.float().mean().item()


def test_cea_tight_mode_meets_the_bar(gp24, oracle_mod, cuda):
    """Tightened convergence tests (flint_cea_set_tolerance(1e-10)): the default-guess solve and
    the learned warm start both land within 1e-6 K and 1e-10 (relative to the largest mass
    fraction) of the oracle's tightened solve from the default guess, and the warm start still
    needs fewer Newton iterations."""
    import torch
    from flint_amd.warmstart import WarmStart, DEFAULT_MODEL
    md, mech = gp24
    ws = WarmStart.load(DEFAULT_MODEL)
    n = 4096
    st = S.cea_states(md.species, md.W, 26, 0, n)
    Yh = S.soa(st["Y"])
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda)
    Y0, T0, p = dev(Yh), dev(st["T0"]), dev(st["p_bar"])
    out = []
    print("[gpu] cea tightened mode", flush=True)
    try:
        mech.cea_tolerance(1e-10)
        for warm in (False, True):
            T, Y = torch.empty_like(T0), torch.empty_like(Y0)
            it, s = torch.empty(n, dtype=torch.int32, device=cuda), torch.empty(n, dtype=torch.int32, device=cuda)
            if warm:
                ws.solve(mech, md, Y0, T0, p, T, Y, iters=it, status=s)
            else:
                mech.equilibrium(T0, Y0, T, Y, p=p, hp=True, iters=it, status=s)
            torch.cuda.synchronize()
            out.append([x.cpu().numpy() for x in (T, Y, it, s)])
    finally:
        mech.cea_tolerance(0.0)
    Tr, Yr, itr, sr = oracle_mod.Oracle(md).cea(st["p_bar"], st["T0"], Yh, hp=True, tol=1e-10)
    assert np.all(sr == 0)
    for T, Y, it, s in out:
        assert np.all(s == 0)
        assert np.max(np.abs(T - Tr)) <= 1e-6
        assert np.max(np.abs(Y - Yr) / np.abs(Yr).max(0)) <= 1e-10
    assert out[1][2].mean() < 0.6 * out[0][2].mean()
    # the default-start solve takes the oracle's iteration count
    assert np.mean(out[0][2] == itr) > 0.99


def test_cea_warm_start_iteration_cap(gp24, cuda):
    """flint_cea_warm_options: with the warm-started pass capped at 4 iterations, the states that
    converge within the cap are those of the uncapped warm start (same iterations, same result),
    and every other state is the default-start solution, solved in the compacted second pass
    (its iteration count adds the 5 capped iterations)."""
    import torch
    from flint_amd.warmstart import WarmStart, DEFAULT_MODEL
    md, mech = gp24
    ws = WarmStart.load(DEFAULT_MODEL)
    n = 4096
    st = S.cea_states(md.species, md.W, 24, 0, n)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda)
    Y0, T0, p = dev(S.soa(st["Y"])), dev(st["T0"]), dev(st["p_bar"])
    out = {}
    print("[gpu] cea warm start iteration cap", flush=True)
    try:
        for cap in (50, 4, None):
            T, Y = torch.empty_like(T0), torch.empty_like(Y0)
            it, s = torch.empty(n, dtype=torch.int32, device=cuda), torch.empty(n, dtype=torch.int32, device=cuda)
            if cap is None:
                mech.equilibrium(T0, Y0, T, Y, p=p, hp=True, iters=it, status=s)
            else:
                mech.warm_options(cap)
                ws.solve(mech, md, Y0, T0, p, T, Y, iters=it, status=s)
            torch.cuda.synchronize()
            out[cap] = [x.cpu().numpy() for x in (T, Y, it, s)]
    finally:
        mech.warm_options()
    (T50, Y50, i50, s50), (T4, Y4, i4, s4), (Td, Yd, idf, sd) = out[50], out[4], out[None]
    assert np.all(s50 == 0) and np.all(s4 == 0) and np.all(sd == 0)
    fast = i50 <= 4
    assert 0.2 * n < fast.sum() < n
    assert np.array_equal(T4[fast], T50[fast]) and np.array_equal(Y4[:, fast], Y50[:, fast])
    assert np.array_equal(i4[fast], i50[fast])
    slow = ~fast
    assert np.array_equal(T4[slow], Td[slow]) and np.array_equal(Y4[:, slow], Yd[:, slow])
    assert np.array_equal(i4[slow], idf[slow] + 5)
    with pytest.raises(Exception):
        mech.warm_options(0)


def test_cea_fused_warm_start_matches_layered(gp24, cuda):
    """The one-kernel warm start (features, MFMA network with LDS-resident weights, Newton start)
    against the same network run layer by layer (flint_mlp_forward between the feature and start
    kernels): both feed the same solver (in the tightened mode), so the equilibria agree to 1e-6 K
    and 1e-10 and the iteration counts are those of (nearly) the same starts."""
    import torch
    from flint_amd.warmstart import WarmStart, DEFAULT_MODEL
    md, mech = gp24
    ws = WarmStart.load(DEFAULT_MODEL)
    n = 4096 + 13
    st = S.cea_states(md.species, md.W, 25, 0, n)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda)
    Y0, T0, p = dev(S.soa(st["Y"])), dev(st["T0"]), dev(st["p_bar"])
    out = {}
    print("[gpu] cea fused vs layered warm start", flush=True)
    try:
        mech.cea_tolerance(1e-10)
        for fused in (True, False):
            mech.warm_options(fused=fused)
            T, Y = torch.empty_like(T0), torch.empty_like(Y0)
            it, s = torch.empty(n, dtype=torch.int32, device=cuda), torch.empty(n, dtype=torch.int32, device=cuda)
            ws.solve(mech, md, Y0, T0, p, T, Y, iters=it, status=s)
            torch.cuda.synchronize()
            out[fused] = [x.cpu().numpy() for x in (T, Y, it, s)]
    finally:
        mech.warm_options()
        mech.cea_tolerance(0.0)
    (Tf, Yf, itf, sf), (Tl, Yl, itl, sl) = out[True], out[False]
    assert np.all(sf == 0) and np.all(sl == 0)
    # under the tightened tests both starts reach the same equilibrium to the workload's bar
    assert np.max(np.abs(Tf - Tl)) <= 1e-6
    assert np.max(np.abs(Yf - Yl) / np.abs(Yl).max(0)) <= 1e-10
    assert np.mean(itf == itl) > 0.95 and abs(itf.mean() - itl.mean()) < 0.05 * itl.mean()
    print(f"[gpu] identical iteration counts {np.mean(itf == itl):.4f}, identical T {np.mean(Tf == Tl):.4f}",
          flush=True)
