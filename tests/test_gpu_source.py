"""HIP source-term, RHS and Jacobian kernels against the oracle and the reference fixtures."""
import numpy as np
import pytest

from conftest import golden, golden_states, normwise
from flint_amd import states as S
from flint_amd.native import Mechanism

pytestmark = pytest.mark.gpu
TOL = 1e-12  # normwise, FP64


@pytest.fixture(scope="module")
def gpu_mechs(mechs, cuda):
    return {n: Mechanism.from_data(md) for n, md in mechs.items()}


def _dev(a, cuda, ld=None):
    import torch
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 1:
        a = a[None, :]
    n = a.shape[1]
    ld = ld or n
    t = torch.full((a.shape[0], ld), np.nan, dtype=torch.float64, device=cuda)
    t[:, :n] = torch.from_numpy(a).to(cuda)
    return t


@pytest.mark.parametrize("mech", ["WD", "TSR-GP-24", "TSR-Rich-31"])
def test_source_matches_reference_fixture(gpu_mechs, mechs, oracle_mod, cuda, mech):
    import torch
    md = mechs[mech]
    g = golden(f"wdot_{mech}.npz")
    st = golden_states(md, g)
    rho, T = S.soa(st["rho"]), st["T"]
    n = T.size
    print(f"[gpu] source {mech} n={n}", flush=True)
    ld = n + 61  # ragged leading dimension
    drho, dT = _dev(rho, cuda, ld), _dev(T, cuda, ld)[0]
    w = torch.zeros((md.ns, ld), dtype=torch.float64, device=cuda)
    st_d = torch.zeros(ld, dtype=torch.int32, device=cuda)
    gpu_mechs[mech].source(drho, dT, w, status=st_d, n=n)
    torch.cuda.synchronize()
    w = w[:, :n].cpu().numpy()
    _, sc = oracle_mod.Oracle(md).wdot(rho, T)
    assert normwise(w, g["wdot"], sc) <= TOL
    assert int(st_d[:n].abs().sum()) == 0


@pytest.mark.parametrize("mech", ["TSR-GP-24", "TSR-Rich-31"])
def test_source_matches_oracle_large(gpu_mechs, mechs, oracle_mod, cuda, mech):
    import torch
    md = mechs[mech]
    n = 50_000 + 17
    st = S.radical_states(md.species, md.W, 21, 0, n, ignition_half=True)
    rho, T = S.soa(st["rho"]), st["T"]
    print(f"[gpu] source large {mech} n={n}", flush=True)
    w = torch.empty((md.ns, n), dtype=torch.float64, device=cuda)
    gpu_mechs[mech].source(_dev(rho, cuda), _dev(T, cuda)[0], w)
    torch.cuda.synchronize()
    wo, sc = oracle_mod.Oracle(md).wdot(rho, T)
    assert normwise(w.cpu().numpy(), wo, sc) <= TOL


def test_wd_heat_release(gpu_mechs, mechs, oracle_mod, cuda):
    import torch
    md = mechs["WD"]
    n = 65536
    st = S.wd_states(md.species, md.W, 1, 0, n)
    rho, T = S.soa(st["rho"]), st["T"]
    print(f"[gpu] wd hrr n={n}", flush=True)
    w = torch.empty((5, n), dtype=torch.float64, device=cuda)
    q = torch.empty(n, dtype=torch.float64, device=cuda)
    gpu_mechs["WD"].source(_dev(rho, cuda), _dev(T, cuda)[0], w, hrr=q)
    torch.cuda.synchronize()
    wo, qo = oracle_mod.Oracle(md).wd_hrr(rho, T)
    assert np.allclose(w.cpu().numpy(), wo, rtol=1e-12, atol=0)
    h_scale = np.abs(wo).sum(0) * np.abs(md.h).max()
    assert np.max(np.abs(q.cpu().numpy() - qo) / h_scale) < 1e-13


@pytest.mark.parametrize("mech", ["TSR-GP-24", "TSR-Rich-31"])
def test_rhs_and_jacobian_match_oracle(gpu_mechs, mechs, oracle_mod, cuda, mech):
    import torch
    md = mechs[mech]
    n = 4096 + 3
    st = S.radical_states(md.species, md.W, 31, 0, n, ignition_half=True)
    Z = np.vstack([S.soa(st["rho"]), st["T"][None, :]])
    Z[3, 5] = -1e-9  # a negative density: clipped, zero derivative column entry
    nz = md.ns + 1
    print(f"[gpu] rhs+jac {mech} n={n}", flush=True)
    F = torch.empty((nz, n), dtype=torch.float64, device=cuda)
    J = torch.empty((nz * nz, n), dtype=torch.float64, device=cuda)
    gpu_mechs[mech].rhs(_dev(Z, cuda), F, J)
    torch.cuda.synchronize()
    Fo, Jo, Ja, Fa = oracle_mod.Oracle(md).jac(Z, fscale=True)
    F, J = F.cpu().numpy(), J.cpu().numpy()
    # species rows and the T row (sum_k |e_k| Fa_k / |rho cv|) both normwise
    assert normwise(F, Fo, Fa) <= TOL
    assert normwise(J, Jo, Ja) <= TOL
    # RHS-only launch agrees with the fused one
    F2 = torch.empty((nz, n), dtype=torch.float64, device=cuda)
    gpu_mechs[mech].rhs(_dev(Z, cuda), F2)
    torch.cuda.synchronize()
    assert np.array_equal(F2.cpu().numpy(), F)


def test_out_of_range_and_empty(gpu_mechs, mechs, cuda):
    import torch
    md = mechs["TSR-Rich-31"]  # tables 500..5000 K
    st = S.radical_states(md.species, md.W, 41, 0, 4)
    Z = np.vstack([S.soa(st["rho"]), np.array([[400.0, 5000.0, 6000.0, np.nan]])])
    nz = md.ns + 1
    print("[gpu] out-of-range", flush=True)
    F = torch.empty((nz, 4), dtype=torch.float64, device=cuda)
    J = torch.empty((nz * nz, 4), dtype=torch.float64, device=cuda)
    stt = torch.zeros(4, dtype=torch.int32, device=cuda)
    gpu_mechs["TSR-Rich-31"].rhs(_dev(Z, cuda), F, J, status=stt)
    torch.cuda.synchronize()
    F, J, stt = F.cpu().numpy(), J.cpu().numpy(), stt.cpu().numpy()
    assert np.all(F[:, [0, 2, 3]] == -1.0) and np.all(J[:, [0, 2, 3]] == 0.0)
    assert stt[0] & 2 and stt[2] & 2 and stt[3] & 2 and not (stt[1] & 2)
    assert np.all(np.isfinite(F[:, 1]))  # T = Tmax uses the guard row
    # n = 0 is a no-op
    gpu_mechs["TSR-Rich-31"].rhs(_dev(Z, cuda), torch.empty((nz, 4), dtype=torch.float64, device=cuda), n=0)
