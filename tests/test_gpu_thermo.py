"""cons2prim / prim2cons kernels (flint_cons2prim, flint_prim2cons) on the GPU.

* against the reference's own outputs (tests/golden/cons2prim_<mech>.npz, flang build of
  Lib_ThermoTransport.f90:130-185) and the oracle: same arithmetic in the same order, so equal to
  round-off (1e-14 relative, equality expected);
* prim -> cons -> prim on 2^20 cells returns p and T to the Newton tolerance;
* a start outside the thermo table gives status 2 and NaN; bad sizes are rejected."""
import numpy as np
import pytest

from conftest import golden
from flint_amd import states as S

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


@pytest.mark.parametrize("mech", ["TSR-GP-24", "TSR-Rich-31"])
def test_cons2prim_matches_reference_outputs(mechs, oracle_mod, cuda, mech):
    import torch
    from flint_amd.native import Mechanism
    md = mechs[mech]
    m = Mechanism.from_data(md)
    g = golden(f"cons2prim_{mech}.npz")
    n, nd = int(g["count"]), int(g["nd"])
    fs = S.flow_states(md.species, md.W, int(g["seed"]), int(g["start"]), n, nd=nd)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda)
    q = dev(fs["prim"])
    c = torch.empty_like(q)
    st = torch.full((n,), -1, dtype=torch.int32, device=cuda)
    print(f"[gpu] prim2cons {mech} n={n} nd={nd}", flush=True)
    m.prim2cons(q, c, status=st)
    torch.cuda.synchronize()
    assert int(st.max()) == 0 and int(st.min()) == 0
    assert _rel(c.cpu().numpy(), g["cons"]) <= 1e-14
    p = torch.empty_like(q)
    T = torch.empty(n, dtype=torch.float64, device=cuda)
    it = torch.empty(n, dtype=torch.int32, device=cuda)
    print(f"[gpu] cons2prim {mech}", flush=True)
    m.cons2prim(dev(g["cons"]), dev(fs["T0"]), p, T_out=T, iters=it, status=st)
    torch.cuda.synchronize()
    assert int(st.max()) == 0
    assert _rel(p.cpu().numpy(), g["prim"]) <= 1e-14
    po, To, ito, sto = oracle_mod.Oracle(md).cons2prim(g["cons"], fs["T0"])
    assert np.array_equal(it.cpu().numpy(), ito) and _rel(T.cpu().numpy(), To) <= 1e-14
    print(f"[gpu] identical to the reference outputs: {np.array_equal(p.cpu().numpy(), g['prim'])}", flush=True)


def test_cons2prim_round_trip_full_size(mechs, cuda):
    import torch
    from flint_amd.native import Mechanism
    md = mechs["TSR-GP-24"]
    m = Mechanism.from_data(md)
    n, nd, ns = 1 << 20, 3, md.ns
    fs = S.flow_states(md.species, md.W, 17, 0, n, nd=nd)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda)
    q = dev(fs["prim"])
    c, p = torch.empty_like(q), torch.empty_like(q)
    T = torch.empty(n, dtype=torch.float64, device=cuda)
    it = torch.empty(n, dtype=torch.int32, device=cuda)
    st = torch.empty(n, dtype=torch.int32, device=cuda)
    print(f"[gpu] prim2cons -> cons2prim n={n}", flush=True)
    m.prim2cons(q, c)
    m.cons2prim(c, dev(fs["T0"]), p, T_out=T, iters=it, status=st)
    torch.cuda.synchronize()
    assert int(st.max()) == 0
    assert torch.equal(p[:ns], q[:ns])
    assert float(((p[ns:ns + nd] - q[ns:ns + nd]).abs() / 300.0).max()) < 1e-15
    assert float((p[-1] / q[-1] - 1.0).abs().max()) < 1e-7
    assert float((T / dev(fs["T"]) - 1.0).abs().max()) < 1e-7
    assert int(it.min()) >= 1 and int(it.max()) <= 5


def test_cons2prim_edges_and_host_form(mechs, cuda):
    import torch
    from flint_amd.native import Mechanism, FlintError
    md = mechs["TSR-Rich-31"]
    m = Mechanism.from_data(md)
    n, nd, ld = 37, 1, 64
    fs = S.flow_states(md.species, md.W, 18, 0, n, nd=nd)
    cons = m.prim2cons_host(fs["prim"])
    T0 = fs["T0"].copy()
    T0[5] = 100.0  # below the thermo table (500 K)
    prim_h, T_h, st_h = m.cons2prim_host(cons, T0)
    assert st_h[5] == 2 and np.isnan(prim_h[-1, 5]) and np.isnan(T_h[5])
    assert np.all(np.delete(st_h, 5) == 0)
    # device form with a padded leading dimension gives the same cells
    cpad = np.zeros((cons.shape[0], ld))
    cpad[:, :n] = cons
    Tpad = np.full(ld, 1000.0)
    Tpad[:n] = T0
    c = torch.from_numpy(cpad).to(cuda)
    p = torch.full_like(c, -7.0)
    m.cons2prim(c, torch.from_numpy(Tpad).to(cuda), p, n=n)
    torch.cuda.synchronize()
    pd = p.cpu().numpy()
    assert np.array_equal(pd[:, :n], prim_h, equal_nan=True) and np.all(pd[:, n:] == -7.0)
    m.cons2prim(c[:, :0], torch.from_numpy(Tpad[:0]).to(cuda), p[:, :0])  # n = 0
    with pytest.raises(FlintError):
        m.cons2prim_host(np.zeros((md.ns + 5, 4)), np.full(4, 1000.0))  # nd = 4
