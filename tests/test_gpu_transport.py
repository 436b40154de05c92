"""Mixture transport by Wilke's rule on the GPU (flint_transport).

The kernel splits sqrt(mu_i / (mu_j + 1e-20)) into sqrt(mu_i) / sqrt(mu_j + 1e-20) and expands
the square of each interaction factor into three constant-matrix products; every term is
positive, so the results differ from the reference's by a few ulps per species.  Tolerance:
1e-13 relative on mu and lambda.
* against the reference's own outputs (tests/golden/transport_<mech>.npz) for GP-24 and Rich-31
  (the compiled 24- and 31-species kernels);
* against the oracle for WD (5 species) and Pelucchi (25 species, the generic loop kernel);
* 2^20 cells: pure-species cells return their species' table values, a sample matches the oracle;
* host form, tables read from transport.dat by flint_mech_load, out-of-table T (status 2, NaN),
  and the error before flint_transport_setup."""
import os

import numpy as np
import pytest

from conftest import golden
from flint_amd import states as S
from flint_amd.transport import species_tables

pytestmark = pytest.mark.gpu
TOL = 1e-13


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def _run(m, r, rho, T, cuda):
    import torch
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda)
    n = T.size
    mu = torch.empty(n, dtype=torch.float64, device=cuda)
    lam = torch.empty_like(mu)
    st = torch.full((n,), -1, dtype=torch.int32, device=cuda)
    m.transport(dev(r), dev(rho), dev(T), mu, lam, status=st)
    torch.cuda.synchronize()
    return mu.cpu().numpy(), lam.cpu().numpy(), st.cpu().numpy()


def _random_cells(md, n, seed):
    rng = np.random.default_rng(seed)
    r = rng.uniform(0.0, 1.0, (md.ns, n)) ** 4 * 0.5
    return r, r.sum(0), rng.uniform(600.0, 2800.0, n)


@pytest.mark.parametrize("mech", ["TSR-GP-24", "TSR-Rich-31"])
def test_wilke_matches_reference_outputs(mechs, cuda, mech):
    from flint_amd.native import Mechanism
    md = mechs[mech]
    g = golden(f"transport_{mech}.npz")
    fs = S.flow_states(md.species, md.W, int(g["seed"]), int(g["start"]), int(g["count"]))
    r = fs["prim"][:md.ns]
    m = Mechanism.from_data(md)
    m.transport_setup(species_tables(md, int(g["table_seed"])))
    print(f"[gpu] transport {mech} n={r.shape[1]}", flush=True)
    mu, lam, st = _run(m, r, r.sum(0), fs["T"], cuda)
    assert (st == 0).all()
    assert _rel(mu, g["mu"]) <= TOL and _rel(lam, g["lam"]) <= TOL


@pytest.mark.parametrize("mech", ["WD", "Pelucchi"])
def test_wilke_matches_oracle_other_species_counts(oracle_mod, cuda, mech):
    from flint_amd.mechdata import mech_data
    from flint_amd.native import Mechanism
    md = mech_data(mech)
    tb = species_tables(md)
    r, rho, T = _random_cells(md, 4096, 5)
    m = Mechanism.from_data(md)
    m.transport_setup(tb)
    print(f"[gpu] transport {mech} ns={md.ns}", flush=True)
    mu, lam, st = _run(m, r, rho, T, cuda)
    omu, olam, ost = oracle_mod.Oracle(md).wilke(tb, r, rho, T)
    assert (st == 0).all() and (ost == 0).all()
    assert _rel(mu, omu) <= TOL and _rel(lam, olam) <= TOL


def test_wilke_full_size(mechs, oracle_mod, cuda):
    from flint_amd.native import Mechanism
    md = mechs["TSR-GP-24"]
    tb = species_tables(md)
    Tt_min, _, mu_tab, k_tab = tb
    n = 1 << 20
    r, rho, T = _random_cells(md, n, 9)
    pure = np.arange(0, n, 97)
    sp = pure % md.ns
    r[:, pure] = 0.0
    r[sp, pure] = 0.3
    rho = r.sum(0)
    m = Mechanism.from_data(md)
    m.transport_setup(tb)
    print(f"[gpu] transport full size n={n}", flush=True)
    mu, lam, st = _run(m, r, rho, T, cuda)
    assert (st == 0).all()
    Tp = T[pure]
    i = np.floor(Tp).astype(int) - Tt_min
    td = Tp - np.floor(Tp)
    mu_s = mu_tab[sp, i] + (mu_tab[sp, i + 1] - mu_tab[sp, i]) * td
    k_s = k_tab[sp, i] + (k_tab[sp, i + 1] - k_tab[sp, i]) * td
    assert _rel(mu[pure], mu_s) <= 1e-14 and _rel(lam[pure], k_s) <= 1e-14
    sample = np.arange(5, n, 251)
    omu, olam, _ = oracle_mod.Oracle(md).wilke(tb, r[:, sample], rho[sample], T[sample])
    assert _rel(mu[sample], omu) <= TOL and _rel(lam[sample], olam) <= TOL


def test_wilke_host_form_folder_and_edges(mechs, oracle_mod, cuda, tmp_path):
    from flint_amd import tables as TB
    from flint_amd.native import FlintError, Mechanism
    md = mechs["TSR-GP-24"]
    tb = species_tables(md)
    r, rho, T = _random_cells(md, 1000, 13)
    T[:3] = [5.0, 15000.0, 20000.0]
    m = Mechanism.from_data(md)
    with pytest.raises(FlintError):
        m.transport_host(r, rho, T)
    m.transport_setup(tb)
    mu, lam, st = m.transport_host(r, rho, T)
    assert list(st[:3]) == [2, 2, 2] and (st[3:] == 0).all()
    assert np.isnan(mu[:3]).all() and np.isnan(lam[:3]).all()
    dmu, dlam, _ = _run(m, r, rho, T, cuda)
    assert np.array_equal(mu[3:], dmu[3:]) and np.array_equal(lam[3:], dlam[3:])
    omu, olam, _ = oracle_mod.Oracle(md).wilke(tb, r[:, 3:], rho[3:], T[3:])
    assert _rel(mu[3:], omu) <= TOL
    folder = str(tmp_path / "native")
    TB.write_native("TSR-GP-24", folder, transport=(tb[2], tb[3]))
    assert os.path.exists(os.path.join(folder, "transport.dat"))
    mf = Mechanism.load(folder)
    fmu, flam, fst = mf.transport_host(r, rho, T)
    assert np.array_equal(fst, st)
    assert np.array_equal(fmu[3:], mu[3:]) and np.array_equal(flam[3:], lam[3:])
