"""The data-driven general source term on the GPU (flint_general_setup, flint_source_general).

* against the reference's own general outputs (tests/golden/general_<mech>.npz) and the oracle:
  <= 1e-12 normwise (FLINT's metric: the scale is the sum of |W_k (ni2 - ni1)| (|rate_fwd| +
  |rate_rev|) over reactions); only log10 and 10** of the Troe factor differ from the host's in
  the last bit;
* a mechanism without compiled kernels (Pelucchi: Arrhenius, Troe and Lindemann reactions) gets
  its source terms from general, as FLINT's Assign_Mechanism falls back to it, and reports the
  entry points that need compiled kernels as unsupported;
* out-of-table temperatures set status bit 0, padded leading dimensions leave the padding
  untouched, bad tables are rejected;
* reactions outside the compact per-reaction record (more reactants, exponents nint() takes to
  0, real fall-off exponents, more distinct efficiency sets than the record form holds) and a
  36-species mechanism (concentrations in LDS) against the oracle."""
import numpy as np
import pytest

from conftest import golden
from flint_amd import states as S

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mech", ["TSR-GP-24", "TSR-Rich-31"])
def test_general_matches_reference_outputs(mechs, oracle_mod, cuda, mech):
    import torch
    from flint_amd.native import Mechanism
    md = mechs[mech]
    m = Mechanism.from_data(md)
    m.general_setup(md.general)
    g = golden(f"general_{mech}.npz")
    n = int(g["count"])
    st = S.radical_states(md.species, md.W, int(g["seed"]), int(g["start"]), n, ignition_half=True)
    rho = S.soa(st["rho"])
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda)
    w = torch.empty((md.ns, n), dtype=torch.float64, device=cuda)
    s = torch.full((n,), -1, dtype=torch.int32, device=cuda)
    print(f"[gpu] general {mech} n={n}", flush=True)
    m.source_general(dev(rho), dev(st["T"]), w, status=s)
    torch.cuda.synchronize()
    assert int(s.max()) == 0 and int(s.min()) == 0
    wo, sc = oracle_mod.Oracle(md).general(rho, st["T"])
    wg = w.cpu().numpy()
    assert np.max(np.abs(wg - g["wdot"]) / np.maximum(sc, 1e-300)) <= 1e-12
    assert np.max(np.abs(wg - wo) / np.maximum(sc, 1e-300)) <= 1e-12
    print(f"[gpu] identical to the reference outputs: {np.mean(wg == g['wdot']):.4f} of entries", flush=True)


def test_general_half_integer_coefficients(oracle_mod, cuda):
    """WD's general tables carry the coefficients 0.5 and 1.5, raised to nint() powers
    (Lib_Chemistry_wdot.f90:140-141, halves away from zero): against the reference's outputs."""
    import torch
    from flint_amd.mechdata import mech_data
    from flint_amd.native import Mechanism
    md = mech_data("WD")
    m = Mechanism.from_data(md)
    m.general_setup(md.general)
    g = golden("general_WD.npz")
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda)
    n = g["T"].size
    w = torch.empty((md.ns, n), dtype=torch.float64, device=cuda)
    print(f"[gpu] general WD n={n}", flush=True)
    m.source_general(dev(g["rho"]), dev(g["T"]), w)
    torch.cuda.synchronize()
    _, sc = oracle_mod.Oracle(md).general(g["rho"], g["T"])
    assert np.max(np.abs(w.cpu().numpy() - g["wdot"]) / np.maximum(sc, 1e-300)) <= 1e-12


def test_general_edges(mechs, oracle_mod, cuda):
    import torch
    from flint_amd.native import Mechanism, FlintError
    md = mechs["TSR-GP-24"]
    m = Mechanism.from_data(md)
    with pytest.raises(FlintError):
        m.source_general(torch.zeros((md.ns, 4), dtype=torch.float64, device=cuda),
                         torch.full((4,), 1000.0, dtype=torch.float64, device=cuda),
                         torch.zeros((md.ns, 4), dtype=torch.float64, device=cuda))  # no setup yet
    na, nt, nl, ni1, ni2, eps = md.general
    with pytest.raises(FlintError):
        m.general_setup((na - 1, nt, nl + 1, ni1, ni2, eps))
    m.general_setup(md.general)
    n, ld = 50, 64
    st = S.radical_states(md.species, md.W, 19, 0, n)
    T = st["T"].copy()
    T[7] = 20000.0  # above the kinetics table (15000 K): status bit 0, row clamped
    rho = np.zeros((md.ns, ld))
    rho[:, :n] = S.soa(st["rho"])
    Tp = np.full(ld, 1000.0)
    Tp[:n] = T
    w = torch.full((md.ns, ld), -7.0, dtype=torch.float64, device=cuda)
    s = torch.zeros(ld, dtype=torch.int32, device=cuda)
    m.source_general(torch.from_numpy(rho).to(cuda), torch.from_numpy(Tp).to(cuda), w, status=s, n=n)
    torch.cuda.synchronize()
    wn, sn = w.cpu().numpy(), s.cpu().numpy()
    assert sn[7] == 1 and np.all(np.delete(sn[:n], 7) == 0) and np.all(wn[:, n:] == -7.0)
    ok = np.arange(n) != 7
    wo, sc = oracle_mod.Oracle(md).general(rho[:, :n][:, ok], T[ok])
    assert np.max(np.abs(wn[:, :n][:, ok] - wo) / np.maximum(sc, 1e-300)) <= 1e-12


def test_mechanism_without_compiled_kernels_uses_general(oracle_mod, cuda):
    import torch
    from flint_amd.mechdata import mech_data
    from flint_amd.native import Mechanism, FlintError
    md = mech_data("Pelucchi")
    assert md.n_lind > 0 and md.n_troe > 0
    m = Mechanism.from_data(md)
    n = 2048
    T = 300.0 + 2600.0 * (np.arange(n) % 997) / 997.0
    rng = np.random.default_rng(5)
    rho = 10.0 ** rng.uniform(-9, -1, (md.ns, n))
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda)
    w = torch.empty((md.ns, n), dtype=torch.float64, device=cuda)
    with pytest.raises(FlintError):
        m.source(dev(rho), dev(T), w)  # no compiled routine and no general tables yet
    with pytest.raises(FlintError):
        m.rhs(dev(np.vstack([rho, T[None]])), torch.empty((md.ns + 1, n), dtype=torch.float64, device=cuda))
    m.general_setup(md.general)
    print(f"[gpu] Pelucchi through general n={n}", flush=True)
    m.source(dev(rho), dev(T), w)
    torch.cuda.synchronize()
    wo, sc = oracle_mod.Oracle(md).general(rho, T)
    assert np.max(np.abs(w.cpu().numpy() - wo) / np.maximum(sc, 1e-300)) <= 1e-12
    # against the reference's own outputs (tests/golden/general_Pelucchi.npz)
    g = golden("general_Pelucchi.npz")
    wg = torch.empty((md.ns, g["T"].size), dtype=torch.float64, device=cuda)
    m.source(dev(g["rho"]), dev(g["T"]), wg)
    torch.cuda.synchronize()
    _, sg = oracle_mod.Oracle(md).general(g["rho"], g["T"])
    assert np.max(np.abs(wg.cpu().numpy() - g["wdot"]) / np.maximum(sg, 1e-300)) <= 1e-12


def _gp24_variant_tables(md):
    """GP-24's general tables altered so that reactions leave the compact form: five reactants,
    an exponent that nint() takes to 0, a real exponent on a Troe reaction, and 20 third bodies
    with efficiency rows of their own (more distinct sets than the compact form holds)."""
    na, nt, nl, ni1, ni2, eps = md.general
    ni1, ni2, eps = ni1.copy(), ni2.copy(), eps.copy()
    ns = md.ns
    ni1[0, :5] = 1.0
    ni1[3, 7] = 0.4
    ni1[na, 2] = 1.5
    rng = np.random.default_rng(3)
    for j in range(10, 30):
        ni1[j, ns] = 1.0
        eps[j, :ns] = np.round(rng.uniform(0.5, 3.0, ns), 2)
    return na, nt, nl, ni1, ni2, eps


def test_general_list_form_reactions(mechs, oracle_mod, cuda):
    """Reactions that do not fit the compact record run from their entry lists: against the
    oracle's general on the same altered tables."""
    import copy
    import torch
    from flint_amd.native import Mechanism
    md = copy.copy(mechs["TSR-GP-24"])
    md.general = _gp24_variant_tables(md)
    m = Mechanism.from_data(md)
    m.general_setup(md.general)
    n = 4096
    st = S.radical_states(md.species, md.W, 41, 0, n, ignition_half=True)
    rho = S.soa(st["rho"])
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda)
    w = torch.empty((md.ns, n), dtype=torch.float64, device=cuda)
    print(f"[gpu] general list-form reactions n={n}", flush=True)
    m.source_general(dev(rho), dev(st["T"]), w)
    torch.cuda.synchronize()
    wo, sc = oracle_mod.Oracle(md).general(rho, st["T"])
    assert np.max(np.abs(w.cpu().numpy() - wo) / np.maximum(sc, 1e-300)) <= 1e-12


def test_general_beyond_32_species(mechs, oracle_mod, cuda):
    """A mechanism of 36 species (GP-24's species scattered, in their own order, among 12 inert
    ones that take part in no reaction) runs the kernel that keeps the concentrations in LDS:
    its rates equal GP-24's from the oracle within the normwise bound and the inert species get
    exactly zero."""
    import copy
    import torch
    from flint_amd.native import Mechanism
    md = mechs["TSR-GP-24"]
    ns, nx = md.ns, 36
    rng = np.random.default_rng(8)
    pos = np.sort(rng.choice(nx, ns, replace=False))  # new index of each GP-24 species
    inert = np.setdiff1d(np.arange(nx), pos)
    k_in = md.species.index("CO2")  # thermo of the inert species
    src = np.full(nx, k_in)
    src[pos] = np.arange(ns)
    big = copy.copy(md)
    big.ns = nx
    big.species = [md.species[s] if i in pos else f"X{i}" for i, s in enumerate(src)]
    big.W = np.ascontiguousarray(md.W[src])
    big.cp, big.h, big.s = (np.ascontiguousarray(a[:, src]) for a in (md.cp, md.h, md.s))
    big.composition = np.zeros((0, nx))
    big.elements = []
    big.desc = dict(md.desc, name="GP24-scattered-36")
    na, nt, nl, ni1, ni2, eps = md.general
    nr = ni1.shape[0]

    def scatter(a):
        out = np.zeros((nr, nx + 1))
        out[:, pos] = a[:, :ns]
        out[:, nx] = a[:, ns]
        return out

    big.general = (na, nt, nl, scatter(ni1), scatter(ni2), scatter(eps))
    m = Mechanism.from_data(big)
    m.general_setup(big.general)
    n = 4096
    st = S.radical_states(md.species, md.W, 43, 0, n, ignition_half=True)
    rho = S.soa(st["rho"])
    rho_big = np.zeros((nx, n))
    rho_big[pos] = rho
    rho_big[inert] = 1e-3 * rho[k_in]
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda)
    w = torch.empty((nx, n), dtype=torch.float64, device=cuda)
    print(f"[gpu] general 36 species n={n}", flush=True)
    m.source_general(dev(rho_big), dev(st["T"]), w)
    torch.cuda.synchronize()
    wg = w.cpu().numpy()
    wo, sc = oracle_mod.Oracle(md).general(rho, st["T"])
    assert np.all(wg[inert] == 0.0)
    assert np.max(np.abs(wg[pos] - wo) / np.maximum(sc, 1e-300)) <= 1e-12
