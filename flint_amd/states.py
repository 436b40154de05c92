"""Seeded synthetic reactor states (counter-based, identical on every host).

Cell ``i`` of seed ``s`` draws its j-th uniform from splitmix64(s, i, j), so any index range can
be regenerated independently; fixtures store (seed, first index, count).

Composition model: unburnt CH4/O2 at equivalence ratio phi (2/phi mol O2 per mol CH4), complete
combustion products (CO2 + 2 H2O + leftover O2 for phi <= 1, CH4 left over for phi > 1), mixed by
a progress variable c and renormalised.  The WD, TSR-GP-24 and TSR-Rich-31 species sets carry
no N2, so the oxidiser is pure O2 (the reference's own batch tests use CH4:1, O2:4,
src/test/CXX/test-batchCXX.cpp:470).  Densities follow rho = p/(R_mix T),
R_mix = sum Y_k 8314.51/W_k (Lib_ThermoTransport.f90:13,217-230).
"""
import numpy as np

RUNIV = 8314.51
P_ATM = 101325.0
_GOLD = np.uint64(0x9E3779B97F4A7C15)


def _splitmix(x):
    x = (x + _GOLD).astype(np.uint64)
    z = x
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)).astype(np.uint64)
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)).astype(np.uint64)
    return z ^ (z >> np.uint64(31))


def uniforms(seed, idx, nstream):
    """[len(idx), nstream] uniforms in [0,1) from (seed, cell index, stream)."""
    idx = np.asarray(idx, dtype=np.uint64)
    with np.errstate(over="ignore"):
        base = (np.uint64(seed) * np.uint64(0xD1B54A32D192ED03)) ^ (idx * np.uint64(nstream + 1))
        u = np.empty((idx.size, nstream))
        for j in range(nstream):
            z = _splitmix(base + np.uint64(j + 1) * _GOLD)
            u[:, j] = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return u


def _blend(species, W, phi, c):
    """Mass fractions (n, ns) of the reactant/product blend."""
    n = phi.size
    k = {s: species.index(s) for s in ("CH4", "O2", "CO2", "H2O")}
    wu = np.zeros((n, len(species)))
    wb = np.zeros_like(wu)
    wu[:, k["CH4"]] = 1.0 * W[k["CH4"]]
    wu[:, k["O2"]] = (2.0 / phi) * W[k["O2"]]
    lean = phi <= 1.0
    wb[:, k["CO2"]] = np.where(lean, 1.0, 1.0 / phi) * W[k["CO2"]]
    wb[:, k["H2O"]] = np.where(lean, 2.0, 2.0 / phi) * W[k["H2O"]]
    wb[:, k["O2"]] = np.where(lean, 2.0 / phi - 2.0, 0.0) * W[k["O2"]]
    wb[:, k["CH4"]] = np.where(lean, 0.0, 1.0 - 1.0 / phi) * W[k["CH4"]]
    yu = wu / wu.sum(axis=1, keepdims=True)
    yb = wb / wb.sum(axis=1, keepdims=True)
    y = (1.0 - c)[:, None] * yu + c[:, None] * yb
    return y / y.sum(axis=1, keepdims=True)


def _to_rho(Y, W, T, p):
    Rmix = (Y * (RUNIV / np.asarray(W))[None, :]).sum(axis=1)
    rho = p / (Rmix * T)
    return (rho[:, None] * Y)


def wd_states(species, W, seed, start, count):
    """WD source-term workload: T~U[1200,2400] K, p = 1 atm, phi~U[0.5,1.5], c~U[0,1]."""
    u = uniforms(seed, np.arange(start, start + count), 3)
    T = 1200.0 + 1200.0 * u[:, 0]
    phi = 0.5 + u[:, 1]
    c = u[:, 2]
    Y = _blend(species, np.asarray(W), phi, c)
    p = np.full(count, P_ATM)
    return dict(rho=_to_rho(Y, W, T, p), T=T, p=p, Y=Y)


RADICALS = ("H", "O", "OH", "HO2", "CH3")


def radical_states(species, W, seed, start, count, T_lo=900.0, T_hi=2800.0, ignition_half=False):
    """Seeded states of the batched GP-24 / Rich-31 workloads.

    T~U[T_lo,T_hi], p log-uniform in [1,50] atm, phi~U[0.5,1.5], c~U[0,1], then each of H, O,
    OH, HO2, CH3 gets a log-uniform 1e-6..1e-3 mass fraction before renormalising.  With
    ``ignition_half`` every odd cell is replaced by a near-ignition state, T~U[1000,1300] K and
    c~U[0,0.05].
    """
    idx = np.arange(start, start + count)
    u = uniforms(seed, idx, 4 + len(RADICALS))
    T = T_lo + (T_hi - T_lo) * u[:, 0]
    p = P_ATM * np.exp(np.log(50.0) * u[:, 1])
    phi = 0.5 + u[:, 2]
    c = u[:, 3]
    if ignition_half:
        odd = (idx % 2) == 1
        T = np.where(odd, 1000.0 + 300.0 * u[:, 0], T)
        c = np.where(odd, 0.05 * u[:, 3], c)
    Y = _blend(species, np.asarray(W), phi, c)
    for j, r in enumerate(RADICALS):
        Y[:, species.index(r)] += 10.0 ** (-6.0 + 3.0 * u[:, 4 + j])
    Y = Y / Y.sum(axis=1, keepdims=True)
    return dict(rho=_to_rho(Y, W, T, p), T=T, p=p, Y=Y)


def cea_states(species, W, seed, start, count):
    """Seeded states of the batched HP-equilibrium workload: the composition and pressure of
    :func:`radical_states` with the unburnt temperature T0~U[300,900] K in place of T.  Returns Y
    (the solver's input amounts, summing to 1), T0 and p in bar (the solver's pressure unit)."""
    st = radical_states(species, W, seed, start, count, T_lo=300.0, T_hi=900.0)
    return dict(Y=st["Y"], T0=st["T"], p_bar=st["p"] / 1e5)


def flow_states(species, W, seed, start, count, nd=3):
    """Seeded CFD cells of the cons2prim / prim2cons workload: the composition and pressure of
    :func:`radical_states` with T~U[600,2800] K, nd velocity components ~U[-300,300] m/s, and a
    Newton start T0 within 2 % of the cell's T (as a CFD step starts from the previous step's
    temperature).  Returns prim [ns+nd+1, count] = (rho_k, u, p), T and T0."""
    st = radical_states(species, W, seed, start, count, T_lo=600.0, T_hi=2800.0)
    u = uniforms(seed + 7919, np.arange(start, start + count), nd + 1)
    vel = -300.0 + 600.0 * u[:, :nd]
    T0 = st["T"] * (1.0 + 0.04 * (u[:, nd] - 0.5))
    prim = np.ascontiguousarray(np.vstack([soa(st["rho"]), soa(vel), st["p"][None, :]]))
    return dict(prim=prim, T=st["T"], T0=T0)


def soa(a):
    """(n, k) -> contiguous (k, n) structure-of-arrays."""
    return np.ascontiguousarray(np.asarray(a).T)
