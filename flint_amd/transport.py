"""Species transport tables for the mixture-transport kernel (Wilke's rule).

FLINT reads mi_tab / k_tab from an optional transport.dat (Load_ThermoTransport.f90:129-174) and
ships neither such a file nor a generator for one.  :func:`species_tables` makes tables of physical
magnitude on the mechanism's thermo temperature grid: a power law mu_k = mu0_k (T / 300)^0.7 with
mu0_k = 1.8e-5 Pa s (W_k / 28)^(1/2) times a seeded factor in [0.8, 1.2], and the modified Eucken
conductivity k_k = mu_k (cp_k(T) + 1.25 R_k) (Poling, Prausnitz and O'Connell, The Properties of
Gases and Liquids, 5th ed., eq. 10-3.3) from the mechanism's cp table.  The mixing rule the kernel
implements does not depend on where the species values come from.
"""
import numpy as np

RUNIV = 8314.51  # J/(kmol K), Lib_ThermoTransport.f90:13


def species_tables(md, seed=7):
    """(Tt_min, Tt_max, mu_tab [ns, nT], k_tab [ns, nT]) on md's thermo rows, temperature fastest
    per species (the reference's mi_tab(Tmin:Tmax, ns) in memory order)."""
    nT, ns = md.cp.shape
    T = md.Th_min + np.arange(nT, dtype=np.float64)
    Tpos = np.maximum(T, 1.0)
    rng = np.random.default_rng(seed)
    mu0 = 1.8e-5 * np.sqrt(np.asarray(md.W) / 28.0) * rng.uniform(0.8, 1.2, ns)
    mu = mu0[:, None] * (Tpos[None, :] / 300.0) ** 0.7
    R = RUNIV / np.asarray(md.W)
    k = mu * (np.asarray(md.cp).T + 1.25 * R[:, None])
    return md.Th_min, md.Th_min + nT - 1, np.ascontiguousarray(mu), np.ascontiguousarray(k)
