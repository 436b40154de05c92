"""FLINT's data-driven general routine (Lib_Chemistry_wdot.f90:96-221) in the oracle.

* the tables derived from the mechanism descriptors (flint_amd.mechanism.general_tables) equal the
  Reaction definition blocks of the chemistry-info.txt files the reference ships for TSR-GP-24
  and TSR-Rich-31 (test/<mech>/INPUT, read where the reference checkout exists);
* oracle/orc_general is bit-identical to the reference's outputs (tests/golden/general_<mech>.npz,
  flang build) and, where it exists, to the reference build itself;
* general differs from the explicit generated routines only through their float32-rounded
  efficiency literals (normwise ~2e-8), so the two are not interchangeable parity targets."""
import os

import numpy as np
import pytest

from conftest import golden
from flint_amd import states as S

REF = "/root/reference/test"
REF_LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "libflint_ref.so")


def _reaction_definition(path, ns):
    L = open(path).read().splitlines()
    k = L.index("Reaction definition") + 1
    rows = [l.split() for l in L[k:] if l.strip()]
    a = np.array([[float(r[2]), float(r[3]), float(r[4])] for r in rows]).reshape(-1, ns + 1, 3)
    types = []
    j = L.index("Reaction type") + 1
    while L[j].strip():
        types.append(L[j].split(None, 1)[1])
        j += 1
    return types, a


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference checkout absent")
@pytest.mark.parametrize("mech", ["TSR-GP-24", "TSR-Rich-31", "Pelucchi"])
def test_general_tables_match_shipped_chemistry_info(mechs, mech):
    from flint_amd.mechdata import mech_data
    md = mech_data(mech)
    na, nt, nl, ni1, ni2, eps = md.general
    types, a = _reaction_definition(os.path.join(REF, mech, "INPUT", "chemistry-info.txt"), md.ns)
    kind = np.array([2 if "Lindemann" in t else (1 if "Troe" in t else 0) for t in types])
    order = np.concatenate([np.nonzero(kind == q)[0] for q in (0, 1, 2)])  # FLINT's per-kind order
    assert (na, nt, nl) == tuple(int((kind == q).sum()) for q in (0, 1, 2))
    assert np.array_equal(ni1, a[order, :, 0]) and np.array_equal(ni2, a[order, :, 1])
    assert np.array_equal(eps, a[order, :, 2])


@pytest.mark.parametrize("mech", ["TSR-GP-24", "TSR-Rich-31"])
def test_general_oracle_matches_reference_outputs(oracle_mod, mechs, mech):
    md = mechs[mech]
    g = golden(f"general_{mech}.npz")
    st = S.radical_states(md.species, md.W, int(g["seed"]), int(g["start"]), int(g["count"]), ignition_half=True)
    o = oracle_mod.Oracle(md)
    w, sc = o.general(S.soa(st["rho"]), st["T"])
    assert np.array_equal(w, g["wdot"])
    we, sce = o.wdot(S.soa(st["rho"]), st["T"])
    d = np.max(np.abs(w - we) / np.maximum(sc, 1e-300))
    assert 1e-12 < d < 1e-6


def test_general_oracle_matches_reference_outputs_pelucchi(oracle_mod):
    from flint_amd.mechdata import mech_data
    g = golden("general_Pelucchi.npz")
    w, _ = oracle_mod.Oracle(mech_data("Pelucchi")).general(g["rho"], g["T"])
    assert np.array_equal(w, g["wdot"])


def test_general_oracle_half_integer_coefficients(oracle_mod):
    """WD through general: its tables hold the stoichiometric coefficients 0.5 and 1.5, raised to
    nint() powers (Lib_Chemistry_wdot.f90:140-141), which round halves away from zero
    (nint(0.5) = 1).  The reference's outputs (tests/golden/general_WD.npz, CO seeded so that
    CO + 0.5 O2 runs) are matched bit for bit; rounding halves to even (0.5 -> 0) misses them by
    O(1) normwise."""
    from flint_amd.mechdata import mech_data
    g = golden("general_WD.npz")
    w, sc = oracle_mod.Oracle(mech_data("WD")).general(g["rho"], g["T"])
    assert np.array_equal(w, g["wdot"])
    assert np.all(np.abs(g["wdot"]).max(1) > 0)


@pytest.mark.skipif(not os.path.exists(REF_LIB), reason="reference build (make -C oracle ref) absent")
def test_general_oracle_equals_reference_build(oracle_mod, mechs):
    from flint_amd.mechdata import mech_data
    for mech in ("TSR-GP-24", "TSR-Rich-31"):
        md = mechs[mech]
        st = S.radical_states(md.species, md.W, 400, 0, 3000, T_lo=600.0, T_hi=4000.0, ignition_half=True)
        rho = S.soa(st["rho"])
        w, _ = oracle_mod.Oracle(md).general(rho, st["T"])
        assert np.array_equal(w, oracle_mod.Reference(md).general(rho, st["T"]))
    # Pelucchi: Arrhenius, Troe and Lindemann reactions, no compiled kernels here
    md = mech_data("Pelucchi")
    T = 300.0 + 2600.0 * (np.arange(2048) % 997) / 997.0
    rho = 10.0 ** np.random.default_rng(5).uniform(-9, -1, (md.ns, 2048))
    w, _ = oracle_mod.Oracle(md).general(rho, T)
    assert np.all(np.isfinite(w)) and np.array_equal(w, oracle_mod.Reference(md).general(rho, T))
