"""The CPU oracle against the reference's own outputs and against what the mathematics fixes."""
import numpy as np
import pytest

from conftest import golden, golden_states, normwise
from flint_amd import states as S
from flint_amd.mechanism import compile_mechanism


@pytest.mark.parametrize("mech", ["WD", "TSR-GP-24", "TSR-Rich-31"])
def test_oracle_matches_reference_fixtures(oracle_mod, mechs, mech):
    """wdot of the C restatement equals the flang-compiled reference routine (tests/golden)."""
    md = mechs[mech]
    g = golden(f"wdot_{mech}.npz")
    st = golden_states(md, g)
    w, sc = oracle_mod.Oracle(md).wdot(S.soa(st["rho"]), st["T"])
    assert normwise(w, g["wdot"], sc) <= 1e-14


def test_float32_literals_matter(oracle_mod, mechs):
    """Planted mistake: third-body efficiencies as exact doubles (the `general` routine's reading,
    Lib_Chemistry_wdot.f90:127-134) instead of float32 literals must fail the fixture check."""
    md = mechs["TSR-GP-24"]
    g = golden("wdot_TSR-GP-24.npz")
    st = golden_states(md, g)
    import copy
    bad = copy.copy(md)
    prog = compile_mechanism("TSR-GP-24")
    for i, m in enumerate(prog.m_sets):
        if m != "sum":
            prog.m_sets[i] = tuple((k, None if e is None else float(f"{e:.6g}")) for k, e in m)
    bad.prog = prog
    w, sc = oracle_mod.Oracle(bad).wdot(S.soa(st["rho"]), st["T"])
    assert normwise(w, g["wdot"], sc) > 1e-12


@pytest.mark.parametrize("mech", ["TSR-GP-24", "TSR-Rich-31"])
def test_mass_and_element_conservation(oracle_mod, mechs, mech):
    md = mechs[mech]
    st = S.radical_states(md.species, md.W, 5, 0, 2000)
    w, sc = oracle_mod.Oracle(md).wdot(S.soa(st["rho"]), st["T"])
    # sum_k wdot_k = 0 to round-off (every reaction conserves mass)
    assert np.max(np.abs(w.sum(axis=0)) / sc.sum(axis=0)) < 1e-14
    # element conservation: sum_k a_ek wdot_k / W_k = 0
    comp = np.array(md.desc["composition"])  # [ne, ns]
    el = comp @ (w / md.W[:, None])
    elsc = np.abs(comp) @ (sc / md.W[:, None])
    assert np.max(np.abs(el) / elsc) < 1e-14


def test_wd_closed_form(oracle_mod, mechs):
    """WD.f90:31-45 evaluated directly from its closed form in numpy."""
    md = mechs["WD"]
    st = S.wd_states(md.species, md.W, 3, 0, 500)
    rho, T = S.soa(st["rho"]), st["T"]
    w, _ = oracle_mod.Oracle(md).wdot(rho, T)
    c = np.maximum(rho, 0) / md.W[:, None]
    Ti = T.astype(int)
    f = T - Ti
    kf = [md.kf[Ti - md.Tk_min, r] + (md.kf[Ti + 1 - md.Tk_min, r] - md.kf[Ti - md.Tk_min, r]) * f for r in range(3)]
    p1 = kf[0] * c[0] ** np.float32(0.70).astype(float) * c[1] ** np.float32(0.80).astype(float)
    p2 = kf[1] * c[3] ** 0.5 * c[4] * c[1] ** 0.25
    p3 = kf[2] * c[2]
    ref = np.stack([-p1, -1.5 * p1 - 0.5 * p2 + 0.5 * p3, p2 - p3, 2 * p1, p1 - p2 + p3]) * md.W[:, None]
    assert np.allclose(w, ref, rtol=1e-13, atol=0)


@pytest.mark.parametrize("mech", ["TSR-GP-24", "TSR-Rich-31"])
def test_jacobian_matches_complex_step(oracle_mod, mechs, mech):
    md = mechs[mech]
    st = S.radical_states(md.species, md.W, 7, 0, 300, ignition_half=True)
    Z = np.vstack([S.soa(st["rho"]), st["T"][None, :]])
    o = oracle_mod.Oracle(md)
    F, J, Ja = o.jac(Z)
    Jc = o.jac_cstep(Z)
    assert normwise(J, Jc, Ja) < 1e-12
    assert np.array_equal(F, o.rhs(Z))


def test_jacobian_planted_mistake_is_caught(oracle_mod, mechs):
    """Planted mistake in the checker: a dropped Troe dF/dFcent term shows up against complex step."""
    md = mechs["TSR-GP-24"]
    st = S.radical_states(md.species, md.W, 8, 0, 200)
    Z = np.vstack([S.soa(st["rho"]), st["T"][None, :]])
    o = oracle_mod.Oracle(md)
    F, J, Ja = o.jac(Z)
    Jc = o.jac_cstep(Z)
    nz = md.ns + 1
    # corrupt the T column as a missing falloff-centre derivative would
    Jbad = J.copy().reshape(nz, nz, -1)
    Jbad[:, nz - 1, :] *= (1 + 1e-6)
    assert normwise(Jbad.reshape(nz * nz, -1), Jc, Ja) > 1e-9


@pytest.mark.parametrize("mech", ["TSR-GP-24", "TSR-Rich-31"])
def test_rhs_conserves_internal_energy(oracle_mod, mechs, mech):
    """Constant volume: d/dt sum_k rho_k e_k(T) = sum_k e_k wdot_k + Tdot sum_k rho_k cv_k = 0."""
    md = mechs[mech]
    st = S.radical_states(md.species, md.W, 9, 0, 500)
    rho, T = S.soa(st["rho"]), st["T"]
    Z = np.vstack([rho, T[None, :]])
    F = oracle_mod.Oracle(md).rhs(Z)
    Ti = T.astype(int)
    f = T - Ti
    r0, r1 = Ti - md.Th_min, Ti + 1 - md.Th_min
    h = md.h[r0] + (md.h[r1] - md.h[r0]) * f[:, None]
    cp = md.cp[r0] + (md.cp[r1] - md.cp[r0]) * f[:, None]
    R = 8314.51 / md.W
    e = (h - R[None, :] * T[:, None]).T
    cv = (cp - R[None, :]).T
    bal = (e * F[:-1]).sum(0) + F[-1] * (rho * cv).sum(0)
    scale = np.abs(e * F[:-1]).sum(0)
    assert np.max(np.abs(bal) / scale) < 1e-13


def test_rhs_out_of_range_temperature(oracle_mod, mechs):
    """Lib_Chemistry_rhs.f90:31-34: F = -1 when T is outside [Tmin, Tmax] or NaN."""
    md = mechs["TSR-GP-24"]
    st = S.radical_states(md.species, md.W, 9, 0, 3)
    Z = np.vstack([S.soa(st["rho"]), np.array([[md.Th_min - 1.0, md.Th_max + 1.0, np.nan]])])
    F = oracle_mod.Oracle(md).rhs(Z)
    assert np.all(F == -1.0)


def _cp_states(md, seed, n):
    """Constant-pressure states (Y_k, T) at p, with a few slightly negative Y (the model does not
    clip them) and temperatures close to both ends of the thermo table and of a 1 K bin."""
    st = S.radical_states(md.species, md.W, seed, 0, n, ignition_half=True)
    X = np.vstack([S.soa(st["Y"]), st["T"][None, :]])
    rng = np.random.default_rng(seed)
    neg = rng.random((md.ns, n)) < 0.05
    X[:md.ns][neg] = -1e-9 * rng.random(neg.sum())
    lo, hi = max(md.Th_min, md.Tk_min, 300), min(md.Th_max, md.Tk_max)
    X[-1, :4] = [lo + 0.25, hi - 1.0 + 1e-6, 1500.0 + 1e-9, 1500.0 + 1.0 - 1e-9]
    return X, st["p"]


@pytest.mark.parametrize("mech", ["TSR-GP-24", "TSR-Rich-31"])
def test_jac_cp_matches_complex_step(oracle_mod, mechs, mech):
    """The constant-pressure Jacobian (the reactor's Newton matrix) against the complex-step
    derivative of the constant-pressure right-hand side, at 1e-12 normwise: every column, the T
    column within the 1 K table bin (the step does not leave it)."""
    md = mechs[mech]
    X, p = _cp_states(md, 11, 400)
    o = oracle_mod.Oracle(md)
    F, J, Fa, Ja = o.jac_cp(X, p, scales=True)
    Jc = o.jac_cp_cstep(X, p)
    # entries that vanish structurally (species with no nonzero rate) come out exactly 0 here and
    # O(1e-50) from the complex step: a 1e-30 1/s floor on the scale
    assert normwise(J, Jc, Ja + 1e-30) < 1e-12
    assert np.array_equal(F, o.rhs_cp(X, p))
    # the scales bound the entries: |J| <= Ja, |F| <= Fa
    assert np.all(np.abs(J) <= Ja * (1 + 1e-12)) and np.all(np.abs(F) <= Fa * (1 + 1e-12))


def test_jac_cp_planted_mistakes_are_caught(oracle_mod, mechs):
    """Planted mistakes in the checker's formulas fail the complex-step comparison: the
    (wdot_k / rho) R_j / R_mix term of dF_k/dY_j dropped (the density's dependence on Y), and the
    sign of the 1/T term of dF_T/dT flipped."""
    md = mechs["TSR-GP-24"]
    X, p = _cp_states(md, 12, 200)
    o = oracle_mod.Oracle(md)
    F, J, Fa, Ja = o.jac_cp(X, p, scales=True)
    Jc = o.jac_cp_cstep(X, p)
    nz, ns = md.ns + 1, md.ns
    R = 8314.51 / md.W
    Rmix = (X[:ns] * R[:, None]).sum(0)
    bad = J.copy().reshape(nz, nz, -1)
    bad[:ns, :ns] -= F[:ns, None, :] * (R[None, :, None] / Rmix[None, None, :])
    assert normwise(bad.reshape(nz * nz, -1), Jc, Ja + 1e-30) > 1e-6
    bad = J.copy().reshape(nz, nz, -1)
    bad[ns, ns] += 2.0 * F[ns] / X[-1]
    assert normwise(bad.reshape(nz * nz, -1), Jc, Ja + 1e-30) > 1e-6


def _thermo(md, T):
    Ti = T.astype(int)
    f = (T - Ti)[:, None]
    r0, r1 = Ti - md.Th_min, Ti + 1 - md.Th_min
    h = md.h[r0] + (md.h[r1] - md.h[r0]) * f
    cp = md.cp[r0] + (md.cp[r1] - md.cp[r0]) * f
    return h.T, cp.T


@pytest.mark.parametrize("mech", ["TSR-GP-24", "TSR-Rich-31"])
def test_rhs_cp_closed_form(oracle_mod, mechs, mech):
    """The constant-pressure right-hand side from its definition, in numpy around the source term
    (which tests/golden pins to the reference routines): rho = p / (R_mix T) with
    R_mix = sum_k Y_k R_k, dY_k/dt = wdot_k(rho Y) / rho, dT/dt = -sum_k h_k wdot_k / (rho cp),
    cp = sum_k Y_k cp_k (tables interpolated in numpy).  A density off by 1 % (a wrong R_mix)
    fails the same comparison."""
    md = mechs[mech]
    X, p = _cp_states(md, 13, 500)
    X, p = X[:, 4:], p[4:]  # table-edge states: covered by the Jacobian test
    o = oracle_mod.Oracle(md)
    F = o.rhs_cp(X, p)
    Y, T = X[:-1], X[-1]
    R = 8314.51 / md.W

    def closed(rho):
        w, sc = o.wdot(rho[None, :] * Y, T)
        h, cp = _thermo(md, T)
        FT = -(h * w).sum(0) / (rho * (Y * cp).sum(0))
        FTs = (np.abs(h) * sc).sum(0) / (rho * (Y * cp).sum(0))
        return w / rho, sc / rho, FT, FTs

    rho = p / ((Y * R[:, None]).sum(0) * T)
    Fk, sk, FT, FTs = closed(rho)
    assert normwise(F[:-1], Fk, sk) < 1e-14
    assert np.max(np.abs(F[-1] - FT) / FTs) < 1e-13
    Fk2, _, _, _ = closed(1.01 * rho)
    assert normwise(F[:-1], Fk2, sk) > 1e-3
