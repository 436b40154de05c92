"""Static FP64 operation counts of the reactor's units of work, from a mechanism's rate program.

The counts are those of the algorithm, one flop per addition, subtraction, multiplication,
division or square root, with the transcendental functions (log10, 10**x) counted apart:

* right-hand side (per evaluation): concentrations c_k = rho Y_k / W_k, the third-body sums
  (TSR-GP-24.f90:28 and its siblings), per reaction the table interpolations a + (b - a) Tdiff
  (Lib_Chemistry_data.f90:36-48, three flops per coefficient), the products of concentrations and
  the falloff blending (Lib_Chemistry_falloff.f90:36-80), the net production sums in reaction
  order, then the energy equation (constant pressure: R_mix, rho, c_p, h, sum h_k wdot_k);
* analytic Jacobian (per evaluation): the derivative of every production term with respect to
  each of its concentration factors and its third body, the species scaling, the
  constant-pressure change of variables and the temperature row;
* inverse of I - cJ by Gauss-Jordan elimination: forming the matrix (2 n^2) and the elimination
  (n steps, each one division, n - 1 scalings and (n - 1) (n - 1) multiply-adds: about 2 n^3);
* Newton correction: one matrix-vector product (2 n^2) plus the iterate update and its RMS norm;
* step: predictor, psi, error estimate and difference-array update over the components.
"""
from dataclasses import dataclass


@dataclass
class ReactorFlops:
    rhs: int          # per right-hand side evaluation (arithmetic)
    rhs_trans: int    # per right-hand side evaluation (log10 and 10**x calls)
    jac: int          # per analytic Jacobian
    inv: int          # per Gauss-Jordan inverse of I - cJ
    newton: int       # per Newton correction (solve + update + norm)
    step: int         # per accepted or rejected step attempt (predictor, error test, D update)

    def total(self, rhs, jac, lu, newton, steps):
        """Flops of a run from its per-cell counters (arrays or totals)."""
        return (rhs * self.rhs + jac * self.jac + lu * self.inv + newton * self.newton + steps * self.step)


def reactor_flops(prog, mode=1):
    ns = prog.ns
    n = ns + 1
    # concentrations (constant pressure: R_mix = sum Y_k R_k, rho = p / (R_mix T), rho_k = rho Y_k)
    rhs = ns if mode == 0 else 2 * ns + 2 + ns + ns
    for m in prog.m_sets:
        if m == "sum":
            rhs += ns - 1
        else:
            rhs += (len(m) - 1) + sum(1 for _, e in m if e is not None)
    trans = 0
    for r in prog.reactions:
        nf = sum(p for _, p in r.fwd)
        nb = sum(p for _, p in r.rev)
        if r.kind in ("arrhenius", "three-body"):
            rhs += 6 + nf + nb + (2 if r.kind == "three-body" else 0)
        elif r.kind == "troe":
            # 4 interpolations, Pr, Pr/(1+Pr), c, n, f1, the exponent, k_inf g F, k_f / k_c
            rhs += 12 + 18 + nf + nb
            trans += 3
        else:
            rhs += 9 + 6 + nf + nb
        rhs += 1  # q = prodf - prodb
    for terms in prog.wdot_terms:
        rhs += 2 * len(terms) + (1 if terms else 0)
    # thermo at T (cp, h) and the energy equation
    rhs += 6 * ns
    rhs += (4 * ns + 3 + ns) if mode == 1 else (2 * ns + 2 * ns + 3 * ns + 2)
    # Jacobian: per production term, d/dc of each factor (coefficient times the other factors,
    # accumulated) and, with a third body, one column per species; dq/dT per term
    jac = 0
    for k, terms in enumerate(prog.wdot_terms):
        for j, _ in terms:
            r = prog.reactions[j]
            nf = sum(p for _, p in r.fwd)
            nb = sum(p for _, p in r.rev)
            jac += 3 * (nf + nb) + 2
            if r.m_set >= 0:
                jac += 2 * ns
    jac += 2 * ns * ns  # species scaling W_k / W_j
    if mode == 1:
        jac += 2 * ns * ns + 6 * ns * ns + 4 * ns + 2 * ns * ns + 8 * n
    else:
        jac += 2 * ns * ns + 6 * n
    inv = 2 * n * n + n * (1 + (n - 1) + 2 * (n - 1) * (n - 1))
    newton = 2 * n * n + 3 * n + 2 * n + 3 * n + 2
    step = 5 * n + 10 * n + 3 * n + 20 * n
    return ReactorFlops(rhs=rhs, rhs_trans=trans, jac=jac, inv=inv, newton=newton, step=step)
