"""HIP code generator: emits one gfx950 source file per mechanism from its compiled rate program.

The emitted code is straight-line FP64 with compile-time stoichiometry, so every concentration,
third-body sum and accumulator lives in VGPRs.  One thread owns one cell; cells are stored
structure-of-arrays (cell fastest) so every state load and every output store of a wavefront is
one contiguous 512-byte segment.

Emitted per mechanism (namespace ``flint_gen_<id>``):
  * ``eval<JAC, CLIP, Sink>``   rates of progress, net production rates, and with JAC the analytic
                                 Jacobian of the constant-volume RHS, written through ``sink``;
  * ``wdot_kernel``              chemistry_source over a batch (Lib_Chemistry_wdot.f90:6-20);
  * ``rhs_jac_kernel``           rhs_native (Lib_Chemistry_rhs.f90:15-52) and its Jacobian.

Arithmetic order follows the reference routine (see flint_amd/mechanism.py); the Jacobian is
the exact derivative of the table-interpolated model (the reference has none).
"""
import os
import re

from .mechanism import compile_mechanism, RUNIV

GEN_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "gen")


def ident(name):
    return "".join(ch if ch.isalnum() else "_" for ch in name).lower()


def _lit(x):
    r = repr(float(x))
    if "e" not in r and "." not in r and "inf" not in r and "nan" not in r:
        r += ".0"
    return r


def _factor_expr(fac):
    """'(c3)*(c5*c5)' as emitted by YTF for [(3,1),(5,2)]."""
    parts = []
    for k, n in fac:
        parts.append("(" + "*".join([f"c{k}"] * n) + ")")
    return parts


def _dprod_expr(fac, j):
    """d/dc_j of prod c_k^n_k (None when j is absent)."""
    terms = []
    found = False
    for k, n in fac:
        if k == j:
            found = True
            if n == 1:
                continue
            terms.append(_lit(float(n)))
            terms.extend([f"c{k}"] * (n - 1))
        else:
            terms.extend([f"c{k}"] * n)
    if not found:
        return None
    return "*".join(terms) if terms else "1.0"


# Launch shape of the source-term/Jacobian kernel: threads per block, the waves per SIMD the
# compiler is held to (0: its choice) and how many per-cell scratch values live in LDS
JAC_OPTS = dict(block=128, waves=0, lds_slots=1 << 30, cache_slots=80)
# Reaction blocks whose kinetics-table loads go out together (_hoist_loads), per mechanism: measured
# on the source-term/Jacobian kernel (GP-24 over 2^20 cells: 2.86 ms one block at a time, 2.22 ms
# by 4, 2.04 by 8, 1.96 by 16, 2.30 by 32 with spills; Rich-31 over 2^18: 1.45, 1.17, 1.24, 1.24)
LOAD_GROUP = {"TSR-GP-24": 16, "TSR-Rich-31": 4}


_LOAD = re.compile(r"^(\s*)(\{ )?const double2 a0 = ld2\(rk0 \+ (\d+)\), a1 = ld2\(rk1 \+ (\d+)\);$")


def _block_end(lines, b0):
    """Index of the line closing the block opened at lines[b0] (a "}" at the same indent)."""
    ind = len(lines[b0]) - len(lines[b0].lstrip(" "))
    k = b0 + 1
    while k < len(lines) and lines[k] != " " * ind + "}":
        k += 1
    return k if k < len(lines) else None


def _hoist_loads(lines, group, counter=None):
    """Table loads ahead of their use: in each run of consecutive reaction blocks (the rates pass
    and every Jacobian column), the kinetics-table pairs of `group` blocks are loaded together
    before the blocks, which then alias them.  Emitted where each block loads its own pair, the
    loads go out one at a time, each waited for before the next: the compiler keeps the order of
    the straight-line code.  Only the placement of the loads changes, not a value."""
    counter = counter if counter is not None else [0]

    def is_rxn(ln):
        st = ln.strip()
        return st.startswith("{  // reaction") or st == "{" or bool(_LOAD.match(ln) and st.startswith("{"))

    out, i = [], 0
    while i < len(lines):
        ln = lines[i]
        st = ln.strip()
        if st.startswith("{") and not st.endswith("}") and is_rxn(ln):
            ind = len(ln) - len(ln.lstrip(" "))
            runs, j = [], i
            while j < len(lines):
                lj = lines[j]
                if not (is_rxn(lj) and len(lj) - len(lj.lstrip(" ")) == ind and not lj.strip().endswith("}")):
                    break
                k = _block_end(lines, j)
                if k is None:
                    break
                runs.append((j, k))
                j = k + 1
            if not runs:
                out.append(ln)
                i += 1
                continue
            for c0 in range(0, len(runs), group):
                hoisted, blocks = [], []
                for b0, b1 in runs[c0:c0 + group]:
                    blk = list(lines[b0:b1 + 1])
                    for t in range(min(2, len(blk))):
                        m = _LOAD.match(blk[t])
                        if m:
                            counter[0] += 1
                            n = counter[0]
                            hoisted.append(f"{' ' * ind}const double2 A0_{n} = ld2(rk0 + {m.group(3)}), "
                                           f"A1_{n} = ld2(rk1 + {m.group(4)});")
                            blk[t] = f"{m.group(1)}{m.group(2) or ''}const double2 a0 = A0_{n}, a1 = A1_{n};"
                            break
                    blocks.extend(blk)
                out.extend(hoisted)
                out.extend(blocks)
            i = runs[-1][1] + 1
        elif st.startswith("{") and not st.endswith("}"):
            k = _block_end(lines, i)
            if k is None:
                out.append(ln)
                i += 1
                continue
            out.append(ln)
            out.extend(_hoist_loads(lines[i + 1:k], group, counter))
            out.append(lines[k])
            i = k + 1
        else:
            out.append(ln)
            i += 1
    return out


def emit(prog):
    mid = ident(prog.name)
    ns = prog.ns
    nz = ns + 1
    na, nt, nl = prog.n_arrh, prog.n_troe, prog.n_lind
    wk = 2 * na + 4 * nt + 3 * nl
    wk += wk % 2
    W = prog.W
    parts = {r.index: [] for r in prog.reactions}
    for k, terms in enumerate(prog.wdot_terms):
        for j, c in terms:
            parts[j].append((k, c))
    first_term = {}
    for k, terms in enumerate(prog.wdot_terms):
        if terms:
            first_term[k] = terms[0][0]
    # per-cell scratch slots (LDS in the Jacobian kernel): dq/dM of every third-body/falloff
    # reaction, the two falloff rate coefficients, and e_k = h_k - R_k T
    slot = {}
    for r in prog.reactions:
        if r.m_set >= 0:
            slot[("D", r.index)] = len(slot)
        if r.kind in ("troe", "lindemann"):
            slot[("K1", r.index)] = len(slot)
            slot[("K2", r.index)] = len(slot)
    # slots past the LDS budget stay in registers
    keep = list(slot)[:JAC_OPTS["lds_slots"]]
    reg = [k for k in slot if k not in keep]
    slot = {k: i for i, k in enumerate(keep)}
    # up to cache_slots in all: the interpolated k_f, k_b of the Arrhenius/three-body reactions
    # that enter the most Jacobian columns, kept from the rates pass so that those columns skip
    # the reaction's two table-row gathers
    def columns(r):
        return len({k for k, _ in r.fwd} | ({k for k, _ in r.rev} if r.reversible else set()))
    cands = sorted((r for r in prog.reactions if r.kind in ("arrhenius", "three-body") and parts[r.index]),
                   key=lambda r: -columns(r))
    cached = set()
    for r in cands:
        if len(slot) + 2 > JAC_OPTS["cache_slots"]:
            break
        slot[("C1", r.index)] = len(slot)
        slot[("C2", r.index)] = len(slot)
        cached.add(r.index)
    nslot = len(slot)

    def S(kind, i):
        if kind == "e":
            return f"e{i}"
        if (kind, i) not in slot:
            return f"v{kind}{i}"
        return f"scr[{slot[(kind, i)]} * SST]"

    L = []
    a = L.append
    a(f"// Generated by flint_amd/codegen.py from the {prog.name} mechanism descriptor. Do not edit.")
    a(f"// {ns} species, {len(prog.reactions)} reactions ({na} Arrhenius/three-body, {nt} Troe, {nl} Lindemann).")
    a('#include "../flint_device.h"')
    a(f"namespace flint_gen_{mid} {{")
    a(f"constexpr int NS = {ns}, NZ = {nz}, WK = {wk}, WH = {2 * ns}, NSLOT = {nslot};")
    a("")
    a("// Rates, net production and (JAC) the analytic Jacobian of the constant-volume RHS for one cell.")
    a("// r[k] = rho_k, T; outputs F[0..NS-1] = wdot, F[NS] = dT/dt (rhs_native); sink(row, col, v).")
    a("// RHS applies rho = max(rho, 0) as rhs_native does (Lib_Chemistry_rhs.f90:37-39) and adds dT/dt.")
    a("// With JAC, scr[s*SST] (s < NSLOT) is per-cell scratch for values reused by the columns.")
    a("// Returns a status: bit 0 T outside the kinetics table, bit 1 T outside the thermo table.")
    a("// With JAC, fout(F) receives F as soon as it is final, so that F need not stay live across the")
    a("// Jacobian columns.")
    a("template <bool JAC, bool RHS, int SST, class Sink, class FOut>")
    a("__device__ __forceinline__ int eval(const double* __restrict__ tabk, int Tk_min, int Tk_max,")
    a("                                    const double* __restrict__ tabh, int Th_min, int Th_max,")
    a("                                    const double (&r)[NS], double T, double (&F)[NZ],")
    a("                                    double* __restrict__ scr, Sink&& sink, FOut&& fout) {")
    a("  int status = 0;")
    a("  if (!(T >= (double)Tk_min && T <= (double)Tk_max)) status |= 1;")
    a("  if (RHS && !(T >= (double)Th_min && T <= (double)Th_max)) {")
    a("    status |= 2;")
    a("#pragma unroll")
    a("    for (int k = 0; k < NZ; ++k) F[k] = -1.0;")
    a("    if (JAC) fout(F);")
    a("    if (JAC) {")
    a("      for (int x = 0; x < NZ; ++x)")
    a("        for (int y = 0; y < NZ; ++y) sink(x, y, 0.0);")
    a("    }")
    a("    return status;")
    a("  }")
    a("  unsigned negmask = 0u;")
    for k in range(ns):
        a(f"  const double r{k} = RHS ? fmax(r[{k}], 0.0) : r[{k}];")
    for k in range(ns):
        a(f"  if (r[{k}] < 0.0) negmask |= (1u << {k});" if ns <= 32 else "")
    for k in range(ns):
        a(f"  const double c{k} = r{k} / {_lit(W[k])};")
    a("  const int T_i = (int)T;")
    a("  const double Td = T - (double)T_i;")
    a("  const int rowk = min(max(T_i, Tk_min), Tk_max) - Tk_min;")
    a("  const double* __restrict__ rk0 = tabk + (long)rowk * WK;")
    a("  const double* __restrict__ rk1 = rk0 + WK;")
    # M sums
    for s, m in enumerate(prog.m_sets):
        if m == "sum":
            expr = "+".join(f"c{k}" for k in range(ns))
        else:
            terms = []
            for k, e in m:
                terms.append(f"c{k}" if e is None else f"c{k}*{_lit(e)}")
            expr = "+".join(terms)
        a(f"  const double M{s} = {expr};")
    for k in range(ns):
        if first_term.get(k) is None:
            a(f"  double w{k} = 0.0;")
        else:
            a(f"  double w{k};")
    if reg:
        a(f"  [[maybe_unused]] double {', '.join(f'v{k}{i}' for k, i in reg)};")
    a("  // JAC: dT-derivative accumulators")
    a("  [[maybe_unused]] double dT[NS] = {};")

    # ---------------- pass 1: rates in reaction order ----------------
    for r in prog.reactions:
        a(f"  {{  // reaction {r.index + 1}: {r.eq}")
        pfac = _factor_expr(r.fwd)
        bfac = _factor_expr(r.rev)
        Pf = "*".join(pfac) if pfac else "1.0"
        Pb = "*".join(bfac) if bfac else "1.0"
        if r.kind in ("arrhenius", "three-body"):
            a(f"    const double2 a0 = ld2(rk0 + {2 * r.table}), a1 = ld2(rk1 + {2 * r.table});")
            a("    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;")
            if r.index in cached:
                a(f"    if (JAC) {{ {S('C1', r.index)} = kf; {S('C2', r.index)} = kb; }}")
            mf = f"*M{r.m_set}" if r.kind == "three-body" else ""
            a(f"    const double pf = kf{''.join('*' + p for p in pfac)}{mf};")
            if r.reversible:
                a(f"    const double pb = kb{''.join('*' + p for p in bfac)}{mf};")
            else:
                a("    const double pb = 0.0;  // irreversible: kb table is identically 0")
            a("    const double q = pf - pb;")
            a("    if (JAC) {")
            a(f"      const double Pf = {Pf}, Pb = {Pb};")
            kbs = "(a1.y - a0.y) * Pb" if r.reversible else "0.0"
            if r.kind == "three-body":
                a(f"      const double dqT = ((a1.x - a0.x) * Pf - {kbs}) * M{r.m_set};")
                a(f"      {S('D', r.index)} = kf * Pf - {'kb * Pb' if r.reversible else '0.0'};")
            else:
                a(f"      const double dqT = (a1.x - a0.x) * Pf - {kbs};")
            for k, c in parts[r.index]:
                a(f"      dT[{k}] += {_lit(c)} * dqT;")
            a("    }")
        else:
            troe = r.kind == "troe"
            base = 2 * na + (4 * r.table if troe else 4 * nt + 3 * r.table)
            if troe:
                a(f"    const double2 x0 = ld2(rk0 + {base}), y0 = ld2(rk0 + {base + 2});")
                a(f"    const double2 x1 = ld2(rk1 + {base}), y1 = ld2(rk1 + {base + 2});")
                a("    const double kinf = x0.x + (x1.x - x0.x) * Td, k0 = x0.y + (x1.y - x0.y) * Td;")
                a("    const double kc = y0.x + (y1.x - y0.x) * Td, Fc = y0.y + (y1.y - y0.y) * Td;")
            else:
                a(f"    const double kinf0 = rk0[{base}], k00 = rk0[{base + 1}], kc0 = rk0[{base + 2}];")
                a(f"    const double kinf1 = rk1[{base}], k01 = rk1[{base + 1}], kc1 = rk1[{base + 2}];")
                a("    const double kinf = kinf0 + (kinf1 - kinf0) * Td, k0 = k00 + (k01 - k00) * Td;")
                a("    const double kc = kc0 + (kc1 - kc0) * Td;")
            a(f"    const double Mx = M{r.m_set};")
            a("    const double Pr = k0 * Mx / kinf;")
            if troe:
                a("    const double lFc = log10(Fc), lPr = log10(Pr);")
                a("    const double cc = -0.4 - 0.67 * lFc, nn = 0.75 - 1.27 * lFc;")
                a("    const double f1 = (lPr + cc) / (nn - 0.14 * (lPr + cc));")
                a("    const double Fb = exp10(lFc / (1.0 + f1 * f1));")
                a("    const double k1 = kinf * (Pr / (1.0 + Pr)) * Fb;")
            else:
                a("    const double k1 = kinf * (Pr / (1.0 + Pr));")
            a("    const double k2 = k1 / kc;")
            a(f"    const double pf = k1{''.join('*' + p for p in pfac)};")
            a(f"    const double pb = k2{''.join('*' + p for p in bfac)};")
            a("    const double q = pf - pb;")
            a("    if (JAC) {")
            a(f"      const double Pf = {Pf}, Pb = {Pb};")
            a("      const double g = Pr / (1.0 + Pr), dg = 1.0 / ((1.0 + Pr) * (1.0 + Pr));")
            if troe:
                a("      const double u = lPr + cc, den = nn - 0.14 * u, qq = 1.0 + f1 * f1;")
                a("      const double dlnF_dPr = -2.0 * lFc * f1 * (nn / (den * den)) / (qq * qq * Pr);")
                a("      const double df1_dL = (-0.67 * nn + 1.27 * u) / (den * den);")
                a("      const double dlnF_dFc = (1.0 / qq - lFc * 2.0 * f1 * df1_dL / (qq * qq)) / Fc;")
                a("      const double dk1_dPr = kinf * Fb * (dg + g * dlnF_dPr);")
                a("      const double dPr_dT = Mx * ((x1.y - x0.y) * kinf - k0 * (x1.x - x0.x)) / (kinf * kinf);")
                a("      const double dk1T = (x1.x - x0.x) * g * Fb + dk1_dPr * dPr_dT + k1 * dlnF_dFc * (y1.y - y0.y);")
                a("      const double dk2T = dk1T / kc - k1 * (y1.x - y0.x) / (kc * kc);")
            else:
                a("      const double dk1_dPr = kinf * dg;")
                a("      const double dPr_dT = Mx * ((k01 - k00) * kinf - k0 * (kinf1 - kinf0)) / (kinf * kinf);")
                a("      const double dk1T = (kinf1 - kinf0) * g + dk1_dPr * dPr_dT;")
                a("      const double dk2T = dk1T / kc - k1 * (kc1 - kc0) / (kc * kc);")
            a("      const double dk1M = dk1_dPr * k0 / kinf;")
            a(f"      {S('D', r.index)} = dk1M * Pf - (dk1M / kc) * Pb;")
            a(f"      {S('K1', r.index)} = k1; {S('K2', r.index)} = k2;")
            a("      const double dqT = dk1T * Pf - dk2T * Pb;")
            for k, c in parts[r.index]:
                a(f"      dT[{k}] += {_lit(c)} * dqT;")
            a("    }")
        for k, c in parts[r.index]:
            if first_term[k] == r.index:
                a(f"    w{k} = {_lit(c)} * q;")
            else:
                a(f"    w{k} = w{k} + {_lit(c)} * q;")
        a("  }")
    for k in range(ns):
        a(f"  F[{k}] = {_lit(W[k])} * w{k};")
    a("  if (!RHS) return status;")
    # ---------------- thermo: dT/dt ----------------
    a("  const int rowh = T_i - Th_min;")
    a("  const double* __restrict__ rh0 = tabh + (long)rowh * WH;")
    a("  const double* __restrict__ rh1 = rh0 + WH;")
    a("  double eiroi = 0.0, rho_cv = 0.0;")
    a("  [[maybe_unused]] double s_de = 0.0, s_dcv = 0.0;")
    a(f"  [[maybe_unused]] double {', '.join(f'e{k}' for k in range(ns))};")
    for k in range(ns):
        Rk = RUNIV / W[k]
        a(f"  {{ const double2 t0 = ld2(rh0 + {2 * k}), t1 = ld2(rh1 + {2 * k});")
        a("    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;")
        a(f"    const double ek = hk - {_lit(Rk)} * T;")
        a(f"    eiroi = eiroi + ek * F[{k}];")
        a(f"    rho_cv = rho_cv + r{k} * (cpk - {_lit(Rk)});")
        a(f"    if (JAC) {{ {S('e', k)} = ek; s_de += ((t1.y - t0.y) - {_lit(Rk)}) * F[{k}]; s_dcv += r{k} * (t1.x - t0.x); }} }}")
    a("  F[NS] = -eiroi / rho_cv;")
    a("  if (JAC) fout(F);")
    a("  if (!JAC) return status;")
    a("  const double Tdot = F[NS], invD = 1.0 / rho_cv;")
    a("  const double* const rk0_ = rk0;")
    # ---------------- T column ----------------
    a("  {  // column NS: d/dT")
    a("    double acc = s_de;")
    for k in range(ns):
        a(f"    {{ const double v = {_lit(W[k])} * dT[{k}]; sink({k}, NS, v); acc += {S('e', k)} * v; }}")
    a("    sink(NS, NS, -(acc + Tdot * s_dcv) * invD);")
    a("  }")
    # ---------------- species columns ----------------
    inv_parts = parts
    for j in range(ns):
        rows = set()
        body = []
        for r in prog.reactions:
            fd = _dprod_expr(r.fwd, j)
            bd = _dprod_expr(r.rev, j) if r.reversible else None
            meff = None
            if r.m_set >= 0:
                m = prog.m_sets[r.m_set]
                if m == "sum":
                    meff = 1.0
                else:
                    for k, e in m:
                        if k == j:
                            meff = 1.0 if e is None else e
            if fd is None and bd is None and meff is None:
                continue
            if not parts[r.index]:
                continue
            terms = []
            if fd is not None or bd is not None:
                if r.kind in ("arrhenius", "three-body") and r.index in cached:
                    body.append("    {")
                    mf = f" * M{r.m_set}" if r.kind == "three-body" else ""
                    if fd is not None:
                        terms.append(f"{S('C1', r.index)}{mf} * {fd}")
                    if bd is not None:
                        terms.append(f"-{S('C2', r.index)}{mf} * {bd}")
                elif r.kind in ("arrhenius", "three-body"):
                    body.append(f"    {{ const double2 a0 = ld2(rk0 + {2 * r.table}), a1 = ld2(rk1 + {2 * r.table});")
                    mf = f" * M{r.m_set}" if r.kind == "three-body" else ""
                    if fd is not None:
                        terms.append(f"(a0.x + (a1.x - a0.x) * Td){mf} * {fd}")
                    if bd is not None:
                        terms.append(f"-(a0.y + (a1.y - a0.y) * Td){mf} * {bd}")
                else:
                    body.append("    {")
                    if fd is not None:
                        terms.append(f"{S('K1', r.index)} * {fd}")
                    if bd is not None:
                        terms.append(f"-{S('K2', r.index)} * {bd}")
            else:
                body.append("    {")
            if meff is not None:
                terms.append(f"{_lit(meff)} * {S('D', r.index)}" if meff != 1.0 else S('D', r.index))
            expr = " + ".join(terms).replace("+ -", "- ")
            body.append(f"      const double dq = {expr};")
            for k, c in parts[r.index]:
                rows.add(k)
                body.append(f"      jc{k} += {_lit(c)} * dq;")
            body.append("    }")
        a(f"  {{  // column {j}: d/drho_{j} ({prog.species[j]})")
        a("    const double* __restrict__ rk0 = opaque(rk0_);")
        a("    const double* __restrict__ rk1 = rk0 + WK;")
        a(f"    const double gj = (negmask >> {j}) & 1u ? 0.0 : {_lit(1.0 / W[j])};")
        for k in sorted(rows):
            a(f"    double jc{k} = 0.0;")
        L.extend(body)
        a("    double acc = 0.0;")
        for k in range(ns):
            if k in rows:
                a(f"    {{ const double v = {_lit(W[k])} * gj * jc{k}; sink({k}, {j}, v); acc += {S('e', k)} * v; }}")
            else:
                a(f"    sink({k}, {j}, 0.0);")
        Rj = RUNIV / W[j]
        a(f"    const double cvj = (ld2(rh0 + {2 * j}).x + (ld2(rh1 + {2 * j}).x - ld2(rh0 + {2 * j}).x) * Td) - {_lit(Rj)};")
        a(f"    sink(NS, {j}, -(acc + Tdot * cvj * (((negmask >> {j}) & 1u) ? 0.0 : 1.0)) * invD);")
        a("  }")
    a("  return status;")
    a("}")
    a("")
    # ---------------- kernels ----------------
    a("// chemistry_source over a batch: rho[k*ld+i], T[i] -> wdot[k*ld+i] (no clipping, as the generated")
    a("// routines; Lib_ChemMech/*.f90).")
    a("__global__ __launch_bounds__(256) void wdot_kernel(long n, long ld, const double* __restrict__ rho,")
    a("    const double* __restrict__ T, double* __restrict__ wdot, int* __restrict__ status,")
    a("    const double* __restrict__ tabk, int Tk_min, int Tk_max) {")
    a("  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;")
    a("  if (i >= n) return;")
    a("  double r[NS], F[NZ];")
    a("#pragma unroll")
    a("  for (int k = 0; k < NS; ++k) r[k] = stream_load(rho + k * ld + i);")
    a("  const int st = eval<false, false, 1>(tabk, Tk_min, Tk_max, nullptr, 0, 0, r, stream_load(T + i), F, nullptr,")
    a("      NullSink{}, NullOut{});")
    a("#pragma unroll")
    a("  for (int k = 0; k < NS; ++k) stream_store(wdot + k * ld + i, F[k]);")
    a("  if (status) status[i] = st;")
    a("}")
    a("")
    a(f"constexpr int JAC_BLOCK = {JAC_OPTS['block']};")
    a("// rhs_native and its Jacobian over a batch: Z[k*ld+i] (rho_1..rho_NS, T) -> F[k*ld+i],")
    a("// J[(row*NZ+col)*ld+i] (J may be null: RHS only).")
    a("template <bool JAC>")
    wpe = JAC_OPTS["waves"]
    att = f" __attribute__((amdgpu_waves_per_eu({wpe}, {wpe})))" if wpe else ""
    a(f"__global__ __launch_bounds__(JAC_BLOCK){att} void rhs_jac_kernel(long n, long ld, const double* __restrict__ Z,")
    a("    double* __restrict__ Fo, double* __restrict__ J, int* __restrict__ status,")
    a("    const double* __restrict__ tabk, int Tk_min, int Tk_max, const double* __restrict__ tabh,")
    a("    int Th_min, int Th_max) {")
    a("  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;")
    a("  if (i >= n) return;")
    a("  double r[NS], F[NZ];")
    a("#pragma unroll")
    a("  for (int k = 0; k < NS; ++k) r[k] = stream_load(Z + k * ld + i);")
    a("  const double T = stream_load(Z + NS * ld + i);")
    a("  double* __restrict__ Ji = J + i;")
    a("  __shared__ double smem[JAC ? NSLOT * JAC_BLOCK : 1];")
    a("  const int st = eval<JAC, true, JAC_BLOCK>(tabk, Tk_min, Tk_max, tabh, Th_min, Th_max, r, T, F, smem + threadIdx.x,")
    a("      [&](int row, int col, double v) { stream_store(Ji + (long)(row * NZ + col) * ld, v); },")
    a("      [&](const double (&Fv)[NZ]) {")
    a("#pragma unroll")
    a("        for (int k = 0; k < NZ; ++k) stream_store(Fo + k * ld + i, Fv[k]);")
    a("      });")
    a("  if (!JAC) {")
    a("#pragma unroll")
    a("    for (int k = 0; k < NZ; ++k) stream_store(Fo + k * ld + i, F[k]);")
    a("  }")
    a("  if (status) status[i] = st;")
    a("}")
    # host copy of the rate program for the reaction-parallel reactor kernel
    kinds = {"arrhenius": 0, "three-body": 1, "troe": 2, "lindemann": 3}
    rx_kind, rx_col, rx_mset, rx_fac = [], [], [], []
    for r in prog.reactions:
        rx_kind.append(kinds[r.kind] | (4 if r.reversible else 0))
        if r.kind in ("arrhenius", "three-body"):
            rx_col.append(2 * r.table)
        elif r.kind == "troe":
            rx_col.append(2 * na + 4 * r.table)
        else:
            rx_col.append(2 * na + 4 * nt + 3 * r.table)
        rx_mset.append(r.m_set)
        for side in (r.fwd, r.rev):
            assert len(side) <= 4
            codes = [k | (n << 8) for k, n in side] + [-1] * (4 - len(side))
            rx_fac.extend(codes)
    wd_ptr, wd_rx, wd_coef = [0], [], []
    for terms in prog.wdot_terms:
        for j, c in terms:
            wd_rx.append(j)
            wd_coef.append(c)
        wd_ptr.append(len(wd_rx))
    ms_ptr, ms_sp, ms_eff = [0], [], []
    for m in prog.m_sets:
        items = [(k, 1.0) for k in range(ns)] if m == "sum" else [(k, 1.0 if e is None else e) for k, e in m]
        for k, e in items:
            ms_sp.append(k)
            ms_eff.append(e)
        ms_ptr.append(len(ms_sp))

    def arr(ty, name, vals):
        vals = vals or [0]
        body = ", ".join(_lit(v) if ty == "double" else str(int(v)) for v in vals)
        a(f"static const {ty} {name}[] = {{{body}}};")
    arr("int", "p_rx_kind", rx_kind)
    arr("int", "p_rx_col", rx_col)
    arr("int", "p_rx_mset", rx_mset)
    arr("int", "p_rx_fac", rx_fac)
    arr("int", "p_wd_ptr", wd_ptr)
    arr("int", "p_wd_rx", wd_rx)
    arr("double", "p_wd_coef", wd_coef)
    arr("int", "p_ms_ptr", ms_ptr)
    arr("int", "p_ms_sp", ms_sp)
    arr("double", "p_ms_eff", ms_eff)
    a(f"const flint_host_prog host_prog = {{{ns}, {len(prog.reactions)}, {len(prog.m_sets)}, {len(wd_rx)}, "
      f"{len(ms_sp)}, p_rx_kind, p_rx_col, p_rx_mset, p_rx_fac, p_wd_ptr, p_wd_rx, p_wd_coef, p_ms_ptr, p_ms_sp, "
      f"p_ms_eff}};")
    # element matrix of the equilibrium solver (species_composition, Lib_CEA_setup.f90:49-52) as a
    # compile-time table, so the per-species element sums fold at compile time
    comp = prog.desc.get("composition") or []
    if comp:
        ne = len(comp)
        rows = ", ".join("{" + ", ".join(repr(float(v)) for v in row) + "}" for row in comp)
        a("// element matrix for the equilibrium kernel: a(e, j) = atoms of element e in species j")
        a("struct Cea {")
        a(f"  static constexpr int NS = {ns}, NE = {ne};")
        a("  __host__ __device__ static constexpr double a(int e, int j) {")
        a(f"    constexpr double t[NE][NS] = {{{rows}}};")
        a("    return t[e][j];")
        a("  }")
        a("};")
        flat = ", ".join(repr(float(v)) for row in comp for v in row)
        a(f"static const double cea_A[] = {{{flat}}};")
    a(f"}}  // namespace flint_gen_{mid}")
    a("")
    if comp:
        a('#include "../flint_cea_kernel.h"')
        a(f"FLINT_REGISTER_MECH_CEA({mid}, flint_gen_{mid})")
    else:
        a(f"FLINT_REGISTER_MECH({mid}, flint_gen_{mid})")
    L = _hoist_loads(L, LOAD_GROUP.get(prog.name, 4))
    return "\n".join(L) + "\n"


def generate(names=("TSR-GP-24", "TSR-Rich-31")):
    os.makedirs(GEN_DIR, exist_ok=True)
    out = []
    for n in names:
        prog = compile_mechanism(n)
        path = os.path.join(GEN_DIR, ident(n) + ".hip")
        src = emit(prog)
        old = open(path).read() if os.path.exists(path) else None
        if old != src:
            with open(path, "w") as f:
                f.write(src)
        out.append(path)
    return out


if __name__ == "__main__":
    for p in generate():
        print(p)
