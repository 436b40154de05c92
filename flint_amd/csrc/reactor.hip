// Batched stiff 0D reactor integration on gfx950: one persistent launch, one 32-lane team per
// cell, work-stealing refill from an atomic cell counter.
//
// Method: the variable-order (1..5) quasi-constant step BDF of Shampine & Reichelt ("The MATLAB
// ODE Suite", SIAM J. Sci. Comput. 18, 1997; kappa = 0), modified Newton with the analytic
// Jacobian refreshed only on Newton failure, RMS error norm of e/(atol + rtol|y|).  It is the same
// algorithm, decision for decision, as the CPU checker in oracle/oracle_bdf.c.  FLINT itself
// integrates through the external OSlo library (test-batchF.f90:81,105), absent from the
// reference checkout.
//
// Layout of one team (lanes l = 0..31 of a half wavefront):
//   lane k < ns owns species k, lane ns owns T: the state component and the 8-row BDF difference
//   column D[.][k] live in VGPRs; row k of the Jacobian sits in a per-team global slot and row k
//   of the LU factors of I - cJ in LDS (stride 33, conflict-free); partial pivoting picks rows by
//   a team argmax and every lane reads the pivot row as an LDS broadcast.
//   The right-hand side is evaluated reaction-parallel (reaction r on lane r mod 32) from the
//   kinetics table row of the cell's temperature (coalesced over the team), with the
//   net production rate of species k summed by lane k in reaction order (the summation order of
//   the reference routine), from per-reaction rates staged in LDS.
//   Modes: 0 = constant volume, state (rho_k, T), rhs_native (Lib_Chemistry_rhs.f90:15-52);
//          1 = constant pressure, state (Y_k, T) at fixed p.
#include "flint_device.h"
#include "flint_reactor.h"

namespace flint_reactor {

constexpr int TEAM = 32;
constexpr int MAXO = 5;
constexpr int NEWTON_MAXITER = 4;
constexpr double MIN_FACTOR = 0.2, MAX_FACTOR = 10.0;

enum { ST_STEPS, ST_REJECT, ST_NFAIL, ST_RHS, ST_JAC, ST_LU, ST_NEWTON, ST_ORDER };

// Team sum, bit-identical on every lane: each butterfly level adds the same two rounded values
// (a + b == b + a exactly).  The register barrier rounds v first; without it the compiler may
// contract a caller's product into the first addition, fma(x_i, y_i, shfl(x_j y_j)), which differs
// between partner lanes, and step sizes, orders and branch decisions would then diverge inside
// one team.
__device__ __forceinline__ double tsum(double v) {
  asm volatile("" : "+v"(v));
#pragma unroll
  for (int m = 16; m >= 1; m >>= 1) v += __shfl_xor(v, m, TEAM);
  return v;
}
__device__ __forceinline__ double tbc(double v, int src) { return __shfl(v, src, TEAM); }
__device__ __forceinline__ int tbci(int v, int src) { return __shfl(v, src, TEAM); }
// All LDS traffic of a team stays inside one wavefront: a wavefront-scope fence orders it.
__device__ __forceinline__ void tsync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

struct Lds {
  double* c;    // [32] concentrations
  double* Mv;   // [nms] third-body sums
  double* aux;  // [32] per-lane shared values (Y_k, e_k or h_k)
  double* g;    // [32] clip derivative per column
  double* q;    // [nr] rates of progress
  double* lu;   // [32][33] LU factors of I - cJ (row = lane, padded)
  double* kfe;  // [nr] effective forward coefficient (incl. M or falloff)   } alias lu while the
  double* kbe;  // [nr]                                                     } Jacobian is built
  double* Dm;   // [nr] dq/dM                                               } (the LU is rebuilt
  double* dqT;  // [nr] dq/dT at fixed concentrations                       }  right after)
  int* piv;     // [32] pivot lane of each elimination step
};

__host__ __device__ inline size_t team_lds_bytes(int nr, int nms) {
  return (size_t)(32 * 3 + (nms > 0 ? nms : 1) + (size_t)nr + 32 * 33) * sizeof(double) + 32 * sizeof(int);
}

__device__ inline Lds carve(char* base, int nr, int nms) {
  Lds L;
  double* d = reinterpret_cast<double*>(base);
  L.c = d; d += 32;
  L.aux = d; d += 32;
  L.g = d; d += 32;
  L.Mv = d; d += (nms > 0 ? nms : 1);
  L.q = d; d += nr;
  L.lu = d; d += 32 * 33;
  L.kfe = L.lu;
  L.kbe = L.lu + nr;
  L.Dm = L.lu + 2 * nr;
  L.dqT = L.lu + 3 * nr;
  L.piv = reinterpret_cast<int*>(d);
  return L;
}

struct RhsOut {
  bool ok;                  // T inside the tables (team-uniform)
  double F;                 // this lane's RHS component
  double xc, g, w, rho;     // clipped state, clip derivative, wdot_k, density
  double cp, h, dcp, dh;    // thermo of species k at T
  double Rmix, rho_cv, cpm; // team-uniform sums
};

// Rates of progress of reaction r (kinds: 0 Arrhenius, 1 three-body, 2 Troe, 3 Lindemann),
// Lib_Chemistry_falloff.f90:12-80 for the falloff forms.
template <bool JAC>
__device__ __forceinline__ void reaction(const Prog& P, int r, const double* __restrict__ rk0, const double* __restrict__ rk1,
                                         double Td, const Lds& L) {
  const int kinfo = P.rx_kind[r];
  const int kind = kinfo & 3;
  const int col = P.rx_col[r];
  const int ms = P.rx_mset[r];
  const double M = ms >= 0 ? L.Mv[ms] : 1.0;
  double Pf = 1.0, Pb = 1.0;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int code = P.rx_fac[r * 8 + s];
    if (code >= 0) {
      const double cv = L.c[code & 255];
      double f = cv;
      for (int e = 1; e < (code >> 8); ++e) f *= cv;
      Pf *= f;
    }
  }
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int code = P.rx_fac[r * 8 + 4 + s];
    if (code >= 0) {
      const double cv = L.c[code & 255];
      double f = cv;
      for (int e = 1; e < (code >> 8); ++e) f *= cv;
      Pb *= f;
    }
  }
  double k1, k2, dk1T = 0.0, dk2T = 0.0, dqM = 0.0;
  if (kind <= 1) {
    const double2 a0 = ld2(rk0 + col), a1 = ld2(rk1 + col);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double Mf = kind == 1 ? M : 1.0;
    k1 = kf * Mf;
    k2 = kb * Mf;
    if (JAC) {
      dk1T = (a1.x - a0.x) * Mf;
      dk2T = (a1.y - a0.y) * Mf;
      if (kind == 1) dqM = kf * Pf - kb * Pb;
    }
  } else {
    const bool troe = kind == 2;
    double kinf, k0, kc, Fc = 1.0, skinf, sk0, skc, sFc = 0.0;
    if (troe) {
      const double2 x0 = ld2(rk0 + col), y0 = ld2(rk0 + col + 2), x1 = ld2(rk1 + col), y1 = ld2(rk1 + col + 2);
      kinf = x0.x + (x1.x - x0.x) * Td; skinf = x1.x - x0.x;
      k0 = x0.y + (x1.y - x0.y) * Td; sk0 = x1.y - x0.y;
      kc = y0.x + (y1.x - y0.x) * Td; skc = y1.x - y0.x;
      Fc = y0.y + (y1.y - y0.y) * Td; sFc = y1.y - y0.y;
    } else {
      kinf = rk0[col] + (rk1[col] - rk0[col]) * Td; skinf = rk1[col] - rk0[col];
      k0 = rk0[col + 1] + (rk1[col + 1] - rk0[col + 1]) * Td; sk0 = rk1[col + 1] - rk0[col + 1];
      kc = rk0[col + 2] + (rk1[col + 2] - rk0[col + 2]) * Td; skc = rk1[col + 2] - rk0[col + 2];
    }
    const double Pr = k0 * M / kinf;
    const double gg = Pr / (1.0 + Pr);
    double Fb = 1.0, lFc = 0.0, f1 = 0.0, cc = 0.0, nn = 0.0, lPr = 0.0;
    if (troe) {
      lFc = log10(Fc);
      lPr = log10(Pr);
      cc = -0.4 - 0.67 * lFc;
      nn = 0.75 - 1.27 * lFc;
      f1 = (lPr + cc) / (nn - 0.14 * (lPr + cc));
      Fb = exp10(lFc / (1.0 + f1 * f1));
    }
    k1 = kinf * gg * Fb;
    k2 = k1 / kc;
    if (JAC) {
      const double dg = 1.0 / ((1.0 + Pr) * (1.0 + Pr));
      double dlnF_dPr = 0.0, dlnF_dFc = 0.0;
      if (troe) {
        const double u = lPr + cc, den = nn - 0.14 * u, qq = 1.0 + f1 * f1;
        dlnF_dPr = -2.0 * lFc * f1 * (nn / (den * den)) / (qq * qq * Pr);
        const double df1_dL = (-0.67 * nn + 1.27 * u) / (den * den);
        dlnF_dFc = (1.0 / qq - lFc * 2.0 * f1 * df1_dL / (qq * qq)) / Fc;
      }
      const double dk1_dPr = kinf * Fb * (dg + gg * dlnF_dPr);
      const double dPr_dT = M * (sk0 * kinf - k0 * skinf) / (kinf * kinf);
      dk1T = skinf * gg * Fb + dk1_dPr * dPr_dT + k1 * dlnF_dFc * sFc;
      dk2T = dk1T / kc - k1 * skc / (kc * kc);
      const double dk1M = dk1_dPr * k0 / kinf;
      dqM = dk1M * Pf - (dk1M / kc) * Pb;
    }
  }
  L.q[r] = k1 * Pf - k2 * Pb;
  if (JAC) {
    L.kfe[r] = k1;
    L.kbe[r] = k2;
    L.Dm[r] = dqM;
    L.dqT[r] = dk1T * Pf - dk2T * Pb;
  }
}

// Team right-hand side; o.ok is false (team-uniform) when T leaves the tables.  With jac the
// per-reaction values the Jacobian needs are staged in LDS as well.  Kept out of line: it is called
// from five places of the integrator and inlining every copy would multiply its register demand.
__device__ __forceinline__ RhsOut team_rhs(const Prog& P, int mode, double p, double x, const Lds& L, bool jac) {
  RhsOut o;
  o.ok = false;
  const int lane = threadIdx.x & (TEAM - 1);
  const int ns = P.ns;
  const double T = tbc(x, ns);
  if (!(T >= (double)P.Th_min && T < (double)P.Th_max && T >= (double)P.Tk_min && T < (double)P.Tk_max)) return o;
  const bool sp = lane < ns;
  const double Wk = sp ? P.W[lane] : 1.0;
  const double Rk = sp ? 8314.51 / Wk : 0.0;
  // constant volume clips rho at 0 as rhs_native does (Lib_Chemistry_rhs.f90:37-39); the
  // constant-pressure model keeps Y as it is, so its right-hand side stays smooth through Y = 0
  const bool clip = mode == 0 && x < 0.0;
  o.xc = sp ? (clip ? 0.0 : x) : 0.0;
  o.g = (sp && !clip) ? 1.0 : 0.0;
  double roi;
  if (mode == 0) {
    roi = o.xc;
    o.rho = 1.0;
    o.Rmix = 0.0;
  } else {
    o.Rmix = tsum(o.xc * Rk);
    o.rho = p / (o.Rmix * T);
    roi = o.rho * o.xc;
  }
  if (sp) L.c[lane] = roi / Wk;
  tsync();
#pragma unroll 1
  for (int s = lane; s < P.nms; s += TEAM) {
    const int t0 = P.ms_ptr[s], t1 = P.ms_ptr[s + 1];
    double m = L.c[P.ms_sp[t0]] * P.ms_eff[t0];
#pragma unroll 1
    for (int t = t0 + 1; t < t1; ++t) m = m + L.c[P.ms_sp[t]] * P.ms_eff[t];
    L.Mv[s] = m;
  }
  tsync();
  const int T_i = (int)T;
  const double Td = T - (double)T_i;
  const double* __restrict__ rk0 = P.tabk + (long)(T_i - P.Tk_min) * P.wk;
  const double* __restrict__ rk1 = rk0 + P.wk;
  if (jac) {
#pragma unroll 1
    for (int r = lane; r < P.nr; r += TEAM) reaction<true>(P, r, rk0, rk1, Td, L);
  } else {
#pragma unroll 1
    for (int r = lane; r < P.nr; r += TEAM) reaction<false>(P, r, rk0, rk1, Td, L);
  }
  tsync();
  double w = 0.0;
  if (sp) {
    const int t0 = P.wd_ptr[lane], t1 = P.wd_ptr[lane + 1];
    if (t1 > t0) {
      double acc = P.wd_coef[t0] * L.q[P.wd_rx[t0]];
#pragma unroll 1
      for (int t = t0 + 1; t < t1; ++t) acc = acc + P.wd_coef[t] * L.q[P.wd_rx[t]];
      w = Wk * acc;
    }
  }
  o.w = w;
  const double* __restrict__ rh0 = P.tabh + (long)(T_i - P.Th_min) * (2 * ns);
  const double* __restrict__ rh1 = rh0 + 2 * ns;
  o.cp = o.h = o.dcp = o.dh = 0.0;
  if (sp) {
    const double2 t0 = ld2(rh0 + 2 * lane), t1 = ld2(rh1 + 2 * lane);
    o.cp = t0.x + (t1.x - t0.x) * Td;
    o.h = t0.y + (t1.y - t0.y) * Td;
    o.dcp = t1.x - t0.x;
    o.dh = t1.y - t0.y;
  }
  double FT;
  if (mode == 0) {
    const double eiroi = tsum((o.h - Rk * T) * w);
    o.rho_cv = tsum(roi * (o.cp - Rk));
    FT = -eiroi / o.rho_cv;
    o.F = sp ? w : 0.0;
  } else {
    const double Q = tsum(o.h * w);
    o.cpm = tsum(o.xc * o.cp);
    FT = -Q / (o.rho * o.cpm);
    o.F = sp ? w / o.rho : 0.0;
  }
  if (lane == ns) o.F = FT;
  o.ok = true;
  return o;
}

// Jacobian of the team's cell, left in the team's global slot Jg (row k = lane k, stride 33),
// after team_rhs<true>.
__device__ __forceinline__ void team_jac(const Prog& P, int mode, double x, const Lds& L, const RhsOut o,
                                                   double* __restrict__ Jg) {
  const int lane = threadIdx.x & (TEAM - 1);
  const int ns = P.ns, nz = ns + 1;
  const bool sp = lane < ns;
  const double T = tbc(x, ns);
  const double FT = tbc(o.F, ns);
  double* myrow = Jg + lane * 33;
  for (int j = 0; j < TEAM; ++j) myrow[j] = 0.0;
  L.g[lane] = o.g;
  if (mode == 1) L.aux[lane] = o.xc;
  tsync();
  double bT = 0.0;
  const double Wk = sp ? P.W[lane] : 1.0;
  if (sp) {
    const int t0 = P.wd_ptr[lane], t1 = P.wd_ptr[lane + 1];
#pragma unroll 1
    for (int t = t0; t < t1; ++t) {
      const int r = P.wd_rx[t];
      const double coef = P.wd_coef[t];
      for (int side = 0; side < 2; ++side) {
        const double kk = side == 0 ? L.kfe[r] : -L.kbe[r];
        for (int s = 0; s < 4; ++s) {
          const int code = P.rx_fac[r * 8 + side * 4 + s];
          if (code < 0) continue;
          const int j = code & 255, pw = code >> 8;
          double d = kk * pw;
          for (int e = 1; e < pw; ++e) d *= L.c[j];
          for (int u = 0; u < 4; ++u) {
            const int cu = P.rx_fac[r * 8 + side * 4 + u];
            if (u == s || cu < 0) continue;
            for (int e = 0; e < (cu >> 8); ++e) d *= L.c[cu & 255];
          }
          myrow[j] += coef * d;
        }
      }
      const int ms = P.rx_mset[r];
      if (ms >= 0) {
        const double Dv = coef * L.Dm[r];
        const double* __restrict__ ef = P.effd + ms * TEAM;
        for (int j = 0; j < ns; ++j) myrow[j] += Dv * ef[j];
      }
      bT += coef * L.dqT[r];
    }
    // species block A_kj = d wdot_k / d rho_j (unclipped), b_k = d wdot_k / dT at fixed rho
    for (int j = 0; j < ns; ++j) myrow[j] *= Wk / P.W[j];
  }
  const double bk = Wk * bT;
  const double Rk = sp ? 8314.51 / Wk : 0.0;
  if (mode == 0) {
    // constant volume: J_kj = g_j A_kj, J_kT = b_k; T row from column sums of e_k J_kj
    if (sp) {
      for (int j = 0; j < ns; ++j) myrow[j] *= L.g[j];
      myrow[ns] = bk;
    }
    const double s_de = tsum(sp ? (o.dh - Rk) * o.w : 0.0);
    const double s_dcv = tsum(sp ? o.xc * o.dcp : 0.0);
    L.aux[lane] = sp ? o.h - Rk * T : 0.0;
    tsync();
    if (lane < nz) {
      double cs = 0.0;
      for (int k = 0; k < ns; ++k) cs += L.aux[k] * Jg[k * 33 + lane];
      const double jt = lane < ns ? -(cs + FT * (o.cp - Rk) * o.g) / o.rho_cv : -(s_de + cs + FT * s_dcv) / o.rho_cv;
      Jg[ns * 33 + lane] = jt;
    }
    tsync();
  } else {
    // constant pressure (Y, T): see oracle/flint_oracle.c jac_cp_cell for the formulas
    const double Rm = o.Rmix, rho = o.rho;
    double AY = 0.0;
    if (sp) {
      for (int j = 0; j < ns; ++j) AY += myrow[j] * L.aux[j];
      // stage dwdot_k/dY_j and dwdot_k/dT for the enthalpy-weighted column sums
      for (int j = 0; j < ns; ++j) myrow[j] = rho * L.g[j] * (myrow[j] - AY * (8314.51 / P.W[j]) / Rm);
      myrow[ns] = bk - rho * AY / T;
    }
    const double s_dhw = tsum(sp ? o.dh * o.w : 0.0);
    const double sYdcp = tsum(sp ? o.xc * o.dcp : 0.0);
    tsync();
    L.aux[lane] = sp ? o.h : 0.0;
    tsync();
    double jt = 0.0;
    if (lane < nz) {
      double ch = 0.0;
      for (int k = 0; k < ns; ++k) ch += L.aux[k] * Jg[k * 33 + lane];
      jt = lane < ns ? -ch / (rho * o.cpm) - FT * o.g * (o.cp / o.cpm - Rk / Rm)
                     : -(s_dhw + ch) / (rho * o.cpm) - FT * (sYdcp / o.cpm - 1.0 / T);
    }
    tsync();
    if (sp) {
      for (int j = 0; j < ns; ++j) myrow[j] = myrow[j] / rho + L.g[j] * (o.w / rho) * (8314.51 / P.W[j]) / Rm;
      myrow[ns] = myrow[ns] / rho + o.w / (rho * T);
    }
    if (lane < nz) Jg[ns * 33 + lane] = jt;
    tsync();
  }
}

// LU of I - c J with partial pivoting in LDS, one row per lane (stride 33, conflict-free); the
// pivot row is read by every lane as an LDS broadcast.  Returns this lane's elimination step.
__device__ __forceinline__ int team_lu(double c, int nz, const Lds& L, const double* __restrict__ Jg) {
  const int lane = threadIdx.x & (TEAM - 1);
  double* arow = L.lu + lane * 33;
  const double* __restrict__ jrow = Jg + lane * 33;
  for (int j = 0; j < TEAM; ++j) arow[j] = (lane == j ? 1.0 : 0.0) - ((lane < nz && j < nz) ? c * jrow[j] : 0.0);
  int mystep = -1;
  tsync();
#pragma unroll 1
  for (int k = 0; k < TEAM; ++k) {
    double best = mystep < 0 ? fabs(arow[k]) : -1.0;
    int bl = lane;
#pragma unroll
    for (int m = 16; m >= 1; m >>= 1) {
      const double ob = __shfl_xor(best, m, TEAM);
      const int ol = __shfl_xor(bl, m, TEAM);
      if (ob > best || (ob == best && ol < bl)) { best = ob; bl = ol; }
    }
    if (lane == bl) mystep = k;
    if (lane == 0) L.piv[k] = bl;
    const double* prow = L.lu + bl * 33;
    if (mystep < 0) {
      const double l = arow[k] / prow[k];
      arow[k] = l;
#pragma unroll 1
      for (int j = k + 1; j < TEAM; ++j) arow[j] -= l * prow[j];
    }
    tsync();
  }
  return mystep;
}

// Solve (I - cJ) x = b with the factors above; b and x are this lane's component.
__device__ __forceinline__ double team_solve(int mystep, double b, const Lds& L) {
  const int lane = threadIdx.x & (TEAM - 1);
  const double* arow = L.lu + lane * 33;
#pragma unroll 1
  for (int k = 0; k < TEAM; ++k) {
    const double zk = tbc(b, L.piv[k]);
    if (mystep > k) b -= arow[k] * zk;
  }
  double s = 0.0, x = 0.0;
#pragma unroll 1
  for (int k = TEAM - 1; k >= 0; --k) {
    const double xk = tbc((b - s) / arow[k], L.piv[k]);
    if (mystep < k) s += arow[k] * xk;
    if (lane == k) x = xk;
  }
  return x;
}

__device__ __forceinline__ double team_rms(double v, double scale, int nz) {
  const int lane = threadIdx.x & (TEAM - 1);
  const double q = lane < nz ? v / scale : 0.0;
  return sqrt(tsum(q * q) / nz);
}

// D[0..order] <- (R U)^T D[0..order] on this lane's column, with R(order, factor) and U = R(order, 1)
// of the method: R[j][l] = prod_{m=1..j} (m - 1 - factor l)/m, so R[0][l] = 1 and R[j][0] = [j == 0].
// Because RU[j][0] = [j == 0] and RU[0][i] = [i == 0], D[0] (the solution itself) is left as it is
// and never enters the differences: v_l = sum_{j>=1} R[j][l] D[j], then D_i = sum_{l>=1} U[l][i] v_l
// for i >= 1.  Mixing D[0] into v_l and relying on sum_l U[l][i] = 0 to remove it again would leave
// an absolute error of eps |y| in every difference, which later step-size increases amplify by
// factor^j into a visible loss of mass after a long run of step halvings.
__device__ __forceinline__ void change_D(double (&D)[MAXO + 3], int order, double factor) {
  double v[MAXO + 1];
#pragma unroll
  for (int l = 1; l <= MAXO; ++l) {
    double prod = 1.0, acc = 0.0;
#pragma unroll
    for (int j = 1; j <= MAXO; ++j) {
      prod *= (j - 1 - factor * l) / j;
      if (j <= order) acc += prod * D[j];
    }
    v[l] = (l <= order) ? acc : 0.0;
  }
#pragma unroll
  for (int i = 1; i <= MAXO; ++i) {
    if (i > order) continue;
    double acc = 0.0;
#pragma unroll
    for (int l = 1; l <= MAXO; ++l) {
      // U[l][i] = prod_{m=1..l} (m - 1 - i)/m: a signed binomial coefficient, exact in double
      double u = 1.0;
#pragma unroll
      for (int m = 1; m <= MAXO; ++m)
        if (m <= l) u *= (double)(m - 1 - i) / m;
      if (l <= order) acc += u * v[l];
    }
    D[i] = acc;
  }
}

// One cell; mirrors bdf_cell of oracle/oracle_bdf.c.  Returns 0 ok, 1 step underflow, 2 too many
// steps, 3 invalid state.
__device__ __forceinline__ int integrate(const Prog& P, const Opts& O, double p, double& y, int (&st)[8], const Lds& L,
                                          double* __restrict__ Jg) {
  const int lane = threadIdx.x & (TEAM - 1);
  const int nz = P.ns + 1;
  const double rtol = O.rtol, atol = O.atol, t_end = O.t_end;
  double gamma[MAXO + 1];
  gamma[0] = 0.0;
#pragma unroll
  for (int k = 1; k <= MAXO; ++k) gamma[k] = gamma[k - 1] + 1.0 / k;
  const double newton_tol = fmax(10.0 * 2.220446049250313e-16 / rtol, fmin(0.03, sqrt(rtol)));
  RhsOut o;
  o = team_rhs(P, O.mode, p, y, L, false);
  if (!o.ok) return 3;
  st[ST_RHS]++;
  const double f0 = o.F;
  double sc = atol + fabs(y) * rtol;
  const double d0 = team_rms(y, sc, nz), d1 = team_rms(f0, sc, nz);
  double h0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
  h0 = fmin(h0, t_end);
  o = team_rhs(P, O.mode, p, y + h0 * f0, L, false);
  if (!o.ok) return 3;
  st[ST_RHS]++;
  const double d2 = team_rms(o.F - f0, sc, nz) / h0;
  const double h1 = (d1 <= 1e-15 && d2 <= 1e-15) ? fmax(1e-6, h0 * 1e-3) : sqrt(0.01 / fmax(d1, d2));
  double h_abs = fmin(100.0 * h0, fmin(h1, t_end));

  double D[MAXO + 3];
#pragma unroll
  for (int j = 0; j < MAXO + 3; ++j) D[j] = 0.0;
  D[0] = y;
  D[1] = f0 * h_abs;
  int order = 1, n_equal = 0;
  bool have_lu = false, current_jac = false;
  int mystep = lane;
  // Jacobian at the initial state
  o = team_rhs(P, O.mode, p, y, L, true);
  if (!o.ok) return 3;
  team_jac(P, O.mode, y, L, o, Jg);
  st[ST_JAC]++;
  double t = 0.0;

  while (t < t_end) {
    if (st[ST_STEPS] >= O.max_steps) return 2;
    const double min_step = 10.0 * fabs(nextafter(t, INFINITY) - t);
    if (h_abs < min_step) {
      change_D(D, order, min_step / h_abs);
      h_abs = min_step;
      n_equal = 0;
    }
    double ynew = 0.0, d = 0.0, t_new = t, err_norm = 0.0, safety = 0.9;
    bool accepted = false;
    while (!accepted) {
      if (h_abs < min_step) return 1;
      t_new = t + h_abs;
      if (t_new - t_end > 0) {
        t_new = t_end;
        change_D(D, order, fabs(t_new - t) / h_abs);
        n_equal = 0;
        have_lu = false;
      }
      const double h = t_new - t;
      h_abs = fabs(h);
      double ypred = 0.0, psi = 0.0;
#pragma unroll
      for (int j = 0; j <= MAXO; ++j)
        if (j <= order) ypred += D[j];
#pragma unroll
      for (int j = 1; j <= MAXO; ++j)
        if (j <= order) psi += D[j] * gamma[j];
      double alpha_o = 0.0;
#pragma unroll
      for (int j = 1; j <= MAXO; ++j)
        if (j == order) alpha_o = gamma[j];
      psi /= alpha_o;
      const double scale = atol + rtol * fabs(ypred);
      const double cc = h / alpha_o;
      bool converged = false;
      int n_iter = 0;
      while (!converged) {
        if (!have_lu) {
          mystep = team_lu(cc, nz, L, Jg);
          st[ST_LU]++;
          have_lu = true;
        }
        ynew = ypred;
        d = 0.0;
        double dy_norm_old = -1.0;
        int k;
        for (k = 0; k < NEWTON_MAXITER; ++k) {
          o = team_rhs(P, O.mode, p, ynew, L, false);
          const bool ok = o.ok;
          st[ST_RHS]++;
          if (!ok || !isfinite(tsum(o.F))) break;
          const double dy = team_solve(mystep, cc * o.F - psi - d, L);
          st[ST_NEWTON]++;
          const double dy_norm = team_rms(dy, scale, nz);
          const double rate = dy_norm_old >= 0.0 ? dy_norm / dy_norm_old : -1.0;
          if (rate >= 0.0 && (rate >= 1.0 || pow(rate, (double)(NEWTON_MAXITER - k)) / (1.0 - rate) * dy_norm > newton_tol))
            break;
          ynew += dy;
          d += dy;
          if (dy_norm == 0.0 || (rate >= 0.0 && rate / (1.0 - rate) * dy_norm < newton_tol)) {
            converged = true;
            break;
          }
          dy_norm_old = dy_norm;
        }
        n_iter = k + 1;
        if (!converged) {
          if (current_jac) break;
          o = team_rhs(P, O.mode, p, ypred, L, true);
          if (!o.ok) break;
          team_jac(P, O.mode, ypred, L, o, Jg);
          st[ST_JAC]++;
          have_lu = false;
          current_jac = true;
        }
      }
      if (!converged) {
        st[ST_NFAIL]++;
        h_abs *= 0.5;
        change_D(D, order, 0.5);
        n_equal = 0;
        have_lu = false;
        continue;
      }
      safety = 0.9 * (2 * NEWTON_MAXITER + 1) / (double)(2 * NEWTON_MAXITER + n_iter);
      const double scale2 = atol + rtol * fabs(ynew);
      sc = scale2;
      err_norm = team_rms(d / (order + 1), scale2, nz);
      if (err_norm > 1.0) {
        st[ST_REJECT]++;
        const double factor = fmax(MIN_FACTOR, safety * pow(err_norm, -1.0 / (order + 1)));
        h_abs *= factor;
        change_D(D, order, factor);
        n_equal = 0;
      } else {
        accepted = true;
      }
    }
    st[ST_STEPS]++;
    n_equal++;
    t = t_new;
    y = ynew;
    current_jac = false;
    // D^{j+1} y_n = D^j y_n - D^j y_{n-1}, d = D^{order+1} y_n
    {
      double Dop1 = 0.0;
#pragma unroll
      for (int j = 0; j < MAXO + 3; ++j)
        if (j == order + 1) Dop1 = D[j];
#pragma unroll
      for (int j = 0; j < MAXO + 3; ++j) {
        if (j == order + 2) D[j] = d - Dop1;
        if (j == order + 1) D[j] = d;
      }
#pragma unroll
      for (int j = MAXO; j >= 0; --j)
        if (j <= order) D[j] += D[j + 1];
    }
    if (n_equal < order + 1) continue;
    double Dm = 0.0, Dp = 0.0;
#pragma unroll
    for (int j = 0; j < MAXO + 3; ++j) {
      if (j == order) Dm = D[j];
      if (j == order + 2) Dp = D[j];
    }
    const double em = order > 1 ? team_rms(Dm / order, sc, nz) : INFINITY;
    const double ep = order < MAXO ? team_rms(Dp / (order + 2), sc, nz) : INFINITY;
    const double norms[3] = {em, err_norm, ep};
    double best = -1.0;
    int which = 1;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double fct = norms[q] == 0.0 ? INFINITY : pow(norms[q], -1.0 / (order + q));
      if (fct > best) { best = fct; which = q; }
    }
    order += which - 1;
    const double factor = fmin(MAX_FACTOR, safety * best);
    h_abs *= factor;
    change_D(D, order, factor);
    n_equal = 0;
    have_lu = false;
  }
  st[ST_ORDER] = order;
  return 0;
}

__global__ void zero_counter(int* c) {
  if (threadIdx.x == 0 && blockIdx.x == 0) c[0] = 0;
}

// Two waves per SIMD: the per-team LDS (LU factors) allows about seven resident waves per CU, and
// at two waves the register budget costs only a small spill (measured faster than one wave at 400
// registers, and than three waves).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void bdf_kernel(Prog P, Opts O, long n, long ld, double* __restrict__ X,
                                                  const double* __restrict__ pres, int* __restrict__ stats,
                                                  int* __restrict__ status, int* __restrict__ counter,
                                                  unsigned team_bytes, double* __restrict__ jac_scratch) {
  extern __shared__ __align__(16) char smem[];
  const int lane = threadIdx.x & (TEAM - 1);
  const int team = threadIdx.x / TEAM;
  const Lds L = carve(smem + (size_t)team * team_bytes, P.nr, P.nms);
  double* __restrict__ Jg = jac_scratch + ((size_t)blockIdx.x * (blockDim.x / TEAM) + team) * (TEAM * 33);
  const int nz = P.ns + 1;
  while (true) {
    int cell = 0;
    if (lane == 0) cell = atomicAdd(counter, 1);
    cell = tbci(cell, 0);
    if (cell >= n) break;
    double y = lane < nz ? X[(long)lane * ld + cell] : 0.0;
    const double p = O.mode == 1 ? pres[cell] : 0.0;
    int st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int rc = integrate(P, O, p, y, st, L, Jg);
    if (lane < nz) X[(long)lane * ld + cell] = y;
    if (stats && lane < 8) {
      int v = st[0];
#pragma unroll
      for (int q = 1; q < 8; ++q)
        if (lane == q) v = st[q];
      stats[(long)lane * ld + cell] = v;
    }
    if (status && lane == 0) status[cell] = rc;
  }
}

// The reactor's right-hand side and (JAC) Jacobian at given states (the same team_rhs/team_jac the
// integrator uses), one team per cell over a grid-stride loop.  Outside the tables F = -1, J = 0.
template <bool JAC>
__global__ __launch_bounds__(256) void jac_kernel(Prog P, int mode, long n, long ld, const double* __restrict__ X,
                                                  const double* __restrict__ pres, double* __restrict__ F,
                                                  double* __restrict__ J, unsigned team_bytes,
                                                  double* __restrict__ jac_scratch) {
  extern __shared__ __align__(16) char smem[];
  const int lane = threadIdx.x & (TEAM - 1);
  const int team = threadIdx.x / TEAM;
  const int tpb = blockDim.x / TEAM;
  const Lds L = carve(smem + (size_t)team * team_bytes, P.nr, P.nms);
  double* __restrict__ Jg = jac_scratch + ((size_t)blockIdx.x * tpb + team) * (TEAM * 33);
  const int nz = P.ns + 1;
  const long stride = (long)gridDim.x * tpb;
  for (long cell = (long)blockIdx.x * tpb + team; cell < n; cell += stride) {
    const double y = lane < nz ? X[(long)lane * ld + cell] : 0.0;
    const double p = mode == 1 ? pres[cell] : 0.0;
    const RhsOut o = team_rhs(P, mode, p, y, L, JAC);
    if (!o.ok) {
      if (lane < nz) {
        F[(long)lane * ld + cell] = -1.0;
        if (JAC)
          for (int c = 0; c < nz; ++c) J[((long)lane * nz + c) * ld + cell] = 0.0;
      }
      continue;
    }
    if (!JAC) {
      if (lane < nz) F[(long)lane * ld + cell] = o.F;
      tsync();
      continue;
    }
    team_jac(P, mode, y, L, o, Jg);
    if (lane < nz) {
      F[(long)lane * ld + cell] = o.F;
      for (int c = 0; c < nz; ++c) J[((long)lane * nz + c) * ld + cell] = Jg[lane * 33 + c];
    }
    tsync();
  }
}

}  // namespace flint_reactor

hipError_t flint_reactor_set_max_lds(int bytes) {
  const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&flint_reactor::bdf_kernel),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (e != hipSuccess) return e;
  const hipError_t e2 = hipFuncSetAttribute(reinterpret_cast<const void*>(&flint_reactor::jac_kernel<true>),
                                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (e2 != hipSuccess) return e2;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(&flint_reactor::jac_kernel<false>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

size_t flint_reactor_team_bytes(int nr, int nms) { return flint_reactor::team_lds_bytes(nr, nms); }

hipError_t flint_reactor_launch(const flint_reactor::Prog& P, const flint_reactor::Opts& O, long n, long ld,
                                double* X, const double* pres, int* stats, int* status, int* counter, int teams_per_block,
                                int blocks, double* jac_scratch, hipStream_t s) {
  const unsigned tb = (unsigned)((flint_reactor::team_lds_bytes(P.nr, P.nms) + 15) & ~size_t(15));
  hipLaunchKernelGGL(flint_reactor::zero_counter, dim3(1), dim3(64), 0, s, counter);
  hipLaunchKernelGGL(flint_reactor::bdf_kernel, dim3(blocks), dim3(teams_per_block * flint_reactor::TEAM),
                     (size_t)tb * teams_per_block, s, P, O, n, ld, X, pres, stats, status, counter, tb, jac_scratch);
  return hipGetLastError();
}

hipError_t flint_reactor_jac_launch(const flint_reactor::Prog& P, int mode, long n, long ld, const double* X,
                                    const double* pres, double* F, double* J, int teams_per_block, int blocks,
                                    double* jac_scratch, hipStream_t s) {
  const unsigned tb = (unsigned)((flint_reactor::team_lds_bytes(P.nr, P.nms) + 15) & ~size_t(15));
  if (J)
    hipLaunchKernelGGL(flint_reactor::jac_kernel<true>, dim3(blocks), dim3(teams_per_block * flint_reactor::TEAM),
                       (size_t)tb * teams_per_block, s, P, mode, n, ld, X, pres, F, J, tb, jac_scratch);
  else
    hipLaunchKernelGGL(flint_reactor::jac_kernel<false>, dim3(blocks), dim3(teams_per_block * flint_reactor::TEAM),
                       (size_t)tb * teams_per_block, s, P, mode, n, ld, X, pres, F, J, tb, jac_scratch);
  return hipGetLastError();
}
