/* CPU oracle of the FLINT per-cell thermochemistry path -- TEST INFRASTRUCTURE ONLY.
 *
 * Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the
 * shipped library.  See flint_oracle.h for layouts and oracle_body.inc for the restated routines.
 *
 * Pinning (what checks each function):
 *   orc_wdot / orc_wd_hrr  against the reference Fortran itself, compiled by oracle/Makefile into
 *                          oracle/_ref (tests/test_oracle_ref.py), and against the committed
 *                          fixtures written from it (tests/golden/); mass conservation
 *                          sum_k wdot_k = 0 and element conservation (tests/test_oracle.py)
 *   orc_rhs                against oracle/_ref's restatement-free closed form of rhs_native
 *                          (Lib_Chemistry_rhs.f90:15-52) and internal-energy conservation
 *   orc_jac                against the complex-step derivative orc_jac_cstep of orc_rhs
 *                          (no Jacobian exists in the reference: that is the mathematics' pin)
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "flint_oracle.h"

#define ORC_MAX_NS 64
#define ORC_MAX_NR 512
#define ORC_RUNIV 8314.51 /* Lib_ThermoTransport.f90:13 */

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* threads of the following parallel loops (the CPU baseline's 1-core and N-core figures) */
void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ---- real instantiation ---- */
#define REAL double
#define SFX(x) x##_r
#define RREAL(x) (x)
#define RLOG10(x) log10(x)
#define REXP10(x) pow(10.0, (x))
#define RPOW(x, e) pow((x), (e))
#define cabs_(x) fabs(x)
#include "oracle_body.inc"
#undef REAL
#undef SFX
#undef RREAL
#undef RLOG10
#undef REXP10
#undef RPOW
#undef cabs_

/* x**n for an integer n, by binary powering (the runtime routine a Fortran compiler calls for
 * real**integer with a run-time exponent) */
static double powi_r(double a, int b) {
  const int recip = b < 0;
  if (recip) b = -b;
  double r = 1.0;
  for (;;) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return recip ? 1.0 / r : r;
}

/* The data-driven general routine, Lib_Chemistry_wdot.f90:96-221: per kind (Arrhenius and
 * three-body, Troe, Lindemann) a loop over its reactions with the dense (ns+1)-row tables ni1,
 * ni2, eps ([reaction][ns+1], last row the third body): coM over all species when ni1(M) > 0,
 * products over the species with a nonzero coefficient (integer powers for Arrhenius, real
 * powers for fall-off), and omegadot(k) += W_k (ni2 - ni1) net for every species.  scale[k] =
 * sum over reactions of |W_k (ni2 - ni1)| (|rate_fwd| + |rate_rev|) (may be NULL). */
void orc_general(const orc_mech *m, int na, int nt, int nl, const double *ni1, const double *ni2,
                 const double *eps, long n, long ld, const double *rho, const double *T, double *wdot,
                 double *scale) {
  const int ns = m->ns, nz = ns + 1;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) {
    double coi[ORC_MAX_NS], om[ORC_MAX_NS], sc[ORC_MAX_NS];
    for (int k = 0; k < ns; ++k) coi[k] = rho[(long)k * ld + i] / m->W[k];
    const double temp = T[i];
    const int T_i = (int)temp;
    const double Tdiff = temp - T_i;
    const long row = T_i - m->Tk_min;
    for (int k = 0; k < ns; ++k) om[k] = sc[k] = 0.0;
    for (int j = 0; j < na + nt + nl; ++j) {
      const double *a1 = ni1 + (long)j * nz, *a2 = ni2 + (long)j * nz, *e = eps + (long)j * nz;
      double coM = 1.0;
      if (a1[ns] > 0.0) {
        coM = 0.0;
        for (int k = 0; k < ns; ++k) coM = coM + coi[k] * e[k];
      }
      double pf = 1.0, pb = 1.0;
      for (int k = 0; k < ns; ++k) {
        if (j < na) {
          if (a1[k] != 0) pf = pf * powi_r(coi[k], (int)lround(a1[k]));
          if (a2[k] != 0) pb = pb * powi_r(coi[k], (int)lround(a2[k]));
        } else {
          if (a1[k] != 0) pf = pf * pow(coi[k], a1[k]);
          if (a2[k] != 0) pb = pb * pow(coi[k], a2[k]);
        }
      }
      double rf, rb;
      if (j < na) {
        rf = interp_r(m->kf, row, m->n_arrh, j, Tdiff) * pf * coM;
        rb = interp_r(m->kb, row, m->n_arrh, j, Tdiff) * pb * coM;
      } else {
        double kinf, k0, kc, k1;
        if (j < na + nt) {
          const int tb = j - na;
          kinf = interp_r(m->troe, row, 4L * m->n_troe, 4L * tb + 0, Tdiff);
          k0 = interp_r(m->troe, row, 4L * m->n_troe, 4L * tb + 1, Tdiff);
          kc = interp_r(m->troe, row, 4L * m->n_troe, 4L * tb + 2, Tdiff);
          const double Fcent = interp_r(m->troe, row, 4L * m->n_troe, 4L * tb + 3, Tdiff);
          const double Pr = k0 * coM / kinf;
          k1 = kinf * (Pr / (1.0 + Pr)) * troeF_r(Pr, Fcent);
        } else {
          const int tb = j - na - nt;
          kinf = interp_r(m->lind, row, 3L * m->n_lind, 3L * tb + 0, Tdiff);
          k0 = interp_r(m->lind, row, 3L * m->n_lind, 3L * tb + 1, Tdiff);
          kc = interp_r(m->lind, row, 3L * m->n_lind, 3L * tb + 2, Tdiff);
          const double Pr = k0 * coM / kinf;
          k1 = kinf * (Pr / (1.0 + Pr));
        }
        rf = k1 * pf;
        rb = k1 / kc * pb;
      }
      const double net = rf - rb;
      for (int k = 0; k < ns; ++k) {
        const double d = a2[k] - a1[k];
        om[k] = om[k] + m->W[k] * d * net;
        sc[k] += fabs(m->W[k] * d) * (fabs(rf) + fabs(rb));
      }
    }
    for (int k = 0; k < ns; ++k) {
      wdot[(long)k * ld + i] = om[k];
      if (scale) scale[(long)k * ld + i] = sc[k];
    }
  }
}

/* ---- complex instantiation (complex-step differentiation) ---- */
#define REAL double _Complex
#define SFX(x) x##_c
#define RREAL(x) creal(x)
#define RLOG10(x) (clog(x) / M_LN10)
#define REXP10(x) cexp((x) * M_LN10)
#define RPOW(x, e) cpow((x), (e))
#define cabs_(x) cabs(x)
#include "oracle_body.inc"
#undef REAL
#undef SFX
#undef RREAL
#undef RLOG10
#undef REXP10
#undef RPOW
#undef cabs_

void orc_wdot(const orc_mech *m, long n, long ld, const double *rho, const double *T, double *wdot,
              double *wdot_scale) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) {
    double roi[ORC_MAX_NS], w[ORC_MAX_NS], sc[ORC_MAX_NS], pf[ORC_MAX_NR], pb[ORC_MAX_NR];
    for (int k = 0; k < m->ns; ++k) roi[k] = rho[k * ld + i];
    if (m->is_wd) {
      wd_cell_r(m, roi, T[i], w);
      for (int k = 0; k < m->ns; ++k) sc[k] = fabs(w[k]);
    } else {
      wdot_cell_r(m, roi, T[i], w, pf, pb, sc);
    }
    for (int k = 0; k < m->ns; ++k) {
      wdot[k * ld + i] = w[k];
      if (wdot_scale) wdot_scale[k * ld + i] = sc[k];
    }
  }
}

/* WD source term plus heat release -sum_k h_k(T) wdot_k (h from the thermo table,
 * Lib_ThermoTransport.f90:77-96 H0i). */
void orc_wd_hrr(const orc_mech *m, long n, long ld, const double *rho, const double *T, double *wdot,
                double *hrr) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) {
    double roi[ORC_MAX_NS], w[ORC_MAX_NS], pf[ORC_MAX_NR], pb[ORC_MAX_NR];
    for (int k = 0; k < m->ns; ++k) roi[k] = rho[k * ld + i];
    if (m->is_wd) wd_cell_r(m, roi, T[i], w);
    else wdot_cell_r(m, roi, T[i], w, pf, pb, 0);
    int T_i = (int)T[i];
    double Tdiff = T[i] - T_i;
    long row = T_i - m->Th_min;
    double q = 0.0;
    for (int k = 0; k < m->ns; ++k) {
      wdot[k * ld + i] = w[k];
      q = q + interp_r(m->h, row, m->ns, k, Tdiff) * w[k];
    }
    hrr[i] = -q;
  }
}

void orc_rhs(const orc_mech *m, long n, long ld, const double *Z, double *F) {
  int nz = m->ns + 1;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) {
    double z[ORC_MAX_NS + 1], f[ORC_MAX_NS + 1];
    for (int k = 0; k < nz; ++k) z[k] = Z[k * ld + i];
    rhs_cell_r(m, z, f);
    for (int k = 0; k < nz; ++k) F[k * ld + i] = f[k];
  }
}

void orc_jac_cstep(const orc_mech *m, long n, long ld, const double *Z, double h, double *jac) {
  int nz = m->ns + 1;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) {
    double _Complex z[ORC_MAX_NS + 1], f[ORC_MAX_NS + 1];
    for (int c = 0; c < nz; ++c) {
      for (int k = 0; k < nz; ++k) z[k] = Z[k * ld + i];
      z[c] += I * h;
      rhs_cell_c(m, z, f);
      for (int r = 0; r < nz; ++r) jac[((long)r * nz + c) * ld + i] = cimag(f[r]) / h;
    }
  }
}

/* ------------------------------------------------------------------------------------------
 * Analytic Jacobian of the constant-volume RHS.  The reference has none (its integrator
 * approximates one internally); this is the exact derivative of the table-interpolated model:
 *   d(a+(b-a)Tdiff)/dT = b-a within the bin, d max(rho,0)/drho = 1 for rho >= 0 else 0.
 * Species block  J_kj = W_k sum_r nu_kr dq_r/dc_j / W_j,  q_r = prodf_r - prodb_r
 * T row          dTdot/drho_j = -(sum_k e_k J_kj + Tdot cv_j) / D
 *                dTdot/dT     = -(sum_k (de_k/dT wdot_k + e_k J_kT) + Tdot sum_k rho_k dcv_k/dT) / D
 * with e_k = h_k - R_k T, cv_k = cp_k - R_k, D = sum_k rho_k cv_k.
 * ------------------------------------------------------------------------------------------ */
typedef struct {
  double k1, k2;     /* effective forward / reverse coefficient (incl. M for three-body) */
  double dk1T, dk2T; /* their T derivatives at fixed concentrations */
  double dk1M, dk2M; /* derivatives with respect to M (falloff) or the bare factor (three-body) */
} orc_rate;

static void falloff_derivs(double kinf, double k0, double kc, double Fcent, double dkinf, double dk0, double dkc,
                           double dFcent, double M, int troe, orc_rate *o) {
  double Pr = k0 * M / kinf;
  double g = Pr / (1.0 + Pr), dg = 1.0 / ((1.0 + Pr) * (1.0 + Pr));
  double F = 1.0, dlnF_dPr = 0.0, dlnF_dFc = 0.0;
  if (troe) {
    double L = log10(Fcent), x = log10(Pr);
    double c = -0.4 - 0.67 * L, n = 0.75 - 1.27 * L, u = x + c, den = n - 0.14 * u;
    double f1 = u / den, q = 1.0 + f1 * f1;
    F = pow(10.0, L / q);
    double df1_du = n / (den * den);
    double df1_dL = (-0.67 * n + 1.27 * u) / (den * den);
    dlnF_dPr = -2.0 * L * f1 * df1_du / (q * q * Pr);
    dlnF_dFc = M_LN10 * (1.0 / q - L * 2.0 * f1 * df1_dL / (q * q)) / (Fcent * M_LN10);
  }
  double k1 = kinf * g * F;
  double dk1_dPr = kinf * F * (dg + g * dlnF_dPr);
  double dPr_dT = M * (dk0 * kinf - k0 * dkinf) / (kinf * kinf);
  double dk1T = dkinf * g * F + dk1_dPr * dPr_dT + k1 * dlnF_dFc * dFcent;
  o->k1 = k1;
  o->k2 = k1 / kc;
  o->dk1T = dk1T;
  o->dk2T = dk1T / kc - k1 * dkc / (kc * kc);
  o->dk1M = dk1_dPr * k0 / kinf;
  o->dk2M = o->dk1M / kc;
}

static double slope(const double *tab, long row, long stride, long col) {
  return tab[(row + 1) * stride + col] - tab[row * stride + col];
}

/* dq/dc_j of prod_{t} c_{sp_t}^{pw_t} times k, accumulated into dq[] with sign */
static void add_dprod(double k, const int *ptr, const int *sp, const int *pw, int r, const double *coi, double sign,
                      double *dq, double *dqabs) {
  for (int t = ptr[r]; t < ptr[r + 1]; ++t) {
    double d = k * pw[t];
    for (int p = 1; p < pw[t]; ++p) d *= coi[sp[t]];
    for (int u = ptr[r]; u < ptr[r + 1]; ++u) {
      if (u == t) continue;
      for (int p = 0; p < pw[u]; ++p) d *= coi[sp[u]];
    }
    dq[sp[t]] += sign * d;
    dqabs[sp[t]] += fabs(d);
  }
}

static double prod_only(const int *ptr, const int *sp, const int *pw, int r, const double *coi) {
  double v = 1.0;
  for (int t = ptr[r]; t < ptr[r + 1]; ++t)
    for (int p = 0; p < pw[t]; ++p) v *= coi[sp[t]];
  return v;
}

/* Species block of the Jacobian at clipped densities roi (all >= 0):
 * A[k*ns+j] = d wdot_k / d rho_j, b[k] = d wdot_k / dT at fixed rho (with |.| scales). */
static void jac_core(const orc_mech *m, const double *nu, const double *roi, double T, double *A, double *Aa,
                     double *b, double *ba) {
  int ns = m->ns;
  double coi[ORC_MAX_NS];
  for (int k = 0; k < ns; ++k) coi[k] = roi[k] / m->W[k];
  int T_i = (int)T;
  double Tdiff = T - T_i;
  long row = T_i - m->Tk_min;
  double dwdc[ORC_MAX_NS][ORC_MAX_NS]; /* [k][j] = sum_r nu_kr dq_r/dc_j */
  double dwdc_abs[ORC_MAX_NS][ORC_MAX_NS];
  double dwdT[ORC_MAX_NS], dwdT_abs[ORC_MAX_NS];
  memset(dwdc, 0, sizeof dwdc);
  memset(dwdc_abs, 0, sizeof dwdc_abs);
  memset(dwdT, 0, sizeof dwdT);
  memset(dwdT_abs, 0, sizeof dwdT_abs);
  for (int r = 0; r < m->nr; ++r) {
    int kind = m->rx_kind[r], tb = m->rx_table[r];
    double M = 1.0;
    if (m->rx_mset[r] >= 0) M = msum_r(m, m->rx_mset[r], coi);
    double Pf = prod_only(m->rx_fwd_ptr, m->rx_fwd_sp, m->rx_fwd_pow, r, coi);
    double Pb = prod_only(m->rx_rev_ptr, m->rx_rev_sp, m->rx_rev_pow, r, coi);
    double dq[ORC_MAX_NS], dqa[ORC_MAX_NS];
    for (int j = 0; j < ns; ++j) { dq[j] = 0.0; dqa[j] = 0.0; }
    double dqT, dqTa, dqM = 0.0, dqMa = 0.0;
    if (kind <= 1) {
      double kf = interp_r(m->kf, row, m->n_arrh, tb, Tdiff), kb = interp_r(m->kb, row, m->n_arrh, tb, Tdiff);
      double dkf = slope(m->kf, row, m->n_arrh, tb), dkb = slope(m->kb, row, m->n_arrh, tb);
      double Mf = (kind == 1) ? M : 1.0;
      add_dprod(kf * Mf, m->rx_fwd_ptr, m->rx_fwd_sp, m->rx_fwd_pow, r, coi, 1.0, dq, dqa);
      add_dprod(kb * Mf, m->rx_rev_ptr, m->rx_rev_sp, m->rx_rev_pow, r, coi, -1.0, dq, dqa);
      dqT = (dkf * Pf - dkb * Pb) * Mf;
      dqTa = (fabs(dkf * Pf) + fabs(dkb * Pb)) * fabs(Mf);
      if (kind == 1) { dqM = kf * Pf - kb * Pb; dqMa = fabs(kf * Pf) + fabs(kb * Pb); }
    } else {
      orc_rate o;
      if (kind == 2) {
        long st = 4L * m->n_troe, c0 = 4L * tb;
        falloff_derivs(interp_r(m->troe, row, st, c0, Tdiff), interp_r(m->troe, row, st, c0 + 1, Tdiff),
                       interp_r(m->troe, row, st, c0 + 2, Tdiff), interp_r(m->troe, row, st, c0 + 3, Tdiff),
                       slope(m->troe, row, st, c0), slope(m->troe, row, st, c0 + 1), slope(m->troe, row, st, c0 + 2),
                       slope(m->troe, row, st, c0 + 3), M, 1, &o);
      } else {
        long st = 3L * m->n_lind, c0 = 3L * tb;
        falloff_derivs(interp_r(m->lind, row, st, c0, Tdiff), interp_r(m->lind, row, st, c0 + 1, Tdiff),
                       interp_r(m->lind, row, st, c0 + 2, Tdiff), 1.0, slope(m->lind, row, st, c0),
                       slope(m->lind, row, st, c0 + 1), slope(m->lind, row, st, c0 + 2), 0.0, M, 0, &o);
      }
      add_dprod(o.k1, m->rx_fwd_ptr, m->rx_fwd_sp, m->rx_fwd_pow, r, coi, 1.0, dq, dqa);
      add_dprod(o.k2, m->rx_rev_ptr, m->rx_rev_sp, m->rx_rev_pow, r, coi, -1.0, dq, dqa);
      dqT = o.dk1T * Pf - o.dk2T * Pb;
      dqTa = fabs(o.dk1T * Pf) + fabs(o.dk2T * Pb);
      dqM = o.dk1M * Pf - o.dk2M * Pb;
      dqMa = fabs(o.dk1M * Pf) + fabs(o.dk2M * Pb);
    }
    if (m->rx_mset[r] >= 0) { /* chain through M = sum_j eff_j c_j */
      int set = m->rx_mset[r];
      if (m->ms_is_sum[set]) {
        for (int j = 0; j < ns; ++j) { dq[j] += dqM; dqa[j] += dqMa; }
      } else {
        for (int t = m->ms_ptr[set]; t < m->ms_ptr[set + 1]; ++t) {
          double e = isnan(m->ms_eff[t]) ? 1.0 : m->ms_eff[t];
          dq[m->ms_sp[t]] += e * dqM;
          dqa[m->ms_sp[t]] += fabs(e * dqMa);
        }
      }
    }
    for (int k = 0; k < ns; ++k) {
      double v = nu[(long)r * ns + k];
      if (v == 0.0) continue;
      for (int j = 0; j < ns; ++j) {
        dwdc[k][j] += v * dq[j];
        dwdc_abs[k][j] += fabs(v) * dqa[j];
      }
      dwdT[k] += v * dqT;
      dwdT_abs[k] += fabs(v) * dqTa;
    }
  }
  for (int k = 0; k < ns; ++k) {
    for (int j = 0; j < ns; ++j) {
      double f = m->W[k] / m->W[j];
      A[k * ns + j] = f * dwdc[k][j];
      Aa[k * ns + j] = fabs(f) * dwdc_abs[k][j];
    }
    b[k] = m->W[k] * dwdT[k];
    ba[k] = m->W[k] * dwdT_abs[k];
  }
}

/* Fabs (optional): the scale of F for the normwise comparison, the sum of the absolute values of
 * the terms each component sums: W_k sum_r |nu_kr| (|prodf_r| + |prodb_r|) for the species and
 * sum_k |e_k| Fabs_k / |D| for T. */
static void jac_cell(const orc_mech *m, const double *nu, const double *Z, double *F, double *J, double *Jabs,
                     double *Fabs) {
  int ns = m->ns, nz = ns + 1;
  double roi[ORC_MAX_NS], dclip[ORC_MAX_NS];
  double T = Z[ns];
  rhs_cell_r(m, Z, F);
  for (int a = 0; a < nz * nz; ++a) { J[a] = 0.0; Jabs[a] = 0.0; }
  if (Fabs)
    for (int k = 0; k < nz; ++k) Fabs[k] = 0.0;
  if (F[ns] == -1.0 && (T < m->Th_min || T > m->Th_max || isnan(T))) return;
  for (int k = 0; k < ns; ++k) {
    roi[k] = Z[k] < 0.0 ? 0.0 : Z[k];
    dclip[k] = Z[k] < 0.0 ? 0.0 : 1.0;
  }
  double A[ORC_MAX_NS * ORC_MAX_NS], Aa[ORC_MAX_NS * ORC_MAX_NS], b[ORC_MAX_NS], ba[ORC_MAX_NS];
  jac_core(m, nu, roi, T, A, Aa, b, ba);
  int T_i = (int)T;
  double Tdiff = T - T_i;
  /* species rows */
  for (int k = 0; k < ns; ++k) {
    for (int j = 0; j < ns; ++j) {
      J[k * nz + j] = dclip[j] * A[k * ns + j];
      Jabs[k * nz + j] = dclip[j] * Aa[k * ns + j];
    }
    J[k * nz + ns] = b[k];
    Jabs[k * nz + ns] = ba[k];
  }
  /* temperature row */
  long rowh = T_i - m->Th_min;
  double D = 0.0, Tdot = F[ns], s_dcv = 0.0, s_de = 0.0, s_de_abs = 0.0;
  double e[ORC_MAX_NS], cv[ORC_MAX_NS];
  for (int k = 0; k < ns; ++k) {
    double Ri = ORC_RUNIV / m->W[k];
    e[k] = interp_r(m->h, rowh, ns, k, Tdiff) - Ri * T;
    cv[k] = interp_r(m->cp, rowh, ns, k, Tdiff) - Ri;
    D += roi[k] * cv[k];
    s_dcv += roi[k] * slope(m->cp, rowh, ns, k);
    s_de += (slope(m->h, rowh, ns, k) - Ri) * F[k];
    s_de_abs += fabs((slope(m->h, rowh, ns, k) - Ri) * F[k]);
  }
  if (Fabs) {
    double wt[ORC_MAX_NS], pf[ORC_MAX_NR], pb[ORC_MAX_NR], qa = 0.0;
    wdot_cell_r(m, roi, T, wt, pf, pb, Fabs);
    for (int k = 0; k < ns; ++k) qa += fabs(e[k]) * Fabs[k];
    Fabs[ns] = qa / fabs(D);
  }
  for (int j = 0; j < nz; ++j) {
    double acc = 0.0, acca = 0.0;
    for (int k = 0; k < ns; ++k) { acc += e[k] * J[k * nz + j]; acca += fabs(e[k]) * Jabs[k * nz + j]; }
    if (j < ns) {
      J[ns * nz + j] = -(acc + Tdot * cv[j] * dclip[j]) / D;
      Jabs[ns * nz + j] = (acca + fabs(Tdot * cv[j])) / fabs(D);
    } else {
      J[ns * nz + ns] = -(s_de + acc + Tdot * s_dcv) / D;
      Jabs[ns * nz + ns] = (s_de_abs + acca + fabs(Tdot * s_dcv)) / fabs(D);
    }
  }
}

void orc_jac(const orc_mech *m, long n, long ld, const double *Z, double *F, double *jac, double *jabs,
             double *fabs_out) {
  int ns = m->ns, nz = ns + 1;
  if (m->is_wd) return; /* WD has fractional orders: no Jacobian program */
  /* nu_kr from the net-production program */
  double *nu = (double *)calloc((size_t)m->nr * ns, sizeof(double));
  for (int k = 0; k < ns; ++k)
    for (int t = m->wd_ptr[k]; t < m->wd_ptr[k + 1]; ++t) nu[(long)m->wd_rx[t] * ns + k] = m->wd_coef[t];
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) {
    double z[ORC_MAX_NS + 1], f[ORC_MAX_NS + 1];
    double J[(ORC_MAX_NS + 1) * (ORC_MAX_NS + 1)], Ja[(ORC_MAX_NS + 1) * (ORC_MAX_NS + 1)], fa[ORC_MAX_NS + 1];
    for (int k = 0; k < nz; ++k) z[k] = Z[k * ld + i];
    jac_cell(m, nu, z, f, J, Ja, fabs_out ? fa : 0);
    for (int k = 0; k < nz; ++k) F[k * ld + i] = f[k];
    if (fabs_out)
      for (int k = 0; k < nz; ++k) fabs_out[k * ld + i] = fa[k];
    for (int a = 0; a < nz * nz; ++a) {
      jac[(long)a * ld + i] = J[a];
      if (jabs) jabs[(long)a * ld + i] = Ja[a];
    }
  }
  free(nu);
}

/* Constant-pressure Jacobian d(dY/dt, dT/dt)/d(Y, T) from the species block at rho_k = rho Y_k:
 *   dF_k/dY_j = g_j [A_kj - (A Y)_k R_j/R + (wdot_k/rho) R_j/R]
 *   dF_k/dT   = (b_k - (A rho)_k / T)/rho + wdot_k/(rho T)
 *   dF_T/dY_j = -sum_k h_k dwdot_k/dY_j /(rho cp) - F_T g_j (cp_j/cp - R_j/R)
 *   dF_T/dT   = -(sum_k h'_k wdot_k + h_k dwdot_k/dT)/(rho cp) - F_T (sum_k Y_k cp'_k/cp - 1/T)
 * with dwdot_k/dY_j = rho g_j (A_kj - (A Y)_k R_j/R), dwdot_k/dT = b_k - (A rho)_k/T, R = R_mix,
 * and g_j = 1 (the constant-pressure model does not clip Y at 0). */
/* Ja, Fa (optional): the scales of J and F for the normwise comparison, each formula above with
 * every term replaced by its absolute value and every inner sum by the sum of |terms|: Aa, ba from
 * jac_core, W_k sum_r |nu_kr| (|prodf_r| + |prodb_r|) for wdot_k. */
static void jac_cp_cell(const orc_mech *m, const double *nu, const double *x, double p, double *F, double *J,
                        double *Fa, double *Ja) {
  int ns = m->ns, nz = ns + 1;
  double T = x[ns];
  rhs_cp_cell_r(m, x, p, F);
  for (int a = 0; a < nz * nz; ++a) J[a] = 0.0;
  if (Ja)
    for (int a = 0; a < nz * nz; ++a) Ja[a] = 0.0;
  if (Fa)
    for (int k = 0; k < nz; ++k) Fa[k] = 0.0;
  if (T < m->Th_min || T > m->Th_max || isnan(T)) return;
  double Yc[ORC_MAX_NS], g[ORC_MAX_NS], R[ORC_MAX_NS], roi[ORC_MAX_NS], Rmix = 0.0;
  for (int k = 0; k < ns; ++k) {
    Yc[k] = x[k];
    g[k] = 1.0;
    R[k] = ORC_RUNIV / m->W[k];
    Rmix += Yc[k] * R[k];
  }
  double rho = p / (Rmix * T);
  for (int k = 0; k < ns; ++k) roi[k] = rho * Yc[k];
  double A[ORC_MAX_NS * ORC_MAX_NS], Aa[ORC_MAX_NS * ORC_MAX_NS], b[ORC_MAX_NS], ba[ORC_MAX_NS];
  jac_core(m, nu, roi, T, A, Aa, b, ba);
  int T_i = (int)T;
  double Tdiff = T - T_i;
  long row = T_i - m->Th_min;
  double h[ORC_MAX_NS], cp[ORC_MAX_NS], dh[ORC_MAX_NS], dcp[ORC_MAX_NS], cpm = 0.0, sYdcp = 0.0;
  for (int k = 0; k < ns; ++k) {
    h[k] = interp_r(m->h, row, ns, k, Tdiff);
    cp[k] = interp_r(m->cp, row, ns, k, Tdiff);
    dh[k] = slope(m->h, row, ns, k);
    dcp[k] = slope(m->cp, row, ns, k);
    cpm += Yc[k] * cp[k];
    sYdcp += Yc[k] * dcp[k];
  }
  double FT = F[ns];
  double colh[ORC_MAX_NS + 1];
  for (int j = 0; j < nz; ++j) colh[j] = 0.0;
  double s_dhw = 0.0;
  for (int k = 0; k < ns; ++k) {
    double AY = 0.0;
    for (int j = 0; j < ns; ++j) AY += A[k * ns + j] * Yc[j];
    double w = F[k] * rho;
    for (int j = 0; j < ns; ++j) {
      double dwdY = rho * g[j] * (A[k * ns + j] - AY * R[j] / Rmix);
      J[k * nz + j] = g[j] * (A[k * ns + j] - AY * R[j] / Rmix + (w / rho) * R[j] / Rmix);
      colh[j] += h[k] * dwdY;
    }
    double dwdT = b[k] - rho * AY / T;
    J[k * nz + ns] = dwdT / rho + w / (rho * T);
    colh[ns] += h[k] * dwdT;
    s_dhw += dh[k] * w;
  }
  for (int j = 0; j < ns; ++j) J[ns * nz + j] = -colh[j] / (rho * cpm) - FT * g[j] * (cp[j] / cpm - R[j] / Rmix);
  J[ns * nz + ns] = -(s_dhw + colh[ns]) / (rho * cpm) - FT * (sYdcp / cpm - 1.0 / T);
  if (!Fa && !Ja) return;
  double wt[ORC_MAX_NS], pf[ORC_MAX_NR], pb[ORC_MAX_NR], wa[ORC_MAX_NS];
  wdot_cell_r(m, roi, T, wt, pf, pb, wa);
  double colha[ORC_MAX_NS + 1], s_dhwa = 0.0, Qa = 0.0, sYdcpa = 0.0, cpa = fabs(cpm);
  for (int j = 0; j < nz; ++j) colha[j] = 0.0;
  double Jk[ORC_MAX_NS + 1];
  for (int k = 0; k < ns; ++k) {
    double AYa = 0.0;
    for (int j = 0; j < ns; ++j) AYa += Aa[k * ns + j] * fabs(Yc[j]);
    for (int j = 0; j < ns; ++j) {
      double dwdYa = rho * (Aa[k * ns + j] + AYa * R[j] / Rmix);
      Jk[j] = Aa[k * ns + j] + AYa * R[j] / Rmix + (wa[k] / rho) * R[j] / Rmix;
      colha[j] += fabs(h[k]) * dwdYa;
    }
    double dwdTa = ba[k] + rho * AYa / T;
    Jk[ns] = dwdTa / rho + wa[k] / (rho * T);
    colha[ns] += fabs(h[k]) * dwdTa;
    s_dhwa += fabs(dh[k]) * wa[k];
    Qa += fabs(h[k]) * wa[k];
    sYdcpa += fabs(Yc[k] * dcp[k]);
    if (Ja)
      for (int j = 0; j < nz; ++j) Ja[k * nz + j] = Jk[j];
    if (Fa) Fa[k] = wa[k] / rho;
  }
  double FTa = Qa / (rho * cpa);
  if (Fa) Fa[ns] = FTa;
  if (Ja) {
    for (int j = 0; j < ns; ++j) Ja[ns * nz + j] = colha[j] / (rho * cpa) + FTa * (fabs(cp[j]) / cpa + R[j] / Rmix);
    Ja[ns * nz + ns] = (s_dhwa + colha[ns]) / (rho * cpa) + FTa * (sYdcpa / cpa + 1.0 / T);
  }
}

double *orc_nu_alloc(const orc_mech *m) {
  int ns = m->ns;
  double *nu = (double *)calloc((size_t)m->nr * ns, sizeof(double));
  for (int k = 0; k < ns; ++k)
    for (int t = m->wd_ptr[k]; t < m->wd_ptr[k + 1]; ++t) nu[(long)m->wd_rx[t] * ns + k] = m->wd_coef[t];
  return nu;
}

/* one cell, mode 0 = constant volume (x = rho_k, T), 1 = constant pressure (x = Y_k, T) */
void orc_rhs_cell(const orc_mech *m, int mode, const double *x, double p, double *F) {
  if (mode == 0) rhs_cell_r(m, x, F);
  else rhs_cp_cell_r(m, x, p, F);
}

void orc_jac_cell(const orc_mech *m, const double *nu, int mode, const double *x, double p, double *F, double *J) {
  if (mode == 0) {
    double Ja[(ORC_MAX_NS + 1) * (ORC_MAX_NS + 1)];
    jac_cell(m, nu, x, F, J, Ja, 0);
  } else {
    jac_cp_cell(m, nu, x, p, F, J, 0, 0);
  }
}

void orc_rhs_cp(const orc_mech *m, long n, long ld, const double *X, const double *p, double *F) {
  int nz = m->ns + 1;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) {
    double z[ORC_MAX_NS + 1], f[ORC_MAX_NS + 1];
    for (int k = 0; k < nz; ++k) z[k] = X[k * ld + i];
    rhs_cp_cell_r(m, z, p[i], f);
    for (int k = 0; k < nz; ++k) F[k * ld + i] = f[k];
  }
}

void orc_jac_cp(const orc_mech *m, long n, long ld, const double *X, const double *p, double *F, double *jac,
                double *fabs_out, double *jabs) {
  int nz = m->ns + 1;
  if (m->is_wd) return;
  double *nu = orc_nu_alloc(m);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) {
    double z[ORC_MAX_NS + 1], f[ORC_MAX_NS + 1], J[(ORC_MAX_NS + 1) * (ORC_MAX_NS + 1)];
    double fa[ORC_MAX_NS + 1], Ja[(ORC_MAX_NS + 1) * (ORC_MAX_NS + 1)];
    for (int k = 0; k < nz; ++k) z[k] = X[k * ld + i];
    jac_cp_cell(m, nu, z, p[i], f, J, fabs_out ? fa : 0, jabs ? Ja : 0);
    for (int k = 0; k < nz; ++k) F[k * ld + i] = f[k];
    for (int a = 0; a < nz * nz; ++a) jac[(long)a * ld + i] = J[a];
    if (fabs_out)
      for (int k = 0; k < nz; ++k) fabs_out[k * ld + i] = fa[k];
    if (jabs)
      for (int a = 0; a < nz * nz; ++a) jabs[(long)a * ld + i] = Ja[a];
  }
  free(nu);
}

void orc_jac_cp_cstep(const orc_mech *m, long n, long ld, const double *X, const double *p, double h, double *jac) {
  int nz = m->ns + 1;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) {
    double _Complex z[ORC_MAX_NS + 1], f[ORC_MAX_NS + 1];
    for (int c = 0; c < nz; ++c) {
      for (int k = 0; k < nz; ++k) z[k] = X[k * ld + i];
      z[c] += I * h;
      rhs_cp_cell_c(m, z, p[i], f);
      for (int r = 0; r < nz; ++r) jac[((long)r * nz + c) * ld + i] = cimag(f[r]) / h;
    }
  }
}
