/* Mixture transport of the FLINT reference -- TEST INFRASTRUCTURE ONLY.
 *
 * Restates co_k_mi_lam_Wilke / co_k_mi_lam_Wilke_expr (Lib_ThermoTransport.f90:407-462) and
 * f_laminarViscosity (:365-404) with co_fiij (:350-362), f_molecularWeight (:200-214) and
 * f_tabT_expr (:531-542), literally: the full ns x ns interaction matrix per cell, in the
 * reference's operation order (gcc on x86-64 emits no FMA without -mfma).  Tables mu_tab, k_tab
 * are [ns][nT] with T fastest, as mi_tab(Tmin:Tmax, ns) in memory; rhoi[s*ld + i].
 */
#include <math.h>
#include <stdlib.h>

#include "flint_oracle.h"

/* f_molecularWeight: 1 / sum(rhoi / sum(rhoi) / W) */
static double wmix(int ns, const double *W, const double *r, long ld) {
  double rho = 0.0, x = 0.0;
  for (int s = 0; s < ns; ++s) rho = rho + r[(long)s * ld];
  for (int s = 0; s < ns; ++s) x = x + (r[(long)s * ld] / rho / W[s]);
  return 1.0 / x;
}

/* co_fiij row i: fi(i,j) = (1 + sqrt(mi(i)/(mi(j)+1d-20)) / (Mi_Mj**0.25))**2 / sqrt(8*(1+Mi_Mj)) */
static double fiij(const double *W, const double *mi, int i, int j) {
  const double q = W[i] / W[j];
  const double t = 1 + sqrt(mi[i] / (mi[j] + 1e-20)) / pow(q, 0.25);
  return t * t / sqrt(8 * (1 + q));
}

/* Species values at T and the Wilke denominators; 0 outside [Tt_min, Tt_max) */
static int species(int ns, const double *W, int Tt_min, int Tt_max, const double *mu_tab, const double *k_tab,
                   const double *r, long ld, double rho, double T, double *X, double *mi, double *ki, double *den) {
  if (!(T >= (double)Tt_min && T < (double)Tt_max)) return 0;
  const long nT = (long)(Tt_max - Tt_min + 1);
  const int Ti = (int)T;
  const double Td = T - (double)Ti;
  const double wm = wmix(ns, W, r, ld);
  for (int s = 0; s < ns; ++s) {
    X[s] = r[(long)s * ld] * wm / (rho * W[s]);
    const double *m = mu_tab + (long)s * nT + (Ti - Tt_min), *k = k_tab + (long)s * nT + (Ti - Tt_min);
    mi[s] = m[0] + (m[1] - m[0]) * Td;
    ki[s] = k[0] + (k[1] - k[0]) * Td;
  }
  for (int i = 0; i < ns; ++i) {
    double d = 0.0;
    for (int j = 0; j < ns; ++j) d = d + X[j] * fiij(W, mi, i, j);
    den[i] = d;
  }
  return 1;
}

void orc_wilke(int ns, const double *W, int Tt_min, int Tt_max, const double *mu_tab, const double *k_tab, long n,
               long ld, const double *rhoi, const double *rho, const double *T, double *mu, double *lam, int *status) {
#pragma omp parallel
  {
    double *buf = malloc(sizeof(double) * 4 * (size_t)ns);
    double *X = buf, *mi = buf + ns, *ki = buf + 2 * ns, *den = buf + 3 * ns;
#pragma omp for schedule(static)
    for (long i = 0; i < n; ++i) {
      if (!species(ns, W, Tt_min, Tt_max, mu_tab, k_tab, rhoi + i, ld, rho[i], T[i], X, mi, ki, den)) {
        mu[i] = lam[i] = NAN;
        if (status) status[i] = 2;
        continue;
      }
      double m = 0.0, k = 0.0;
      for (int s = 0; s < ns; ++s) {
        const double inv = 1 / den[s];
        m = m + X[s] * mi[s] * inv;
        k = k + X[s] * ki[s] * inv;
      }
      mu[i] = m;
      lam[i] = k;
      if (status) status[i] = 0;
    }
    free(buf);
  }
}

void orc_laminar_viscosity(int ns, const double *W, int Tt_min, int Tt_max, const double *mu_tab, long n, long ld,
                           const double *rhoi, const double *p, const double *rho, const double *Rtot, double *mu) {
#pragma omp parallel
  {
    double *buf = malloc(sizeof(double) * 4 * (size_t)ns);
    double *X = buf, *mi = buf + ns, *ki = buf + 2 * ns, *den = buf + 3 * ns;
#pragma omp for schedule(static)
    for (long i = 0; i < n; ++i) {
      const double T = p[i] / (Rtot[i] * rho[i]);
      if (!species(ns, W, Tt_min, Tt_max, mu_tab, mu_tab, rhoi + i, ld, rho[i], T, X, mi, ki, den)) {
        mu[i] = NAN;
        continue;
      }
      double m = 0.0;
      for (int s = 0; s < ns; ++s) m = m + X[s] * mi[s] / den[s];
      mu[i] = m;
    }
    free(buf);
  }
}
