/* Conservative <-> primitive variables of the FLINT reference -- TEST INFRASTRUCTURE ONLY.
 *
 * Restates cons2prim and prim2cons (Lib_ThermoTransport.f90:130-185) with the helpers they call:
 * f_Rtot (:217-230), f_cp (:251-271), H / E / E0 (:40-74), f_tab (:505-514) and f_tabT /
 * f_tabT_expr (:516-542), in the reference's operation order (no contraction: gcc on x86-64 emits
 * no FMA without -mfma).  Layout: cons[k*ld + i], k = 0..np-1 = (rho_1..rho_ns, rho u_1..rho u_nd,
 * rho E); prim[k*ld + i] = (rho_1..rho_ns, u_1..u_nd, p).
 */
#include <math.h>

#include "flint_oracle.h"

#define RUNIV 8314.51 /* Runiv, Lib_ThermoTransport.f90:13; Ri_tab = Runiv/wm_tab (Load_ThermoTransport.f90:54) */

/* f_tabT: Tint = idint(T), tab(Tint) + (tab(Tint+1) - tab(Tint)) * (T - Tint).  The reference
 * reads outside its table for T outside [Tmin, Tmax); here that returns 0 (a status). */
static int tabT(const orc_mech *m, const double *tab, double T, int s, double *v) {
  if (!(T >= (double)m->Th_min && T < (double)m->Th_max)) return 0;
  const int Ti = (int)T;
  const double Td = T - (double)Ti;
  const double a = tab[(long)(Ti - m->Th_min) * m->ns + s];
  const double b = tab[(long)(Ti + 1 - m->Th_min) * m->ns + s];
  *v = a + (b - a) * Td;
  return 1;
}

/* NORM2 of the velocity as flang evaluates it (the reference build's compiler; matched bit for
 * bit on 20,000 random 3-vectors): m sqrt(1 + sum over the others of (x/m)^2), with the running
 * maximum m rescaling the sum when it grows. */
static double norm2(const double *u, int nd) {
  double mx = 0.0, s = 0.0;
  for (int d = 0; d < nd; ++d) {
    const double ax = fabs(u[d]);
    if (mx == 0.0) {
      mx = ax;
    } else if (ax > mx) {
      const double t = mx / ax, tsq = t * t;
      s = s * tsq;
      s = s + tsq;
      mx = ax;
    } else {
      const double t = ax / mx;
      s = s + t * t;
    }
  }
  return mx * sqrt(1.0 + s);
}

/* sum(r) and f_Rtot(r) of one cell */
static void rho_R(const orc_mech *m, const double *r, double *rho, double *R) {
  double a = 0.0;
  for (int s = 0; s < m->ns; ++s) a = a + r[s];
  double b = 0.0;
  for (int s = 0; s < m->ns; ++s) b = b + (RUNIV / m->W[s]) * r[s] / a;
  *rho = a;
  *R = b;
}

/* E(p, r) = H(p, r) - p / sum(r), H from h_tab at T = p / (Rtot rho) (f_tab) */
static int energy_E(const orc_mech *m, double p, const double *r, double *e) {
  double rho, R;
  rho_R(m, r, &rho, &R);
  const double T = p / (R * rho);
  double ent = 0.0;
  for (int s = 0; s < m->ns; ++s) {
    double h;
    if (!tabT(m, m->h, T, s, &h)) return 0;
    ent = ent + h * r[s] / rho;
  }
  *e = ent - p / rho;
  return 1;
}

/* cons2prim per cell; iters = Newton iterations, status 0, or 2 when a table lookup leaves the
 * thermo table (outputs NaN from there). */
void orc_cons2prim(const orc_mech *m, long n, long ld, int np, const double *cons, const double *T,
                   double *prim, double *Tout, int *iters, int *status) {
  const int ns = m->ns, nd = np - ns - 1;
#pragma omp parallel for schedule(dynamic, 256)
  for (long i = 0; i < n; ++i) {
    double r[64], u[8];
    for (int s = 0; s < ns; ++s) r[s] = cons[(long)s * ld + i];
    double rho, Rgas;
    rho_R(m, r, &rho, &Rgas);
    for (int s = 0; s < ns; ++s) prim[(long)s * ld + i] = r[s];
    for (int d = 0; d < nd; ++d) {
      u[d] = cons[(long)(ns + d) * ld + i] / rho;
      prim[(long)(ns + d) * ld + i] = u[d];
    }
    const double vel = norm2(u, nd);
    const double energy = cons[(long)(np - 1) * ld + i] / rho - 0.5 * vel * vel;
    double diff = 1.0e3, Ti = T[i];
    int it = 1, st = 0;
    while (fabs(diff / energy) >= 1e-6 && it <= 20) {
      const double p_it = rho * Rgas * Ti;
      double cp = 0.0;
      for (int s = 0; s < ns && !st; ++s) {
        double c;
        if (!tabT(m, m->cp, Ti, s, &c)) st = 2;
        else cp = cp + r[s] / rho * c;
      }
      double e_it = 0.0;
      if (!st && !energy_E(m, p_it, r, &e_it)) st = 2;
      if (st) break;
      const double cv = cp - Rgas;
      diff = fabs(e_it - energy);
      Ti = Ti + (energy - e_it) / cv;
      it = it + 1;
    }
    prim[(long)(np - 1) * ld + i] = st ? NAN : rho * Rgas * Ti;
    if (Tout) Tout[i] = st ? NAN : Ti;
    if (iters) iters[i] = it - 1;
    if (status) status[i] = st;
  }
}

/* prim2cons per cell (the inverse map); status 2 when T = p / (R rho) is outside the table */
void orc_prim2cons(const orc_mech *m, long n, long ld, int np, const double *prim, double *cons, int *status) {
  const int ns = m->ns, nd = np - ns - 1;
  for (long i = 0; i < n; ++i) {
    double r[64], u[8];
    for (int s = 0; s < ns; ++s) r[s] = prim[(long)s * ld + i];
    double rho, Rgas;
    rho_R(m, r, &rho, &Rgas);
    for (int d = 0; d < nd; ++d) u[d] = prim[(long)(ns + d) * ld + i];
    const double vel = norm2(u, nd);
    for (int s = 0; s < ns; ++s) cons[(long)s * ld + i] = r[s];
    for (int d = 0; d < nd; ++d) cons[(long)(ns + d) * ld + i] = rho * u[d];
    double e = 0.0;
    const int ok = energy_E(m, prim[(long)(np - 1) * ld + i], r, &e);
    cons[(long)(np - 1) * ld + i] = ok ? rho * (e + 0.5 * vel * vel) : NAN;
    if (status) status[i] = ok ? 0 : 2;
  }
}
