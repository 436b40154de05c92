/* CEA-style chemical equilibrium, CPU restatement -- TEST INFRASTRUCTURE ONLY.
 *
 * Restates the reference's CEA path for gas-phase mixtures (no condensed species, no ions,
 * Trace = 0), which is the path every FLINT mechanism takes:
 *   CEA_initialize_global   Lib_CEA_setup.f90:6-99    (A = species_composition, Size)
 *   CEA_initialize_local    Lib_CEA_setup.f90:102-155 (Vv, b0, Hsub0, Bcheck)
 *   CEA_solve               Lib_CEA_solver.f90:6-856  (reduced Newton on pi_i, dln n, dln T;
 *                                                      damping lambda; convergence tests)
 *   MATRIX                  Lib_CEA_solver.f90:857-981
 *   GAUSS                   Lib_CEA_solver.f90:983-1088
 *   COMP_THERMO_QUANTS      Lib_ThermoTransport.f90:100-115 (cp/R, h/RT, s/R from the tables)
 * with the reference's operation order and its single-precision literals (Ennl = -2.3025851,
 * 0.5E-5, smalno = 1.E-6, smnol = -13.815511, bigno = 1.E+25).  Branches that only condensed
 * species, ions or a singular basis after three retries reach (component switching,
 * Lib_CEA_solver.f90:725-848, which also rewrites the reference's global element matrix) are
 * reported with status 2 instead of being taken.
 *
 * Tightened mode (c->tol > 0, no counterpart in the reference): the convergence tests of
 * Lib_CEA_solver.f90:345-364 (relative corrections of n and of the element potentials <= 0.5e-5,
 * |dln T| <= 1e-4, element residual <= 1e-6 max b0) all at tol, so that the converged state no
 * longer depends on the starting point (a warm start and the default guess land within round-off
 * of each other); everything else is the reference's iteration.
 * Status: 0 converged, 1 more than maxitn = 50 iterations (the reference leaves T_out unset),
 * 2 unsupported branch, 3 temperature outside the thermo table. */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "flint_oracle.h"

#define CEA_MAXE 8
#define CEA_MAXS 64
#define CEA_RR 8314.51                 /* Rr, Lib_CEA_data.f90:65 */
#define CEA_SIZE 18.420681             /* Size, Lib_CEA_setup.f90:97 */

/* COMP_THERMO_QUANTS: cp/R, h/(R T), s/R per species at T (f_tabT interpolation) */
static int thermo(const orc_mech *m, double T, double *Cp, double *H0, double *S) {
  if (!(T >= m->Th_min) || !(T < m->Th_max)) return 3;
  int Ti = (int)T;
  double Tdiff = T - Ti;
  long r0 = (long)(Ti - m->Th_min) * m->ns, r1 = r0 + m->ns;
  for (int j = 0; j < m->ns; ++j) {
    double cp = m->cp[r0 + j] + (m->cp[r1 + j] - m->cp[r0 + j]) * Tdiff;
    double h = m->h[r0 + j] + (m->h[r1 + j] - m->h[r0 + j]) * Tdiff;
    double s = m->s[r0 + j] + (m->s[r1 + j] - m->s[r0 + j]) * Tdiff;
    Cp[j] = cp * m->W[j] / CEA_RR;
    H0[j] = h * m->W[j] / (CEA_RR * T);
    S[j] = s * m->W[j] / CEA_RR;
  }
  return 0;
}

/* GAUSS: scaled partial pivoting on the row with the smallest max |G(i,j>nn)|/|G(i,nn)|.
 * G is [imat][imat+1] row-major (1-based indices in the comments).  Returns Msing (0 = ok). */
static int gauss(double G[][CEA_MAXE + 4], int imat, double *X) {
  const double bigno = (double)1.E+25f;
  double coefx[CEA_MAXE + 4];
  int imatp1 = imat + 1;
  for (int nn = 0; nn < imat; ++nn) {
    if (nn != imat - 1) {
      for (int i = nn; i < imat; ++i) {
        coefx[i] = bigno;
        if (G[i][nn] != 0.0) {
          coefx[i] = 0.0;
          for (int j = nn + 1; j < imatp1; ++j) coefx[i] = fmax(coefx[i], fabs(G[i][j]));
          double tmp = fabs(G[i][nn]);
          if (bigno * tmp > coefx[i]) coefx[i] = coefx[i] / tmp;
          else coefx[i] = bigno;
        }
      }
      double tmp = bigno;
      int ip = -1;
      for (int j = nn; j < imat; ++j)
        if (coefx[j] < tmp) { tmp = coefx[j]; ip = j; }
      if (ip < 0) return nn + 1;
      if (ip != nn)
        for (int j = nn; j < imatp1; ++j) { double t = G[ip][j]; G[ip][j] = G[nn][j]; G[nn][j] = t; }
    } else if (G[nn][nn] == 0.0) {
      return nn + 1;
    }
    double tmp = G[nn][nn];
    if (tmp == 0.0) return nn + 1;
    for (int j = nn + 1; j < imatp1; ++j) G[nn][j] = G[nn][j] / tmp;
    if (nn + 1 != imatp1)
      for (int i = nn + 1; i < imat; ++i)
        for (int j = nn + 1; j < imatp1; ++j) G[i][j] = G[i][j] - G[i][nn] * G[nn][j];
  }
  for (int k = imat - 1; k >= 0; --k) {
    double tmp = 0.0;
    for (int i = k + 1; i < imat; ++i) tmp = tmp + G[k][i] * X[i];
    X[k] = G[k][imatp1 - 1] - tmp;
  }
  return 0;
}

int orc_cea_cell(const orc_mech *m, const orc_cea_data *c, int hp, double p_in, double T_in, const double *rho_in,
                 const double *guess, double *T_out, double *Y_out, int *iters) {
  const int ns = m->ns, nlm = c->ne, ng = ns;
  const int vol = !hp;
  double Cp[CEA_MAXS], H0[CEA_MAXS], S[CEA_MAXS], Mu[CEA_MAXS], En[CEA_MAXS], Enln[CEA_MAXS], Deln[CEA_MAXS];
  double B0[CEA_MAXE], X[CEA_MAXE + 4];
  double G[CEA_MAXE + 3][CEA_MAXE + 4];
  *iters = 0;
  if (ns > CEA_MAXS || nlm > CEA_MAXE) return 2;

  /* CEA_initialize_local (Lib_CEA_setup.f90:102-155) */
  double sumr = 0.0;
  for (int n = 0; n < ns; ++n) sumr += rho_in[n];
  const double Vv = 1e05 / sumr;
  double Wp = 0.0, hpp = 0.0;
  for (int j = 0; j < nlm; ++j) B0[j] = 0.0;
  if (thermo(m, T_in, Cp, H0, S)) return 3;
  for (int n = 0; n < ns; ++n) {
    double rm = 0.0;
    for (int jj = 0; jj < nlm; ++jj) rm = rm + c->A[jj * ns + n] * c->elw[jj];
    if (rho_in[n] <= 1e-15) continue;
    Wp = Wp + rho_in[n];
    /* H0i(n,Tt)*wm_tab(n)/Rr: the per-mass table value, interpolated */
    int Ti = (int)T_in;
    double Tdiff = T_in - Ti;
    long r0 = (long)(Ti - m->Th_min) * ns, r1 = r0 + ns;
    double hi = m->h[r0 + n] + (m->h[r1 + n] - m->h[r0 + n]) * Tdiff;
    double enth = hi * m->W[n] / CEA_RR;
    if (vol) enth = enth - T_in;
    hpp = hpp + enth * rho_in[n] / rm;
    for (int j = 0; j < nlm; ++j) B0[j] = c->A[j * ns + n] * rho_in[n] / rm + B0[j];
  }
  const double Hsub0 = hpp;
  if (Wp != 0.0)
    for (int j = 0; j < nlm; ++j) B0[j] = B0[j] / Wp;
  double bigb = 0.0;
  for (int j = 0; j < nlm; ++j) bigb = fmax(bigb, fabs(B0[j]));
  /* tightened mode: the element residual relative to the largest b0 at tol as well */
  const double Bcheck = c->tol > 0.0 ? bigb * c->tol : bigb * .000001;

  /* CEA_solve: initial estimates (Lib_CEA_solver.f90:45-62) */
  double Pp = p_in, Tt = T_in;
  double Enn = .1, Ennl = (double)-2.3025851f, Sumn = Enn;
  double xi = (double)ng;
  xi = Enn / xi;
  double xln = log(xi);
  for (int j = 0; j < ng; ++j) { En[j] = xi; Enln[j] = xln; }
  if (guess) {
    /* warm start: T and ln n_j from the caller; the totals follow from them */
    Tt = guess[0];
    Sumn = 0.0;
    for (int j = 0; j < ng; ++j) {
      Enln[j] = guess[1 + j];
      En[j] = exp(Enln[j]);
      Sumn += En[j];
    }
    Enn = Sumn;
    Ennl = log(Enn);
  }
  int ixsing = 0, numb = 0;
  const int maxitn = 50;
  const double siz9 = CEA_SIZE - 9.2103404;
  double tsize = CEA_SIZE;
  const double smalno = (double)1.E-6f, smnol = (double)-13.815511f;
  int convg = 0;

  /* label 400 */
  double Tln = log(Tt);
  if (vol) Pp = CEA_RR * Enn * Tt / Vv;
  if (thermo(m, Tt, Cp, H0, S)) return 3;
  double Tm = log(Pp / Enn);
  for (;;) {
    /* label 500 */
    double Cpsum = 0.0;
    for (int j = 0; j < ng; ++j) Cpsum = Cpsum + En[j] * Cp[j];
    numb = numb + 1;
    /* MATRIX (Lib_CEA_solver.f90:857-981), Npr = 0, Tp = .false., Hp = .true. */
    const int iq = nlm, Iq1 = iq, iq2 = iq + 1, iq3 = iq + 2; /* 0-based columns */
    int kmat = iq3 + 1, imat = kmat - 1;                      /* counts */
    for (int i = 0; i < imat; ++i)
      for (int k = 0; k < kmat; ++k) G[i][k] = 0.0;
    G[iq2][Iq1] = 0.0;
    double Hsum = 0.0;
    for (int j = 0; j < ng; ++j) {
      Mu[j] = H0[j] - S[j] + Enln[j] + Tm;
      if (En[j] != 0.0) {
        double h = H0[j] * En[j];
        double f = Mu[j] * En[j];
        double term1 = h;
        for (int i = 0; i < nlm; ++i) {
          double aij = c->A[i * ns + j];
          if (aij != 0.0) {
            double term = aij * En[j];
            for (int k = i; k < nlm; ++k) G[i][k] = G[i][k] + c->A[k * ns + j] * term;
            G[i][Iq1] = G[i][Iq1] + term;
            G[i][iq2] = G[i][iq2] + aij * term1;
            G[i][iq3] = G[i][iq3] + aij * f;
          }
        }
        G[iq2][iq2] = G[iq2][iq2] + H0[j] * h;
        G[iq2][iq3] = G[iq2][iq3] + H0[j] * f;
        G[Iq1][iq3] = G[Iq1][iq3] + f;
        G[Iq1][iq2] = G[Iq1][iq2] + term1;
      }
    }
    Hsum = Hsum + G[Iq1][iq2];
    G[Iq1][Iq1] = Sumn - Enn;
    for (int i = 0; i <= iq2; ++i)
      for (int j = i; j <= iq2; ++j) G[j][i] = G[i][j];
    for (int i = 0; i < nlm; ++i) G[i][kmat - 1] = G[i][kmat - 1] + B0[i] - G[i][Iq1];
    G[Iq1][kmat - 1] = G[Iq1][kmat - 1] + Enn - Sumn;
    double energyl = Hsub0 / Tt - Hsum;
    G[iq2][iq3] = G[iq2][iq3] + energyl;
    G[iq2][iq2] = G[iq2][iq2] + Cpsum;
    if (vol) {
      /* constant-volume matrix (Lib_CEA_solver.f90:964-979) */
      for (int i = 0; i < iq; ++i) {
        G[Iq1][i] = G[iq2][i] - G[Iq1][i];
        G[i][Iq1] = G[i][iq2] - G[i][Iq1];
        G[i][iq2] = G[i][iq3];
      }
      G[Iq1][Iq1] = G[iq2][iq2] - G[Iq1][iq2] - G[iq2][Iq1];
      G[Iq1][iq2] = G[iq2][iq3] - G[Iq1][iq3];
      G[Iq1][iq2] = G[Iq1][iq2] + Enn;
      kmat = imat;
      imat = imat - 1;
    }
    /* GAUSS writes X(1..imat); the augmented column is kmat (= imat + 1) */
    int msing = gauss(G, imat, X);
    if (msing == 0) {
      /* corrections (Lib_CEA_solver.f90:193-307) */
      if (vol) X[iq2] = X[Iq1];
      double dlnt = X[iq2];
      double sum = X[Iq1];
      if (vol) {
        X[Iq1] = 0.0;
        sum = -dlnt;
      }
      for (int j = 0; j < ng; ++j) {
        Deln[j] = -Mu[j] + H0[j] * dlnt + sum;
        for (int k = 0; k < nlm; ++k) Deln[j] = Deln[j] + c->A[k * ns + j] * X[k];
      }
      double ambda = 1.0, ambda1 = 1.0;
      sum = fmax(fabs(X[Iq1]), fabs(dlnt));
      sum = sum * 5.;
      for (int j = 0; j < ng; ++j) {
        if (Deln[j] > 0.0) {
          if ((Enln[j] - Ennl + CEA_SIZE) <= 0.0) {
            double sum1 = fabs(Deln[j] - X[Iq1]);
            if (sum1 >= siz9) {
              sum1 = fabs(-9.2103404 - Enln[j] + Ennl) / sum1;
              if (sum1 < ambda1) ambda1 = sum1;
            }
          } else if (Deln[j] > sum) {
            sum = Deln[j];
          }
        }
      }
      if (sum > 2.0) ambda = 2.0 / sum;
      if (ambda1 <= ambda) ambda = ambda1;
      double Totn = 0.0;
      for (int j = 0; j < ng; ++j) Enln[j] = Enln[j] + ambda * Deln[j];
      for (int j = 0; j < ng; ++j) {
        En[j] = 0.0;
        if ((Enln[j] - Ennl + tsize) > 0.0) {
          En[j] = exp(Enln[j]);
          Totn = Totn + En[j];
        }
      }
      Sumn = Totn;
      Tln = Tln + ambda * dlnt;
      Tt = exp(Tln);
      if (thermo(m, Tt, Cp, H0, S)) { *iters = numb; return 3; }
      if (vol) {
        Enn = Sumn;
        Ennl = log(Enn);
        Pp = CEA_RR * Tt * Enn / Vv;
      } else {
        Ennl = Ennl + ambda * X[Iq1];
        Enn = exp(Ennl);
      }
      Tm = log(Pp / Enn);
      /* convergence tests (Lib_CEA_solver.f90:311-368) */
      if (numb > maxitn) { *iters = numb; return 1; }
      sum = (X[Iq1] * Enn / Totn);
      if (fabs(sum) > (c->tol > 0.0 ? c->tol : (double)0.5E-5f)) continue;
      int again = 0;
      for (int j = 0; j < ng && !again; ++j)
        if (fabs(Deln[j]) * En[j] / Totn > (c->tol > 0.0 ? c->tol : 0.5e-5)) again = 1;
      if (again) continue;
      if (fabs(dlnt) > (c->tol > 0.0 ? c->tol : 1.e-04)) continue;
      for (int i = 0; i < nlm && !again; ++i) {
        if (fabs(B0[i]) >= 1.e-06) {
          double s = 0.0;
          for (int j = 0; j < ng; ++j) s = s + En[j] * c->A[i * ns + j];
          if (fabs(B0[i] - s) > Bcheck) again = 1;
        }
      }
      if (again) continue;
      convg = 1;
      break;
    }
    /* singular matrix (Lib_CEA_solver.f90:436-536) */
    ixsing = ixsing + 1;
    if (ixsing > 8) { *iters = numb; return 2; }
    tsize = 80.0;
    if (ixsing >= 3 && msing < Iq1 + 1) { *iters = numb; return 2; } /* component switch */
    for (int j = 0; j < ng; ++j)
      if (En[j] == 0.0) { En[j] = smalno; Enln[j] = smnol; }
  }
  (void)convg;
  *iters = numb;
  *T_out = Tt;
  for (int j = 0; j < ns; ++j) Y_out[j] = En[j] * m->W[j];
  return 0;
}

void orc_cea(const orc_mech *m, const orc_cea_data *c, long n, long ld, int hp, const double *p, const double *T,
             const double *rho, const double *guess, double *T_out, double *Y_out, int *iters, int *status) {
  const int ns = m->ns;
#pragma omp parallel for schedule(dynamic, 16)
  for (long i = 0; i < n; ++i) {
    double r[CEA_MAXS], y[CEA_MAXS], g[CEA_MAXS + 1], to = NAN;
    for (int k = 0; k < ns; ++k) r[k] = rho[k * ld + i];
    if (guess)
      for (int k = 0; k <= ns; ++k) g[k] = guess[k * ld + i];
    int it = 0;
    int rc = orc_cea_cell(m, c, hp, p[i], T[i], r, guess ? g : NULL, &to, y, &it);
    if (rc != 0) {
      to = NAN;
      for (int k = 0; k < ns; ++k) y[k] = NAN;
    }
    T_out[i] = to;
    for (int k = 0; k < ns; ++k) Y_out[k * ld + i] = y[k];
    iters[i] = it;
    status[i] = rc;
  }
}
