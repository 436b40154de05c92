/* CPU BDF integrator of the 0D reactor -- TEST INFRASTRUCTURE ONLY (checker and CPU baseline).
 *
 * FLINT advances its reactor through the OSlo library (setup_odesolver/run_odesolver,
 * test-batchF.f90:81,105), which is not part of the reference checkout: integrator parity is
 * unpinned.  This file and the GPU kernel implement one published method, the variable-order
 * (1..5), quasi-constant step size BDF of Shampine & Reichelt, "The MATLAB ODE Suite",
 * SIAM J. Sci. Comput. 18 (1997) (the NDF/BDF family with kappa = 0), with a modified Newton
 * iteration (at most 4 iterations, convergence-rate test), the analytic Jacobian refreshed only
 * when Newton fails, error control in the RMS norm of e / (atol + rtol |y|), and the step/order
 * selection of that paper.  The GPU path must reproduce this algorithm; the oracle is pinned by
 * tight reference solutions (scipy Radau at rtol 1e-12, tests/test_bdf.py) and by invariants
 * (mass, elements, and internal energy / enthalpy).  The step-size factor series of Newton
 * policy 1 (step_pow) is pinned against libm's pow (tests/test_bdf.py). */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "flint_oracle.h"

#define MAXO 5
#define NEWTON_MAXITER 4
#define MIN_FACTOR 0.2
#define MAX_FACTOR 10.0
#define NZMAX 65

enum { ST_STEPS, ST_REJECT, ST_NFAIL, ST_RHS, ST_JAC, ST_LU, ST_NEWTON, ST_ORDER };

/* Step-size factor x^e of the controller.  method 0 takes scipy's pow.  method 1 (Newton policy 1)
 * takes the GPU kernel's evaluation, the same IEEE double operations in the same order (explicit
 * fma): x = m 2^k with m in [sqrt(1/2), sqrt(2)), log2 x = k + 2 atanh(s) / ln 2 with
 * s = (m - 1)/(m + 1) (series to s^13), 2^t = 2^n e^(f ln 2) with n = rint(t) (series to f^10);
 * about 1e-12 relative, and the kernel's factors bit for bit.  Under method 0 the oracle stays on
 * scipy's pow while the kernel keeps its series, so the two differ there by about 1e-12. */
static double step_pow(double x, double e, int method) {
  if (!method) return pow(x, e);
  if (x == 0.0) return INFINITY;
  if (!(x <= DBL_MAX)) return x == INFINITY ? 0.0 : x;
  int k;
  double m = frexp(x, &k);
  if (m < 0.70710678118654752) {
    m = m * 2.0;
    k = k - 1;
  }
  const double s = (m - 1.0) / (m + 1.0), s2 = s * s;
  double p = 1.0 / 13.0;
  p = fma(p, s2, 1.0 / 11.0);
  p = fma(p, s2, 1.0 / 9.0);
  p = fma(p, s2, 1.0 / 7.0);
  p = fma(p, s2, 1.0 / 5.0);
  p = fma(p, s2, 1.0 / 3.0);
  p = fma(p, s2, 1.0);
  const double l2 = fma(2.0 * s * p, 1.4426950408889634, (double)k);
  const double t = fmin(fmax(e * l2, -1074.0), 1024.0);
  const double n = rint(t), y = (t - n) * 0.6931471805599453;
  double q = 1.0 / 3628800.0;
  q = fma(q, y, 1.0 / 362880.0);
  q = fma(q, y, 1.0 / 40320.0);
  q = fma(q, y, 1.0 / 5040.0);
  q = fma(q, y, 1.0 / 720.0);
  q = fma(q, y, 1.0 / 120.0);
  q = fma(q, y, 1.0 / 24.0);
  q = fma(q, y, 1.0 / 6.0);
  q = fma(q, y, 0.5);
  q = fma(q, y, 1.0);
  q = fma(q, y, 1.0);
  return ldexp(q, (int)n);
}

/* exported for tests/test_bdf.py, which pins the series against pow */
double orc_step_pow(double x, double e, int method) { return step_pow(x, e, method); }

static double rms(const double *v, const double *scale, int n) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) {
    double q = v[i] / scale[i];
    s += q * q;
  }
  return sqrt(s / n);
}

/* R(order, factor): M[0][*] = 1, M[i][j] = (i - 1 - factor j)/i, cumulative product down rows.
 * rcp: the division by i taken as a product with the constant 1/i, as the GPU kernel does (Newton
 * policy 1); scipy divides (policy 0, whose step sequence is pinned to scipy's). */
static void compute_R(int order, double factor, double R[MAXO + 1][MAXO + 1], int rcp) {
  for (int j = 0; j <= order; ++j) R[0][j] = 1.0;
  for (int i = 1; i <= order; ++i)
    for (int j = 0; j <= order; ++j) {
      double mij = (j == 0) ? 0.0 : (rcp ? (i - 1 - factor * j) * (1.0 / i) : (i - 1 - factor * j) / i);
      R[i][j] = R[i - 1][j] * mij;
    }
}

/* D[0..order] <- (R U)^T D[0..order] : rescales the difference array to step factor*h */
static void change_D(double D[MAXO + 3][NZMAX], int n, int order, double factor, int rcp) {
  double R[MAXO + 1][MAXO + 1], U[MAXO + 1][MAXO + 1], RU[MAXO + 1][MAXO + 1];
  compute_R(order, factor, R, rcp);
  compute_R(order, 1.0, U, rcp);
  for (int i = 0; i <= order; ++i)
    for (int j = 0; j <= order; ++j) {
      double s = 0.0;
      for (int l = 0; l <= order; ++l) s += R[i][l] * U[l][j];
      RU[i][j] = s;
    }
  double tmp[MAXO + 1][NZMAX];
  for (int i = 0; i <= order; ++i)
    for (int c = 0; c < n; ++c) {
      double s = 0.0;
      for (int j = 0; j <= order; ++j) s += RU[j][i] * D[j][c];
      tmp[i][c] = s;
    }
  for (int i = 0; i <= order; ++i) memcpy(D[i], tmp[i], sizeof(double) * n);
}

/* LU with partial pivoting of A (n x n row-major, in place); piv[k] = row chosen at step k */
static int lu_factor(double *A, int n, int *piv) {
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = fabs(A[k * n + k]);
    for (int i = k + 1; i < n; ++i)
      if (fabs(A[i * n + k]) > best) { best = fabs(A[i * n + k]); p = i; }
    piv[k] = p;
    if (best == 0.0) return 1;
    if (p != k)
      for (int j = 0; j < n; ++j) { double t = A[k * n + j]; A[k * n + j] = A[p * n + j]; A[p * n + j] = t; }
    double inv = 1.0 / A[k * n + k];
    for (int i = k + 1; i < n; ++i) {
      double l = A[i * n + k] * inv;
      A[i * n + k] = l;
      for (int j = k + 1; j < n; ++j) A[i * n + j] -= l * A[k * n + j];
    }
  }
  return 0;
}

/* rows were swapped whole (multipliers included), so apply every interchange first, then L, U */
static void lu_solve(const double *A, int n, const int *piv, double *b) {
  for (int k = 0; k < n; ++k)
    if (piv[k] != k) { double t = b[k]; b[k] = b[piv[k]]; b[piv[k]] = t; }
  for (int k = 0; k < n; ++k)
    for (int i = k + 1; i < n; ++i) b[i] -= A[i * n + k] * b[k];
  for (int k = n - 1; k >= 0; --k) {
    double s = b[k];
    for (int j = k + 1; j < n; ++j) s -= A[k * n + j] * b[j];
    b[k] = s / A[k * n + k];
  }
}

static int finite_vec(const double *v, int n) {
  for (int i = 0; i < n; ++i)
    if (!isfinite(v[i])) return 0;
  return 1;
}

typedef struct {
  const orc_mech *m;
  const double *nu;
  int mode;
  double p;
  int *st;
} ctx_t;

/* rhs with the out-of-table convention turned into a failure flag (F = -1 in rhs_native) */
static int fun(ctx_t *c, const double *x, double *F) {
  int nz = c->m->ns + 1;
  double T = x[nz - 1];
  c->st[ST_RHS]++;
  if (!(T >= c->m->Th_min && T < c->m->Th_max) || !(T >= c->m->Tk_min && T < c->m->Tk_max)) return 1;
  orc_rhs_cell(c->m, c->mode, x, c->p, F);
  return !finite_vec(F, nz);
}

static void jac(ctx_t *c, const double *x, double *J) {
  double F[NZMAX];
  c->st[ST_JAC]++;
  orc_jac_cell(c->m, c->nu, c->mode, x, c->p, F, J);
}

/* Newton policies of the modified Newton iteration:
 *   0  as scipy's BDF (Shampine & Reichelt): at most 4 iterations, the convergence rate measured
 *      within the step (so at least two iterations), converged when rate/(1-rate) |dy| < newton_tol,
 *      the iteration matrix refactored at every change of h or order;
 *   1  as SUNDIALS CVODE's BDF (the integrator FLINT's OSlo wraps as 'cvode', test-batchF.f90:22-26):
 *      at most 3 iterations, the rate estimate carried over from earlier steps (crate, reset to 1
 *      when the matrix is refactored, decayed by 0.3 per iteration), converged when
 *      |dy| min(1, crate) C_q <= 0.005 with C_q = 1/(q+1) the error-test constant, diverged when
 *      |dy| > 2 |dy_prev|; the factorised I - c_lu J is kept while |c/c_lu - 1| <= 0.3 and fewer
 *      than 20 converged attempts old, its corrections scaled by 2/(1 + c/c_lu).  CVODE's own
 *      coefficient is 0.1; with it the Newton error enters the error estimate and the Rich-31
 *      reactors reject 25x more steps, so the iteration is held 20x tighter. */
#define CV_MAXCOR 3
#define CV_DGMAX 0.3
#define CV_MSBP 20
#define CV_CRDOWN 0.3
#define CV_RDIV 2.0
#define CV_NLSCOEF 0.005

/* Returns 0 on success, 1 step size underflow, 2 too many steps, 3 invalid state. */
static int bdf_cell(const orc_mech *m, const double *nu, int mode, double *y, double p, double rtol, double atol,
                    double t_end, int max_steps, int *st, int method) {
  const int n = m->ns + 1;
  ctx_t c = {m, nu, mode, p, st};
  double gamma[MAXO + 1], alpha[MAXO + 1], err_c[MAXO + 2];
  gamma[0] = 0.0;
  for (int k = 1; k <= MAXO; ++k) gamma[k] = gamma[k - 1] + 1.0 / k;
  for (int k = 0; k <= MAXO; ++k) alpha[k] = gamma[k];
  for (int k = 0; k <= MAXO + 1; ++k) err_c[k] = 1.0 / (k + 1);
  const double newton_tol = fmax(10.0 * DBL_EPSILON / rtol, fmin(0.03, sqrt(rtol)));

  double f0[NZMAX], f1[NZMAX], y1[NZMAX], scale[NZMAX];
  if (fun(&c, y, f0)) return 3;
  /* initial step (Hairer-Norsett-Wanner II.4, as in scipy select_initial_step) */
  for (int i = 0; i < n; ++i) scale[i] = atol + fabs(y[i]) * rtol;
  double d0 = rms(y, scale, n), d1 = rms(f0, scale, n), h0;
  h0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
  h0 = fmin(h0, t_end);
  for (int i = 0; i < n; ++i) y1[i] = y[i] + h0 * f0[i];
  if (fun(&c, y1, f1)) return 3;
  double diff[NZMAX];
  for (int i = 0; i < n; ++i) diff[i] = f1[i] - f0[i];
  double d2 = rms(diff, scale, n) / h0, h1;
  if (d1 <= 1e-15 && d2 <= 1e-15) h1 = fmax(1e-6, h0 * 1e-3);
  else h1 = pow(0.01 / fmax(d1, d2), 0.5);
  double h_abs = fmin(100.0 * h0, fmin(h1, t_end));

  static __thread double D[MAXO + 3][NZMAX];
  double J[NZMAX * NZMAX], LU[NZMAX * NZMAX];
  int piv[NZMAX];
  memset(D, 0, sizeof D);
  for (int i = 0; i < n; ++i) { D[0][i] = y[i]; D[1][i] = f0[i] * h_abs; }
  int order = 1, n_equal = 0, have_lu = 0;
  jac(&c, y, J);
  int current_jac = 0;
  double t = 0.0;
  /* method 1: coefficient of the factorised matrix, steps since, carried rate estimate */
  double c_lu = 0.0, crate = 1.0;
  int lu_age = 0;

  while (t < t_end) {
    if (st[ST_STEPS] >= max_steps) return 2;
    double min_step = 10.0 * fabs(nextafter(t, INFINITY) - t);
    if (h_abs < min_step) {
      change_D(D, n, order, min_step / h_abs, method);
      h_abs = min_step;
      n_equal = 0;
    }
    double ynew[NZMAX], d[NZMAX], ypred[NZMAX], psi[NZMAX], fv[NZMAX], dy[NZMAX];
    double t_new, err_norm = 0.0, safety = 0.9;
    int accepted = 0;
    while (!accepted) {
      if (h_abs < min_step) return 1;
      t_new = t + h_abs;
      if (t_new - t_end > 0) {
        t_new = t_end;
        change_D(D, n, order, fabs(t_new - t) / h_abs, method);
        n_equal = 0;
        if (method == 0) have_lu = 0;
      }
      double h = t_new - t;
      h_abs = fabs(h);
      for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int j = 0; j <= order; ++j) s += D[j][i];
        ypred[i] = s;
        scale[i] = atol + rtol * fabs(ypred[i]);
        double q = 0.0;
        for (int j = 1; j <= order; ++j) q += D[j][i] * gamma[j];
        psi[i] = q / alpha[order];
      }
      int converged = 0, n_iter = 0;
      double cc = h / alpha[order];
      while (!converged) {
        if (method == 1 && have_lu && (fabs(cc / c_lu - 1.0) > CV_DGMAX || lu_age >= CV_MSBP)) have_lu = 0;
        if (!have_lu) {
          for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) LU[i * n + j] = (i == j ? 1.0 : 0.0) - cc * J[i * n + j];
          st[ST_LU]++;
          lu_factor(LU, n, piv);
          have_lu = 1;
          c_lu = cc;
          lu_age = 0;
          crate = 1.0;
        }
        const double corr = method == 1 ? 2.0 / (1.0 + cc / c_lu) : 1.0;
        /* modified Newton on  y - ypred - c f(y) + psi = 0, accumulated correction d */
        for (int i = 0; i < n; ++i) { ynew[i] = ypred[i]; d[i] = 0.0; }
        double dy_norm_old = -1.0;
        converged = 0;
        int k;
        const int maxit = method == 1 ? CV_MAXCOR : NEWTON_MAXITER;
        for (k = 0; k < maxit; ++k) {
          if (fun(&c, ynew, fv)) break;
          for (int i = 0; i < n; ++i) dy[i] = cc * fv[i] - psi[i] - d[i];
          lu_solve(LU, n, piv, dy);
          if (method == 1)
            for (int i = 0; i < n; ++i) dy[i] *= corr;
          st[ST_NEWTON]++;
          double dy_norm = rms(dy, scale, n);
          if (method == 1) {
            if (k > 0) {
              if (dy_norm > CV_RDIV * dy_norm_old) break;
              crate = fmax(CV_CRDOWN * crate, dy_norm / dy_norm_old);
            }
            for (int i = 0; i < n; ++i) { ynew[i] += dy[i]; d[i] += dy[i]; }
            if (dy_norm * fmin(1.0, crate) * err_c[order] <= CV_NLSCOEF) {
              converged = 1;
              break;
            }
            dy_norm_old = dy_norm;
            continue;
          }
          double rate = -1.0;
          if (dy_norm_old >= 0.0) rate = dy_norm / dy_norm_old;
          if (rate >= 0.0 && (rate >= 1.0 || pow(rate, NEWTON_MAXITER - k) / (1.0 - rate) * dy_norm > newton_tol))
            break;
          for (int i = 0; i < n; ++i) { ynew[i] += dy[i]; d[i] += dy[i]; }
          if (dy_norm == 0.0 || (rate >= 0.0 && rate / (1.0 - rate) * dy_norm < newton_tol)) {
            converged = 1;
            break;
          }
          dy_norm_old = dy_norm;
        }
        n_iter = k + 1;
        if (!converged) {
          if (current_jac) break;
          jac(&c, ypred, J);
          have_lu = 0;
          current_jac = 1;
        }
      }
      if (!converged) {
        st[ST_NFAIL]++;
        h_abs *= 0.5;
        change_D(D, n, order, 0.5, method);
        n_equal = 0;
        if (method == 0) have_lu = 0;
        continue;
      }
      if (method == 1) lu_age++;
      safety = 0.9 * (2 * NEWTON_MAXITER + 1) / (double)(2 * NEWTON_MAXITER + n_iter);
      for (int i = 0; i < n; ++i) {
        scale[i] = atol + rtol * fabs(ynew[i]);
        dy[i] = err_c[order] * d[i];
      }
      err_norm = rms(dy, scale, n);
      if (err_norm > 1.0) {
        st[ST_REJECT]++;
        double factor = fmax(MIN_FACTOR, safety * step_pow(err_norm, -1.0 / (order + 1), method));
        h_abs *= factor;
        change_D(D, n, order, factor, method);
        n_equal = 0;
      } else {
        accepted = 1;
      }
    }
    st[ST_STEPS]++;
    n_equal++;
    t = t_new;
    for (int i = 0; i < n; ++i) y[i] = ynew[i];
    current_jac = 0;
    /* D^{j+1} y_n = D^j y_n - D^j y_{n-1}; d = D^{order+1} y_n */
    for (int i = 0; i < n; ++i) {
      D[order + 2][i] = d[i] - D[order + 1][i];
      D[order + 1][i] = d[i];
    }
    for (int j = order; j >= 0; --j)
      for (int i = 0; i < n; ++i) D[j][i] += D[j + 1][i];
    if (n_equal < order + 1) continue;
    double em = INFINITY, ep = INFINITY, tmp[NZMAX];
    if (order > 1) {
      for (int i = 0; i < n; ++i) tmp[i] = err_c[order - 1] * D[order][i];
      em = rms(tmp, scale, n);
    }
    if (order < MAXO) {
      for (int i = 0; i < n; ++i) tmp[i] = err_c[order + 1] * D[order + 2][i];
      ep = rms(tmp, scale, n);
    }
    double norms[3] = {em, err_norm, ep}, best = -1.0;
    int which = 1;
    for (int q = 0; q < 3; ++q) {
      double fct = (norms[q] == 0.0) ? INFINITY : step_pow(norms[q], -1.0 / (order + q), method);
      if (fct > best) { best = fct; which = q; }
    }
    order += which - 1;
    double factor = fmin(MAX_FACTOR, safety * best);
    h_abs *= factor;
    change_D(D, n, order, factor, method);
    n_equal = 0;
    if (method == 0) have_lu = 0;
  }
  st[ST_ORDER] = order;
  return 0;
}

/* Advance n cells (SoA X, leading dimension ld) from t = 0 to t_end.  mode 0: X = (rho_k, T),
 * constant volume (rhs_native); mode 1: X = (Y_k, T), constant pressure p[i].  stats[i*8 + q]
 * counts steps, rejected steps, Newton failures, RHS, Jacobians, LU, Newton iterations, final
 * order; status[i] is 0 on success. */
void orc_bdf_method(const orc_mech *m, long n, long ld, int mode, double *X, const double *p, double rtol,
                    double atol, double t_end, int max_steps, int *stats, int *status, int method) {
  const int nz = m->ns + 1;
  double *nu = orc_nu_alloc(m);
#pragma omp parallel for schedule(dynamic, 16)
  for (long i = 0; i < n; ++i) {
    double y[NZMAX];
    int st[8] = {0};
    for (int k = 0; k < nz; ++k) y[k] = X[k * ld + i];
    int rc = bdf_cell(m, nu, mode, y, p ? p[i] : 0.0, rtol, atol, t_end, max_steps, st, method);
    for (int k = 0; k < nz; ++k) X[k * ld + i] = y[k];
    for (int q = 0; q < 8; ++q) stats[i * 8 + q] = st[q];
    status[i] = rc;
  }
  free(nu);
}

void orc_bdf(const orc_mech *m, long n, long ld, int mode, double *X, const double *p, double rtol, double atol,
             double t_end, int max_steps, int *stats, int *status) {
  orc_bdf_method(m, n, ld, mode, X, p, rtol, atol, t_end, max_steps, stats, status, 0);
}
