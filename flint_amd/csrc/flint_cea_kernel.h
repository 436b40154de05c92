// Batched CEA-style chemical equilibrium on gfx950: one thread per state (kernel template,
// instantiated per mechanism by the generated translation units).
//
// The reference's equilibrium solver for gas mixtures (CEA_initialize_local,
// Lib_CEA_setup.f90:102-155; CEA_solve with MATRIX and GAUSS, Lib_CEA_solver.f90:6-1088;
// COMP_THERMO_QUANTS, Lib_ThermoTransport.f90:100-115): Newton on the element potentials pi_i,
// dln n and dln T with the reference's damping, trace cut-off and convergence tests, HP (fixed
// enthalpy and pressure) or UV (fixed internal energy and volume).  Every loop runs over
// compile-time species and element counts and the element matrix is a compile-time table of the
// mechanism (zeros and ones fold away), so the per-species state (n_j, ln n_j, h_j/RT, mu_j/RT),
// the (ne+2) x (ne+3) Newton matrix and its row interchanges stay in registers.
// Arithmetic follows the reference's operation order with contraction off (scoped to each
// function); only exp/log differ from the host's in the last bit.
//
// Tightened mode (P.tol > 0, no counterpart in the reference): the relative-correction, |dln T|
// and element-residual tests of Lib_CEA_solver.f90:345-364 all at tol, so that the converged state
// does not depend on the start (the learned warm start and the default guess agree to round-off).
// Status per state: 0 converged, 1 more than P.max_iter iterations (the reference's maxitn = 50,
// Lib_CEA_solver.f90:68,318), 2 singular Newton matrix after the reference's retries (its
// component switch, Lib_CEA_solver.f90:725-848, is not taken), 3 temperature outside the thermo
// table.
#pragma once
#include "flint_device.h"
#include "flint_cea.h"


namespace flint_cea {

constexpr double RR = 8314.51;      // Rr, Lib_CEA_data.f90:65
constexpr double SIZE = 18.420681;  // Size, Lib_CEA_setup.f90:97

// cp/R (summed into Cpsum with the current n_j), h/RT and h/RT - s/R of every species at T.
template <int NS>
__device__ __forceinline__ bool thermo(const CeaProg& P, const double* __restrict__ W, double T, const double (&En)[NS],
                                       double (&H0)[NS], double (&G0)[NS], double& Cpsum) {
#pragma clang fp contract(off)
  if (!(T >= (double)P.Th_min) || !(T < (double)P.Th_max)) return false;
  const int Ti = (int)T;
  const double Td = T - (double)Ti;
  const double* __restrict__ r0 = P.tabh + (long)(Ti - P.Th_min) * (2 * NS);
  const double* __restrict__ r1 = r0 + 2 * NS;
  const double* __restrict__ s0 = P.tabs + (long)(Ti - P.Th_min) * NS;
  const double* __restrict__ s1 = s0 + NS;
  Cpsum = 0.0;
  // the table rows of eight species read before their arithmetic: read as it consumes them, each
  // load is waited for before the next one goes out
  constexpr int G = 8;
  double2 ga[G], gb[G];
  double gs0[G], gs1[G];
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    if (j % G == 0) {
#pragma unroll
      for (int u = 0; u < G; ++u) {
        const int jj = j + u < NS ? j + u : NS - 1;
        ga[u] = ld2(r0 + 2 * jj);
        gb[u] = ld2(r1 + 2 * jj);
        gs0[u] = s0[jj];
        gs1[u] = s1[jj];
      }
    }
    const double2 a = ga[j % G], b = gb[j % G];
    const double cp = a.x + (b.x - a.x) * Td;
    const double h = a.y + (b.y - a.y) * Td;
    const double s = gs0[j % G] + (gs1[j % G] - gs0[j % G]) * Td;
    const double Cp = cp * W[j] / RR;
    H0[j] = h * W[j] / (RR * T);
    G0[j] = H0[j] - s * W[j] / RR;
    Cpsum = Cpsum + En[j] * Cp;
  }
  return true;
}

// GAUSS: elimination with the row of smallest max|G(i,j>nn)|/|G(i,nn)| as pivot, row division,
// back substitution.  Returns 0 or the 1-based index of the singular step.
template <int IM>
__device__ __forceinline__ int gauss(double (&G)[IM][IM + 1], double (&X)[IM]) {
#pragma clang fp contract(off)
  const double bigno = (double)1.E+25f;
#pragma unroll
  for (int nn = 0; nn < IM; ++nn) {
    if (nn != IM - 1) {
      double coefx[IM];
#pragma unroll
      for (int i = nn; i < IM; ++i) {
        coefx[i] = bigno;
        if (G[i][nn] != 0.0) {
          double c = 0.0;
#pragma unroll
          for (int j = nn + 1; j < IM + 1; ++j) c = fmax(c, fabs(G[i][j]));
          const double tmp = fabs(G[i][nn]);
          coefx[i] = (bigno * tmp > c) ? c / tmp : bigno;
        }
      }
      double tmp = bigno;
      int ip = -1;
#pragma unroll
      for (int j = nn; j < IM; ++j)
        if (coefx[j] < tmp) {
          tmp = coefx[j];
          ip = j;
        }
      if (ip < 0) return nn + 1;
      // interchange rows ip and nn with selects only, so that every index stays static
#pragma unroll
      for (int j = nn; j < IM + 1; ++j) {
        const double old = G[nn][j];
        double piv = old;
#pragma unroll
        for (int r = nn + 1; r < IM; ++r) piv = (r == ip) ? G[r][j] : piv;
#pragma unroll
        for (int r = nn + 1; r < IM; ++r) G[r][j] = (r == ip) ? old : G[r][j];
        G[nn][j] = piv;
      }
    } else if (G[nn][nn] == 0.0) {
      return nn + 1;
    }
    const double tmp = G[nn][nn];
    if (tmp == 0.0) return nn + 1;
#pragma unroll
    for (int j = nn + 1; j < IM + 1; ++j) G[nn][j] = G[nn][j] / tmp;
#pragma unroll
    for (int i = nn + 1; i < IM; ++i)
#pragma unroll
      for (int j = nn + 1; j < IM + 1; ++j) G[i][j] = G[i][j] - G[i][nn] * G[nn][j];
  }
#pragma unroll
  for (int k = IM - 1; k >= 0; --k) {
    double tmp = 0.0;
#pragma unroll
    for (int i = k + 1; i < IM; ++i) tmp = tmp + G[k][i] * X[i];
    X[k] = G[k][IM] - tmp;
  }
  return 0;
}

template <class M, bool VOL>
__global__ __launch_bounds__(64) void cea_kernel(CeaProg P, long n, long ld, const double* __restrict__ pin,
                                                 const double* __restrict__ Tin, const double* __restrict__ rho,
                                                 const double* __restrict__ guess, double* __restrict__ Tout,
                                                 double* __restrict__ Yout, int* __restrict__ iters,
                                                 int* __restrict__ status) {
#pragma clang fp contract(off)
  constexpr int NS = M::NS, NE = M::NE;
  __shared__ double sW[NS], sRm[NS];
  for (int t = threadIdx.x; t < NS; t += blockDim.x) {
    sW[t] = P.W[t];
    sRm[t] = P.rm[t];
  }
  __syncthreads();
  long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (P.idx) {
    if (c >= (long)*P.count) return;
    c = P.idx[c];
  } else if (c >= n) {
    return;
  }
  if (P.redo && status[c] == 0) return;

  double En[NS], Enln[NS], H0[NS], G0[NS];
  double B0[NE];
  int rc = 0, numb = 0;
  double Tt = Tin[c];

  // CEA_initialize_local: Vv, b0 (normalised by the input mass), Hsub0 (not normalised)
  double Vv, Hsub0, Bcheck;
  {
    const double T0 = Tt;
    if (!(T0 >= (double)P.Th_min) || !(T0 < (double)P.Th_max)) {
      rc = 3;
    } else {
      const int Ti = (int)T0;
      const double Td = T0 - (double)Ti;
      const double* __restrict__ r0 = P.tabh + (long)(Ti - P.Th_min) * (2 * NS);
      const double* __restrict__ r1 = r0 + 2 * NS;
      double sumr = 0.0, Wp = 0.0, hpp = 0.0;
#pragma unroll
      for (int k = 0; k < NE; ++k) B0[k] = 0.0;
#pragma unroll
      for (int j = 0; j < NS; ++j) sumr += rho[(long)j * ld + c];
      Vv = 1e05 / sumr;
#pragma unroll
      for (int j = 0; j < NS; ++j) {
        const double pj = rho[(long)j * ld + c];
        const double h0j = r0[2 * j + 1], h1j = r1[2 * j + 1];  // read unguarded (in bounds)
        if (pj <= 1e-15) continue;
        Wp = Wp + pj;
        const double hi = h0j + (h1j - h0j) * Td;
        double enth = hi * sW[j] / RR;
        if (VOL) enth = enth - T0;
        hpp = hpp + enth * pj / sRm[j];
#pragma unroll
        for (int k = 0; k < NE; ++k) B0[k] = M::a(k, j) * pj / sRm[j] + B0[k];
      }
      Hsub0 = hpp;
      if (Wp != 0.0)
#pragma unroll
        for (int k = 0; k < NE; ++k) B0[k] = B0[k] / Wp;
      double bigb = 0.0;
#pragma unroll
      for (int k = 0; k < NE; ++k) bigb = fmax(bigb, fabs(B0[k]));
      Bcheck = P.tol > 0.0 ? bigb * P.tol : bigb * .000001;
    }
  }

  double Enn = .1, Ennl = (double)-2.3025851f, Sumn = Enn;
  if (rc == 0) {
    const double xi = Enn / (double)NS;
    const double xln = log(xi);
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      En[j] = xi;
      Enln[j] = xln;
    }
    if (guess) {
      Tt = guess[c];
      Sumn = 0.0;
#pragma unroll
      for (int j = 0; j < NS; ++j) {
        Enln[j] = guess[(long)(j + 1) * ld + c];
        En[j] = exp(Enln[j]);
        Sumn += En[j];
      }
      Enn = Sumn;
      Ennl = log(Enn);
    }
  }

  if (rc == 0) {
    const double pp_in = pin[c];
    const double siz9 = SIZE - 9.2103404;
    const double smalno = (double)1.E-6f, smnol = (double)-13.815511f;
    double tsize = SIZE;
    int ixsing = 0;
    double Tln = log(Tt);
    double Pp = VOL ? RR * Enn * Tt / Vv : pp_in;
    double Cpsum;
    if (!thermo<NS>(P, sW, Tt, En, H0, G0, Cpsum)) rc = 3;
    double Tm = log(Pp / Enn);
    constexpr int Iq1 = NE, iq2 = NE + 1, iq3 = NE + 2;
    constexpr int IM = VOL ? NE + 1 : NE + 2;
    while (rc == 0) {
      numb = numb + 1;
      // MATRIX; G0 becomes mu_j/RT for this iteration
      double Gm[NE + 3][NE + 3];
#pragma unroll
      for (int i = 0; i < NE + 3; ++i)
#pragma unroll
        for (int k = 0; k < NE + 3; ++k) Gm[i][k] = 0.0;
#pragma unroll
      for (int j = 0; j < NS; ++j) {
        G0[j] = G0[j] + Enln[j] + Tm;
        if (En[j] != 0.0) {
          const double h = H0[j] * En[j];
          const double f = G0[j] * En[j];
#pragma unroll
          for (int i = 0; i < NE; ++i) {
            const double aij = M::a(i, j);
            if (aij != 0.0) {
              const double term = aij * En[j];
#pragma unroll
              for (int k = i; k < NE; ++k) Gm[i][k] = Gm[i][k] + M::a(k, j) * term;
              Gm[i][Iq1] = Gm[i][Iq1] + term;
              Gm[i][iq2] = Gm[i][iq2] + aij * h;
              Gm[i][iq3] = Gm[i][iq3] + aij * f;
            }
          }
          Gm[iq2][iq2] = Gm[iq2][iq2] + H0[j] * h;
          Gm[iq2][iq3] = Gm[iq2][iq3] + H0[j] * f;
          Gm[Iq1][iq3] = Gm[Iq1][iq3] + f;
          Gm[Iq1][iq2] = Gm[Iq1][iq2] + h;
        }
      }
      const double Hsum = 0.0 + Gm[Iq1][iq2];
      Gm[Iq1][Iq1] = Sumn - Enn;
#pragma unroll
      for (int i = 0; i <= iq2; ++i)
#pragma unroll
        for (int j = i; j <= iq2; ++j) Gm[j][i] = Gm[i][j];
#pragma unroll
      for (int i = 0; i < NE; ++i) Gm[i][iq3] = Gm[i][iq3] + B0[i] - Gm[i][Iq1];
      Gm[Iq1][iq3] = Gm[Iq1][iq3] + Enn - Sumn;
      const double energyl = Hsub0 / Tt - Hsum;
      Gm[iq2][iq3] = Gm[iq2][iq3] + energyl;
      Gm[iq2][iq2] = Gm[iq2][iq2] + Cpsum;
      double G[IM][IM + 1];
      if (VOL) {
        // constant-volume matrix (Lib_CEA_solver.f90:964-979)
#pragma unroll
        for (int i = 0; i < NE; ++i) {
          Gm[Iq1][i] = Gm[iq2][i] - Gm[Iq1][i];
          Gm[i][Iq1] = Gm[i][iq2] - Gm[i][Iq1];
          Gm[i][iq2] = Gm[i][iq3];
        }
        Gm[Iq1][Iq1] = Gm[iq2][iq2] - Gm[Iq1][iq2] - Gm[iq2][Iq1];
        Gm[Iq1][iq2] = Gm[iq2][iq3] - Gm[Iq1][iq3];
        Gm[Iq1][iq2] = Gm[Iq1][iq2] + Enn;
      }
#pragma unroll
      for (int i = 0; i < IM; ++i)
#pragma unroll
        for (int k = 0; k < IM + 1; ++k) G[i][k] = Gm[i][k];
      double X[IM];
      const int msing = gauss<IM>(G, X);
      if (msing == 0) {
        const double dlnt = VOL ? X[Iq1] : X[IM - 1];
        const double xq1 = VOL ? 0.0 : X[Iq1];
        const double sum0 = VOL ? -dlnt : X[Iq1];
        // pass 1: the control factor lambda from the corrections
        double ambda = 1.0, ambda1 = 1.0;
        double sum = fmax(fabs(xq1), fabs(dlnt));
        sum = sum * 5.;
#pragma unroll
        for (int j = 0; j < NS; ++j) {
          double dl = -G0[j] + H0[j] * dlnt + sum0;
#pragma unroll
          for (int k = 0; k < NE; ++k) dl = dl + M::a(k, j) * X[k];
          if (dl > 0.0) {
            if ((Enln[j] - Ennl + SIZE) <= 0.0) {
              double sum1 = fabs(dl - xq1);
              if (sum1 >= siz9) {
                sum1 = fabs(-9.2103404 - Enln[j] + Ennl) / sum1;
                if (sum1 < ambda1) ambda1 = sum1;
              }
            } else if (dl > sum) {
              sum = dl;
            }
          }
        }
        if (sum > 2.0) ambda = 2.0 / sum;
        if (ambda1 <= ambda) ambda = ambda1;
        // pass 2: apply (the same corrections, recomputed), keep max |dln n_j| n_j for the test
        double Totn = 0.0, dmax = 0.0;
#pragma unroll
        for (int j = 0; j < NS; ++j) {
          double dl = -G0[j] + H0[j] * dlnt + sum0;
#pragma unroll
          for (int k = 0; k < NE; ++k) dl = dl + M::a(k, j) * X[k];
          Enln[j] = Enln[j] + ambda * dl;
          En[j] = 0.0;
          if ((Enln[j] - Ennl + tsize) > 0.0) {
            En[j] = exp(Enln[j]);
            Totn = Totn + En[j];
          }
          dmax = fmax(dmax, fabs(dl) * En[j]);
        }
        Sumn = Totn;
        Tln = Tln + ambda * dlnt;
        Tt = exp(Tln);
        if (!thermo<NS>(P, sW, Tt, En, H0, G0, Cpsum)) {
          rc = 3;
          break;
        }
        if (VOL) {
          Enn = Sumn;
          Ennl = log(Enn);
          Pp = RR * Tt * Enn / Vv;
        } else {
          Ennl = Ennl + ambda * xq1;
          Enn = exp(Ennl);
        }
        Tm = log(Pp / Enn);
        if (numb > P.max_iter) {
          rc = 1;
          break;
        }
        if (fabs(xq1 * Enn / Totn) > (P.tol > 0.0 ? P.tol : (double)0.5E-5f)) continue;
        if (dmax / Totn > (P.tol > 0.0 ? P.tol : 0.5e-5)) continue;
        if (fabs(dlnt) > (P.tol > 0.0 ? P.tol : 1.e-04)) continue;
        bool again = false;
#pragma unroll
        for (int i = 0; i < NE; ++i)
          if (fabs(B0[i]) >= 1.e-06) {
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < NS; ++j) s = s + En[j] * M::a(i, j);
            if (fabs(B0[i] - s) > Bcheck) again = true;
          }
        if (again) continue;
        break;  // converged
      }
      // singular matrix: retry with the zero species revived at 1e-6 (Lib_CEA_solver.f90:436-536)
      ixsing = ixsing + 1;
      tsize = 80.0;
      if (ixsing > 8 || (ixsing >= 3 && msing <= NE)) {
        rc = 2;
        break;
      }
#pragma unroll
      for (int j = 0; j < NS; ++j)
        if (En[j] == 0.0) {
          En[j] = smalno;
          Enln[j] = smnol;
        }
      if (!thermo<NS>(P, sW, Tt, En, H0, G0, Cpsum)) rc = 3;
    }
  }
  if (rc == 0) {
    Tout[c] = Tt;
#pragma unroll
    for (int j = 0; j < NS; ++j) Yout[(long)j * ld + c] = En[j] * sW[j];
  } else {
    Tout[c] = __builtin_nan("");
#pragma unroll
    for (int j = 0; j < NS; ++j) Yout[(long)j * ld + c] = __builtin_nan("");
  }
  if (iters) iters[c] = P.redo ? iters[c] + numb : numb;
  if (status) status[c] = rc;
}


// Launcher for mechanism M (instantiated by the generated mechanism translation units).
template <class M>
hipError_t launch(const CeaProg& P, bool hp, long n, long ld, const double* p, const double* T, const double* rho,
                  const double* guess, double* Tout, double* Yout, int* iters, int* status, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const unsigned blocks = (unsigned)((n + 63) / 64);
  if (hp)
    hipLaunchKernelGGL((cea_kernel<M, false>), dim3(blocks), dim3(64), 0, s, P, n, ld, p, T, rho, guess, Tout, Yout,
                       iters, status);
  else
    hipLaunchKernelGGL((cea_kernel<M, true>), dim3(blocks), dim3(64), 0, s, P, n, ld, p, T, rho, guess, Tout, Yout,
                       iters, status);
  return hipGetLastError();
}

}  // namespace flint_cea

