// Learned constant-pressure reactor step with the plain BDF as fallback (no counterpart in the
// FLINT reference).
//
// Per cell (one thread): features (log max(Y_k, ymin), T, log p) standardised, and for a network
// of 2 ns + 3 inputs also the rate features asinh(dt dY_k/dt / yfloor), asinh(dt dT/dt / tfloor)
// with the right-hand side at the old state -> MLP (mlp.hip)
// -> o de-standardised, the increments in units of the local rate: dY_k = o_k (dt |dY_k/dt| + yfloor),
// dT = o_T (dt |dT/dt| + tfloor), with the right-hand side at the old state, so that a cell whose
// chemistry is nearly frozen gets a nearly zero increment whatever the network's relative error.
// The raw prediction Y'_k = max(Y_k + dY_k, 0) is rejected, and the cell handed to the BDF kernel,
// when it moves an element mass fraction
// Z_e = W_e sum_k a_ek Y_k / W_k or the total mass by more than max_residual, or predicts
// |dT| > max_dT.  An accepted prediction is projected onto the conserved quantities exactly:
// Y = Y' + D C^T (C D C^T)^-1 (C Y0 - C Y') with C the element rows a_ek / W_k and a row of ones,
// D = diag(Y'), so element amounts and total mass are those of the old state to round-off; T then
// solves sum_k Y_k h_k(T) = sum_k Y0_k h_k(T0) (the enthalpy the adiabatic constant-pressure
// reactor conserves) by Newton on the tabulated h_k, and the cell still falls back when that T
// differs from the network's own T0 + dT by more than max_Tmis (the composition and temperature
// outputs disagree about the energy released).
#include "flint_device.h"
#include "flint_surrogate.h"

namespace flint_sur {

constexpr int MAXE = 4;

template <int NS, bool RATES>
__global__ __launch_bounds__(256) void features_kernel(SurProg P, long n, long ld, const double* __restrict__ X,
                                                       const double* __restrict__ pres, float* __restrict__ F) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  constexpr int NIN = RATES ? 2 * NS + 3 : NS + 2;
  const double* __restrict__ mu = P.scal;
  const double* __restrict__ sd = P.scal + NIN;
  float* __restrict__ f = F + c * NIN;
#pragma unroll
  for (int k = 0; k < NS; ++k) f[k] = (float)((log(fmax(X[(long)k * ld + c], P.ymin)) - mu[k]) / sd[k]);
  f[NS] = (float)((X[(long)NS * ld + c] - mu[NS]) / sd[NS]);
  f[NS + 1] = (float)((log(pres[c]) - mu[NS + 1]) / sd[NS + 1]);
  if (RATES) {
#pragma unroll
    for (int k = 0; k <= NS; ++k) {
      const double fl = k < NS ? P.yfloor : P.tfloor;
      const double r = asinh(P.dt * P.F[(long)k * P.ldF + c] / fl);
      f[NS + 2 + k] = (float)((r - mu[NS + 2 + k]) / sd[NS + 2 + k]);
    }
  }
}

// mixture enthalpy sum_k Y_k h_k(T) and its slope within the table bin
template <int NS>
__device__ __forceinline__ bool hmix(const SurProg& P, const double (&Y)[NS], double T, double& h, double& dh) {
  if (!(T >= (double)P.Th_min) || !(T < (double)P.Th_max)) return false;
  const int Ti = (int)T;
  const double Td = T - (double)Ti;
  const double* __restrict__ r0 = P.tabh + (long)(Ti - P.Th_min) * (2 * NS);
  const double* __restrict__ r1 = r0 + 2 * NS;
  h = 0.0;
  dh = 0.0;
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    const double a = r0[2 * k + 1], b = r1[2 * k + 1];
    h += Y[k] * (a + (b - a) * Td);
    dh += Y[k] * (b - a);
  }
  return true;
}

// The network's increments applied to X (read only): the prediction, projected onto the conserved
// element and mass totals, with T from the conserved enthalpy, goes to Xp; flag[c] = 1 marks a
// cell that fails the plausibility tests (|dT|, raw element residual, T from enthalpy).
template <int NS>
__global__ __launch_bounds__(128) void apply_kernel(SurProg P, long n, long ld, const double* __restrict__ X,
                                                    double* __restrict__ Xp, const float* __restrict__ O,
                                                    int* __restrict__ flag) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const int ne = P.ne;
  const double* __restrict__ om = P.scal + 2 * P.nin;
  const double* __restrict__ os = om + (NS + 1);
  double Y0[NS], Y[NS];
  const double T0 = X[(long)NS * ld + c];
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    Y0[k] = X[(long)k * ld + c];
    const double o = (double)O[c * (NS + 1) + k] * os[k] + om[k];
    Y[k] = fmax(Y0[k] + o * (P.dt * fabs(P.F[(long)k * P.ldF + c]) + P.yfloor), 0.0);
  }
  const double dT = ((double)O[c * (NS + 1) + NS] * os[NS] + om[NS]) *
                    (P.dt * fabs(P.F[(long)NS * P.ldF + c]) + P.tfloor);
  // residuals of the raw prediction: element mass fractions and total mass
  double res = 0.0, s0 = 0.0, s1 = 0.0;
  double r[MAXE + 1];
#pragma unroll
  for (int e = 0; e < MAXE; ++e) {
    double z0 = 0.0, z1 = 0.0;
    if (e < ne) {
#pragma unroll
      for (int k = 0; k < NS; ++k) {
        const double a = P.comp[e * NS + k] / P.W[k];
        z0 += a * Y0[k];
        z1 += a * Y[k];
      }
      res = fmax(res, fabs(z1 - z0) * P.elw[e]);
    }
    r[e] = z0 - z1;
  }
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    s0 += Y0[k];
    s1 += Y[k];
  }
  res = fmax(res, fabs(s1 - s0));
  r[ne < MAXE ? ne : MAXE] = s0 - s1;
  bool fall = !(fabs(dT) <= P.max_dT) || !(res <= P.max_res);
  double T = T0 + dT;
  // weighted least-norm correction onto C Y = C Y0 (ne element rows and the mass row), applied
  // three times: each pass removes what the previous one left of the residual (the normal matrix
  // is ill-conditioned when an element is nearly absent, and two passes left up to 1e-12 of total
  // mass)
#pragma unroll 1
  for (int pass = 0; pass < 3 && !fall; ++pass) {
    if (pass > 0) {
#pragma unroll
      for (int e = 0; e < MAXE; ++e) {
        double z0 = 0.0, z1 = 0.0;
        if (e < ne) {
#pragma unroll
          for (int k = 0; k < NS; ++k) {
            const double a = P.comp[e * NS + k] / P.W[k];
            z0 += a * Y0[k];
            z1 += a * Y[k];
          }
        }
        r[e] = z0 - z1;
      }
      s1 = 0.0;
#pragma unroll
      for (int k = 0; k < NS; ++k) s1 += Y[k];
      r[ne < MAXE ? ne : MAXE] = s0 - s1;
    }
    const int nc = ne + 1;
    double M[MAXE + 1][MAXE + 2];
#pragma unroll
    for (int i = 0; i <= MAXE; ++i)
#pragma unroll
      for (int j = 0; j <= MAXE + 1; ++j) M[i][j] = 0.0;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      const double dk = fmax(Y[k], 0.0);
      double ck[MAXE + 1];
#pragma unroll
      for (int i = 0; i <= MAXE; ++i) ck[i] = i < ne ? P.comp[i * NS + k] / P.W[k] : (i == ne ? 1.0 : 0.0);
#pragma unroll
      for (int i = 0; i <= MAXE; ++i)
#pragma unroll
        for (int j = 0; j <= MAXE; ++j) M[i][j] += ck[i] * dk * ck[j];
    }
#pragma unroll
    for (int i = 0; i <= MAXE; ++i) {
      M[i][MAXE + 1] = i < nc ? r[i] : 0.0;
      if (i >= nc) M[i][i] = 1.0;
    }
    // Gaussian elimination with partial pivoting on the (ne+1) system (padded to MAXE+1)
#pragma unroll
    for (int k = 0; k <= MAXE; ++k) {
      int p = k;
      double best = fabs(M[k][k]);
#pragma unroll
      for (int i = k + 1; i <= MAXE; ++i)
        if (fabs(M[i][k]) > best) {
          best = fabs(M[i][k]);
          p = i;
        }
#pragma unroll
      for (int j = 0; j <= MAXE + 1; ++j) {
        double pv = M[k][j];
#pragma unroll
        for (int i = k + 1; i <= MAXE; ++i) pv = (i == p) ? M[i][j] : pv;
#pragma unroll
        for (int i = k + 1; i <= MAXE; ++i) M[i][j] = (i == p) ? M[k][j] : M[i][j];
        M[k][j] = pv;
      }
      if (M[k][k] == 0.0) {
        fall = true;
        break;
      }
#pragma unroll
      for (int i = k + 1; i <= MAXE; ++i) {
        const double l = M[i][k] / M[k][k];
#pragma unroll
        for (int j = k; j <= MAXE + 1; ++j) M[i][j] -= l * M[k][j];
      }
    }
    double lam[MAXE + 1];
#pragma unroll
    for (int i = MAXE; i >= 0; --i) {
      double s = M[i][MAXE + 1];
#pragma unroll
      for (int j = i + 1; j <= MAXE; ++j) s -= M[i][j] * lam[j];
      lam[i] = s / M[i][i];
    }
    if (!fall) {
#pragma unroll
      for (int k = 0; k < NS; ++k) {
        double ct = 0.0;
#pragma unroll
        for (int i = 0; i <= MAXE; ++i)
          if (i < nc) ct += (i < ne ? P.comp[i * NS + k] / P.W[k] : 1.0) * lam[i];
        Y[k] += fmax(Y[k], 0.0) * ct;
      }
    }
  }
  if (!fall) {
    {
      // temperature from the conserved mixture enthalpy
      double h0, dh0;
      if (!hmix<NS>(P, Y0, T0, h0, dh0)) fall = true;
      bool conv = false;
#pragma unroll 1
      for (int it = 0; it < 30 && !fall; ++it) {
        double h, dh;
        if (!hmix<NS>(P, Y, T, h, dh) || !(dh > 0.0)) {
          fall = true;
          break;
        }
        const double step = (h - h0) / dh;
        T -= step;
        if (fabs(step) <= 1e-10 * T) {
          conv = true;
          break;
        }
      }
      if (!conv || !(fabs(T - (T0 + dT)) <= P.max_Tmis)) fall = true;
    }
  }
  flag[c] = fall ? 1 : 0;
  if (fall) return;
#pragma unroll
  for (int k = 0; k < NS; ++k) Xp[(long)k * ld + c] = Y[k];
  Xp[(long)NS * ld + c] = T;
}

__global__ void zero_int(int* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// gather (dir = 0) the listed cells into a compact SoA block, or scatter (dir = 1) them back
__global__ void gather_kernel(int dir, int nrow, long ld, double* __restrict__ X, long nf, long ldg,
                              double* __restrict__ Xg, const int* __restrict__ idx, const double* __restrict__ p,
                              double* __restrict__ pg) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nf) return;
  const long c = idx[i];
  for (int r = 0; r < nrow; ++r) {
    if (dir == 0)
      Xg[(long)r * ldg + i] = X[(long)r * ld + c];
    else
      X[(long)r * ld + c] = Xg[(long)r * ldg + i];
  }
  if (dir == 0 && p) pg[i] = p[c];
}

__global__ void scatter_int(int nrow, long ld, int* __restrict__ dst, long nf, long ldg, const int* __restrict__ src,
                            const int* __restrict__ idx) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nf) return;
  const long c = idx[i];
  for (int r = 0; r < nrow; ++r) dst[(long)r * ld + c] = src[(long)r * ldg + i];
}

}  // namespace flint_sur

hipError_t flint_sur_features(const flint_sur::SurProg& P, long n, long ld, const double* X, const double* p, float* F,
                              hipStream_t s) {
  const unsigned b = (unsigned)((n + 255) / 256);
  const bool rates = P.nin == 2 * P.ns + 3;
  if (!rates && P.nin != P.ns + 2) return hipErrorInvalidValue;
  if (P.ns == 31 && rates)
    hipLaunchKernelGGL((flint_sur::features_kernel<31, true>), dim3(b), dim3(256), 0, s, P, n, ld, X, p, F);
  else if (P.ns == 31)
    hipLaunchKernelGGL((flint_sur::features_kernel<31, false>), dim3(b), dim3(256), 0, s, P, n, ld, X, p, F);
  else if (P.ns == 24 && rates)
    hipLaunchKernelGGL((flint_sur::features_kernel<24, true>), dim3(b), dim3(256), 0, s, P, n, ld, X, p, F);
  else if (P.ns == 24)
    hipLaunchKernelGGL((flint_sur::features_kernel<24, false>), dim3(b), dim3(256), 0, s, P, n, ld, X, p, F);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t flint_sur_apply(const flint_sur::SurProg& P, long n, long ld, const double* X, double* Xp, const float* O,
                           int* flag, hipStream_t s) {
  const unsigned b = (unsigned)((n + 127) / 128);
  if (P.ns == 31)
    hipLaunchKernelGGL(flint_sur::apply_kernel<31>, dim3(b), dim3(128), 0, s, P, n, ld, X, Xp, O, flag);
  else if (P.ns == 24)
    hipLaunchKernelGGL(flint_sur::apply_kernel<24>, dim3(b), dim3(128), 0, s, P, n, ld, X, Xp, O, flag);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t flint_sur_zero(int* p, int n, hipStream_t s) {
  hipLaunchKernelGGL(flint_sur::zero_int, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, p, n);
  return hipGetLastError();
}

hipError_t flint_sur_gather(int dir, int nrow, long ld, double* X, long nf, long ldg, double* Xg, const int* idx,
                            const double* p, double* pg, hipStream_t s) {
  if (nf <= 0) return hipSuccess;
  hipLaunchKernelGGL(flint_sur::gather_kernel, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, s, dir, nrow, ld, X, nf,
                     ldg, Xg, idx, p, pg);
  return hipGetLastError();
}

hipError_t flint_sur_scatter_int(int nrow, long ld, int* dst, long nf, long ldg, const int* src, const int* idx,
                                 hipStream_t s) {
  if (nf <= 0) return hipSuccess;
  hipLaunchKernelGGL(flint_sur::scatter_int, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, s, nrow, ld, dst, nf, ldg,
                     src, idx);
  return hipGetLastError();
}
