// The data-driven general source term: FLINT's general routine (Lib_Chemistry_wdot.f90:96-221),
// which covers any mechanism from its tables alone, one cell per thread.
//
// Per reaction of each kind (Arrhenius and three-body, Troe, Lindemann), in FLINT's order: the
// third-body concentration coM = sum_k coi_k eps_k when the reaction has one (else 1), the
// products of the concentrations (integer powers, by binary powering, for Arrhenius; real powers
// for fall-off, where integer-valued exponents are applied the same way), the rates from the
// 1 K tables (f_kf / f_kb; f_k_troe / f_k_lindemann, Lib_Chemistry_falloff.f90:12-58) and
// omegadot_k += W_k (ni2 - ni1) net.  The host keeps only the nonzero table entries (adding the
// zero terms of FLINT's dense loops leaves every finite sum unchanged), in species order, and the
// kernel keeps the reference's operation order with contraction off.  The concentrations and the
// running sums sit in LDS, one column per thread, so the kernel runs for any species count.
#include "flint_device.h"
#include "flint_general.h"

namespace flint_general_k {

using flint_general::GenProg;
constexpr int BLOCK = 128;

// x**n for an integer-valued n
__device__ __forceinline__ double powi(double a, int b) {
#pragma clang fp contract(off)
  const bool recip = b < 0;
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

__device__ __forceinline__ double factor(double c, double e, bool integer) {
  if (integer || (e == rint(e) && fabs(e) <= 64.0)) return powi(c, (int)e);
  return pow(c, e);
}

__device__ __forceinline__ double interp(const double* __restrict__ r0, const double* __restrict__ r1, int col,
                                         double Td) {
#pragma clang fp contract(off)
  const double a = r0[col], b = r1[col];
  return a + (b - a) * Td;
}

__global__ __launch_bounds__(BLOCK) void general_kernel(GenProg P, const double* __restrict__ tabk, int Tk_min,
                                                        int Tk_max, const double* __restrict__ W, long n, long ld,
                                                        const double* __restrict__ rho, const double* __restrict__ T,
                                                        double* __restrict__ wdot, int* __restrict__ status) {
#pragma clang fp contract(off)
  extern __shared__ double smem[];
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int ns = P.ns, B = blockDim.x;
  double* __restrict__ co = smem + threadIdx.x;           // co[k * B]
  double* __restrict__ om = smem + (long)ns * B + threadIdx.x;  // om[k * B]
  for (int k = 0; k < ns; ++k) {
    co[k * B] = rho[(long)k * ld + i] / W[k];
    om[k * B] = 0.0;
  }
  const double temp = T[i];
  int st = 0;
  if (!(temp >= (double)Tk_min && temp <= (double)Tk_max)) st = 1;
  const int T_i = (int)temp;
  const double Td = temp - (double)T_i;
  const int row = min(max(T_i, Tk_min), Tk_max) - Tk_min;
  const double* __restrict__ r0 = tabk + (long)row * P.wk;
  const double* __restrict__ r1 = r0 + P.wk;
  for (int j = 0; j < P.nr; ++j) {
    const int kind = P.kind[j] & 3;
    double coM = 1.0;
    if (P.kind[j] & 4) {
      coM = 0.0;
      for (int t = P.m_ptr[j]; t < P.m_ptr[j + 1]; ++t) coM = coM + co[P.m_sp[t] * B] * P.m_eps[t];
    }
    double pf = 1.0, pb = 1.0;
    for (int t = P.f_ptr[j]; t < P.f_ptr[j + 1]; ++t) pf = pf * factor(co[P.f_sp[t] * B], P.f_exp[t], kind == 0);
    for (int t = P.r_ptr[j]; t < P.r_ptr[j + 1]; ++t) pb = pb * factor(co[P.r_sp[t] * B], P.r_exp[t], kind == 0);
    double rf, rb;
    if (kind == 0) {
      rf = interp(r0, r1, 2 * j, Td) * pf * coM;
      rb = interp(r0, r1, 2 * j + 1, Td) * pb * coM;
    } else if (kind == 1) {
      const int c = 2 * P.na + 4 * (j - P.na);
      const double kinf = interp(r0, r1, c, Td), k0 = interp(r0, r1, c + 1, Td);
      const double kc = interp(r0, r1, c + 2, Td), Fcent = interp(r0, r1, c + 3, Td);
      const double Pr = k0 * coM / kinf;
      const double lFc = log10(Fcent), lPr = log10(Pr);
      const double cc = -0.4 - 0.67 * lFc, nn = 0.75 - 1.27 * lFc;
      const double f1 = (lPr + cc) / (nn - 0.14 * (lPr + cc));
      const double k1 = kinf * (Pr / (1.0 + Pr)) * pow(10.0, lFc / (1.0 + f1 * f1));
      rf = k1 * pf;
      rb = k1 / kc * pb;
    } else {
      const int c = 2 * P.na + 4 * P.nt + 3 * (j - P.na - P.nt);
      const double kinf = interp(r0, r1, c, Td), k0 = interp(r0, r1, c + 1, Td), kc = interp(r0, r1, c + 2, Td);
      const double Pr = k0 * coM / kinf;
      const double k1 = kinf * (Pr / (1.0 + Pr));
      rf = k1 * pf;
      rb = k1 / kc * pb;
    }
    const double net = rf - rb;
    for (int t = P.o_ptr[j]; t < P.o_ptr[j + 1]; ++t) {
      double* o = om + P.o_sp[t] * B;
      *o = *o + P.o_wd[t] * net;
    }
  }
  for (int k = 0; k < ns; ++k) wdot[(long)k * ld + i] = om[k * B];
  if (status) status[i] = st;
}

}  // namespace flint_general_k

hipError_t flint_general_launch(const flint_general::GenProg& P, const double* tabk, int Tk_min, int Tk_max,
                                const double* W, long n, long ld, const double* rho, const double* T, double* wdot,
                                int* status, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int B = flint_general_k::BLOCK;
  const size_t lds = (size_t)2 * P.ns * B * sizeof(double);
  if (lds > 65536) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&flint_general_k::general_kernel),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(flint_general_k::general_kernel, dim3((unsigned)((n + B - 1) / B)), dim3(B), lds, s, P, tabk,
                     Tk_min, Tk_max, W, n, ld, rho, T, wdot, status);
  return hipGetLastError();
}
