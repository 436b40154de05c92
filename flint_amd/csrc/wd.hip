// Hand-written Westbrook-Dryer global mechanism (5 species, 3 irreversible reactions) with the
// heat release rate, one thread per cell, SoA state.
//
// Follows Lib_ChemMech/WD.f90:6-47 operation by operation: in-place clip rho = max(rho, 0)
// (:18-21), species index semantics of the routine's own comment [CH4, O2, CO2, H2O, CO]
// (:28) applied to phase.txt-ordered arrays, default-REAL exponents 0.70 and 0.80 (float32
// values promoted to double), exact 0.5 / 0.25 / 1.5 / 2.
// Heat release = -sum_k h_k(T) wdot_k with h from the thermo table (Lib_ThermoTransport.f90:77-96).
#include "flint_device.h"

namespace flint_wd {
constexpr int NS = 5;

__global__ __launch_bounds__(256) void wdot_hrr_kernel(long n, long ld, double* __restrict__ rho,
                                                       const double* __restrict__ T, double* __restrict__ wdot,
                                                       double* __restrict__ hrr, int* __restrict__ status,
                                                       const double* __restrict__ W, const double* __restrict__ tabk,
                                                       int Tk_min, int Tk_max, int wk, const double* __restrict__ tabh,
                                                       int Th_min, int Th_max, int write_back) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double temp = T[i];
  double roi[NS], coi[NS];
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    const double v = rho[k * ld + i];
    roi[k] = fmax(v, 0.0);
    if (write_back && v < 0.0) rho[k * ld + i] = 0.0;  // intent(inout) clip of WD.f90:19
    coi[k] = roi[k] / W[k];
  }
  int st = 0;
  if (!(temp >= (double)Tk_min && temp <= (double)Tk_max)) st |= 1;
  const int T_i = (int)temp;
  const double Td = temp - (double)T_i;
  const double* __restrict__ r0 = tabk + (long)(min(max(T_i, Tk_min), Tk_max) - Tk_min) * wk;
  const double* __restrict__ r1 = r0 + wk;
  const double kf1 = r0[0] + (r1[0] - r0[0]) * Td;
  const double kf2 = r0[2] + (r1[2] - r0[2]) * Td;
  const double kf3 = r0[4] + (r1[4] - r0[4]) * Td;
  const double e07 = (double)0.70f, e08 = (double)0.80f;
  const double prod1 = kf1 * pow(coi[0], e07) * pow(coi[1], e08);
  const double prod2 = kf2 * pow(coi[3], 0.5) * coi[4] * pow(coi[1], 0.25);
  const double prod3 = kf3 * coi[2];
  double w[NS];
  w[0] = W[0] * (-prod1);
  w[1] = W[1] * (-1.5 * prod1 - 0.5 * prod2 + 0.5 * prod3);
  w[2] = W[2] * (prod2 - prod3);
  w[3] = W[3] * (2.0 * prod1);
  w[4] = W[4] * (prod1 - prod2 + prod3);
#pragma unroll
  for (int k = 0; k < NS; ++k) wdot[k * ld + i] = w[k];
  if (hrr) {
    if (!(temp >= (double)Th_min && temp <= (double)Th_max)) st |= 2;
    const double* __restrict__ h0 = tabh + (long)(min(max(T_i, Th_min), Th_max) - Th_min) * (2 * NS);
    const double* __restrict__ h1 = h0 + 2 * NS;
    double q = 0.0;
#pragma unroll
    for (int k = 0; k < NS; ++k) q = q + (h0[2 * k + 1] + (h1[2 * k + 1] - h0[2 * k + 1]) * Td) * w[k];
    hrr[i] = -q;
  }
  if (status) status[i] = st;
}
}  // namespace flint_wd

hipError_t flint_wd_launch(long n, long ld, double* rho, const double* T, double* wdot, double* hrr, int* status,
                           const double* W, const double* tabk, int Tk_min, int Tk_max, int wk, const double* tabh,
                           int Th_min, int Th_max, int write_back, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(flint_wd::wdot_hrr_kernel, dim3(blocks), dim3(256), 0, s, n, ld, rho, T, wdot, hrr, status, W,
                     tabk, Tk_min, Tk_max, wk, tabh, Th_min, Th_max, write_back);
  return hipGetLastError();
}
