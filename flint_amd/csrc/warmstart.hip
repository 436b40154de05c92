// Learned equilibrium warm start: the network's input features and the Newton start built from its
// outputs, one state per thread (no counterpart in the FLINT reference).
#include "flint_warm.h"

namespace flint_ws_k {

__global__ __launch_bounds__(256) void features_kernel(long n, long ld, int ns, const double* __restrict__ Y0,
                                                       const double* __restrict__ T0, const double* __restrict__ p,
                                                       const double* __restrict__ tabh, int Th_min, int Th_max,
                                                       const double* __restrict__ scal, float* __restrict__ F) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const double* __restrict__ mean = scal;
  const double* __restrict__ scale = scal + (ns + 2);
  const double T = T0[c];
  // the table row of T, kept inside the table (a state outside it fails in the solver anyway)
  double Tc = fmin(fmax(T, (double)Th_min), (double)Th_max - 1.0);
  if (!(Tc == Tc)) Tc = (double)Th_min;
  const int Ti = (int)floor(Tc);
  const double d = T - (double)Ti;
  const double* __restrict__ r0 = tabh + (long)(Ti - Th_min) * (2 * ns);
  const double* __restrict__ r1 = r0 + 2 * ns;
  float* __restrict__ f = F + c * (ns + 2);
  double h0 = 0.0;
  for (int k = 0; k < ns; ++k) {
    const double y = Y0[(long)k * ld + c];
    const double a = r0[2 * k + 1], b = r1[2 * k + 1];
    h0 += (a + (b - a) * d) * y;
    f[k] = (float)((y - mean[k]) / scale[k]);
  }
  f[ns] = (float)((h0 - mean[ns]) / scale[ns]);
  f[ns + 1] = (float)((log(p[c]) - mean[ns + 1]) / scale[ns + 1]);
}

__global__ __launch_bounds__(256) void guess_kernel(long n, long ld, int ns, const float* __restrict__ O,
                                                    const double* __restrict__ scal, double* __restrict__ G) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const double* __restrict__ mean = scal + 2 * (ns + 2);
  const double* __restrict__ scale = mean + (ns + 1);
  for (int j = 0; j <= ns; ++j) G[(long)j * ld + c] = (double)O[c * (ns + 1) + j] * scale[j] + mean[j];
}

}  // namespace flint_ws_k

hipError_t flint_ws_features(long n, long ld, int ns, const double* Y0, const double* T0, const double* p,
                             const double* tabh, int Th_min, int Th_max, const double* scal, float* F, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(flint_ws_k::features_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, ld, ns, Y0, T0,
                     p, tabh, Th_min, Th_max, scal, F);
  return hipGetLastError();
}

hipError_t flint_ws_guess(long n, long ld, int ns, const float* O, const double* scal, double* G, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(flint_ws_k::guess_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, ld, ns, O, scal,
                     G);
  return hipGetLastError();
}
