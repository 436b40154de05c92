// Kernels of the learned equilibrium warm start (warmstart.hip), used by flint_cea_solve_warm.
#pragma once
#include <hip/hip_runtime.h>

// Standardised network inputs F [n][ns+2] (float) = ((Y0_k, h0, ln p) - mean) / scale, with
// h0 = sum_k Y0_k h_k(T0) from the thermo table tabh [T][ns][2] (cp, h); scal = in_mean[ns+2],
// in_scale[ns+2], out_mean[ns+1], out_scale[ns+1].
hipError_t flint_ws_features(long n, long ld, int ns, const double* Y0, const double* T0, const double* p,
                             const double* tabh, int Th_min, int Th_max, const double* scal, float* F, hipStream_t s);

// Newton start G [ns+1][ld] = (T, ln n_1..ln n_ns) from the network outputs O [n][ns+1] (float).
hipError_t flint_ws_guess(long n, long ld, int ns, const float* O, const double* scal, double* G, hipStream_t s);
