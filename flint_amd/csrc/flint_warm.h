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

// The states whose status is nonzero: idx[0 .. *count) (in no particular order; count is cleared
// by a kernel first, so the call can be captured in a graph).
hipError_t flint_ws_collect(long n, const int* status, int* idx, int* count, hipStream_t s);

// The network shapes flint_ws_fused takes: 3 layers, d0 <= 32 inputs, two hidden layers of 128, and
// dL <= 32 outputs.
inline bool flint_ws_fused_fits(const int* dims) {
  return dims[0] >= 1 && dims[0] <= 32 && dims[1] == 128 && dims[2] == 128 && dims[3] >= 1 && dims[3] <= 32;
}

// The network's weights rearranged for flint_ws_fused (flint_ws_packed_bytes of device memory):
// dims[0..3] the layer widths, WT[l] the BF16 weights of layer l transposed ([dout][din]).
size_t flint_ws_packed_bytes();
hipError_t flint_ws_pack(const int* dims, const __bf16* const* WT, void* packed, hipStream_t s);

// flint_ws_features, the network's forward pass and flint_ws_guess in one kernel (G as in
// flint_ws_guess): packed from flint_ws_pack, b[l] the FP32 bias of layer l; at most `blocks`
// blocks of 8 wavefronts.
hipError_t flint_ws_fused(long n, long ld, int ns, const double* Y0, const double* T0, const double* p,
                          const double* tabh, int Th_min, int Th_max, const double* scal, const int* dims,
                          const void* packed, const float* const* b, double* G, int blocks, hipStream_t s);
