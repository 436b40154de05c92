// Learned equilibrium warm start: the network's input features and the Newton start built from its
// outputs, one state per thread (no counterpart in the FLINT reference).
#include "flint_warm.h"

#include <algorithm>

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

// ---- the whole warm start of a 3-layer network d0 -> 32 H1T -> 32 H2T -> dL in one kernel ----
//
// Each wavefront takes 32 states at a time and computes the transposed products
// H^T = W^T X^T on v_mfma_f32_32x32x16_bf16, so that an accumulator tile (lane l: state l % 32,
// features 8 (r / 4) + 4 (l / 32) + r % 4 of the tile for r = 0..15) is, after bias, GELU and
// the BF16 rounding, already the B operand of the next layer: the 16 features of K-step
// (tile, q) are those of r = 8q..8q+7, and the weights are staged in LDS with their input index
// permuted to match, one 16-byte fragment per (output tile, K-step, lane).  The first layer's
// K-steps take the features in their natural order (k = 16 s + 8 (l / 32) + e), as the input of
// the layered forward pass does.  Features, biases, GELU (erf form), the BF16 roundings and the
// guess follow features_kernel, the forward pass and guess_kernel value for value; only the
// order of the per-K-step dot products differs (the hardware sums the same 16 products).
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }

// input index of element e of lane half h in K-step ks of a chained layer
__device__ __forceinline__ int chained_k(int ks, int h, int e) {
  return 32 * (ks >> 1) + 16 * (ks & 1) + 8 * (e >> 2) + 4 * h + (e & 3);
}

// Fragments of the three layers' weights in the order the fused kernel reads them: per layer,
// (output tile, K-step, lane) -> 8 BF16 of row 32 tile + lane % 32 of WT at the input indices of
// that lane's half (natural for the first layer, chained_k for the others), zero outside.
template <int H1T, int H2T>
__global__ __launch_bounds__(256) void pack_kernel(int d0, int dL, const __bf16* __restrict__ WT1,
                                                   const __bf16* __restrict__ WT2, const __bf16* __restrict__ WT3,
                                                   bf16x8* __restrict__ out) {
  constexpr int H1 = 32 * H1T, H2 = 32 * H2T, KS1 = 2, KS2 = 2 * H1T, KS3 = 2 * H2T;
  constexpr int N1 = H1T * KS1 * 64, N2 = H2T * KS2 * 64, N3 = KS3 * 64;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= N1 + N2 + N3) return;
  const int ln = q & 63, h = ln >> 5;
  bf16x8 v;
  if (q < N1) {
    const int ks = (q >> 6) % KS1, mt = q / (64 * KS1), i = 32 * mt + (ln & 31);
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int k = 16 * ks + 8 * h + e;
      v[e] = k < d0 ? WT1[(long)i * d0 + k] : (__bf16)0.0f;
    }
  } else if (q < N1 + N2) {
    const int r = q - N1, ks = (r >> 6) % KS2, mt = r / (64 * KS2), i = 32 * mt + (ln & 31);
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = WT2[(long)i * H1 + chained_k(ks, h, e)];
  } else {
    const int ks = (q - N1 - N2) >> 6, i = ln & 31;
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = i < dL ? WT3[(long)i * H2 + chained_k(ks, h, e)] : (__bf16)0.0f;
  }
  out[q] = v;
}

template <int H1T, int H2T>
__global__ __launch_bounds__(512) void fused_kernel(long n, long ld, int ns, const double* __restrict__ Y0,
                                                    const double* __restrict__ T0, const double* __restrict__ p,
                                                    const double* __restrict__ tabh, int Th_min, int Th_max,
                                                    const double* __restrict__ scal, int d0, int dL,
                                                    const bf16x8* __restrict__ wpk, const float* __restrict__ b1,
                                                    const float* __restrict__ b2, const float* __restrict__ b3,
                                                    double* __restrict__ G) {
  constexpr int H1 = 32 * H1T, H2 = 32 * H2T, KS1 = 2, KS2 = 2 * H1T, KS3 = 2 * H2T;
  constexpr int N1 = H1T * KS1 * 64, N2 = H2T * KS2 * 64, N3 = KS3 * 64;
  __shared__ bf16x8 wl[N1 + N2 + N3];
  __shared__ float sb1[H1], sb2[H2], sb3[32];
  // the output scaling in LDS as well: global loads between the stores of G would each wait for the
  // stores before them (one memory counter for both on this GPU)
  __shared__ double som[32], sos[32];
  const bf16x8* w1 = wl;
  const bf16x8* w2 = wl + N1;
  const bf16x8* w3 = wl + N1 + N2;
  for (int q = threadIdx.x; q < N1 + N2 + N3; q += blockDim.x) wl[q] = wpk[q];
  for (int q = threadIdx.x; q < H1; q += blockDim.x) sb1[q] = b1[q];
  for (int q = threadIdx.x; q < H2; q += blockDim.x) sb2[q] = b2[q];
  for (int q = threadIdx.x; q < 32; q += blockDim.x) sb3[q] = q < dL ? b3[q] : 0.0f;
  for (int q = threadIdx.x; q < 32; q += blockDim.x) {
    som[q] = q < dL ? scal[2 * (ns + 2) + q] : 0.0;
    sos[q] = q < dL ? scal[2 * (ns + 2) + (ns + 1) + q] : 0.0;
  }
  __syncthreads();

  const int lane = threadIdx.x & 63, h = lane >> 5;
  const int wpb = blockDim.x >> 6;
  const double* __restrict__ imean = scal;
  const double* __restrict__ iscale = scal + (ns + 2);
  for (long tile = (long)blockIdx.x * wpb + (threadIdx.x >> 6); tile * 32 < n; tile += (long)gridDim.x * wpb) {
    const long c = tile * 32 + (lane & 31);
    const bool ok = c < n;
    // the LDS and scaling reads below are the same in every tile; an index the optimiser cannot see
    // through keeps it from hoisting them out of the loop (hundreds of registers)
    int z = 0;
    asm volatile("" : "+v"(z));
    const int zl = lane + z;
    // features (features_kernel)
    bf16x8 xb[KS1];
    {
      double h0 = 0.0, lp = 0.0;
      if (ok) {
        const double T = T0[c];
        double Tc = fmin(fmax(T, (double)Th_min), (double)Th_max - 1.0);
        if (!(Tc == Tc)) Tc = (double)Th_min;
        const int Ti = (int)floor(Tc);
        const double d = T - (double)Ti;
        const double* __restrict__ r0 = tabh + (long)(Ti - Th_min) * (2 * ns);
        const double* __restrict__ r1 = r0 + 2 * ns;
        for (int k = 0; k < ns; ++k) {
          const double a = r0[2 * k + 1], b = r1[2 * k + 1];
          h0 += (a + (b - a) * d) * Y0[(long)k * ld + c];
        }
        lp = log(p[c]);
      }
#pragma unroll
      for (int s = 0; s < KS1; ++s)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int k = 16 * s + 8 * h + e;
          float f = 0.0f;
          if (ok && k < d0) {
            if (k < ns)
              f = (float)((Y0[(long)(k + z) * ld + c] - imean[k + z]) / iscale[k + z]);
            else if (k == ns)
              f = (float)((h0 - imean[ns + z]) / iscale[ns + z]);
            else
              f = (float)((lp - imean[ns + 1 + z]) / iscale[ns + 1 + z]);
          }
          xb[s][e] = (__bf16)f;
        }
    }
    // layer 1, each output tile finished (bias, GELU, BF16) as soon as it is summed
    bf16x8 hb1[KS2];
#pragma unroll
    for (int mt = 0; mt < H1T; ++mt) {
      f32x16 acc = f32x16{};
#pragma unroll
      for (int s = 0; s < KS1; ++s)
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w1[(mt * KS1 + s) * 64 + zl], xb[s], acc, 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int i = 32 * mt + 8 * (r >> 2) + 4 * h + (r & 3);
        hb1[2 * mt + (r >> 3)][r & 7] = (__bf16)gelu(acc[r] + sb1[i + z]);
      }
    }
    // layer 2
    bf16x8 hb2[KS3];
#pragma unroll
    for (int mt = 0; mt < H2T; ++mt) {
      f32x16 acc = f32x16{};
#pragma unroll
      for (int ks = 0; ks < KS2; ++ks)
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w2[(mt * KS2 + ks) * 64 + zl], hb1[ks], acc, 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int i = 32 * mt + 8 * (r >> 2) + 4 * h + (r & 3);
        hb2[2 * mt + (r >> 3)][r & 7] = (__bf16)gelu(acc[r] + sb2[i + z]);
      }
    }
    // output layer and the Newton start (guess_kernel)
    f32x16 o = f32x16{};
#pragma unroll
    for (int ks = 0; ks < KS3; ++ks) o = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w3[ks * 64 + zl], hb2[ks], o, 0, 0, 0);
    if (ok) {
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int j = 8 * (r >> 2) + 4 * h + (r & 3);
        if (j < dL) G[(long)j * ld + c] = (double)(o[r] + sb3[j + z]) * sos[j + z] + som[j + z];
      }
    }
  }
}

__global__ void zero_kernel(int* count) {
  if (threadIdx.x == 0 && blockIdx.x == 0) count[0] = 0;
}

// one atomic per block of 1024 states: every state of a batch may be flagged, and atomics on
// one address serialise (one per wavefront still took 48 us for 2^18 states)
constexpr int COLLECT_BLOCK = 1024;
__global__ __launch_bounds__(COLLECT_BLOCK) void collect_kernel(long n, const int* __restrict__ status,
                                                                int* __restrict__ idx, int* __restrict__ count) {
  __shared__ int wbase[COLLECT_BLOCK / 64];
  __shared__ int bbase;
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const bool f = c < n && status[c] != 0;
  const unsigned long long mask = __ballot(f);
  if (lane == 0) wbase[w] = __popcll(mask);
  __syncthreads();
  if (threadIdx.x == 0) {
    int tot = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
      const int t = wbase[i];
      wbase[i] = tot;
      tot += t;
    }
    bbase = tot ? atomicAdd(count, tot) : 0;
  }
  __syncthreads();
  if (f) idx[bbase + wbase[w] + __popcll(mask & ((1ull << lane) - 1))] = (int)c;
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

hipError_t flint_ws_collect(long n, const int* status, int* idx, int* count, hipStream_t s) {
  hipLaunchKernelGGL(flint_ws_k::zero_kernel, dim3(1), dim3(64), 0, s, count);
  if (n > 0)
    hipLaunchKernelGGL(flint_ws_k::collect_kernel, dim3((unsigned)((n + flint_ws_k::COLLECT_BLOCK - 1) /
                                                               flint_ws_k::COLLECT_BLOCK)),
                       dim3(flint_ws_k::COLLECT_BLOCK), 0, s, n, status, idx, count);
  return hipGetLastError();
}

size_t flint_ws_packed_bytes() { return (size_t)(4 * 2 + 4 * 8 + 8) * 64 * 16; }

hipError_t flint_ws_pack(const int* dims, const __bf16* const* WT, void* packed, hipStream_t s) {
  if (!flint_ws_fused_fits(dims)) return hipErrorInvalidValue;
  const int total = (int)(flint_ws_packed_bytes() / 16);
  hipLaunchKernelGGL((flint_ws_k::pack_kernel<4, 4>), dim3((total + 255) / 256), dim3(256), 0, s, dims[0], dims[3], WT[0],
                     WT[1], WT[2], static_cast<flint_ws_k::bf16x8*>(packed));
  return hipGetLastError();
}

hipError_t flint_ws_fused(long n, long ld, int ns, const double* Y0, const double* T0, const double* p,
                          const double* tabh, int Th_min, int Th_max, const double* scal, const int* dims,
                          const void* packed, const float* const* b, double* G, int blocks, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (!flint_ws_fused_fits(dims)) return hipErrorInvalidValue;
  const long tiles = (n + 31) / 32;
  const long need = (tiles + 7) / 8;
  const unsigned nb = (unsigned)std::max<long>(1, std::min<long>(blocks, need));
  hipLaunchKernelGGL((flint_ws_k::fused_kernel<4, 4>), dim3(nb), dim3(512), 0, s, n, ld, ns, Y0, T0, p, tabh, Th_min,
                     Th_max, scal, dims[0], dims[3], static_cast<const flint_ws_k::bf16x8*>(packed), b[0], b[1], b[2],
                     G);
  return hipGetLastError();
}
