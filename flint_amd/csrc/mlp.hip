// Multilayer perceptron on gfx950 MFMA: BF16 weights and activations, FP32 accumulation, FP32
// master weights with Adam.  Used as the learned reactor step and as the equilibrium warm start
// (no counterpart in the FLINT reference).
//
// Every product is one kernel: a wavefront computes a 32 x 32 output tile with
// v_mfma_f32_32x32x16_bf16 (lane l holds row/column l % 32 and k = 8 (l / 32) + e of the A and B
// fragments; accumulator r is row 8 (r / 4) + 4 (l / 32) + r % 4 of column l % 32), walking K in
// steps of 16 with operands read through (row, column) strides, so the forward pass, the input
// gradient (W transposed) and the weight gradient (activations transposed) share it.  The
// epilogue adds the bias and applies GELU (erf form) or its derivative in FP32.
//
// Layout: activations row-major [n][features]; layer l has W_l [d_l][d_{l+1}] (row-major) and b_l,
// stored flat per layer (W then b) in the FP32 master vector.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/flint_amd.h"
#include "flint_internal.h"
#include "flint_mlp.h"

namespace flint_mlp_k {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

enum { EPI_LINEAR = 0, EPI_GELU = 1, EPI_DGELU = 2, EPI_STORE = 3 };

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float dgelu(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// C[M][N] = A[M][K] B[K][N] with A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn].
template <int EPI>
__global__ __launch_bounds__(64) void gemm_kernel(int M, int N, int K, const __bf16* __restrict__ A, long sam, long sak,
                                                  const __bf16* __restrict__ B, long sbk, long sbn,
                                                  const float* __restrict__ bias, const float* __restrict__ Zaux,
                                                  float* __restrict__ O32, __bf16* __restrict__ O16, long ldo) {
  const int l = threadIdx.x;
  const int m0 = blockIdx.x * 32, n0 = blockIdx.y * 32;
  const int am = m0 + (l & 31), bn = n0 + (l & 31), kk = 8 * (l >> 5);
  f32x16 acc = {};
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x8 a, b;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int k = k0 + kk + e;
      a[e] = (am < M && k < K) ? A[(long)am * sam + (long)k * sak] : (__bf16)0.0f;
      b[e] = (bn < N && k < K) ? B[(long)k * sbk + (long)bn * sbn] : (__bf16)0.0f;
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int m = m0 + 8 * (r >> 2) + 4 * (l >> 5) + (r & 3);
    const int n = n0 + (l & 31);
    if (m >= M || n >= N) continue;
    const long o = (long)m * ldo + n;
    float v = acc[r];
    if (EPI == EPI_LINEAR) {
      if (bias) v += bias[n];
      if (O32) O32[o] = v;
      if (O16) O16[o] = (__bf16)v;
    } else if (EPI == EPI_GELU) {
      v += bias[n];
      if (O32) O32[o] = v;
      O16[o] = (__bf16)gelu(v);
    } else if (EPI == EPI_DGELU) {
      v *= dgelu(Zaux[o]);
      if (O32) O32[o] = v;
      O16[o] = (__bf16)v;
    } else {
      O32[o] = v;
    }
  }
}

// Forward product for K % 8 == 0: A [M][K] row-major and the transposed weights WT [N][K], so that
// every fragment is one 16-byte load; a wavefront computes a 32 x (32 NT) strip, reusing its A
// fragment for NT tiles.  Same epilogues as gemm_kernel.
template <int EPI, int NT>
__global__ __launch_bounds__(64) void gemm_fwd_kernel(int M, int N, int K, const __bf16* __restrict__ A,
                                                      const __bf16* __restrict__ WT, const float* __restrict__ bias,
                                                      float* __restrict__ O32, __bf16* __restrict__ O16, long ldo) {
  const int l = threadIdx.x;
  const int m0 = blockIdx.x * 32, n0 = blockIdx.y * 32 * NT;
  const int am = m0 + (l & 31), kk = 8 * (l >> 5);
  f32x16 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = f32x16{};
  const bf16x8 zero = {};
  for (int k0 = 0; k0 < K; k0 += 16) {
    const int k = k0 + kk;
    const bf16x8 a = (am < M && k < K) ? *reinterpret_cast<const bf16x8*>(A + (long)am * K + k) : zero;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int bn = n0 + 32 * t + (l & 31);
      const bf16x8 b = (bn < N && k < K) ? *reinterpret_cast<const bf16x8*>(WT + (long)bn * K + k) : zero;
      acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int m = m0 + 8 * (r >> 2) + 4 * (l >> 5) + (r & 3);
      const int n = n0 + 32 * t + (l & 31);
      if (m >= M || n >= N) continue;
      const long o = (long)m * ldo + n;
      float v = acc[t][r];
      if (EPI == EPI_LINEAR) {
        if (bias) v += bias[n];
        if (O32) O32[o] = v;
        if (O16) O16[o] = (__bf16)v;
      } else {
        v += bias[n];
        if (O32) O32[o] = v;
        O16[o] = (__bf16)gelu(v);
      }
    }
  }
}

// Forward product with the weights staged in LDS, for K <= 256, N % 32 == 0 and N <= 256
// (NT = N / 32): each block copies WT [N][K] into LDS once, zero-filled to a multiple of 16 columns
// and padded by 16 bytes per row (so that the 32 rows a fragment read touches start in different
// banks); each wavefront then computes 32 x N output strips over a grid-stride loop.  All A
// fragments of a strip (up to 16 k-steps) are loaded into registers before its MFMAs, so the loads
// overlap each other and the other wavefronts' MFMAs; every B fragment comes from LDS.  AL: K % 8
// == 0, one 16-byte load per A fragment (otherwise element loads, zero beyond K).  The same k-steps,
// zero tail and MFMA order per output tile as gemm_kernel and gemm_fwd_kernel (bit-identical
// results), the same epilogues.
template <bool AL>
__device__ __forceinline__ bf16x8 a_fragment(const __bf16* __restrict__ A, int M, int K, int am, int k) {
  bf16x8 v = {};
  if (am >= M) return v;
  if (AL) {
    if (k < K) v = *reinterpret_cast<const bf16x8*>(A + (long)am * K + k);
  } else {
#pragma unroll
    for (int e = 0; e < 8; ++e)
      if (k + e < K) v[e] = A[(long)am * K + k + e];
  }
  return v;
}

template <int EPI, int NT, bool AL>
__global__ __launch_bounds__(512) void gemm_fwd_lds_kernel(int M, int N, int K, const __bf16* __restrict__ A,
                                                          const __bf16* __restrict__ WT,
                                                          const float* __restrict__ bias, float* __restrict__ O32,
                                                          __bf16* __restrict__ O16, long ldo) {
  constexpr int KSMAX = 16;  // k-steps of 16: K <= 256
  extern __shared__ __align__(16) char smem_raw[];
  __bf16* ws = reinterpret_cast<__bf16*>(smem_raw);
  const int K16 = (K + 15) & ~15, KP = K16 + 8, ks = K16 / 16;
  if (AL) {
    const int kc = K16 / 8, kv = K / 8;
    for (int i = threadIdx.x; i < N * kc; i += blockDim.x) {
      const int r = i / kc, c = i - r * kc;
      bf16x8 v = {};
      if (c < kv) v = *reinterpret_cast<const bf16x8*>(WT + (long)r * K + 8 * c);
      *reinterpret_cast<bf16x8*>(ws + r * KP + 8 * c) = v;
    }
  } else {
    for (int i = threadIdx.x; i < N * K16; i += blockDim.x) {
      const int r = i / K16, c = i - r * K16;
      ws[r * KP + c] = c < K ? WT[(long)r * K + c] : (__bf16)0.0f;
    }
  }
  __syncthreads();
  const int l = threadIdx.x & 63, nw = blockDim.x >> 6;
  const int kk = 8 * (l >> 5);
  const int tiles = (M + 31) / 32;
  for (int tile = blockIdx.x * nw + (threadIdx.x >> 6); tile < tiles; tile += gridDim.x * nw) {
    const int m0 = tile * 32, am = m0 + (l & 31);
    bf16x8 af[KSMAX];
#pragma unroll
    for (int q = 0; q < KSMAX; ++q)
      if (q < ks) af[q] = a_fragment<AL>(A, M, K, am, 16 * q + kk);
    // the column tiles in groups of at most 4 (64 accumulator registers), reusing the A fragments
    constexpr int G = NT < 4 ? NT : 4;
#pragma unroll 1
    for (int t0 = 0; t0 < NT; t0 += G) {
    f32x16 acc[G];
#pragma unroll
    for (int t = 0; t < G; ++t) acc[t] = f32x16{};
#pragma unroll
    for (int q = 0; q < KSMAX; ++q) {
      if (q < ks) {
#pragma unroll
        for (int t = 0; t < G; ++t) {
          const bf16x8 b = *reinterpret_cast<const bf16x8*>(ws + (32 * (t0 + t) + (l & 31)) * KP + 16 * q + kk);
          acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[q], b, acc[t], 0, 0, 0);
        }
      }
    }
    // the bias of this lane's columns is loaded once, before the stores: a load waited for between
    // stores would wait for the stores as well
    float bv[G];
#pragma unroll
    for (int t = 0; t < G; ++t) bv[t] = bias ? bias[32 * (t0 + t) + (l & 31)] : 0.0f;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int m = m0 + 8 * (r >> 2) + 4 * (l >> 5) + (r & 3);
      if (m >= M) continue;
      const long o0 = (long)m * ldo + 32 * t0 + (l & 31);
#pragma unroll
      for (int t = 0; t < G; ++t) {
        const long o = o0 + 32 * t;
        float v = acc[t][r];
        if (EPI == EPI_LINEAR) {
          if (bias) v += bv[t];
          if (O32) O32[o] = v;
          if (O16) O16[o] = (__bf16)v;
        } else {
          v += bv[t];
          if (O32) O32[o] = v;
          O16[o] = (__bf16)gelu(v);
        }
      }
    }
    }
  }
}

// WT_l [dout][din] from W_l [din][dout] (bf16), all layers
__global__ void transpose_w16(int din, int dout, const __bf16* __restrict__ w, __bf16* __restrict__ wt) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)din * dout) return;
  const int r = (int)(i / dout), c = (int)(i % dout);
  wt[(long)c * din + r] = w[i];
}

__global__ void to_bf16(long n, const float* __restrict__ x, __bf16* __restrict__ y) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (__bf16)x[i];
}

// mean squared error over n x d outputs: dOut = 2 (out - t) / (n d); loss accumulates sum / (n d)
__global__ void mse_kernel(long n, int d, const float* __restrict__ out, const float* __restrict__ t,
                           float* __restrict__ g32, __bf16* __restrict__ g16, float* __restrict__ loss) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  float e2 = 0.0f;
  if (i < n * d) {
    const float r = out[i] - t[i];
    const float g = 2.0f * r / (float)((double)n * d);
    g32[i] = g;
    g16[i] = (__bf16)g;
    e2 = r * r;
  }
  for (int m = 32; m >= 1; m >>= 1) e2 += __shfl_xor(e2, m, 64);
  if (loss && (threadIdx.x & 63) == 0) atomicAdd(loss, e2 / (float)((double)n * d));
}

__global__ void zero_f32(long n, float* x) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0.0f;
}

// column sums of g [n][d] (bias gradient): one block per column block of 64
__global__ __launch_bounds__(256) void colsum_kernel(long n, int d, const float* __restrict__ g,
                                                     float* __restrict__ out) {
  __shared__ float part[4][64];
  const int c = blockIdx.x * 64 + (threadIdx.x & 63);
  const int row0 = threadIdx.x >> 6;
  float s = 0.0f;
  if (c < d)
    for (long r = row0; r < n; r += 4) s += g[r * d + c];
  part[row0][threadIdx.x & 63] = s;
  __syncthreads();
  if (threadIdx.x < 64 && c < d) out[c] = part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] + part[3][threadIdx.x];
}

// Adam (Kingma & Ba 2015) on the FP32 master vector; refreshes the BF16 copy of the weights
__global__ void adam_kernel(long np, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, __bf16* __restrict__ w16, const int* __restrict__ w16_index,
                            float lr, float b1, float b2, float eps, float c1, float c2) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  const int j = w16_index[i];
  if (j >= 0) w16[j] = (__bf16)wi;
}

__global__ void sync_bf16(long np, const float* __restrict__ w, __bf16* __restrict__ w16,
                          const int* __restrict__ w16_index) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const int j = w16_index[i];
  if (j >= 0) w16[j] = (__bf16)w[i];
}

template <int EPI>
hipError_t gemm(int M, int N, int K, const __bf16* A, long sam, long sak, const __bf16* B, long sbk, long sbn,
                const float* bias, const float* Zaux, float* O32, __bf16* O16, long ldo, hipStream_t s) {
  dim3 grid((M + 31) / 32, (N + 31) / 32);
  hipLaunchKernelGGL(gemm_kernel<EPI>, grid, dim3(64), 0, s, M, N, K, A, sam, sak, B, sbk, sbn, bias, Zaux, O32, O16,
                     ldo);
  return hipGetLastError();
}

inline unsigned nb(long n, int b) { return (unsigned)((n + b - 1) / b); }

// blocks of the LDS-staged forward product: as many as fit on the GPU at once (LDS-limited), at
// most one per 8 output strips
template <int EPI, int NT, bool AL>
hipError_t gemm_fwd_lds(int M, int K, const __bf16* A, const __bf16* WT, const float* bias, float* O32, __bf16* O16,
                        long ldo, hipStream_t s) {
  constexpr int N = 32 * NT;
  const size_t lds = (size_t)N * (((K + 15) & ~15) + 8) * sizeof(__bf16);
  static int cus = 0;
  static size_t lds_set = 0;
  if (!cus) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
  }
  if (lds > lds_set) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&gemm_fwd_lds_kernel<EPI, NT, AL>),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    lds_set = lds;
  }
  const int per_cu = (int)std::max<size_t>(1, std::min<size_t>(4, (160 * 1024) / lds));
  const long strips = (M + 31) / 32;
  const int blocks = (int)std::max<long>(1, std::min<long>((long)cus * per_cu, (strips + 7) / 8));
  hipLaunchKernelGGL((gemm_fwd_lds_kernel<EPI, NT, AL>), dim3(blocks), dim3(512), lds, s, M, N, K, A, WT, bias, O32,
                     O16, ldo);
  return hipGetLastError();
}

template <int EPI, bool AL>
hipError_t gemm_fwd_lds_n(int M, int N, int K, const __bf16* A, const __bf16* WT, const float* bias, float* O32,
                          __bf16* O16, long ldo, hipStream_t s) {
  switch (N / 32) {
    case 1: return gemm_fwd_lds<EPI, 1, AL>(M, K, A, WT, bias, O32, O16, ldo, s);
    case 2: return gemm_fwd_lds<EPI, 2, AL>(M, K, A, WT, bias, O32, O16, ldo, s);
    case 4: return gemm_fwd_lds<EPI, 4, AL>(M, K, A, WT, bias, O32, O16, ldo, s);
    default: return gemm_fwd_lds<EPI, 8, AL>(M, K, A, WT, bias, O32, O16, ldo, s);
  }
}

// true when the LDS-staged forward product takes the shape
inline bool lds_shape(int N, int K, long ldo) {
  return K >= 1 && K <= 256 && (N == 32 || N == 64 || N == 128 || N == 256) && ldo >= N;
}

template <int EPI>
hipError_t gemm_fwd(int M, int N, int K, const __bf16* A, const __bf16* WT, const float* bias, float* O32,
                    __bf16* O16, long ldo, hipStream_t s) {
  if (lds_shape(N, K, ldo)) return gemm_fwd_lds_n<EPI, true>(M, N, K, A, WT, bias, O32, O16, ldo, s);
  if (N <= 32) {
    hipLaunchKernelGGL((gemm_fwd_kernel<EPI, 1>), dim3(nb(M, 32), 1), dim3(64), 0, s, M, N, K, A, WT, bias, O32, O16, ldo);
  } else if (N <= 64) {
    hipLaunchKernelGGL((gemm_fwd_kernel<EPI, 2>), dim3(nb(M, 32), 1), dim3(64), 0, s, M, N, K, A, WT, bias, O32, O16, ldo);
  } else {
    hipLaunchKernelGGL((gemm_fwd_kernel<EPI, 4>), dim3(nb(M, 32), nb(N, 128)), dim3(64), 0, s, M, N, K, A, WT, bias, O32,
                       O16, ldo);
  }
  return hipGetLastError();
}

}  // namespace flint_mlp_k

using namespace flint_mlp_k;

#define HIPCHK(expr)                                                                                      \
  do {                                                                                                    \
    hipError_t e_ = (expr);                                                                               \
    if (e_ != hipSuccess) return flint_fail(FLINT_EHIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

static int dalloc(void** p, size_t bytes) {
  HIPCHK(hipMalloc(p, bytes > 0 ? bytes : 16));
  return FLINT_OK;
}

flint_mlp::~flint_mlp() {
  for (void* p : {(void*)w32, (void*)g32, (void*)m32, (void*)v32, (void*)w16, (void*)w16t, (void*)w16_index, (void*)loss})
    if (p) (void)hipFree(p);
  for (void* p : ws) (void)hipFree(p);
}

// per-batch workspace: a16[l] (input of layer l, bf16), z32[l] (pre-activation), dz32/dz16 (gradients)
int flint_mlp::reserve(long n, bool train) {
  if (n <= cap && (!train || cap_train)) return FLINT_OK;
  const long c = std::max<long>(n, cap);
  for (void* p : ws) (void)hipFree(p);
  ws.clear();
  a16.assign(L + 1, nullptr);
  z32.assign(L, nullptr);
  dz32.assign(L, nullptr);
  dz16.assign(L, nullptr);
  for (int l = 0; l <= L; ++l) {
    void* p;
    if (int rc = dalloc(&p, (size_t)c * dims[l] * sizeof(__bf16))) return rc;
    ws.push_back(p);
    a16[l] = (__bf16*)p;
  }
  for (int l = 0; l < L; ++l) {
    void* p;
    if (int rc = dalloc(&p, (size_t)c * dims[l + 1] * sizeof(float))) return rc;
    ws.push_back(p);
    z32[l] = (float*)p;
    if (train) {
      if (int rc = dalloc(&p, (size_t)c * dims[l + 1] * sizeof(float))) return rc;
      ws.push_back(p);
      dz32[l] = (float*)p;
      if (int rc = dalloc(&p, (size_t)c * dims[l + 1] * sizeof(__bf16))) return rc;
      ws.push_back(p);
      dz16[l] = (__bf16*)p;
    }
  }
  cap = c;
  cap_train = train;
  return FLINT_OK;
}

// the transposed BF16 weights w16t, rebuilt after a weight update
int flint_mlp::refresh_wt(hipStream_t s) {
  if (wt_stale) {
    for (int l = 0; l < L; ++l)
      hipLaunchKernelGGL(transpose_w16, dim3(nb((long)dims[l] * dims[l + 1], 256)), dim3(256), 0, s, dims[l], dims[l + 1],
                         w16 + w16_off[l], w16t + w16t_off[l]);
    HIPCHK(hipGetLastError());
    wt_stale = false;
  }
  return FLINT_OK;
}

// forward from a16[0]; the last layer writes FP32 to out (z32[L-1] when out is null)
int flint_mlp::forward(long n, float* out, hipStream_t s, bool keep) {
  if (int rc = refresh_wt(s)) return rc;
  for (int l = 0; l < L; ++l) {
    const int din = dims[l], dout = dims[l + 1];
    const __bf16* W = w16 + w16_off[l];
    const __bf16* WT = w16t + w16t_off[l];
    const float* b = w32 + b_off[l];
    const bool fast = din % 8 == 0;
    hipError_t e;
    if (!fast && lds_shape(dout, din, dout)) {
      // K not a multiple of 8 (the input layer): element loads of A, weights zero-filled in LDS
      if (l + 1 < L)
        e = gemm_fwd_lds_n<EPI_GELU, false>((int)n, dout, din, a16[l], WT, b, keep ? z32[l] : nullptr, a16[l + 1],
                                            dout, s);
      else
        e = gemm_fwd_lds_n<EPI_LINEAR, false>((int)n, dout, din, a16[l], WT, b, out ? out : z32[l], nullptr, dout, s);
    } else if (l + 1 < L)
      e = fast ? gemm_fwd<EPI_GELU>((int)n, dout, din, a16[l], WT, b, keep ? z32[l] : nullptr, a16[l + 1], dout, s)
               : gemm<EPI_GELU>((int)n, dout, din, a16[l], din, 1, W, dout, 1, b, nullptr, keep ? z32[l] : nullptr,
                                a16[l + 1], dout, s);
    else
      e = fast ? gemm_fwd<EPI_LINEAR>((int)n, dout, din, a16[l], WT, b, out ? out : z32[l], nullptr, dout, s)
               : gemm<EPI_LINEAR>((int)n, dout, din, a16[l], din, 1, W, dout, 1, b, nullptr, out ? out : z32[l],
                                  nullptr, dout, s);
    if (e != hipSuccess) return flint_fail(FLINT_EHIP, std::string("mlp forward: ") + hipGetErrorString(e));
  }
  return FLINT_OK;
}

// gradients of the mean squared error into g32 (a16[0] holds the input, t the targets)
int flint_mlp::backward(long n, const float* t, hipStream_t s) {
  const int dL = dims[L];
  hipLaunchKernelGGL(zero_f32, dim3(1), dim3(64), 0, s, 1L, loss);
  hipLaunchKernelGGL(mse_kernel, dim3(nb(n * dL, 256)), dim3(256), 0, s, n, dL, z32[L - 1], t, dz32[L - 1],
                     dz16[L - 1], loss);
  for (int l = L - 1; l >= 0; --l) {
    const int din = dims[l], dout = dims[l + 1];
    hipError_t e = gemm<EPI_STORE>(din, dout, (int)n, a16[l], 1, din, dz16[l], dout, 1, nullptr, nullptr,
                                   g32 + w_off[l], nullptr, dout, s);
    if (e != hipSuccess) return flint_fail(FLINT_EHIP, std::string("mlp dW: ") + hipGetErrorString(e));
    hipLaunchKernelGGL(colsum_kernel, dim3(nb(dout, 64)), dim3(256), 0, s, n, dout, dz32[l], g32 + b_off[l]);
    if (l > 0) {
      e = gemm<EPI_DGELU>((int)n, din, dout, dz16[l], dout, 1, w16 + w16_off[l], 1, dout, nullptr, z32[l - 1],
                          dz32[l - 1], dz16[l - 1], din, s);
      if (e != hipSuccess) return flint_fail(FLINT_EHIP, std::string("mlp dX: ") + hipGetErrorString(e));
    }
  }
  HIPCHK(hipGetLastError());
  return FLINT_OK;
}

int flint_mlp::load_input(long n, const float* X, hipStream_t s) {
  hipLaunchKernelGGL(to_bf16, dim3(nb(n * dims[0], 256)), dim3(256), 0, s, n * dims[0], X, a16[0]);
  HIPCHK(hipGetLastError());
  return FLINT_OK;
}

extern "C" int flint_mlp_create(int nlayers, const int* dims, unsigned long long seed, flint_mlp** out) {
  if (!out || !dims || nlayers < 1 || nlayers > 16) return flint_fail(FLINT_EARG, "bad layer count");
  *out = nullptr;
  for (int l = 0; l <= nlayers; ++l)
    if (dims[l] < 1 || dims[l] > 4096) return flint_fail(FLINT_EARG, "layer widths must be in 1..4096");
  static std::atomic<unsigned long long> next_uid{1};
  auto* m = new flint_mlp();
  m->uid = next_uid++;
  m->L = nlayers;
  m->dims.assign(dims, dims + nlayers + 1);
  long np = 0, nw = 0;
  long nwt = 0;
  for (int l = 0; l < nlayers; ++l) {
    m->w16t_off.push_back(nwt);  // 16-byte aligned layer starts for the transposed copy
    nwt += ((long)dims[l] * dims[l + 1] + 7) / 8 * 8;
    m->w_off.push_back(np);
    m->w16_off.push_back(nw);
    np += (long)dims[l] * dims[l + 1];
    nw += (long)dims[l] * dims[l + 1];
    m->b_off.push_back(np);
    np += dims[l + 1];
  }
  m->np = np;
  m->nw = nw;
  // Glorot-uniform weights from a counter-based generator (splitmix64 of seed and index), zero bias
  std::vector<float> w(np, 0.0f);
  std::vector<int> idx(np, -1);
  for (int l = 0; l < nlayers; ++l) {
    const double lim = std::sqrt(6.0 / (dims[l] + dims[l + 1]));
    for (long i = 0; i < (long)dims[l] * dims[l + 1]; ++i) {
      uint64_t z = seed * 0x9E3779B97F4A7C15ull + (uint64_t)(m->w_off[l] + i) + 1;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      z ^= z >> 31;
      const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
      w[m->w_off[l] + i] = (float)((2.0 * u - 1.0) * lim);
      idx[m->w_off[l] + i] = (int)(m->w16_off[l] + i);
    }
  }
  int rc = 0;
  rc |= dalloc((void**)&m->w32, np * sizeof(float));
  rc |= dalloc((void**)&m->g32, np * sizeof(float));
  rc |= dalloc((void**)&m->m32, np * sizeof(float));
  rc |= dalloc((void**)&m->v32, np * sizeof(float));
  rc |= dalloc((void**)&m->w16, nw * sizeof(__bf16));
  rc |= dalloc((void**)&m->w16t, nwt * sizeof(__bf16));
  rc |= dalloc((void**)&m->w16_index, np * sizeof(int));
  rc |= dalloc((void**)&m->loss, sizeof(float));
  if (rc) {
    delete m;
    return rc;
  }
  std::vector<float> zeros(np, 0.0f);
  hipError_t e = hipMemcpy(m->w32, w.data(), np * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(m->m32, zeros.data(), np * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(m->v32, zeros.data(), np * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(m->w16_index, idx.data(), np * sizeof(int), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    delete m;
    return flint_fail(FLINT_EHIP, std::string("mlp upload: ") + hipGetErrorString(e));
  }
  hipLaunchKernelGGL(sync_bf16, dim3(nb(np, 256)), dim3(256), 0, 0, np, m->w32, m->w16, m->w16_index);
  e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    delete m;
    return flint_fail(FLINT_EHIP, std::string("mlp init: ") + hipGetErrorString(e));
  }
  *out = m;
  return FLINT_OK;
}

extern "C" void flint_mlp_destroy(flint_mlp* m) { delete m; }

extern "C" int flint_mlp_num_params(const flint_mlp* m, long* count) {
  if (!m || !count) return flint_fail(FLINT_EARG, "null argument");
  *count = m->np;
  return FLINT_OK;
}

extern "C" int flint_mlp_get_params(flint_mlp* m, float* host) {
  if (!m || !host) return flint_fail(FLINT_EARG, "null argument");
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipMemcpy(host, m->w32, m->np * sizeof(float), hipMemcpyDeviceToHost));
  return FLINT_OK;
}

extern "C" int flint_mlp_set_params(flint_mlp* m, const float* host) {
  if (!m || !host) return flint_fail(FLINT_EARG, "null argument");
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipMemcpy(m->w32, host, m->np * sizeof(float), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(sync_bf16, dim3(nb(m->np, 256)), dim3(256), 0, 0, m->np, m->w32, m->w16, m->w16_index);
  HIPCHK(hipDeviceSynchronize());
  m->wt_stale = true;
  m->version += 1;
  m->adam_t = 0;
  std::vector<float> zeros(m->np, 0.0f);
  HIPCHK(hipMemcpy(m->m32, zeros.data(), m->np * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(m->v32, zeros.data(), m->np * sizeof(float), hipMemcpyHostToDevice));
  return FLINT_OK;
}

extern "C" int flint_mlp_forward(flint_mlp* m, long n, const float* X, float* Y, void* stream) {
  if (!m || n < 0) return flint_fail(FLINT_EARG, "bad argument");
  if (n == 0) return FLINT_OK;
  if (!X || !Y) return flint_fail(FLINT_EARG, "null array");
  hipStream_t s = (hipStream_t)stream;
  if (int rc = m->reserve(n, false)) return rc;
  if (int rc = m->load_input(n, X, s)) return rc;
  return m->forward(n, Y, s, false);
}

extern "C" int flint_mlp_gradients(flint_mlp* m, long n, const float* X, const float* T, float* grad, float* loss,
                                   void* stream) {
  if (!m || n <= 0 || !X || !T) return flint_fail(FLINT_EARG, "bad argument");
  hipStream_t s = (hipStream_t)stream;
  if (int rc = m->reserve(n, true)) return rc;
  if (int rc = m->load_input(n, X, s)) return rc;
  if (int rc = m->forward(n, nullptr, s)) return rc;
  if (int rc = m->backward(n, T, s)) return rc;
  if (grad) HIPCHK(hipMemcpyAsync(grad, m->g32, m->np * sizeof(float), hipMemcpyDeviceToDevice, s));
  if (loss) HIPCHK(hipMemcpyAsync(loss, m->loss, sizeof(float), hipMemcpyDeviceToDevice, s));
  return FLINT_OK;
}

extern "C" int flint_mlp_train_step(flint_mlp* m, long n, const float* X, const float* T, float lr, float* loss,
                                    void* stream) {
  if (!m || n <= 0 || !X || !T || !(lr > 0.0f)) return flint_fail(FLINT_EARG, "bad argument");
  hipStream_t s = (hipStream_t)stream;
  if (int rc = flint_mlp_gradients(m, n, X, T, nullptr, loss, stream)) return rc;
  m->adam_t += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - std::pow(b1, (float)m->adam_t), c2 = 1.0f - std::pow(b2, (float)m->adam_t);
  hipLaunchKernelGGL(adam_kernel, dim3(nb(m->np, 256)), dim3(256), 0, s, m->np, m->w32, m->g32, m->m32, m->v32, m->w16,
                     m->w16_index, lr, b1, b2, 1e-8f, c1, c2);
  HIPCHK(hipGetLastError());
  m->wt_stale = true;
  m->version += 1;
  return FLINT_OK;
}
