// Device-side helpers shared by the hand-written and the generated gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "flint_registry.h"
#include "flint_reactor.h"

// 16-byte load of two adjacent table doubles (tables are padded so every pair is aligned).
__device__ __forceinline__ double2 ld2(const double* __restrict__ p) {
  return *reinterpret_cast<const double2*>(p);
}

// Output that is written once and not read again by the kernel: non-temporal vector store,
// so the multi-GB Jacobian stream does not evict the rate tables from L2.
__device__ __forceinline__ void stream_store(double* p, double v) { __builtin_nontemporal_store(v, p); }
// Input that is read once: non-temporal load, so the state stream does not evict the tables the
// kernel gathers from L2.
__device__ __forceinline__ double stream_load(const double* p) { return __builtin_nontemporal_load(p); }

// Returns p unchanged but opaque to the optimiser, so loads through it are issued where they are
// written instead of being merged with earlier loads of the same table row (which would keep
// hundreds of values live across the Jacobian columns).
__device__ __forceinline__ const double* opaque(const double* p) {
  unsigned long long v = reinterpret_cast<unsigned long long>(p);
  asm volatile("" : "+v"(v));
  return reinterpret_cast<const double*>(v);
}

struct NullSink {
  __device__ __forceinline__ void operator()(int, int, double) const {}
};

struct NullOut {
  template <class A>
  __device__ __forceinline__ void operator()(const A&) const {}
};

#define FLINT_CHECK_LAUNCH() hipGetLastError()

// Registers the launchers of one generated mechanism with the C-ABI dispatcher.
#define FLINT_REGISTER_MECH(ID, NSPACE)                                                                 \
  namespace {                                                                                           \
  hipError_t ID##_wdot(long n, long ld, const double* rho, const double* T, double* w, int* st,         \
                       const double* tabk, int Tk_min, int Tk_max, hipStream_t s) {                     \
    if (n <= 0) return hipSuccess;                                                                      \
    const unsigned blocks = (unsigned)((n + 255) / 256);                                                \
    hipLaunchKernelGGL(NSPACE::wdot_kernel, dim3(blocks), dim3(256), 0, s, n, ld, rho, T, w, st, tabk,  \
                       Tk_min, Tk_max);                                                                 \
    return hipGetLastError();                                                                           \
  }                                                                                                     \
  hipError_t ID##_rhs_jac(long n, long ld, const double* Z, double* F, double* J, int* st,              \
                          const double* tabk, int Tk_min, int Tk_max, const double* tabh, int Th_min,   \
                          int Th_max, hipStream_t s) {                                                  \
    if (n <= 0) return hipSuccess;                                                                      \
    const int B = NSPACE::JAC_BLOCK;                                                                    \
    const unsigned blocks = (unsigned)((n + B - 1) / B);                                                \
    if (J)                                                                                              \
      hipLaunchKernelGGL(NSPACE::rhs_jac_kernel<true>, dim3(blocks), dim3(B), 0, s, n, ld, Z, F, J,     \
                         st, tabk, Tk_min, Tk_max, tabh, Th_min, Th_max);                               \
    else                                                                                                \
      hipLaunchKernelGGL(NSPACE::rhs_jac_kernel<false>, dim3(blocks), dim3(B), 0, s, n, ld, Z, F, J,    \
                         st, tabk, Tk_min, Tk_max, tabh, Th_min, Th_max);                               \
    return hipGetLastError();                                                                           \
  }                                                                                                     \
  const flint_mech_ops ID##_ops = {#ID,         NSPACE::NS,          NSPACE::WK, ID##_wdot, ID##_rhs_jac, \
                                   &NSPACE::host_prog, nullptr, 0,          nullptr};                        \
  const int ID##_reg = flint_register_mech(&ID##_ops);                                                  \
  }

// As FLINT_REGISTER_MECH, for a mechanism that also has an equilibrium kernel (NSPACE::Cea).
#define FLINT_REGISTER_MECH_CEA(ID, NSPACE)                                                             \
  namespace {                                                                                           \
  hipError_t ID##_wdot(long n, long ld, const double* rho, const double* T, double* w, int* st,         \
                       const double* tabk, int Tk_min, int Tk_max, hipStream_t s) {                     \
    if (n <= 0) return hipSuccess;                                                                      \
    const unsigned blocks = (unsigned)((n + 255) / 256);                                                \
    hipLaunchKernelGGL(NSPACE::wdot_kernel, dim3(blocks), dim3(256), 0, s, n, ld, rho, T, w, st, tabk,  \
                       Tk_min, Tk_max);                                                                 \
    return hipGetLastError();                                                                           \
  }                                                                                                     \
  hipError_t ID##_rhs_jac(long n, long ld, const double* Z, double* F, double* J, int* st,              \
                          const double* tabk, int Tk_min, int Tk_max, const double* tabh, int Th_min,   \
                          int Th_max, hipStream_t s) {                                                  \
    if (n <= 0) return hipSuccess;                                                                      \
    const int B = NSPACE::JAC_BLOCK;                                                                    \
    const unsigned blocks = (unsigned)((n + B - 1) / B);                                                \
    if (J)                                                                                              \
      hipLaunchKernelGGL(NSPACE::rhs_jac_kernel<true>, dim3(blocks), dim3(B), 0, s, n, ld, Z, F, J,     \
                         st, tabk, Tk_min, Tk_max, tabh, Th_min, Th_max);                               \
    else                                                                                                \
      hipLaunchKernelGGL(NSPACE::rhs_jac_kernel<false>, dim3(blocks), dim3(B), 0, s, n, ld, Z, F, J,    \
                         st, tabk, Tk_min, Tk_max, tabh, Th_min, Th_max);                               \
    return hipGetLastError();                                                                           \
  }                                                                                                     \
  hipError_t ID##_cea(const flint_cea::CeaProg& P, bool hp, long n, long ld, const double* p,           \
                      const double* T, const double* rho, const double* guess, double* To, double* Yo,  \
                      int* it, int* st, hipStream_t s) {                                                \
    return flint_cea::launch<NSPACE::Cea>(P, hp, n, ld, p, T, rho, guess, To, Yo, it, st, s);           \
  }                                                                                                     \
  const flint_mech_ops ID##_ops = {#ID,           NSPACE::NS,       NSPACE::WK, ID##_wdot, ID##_rhs_jac, \
                                   &NSPACE::host_prog, ID##_cea, NSPACE::Cea::NE, NSPACE::cea_A};        \
  const int ID##_reg = flint_register_mech(&ID##_ops);                                                  \
  }
