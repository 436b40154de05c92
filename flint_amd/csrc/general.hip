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
// kernel keeps the reference's operation order with contraction off.
//
// The program is read by the whole wavefront at uniform addresses (scalar loads): one compact
// record per reaction, or its entry lists when it does not fit the record.  The third-body sums
// are formed once per distinct efficiency set (reactions with equal rows get bit-identical sums).
// Up to 32 species the concentrations sit in registers; beyond that in LDS.  The sums omegadot_k
// sit in LDS in either case.
#include "flint_device.h"
#include "flint_general.h"

namespace flint_general_k {

using flint_general::Ent;
using flint_general::GenProg;
using flint_general::Hdr;
using flint_general::kMaxSets;
using flint_general::kRealExp;
using flint_general::Rec;
constexpr int BLOCK = 128;

// x**n for an integer-valued n >= 1 by binary powering (n = 1 returns x, as the product 1 * x does)
__device__ __forceinline__ double powi_pos(double a, int b) {
#pragma clang fp contract(off)
  if (b == 1) return a;
  double r = 1.0;
  for (;;) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return r;
}

// One concentration factor c**e of a forward or reverse product.  The exponent is wave-uniform
// (it comes from the program), so each branch here is a scalar branch.
__device__ __noinline__ double factor_slow(double c, const Ent& x) {
  if (x.e == kRealExp) return pow(c, x.v);
  if (x.e == 0) return 1.0;
  return 1.0 / powi_pos(c, -x.e);
}

__device__ __forceinline__ double factor(double c, const Ent& x) {
  if (x.e > 0) return powi_pos(c, x.e);
  return factor_slow(c, x);
}

__device__ __forceinline__ double lerp(double a, double b, double Td) {
#pragma clang fp contract(off)
  return a + (b - a) * Td;
}

__device__ __forceinline__ double interp(const double* __restrict__ r0, const double* __restrict__ r1, int col,
                                         double Td) {
#pragma clang fp contract(off)
  const double a = r0[col], b = r1[col];
  return a + (b - a) * Td;
}

// The reaction loop over a cell, shared by both storage forms of the concentrations (c) and the
// running sums (add).  Every list index is wave-uniform: it comes from the program, not the cell.
// Columns c and c + 1 of a kinetics-table row (8-byte aligned: Lindemann triples may start at an
// odd column).
__device__ __forceinline__ double2 ld2u(const double* __restrict__ p) {
  double2 v;
  __builtin_memcpy(&v, p, sizeof(v));
  return v;
}

// The rates of one reaction from its kinetics-table columns (two rows bracketing T); a and b hold
// columns c and c + 1 of the two rows, loaded ahead by the caller.
__device__ __forceinline__ void rates(int kind, int c, double coM, double pf, double pb, double2 a, double2 b,
                                      const double* __restrict__ r0, const double* __restrict__ r1, double Td,
                                      double& rf, double& rb) {
#pragma clang fp contract(off)
  if (kind == 0) {
    rf = lerp(a.x, b.x, Td) * pf * coM;
    rb = lerp(a.y, b.y, Td) * pb * coM;
  } else if (kind == 1) {
    // Troe quadruples start at even columns of rows of even width: 16-byte loads
    const double2 a0 = a, b0 = b, a1 = ld2(r0 + c + 2), b1 = ld2(r1 + c + 2);
    const double kinf = lerp(a0.x, b0.x, Td), k0 = lerp(a0.y, b0.y, Td);
    const double kc = lerp(a1.x, b1.x, Td), Fcent = lerp(a1.y, b1.y, Td);
    const double Pr = k0 * coM / kinf;
    const double lFc = log10(Fcent), lPr = log10(Pr);
    const double cc = -0.4 - 0.67 * lFc, nn = 0.75 - 1.27 * lFc;
    const double f1 = (lPr + cc) / (nn - 0.14 * (lPr + cc));
    const double k1 = kinf * (Pr / (1.0 + Pr)) * pow(10.0, lFc / (1.0 + f1 * f1));
    rf = k1 * pf;
    rb = k1 / kc * pb;
  } else {
    const double kinf = lerp(a.x, b.x, Td), k0 = lerp(a.y, b.y, Td), kc = interp(r0, r1, c + 2, Td);
    const double Pr = k0 * coM / kinf;
    const double k1 = kinf * (Pr / (1.0 + Pr));
    rf = k1 * pf;
    rb = k1 / kc * pb;
  }
}

// A reaction from its entry lists (any stoichiometry and exponents).  Inlined: a call would take
// the address of the cell arrays and send them to scratch memory.
template <class Cell>
__device__ __forceinline__ void list_reaction(const GenProg& P, int j, Cell& S, double2 a, double2 b,
                                              const double* __restrict__ r0, const double* __restrict__ r1,
                                              double Td) {
#pragma clang fp contract(off)
  const Ent* __restrict__ E = P.ent;
  const Hdr h = P.hdr[j];
  double coM = 1.0;
  if (h.kind & 4) {
    coM = 0.0;
    for (int t = h.m0; t < h.f0; ++t) {
      const Ent x = E[t];
      coM = coM + S.c(x.sp) * x.v;
    }
  }
  double pf = 1.0, pb = 1.0;
  for (int t = h.f0; t < h.r0; ++t) {
    const Ent x = E[t];
    pf = pf * factor(S.c(x.sp), x);
  }
  for (int t = h.r0; t < h.o0; ++t) {
    const Ent x = E[t];
    pb = pb * factor(S.c(x.sp), x);
  }
  double rf, rb;
  rates(h.kind & 3, h.col, coM, pf, pb, a, b, r0, r1, Td, rf, rb);
  const double net = rf - rb;
  for (int t = h.o0; t < h.end; ++t) {
    const Ent x = E[t];
    S.add(x.sp, x.v * net);
  }
}

__device__ __forceinline__ int byte_of(unsigned w, int q) { return (int)((w >> (8 * q)) & 255u); }

template <class Cell>
__device__ __forceinline__ void reactions(const GenProg& P, Cell& S, const double* __restrict__ r0,
                                          const double* __restrict__ r1, double Td) {
#pragma clang fp contract(off)
  // third-body concentrations, one per distinct efficiency set, each summed in species order
  double msum[kMaxSets];
#pragma unroll
  for (int q = 0; q < kMaxSets; ++q) {
    double m = 0.0;
    if (q < P.nsets)
      for (int t = P.set_ptr[q]; t < P.set_ptr[q + 1]; ++t) {
        const Ent x = P.set_ent[t];
        m = m + S.c(x.sp) * x.v;
      }
    msum[q] = m;
  }
  // Software pipeline over the reactions: the table columns of reaction j + 1 are requested
  // before reaction j is evaluated (vector loads complete in order, so waiting for j's does not
  // wait for them), and its record is read meanwhile.
  Rec R = P.rec[0];
  int c1 = P.col[1];
  double2 a = ld2u(r0 + R.col), b = ld2u(r1 + R.col);
  for (int j = 0; j < P.nr; ++j) {
    const double2 an = ld2u(r0 + c1), bn = ld2u(r1 + c1);
    const int c2 = P.col[j + 2];
    const Rec Rn = P.rec[j + 1];
    if (R.kind & 8) {
      list_reaction(P, j, S, a, b, r0, r1, Td);
    } else {
      const double coM = (R.kind & 4) ? msum[R.mset & (kMaxSets - 1)] : 1.0;
      const int nf = R.cnt & 15, nrv = (R.cnt >> 4) & 15, no = R.cnt >> 8;
      double pf = 1.0, pb = 1.0;
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (q < nf) pf = pf * powi_pos(S.c(byte_of(R.fsp, q)), byte_of(R.fexp, q));
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (q < nrv) pb = pb * powi_pos(S.c(byte_of(R.rsp, q)), byte_of(R.rexp, q));
      double rf, rb;
      rates(R.kind & 3, R.col, coM, pf, pb, a, b, r0, r1, Td, rf, rb);
      const double net = rf - rb;
#pragma unroll
      for (int q = 0; q < 8; ++q)
        if (q < no) S.add(byte_of(q < 4 ? R.osp0 : R.osp1, q & 3), R.wd[q] * net);
    }
    R = Rn;
    a = an;
    b = bn;
    c1 = c2;
  }
}

// The running sums omegadot_k of a cell: one LDS column per thread, om[k * BLOCK], accumulated by
// no-return LDS adds (ds_add_f64).  Each is issued without a wait, and one thread's adds to one
// address execute in program order, so every sum is formed in the reference's order.
__device__ __forceinline__ void lds_add(double* om, int k, double v) {
  __hip_atomic_fetch_add(om + k * BLOCK, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Up to 16 NCH species: the concentrations of a cell in registers, indexed by the wave-uniform
// species numbers of the program (indexed register moves, no memory traffic).  They come in
// chunks of 16, the size the compiler keeps in registers under dynamic indexing.
template <int NCH>
struct RegCell {
  double (&c0)[16];
  double (&c1)[16];
  double* om;
  // each chunk is read on its own and the values selected (a pointer select between the arrays
  // would send them to scratch memory)
  __device__ __forceinline__ double c(int k) const {
    const double a = c0[k & 15];
    if (NCH == 1) return a;
    const double b = c1[k & 15];
    return k < 16 ? a : b;
  }
  __device__ __forceinline__ void add(int k, double v) { lds_add(om, k, v); }
};

// Any species count: the concentrations in a second LDS column per thread.
struct LdsCell {
  const double* co;
  double* om;
  __device__ __forceinline__ double c(int k) const { return co[k * BLOCK]; }
  __device__ __forceinline__ void add(int k, double v) { lds_add(om, k, v); }
};

struct Row {
  const double *r0, *r1;
  double Td;
  int st;
};

__device__ __forceinline__ Row table_row(const GenProg& P, const double* __restrict__ tabk, int Tk_min, int Tk_max,
                                         double temp) {
  Row R;
  R.st = (temp >= (double)Tk_min && temp <= (double)Tk_max) ? 0 : 1;
  const int T_i = (int)temp;
  R.Td = temp - (double)T_i;
  const int row = min(max(T_i, Tk_min), Tk_max) - Tk_min;
  R.r0 = tabk + (long)row * P.wk;
  R.r1 = R.r0 + P.wk;
  return R;
}

template <int NCH>
__global__ __launch_bounds__(BLOCK) void general_reg_kernel(GenProg P, const double* __restrict__ tabk, int Tk_min,
                                                            int Tk_max, const double* __restrict__ W, long n, long ld,
                                                            const double* __restrict__ rho,
                                                            const double* __restrict__ T, double* __restrict__ wdot,
                                                            int* __restrict__ status) {
  extern __shared__ double smem[];
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int ns = P.ns;
  double c0[16], c1[16];
  double* om = smem + threadIdx.x;
  // the densities read unguarded (rows clamped into the state) and the guards applied to the
  // values: a guarded load is a branch with its own wait, and the species' loads would go out one
  // after another
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int k0 = min(k, ns - 1);
    const double v0 = stream_load(rho + (long)k0 * ld + i) / W[k0];
    c0[k] = k < ns ? v0 : 0.0;
    if (NCH > 1) {
      const int k1 = min(k + 16, ns - 1);
      const double v1 = stream_load(rho + (long)k1 * ld + i) / W[k1];
      c1[k] = k + 16 < ns ? v1 : 0.0;
    }
  }
  for (int k = 0; k < ns; ++k) om[k * BLOCK] = 0.0;
  RegCell<NCH> S{c0, c1, om};
  const Row R = table_row(P, tabk, Tk_min, Tk_max, stream_load(T + i));
  reactions(P, S, R.r0, R.r1, R.Td);
  for (int k = 0; k < ns; ++k) stream_store(wdot + (long)k * ld + i, om[k * BLOCK]);
  if (status) status[i] = R.st;
}

__global__ __launch_bounds__(BLOCK) void general_lds_kernel(GenProg P, const double* __restrict__ tabk, int Tk_min,
                                                            int Tk_max, const double* __restrict__ W, long n, long ld,
                                                            const double* __restrict__ rho,
                                                            const double* __restrict__ T, double* __restrict__ wdot,
                                                            int* __restrict__ status) {
  extern __shared__ double smem[];
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int ns = P.ns;
  double* co = smem + (long)ns * BLOCK + threadIdx.x;
  double* om = smem + threadIdx.x;
  for (int k = 0; k < ns; ++k) {
    co[k * BLOCK] = stream_load(rho + (long)k * ld + i) / W[k];
    om[k * BLOCK] = 0.0;
  }
  LdsCell S{co, om};
  const Row R = table_row(P, tabk, Tk_min, Tk_max, stream_load(T + i));
  reactions(P, S, R.r0, R.r1, R.Td);
  for (int k = 0; k < ns; ++k) stream_store(wdot + (long)k * ld + i, om[k * BLOCK]);
  if (status) status[i] = R.st;
}

}  // namespace flint_general_k

hipError_t flint_general_launch(const flint_general::GenProg& P, const double* tabk, int Tk_min, int Tk_max,
                                const double* W, long n, long ld, const double* rho, const double* T, double* wdot,
                                int* status, hipStream_t s) {
  using namespace flint_general_k;
  if (n <= 0) return hipSuccess;
  const dim3 grid((unsigned)((n + BLOCK - 1) / BLOCK));
  const size_t om_bytes = (size_t)P.ns * BLOCK * sizeof(double);
  if (P.ns <= 16) {
    hipLaunchKernelGGL(general_reg_kernel<1>, grid, dim3(BLOCK), om_bytes, s, P, tabk, Tk_min, Tk_max, W, n, ld,
                       rho, T, wdot, status);
  } else if (P.ns <= 32) {
    hipLaunchKernelGGL(general_reg_kernel<2>, grid, dim3(BLOCK), om_bytes, s, P, tabk, Tk_min, Tk_max, W, n, ld,
                       rho, T, wdot, status);
  } else {
    const size_t lds = 2 * om_bytes;  // at most 64 species: 128 KiB
    if (lds > 65536) {
      const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&general_lds_kernel),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(general_lds_kernel, grid, dim3(BLOCK), lds, s, P, tabk, Tk_min, Tk_max, W, n, ld, rho, T,
                       wdot, status);
  }
  return hipGetLastError();
}
