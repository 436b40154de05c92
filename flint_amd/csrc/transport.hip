// Mixture transport of a cell by Wilke's rule: co_k_mi_lam_Wilke / co_k_mi_lam_Wilke_expr
// (Lib_ThermoTransport.f90:407-462) with co_fiij (:350-362), f_molecularWeight (:200-214) and
// f_tabT_expr (:531-542).  One thread per cell.
//
// The reference forms every interaction factor
//   phi_ij = (1 + sqrt(mu_i / (mu_j + 1e-20)) / (W_i/W_j)^(1/4))^2 / sqrt(8 (1 + W_i/W_j))
// with a division and two square roots, ns^2 times per cell.  Here the square root is split,
// sqrt(mu_i / (mu_j + 1e-20)) = s_i r_j with s_i = sqrt(mu_i), r_j = 1/sqrt(mu_j + 1e-20), and the
// square expanded, so the Wilke denominator becomes three matrix-vector products with constant
// per-mechanism matrices D, 2 D c, D c^2 (c = (W_i/W_j)^(-1/4), D = 1/sqrt(8 (1 + W_i/W_j))):
//   den_i = sum_j X_j D_ij phi-expanded = (D X)_i + s_i (2Dc (X r))_i + s_i^2 (Dc^2 (X r^2))_i.
// Every term is positive, so the rewrite costs a few ulps, not cancellation: 3 ns^2 FMAs per cell
// instead of ns^2 divisions and square roots.  The matrices are read at wave-uniform addresses
// (scalar loads); the per-cell vectors stay in registers for the compiled species counts (the
// densities are read once), and a generic loop form serves any other count up to
// FLINT_TRANSPORT_MAXNS.
#include "flint_device.h"
#include "flint_transport.h"

namespace flint_tr {

// f_tabT_expr of the (mu, k) pair of species s: Vij + (Viij - Vij) * Tdiff
__device__ __forceinline__ double2 interp(const double* __restrict__ r0, int ns, int s, double Td) {
#pragma clang fp contract(off)
  const double2 a = ld2(r0 + 2 * s), b = ld2(r0 + 2 * ns + 2 * s);
  return make_double2(a.x + (b.x - a.x) * Td, a.y + (b.y - a.y) * Td);
}

__device__ __forceinline__ bool in_table(const TrProg& P, double T) {
  return T >= (double)P.Tt_min && T < (double)P.Tt_max;
}

// 1 / Wmtot = sum_s rhoi_s / sum(rhoi) / W_s (f_molecularWeight)
__device__ __forceinline__ double inv_wmix(const double* __restrict__ c, long ld, int ns,
                                           const double* __restrict__ W) {
#pragma clang fp contract(off)
  double a = 0.0;
  for (int s = 0; s < ns; ++s) a = a + c[(long)s * ld];
  double b = 0.0;
  for (int s = 0; s < ns; ++s) b = b + c[(long)s * ld] / a / W[s];
  return b;
}

template <int NS>
__global__ __launch_bounds__(256) void wilke_kernel(TrProg P, long n, long ld, const double* __restrict__ rhoi,
                                                    const double* __restrict__ rho, const double* __restrict__ T,
                                                    double* __restrict__ mu, double* __restrict__ lam,
                                                    int* __restrict__ status) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int ns = NS > 0 ? NS : P.ns;
  const double Tc = T[i];
  if (!in_table(P, Tc)) {
    mu[i] = __builtin_nan("");
    lam[i] = __builtin_nan("");
    if (status) status[i] = 2;
    return;
  }
  const int Ti = (int)Tc;
  const double Td = Tc - (double)Ti;
  const double* __restrict__ r0 = P.tab + (long)(Ti - P.Tt_min) * (2 * ns);
  const double* __restrict__ c = rhoi + i;
  const double* __restrict__ W = P.W;
  const double* __restrict__ M0 = P.M;
  const double* __restrict__ M1 = P.M + ns * ns;
  const double* __restrict__ M2 = P.M + 2 * ns * ns;
  const double rc = rho[i];
  double m_acc = 0.0, k_acc = 0.0;
  if constexpr (NS > 0) {
    // the cell's mole fractions, also in LDS for the row loop's runtime index (each thread reads
    // only its own column, so no barrier); the densities are read from HBM once
    __shared__ double sx[NS * 256];
    double u[NS], v[NS], w[NS];
#pragma unroll
    for (int j = 0; j < NS; ++j) u[j] = c[(long)j * ld];
    // Xi = rhoi * Wmtot / (rho * W) with 1 / Wmtot = sum_s rhoi_s / sum(rhoi) / W_s
    double ra = 0.0, rb = 0.0;
#pragma unroll
    for (int j = 0; j < NS; ++j) ra = ra + u[j];
#pragma unroll
    for (int j = 0; j < NS; ++j) rb = rb + u[j] / ra / W[j];
    const double wm = 1.0 / rb;
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      const double X = u[j] * wm / (rc * W[j]);
      sx[j * 256 + threadIdx.x] = X;
      const double2 mk = interp(r0, NS, j, Td);
      const double r = 1.0 / sqrt(mk.x + 1e-20);
      u[j] = X;
      v[j] = X * r;
      w[j] = v[j] * r;
    }
    // rows four at a time: a runtime loop keeps the scalar-loaded matrix rows from all being
    // hoisted into registers, and four rows keep enough loads in flight (0.37 ms one row at a
    // time, 0.30 ms two, 0.28 ms four, per 2^20 24-species cells); u, v, w are indexed by the
    // unrolled j only
#pragma unroll 4
    for (int a = 0; a < NS; ++a) {
      double A = 0.0, B = 0.0, C = 0.0;
#pragma unroll
      for (int j = 0; j < NS; ++j) {
        A = fma(M0[a * NS + j], u[j], A);
        B = fma(M1[a * NS + j], v[j], B);
        C = fma(M2[a * NS + j], w[j], C);
      }
      const double2 mk = interp(r0, NS, a, Td);
      const double s = sqrt(mk.x);
      const double inv = 1.0 / fma(s, fma(s, C, B), A);
      const double Xa = sx[a * 256 + threadIdx.x];
      m_acc = fma(Xa * mk.x, inv, m_acc);
      k_acc = fma(Xa * mk.y, inv, k_acc);
    }
  } else {
    const double wm = 1.0 / inv_wmix(c, ld, ns, W);
    for (int a = 0; a < ns; ++a) {
      double A = 0.0, B = 0.0, C = 0.0;
      for (int j = 0; j < ns; ++j) {
        const double X = c[(long)j * ld] * wm / (rc * W[j]);
        const double r = 1.0 / sqrt(interp(r0, ns, j, Td).x + 1e-20);
        A = fma(M0[a * ns + j], X, A);
        B = fma(M1[a * ns + j], X * r, B);
        C = fma(M2[a * ns + j], X * r * r, C);
      }
      const double2 mk = interp(r0, ns, a, Td);
      const double s = sqrt(mk.x);
      const double inv = 1.0 / fma(s, fma(s, C, B), A);
      const double Xa = c[(long)a * ld] * wm / (rc * W[a]);
      m_acc = fma(Xa * mk.x, inv, m_acc);
      k_acc = fma(Xa * mk.y, inv, k_acc);
    }
  }
  stream_store(mu + i, m_acc);
  stream_store(lam + i, k_acc);
  if (status) status[i] = 0;
}

}  // namespace flint_tr

hipError_t flint_transport_wilke(const flint_tr::TrProg& P, long n, long ld, const double* rhoi, const double* rho,
                                 const double* T, double* mu, double* lam, int* status, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (P.ns < 1 || P.ns > FLINT_TRANSPORT_MAXNS) return hipErrorInvalidValue;
  const dim3 grid((unsigned)((n + 255) / 256)), block(256);
  switch (P.ns) {
    case 5:
      hipLaunchKernelGGL(flint_tr::wilke_kernel<5>, grid, block, 0, s, P, n, ld, rhoi, rho, T, mu, lam, status);
      break;
    case 24:
      hipLaunchKernelGGL(flint_tr::wilke_kernel<24>, grid, block, 0, s, P, n, ld, rhoi, rho, T, mu, lam, status);
      break;
    case 31:
      hipLaunchKernelGGL(flint_tr::wilke_kernel<31>, grid, block, 0, s, P, n, ld, rhoi, rho, T, mu, lam, status);
      break;
    default:
      hipLaunchKernelGGL(flint_tr::wilke_kernel<0>, grid, block, 0, s, P, n, ld, rhoi, rho, T, mu, lam, status);
  }
  return hipGetLastError();
}
