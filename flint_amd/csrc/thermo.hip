// Conservative <-> primitive variables of a CFD cell: the batched cons2prim (the Newton solve for
// T from the internal energy that precedes every source-term call in a CFD step) and prim2cons,
// one thread per cell.
//
// cons2prim and prim2cons follow Lib_ThermoTransport.f90:130-185 with the helpers they call:
// f_Rtot (:217-230), f_cp (:251-271), H / E / E0 (:40-74), f_tab (:505-514) and f_tabT (:516-542).
// The arithmetic keeps the reference's operation order with contraction off, including the
// temperature f_tab recomputes from the pressure, T' = p / (R rho), for the enthalpy.  Layout:
// cons[k*ld + i], k = (rho_1..rho_ns, rho u_1..rho u_nd, rho E); prim[k*ld + i] =
// (rho_1..rho_ns, u_1..u_nd, p).  The table lookups are gathers of one row pair per cell and
// species.  Up to 32 species cons2prim keeps the cell's densities in registers for the Newton
// loop's sums; beyond that it rereads them and runs for any species count.
#include "flint_device.h"
#include "flint_thermo.h"

namespace flint_thermo_k {

constexpr double RUNIV = 8314.51;  // Runiv, Lib_ThermoTransport.f90:13 (Ri_tab = Runiv/wm_tab)

struct Tab {
  const double* tabh;  // [T][ns][2] (cp, h) per mass, rows Th_min..Th_max
  int Th_min, Th_max, ns;
};

// f_tabT of column c (0: cp, 1: h) of species s; false outside [Th_min, Th_max)
__device__ __forceinline__ bool tabT(const Tab& t, double T, int s, int c, double& v) {
#pragma clang fp contract(off)
  if (!(T >= (double)t.Th_min && T < (double)t.Th_max)) return false;
  const int Ti = (int)T;
  const double Td = T - (double)Ti;
  const double* __restrict__ r0 = t.tabh + (long)(Ti - t.Th_min) * (2 * t.ns) + 2 * s + c;
  const double a = r0[0], b = r0[2 * t.ns];
  v = a + (b - a) * Td;
  return true;
}

// sum(r), f_Rtot(r) and the sum of squares of the velocity of cell i
__device__ __forceinline__ void rho_R(const double* __restrict__ sRi, int ns, const double* __restrict__ c, long ld,
                                      double& rho, double& R) {
#pragma clang fp contract(off)
  double a = 0.0;
  for (int s = 0; s < ns; ++s) a = a + c[(long)s * ld];
  double b = 0.0;
  for (int s = 0; s < ns; ++s) b = b + sRi[s] * c[(long)s * ld] / a;
  rho = a;
  R = b;
}

// NORM2 as the reference build (flang) evaluates it: m sqrt(1 + sum over the others of (x/m)^2),
// the running maximum m rescaling the sum when it grows; x(d) = v[d * ld] * scale
__device__ __forceinline__ double norm2(const double* __restrict__ v, long ld, int nd, double rho, bool divide) {
#pragma clang fp contract(off)
  double mx = 0.0, s = 0.0;
  for (int d = 0; d < nd; ++d) {
    const double x = divide ? v[(long)d * ld] / rho : v[(long)d * ld];
    const double ax = fabs(x);
    if (mx == 0.0) {
      mx = ax;
    } else if (ax > mx) {
      const double t = mx / ax, tsq = t * t;
      s = s * tsq;
      s = s + tsq;
      mx = ax;
    } else {
      const double t = ax / mx;
      s = s + t * t;
    }
  }
  return mx * sqrt(1.0 + s);
}

// E(p, r) = H(p, r) - p / sum(r): enthalpy from h_tab at T' = p / (R rho) (f_tab)
__device__ __forceinline__ bool energy_E(const Tab& t, const double* __restrict__ c, long ld, double rho, double R,
                                         double p, double& e) {
#pragma clang fp contract(off)
  const double T = p / (R * rho);
  double ent = 0.0;
  for (int s = 0; s < t.ns; ++s) {
    double h;
    if (!tabT(t, T, s, 1, h)) return false;
    ent = ent + h * c[(long)s * ld] / rho;
  }
  e = ent - p / rho;
  return true;
}

// The species loops of cons2prim over the cell's densities: NSM > 0 holds them in registers (read
// once from memory with non-temporal loads, so that the state stream does not evict the table rows
// the Newton loop gathers), every loop unrolled over NSM with static indices; NSM = 0 rereads them
// from memory and serves any species count.
template <int NSM>
struct Dens {
  double r[NSM > 0 ? NSM : 1];
  const double* __restrict__ c;
  long ld;
  __device__ __forceinline__ double operator()(int s) const { return NSM > 0 ? r[s] : c[(long)s * ld]; }
};

template <int NSM, class F>
__device__ __forceinline__ void for_species(int ns, F&& f) {
  if constexpr (NSM > 0) {
#pragma unroll
    for (int s = 0; s < NSM; ++s)
      if (s < ns) f(s);
  } else {
    for (int s = 0; s < ns; ++s) f(s);
  }
}

template <int NSM>
__global__ __launch_bounds__(256) void cons2prim_kernel(Tab t, const double* __restrict__ W, long n, long ld, int nd,
                                                        const double* __restrict__ cons, const double* __restrict__ Tin,
                                                        double* __restrict__ prim, double* __restrict__ Tout,
                                                        int* __restrict__ iters, int* __restrict__ status) {
#pragma clang fp contract(off)
  __shared__ double sRi[FLINT_THERMO_MAXNS];
  for (int s = threadIdx.x; s < t.ns; s += blockDim.x) sRi[s] = RUNIV / W[s];
  __syncthreads();
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int ns = t.ns, np = ns + nd + 1;
  const double* __restrict__ c = cons + i;
  Dens<NSM> den;
  den.c = c;
  den.ld = ld;
  if constexpr (NSM > 0) {
#pragma unroll
    for (int s = 0; s < NSM; ++s) den.r[s] = s < ns ? __builtin_nontemporal_load(c + (long)s * ld) : 0.0;
  }
  // sum(r) and f_Rtot(r), in the order of rho_R
  double rho = 0.0, Rgas = 0.0;
  for_species<NSM>(ns, [&](int s) { rho = rho + den(s); });
  for_species<NSM>(ns, [&](int s) { Rgas = Rgas + sRi[s] * den(s) / rho; });
  // outputs bypass L2 (streaming stores), which keeps the table rows the Newton loop gathers there
  for_species<NSM>(ns, [&](int s) { stream_store(prim + (long)s * ld + i, den(s)); });
  for (int d = 0; d < nd; ++d) stream_store(prim + (long)(ns + d) * ld + i, c[(long)(ns + d) * ld] / rho);
  const double vel = norm2(c + (long)ns * ld, ld, nd, rho, true);
  const double energy = c[(long)(np - 1) * ld] / rho - 0.5 * vel * vel;
  // Newton on T (Lib_ThermoTransport.f90:153-161).  f_cp reads cp at T and E reads h at
  // T' = p / (R rho), which falls in the same table row as T (T' = T to an ulp) but for rare
  // cells at a row boundary: one 16-byte (cp, h) load per row serves both sums then.
  double diff = 1.0e3, T = Tin[i];
  int it = 1, st = 0;
  while (fabs(diff / energy) >= 1e-6 && it <= 20) {
    const double p_it = rho * Rgas * T;
    const double T2 = p_it / (Rgas * rho);
    if (!(T >= (double)t.Th_min && T < (double)t.Th_max) || !(T2 >= (double)t.Th_min && T2 < (double)t.Th_max)) {
      st = 2;
      break;
    }
    const int Ti = (int)T, Ti2 = (int)T2;
    const double Td = T - (double)Ti, Td2 = T2 - (double)Ti2;
    const double* __restrict__ r0 = t.tabh + (long)(Ti - t.Th_min) * (2 * ns);
    const double* __restrict__ r2 = t.tabh + (long)(Ti2 - t.Th_min) * (2 * ns);
    double cp = 0.0, ent = 0.0;
    if (Ti2 == Ti) {
      for_species<NSM>(ns, [&](int s) {
        const double2 a = ld2(r0 + 2 * s), b = ld2(r0 + 2 * ns + 2 * s);
        const double r = den(s);
        cp = cp + r / rho * (a.x + (b.x - a.x) * Td);
        ent = ent + (a.y + (b.y - a.y) * Td2) * r / rho;
      });
    } else {
      for_species<NSM>(ns, [&](int s) {
        const double2 a = ld2(r0 + 2 * s), b = ld2(r0 + 2 * ns + 2 * s);
        const double2 a2 = ld2(r2 + 2 * s), b2 = ld2(r2 + 2 * ns + 2 * s);
        const double r = den(s);
        cp = cp + r / rho * (a.x + (b.x - a.x) * Td);
        ent = ent + (a2.y + (b2.y - a2.y) * Td2) * r / rho;
      });
    }
    const double e_it = ent - p_it / rho;
    const double cv = cp - Rgas;
    diff = fabs(e_it - energy);
    T = T + (energy - e_it) / cv;
    it = it + 1;
  }
  stream_store(prim + (long)(np - 1) * ld + i, st ? __builtin_nan("") : rho * Rgas * T);
  if (Tout) stream_store(Tout + i, st ? __builtin_nan("") : T);
  if (iters) iters[i] = it - 1;
  if (status) status[i] = st;
}

__global__ __launch_bounds__(256) void prim2cons_kernel(Tab t, const double* __restrict__ W, long n, long ld, int nd,
                                                        const double* __restrict__ prim, double* __restrict__ cons,
                                                        int* __restrict__ status) {
#pragma clang fp contract(off)
  __shared__ double sRi[FLINT_THERMO_MAXNS];
  for (int s = threadIdx.x; s < t.ns; s += blockDim.x) sRi[s] = RUNIV / W[s];
  __syncthreads();
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int ns = t.ns, np = ns + nd + 1;
  const double* __restrict__ q = prim + i;
  double rho, Rgas;
  rho_R(sRi, ns, q, ld, rho, Rgas);
  const double vel = norm2(q + (long)ns * ld, ld, nd, rho, false);
  for (int s = 0; s < ns; ++s) cons[(long)s * ld + i] = q[(long)s * ld];
  for (int d = 0; d < nd; ++d) cons[(long)(ns + d) * ld + i] = rho * q[(long)(ns + d) * ld];
  double e;
  const bool ok = energy_E(t, q, ld, rho, Rgas, q[(long)(np - 1) * ld], e);
  cons[(long)(np - 1) * ld + i] = ok ? rho * (e + 0.5 * vel * vel) : __builtin_nan("");
  if (status) status[i] = ok ? 0 : 2;
}

}  // namespace flint_thermo_k

hipError_t flint_thermo_cons2prim(const double* tabh, int Th_min, int Th_max, int ns, const double* W, long n,
                                  long ld, int nd, const double* cons, const double* T, double* prim, double* T_out,
                                  int* iters, int* status, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (ns < 1 || ns > FLINT_THERMO_MAXNS) return hipErrorInvalidValue;
  const flint_thermo_k::Tab t{tabh, Th_min, Th_max, ns};
if (t.ns <= 16)
    hipLaunchKernelGGL(flint_thermo_k::cons2prim_kernel<16>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, t, W,
                       n, ld, nd, cons, T, prim, T_out, iters, status);
  else if (t.ns <= 32)
    hipLaunchKernelGGL(flint_thermo_k::cons2prim_kernel<32>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, t, W,
                       n, ld, nd, cons, T, prim, T_out, iters, status);
  else
    hipLaunchKernelGGL(flint_thermo_k::cons2prim_kernel<0>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, t, W,
                       n, ld, nd, cons, T, prim, T_out, iters, status);
  return hipGetLastError();
}

hipError_t flint_thermo_prim2cons(const double* tabh, int Th_min, int Th_max, int ns, const double* W, long n,
                                  long ld, int nd, const double* prim, double* cons, int* status, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (ns < 1 || ns > FLINT_THERMO_MAXNS) return hipErrorInvalidValue;
  const flint_thermo_k::Tab t{tabh, Th_min, Th_max, ns};
  hipLaunchKernelGGL(flint_thermo_k::prim2cons_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, t, W, n, ld,
                     nd, prim, cons, status);
  return hipGetLastError();
}
