// Host side of the batched equilibrium solver: per-handle device data (CEA_initialize_global's
// part, Lib_CEA_setup.f90:6-99) and the C-ABI entry points of include/flint_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/flint_amd.h"
#include "flint_cea.h"
#include "flint_internal.h"

struct flint_cea_data {
  flint_cea::CeaProg P{};
  std::vector<void*> bufs;
};

#define HIPCHK(expr)                                                                                      \
  do {                                                                                                    \
    hipError_t e_ = (expr);                                                                               \
    if (e_ != hipSuccess) return flint_fail(FLINT_EHIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

void flint_mech_release_cea(flint_mech* m) {
  if (!m->cea) return;
  for (void* p : m->cea->bufs) (void)hipFree(p);
  delete m->cea;
  m->cea = nullptr;
}

static int upload(flint_cea_data* c, const std::vector<double>& v, const double** dst) {
  void* p = nullptr;
  HIPCHK(hipMalloc(&p, std::max<size_t>(v.size(), 1) * sizeof(double)));
  c->bufs.push_back(p);
  HIPCHK(hipMemcpy(p, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
  *dst = static_cast<const double*>(p);
  return FLINT_OK;
}

extern "C" int flint_cea_setup(flint_mech* m, const double* s, const double* el_weight) {
  if (!m || !s || !el_weight) return flint_fail(FLINT_EARG, "null argument");
  if (m->ne <= 0 || m->comp.empty()) return flint_fail(FLINT_EARG, "the mechanism has no species composition");
  const int ns = m->ns, ne = m->ne;
  if (!m->ops || !m->ops->cea) return flint_fail(FLINT_EUNSUPPORTED, "no compiled equilibrium kernel for this mechanism");
  if (m->ops->cea_ne != ne) return flint_fail(FLINT_EARG, "element count differs from the compiled mechanism");
  for (int j = 0; j < ns; ++j)
    for (int e = 0; e < ne; ++e)
      if (m->comp[(size_t)j * ne + e] != m->ops->cea_A[(size_t)e * ns + j])
        return flint_fail(FLINT_EARG, "species composition differs from the compiled mechanism");
  const int nTh = m->Th_max - m->Th_min + 1;
  for (int e = 0; e < ne; ++e)
    if (!(el_weight[e] > 0.0)) return flint_fail(FLINT_EARG, "atomic weights must be positive");
  flint_mech_release_cea(m);
  auto* c = new flint_cea_data();
  m->cea = c;
  std::vector<double> rm(ns, 0.0), tabs((size_t)(nTh + 1) * ns);
  // species weight from the elements, in the reference's summation order (Lib_CEA_setup.f90:126-130)
  for (int j = 0; j < ns; ++j) {
    double r = 0.0;
    for (int e = 0; e < ne; ++e) r = r + m->comp[(size_t)j * ne + e] * el_weight[e];
    rm[j] = r;
  }
  for (int t = 0; t < nTh; ++t)
    for (int k = 0; k < ns; ++k) tabs[(size_t)t * ns + k] = s[(size_t)k * nTh + t];
  for (int k = 0; k < ns; ++k) tabs[(size_t)nTh * ns + k] = tabs[(size_t)(nTh - 1) * ns + k];
  flint_cea::CeaProg& P = c->P;
  int rc = 0;
  rc |= upload(c, rm, &P.rm);
  rc |= upload(c, tabs, &P.tabs);
  if (rc) {
    flint_mech_release_cea(m);
    return rc;
  }
  P.tabh = m->d_tabh;
  P.W = m->d_W;
  P.Th_min = m->Th_min;
  P.Th_max = m->Th_max;
  return FLINT_OK;
}

extern "C" int flint_cea_solve(flint_mech* m, long n, long ld, int hp, const double* p, const double* T,
                               const double* rho, const double* guess, double* T_out, double* Y_out, int* iters,
                               int* status, void* stream) {
  if (!m) return flint_fail(FLINT_EARG, "null mechanism");
  if (!m->cea) return flint_fail(FLINT_EARG, "flint_cea_setup has not been called for this mechanism");
  if (n < 0 || ld < n) return flint_fail(FLINT_EARG, "need 0 <= n <= ld");
  if (hp != 0 && hp != 1) return flint_fail(FLINT_EARG, "hp must be 0 (UV) or 1 (HP)");
  if (n == 0) return FLINT_OK;
  if (!T || !rho || !T_out || !Y_out || (hp && !p)) return flint_fail(FLINT_EARG, "null array");
  hipError_t e = m->ops->cea(m->cea->P, hp == 1, n, ld, p, T, rho, guess, T_out, Y_out, iters, status,
                             (hipStream_t)stream);
  if (e != hipSuccess) return flint_fail(FLINT_EHIP, std::string("equilibrium launch: ") + hipGetErrorString(e));
  return FLINT_OK;
}

extern "C" int flint_cea_solve_host(flint_mech* m, long n, int hp, const double* p, const double* T,
                                    const double* rho, double* T_out, double* Y_out, int* status) {
  if (!m) return flint_fail(FLINT_EARG, "null mechanism");
  if (n < 0) return flint_fail(FLINT_EARG, "need n >= 0");
  if (n == 0) return FLINT_OK;
  if (!T || !rho || !T_out || !Y_out || (hp == 1 && !p)) return flint_fail(FLINT_EARG, "null array");
  std::lock_guard<std::mutex> g(m->mu);
  const size_t ns = m->ns, nn = (size_t)n;
  // p, T, rho, T_out, Y_out, status (ints in a double-sized slot)
  if (int rc = flint_stage(m, (2 * ns + 4) * nn)) return rc;
  double* d_p = m->d_stage;
  double* d_T = d_p + nn;
  double* d_rho = d_T + nn;
  double* d_To = d_rho + ns * nn;
  double* d_Yo = d_To + nn;
  int* d_st = reinterpret_cast<int*>(d_Yo + ns * nn);
  if (hp == 1) HIPCHK(hipMemcpyAsync(d_p, p, nn * sizeof(double), hipMemcpyHostToDevice, m->stream));
  HIPCHK(hipMemcpyAsync(d_T, T, nn * sizeof(double), hipMemcpyHostToDevice, m->stream));
  HIPCHK(hipMemcpyAsync(d_rho, rho, ns * nn * sizeof(double), hipMemcpyHostToDevice, m->stream));
  if (int rc = flint_cea_solve(m, n, n, hp, hp == 1 ? d_p : nullptr, d_T, d_rho, nullptr, d_To, d_Yo, nullptr, d_st,
                               m->stream))
    return rc;
  HIPCHK(hipMemcpyAsync(T_out, d_To, nn * sizeof(double), hipMemcpyDeviceToHost, m->stream));
  HIPCHK(hipMemcpyAsync(Y_out, d_Yo, ns * nn * sizeof(double), hipMemcpyDeviceToHost, m->stream));
  if (status) HIPCHK(hipMemcpyAsync(status, d_st, nn * sizeof(int), hipMemcpyDeviceToHost, m->stream));
  HIPCHK(hipStreamSynchronize(m->stream));
  return FLINT_OK;
}
