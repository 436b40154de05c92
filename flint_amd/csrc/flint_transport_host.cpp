// C ABI of the mixture-transport path: species tables in, mixture viscosity and conductivity out.
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/flint_amd.h"
#include "flint_internal.h"
#include "flint_transport.h"

struct flint_transport_data {
  int Tt_min = 0, Tt_max = 0;
  double* d_tab = nullptr;  // [T][ns][2] (mu, k)
  double* d_M = nullptr;    // [3][ns][ns]
};

#define HIPCHK(expr)                                                                                  \
  do {                                                                                                \
    hipError_t e_ = (expr);                                                                           \
    if (e_ != hipSuccess) return flint_fail(FLINT_EHIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

void flint_mech_release_transport(flint_mech* m) {
  if (!m || !m->tr) return;
  if (m->tr->d_tab) (void)hipFree(m->tr->d_tab);
  if (m->tr->d_M) (void)hipFree(m->tr->d_M);
  delete m->tr;
  m->tr = nullptr;
}

extern "C" int flint_transport_setup(flint_mech* m, int Tt_min, int Tt_max, const double* mu_tab,
                                     const double* k_tab) {
  if (!m) return flint_fail(FLINT_EARG, "null mechanism");
  if (!mu_tab || !k_tab) return flint_fail(FLINT_EARG, "null table");
  if (Tt_max <= Tt_min) return flint_fail(FLINT_EARG, "need Tt_max > Tt_min");
  if (m->ns > FLINT_TRANSPORT_MAXNS) return flint_fail(FLINT_EUNSUPPORTED, "more species than the kernel holds");
  const int ns = m->ns;
  const size_t nT = (size_t)(Tt_max - Tt_min + 1);
  // row-major (mu, k) pairs so that one cell's interpolation reads two contiguous rows
  std::vector<double> tab(nT * ns * 2);
  for (size_t t = 0; t < nT; ++t)
    for (int s = 0; s < ns; ++s) {
      tab[(t * ns + s) * 2 + 0] = mu_tab[(size_t)s * nT + t];
      tab[(t * ns + s) * 2 + 1] = k_tab[(size_t)s * nT + t];
    }
  // the molecular-weight factors of co_fiij (Lib_ThermoTransport.f90:357-358), per species pair
  std::vector<double> M((size_t)3 * ns * ns);
  for (int i = 0; i < ns; ++i)
    for (int j = 0; j < ns; ++j) {
      const double q = m->W[i] / m->W[j];
      const double D = 1.0 / std::sqrt(8 * (1 + q));
      const double c = 1.0 / std::pow(q, 0.25);
      M[(size_t)i * ns + j] = D;
      M[(size_t)ns * ns + (size_t)i * ns + j] = 2.0 * D * c;
      M[(size_t)2 * ns * ns + (size_t)i * ns + j] = D * c * c;
    }
  std::lock_guard<std::mutex> g(m->mu);
  flint_mech_release_transport(m);
  m->tr = new flint_transport_data;
  m->tr->Tt_min = Tt_min;
  m->tr->Tt_max = Tt_max;
  HIPCHK(hipMalloc(&m->tr->d_tab, tab.size() * sizeof(double)));
  HIPCHK(hipMalloc(&m->tr->d_M, M.size() * sizeof(double)));
  HIPCHK(hipMemcpy(m->tr->d_tab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(m->tr->d_M, M.data(), M.size() * sizeof(double), hipMemcpyHostToDevice));
  return FLINT_OK;
}

extern "C" int flint_transport(const flint_mech* m, long n, long ld, const double* rhoi, const double* rho,
                               const double* T, double* mu, double* lam, int* status, void* stream) {
  if (!m) return flint_fail(FLINT_EARG, "null mechanism");
  if (n < 0 || ld < n) return flint_fail(FLINT_EARG, "need 0 <= n <= ld");
  if (!m->tr) return flint_fail(FLINT_EUNSUPPORTED, "no transport tables: call flint_transport_setup first");
  if (n == 0) return FLINT_OK;
  if (!rhoi || !rho || !T || !mu || !lam) return flint_fail(FLINT_EARG, "null array");
  const flint_tr::TrProg P{m->tr->d_tab, m->tr->d_M, m->d_W, m->tr->Tt_min, m->tr->Tt_max, m->ns};
  hipError_t e = flint_transport_wilke(P, n, ld, rhoi, rho, T, mu, lam, status, (hipStream_t)stream);
  if (e != hipSuccess) return flint_fail(FLINT_EHIP, hipGetErrorString(e));
  return FLINT_OK;
}

extern "C" int flint_transport_host(flint_mech* m, long n, const double* rhoi, const double* rho, const double* T,
                                    double* mu, double* lam, int* status) {
  if (!m) return flint_fail(FLINT_EARG, "null mechanism");
  if (n < 0) return flint_fail(FLINT_EARG, "need n >= 0");
  if (!m->tr) return flint_fail(FLINT_EUNSUPPORTED, "no transport tables: call flint_transport_setup first");
  if (n == 0) return FLINT_OK;
  if (!rhoi || !rho || !T || !mu || !lam) return flint_fail(FLINT_EARG, "null array");
  std::lock_guard<std::mutex> g(m->mu);
  const size_t ns = m->ns;
  if (int rc = flint_stage(m, (ns + 5) * (size_t)n)) return rc;
  double* d_r = m->d_stage;
  double* d_rho = d_r + ns * n;
  double* d_T = d_rho + n;
  double* d_mu = d_T + n;
  double* d_k = d_mu + n;
  int* d_st = reinterpret_cast<int*>(d_k + n);
  HIPCHK(hipMemcpyAsync(d_r, rhoi, ns * n * sizeof(double), hipMemcpyHostToDevice, m->stream));
  HIPCHK(hipMemcpyAsync(d_rho, rho, n * sizeof(double), hipMemcpyHostToDevice, m->stream));
  HIPCHK(hipMemcpyAsync(d_T, T, n * sizeof(double), hipMemcpyHostToDevice, m->stream));
  if (int rc = flint_transport(m, n, n, d_r, d_rho, d_T, d_mu, d_k, d_st, m->stream)) return rc;
  HIPCHK(hipMemcpyAsync(mu, d_mu, n * sizeof(double), hipMemcpyDeviceToHost, m->stream));
  HIPCHK(hipMemcpyAsync(lam, d_k, n * sizeof(double), hipMemcpyDeviceToHost, m->stream));
  if (status) HIPCHK(hipMemcpyAsync(status, d_st, n * sizeof(int), hipMemcpyDeviceToHost, m->stream));
  HIPCHK(hipStreamSynchronize(m->stream));
  return FLINT_OK;
}
