// Host side of the batched equilibrium solver: per-handle device data (CEA_initialize_global's
// part, Lib_CEA_setup.f90:6-99) and the C-ABI entry points of include/flint_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/flint_amd.h"
#include "flint_cea.h"
#include "flint_internal.h"
#include "flint_mlp.h"
#include "flint_warm.h"

struct flint_cea_data {
  flint_cea::CeaProg P{};
  std::vector<void*> bufs;
  // learned warm start workspace (flint_cea_solve_warm), grown on demand
  long wcap = 0;
  float *F = nullptr, *O = nullptr;
  double *G = nullptr, *scal = nullptr;
  int* st = nullptr;
  int* redo_idx = nullptr;  // [ld + 1]: compacted list of the states to solve again, then its length
  int warm_iters = 6;       // iteration cap of the warm-started pass (flint_cea_warm_options)
  int cus = 0;              // compute units of the device (grid of the fused warm-start kernel)
  bool warm_fused = true;   // features, network and start in one kernel where the network fits
  void* wpk = nullptr;      // the network's weights packed for that kernel ...
  unsigned long long pk_uid = 0, pk_version = 0;  // ... from this network at this weight version
  std::vector<void*> wbufs;
};

#define HIPCHK(expr)                                                                                      \
  do {                                                                                                    \
    hipError_t e_ = (expr);                                                                               \
    if (e_ != hipSuccess) return flint_fail(FLINT_EHIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

void flint_mech_release_cea(flint_mech* m) {
  if (!m->cea) return;
  for (void* p : m->cea->bufs) (void)hipFree(p);
  for (void* p : m->cea->wbufs) (void)hipFree(p);
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
  P.max_iter = 50;
  P.tol = 0.0;
  return FLINT_OK;
}

extern "C" int flint_cea_set_tolerance(flint_mech* m, double tol) {
  if (!m) return flint_fail(FLINT_EARG, "null mechanism");
  if (!m->cea) return flint_fail(FLINT_EARG, "flint_cea_setup has not been called for this mechanism");
  if (!(tol >= 0.0) || tol > 1e-4) return flint_fail(FLINT_EARG, "tolerance must be 0 (the reference's tests) or in (0, 1e-4]");
  m->cea->P.tol = tol;
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

template <class T>
static int wgrab(flint_cea_data* c, T** p, size_t count) {
  void* q = nullptr;
  HIPCHK(hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)));
  c->wbufs.push_back(q);
  *p = static_cast<T*>(q);
  return FLINT_OK;
}

extern "C" int flint_cea_solve_warm(flint_mech* m, flint_mlp* net, long n, long ld, const double* p, const double* T,
                                    const double* rho, const double* scaling, double* T_out, double* Y_out,
                                    int* iters, int* status, void* stream) {
  if (!m || !net || !scaling) return flint_fail(FLINT_EARG, "null argument");
  if (!m->cea) return flint_fail(FLINT_EARG, "flint_cea_setup has not been called for this mechanism");
  if (n < 0 || ld < n) return flint_fail(FLINT_EARG, "need 0 <= n <= ld");
  if (n == 0) return FLINT_OK;
  if (!p || !T || !rho || !T_out || !Y_out) return flint_fail(FLINT_EARG, "null array");
  const int ns = m->ns;
  if (net->dims.front() != ns + 2 || net->dims.back() != ns + 1)
    return flint_fail(FLINT_EARG, "network widths must be ns+2 inputs and ns+1 outputs");
  flint_cea_data* c = m->cea;
  if (c->wcap < ld) {
    for (void* q : c->wbufs) (void)hipFree(q);
    c->wbufs.clear();
    c->wcap = 0;
    int rc = 0;
    rc |= wgrab(c, &c->F, (size_t)ld * (ns + 2));
    rc |= wgrab(c, &c->O, (size_t)ld * (ns + 1));
    rc |= wgrab(c, &c->G, (size_t)ld * (ns + 1));
    rc |= wgrab(c, &c->scal, (size_t)4 * ns + 6);
    rc |= wgrab(c, &c->st, (size_t)ld);
    rc |= wgrab(c, &c->redo_idx, (size_t)ld + 1);
    if (rc) return rc;
    c->wcap = ld;
  }
  hipStream_t s = (hipStream_t)stream;
  HIPCHK(hipMemcpyAsync(c->scal, scaling, (size_t)(4 * ns + 6) * sizeof(double), hipMemcpyHostToDevice, s));
  if (c->warm_fused && net->L == 3 && flint_ws_fused_fits(net->dims.data())) {
    // features, the network and the Newton start in one kernel
    if (!c->cus) {
      int dev = 0;
      HIPCHK(hipGetDevice(&dev));
      HIPCHK(hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, dev));
    }
    if (!c->wpk) {
      HIPCHK(hipMalloc(&c->wpk, flint_ws_packed_bytes()));
      c->bufs.push_back(c->wpk);
    }
    if (c->pk_uid != net->uid || c->pk_version != net->version) {
      if (int rc = net->refresh_wt(s)) return rc;
      const __bf16* WT[3] = {net->w16t + net->w16t_off[0], net->w16t + net->w16t_off[1],
                             net->w16t + net->w16t_off[2]};
      HIPCHK(flint_ws_pack(net->dims.data(), WT, c->wpk, s));
      c->pk_uid = net->uid;
      c->pk_version = net->version;
    }
    const float* b[3] = {net->w32 + net->b_off[0], net->w32 + net->b_off[1], net->w32 + net->b_off[2]};
    HIPCHK(flint_ws_fused(n, ld, ns, rho, T, p, c->P.tabh, m->Th_min, m->Th_max, c->scal, net->dims.data(), c->wpk, b,
                          c->G, c->cus, s));
  } else {
    HIPCHK(flint_ws_features(n, ld, ns, rho, T, p, c->P.tabh, m->Th_min, m->Th_max, c->scal, c->F, s));
    if (int rc = flint_mlp_forward(net, n, c->F, c->O, s)) return rc;
    HIPCHK(flint_ws_guess(n, ld, ns, c->O, c->scal, c->G, s));
  }
  int* st = status ? status : c->st;
  // the warm-started pass, with the iteration cap set by flint_cea_warm_options
  flint_cea::CeaProg Wp = c->P;
  Wp.max_iter = c->warm_iters;
  hipError_t e = m->ops->cea(Wp, true, n, ld, p, T, rho, c->G, T_out, Y_out, iters, st, s);
  if (e != hipSuccess) return flint_fail(FLINT_EHIP, std::string("equilibrium launch: ") + hipGetErrorString(e));
  // the states that did not converge, compacted, again from the default start with the full cap
  int* count = c->redo_idx + ld;
  HIPCHK(flint_ws_collect(n, st, c->redo_idx, count, s));
  flint_cea::CeaProg R = c->P;
  R.redo = 1;
  R.idx = c->redo_idx;
  R.count = count;
  e = m->ops->cea(R, true, n, ld, p, T, rho, nullptr, T_out, Y_out, iters, st, s);
  if (e != hipSuccess) return flint_fail(FLINT_EHIP, std::string("equilibrium launch: ") + hipGetErrorString(e));
  return FLINT_OK;
}

extern "C" int flint_cea_warm_options(flint_mech* m, int max_iter, int fused) {
  if (!m) return flint_fail(FLINT_EARG, "null mechanism");
  if (!m->cea) return flint_fail(FLINT_EARG, "flint_cea_setup has not been called for this mechanism");
  if (max_iter < 1 || max_iter > 50) return flint_fail(FLINT_EARG, "the warm-pass iteration cap must be in 1..50");
  if (fused != 0 && fused != 1) return flint_fail(FLINT_EARG, "fused must be 0 or 1");
  m->cea->warm_iters = max_iter;
  m->cea->warm_fused = fused == 1;
  return FLINT_OK;
}
