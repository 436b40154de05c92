// Host side of the learned reactor step: buffers, the MLP call, compaction of the rejected cells
// and the plain BDF on them (C-ABI entry point of include/flint_amd.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/flint_amd.h"
#include "flint_internal.h"
#include "flint_mlp.h"
#include "flint_reactor.h"
#include "flint_surrogate.h"

struct flint_sur_data {
  long cap = 0;
  float *F = nullptr, *O = nullptr;
  int *flag = nullptr, *idx = nullptr, *count = nullptr, *stg = nullptr, *sg = nullptr, *ord = nullptr;
  double *Xg = nullptr, *pg = nullptr, *scal = nullptr, *comp = nullptr, *elw = nullptr, *Fr = nullptr, *Xp = nullptr;
  std::vector<void*> all;
};

#define HIPCHK(expr)                                                                                      \
  do {                                                                                                    \
    hipError_t e_ = (expr);                                                                               \
    if (e_ != hipSuccess) return flint_fail(FLINT_EHIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

void flint_mech_release_sur(flint_mech* m) {
  if (!m->sur) return;
  for (void* p : m->sur->all) (void)hipFree(p);
  delete m->sur;
  m->sur = nullptr;
}

template <class T>
static int grab(flint_sur_data* d, T** p, size_t count) {
  void* q = nullptr;
  HIPCHK(hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)));
  d->all.push_back(q);
  *p = static_cast<T*>(q);
  return FLINT_OK;
}

static int ensure(flint_mech* m, long n) {
  if (m->sur && m->sur->cap >= n) return FLINT_OK;
  // n here is the leading dimension of the caller's state (>= the cell count)
  flint_mech_release_sur(m);
  auto* d = new flint_sur_data();
  m->sur = d;
  const int ns = m->ns, ne = m->ne;
  int rc = 0;
  rc |= grab(d, &d->F, (size_t)n * (2 * ns + 3));
  rc |= grab(d, &d->O, (size_t)n * (ns + 1));
  rc |= grab(d, &d->flag, n);
  rc |= grab(d, &d->idx, n);
  rc |= grab(d, &d->ord, (size_t)n + FLINT_ORDER_BUCKETS);  // fallback work order, then class counts
  rc |= grab(d, &d->count, 64);
  rc |= grab(d, &d->stg, (size_t)8 * n);
  rc |= grab(d, &d->sg, n);
  rc |= grab(d, &d->Xg, (size_t)(ns + 1) * n);
  rc |= grab(d, &d->pg, n);
  rc |= grab(d, &d->scal, (size_t)2 * (2 * ns + 3) + 2 * (ns + 1));
  rc |= grab(d, &d->comp, (size_t)ne * ns);
  rc |= grab(d, &d->elw, ne);
  rc |= grab(d, &d->Fr, (size_t)(ns + 1) * n);
  rc |= grab(d, &d->Xp, (size_t)(ns + 1) * n);
  if (rc) {
    flint_mech_release_sur(m);
    return rc;
  }
  std::vector<double> comp((size_t)ne * ns);
  for (int e = 0; e < ne; ++e)
    for (int k = 0; k < ns; ++k) comp[(size_t)e * ns + k] = m->comp[(size_t)k * ne + e];
  HIPCHK(hipMemcpy(d->comp, comp.data(), comp.size() * sizeof(double), hipMemcpyHostToDevice));
  d->cap = n;
  return FLINT_OK;
}

extern "C" int flint_reactor_advance_mlp(flint_mech* m, flint_mlp* net, long n, long ld, double* X, const double* p,
                                         double t_end, double rtol, double atol, const double* scaling,
                                         const double* el_weight, double ymin, double yfloor, double tfloor,
                                         double max_dT, double max_residual, double max_T_mismatch,
                                         double max_newton_dY, double max_newton_dT, int* fallback, long* n_fallback,
                                         double* check, int* stats, int* status, void* stream) {
  if (!m || !net || !scaling || !el_weight) return flint_fail(FLINT_EARG, "null argument");
  if (n < 0 || ld < n) return flint_fail(FLINT_EARG, "need 0 <= n <= ld");
  if (n_fallback) *n_fallback = 0;
  if (n == 0) return FLINT_OK;
  if (!X || !p) return flint_fail(FLINT_EARG, "null state or pressure");
  const int ns = m->ns, ne = m->ne;
  if (ns != 24 && ns != 31) return flint_fail(FLINT_EUNSUPPORTED, "the learned step supports the generated mechanisms");
  if (ne < 1 || ne > 4 || m->comp.empty()) return flint_fail(FLINT_EARG, "the learned step needs 1..4 elements");
  const int nin = net->dims.front();
  if ((nin != ns + 2 && nin != 2 * ns + 3) || net->dims.back() != ns + 1)
    return flint_fail(FLINT_EARG, "network widths must be ns+2 or 2ns+3 inputs and ns+1 outputs");
  if (!(t_end > 0.0) || !(rtol > 0.0) || !(atol > 0.0) || !(ymin > 0.0) || !(yfloor >= 0.0) || !(tfloor >= 0.0) ||
      !(max_dT >= 0.0) || !(max_residual >= 0.0) || !(max_T_mismatch >= 0.0) || !(max_newton_dY >= 0.0) ||
      !(max_newton_dT >= 0.0))
    return flint_fail(FLINT_EARG, "need t_end, rtol, atol, ymin > 0 and non-negative floors and thresholds");
  if (int rc = ensure(m, ld)) return rc;
  flint_sur_data* d = m->sur;
  hipStream_t s = (hipStream_t)stream;
  HIPCHK(hipMemcpyAsync(d->scal, scaling, (size_t)(2 * nin + 2 * (ns + 1)) * sizeof(double), hipMemcpyHostToDevice,
                        s));
  HIPCHK(hipMemcpyAsync(d->elw, el_weight, (size_t)ne * sizeof(double), hipMemcpyHostToDevice, s));
  if (int rc = flint_reactor_jacobian(m, n, ld, 1, X, p, d->Fr, nullptr, stream)) return rc;
  flint_sur::SurProg P{ns,    ne,   nin,    d->comp, m->d_W, d->elw,       m->d_tabh,   m->Th_min, m->Th_max, d->scal,
                       d->Fr, ld,   t_end,  ymin,    yfloor, tfloor, max_dT, max_residual, max_T_mismatch};
  HIPCHK(flint_sur_features(P, n, ld, X, p, d->F, s));
  if (int rc = flint_mlp_forward(net, n, d->F, d->O, s)) return rc;
  int* flag = fallback ? fallback : d->flag;
  HIPCHK(flint_sur_apply(P, n, ld, X, d->Xp, d->O, flag, s));
  // the FP64 judge: one implicit-step Newton correction from each prediction
  HIPCHK(flint_sur_zero(d->count, 1, s));
  if (int rc = flint_reactor_verify(m, n, ld, X, d->Xp, ld, d->Fr, ld, p, t_end, max_newton_dY, max_newton_dT, flag,
                                    d->idx, d->count, check, stream))
    return rc;
  int nf = 0;
  HIPCHK(hipMemcpyAsync(&nf, d->count, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (status) HIPCHK(hipMemsetAsync(status, 0, (size_t)n * sizeof(int), s));
  if (stats)
    for (int r = 0; r < 8; ++r) HIPCHK(hipMemsetAsync(stats + (size_t)r * ld, 0, (size_t)n * sizeof(int), s));
  if (nf > 0) {
    // the fallback cells by descending |dT/dt| at the old state: the igniting cells, which take the
    // most BDF steps, start first, so the launch does not wait on one of them started last
    HIPCHK(flint_reactor_order(nf, d->Fr + (size_t)ns * ld, d->idx, d->ord, d->ord + nf, s));
    HIPCHK(flint_sur_gather(0, ns + 1, ld, X, nf, nf, d->Xg, d->ord, p, d->pg, s));
    if (int rc = flint_reactor_advance_as_given(m, nf, nf, 1, d->Xg, d->pg, t_end, rtol, atol, 0,
                                                stats ? d->stg : nullptr, d->sg, stream))
      return rc;
    HIPCHK(flint_sur_gather(1, ns + 1, ld, X, nf, nf, d->Xg, d->ord, nullptr, nullptr, s));
    if (status) HIPCHK(flint_sur_scatter_int(1, n, status, nf, nf, d->sg, d->ord, s));
    if (stats) HIPCHK(flint_sur_scatter_int(8, ld, stats, nf, nf, d->stg, d->ord, s));
  }
  if (n_fallback) *n_fallback = nf;
  return FLINT_OK;
}
