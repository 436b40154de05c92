// Host side of the data-driven general source term (general.hip): the C-ABI entry points
// flint_general_setup and flint_source_general of include/flint_amd.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../include/flint_amd.h"
#include "flint_general.h"
#include "flint_internal.h"

struct flint_general_data {
  flint_general::GenProg P{};
  std::vector<void*> bufs;
};

#define HIPCHK(expr)                                                                                      \
  do {                                                                                                    \
    hipError_t e_ = (expr);                                                                               \
    if (e_ != hipSuccess) return flint_fail(FLINT_EHIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

void flint_mech_release_general(flint_mech* m) {
  if (!m->gen) return;
  for (void* p : m->gen->bufs) (void)hipFree(p);
  delete m->gen;
  m->gen = nullptr;
}

template <class T>
static int put(flint_general_data* g, const std::vector<T>& v, const T** dst) {
  void* p = nullptr;
  HIPCHK(hipMalloc(&p, std::max<size_t>(v.size(), 1) * sizeof(T)));
  g->bufs.push_back(p);
  if (!v.empty()) HIPCHK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  *dst = static_cast<const T*>(p);
  return FLINT_OK;
}

extern "C" int flint_general_setup(flint_mech* m, int na, int nt, int nl, const double* ni1, const double* ni2,
                                   const double* eps) {
  if (!m || !ni1 || !ni2 || !eps) return flint_fail(FLINT_EARG, "null argument");
  if (na != m->n_arrh || nt != m->n_troe || nl != m->n_lind)
    return flint_fail(FLINT_EARG, "reaction counts differ from the mechanism's rate tables");
  const int ns = m->ns, nz = ns + 1, nr = na + nt + nl;
  if (ns > 64) return flint_fail(FLINT_EUNSUPPORTED, "more than 64 species");
  std::vector<int> kind(nr), fp(1, 0), fs, rp(1, 0), rs, mp(1, 0), ms, op(1, 0), os;
  std::vector<double> fe, re, me, ow;
  for (int j = 0; j < nr; ++j) {
    const double* a1 = ni1 + (size_t)j * nz;
    const double* a2 = ni2 + (size_t)j * nz;
    const double* e = eps + (size_t)j * nz;
    const int k = j < na ? 0 : (j < na + nt ? 1 : 2);
    kind[j] = k | (a1[ns] > 0.0 ? 4 : 0);
    for (int s = 0; s < ns; ++s) {
      if (!std::isfinite(a1[s]) || !std::isfinite(a2[s]) || !std::isfinite(e[s]))
        return flint_fail(FLINT_EARG, "non-finite table entry");
      if (std::fabs(a1[s]) > 64.0 || std::fabs(a2[s]) > 64.0)
        return flint_fail(FLINT_EARG, "stoichiometric coefficient beyond 64");
      if (a1[s] != 0.0) {
        fs.push_back(s);
        fe.push_back(k == 0 ? std::round(a1[s]) : a1[s]);  // nint(): halves away from zero (Lib_Chemistry_wdot.f90:140-141)
      }
      if (a2[s] != 0.0) {
        rs.push_back(s);
        re.push_back(k == 0 ? std::round(a2[s]) : a2[s]);
      }
      if (e[s] != 0.0) {
        ms.push_back(s);
        me.push_back(e[s]);
      }
      const double d = a2[s] - a1[s];
      if (d != 0.0) {
        os.push_back(s);
        ow.push_back(m->W[s] * d);
      }
    }
    fp.push_back((int)fs.size());
    rp.push_back((int)rs.size());
    mp.push_back((int)ms.size());
    op.push_back((int)os.size());
  }
  flint_mech_release_general(m);
  auto* g = new flint_general_data();
  m->gen = g;
  flint_general::GenProg& P = g->P;
  P.ns = ns;
  P.nr = nr;
  P.na = na;
  P.nt = nt;
  P.nl = nl;
  P.wk = m->wk;
  int rc = 0;
  rc |= put(g, kind, &P.kind);
  rc |= put(g, fp, &P.f_ptr);
  rc |= put(g, fs, &P.f_sp);
  rc |= put(g, fe, &P.f_exp);
  rc |= put(g, rp, &P.r_ptr);
  rc |= put(g, rs, &P.r_sp);
  rc |= put(g, re, &P.r_exp);
  rc |= put(g, mp, &P.m_ptr);
  rc |= put(g, ms, &P.m_sp);
  rc |= put(g, me, &P.m_eps);
  rc |= put(g, op, &P.o_ptr);
  rc |= put(g, os, &P.o_sp);
  rc |= put(g, ow, &P.o_wd);
  if (rc) {
    flint_mech_release_general(m);
    return rc;
  }
  return FLINT_OK;
}

extern "C" int flint_source_general(const flint_mech* m, long n, long ld, const double* rho, const double* T,
                                    double* wdot, int* status, void* stream) {
  if (!m) return flint_fail(FLINT_EARG, "null mechanism");
  if (!m->gen) return flint_fail(FLINT_EARG, "flint_general_setup has not been called for this mechanism");
  if (n < 0 || ld < n) return flint_fail(FLINT_EARG, "need 0 <= n <= ld");
  if (n == 0) return FLINT_OK;
  if (!rho || !T || !wdot) return flint_fail(FLINT_EARG, "null array");
  hipError_t e = flint_general_launch(m->gen->P, m->d_tabk, m->Tk_min, m->Tk_max, m->d_W, n, ld, rho, T, wdot, status,
                                      (hipStream_t)stream);
  if (e != hipSuccess) return flint_fail(FLINT_EHIP, std::string("general launch: ") + hipGetErrorString(e));
  return FLINT_OK;
}
