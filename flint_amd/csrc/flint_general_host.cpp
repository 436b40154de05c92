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
  std::vector<flint_general::Hdr> hdr(nr);
  std::vector<flint_general::Ent> ent;
  auto factor_entry = [](int s, double x, bool arrhenius) {
    // nint() for Arrhenius (halves away from zero, Lib_Chemistry_wdot.f90:140-141); a real exponent
    // of a fall-off reaction keeps pow unless it is integer-valued
    const double e = arrhenius ? std::round(x) : x;
    return flint_general::Ent{e, s, e == std::rint(e) ? (int)e : flint_general::kRealExp};
  };
  for (int j = 0; j < nr; ++j) {
    const double* a1 = ni1 + (size_t)j * nz;
    const double* a2 = ni2 + (size_t)j * nz;
    const double* e = eps + (size_t)j * nz;
    const int k = j < na ? 0 : (j < na + nt ? 1 : 2);
    for (int s = 0; s < ns; ++s) {
      if (!std::isfinite(a1[s]) || !std::isfinite(a2[s]) || !std::isfinite(e[s]))
        return flint_fail(FLINT_EARG, "non-finite table entry");
      if (std::fabs(a1[s]) > 64.0 || std::fabs(a2[s]) > 64.0)
        return flint_fail(FLINT_EARG, "stoichiometric coefficient beyond 64");
    }
    flint_general::Hdr& h = hdr[j];
    h.kind = k | (a1[ns] > 0.0 ? 4 : 0);
    h.col = k == 0 ? 2 * j : (k == 1 ? 2 * na + 4 * (j - na) : 2 * na + 4 * nt + 3 * (j - na - nt));
    h.pad = 0;
    h.m0 = (int)ent.size();
    for (int s = 0; s < ns; ++s)
      if (e[s] != 0.0) ent.push_back({e[s], s, 0});
    h.f0 = (int)ent.size();
    for (int s = 0; s < ns; ++s)
      if (a1[s] != 0.0) ent.push_back(factor_entry(s, a1[s], k == 0));
    h.r0 = (int)ent.size();
    for (int s = 0; s < ns; ++s)
      if (a2[s] != 0.0) ent.push_back(factor_entry(s, a2[s], k == 0));
    h.o0 = (int)ent.size();
    for (int s = 0; s < ns; ++s) {
      const double d = a2[s] - a1[s];
      if (d != 0.0) ent.push_back({m->W[s] * d, s, 0});
    }
    h.end = (int)ent.size();
  }
  // compact records and the distinct efficiency sets (equal rows give bit-identical sums)
  std::vector<flint_general::Rec> rec(nr);
  std::vector<int> set_ptr(1, 0), set_first;
  std::vector<flint_general::Ent> set_ent;
  for (int j = 0; j < nr; ++j) {
    const flint_general::Hdr& h = hdr[j];
    flint_general::Rec& R = rec[j];
    R = flint_general::Rec{};
    R.kind = h.kind;
    R.col = h.col;
    R.mset = -1;
    const int nf = h.r0 - h.f0, nrv = h.o0 - h.r0, no = h.end - h.o0;
    bool fits = nf <= 4 && nrv <= 4 && no <= 8;
    for (int t = h.f0; t < h.o0 && fits; ++t) fits = ent[t].e >= 1 && ent[t].e <= 255;
    if (fits && (h.kind & 4)) {
      int found = -1;
      for (size_t q = 0; q + 1 < set_ptr.size() && found < 0; ++q) {
        const int a = set_ptr[q], n1 = set_ptr[q + 1] - a;
        if (n1 != h.f0 - h.m0) continue;
        bool same = true;
        for (int t = 0; t < n1 && same; ++t)
          same = set_ent[a + t].sp == ent[h.m0 + t].sp && set_ent[a + t].v == ent[h.m0 + t].v;
        if (same) found = (int)q;
      }
      if (found < 0 && (int)set_ptr.size() - 1 < flint_general::kMaxSets) {
        set_ent.insert(set_ent.end(), ent.begin() + h.m0, ent.begin() + h.f0);
        set_ptr.push_back((int)set_ent.size());
        found = (int)set_ptr.size() - 2;
      }
      if (found < 0) fits = false;
      R.mset = found;
    }
    if (!fits) {
      R.kind |= 8;
      continue;
    }
    R.cnt = nf | nrv << 4 | no << 8;
    for (int q = 0; q < nf; ++q) {
      R.fsp |= (unsigned)ent[h.f0 + q].sp << (8 * q);
      R.fexp |= (unsigned)ent[h.f0 + q].e << (8 * q);
    }
    for (int q = 0; q < nrv; ++q) {
      R.rsp |= (unsigned)ent[h.r0 + q].sp << (8 * q);
      R.rexp |= (unsigned)ent[h.r0 + q].e << (8 * q);
    }
    for (int q = 0; q < no; ++q) {
      (q < 4 ? R.osp0 : R.osp1) |= (unsigned)ent[h.o0 + q].sp << (8 * (q & 3));
      R.wd[q] = ent[h.o0 + q].v;
    }
  }
  if (set_ent.empty()) set_ent.push_back({0.0, 0, 0});
  std::vector<int> cols(nr + 2, 0);
  for (int j = 0; j < nr; ++j) cols[j] = rec[j].col;
  rec.push_back(flint_general::Rec{});  // read ahead of the last reaction, never used
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
  rc |= put(g, hdr, &P.hdr);
  rc |= put(g, ent, &P.ent);
  rc |= put(g, rec, &P.rec);
  rc |= put(g, cols, &P.col);
  P.nsets = (int)set_ptr.size() - 1;
  rc |= put(g, set_ptr, &P.set_ptr);
  rc |= put(g, set_ent, &P.set_ent);
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
