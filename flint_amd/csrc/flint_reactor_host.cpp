// Host side of the batched reactor: device copy of the rate program, launch-shape autotuner,
// hipGraph capture/replay, and the C-ABI entry points of include/flint_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/flint_amd.h"
#include "flint_internal.h"
#include "flint_reactor.h"

hipError_t flint_reactor_set_max_lds(int bytes);

struct flint_extra {
  flint_reactor::Prog prog{};
  std::vector<void*> bufs;  // device allocations backing prog
  int* d_counter = nullptr;
  double* d_jac = nullptr;  // per-team Jacobian slots (32 x 33 doubles each)
  size_t jac_teams = 0;
  int teams_per_block = 4;
  int blocks = 0;
  bool tuned = false;
  int newton = 1;  // Newton policy (flint_reactor_set_newton)
  double* d_chkF = nullptr;  // right-hand side at the old states (flint_reactor_check_step, work order)
  size_t chk_cap = 0;
  int order = 0;              // work order policy (flint_reactor_set_order)
  int* d_order = nullptr;     // [order_cap] cell permutation, then FLINT_ORDER_BUCKETS class counts
  size_t order_cap = 0;
  // captured graph of the last flint_reactor_advance_graph call
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  std::vector<char> gkey;
};

#define HIPCHK(expr)                                                                                      \
  do {                                                                                                    \
    hipError_t e_ = (expr);                                                                               \
    if (e_ != hipSuccess) return flint_fail(FLINT_EHIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

void flint_mech_release_extra(flint_mech* m) {
  flint_extra* x = m->extra;
  if (!x) return;
  if (x->gexec) (void)hipGraphExecDestroy(x->gexec);
  if (x->graph) (void)hipGraphDestroy(x->graph);
  for (void* p : x->bufs) (void)hipFree(p);
  if (x->d_counter) (void)hipFree(x->d_counter);
  if (x->d_jac) (void)hipFree(x->d_jac);
  if (x->d_chkF) (void)hipFree(x->d_chkF);
  if (x->d_order) (void)hipFree(x->d_order);
  delete x;
  m->extra = nullptr;
}

template <class T>
static int upload(flint_extra* x, const T* src, size_t count, const T** dst) {
  void* p = nullptr;
  HIPCHK(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
  x->bufs.push_back(p);
  if (count) HIPCHK(hipMemcpy(p, src, count * sizeof(T), hipMemcpyHostToDevice));
  *dst = static_cast<const T*>(p);
  return FLINT_OK;
}

// Re-encodes the rate program for the integrator's right-hand side:
//  * reactions in rounds of 32 (slot 32*round + lane), grouped by kind so that a round runs one
//    code path; each slot holds its kind, third-body set, table column, reaction index and up to
//    three reactant and three product factor slots (5 bits each, 31 = a constant 1);
//  * third-body efficiencies as a dense [ns][nms] table;
//  * the net production sum of species k split into chunks of at most CW terms in reaction order,
//    chunks spread over (round, lane) slots, and per species the list of its chunk slots.
//  * for the Jacobian, each slot's net-production terms (species, coefficient; at most six) and the
//    index of its dq/dM slot, and per species the list of its terms whose reaction has a
//    third-body or falloff concentration M.
static int encode_rhs(const flint_host_prog* hp, flint_reactor::Prog& P, std::vector<int>& rrec,
                      std::vector<float>& effT, std::vector<int>& cterm, std::vector<int>& kpart,
                      std::vector<int>& rprod, std::vector<int>& tbe, std::vector<int>& tbptr) {
  const int nr = hp->nr, ns = hp->ns, nms = hp->nms;
  if (ns > 31) return flint_fail(FLINT_EUNSUPPORTED, "the reactor kernel supports ns <= 31");
  // Arrhenius and three-body reactions (one code path) fill the first wavefront rounds; the
  // falloff reactions (Troe, Lindemann: logarithms, powers and divisions) start a round of their
  // own, so that their code runs in as few rounds as they need.
  // Within each group the reactions with the most net-production terms and factors come first, so
  // that the rounds after the first carry fewer of the Jacobian's scatter iterations (a round runs
  // as many as its busiest reaction needs).  The slot order changes no rate: each reaction's rate
  // goes to its own index, and the production sums run in reaction order.
  std::vector<int> nterm_of(nr, 0), nfac_of(nr, 0);
  for (int t = 0; t < hp->wd_ptr[ns]; ++t) nterm_of[hp->wd_rx[t]]++;
  for (int r = 0; r < nr; ++r)
    for (int f = 0; f < 8; ++f)
      if (hp->rx_fac[r * 8 + f] >= 0) nfac_of[r] += hp->rx_fac[r * 8 + f] >> 8;
  auto busier = [&](int a, int b) {
    if (nterm_of[a] != nterm_of[b]) return nterm_of[a] > nterm_of[b];
    if (nfac_of[a] != nfac_of[b]) return nfac_of[a] > nfac_of[b];
    return a < b;
  };
  std::vector<int> order;
  for (int kind = 0; kind < 2; ++kind)
    for (int r = 0; r < nr; ++r)
      if ((hp->rx_kind[r] & 3) == kind) order.push_back(r);
  std::sort(order.begin(), order.end(), busier);
  const size_t n_plain = order.size();
  if ((int)n_plain < nr)
    while (order.size() % 64) order.push_back(-1);
  const size_t f0 = order.size();
  for (int kind = 2; kind < 4; ++kind)
    for (int r = 0; r < nr; ++r)
      if ((hp->rx_kind[r] & 3) == kind) order.push_back(r);
  std::sort(order.begin() + f0, order.end(), [&](int a, int b) {
    if ((hp->rx_kind[a] & 3) != (hp->rx_kind[b] & 3)) return (hp->rx_kind[a] & 3) < (hp->rx_kind[b] & 3);
    return busier(a, b);
  });
  P.R = ((int)order.size() + 31) / 32;
  rrec.assign((size_t)P.R * 32 * 4, 0);
  for (int s = 0; s < P.R * 32; ++s) {
    int* rec = &rrec[(size_t)s * 4];
    if (s >= (int)order.size() || order[s] < 0) {
      rec[0] = 7;
      rec[2] = nr;
      rec[3] = 0x3fffffff;  // every factor slot 31
      continue;
    }
    const int r = order[s];
    rec[0] = (hp->rx_kind[r] & 3) | ((hp->rx_mset[r] + 1) << 3);
    rec[1] = hp->rx_col[r];
    rec[2] = r;
    unsigned packed = 0;
    for (int side = 0; side < 2; ++side) {
      int slots[3] = {31, 31, 31}, nf = 0;
      for (int f = 0; f < 4; ++f) {
        const int code = hp->rx_fac[r * 8 + side * 4 + f];
        if (code < 0) continue;
        for (int e = 0; e < (code >> 8); ++e) {
          if (nf == 3) return flint_fail(FLINT_EUNSUPPORTED, "the reactor kernel supports at most 3 factors per side");
          slots[nf++] = code & 255;
        }
      }
      for (int f = 0; f < 3; ++f) packed |= (unsigned)slots[f] << (5 * (3 * side + f));
    }
    rec[3] = (int)packed;
  }
  const int nmsp = std::max(nms, 1);
  effT.assign((size_t)ns * nmsp, 0.0f);
  for (int s = 0; s < nms; ++s)
    for (int t = hp->ms_ptr[s]; t < hp->ms_ptr[s + 1]; ++t) {
      effT[(size_t)hp->ms_sp[t] * nmsp + s] = (float)hp->ms_eff[t];
      if ((double)(float)hp->ms_eff[t] != hp->ms_eff[t])
        return flint_fail(FLINT_EUNSUPPORTED, "non-float32 third-body efficiency");
    }
  // production-sum chunks
  P.CW = 16;
  std::vector<std::pair<int, int>> chunks;  // (species, first term)
  for (int k = 0; k < ns; ++k)
    for (int t = hp->wd_ptr[k]; t < hp->wd_ptr[k + 1]; t += P.CW) chunks.push_back({k, t});
  const int nch = (int)chunks.size();
  P.C = std::max(1, (nch + 31) / 32);
  cterm.assign((size_t)P.C * P.CW * 32 * 2, 0);
  for (int i = 0; i < P.C * P.CW * 32; ++i) cterm[(size_t)i * 2] = nr;  // empty term: q[nr] = 0
  std::vector<std::vector<int>> slots_of(ns);
  for (int c = 0; c < nch; ++c) {
    const int k = chunks[c].first, t0 = chunks[c].second, t1 = std::min(t0 + P.CW, hp->wd_ptr[k + 1]);
    const int round = c / 32, lane = c % 32;
    for (int t = t0; t < t1; ++t) {
      const size_t at = (((size_t)round * P.CW + (t - t0)) * 32 + lane) * 2;
      const float coef = (float)hp->wd_coef[t];
      if ((double)coef != hp->wd_coef[t]) return flint_fail(FLINT_EUNSUPPORTED, "non-float32 stoichiometric coefficient");
      cterm[at] = hp->wd_rx[t];
      std::memcpy(&cterm[at + 1], &coef, sizeof coef);
    }
    slots_of[k].push_back(round * 32 + lane);
  }
  // dq/dM slots: the reactions with a concentration M, in reaction order
  std::vector<int> dqslot(nr, -1);
  P.nmr = 0;
  for (int r = 0; r < nr; ++r)
    if (hp->rx_mset[r] >= 0) dqslot[r] = P.nmr++;
  // per reaction its net-production terms (species order), as the Jacobian scatters them
  std::vector<std::vector<std::pair<int, float>>> terms_of(nr);
  for (int k = 0; k < ns; ++k)
    for (int t = hp->wd_ptr[k]; t < hp->wd_ptr[k + 1]; ++t)
      terms_of[hp->wd_rx[t]].push_back({k, (float)hp->wd_coef[t]});
  rprod.assign((size_t)P.R * 32 * 8, 0);
  for (int s = 0; s < P.R * 32; ++s) {
    int* pr = &rprod[(size_t)s * 8];
    if (s >= (int)order.size() || order[s] < 0) continue;  // empty slot: no terms
    const int r = order[s];
    const auto& tl = terms_of[r];
    if (tl.size() > 6) return flint_fail(FLINT_EUNSUPPORTED, "the reactor kernel supports at most 6 species per reaction");
    unsigned sk = 0;
    for (size_t i = 0; i < tl.size(); ++i) {
      sk |= (unsigned)tl[i].first << (5 * i);
      std::memcpy(&pr[2 + i], &tl[i].second, sizeof(float));
    }
    pr[0] = (int)sk;
    pr[1] = (int)tl.size() | ((dqslot[r] + 1) << 3);
  }
  // per species its terms with M (species order, reaction order within), for the third-body columns
  tbptr.assign((size_t)ns + 1, 0);
  tbe.clear();
  for (int k = 0; k < ns; ++k) {
    tbptr[k] = (int)(tbe.size() / 4);
    for (int t = hp->wd_ptr[k]; t < hp->wd_ptr[k + 1]; ++t) {
      const int r = hp->wd_rx[t];
      if (hp->rx_mset[r] < 0) continue;
      const float coef = (float)hp->wd_coef[t];
      int cb;
      std::memcpy(&cb, &coef, sizeof cb);
      tbe.insert(tbe.end(), {k, cb, hp->rx_mset[r], dqslot[r]});
    }
  }
  tbptr[ns] = (int)(tbe.size() / 4);
  P.ntb = tbptr[ns];
  if (tbe.empty()) tbe.assign(4, 0);
  P.MAXP = 1;
  for (int k = 0; k < ns; ++k) P.MAXP = std::max<int>(P.MAXP, (int)slots_of[k].size());
  kpart.assign((size_t)P.MAXP * 32, -1);
  for (int k = 0; k < ns; ++k)
    for (size_t i = 0; i < slots_of[k].size(); ++i) kpart[i * 32 + k] = slots_of[k][i];
  return FLINT_OK;
}

static int build_extra(flint_mech* m) {
  if (m->is_wd || !m->ops || !m->ops->prog)
    return flint_fail(FLINT_EUNSUPPORTED, "the reactor kernel needs a generated mechanism (not WD)");
  if (m->ns + 1 > 32) return flint_fail(FLINT_EUNSUPPORTED, "the reactor kernel supports ns + 1 <= 32");
  const flint_host_prog* hp = m->ops->prog;
  auto* x = new flint_extra();
  m->extra = x;
  flint_reactor::Prog& P = x->prog;
  P.ns = hp->ns;
  P.nr = hp->nr;
  P.nms = hp->nms;
  std::vector<int> rrec, kpart;
  std::vector<float> effT;
  std::vector<int> cterm, rprod, tbe, tbptr;
  if (int rc = encode_rhs(hp, P, rrec, effT, cterm, kpart, rprod, tbe, tbptr)) return rc;
  int rc = 0;
  rc |= upload(x, reinterpret_cast<const int4*>(rrec.data()), rrec.size() / 4, &P.rrec);
  rc |= upload(x, effT.data(), effT.size(), &P.effT);
  rc |= upload(x, reinterpret_cast<const int2*>(cterm.data()), cterm.size() / 2, &P.cterm);
  rc |= upload(x, kpart.data(), kpart.size(), &P.kpart);
  rc |= upload(x, reinterpret_cast<const int4*>(rprod.data()), rprod.size() / 4, &P.rprod);
  rc |= upload(x, reinterpret_cast<const int4*>(tbe.data()), tbe.size() / 4, &P.tbe);
  rc |= upload(x, tbptr.data(), tbptr.size(), &P.tbptr);
  if (rc) return rc;
  P.W = m->d_W;
  P.tabk = m->d_tabk;
  P.Tk_min = m->Tk_min;
  P.Tk_max = m->Tk_max;
  P.wk = m->wk;
  P.tabh = m->d_tabh;
  P.Th_min = m->Th_min;
  P.Th_max = m->Th_max;
  HIPCHK(hipMalloc(&x->d_counter, 64 * sizeof(int)));
  // default launch: four-team blocks (the program is shared per block), as many as are resident per CU
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, m->device));
  const size_t tb = (flint_reactor_team_bytes(P) + 15) & ~size_t(15);
  const size_t lds_cu = prop.maxSharedMemoryPerMultiProcessor ? prop.maxSharedMemoryPerMultiProcessor : 65536;
  const size_t sh = flint_reactor_shared_bytes(P);
  const int per_cu = std::max<int>(1, (int)(std::min<size_t>(lds_cu, 160 * 1024) / (sh + tb * x->teams_per_block)));
  // a team is one wavefront, and the kernel's register budget allows 8 resident wavefronts per CU
  x->blocks = prop.multiProcessorCount * std::max(1, std::min(per_cu, 8 / x->teams_per_block));
  HIPCHK(flint_reactor_set_max_lds((int)std::min<size_t>(lds_cu, 160 * 1024)));
  return FLINT_OK;
}

// The reactor program of the handle, built on first use; a failed build leaves no partial state.
static int ensure_extra(flint_mech* m) {
  if (m->extra) return FLINT_OK;
  const int rc = build_extra(m);
  if (rc) flint_mech_release_extra(m);
  return rc;
}

int flint_reactor_prepare(flint_mech* m) {
  if (m->is_wd || !m->ops || !m->ops->prog) return FLINT_OK;
  const int rc = ensure_extra(m);
  return rc == FLINT_EUNSUPPORTED ? FLINT_OK : rc;  // the other entry points stay usable
}

static int check_args(const flint_mech* m, long n, long ld, int mode, const double* X, const double* p, double t_end,
                      double rtol, double atol) {
  if (!m) return flint_fail(FLINT_EARG, "null mechanism");
  if (n < 0 || ld < n) return flint_fail(FLINT_EARG, "need 0 <= n <= ld");
  if (mode != 0 && mode != 1) return flint_fail(FLINT_EARG, "mode must be 0 (constant volume) or 1 (constant pressure)");
  if (n > 0 && !X) return flint_fail(FLINT_EARG, "null state");
  if (mode == 1 && n > 0 && !p) return flint_fail(FLINT_EARG, "constant pressure needs p");
  if (!(t_end > 0.0) || !(rtol > 0.0) || !(atol > 0.0)) return flint_fail(FLINT_EARG, "need t_end, rtol, atol > 0");
  return FLINT_OK;
}

// Per-team Jacobian slots for a launch of tpb x blocks teams.  Grows only outside graph capture:
// the capture path sizes the buffer before capturing.
static int ensure_jac(flint_extra* x, int tpb, int blocks) {
  const size_t teams = (size_t)tpb * blocks;
  if (teams > x->jac_teams) {
    // a captured graph holds the old slot address: drop it before the buffer moves
    if (x->gexec) (void)hipGraphExecDestroy(x->gexec);
    if (x->graph) (void)hipGraphDestroy(x->graph);
    x->gexec = nullptr;
    x->graph = nullptr;
    x->gkey.clear();
    if (x->d_jac) HIPCHK(hipFree(x->d_jac));
    x->d_jac = nullptr;
    x->jac_teams = 0;
    HIPCHK(hipMalloc(&x->d_jac, teams * 32 * 33 * sizeof(double)));
    x->jac_teams = teams;
  }
  return FLINT_OK;
}

// Right-hand-side buffer [(ns+1) x ld] (d_chkF) for the trapezoidal check and the work order.
static int ensure_rhs_buf(flint_mech* m, long ld) {
  flint_extra* x = m->extra;
  const size_t need = (size_t)(m->ns + 1) * ld;
  if (x->chk_cap < need) {
    if (x->gexec) (void)hipGraphExecDestroy(x->gexec);  // a captured graph may hold the old buffer
    if (x->graph) (void)hipGraphDestroy(x->graph);
    x->gexec = nullptr;
    x->graph = nullptr;
    x->gkey.clear();
    if (x->d_chkF) HIPCHK(hipFree(x->d_chkF));
    x->d_chkF = nullptr;
    x->chk_cap = 0;
    HIPCHK(hipMalloc(&x->d_chkF, need * sizeof(double)));
    x->chk_cap = need;
  }
  return FLINT_OK;
}

// Buffers of the work order for n cells of leading dimension ld (allocates; not under capture).
static int ensure_order(flint_mech* m, long n, long ld) {
  flint_extra* x = m->extra;
  if (int rc = ensure_rhs_buf(m, ld)) return rc;
  if (x->order_cap < (size_t)n) {
    if (x->gexec) (void)hipGraphExecDestroy(x->gexec);
    if (x->graph) (void)hipGraphDestroy(x->graph);
    x->gexec = nullptr;
    x->graph = nullptr;
    x->gkey.clear();
    if (x->d_order) HIPCHK(hipFree(x->d_order));
    x->d_order = nullptr;
    x->order_cap = 0;
    HIPCHK(hipMalloc(&x->d_order, ((size_t)n + FLINT_ORDER_BUCKETS) * sizeof(int)));
    x->order_cap = (size_t)n;
  }
  return FLINT_OK;
}

// Enqueues the work order of policy 1: dX/dt at the initial states, then the cells by descending
// |dT/dt| (the stiffest, costliest cells first).  Returns the permutation (null for policy 0).
static int enqueue_order(flint_mech* m, long n, long ld, int mode, const double* X, const double* p, hipStream_t s,
                         const int** order) {
  flint_extra* x = m->extra;
  *order = nullptr;
  if (x->order == 0) return FLINT_OK;
  const int tpb = x->teams_per_block;
  const int blocks = (int)std::min<long>(x->blocks, (n + tpb - 1) / tpb);
  hipError_t e = flint_reactor_jac_launch(x->prog, mode, n, ld, X, p, x->d_chkF, nullptr, tpb, blocks, x->d_jac, s);
  if (e == hipSuccess)
    e = flint_reactor_order(n, x->d_chkF + (size_t)m->ns * ld, nullptr, x->d_order, x->d_order + x->order_cap, s);
  if (e != hipSuccess) return flint_fail(FLINT_EHIP, std::string("work order: ") + hipGetErrorString(e));
  *order = x->d_order;
  return FLINT_OK;
}

static int launch(flint_mech* m, long n, long ld, int mode, double* X, const double* p, double t_end, double rtol,
                  double atol, int max_steps, int* stats, int* status, hipStream_t s, int tpb, int blocks,
                  const int* order = nullptr) {
  flint_extra* x = m->extra;
  if (int rc = ensure_jac(x, tpb, blocks)) return rc;
  flint_reactor::Opts O{mode, rtol, atol, t_end, max_steps > 0 ? max_steps : 20000, x->newton};
  hipError_t e =
      flint_reactor_launch(x->prog, O, n, ld, X, p, stats, status, x->d_counter, tpb, blocks, x->d_jac, order, s);
  if (e != hipSuccess) return flint_fail(FLINT_EHIP, std::string("reactor launch: ") + hipGetErrorString(e));
  return FLINT_OK;
}

extern "C" int flint_reactor_advance(flint_mech* m, long n, long ld, int mode, double* X, const double* p, double t_end,
                                     double rtol, double atol, int max_steps, int* stats, int* status, void* stream) {
  if (int rc = check_args(m, n, ld, mode, X, p, t_end, rtol, atol)) return rc;
  if (n == 0) return FLINT_OK;
  if (int rc = ensure_extra(m)) return rc;
  flint_extra* x = m->extra;
  const int* order = nullptr;
  if (x->order) {
    if (int rc = ensure_jac(x, x->teams_per_block, x->blocks)) return rc;
    if (int rc = ensure_order(m, n, ld)) return rc;
    if (int rc = enqueue_order(m, n, ld, mode, X, p, (hipStream_t)stream, &order)) return rc;
  }
  return launch(m, n, ld, mode, X, p, t_end, rtol, atol, max_steps, stats, status, (hipStream_t)stream,
                x->teams_per_block, x->blocks, order);
}

int flint_reactor_advance_as_given(flint_mech* m, long n, long ld, int mode, double* X, const double* p, double t_end,
                                   double rtol, double atol, int max_steps, int* stats, int* status, void* stream) {
  if (int rc = check_args(m, n, ld, mode, X, p, t_end, rtol, atol)) return rc;
  if (n == 0) return FLINT_OK;
  if (int rc = ensure_extra(m)) return rc;
  return launch(m, n, ld, mode, X, p, t_end, rtol, atol, max_steps, stats, status, (hipStream_t)stream,
                m->extra->teams_per_block, m->extra->blocks);
}

// No counterpart in the reference: flint_reactor_advance for callers that cannot allocate device
// memory (code under graph capture, serving loops).  The Jacobian slots and the work counter come
// from the library's static device scratch (flint_reactor_scratch); the handle's program was
// uploaded when the handle was created.  Nothing here calls hipMalloc or hipFree.
extern "C" int flint_reactor_advance_noalloc(flint_mech* m, long n, long ld, int mode, double* X, const double* p,
                                             double t_end, double rtol, double atol, int max_steps, int* stats,
                                             int* status, void* stream) {
  if (int rc = check_args(m, n, ld, mode, X, p, t_end, rtol, atol)) return rc;
  if (n == 0) return FLINT_OK;
  flint_extra* x = m->extra;
  if (!x) return flint_fail(FLINT_EUNSUPPORTED, "the reactor kernel needs a generated mechanism (not WD)");
  const long teams = (long)x->teams_per_block * x->blocks;
  if (teams > FLINT_SCRATCH_TEAMS)
    return flint_fail(FLINT_EARG, "launch shape of " + std::to_string(teams) + " teams exceeds the " +
                                      std::to_string(FLINT_SCRATCH_TEAMS) +
                                      " Jacobian slots of the allocation-free scratch (flint_reactor_set_launch)");
  int* counter = nullptr;
  double* slots = nullptr;
  HIPCHK(flint_reactor_scratch(&counter, &slots));
  flint_reactor::Opts O{mode, rtol, atol, t_end, max_steps > 0 ? max_steps : 20000, x->newton};
  hipError_t e = flint_reactor_launch(x->prog, O, n, ld, X, p, stats, status, counter, x->teams_per_block, x->blocks,
                                      slots, nullptr, (hipStream_t)stream);
  if (e != hipSuccess) return flint_fail(FLINT_EHIP, std::string("reactor launch: ") + hipGetErrorString(e));
  return FLINT_OK;
}

extern "C" int flint_reactor_jacobian(flint_mech* m, long n, long ld, int mode, const double* X, const double* p,
                                      double* F, double* J, void* stream) {
  if (int rc = check_args(m, n, ld, mode, X, p, 1.0, 1.0, 1.0)) return rc;
  if (n > 0 && !F) return flint_fail(FLINT_EARG, "null output");
  if (n == 0) return FLINT_OK;
  if (int rc = ensure_extra(m)) return rc;
  flint_extra* x = m->extra;
  const int tpb = x->teams_per_block;
  const int blocks = (int)std::min<long>(x->blocks, (n + tpb - 1) / tpb);
  if (int rc = ensure_jac(x, tpb, blocks)) return rc;
  hipError_t e = flint_reactor_jac_launch(x->prog, mode, n, ld, X, p, F, J, tpb, blocks, x->d_jac, (hipStream_t)stream);
  if (e != hipSuccess) return flint_fail(FLINT_EHIP, std::string("reactor jacobian launch: ") + hipGetErrorString(e));
  return FLINT_OK;
}

int flint_reactor_verify(flint_mech* m, long n, long ld, double* X, const double* Xp, long ldp, const double* F0,
                         long ldf, const double* p, double dt, double tol_y, double tol_t, int* flag, int* idx,
                         int* count, double* dmax, void* stream) {
  if (int rc = check_args(m, n, ld, 1, X, p, dt, 1.0, 1.0)) return rc;
  if (n == 0) return FLINT_OK;
  if (!Xp || !flag || !idx || !count || ldp < n) return flint_fail(FLINT_EARG, "null or short check buffers");
  if (int rc = ensure_extra(m)) return rc;
  flint_extra* x = m->extra;
  const int tpb = x->teams_per_block;
  const int blocks = (int)std::min<long>(x->blocks, (n + tpb - 1) / tpb);
  if (int rc = ensure_jac(x, tpb, blocks)) return rc;
  if (F0 && ldf < n) return flint_fail(FLINT_EARG, "short right-hand side buffer");
  hipError_t e = flint_reactor_verify_launch(x->prog, n, ld, X, Xp, ldp, F0, ldf, p, dt, tol_y, tol_t, 1, flag, idx,
                                             count, dmax, tpb, blocks, x->d_jac, (hipStream_t)stream);
  if (e != hipSuccess) return flint_fail(FLINT_EHIP, std::string("reactor check launch: ") + hipGetErrorString(e));
  return FLINT_OK;
}

extern "C" int flint_reactor_check_step(flint_mech* m, long n, long ld, const double* X, const double* Xp,
                                        const double* p, double dt, double tol_y, double tol_t, int* flag,
                                        double* check, void* stream) {
  if (int rc = check_args(m, n, ld, 1, X, p, dt, 1.0, 1.0)) return rc;
  if (n == 0) return FLINT_OK;
  if (!Xp || !flag) return flint_fail(FLINT_EARG, "null candidate state or flag array");
  if (!(tol_y >= 0.0) || !(tol_t >= 0.0)) return flint_fail(FLINT_EARG, "need non-negative tolerances");
  if (int rc = ensure_extra(m)) return rc;
  flint_extra* x = m->extra;
  const int tpb = x->teams_per_block;
  const int blocks = (int)std::min<long>(x->blocks, (n + tpb - 1) / tpb);
  if (int rc = ensure_jac(x, tpb, blocks)) return rc;
  if (int rc = ensure_rhs_buf(m, ld)) return rc;
  hipStream_t s = (hipStream_t)stream;
  // F at the old states (the trapezoidal correction), then every cell judged (flag cleared by the kernel)
  hipError_t e = flint_reactor_jac_launch(x->prog, 1, n, ld, X, p, x->d_chkF, nullptr, tpb, blocks, x->d_jac, s);
  if (e == hipSuccess) e = flint_reactor_fill_int(flag, n, 0, s);
  if (e == hipSuccess)
    e = flint_reactor_verify_launch(x->prog, n, ld, const_cast<double*>(X), Xp, ld, x->d_chkF, ld, p, dt, tol_y,
                                    tol_t, 0, flag, nullptr, nullptr, check, tpb, blocks, x->d_jac, s);
  if (e != hipSuccess) return flint_fail(FLINT_EHIP, std::string("reactor check launch: ") + hipGetErrorString(e));
  return FLINT_OK;
}

extern "C" int flint_reactor_set_launch(flint_mech* m, int teams_per_block, int blocks) {
  if (!m) return flint_fail(FLINT_EARG, "null mechanism");
  if (int rc = ensure_extra(m)) return rc;
  if (teams_per_block < 1 || teams_per_block > 8 || blocks < 1) return flint_fail(FLINT_EARG, "bad launch shape");
  m->extra->teams_per_block = teams_per_block;
  m->extra->blocks = blocks;
  m->extra->tuned = true;
  return FLINT_OK;
}

extern "C" int flint_reactor_set_newton(flint_mech* m, int policy) {
  if (!m) return flint_fail(FLINT_EARG, "null mechanism");
  if (policy != 0 && policy != 1) return flint_fail(FLINT_EARG, "Newton policy must be 0 (scipy) or 1 (CVODE)");
  if (int rc = ensure_extra(m)) return rc;
  if (m->extra->newton != policy) {
    // a captured graph holds the old options
    flint_extra* x = m->extra;
    if (x->gexec) (void)hipGraphExecDestroy(x->gexec);
    if (x->graph) (void)hipGraphDestroy(x->graph);
    x->gexec = nullptr;
    x->graph = nullptr;
    x->gkey.clear();
  }
  m->extra->newton = policy;
  return FLINT_OK;
}

extern "C" int flint_reactor_set_order(flint_mech* m, int policy) {
  if (!m) return flint_fail(FLINT_EARG, "null mechanism");
  if (policy != 0 && policy != 1) return flint_fail(FLINT_EARG, "work order must be 0 (as given) or 1 (by |dT/dt|)");
  if (int rc = ensure_extra(m)) return rc;
  flint_extra* x = m->extra;
  if (x->order != policy) {
    if (x->gexec) (void)hipGraphExecDestroy(x->gexec);  // a captured graph holds the old work order
    if (x->graph) (void)hipGraphDestroy(x->graph);
    x->gexec = nullptr;
    x->graph = nullptr;
    x->gkey.clear();
  }
  x->order = policy;
  return FLINT_OK;
}

extern "C" int flint_reactor_get_launch(flint_mech* m, int* teams_per_block, int* blocks) {
  if (!m) return flint_fail(FLINT_EARG, "null mechanism");
  if (int rc = ensure_extra(m)) return rc;
  if (teams_per_block) *teams_per_block = m->extra->teams_per_block;
  if (blocks) *blocks = m->extra->blocks;
  return FLINT_OK;
}

// Autotuner: times every candidate launch shape on a copy of the sample (the caller's data is not
// modified) and keeps the fastest for this handle.  The sample's cells are handed out by descending
// |dT/dt| in every candidate launch (the order of policy 1): a sample holds fewer cells per team
// than the full batch, and in the caller's order one stiff cell started late would set its time.
extern "C" int flint_reactor_autotune(flint_mech* m, long n, long ld, int mode, const double* X, const double* p,
                                      double t_end, double rtol, double atol, double* best_ms) {
  if (int rc = check_args(m, n, ld, mode, X, p, t_end, rtol, atol)) return rc;
  if (n == 0) return FLINT_OK;
  if (int rc = ensure_extra(m)) return rc;
  flint_extra* x = m->extra;
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, m->device));
  const int cus = prop.multiProcessorCount;
  const size_t tb = (flint_reactor_team_bytes(x->prog) + 15) & ~size_t(15);
  const size_t lds_cu = std::min<size_t>(prop.maxSharedMemoryPerMultiProcessor ? prop.maxSharedMemoryPerMultiProcessor : 65536,
                                         160 * 1024);
  const int nz = m->ns + 1;
  double* work = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = FLINT_OK;
  // records the first failure; every later step is skipped and the resources are released below
  auto hip = [&rc](hipError_t e, const char* what) {
    if (e != hipSuccess && rc == FLINT_OK) rc = flint_fail(FLINT_EHIP, std::string("autotune: ") + what + ": " + hipGetErrorString(e));
    return rc == FLINT_OK;
  };
  auto copy_sample = [&]() {
    return hip(hipMemcpy2DAsync(work, n * sizeof(double), X, ld * sizeof(double), n * sizeof(double), nz,
                                hipMemcpyDeviceToDevice, s),
               "sample copy");
  };
  float best = 1e30f;
  int best_tpb = x->teams_per_block, best_blocks = x->blocks;
  const int* order = nullptr;
  if (hip(hipMalloc(&work, (size_t)nz * n * sizeof(double)), "hipMalloc") &&
      hip(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "stream") && hip(hipEventCreate(&e0), "event") &&
      hip(hipEventCreate(&e1), "event")) {
    rc = ensure_jac(x, x->teams_per_block, x->blocks);
    if (!rc) rc = ensure_order(m, n, n);
    if (!rc && copy_sample()) {
      const int keep = x->order;
      x->order = 1;
      rc = enqueue_order(m, n, n, mode, work, p, s, &order);
      x->order = keep;
    }
  }
  for (int tpb : {1, 2, 4, 8}) {
    const int fit = (int)(lds_cu / (flint_reactor_shared_bytes(x->prog) + tb * tpb));
    for (int per_cu : {1, 2, 3, 4, 6, 8}) {
      if (rc || fit < 1 || per_cu > fit || per_cu * tpb > 8) continue;
      const int blocks = cus * per_cu;
      int lrc = 0;
      for (int rep = 0; rep < 2 && !lrc && !rc; ++rep) {
        if (!copy_sample() || !hip(hipEventRecord(e0, s), "event")) break;
        lrc = launch(m, n, n, mode, work, p, t_end, rtol, atol, 0, nullptr, nullptr, s, tpb, blocks, order);
        hip(hipEventRecord(e1, s), "event");
        hip(hipStreamSynchronize(s), "synchronize");
      }
      float ms = 0.f;
      if (lrc || rc || !hip(hipEventElapsedTime(&ms, e0, e1), "timing")) continue;  // a shape that fails to launch is skipped
      if (ms < best) {
        best = ms;
        best_tpb = tpb;
        best_blocks = blocks;
      }
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  if (work) (void)hipFree(work);
  if (rc) return rc;
  x->teams_per_block = best_tpb;
  x->blocks = best_blocks;
  x->tuned = true;
  if (best_ms) *best_ms = best;
  return FLINT_OK;
}

// Same as flint_reactor_advance, through a hipGraph captured on the first call with these
// arguments and replayed on later identical calls (a CFD time loop on fixed buffers).
extern "C" int flint_reactor_advance_graph(flint_mech* m, long n, long ld, int mode, double* X, const double* p,
                                           double t_end, double rtol, double atol, int max_steps, int* stats, int* status,
                                           void* stream) {
  if (int rc = check_args(m, n, ld, mode, X, p, t_end, rtol, atol)) return rc;
  if (n == 0) return FLINT_OK;
  if (int rc = ensure_extra(m)) return rc;
  flint_extra* x = m->extra;
  struct Key {
    long n, ld;
    int mode;
    double* X;
    const double* p;
    double t_end, rtol, atol;
    int max_steps;
    int* stats;
    int* status;
    int tpb, blocks, order;
  } key{n, ld, mode, X, p, t_end, rtol, atol, max_steps, stats, status, x->teams_per_block, x->blocks, x->order};
  std::vector<char> kb(reinterpret_cast<char*>(&key), reinterpret_cast<char*>(&key) + sizeof key);
  hipStream_t s = (hipStream_t)stream;
  if (!x->gexec || kb != x->gkey) {
    if (int rc = ensure_jac(x, x->teams_per_block, x->blocks)) return rc;
    if (x->order)
      if (int rc = ensure_order(m, n, ld)) return rc;
    if (x->gexec) (void)hipGraphExecDestroy(x->gexec);
    if (x->graph) (void)hipGraphDestroy(x->graph);
    x->gexec = nullptr;
    x->graph = nullptr;
    hipStream_t cs;
    HIPCHK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    HIPCHK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    const int* order = nullptr;
    int rc = enqueue_order(m, n, ld, mode, X, p, cs, &order);
    if (!rc)
      rc = launch(m, n, ld, mode, X, p, t_end, rtol, atol, max_steps, stats, status, cs, x->teams_per_block,
                  x->blocks, order);
    hipGraph_t g = nullptr;
    hipError_t ee = hipStreamEndCapture(cs, &g);
    (void)hipStreamDestroy(cs);
    if (rc) return rc;
    if (ee != hipSuccess) return flint_fail(FLINT_EHIP, std::string("capture: ") + hipGetErrorString(ee));
    x->graph = g;
    HIPCHK(hipGraphInstantiate(&x->gexec, g, nullptr, nullptr, 0));
    x->gkey = kb;
  }
  HIPCHK(hipGraphLaunch(x->gexec, s));
  return FLINT_OK;
}

extern "C" int flint_reactor_advance_host(flint_mech* m, long n, int mode, double* X, const double* p, double t_end,
                                          double rtol, double atol, int max_steps, int* status) {
  if (int rc = check_args(m, n, n, mode, X, p, t_end, rtol, atol)) return rc;
  if (n == 0) return FLINT_OK;
  std::lock_guard<std::mutex> g(m->mu);
  const size_t nz = m->ns + 1, nn = (size_t)n;
  // X, p, status (ints in a double-sized slot)
  if (int rc = flint_stage(m, (nz + 2) * nn)) return rc;
  double* d_X = m->d_stage;
  double* d_p = d_X + nz * nn;
  int* d_st = reinterpret_cast<int*>(d_p + nn);
  HIPCHK(hipMemcpyAsync(d_X, X, nz * nn * sizeof(double), hipMemcpyHostToDevice, m->stream));
  if (mode == 1) HIPCHK(hipMemcpyAsync(d_p, p, nn * sizeof(double), hipMemcpyHostToDevice, m->stream));
  if (int rc = flint_reactor_advance(m, n, n, mode, d_X, mode == 1 ? d_p : nullptr, t_end, rtol, atol, max_steps,
                                     nullptr, d_st, m->stream))
    return rc;
  HIPCHK(hipMemcpyAsync(X, d_X, nz * nn * sizeof(double), hipMemcpyDeviceToHost, m->stream));
  if (status) HIPCHK(hipMemcpyAsync(status, d_st, nn * sizeof(int), hipMemcpyDeviceToHost, m->stream));
  HIPCHK(hipStreamSynchronize(m->stream));
  return FLINT_OK;
}
