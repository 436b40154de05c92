// C ABI of the library (include/flint_amd.h): mechanism handles, device table layout, dispatch
// to the generated / hand-written kernels, FLINT native-file reader.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/flint_amd.h"
#include "flint_internal.h"
#include "flint_thermo.h"
#include "flint_registry.h"

// ------------------------------------------------------------------------------------------------
// errors and registry
// ------------------------------------------------------------------------------------------------
static thread_local std::string g_err;

int flint_fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

extern "C" const char* flint_last_error(void) { return g_err.c_str(); }

static const flint_mech_ops* g_ops[64];
static int g_nops = 0;

int flint_register_mech(const flint_mech_ops* ops) {
  if (g_nops < 64) g_ops[g_nops++] = ops;
  return g_nops;
}

const flint_mech_ops* flint_find_mech(const char* id) {
  for (int i = 0; i < g_nops; ++i)
    if (std::strcmp(g_ops[i]->id, id) == 0) return g_ops[i];
  return nullptr;
}

std::string flint_mech_id(const std::string& name) {
  std::string id;
  for (char c : name) id += std::isalnum((unsigned char)c) ? (char)std::tolower((unsigned char)c) : '_';
  return id;
}

hipError_t flint_wd_launch(long n, long ld, double* rho, const double* T, double* wdot, double* hrr, int* status,
                           const double* W, const double* tabk, int Tk_min, int Tk_max, int wk, const double* tabh,
                           int Th_min, int Th_max, int write_back, hipStream_t s);

#define HIPCHK(expr)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) return flint_fail(FLINT_EHIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

// ------------------------------------------------------------------------------------------------
// mechanism construction
// ------------------------------------------------------------------------------------------------
extern "C" int flint_mech_create(const char* mech_name, int ns, const double* W, int Tk_min, int Tk_max, int n_arrh,
                                 const double* kf, const double* kb, int n_troe, const double* troe, int n_lind,
                                 const double* lind, int Th_min, int Th_max, const double* cp, const double* h, int ne,
                                 const double* comp, flint_mech** out) {
  if (!out || !mech_name || !W || !kf || !kb || !cp || !h) return flint_fail(FLINT_EARG, "null argument");
  *out = nullptr;
  if (ns <= 0 || ns > 64 || Tk_max <= Tk_min || Th_max <= Th_min || n_arrh < 0 || n_troe < 0 || n_lind < 0)
    return flint_fail(FLINT_EARG, "invalid sizes");
  if ((n_troe > 0 && !troe) || (n_lind > 0 && !lind)) return flint_fail(FLINT_EARG, "missing falloff tables");
  auto m = new flint_mech();
  m->name = mech_name;
  m->id = flint_mech_id(mech_name);
  m->is_wd = (m->id == "wd");
  // A mechanism without compiled kernels is kept (FLINT's Assign_Mechanism falls back to the
  // data-driven general routine for it): its source terms then come from flint_source_general
  // once flint_general_setup has run, and the entry points that need compiled kernels (RHS,
  // Jacobian, reactors, equilibrium) report FLINT_EUNSUPPORTED.
  m->ops = m->is_wd ? nullptr : flint_find_mech(m->id.c_str());
  if (m->ops && m->ops->ns != ns) {
    delete m;
    return flint_fail(FLINT_EARG, "species count does not match the compiled mechanism");
  }
  if (m->is_wd && (ns != 5 || n_arrh < 3)) {
    delete m;
    return flint_fail(FLINT_EARG, "WD needs 5 species and 3 reactions");
  }
  m->ns = ns;
  m->Tk_min = Tk_min;
  m->Tk_max = Tk_max;
  m->Th_min = Th_min;
  m->Th_max = Th_max;
  m->n_arrh = n_arrh;
  m->n_troe = n_troe;
  m->n_lind = n_lind;
  m->W.assign(W, W + ns);
  m->ne = ne;
  if (ne > 0 && comp) m->comp.assign(comp, comp + (size_t)ne * ns);
  const long nTk = (long)Tk_max - Tk_min + 1, nTh = (long)Th_max - Th_min + 1;
  int wk = 2 * n_arrh + 4 * n_troe + 3 * n_lind;
  wk += wk & 1;
  if (m->ops && m->ops->wk != wk) {
    delete m;
    return flint_fail(FLINT_EARG, "reaction counts do not match the compiled mechanism");
  }
  m->wk = wk;
  // kinetics: row-major [nTk+1][wk], interleaved (kf,kb) pairs then Troe quadruples, then Lindemann
  m->h_tabk.assign((size_t)(nTk + 1) * wk, 0.0);
  for (long t = 0; t < nTk; ++t) {
    double* row = &m->h_tabk[(size_t)t * wk];
    for (int a = 0; a < n_arrh; ++a) {
      row[2 * a] = kf[(size_t)a * nTk + t];
      row[2 * a + 1] = kb[(size_t)a * nTk + t];
    }
    for (int q = 0; q < n_troe; ++q)
      for (int v = 0; v < 4; ++v) row[2 * n_arrh + 4 * q + v] = troe[((size_t)v * n_troe + q) * nTk + t];
    for (int q = 0; q < n_lind; ++q)
      for (int v = 0; v < 3; ++v) row[2 * n_arrh + 4 * n_troe + 3 * q + v] = lind[((size_t)v * n_lind + q) * nTk + t];
  }
  std::copy(m->h_tabk.begin() + (size_t)(nTk - 1) * wk, m->h_tabk.begin() + (size_t)nTk * wk,
            m->h_tabk.begin() + (size_t)nTk * wk);
  // thermo: row-major [nTh+1][2 ns], (cp, h) pairs
  m->h_tabh.assign((size_t)(nTh + 1) * 2 * ns, 0.0);
  for (long t = 0; t < nTh; ++t)
    for (int k = 0; k < ns; ++k) {
      m->h_tabh[(size_t)t * 2 * ns + 2 * k] = cp[(size_t)k * nTh + t];
      m->h_tabh[(size_t)t * 2 * ns + 2 * k + 1] = h[(size_t)k * nTh + t];
    }
  std::copy(m->h_tabh.begin() + (size_t)(nTh - 1) * 2 * ns, m->h_tabh.begin() + (size_t)nTh * 2 * ns,
            m->h_tabh.begin() + (size_t)nTh * 2 * ns);
  m->t_cp.assign(cp, cp + (size_t)ns * nTh);
  m->t_h.assign(h, h + (size_t)ns * nTh);
  m->t_s.assign((size_t)ns * nTh, 0.0);
  m->t_dcpi.assign((size_t)ns * nTh, 0.0);
  for (int k = 0; k < ns; ++k) m->species.push_back("S" + std::to_string(k + 1));
  int rc = flint_mech_upload(m);
  if (!rc) rc = flint_reactor_prepare(m);
  if (rc) {
    flint_mech_destroy(m);
    return rc;
  }
  *out = m;
  return FLINT_OK;
}

int flint_mech_upload(flint_mech* m) {
  HIPCHK(hipGetDevice(&m->device));
  HIPCHK(hipMalloc(&m->d_tabk, m->h_tabk.size() * sizeof(double)));
  HIPCHK(hipMalloc(&m->d_tabh, m->h_tabh.size() * sizeof(double)));
  // padded to 64 entries so kernels may read W[j] for any j of a 32/64-wide unrolled loop
  std::vector<double> Wp(std::max<size_t>(m->W.size(), 64), 1.0);
  std::copy(m->W.begin(), m->W.end(), Wp.begin());
  HIPCHK(hipMalloc(&m->d_W, Wp.size() * sizeof(double)));
  HIPCHK(hipMemcpy(m->d_tabk, m->h_tabk.data(), m->h_tabk.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(m->d_tabh, m->h_tabh.data(), m->h_tabh.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(m->d_W, Wp.data(), Wp.size() * sizeof(double), hipMemcpyHostToDevice));
  return FLINT_OK;
}

extern "C" void flint_mech_destroy(flint_mech* m) {
  if (!m) return;
  flint_mech_release_extra(m);
  flint_mech_release_cea(m);
  flint_mech_release_sur(m);
  flint_mech_release_transport(m);
  flint_mech_release_general(m);
  if (m->d_tabk) (void)hipFree(m->d_tabk);
  if (m->d_tabh) (void)hipFree(m->d_tabh);
  if (m->d_W) (void)hipFree(m->d_W);
  if (m->d_stage) (void)hipFree(m->d_stage);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

extern "C" int flint_mech_thermo(const flint_mech* m, double* W, double* cp, double* h, double* s, double* dcpi) {
  if (!m) return flint_fail(FLINT_EARG, "null mechanism");
  if (W) std::copy(m->W.begin(), m->W.end(), W);
  if (cp) std::copy(m->t_cp.begin(), m->t_cp.end(), cp);
  if (h) std::copy(m->t_h.begin(), m->t_h.end(), h);
  if (s) std::copy(m->t_s.begin(), m->t_s.end(), s);
  if (dcpi) std::copy(m->t_dcpi.begin(), m->t_dcpi.end(), dcpi);
  return FLINT_OK;
}

extern "C" int flint_mech_names(const flint_mech* m, int kind, char* buf, int len, int* count) {
  if (!m) return flint_fail(FLINT_EARG, "null mechanism");
  if (kind != 0 && kind != 1) return flint_fail(FLINT_EARG, "kind must be 0 (species) or 1 (elements)");
  const std::vector<std::string>& v = kind == 0 ? m->species : m->elements;
  if (count) *count = (int)v.size();
  if (!buf) return FLINT_OK;
  if (len <= 0) return flint_fail(FLINT_EARG, "need len > 0");
  for (size_t i = 0; i < v.size(); ++i) {
    char* d = buf + i * (size_t)len;
    std::fill(d, d + len, ' ');
    std::copy(v[i].begin(), v[i].begin() + std::min<size_t>(v[i].size(), (size_t)len), d);
  }
  return FLINT_OK;
}

extern "C" int flint_mech_composition(const flint_mech* m, double* comp, double* el_weight) {
  if (!m) return flint_fail(FLINT_EARG, "null mechanism");
  if (comp) std::copy(m->comp.begin(), m->comp.end(), comp);
  if (el_weight) {
    for (size_t e = 0; e < (size_t)m->ne; ++e) el_weight[e] = e < m->el_weight.size() ? m->el_weight[e] : 0.0;
  }
  return FLINT_OK;
}

extern "C" int flint_mech_query(const flint_mech* m, int* ns, int* Tk_min, int* Tk_max, int* Th_min, int* Th_max) {
  if (!m) return flint_fail(FLINT_EARG, "null mechanism");
  if (ns) *ns = m->ns;
  if (Tk_min) *Tk_min = m->Tk_min;
  if (Tk_max) *Tk_max = m->Tk_max;
  if (Th_min) *Th_min = m->Th_min;
  if (Th_max) *Th_max = m->Th_max;
  return FLINT_OK;
}

// ------------------------------------------------------------------------------------------------
// batched device entry points
// ------------------------------------------------------------------------------------------------
static int check_batch(const flint_mech* m, long n, long ld) {
  if (!m) return flint_fail(FLINT_EARG, "null mechanism");
  if (n < 0 || ld < n) return flint_fail(FLINT_EARG, "need 0 <= n <= ld");
  return FLINT_OK;
}

extern "C" int flint_source(const flint_mech* m, long n, long ld, const double* rho, const double* T, double* wdot,
                            double* hrr, int* status, void* stream) {
  if (int rc = check_batch(m, n, ld)) return rc;
  if (n == 0) return FLINT_OK;
  if (!rho || !T || !wdot) return flint_fail(FLINT_EARG, "null array");
  hipStream_t s = (hipStream_t)stream;
  hipError_t e;
  if (m->is_wd) {
    e = flint_wd_launch(n, ld, const_cast<double*>(rho), T, wdot, hrr, status, m->d_W, m->d_tabk, m->Tk_min, m->Tk_max,
                        m->wk, m->d_tabh, m->Th_min, m->Th_max, 0, s);
  } else if (!m->ops) {
    // chemistry_source => general (Lib_Chemistry_wdot.f90:86-89)
    if (hrr) return flint_fail(FLINT_EUNSUPPORTED, "heat release output is provided for WD only");
    if (!m->gen)
      return flint_fail(FLINT_EMECH, "no compiled kernels for mechanism " + m->name +
                                         " and no general tables (flint_general_setup)");
    return flint_source_general(m, n, ld, rho, T, wdot, status, stream);
  } else {
    if (hrr) return flint_fail(FLINT_EUNSUPPORTED, "heat release output is provided for WD only");
    e = m->ops->wdot(n, ld, rho, T, wdot, status, m->d_tabk, m->Tk_min, m->Tk_max, s);
  }
  if (e != hipSuccess) return flint_fail(FLINT_EHIP, hipGetErrorString(e));
  return FLINT_OK;
}

extern "C" int flint_rhs(const flint_mech* m, long n, long ld, const double* Z, double* F, double* J, int* status,
                         void* stream) {
  if (int rc = check_batch(m, n, ld)) return rc;
  if (n == 0) return FLINT_OK;
  if (!Z || !F) return flint_fail(FLINT_EARG, "null array");
  if (m->is_wd) return flint_fail(FLINT_EUNSUPPORTED, "rhs/Jacobian kernels are generated for YTF mechanisms only");
  if (!m->ops) return flint_fail(FLINT_EUNSUPPORTED, "no compiled rhs/Jacobian kernel for mechanism " + m->name);
  hipError_t e = m->ops->rhs_jac(n, ld, Z, F, J, status, m->d_tabk, m->Tk_min, m->Tk_max, m->d_tabh, m->Th_min,
                                 m->Th_max, (hipStream_t)stream);
  if (e != hipSuccess) return flint_fail(FLINT_EHIP, hipGetErrorString(e));
  return FLINT_OK;
}

static int check_thermo(const flint_mech* m, long n, long ld, int nd) {
  if (int rc = check_batch(m, n, ld)) return rc;
  if (nd < 0 || nd > 3) return flint_fail(FLINT_EARG, "need 0 <= nd <= 3 velocity components");
  if (m->ns > FLINT_THERMO_MAXNS) return flint_fail(FLINT_EUNSUPPORTED, "more species than the kernel holds");
  return FLINT_OK;
}

extern "C" int flint_cons2prim(const flint_mech* m, long n, long ld, int nd, const double* cons, const double* T,
                               double* prim, double* T_out, int* iters, int* status, void* stream) {
  if (int rc = check_thermo(m, n, ld, nd)) return rc;
  if (n == 0) return FLINT_OK;
  if (!cons || !T || !prim) return flint_fail(FLINT_EARG, "null array");
  hipError_t e = flint_thermo_cons2prim(m->d_tabh, m->Th_min, m->Th_max, m->ns, m->d_W, n, ld, nd, cons, T, prim, T_out,
                                        iters, status, (hipStream_t)stream);
  if (e != hipSuccess) return flint_fail(FLINT_EHIP, hipGetErrorString(e));
  return FLINT_OK;
}

extern "C" int flint_prim2cons(const flint_mech* m, long n, long ld, int nd, const double* prim, double* cons,
                               int* status, void* stream) {
  if (int rc = check_thermo(m, n, ld, nd)) return rc;
  if (n == 0) return FLINT_OK;
  if (!prim || !cons) return flint_fail(FLINT_EARG, "null array");
  hipError_t e = flint_thermo_prim2cons(m->d_tabh, m->Th_min, m->Th_max, m->ns, m->d_W, n, ld, nd, prim, cons, status,
                                        (hipStream_t)stream);
  if (e != hipSuccess) return flint_fail(FLINT_EHIP, hipGetErrorString(e));
  return FLINT_OK;
}

// ------------------------------------------------------------------------------------------------
// host-memory convenience
// ------------------------------------------------------------------------------------------------
int flint_stage(flint_mech* m, size_t doubles) {
  if (!m->stream) HIPCHK(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
  if (m->stage_cap < doubles) {
    if (m->d_stage) HIPCHK(hipFree(m->d_stage));
    m->d_stage = nullptr;
    HIPCHK(hipMalloc(&m->d_stage, doubles * sizeof(double)));
    m->stage_cap = doubles;
  }
  return FLINT_OK;
}

extern "C" int flint_source_host(flint_mech* m, long n, const double* rho, const double* T, double* wdot) {
  if (int rc = check_batch(m, n, n)) return rc;
  if (n == 0) return FLINT_OK;
  std::lock_guard<std::mutex> g(m->mu);
  const size_t ns = m->ns;
  if (int rc = flint_stage(m, (2 * ns + 1) * (size_t)n)) return rc;
  double* d_rho = m->d_stage;
  double* d_T = d_rho + ns * n;
  double* d_w = d_T + n;
  HIPCHK(hipMemcpyAsync(d_rho, rho, ns * n * sizeof(double), hipMemcpyHostToDevice, m->stream));
  HIPCHK(hipMemcpyAsync(d_T, T, n * sizeof(double), hipMemcpyHostToDevice, m->stream));
  if (int rc = flint_source(m, n, n, d_rho, d_T, d_w, nullptr, nullptr, m->stream)) return rc;
  HIPCHK(hipMemcpyAsync(wdot, d_w, ns * n * sizeof(double), hipMemcpyDeviceToHost, m->stream));
  HIPCHK(hipStreamSynchronize(m->stream));
  return FLINT_OK;
}

extern "C" int flint_source_general_host(flint_mech* m, long n, const double* rho, const double* T, double* wdot) {
  if (int rc = check_batch(m, n, n)) return rc;
  if (n == 0) return FLINT_OK;
  std::lock_guard<std::mutex> g(m->mu);
  const size_t ns = m->ns;
  if (int rc = flint_stage(m, (2 * ns + 1) * (size_t)n)) return rc;
  double* d_rho = m->d_stage;
  double* d_T = d_rho + ns * n;
  double* d_w = d_T + n;
  HIPCHK(hipMemcpyAsync(d_rho, rho, ns * n * sizeof(double), hipMemcpyHostToDevice, m->stream));
  HIPCHK(hipMemcpyAsync(d_T, T, n * sizeof(double), hipMemcpyHostToDevice, m->stream));
  if (int rc = flint_source_general(m, n, n, d_rho, d_T, d_w, nullptr, m->stream)) return rc;
  HIPCHK(hipMemcpyAsync(wdot, d_w, ns * n * sizeof(double), hipMemcpyDeviceToHost, m->stream));
  HIPCHK(hipStreamSynchronize(m->stream));
  return FLINT_OK;
}

extern "C" int flint_cons2prim_host(flint_mech* m, long n, int nd, const double* cons, const double* T, double* prim,
                                    double* T_out, int* status) {
  if (int rc = check_thermo(m, n, n, nd)) return rc;
  if (n == 0) return FLINT_OK;
  if (!cons || !T || !prim) return flint_fail(FLINT_EARG, "null array");
  std::lock_guard<std::mutex> g(m->mu);
  const size_t np = m->ns + nd + 1;
  if (int rc = flint_stage(m, (2 * np + 3) * (size_t)n)) return rc;
  double* d_c = m->d_stage;
  double* d_p = d_c + np * n;
  double* d_T = d_p + np * n;
  double* d_To = d_T + n;
  int* d_st = reinterpret_cast<int*>(d_To + n);
  HIPCHK(hipMemcpyAsync(d_c, cons, np * n * sizeof(double), hipMemcpyHostToDevice, m->stream));
  HIPCHK(hipMemcpyAsync(d_T, T, n * sizeof(double), hipMemcpyHostToDevice, m->stream));
  if (int rc = flint_cons2prim(m, n, n, nd, d_c, d_T, d_p, d_To, nullptr, d_st, m->stream)) return rc;
  HIPCHK(hipMemcpyAsync(prim, d_p, np * n * sizeof(double), hipMemcpyDeviceToHost, m->stream));
  if (T_out) HIPCHK(hipMemcpyAsync(T_out, d_To, n * sizeof(double), hipMemcpyDeviceToHost, m->stream));
  if (status) HIPCHK(hipMemcpyAsync(status, d_st, n * sizeof(int), hipMemcpyDeviceToHost, m->stream));
  HIPCHK(hipStreamSynchronize(m->stream));
  return FLINT_OK;
}

extern "C" int flint_prim2cons_host(flint_mech* m, long n, int nd, const double* prim, double* cons) {
  if (int rc = check_thermo(m, n, n, nd)) return rc;
  if (n == 0) return FLINT_OK;
  if (!prim || !cons) return flint_fail(FLINT_EARG, "null array");
  std::lock_guard<std::mutex> g(m->mu);
  const size_t np = m->ns + nd + 1;
  if (int rc = flint_stage(m, 2 * np * (size_t)n)) return rc;
  double* d_q = m->d_stage;
  double* d_c = d_q + np * n;
  HIPCHK(hipMemcpyAsync(d_q, prim, np * n * sizeof(double), hipMemcpyHostToDevice, m->stream));
  if (int rc = flint_prim2cons(m, n, n, nd, d_q, d_c, nullptr, m->stream)) return rc;
  HIPCHK(hipMemcpyAsync(cons, d_c, np * n * sizeof(double), hipMemcpyDeviceToHost, m->stream));
  HIPCHK(hipStreamSynchronize(m->stream));
  return FLINT_OK;
}

extern "C" int flint_rhs_host(flint_mech* m, long n, const double* Z, double* F, double* J) {
  if (int rc = check_batch(m, n, n)) return rc;
  if (n == 0) return FLINT_OK;
  std::lock_guard<std::mutex> g(m->mu);
  const size_t nz = m->ns + 1;
  if (int rc = flint_stage(m, (2 * nz + (J ? nz * nz : 0)) * (size_t)n)) return rc;
  double* d_Z = m->d_stage;
  double* d_F = d_Z + nz * n;
  double* d_J = J ? d_F + nz * n : nullptr;
  HIPCHK(hipMemcpyAsync(d_Z, Z, nz * n * sizeof(double), hipMemcpyHostToDevice, m->stream));
  if (int rc = flint_rhs(m, n, n, d_Z, d_F, d_J, nullptr, m->stream)) return rc;
  HIPCHK(hipMemcpyAsync(F, d_F, nz * n * sizeof(double), hipMemcpyDeviceToHost, m->stream));
  if (J) HIPCHK(hipMemcpyAsync(J, d_J, nz * nz * n * sizeof(double), hipMemcpyDeviceToHost, m->stream));
  HIPCHK(hipStreamSynchronize(m->stream));
  return FLINT_OK;
}
