// Internal definition of the opaque mechanism handle shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "flint_registry.h"

struct flint_extra;     // per-handle state of the reactor path
struct flint_general_data;  // program of the data-driven general source term
struct flint_cea_data;  // per-handle state of the equilibrium path
struct flint_sur_data;  // per-handle buffers of the learned reactor step
struct flint_transport_data;  // species transport tables and the Wilke weight-factor matrices

struct flint_mech {
  std::string name, id;
  int ns = 0, ne = 0;
  int Tk_min = 0, Tk_max = 0, Th_min = 0, Th_max = 0;
  int n_arrh = 0, n_troe = 0, n_lind = 0, wk = 0;
  bool is_wd = false;
  int device = 0;
  const flint_mech_ops* ops = nullptr;
  std::vector<double> W, comp;       // comp[ne*ns], species_composition(ne, ns) column-major
  std::vector<double> h_tabk, h_tabh;  // host copies of the device tables
  // what FLINT's loaders leave in FLINT_Lib_Thermodynamic (Load_ThermoTransport.f90:30-126):
  // species and element names, per-species thermo tables (Th_min..Th_max) x ns column-major,
  // and the element weights of CEA_initialize_global (0 for an unknown symbol)
  std::vector<std::string> species, elements;
  std::vector<double> t_cp, t_h, t_s, t_dcpi, el_weight;
  double* d_tabk = nullptr;
  double* d_tabh = nullptr;
  double* d_W = nullptr;
  // host-pointer convenience calls
  hipStream_t stream = nullptr;
  double* d_stage = nullptr;
  size_t stage_cap = 0;
  std::mutex mu;
  flint_extra* extra = nullptr;
  flint_general_data* gen = nullptr;
  flint_cea_data* cea = nullptr;
  flint_sur_data* sur = nullptr;
  flint_transport_data* tr = nullptr;
};

int flint_fail(int code, const std::string& msg);
std::string flint_mech_id(const std::string& name);
int flint_mech_upload(flint_mech* m);
int flint_stage(flint_mech* m, size_t doubles);
void flint_mech_release_extra(flint_mech* m);
// The implicit-step check of predicted states (verify_kernel): cells with flag 0 whose
// backward-Euler (and, F0 = F(X) given, trapezoidal) Newton corrections over dt are within
// (tol_y, tol_t) take Xp into X; the others get flag 1 and are appended to idx (count must start
// at zero).
int flint_reactor_verify(flint_mech* m, long n, long ld, double* X, const double* Xp, long ldp, const double* F0,
                         long ldf, const double* p, double dt, double tol_y, double tol_t, int* flag, int* idx,
                         int* count, double* dmax, void* stream);
// flint_reactor_advance with the cells taken in the order given (whatever the handle's work-order
// policy): for callers that have ordered the cells themselves.
int flint_reactor_advance_as_given(flint_mech* m, long n, long ld, int mode, double* X, const double* p, double t_end,
                                   double rtol, double atol, int max_steps, int* stats, int* status, void* stream);
int flint_reactor_prepare(flint_mech* m);  // reactor program of a generated mechanism (no-op otherwise)
void flint_mech_release_cea(flint_mech* m);
void flint_mech_release_sur(flint_mech* m);
void flint_mech_release_general(flint_mech* m);
void flint_mech_release_transport(flint_mech* m);
double flint_atomic_weight(const char* symbol);
