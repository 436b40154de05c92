// Registry of compiled mechanisms: each generated translation unit registers its launchers at
// load time; the C-ABI looks them up by mechanism id (FLINT's Assign_Mechanism name, lower-cased
// with non-alphanumerics as '_', Lib_Chemistry_wdot.f90:24-92).
#pragma once
#include <hip/hip_runtime.h>

struct flint_host_prog;
namespace flint_cea {
struct CeaProg;
}
typedef hipError_t (*flint_cea_fn)(const flint_cea::CeaProg& P, bool hp, long n, long ld, const double* p,
                                   const double* T, const double* rho, const double* guess, double* Tout,
                                   double* Yout, int* iters, int* status, hipStream_t s);

struct flint_mech_ops {
  const char* id;
  int ns;
  int wk;  // kinetics table row width (doubles)
  hipError_t (*wdot)(long n, long ld, const double* rho, const double* T, double* wdot, int* status,
                     const double* tabk, int Tk_min, int Tk_max, hipStream_t s);
  hipError_t (*rhs_jac)(long n, long ld, const double* Z, double* F, double* J, int* status, const double* tabk,
                        int Tk_min, int Tk_max, const double* tabh, int Th_min, int Th_max, hipStream_t s);
  const flint_host_prog* prog;  // rate program for the reaction-parallel reactor kernel
  flint_cea_fn cea;              // equilibrium kernel (null if the mechanism has no composition)
  int cea_ne;                    // element count of the compiled element matrix
  const double* cea_A;           // [ne][ns] compiled element matrix
};

int flint_register_mech(const flint_mech_ops* ops);
const flint_mech_ops* flint_find_mech(const char* id);
