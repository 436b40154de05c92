// Device program of the data-driven general source term (general.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace flint_general {

// One reaction per entry, in FLINT's order (Arrhenius and three-body, then Troe, then Lindemann),
// with the nonzero entries of its dense (ns+1)-row tables as lists in species order.
struct GenProg {
  int ns, nr, na, nt, nl, wk;
  const int* kind;       // [nr] 0 Arrhenius, 1 Troe, 2 Lindemann; | 4 when ni1(M) > 0 (third body)
  const int* f_ptr;      // [nr+1] forward factors: f_sp, f_exp
  const int* f_sp;
  const double* f_exp;   // integer-valued for Arrhenius (applied as integer powers)
  const int* r_ptr;      // [nr+1] reverse factors
  const int* r_sp;
  const double* r_exp;
  const int* m_ptr;      // [nr+1] third-body efficiencies (nonzero ones)
  const int* m_sp;
  const double* m_eps;
  const int* o_ptr;      // [nr+1] production terms: species, W_k (ni2 - ni1)
  const int* o_sp;
  const double* o_wd;
};

}  // namespace flint_general

hipError_t flint_general_launch(const flint_general::GenProg& P, const double* tabk, int Tk_min, int Tk_max,
                                const double* W, long n, long ld, const double* rho, const double* T, double* wdot,
                                int* status, hipStream_t s);
