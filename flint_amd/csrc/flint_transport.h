// Mixture viscosity and thermal conductivity by Wilke's rule (transport.hip).
#pragma once
#include <hip/hip_runtime.h>

#define FLINT_TRANSPORT_MAXNS 64

namespace flint_tr {

struct TrProg {
  const double* tab;  // [T][ns][2] (mu, k) per species, rows Tt_min..Tt_max
  const double* M;    // [3][ns][ns] per species pair (i row, j column): D, 2 D c, D c^2 with
                      // D = 1/sqrt(8 (1 + W_i/W_j)) and c = 1/(W_i/W_j)^(1/4)
  const double* W;    // [ns] molecular weights
  int Tt_min, Tt_max, ns;
};

}  // namespace flint_tr

// mu, lam [n] of cells rhoi [ns][ld], rho [n], T [n] (status may be null: 0, or 2 outside the table)
hipError_t flint_transport_wilke(const flint_tr::TrProg& P, long n, long ld, const double* rhoi, const double* rho,
                                 const double* T, double* mu, double* lam, int* status, hipStream_t s);
