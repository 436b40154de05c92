// Device data of the batched equilibrium kernel (flint_cea_kernel.h).
#pragma once
#include <hip/hip_runtime.h>

namespace flint_cea {

struct CeaProg {
  const double* tabh;  // [nT+1][ns][2] (cp, h) per mass, the mechanism's thermo table
  const double* tabs;  // [nT+1][ns] s per mass
  const double* W;     // [ns] molecular weights (wm_tab)
  const double* rm;    // [ns] species weights from the elements' atomic weights
  int Th_min, Th_max;
  int redo;  // nonzero: solve only the states whose status is nonzero, from the default guess, adding
             // to their iteration counts (the second pass of a warm-started batch)
  int max_iter;        // iterations before status 1 (50 in the reference, Lib_CEA_solver.f90)
  double tol;          // 0: the reference's convergence tests; > 0: every test at tol (tightened mode)
  const int* idx;      // non-null: thread t solves state idx[t], for t < *count (a compacted redo list)
  const int* count;
};

}  // namespace flint_cea

