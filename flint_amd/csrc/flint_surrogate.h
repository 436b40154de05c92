// Device data of the learned reactor step (surrogate.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace flint_sur {

struct SurProg {
  int ns, ne;
  int nin;             // network inputs: ns + 2 (state) or 2 ns + 3 (state and rates)
  const double* comp;  // [ne][ns] atoms of element e in species k
  const double* W;     // [ns] molecular weights
  const double* elw;   // [ne] atomic weights (element mass fractions for the residual test)
  const double* tabh;  // [nT+1][ns][2] (cp, h) per mass
  int Th_min, Th_max;
  const double* scal;  // in_mean[nin], in_scale[nin], out_mean[ns+1], out_scale[ns+1]
  const double* F;     // [ns+1][ldF] right-hand side dX/dt at the old state
  long ldF;
  double dt, ymin, yfloor, tfloor, max_dT, max_res, max_Tmis;
};

}  // namespace flint_sur

hipError_t flint_sur_features(const flint_sur::SurProg& P, long n, long ld, const double* X, const double* p, float* F,
                              hipStream_t s);
hipError_t flint_sur_apply(const flint_sur::SurProg& P, long n, long ld, const double* X, double* Xp, const float* O,
                           int* flag, hipStream_t s);
hipError_t flint_sur_zero(int* p, int n, hipStream_t s);
hipError_t flint_sur_gather(int dir, int nrow, long ld, double* X, long nf, long ldg, double* Xg, const int* idx,
                            const double* p, double* pg, hipStream_t s);
hipError_t flint_sur_scatter_int(int nrow, long ld, int* dst, long nf, long ldg, const int* src, const int* idx,
                                 hipStream_t s);
