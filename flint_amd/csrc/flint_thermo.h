// Kernels of the conservative <-> primitive variable maps (thermo.hip).
#pragma once
#include <hip/hip_runtime.h>

#define FLINT_THERMO_MAXNS 64

// cons2prim over n cells (cons, prim [ns+nd+1][ld]; T the Newton start; T_out, iters, status may be
// null).  tabh [T][ns][2] (cp, h) per mass with rows Th_min..Th_max, W [ns] molecular weights.
hipError_t flint_thermo_cons2prim(const double* tabh, int Th_min, int Th_max, int ns, const double* W, long n,
                                  long ld, int nd, const double* cons, const double* T, double* prim, double* T_out,
                                  int* iters, int* status, hipStream_t s);

// prim2cons over n cells (status may be null)
hipError_t flint_thermo_prim2cons(const double* tabh, int Th_min, int Th_max, int ns, const double* W, long n,
                                  long ld, int nd, const double* prim, double* cons, int* status, hipStream_t s);
