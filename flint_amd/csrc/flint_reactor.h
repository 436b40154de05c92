// Device-side rate program and options of the batched reactor kernel (reactor.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace flint_reactor {

// Rate program in device memory (built from the host program a generated mechanism registers).
struct Prog {
  int ns, nr, nms;
  // The program encoded for the integrator (flint_reactor_host.cpp encode_rhs, reactor.hip):
  int R;                  // reaction rounds: slot 32*round + lane
  const int4* rrec;       // [R*32] {kind | (mset + 1) << 3 (kind 7: empty), column, reaction, factor slots}
  const float* effT;      // [ns][nms] efficiency of species j in third-body set s (0 if absent)
  int C, CW, MAXP;        // production-sum chunk rounds, chunk width, most chunks of one species
  const int2* cterm;      // [C][CW][32] {reaction (nr: empty), coefficient bits} of chunk (round, lane)
  const int* kpart;       // [MAXP][32] partial-sum slots (32*round + lane) of species k in order, -1 none
  const int4* rprod;      // [R*32][2] net-production terms of the slot's reaction: {species (5 bits each),
                          // count | (dq/dM slot + 1) << 3, coefficient bits x 6}
  int nmr;                // reactions with a concentration M (the dq/dM slots)
  int ntb;                // entries of tbe
  const int4* tbe;        // [ntb] terms whose reaction has M, by species: {species, coefficient bits, mset, dq slot}
  const int* tbptr;       // [ns+1] first tbe entry of species k
  const double* W;        // [ns]
  const double* tabk;
  int Tk_min, Tk_max, wk;
  const double* tabh;
  int Th_min, Th_max;
};

struct Opts {
  int mode;  // 0 constant volume (rho_k, T), 1 constant pressure (Y_k, T)
  double rtol, atol, t_end;
  int max_steps;
  int newton;  // Newton policy: 0 as scipy's BDF, 1 as SUNDIALS CVODE (reactor.hip)
};

}  // namespace flint_reactor

size_t flint_reactor_team_bytes(const flint_reactor::Prog& P);
size_t flint_reactor_shared_bytes(const flint_reactor::Prog& P);  // block-shared program in LDS
hipError_t flint_reactor_launch(const flint_reactor::Prog& P, const flint_reactor::Opts& O, long n, long ld,
                                double* X, const double* pres, int* stats, int* status, int* counter,
                                int teams_per_block, int blocks, double* jac_scratch, const int* order, hipStream_t s);
// Work order by descending |key[cell]| (cells idx[0..n) or, with idx null, 0..n-1) into out[n]; hist
// is scratch of FLINT_ORDER_BUCKETS ints.  Kernels only (no memset node: graph-safe).
constexpr int FLINT_ORDER_BUCKETS = 4096;
hipError_t flint_reactor_order(long n, const double* key, const int* idx, int* out, int* hist, hipStream_t s);

// Allocation-free reactor launches (flint_reactor_advance_noalloc): the per-team Jacobian slots
// and the work counter live in a file-scope device array of reactor.hip with room for
// FLINT_SCRATCH_TEAMS teams.  flint_reactor_scratch returns that array's address on the current
// device (looked up once per device) as (counter, Jacobian slots).
constexpr int FLINT_SCRATCH_TEAMS = 2048;
hipError_t flint_reactor_scratch(int** counter, double** jac_slots);

hipError_t flint_reactor_fill_int(int* p, long n, int v, hipStream_t s);
hipError_t flint_reactor_verify_launch(const flint_reactor::Prog& P, long n, long ld, double* X, const double* Xp,
                                       long ldp, const double* F0, long ldf, const double* pres, double dt,
                                       double tol_y, double tol_t, int commit, int* flag, int* idx, int* count,
                                       double* dmax, int teams_per_block, int blocks, double* jac_scratch,
                                       hipStream_t s);
hipError_t flint_reactor_jac_launch(const flint_reactor::Prog& P, int mode, long n, long ld, const double* X,
                                    const double* pres, double* F, double* J, int teams_per_block, int blocks,
                                    double* jac_scratch, hipStream_t s);

// Host copy of a generated mechanism's rate program (emitted by flint_amd/codegen.py).
struct flint_host_prog {
  int ns, nr, nms, nterm, nmterm;
  const int* rx_kind;
  const int* rx_col;
  const int* rx_mset;
  const int* rx_fac;
  const int* wd_ptr;
  const int* wd_rx;
  const double* wd_coef;
  const int* ms_ptr;
  const int* ms_sp;
  const double* ms_eff;
};
