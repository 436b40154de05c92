// Device-side rate program and options of the batched reactor kernel (reactor.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace flint_reactor {

// Rate program in device memory (built from the host program a generated mechanism registers).
struct Prog {
  int ns, nr, nms;
  const int* rx_kind;     // [nr] 0 Arrhenius, 1 three-body, 2 Troe, 3 Lindemann
  const int* rx_col;      // [nr] first column of the reaction in a kinetics table row
  const int* rx_mset;     // [nr] third-body sum used, -1 none
  const int* rx_fac;      // [nr*8] forward then reverse factors: species | power << 8, -1 unused
  const int* wd_ptr;      // [ns+1] net-production terms of species k
  const int* wd_rx;       // reaction of each term, in reaction order
  const double* wd_coef;  // float32-valued stoichiometric difference of each term
  const int* ms_ptr;      // [nms+1] terms of each third-body sum
  const int* ms_sp;
  const double* ms_eff;   // efficiency (1 for a bare term)
  const double* effd;     // [nms*32] dense efficiencies, 0 where a species is absent
  // The same program re-encoded for the right-hand side of the integrator (see reactor.hip):
  int R;                  // reaction rounds: slot 32*round + lane
  const int4* rrec;       // [R*32] {kind | (mset + 1) << 3 (kind 7: empty), column, reaction, factor slots}
  const float* effT;      // [ns][nms] efficiency of species j in third-body set s (0 if absent)
  int C, CW, MAXP;        // production-sum chunk rounds, chunk width, most chunks of one species
  const int2* cterm;      // [C][CW][32] {reaction (nr: empty), coefficient bits} of chunk (round, lane)
  const int* kpart;       // [MAXP][32] partial-sum slots (32*round + lane) of species k in order, -1 none
  const int4* jterm;      // [nterm] terms of wd_ptr order: {reaction, coefficient bits, mset, factor slots}
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
};

}  // namespace flint_reactor

size_t flint_reactor_team_bytes(const flint_reactor::Prog& P);
size_t flint_reactor_shared_bytes(const flint_reactor::Prog& P);  // block-shared program in LDS
hipError_t flint_reactor_launch(const flint_reactor::Prog& P, const flint_reactor::Opts& O, long n, long ld,
                                double* X, const double* pres, int* stats, int* status, int* counter,
                                int teams_per_block, int blocks, double* jac_scratch, hipStream_t s);

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
