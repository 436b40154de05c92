// Batched stiff 0D reactor integration on gfx950: one persistent launch, one 32-lane team per
// cell, work-stealing refill from an atomic cell counter.
//
// Method: the variable-order (1..5) quasi-constant step BDF of Shampine & Reichelt ("The MATLAB
// ODE Suite", SIAM J. Sci. Comput. 18, 1997; kappa = 0), modified Newton with the analytic
// Jacobian refreshed only on Newton failure, RMS error norm of e/(atol + rtol|y|).  Two Newton
// policies (Opts::newton):
//   0  scipy's: at most 4 iterations, the convergence rate measured within the step, converged
//      when rate/(1-rate)|dy| < newton_tol; I - cJ refactored whenever h or the order changes;
//   1  SUNDIALS CVODE's (the BDF behind FLINT's OSlo 'cvode' choice, test-batchF.f90:22-26): at
//      most 3 iterations, the rate estimate carried over between steps (reset to 1 with a new
//      matrix, decayed by 0.3 per iteration), converged when |dy| min(1, rate)/(q+1) <= 0.005,
//      diverged when |dy| > 2|dy_prev|; the factorised I - c_lu J kept while |c/c_lu - 1| <= 0.3
//      and younger than 20 converged attempts, its corrections scaled by 2/(1 + c/c_lu).
// Either is the same algorithm, decision for decision, as the CPU checker in oracle/oracle_bdf.c.  FLINT itself
// integrates through the external OSlo library (test-batchF.f90:81,105), absent from the
// reference checkout.
//
// Layout of one team (lanes l = 0..31 of a half wavefront):
//   lane k < ns owns species k, lane ns owns T: the state component and the 8-row BDF difference
//   column D[.][k] live in VGPRs and LDS; row k of the Jacobian is accumulated in LDS, kept in a
//   per-team global slot, and row k of the inverse of I - cJ sits in LDS (stride 33, conflict-free).  The inverse comes from an
//   in-place Gauss-Jordan elimination with partial pivoting (team argmax, pivot row read as an LDS
//   broadcast), so that each Newton correction is one matrix-vector product instead of two
//   32-step triangular sweeps.
//   The right-hand side is evaluated reaction-parallel: reactions sit in rounds of 32, grouped by
//   kind so that a round runs one code path, each lane reading its reaction's kinetics-table
//   entries for the cell's temperature; net production rates are summed in chunks of at most 16
//   terms (in reaction order) spread over the lanes, then per species over its chunks.  Every
//   program load of a phase is independent of the others, so a phase costs one memory latency
//   instead of one per term.
//   Modes: 0 = constant volume, state (rho_k, T), rhs_native (Lib_Chemistry_rhs.f90:15-52);
//          1 = constant pressure, state (Y_k, T) at fixed p.
#include <mutex>
#include <vector>

#include "flint_device.h"
#include "flint_reactor.h"

namespace flint_reactor {

constexpr int TEAM = 32;  // component lanes of a cell: lane k < ns species k, lane ns temperature
constexpr int WAVE = 64;  // a cell's team is a whole wavefront: two 32-lane halves
constexpr int MAXO = 5;
constexpr int NEWTON_MAXITER = 4;
constexpr int CW = 16;  // production-sum chunk width (the host encoder uses the same value)
constexpr double MIN_FACTOR = 0.2, MAX_FACTOR = 10.0;
// Newton policy 1 (CVODE's): iterations, matrix-reuse bounds, rate decay, divergence ratio, and
// the convergence coefficient (CVODE's 0.1 tightened, see oracle/oracle_bdf.c)
constexpr int CV_MAXCOR = 3, CV_MSBP = 20;
constexpr double CV_DGMAX = 0.3, CV_CRDOWN = 0.3, CV_RDIV = 2.0, CV_NLSCOEF = 0.005;

enum { ST_STEPS, ST_REJECT, ST_NFAIL, ST_RHS, ST_JAC, ST_LU, ST_NEWTON, ST_ORDER };

// Cross-lane exchanges inside a 32-lane team without the LDS crossbar where possible: DPP quad
// permutations (lane ^ 1, lane ^ 2) and row mirrors (lane i <-> 7 - i within 8, i <-> 15 - i within
// 16), and for lane ^ 16 a permlane16 swap (sums, maxima) or a bit-mode swizzle (where the partner
// must be told apart), all verified on gfx950.  Each level pairs disjoint groups, so five levels
// combine all 32 lanes.
template <int CTRL>
__device__ __forceinline__ double dpp(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
template <int CTRL>
__device__ __forceinline__ int dppi(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}
__device__ __forceinline__ double swz16(double v) {
  const int lo = __builtin_amdgcn_ds_swizzle(__double2loint(v), 0x401F);
  const int hi = __builtin_amdgcn_ds_swizzle(__double2hiint(v), 0x401F);
  return __hiloint2double(hi, lo);
}
// (this lane's value, lane ^ 16's value) in some order, the same two values on both lanes: a
// v_permlane16_swap of v with itself (lanes of even 16-lane rows receive their partner in the
// second value, lanes of odd rows in the first; verified on gfx950).  A VALU exchange, where the
// swizzle goes through the LDS crossbar; sums and maxima need no order.
__device__ __forceinline__ double2 pair16(double v) {
  const auto a = __builtin_amdgcn_permlane16_swap((unsigned)__double2loint(v), (unsigned)__double2loint(v), false, false);
  const auto b = __builtin_amdgcn_permlane16_swap((unsigned)__double2hiint(v), (unsigned)__double2hiint(v), false, false);
  return make_double2(__hiloint2double((int)b[0], (int)a[0]), __hiloint2double((int)b[1], (int)a[1]));
}
constexpr int DPP_X1 = 0xB1, DPP_X2 = 0x4E, DPP_HMIR = 0x141, DPP_MIR = 0x140;

// Team sum, bit-identical on every lane: each level adds the same two rounded values (a + b ==
// b + a exactly).  The register barrier (an empty asm the compiler cannot see through, not volatile
// so that it does not bound the instruction scheduler) rounds v first; without it the compiler may contract a
// caller's product into the first addition, fma(x_i, y_i, partner's x_j y_j), which differs between
// partner lanes, and step sizes, orders and branch decisions would then diverge inside one team.
// Every call site runs with all 32 lanes of the team active.
__device__ __forceinline__ double tsum(double v) {
  asm("" : "+v"(v));
  v += dpp<DPP_X1>(v);
  v += dpp<DPP_X2>(v);
  v += dpp<DPP_HMIR>(v);
  v += dpp<DPP_MIR>(v);
  const double2 q = pair16(v);
  return q.x + q.y;  // v + (lane ^ 16's v), in either order the same bits
}

// Team argmax of (best, lane), ties to the lower lane: the same pairing levels as tsum, with
// branch-free selects.
__device__ __forceinline__ void targmax(double& best, int& bl) {
#define FLINT_ARGMAX_LEVEL(OPD, OPI)                              \
  {                                                               \
    const double ob = OPD;                                        \
    const int ol = OPI;                                           \
    const bool take = (ob > best) | ((ob == best) & (ol < bl));   \
    best = take ? ob : best;                                      \
    bl = take ? ol : bl;                                          \
  }
  FLINT_ARGMAX_LEVEL(dpp<DPP_X1>(best), dppi<DPP_X1>(bl))
  FLINT_ARGMAX_LEVEL(dpp<DPP_X2>(best), dppi<DPP_X2>(bl))
  FLINT_ARGMAX_LEVEL(dpp<DPP_HMIR>(best), dppi<DPP_HMIR>(bl))
  FLINT_ARGMAX_LEVEL(dpp<DPP_MIR>(best), dppi<DPP_MIR>(bl))
  FLINT_ARGMAX_LEVEL(swz16(best), __builtin_amdgcn_ds_swizzle(bl, 0x401F))
#undef FLINT_ARGMAX_LEVEL
}

// Pivot search of the eliminations: the team lane with the largest |v| among the lanes with
// eligible set, as one 32-bit unsigned maximum.  The key keeps |v|'s exponent and its 15 leading
// mantissa bits (v's high word with the low five bits cleared) above 32 - lane, so that values
// within a factor 1 + 2^-15 of each other count as equal and go to the lower lane; an ineligible
// lane has key 0.  Every lane ends with the same key, and the decoded lane.
__device__ __forceinline__ int tpivot(double v, bool eligible) {
  const int lane = threadIdx.x & 31;
  const unsigned hi = (unsigned)__double2hiint(v) & 0x7fffffffu;
  unsigned key = eligible ? (hi & ~31u) + (unsigned)(32 - lane) : 0u;
  key = max(key, (unsigned)dppi<DPP_X1>((int)key));
  key = max(key, (unsigned)dppi<DPP_X2>((int)key));
  key = max(key, (unsigned)dppi<DPP_HMIR>((int)key));
  key = max(key, (unsigned)dppi<DPP_MIR>((int)key));
  const auto kp = __builtin_amdgcn_permlane16_swap(key, key, false, false);
  key = max((unsigned)kp[0], (unsigned)kp[1]);
  return 32 - (int)(key & 31u) - ((key & 31u) == 0u ? 32 : 0);
}

// Both halves of the wavefront hold the same component values (component lane k = lane k and
// lane 32 + k) and run the same control flow; they split the reaction, chunk, elimination and
// solve work.  half(): which half this lane is in.
__device__ __forceinline__ int half() { return (threadIdx.x >> 5) & 1; }
// Component src of a value both halves hold alike: a scalar read of lane src.
__device__ __forceinline__ double tbc(double v, int src) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
  return __hiloint2double(hi, lo);
}
// (lower half's value, upper half's value) of component lane l at both lanes l and 32 + l
// (v_permlane32_swap; verified on gfx950).
__device__ __forceinline__ double2 halves(double v) {
  const auto a = __builtin_amdgcn_permlane32_swap((unsigned)__double2loint(v), (unsigned)__double2loint(v), false, false);
  const auto b = __builtin_amdgcn_permlane32_swap((unsigned)__double2hiint(v), (unsigned)__double2hiint(v), false, false);
  return make_double2(__hiloint2double((int)b[0], (int)a[0]), __hiloint2double((int)b[1], (int)a[1]));
}
// All LDS traffic of a team stays inside one wavefront: a wavefront-scope fence orders it.
__device__ __forceinline__ void tsync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// The lane's index in the wavefront, read where it is used.  threadIdx.x is live through the whole
// kernel, so lane-derived LDS addresses computed from it are loop invariants that the compiler
// hoists out of the integrator's loop and then has to keep: across an evaluation they spilled to
// scratch.  A fresh read (an opaque v_mbcnt pair) makes them cheap values computed again per call.
__device__ __forceinline__ int wave_lane() {
  int l;
  asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(l));
  return l;
}

struct Lds {
  // block-shared copy of the program (read-only after the kernel's first barrier)
  const int4* rrec;   // [R*32] reaction records
  const float* effT;  // [ns][nms] third-body efficiencies
  const int* kpart;   // [MAXP][32] chunk slots of each species
  const int2* cterm;  // [C][CW][32] production-sum chunk terms
  const double* W;    // [32] molar masses (1 beyond ns)
  const double* Rk;   // [32] specific gas constants 8314.51 / W_k (0 beyond ns)
  const double* rcp;  // [32] 1/k for k >= 1, and 1/(ns+1) at 0
  const double* Wi;   // [32] 1 / W_k (1 beyond ns)
  const int4* rprod;  // [R*32][2] net-production terms of each slot's reaction (Jacobian scatter)
  const int4* tbe;    // [ntb] terms with a concentration M, by species
  const int* tbptr;   // [ns+1]
  // per team
  double* c;    // [32] concentrations; slot 31 (never a species) holds 1 for empty factors
  double* Mv;   // [nms] third-body sums
  double* aux;  // [32] per-lane shared values (Y_k, e_k or h_k)
  double* g;    // [32] clip derivative per column
  double* q;    // [nr + 1] rates of progress; q[nr] = 0 for empty sum terms
  double* part; // [C*32] partial production sums of the chunks
  double* vb;   // [32] right-hand side of a Newton correction, in pivot order
  double* vx;   // [32] its solution
  double* D;    // [8][32] BDF difference array (row j, component k); kept in LDS to spare registers
  double* lu;   // [32][33] inverse of I - cJ (row = lane, padded); Jacobian rows while one is built
  double* dq;   // [nmr] dq/dM of the reactions with a concentration M (while a Jacobian is built)
  double* sx;   // [8] team scalars: [0] the Newton correction scale of the current matrix
};

__host__ __device__ inline size_t team_lds_bytes(const Prog& P) {
  return (size_t)(32 * 5 + 8 * 32 + (P.nms > 0 ? P.nms : 1) + (size_t)P.nr + 1 + (size_t)P.C * 32 + 32 * 33 +
                  (P.nmr > 0 ? P.nmr : 1) + 8) *
         sizeof(double);
}

// Block-shared program (16-byte multiples): records, efficiencies, chunk slots, chunk terms, molar
// masses, gas constants and small reciprocals, and the Jacobian's term lists.  Everything the right-hand side reads besides the state and the
// tables sits here, so that an evaluation waits on global memory only for its table rows.
struct SharedLayout {
  size_t eff, kp, ct, w, rp, tb, tbp, total;
};
__host__ __device__ inline SharedLayout shared_layout(const Prog& P) {
  SharedLayout o;
  o.eff = (size_t)P.R * 32 * 16;
  o.kp = o.eff + ((((size_t)P.ns * (P.nms > 0 ? P.nms : 1) * 4) + 15) & ~size_t(15));
  o.ct = o.kp + ((((size_t)P.MAXP * 32 * 4) + 15) & ~size_t(15));
  o.w = o.ct + (size_t)P.C * CW * TEAM * 8;
  o.rp = o.w + 4 * TEAM * 8;
  o.tb = o.rp + (size_t)P.R * 32 * 32;
  o.tbp = o.tb + (size_t)(P.ntb > 0 ? P.ntb : 1) * 16;
  o.total = o.tbp + ((((size_t)P.ns + 1) * 4 + 15) & ~size_t(15));
  return o;
}
__host__ __device__ inline size_t shared_prog_bytes(const Prog& P) { return shared_layout(P).total; }

// Copies the program into the block's LDS (all threads of the block; a block barrier follows).
__device__ inline void load_shared_prog(const Prog& P, char* base) {
  const SharedLayout o = shared_layout(P);
  int4* rec = reinterpret_cast<int4*>(base);
  for (int i = threadIdx.x; i < P.R * 32; i += blockDim.x) rec[i] = P.rrec[i];
  float* eff = reinterpret_cast<float*>(base + o.eff);
  for (int i = threadIdx.x; i < P.ns * P.nms; i += blockDim.x) eff[i] = P.effT[i];
  int* kp = reinterpret_cast<int*>(base + o.kp);
  for (int i = threadIdx.x; i < P.MAXP * 32; i += blockDim.x) kp[i] = P.kpart[i];
  int2* ct = reinterpret_cast<int2*>(base + o.ct);
  for (int i = threadIdx.x; i < P.C * CW * TEAM; i += blockDim.x) ct[i] = P.cterm[i];
  double* w = reinterpret_cast<double*>(base + o.w);
  for (int i = threadIdx.x; i < TEAM; i += blockDim.x) {
    const bool sp = i < P.ns;
    w[i] = sp ? P.W[i] : 1.0;
    w[TEAM + i] = sp ? 8314.51 / P.W[i] : 0.0;
    // reciprocals the step control multiplies by: 1/k (slot k >= 1) and 1/nz (slot 0)
    w[2 * TEAM + i] = i == 0 ? 1.0 / (P.ns + 1) : 1.0 / i;
    w[3 * TEAM + i] = sp ? 1.0 / P.W[i] : 1.0;
  }
  int4* rp = reinterpret_cast<int4*>(base + o.rp);
  for (int i = threadIdx.x; i < P.R * 32 * 2; i += blockDim.x) rp[i] = P.rprod[i];
  int4* tb = reinterpret_cast<int4*>(base + o.tb);
  for (int i = threadIdx.x; i < (P.ntb > 0 ? P.ntb : 1); i += blockDim.x) tb[i] = P.tbe[i];
  int* tbp = reinterpret_cast<int*>(base + o.tbp);
  for (int i = threadIdx.x; i <= P.ns; i += blockDim.x) tbp[i] = P.tbptr[i];
}

__device__ inline Lds carve(char* shared, char* base, const Prog& P) {
  Lds L;
  const SharedLayout o = shared_layout(P);
  L.rrec = reinterpret_cast<const int4*>(shared);
  L.effT = reinterpret_cast<const float*>(shared + o.eff);
  L.kpart = reinterpret_cast<const int*>(shared + o.kp);
  L.cterm = reinterpret_cast<const int2*>(shared + o.ct);
  L.W = reinterpret_cast<const double*>(shared + o.w);
  L.Rk = L.W + TEAM;
  L.rcp = L.W + 2 * TEAM;
  L.Wi = L.W + 3 * TEAM;
  L.rprod = reinterpret_cast<const int4*>(shared + o.rp);
  L.tbe = reinterpret_cast<const int4*>(shared + o.tb);
  L.tbptr = reinterpret_cast<const int*>(shared + o.tbp);
  double* d = reinterpret_cast<double*>(base);
  L.c = d; d += 32;
  L.aux = d; d += 32;
  L.g = d; d += 32;
  L.vb = d; d += 32;
  L.vx = d; d += 32;
  L.D = d; d += 8 * 32;
  L.Mv = d; d += (P.nms > 0 ? P.nms : 1);
  L.q = d; d += P.nr + 1;
  L.part = d; d += P.C * 32;
  L.lu = d; d += 32 * 33;
  L.dq = d; d += (P.nmr > 0 ? P.nmr : 1);
  L.sx = d;
  return L;
}

struct RhsOut {
  bool ok;                  // T inside the tables (team-uniform)
  double F;                 // this lane's RHS component
  double xc, g, w, rho;     // clipped state, clip derivative, wdot_k, density
  double cp, h, dcp, dh;    // thermo of species k at T
  double Rmix, rho_cv, cpm; // team-uniform sums
};

// No-return add to a Jacobian entry in the team's LDS (ds_add_f64).
__device__ __forceinline__ void jadd(double* p, double v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}

// The kinetics-table values record rec reads: rows T_i (a0, b0) and T_i + 1 (a1, b1), columns col
// to col + 3 as its kind uses them.  Loaded ahead for the first round of rates, whose latency then
// overlaps the third-body sums.
struct Tab {
  double2 a0, a1, b0, b1;
};
__device__ __forceinline__ Tab tab_load(const int4 rec, const double* __restrict__ rk0, const double* __restrict__ rk1) {
  Tab t;
  const int kind = rec.x & 7;
  const int col = rec.y;
  t.a0 = t.a1 = t.b0 = t.b1 = make_double2(0.0, 0.0);
  if (kind <= 1) {
    t.a0 = ld2(rk0 + col);
    t.a1 = ld2(rk1 + col);
  } else if (kind == 2) {
    t.a0 = ld2(rk0 + col);
    t.b0 = ld2(rk0 + col + 2);
    t.a1 = ld2(rk1 + col);
    t.b1 = ld2(rk1 + col + 2);
  } else if (kind == 3) {
    t.a0 = make_double2(rk0[col], rk0[col + 1]);
    t.b0.x = rk0[col + 2];
    t.a1 = make_double2(rk1[col], rk1[col + 1]);
    t.b1.x = rk1[col + 2];
  }
  return t;
}

// Rates of progress of the reaction in record rec (kinds: 0 Arrhenius, 1 three-body, 2 Troe,
// 3 Lindemann), Lib_Chemistry_falloff.f90:12-80 for the falloff forms.  With JAC the reaction's
// derivatives are scattered into the Jacobian rows in L.lu (zeroed by the caller): slot is the
// record's index, which also indexes its net-production terms L.rprod.  With PRE the table values
// come from tv (tab_load) instead of being loaded here.
template <bool JAC, bool PRE = false>
__device__ __forceinline__ void reaction(const int4 rec, const double* __restrict__ rk0, const double* __restrict__ rk1,
                                         double Td, const Lds& L, int slot, const Tab tv = Tab{}) {
  const int kind = rec.x & 7;
  const int ms = (rec.x >> 3) - 1;
  const int col = rec.y;
  const int r = rec.z;
  const unsigned fs = (unsigned)rec.w;
  const double M = ms >= 0 ? L.Mv[ms] : 1.0;
  const double Pf = L.c[fs & 31] * L.c[(fs >> 5) & 31] * L.c[(fs >> 10) & 31];
  const double Pb = L.c[(fs >> 15) & 31] * L.c[(fs >> 20) & 31] * L.c[(fs >> 25) & 31];
  double k1, k2, dk1T = 0.0, dk2T = 0.0, dqM = 0.0;
  if (kind <= 1) {
    const double2 a0 = PRE ? tv.a0 : ld2(rk0 + col), a1 = PRE ? tv.a1 : ld2(rk1 + col);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double Mf = kind == 1 ? M : 1.0;
    k1 = kf * Mf;
    k2 = kb * Mf;
    if (JAC) {
      dk1T = (a1.x - a0.x) * Mf;
      dk2T = (a1.y - a0.y) * Mf;
      if (kind == 1) dqM = kf * Pf - kb * Pb;
    }
  } else {
    const bool troe = kind == 2;
    double kinf, k0, kc, Fc = 1.0, skinf, sk0, skc, sFc = 0.0;
    if (troe) {
      const double2 x0 = PRE ? tv.a0 : ld2(rk0 + col), y0 = PRE ? tv.b0 : ld2(rk0 + col + 2);
      const double2 x1 = PRE ? tv.a1 : ld2(rk1 + col), y1 = PRE ? tv.b1 : ld2(rk1 + col + 2);
      kinf = x0.x + (x1.x - x0.x) * Td; skinf = x1.x - x0.x;
      k0 = x0.y + (x1.y - x0.y) * Td; sk0 = x1.y - x0.y;
      kc = y0.x + (y1.x - y0.x) * Td; skc = y1.x - y0.x;
      Fc = y0.y + (y1.y - y0.y) * Td; sFc = y1.y - y0.y;
    } else {
      const double u0 = PRE ? tv.a0.x : rk0[col], u1 = PRE ? tv.a1.x : rk1[col];
      const double v0 = PRE ? tv.a0.y : rk0[col + 1], v1 = PRE ? tv.a1.y : rk1[col + 1];
      const double w0 = PRE ? tv.b0.x : rk0[col + 2], w1 = PRE ? tv.b1.x : rk1[col + 2];
      kinf = u0 + (u1 - u0) * Td; skinf = u1 - u0;
      k0 = v0 + (v1 - v0) * Td; sk0 = v1 - v0;
      kc = w0 + (w1 - w0) * Td; skc = w1 - w0;
    }
    const double Pr = k0 * M / kinf;
    const double gg = Pr / (1.0 + Pr);
    double Fb = 1.0, lFc = 0.0, f1 = 0.0, cc = 0.0, nn = 0.0, lPr = 0.0;
    if (troe) {
      lFc = log10(Fc);
      lPr = log10(Pr);
      cc = -0.4 - 0.67 * lFc;
      nn = 0.75 - 1.27 * lFc;
      f1 = (lPr + cc) / (nn - 0.14 * (lPr + cc));
      Fb = exp10(lFc / (1.0 + f1 * f1));
    }
    k1 = kinf * gg * Fb;
    k2 = k1 / kc;
    if (JAC) {
      const double dg = 1.0 / ((1.0 + Pr) * (1.0 + Pr));
      double dlnF_dPr = 0.0, dlnF_dFc = 0.0;
      if (troe) {
        const double u = lPr + cc, den = nn - 0.14 * u, qq = 1.0 + f1 * f1;
        dlnF_dPr = -2.0 * lFc * f1 * (nn / (den * den)) / (qq * qq * Pr);
        const double df1_dL = (-0.67 * nn + 1.27 * u) / (den * den);
        dlnF_dFc = (1.0 / qq - lFc * 2.0 * f1 * df1_dL / (qq * qq)) / Fc;
      }
      const double dk1_dPr = kinf * Fb * (dg + gg * dlnF_dPr);
      const double dPr_dT = M * (sk0 * kinf - k0 * skinf) / (kinf * kinf);
      dk1T = skinf * gg * Fb + dk1_dPr * dPr_dT + k1 * dlnF_dFc * sFc;
      dk2T = dk1T / kc - k1 * skc / (kc * kc);
      const double dk1M = dk1_dPr * k0 / kinf;
      dqM = dk1M * Pf - (dk1M / kc) * Pb;
    }
  }
  L.q[r] = k1 * Pf - k2 * Pb;
  if (JAC) {
    // dq/dc at each factor slot (a repeated species gets one term per slot, the power rule; slot
    // 31 is the constant 1 and has no column), times each net-production coefficient nu_kr, added
    // to row k.  Several reactions of one round may add to one entry, hence the atomic adds.
    // dq/dT goes to the row's spare column 32, dq/dM to the reaction's slot in L.dq.
    const int a0 = fs & 31, a1 = (fs >> 5) & 31, a2 = (fs >> 10) & 31;
    const int b0 = (fs >> 15) & 31, b1 = (fs >> 20) & 31, b2 = (fs >> 25) & 31;
    const double ca0 = L.c[a0], ca1 = L.c[a1], ca2 = L.c[a2];
    const double cb0 = L.c[b0], cb1 = L.c[b1], cb2 = L.c[b2];
    const double da0 = k1 * (ca1 * ca2), da1 = k1 * (ca0 * ca2), da2 = k1 * (ca0 * ca1);
    const double db0 = k2 * (cb1 * cb2), db1 = k2 * (cb0 * cb2), db2 = k2 * (cb0 * cb1);
    const double dqT = dk1T * Pf - dk2T * Pb;
    const int4 p0 = L.rprod[2 * slot], p1 = L.rprod[2 * slot + 1];
    const int cnt = p0.y & 7, dqs = (p0.y >> 3) - 1;
    if (dqs >= 0) L.dq[dqs] = dqM;
    const int cbits[6] = {p0.z, p0.w, p1.x, p1.y, p1.z, p1.w};
#pragma unroll
    for (int t = 0; t < 6; ++t) {
      if (t < cnt) {
        const int k = ((unsigned)p0.x >> (5 * t)) & 31;
        const double nu = (double)__int_as_float(cbits[t]);
        double* row = L.lu + k * 33;
        if (a0 != 31) jadd(row + a0, nu * da0);
        if (a1 != 31) jadd(row + a1, nu * da1);
        if (a2 != 31) jadd(row + a2, nu * da2);
        if (b0 != 31) jadd(row + b0, -(nu * db0));
        if (b1 != 31) jadd(row + b1, -(nu * db1));
        if (b2 != 31) jadd(row + b2, -(nu * db2));
        jadd(row + 32, nu * dqT);
      }
    }
  }
}

// Record s of the program, or an empty one (kind 7) past the end: the read unguarded (index
// clamped into the table), the guard a select.
__device__ __forceinline__ int4 rec_at(const Lds& L, int s, int nslot) {
  const int4 r = L.rrec[min(s, nslot - 1)];
  return s < nslot ? r : make_int4(7, 0, 0, 0);
}

// Team right-hand side; o.ok is false (team-uniform) when T leaves the tables.  With jac the
// reactions' derivatives are accumulated in the Jacobian rows in LDS as well (team_jac finishes
// them); the rows replace the inverse of I - cJ there, which a new Jacobian always renews.  jac is
// a template argument so that the plain evaluation, nearly all of the integrator's, is compiled
// apart from the scatter and does not carry its register needs.
template <bool jac>
__device__ __forceinline__ RhsOut team_rhs(const Prog& P, int mode, double p, double x, const Lds& L) {
  RhsOut o;
  o.ok = false;
  const int wlf = wave_lane();
  const int lane = wlf & (TEAM - 1);
  const int ns = P.ns;
  const double T = tbc(x, ns);
  if (!(T >= (double)P.Th_min && T < (double)P.Th_max && T >= (double)P.Tk_min && T < (double)P.Tk_max)) return o;
  const bool sp = lane < ns;
  const double Rk = L.Rk[lane];
  // constant volume clips rho at 0 as rhs_native does (Lib_Chemistry_rhs.f90:37-39); the
  // constant-pressure model keeps Y as it is, so its right-hand side stays smooth through Y = 0
  const bool clip = mode == 0 && x < 0.0;
  o.xc = sp ? (clip ? 0.0 : x) : 0.0;
  o.g = (sp && !clip) ? 1.0 : 0.0;
  double roi;
  if (mode == 0) {
    roi = o.xc;
    o.rho = 1.0;
    o.Rmix = 0.0;
  } else {
    o.Rmix = tsum(o.xc * Rk);
    o.rho = p / (o.Rmix * T);
    roi = o.rho * o.xc;
  }
  const double wi = L.Wi[lane];  // read unguarded (32 entries), used under the guard
  L.c[lane] = sp ? roi * wi : (lane == TEAM - 1 ? 1.0 : 0.0);
  if (lane == 0) L.q[P.nr] = 0.0;
  if (jac) {
    const int wl0 = wlf;
    for (int i = wl0; i < 32 * 33; i += WAVE) L.lu[i] = 0.0;
  }
  const int T_i = (int)T;
  const double Td = T - (double)T_i;
  const double* __restrict__ rk0 = P.tabk + (long)(T_i - P.Tk_min) * P.wk;
  const double* __restrict__ rk1 = rk0 + P.wk;
  // the first round's records (and, for a plain evaluation, their table values) do not depend on
  // the sums below: issue them now
  const int wl = wlf;
  const int nslot = P.R * TEAM;
  int4 rec = rec_at(L, wl, nslot);
  Tab tv0;
  if (!jac) tv0 = tab_load(rec, rk0, rk1);  // the first round's table values, ahead
  tsync();
  // third-body sums M_s = sum_j eff_sj c_j.  Up to 16 sets: set s = wl & 15 in four lane groups
  // g = wl / 16, group g summing the species j = g (mod 4), the four partial sums then added as
  // (g0 + g1) + (g2 + g3) (the same value on every lane of set s).  More sets: one lane per set.
  if (P.nms <= 16) {
    const int nms = P.nms, g = wl >> 4, sset = wl & 15;
    double m = 0.0;
    if (sset < nms) {
      // all sixteen operands read first (the efficiency row index clamped into the table), the
      // species guard a select: guarded loads would wait one after another
      float ev[8];
      double cv8[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const int j = g + 4 * t;
        ev[t] = L.effT[min(j, ns - 1) * nms + sset];
        cv8[t] = L.c[j];
      }
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const double u = m + (double)ev[t] * cv8[t];
        m = g + 4 * t < ns ? u : m;
      }
    }
    const double2 mq = pair16(m);
    m = mq.x + mq.y;
    const double2 hm = halves(m);
    if (wl < nms) L.Mv[wl] = hm.x + hm.y;
  } else if (lane < P.nms) {
    const int nms = P.nms;
    double m = 0.0;
#pragma unroll 8
    for (int j = 0; j < ns; ++j) m = m + (double)L.effT[j * nms + lane] * L.c[j];
    L.Mv[lane] = m;
  }
  tsync();
  // rates of progress, one reaction per lane and round; the next round's record is loaded ahead
  int i0 = 0;
  if (!jac) {  // the first round from the values loaded ahead
    const int4 cur = rec;
    rec = rec_at(L, WAVE + wl, nslot);
    if ((cur.x & 7) != 7) reaction<false, true>(cur, rk0, rk1, Td, L, 0, tv0);
    i0 = 1;
  }
#pragma unroll 1
  for (int i = i0; i * WAVE < nslot; ++i) {
    const int4 cur = rec;
    const int nxt = (i + 1) * WAVE + wl;
    rec = rec_at(L, nxt, nslot);
    if ((cur.x & 7) != 7) {
      if (jac) reaction<true>(cur, rk0, rk1, Td, L, i * WAVE + wl);
      else reaction<false>(cur, rk0, rk1, Td, L, 0);
    }
  }
  tsync();
  // net production: chunk sums in reaction order, then the chunks of species k in order
#pragma unroll 1
  for (int c2 = 0; c2 < P.C; c2 += 2) {
    const int c = c2 + ((wlf >> 5) & 1);
    if (c < P.C) {
      const int2* __restrict__ ct = L.cterm + c * CW * TEAM + lane;
      int2 tm[CW];
#pragma unroll
      for (int t = 0; t < CW; ++t) tm[t] = ct[t * TEAM];
      // every rate gathered before the sum: gathered inside it, each one is waited for in turn
      double qv[CW];
#pragma unroll
      for (int t = 0; t < CW; ++t) qv[t] = L.q[tm[t].x];
      double acc = 0.0;
#pragma unroll
      for (int t = 0; t < CW; ++t) acc = acc + (double)__int_as_float(tm[t].y) * qv[t];
      L.part[c * TEAM + lane] = acc;
    }
  }
  // thermo rows, only needed at the end: loaded here, their latency hidden by the production sums
  // (held through the reaction rounds they would cost those rounds registers)
  const double* __restrict__ rh0 = P.tabh + (long)(T_i - P.Th_min) * (2 * ns);
  const double* __restrict__ rh1 = rh0 + 2 * ns;
  double2 th0 = make_double2(0.0, 0.0), th1 = th0;
  if (sp) {
    th0 = ld2(rh0 + 2 * lane);
    th1 = ld2(rh1 + 2 * lane);
  }
  tsync();
  double w = 0.0;
  if (sp) {
    double acc = 0.0;
    if (P.MAXP <= 8) {
      // all slots read at once, then all partial sums (slot -1 reads slot 0 and is not added):
      // two LDS round trips instead of two per chunk, the same additions in the same order
      int sl[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int v = L.kpart[min(u, P.MAXP - 1) * TEAM + lane];  // clamped into the table
        sl[u] = u < P.MAXP ? v : -1;
      }
      double pv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) pv[u] = L.part[max(sl[u], 0)];
#pragma unroll
      for (int u = 0; u < 8; ++u) acc = sl[u] >= 0 ? acc + pv[u] : acc;
    } else {
#pragma unroll 1
      for (int i = 0; i < P.MAXP; ++i) {
        const int slot = L.kpart[i * TEAM + lane];
        if (slot < 0) break;
        acc = acc + L.part[slot];
      }
    }
    w = L.W[lane] * acc;
  }
  o.w = w;
  o.cp = o.h = o.dcp = o.dh = 0.0;
  if (sp) {
    o.cp = th0.x + (th1.x - th0.x) * Td;
    o.h = th0.y + (th1.y - th0.y) * Td;
    o.dcp = th1.x - th0.x;
    o.dh = th1.y - th0.y;
  }
  double FT;
  if (mode == 0) {
    const double eiroi = tsum((o.h - Rk * T) * w);
    o.rho_cv = tsum(roi * (o.cp - Rk));
    FT = -eiroi / o.rho_cv;
    o.F = sp ? w : 0.0;
  } else {
    const double Q = tsum(o.h * w);
    o.cpm = tsum(o.xc * o.cp);
    const double irho = 1.0 / o.rho;
    FT = -Q * irho / o.cpm;
    o.F = sp ? w * irho : 0.0;
  }
  if (lane == ns) o.F = FT;
  o.ok = true;
  return o;
}

// Jacobian of the team's cell, after team_rhs with jac (which left the reactions' dq/dc terms,
// times nu_kr, in row k of L.lu, their dq/dT sums in column 32, and dq/dM of the reactions with a
// concentration M in L.dq).  The third-body columns are added here: row k gets
// sum over its terms with M of nu_kr dq/dM eff_j, column j by lane j, rows 2i and 2i + 1 by the two
// halves.  Then row k (lane k) is scaled and completed, the temperature row formed from column
// sums, and the rows stored to the team's global slot Jg (row k = lane k, stride 33).
__device__ __forceinline__ void team_jac(const Prog& P, int mode, double x, const Lds& L, const RhsOut o,
                                         double* __restrict__ Jg) {
  const int lane = threadIdx.x & (TEAM - 1);
  const int ns = P.ns, nz = ns + 1;
  const bool sp = lane < ns;
  const double T = tbc(x, ns);
  const double FT = tbc(o.F, ns);
  // rows are completed by the lower half (the upper half joins the team sums with identical values)
  const bool h0 = half() == 0;
  double* row = L.lu + lane * 33;
  L.g[lane] = o.g;
  if (mode == 1) L.aux[lane] = o.xc;
  tsync();
  {
    const int h = half();
#pragma unroll 1
    for (int i = 0; 2 * i < ns; ++i) {
      const int k = 2 * i + h;
      const int e0 = k < ns ? L.tbptr[k] : 0, e1 = k < ns ? L.tbptr[k + 1] : 0;
      if (e1 > e0) {
        double acc = 0.0;
#pragma unroll 2
        for (int e = e0; e < e1; ++e) {
          const int4 te = L.tbe[e];
          const double Dv = (double)__int_as_float(te.y) * L.dq[te.w];
          if (sp) acc += Dv * (double)L.effT[lane * P.nms + te.z];
        }
        if (sp) L.lu[k * 33 + lane] += acc;
      }
    }
  }
  tsync();
  // each row loop below touches only this lane's row and reads per-column values kept elsewhere
  // (restrict-qualified within its block, so that its loads need not wait for its stores)
  double bT = 0.0;
  const double Wk = L.W[lane];
  const double Rk = L.Rk[lane];
  if (mode == 0) {
    // constant volume: species block A_kj = d wdot_k / d rho_j (unclipped) = (W_k / W_j) times the
    // concentration derivative, J_kj = g_j A_kj, J_kT = b_k = W_k d wdot_k / dT at fixed rho; the
    // T row from column sums of e_k J_kj
    if (sp && h0) {
      double* __restrict__ rw = row;
      const double* __restrict__ gj = L.g;
      const double* __restrict__ Wj = L.W;
      bT = rw[32];
#pragma unroll 4
      for (int j = 0; j < ns; ++j) rw[j] = rw[j] * (Wk / Wj[j]) * gj[j];
      rw[ns] = Wk * bT;
    }
    const double s_de = tsum(sp ? (o.dh - Rk) * o.w : 0.0);
    const double s_dcv = tsum(sp ? o.xc * o.dcp : 0.0);
    L.aux[lane] = sp ? o.h - Rk * T : 0.0;
    tsync();
    double jt = 0.0;
    if (lane < nz) {
      double cs = 0.0;
      for (int k = 0; k < ns; ++k) cs += L.aux[k] * L.lu[k * 33 + lane];
      jt = lane < ns ? -(cs + FT * (o.cp - Rk) * o.g) / o.rho_cv : -(s_de + cs + FT * s_dcv) / o.rho_cv;
    }
    tsync();
    if (lane < nz) L.lu[ns * 33 + lane] = jt;
    tsync();
  } else {
    // constant pressure (Y, T): see oracle/flint_oracle.c jac_cp_cell for the formulas
    const double Rm = o.Rmix, rho = o.rho;
    const double iRm = 1.0 / Rm, irho = 1.0 / rho;
    double AY = 0.0;
    if (sp && h0) {
      double* __restrict__ rw = row;
      const double* __restrict__ gj = L.g;
      const double* __restrict__ Rj = L.Rk;
      const double* __restrict__ Wj = L.W;
      const double* __restrict__ Yj = L.aux;
      bT = rw[32];
#pragma unroll 4
      for (int j = 0; j < ns; ++j) {
        const double a = rw[j] * (Wk / Wj[j]);
        rw[j] = a;
        AY += a * Yj[j];
      }
      // stage dwdot_k/dY_j and dwdot_k/dT for the enthalpy-weighted column sums
      const double t = AY * iRm;
#pragma unroll 4
      for (int j = 0; j < ns; ++j) rw[j] = rho * gj[j] * (rw[j] - t * Rj[j]);
      rw[ns] = Wk * bT - rho * AY / T;
    }
    const double s_dhw = tsum(sp ? o.dh * o.w : 0.0);
    const double sYdcp = tsum(sp ? o.xc * o.dcp : 0.0);
    tsync();
    L.aux[lane] = sp ? o.h : 0.0;
    tsync();
    double jt = 0.0;
    if (lane < nz) {
      double ch = 0.0;
      for (int k = 0; k < ns; ++k) ch += L.aux[k] * L.lu[k * 33 + lane];
      jt = lane < ns ? -ch / (rho * o.cpm) - FT * o.g * (o.cp / o.cpm - Rk / Rm)
                     : -(s_dhw + ch) / (rho * o.cpm) - FT * (sYdcp / o.cpm - 1.0 / T);
    }
    tsync();
    if (sp && h0) {
      double* __restrict__ rw = row;
      const double* __restrict__ gj = L.g;
      const double* __restrict__ Rj = L.Rk;
      const double u = o.w * irho * iRm;
#pragma unroll 4
      for (int j = 0; j < ns; ++j) rw[j] = rw[j] * irho + gj[j] * u * Rj[j];
      rw[ns] = rw[ns] * irho + o.w / (rho * T);
    }
    if (lane < nz) L.lu[ns * 33 + lane] = jt;
    tsync();
  }
  // the rows go to the global slot, where the inverse reads them row by row (the same lane)
  if (h0) {
    double* __restrict__ jrow = Jg + lane * 33;
#pragma unroll
    for (int j = 0; j < TEAM; ++j) jrow[j] = row[j];
  }
  tsync();
}

// Inverse of A = I - c J by Gauss-Jordan elimination with partial pivoting and no row exchanges.
// Row i lives in registers of component lane i: columns 0..15 in the lower half, 16..31 in the
// upper half.  Step k takes the unused row p with the largest |a_pk| (tpivot: ties, and values
// within 2^-15 of each other, to the lower lane), row p is published in LDS already scaled by
// 1 / a_pk (its column-k slot carrying 1 / a_pk itself), and
//   row p:        a_pj <- a_pj / a_pk (j != k),   a_pk <- 1 / a_pk
//   rows i != p:  a_ij <- a_ij - a_ik a_pj (j != k, a_pj scaled),   a_ik <- -a_ik / a_pk
// (the column-k storage takes over the identity column p, as in the classic in-place scheme).
// Both halves rotate their 16 registers by one after every step, so that the current column k is
// always register 0 of half k / 16 and every register index stays static (a dynamic one would
// move the row to scratch); the pivot row is published in the same rotated order.
// Afterwards lane i holds row step(i) of A^-1 with its columns in pivot order,
// A^-1[step(i)][p_m] = a_im (p_m: the row chosen at step m); the rows go to LDS (stride 33) for
// the solves.  Rows and columns beyond nz are the identity and take no part.  Returns step(lane).
__device__ __forceinline__ int team_inv(double c, int nz, const Lds& L, const double* __restrict__ Jg) {
  constexpr int HC = TEAM / 2;  // columns per half
  const int wlf = wave_lane();
  const int lane = wlf & (TEAM - 1);
  const int h = (wlf >> 5) & 1, c0 = HC * h;
  const double* __restrict__ jrow = Jg + lane * 33;
  double r[HC];
  double jv[HC];  // the slot row read unguarded (it is always in bounds), the guard a select
#pragma unroll
  for (int j = 0; j < HC; ++j) jv[j] = jrow[c0 + j];
#pragma unroll
  for (int j = 0; j < HC; ++j)
    r[j] = (lane == c0 + j ? 1.0 : 0.0) - ((lane < nz && c0 + j < nz) ? c * jv[j] : 0.0);
  double* __restrict__ pb = L.vx;  // the pivot row of the current step, rotated like the registers
  int mystep = -1;
#pragma unroll 1
  for (int k = 0; k < nz; ++k) {
    const int hk = k / HC;
    const double2 hv = halves(r[0]);  // column k is register 0 of half hk
    const double rk = hk == 0 ? hv.x : hv.y;
    double iv = 1.0 / rk;  // used by the pivot row only; computed by all while the search runs
    asm volatile("" : "+v"(iv));  // (kept ahead of the search rather than sunk into the pivot branch)
    const bool me = lane == tpivot(rk, mystep < 0);
    if (me) {
      mystep = k;
      pb[c0] = h == hk ? iv : r[0] * iv;
#pragma unroll
      for (int j = 1; j < HC; ++j) pb[c0 + j] = r[j] * iv;
    }
    tsync();
    // the pivot row's half read whole before the update: read as the update consumes it, the
    // compiler reuses a few registers and waits for every pair of values
    double pv[HC];
#pragma unroll
    for (int j = 0; j < HC; ++j) pv[j] = pb[c0 + j];
    const double inv = pb[HC * hk];
    // pivot lane: a_pj / a_pk; other lanes: a_ij - a_ik (a_pj / a_pk); one fused form for both,
    // rotating by one register on the way
    const double keep = me ? 0.0 : 1.0, fm = me ? -1.0 : rk;
    const double nk = me ? inv : -rk * inv;
    const double r0 = h == hk ? nk : keep * r[0] - fm * pv[0];
#pragma unroll
    for (int j = 0; j < HC - 1; ++j) r[j] = keep * r[j + 1] - fm * pv[j + 1];
    r[HC - 1] = r0;
    tsync();
  }
  // undo the rotation by nz columns: register j holds column c0 + (j + nz) mod 16
  double* arow = L.lu + lane * 33;
#pragma unroll
  for (int j = 0; j < HC; ++j) arow[c0 + ((j + nz) & (HC - 1))] = r[j];
  tsync();
  return mystep < 0 ? lane : mystep;
}

// x = A^-1 b with the rows above; b and x are this lane's component.  b is placed in pivot order
// (vb[step(l)] = b_l), each half forms the dot product over its 16 columns, the halves' partial
// sums are added (lower + upper in both) and the result goes to its step's slot.
__device__ __forceinline__ double team_solve(int mystep, double b, int nz, const Lds& L) {
  constexpr int HC = TEAM / 2;
  const int lane = threadIdx.x & (TEAM - 1);
  const int h = half(), c0 = HC * h;
  const double* arow = L.lu + lane * 33;
  L.vb[mystep] = b;
  tsync();
  // all operands read first and the guard applied as a select: a guarded load becomes a branch
  // with its own wait, and sixteen of them would run one after another
  double av[HC], bv[HC];
#pragma unroll
  for (int m = 0; m < HC; ++m) {
    av[m] = arow[c0 + m];
    bv[m] = L.vb[c0 + m];
  }
  double yp = 0.0;
#pragma unroll
  for (int m = 0; m < HC; ++m) {
    const double t = yp + av[m] * bv[m];
    yp = c0 + m < nz ? t : yp;
  }
  const double2 yv = halves(yp);
  const double y = lane < nz ? yv.x + yv.y : b;
  L.vx[mystep] = y;
  tsync();
  const double x = L.vx[lane];
  tsync();
  return x;
}

// Two solves, (I - c1 J) x1 = b1 and (I - c2 J) x2 = b2: system 1 in the lower half of the
// wavefront, system 2 in the upper half, each lane holding one full row (32 columns in registers,
// rotated by one per step so that column k is always register 0 and every index stays static).
// Gauss-Jordan elimination of the augmented systems with partial pivoting, each half on its own:
// the pivot search, the publication of the pivot row and the row updates all stay within a half,
// so one instruction stream serves both systems.  The pivot row is left unscaled (it stands for
// the classic scaled row times a_pk, a factor the later steps carry along linearly and the end
// divides out), and no inverse is kept: row p_k, the pivot of step k, ends as
// x_k = beta_{p_k} / a_{p_k k}.  The pivot rows go through L.vx (system 1) and L.vb (system 2),
// their right-hand sides through L.g[0..1].  Every lane receives x1 and x2 of its component.
__device__ __forceinline__ void team_solve2(double c1, double b1, double c2, double b2, int nz, const Lds& L,
                                            const double* __restrict__ Jg, double& x1, double& x2) {
  const int lane = threadIdx.x & (TEAM - 1);
  const int h = half();
  const double c = h == 0 ? c1 : c2;
  const double* __restrict__ jrow = Jg + lane * 33;
  double r[TEAM];
#pragma unroll
  for (int j = 0; j < TEAM; ++j) r[j] = jrow[j];  // unguarded (always in bounds), guard below
#pragma unroll
  for (int j = 0; j < TEAM; ++j) {
    const double jv = (lane < nz && j < nz) ? r[j] : 0.0;
    r[j] = (lane == j ? 1.0 : 0.0) - c * jv;
  }
  double be = lane < nz ? (h == 0 ? b1 : b2) : 0.0;
  double* __restrict__ pb = h == 0 ? L.vx : L.vb;
  double* __restrict__ pbe = L.g + h;
  int my = -1;
  double myiv = 1.0;
#pragma unroll 1
  for (int k = 0; k < nz; ++k) {
    const double rk = r[0];
    double iv = 1.0 / rk;  // used by the pivot row only; computed by all while the search runs
    asm volatile("" : "+v"(iv));
    const bool me = lane == tpivot(rk, my < 0);
    if (me) {
      my = k;
      myiv = iv;
      pb[0] = iv;
#pragma unroll
      for (int j = 1; j < TEAM; ++j) pb[j] = r[j];
      *pbe = be;
    }
    tsync();
    // f = a_ik / a_pk (0 on the pivot row, which keeps its values); column k itself is not needed
    const double f = me ? 0.0 : rk * pb[0];
    be = be - f * *pbe;
    // the pivot row read eight values at a time ahead of their use (one at a time, each read is
    // waited for; all 32 at once do not fit beside the rows)
#pragma unroll
    for (int j0 = 0; j0 < TEAM; j0 += 8) {
      double pv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) pv[u] = pb[min(j0 + u + 1, TEAM - 1)];
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (j0 + u < TEAM - 1) r[j0 + u] = r[j0 + u + 1] - f * pv[u];
    }
    r[TEAM - 1] = 0.0;
    tsync();
  }
  if (lane < nz) pb[my < 0 ? lane : my] = be * myiv;  // (a row that never pivoted: singular system)
  tsync();
  x1 = L.vx[lane];
  x2 = L.vb[lane];
  tsync();
}

// Step-size factor x^e of the controller (x >= 0, -1 <= e < 0): x = m 2^k with m in
// [sqrt(1/2), sqrt(2)), log2 x = k + 2 atanh(s) / ln 2 with s = (m - 1)/(m + 1) (|s| <= 0.172, the
// series to s^13), then 2^t = 2^n e^(f ln 2) with n = rint(t), |f| <= 1/2 (the series to f^10):
// about 1e-12 relative, which is more than a step-size factor needs, in about 100 instructions
// against about 250 for an FP64 pow.  Every operation is an IEEE double operation written out
// (explicit fma), so the CPU BDF computes the same factors bit for bit.  x = 0 gives +inf,
// x = +inf gives 0, as pow does.
__device__ __forceinline__ double step_pow(double x, double e) {
  if (x == 0.0) return INFINITY;
  if (!(x <= 1.7976931348623157e308)) return x == INFINITY ? 0.0 : x;
  int k;
  double m = frexp(x, &k);
  if (m < 0.70710678118654752) {
    m = m * 2.0;
    k = k - 1;
  }
  const double s = (m - 1.0) / (m + 1.0), s2 = s * s;
  double p = 1.0 / 13.0;
  p = fma(p, s2, 1.0 / 11.0);
  p = fma(p, s2, 1.0 / 9.0);
  p = fma(p, s2, 1.0 / 7.0);
  p = fma(p, s2, 1.0 / 5.0);
  p = fma(p, s2, 1.0 / 3.0);
  p = fma(p, s2, 1.0);
  const double l2 = fma(2.0 * s * p, 1.4426950408889634, (double)k);
  const double t = fmin(fmax(e * l2, -1074.0), 1024.0);
  const double n = rint(t), y = (t - n) * 0.6931471805599453;
  double q = 1.0 / 3628800.0;
  q = fma(q, y, 1.0 / 362880.0);
  q = fma(q, y, 1.0 / 40320.0);
  q = fma(q, y, 1.0 / 5040.0);
  q = fma(q, y, 1.0 / 720.0);
  q = fma(q, y, 1.0 / 120.0);
  q = fma(q, y, 1.0 / 24.0);
  q = fma(q, y, 1.0 / 6.0);
  q = fma(q, y, 0.5);
  q = fma(q, y, 1.0);
  q = fma(q, y, 1.0);
  return ldexp(q, (int)n);
}

// RMS over the team's nz components of v times the reciprocal weight iscale (products by
// reciprocals instead of divisions: the step control runs this several times per step).
__device__ __forceinline__ double team_rms(double v, double iscale, int nz, const Lds& L) {
  const int lane = threadIdx.x & (TEAM - 1);
  const double q = lane < nz ? v * iscale : 0.0;
  return sqrt(tsum(q * q) * L.rcp[0]);
}

// This lane's column of the BDF difference array in LDS (both halves hold the same values).
struct Dcol {
  double* p;
  __device__ __forceinline__ double& operator[](int j) const { return p[j * TEAM]; }
};

// D[0..order] <- (R U)^T D[0..order] on this lane's column, with R(order, factor) and U = R(order, 1)
// of the method: R[j][l] = prod_{m=1..j} (m - 1 - factor l)/m, so R[0][l] = 1 and R[j][0] = [j == 0].
// Because RU[j][0] = [j == 0] and RU[0][i] = [i == 0], D[0] (the solution itself) is left as it is
// and never enters the differences: v_l = sum_{j>=1} R[j][l] D[j], then D_i = sum_{l>=1} U[l][i] v_l
// for i >= 1.  Mixing D[0] into v_l and relying on sum_l U[l][i] = 0 to remove it again would leave
// an absolute error of eps |y| in every difference, which later step-size increases amplify by
// factor^j into a visible loss of mass after a long run of step halvings.
__device__ __forceinline__ void change_D(const Dcol D, int order, double factor) {
  double dv[MAXO + 1];  // every row read at once (in bounds), the order's guard a select
#pragma unroll
  for (int j = 1; j <= MAXO; ++j) dv[j] = D[j];
  double v[MAXO + 1];
#pragma unroll
  for (int l = 1; l <= MAXO; ++l) {
    double prod = 1.0, acc = 0.0;
#pragma unroll
    for (int j = 1; j <= MAXO; ++j) {
      prod *= (j - 1 - factor * l) * (1.0 / j);  // 1/j folds to a constant: no division
      const double t = acc + prod * dv[j];
      acc = j <= order ? t : acc;
    }
    v[l] = (l <= order) ? acc : 0.0;
  }
#pragma unroll
  for (int i = 1; i <= MAXO; ++i) {
    if (i > order) continue;
    double acc = 0.0;
#pragma unroll
    for (int l = 1; l <= MAXO; ++l) {
      // U[l][i] = prod_{m=1..l} (m - 1 - i)/m: a signed binomial coefficient, exact in double
      double u = 1.0;
#pragma unroll
      for (int m = 1; m <= MAXO; ++m)
        if (m <= l) u *= (double)(m - 1 - i) / m;
      if (l <= order) acc += u * v[l];
    }
    D[i] = acc;
  }
}

// A decision of the step control, taken as the first lane's.  Every lane computes the same bits
// (team sums are bit-identical on all lanes and both halves hold the same values), but the compiler
// cannot know it: with the value read from one lane the branch is a scalar branch with no exec-mask
// save and restore, and the state it sets (order, phase, counters) lives in scalar registers,
// which leaves vector registers to the evaluations.  Only called with the whole wavefront active.
__device__ __forceinline__ bool unib(bool b) { return __builtin_amdgcn_readfirstlane((int)b) != 0; }

// One cell; mirrors bdf_cell of oracle/oracle_bdf.c decision for decision.  Returns 0 ok, 1 step
// underflow, 2 too many steps, 3 invalid state.
//
// The integrator is written as a state machine around one right-hand-side evaluation site (and a
// single Jacobian, inverse and solve site): every place where the method needs f(x) sets x and the
// phase that consumes the result, which keeps the kernel's code small.
__device__ __forceinline__ int integrate(const Prog& P, const Opts& O, double p, double& y, int (&st)[8], const Lds& L,
                                          double* __restrict__ Jg) {
  const int lane = threadIdx.x & (TEAM - 1);
  const int nz = P.ns + 1;
  const double rtol = O.rtol, atol = O.atol, t_end = O.t_end;
  double gamma[MAXO + 1];
  gamma[0] = 0.0;
#pragma unroll
  for (int k = 1; k <= MAXO; ++k) gamma[k] = gamma[k - 1] + 1.0 / k;
  const double newton_tol = fmax(10.0 * 2.220446049250313e-16 / rtol, fmin(0.03, sqrt(rtol)));

  enum { PH_F0, PH_F1, PH_J0, PH_NEWTON, PH_JREF };
  enum { GO_EVAL, GO_STEP, GO_ATTEMPT, GO_NEWTON, GO_HALVE, GO_ACCEPT };
  // start-up values
  double f0 = 0.0, isc = 0.0, h0 = 0.0, d1 = 0.0;  // isc: 1 / the error weight of the last accepted state
  // method state
  const Dcol D{L.D + lane};
#pragma unroll
  for (int j = 0; j < MAXO + 3; ++j) D[j] = 0.0;
  int order = 1, n_equal = 0;
  bool have_lu = false, current_jac = false;
  int mystep = lane;
  double t = 0.0, h_abs = 0.0, min_step = 0.0;
  // current step attempt and Newton iteration
  double t_new = 0.0, ypred = 0.0, psi = 0.0, iscale = 0.0, cc = 0.0, ynew = 0.0, d = 0.0;
  double dy_norm_old = -1.0, err_norm = 0.0, safety = 0.9;
  int k_it = 0;
  // Newton policy 1: coefficient of the factorised matrix, its age, the carried rate estimate
  const bool cv = O.newton == 1;
  double c_lu = 0.0, crate = 1.0;
  int lu_age = 0;

  int phase = PH_F0;
  double xe = y;
  bool wj = false;
  while (true) {
    const RhsOut o = wj ? team_rhs<true>(P, O.mode, p, xe, L) : team_rhs<false>(P, O.mode, p, xe, L);
    if (wj) {
      if (o.ok) {
        team_jac(P, O.mode, xe, L, o, Jg);
        st[ST_JAC]++;
      }
    } else if (o.ok || phase == PH_NEWTON) {
      st[ST_RHS]++;
    }
    int go = GO_EVAL;
    if (phase == PH_F0) {
      if (!o.ok) return 3;
      f0 = o.F;
      isc = 1.0 / (atol + fabs(y) * rtol);
      const double d0 = team_rms(y, isc, nz, L);
      d1 = team_rms(f0, isc, nz, L);
      h0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
      h0 = fmin(h0, t_end);
      xe = y + h0 * f0;
      phase = PH_F1;
    } else if (phase == PH_F1) {
      if (!o.ok) return 3;
      const double d2 = team_rms(o.F - f0, isc, nz, L) / h0;
      const double h1 = (d1 <= 1e-15 && d2 <= 1e-15) ? fmax(1e-6, h0 * 1e-3) : sqrt(0.01 / fmax(d1, d2));
      h_abs = fmin(100.0 * h0, fmin(h1, t_end));
      D[0] = y;
      D[1] = f0 * h_abs;
      // Jacobian at the initial state
      xe = y;
      wj = true;
      phase = PH_J0;
    } else if (phase == PH_J0) {
      if (!o.ok) return 3;
      t = 0.0;
      go = GO_STEP;
    } else if (phase == PH_JREF) {
      if (!o.ok) {
        go = GO_HALVE;
      } else {
        have_lu = false;
        current_jac = true;
        go = GO_NEWTON;
      }
    } else {  // PH_NEWTON: f at the current iterate ynew
      bool conv = false, fail = false;
      int n_iter = k_it + 1;
      if (!o.ok || unib(!isfinite(tsum(o.F)))) {
        fail = true;
      } else if (cv) {
        const double dy = team_solve(mystep, cc * o.F - psi - d, nz, L) * L.sx[0];
        st[ST_NEWTON]++;
        const double dy_norm = team_rms(dy, iscale, nz, L);
        if (k_it > 0 && unib(dy_norm > CV_RDIV * dy_norm_old)) {
          fail = true;
        } else {
          if (k_it > 0) crate = fmax(CV_CRDOWN * crate, dy_norm / dy_norm_old);
          ynew += dy;
          d += dy;
          if (unib(dy_norm * fmin(1.0, crate) * (1.0 / (order + 1)) <= CV_NLSCOEF)) {
            conv = true;
          } else {
            dy_norm_old = dy_norm;
            ++k_it;
            if (k_it < CV_MAXCOR) {
              xe = ynew;
            } else {
              fail = true;
              n_iter = k_it + 1;
            }
          }
        }
      } else {
        const double dy = team_solve(mystep, cc * o.F - psi - d, nz, L);
        st[ST_NEWTON]++;
        const double dy_norm = team_rms(dy, iscale, nz, L);
        const double rate = dy_norm_old >= 0.0 ? dy_norm / dy_norm_old : -1.0;
        if (unib(rate >= 0.0 && (rate >= 1.0 || pow(rate, (double)(NEWTON_MAXITER - k_it)) / (1.0 - rate) * dy_norm > newton_tol))) {
          fail = true;
        } else {
          ynew += dy;
          d += dy;
          if (unib(dy_norm == 0.0 || (rate >= 0.0 && rate / (1.0 - rate) * dy_norm < newton_tol))) {
            conv = true;
          } else {
            dy_norm_old = dy_norm;
            ++k_it;
            if (k_it < NEWTON_MAXITER) {
              xe = ynew;  // next Newton iteration
            } else {
              fail = true;
              n_iter = k_it + 1;
            }
          }
        }
      }
      if (fail) {
        if (!current_jac) {
          xe = ypred;  // refresh the Jacobian at the predictor and iterate again
          wj = true;
          phase = PH_JREF;
        } else {
          go = GO_HALVE;
        }
      } else if (conv) {
        if (cv) lu_age++;
        safety = 0.9 * (2 * NEWTON_MAXITER + 1) / (double)(2 * NEWTON_MAXITER + n_iter);
        isc = 1.0 / (atol + rtol * fabs(ynew));
        err_norm = team_rms(d * L.rcp[order + 1], isc, nz, L);
        if (unib(err_norm > 1.0)) {
          st[ST_REJECT]++;
          const double factor = fmax(MIN_FACTOR, safety * step_pow(err_norm, -L.rcp[order + 1]));
          h_abs *= factor;
          change_D(D, order, factor);
          n_equal = 0;
          go = GO_ATTEMPT;
        } else {
          go = GO_ACCEPT;
        }
      }
    }
    // control flow between evaluations
    while (go != GO_EVAL) {
      if (go == GO_HALVE) {  // Newton failed with a current Jacobian
        st[ST_NFAIL]++;
        h_abs *= 0.5;
        change_D(D, order, 0.5);
        n_equal = 0;
        if (!cv) have_lu = false;
        go = GO_ATTEMPT;
      } else if (go == GO_ACCEPT) {
        st[ST_STEPS]++;
        n_equal++;
        t = t_new;
        y = ynew;
        current_jac = false;
        // D^{j+1} y_n = D^j y_n - D^j y_{n-1}, d = D^{order+1} y_n
        {
          const double Dop1 = D[order + 1];
          D[order + 2] = d - Dop1;
          D[order + 1] = d;
#pragma unroll 1
          for (int j = order; j >= 0; --j) D[j] += D[j + 1];
        }
        go = GO_STEP;
        if (n_equal < order + 1) continue;
        const double Dm = D[order], Dp = D[order + 2];
        const double em = order > 1 ? team_rms(Dm * L.rcp[order], isc, nz, L) : INFINITY;
        const double ep = order < MAXO ? team_rms(Dp * L.rcp[order + 2], isc, nz, L) : INFINITY;
        const double norms[3] = {em, err_norm, ep};
        double best = -1.0;
        int which = 1;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          const double fct = norms[q] == 0.0 ? INFINITY : step_pow(norms[q], -L.rcp[order + q]);
          if (unib(fct > best)) { best = fct; which = q; }
        }
        order += which - 1;
        const double factor = fmin(MAX_FACTOR, safety * best);
        h_abs *= factor;
        change_D(D, order, factor);
        n_equal = 0;
        if (!cv) have_lu = false;
      } else if (go == GO_STEP) {
        if (unib(!(t < t_end))) {
          st[ST_ORDER] = order;
          return 0;
        }
        if (st[ST_STEPS] >= O.max_steps) return 2;
        min_step = 10.0 * fabs(nextafter(t, INFINITY) - t);
        if (unib(h_abs < min_step)) {
          change_D(D, order, min_step / h_abs);
          h_abs = min_step;
          n_equal = 0;
        }
        safety = 0.9;
        err_norm = 0.0;
        go = GO_ATTEMPT;
      } else if (go == GO_ATTEMPT) {
        if (unib(h_abs < min_step)) return 1;
        t_new = t + h_abs;
        if (unib(t_new - t_end > 0)) {
          t_new = t_end;
          change_D(D, order, fabs(t_new - t) / h_abs);
          n_equal = 0;
          if (!cv) have_lu = false;
        }
        const double h = t_new - t;
        h_abs = fabs(h);
        ypred = 0.0;
        psi = 0.0;
        double dv[MAXO + 1];  // every row read at once, the order's guard a select
#pragma unroll
        for (int j = 0; j <= MAXO; ++j) dv[j] = D[j];
#pragma unroll
        for (int j = 0; j <= MAXO; ++j) {
          const double t = ypred + dv[j];
          ypred = j <= order ? t : ypred;
        }
#pragma unroll
        for (int j = 1; j <= MAXO; ++j) {
          const double t = psi + dv[j] * gamma[j];
          psi = j <= order ? t : psi;
        }
        double alpha_o = 0.0;
#pragma unroll
        for (int j = 1; j <= MAXO; ++j)
          if (j == order) alpha_o = gamma[j];
        psi /= alpha_o;
        iscale = 1.0 / (atol + rtol * fabs(ypred));
        cc = h / alpha_o;
        go = GO_NEWTON;
      } else {  // GO_NEWTON: (re)start the Newton iteration from the predictor
        if (cv && have_lu && (unib(fabs(cc / c_lu - 1.0) > CV_DGMAX) || lu_age >= CV_MSBP)) have_lu = false;
        if (!have_lu) {
          mystep = team_inv(cc, nz, L, Jg);
          st[ST_LU]++;
          have_lu = true;
          c_lu = cc;
          lu_age = 0;
          crate = 1.0;
        }
        // CVODE's correction scale for a matrix factored at c_lu, fixed for this attempt's iterations
        if (cv) L.sx[0] = 2.0 / (1.0 + cc / c_lu);
        ynew = ypred;
        d = 0.0;
        dy_norm_old = -1.0;
        k_it = 0;
        xe = ynew;
        wj = false;
        phase = PH_NEWTON;
        go = GO_EVAL;
      }
    }
  }
}

__global__ void fill_int(int* p, long n, int v) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void zero_counter(int* c) {
  if (threadIdx.x == 0 && blockIdx.x == 0) c[0] = 0;
}

// Work order of the persistent kernel: cells by descending |key| (|dT/dt| at the initial state),
// a counting sort over ORDER_BUCKETS classes of the key's binary exponent and leading mantissa
// bit (a factor of sqrt(2) per class).  Cells are independent, so the order changes no result;
// handing the costliest cells out first shortens the tail that one slow cell would otherwise set
// (longest-processing-time-first).  Within a class the order follows the atomics.
constexpr int ORDER_BUCKETS = 4096;
__device__ __forceinline__ int order_bucket(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(fabs(v));  // sign bit 0
  return ORDER_BUCKETS - 1 - (int)(b >> 51);
}
// Each block counts its chunk of ORDER_CHUNK cells in LDS first: cells of similar |key| (all of
// them, at t = 0) would otherwise serialise on one global counter.
constexpr int ORDER_CHUNK = 4096, ORDER_THREADS = 256;
__device__ __forceinline__ void order_count(long n, const double* __restrict__ key, const int* __restrict__ idx,
                                            int* lh) {
  for (int b = threadIdx.x; b < ORDER_BUCKETS; b += ORDER_THREADS) lh[b] = 0;
  __syncthreads();
  const long i0 = (long)blockIdx.x * ORDER_CHUNK, i1 = i0 + ORDER_CHUNK < n ? i0 + ORDER_CHUNK : n;
  for (long i = i0 + threadIdx.x; i < i1; i += ORDER_THREADS) atomicAdd(&lh[order_bucket(key[idx ? idx[i] : i])], 1);
  __syncthreads();
}
__global__ __launch_bounds__(ORDER_THREADS) void order_hist(long n, const double* __restrict__ key,
                                                            const int* __restrict__ idx, int* __restrict__ hist) {
  __shared__ int lh[ORDER_BUCKETS];
  order_count(n, key, idx, lh);
  for (int b = threadIdx.x; b < ORDER_BUCKETS; b += ORDER_THREADS)
    if (lh[b]) atomicAdd(&hist[b], lh[b]);
}
// exclusive prefix sum of the class counts in place (one block of 1024 threads, 4 classes each)
__global__ __launch_bounds__(1024) void order_scan(int* __restrict__ hist) {
  __shared__ int s[1024];
  const int t = threadIdx.x;
  int v[4], sum = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    v[j] = hist[4 * t + j];
    sum += v[j];
  }
  s[t] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int x = t >= off ? s[t - off] : 0;
    __syncthreads();
    s[t] += x;
    __syncthreads();
  }
  int run = s[t] - sum;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    hist[4 * t + j] = run;
    run += v[j];
  }
}
// The block reserves a range of each class's positions (hist holds the classes' next free
// positions), then hands them out to its cells.
__global__ __launch_bounds__(ORDER_THREADS) void order_scatter(long n, const double* __restrict__ key,
                                                               const int* __restrict__ idx, int* __restrict__ hist,
                                                               int* __restrict__ out) {
  __shared__ int lh[ORDER_BUCKETS];
  order_count(n, key, idx, lh);
  for (int b = threadIdx.x; b < ORDER_BUCKETS; b += ORDER_THREADS)
    if (lh[b]) lh[b] = atomicAdd(&hist[b], lh[b]);
  __syncthreads();
  const long i0 = (long)blockIdx.x * ORDER_CHUNK, i1 = i0 + ORDER_CHUNK < n ? i0 + ORDER_CHUNK : n;
  for (long i = i0 + threadIdx.x; i < i1; i += ORDER_THREADS) {
    const int cell = idx ? idx[i] : (int)i;
    out[atomicAdd(&lh[order_bucket(key[cell])], 1)] = cell;
  }
}

// Two waves per SIMD: the per-team LDS (LU factors) allows about seven resident waves per CU, and
// at two waves the register budget costs only a small spill (measured faster than one wave at 400
// registers, and than three waves).
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2, 2))) void bdf_kernel(Prog P, Opts O, long n, long ld, double* __restrict__ X,
                                                  const double* __restrict__ pres, int* __restrict__ stats,
                                                  int* __restrict__ status, int* __restrict__ counter,
                                                  unsigned team_bytes, double* __restrict__ jac_scratch,
                                                  const int* __restrict__ order) {
  extern __shared__ __align__(16) char smem[];
  const int lane = threadIdx.x & (TEAM - 1);
  const bool h0 = half() == 0;
  // the wavefront's team index, in a scalar register: the team's LDS addresses derived from it are
  // scalars too (the compiler cannot see that threadIdx.x / 64 is uniform)
  const int team = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
  const size_t sh = shared_prog_bytes(P);
  load_shared_prog(P, smem);
  __syncthreads();
  const Lds L = carve(smem, smem + sh + (size_t)team * team_bytes, P);
  double* __restrict__ Jg = jac_scratch + ((size_t)blockIdx.x * (blockDim.x / WAVE) + team) * (TEAM * 33);
  const int nz = P.ns + 1;
  while (true) {
    int cell = 0;
    if ((threadIdx.x & (WAVE - 1)) == 0) cell = atomicAdd(counter, 1);
    cell = __builtin_amdgcn_readlane(cell, 0);
    if (cell >= n) break;
    if (order) cell = order[cell];  // the i-th cell handed out (a permutation of 0..n-1)
    double y = lane < nz ? X[(long)lane * ld + cell] : 0.0;
    const double p = O.mode == 1 ? pres[cell] : 0.0;
    int st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int rc = integrate(P, O, p, y, st, L, Jg);
    if (h0 && lane < nz) X[(long)lane * ld + cell] = y;
    if (stats && h0 && lane < 8) {
      int v = st[0];
#pragma unroll
      for (int q = 1; q < 8; ++q)
        if (lane == q) v = st[q];
      stats[(long)lane * ld + cell] = v;
    }
    if (status && h0 && lane == 0) status[cell] = rc;
  }
}

// The reactor's right-hand side and (JAC) Jacobian at given states (the same team_rhs/team_jac the
// integrator uses), one team per cell over a grid-stride loop.  Outside the tables F = -1, J = 0.
template <bool JAC>
__global__ __launch_bounds__(512) void jac_kernel(Prog P, int mode, long n, long ld, const double* __restrict__ X,
                                                  const double* __restrict__ pres, double* __restrict__ F,
                                                  double* __restrict__ J, unsigned team_bytes,
                                                  double* __restrict__ jac_scratch) {
  extern __shared__ __align__(16) char smem[];
  const int lane = threadIdx.x & (TEAM - 1);
  const bool h0 = half() == 0;
  const int team = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);  // scalar: see bdf_kernel
  const int tpb = blockDim.x / WAVE;
  const size_t sh = shared_prog_bytes(P);
  load_shared_prog(P, smem);
  __syncthreads();
  const Lds L = carve(smem, smem + sh + (size_t)team * team_bytes, P);
  double* __restrict__ Jg = jac_scratch + ((size_t)blockIdx.x * tpb + team) * (TEAM * 33);
  const int nz = P.ns + 1;
  const long stride = (long)gridDim.x * tpb;
  for (long cell = (long)blockIdx.x * tpb + team; cell < n; cell += stride) {
    const double y = lane < nz ? X[(long)lane * ld + cell] : 0.0;
    const double p = mode == 1 ? pres[cell] : 0.0;
    const RhsOut o = team_rhs<JAC>(P, mode, p, y, L);
    if (!o.ok) {
      if (h0 && lane < nz) {
        F[(long)lane * ld + cell] = -1.0;
        if (JAC)
          for (int c = 0; c < nz; ++c) J[((long)lane * nz + c) * ld + cell] = 0.0;
      }
      continue;
    }
    if (!JAC) {
      if (h0 && lane < nz) F[(long)lane * ld + cell] = o.F;
      tsync();
      continue;
    }
    team_jac(P, mode, y, L, o, Jg);
    if (h0 && lane < nz) {
      F[(long)lane * ld + cell] = o.F;
      for (int c = 0; c < nz; ++c) J[((long)lane * nz + c) * ld + cell] = Jg[lane * 33 + c];
    }
    tsync();
  }
}

// The check of a predicted step Xp over dt from X: Newton corrections from the prediction, with F
// and J at Xp, towards the backward-Euler step and (F0 = F(X) given) the trapezoidal step,
//   delta_BE = -(I - dt J)^-1 (Xp - X - dt F(Xp)),
//   delta_TR = -(I - dt/2 J)^-1 (Xp - X - dt/2 (F0 + F(Xp))).
// A cell passes when |delta_k| <= tol_y for every species and |delta_T| <= tol_t in both.  Stiff
// modes enter delta_BE damped by 1 / (1 - dt lambda), so a prediction on the slow manifold passes
// it; the second-order trapezoidal step rejects a prediction that merely agrees with backward
// Euler's own first-order error.  A passing cell takes Xp into X; a failing one, or one flagged on
// entry, keeps X, gets flag 1 and is appended to idx (idx may be null: nothing is listed; commit
// 0: X is never written).  dmax (optional): the larger of the two corrections' max |delta| over
// species and |delta_T| of each checked cell.
__global__ __launch_bounds__(512) void verify_kernel(Prog P, long n, long ld, double* __restrict__ X,
                                                     const double* __restrict__ Xp, long ldp,
                                                     const double* __restrict__ F0, long ldf,
                                                     const double* __restrict__ pres, double dt, double tol_y,
                                                     double tol_t, int commit, int* __restrict__ flag,
                                                     int* __restrict__ idx, int* __restrict__ count,
                                                     double* __restrict__ dmax, unsigned team_bytes,
                                                     double* __restrict__ jac_scratch) {
  extern __shared__ __align__(16) char smem[];
  const int lane = threadIdx.x & (TEAM - 1);
  const bool h0 = half() == 0;
  const int team = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);  // scalar: see bdf_kernel
  const int tpb = blockDim.x / WAVE;
  const size_t sh = shared_prog_bytes(P);
  load_shared_prog(P, smem);
  __syncthreads();
  const Lds L = carve(smem, smem + sh + (size_t)team * team_bytes, P);
  double* __restrict__ Jg = jac_scratch + ((size_t)blockIdx.x * tpb + team) * (TEAM * 33);
  const int nz = P.ns + 1;
  const long stride = (long)gridDim.x * tpb;
  for (long cell = (long)blockIdx.x * tpb + team; cell < n; cell += stride) {
    bool pass = flag[cell] == 0;
    if (pass) {
      const double y0 = lane < nz ? X[(long)lane * ld + cell] : 0.0;
      const double yp = lane < nz ? Xp[(long)lane * ldp + cell] : 0.0;
      const RhsOut o = team_rhs<true>(P, 1, pres[cell], yp, L);
      pass = o.ok;
      if (pass) {
        team_jac(P, 1, yp, L, o, Jg);
        const double G = lane < nz ? yp - y0 - dt * o.F : 0.0;
        const double f0 = (F0 && lane < nz) ? F0[(long)lane * ldf + cell] : 0.0;
        const double G2 = lane < nz ? yp - y0 - 0.5 * dt * (f0 + o.F) : 0.0;
        double d1, d2;
        team_solve2(dt, -G, 0.5 * dt, -G2, nz, L, Jg, d1, d2);
        double d = fabs(d1);
        if (F0) d = fmax(d, fabs(d2));
        double ey = lane < P.ns ? d : 0.0, et = lane == P.ns ? d : 0.0;
        const double bad = tsum((lane < nz && !isfinite(d)) ? 1.0 : 0.0);
        int l1 = lane, l2 = lane;
        targmax(ey, l1);
        targmax(et, l2);
        pass = bad == 0.0 && ey <= tol_y && et <= tol_t;
        if (dmax && lane == 0) {
          dmax[cell] = bad == 0.0 ? ey : INFINITY;
          dmax[ld + cell] = bad == 0.0 ? et : INFINITY;
        }
        if (pass && commit && h0 && lane < nz) X[(long)lane * ld + cell] = yp;
      }
      tsync();
    }
    if (threadIdx.x % WAVE == 0) {
      flag[cell] = pass ? 0 : 1;
      if (!pass && idx) idx[atomicAdd(count, 1)] = (int)cell;
    }
  }
}

}  // namespace flint_reactor

hipError_t flint_reactor_set_max_lds(int bytes) {
  const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&flint_reactor::bdf_kernel),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (e != hipSuccess) return e;
  const hipError_t e2 = hipFuncSetAttribute(reinterpret_cast<const void*>(&flint_reactor::jac_kernel<true>),
                                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (e2 != hipSuccess) return e2;
  const hipError_t e3 = hipFuncSetAttribute(reinterpret_cast<const void*>(&flint_reactor::jac_kernel<false>),
                                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (e3 != hipSuccess) return e3;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(&flint_reactor::verify_kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

// Scratch of the allocation-free launches: 64 ints of work counter (256 bytes), then
// FLINT_SCRATCH_TEAMS Jacobian slots of 32 x 33 doubles (17.3 MB).  Each GPU holds its own copy
// for the life of the process; every allocation-free launch on a GPU shares it, so two such
// launches that overlap on one GPU race.
static __device__ __align__(256) double scratch[32 + (size_t)FLINT_SCRATCH_TEAMS * 32 * 33];

hipError_t flint_reactor_scratch(int** counter, double** jac_slots) {
  static std::mutex mu;
  static std::vector<double*> addr;  // per device; the host-side name of `scratch` is no address
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> g(mu);
  if ((int)addr.size() <= dev) addr.resize(dev + 1, nullptr);
  if (!addr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[dev] = static_cast<double*>(p);
  }
  *counter = reinterpret_cast<int*>(addr[dev]);
  *jac_slots = addr[dev] + 32;
  return hipSuccess;
}

size_t flint_reactor_team_bytes(const flint_reactor::Prog& P) { return flint_reactor::team_lds_bytes(P); }
size_t flint_reactor_shared_bytes(const flint_reactor::Prog& P) { return flint_reactor::shared_prog_bytes(P); }

hipError_t flint_reactor_launch(const flint_reactor::Prog& P, const flint_reactor::Opts& O, long n, long ld,
                                double* X, const double* pres, int* stats, int* status, int* counter, int teams_per_block,
                                int blocks, double* jac_scratch, const int* order, hipStream_t s) {
  const unsigned tb = (unsigned)((flint_reactor::team_lds_bytes(P) + 15) & ~size_t(15));
  hipLaunchKernelGGL(flint_reactor::zero_counter, dim3(1), dim3(64), 0, s, counter);
  hipLaunchKernelGGL(flint_reactor::bdf_kernel, dim3(blocks), dim3(teams_per_block * flint_reactor::WAVE),
                     flint_reactor::shared_prog_bytes(P) + (size_t)tb * teams_per_block, s, P, O, n, ld, X, pres, stats, status, counter, tb, jac_scratch, order);
  return hipGetLastError();
}

hipError_t flint_reactor_order(long n, const double* key, const int* idx, int* out, int* hist, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const unsigned g = (unsigned)((n + flint_reactor::ORDER_CHUNK - 1) / flint_reactor::ORDER_CHUNK);
  const unsigned t = flint_reactor::ORDER_THREADS;
  hipLaunchKernelGGL(flint_reactor::fill_int, dim3(flint_reactor::ORDER_BUCKETS / 256), dim3(256), 0, s, hist,
                     (long)flint_reactor::ORDER_BUCKETS, 0);
  hipLaunchKernelGGL(flint_reactor::order_hist, dim3(g), dim3(t), 0, s, n, key, idx, hist);
  hipLaunchKernelGGL(flint_reactor::order_scan, dim3(1), dim3(1024), 0, s, hist);
  hipLaunchKernelGGL(flint_reactor::order_scatter, dim3(g), dim3(t), 0, s, n, key, idx, hist, out);
  return hipGetLastError();
}

hipError_t flint_reactor_fill_int(int* p, long n, int v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(flint_reactor::fill_int, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t flint_reactor_verify_launch(const flint_reactor::Prog& P, long n, long ld, double* X, const double* Xp,
                                       long ldp, const double* F0, long ldf, const double* pres, double dt,
                                       double tol_y, double tol_t, int commit, int* flag, int* idx, int* count,
                                       double* dmax, int teams_per_block, int blocks, double* jac_scratch,
                                       hipStream_t s) {
  const unsigned tb = (unsigned)((flint_reactor::team_lds_bytes(P) + 15) & ~size_t(15));
  hipLaunchKernelGGL(flint_reactor::verify_kernel, dim3(blocks), dim3(teams_per_block * flint_reactor::WAVE),
                     flint_reactor::shared_prog_bytes(P) + (size_t)tb * teams_per_block, s, P, n, ld, X, Xp, ldp,
                     F0, ldf, pres, dt, tol_y, tol_t, commit, flag, idx, count, dmax, tb, jac_scratch);
  return hipGetLastError();
}

hipError_t flint_reactor_jac_launch(const flint_reactor::Prog& P, int mode, long n, long ld, const double* X,
                                    const double* pres, double* F, double* J, int teams_per_block, int blocks,
                                    double* jac_scratch, hipStream_t s) {
  const unsigned tb = (unsigned)((flint_reactor::team_lds_bytes(P) + 15) & ~size_t(15));
  if (J)
    hipLaunchKernelGGL(flint_reactor::jac_kernel<true>, dim3(blocks), dim3(teams_per_block * flint_reactor::WAVE),
                       flint_reactor::shared_prog_bytes(P) + (size_t)tb * teams_per_block, s, P, mode, n, ld, X, pres, F, J, tb, jac_scratch);
  else
    hipLaunchKernelGGL(flint_reactor::jac_kernel<false>, dim3(blocks), dim3(teams_per_block * flint_reactor::WAVE),
                       flint_reactor::shared_prog_bytes(P) + (size_t)tb * teams_per_block, s, P, mode, n, ld, X, pres, F, J, tb, jac_scratch);
  return hipGetLastError();
}
