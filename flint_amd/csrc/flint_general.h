// Device program of the data-driven general source term (general.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace flint_general {

// One list entry of a reaction: the species, the integer exponent of a concentration factor
// (when the exponent is integer-valued, applied by binary powering) and the real coefficient
// (third-body efficiency, real exponent, or W_k (ni2 - ni1) of a production term).
struct Ent {
  double v;
  int sp;
  int e;  // integer exponent; kRealExp when the factor needs pow(c, v)
};
constexpr int kRealExp = 1 << 30;

// One reaction, in FLINT's order (Arrhenius and three-body, then Troe, then Lindemann): its kind
// and the ranges of its entries in one array, third-body efficiencies [m0, f0), forward factors
// [f0, r0), reverse factors [r0, o0) and production terms [o0, end), each in species order, with
// the nonzero entries of the dense (ns+1)-row tables only.
struct Hdr {
  int kind;  // 0 Arrhenius, 1 Troe, 2 Lindemann; | 4 when ni1(M) > 0 (third body)
  int m0, f0, r0, o0, end;
  int col;   // first column of the reaction's rates in a kinetics-table row
  int pad;
};

// The compact form of a reaction, read with two 64-byte scalar loads: up to 4 forward and 4
// reverse factors with positive integer exponents and up to 8 production terms, species and
// exponents packed one byte each, and the third-body concentration taken from its efficiency set.
// A reaction that does not fit (more terms, real or non-positive exponents, a 17th set) keeps its
// entry lists (kind | 8).
struct alignas(16) Rec {
  int kind;  // Hdr::kind | 8 when the reaction runs from its entry lists
  int col;
  int mset;  // efficiency set of the third body
  int cnt;   // nf | nr << 4 | no << 8
  unsigned fsp, rsp, osp0, osp1;
  unsigned fexp, rexp;
  int pad[6];
  double wd[8];
};
constexpr int kMaxSets = 16;

struct GenProg {
  int ns, nr, na, nt, nl, wk;
  const Hdr* hdr;  // [nr]
  const Ent* ent;
  const Rec* rec;  // [nr + 1], the last an empty record read ahead by the loop
  const int* col;  // [nr + 2] Rec::col, then two zeros read ahead
  int nsets;
  const int* set_ptr;  // [nsets+1] efficiency sets: entries (species, eps) in set_ent
  const Ent* set_ent;
};

}  // namespace flint_general

hipError_t flint_general_launch(const flint_general::GenProg& P, const double* tabk, int Tk_min, int Tk_max,
                                const double* W, long n, long ld, const double* rho, const double* T, double* wdot,
                                int* status, hipStream_t s);
