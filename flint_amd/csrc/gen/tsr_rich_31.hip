// Generated by flint_amd/codegen.py from the TSR-Rich-31 mechanism descriptor. Do not edit.
// 31 species, 197 reactions (185 Arrhenius/three-body, 12 Troe, 0 Lindemann).
#include "../flint_device.h"
namespace flint_gen_tsr_rich_31 {
constexpr int NS = 31, NZ = 32, WK = 418, WH = 62, NSLOT = 79;

// Rates, net production and (JAC) the analytic Jacobian of the constant-volume RHS for one cell.
// r[k] = rho_k, T; outputs F[0..NS-1] = wdot, F[NS] = dT/dt (rhs_native); sink(row, col, v).
// RHS applies rho = max(rho, 0) as rhs_native does (Lib_Chemistry_rhs.f90:37-39) and adds dT/dt.
// With JAC, scr[s*SST] (s < NSLOT) is per-cell scratch for values reused by the columns.
// Returns a status: bit 0 T outside the kinetics table, bit 1 T outside the thermo table.
// With JAC, fout(F) receives F as soon as it is final, so that F need not stay live across the
// Jacobian columns.
template <bool JAC, bool RHS, int SST, class Sink, class FOut>
__device__ __forceinline__ int eval(const double* __restrict__ tabk, int Tk_min, int Tk_max,
                                    const double* __restrict__ tabh, int Th_min, int Th_max,
                                    const double (&r)[NS], double T, double (&F)[NZ],
                                    double* __restrict__ scr, Sink&& sink, FOut&& fout) {
  int status = 0;
  if (!(T >= (double)Tk_min && T <= (double)Tk_max)) status |= 1;
  if (RHS && !(T >= (double)Th_min && T <= (double)Th_max)) {
    status |= 2;
#pragma unroll
    for (int k = 0; k < NZ; ++k) F[k] = -1.0;
    if (JAC) fout(F);
    if (JAC) {
      for (int x = 0; x < NZ; ++x)
        for (int y = 0; y < NZ; ++y) sink(x, y, 0.0);
    }
    return status;
  }
  unsigned negmask = 0u;
  const double r0 = RHS ? fmax(r[0], 0.0) : r[0];
  const double r1 = RHS ? fmax(r[1], 0.0) : r[1];
  const double r2 = RHS ? fmax(r[2], 0.0) : r[2];
  const double r3 = RHS ? fmax(r[3], 0.0) : r[3];
  const double r4 = RHS ? fmax(r[4], 0.0) : r[4];
  const double r5 = RHS ? fmax(r[5], 0.0) : r[5];
  const double r6 = RHS ? fmax(r[6], 0.0) : r[6];
  const double r7 = RHS ? fmax(r[7], 0.0) : r[7];
  const double r8 = RHS ? fmax(r[8], 0.0) : r[8];
  const double r9 = RHS ? fmax(r[9], 0.0) : r[9];
  const double r10 = RHS ? fmax(r[10], 0.0) : r[10];
  const double r11 = RHS ? fmax(r[11], 0.0) : r[11];
  const double r12 = RHS ? fmax(r[12], 0.0) : r[12];
  const double r13 = RHS ? fmax(r[13], 0.0) : r[13];
  const double r14 = RHS ? fmax(r[14], 0.0) : r[14];
  const double r15 = RHS ? fmax(r[15], 0.0) : r[15];
  const double r16 = RHS ? fmax(r[16], 0.0) : r[16];
  const double r17 = RHS ? fmax(r[17], 0.0) : r[17];
  const double r18 = RHS ? fmax(r[18], 0.0) : r[18];
  const double r19 = RHS ? fmax(r[19], 0.0) : r[19];
  const double r20 = RHS ? fmax(r[20], 0.0) : r[20];
  const double r21 = RHS ? fmax(r[21], 0.0) : r[21];
  const double r22 = RHS ? fmax(r[22], 0.0) : r[22];
  const double r23 = RHS ? fmax(r[23], 0.0) : r[23];
  const double r24 = RHS ? fmax(r[24], 0.0) : r[24];
  const double r25 = RHS ? fmax(r[25], 0.0) : r[25];
  const double r26 = RHS ? fmax(r[26], 0.0) : r[26];
  const double r27 = RHS ? fmax(r[27], 0.0) : r[27];
  const double r28 = RHS ? fmax(r[28], 0.0) : r[28];
  const double r29 = RHS ? fmax(r[29], 0.0) : r[29];
  const double r30 = RHS ? fmax(r[30], 0.0) : r[30];
  if (r[0] < 0.0) negmask |= (1u << 0);
  if (r[1] < 0.0) negmask |= (1u << 1);
  if (r[2] < 0.0) negmask |= (1u << 2);
  if (r[3] < 0.0) negmask |= (1u << 3);
  if (r[4] < 0.0) negmask |= (1u << 4);
  if (r[5] < 0.0) negmask |= (1u << 5);
  if (r[6] < 0.0) negmask |= (1u << 6);
  if (r[7] < 0.0) negmask |= (1u << 7);
  if (r[8] < 0.0) negmask |= (1u << 8);
  if (r[9] < 0.0) negmask |= (1u << 9);
  if (r[10] < 0.0) negmask |= (1u << 10);
  if (r[11] < 0.0) negmask |= (1u << 11);
  if (r[12] < 0.0) negmask |= (1u << 12);
  if (r[13] < 0.0) negmask |= (1u << 13);
  if (r[14] < 0.0) negmask |= (1u << 14);
  if (r[15] < 0.0) negmask |= (1u << 15);
  if (r[16] < 0.0) negmask |= (1u << 16);
  if (r[17] < 0.0) negmask |= (1u << 17);
  if (r[18] < 0.0) negmask |= (1u << 18);
  if (r[19] < 0.0) negmask |= (1u << 19);
  if (r[20] < 0.0) negmask |= (1u << 20);
  if (r[21] < 0.0) negmask |= (1u << 21);
  if (r[22] < 0.0) negmask |= (1u << 22);
  if (r[23] < 0.0) negmask |= (1u << 23);
  if (r[24] < 0.0) negmask |= (1u << 24);
  if (r[25] < 0.0) negmask |= (1u << 25);
  if (r[26] < 0.0) negmask |= (1u << 26);
  if (r[27] < 0.0) negmask |= (1u << 27);
  if (r[28] < 0.0) negmask |= (1u << 28);
  if (r[29] < 0.0) negmask |= (1u << 29);
  if (r[30] < 0.0) negmask |= (1u << 30);
  const double c0 = r0 / 71.099;
  const double c1 = r1 / 43.045;
  const double c2 = r2 / 1.008;
  const double c3 = r3 / 16.043;
  const double c4 = r4 / 26.038;
  const double c5 = r5 / 39.057;
  const double c6 = r6 / 27.046;
  const double c7 = r7 / 15.999;
  const double c8 = r8 / 34.014;
  const double c9 = r9 / 28.01;
  const double c10 = r10 / 30.026;
  const double c11 = r11 / 30.07;
  const double c12 = r12 / 42.037;
  const double c13 = r13 / 43.089;
  const double c14 = r14 / 18.015;
  const double c15 = r15 / 31.998;
  const double c16 = r16 / 14.027;
  const double c17 = r17 / 29.018;
  const double c18 = r18 / 15.035;
  const double c19 = r19 / 2.016;
  const double c20 = r20 / 43.045;
  const double c21 = r21 / 58.08;
  const double c22 = r22 / 44.053;
  const double c23 = r23 / 43.045;
  const double c24 = r24 / 33.006;
  const double c25 = r25 / 28.054;
  const double c26 = r26 / 13.019;
  const double c27 = r27 / 17.007;
  const double c28 = r28 / 29.062;
  const double c29 = r29 / 14.027;
  const double c30 = r30 / 44.009;
  const int T_i = (int)T;
  const double Td = T - (double)T_i;
  const int rowk = min(max(T_i, Tk_min), Tk_max) - Tk_min;
  const double* __restrict__ rk0 = tabk + (long)rowk * WK;
  const double* __restrict__ rk1 = rk0 + WK;
  const double M0 = c0+c1+c2+c3*2.0+c4+c5+c6+c7+c8+c9*1.75+c10+c11*3.0+c12+c13+c14*15.399999618530273+c15+c16+c17+c18+c19*2.4000000953674316+c20+c21+c22+c23+c24+c25+c26+c27+c28+c29+c30*3.5999999046325684;
  const double M1 = c0+c1+c2+c3*2.0+c4+c5+c6+c7+c8+c9*1.5+c10+c11*3.0+c12+c13+c14*6.0+c15+c16+c17+c18+c19*2.0+c20+c21+c22+c23+c24+c25+c26+c27+c28+c29+c30*2.0;
  const double M2 = c0+c1+c2+c3*2.0+c4+c5+c6+c7+c8+c9*1.5+c10+c11*3.0+c12+c13+c14*6.0+c15*6.0+c16+c17+c18+c19*2.0+c20+c21+c22+c23+c24+c25+c26+c27+c28+c29+c30*3.5;
  const double M3 = c0+c1+c2+c3+c4+c5+c6+c7+c8+c9*0.75+c10+c11*1.5+c12+c13+c16+c17+c18+c19+c20+c21+c22+c23+c24+c25+c26+c27+c28+c29+c30*1.5;
  const double M4 = c15;
  const double M5 = c14;
  const double M6 = c0+c1+c2+c3*2.0+c4+c5+c6+c7+c8+c9+c10+c11*3.0+c12+c13+c15+c16+c17+c18+c20+c21+c22+c23+c24+c25+c26+c27+c28+c29;
  const double M7 = c19;
  const double M8 = c30;
  const double M9 = c0+c1+c2+c3*2.0+c4+c5+c6+c7+c8+c9+c10+c11*3.0+c12+c13+c14*3.6500000953674316+c15+c16+c17+c18+c19*0.7300000190734863+c20+c21+c22+c23+c24+c25+c26+c27+c28+c29+c30;
  const double M10 = c0+c1+c2+c3*2.0+c4+c5+c6+c7+c8+c9*1.5+c10+c11*3.0+c12+c13+c15+c16+c17+c18+c19*2.0+c20+c21+c22+c23+c24+c25+c26+c27+c28+c29+c30*2.0;
  const double M11 = c0+c1+c2+c3+c4+c5+c6+c7+c8+c9+c10+c11+c12+c13+c14+c15+c16+c17+c18+c19+c20+c21+c22+c23+c24+c25+c26+c27+c28+c29+c30;
  double w0;
  double w1;
  double w2;
  double w3;
  double w4;
  double w5;
  double w6;
  double w7;
  double w8;
  double w9;
  double w10;
  double w11;
  double w12;
  double w13;
  double w14;
  double w15;
  double w16;
  double w17;
  double w18;
  double w19;
  double w20;
  double w21;
  double w22;
  double w23;
  double w24;
  double w25;
  double w26;
  double w27;
  double w28;
  double w29;
  double w30;
  // JAC: dT-derivative accumulators
  [[maybe_unused]] double dT[NS] = {};
  {  // reaction 1: 2 O + M <=> O2 + M
    const double2 a0 = ld2(rk0 + 0), a1 = ld2(rk1 + 0);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c7*c7)*M0;
    const double pb = kb*(c15)*M0;
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c7*c7), Pb = (c15);
      const double dqT = ((a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb) * M0;
      scr[0 * SST] = kf * Pf - kb * Pb;
      dT[7] += -2.0 * dqT;
      dT[15] += 1.0 * dqT;
    }
    w7 = -2.0 * q;
    w15 = 1.0 * q;
  }
  {  // reaction 2: H + O + M <=> OH + M
    const double2 a0 = ld2(rk0 + 2), a1 = ld2(rk1 + 2);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c2)*(c7)*M1;
    const double pb = kb*(c27)*M1;
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c2)*(c7), Pb = (c27);
      const double dqT = ((a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb) * M1;
      scr[1 * SST] = kf * Pf - kb * Pb;
      dT[2] += -1.0 * dqT;
      dT[7] += -1.0 * dqT;
      dT[27] += 1.0 * dqT;
    }
    w2 = -1.0 * q;
    w7 = w7 + -1.0 * q;
    w27 = 1.0 * q;
  }
  {  // reaction 3: H2 + O <=> H + OH
    const double2 a0 = ld2(rk0 + 4), a1 = ld2(rk1 + 4);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    if (JAC) { scr[53 * SST] = kf; scr[54 * SST] = kb; }
    const double pf = kf*(c19)*(c7);
    const double pb = kb*(c2)*(c27);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c19)*(c7), Pb = (c2)*(c27);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += 1.0 * dqT;
      dT[7] += -1.0 * dqT;
      dT[19] += -1.0 * dqT;
      dT[27] += 1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w7 = w7 + -1.0 * q;
    w19 = -1.0 * q;
    w27 = w27 + 1.0 * q;
  }
  {  // reaction 4: HO2 + O <=> O2 + OH
    const double2 a0 = ld2(rk0 + 6), a1 = ld2(rk1 + 6);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    if (JAC) { scr[55 * SST] = kf; scr[56 * SST] = kb; }
    const double pf = kf*(c24)*(c7);
    const double pb = kb*(c15)*(c27);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c24)*(c7), Pb = (c15)*(c27);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[7] += -1.0 * dqT;
      dT[15] += 1.0 * dqT;
      dT[24] += -1.0 * dqT;
      dT[27] += 1.0 * dqT;
    }
    w7 = w7 + -1.0 * q;
    w15 = w15 + 1.0 * q;
    w24 = -1.0 * q;
    w27 = w27 + 1.0 * q;
  }
  {  // reaction 5: H2O2 + O <=> HO2 + OH
    const double2 a0 = ld2(rk0 + 8), a1 = ld2(rk1 + 8);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    if (JAC) { scr[57 * SST] = kf; scr[58 * SST] = kb; }
    const double pf = kf*(c8)*(c7);
    const double pb = kb*(c24)*(c27);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c8)*(c7), Pb = (c24)*(c27);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[7] += -1.0 * dqT;
      dT[8] += -1.0 * dqT;
      dT[24] += 1.0 * dqT;
      dT[27] += 1.0 * dqT;
    }
    w7 = w7 + -1.0 * q;
    w8 = -1.0 * q;
    w24 = w24 + 1.0 * q;
    w27 = w27 + 1.0 * q;
  }
  {  // reaction 6: CH + O <=> CO + H
    const double2 a0 = ld2(rk0 + 10), a1 = ld2(rk1 + 10);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    if (JAC) { scr[59 * SST] = kf; scr[60 * SST] = kb; }
    const double pf = kf*(c26)*(c7);
    const double pb = kb*(c9)*(c2);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c26)*(c7), Pb = (c9)*(c2);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += 1.0 * dqT;
      dT[7] += -1.0 * dqT;
      dT[9] += 1.0 * dqT;
      dT[26] += -1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w7 = w7 + -1.0 * q;
    w9 = 1.0 * q;
    w26 = -1.0 * q;
  }
  {  // reaction 7: CH2 + O <=> H + HCO
    const double2 a0 = ld2(rk0 + 12), a1 = ld2(rk1 + 12);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    if (JAC) { scr[61 * SST] = kf; scr[62 * SST] = kb; }
    const double pf = kf*(c29)*(c7);
    const double pb = kb*(c2)*(c17);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c29)*(c7), Pb = (c2)*(c17);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += 1.0 * dqT;
      dT[7] += -1.0 * dqT;
      dT[17] += 1.0 * dqT;
      dT[29] += -1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w7 = w7 + -1.0 * q;
    w17 = 1.0 * q;
    w29 = -1.0 * q;
  }
  {  // reaction 8: CH2(S) + O <=> CO + H2
    const double2 a0 = ld2(rk0 + 14), a1 = ld2(rk1 + 14);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    if (JAC) { scr[63 * SST] = kf; scr[64 * SST] = kb; }
    const double pf = kf*(c16)*(c7);
    const double pb = kb*(c9)*(c19);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c16)*(c7), Pb = (c9)*(c19);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[7] += -1.0 * dqT;
      dT[9] += 1.0 * dqT;
      dT[16] += -1.0 * dqT;
      dT[19] += 1.0 * dqT;
    }
    w7 = w7 + -1.0 * q;
    w9 = w9 + 1.0 * q;
    w16 = -1.0 * q;
    w19 = w19 + 1.0 * q;
  }
  {  // reaction 9: CH2(S) + O <=> H + HCO
    const double2 a0 = ld2(rk0 + 16), a1 = ld2(rk1 + 16);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    if (JAC) { scr[65 * SST] = kf; scr[66 * SST] = kb; }
    const double pf = kf*(c16)*(c7);
    const double pb = kb*(c2)*(c17);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c16)*(c7), Pb = (c2)*(c17);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += 1.0 * dqT;
      dT[7] += -1.0 * dqT;
      dT[16] += -1.0 * dqT;
      dT[17] += 1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w7 = w7 + -1.0 * q;
    w16 = w16 + -1.0 * q;
    w17 = w17 + 1.0 * q;
  }
  {  // reaction 10: CH3 + O <=> CH2O + H
    const double2 a0 = ld2(rk0 + 18), a1 = ld2(rk1 + 18);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    if (JAC) { scr[67 * SST] = kf; scr[68 * SST] = kb; }
    const double pf = kf*(c18)*(c7);
    const double pb = kb*(c10)*(c2);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c18)*(c7), Pb = (c10)*(c2);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += 1.0 * dqT;
      dT[7] += -1.0 * dqT;
      dT[10] += 1.0 * dqT;
      dT[18] += -1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w7 = w7 + -1.0 * q;
    w10 = 1.0 * q;
    w18 = -1.0 * q;
  }
  {  // reaction 11: CH4 + O <=> CH3 + OH
    const double2 a0 = ld2(rk0 + 20), a1 = ld2(rk1 + 20);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    if (JAC) { scr[69 * SST] = kf; scr[70 * SST] = kb; }
    const double pf = kf*(c3)*(c7);
    const double pb = kb*(c18)*(c27);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c3)*(c7), Pb = (c18)*(c27);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[3] += -1.0 * dqT;
      dT[7] += -1.0 * dqT;
      dT[18] += 1.0 * dqT;
      dT[27] += 1.0 * dqT;
    }
    w3 = -1.0 * q;
    w7 = w7 + -1.0 * q;
    w18 = w18 + 1.0 * q;
    w27 = w27 + 1.0 * q;
  }
  {  // reaction 12: CO + O + M <=> CO2 + M
    const double2 a0 = ld2(rk0 + 22), a1 = ld2(rk1 + 22);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c9)*(c7)*M2;
    const double pb = kb*(c30)*M2;
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c9)*(c7), Pb = (c30);
      const double dqT = ((a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb) * M2;
      scr[2 * SST] = kf * Pf - kb * Pb;
      dT[7] += -1.0 * dqT;
      dT[9] += -1.0 * dqT;
      dT[30] += 1.0 * dqT;
    }
    w7 = w7 + -1.0 * q;
    w9 = w9 + -1.0 * q;
    w30 = 1.0 * q;
  }
  {  // reaction 13: HCO + O <=> CO + OH
    const double2 a0 = ld2(rk0 + 24), a1 = ld2(rk1 + 24);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    if (JAC) { scr[71 * SST] = kf; scr[72 * SST] = kb; }
    const double pf = kf*(c17)*(c7);
    const double pb = kb*(c9)*(c27);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c17)*(c7), Pb = (c9)*(c27);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[7] += -1.0 * dqT;
      dT[9] += 1.0 * dqT;
      dT[17] += -1.0 * dqT;
      dT[27] += 1.0 * dqT;
    }
    w7 = w7 + -1.0 * q;
    w9 = w9 + 1.0 * q;
    w17 = w17 + -1.0 * q;
    w27 = w27 + 1.0 * q;
  }
  {  // reaction 14: HCO + O <=> CO2 + H
    const double2 a0 = ld2(rk0 + 26), a1 = ld2(rk1 + 26);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    if (JAC) { scr[73 * SST] = kf; scr[74 * SST] = kb; }
    const double pf = kf*(c17)*(c7);
    const double pb = kb*(c30)*(c2);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c17)*(c7), Pb = (c30)*(c2);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += 1.0 * dqT;
      dT[7] += -1.0 * dqT;
      dT[17] += -1.0 * dqT;
      dT[30] += 1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w7 = w7 + -1.0 * q;
    w17 = w17 + -1.0 * q;
    w30 = w30 + 1.0 * q;
  }
  {  // reaction 15: CH2O + O <=> HCO + OH
    const double2 a0 = ld2(rk0 + 28), a1 = ld2(rk1 + 28);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    if (JAC) { scr[75 * SST] = kf; scr[76 * SST] = kb; }
    const double pf = kf*(c10)*(c7);
    const double pb = kb*(c17)*(c27);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c10)*(c7), Pb = (c17)*(c27);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[7] += -1.0 * dqT;
      dT[10] += -1.0 * dqT;
      dT[17] += 1.0 * dqT;
      dT[27] += 1.0 * dqT;
    }
    w7 = w7 + -1.0 * q;
    w10 = w10 + -1.0 * q;
    w17 = w17 + 1.0 * q;
    w27 = w27 + 1.0 * q;
  }
  {  // reaction 16: C2H2 + O <=> CH2 + CO
    const double2 a0 = ld2(rk0 + 30), a1 = ld2(rk1 + 30);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    if (JAC) { scr[77 * SST] = kf; scr[78 * SST] = kb; }
    const double pf = kf*(c4)*(c7);
    const double pb = kb*(c29)*(c9);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c4)*(c7), Pb = (c29)*(c9);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[4] += -1.0 * dqT;
      dT[7] += -1.0 * dqT;
      dT[9] += 1.0 * dqT;
      dT[29] += 1.0 * dqT;
    }
    w4 = -1.0 * q;
    w7 = w7 + -1.0 * q;
    w9 = w9 + 1.0 * q;
    w29 = w29 + 1.0 * q;
  }
  {  // reaction 17: C2H3 + O <=> CH2CO + H
    const double2 a0 = ld2(rk0 + 32), a1 = ld2(rk1 + 32);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c6)*(c7);
    const double pb = kb*(c12)*(c2);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c6)*(c7), Pb = (c12)*(c2);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += 1.0 * dqT;
      dT[6] += -1.0 * dqT;
      dT[7] += -1.0 * dqT;
      dT[12] += 1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w6 = -1.0 * q;
    w7 = w7 + -1.0 * q;
    w12 = 1.0 * q;
  }
  {  // reaction 18: C2H4 + O <=> CH3 + HCO
    const double2 a0 = ld2(rk0 + 34), a1 = ld2(rk1 + 34);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c25)*(c7);
    const double pb = kb*(c18)*(c17);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c25)*(c7), Pb = (c18)*(c17);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[7] += -1.0 * dqT;
      dT[17] += 1.0 * dqT;
      dT[18] += 1.0 * dqT;
      dT[25] += -1.0 * dqT;
    }
    w7 = w7 + -1.0 * q;
    w17 = w17 + 1.0 * q;
    w18 = w18 + 1.0 * q;
    w25 = -1.0 * q;
  }
  {  // reaction 19: C2H5 + O <=> CH2O + CH3
    const double2 a0 = ld2(rk0 + 36), a1 = ld2(rk1 + 36);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c28)*(c7);
    const double pb = kb*(c10)*(c18);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c28)*(c7), Pb = (c10)*(c18);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[7] += -1.0 * dqT;
      dT[10] += 1.0 * dqT;
      dT[18] += 1.0 * dqT;
      dT[28] += -1.0 * dqT;
    }
    w7 = w7 + -1.0 * q;
    w10 = w10 + 1.0 * q;
    w18 = w18 + 1.0 * q;
    w28 = -1.0 * q;
  }
  {  // reaction 20: C2H6 + O <=> C2H5 + OH
    const double2 a0 = ld2(rk0 + 38), a1 = ld2(rk1 + 38);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c11)*(c7);
    const double pb = kb*(c28)*(c27);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c11)*(c7), Pb = (c28)*(c27);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[7] += -1.0 * dqT;
      dT[11] += -1.0 * dqT;
      dT[27] += 1.0 * dqT;
      dT[28] += 1.0 * dqT;
    }
    w7 = w7 + -1.0 * q;
    w11 = -1.0 * q;
    w27 = w27 + 1.0 * q;
    w28 = w28 + 1.0 * q;
  }
  {  // reaction 21: CH2CO + O <=> CH2 + CO2
    const double2 a0 = ld2(rk0 + 40), a1 = ld2(rk1 + 40);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c12)*(c7);
    const double pb = kb*(c29)*(c30);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c12)*(c7), Pb = (c29)*(c30);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[7] += -1.0 * dqT;
      dT[12] += -1.0 * dqT;
      dT[29] += 1.0 * dqT;
      dT[30] += 1.0 * dqT;
    }
    w7 = w7 + -1.0 * q;
    w12 = w12 + -1.0 * q;
    w29 = w29 + 1.0 * q;
    w30 = w30 + 1.0 * q;
  }
  {  // reaction 22: CO + O2 <=> CO2 + O
    const double2 a0 = ld2(rk0 + 42), a1 = ld2(rk1 + 42);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c9)*(c15);
    const double pb = kb*(c30)*(c7);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c9)*(c15), Pb = (c30)*(c7);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[7] += 1.0 * dqT;
      dT[9] += -1.0 * dqT;
      dT[15] += -1.0 * dqT;
      dT[30] += 1.0 * dqT;
    }
    w7 = w7 + 1.0 * q;
    w9 = w9 + -1.0 * q;
    w15 = w15 + -1.0 * q;
    w30 = w30 + 1.0 * q;
  }
  {  // reaction 23: CH2O + O2 <=> HCO + HO2
    const double2 a0 = ld2(rk0 + 44), a1 = ld2(rk1 + 44);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c10)*(c15);
    const double pb = kb*(c17)*(c24);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c10)*(c15), Pb = (c17)*(c24);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[10] += -1.0 * dqT;
      dT[15] += -1.0 * dqT;
      dT[17] += 1.0 * dqT;
      dT[24] += 1.0 * dqT;
    }
    w10 = w10 + -1.0 * q;
    w15 = w15 + -1.0 * q;
    w17 = w17 + 1.0 * q;
    w24 = w24 + 1.0 * q;
  }
  {  // reaction 24: H + O2 + M <=> HO2 + M
    const double2 a0 = ld2(rk0 + 46), a1 = ld2(rk1 + 46);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c2)*(c15)*M3;
    const double pb = kb*(c24)*M3;
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c2)*(c15), Pb = (c24);
      const double dqT = ((a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb) * M3;
      scr[3 * SST] = kf * Pf - kb * Pb;
      dT[2] += -1.0 * dqT;
      dT[15] += -1.0 * dqT;
      dT[24] += 1.0 * dqT;
    }
    w2 = w2 + -1.0 * q;
    w15 = w15 + -1.0 * q;
    w24 = w24 + 1.0 * q;
  }
  {  // reaction 25: H + O2 + O2 <=> HO2 + O2
    const double2 a0 = ld2(rk0 + 48), a1 = ld2(rk1 + 48);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c2)*(c15)*M4;
    const double pb = kb*(c24)*M4;
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c2)*(c15), Pb = (c24);
      const double dqT = ((a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb) * M4;
      scr[4 * SST] = kf * Pf - kb * Pb;
      dT[2] += -1.0 * dqT;
      dT[15] += -1.0 * dqT;
      dT[24] += 1.0 * dqT;
    }
    w2 = w2 + -1.0 * q;
    w15 = w15 + -1.0 * q;
    w24 = w24 + 1.0 * q;
  }
  {  // reaction 26: H + O2 + H2O <=> HO2 + H2O
    const double2 a0 = ld2(rk0 + 50), a1 = ld2(rk1 + 50);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c2)*(c15)*M5;
    const double pb = kb*(c24)*M5;
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c2)*(c15), Pb = (c24);
      const double dqT = ((a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb) * M5;
      scr[5 * SST] = kf * Pf - kb * Pb;
      dT[2] += -1.0 * dqT;
      dT[15] += -1.0 * dqT;
      dT[24] += 1.0 * dqT;
    }
    w2 = w2 + -1.0 * q;
    w15 = w15 + -1.0 * q;
    w24 = w24 + 1.0 * q;
  }
  {  // reaction 27: H + O2 <=> O + OH
    const double2 a0 = ld2(rk0 + 52), a1 = ld2(rk1 + 52);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c2)*(c15);
    const double pb = kb*(c7)*(c27);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c2)*(c15), Pb = (c7)*(c27);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += -1.0 * dqT;
      dT[7] += 1.0 * dqT;
      dT[15] += -1.0 * dqT;
      dT[27] += 1.0 * dqT;
    }
    w2 = w2 + -1.0 * q;
    w7 = w7 + 1.0 * q;
    w15 = w15 + -1.0 * q;
    w27 = w27 + 1.0 * q;
  }
  {  // reaction 28: 2 H + M <=> H2 + M
    const double2 a0 = ld2(rk0 + 54), a1 = ld2(rk1 + 54);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c2*c2)*M6;
    const double pb = kb*(c19)*M6;
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c2*c2), Pb = (c19);
      const double dqT = ((a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb) * M6;
      scr[6 * SST] = kf * Pf - kb * Pb;
      dT[2] += -2.0 * dqT;
      dT[19] += 1.0 * dqT;
    }
    w2 = w2 + -2.0 * q;
    w19 = w19 + 1.0 * q;
  }
  {  // reaction 29: 2 H + H2 <=> H2 + H2
    const double2 a0 = ld2(rk0 + 56), a1 = ld2(rk1 + 56);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c2*c2)*M7;
    const double pb = kb*(c19)*M7;
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c2*c2), Pb = (c19);
      const double dqT = ((a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb) * M7;
      scr[7 * SST] = kf * Pf - kb * Pb;
      dT[2] += -2.0 * dqT;
      dT[19] += 1.0 * dqT;
    }
    w2 = w2 + -2.0 * q;
    w19 = w19 + 1.0 * q;
  }
  {  // reaction 30: 2 H + H2O <=> H2 + H2O
    const double2 a0 = ld2(rk0 + 58), a1 = ld2(rk1 + 58);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c2*c2)*M5;
    const double pb = kb*(c19)*M5;
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c2*c2), Pb = (c19);
      const double dqT = ((a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb) * M5;
      scr[8 * SST] = kf * Pf - kb * Pb;
      dT[2] += -2.0 * dqT;
      dT[19] += 1.0 * dqT;
    }
    w2 = w2 + -2.0 * q;
    w19 = w19 + 1.0 * q;
  }
  {  // reaction 31: 2 H + CO2 <=> H2 + CO2
    const double2 a0 = ld2(rk0 + 60), a1 = ld2(rk1 + 60);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c2*c2)*M8;
    const double pb = kb*(c19)*M8;
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c2*c2), Pb = (c19);
      const double dqT = ((a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb) * M8;
      scr[9 * SST] = kf * Pf - kb * Pb;
      dT[2] += -2.0 * dqT;
      dT[19] += 1.0 * dqT;
    }
    w2 = w2 + -2.0 * q;
    w19 = w19 + 1.0 * q;
  }
  {  // reaction 32: H + OH + M <=> H2O + M
    const double2 a0 = ld2(rk0 + 62), a1 = ld2(rk1 + 62);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c2)*(c27)*M9;
    const double pb = kb*(c14)*M9;
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c2)*(c27), Pb = (c14);
      const double dqT = ((a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb) * M9;
      scr[10 * SST] = kf * Pf - kb * Pb;
      dT[2] += -1.0 * dqT;
      dT[14] += 1.0 * dqT;
      dT[27] += -1.0 * dqT;
    }
    w2 = w2 + -1.0 * q;
    w14 = 1.0 * q;
    w27 = w27 + -1.0 * q;
  }
  {  // reaction 33: H + HO2 <=> H2O + O
    const double2 a0 = ld2(rk0 + 64), a1 = ld2(rk1 + 64);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c2)*(c24);
    const double pb = kb*(c14)*(c7);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c2)*(c24), Pb = (c14)*(c7);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += -1.0 * dqT;
      dT[7] += 1.0 * dqT;
      dT[14] += 1.0 * dqT;
      dT[24] += -1.0 * dqT;
    }
    w2 = w2 + -1.0 * q;
    w7 = w7 + 1.0 * q;
    w14 = w14 + 1.0 * q;
    w24 = w24 + -1.0 * q;
  }
  {  // reaction 34: H + HO2 <=> H2 + O2
    const double2 a0 = ld2(rk0 + 66), a1 = ld2(rk1 + 66);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c2)*(c24);
    const double pb = kb*(c19)*(c15);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c2)*(c24), Pb = (c19)*(c15);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += -1.0 * dqT;
      dT[15] += 1.0 * dqT;
      dT[19] += 1.0 * dqT;
      dT[24] += -1.0 * dqT;
    }
    w2 = w2 + -1.0 * q;
    w15 = w15 + 1.0 * q;
    w19 = w19 + 1.0 * q;
    w24 = w24 + -1.0 * q;
  }
  {  // reaction 35: H + HO2 <=> 2 OH
    const double2 a0 = ld2(rk0 + 68), a1 = ld2(rk1 + 68);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c2)*(c24);
    const double pb = kb*(c27*c27);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c2)*(c24), Pb = (c27*c27);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += -1.0 * dqT;
      dT[24] += -1.0 * dqT;
      dT[27] += 2.0 * dqT;
    }
    w2 = w2 + -1.0 * q;
    w24 = w24 + -1.0 * q;
    w27 = w27 + 2.0 * q;
  }
  {  // reaction 36: H + H2O2 <=> H2 + HO2
    const double2 a0 = ld2(rk0 + 70), a1 = ld2(rk1 + 70);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c2)*(c8);
    const double pb = kb*(c19)*(c24);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c2)*(c8), Pb = (c19)*(c24);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += -1.0 * dqT;
      dT[8] += -1.0 * dqT;
      dT[19] += 1.0 * dqT;
      dT[24] += 1.0 * dqT;
    }
    w2 = w2 + -1.0 * q;
    w8 = w8 + -1.0 * q;
    w19 = w19 + 1.0 * q;
    w24 = w24 + 1.0 * q;
  }
  {  // reaction 37: H + H2O2 <=> H2O + OH
    const double2 a0 = ld2(rk0 + 72), a1 = ld2(rk1 + 72);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c2)*(c8);
    const double pb = kb*(c14)*(c27);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c2)*(c8), Pb = (c14)*(c27);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += -1.0 * dqT;
      dT[8] += -1.0 * dqT;
      dT[14] += 1.0 * dqT;
      dT[27] += 1.0 * dqT;
    }
    w2 = w2 + -1.0 * q;
    w8 = w8 + -1.0 * q;
    w14 = w14 + 1.0 * q;
    w27 = w27 + 1.0 * q;
  }
  {  // reaction 38: CH2 + H (+M) <=> CH3 (+M)
    const double2 x0 = ld2(rk0 + 370), y0 = ld2(rk0 + 372);
    const double2 x1 = ld2(rk1 + 370), y1 = ld2(rk1 + 372);
    const double kinf = x0.x + (x1.x - x0.x) * Td, k0 = x0.y + (x1.y - x0.y) * Td;
    const double kc = y0.x + (y1.x - y0.x) * Td, Fc = y0.y + (y1.y - y0.y) * Td;
    const double Mx = M1;
    const double Pr = k0 * Mx / kinf;
    const double lFc = log10(Fc), lPr = log10(Pr);
    const double cc = -0.4 - 0.67 * lFc, nn = 0.75 - 1.27 * lFc;
    const double f1 = (lPr + cc) / (nn - 0.14 * (lPr + cc));
    const double Fb = exp10(lFc / (1.0 + f1 * f1));
    const double k1 = kinf * (Pr / (1.0 + Pr)) * Fb;
    const double k2 = k1 / kc;
    const double pf = k1*(c29)*(c2);
    const double pb = k2*(c18);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c29)*(c2), Pb = (c18);
      const double g = Pr / (1.0 + Pr), dg = 1.0 / ((1.0 + Pr) * (1.0 + Pr));
      const double u = lPr + cc, den = nn - 0.14 * u, qq = 1.0 + f1 * f1;
      const double dlnF_dPr = -2.0 * lFc * f1 * (nn / (den * den)) / (qq * qq * Pr);
      const double df1_dL = (-0.67 * nn + 1.27 * u) / (den * den);
      const double dlnF_dFc = (1.0 / qq - lFc * 2.0 * f1 * df1_dL / (qq * qq)) / Fc;
      const double dk1_dPr = kinf * Fb * (dg + g * dlnF_dPr);
      const double dPr_dT = Mx * ((x1.y - x0.y) * kinf - k0 * (x1.x - x0.x)) / (kinf * kinf);
      const double dk1T = (x1.x - x0.x) * g * Fb + dk1_dPr * dPr_dT + k1 * dlnF_dFc * (y1.y - y0.y);
      const double dk2T = dk1T / kc - k1 * (y1.x - y0.x) / (kc * kc);
      const double dk1M = dk1_dPr * k0 / kinf;
      scr[11 * SST] = dk1M * Pf - (dk1M / kc) * Pb;
      scr[12 * SST] = k1; scr[13 * SST] = k2;
      const double dqT = dk1T * Pf - dk2T * Pb;
      dT[2] += -1.0 * dqT;
      dT[18] += 1.0 * dqT;
      dT[29] += -1.0 * dqT;
    }
    w2 = w2 + -1.0 * q;
    w18 = w18 + 1.0 * q;
    w29 = w29 + -1.0 * q;
  }
  {  // reaction 39: CH2(S) + H <=> CH + H2
    const double2 a0 = ld2(rk0 + 74), a1 = ld2(rk1 + 74);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c16)*(c2);
    const double pb = kb*(c26)*(c19);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c16)*(c2), Pb = (c26)*(c19);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += -1.0 * dqT;
      dT[16] += -1.0 * dqT;
      dT[19] += 1.0 * dqT;
      dT[26] += 1.0 * dqT;
    }
    w2 = w2 + -1.0 * q;
    w16 = w16 + -1.0 * q;
    w19 = w19 + 1.0 * q;
    w26 = w26 + 1.0 * q;
  }
  {  // reaction 40: CH3 + H (+M) <=> CH4 (+M)
    const double2 x0 = ld2(rk0 + 374), y0 = ld2(rk0 + 376);
    const double2 x1 = ld2(rk1 + 374), y1 = ld2(rk1 + 376);
    const double kinf = x0.x + (x1.x - x0.x) * Td, k0 = x0.y + (x1.y - x0.y) * Td;
    const double kc = y0.x + (y1.x - y0.x) * Td, Fc = y0.y + (y1.y - y0.y) * Td;
    const double Mx = M1;
    const double Pr = k0 * Mx / kinf;
    const double lFc = log10(Fc), lPr = log10(Pr);
    const double cc = -0.4 - 0.67 * lFc, nn = 0.75 - 1.27 * lFc;
    const double f1 = (lPr + cc) / (nn - 0.14 * (lPr + cc));
    const double Fb = exp10(lFc / (1.0 + f1 * f1));
    const double k1 = kinf * (Pr / (1.0 + Pr)) * Fb;
    const double k2 = k1 / kc;
    const double pf = k1*(c18)*(c2);
    const double pb = k2*(c3);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c18)*(c2), Pb = (c3);
      const double g = Pr / (1.0 + Pr), dg = 1.0 / ((1.0 + Pr) * (1.0 + Pr));
      const double u = lPr + cc, den = nn - 0.14 * u, qq = 1.0 + f1 * f1;
      const double dlnF_dPr = -2.0 * lFc * f1 * (nn / (den * den)) / (qq * qq * Pr);
      const double df1_dL = (-0.67 * nn + 1.27 * u) / (den * den);
      const double dlnF_dFc = (1.0 / qq - lFc * 2.0 * f1 * df1_dL / (qq * qq)) / Fc;
      const double dk1_dPr = kinf * Fb * (dg + g * dlnF_dPr);
      const double dPr_dT = Mx * ((x1.y - x0.y) * kinf - k0 * (x1.x - x0.x)) / (kinf * kinf);
      const double dk1T = (x1.x - x0.x) * g * Fb + dk1_dPr * dPr_dT + k1 * dlnF_dFc * (y1.y - y0.y);
      const double dk2T = dk1T / kc - k1 * (y1.x - y0.x) / (kc * kc);
      const double dk1M = dk1_dPr * k0 / kinf;
      scr[14 * SST] = dk1M * Pf - (dk1M / kc) * Pb;
      scr[15 * SST] = k1; scr[16 * SST] = k2;
      const double dqT = dk1T * Pf - dk2T * Pb;
      dT[2] += -1.0 * dqT;
      dT[3] += 1.0 * dqT;
      dT[18] += -1.0 * dqT;
    }
    w2 = w2 + -1.0 * q;
    w3 = w3 + 1.0 * q;
    w18 = w18 + -1.0 * q;
  }
  {  // reaction 41: CH4 + H <=> CH3 + H2
    const double2 a0 = ld2(rk0 + 76), a1 = ld2(rk1 + 76);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c3)*(c2);
    const double pb = kb*(c18)*(c19);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c3)*(c2), Pb = (c18)*(c19);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += -1.0 * dqT;
      dT[3] += -1.0 * dqT;
      dT[18] += 1.0 * dqT;
      dT[19] += 1.0 * dqT;
    }
    w2 = w2 + -1.0 * q;
    w3 = w3 + -1.0 * q;
    w18 = w18 + 1.0 * q;
    w19 = w19 + 1.0 * q;
  }
  {  // reaction 42: H + HCO (+M) <=> CH2O (+M)
    const double2 x0 = ld2(rk0 + 378), y0 = ld2(rk0 + 380);
    const double2 x1 = ld2(rk1 + 378), y1 = ld2(rk1 + 380);
    const double kinf = x0.x + (x1.x - x0.x) * Td, k0 = x0.y + (x1.y - x0.y) * Td;
    const double kc = y0.x + (y1.x - y0.x) * Td, Fc = y0.y + (y1.y - y0.y) * Td;
    const double Mx = M1;
    const double Pr = k0 * Mx / kinf;
    const double lFc = log10(Fc), lPr = log10(Pr);
    const double cc = -0.4 - 0.67 * lFc, nn = 0.75 - 1.27 * lFc;
    const double f1 = (lPr + cc) / (nn - 0.14 * (lPr + cc));
    const double Fb = exp10(lFc / (1.0 + f1 * f1));
    const double k1 = kinf * (Pr / (1.0 + Pr)) * Fb;
    const double k2 = k1 / kc;
    const double pf = k1*(c2)*(c17);
    const double pb = k2*(c10);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c2)*(c17), Pb = (c10);
      const double g = Pr / (1.0 + Pr), dg = 1.0 / ((1.0 + Pr) * (1.0 + Pr));
      const double u = lPr + cc, den = nn - 0.14 * u, qq = 1.0 + f1 * f1;
      const double dlnF_dPr = -2.0 * lFc * f1 * (nn / (den * den)) / (qq * qq * Pr);
      const double df1_dL = (-0.67 * nn + 1.27 * u) / (den * den);
      const double dlnF_dFc = (1.0 / qq - lFc * 2.0 * f1 * df1_dL / (qq * qq)) / Fc;
      const double dk1_dPr = kinf * Fb * (dg + g * dlnF_dPr);
      const double dPr_dT = Mx * ((x1.y - x0.y) * kinf - k0 * (x1.x - x0.x)) / (kinf * kinf);
      const double dk1T = (x1.x - x0.x) * g * Fb + dk1_dPr * dPr_dT + k1 * dlnF_dFc * (y1.y - y0.y);
      const double dk2T = dk1T / kc - k1 * (y1.x - y0.x) / (kc * kc);
      const double dk1M = dk1_dPr * k0 / kinf;
      scr[17 * SST] = dk1M * Pf - (dk1M / kc) * Pb;
      scr[18 * SST] = k1; scr[19 * SST] = k2;
      const double dqT = dk1T * Pf - dk2T * Pb;
      dT[2] += -1.0 * dqT;
      dT[10] += 1.0 * dqT;
      dT[17] += -1.0 * dqT;
    }
    w2 = w2 + -1.0 * q;
    w10 = w10 + 1.0 * q;
    w17 = w17 + -1.0 * q;
  }
  {  // reaction 43: H + HCO <=> CO + H2
    const double2 a0 = ld2(rk0 + 78), a1 = ld2(rk1 + 78);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c2)*(c17);
    const double pb = kb*(c9)*(c19);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c2)*(c17), Pb = (c9)*(c19);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += -1.0 * dqT;
      dT[9] += 1.0 * dqT;
      dT[17] += -1.0 * dqT;
      dT[19] += 1.0 * dqT;
    }
    w2 = w2 + -1.0 * q;
    w9 = w9 + 1.0 * q;
    w17 = w17 + -1.0 * q;
    w19 = w19 + 1.0 * q;
  }
  {  // reaction 44: CH2O + H <=> H2 + HCO
    const double2 a0 = ld2(rk0 + 80), a1 = ld2(rk1 + 80);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c10)*(c2);
    const double pb = kb*(c19)*(c17);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c10)*(c2), Pb = (c19)*(c17);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += -1.0 * dqT;
      dT[10] += -1.0 * dqT;
      dT[17] += 1.0 * dqT;
      dT[19] += 1.0 * dqT;
    }
    w2 = w2 + -1.0 * q;
    w10 = w10 + -1.0 * q;
    w17 = w17 + 1.0 * q;
    w19 = w19 + 1.0 * q;
  }
  {  // reaction 45: C2H2 + H (+M) <=> C2H3 (+M)
    const double2 x0 = ld2(rk0 + 382), y0 = ld2(rk0 + 384);
    const double2 x1 = ld2(rk1 + 382), y1 = ld2(rk1 + 384);
    const double kinf = x0.x + (x1.x - x0.x) * Td, k0 = x0.y + (x1.y - x0.y) * Td;
    const double kc = y0.x + (y1.x - y0.x) * Td, Fc = y0.y + (y1.y - y0.y) * Td;
    const double Mx = M1;
    const double Pr = k0 * Mx / kinf;
    const double lFc = log10(Fc), lPr = log10(Pr);
    const double cc = -0.4 - 0.67 * lFc, nn = 0.75 - 1.27 * lFc;
    const double f1 = (lPr + cc) / (nn - 0.14 * (lPr + cc));
    const double Fb = exp10(lFc / (1.0 + f1 * f1));
    const double k1 = kinf * (Pr / (1.0 + Pr)) * Fb;
    const double k2 = k1 / kc;
    const double pf = k1*(c4)*(c2);
    const double pb = k2*(c6);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c4)*(c2), Pb = (c6);
      const double g = Pr / (1.0 + Pr), dg = 1.0 / ((1.0 + Pr) * (1.0 + Pr));
      const double u = lPr + cc, den = nn - 0.14 * u, qq = 1.0 + f1 * f1;
      const double dlnF_dPr = -2.0 * lFc * f1 * (nn / (den * den)) / (qq * qq * Pr);
      const double df1_dL = (-0.67 * nn + 1.27 * u) / (den * den);
      const double dlnF_dFc = (1.0 / qq - lFc * 2.0 * f1 * df1_dL / (qq * qq)) / Fc;
      const double dk1_dPr = kinf * Fb * (dg + g * dlnF_dPr);
      const double dPr_dT = Mx * ((x1.y - x0.y) * kinf - k0 * (x1.x - x0.x)) / (kinf * kinf);
      const double dk1T = (x1.x - x0.x) * g * Fb + dk1_dPr * dPr_dT + k1 * dlnF_dFc * (y1.y - y0.y);
      const double dk2T = dk1T / kc - k1 * (y1.x - y0.x) / (kc * kc);
      const double dk1M = dk1_dPr * k0 / kinf;
      scr[20 * SST] = dk1M * Pf - (dk1M / kc) * Pb;
      scr[21 * SST] = k1; scr[22 * SST] = k2;
      const double dqT = dk1T * Pf - dk2T * Pb;
      dT[2] += -1.0 * dqT;
      dT[4] += -1.0 * dqT;
      dT[6] += 1.0 * dqT;
    }
    w2 = w2 + -1.0 * q;
    w4 = w4 + -1.0 * q;
    w6 = w6 + 1.0 * q;
  }
  {  // reaction 46: C2H3 + H (+M) <=> C2H4 (+M)
    const double2 x0 = ld2(rk0 + 386), y0 = ld2(rk0 + 388);
    const double2 x1 = ld2(rk1 + 386), y1 = ld2(rk1 + 388);
    const double kinf = x0.x + (x1.x - x0.x) * Td, k0 = x0.y + (x1.y - x0.y) * Td;
    const double kc = y0.x + (y1.x - y0.x) * Td, Fc = y0.y + (y1.y - y0.y) * Td;
    const double Mx = M1;
    const double Pr = k0 * Mx / kinf;
    const double lFc = log10(Fc), lPr = log10(Pr);
    const double cc = -0.4 - 0.67 * lFc, nn = 0.75 - 1.27 * lFc;
    const double f1 = (lPr + cc) / (nn - 0.14 * (lPr + cc));
    const double Fb = exp10(lFc / (1.0 + f1 * f1));
    const double k1 = kinf * (Pr / (1.0 + Pr)) * Fb;
    const double k2 = k1 / kc;
    const double pf = k1*(c6)*(c2);
    const double pb = k2*(c25);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c6)*(c2), Pb = (c25);
      const double g = Pr / (1.0 + Pr), dg = 1.0 / ((1.0 + Pr) * (1.0 + Pr));
      const double u = lPr + cc, den = nn - 0.14 * u, qq = 1.0 + f1 * f1;
      const double dlnF_dPr = -2.0 * lFc * f1 * (nn / (den * den)) / (qq * qq * Pr);
      const double df1_dL = (-0.67 * nn + 1.27 * u) / (den * den);
      const double dlnF_dFc = (1.0 / qq - lFc * 2.0 * f1 * df1_dL / (qq * qq)) / Fc;
      const double dk1_dPr = kinf * Fb * (dg + g * dlnF_dPr);
      const double dPr_dT = Mx * ((x1.y - x0.y) * kinf - k0 * (x1.x - x0.x)) / (kinf * kinf);
      const double dk1T = (x1.x - x0.x) * g * Fb + dk1_dPr * dPr_dT + k1 * dlnF_dFc * (y1.y - y0.y);
      const double dk2T = dk1T / kc - k1 * (y1.x - y0.x) / (kc * kc);
      const double dk1M = dk1_dPr * k0 / kinf;
      scr[23 * SST] = dk1M * Pf - (dk1M / kc) * Pb;
      scr[24 * SST] = k1; scr[25 * SST] = k2;
      const double dqT = dk1T * Pf - dk2T * Pb;
      dT[2] += -1.0 * dqT;
      dT[6] += -1.0 * dqT;
      dT[25] += 1.0 * dqT;
    }
    w2 = w2 + -1.0 * q;
    w6 = w6 + -1.0 * q;
    w25 = w25 + 1.0 * q;
  }
  {  // reaction 47: C2H3 + H <=> C2H2 + H2
    const double2 a0 = ld2(rk0 + 82), a1 = ld2(rk1 + 82);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c6)*(c2);
    const double pb = kb*(c4)*(c19);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c6)*(c2), Pb = (c4)*(c19);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += -1.0 * dqT;
      dT[4] += 1.0 * dqT;
      dT[6] += -1.0 * dqT;
      dT[19] += 1.0 * dqT;
    }
    w2 = w2 + -1.0 * q;
    w4 = w4 + 1.0 * q;
    w6 = w6 + -1.0 * q;
    w19 = w19 + 1.0 * q;
  }
  {  // reaction 48: C2H4 + H (+M) <=> C2H5 (+M)
    const double2 x0 = ld2(rk0 + 390), y0 = ld2(rk0 + 392);
    const double2 x1 = ld2(rk1 + 390), y1 = ld2(rk1 + 392);
    const double kinf = x0.x + (x1.x - x0.x) * Td, k0 = x0.y + (x1.y - x0.y) * Td;
    const double kc = y0.x + (y1.x - y0.x) * Td, Fc = y0.y + (y1.y - y0.y) * Td;
    const double Mx = M1;
    const double Pr = k0 * Mx / kinf;
    const double lFc = log10(Fc), lPr = log10(Pr);
    const double cc = -0.4 - 0.67 * lFc, nn = 0.75 - 1.27 * lFc;
    const double f1 = (lPr + cc) / (nn - 0.14 * (lPr + cc));
    const double Fb = exp10(lFc / (1.0 + f1 * f1));
    const double k1 = kinf * (Pr / (1.0 + Pr)) * Fb;
    const double k2 = k1 / kc;
    const double pf = k1*(c25)*(c2);
    const double pb = k2*(c28);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c25)*(c2), Pb = (c28);
      const double g = Pr / (1.0 + Pr), dg = 1.0 / ((1.0 + Pr) * (1.0 + Pr));
      const double u = lPr + cc, den = nn - 0.14 * u, qq = 1.0 + f1 * f1;
      const double dlnF_dPr = -2.0 * lFc * f1 * (nn / (den * den)) / (qq * qq * Pr);
      const double df1_dL = (-0.67 * nn + 1.27 * u) / (den * den);
      const double dlnF_dFc = (1.0 / qq - lFc * 2.0 * f1 * df1_dL / (qq * qq)) / Fc;
      const double dk1_dPr = kinf * Fb * (dg + g * dlnF_dPr);
      const double dPr_dT = Mx * ((x1.y - x0.y) * kinf - k0 * (x1.x - x0.x)) / (kinf * kinf);
      const double dk1T = (x1.x - x0.x) * g * Fb + dk1_dPr * dPr_dT + k1 * dlnF_dFc * (y1.y - y0.y);
      const double dk2T = dk1T / kc - k1 * (y1.x - y0.x) / (kc * kc);
      const double dk1M = dk1_dPr * k0 / kinf;
      scr[26 * SST] = dk1M * Pf - (dk1M / kc) * Pb;
      scr[27 * SST] = k1; scr[28 * SST] = k2;
      const double dqT = dk1T * Pf - dk2T * Pb;
      dT[2] += -1.0 * dqT;
      dT[25] += -1.0 * dqT;
      dT[28] += 1.0 * dqT;
    }
    w2 = w2 + -1.0 * q;
    w25 = w25 + -1.0 * q;
    w28 = w28 + 1.0 * q;
  }
  {  // reaction 49: C2H4 + H <=> C2H3 + H2
    const double2 a0 = ld2(rk0 + 84), a1 = ld2(rk1 + 84);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c25)*(c2);
    const double pb = kb*(c6)*(c19);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c25)*(c2), Pb = (c6)*(c19);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += -1.0 * dqT;
      dT[6] += 1.0 * dqT;
      dT[19] += 1.0 * dqT;
      dT[25] += -1.0 * dqT;
    }
    w2 = w2 + -1.0 * q;
    w6 = w6 + 1.0 * q;
    w19 = w19 + 1.0 * q;
    w25 = w25 + -1.0 * q;
  }
  {  // reaction 50: C2H5 + H (+M) <=> C2H6 (+M)
    const double2 x0 = ld2(rk0 + 394), y0 = ld2(rk0 + 396);
    const double2 x1 = ld2(rk1 + 394), y1 = ld2(rk1 + 396);
    const double kinf = x0.x + (x1.x - x0.x) * Td, k0 = x0.y + (x1.y - x0.y) * Td;
    const double kc = y0.x + (y1.x - y0.x) * Td, Fc = y0.y + (y1.y - y0.y) * Td;
    const double Mx = M1;
    const double Pr = k0 * Mx / kinf;
    const double lFc = log10(Fc), lPr = log10(Pr);
    const double cc = -0.4 - 0.67 * lFc, nn = 0.75 - 1.27 * lFc;
    const double f1 = (lPr + cc) / (nn - 0.14 * (lPr + cc));
    const double Fb = exp10(lFc / (1.0 + f1 * f1));
    const double k1 = kinf * (Pr / (1.0 + Pr)) * Fb;
    const double k2 = k1 / kc;
    const double pf = k1*(c28)*(c2);
    const double pb = k2*(c11);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c28)*(c2), Pb = (c11);
      const double g = Pr / (1.0 + Pr), dg = 1.0 / ((1.0 + Pr) * (1.0 + Pr));
      const double u = lPr + cc, den = nn - 0.14 * u, qq = 1.0 + f1 * f1;
      const double dlnF_dPr = -2.0 * lFc * f1 * (nn / (den * den)) / (qq * qq * Pr);
      const double df1_dL = (-0.67 * nn + 1.27 * u) / (den * den);
      const double dlnF_dFc = (1.0 / qq - lFc * 2.0 * f1 * df1_dL / (qq * qq)) / Fc;
      const double dk1_dPr = kinf * Fb * (dg + g * dlnF_dPr);
      const double dPr_dT = Mx * ((x1.y - x0.y) * kinf - k0 * (x1.x - x0.x)) / (kinf * kinf);
      const double dk1T = (x1.x - x0.x) * g * Fb + dk1_dPr * dPr_dT + k1 * dlnF_dFc * (y1.y - y0.y);
      const double dk2T = dk1T / kc - k1 * (y1.x - y0.x) / (kc * kc);
      const double dk1M = dk1_dPr * k0 / kinf;
      scr[29 * SST] = dk1M * Pf - (dk1M / kc) * Pb;
      scr[30 * SST] = k1; scr[31 * SST] = k2;
      const double dqT = dk1T * Pf - dk2T * Pb;
      dT[2] += -1.0 * dqT;
      dT[11] += 1.0 * dqT;
      dT[28] += -1.0 * dqT;
    }
    w2 = w2 + -1.0 * q;
    w11 = w11 + 1.0 * q;
    w28 = w28 + -1.0 * q;
  }
  {  // reaction 51: C2H5 + H <=> C2H4 + H2
    const double2 a0 = ld2(rk0 + 86), a1 = ld2(rk1 + 86);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c28)*(c2);
    const double pb = kb*(c25)*(c19);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c28)*(c2), Pb = (c25)*(c19);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += -1.0 * dqT;
      dT[19] += 1.0 * dqT;
      dT[25] += 1.0 * dqT;
      dT[28] += -1.0 * dqT;
    }
    w2 = w2 + -1.0 * q;
    w19 = w19 + 1.0 * q;
    w25 = w25 + 1.0 * q;
    w28 = w28 + -1.0 * q;
  }
  {  // reaction 52: C2H6 + H <=> C2H5 + H2
    const double2 a0 = ld2(rk0 + 88), a1 = ld2(rk1 + 88);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c11)*(c2);
    const double pb = kb*(c28)*(c19);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c11)*(c2), Pb = (c28)*(c19);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += -1.0 * dqT;
      dT[11] += -1.0 * dqT;
      dT[19] += 1.0 * dqT;
      dT[28] += 1.0 * dqT;
    }
    w2 = w2 + -1.0 * q;
    w11 = w11 + -1.0 * q;
    w19 = w19 + 1.0 * q;
    w28 = w28 + 1.0 * q;
  }
  {  // reaction 53: CH2CO + H <=> CH3 + CO
    const double2 a0 = ld2(rk0 + 90), a1 = ld2(rk1 + 90);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c12)*(c2);
    const double pb = kb*(c18)*(c9);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c12)*(c2), Pb = (c18)*(c9);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += -1.0 * dqT;
      dT[9] += 1.0 * dqT;
      dT[12] += -1.0 * dqT;
      dT[18] += 1.0 * dqT;
    }
    w2 = w2 + -1.0 * q;
    w9 = w9 + 1.0 * q;
    w12 = w12 + -1.0 * q;
    w18 = w18 + 1.0 * q;
  }
  {  // reaction 54: CO + H2 (+M) <=> CH2O (+M)
    const double2 x0 = ld2(rk0 + 398), y0 = ld2(rk0 + 400);
    const double2 x1 = ld2(rk1 + 398), y1 = ld2(rk1 + 400);
    const double kinf = x0.x + (x1.x - x0.x) * Td, k0 = x0.y + (x1.y - x0.y) * Td;
    const double kc = y0.x + (y1.x - y0.x) * Td, Fc = y0.y + (y1.y - y0.y) * Td;
    const double Mx = M1;
    const double Pr = k0 * Mx / kinf;
    const double lFc = log10(Fc), lPr = log10(Pr);
    const double cc = -0.4 - 0.67 * lFc, nn = 0.75 - 1.27 * lFc;
    const double f1 = (lPr + cc) / (nn - 0.14 * (lPr + cc));
    const double Fb = exp10(lFc / (1.0 + f1 * f1));
    const double k1 = kinf * (Pr / (1.0 + Pr)) * Fb;
    const double k2 = k1 / kc;
    const double pf = k1*(c9)*(c19);
    const double pb = k2*(c10);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c9)*(c19), Pb = (c10);
      const double g = Pr / (1.0 + Pr), dg = 1.0 / ((1.0 + Pr) * (1.0 + Pr));
      const double u = lPr + cc, den = nn - 0.14 * u, qq = 1.0 + f1 * f1;
      const double dlnF_dPr = -2.0 * lFc * f1 * (nn / (den * den)) / (qq * qq * Pr);
      const double df1_dL = (-0.67 * nn + 1.27 * u) / (den * den);
      const double dlnF_dFc = (1.0 / qq - lFc * 2.0 * f1 * df1_dL / (qq * qq)) / Fc;
      const double dk1_dPr = kinf * Fb * (dg + g * dlnF_dPr);
      const double dPr_dT = Mx * ((x1.y - x0.y) * kinf - k0 * (x1.x - x0.x)) / (kinf * kinf);
      const double dk1T = (x1.x - x0.x) * g * Fb + dk1_dPr * dPr_dT + k1 * dlnF_dFc * (y1.y - y0.y);
      const double dk2T = dk1T / kc - k1 * (y1.x - y0.x) / (kc * kc);
      const double dk1M = dk1_dPr * k0 / kinf;
      scr[32 * SST] = dk1M * Pf - (dk1M / kc) * Pb;
      scr[33 * SST] = k1; scr[34 * SST] = k2;
      const double dqT = dk1T * Pf - dk2T * Pb;
      dT[9] += -1.0 * dqT;
      dT[10] += 1.0 * dqT;
      dT[19] += -1.0 * dqT;
    }
    w9 = w9 + -1.0 * q;
    w10 = w10 + 1.0 * q;
    w19 = w19 + -1.0 * q;
  }
  {  // reaction 55: H2 + OH <=> H + H2O
    const double2 a0 = ld2(rk0 + 92), a1 = ld2(rk1 + 92);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c19)*(c27);
    const double pb = kb*(c2)*(c14);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c19)*(c27), Pb = (c2)*(c14);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += 1.0 * dqT;
      dT[14] += 1.0 * dqT;
      dT[19] += -1.0 * dqT;
      dT[27] += -1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w14 = w14 + 1.0 * q;
    w19 = w19 + -1.0 * q;
    w27 = w27 + -1.0 * q;
  }
  {  // reaction 56: 2 OH (+M) <=> H2O2 (+M)
    const double2 x0 = ld2(rk0 + 402), y0 = ld2(rk0 + 404);
    const double2 x1 = ld2(rk1 + 402), y1 = ld2(rk1 + 404);
    const double kinf = x0.x + (x1.x - x0.x) * Td, k0 = x0.y + (x1.y - x0.y) * Td;
    const double kc = y0.x + (y1.x - y0.x) * Td, Fc = y0.y + (y1.y - y0.y) * Td;
    const double Mx = M1;
    const double Pr = k0 * Mx / kinf;
    const double lFc = log10(Fc), lPr = log10(Pr);
    const double cc = -0.4 - 0.67 * lFc, nn = 0.75 - 1.27 * lFc;
    const double f1 = (lPr + cc) / (nn - 0.14 * (lPr + cc));
    const double Fb = exp10(lFc / (1.0 + f1 * f1));
    const double k1 = kinf * (Pr / (1.0 + Pr)) * Fb;
    const double k2 = k1 / kc;
    const double pf = k1*(c27*c27);
    const double pb = k2*(c8);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c27*c27), Pb = (c8);
      const double g = Pr / (1.0 + Pr), dg = 1.0 / ((1.0 + Pr) * (1.0 + Pr));
      const double u = lPr + cc, den = nn - 0.14 * u, qq = 1.0 + f1 * f1;
      const double dlnF_dPr = -2.0 * lFc * f1 * (nn / (den * den)) / (qq * qq * Pr);
      const double df1_dL = (-0.67 * nn + 1.27 * u) / (den * den);
      const double dlnF_dFc = (1.0 / qq - lFc * 2.0 * f1 * df1_dL / (qq * qq)) / Fc;
      const double dk1_dPr = kinf * Fb * (dg + g * dlnF_dPr);
      const double dPr_dT = Mx * ((x1.y - x0.y) * kinf - k0 * (x1.x - x0.x)) / (kinf * kinf);
      const double dk1T = (x1.x - x0.x) * g * Fb + dk1_dPr * dPr_dT + k1 * dlnF_dFc * (y1.y - y0.y);
      const double dk2T = dk1T / kc - k1 * (y1.x - y0.x) / (kc * kc);
      const double dk1M = dk1_dPr * k0 / kinf;
      scr[35 * SST] = dk1M * Pf - (dk1M / kc) * Pb;
      scr[36 * SST] = k1; scr[37 * SST] = k2;
      const double dqT = dk1T * Pf - dk2T * Pb;
      dT[8] += 1.0 * dqT;
      dT[27] += -2.0 * dqT;
    }
    w8 = w8 + 1.0 * q;
    w27 = w27 + -2.0 * q;
  }
  {  // reaction 57: 2 OH <=> H2O + O
    const double2 a0 = ld2(rk0 + 94), a1 = ld2(rk1 + 94);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c27*c27);
    const double pb = kb*(c14)*(c7);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c27*c27), Pb = (c14)*(c7);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[7] += 1.0 * dqT;
      dT[14] += 1.0 * dqT;
      dT[27] += -2.0 * dqT;
    }
    w7 = w7 + 1.0 * q;
    w14 = w14 + 1.0 * q;
    w27 = w27 + -2.0 * q;
  }
  {  // reaction 58: HO2 + OH <=> H2O + O2
    const double2 a0 = ld2(rk0 + 96), a1 = ld2(rk1 + 96);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c24)*(c27);
    const double pb = kb*(c14)*(c15);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c24)*(c27), Pb = (c14)*(c15);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[14] += 1.0 * dqT;
      dT[15] += 1.0 * dqT;
      dT[24] += -1.0 * dqT;
      dT[27] += -1.0 * dqT;
    }
    w14 = w14 + 1.0 * q;
    w15 = w15 + 1.0 * q;
    w24 = w24 + -1.0 * q;
    w27 = w27 + -1.0 * q;
  }
  {  // reaction 59: H2O2 + OH <=> H2O + HO2
    const double2 a0 = ld2(rk0 + 98), a1 = ld2(rk1 + 98);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c8)*(c27);
    const double pb = kb*(c14)*(c24);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c8)*(c27), Pb = (c14)*(c24);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[8] += -1.0 * dqT;
      dT[14] += 1.0 * dqT;
      dT[24] += 1.0 * dqT;
      dT[27] += -1.0 * dqT;
    }
    w8 = w8 + -1.0 * q;
    w14 = w14 + 1.0 * q;
    w24 = w24 + 1.0 * q;
    w27 = w27 + -1.0 * q;
  }
  {  // reaction 60: H2O2 + OH <=> H2O + HO2
    const double2 a0 = ld2(rk0 + 100), a1 = ld2(rk1 + 100);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c8)*(c27);
    const double pb = kb*(c14)*(c24);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c8)*(c27), Pb = (c14)*(c24);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[8] += -1.0 * dqT;
      dT[14] += 1.0 * dqT;
      dT[24] += 1.0 * dqT;
      dT[27] += -1.0 * dqT;
    }
    w8 = w8 + -1.0 * q;
    w14 = w14 + 1.0 * q;
    w24 = w24 + 1.0 * q;
    w27 = w27 + -1.0 * q;
  }
  {  // reaction 61: CH + OH <=> H + HCO
    const double2 a0 = ld2(rk0 + 102), a1 = ld2(rk1 + 102);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c26)*(c27);
    const double pb = kb*(c2)*(c17);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c26)*(c27), Pb = (c2)*(c17);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += 1.0 * dqT;
      dT[17] += 1.0 * dqT;
      dT[26] += -1.0 * dqT;
      dT[27] += -1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w17 = w17 + 1.0 * q;
    w26 = w26 + -1.0 * q;
    w27 = w27 + -1.0 * q;
  }
  {  // reaction 62: CH2 + OH <=> CH2O + H
    const double2 a0 = ld2(rk0 + 104), a1 = ld2(rk1 + 104);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c29)*(c27);
    const double pb = kb*(c10)*(c2);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c29)*(c27), Pb = (c10)*(c2);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += 1.0 * dqT;
      dT[10] += 1.0 * dqT;
      dT[27] += -1.0 * dqT;
      dT[29] += -1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w10 = w10 + 1.0 * q;
    w27 = w27 + -1.0 * q;
    w29 = w29 + -1.0 * q;
  }
  {  // reaction 63: CH2 + OH <=> CH + H2O
    const double2 a0 = ld2(rk0 + 106), a1 = ld2(rk1 + 106);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c29)*(c27);
    const double pb = kb*(c26)*(c14);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c29)*(c27), Pb = (c26)*(c14);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[14] += 1.0 * dqT;
      dT[26] += 1.0 * dqT;
      dT[27] += -1.0 * dqT;
      dT[29] += -1.0 * dqT;
    }
    w14 = w14 + 1.0 * q;
    w26 = w26 + 1.0 * q;
    w27 = w27 + -1.0 * q;
    w29 = w29 + -1.0 * q;
  }
  {  // reaction 64: CH2(S) + OH <=> CH2O + H
    const double2 a0 = ld2(rk0 + 108), a1 = ld2(rk1 + 108);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c16)*(c27);
    const double pb = kb*(c10)*(c2);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c16)*(c27), Pb = (c10)*(c2);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += 1.0 * dqT;
      dT[10] += 1.0 * dqT;
      dT[16] += -1.0 * dqT;
      dT[27] += -1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w10 = w10 + 1.0 * q;
    w16 = w16 + -1.0 * q;
    w27 = w27 + -1.0 * q;
  }
  {  // reaction 65: CH3 + OH <=> CH2 + H2O
    const double2 a0 = ld2(rk0 + 110), a1 = ld2(rk1 + 110);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c18)*(c27);
    const double pb = kb*(c29)*(c14);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c18)*(c27), Pb = (c29)*(c14);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[14] += 1.0 * dqT;
      dT[18] += -1.0 * dqT;
      dT[27] += -1.0 * dqT;
      dT[29] += 1.0 * dqT;
    }
    w14 = w14 + 1.0 * q;
    w18 = w18 + -1.0 * q;
    w27 = w27 + -1.0 * q;
    w29 = w29 + 1.0 * q;
  }
  {  // reaction 66: CH3 + OH <=> CH2(S) + H2O
    const double2 a0 = ld2(rk0 + 112), a1 = ld2(rk1 + 112);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c18)*(c27);
    const double pb = kb*(c16)*(c14);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c18)*(c27), Pb = (c16)*(c14);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[14] += 1.0 * dqT;
      dT[16] += 1.0 * dqT;
      dT[18] += -1.0 * dqT;
      dT[27] += -1.0 * dqT;
    }
    w14 = w14 + 1.0 * q;
    w16 = w16 + 1.0 * q;
    w18 = w18 + -1.0 * q;
    w27 = w27 + -1.0 * q;
  }
  {  // reaction 67: CH4 + OH <=> CH3 + H2O
    const double2 a0 = ld2(rk0 + 114), a1 = ld2(rk1 + 114);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c3)*(c27);
    const double pb = kb*(c18)*(c14);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c3)*(c27), Pb = (c18)*(c14);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[3] += -1.0 * dqT;
      dT[14] += 1.0 * dqT;
      dT[18] += 1.0 * dqT;
      dT[27] += -1.0 * dqT;
    }
    w3 = w3 + -1.0 * q;
    w14 = w14 + 1.0 * q;
    w18 = w18 + 1.0 * q;
    w27 = w27 + -1.0 * q;
  }
  {  // reaction 68: CO + OH <=> CO2 + H
    const double2 a0 = ld2(rk0 + 116), a1 = ld2(rk1 + 116);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c9)*(c27);
    const double pb = kb*(c30)*(c2);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c9)*(c27), Pb = (c30)*(c2);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += 1.0 * dqT;
      dT[9] += -1.0 * dqT;
      dT[27] += -1.0 * dqT;
      dT[30] += 1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w9 = w9 + -1.0 * q;
    w27 = w27 + -1.0 * q;
    w30 = w30 + 1.0 * q;
  }
  {  // reaction 69: HCO + OH <=> CO + H2O
    const double2 a0 = ld2(rk0 + 118), a1 = ld2(rk1 + 118);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c17)*(c27);
    const double pb = kb*(c9)*(c14);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c17)*(c27), Pb = (c9)*(c14);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[9] += 1.0 * dqT;
      dT[14] += 1.0 * dqT;
      dT[17] += -1.0 * dqT;
      dT[27] += -1.0 * dqT;
    }
    w9 = w9 + 1.0 * q;
    w14 = w14 + 1.0 * q;
    w17 = w17 + -1.0 * q;
    w27 = w27 + -1.0 * q;
  }
  {  // reaction 70: CH2O + OH <=> H2O + HCO
    const double2 a0 = ld2(rk0 + 120), a1 = ld2(rk1 + 120);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c10)*(c27);
    const double pb = kb*(c14)*(c17);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c10)*(c27), Pb = (c14)*(c17);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[10] += -1.0 * dqT;
      dT[14] += 1.0 * dqT;
      dT[17] += 1.0 * dqT;
      dT[27] += -1.0 * dqT;
    }
    w10 = w10 + -1.0 * q;
    w14 = w14 + 1.0 * q;
    w17 = w17 + 1.0 * q;
    w27 = w27 + -1.0 * q;
  }
  {  // reaction 71: C2H2 + OH <=> CH2CO + H
    const double2 a0 = ld2(rk0 + 122), a1 = ld2(rk1 + 122);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c4)*(c27);
    const double pb = kb*(c12)*(c2);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c4)*(c27), Pb = (c12)*(c2);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += 1.0 * dqT;
      dT[4] += -1.0 * dqT;
      dT[12] += 1.0 * dqT;
      dT[27] += -1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w4 = w4 + -1.0 * q;
    w12 = w12 + 1.0 * q;
    w27 = w27 + -1.0 * q;
  }
  {  // reaction 72: C2H2 + OH <=> CH3 + CO
    const double2 a0 = ld2(rk0 + 124), a1 = ld2(rk1 + 124);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c4)*(c27);
    const double pb = kb*(c18)*(c9);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c4)*(c27), Pb = (c18)*(c9);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[4] += -1.0 * dqT;
      dT[9] += 1.0 * dqT;
      dT[18] += 1.0 * dqT;
      dT[27] += -1.0 * dqT;
    }
    w4 = w4 + -1.0 * q;
    w9 = w9 + 1.0 * q;
    w18 = w18 + 1.0 * q;
    w27 = w27 + -1.0 * q;
  }
  {  // reaction 73: C2H3 + OH <=> C2H2 + H2O
    const double2 a0 = ld2(rk0 + 126), a1 = ld2(rk1 + 126);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c6)*(c27);
    const double pb = kb*(c4)*(c14);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c6)*(c27), Pb = (c4)*(c14);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[4] += 1.0 * dqT;
      dT[6] += -1.0 * dqT;
      dT[14] += 1.0 * dqT;
      dT[27] += -1.0 * dqT;
    }
    w4 = w4 + 1.0 * q;
    w6 = w6 + -1.0 * q;
    w14 = w14 + 1.0 * q;
    w27 = w27 + -1.0 * q;
  }
  {  // reaction 74: C2H4 + OH <=> C2H3 + H2O
    const double2 a0 = ld2(rk0 + 128), a1 = ld2(rk1 + 128);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c25)*(c27);
    const double pb = kb*(c6)*(c14);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c25)*(c27), Pb = (c6)*(c14);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[6] += 1.0 * dqT;
      dT[14] += 1.0 * dqT;
      dT[25] += -1.0 * dqT;
      dT[27] += -1.0 * dqT;
    }
    w6 = w6 + 1.0 * q;
    w14 = w14 + 1.0 * q;
    w25 = w25 + -1.0 * q;
    w27 = w27 + -1.0 * q;
  }
  {  // reaction 75: C2H6 + OH <=> C2H5 + H2O
    const double2 a0 = ld2(rk0 + 130), a1 = ld2(rk1 + 130);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c11)*(c27);
    const double pb = kb*(c28)*(c14);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c11)*(c27), Pb = (c28)*(c14);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[11] += -1.0 * dqT;
      dT[14] += 1.0 * dqT;
      dT[27] += -1.0 * dqT;
      dT[28] += 1.0 * dqT;
    }
    w11 = w11 + -1.0 * q;
    w14 = w14 + 1.0 * q;
    w27 = w27 + -1.0 * q;
    w28 = w28 + 1.0 * q;
  }
  {  // reaction 76: 2 HO2 <=> H2O2 + O2
    const double2 a0 = ld2(rk0 + 132), a1 = ld2(rk1 + 132);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c24*c24);
    const double pb = kb*(c8)*(c15);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c24*c24), Pb = (c8)*(c15);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[8] += 1.0 * dqT;
      dT[15] += 1.0 * dqT;
      dT[24] += -2.0 * dqT;
    }
    w8 = w8 + 1.0 * q;
    w15 = w15 + 1.0 * q;
    w24 = w24 + -2.0 * q;
  }
  {  // reaction 77: 2 HO2 <=> H2O2 + O2
    const double2 a0 = ld2(rk0 + 134), a1 = ld2(rk1 + 134);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c24*c24);
    const double pb = kb*(c8)*(c15);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c24*c24), Pb = (c8)*(c15);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[8] += 1.0 * dqT;
      dT[15] += 1.0 * dqT;
      dT[24] += -2.0 * dqT;
    }
    w8 = w8 + 1.0 * q;
    w15 = w15 + 1.0 * q;
    w24 = w24 + -2.0 * q;
  }
  {  // reaction 78: CH2 + HO2 <=> CH2O + OH
    const double2 a0 = ld2(rk0 + 136), a1 = ld2(rk1 + 136);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c29)*(c24);
    const double pb = kb*(c10)*(c27);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c29)*(c24), Pb = (c10)*(c27);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[10] += 1.0 * dqT;
      dT[24] += -1.0 * dqT;
      dT[27] += 1.0 * dqT;
      dT[29] += -1.0 * dqT;
    }
    w10 = w10 + 1.0 * q;
    w24 = w24 + -1.0 * q;
    w27 = w27 + 1.0 * q;
    w29 = w29 + -1.0 * q;
  }
  {  // reaction 79: CH3 + HO2 <=> CH4 + O2
    const double2 a0 = ld2(rk0 + 138), a1 = ld2(rk1 + 138);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c18)*(c24);
    const double pb = kb*(c3)*(c15);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c18)*(c24), Pb = (c3)*(c15);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[3] += 1.0 * dqT;
      dT[15] += 1.0 * dqT;
      dT[18] += -1.0 * dqT;
      dT[24] += -1.0 * dqT;
    }
    w3 = w3 + 1.0 * q;
    w15 = w15 + 1.0 * q;
    w18 = w18 + -1.0 * q;
    w24 = w24 + -1.0 * q;
  }
  {  // reaction 80: CO + HO2 <=> CO2 + OH
    const double2 a0 = ld2(rk0 + 140), a1 = ld2(rk1 + 140);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c9)*(c24);
    const double pb = kb*(c30)*(c27);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c9)*(c24), Pb = (c30)*(c27);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[9] += -1.0 * dqT;
      dT[24] += -1.0 * dqT;
      dT[27] += 1.0 * dqT;
      dT[30] += 1.0 * dqT;
    }
    w9 = w9 + -1.0 * q;
    w24 = w24 + -1.0 * q;
    w27 = w27 + 1.0 * q;
    w30 = w30 + 1.0 * q;
  }
  {  // reaction 81: CH2O + HO2 <=> H2O2 + HCO
    const double2 a0 = ld2(rk0 + 142), a1 = ld2(rk1 + 142);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c10)*(c24);
    const double pb = kb*(c8)*(c17);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c10)*(c24), Pb = (c8)*(c17);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[8] += 1.0 * dqT;
      dT[10] += -1.0 * dqT;
      dT[17] += 1.0 * dqT;
      dT[24] += -1.0 * dqT;
    }
    w8 = w8 + 1.0 * q;
    w10 = w10 + -1.0 * q;
    w17 = w17 + 1.0 * q;
    w24 = w24 + -1.0 * q;
  }
  {  // reaction 82: CH + O2 <=> HCO + O
    const double2 a0 = ld2(rk0 + 144), a1 = ld2(rk1 + 144);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c26)*(c15);
    const double pb = kb*(c17)*(c7);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c26)*(c15), Pb = (c17)*(c7);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[7] += 1.0 * dqT;
      dT[15] += -1.0 * dqT;
      dT[17] += 1.0 * dqT;
      dT[26] += -1.0 * dqT;
    }
    w7 = w7 + 1.0 * q;
    w15 = w15 + -1.0 * q;
    w17 = w17 + 1.0 * q;
    w26 = w26 + -1.0 * q;
  }
  {  // reaction 83: CH + H2 <=> CH2 + H
    const double2 a0 = ld2(rk0 + 146), a1 = ld2(rk1 + 146);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c26)*(c19);
    const double pb = kb*(c29)*(c2);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c26)*(c19), Pb = (c29)*(c2);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += 1.0 * dqT;
      dT[19] += -1.0 * dqT;
      dT[26] += -1.0 * dqT;
      dT[29] += 1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w19 = w19 + -1.0 * q;
    w26 = w26 + -1.0 * q;
    w29 = w29 + 1.0 * q;
  }
  {  // reaction 84: CH + H2O <=> CH2O + H
    const double2 a0 = ld2(rk0 + 148), a1 = ld2(rk1 + 148);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c26)*(c14);
    const double pb = kb*(c10)*(c2);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c26)*(c14), Pb = (c10)*(c2);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += 1.0 * dqT;
      dT[10] += 1.0 * dqT;
      dT[14] += -1.0 * dqT;
      dT[26] += -1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w10 = w10 + 1.0 * q;
    w14 = w14 + -1.0 * q;
    w26 = w26 + -1.0 * q;
  }
  {  // reaction 85: CH + CH2 <=> C2H2 + H
    const double2 a0 = ld2(rk0 + 150), a1 = ld2(rk1 + 150);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c26)*(c29);
    const double pb = kb*(c4)*(c2);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c26)*(c29), Pb = (c4)*(c2);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += 1.0 * dqT;
      dT[4] += 1.0 * dqT;
      dT[26] += -1.0 * dqT;
      dT[29] += -1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w4 = w4 + 1.0 * q;
    w26 = w26 + -1.0 * q;
    w29 = w29 + -1.0 * q;
  }
  {  // reaction 86: CH + CH3 <=> C2H3 + H
    const double2 a0 = ld2(rk0 + 152), a1 = ld2(rk1 + 152);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c26)*(c18);
    const double pb = kb*(c6)*(c2);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c26)*(c18), Pb = (c6)*(c2);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += 1.0 * dqT;
      dT[6] += 1.0 * dqT;
      dT[18] += -1.0 * dqT;
      dT[26] += -1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w6 = w6 + 1.0 * q;
    w18 = w18 + -1.0 * q;
    w26 = w26 + -1.0 * q;
  }
  {  // reaction 87: CH + CH4 <=> C2H4 + H
    const double2 a0 = ld2(rk0 + 154), a1 = ld2(rk1 + 154);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c26)*(c3);
    const double pb = kb*(c25)*(c2);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c26)*(c3), Pb = (c25)*(c2);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += 1.0 * dqT;
      dT[3] += -1.0 * dqT;
      dT[25] += 1.0 * dqT;
      dT[26] += -1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w3 = w3 + -1.0 * q;
    w25 = w25 + 1.0 * q;
    w26 = w26 + -1.0 * q;
  }
  {  // reaction 88: CH + CO2 <=> CO + HCO
    const double2 a0 = ld2(rk0 + 156), a1 = ld2(rk1 + 156);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c26)*(c30);
    const double pb = kb*(c9)*(c17);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c26)*(c30), Pb = (c9)*(c17);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[9] += 1.0 * dqT;
      dT[17] += 1.0 * dqT;
      dT[26] += -1.0 * dqT;
      dT[30] += -1.0 * dqT;
    }
    w9 = w9 + 1.0 * q;
    w17 = w17 + 1.0 * q;
    w26 = w26 + -1.0 * q;
    w30 = w30 + -1.0 * q;
  }
  {  // reaction 89: CH + CH2O <=> CH2CO + H
    const double2 a0 = ld2(rk0 + 158), a1 = ld2(rk1 + 158);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c26)*(c10);
    const double pb = kb*(c12)*(c2);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c26)*(c10), Pb = (c12)*(c2);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += 1.0 * dqT;
      dT[10] += -1.0 * dqT;
      dT[12] += 1.0 * dqT;
      dT[26] += -1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w10 = w10 + -1.0 * q;
    w12 = w12 + 1.0 * q;
    w26 = w26 + -1.0 * q;
  }
  {  // reaction 90: CH2 + O2 <=> HCO + OH
    const double2 a0 = ld2(rk0 + 160), a1 = ld2(rk1 + 160);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c29)*(c15);
    const double pb = kb*(c17)*(c27);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c29)*(c15), Pb = (c17)*(c27);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[15] += -1.0 * dqT;
      dT[17] += 1.0 * dqT;
      dT[27] += 1.0 * dqT;
      dT[29] += -1.0 * dqT;
    }
    w15 = w15 + -1.0 * q;
    w17 = w17 + 1.0 * q;
    w27 = w27 + 1.0 * q;
    w29 = w29 + -1.0 * q;
  }
  {  // reaction 91: CH2 + H2 <=> CH3 + H
    const double2 a0 = ld2(rk0 + 162), a1 = ld2(rk1 + 162);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c29)*(c19);
    const double pb = kb*(c18)*(c2);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c29)*(c19), Pb = (c18)*(c2);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += 1.0 * dqT;
      dT[18] += 1.0 * dqT;
      dT[19] += -1.0 * dqT;
      dT[29] += -1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w18 = w18 + 1.0 * q;
    w19 = w19 + -1.0 * q;
    w29 = w29 + -1.0 * q;
  }
  {  // reaction 92: 2 CH2 <=> C2H2 + H2
    const double2 a0 = ld2(rk0 + 164), a1 = ld2(rk1 + 164);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c29*c29);
    const double pb = kb*(c4)*(c19);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c29*c29), Pb = (c4)*(c19);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[4] += 1.0 * dqT;
      dT[19] += 1.0 * dqT;
      dT[29] += -2.0 * dqT;
    }
    w4 = w4 + 1.0 * q;
    w19 = w19 + 1.0 * q;
    w29 = w29 + -2.0 * q;
  }
  {  // reaction 93: CH2 + CH3 <=> C2H4 + H
    const double2 a0 = ld2(rk0 + 166), a1 = ld2(rk1 + 166);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c29)*(c18);
    const double pb = kb*(c25)*(c2);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c29)*(c18), Pb = (c25)*(c2);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += 1.0 * dqT;
      dT[18] += -1.0 * dqT;
      dT[25] += 1.0 * dqT;
      dT[29] += -1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w18 = w18 + -1.0 * q;
    w25 = w25 + 1.0 * q;
    w29 = w29 + -1.0 * q;
  }
  {  // reaction 94: CH2 + CH4 <=> 2 CH3
    const double2 a0 = ld2(rk0 + 168), a1 = ld2(rk1 + 168);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c29)*(c3);
    const double pb = kb*(c18*c18);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c29)*(c3), Pb = (c18*c18);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[3] += -1.0 * dqT;
      dT[18] += 2.0 * dqT;
      dT[29] += -1.0 * dqT;
    }
    w3 = w3 + -1.0 * q;
    w18 = w18 + 2.0 * q;
    w29 = w29 + -1.0 * q;
  }
  {  // reaction 95: CH2 + CO (+M) <=> CH2CO (+M)
    const double2 x0 = ld2(rk0 + 406), y0 = ld2(rk0 + 408);
    const double2 x1 = ld2(rk1 + 406), y1 = ld2(rk1 + 408);
    const double kinf = x0.x + (x1.x - x0.x) * Td, k0 = x0.y + (x1.y - x0.y) * Td;
    const double kc = y0.x + (y1.x - y0.x) * Td, Fc = y0.y + (y1.y - y0.y) * Td;
    const double Mx = M1;
    const double Pr = k0 * Mx / kinf;
    const double lFc = log10(Fc), lPr = log10(Pr);
    const double cc = -0.4 - 0.67 * lFc, nn = 0.75 - 1.27 * lFc;
    const double f1 = (lPr + cc) / (nn - 0.14 * (lPr + cc));
    const double Fb = exp10(lFc / (1.0 + f1 * f1));
    const double k1 = kinf * (Pr / (1.0 + Pr)) * Fb;
    const double k2 = k1 / kc;
    const double pf = k1*(c29)*(c9);
    const double pb = k2*(c12);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c29)*(c9), Pb = (c12);
      const double g = Pr / (1.0 + Pr), dg = 1.0 / ((1.0 + Pr) * (1.0 + Pr));
      const double u = lPr + cc, den = nn - 0.14 * u, qq = 1.0 + f1 * f1;
      const double dlnF_dPr = -2.0 * lFc * f1 * (nn / (den * den)) / (qq * qq * Pr);
      const double df1_dL = (-0.67 * nn + 1.27 * u) / (den * den);
      const double dlnF_dFc = (1.0 / qq - lFc * 2.0 * f1 * df1_dL / (qq * qq)) / Fc;
      const double dk1_dPr = kinf * Fb * (dg + g * dlnF_dPr);
      const double dPr_dT = Mx * ((x1.y - x0.y) * kinf - k0 * (x1.x - x0.x)) / (kinf * kinf);
      const double dk1T = (x1.x - x0.x) * g * Fb + dk1_dPr * dPr_dT + k1 * dlnF_dFc * (y1.y - y0.y);
      const double dk2T = dk1T / kc - k1 * (y1.x - y0.x) / (kc * kc);
      const double dk1M = dk1_dPr * k0 / kinf;
      scr[38 * SST] = dk1M * Pf - (dk1M / kc) * Pb;
      scr[39 * SST] = k1; scr[40 * SST] = k2;
      const double dqT = dk1T * Pf - dk2T * Pb;
      dT[9] += -1.0 * dqT;
      dT[12] += 1.0 * dqT;
      dT[29] += -1.0 * dqT;
    }
    w9 = w9 + -1.0 * q;
    w12 = w12 + 1.0 * q;
    w29 = w29 + -1.0 * q;
  }
  {  // reaction 96: CH2(S) + O2 <=> CO + H + OH
    const double2 a0 = ld2(rk0 + 170), a1 = ld2(rk1 + 170);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    if (JAC) { scr[51 * SST] = kf; scr[52 * SST] = kb; }
    const double pf = kf*(c16)*(c15);
    const double pb = kb*(c9)*(c2)*(c27);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c16)*(c15), Pb = (c9)*(c2)*(c27);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += 1.0 * dqT;
      dT[9] += 1.0 * dqT;
      dT[15] += -1.0 * dqT;
      dT[16] += -1.0 * dqT;
      dT[27] += 1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w9 = w9 + 1.0 * q;
    w15 = w15 + -1.0 * q;
    w16 = w16 + -1.0 * q;
    w27 = w27 + 1.0 * q;
  }
  {  // reaction 97: CH2(S) + O2 <=> CO + H2O
    const double2 a0 = ld2(rk0 + 172), a1 = ld2(rk1 + 172);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c16)*(c15);
    const double pb = kb*(c9)*(c14);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c16)*(c15), Pb = (c9)*(c14);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[9] += 1.0 * dqT;
      dT[14] += 1.0 * dqT;
      dT[15] += -1.0 * dqT;
      dT[16] += -1.0 * dqT;
    }
    w9 = w9 + 1.0 * q;
    w14 = w14 + 1.0 * q;
    w15 = w15 + -1.0 * q;
    w16 = w16 + -1.0 * q;
  }
  {  // reaction 98: CH2(S) + H2 <=> CH3 + H
    const double2 a0 = ld2(rk0 + 174), a1 = ld2(rk1 + 174);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c16)*(c19);
    const double pb = kb*(c18)*(c2);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c16)*(c19), Pb = (c18)*(c2);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += 1.0 * dqT;
      dT[16] += -1.0 * dqT;
      dT[18] += 1.0 * dqT;
      dT[19] += -1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w16 = w16 + -1.0 * q;
    w18 = w18 + 1.0 * q;
    w19 = w19 + -1.0 * q;
  }
  {  // reaction 99: CH2(S) + H2O <=> CH2 + H2O
    const double2 a0 = ld2(rk0 + 176), a1 = ld2(rk1 + 176);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c16)*(c14);
    const double pb = kb*(c29)*(c14);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c16)*(c14), Pb = (c29)*(c14);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[14] += 0.0 * dqT;
      dT[16] += -1.0 * dqT;
      dT[29] += 1.0 * dqT;
    }
    w14 = w14 + 0.0 * q;
    w16 = w16 + -1.0 * q;
    w29 = w29 + 1.0 * q;
  }
  {  // reaction 100: CH2(S) + CH3 <=> C2H4 + H
    const double2 a0 = ld2(rk0 + 178), a1 = ld2(rk1 + 178);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c16)*(c18);
    const double pb = kb*(c25)*(c2);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c16)*(c18), Pb = (c25)*(c2);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += 1.0 * dqT;
      dT[16] += -1.0 * dqT;
      dT[18] += -1.0 * dqT;
      dT[25] += 1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w16 = w16 + -1.0 * q;
    w18 = w18 + -1.0 * q;
    w25 = w25 + 1.0 * q;
  }
  {  // reaction 101: CH2(S) + CH4 <=> 2 CH3
    const double2 a0 = ld2(rk0 + 180), a1 = ld2(rk1 + 180);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c16)*(c3);
    const double pb = kb*(c18*c18);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c16)*(c3), Pb = (c18*c18);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[3] += -1.0 * dqT;
      dT[16] += -1.0 * dqT;
      dT[18] += 2.0 * dqT;
    }
    w3 = w3 + -1.0 * q;
    w16 = w16 + -1.0 * q;
    w18 = w18 + 2.0 * q;
  }
  {  // reaction 102: CH2(S) + CO <=> CH2 + CO
    const double2 a0 = ld2(rk0 + 182), a1 = ld2(rk1 + 182);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c16)*(c9);
    const double pb = kb*(c29)*(c9);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c16)*(c9), Pb = (c29)*(c9);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[9] += 0.0 * dqT;
      dT[16] += -1.0 * dqT;
      dT[29] += 1.0 * dqT;
    }
    w9 = w9 + 0.0 * q;
    w16 = w16 + -1.0 * q;
    w29 = w29 + 1.0 * q;
  }
  {  // reaction 103: CH2(S) + CO2 <=> CH2 + CO2
    const double2 a0 = ld2(rk0 + 184), a1 = ld2(rk1 + 184);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c16)*(c30);
    const double pb = kb*(c29)*(c30);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c16)*(c30), Pb = (c29)*(c30);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[16] += -1.0 * dqT;
      dT[29] += 1.0 * dqT;
      dT[30] += 0.0 * dqT;
    }
    w16 = w16 + -1.0 * q;
    w29 = w29 + 1.0 * q;
    w30 = w30 + 0.0 * q;
  }
  {  // reaction 104: CH2(S) + CO2 <=> CH2O + CO
    const double2 a0 = ld2(rk0 + 186), a1 = ld2(rk1 + 186);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c16)*(c30);
    const double pb = kb*(c10)*(c9);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c16)*(c30), Pb = (c10)*(c9);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[9] += 1.0 * dqT;
      dT[10] += 1.0 * dqT;
      dT[16] += -1.0 * dqT;
      dT[30] += -1.0 * dqT;
    }
    w9 = w9 + 1.0 * q;
    w10 = w10 + 1.0 * q;
    w16 = w16 + -1.0 * q;
    w30 = w30 + -1.0 * q;
  }
  {  // reaction 105: C2H6 + CH2(S) <=> C2H5 + CH3
    const double2 a0 = ld2(rk0 + 188), a1 = ld2(rk1 + 188);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c11)*(c16);
    const double pb = kb*(c28)*(c18);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c11)*(c16), Pb = (c28)*(c18);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[11] += -1.0 * dqT;
      dT[16] += -1.0 * dqT;
      dT[18] += 1.0 * dqT;
      dT[28] += 1.0 * dqT;
    }
    w11 = w11 + -1.0 * q;
    w16 = w16 + -1.0 * q;
    w18 = w18 + 1.0 * q;
    w28 = w28 + 1.0 * q;
  }
  {  // reaction 106: CH3 + O2 <=> CH2O + OH
    const double2 a0 = ld2(rk0 + 190), a1 = ld2(rk1 + 190);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c18)*(c15);
    const double pb = kb*(c10)*(c27);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c18)*(c15), Pb = (c10)*(c27);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[10] += 1.0 * dqT;
      dT[15] += -1.0 * dqT;
      dT[18] += -1.0 * dqT;
      dT[27] += 1.0 * dqT;
    }
    w10 = w10 + 1.0 * q;
    w15 = w15 + -1.0 * q;
    w18 = w18 + -1.0 * q;
    w27 = w27 + 1.0 * q;
  }
  {  // reaction 107: CH3 + H2O2 <=> CH4 + HO2
    const double2 a0 = ld2(rk0 + 192), a1 = ld2(rk1 + 192);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c18)*(c8);
    const double pb = kb*(c3)*(c24);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c18)*(c8), Pb = (c3)*(c24);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[3] += 1.0 * dqT;
      dT[8] += -1.0 * dqT;
      dT[18] += -1.0 * dqT;
      dT[24] += 1.0 * dqT;
    }
    w3 = w3 + 1.0 * q;
    w8 = w8 + -1.0 * q;
    w18 = w18 + -1.0 * q;
    w24 = w24 + 1.0 * q;
  }
  {  // reaction 108: 2 CH3 (+M) <=> C2H6 (+M)
    const double2 x0 = ld2(rk0 + 410), y0 = ld2(rk0 + 412);
    const double2 x1 = ld2(rk1 + 410), y1 = ld2(rk1 + 412);
    const double kinf = x0.x + (x1.x - x0.x) * Td, k0 = x0.y + (x1.y - x0.y) * Td;
    const double kc = y0.x + (y1.x - y0.x) * Td, Fc = y0.y + (y1.y - y0.y) * Td;
    const double Mx = M1;
    const double Pr = k0 * Mx / kinf;
    const double lFc = log10(Fc), lPr = log10(Pr);
    const double cc = -0.4 - 0.67 * lFc, nn = 0.75 - 1.27 * lFc;
    const double f1 = (lPr + cc) / (nn - 0.14 * (lPr + cc));
    const double Fb = exp10(lFc / (1.0 + f1 * f1));
    const double k1 = kinf * (Pr / (1.0 + Pr)) * Fb;
    const double k2 = k1 / kc;
    const double pf = k1*(c18*c18);
    const double pb = k2*(c11);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c18*c18), Pb = (c11);
      const double g = Pr / (1.0 + Pr), dg = 1.0 / ((1.0 + Pr) * (1.0 + Pr));
      const double u = lPr + cc, den = nn - 0.14 * u, qq = 1.0 + f1 * f1;
      const double dlnF_dPr = -2.0 * lFc * f1 * (nn / (den * den)) / (qq * qq * Pr);
      const double df1_dL = (-0.67 * nn + 1.27 * u) / (den * den);
      const double dlnF_dFc = (1.0 / qq - lFc * 2.0 * f1 * df1_dL / (qq * qq)) / Fc;
      const double dk1_dPr = kinf * Fb * (dg + g * dlnF_dPr);
      const double dPr_dT = Mx * ((x1.y - x0.y) * kinf - k0 * (x1.x - x0.x)) / (kinf * kinf);
      const double dk1T = (x1.x - x0.x) * g * Fb + dk1_dPr * dPr_dT + k1 * dlnF_dFc * (y1.y - y0.y);
      const double dk2T = dk1T / kc - k1 * (y1.x - y0.x) / (kc * kc);
      const double dk1M = dk1_dPr * k0 / kinf;
      scr[41 * SST] = dk1M * Pf - (dk1M / kc) * Pb;
      scr[42 * SST] = k1; scr[43 * SST] = k2;
      const double dqT = dk1T * Pf - dk2T * Pb;
      dT[11] += 1.0 * dqT;
      dT[18] += -2.0 * dqT;
    }
    w11 = w11 + 1.0 * q;
    w18 = w18 + -2.0 * q;
  }
  {  // reaction 109: 2 CH3 <=> C2H5 + H
    const double2 a0 = ld2(rk0 + 194), a1 = ld2(rk1 + 194);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c18*c18);
    const double pb = kb*(c28)*(c2);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c18*c18), Pb = (c28)*(c2);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += 1.0 * dqT;
      dT[18] += -2.0 * dqT;
      dT[28] += 1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w18 = w18 + -2.0 * q;
    w28 = w28 + 1.0 * q;
  }
  {  // reaction 110: CH3 + HCO <=> CH4 + CO
    const double2 a0 = ld2(rk0 + 196), a1 = ld2(rk1 + 196);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c18)*(c17);
    const double pb = kb*(c3)*(c9);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c18)*(c17), Pb = (c3)*(c9);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[3] += 1.0 * dqT;
      dT[9] += 1.0 * dqT;
      dT[17] += -1.0 * dqT;
      dT[18] += -1.0 * dqT;
    }
    w3 = w3 + 1.0 * q;
    w9 = w9 + 1.0 * q;
    w17 = w17 + -1.0 * q;
    w18 = w18 + -1.0 * q;
  }
  {  // reaction 111: CH2O + CH3 <=> CH4 + HCO
    const double2 a0 = ld2(rk0 + 198), a1 = ld2(rk1 + 198);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c10)*(c18);
    const double pb = kb*(c3)*(c17);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c10)*(c18), Pb = (c3)*(c17);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[3] += 1.0 * dqT;
      dT[10] += -1.0 * dqT;
      dT[17] += 1.0 * dqT;
      dT[18] += -1.0 * dqT;
    }
    w3 = w3 + 1.0 * q;
    w10 = w10 + -1.0 * q;
    w17 = w17 + 1.0 * q;
    w18 = w18 + -1.0 * q;
  }
  {  // reaction 112: C2H4 + CH3 <=> C2H3 + CH4
    const double2 a0 = ld2(rk0 + 200), a1 = ld2(rk1 + 200);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c25)*(c18);
    const double pb = kb*(c6)*(c3);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c25)*(c18), Pb = (c6)*(c3);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[3] += 1.0 * dqT;
      dT[6] += 1.0 * dqT;
      dT[18] += -1.0 * dqT;
      dT[25] += -1.0 * dqT;
    }
    w3 = w3 + 1.0 * q;
    w6 = w6 + 1.0 * q;
    w18 = w18 + -1.0 * q;
    w25 = w25 + -1.0 * q;
  }
  {  // reaction 113: C2H6 + CH3 <=> C2H5 + CH4
    const double2 a0 = ld2(rk0 + 202), a1 = ld2(rk1 + 202);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c11)*(c18);
    const double pb = kb*(c28)*(c3);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c11)*(c18), Pb = (c28)*(c3);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[3] += 1.0 * dqT;
      dT[11] += -1.0 * dqT;
      dT[18] += -1.0 * dqT;
      dT[28] += 1.0 * dqT;
    }
    w3 = w3 + 1.0 * q;
    w11 = w11 + -1.0 * q;
    w18 = w18 + -1.0 * q;
    w28 = w28 + 1.0 * q;
  }
  {  // reaction 114: HCO + H2O <=> CO + H + H2O
    const double2 a0 = ld2(rk0 + 204), a1 = ld2(rk1 + 204);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c17)*M5;
    const double pb = kb*(c9)*(c2)*M5;
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c17), Pb = (c9)*(c2);
      const double dqT = ((a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb) * M5;
      scr[44 * SST] = kf * Pf - kb * Pb;
      dT[2] += 1.0 * dqT;
      dT[9] += 1.0 * dqT;
      dT[17] += -1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w9 = w9 + 1.0 * q;
    w17 = w17 + -1.0 * q;
  }
  {  // reaction 115: HCO + M <=> CO + H + M
    const double2 a0 = ld2(rk0 + 206), a1 = ld2(rk1 + 206);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c17)*M10;
    const double pb = kb*(c9)*(c2)*M10;
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c17), Pb = (c9)*(c2);
      const double dqT = ((a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb) * M10;
      scr[45 * SST] = kf * Pf - kb * Pb;
      dT[2] += 1.0 * dqT;
      dT[9] += 1.0 * dqT;
      dT[17] += -1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w9 = w9 + 1.0 * q;
    w17 = w17 + -1.0 * q;
  }
  {  // reaction 116: HCO + O2 <=> CO + HO2
    const double2 a0 = ld2(rk0 + 208), a1 = ld2(rk1 + 208);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c17)*(c15);
    const double pb = kb*(c9)*(c24);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c17)*(c15), Pb = (c9)*(c24);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[9] += 1.0 * dqT;
      dT[15] += -1.0 * dqT;
      dT[17] += -1.0 * dqT;
      dT[24] += 1.0 * dqT;
    }
    w9 = w9 + 1.0 * q;
    w15 = w15 + -1.0 * q;
    w17 = w17 + -1.0 * q;
    w24 = w24 + 1.0 * q;
  }
  {  // reaction 117: C2H3 + O2 <=> CH2O + HCO
    const double2 a0 = ld2(rk0 + 210), a1 = ld2(rk1 + 210);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c6)*(c15);
    const double pb = kb*(c10)*(c17);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c6)*(c15), Pb = (c10)*(c17);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[6] += -1.0 * dqT;
      dT[10] += 1.0 * dqT;
      dT[15] += -1.0 * dqT;
      dT[17] += 1.0 * dqT;
    }
    w6 = w6 + -1.0 * q;
    w10 = w10 + 1.0 * q;
    w15 = w15 + -1.0 * q;
    w17 = w17 + 1.0 * q;
  }
  {  // reaction 118: C2H4 (+M) <=> C2H2 + H2 (+M)
    const double2 x0 = ld2(rk0 + 414), y0 = ld2(rk0 + 416);
    const double2 x1 = ld2(rk1 + 414), y1 = ld2(rk1 + 416);
    const double kinf = x0.x + (x1.x - x0.x) * Td, k0 = x0.y + (x1.y - x0.y) * Td;
    const double kc = y0.x + (y1.x - y0.x) * Td, Fc = y0.y + (y1.y - y0.y) * Td;
    const double Mx = M1;
    const double Pr = k0 * Mx / kinf;
    const double lFc = log10(Fc), lPr = log10(Pr);
    const double cc = -0.4 - 0.67 * lFc, nn = 0.75 - 1.27 * lFc;
    const double f1 = (lPr + cc) / (nn - 0.14 * (lPr + cc));
    const double Fb = exp10(lFc / (1.0 + f1 * f1));
    const double k1 = kinf * (Pr / (1.0 + Pr)) * Fb;
    const double k2 = k1 / kc;
    const double pf = k1*(c25);
    const double pb = k2*(c4)*(c19);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c25), Pb = (c4)*(c19);
      const double g = Pr / (1.0 + Pr), dg = 1.0 / ((1.0 + Pr) * (1.0 + Pr));
      const double u = lPr + cc, den = nn - 0.14 * u, qq = 1.0 + f1 * f1;
      const double dlnF_dPr = -2.0 * lFc * f1 * (nn / (den * den)) / (qq * qq * Pr);
      const double df1_dL = (-0.67 * nn + 1.27 * u) / (den * den);
      const double dlnF_dFc = (1.0 / qq - lFc * 2.0 * f1 * df1_dL / (qq * qq)) / Fc;
      const double dk1_dPr = kinf * Fb * (dg + g * dlnF_dPr);
      const double dPr_dT = Mx * ((x1.y - x0.y) * kinf - k0 * (x1.x - x0.x)) / (kinf * kinf);
      const double dk1T = (x1.x - x0.x) * g * Fb + dk1_dPr * dPr_dT + k1 * dlnF_dFc * (y1.y - y0.y);
      const double dk2T = dk1T / kc - k1 * (y1.x - y0.x) / (kc * kc);
      const double dk1M = dk1_dPr * k0 / kinf;
      scr[46 * SST] = dk1M * Pf - (dk1M / kc) * Pb;
      scr[47 * SST] = k1; scr[48 * SST] = k2;
      const double dqT = dk1T * Pf - dk2T * Pb;
      dT[4] += 1.0 * dqT;
      dT[19] += 1.0 * dqT;
      dT[25] += -1.0 * dqT;
    }
    w4 = w4 + 1.0 * q;
    w19 = w19 + 1.0 * q;
    w25 = w25 + -1.0 * q;
  }
  {  // reaction 119: C2H5 + O2 <=> C2H4 + HO2
    const double2 a0 = ld2(rk0 + 212), a1 = ld2(rk1 + 212);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c28)*(c15);
    const double pb = kb*(c25)*(c24);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c28)*(c15), Pb = (c25)*(c24);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[15] += -1.0 * dqT;
      dT[24] += 1.0 * dqT;
      dT[25] += 1.0 * dqT;
      dT[28] += -1.0 * dqT;
    }
    w15 = w15 + -1.0 * q;
    w24 = w24 + 1.0 * q;
    w25 = w25 + 1.0 * q;
    w28 = w28 + -1.0 * q;
  }
  {  // reaction 120: CH3 + OH <=> CH2O + H2
    const double2 a0 = ld2(rk0 + 214), a1 = ld2(rk1 + 214);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c18)*(c27);
    const double pb = kb*(c10)*(c19);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c18)*(c27), Pb = (c10)*(c19);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[10] += 1.0 * dqT;
      dT[18] += -1.0 * dqT;
      dT[19] += 1.0 * dqT;
      dT[27] += -1.0 * dqT;
    }
    w10 = w10 + 1.0 * q;
    w18 = w18 + -1.0 * q;
    w19 = w19 + 1.0 * q;
    w27 = w27 + -1.0 * q;
  }
  {  // reaction 121: C2H3 + C2H5 => 2 C2H4
    const double2 a0 = ld2(rk0 + 216), a1 = ld2(rk1 + 216);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c6)*(c28);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c6)*(c28), Pb = (c25*c25);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[6] += -1.0 * dqT;
      dT[25] += 2.0 * dqT;
      dT[28] += -1.0 * dqT;
    }
    w6 = w6 + -1.0 * q;
    w25 = w25 + 2.0 * q;
    w28 = w28 + -1.0 * q;
  }
  {  // reaction 122: 2 C2H4 => C2H3 + C2H5
    const double2 a0 = ld2(rk0 + 218), a1 = ld2(rk1 + 218);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c25*c25);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c25*c25), Pb = (c6)*(c28);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[6] += 1.0 * dqT;
      dT[25] += -2.0 * dqT;
      dT[28] += 1.0 * dqT;
    }
    w6 = w6 + 1.0 * q;
    w25 = w25 + -2.0 * q;
    w28 = w28 + 1.0 * q;
  }
  {  // reaction 123: C2H4 + O => C2H3 + OH
    const double2 a0 = ld2(rk0 + 220), a1 = ld2(rk1 + 220);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c25)*(c7);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c25)*(c7), Pb = (c6)*(c27);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[6] += 1.0 * dqT;
      dT[7] += -1.0 * dqT;
      dT[25] += -1.0 * dqT;
      dT[27] += 1.0 * dqT;
    }
    w6 = w6 + 1.0 * q;
    w7 = w7 + -1.0 * q;
    w25 = w25 + -1.0 * q;
    w27 = w27 + 1.0 * q;
  }
  {  // reaction 124: C2H3 + OH => C2H4 + O
    const double2 a0 = ld2(rk0 + 222), a1 = ld2(rk1 + 222);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c6)*(c27);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c6)*(c27), Pb = (c25)*(c7);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[6] += -1.0 * dqT;
      dT[7] += 1.0 * dqT;
      dT[25] += 1.0 * dqT;
      dT[27] += -1.0 * dqT;
    }
    w6 = w6 + -1.0 * q;
    w7 = w7 + 1.0 * q;
    w25 = w25 + 1.0 * q;
    w27 = w27 + -1.0 * q;
  }
  {  // reaction 125: C2H2 + O2 => 2 HCO
    const double2 a0 = ld2(rk0 + 224), a1 = ld2(rk1 + 224);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c4)*(c15);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c4)*(c15), Pb = (c17*c17);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[4] += -1.0 * dqT;
      dT[15] += -1.0 * dqT;
      dT[17] += 2.0 * dqT;
    }
    w4 = w4 + -1.0 * q;
    w15 = w15 + -1.0 * q;
    w17 = w17 + 2.0 * q;
  }
  {  // reaction 126: 2 HCO => C2H2 + O2
    const double2 a0 = ld2(rk0 + 226), a1 = ld2(rk1 + 226);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c17*c17);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c17*c17), Pb = (c4)*(c15);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[4] += 1.0 * dqT;
      dT[15] += 1.0 * dqT;
      dT[17] += -2.0 * dqT;
    }
    w4 = w4 + 1.0 * q;
    w15 = w15 + 1.0 * q;
    w17 = w17 + -2.0 * q;
  }
  {  // reaction 127: CH + O2 => CO + OH
    const double2 a0 = ld2(rk0 + 228), a1 = ld2(rk1 + 228);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c26)*(c15);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c26)*(c15), Pb = (c9)*(c27);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[9] += 1.0 * dqT;
      dT[15] += -1.0 * dqT;
      dT[26] += -1.0 * dqT;
      dT[27] += 1.0 * dqT;
    }
    w9 = w9 + 1.0 * q;
    w15 = w15 + -1.0 * q;
    w26 = w26 + -1.0 * q;
    w27 = w27 + 1.0 * q;
  }
  {  // reaction 128: CO + OH => CH + O2
    const double2 a0 = ld2(rk0 + 230), a1 = ld2(rk1 + 230);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c9)*(c27);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c9)*(c27), Pb = (c26)*(c15);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[9] += -1.0 * dqT;
      dT[15] += 1.0 * dqT;
      dT[26] += 1.0 * dqT;
      dT[27] += -1.0 * dqT;
    }
    w9 = w9 + -1.0 * q;
    w15 = w15 + 1.0 * q;
    w26 = w26 + 1.0 * q;
    w27 = w27 + -1.0 * q;
  }
  {  // reaction 129: C2H3 + O2 => C2H2 + HO2
    const double2 a0 = ld2(rk0 + 232), a1 = ld2(rk1 + 232);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c6)*(c15);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c6)*(c15), Pb = (c4)*(c24);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[4] += 1.0 * dqT;
      dT[6] += -1.0 * dqT;
      dT[15] += -1.0 * dqT;
      dT[24] += 1.0 * dqT;
    }
    w4 = w4 + 1.0 * q;
    w6 = w6 + -1.0 * q;
    w15 = w15 + -1.0 * q;
    w24 = w24 + 1.0 * q;
  }
  {  // reaction 130: C2H2 + HO2 => C2H3 + O2
    const double2 a0 = ld2(rk0 + 234), a1 = ld2(rk1 + 234);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c4)*(c24);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c4)*(c24), Pb = (c6)*(c15);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[4] += -1.0 * dqT;
      dT[6] += 1.0 * dqT;
      dT[15] += 1.0 * dqT;
      dT[24] += -1.0 * dqT;
    }
    w4 = w4 + -1.0 * q;
    w6 = w6 + 1.0 * q;
    w15 = w15 + 1.0 * q;
    w24 = w24 + -1.0 * q;
  }
  {  // reaction 131: CH2CO + O => 2 HCO
    const double2 a0 = ld2(rk0 + 236), a1 = ld2(rk1 + 236);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c12)*(c7);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c12)*(c7), Pb = (c17*c17);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[7] += -1.0 * dqT;
      dT[12] += -1.0 * dqT;
      dT[17] += 2.0 * dqT;
    }
    w7 = w7 + -1.0 * q;
    w12 = w12 + -1.0 * q;
    w17 = w17 + 2.0 * q;
  }
  {  // reaction 132: 2 HCO => CH2CO + O
    const double2 a0 = ld2(rk0 + 238), a1 = ld2(rk1 + 238);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c17*c17);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c17*c17), Pb = (c12)*(c7);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[7] += 1.0 * dqT;
      dT[12] += 1.0 * dqT;
      dT[17] += -2.0 * dqT;
    }
    w7 = w7 + 1.0 * q;
    w12 = w12 + 1.0 * q;
    w17 = w17 + -2.0 * q;
  }
  {  // reaction 133: CH2CO + OH => CH2O + HCO
    const double2 a0 = ld2(rk0 + 240), a1 = ld2(rk1 + 240);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c12)*(c27);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c12)*(c27), Pb = (c10)*(c17);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[10] += 1.0 * dqT;
      dT[12] += -1.0 * dqT;
      dT[17] += 1.0 * dqT;
      dT[27] += -1.0 * dqT;
    }
    w10 = w10 + 1.0 * q;
    w12 = w12 + -1.0 * q;
    w17 = w17 + 1.0 * q;
    w27 = w27 + -1.0 * q;
  }
  {  // reaction 134: CH2O + HCO => CH2CO + OH
    const double2 a0 = ld2(rk0 + 242), a1 = ld2(rk1 + 242);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c10)*(c17);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c10)*(c17), Pb = (c12)*(c27);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[10] += -1.0 * dqT;
      dT[12] += 1.0 * dqT;
      dT[17] += -1.0 * dqT;
      dT[27] += 1.0 * dqT;
    }
    w10 = w10 + -1.0 * q;
    w12 = w12 + 1.0 * q;
    w17 = w17 + -1.0 * q;
    w27 = w27 + 1.0 * q;
  }
  {  // reaction 135: C2H6 + O2 => C2H5 + HO2
    const double2 a0 = ld2(rk0 + 244), a1 = ld2(rk1 + 244);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c11)*(c15);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c11)*(c15), Pb = (c28)*(c24);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[11] += -1.0 * dqT;
      dT[15] += -1.0 * dqT;
      dT[24] += 1.0 * dqT;
      dT[28] += 1.0 * dqT;
    }
    w11 = w11 + -1.0 * q;
    w15 = w15 + -1.0 * q;
    w24 = w24 + 1.0 * q;
    w28 = w28 + 1.0 * q;
  }
  {  // reaction 136: C2H5 + HO2 => C2H6 + O2
    const double2 a0 = ld2(rk0 + 246), a1 = ld2(rk1 + 246);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c28)*(c24);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c28)*(c24), Pb = (c11)*(c15);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[11] += 1.0 * dqT;
      dT[15] += 1.0 * dqT;
      dT[24] += -1.0 * dqT;
      dT[28] += -1.0 * dqT;
    }
    w11 = w11 + 1.0 * q;
    w15 = w15 + 1.0 * q;
    w24 = w24 + -1.0 * q;
    w28 = w28 + -1.0 * q;
  }
  {  // reaction 137: C2H6 + HO2 => C2H5 + H2O2
    const double2 a0 = ld2(rk0 + 248), a1 = ld2(rk1 + 248);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c11)*(c24);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c11)*(c24), Pb = (c28)*(c8);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[8] += 1.0 * dqT;
      dT[11] += -1.0 * dqT;
      dT[24] += -1.0 * dqT;
      dT[28] += 1.0 * dqT;
    }
    w8 = w8 + 1.0 * q;
    w11 = w11 + -1.0 * q;
    w24 = w24 + -1.0 * q;
    w28 = w28 + 1.0 * q;
  }
  {  // reaction 138: C2H5 + H2O2 => C2H6 + HO2
    const double2 a0 = ld2(rk0 + 250), a1 = ld2(rk1 + 250);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c28)*(c8);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c28)*(c8), Pb = (c11)*(c24);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[8] += -1.0 * dqT;
      dT[11] += 1.0 * dqT;
      dT[24] += 1.0 * dqT;
      dT[28] += -1.0 * dqT;
    }
    w8 = w8 + -1.0 * q;
    w11 = w11 + 1.0 * q;
    w24 = w24 + 1.0 * q;
    w28 = w28 + -1.0 * q;
  }
  {  // reaction 139: C2H4 + C2H6 => 2 C2H5
    const double2 a0 = ld2(rk0 + 252), a1 = ld2(rk1 + 252);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c25)*(c11);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c25)*(c11), Pb = (c28*c28);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[11] += -1.0 * dqT;
      dT[25] += -1.0 * dqT;
      dT[28] += 2.0 * dqT;
    }
    w11 = w11 + -1.0 * q;
    w25 = w25 + -1.0 * q;
    w28 = w28 + 2.0 * q;
  }
  {  // reaction 140: 2 C2H5 => C2H4 + C2H6
    const double2 a0 = ld2(rk0 + 254), a1 = ld2(rk1 + 254);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c28*c28);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c28*c28), Pb = (c25)*(c11);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[11] += 1.0 * dqT;
      dT[25] += 1.0 * dqT;
      dT[28] += -2.0 * dqT;
    }
    w11 = w11 + 1.0 * q;
    w25 = w25 + 1.0 * q;
    w28 = w28 + -2.0 * q;
  }
  {  // reaction 141: C2H3 + CH3 => C2H2 + CH4
    const double2 a0 = ld2(rk0 + 256), a1 = ld2(rk1 + 256);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c6)*(c18);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c6)*(c18), Pb = (c4)*(c3);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[3] += 1.0 * dqT;
      dT[4] += 1.0 * dqT;
      dT[6] += -1.0 * dqT;
      dT[18] += -1.0 * dqT;
    }
    w3 = w3 + 1.0 * q;
    w4 = w4 + 1.0 * q;
    w6 = w6 + -1.0 * q;
    w18 = w18 + -1.0 * q;
  }
  {  // reaction 142: C2H2 + CH4 => C2H3 + CH3
    const double2 a0 = ld2(rk0 + 258), a1 = ld2(rk1 + 258);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c4)*(c3);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c4)*(c3), Pb = (c6)*(c18);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[3] += -1.0 * dqT;
      dT[4] += -1.0 * dqT;
      dT[6] += 1.0 * dqT;
      dT[18] += 1.0 * dqT;
    }
    w3 = w3 + -1.0 * q;
    w4 = w4 + -1.0 * q;
    w6 = w6 + 1.0 * q;
    w18 = w18 + 1.0 * q;
  }
  {  // reaction 143: C2H5 + CH3 => C2H4 + CH4
    const double2 a0 = ld2(rk0 + 260), a1 = ld2(rk1 + 260);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c28)*(c18);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c28)*(c18), Pb = (c25)*(c3);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[3] += 1.0 * dqT;
      dT[18] += -1.0 * dqT;
      dT[25] += 1.0 * dqT;
      dT[28] += -1.0 * dqT;
    }
    w3 = w3 + 1.0 * q;
    w18 = w18 + -1.0 * q;
    w25 = w25 + 1.0 * q;
    w28 = w28 + -1.0 * q;
  }
  {  // reaction 144: C2H4 + CH4 => C2H5 + CH3
    const double2 a0 = ld2(rk0 + 262), a1 = ld2(rk1 + 262);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c25)*(c3);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c25)*(c3), Pb = (c28)*(c18);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[3] += -1.0 * dqT;
      dT[18] += 1.0 * dqT;
      dT[25] += -1.0 * dqT;
      dT[28] += 1.0 * dqT;
    }
    w3 = w3 + -1.0 * q;
    w18 = w18 + 1.0 * q;
    w25 = w25 + -1.0 * q;
    w28 = w28 + 1.0 * q;
  }
  {  // reaction 145: CH3CO + M => CH3 + CO + M
    const double2 a0 = ld2(rk0 + 264), a1 = ld2(rk1 + 264);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c20)*M11;
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c20), Pb = (c18)*(c9);
      const double dqT = ((a1.x - a0.x) * Pf - 0.0) * M11;
      scr[49 * SST] = kf * Pf - 0.0;
      dT[9] += 1.0 * dqT;
      dT[18] += 1.0 * dqT;
      dT[20] += -1.0 * dqT;
    }
    w9 = w9 + 1.0 * q;
    w18 = w18 + 1.0 * q;
    w20 = -1.0 * q;
  }
  {  // reaction 146: CH3 + CO + M => CH3CO + M
    const double2 a0 = ld2(rk0 + 266), a1 = ld2(rk1 + 266);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c18)*(c9)*M11;
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c18)*(c9), Pb = (c20);
      const double dqT = ((a1.x - a0.x) * Pf - 0.0) * M11;
      scr[50 * SST] = kf * Pf - 0.0;
      dT[9] += -1.0 * dqT;
      dT[18] += -1.0 * dqT;
      dT[20] += 1.0 * dqT;
    }
    w9 = w9 + -1.0 * q;
    w18 = w18 + -1.0 * q;
    w20 = w20 + 1.0 * q;
  }
  {  // reaction 147: C2H4 + O2 => C2H3 + HO2
    const double2 a0 = ld2(rk0 + 268), a1 = ld2(rk1 + 268);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c25)*(c15);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c25)*(c15), Pb = (c6)*(c24);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[6] += 1.0 * dqT;
      dT[15] += -1.0 * dqT;
      dT[24] += 1.0 * dqT;
      dT[25] += -1.0 * dqT;
    }
    w6 = w6 + 1.0 * q;
    w15 = w15 + -1.0 * q;
    w24 = w24 + 1.0 * q;
    w25 = w25 + -1.0 * q;
  }
  {  // reaction 148: C2H3 + HO2 => C2H4 + O2
    const double2 a0 = ld2(rk0 + 270), a1 = ld2(rk1 + 270);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c6)*(c24);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c6)*(c24), Pb = (c25)*(c15);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[6] += -1.0 * dqT;
      dT[15] += 1.0 * dqT;
      dT[24] += -1.0 * dqT;
      dT[25] += 1.0 * dqT;
    }
    w6 = w6 + -1.0 * q;
    w15 = w15 + 1.0 * q;
    w24 = w24 + -1.0 * q;
    w25 = w25 + 1.0 * q;
  }
  {  // reaction 149: C2H4 + HO2 => C2H3 + H2O2
    const double2 a0 = ld2(rk0 + 272), a1 = ld2(rk1 + 272);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c25)*(c24);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c25)*(c24), Pb = (c6)*(c8);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[6] += 1.0 * dqT;
      dT[8] += 1.0 * dqT;
      dT[24] += -1.0 * dqT;
      dT[25] += -1.0 * dqT;
    }
    w6 = w6 + 1.0 * q;
    w8 = w8 + 1.0 * q;
    w24 = w24 + -1.0 * q;
    w25 = w25 + -1.0 * q;
  }
  {  // reaction 150: C2H3 + H2O2 => C2H4 + HO2
    const double2 a0 = ld2(rk0 + 274), a1 = ld2(rk1 + 274);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c6)*(c8);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c6)*(c8), Pb = (c25)*(c24);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[6] += -1.0 * dqT;
      dT[8] += -1.0 * dqT;
      dT[24] += 1.0 * dqT;
      dT[25] += 1.0 * dqT;
    }
    w6 = w6 + -1.0 * q;
    w8 = w8 + -1.0 * q;
    w24 = w24 + 1.0 * q;
    w25 = w25 + 1.0 * q;
  }
  {  // reaction 151: IC3H7 => C2H4 + CH3
    const double2 a0 = ld2(rk0 + 276), a1 = ld2(rk1 + 276);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c13);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c13), Pb = (c25)*(c18);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[13] += -1.0 * dqT;
      dT[18] += 1.0 * dqT;
      dT[25] += 1.0 * dqT;
    }
    w13 = -1.0 * q;
    w18 = w18 + 1.0 * q;
    w25 = w25 + 1.0 * q;
  }
  {  // reaction 152: C2H4 + CH3 => IC3H7
    const double2 a0 = ld2(rk0 + 278), a1 = ld2(rk1 + 278);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c25)*(c18);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c25)*(c18), Pb = (c13);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[13] += 1.0 * dqT;
      dT[18] += -1.0 * dqT;
      dT[25] += -1.0 * dqT;
    }
    w13 = w13 + 1.0 * q;
    w18 = w18 + -1.0 * q;
    w25 = w25 + -1.0 * q;
  }
  {  // reaction 153: CH3CO + O => CH2CO + OH
    const double2 a0 = ld2(rk0 + 280), a1 = ld2(rk1 + 280);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c20)*(c7);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c20)*(c7), Pb = (c12)*(c27);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[7] += -1.0 * dqT;
      dT[12] += 1.0 * dqT;
      dT[20] += -1.0 * dqT;
      dT[27] += 1.0 * dqT;
    }
    w7 = w7 + -1.0 * q;
    w12 = w12 + 1.0 * q;
    w20 = w20 + -1.0 * q;
    w27 = w27 + 1.0 * q;
  }
  {  // reaction 154: CH2CO + OH => CH3CO + O
    const double2 a0 = ld2(rk0 + 282), a1 = ld2(rk1 + 282);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c12)*(c27);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c12)*(c27), Pb = (c20)*(c7);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[7] += 1.0 * dqT;
      dT[12] += -1.0 * dqT;
      dT[20] += 1.0 * dqT;
      dT[27] += -1.0 * dqT;
    }
    w7 = w7 + 1.0 * q;
    w12 = w12 + -1.0 * q;
    w20 = w20 + 1.0 * q;
    w27 = w27 + -1.0 * q;
  }
  {  // reaction 155: CH3 + CH3CO => CH2CO + CH4
    const double2 a0 = ld2(rk0 + 284), a1 = ld2(rk1 + 284);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c18)*(c20);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c18)*(c20), Pb = (c12)*(c3);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[3] += 1.0 * dqT;
      dT[12] += 1.0 * dqT;
      dT[18] += -1.0 * dqT;
      dT[20] += -1.0 * dqT;
    }
    w3 = w3 + 1.0 * q;
    w12 = w12 + 1.0 * q;
    w18 = w18 + -1.0 * q;
    w20 = w20 + -1.0 * q;
  }
  {  // reaction 156: CH2CO + CH4 => CH3 + CH3CO
    const double2 a0 = ld2(rk0 + 286), a1 = ld2(rk1 + 286);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c12)*(c3);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c12)*(c3), Pb = (c18)*(c20);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[3] += -1.0 * dqT;
      dT[12] += -1.0 * dqT;
      dT[18] += 1.0 * dqT;
      dT[20] += 1.0 * dqT;
    }
    w3 = w3 + -1.0 * q;
    w12 = w12 + -1.0 * q;
    w18 = w18 + 1.0 * q;
    w20 = w20 + 1.0 * q;
  }
  {  // reaction 157: C2H4 + O => CH2CHO + H
    const double2 a0 = ld2(rk0 + 288), a1 = ld2(rk1 + 288);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c25)*(c7);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c25)*(c7), Pb = (c1)*(c2);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[1] += 1.0 * dqT;
      dT[2] += 1.0 * dqT;
      dT[7] += -1.0 * dqT;
      dT[25] += -1.0 * dqT;
    }
    w1 = 1.0 * q;
    w2 = w2 + 1.0 * q;
    w7 = w7 + -1.0 * q;
    w25 = w25 + -1.0 * q;
  }
  {  // reaction 158: CH2CHO + H => C2H4 + O
    const double2 a0 = ld2(rk0 + 290), a1 = ld2(rk1 + 290);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c1)*(c2);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c1)*(c2), Pb = (c25)*(c7);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[1] += -1.0 * dqT;
      dT[2] += -1.0 * dqT;
      dT[7] += 1.0 * dqT;
      dT[25] += 1.0 * dqT;
    }
    w1 = w1 + -1.0 * q;
    w2 = w2 + -1.0 * q;
    w7 = w7 + 1.0 * q;
    w25 = w25 + 1.0 * q;
  }
  {  // reaction 159: C2H6 + CH => C2H5 + CH2
    const double2 a0 = ld2(rk0 + 292), a1 = ld2(rk1 + 292);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c11)*(c26);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c11)*(c26), Pb = (c28)*(c29);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[11] += -1.0 * dqT;
      dT[26] += -1.0 * dqT;
      dT[28] += 1.0 * dqT;
      dT[29] += 1.0 * dqT;
    }
    w11 = w11 + -1.0 * q;
    w26 = w26 + -1.0 * q;
    w28 = w28 + 1.0 * q;
    w29 = w29 + 1.0 * q;
  }
  {  // reaction 160: C2H5 + CH2 => C2H6 + CH
    const double2 a0 = ld2(rk0 + 294), a1 = ld2(rk1 + 294);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c28)*(c29);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c28)*(c29), Pb = (c11)*(c26);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[11] += 1.0 * dqT;
      dT[26] += 1.0 * dqT;
      dT[28] += -1.0 * dqT;
      dT[29] += -1.0 * dqT;
    }
    w11 = w11 + 1.0 * q;
    w26 = w26 + 1.0 * q;
    w28 = w28 + -1.0 * q;
    w29 = w29 + -1.0 * q;
  }
  {  // reaction 161: C2H6 + CH2 => C2H5 + CH3
    const double2 a0 = ld2(rk0 + 296), a1 = ld2(rk1 + 296);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c11)*(c29);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c11)*(c29), Pb = (c28)*(c18);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[11] += -1.0 * dqT;
      dT[18] += 1.0 * dqT;
      dT[28] += 1.0 * dqT;
      dT[29] += -1.0 * dqT;
    }
    w11 = w11 + -1.0 * q;
    w18 = w18 + 1.0 * q;
    w28 = w28 + 1.0 * q;
    w29 = w29 + -1.0 * q;
  }
  {  // reaction 162: C2H5 + CH3 => C2H6 + CH2
    const double2 a0 = ld2(rk0 + 298), a1 = ld2(rk1 + 298);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c28)*(c18);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c28)*(c18), Pb = (c11)*(c29);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[11] += 1.0 * dqT;
      dT[18] += -1.0 * dqT;
      dT[28] += -1.0 * dqT;
      dT[29] += 1.0 * dqT;
    }
    w11 = w11 + 1.0 * q;
    w18 = w18 + -1.0 * q;
    w28 = w28 + -1.0 * q;
    w29 = w29 + 1.0 * q;
  }
  {  // reaction 163: C2H4 + OH => CH2O + CH3
    const double2 a0 = ld2(rk0 + 300), a1 = ld2(rk1 + 300);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c25)*(c27);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c25)*(c27), Pb = (c10)*(c18);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[10] += 1.0 * dqT;
      dT[18] += 1.0 * dqT;
      dT[25] += -1.0 * dqT;
      dT[27] += -1.0 * dqT;
    }
    w10 = w10 + 1.0 * q;
    w18 = w18 + 1.0 * q;
    w25 = w25 + -1.0 * q;
    w27 = w27 + -1.0 * q;
  }
  {  // reaction 164: CH2O + CH3 => C2H4 + OH
    const double2 a0 = ld2(rk0 + 302), a1 = ld2(rk1 + 302);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c10)*(c18);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c10)*(c18), Pb = (c25)*(c27);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[10] += -1.0 * dqT;
      dT[18] += -1.0 * dqT;
      dT[25] += 1.0 * dqT;
      dT[27] += 1.0 * dqT;
    }
    w10 = w10 + -1.0 * q;
    w18 = w18 + -1.0 * q;
    w25 = w25 + 1.0 * q;
    w27 = w27 + 1.0 * q;
  }
  {  // reaction 165: C2H4O1-2 => CH3 + HCO
    const double2 a0 = ld2(rk0 + 304), a1 = ld2(rk1 + 304);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c22);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c22), Pb = (c18)*(c17);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[17] += 1.0 * dqT;
      dT[18] += 1.0 * dqT;
      dT[22] += -1.0 * dqT;
    }
    w17 = w17 + 1.0 * q;
    w18 = w18 + 1.0 * q;
    w22 = -1.0 * q;
  }
  {  // reaction 166: CH3 + HCO => C2H4O1-2
    const double2 a0 = ld2(rk0 + 306), a1 = ld2(rk1 + 306);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c18)*(c17);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c18)*(c17), Pb = (c22);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[17] += -1.0 * dqT;
      dT[18] += -1.0 * dqT;
      dT[22] += 1.0 * dqT;
    }
    w17 = w17 + -1.0 * q;
    w18 = w18 + -1.0 * q;
    w22 = w22 + 1.0 * q;
  }
  {  // reaction 167: C2H4O1-2 + OH => C2H3O1-2 + H2O
    const double2 a0 = ld2(rk0 + 308), a1 = ld2(rk1 + 308);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c22)*(c27);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c22)*(c27), Pb = (c23)*(c14);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[14] += 1.0 * dqT;
      dT[22] += -1.0 * dqT;
      dT[23] += 1.0 * dqT;
      dT[27] += -1.0 * dqT;
    }
    w14 = w14 + 1.0 * q;
    w22 = w22 + -1.0 * q;
    w23 = 1.0 * q;
    w27 = w27 + -1.0 * q;
  }
  {  // reaction 168: C2H3O1-2 + H2O => C2H4O1-2 + OH
    const double2 a0 = ld2(rk0 + 310), a1 = ld2(rk1 + 310);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c23)*(c14);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c23)*(c14), Pb = (c22)*(c27);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[14] += -1.0 * dqT;
      dT[22] += 1.0 * dqT;
      dT[23] += -1.0 * dqT;
      dT[27] += 1.0 * dqT;
    }
    w14 = w14 + -1.0 * q;
    w22 = w22 + 1.0 * q;
    w23 = w23 + -1.0 * q;
    w27 = w27 + 1.0 * q;
  }
  {  // reaction 169: C2H4O1-2 + H => C2H3O1-2 + H2
    const double2 a0 = ld2(rk0 + 312), a1 = ld2(rk1 + 312);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c22)*(c2);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c22)*(c2), Pb = (c23)*(c19);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[2] += -1.0 * dqT;
      dT[19] += 1.0 * dqT;
      dT[22] += -1.0 * dqT;
      dT[23] += 1.0 * dqT;
    }
    w2 = w2 + -1.0 * q;
    w19 = w19 + 1.0 * q;
    w22 = w22 + -1.0 * q;
    w23 = w23 + 1.0 * q;
  }
  {  // reaction 170: C2H3O1-2 + H2 => C2H4O1-2 + H
    const double2 a0 = ld2(rk0 + 314), a1 = ld2(rk1 + 314);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c23)*(c19);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c23)*(c19), Pb = (c22)*(c2);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[2] += 1.0 * dqT;
      dT[19] += -1.0 * dqT;
      dT[22] += 1.0 * dqT;
      dT[23] += -1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w19 = w19 + -1.0 * q;
    w22 = w22 + 1.0 * q;
    w23 = w23 + -1.0 * q;
  }
  {  // reaction 171: C2H4O1-2 + HO2 => C2H3O1-2 + H2O2
    const double2 a0 = ld2(rk0 + 316), a1 = ld2(rk1 + 316);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c22)*(c24);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c22)*(c24), Pb = (c23)*(c8);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[8] += 1.0 * dqT;
      dT[22] += -1.0 * dqT;
      dT[23] += 1.0 * dqT;
      dT[24] += -1.0 * dqT;
    }
    w8 = w8 + 1.0 * q;
    w22 = w22 + -1.0 * q;
    w23 = w23 + 1.0 * q;
    w24 = w24 + -1.0 * q;
  }
  {  // reaction 172: C2H3O1-2 + H2O2 => C2H4O1-2 + HO2
    const double2 a0 = ld2(rk0 + 318), a1 = ld2(rk1 + 318);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c23)*(c8);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c23)*(c8), Pb = (c22)*(c24);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[8] += -1.0 * dqT;
      dT[22] += 1.0 * dqT;
      dT[23] += -1.0 * dqT;
      dT[24] += 1.0 * dqT;
    }
    w8 = w8 + -1.0 * q;
    w22 = w22 + 1.0 * q;
    w23 = w23 + -1.0 * q;
    w24 = w24 + 1.0 * q;
  }
  {  // reaction 173: C2H4O1-2 + CH3 => C2H3O1-2 + CH4
    const double2 a0 = ld2(rk0 + 320), a1 = ld2(rk1 + 320);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c22)*(c18);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c22)*(c18), Pb = (c23)*(c3);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[3] += 1.0 * dqT;
      dT[18] += -1.0 * dqT;
      dT[22] += -1.0 * dqT;
      dT[23] += 1.0 * dqT;
    }
    w3 = w3 + 1.0 * q;
    w18 = w18 + -1.0 * q;
    w22 = w22 + -1.0 * q;
    w23 = w23 + 1.0 * q;
  }
  {  // reaction 174: C2H3O1-2 + CH4 => C2H4O1-2 + CH3
    const double2 a0 = ld2(rk0 + 322), a1 = ld2(rk1 + 322);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c23)*(c3);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c23)*(c3), Pb = (c22)*(c18);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[3] += -1.0 * dqT;
      dT[18] += 1.0 * dqT;
      dT[22] += 1.0 * dqT;
      dT[23] += -1.0 * dqT;
    }
    w3 = w3 + -1.0 * q;
    w18 = w18 + 1.0 * q;
    w22 = w22 + 1.0 * q;
    w23 = w23 + -1.0 * q;
  }
  {  // reaction 175: C3H6O1-3 => C2H4 + CH2O
    const double2 a0 = ld2(rk0 + 324), a1 = ld2(rk1 + 324);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c21);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c21), Pb = (c25)*(c10);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[10] += 1.0 * dqT;
      dT[21] += -1.0 * dqT;
      dT[25] += 1.0 * dqT;
    }
    w10 = w10 + 1.0 * q;
    w21 = -1.0 * q;
    w25 = w25 + 1.0 * q;
  }
  {  // reaction 176: C2H4 + CH2O => C3H6O1-3
    const double2 a0 = ld2(rk0 + 326), a1 = ld2(rk1 + 326);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c25)*(c10);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c25)*(c10), Pb = (c21);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[10] += -1.0 * dqT;
      dT[21] += 1.0 * dqT;
      dT[25] += -1.0 * dqT;
    }
    w10 = w10 + -1.0 * q;
    w21 = w21 + 1.0 * q;
    w25 = w25 + -1.0 * q;
  }
  {  // reaction 177: C3H6O1-3 + OH => C2H3 + CH2O + H2O
    const double2 a0 = ld2(rk0 + 328), a1 = ld2(rk1 + 328);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c21)*(c27);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c21)*(c27), Pb = (c6)*(c10)*(c14);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[6] += 1.0 * dqT;
      dT[10] += 1.0 * dqT;
      dT[14] += 1.0 * dqT;
      dT[21] += -1.0 * dqT;
      dT[27] += -1.0 * dqT;
    }
    w6 = w6 + 1.0 * q;
    w10 = w10 + 1.0 * q;
    w14 = w14 + 1.0 * q;
    w21 = w21 + -1.0 * q;
    w27 = w27 + -1.0 * q;
  }
  {  // reaction 178: C3H6O1-3 + O => C2H3 + CH2O + OH
    const double2 a0 = ld2(rk0 + 330), a1 = ld2(rk1 + 330);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c21)*(c7);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c21)*(c7), Pb = (c6)*(c10)*(c27);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[6] += 1.0 * dqT;
      dT[7] += -1.0 * dqT;
      dT[10] += 1.0 * dqT;
      dT[21] += -1.0 * dqT;
      dT[27] += 1.0 * dqT;
    }
    w6 = w6 + 1.0 * q;
    w7 = w7 + -1.0 * q;
    w10 = w10 + 1.0 * q;
    w21 = w21 + -1.0 * q;
    w27 = w27 + 1.0 * q;
  }
  {  // reaction 179: C3H6O1-3 + H => C2H3 + CH2O + H2
    const double2 a0 = ld2(rk0 + 332), a1 = ld2(rk1 + 332);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c21)*(c2);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c21)*(c2), Pb = (c6)*(c10)*(c19);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[2] += -1.0 * dqT;
      dT[6] += 1.0 * dqT;
      dT[10] += 1.0 * dqT;
      dT[19] += 1.0 * dqT;
      dT[21] += -1.0 * dqT;
    }
    w2 = w2 + -1.0 * q;
    w6 = w6 + 1.0 * q;
    w10 = w10 + 1.0 * q;
    w19 = w19 + 1.0 * q;
    w21 = w21 + -1.0 * q;
  }
  {  // reaction 180: C3H6O1-3 + HO2 => C2H3 + CH2O + H2O2
    const double2 a0 = ld2(rk0 + 334), a1 = ld2(rk1 + 334);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c21)*(c24);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c21)*(c24), Pb = (c6)*(c10)*(c8);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[6] += 1.0 * dqT;
      dT[8] += 1.0 * dqT;
      dT[10] += 1.0 * dqT;
      dT[21] += -1.0 * dqT;
      dT[24] += -1.0 * dqT;
    }
    w6 = w6 + 1.0 * q;
    w8 = w8 + 1.0 * q;
    w10 = w10 + 1.0 * q;
    w21 = w21 + -1.0 * q;
    w24 = w24 + -1.0 * q;
  }
  {  // reaction 181: C3H6O1-3 + CH3 => C2H3 + CH2O + CH4
    const double2 a0 = ld2(rk0 + 336), a1 = ld2(rk1 + 336);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c21)*(c18);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c21)*(c18), Pb = (c6)*(c10)*(c3);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[3] += 1.0 * dqT;
      dT[6] += 1.0 * dqT;
      dT[10] += 1.0 * dqT;
      dT[18] += -1.0 * dqT;
      dT[21] += -1.0 * dqT;
    }
    w3 = w3 + 1.0 * q;
    w6 = w6 + 1.0 * q;
    w10 = w10 + 1.0 * q;
    w18 = w18 + -1.0 * q;
    w21 = w21 + -1.0 * q;
  }
  {  // reaction 182: C3H3 + O2 => CH2CO + HCO
    const double2 a0 = ld2(rk0 + 338), a1 = ld2(rk1 + 338);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c5)*(c15);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c5)*(c15), Pb = (c12)*(c17);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[5] += -1.0 * dqT;
      dT[12] += 1.0 * dqT;
      dT[15] += -1.0 * dqT;
      dT[17] += 1.0 * dqT;
    }
    w5 = -1.0 * q;
    w12 = w12 + 1.0 * q;
    w15 = w15 + -1.0 * q;
    w17 = w17 + 1.0 * q;
  }
  {  // reaction 183: CH2CO + HCO => C3H3 + O2
    const double2 a0 = ld2(rk0 + 340), a1 = ld2(rk1 + 340);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c12)*(c17);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c12)*(c17), Pb = (c5)*(c15);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[5] += 1.0 * dqT;
      dT[12] += -1.0 * dqT;
      dT[15] += 1.0 * dqT;
      dT[17] += -1.0 * dqT;
    }
    w5 = w5 + 1.0 * q;
    w12 = w12 + -1.0 * q;
    w15 = w15 + 1.0 * q;
    w17 = w17 + -1.0 * q;
  }
  {  // reaction 184: C2H3O1-2 => CH3CO
    const double2 a0 = ld2(rk0 + 342), a1 = ld2(rk1 + 342);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c23);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c23), Pb = (c20);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[20] += 1.0 * dqT;
      dT[23] += -1.0 * dqT;
    }
    w20 = w20 + 1.0 * q;
    w23 = w23 + -1.0 * q;
  }
  {  // reaction 185: CH3CO => C2H3O1-2
    const double2 a0 = ld2(rk0 + 344), a1 = ld2(rk1 + 344);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c20);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c20), Pb = (c23);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[20] += -1.0 * dqT;
      dT[23] += 1.0 * dqT;
    }
    w20 = w20 + -1.0 * q;
    w23 = w23 + 1.0 * q;
  }
  {  // reaction 186: C2H3O1-2 => CH2CHO
    const double2 a0 = ld2(rk0 + 346), a1 = ld2(rk1 + 346);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c23);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c23), Pb = (c1);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[1] += 1.0 * dqT;
      dT[23] += -1.0 * dqT;
    }
    w1 = w1 + 1.0 * q;
    w23 = w23 + -1.0 * q;
  }
  {  // reaction 187: CH2CHO => C2H3O1-2
    const double2 a0 = ld2(rk0 + 348), a1 = ld2(rk1 + 348);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c1);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c1), Pb = (c23);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[1] += -1.0 * dqT;
      dT[23] += 1.0 * dqT;
    }
    w1 = w1 + -1.0 * q;
    w23 = w23 + 1.0 * q;
  }
  {  // reaction 188: CH2CHO => CH2CO + H
    const double2 a0 = ld2(rk0 + 350), a1 = ld2(rk1 + 350);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c1);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c1), Pb = (c12)*(c2);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[1] += -1.0 * dqT;
      dT[2] += 1.0 * dqT;
      dT[12] += 1.0 * dqT;
    }
    w1 = w1 + -1.0 * q;
    w2 = w2 + 1.0 * q;
    w12 = w12 + 1.0 * q;
  }
  {  // reaction 189: CH2CO + H => CH2CHO
    const double2 a0 = ld2(rk0 + 352), a1 = ld2(rk1 + 352);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c12)*(c2);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c12)*(c2), Pb = (c1);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[1] += 1.0 * dqT;
      dT[2] += -1.0 * dqT;
      dT[12] += -1.0 * dqT;
    }
    w1 = w1 + 1.0 * q;
    w2 = w2 + -1.0 * q;
    w12 = w12 + -1.0 * q;
  }
  {  // reaction 190: CH2CHO + O2 => CH2CO + HO2
    const double2 a0 = ld2(rk0 + 354), a1 = ld2(rk1 + 354);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c1)*(c15);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c1)*(c15), Pb = (c12)*(c24);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[1] += -1.0 * dqT;
      dT[12] += 1.0 * dqT;
      dT[15] += -1.0 * dqT;
      dT[24] += 1.0 * dqT;
    }
    w1 = w1 + -1.0 * q;
    w12 = w12 + 1.0 * q;
    w15 = w15 + -1.0 * q;
    w24 = w24 + 1.0 * q;
  }
  {  // reaction 191: CH2CO + HO2 => CH2CHO + O2
    const double2 a0 = ld2(rk0 + 356), a1 = ld2(rk1 + 356);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c12)*(c24);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c12)*(c24), Pb = (c1)*(c15);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[1] += 1.0 * dqT;
      dT[12] += -1.0 * dqT;
      dT[15] += 1.0 * dqT;
      dT[24] += -1.0 * dqT;
    }
    w1 = w1 + 1.0 * q;
    w12 = w12 + -1.0 * q;
    w15 = w15 + 1.0 * q;
    w24 = w24 + -1.0 * q;
  }
  {  // reaction 192: CH2CO + OH => CH3 + CO2
    const double2 a0 = ld2(rk0 + 358), a1 = ld2(rk1 + 358);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c12)*(c27);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c12)*(c27), Pb = (c18)*(c30);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[12] += -1.0 * dqT;
      dT[18] += 1.0 * dqT;
      dT[27] += -1.0 * dqT;
      dT[30] += 1.0 * dqT;
    }
    w12 = w12 + -1.0 * q;
    w18 = w18 + 1.0 * q;
    w27 = w27 + -1.0 * q;
    w30 = w30 + 1.0 * q;
  }
  {  // reaction 193: CH3 + CO2 => CH2CO + OH
    const double2 a0 = ld2(rk0 + 360), a1 = ld2(rk1 + 360);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c18)*(c30);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c18)*(c30), Pb = (c12)*(c27);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[12] += 1.0 * dqT;
      dT[18] += -1.0 * dqT;
      dT[27] += 1.0 * dqT;
      dT[30] += -1.0 * dqT;
    }
    w12 = w12 + 1.0 * q;
    w18 = w18 + -1.0 * q;
    w27 = w27 + 1.0 * q;
    w30 = w30 + -1.0 * q;
  }
  {  // reaction 194: IC3H7CO => CO + IC3H7
    const double2 a0 = ld2(rk0 + 362), a1 = ld2(rk1 + 362);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c0);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c0), Pb = (c9)*(c13);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[0] += -1.0 * dqT;
      dT[9] += 1.0 * dqT;
      dT[13] += 1.0 * dqT;
    }
    w0 = -1.0 * q;
    w9 = w9 + 1.0 * q;
    w13 = w13 + 1.0 * q;
  }
  {  // reaction 195: CO + IC3H7 => IC3H7CO
    const double2 a0 = ld2(rk0 + 364), a1 = ld2(rk1 + 364);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c9)*(c13);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c9)*(c13), Pb = (c0);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[0] += 1.0 * dqT;
      dT[9] += -1.0 * dqT;
      dT[13] += -1.0 * dqT;
    }
    w0 = w0 + 1.0 * q;
    w9 = w9 + -1.0 * q;
    w13 = w13 + -1.0 * q;
  }
  {  // reaction 196: C2H4 + HO2 => C2H4O1-2 + OH
    const double2 a0 = ld2(rk0 + 366), a1 = ld2(rk1 + 366);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c25)*(c24);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c25)*(c24), Pb = (c22)*(c27);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[22] += 1.0 * dqT;
      dT[24] += -1.0 * dqT;
      dT[25] += -1.0 * dqT;
      dT[27] += 1.0 * dqT;
    }
    w22 = w22 + 1.0 * q;
    w24 = w24 + -1.0 * q;
    w25 = w25 + -1.0 * q;
    w27 = w27 + 1.0 * q;
  }
  {  // reaction 197: C2H4O1-2 + OH => C2H4 + HO2
    const double2 a0 = ld2(rk0 + 368), a1 = ld2(rk1 + 368);
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c22)*(c27);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c22)*(c27), Pb = (c25)*(c24);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[22] += -1.0 * dqT;
      dT[24] += 1.0 * dqT;
      dT[25] += 1.0 * dqT;
      dT[27] += -1.0 * dqT;
    }
    w22 = w22 + -1.0 * q;
    w24 = w24 + 1.0 * q;
    w25 = w25 + 1.0 * q;
    w27 = w27 + -1.0 * q;
  }
  F[0] = 71.099 * w0;
  F[1] = 43.045 * w1;
  F[2] = 1.008 * w2;
  F[3] = 16.043 * w3;
  F[4] = 26.038 * w4;
  F[5] = 39.057 * w5;
  F[6] = 27.046 * w6;
  F[7] = 15.999 * w7;
  F[8] = 34.014 * w8;
  F[9] = 28.01 * w9;
  F[10] = 30.026 * w10;
  F[11] = 30.07 * w11;
  F[12] = 42.037 * w12;
  F[13] = 43.089 * w13;
  F[14] = 18.015 * w14;
  F[15] = 31.998 * w15;
  F[16] = 14.027 * w16;
  F[17] = 29.018 * w17;
  F[18] = 15.035 * w18;
  F[19] = 2.016 * w19;
  F[20] = 43.045 * w20;
  F[21] = 58.08 * w21;
  F[22] = 44.053 * w22;
  F[23] = 43.045 * w23;
  F[24] = 33.006 * w24;
  F[25] = 28.054 * w25;
  F[26] = 13.019 * w26;
  F[27] = 17.007 * w27;
  F[28] = 29.062 * w28;
  F[29] = 14.027 * w29;
  F[30] = 44.009 * w30;
  if (!RHS) return status;
  const int rowh = T_i - Th_min;
  const double* __restrict__ rh0 = tabh + (long)rowh * WH;
  const double* __restrict__ rh1 = rh0 + WH;
  double eiroi = 0.0, rho_cv = 0.0;
  [[maybe_unused]] double s_de = 0.0, s_dcv = 0.0;
  [[maybe_unused]] double e0, e1, e2, e3, e4, e5, e6, e7, e8, e9, e10, e11, e12, e13, e14, e15, e16, e17, e18, e19, e20, e21, e22, e23, e24, e25, e26, e27, e28, e29, e30;
  { const double2 t0 = ld2(rh0 + 0), t1 = ld2(rh1 + 0);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 116.94271368092377 * T;
    eiroi = eiroi + ek * F[0];
    rho_cv = rho_cv + r0 * (cpk - 116.94271368092377);
    if (JAC) { e0 = ek; s_de += ((t1.y - t0.y) - 116.94271368092377) * F[0]; s_dcv += r0 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 2), t1 = ld2(rh1 + 2);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 193.15855500058078 * T;
    eiroi = eiroi + ek * F[1];
    rho_cv = rho_cv + r1 * (cpk - 193.15855500058078);
    if (JAC) { e1 = ek; s_de += ((t1.y - t0.y) - 193.15855500058078) * F[1]; s_dcv += r1 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 4), t1 = ld2(rh1 + 4);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 8248.521825396825 * T;
    eiroi = eiroi + ek * F[2];
    rho_cv = rho_cv + r2 * (cpk - 8248.521825396825);
    if (JAC) { e2 = ek; s_de += ((t1.y - t0.y) - 8248.521825396825) * F[2]; s_dcv += r2 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 6), t1 = ld2(rh1 + 6);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 518.264040391448 * T;
    eiroi = eiroi + ek * F[3];
    rho_cv = rho_cv + r3 * (cpk - 518.264040391448);
    if (JAC) { e3 = ek; s_de += ((t1.y - t0.y) - 518.264040391448) * F[3]; s_dcv += r3 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 8), t1 = ld2(rh1 + 8);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 319.3221445579538 * T;
    eiroi = eiroi + ek * F[4];
    rho_cv = rho_cv + r4 * (cpk - 319.3221445579538);
    if (JAC) { e4 = ek; s_de += ((t1.y - t0.y) - 319.3221445579538) * F[4]; s_dcv += r4 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 10), t1 = ld2(rh1 + 10);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 212.8814297053025 * T;
    eiroi = eiroi + ek * F[5];
    rho_cv = rho_cv + r5 * (cpk - 212.8814297053025);
    if (JAC) { e5 = ek; s_de += ((t1.y - t0.y) - 212.8814297053025) * F[5]; s_dcv += r5 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 12), t1 = ld2(rh1 + 12);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 307.4210604155883 * T;
    eiroi = eiroi + ek * F[6];
    rho_cv = rho_cv + r6 * (cpk - 307.4210604155883);
    if (JAC) { e6 = ek; s_de += ((t1.y - t0.y) - 307.4210604155883) * F[6]; s_dcv += r6 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 14), t1 = ld2(rh1 + 14);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 519.689355584724 * T;
    eiroi = eiroi + ek * F[7];
    rho_cv = rho_cv + r7 * (cpk - 519.689355584724);
    if (JAC) { e7 = ek; s_de += ((t1.y - t0.y) - 519.689355584724) * F[7]; s_dcv += r7 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 16), t1 = ld2(rh1 + 16);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 244.44375845240194 * T;
    eiroi = eiroi + ek * F[8];
    rho_cv = rho_cv + r8 * (cpk - 244.44375845240194);
    if (JAC) { e8 = ek; s_de += ((t1.y - t0.y) - 244.44375845240194) * F[8]; s_dcv += r8 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 18), t1 = ld2(rh1 + 18);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 296.84077115315955 * T;
    eiroi = eiroi + ek * F[9];
    rho_cv = rho_cv + r9 * (cpk - 296.84077115315955);
    if (JAC) { e9 = ek; s_de += ((t1.y - t0.y) - 296.84077115315955) * F[9]; s_dcv += r9 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 20), t1 = ld2(rh1 + 20);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 276.9103443682142 * T;
    eiroi = eiroi + ek * F[10];
    rho_cv = rho_cv + r10 * (cpk - 276.9103443682142);
    if (JAC) { e10 = ek; s_de += ((t1.y - t0.y) - 276.9103443682142) * F[10]; s_dcv += r10 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 22), t1 = ld2(rh1 + 22);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 276.50515463917526 * T;
    eiroi = eiroi + ek * F[11];
    rho_cv = rho_cv + r11 * (cpk - 276.50515463917526);
    if (JAC) { e11 = ek; s_de += ((t1.y - t0.y) - 276.50515463917526) * F[11]; s_dcv += r11 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 24), t1 = ld2(rh1 + 24);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 197.79027999143614 * T;
    eiroi = eiroi + ek * F[12];
    rho_cv = rho_cv + r12 * (cpk - 197.79027999143614);
    if (JAC) { e12 = ek; s_de += ((t1.y - t0.y) - 197.79027999143614) * F[12]; s_dcv += r12 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 26), t1 = ld2(rh1 + 26);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 192.96131263199425 * T;
    eiroi = eiroi + ek * F[13];
    rho_cv = rho_cv + r13 * (cpk - 192.96131263199425);
    if (JAC) { e13 = ek; s_de += ((t1.y - t0.y) - 192.96131263199425) * F[13]; s_dcv += r13 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 28), t1 = ld2(rh1 + 28);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 461.53261171246186 * T;
    eiroi = eiroi + ek * F[14];
    rho_cv = rho_cv + r14 * (cpk - 461.53261171246186);
    if (JAC) { e14 = ek; s_de += ((t1.y - t0.y) - 461.53261171246186) * F[14]; s_dcv += r14 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 30), t1 = ld2(rh1 + 30);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 259.844677792362 * T;
    eiroi = eiroi + ek * F[15];
    rho_cv = rho_cv + r15 * (cpk - 259.844677792362);
    if (JAC) { e15 = ek; s_de += ((t1.y - t0.y) - 259.844677792362) * F[15]; s_dcv += r15 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 32), t1 = ld2(rh1 + 32);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 592.7504099237186 * T;
    eiroi = eiroi + ek * F[16];
    rho_cv = rho_cv + r16 * (cpk - 592.7504099237186);
    if (JAC) { e16 = ek; s_de += ((t1.y - t0.y) - 592.7504099237186) * F[16]; s_dcv += r16 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 34), t1 = ld2(rh1 + 34);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 286.52939554759115 * T;
    eiroi = eiroi + ek * F[17];
    rho_cv = rho_cv + r17 * (cpk - 286.52939554759115);
    if (JAC) { e17 = ek; s_de += ((t1.y - t0.y) - 286.52939554759115) * F[17]; s_dcv += r17 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 36), t1 = ld2(rh1 + 36);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 553.0103092783505 * T;
    eiroi = eiroi + ek * F[18];
    rho_cv = rho_cv + r18 * (cpk - 553.0103092783505);
    if (JAC) { e18 = ek; s_de += ((t1.y - t0.y) - 553.0103092783505) * F[18]; s_dcv += r18 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 38), t1 = ld2(rh1 + 38);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 4124.2609126984125 * T;
    eiroi = eiroi + ek * F[19];
    rho_cv = rho_cv + r19 * (cpk - 4124.2609126984125);
    if (JAC) { e19 = ek; s_de += ((t1.y - t0.y) - 4124.2609126984125) * F[19]; s_dcv += r19 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 40), t1 = ld2(rh1 + 40);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 193.15855500058078 * T;
    eiroi = eiroi + ek * F[20];
    rho_cv = rho_cv + r20 * (cpk - 193.15855500058078);
    if (JAC) { e20 = ek; s_de += ((t1.y - t0.y) - 193.15855500058078) * F[20]; s_dcv += r20 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 42), t1 = ld2(rh1 + 42);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 143.15616391184574 * T;
    eiroi = eiroi + ek * F[21];
    rho_cv = rho_cv + r21 * (cpk - 143.15616391184574);
    if (JAC) { e21 = ek; s_de += ((t1.y - t0.y) - 143.15616391184574) * F[21]; s_dcv += r21 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 44), t1 = ld2(rh1 + 44);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 188.7387919097451 * T;
    eiroi = eiroi + ek * F[22];
    rho_cv = rho_cv + r22 * (cpk - 188.7387919097451);
    if (JAC) { e22 = ek; s_de += ((t1.y - t0.y) - 188.7387919097451) * F[22]; s_dcv += r22 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 46), t1 = ld2(rh1 + 46);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 193.15855500058078 * T;
    eiroi = eiroi + ek * F[23];
    rho_cv = rho_cv + r23 * (cpk - 193.15855500058078);
    if (JAC) { e23 = ek; s_de += ((t1.y - t0.y) - 193.15855500058078) * F[23]; s_dcv += r23 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 48), t1 = ld2(rh1 + 48);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 251.90904683996848 * T;
    eiroi = eiroi + ek * F[24];
    rho_cv = rho_cv + r24 * (cpk - 251.90904683996848);
    if (JAC) { e24 = ek; s_de += ((t1.y - t0.y) - 251.90904683996848) * F[24]; s_dcv += r24 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 50), t1 = ld2(rh1 + 50);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 296.3752049618593 * T;
    eiroi = eiroi + ek * F[25];
    rho_cv = rho_cv + r25 * (cpk - 296.3752049618593);
    if (JAC) { e25 = ek; s_de += ((t1.y - t0.y) - 296.3752049618593) * F[25]; s_dcv += r25 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 52), t1 = ld2(rh1 + 52);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 638.6442891159076 * T;
    eiroi = eiroi + ek * F[26];
    rho_cv = rho_cv + r26 * (cpk - 638.6442891159076);
    if (JAC) { e26 = ek; s_de += ((t1.y - t0.y) - 638.6442891159076) * F[26]; s_dcv += r26 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 54), t1 = ld2(rh1 + 54);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 488.8875169048039 * T;
    eiroi = eiroi + ek * F[27];
    rho_cv = rho_cv + r27 * (cpk - 488.8875169048039);
    if (JAC) { e27 = ek; s_de += ((t1.y - t0.y) - 488.8875169048039) * F[27]; s_dcv += r27 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 56), t1 = ld2(rh1 + 56);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 286.0955887413117 * T;
    eiroi = eiroi + ek * F[28];
    rho_cv = rho_cv + r28 * (cpk - 286.0955887413117);
    if (JAC) { e28 = ek; s_de += ((t1.y - t0.y) - 286.0955887413117) * F[28]; s_dcv += r28 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 58), t1 = ld2(rh1 + 58);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 592.7504099237186 * T;
    eiroi = eiroi + ek * F[29];
    rho_cv = rho_cv + r29 * (cpk - 592.7504099237186);
    if (JAC) { e29 = ek; s_de += ((t1.y - t0.y) - 592.7504099237186) * F[29]; s_dcv += r29 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 60), t1 = ld2(rh1 + 60);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 188.92749210388786 * T;
    eiroi = eiroi + ek * F[30];
    rho_cv = rho_cv + r30 * (cpk - 188.92749210388786);
    if (JAC) { e30 = ek; s_de += ((t1.y - t0.y) - 188.92749210388786) * F[30]; s_dcv += r30 * (t1.x - t0.x); } }
  F[NS] = -eiroi / rho_cv;
  if (JAC) fout(F);
  if (!JAC) return status;
  const double Tdot = F[NS], invD = 1.0 / rho_cv;
  const double* const rk0_ = rk0;
  {  // column NS: d/dT
    double acc = s_de;
    { const double v = 71.099 * dT[0]; sink(0, NS, v); acc += e0 * v; }
    { const double v = 43.045 * dT[1]; sink(1, NS, v); acc += e1 * v; }
    { const double v = 1.008 * dT[2]; sink(2, NS, v); acc += e2 * v; }
    { const double v = 16.043 * dT[3]; sink(3, NS, v); acc += e3 * v; }
    { const double v = 26.038 * dT[4]; sink(4, NS, v); acc += e4 * v; }
    { const double v = 39.057 * dT[5]; sink(5, NS, v); acc += e5 * v; }
    { const double v = 27.046 * dT[6]; sink(6, NS, v); acc += e6 * v; }
    { const double v = 15.999 * dT[7]; sink(7, NS, v); acc += e7 * v; }
    { const double v = 34.014 * dT[8]; sink(8, NS, v); acc += e8 * v; }
    { const double v = 28.01 * dT[9]; sink(9, NS, v); acc += e9 * v; }
    { const double v = 30.026 * dT[10]; sink(10, NS, v); acc += e10 * v; }
    { const double v = 30.07 * dT[11]; sink(11, NS, v); acc += e11 * v; }
    { const double v = 42.037 * dT[12]; sink(12, NS, v); acc += e12 * v; }
    { const double v = 43.089 * dT[13]; sink(13, NS, v); acc += e13 * v; }
    { const double v = 18.015 * dT[14]; sink(14, NS, v); acc += e14 * v; }
    { const double v = 31.998 * dT[15]; sink(15, NS, v); acc += e15 * v; }
    { const double v = 14.027 * dT[16]; sink(16, NS, v); acc += e16 * v; }
    { const double v = 29.018 * dT[17]; sink(17, NS, v); acc += e17 * v; }
    { const double v = 15.035 * dT[18]; sink(18, NS, v); acc += e18 * v; }
    { const double v = 2.016 * dT[19]; sink(19, NS, v); acc += e19 * v; }
    { const double v = 43.045 * dT[20]; sink(20, NS, v); acc += e20 * v; }
    { const double v = 58.08 * dT[21]; sink(21, NS, v); acc += e21 * v; }
    { const double v = 44.053 * dT[22]; sink(22, NS, v); acc += e22 * v; }
    { const double v = 43.045 * dT[23]; sink(23, NS, v); acc += e23 * v; }
    { const double v = 33.006 * dT[24]; sink(24, NS, v); acc += e24 * v; }
    { const double v = 28.054 * dT[25]; sink(25, NS, v); acc += e25 * v; }
    { const double v = 13.019 * dT[26]; sink(26, NS, v); acc += e26 * v; }
    { const double v = 17.007 * dT[27]; sink(27, NS, v); acc += e27 * v; }
    { const double v = 29.062 * dT[28]; sink(28, NS, v); acc += e28 * v; }
    { const double v = 14.027 * dT[29]; sink(29, NS, v); acc += e29 * v; }
    { const double v = 44.009 * dT[30]; sink(30, NS, v); acc += e30 * v; }
    sink(NS, NS, -(acc + Tdot * s_dcv) * invD);
  }
  {  // column 0: d/drho_0 (IC3H7CO)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 0) & 1u ? 0.0 : 0.014064895427502496;
    double jc0 = 0.0;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc10 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc13 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc17 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc24 = 0.0;
    double jc25 = 0.0;
    double jc27 = 0.0;
    double jc28 = 0.0;
    double jc29 = 0.0;
    double jc30 = 0.0;
    {
      const double dq = scr[0 * SST];
      jc7 += -2.0 * dq;
      jc15 += 1.0 * dq;
    }
    {
      const double dq = scr[1 * SST];
      jc2 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc7 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc2 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    {
      const double dq = scr[11 * SST];
      jc2 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc2 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[17 * SST];
      jc2 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[20 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
    }
    {
      const double dq = scr[23 * SST];
      jc2 += -1.0 * dq;
      jc6 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = scr[29 * SST];
      jc2 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc9 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[35 * SST];
      jc8 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc9 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[41 * SST];
      jc11 += 1.0 * dq;
      jc18 += -2.0 * dq;
    }
    {
      const double dq = scr[45 * SST];
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[46 * SST];
      jc4 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    {
      const double dq = scr[49 * SST];
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc20 += -1.0 * dq;
    }
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 362), a1 = ld2(rk1 + 362);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 1.0;
      jc0 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc13 += 1.0 * dq;
    }
    double acc = 0.0;
    { const double v = 71.099 * gj * jc0; sink(0, 0, v); acc += e0 * v; }
    sink(1, 0, 0.0);
    { const double v = 1.008 * gj * jc2; sink(2, 0, v); acc += e2 * v; }
    { const double v = 16.043 * gj * jc3; sink(3, 0, v); acc += e3 * v; }
    { const double v = 26.038 * gj * jc4; sink(4, 0, v); acc += e4 * v; }
    sink(5, 0, 0.0);
    { const double v = 27.046 * gj * jc6; sink(6, 0, v); acc += e6 * v; }
    { const double v = 15.999 * gj * jc7; sink(7, 0, v); acc += e7 * v; }
    { const double v = 34.014 * gj * jc8; sink(8, 0, v); acc += e8 * v; }
    { const double v = 28.01 * gj * jc9; sink(9, 0, v); acc += e9 * v; }
    { const double v = 30.026 * gj * jc10; sink(10, 0, v); acc += e10 * v; }
    { const double v = 30.07 * gj * jc11; sink(11, 0, v); acc += e11 * v; }
    { const double v = 42.037 * gj * jc12; sink(12, 0, v); acc += e12 * v; }
    { const double v = 43.089 * gj * jc13; sink(13, 0, v); acc += e13 * v; }
    { const double v = 18.015 * gj * jc14; sink(14, 0, v); acc += e14 * v; }
    { const double v = 31.998 * gj * jc15; sink(15, 0, v); acc += e15 * v; }
    sink(16, 0, 0.0);
    { const double v = 29.018 * gj * jc17; sink(17, 0, v); acc += e17 * v; }
    { const double v = 15.035 * gj * jc18; sink(18, 0, v); acc += e18 * v; }
    { const double v = 2.016 * gj * jc19; sink(19, 0, v); acc += e19 * v; }
    { const double v = 43.045 * gj * jc20; sink(20, 0, v); acc += e20 * v; }
    sink(21, 0, 0.0);
    sink(22, 0, 0.0);
    sink(23, 0, 0.0);
    { const double v = 33.006 * gj * jc24; sink(24, 0, v); acc += e24 * v; }
    { const double v = 28.054 * gj * jc25; sink(25, 0, v); acc += e25 * v; }
    sink(26, 0, 0.0);
    { const double v = 17.007 * gj * jc27; sink(27, 0, v); acc += e27 * v; }
    { const double v = 29.062 * gj * jc28; sink(28, 0, v); acc += e28 * v; }
    { const double v = 14.027 * gj * jc29; sink(29, 0, v); acc += e29 * v; }
    { const double v = 44.009 * gj * jc30; sink(30, 0, v); acc += e30 * v; }
    const double cvj = (ld2(rh0 + 0).x + (ld2(rh1 + 0).x - ld2(rh0 + 0).x) * Td) - 116.94271368092377;
    sink(NS, 0, -(acc + Tdot * cvj * (((negmask >> 0) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 1: d/drho_1 (CH2CHO)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 1) & 1u ? 0.0 : 0.023231501916598906;
    double jc1 = 0.0;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc10 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc17 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc23 = 0.0;
    double jc24 = 0.0;
    double jc25 = 0.0;
    double jc27 = 0.0;
    double jc28 = 0.0;
    double jc29 = 0.0;
    double jc30 = 0.0;
    {
      const double dq = scr[0 * SST];
      jc7 += -2.0 * dq;
      jc15 += 1.0 * dq;
    }
    {
      const double dq = scr[1 * SST];
      jc2 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc7 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc2 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    {
      const double dq = scr[11 * SST];
      jc2 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc2 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[17 * SST];
      jc2 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[20 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
    }
    {
      const double dq = scr[23 * SST];
      jc2 += -1.0 * dq;
      jc6 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = scr[29 * SST];
      jc2 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc9 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[35 * SST];
      jc8 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc9 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[41 * SST];
      jc11 += 1.0 * dq;
      jc18 += -2.0 * dq;
    }
    {
      const double dq = scr[45 * SST];
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[46 * SST];
      jc4 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    {
      const double dq = scr[49 * SST];
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc20 += -1.0 * dq;
    }
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 290), a1 = ld2(rk1 + 290);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c2;
      jc1 += -1.0 * dq;
      jc2 += -1.0 * dq;
      jc7 += 1.0 * dq;
      jc25 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 348), a1 = ld2(rk1 + 348);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 1.0;
      jc1 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 350), a1 = ld2(rk1 + 350);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 1.0;
      jc1 += -1.0 * dq;
      jc2 += 1.0 * dq;
      jc12 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 354), a1 = ld2(rk1 + 354);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c15;
      jc1 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    double acc = 0.0;
    sink(0, 1, 0.0);
    { const double v = 43.045 * gj * jc1; sink(1, 1, v); acc += e1 * v; }
    { const double v = 1.008 * gj * jc2; sink(2, 1, v); acc += e2 * v; }
    { const double v = 16.043 * gj * jc3; sink(3, 1, v); acc += e3 * v; }
    { const double v = 26.038 * gj * jc4; sink(4, 1, v); acc += e4 * v; }
    sink(5, 1, 0.0);
    { const double v = 27.046 * gj * jc6; sink(6, 1, v); acc += e6 * v; }
    { const double v = 15.999 * gj * jc7; sink(7, 1, v); acc += e7 * v; }
    { const double v = 34.014 * gj * jc8; sink(8, 1, v); acc += e8 * v; }
    { const double v = 28.01 * gj * jc9; sink(9, 1, v); acc += e9 * v; }
    { const double v = 30.026 * gj * jc10; sink(10, 1, v); acc += e10 * v; }
    { const double v = 30.07 * gj * jc11; sink(11, 1, v); acc += e11 * v; }
    { const double v = 42.037 * gj * jc12; sink(12, 1, v); acc += e12 * v; }
    sink(13, 1, 0.0);
    { const double v = 18.015 * gj * jc14; sink(14, 1, v); acc += e14 * v; }
    { const double v = 31.998 * gj * jc15; sink(15, 1, v); acc += e15 * v; }
    sink(16, 1, 0.0);
    { const double v = 29.018 * gj * jc17; sink(17, 1, v); acc += e17 * v; }
    { const double v = 15.035 * gj * jc18; sink(18, 1, v); acc += e18 * v; }
    { const double v = 2.016 * gj * jc19; sink(19, 1, v); acc += e19 * v; }
    { const double v = 43.045 * gj * jc20; sink(20, 1, v); acc += e20 * v; }
    sink(21, 1, 0.0);
    sink(22, 1, 0.0);
    { const double v = 43.045 * gj * jc23; sink(23, 1, v); acc += e23 * v; }
    { const double v = 33.006 * gj * jc24; sink(24, 1, v); acc += e24 * v; }
    { const double v = 28.054 * gj * jc25; sink(25, 1, v); acc += e25 * v; }
    sink(26, 1, 0.0);
    { const double v = 17.007 * gj * jc27; sink(27, 1, v); acc += e27 * v; }
    { const double v = 29.062 * gj * jc28; sink(28, 1, v); acc += e28 * v; }
    { const double v = 14.027 * gj * jc29; sink(29, 1, v); acc += e29 * v; }
    { const double v = 44.009 * gj * jc30; sink(30, 1, v); acc += e30 * v; }
    const double cvj = (ld2(rh0 + 2).x + (ld2(rh1 + 2).x - ld2(rh0 + 2).x) * Td) - 193.15855500058078;
    sink(NS, 1, -(acc + Tdot * cvj * (((negmask >> 1) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 2: d/drho_2 (H)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 2) & 1u ? 0.0 : 0.9920634920634921;
    double jc1 = 0.0;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc10 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc17 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc21 = 0.0;
    double jc22 = 0.0;
    double jc23 = 0.0;
    double jc24 = 0.0;
    double jc25 = 0.0;
    double jc26 = 0.0;
    double jc27 = 0.0;
    double jc28 = 0.0;
    double jc29 = 0.0;
    double jc30 = 0.0;
    {
      const double dq = scr[0 * SST];
      jc7 += -2.0 * dq;
      jc15 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 2), a1 = ld2(rk1 + 2);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M1 * c7 + scr[1 * SST];
      jc2 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = -scr[54 * SST] * c27;
      jc2 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc19 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = -scr[60 * SST] * c9;
      jc2 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc26 += -1.0 * dq;
    }
    {
      const double dq = -scr[62 * SST] * c17;
      jc2 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = -scr[66 * SST] * c17;
      jc2 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc16 += -1.0 * dq;
      jc17 += 1.0 * dq;
    }
    {
      const double dq = -scr[68 * SST] * c10;
      jc2 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc7 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    {
      const double dq = -scr[74 * SST] * c30;
      jc2 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc17 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 32), a1 = ld2(rk1 + 32);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c12;
      jc2 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc12 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 46), a1 = ld2(rk1 + 46);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M3 * c15 + scr[3 * SST];
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 48), a1 = ld2(rk1 + 48);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M4 * c15;
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 50), a1 = ld2(rk1 + 50);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M5 * c15;
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 52), a1 = ld2(rk1 + 52);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c15;
      jc2 += -1.0 * dq;
      jc7 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 54), a1 = ld2(rk1 + 54);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M6 * 2.0*c2 + scr[6 * SST];
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 56), a1 = ld2(rk1 + 56);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M7 * 2.0*c2;
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 58), a1 = ld2(rk1 + 58);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M5 * 2.0*c2;
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 60), a1 = ld2(rk1 + 60);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M8 * 2.0*c2;
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 62), a1 = ld2(rk1 + 62);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M9 * c27 + scr[10 * SST];
      jc2 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 64), a1 = ld2(rk1 + 64);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c24;
      jc2 += -1.0 * dq;
      jc7 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc24 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 66), a1 = ld2(rk1 + 66);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c24;
      jc2 += -1.0 * dq;
      jc15 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc24 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 68), a1 = ld2(rk1 + 68);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c24;
      jc2 += -1.0 * dq;
      jc24 += -1.0 * dq;
      jc27 += 2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 70), a1 = ld2(rk1 + 70);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c8;
      jc2 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc24 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 72), a1 = ld2(rk1 + 72);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c8;
      jc2 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[12 * SST] * c29 + scr[11 * SST];
      jc2 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 74), a1 = ld2(rk1 + 74);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c16;
      jc2 += -1.0 * dq;
      jc16 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc26 += 1.0 * dq;
    }
    {
      const double dq = scr[15 * SST] * c18 + scr[14 * SST];
      jc2 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 76), a1 = ld2(rk1 + 76);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c3;
      jc2 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[18 * SST] * c17 + scr[17 * SST];
      jc2 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 78), a1 = ld2(rk1 + 78);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c17;
      jc2 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
      jc19 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 80), a1 = ld2(rk1 + 80);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c10;
      jc2 += -1.0 * dq;
      jc10 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[21 * SST] * c4 + scr[20 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
    }
    {
      const double dq = scr[24 * SST] * c6 + scr[23 * SST];
      jc2 += -1.0 * dq;
      jc6 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 82), a1 = ld2(rk1 + 82);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c6;
      jc2 += -1.0 * dq;
      jc4 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[27 * SST] * c25 + scr[26 * SST];
      jc2 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 84), a1 = ld2(rk1 + 84);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c25;
      jc2 += -1.0 * dq;
      jc6 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    {
      const double dq = scr[30 * SST] * c28 + scr[29 * SST];
      jc2 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 86), a1 = ld2(rk1 + 86);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c28;
      jc2 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 88), a1 = ld2(rk1 + 88);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c11;
      jc2 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc28 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 90), a1 = ld2(rk1 + 90);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c12;
      jc2 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc12 += -1.0 * dq;
      jc18 += 1.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc9 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 92), a1 = ld2(rk1 + 92);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c14;
      jc2 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc19 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    {
      const double dq = scr[35 * SST];
      jc8 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 102), a1 = ld2(rk1 + 102);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c17;
      jc2 += 1.0 * dq;
      jc17 += 1.0 * dq;
      jc26 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 104), a1 = ld2(rk1 + 104);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c10;
      jc2 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc27 += -1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 108), a1 = ld2(rk1 + 108);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c10;
      jc2 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc16 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 116), a1 = ld2(rk1 + 116);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c30;
      jc2 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc27 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 122), a1 = ld2(rk1 + 122);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c12;
      jc2 += 1.0 * dq;
      jc4 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 146), a1 = ld2(rk1 + 146);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c29;
      jc2 += 1.0 * dq;
      jc19 += -1.0 * dq;
      jc26 += -1.0 * dq;
      jc29 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 148), a1 = ld2(rk1 + 148);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c10;
      jc2 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc14 += -1.0 * dq;
      jc26 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 150), a1 = ld2(rk1 + 150);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c4;
      jc2 += 1.0 * dq;
      jc4 += 1.0 * dq;
      jc26 += -1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 152), a1 = ld2(rk1 + 152);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c6;
      jc2 += 1.0 * dq;
      jc6 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc26 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 154), a1 = ld2(rk1 + 154);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c25;
      jc2 += 1.0 * dq;
      jc3 += -1.0 * dq;
      jc25 += 1.0 * dq;
      jc26 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 158), a1 = ld2(rk1 + 158);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c12;
      jc2 += 1.0 * dq;
      jc10 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc26 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 162), a1 = ld2(rk1 + 162);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c18;
      jc2 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc19 += -1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 166), a1 = ld2(rk1 + 166);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c25;
      jc2 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc25 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc9 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = -scr[52 * SST] * c9*c27;
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc16 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 174), a1 = ld2(rk1 + 174);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c18;
      jc2 += 1.0 * dq;
      jc16 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 178), a1 = ld2(rk1 + 178);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c25;
      jc2 += 1.0 * dq;
      jc16 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    {
      const double dq = scr[41 * SST];
      jc11 += 1.0 * dq;
      jc18 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 194), a1 = ld2(rk1 + 194);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c28;
      jc2 += 1.0 * dq;
      jc18 += -2.0 * dq;
      jc28 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 204), a1 = ld2(rk1 + 204);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M5 * c9;
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 206), a1 = ld2(rk1 + 206);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M10 * c9 + scr[45 * SST];
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[46 * SST];
      jc4 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    {
      const double dq = scr[49 * SST];
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc20 += -1.0 * dq;
    }
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 290), a1 = ld2(rk1 + 290);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c1;
      jc1 += -1.0 * dq;
      jc2 += -1.0 * dq;
      jc7 += 1.0 * dq;
      jc25 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 312), a1 = ld2(rk1 + 312);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c22;
      jc2 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 332), a1 = ld2(rk1 + 332);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c21;
      jc2 += -1.0 * dq;
      jc6 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc21 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 352), a1 = ld2(rk1 + 352);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c12;
      jc1 += 1.0 * dq;
      jc2 += -1.0 * dq;
      jc12 += -1.0 * dq;
    }
    double acc = 0.0;
    sink(0, 2, 0.0);
    { const double v = 43.045 * gj * jc1; sink(1, 2, v); acc += e1 * v; }
    { const double v = 1.008 * gj * jc2; sink(2, 2, v); acc += e2 * v; }
    { const double v = 16.043 * gj * jc3; sink(3, 2, v); acc += e3 * v; }
    { const double v = 26.038 * gj * jc4; sink(4, 2, v); acc += e4 * v; }
    sink(5, 2, 0.0);
    { const double v = 27.046 * gj * jc6; sink(6, 2, v); acc += e6 * v; }
    { const double v = 15.999 * gj * jc7; sink(7, 2, v); acc += e7 * v; }
    { const double v = 34.014 * gj * jc8; sink(8, 2, v); acc += e8 * v; }
    { const double v = 28.01 * gj * jc9; sink(9, 2, v); acc += e9 * v; }
    { const double v = 30.026 * gj * jc10; sink(10, 2, v); acc += e10 * v; }
    { const double v = 30.07 * gj * jc11; sink(11, 2, v); acc += e11 * v; }
    { const double v = 42.037 * gj * jc12; sink(12, 2, v); acc += e12 * v; }
    sink(13, 2, 0.0);
    { const double v = 18.015 * gj * jc14; sink(14, 2, v); acc += e14 * v; }
    { const double v = 31.998 * gj * jc15; sink(15, 2, v); acc += e15 * v; }
    { const double v = 14.027 * gj * jc16; sink(16, 2, v); acc += e16 * v; }
    { const double v = 29.018 * gj * jc17; sink(17, 2, v); acc += e17 * v; }
    { const double v = 15.035 * gj * jc18; sink(18, 2, v); acc += e18 * v; }
    { const double v = 2.016 * gj * jc19; sink(19, 2, v); acc += e19 * v; }
    { const double v = 43.045 * gj * jc20; sink(20, 2, v); acc += e20 * v; }
    { const double v = 58.08 * gj * jc21; sink(21, 2, v); acc += e21 * v; }
    { const double v = 44.053 * gj * jc22; sink(22, 2, v); acc += e22 * v; }
    { const double v = 43.045 * gj * jc23; sink(23, 2, v); acc += e23 * v; }
    { const double v = 33.006 * gj * jc24; sink(24, 2, v); acc += e24 * v; }
    { const double v = 28.054 * gj * jc25; sink(25, 2, v); acc += e25 * v; }
    { const double v = 13.019 * gj * jc26; sink(26, 2, v); acc += e26 * v; }
    { const double v = 17.007 * gj * jc27; sink(27, 2, v); acc += e27 * v; }
    { const double v = 29.062 * gj * jc28; sink(28, 2, v); acc += e28 * v; }
    { const double v = 14.027 * gj * jc29; sink(29, 2, v); acc += e29 * v; }
    { const double v = 44.009 * gj * jc30; sink(30, 2, v); acc += e30 * v; }
    const double cvj = (ld2(rh0 + 4).x + (ld2(rh1 + 4).x - ld2(rh0 + 4).x) * Td) - 8248.521825396825;
    sink(NS, 2, -(acc + Tdot * cvj * (((negmask >> 2) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 3: d/drho_3 (CH4)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 3) & 1u ? 0.0 : 0.06233248145608677;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc10 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc17 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc22 = 0.0;
    double jc23 = 0.0;
    double jc24 = 0.0;
    double jc25 = 0.0;
    double jc26 = 0.0;
    double jc27 = 0.0;
    double jc28 = 0.0;
    double jc29 = 0.0;
    double jc30 = 0.0;
    {
      const double dq = 2.0 * scr[0 * SST];
      jc7 += -2.0 * dq;
      jc15 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[1 * SST];
      jc2 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[69 * SST] * c7;
      jc3 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[2 * SST];
      jc7 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[6 * SST];
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[10 * SST];
      jc2 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[11 * SST];
      jc2 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = -scr[16 * SST] * 1.0 + 2.0 * scr[14 * SST];
      jc2 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 76), a1 = ld2(rk1 + 76);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c2;
      jc2 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[17 * SST];
      jc2 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[20 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[23 * SST];
      jc2 += -1.0 * dq;
      jc6 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[26 * SST];
      jc2 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[29 * SST];
      jc2 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[32 * SST];
      jc9 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[35 * SST];
      jc8 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 114), a1 = ld2(rk1 + 114);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc3 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 138), a1 = ld2(rk1 + 138);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c15;
      jc3 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc24 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 154), a1 = ld2(rk1 + 154);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c26;
      jc2 += 1.0 * dq;
      jc3 += -1.0 * dq;
      jc25 += 1.0 * dq;
      jc26 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 168), a1 = ld2(rk1 + 168);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c29;
      jc3 += -1.0 * dq;
      jc18 += 2.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[38 * SST];
      jc9 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 180), a1 = ld2(rk1 + 180);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c16;
      jc3 += -1.0 * dq;
      jc16 += -1.0 * dq;
      jc18 += 2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 192), a1 = ld2(rk1 + 192);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c24;
      jc3 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[41 * SST];
      jc11 += 1.0 * dq;
      jc18 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 196), a1 = ld2(rk1 + 196);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c9;
      jc3 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 198), a1 = ld2(rk1 + 198);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c17;
      jc3 += 1.0 * dq;
      jc10 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 200), a1 = ld2(rk1 + 200);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c6;
      jc3 += 1.0 * dq;
      jc6 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 202), a1 = ld2(rk1 + 202);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c28;
      jc3 += 1.0 * dq;
      jc11 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[45 * SST];
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[46 * SST];
      jc4 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 258), a1 = ld2(rk1 + 258);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c4;
      jc3 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
      jc18 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 262), a1 = ld2(rk1 + 262);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c25;
      jc3 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = scr[49 * SST];
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc20 += -1.0 * dq;
    }
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 286), a1 = ld2(rk1 + 286);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c12;
      jc3 += -1.0 * dq;
      jc12 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 322), a1 = ld2(rk1 + 322);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c23;
      jc3 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc22 += 1.0 * dq;
      jc23 += -1.0 * dq;
    }
    double acc = 0.0;
    sink(0, 3, 0.0);
    sink(1, 3, 0.0);
    { const double v = 1.008 * gj * jc2; sink(2, 3, v); acc += e2 * v; }
    { const double v = 16.043 * gj * jc3; sink(3, 3, v); acc += e3 * v; }
    { const double v = 26.038 * gj * jc4; sink(4, 3, v); acc += e4 * v; }
    sink(5, 3, 0.0);
    { const double v = 27.046 * gj * jc6; sink(6, 3, v); acc += e6 * v; }
    { const double v = 15.999 * gj * jc7; sink(7, 3, v); acc += e7 * v; }
    { const double v = 34.014 * gj * jc8; sink(8, 3, v); acc += e8 * v; }
    { const double v = 28.01 * gj * jc9; sink(9, 3, v); acc += e9 * v; }
    { const double v = 30.026 * gj * jc10; sink(10, 3, v); acc += e10 * v; }
    { const double v = 30.07 * gj * jc11; sink(11, 3, v); acc += e11 * v; }
    { const double v = 42.037 * gj * jc12; sink(12, 3, v); acc += e12 * v; }
    sink(13, 3, 0.0);
    { const double v = 18.015 * gj * jc14; sink(14, 3, v); acc += e14 * v; }
    { const double v = 31.998 * gj * jc15; sink(15, 3, v); acc += e15 * v; }
    { const double v = 14.027 * gj * jc16; sink(16, 3, v); acc += e16 * v; }
    { const double v = 29.018 * gj * jc17; sink(17, 3, v); acc += e17 * v; }
    { const double v = 15.035 * gj * jc18; sink(18, 3, v); acc += e18 * v; }
    { const double v = 2.016 * gj * jc19; sink(19, 3, v); acc += e19 * v; }
    { const double v = 43.045 * gj * jc20; sink(20, 3, v); acc += e20 * v; }
    sink(21, 3, 0.0);
    { const double v = 44.053 * gj * jc22; sink(22, 3, v); acc += e22 * v; }
    { const double v = 43.045 * gj * jc23; sink(23, 3, v); acc += e23 * v; }
    { const double v = 33.006 * gj * jc24; sink(24, 3, v); acc += e24 * v; }
    { const double v = 28.054 * gj * jc25; sink(25, 3, v); acc += e25 * v; }
    { const double v = 13.019 * gj * jc26; sink(26, 3, v); acc += e26 * v; }
    { const double v = 17.007 * gj * jc27; sink(27, 3, v); acc += e27 * v; }
    { const double v = 29.062 * gj * jc28; sink(28, 3, v); acc += e28 * v; }
    { const double v = 14.027 * gj * jc29; sink(29, 3, v); acc += e29 * v; }
    { const double v = 44.009 * gj * jc30; sink(30, 3, v); acc += e30 * v; }
    const double cvj = (ld2(rh0 + 6).x + (ld2(rh1 + 6).x - ld2(rh0 + 6).x) * Td) - 518.264040391448;
    sink(NS, 3, -(acc + Tdot * cvj * (((negmask >> 3) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 4: d/drho_4 (C2H2)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 4) & 1u ? 0.0 : 0.038405407481373374;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc10 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc17 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc24 = 0.0;
    double jc25 = 0.0;
    double jc26 = 0.0;
    double jc27 = 0.0;
    double jc28 = 0.0;
    double jc29 = 0.0;
    double jc30 = 0.0;
    {
      const double dq = scr[0 * SST];
      jc7 += -2.0 * dq;
      jc15 += 1.0 * dq;
    }
    {
      const double dq = scr[1 * SST];
      jc2 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc7 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    {
      const double dq = scr[77 * SST] * c7;
      jc4 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc29 += 1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc2 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    {
      const double dq = scr[11 * SST];
      jc2 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc2 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[17 * SST];
      jc2 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[21 * SST] * c2 + scr[20 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
    }
    {
      const double dq = scr[23 * SST];
      jc2 += -1.0 * dq;
      jc6 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 82), a1 = ld2(rk1 + 82);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c19;
      jc2 += -1.0 * dq;
      jc4 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = scr[29 * SST];
      jc2 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc9 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[35 * SST];
      jc8 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 122), a1 = ld2(rk1 + 122);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc2 += 1.0 * dq;
      jc4 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 124), a1 = ld2(rk1 + 124);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc4 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 126), a1 = ld2(rk1 + 126);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c14;
      jc4 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 150), a1 = ld2(rk1 + 150);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc4 += 1.0 * dq;
      jc26 += -1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 164), a1 = ld2(rk1 + 164);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c19;
      jc4 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc29 += -2.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc9 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[41 * SST];
      jc11 += 1.0 * dq;
      jc18 += -2.0 * dq;
    }
    {
      const double dq = scr[45 * SST];
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = -scr[48 * SST] * c19 + scr[46 * SST];
      jc4 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 224), a1 = ld2(rk1 + 224);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c15;
      jc4 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc17 += 2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 234), a1 = ld2(rk1 + 234);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c24;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 258), a1 = ld2(rk1 + 258);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c3;
      jc3 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
      jc18 += 1.0 * dq;
    }
    {
      const double dq = scr[49 * SST];
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc20 += -1.0 * dq;
    }
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    double acc = 0.0;
    sink(0, 4, 0.0);
    sink(1, 4, 0.0);
    { const double v = 1.008 * gj * jc2; sink(2, 4, v); acc += e2 * v; }
    { const double v = 16.043 * gj * jc3; sink(3, 4, v); acc += e3 * v; }
    { const double v = 26.038 * gj * jc4; sink(4, 4, v); acc += e4 * v; }
    sink(5, 4, 0.0);
    { const double v = 27.046 * gj * jc6; sink(6, 4, v); acc += e6 * v; }
    { const double v = 15.999 * gj * jc7; sink(7, 4, v); acc += e7 * v; }
    { const double v = 34.014 * gj * jc8; sink(8, 4, v); acc += e8 * v; }
    { const double v = 28.01 * gj * jc9; sink(9, 4, v); acc += e9 * v; }
    { const double v = 30.026 * gj * jc10; sink(10, 4, v); acc += e10 * v; }
    { const double v = 30.07 * gj * jc11; sink(11, 4, v); acc += e11 * v; }
    { const double v = 42.037 * gj * jc12; sink(12, 4, v); acc += e12 * v; }
    sink(13, 4, 0.0);
    { const double v = 18.015 * gj * jc14; sink(14, 4, v); acc += e14 * v; }
    { const double v = 31.998 * gj * jc15; sink(15, 4, v); acc += e15 * v; }
    sink(16, 4, 0.0);
    { const double v = 29.018 * gj * jc17; sink(17, 4, v); acc += e17 * v; }
    { const double v = 15.035 * gj * jc18; sink(18, 4, v); acc += e18 * v; }
    { const double v = 2.016 * gj * jc19; sink(19, 4, v); acc += e19 * v; }
    { const double v = 43.045 * gj * jc20; sink(20, 4, v); acc += e20 * v; }
    sink(21, 4, 0.0);
    sink(22, 4, 0.0);
    sink(23, 4, 0.0);
    { const double v = 33.006 * gj * jc24; sink(24, 4, v); acc += e24 * v; }
    { const double v = 28.054 * gj * jc25; sink(25, 4, v); acc += e25 * v; }
    { const double v = 13.019 * gj * jc26; sink(26, 4, v); acc += e26 * v; }
    { const double v = 17.007 * gj * jc27; sink(27, 4, v); acc += e27 * v; }
    { const double v = 29.062 * gj * jc28; sink(28, 4, v); acc += e28 * v; }
    { const double v = 14.027 * gj * jc29; sink(29, 4, v); acc += e29 * v; }
    { const double v = 44.009 * gj * jc30; sink(30, 4, v); acc += e30 * v; }
    const double cvj = (ld2(rh0 + 8).x + (ld2(rh1 + 8).x - ld2(rh0 + 8).x) * Td) - 319.3221445579538;
    sink(NS, 4, -(acc + Tdot * cvj * (((negmask >> 4) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 5: d/drho_5 (C3H3)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 5) & 1u ? 0.0 : 0.02560360498758225;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc5 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc10 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc17 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc24 = 0.0;
    double jc25 = 0.0;
    double jc27 = 0.0;
    double jc28 = 0.0;
    double jc29 = 0.0;
    double jc30 = 0.0;
    {
      const double dq = scr[0 * SST];
      jc7 += -2.0 * dq;
      jc15 += 1.0 * dq;
    }
    {
      const double dq = scr[1 * SST];
      jc2 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc7 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc2 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    {
      const double dq = scr[11 * SST];
      jc2 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc2 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[17 * SST];
      jc2 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[20 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
    }
    {
      const double dq = scr[23 * SST];
      jc2 += -1.0 * dq;
      jc6 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = scr[29 * SST];
      jc2 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc9 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[35 * SST];
      jc8 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc9 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[41 * SST];
      jc11 += 1.0 * dq;
      jc18 += -2.0 * dq;
    }
    {
      const double dq = scr[45 * SST];
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[46 * SST];
      jc4 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    {
      const double dq = scr[49 * SST];
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc20 += -1.0 * dq;
    }
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 338), a1 = ld2(rk1 + 338);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c15;
      jc5 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc17 += 1.0 * dq;
    }
    double acc = 0.0;
    sink(0, 5, 0.0);
    sink(1, 5, 0.0);
    { const double v = 1.008 * gj * jc2; sink(2, 5, v); acc += e2 * v; }
    { const double v = 16.043 * gj * jc3; sink(3, 5, v); acc += e3 * v; }
    { const double v = 26.038 * gj * jc4; sink(4, 5, v); acc += e4 * v; }
    { const double v = 39.057 * gj * jc5; sink(5, 5, v); acc += e5 * v; }
    { const double v = 27.046 * gj * jc6; sink(6, 5, v); acc += e6 * v; }
    { const double v = 15.999 * gj * jc7; sink(7, 5, v); acc += e7 * v; }
    { const double v = 34.014 * gj * jc8; sink(8, 5, v); acc += e8 * v; }
    { const double v = 28.01 * gj * jc9; sink(9, 5, v); acc += e9 * v; }
    { const double v = 30.026 * gj * jc10; sink(10, 5, v); acc += e10 * v; }
    { const double v = 30.07 * gj * jc11; sink(11, 5, v); acc += e11 * v; }
    { const double v = 42.037 * gj * jc12; sink(12, 5, v); acc += e12 * v; }
    sink(13, 5, 0.0);
    { const double v = 18.015 * gj * jc14; sink(14, 5, v); acc += e14 * v; }
    { const double v = 31.998 * gj * jc15; sink(15, 5, v); acc += e15 * v; }
    sink(16, 5, 0.0);
    { const double v = 29.018 * gj * jc17; sink(17, 5, v); acc += e17 * v; }
    { const double v = 15.035 * gj * jc18; sink(18, 5, v); acc += e18 * v; }
    { const double v = 2.016 * gj * jc19; sink(19, 5, v); acc += e19 * v; }
    { const double v = 43.045 * gj * jc20; sink(20, 5, v); acc += e20 * v; }
    sink(21, 5, 0.0);
    sink(22, 5, 0.0);
    sink(23, 5, 0.0);
    { const double v = 33.006 * gj * jc24; sink(24, 5, v); acc += e24 * v; }
    { const double v = 28.054 * gj * jc25; sink(25, 5, v); acc += e25 * v; }
    sink(26, 5, 0.0);
    { const double v = 17.007 * gj * jc27; sink(27, 5, v); acc += e27 * v; }
    { const double v = 29.062 * gj * jc28; sink(28, 5, v); acc += e28 * v; }
    { const double v = 14.027 * gj * jc29; sink(29, 5, v); acc += e29 * v; }
    { const double v = 44.009 * gj * jc30; sink(30, 5, v); acc += e30 * v; }
    const double cvj = (ld2(rh0 + 10).x + (ld2(rh1 + 10).x - ld2(rh0 + 10).x) * Td) - 212.8814297053025;
    sink(NS, 5, -(acc + Tdot * cvj * (((negmask >> 5) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 6: d/drho_6 (C2H3)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 6) & 1u ? 0.0 : 0.03697404422095689;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc10 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc17 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc24 = 0.0;
    double jc25 = 0.0;
    double jc26 = 0.0;
    double jc27 = 0.0;
    double jc28 = 0.0;
    double jc29 = 0.0;
    double jc30 = 0.0;
    {
      const double dq = scr[0 * SST];
      jc7 += -2.0 * dq;
      jc15 += 1.0 * dq;
    }
    {
      const double dq = scr[1 * SST];
      jc2 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc7 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 32), a1 = ld2(rk1 + 32);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c7;
      jc2 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc12 += 1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc2 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    {
      const double dq = scr[11 * SST];
      jc2 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc2 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[17 * SST];
      jc2 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = -scr[22 * SST] * 1.0 + scr[20 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
    }
    {
      const double dq = scr[24 * SST] * c2 + scr[23 * SST];
      jc2 += -1.0 * dq;
      jc6 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 82), a1 = ld2(rk1 + 82);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c2;
      jc2 += -1.0 * dq;
      jc4 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 84), a1 = ld2(rk1 + 84);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c19;
      jc2 += -1.0 * dq;
      jc6 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    {
      const double dq = scr[29 * SST];
      jc2 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc9 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[35 * SST];
      jc8 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 126), a1 = ld2(rk1 + 126);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc4 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 128), a1 = ld2(rk1 + 128);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c14;
      jc6 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc25 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 152), a1 = ld2(rk1 + 152);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc6 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc26 += -1.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc9 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[41 * SST];
      jc11 += 1.0 * dq;
      jc18 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 200), a1 = ld2(rk1 + 200);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c3;
      jc3 += 1.0 * dq;
      jc6 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc25 += -1.0 * dq;
    }
    {
      const double dq = scr[45 * SST];
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 210), a1 = ld2(rk1 + 210);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c15;
      jc6 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc17 += 1.0 * dq;
    }
    {
      const double dq = scr[46 * SST];
      jc4 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 216), a1 = ld2(rk1 + 216);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c28;
      jc6 += -1.0 * dq;
      jc25 += 2.0 * dq;
      jc28 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 222), a1 = ld2(rk1 + 222);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc6 += -1.0 * dq;
      jc7 += 1.0 * dq;
      jc25 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 232), a1 = ld2(rk1 + 232);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c15;
      jc4 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 256), a1 = ld2(rk1 + 256);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc3 += 1.0 * dq;
      jc4 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[49 * SST];
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc20 += -1.0 * dq;
    }
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 270), a1 = ld2(rk1 + 270);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c24;
      jc6 += -1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 274), a1 = ld2(rk1 + 274);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c8;
      jc6 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc24 += 1.0 * dq;
      jc25 += 1.0 * dq;
    }
    double acc = 0.0;
    sink(0, 6, 0.0);
    sink(1, 6, 0.0);
    { const double v = 1.008 * gj * jc2; sink(2, 6, v); acc += e2 * v; }
    { const double v = 16.043 * gj * jc3; sink(3, 6, v); acc += e3 * v; }
    { const double v = 26.038 * gj * jc4; sink(4, 6, v); acc += e4 * v; }
    sink(5, 6, 0.0);
    { const double v = 27.046 * gj * jc6; sink(6, 6, v); acc += e6 * v; }
    { const double v = 15.999 * gj * jc7; sink(7, 6, v); acc += e7 * v; }
    { const double v = 34.014 * gj * jc8; sink(8, 6, v); acc += e8 * v; }
    { const double v = 28.01 * gj * jc9; sink(9, 6, v); acc += e9 * v; }
    { const double v = 30.026 * gj * jc10; sink(10, 6, v); acc += e10 * v; }
    { const double v = 30.07 * gj * jc11; sink(11, 6, v); acc += e11 * v; }
    { const double v = 42.037 * gj * jc12; sink(12, 6, v); acc += e12 * v; }
    sink(13, 6, 0.0);
    { const double v = 18.015 * gj * jc14; sink(14, 6, v); acc += e14 * v; }
    { const double v = 31.998 * gj * jc15; sink(15, 6, v); acc += e15 * v; }
    sink(16, 6, 0.0);
    { const double v = 29.018 * gj * jc17; sink(17, 6, v); acc += e17 * v; }
    { const double v = 15.035 * gj * jc18; sink(18, 6, v); acc += e18 * v; }
    { const double v = 2.016 * gj * jc19; sink(19, 6, v); acc += e19 * v; }
    { const double v = 43.045 * gj * jc20; sink(20, 6, v); acc += e20 * v; }
    sink(21, 6, 0.0);
    sink(22, 6, 0.0);
    sink(23, 6, 0.0);
    { const double v = 33.006 * gj * jc24; sink(24, 6, v); acc += e24 * v; }
    { const double v = 28.054 * gj * jc25; sink(25, 6, v); acc += e25 * v; }
    { const double v = 13.019 * gj * jc26; sink(26, 6, v); acc += e26 * v; }
    { const double v = 17.007 * gj * jc27; sink(27, 6, v); acc += e27 * v; }
    { const double v = 29.062 * gj * jc28; sink(28, 6, v); acc += e28 * v; }
    { const double v = 14.027 * gj * jc29; sink(29, 6, v); acc += e29 * v; }
    { const double v = 44.009 * gj * jc30; sink(30, 6, v); acc += e30 * v; }
    const double cvj = (ld2(rh0 + 12).x + (ld2(rh1 + 12).x - ld2(rh0 + 12).x) * Td) - 307.4210604155883;
    sink(NS, 6, -(acc + Tdot * cvj * (((negmask >> 6) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 7: d/drho_7 (O)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 7) & 1u ? 0.0 : 0.06250390649415588;
    double jc1 = 0.0;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc10 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc17 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc21 = 0.0;
    double jc24 = 0.0;
    double jc25 = 0.0;
    double jc26 = 0.0;
    double jc27 = 0.0;
    double jc28 = 0.0;
    double jc29 = 0.0;
    double jc30 = 0.0;
    { const double2 a0 = ld2(rk0 + 0), a1 = ld2(rk1 + 0);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M0 * 2.0*c7 + scr[0 * SST];
      jc7 += -2.0 * dq;
      jc15 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 2), a1 = ld2(rk1 + 2);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M1 * c2 + scr[1 * SST];
      jc2 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[53 * SST] * c19;
      jc2 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc19 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[55 * SST] * c24;
      jc7 += -1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[57 * SST] * c8;
      jc7 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc24 += 1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[59 * SST] * c26;
      jc2 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc26 += -1.0 * dq;
    }
    {
      const double dq = scr[61 * SST] * c29;
      jc2 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[63 * SST] * c16;
      jc7 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc16 += -1.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[65 * SST] * c16;
      jc2 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc16 += -1.0 * dq;
      jc17 += 1.0 * dq;
    }
    {
      const double dq = scr[67 * SST] * c18;
      jc2 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[69 * SST] * c3;
      jc3 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc27 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 22), a1 = ld2(rk1 + 22);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M2 * c9 + scr[2 * SST];
      jc7 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    {
      const double dq = scr[71 * SST] * c17;
      jc7 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[73 * SST] * c17;
      jc2 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc17 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    {
      const double dq = scr[75 * SST] * c10;
      jc7 += -1.0 * dq;
      jc10 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[77 * SST] * c4;
      jc4 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc29 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 32), a1 = ld2(rk1 + 32);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c6;
      jc2 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc12 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 34), a1 = ld2(rk1 + 34);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c25;
      jc7 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 36), a1 = ld2(rk1 + 36);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c28;
      jc7 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 38), a1 = ld2(rk1 + 38);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c11;
      jc7 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc27 += 1.0 * dq;
      jc28 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 40), a1 = ld2(rk1 + 40);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c12;
      jc7 += -1.0 * dq;
      jc12 += -1.0 * dq;
      jc29 += 1.0 * dq;
      jc30 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 42), a1 = ld2(rk1 + 42);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c30;
      jc7 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 52), a1 = ld2(rk1 + 52);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c27;
      jc2 += -1.0 * dq;
      jc7 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc2 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 64), a1 = ld2(rk1 + 64);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c14;
      jc2 += -1.0 * dq;
      jc7 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc24 += -1.0 * dq;
    }
    {
      const double dq = scr[11 * SST];
      jc2 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc2 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[17 * SST];
      jc2 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[20 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
    }
    {
      const double dq = scr[23 * SST];
      jc2 += -1.0 * dq;
      jc6 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = scr[29 * SST];
      jc2 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc9 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[35 * SST];
      jc8 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 94), a1 = ld2(rk1 + 94);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c14;
      jc7 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 144), a1 = ld2(rk1 + 144);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c17;
      jc7 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc26 += -1.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc9 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[41 * SST];
      jc11 += 1.0 * dq;
      jc18 += -2.0 * dq;
    }
    {
      const double dq = scr[45 * SST];
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[46 * SST];
      jc4 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 220), a1 = ld2(rk1 + 220);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c25;
      jc6 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 236), a1 = ld2(rk1 + 236);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c12;
      jc7 += -1.0 * dq;
      jc12 += -1.0 * dq;
      jc17 += 2.0 * dq;
    }
    {
      const double dq = scr[49 * SST];
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc20 += -1.0 * dq;
    }
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 280), a1 = ld2(rk1 + 280);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c20;
      jc7 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc20 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 288), a1 = ld2(rk1 + 288);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c25;
      jc1 += 1.0 * dq;
      jc2 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 330), a1 = ld2(rk1 + 330);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c21;
      jc6 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc21 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    double acc = 0.0;
    sink(0, 7, 0.0);
    { const double v = 43.045 * gj * jc1; sink(1, 7, v); acc += e1 * v; }
    { const double v = 1.008 * gj * jc2; sink(2, 7, v); acc += e2 * v; }
    { const double v = 16.043 * gj * jc3; sink(3, 7, v); acc += e3 * v; }
    { const double v = 26.038 * gj * jc4; sink(4, 7, v); acc += e4 * v; }
    sink(5, 7, 0.0);
    { const double v = 27.046 * gj * jc6; sink(6, 7, v); acc += e6 * v; }
    { const double v = 15.999 * gj * jc7; sink(7, 7, v); acc += e7 * v; }
    { const double v = 34.014 * gj * jc8; sink(8, 7, v); acc += e8 * v; }
    { const double v = 28.01 * gj * jc9; sink(9, 7, v); acc += e9 * v; }
    { const double v = 30.026 * gj * jc10; sink(10, 7, v); acc += e10 * v; }
    { const double v = 30.07 * gj * jc11; sink(11, 7, v); acc += e11 * v; }
    { const double v = 42.037 * gj * jc12; sink(12, 7, v); acc += e12 * v; }
    sink(13, 7, 0.0);
    { const double v = 18.015 * gj * jc14; sink(14, 7, v); acc += e14 * v; }
    { const double v = 31.998 * gj * jc15; sink(15, 7, v); acc += e15 * v; }
    { const double v = 14.027 * gj * jc16; sink(16, 7, v); acc += e16 * v; }
    { const double v = 29.018 * gj * jc17; sink(17, 7, v); acc += e17 * v; }
    { const double v = 15.035 * gj * jc18; sink(18, 7, v); acc += e18 * v; }
    { const double v = 2.016 * gj * jc19; sink(19, 7, v); acc += e19 * v; }
    { const double v = 43.045 * gj * jc20; sink(20, 7, v); acc += e20 * v; }
    { const double v = 58.08 * gj * jc21; sink(21, 7, v); acc += e21 * v; }
    sink(22, 7, 0.0);
    sink(23, 7, 0.0);
    { const double v = 33.006 * gj * jc24; sink(24, 7, v); acc += e24 * v; }
    { const double v = 28.054 * gj * jc25; sink(25, 7, v); acc += e25 * v; }
    { const double v = 13.019 * gj * jc26; sink(26, 7, v); acc += e26 * v; }
    { const double v = 17.007 * gj * jc27; sink(27, 7, v); acc += e27 * v; }
    { const double v = 29.062 * gj * jc28; sink(28, 7, v); acc += e28 * v; }
    { const double v = 14.027 * gj * jc29; sink(29, 7, v); acc += e29 * v; }
    { const double v = 44.009 * gj * jc30; sink(30, 7, v); acc += e30 * v; }
    const double cvj = (ld2(rh0 + 14).x + (ld2(rh1 + 14).x - ld2(rh0 + 14).x) * Td) - 519.689355584724;
    sink(NS, 7, -(acc + Tdot * cvj * (((negmask >> 7) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 8: d/drho_8 (H2O2)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 8) & 1u ? 0.0 : 0.029399658963956014;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc10 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc17 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc22 = 0.0;
    double jc23 = 0.0;
    double jc24 = 0.0;
    double jc25 = 0.0;
    double jc27 = 0.0;
    double jc28 = 0.0;
    double jc29 = 0.0;
    double jc30 = 0.0;
    {
      const double dq = scr[0 * SST];
      jc7 += -2.0 * dq;
      jc15 += 1.0 * dq;
    }
    {
      const double dq = scr[1 * SST];
      jc2 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[57 * SST] * c7;
      jc7 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc24 += 1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc7 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc2 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 70), a1 = ld2(rk1 + 70);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c2;
      jc2 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc24 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 72), a1 = ld2(rk1 + 72);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c2;
      jc2 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[11 * SST];
      jc2 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc2 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[17 * SST];
      jc2 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[20 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
    }
    {
      const double dq = scr[23 * SST];
      jc2 += -1.0 * dq;
      jc6 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = scr[29 * SST];
      jc2 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc9 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = -scr[37 * SST] * 1.0 + scr[35 * SST];
      jc8 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 98), a1 = ld2(rk1 + 98);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc8 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc24 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 100), a1 = ld2(rk1 + 100);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc8 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc24 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 132), a1 = ld2(rk1 + 132);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c15;
      jc8 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 134), a1 = ld2(rk1 + 134);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c15;
      jc8 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 142), a1 = ld2(rk1 + 142);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c17;
      jc8 += 1.0 * dq;
      jc10 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc24 += -1.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc9 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 192), a1 = ld2(rk1 + 192);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc3 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    {
      const double dq = scr[41 * SST];
      jc11 += 1.0 * dq;
      jc18 += -2.0 * dq;
    }
    {
      const double dq = scr[45 * SST];
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[46 * SST];
      jc4 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 250), a1 = ld2(rk1 + 250);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c28;
      jc8 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc24 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    {
      const double dq = scr[49 * SST];
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc20 += -1.0 * dq;
    }
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 274), a1 = ld2(rk1 + 274);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c6;
      jc6 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc24 += 1.0 * dq;
      jc25 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 318), a1 = ld2(rk1 + 318);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c23;
      jc8 += -1.0 * dq;
      jc22 += 1.0 * dq;
      jc23 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    double acc = 0.0;
    sink(0, 8, 0.0);
    sink(1, 8, 0.0);
    { const double v = 1.008 * gj * jc2; sink(2, 8, v); acc += e2 * v; }
    { const double v = 16.043 * gj * jc3; sink(3, 8, v); acc += e3 * v; }
    { const double v = 26.038 * gj * jc4; sink(4, 8, v); acc += e4 * v; }
    sink(5, 8, 0.0);
    { const double v = 27.046 * gj * jc6; sink(6, 8, v); acc += e6 * v; }
    { const double v = 15.999 * gj * jc7; sink(7, 8, v); acc += e7 * v; }
    { const double v = 34.014 * gj * jc8; sink(8, 8, v); acc += e8 * v; }
    { const double v = 28.01 * gj * jc9; sink(9, 8, v); acc += e9 * v; }
    { const double v = 30.026 * gj * jc10; sink(10, 8, v); acc += e10 * v; }
    { const double v = 30.07 * gj * jc11; sink(11, 8, v); acc += e11 * v; }
    { const double v = 42.037 * gj * jc12; sink(12, 8, v); acc += e12 * v; }
    sink(13, 8, 0.0);
    { const double v = 18.015 * gj * jc14; sink(14, 8, v); acc += e14 * v; }
    { const double v = 31.998 * gj * jc15; sink(15, 8, v); acc += e15 * v; }
    sink(16, 8, 0.0);
    { const double v = 29.018 * gj * jc17; sink(17, 8, v); acc += e17 * v; }
    { const double v = 15.035 * gj * jc18; sink(18, 8, v); acc += e18 * v; }
    { const double v = 2.016 * gj * jc19; sink(19, 8, v); acc += e19 * v; }
    { const double v = 43.045 * gj * jc20; sink(20, 8, v); acc += e20 * v; }
    sink(21, 8, 0.0);
    { const double v = 44.053 * gj * jc22; sink(22, 8, v); acc += e22 * v; }
    { const double v = 43.045 * gj * jc23; sink(23, 8, v); acc += e23 * v; }
    { const double v = 33.006 * gj * jc24; sink(24, 8, v); acc += e24 * v; }
    { const double v = 28.054 * gj * jc25; sink(25, 8, v); acc += e25 * v; }
    sink(26, 8, 0.0);
    { const double v = 17.007 * gj * jc27; sink(27, 8, v); acc += e27 * v; }
    { const double v = 29.062 * gj * jc28; sink(28, 8, v); acc += e28 * v; }
    { const double v = 14.027 * gj * jc29; sink(29, 8, v); acc += e29 * v; }
    { const double v = 44.009 * gj * jc30; sink(30, 8, v); acc += e30 * v; }
    const double cvj = (ld2(rh0 + 16).x + (ld2(rh1 + 16).x - ld2(rh0 + 16).x) * Td) - 244.44375845240194;
    sink(NS, 8, -(acc + Tdot * cvj * (((negmask >> 8) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 9: d/drho_9 (CO)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 9) & 1u ? 0.0 : 0.035701535166012134;
    double jc0 = 0.0;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc10 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc13 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc17 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc24 = 0.0;
    double jc25 = 0.0;
    double jc26 = 0.0;
    double jc27 = 0.0;
    double jc28 = 0.0;
    double jc29 = 0.0;
    double jc30 = 0.0;
    {
      const double dq = 1.75 * scr[0 * SST];
      jc7 += -2.0 * dq;
      jc15 += 1.0 * dq;
    }
    {
      const double dq = 1.5 * scr[1 * SST];
      jc2 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = -scr[60 * SST] * c2;
      jc2 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc26 += -1.0 * dq;
    }
    {
      const double dq = -scr[64 * SST] * c19;
      jc7 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc16 += -1.0 * dq;
      jc19 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 22), a1 = ld2(rk1 + 22);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M2 * c7 + 1.5 * scr[2 * SST];
      jc7 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    {
      const double dq = -scr[72 * SST] * c27;
      jc7 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = -scr[78 * SST] * c29;
      jc4 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc29 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 42), a1 = ld2(rk1 + 42);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c15;
      jc7 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    {
      const double dq = 0.75 * scr[3 * SST];
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc2 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    {
      const double dq = 1.5 * scr[11 * SST];
      jc2 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = 1.5 * scr[14 * SST];
      jc2 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = 1.5 * scr[17 * SST];
      jc2 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 78), a1 = ld2(rk1 + 78);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c19;
      jc2 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = 1.5 * scr[20 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
    }
    {
      const double dq = 1.5 * scr[23 * SST];
      jc2 += -1.0 * dq;
      jc6 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    {
      const double dq = 1.5 * scr[26 * SST];
      jc2 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = 1.5 * scr[29 * SST];
      jc2 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 90), a1 = ld2(rk1 + 90);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c18;
      jc2 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc12 += -1.0 * dq;
      jc18 += 1.0 * dq;
    }
    {
      const double dq = scr[33 * SST] * c19 + 1.5 * scr[32 * SST];
      jc9 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = 1.5 * scr[35 * SST];
      jc8 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 116), a1 = ld2(rk1 + 116);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc2 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc27 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 118), a1 = ld2(rk1 + 118);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c14;
      jc9 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc17 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 124), a1 = ld2(rk1 + 124);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c18;
      jc4 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 140), a1 = ld2(rk1 + 140);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c24;
      jc9 += -1.0 * dq;
      jc24 += -1.0 * dq;
      jc27 += 1.0 * dq;
      jc30 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 156), a1 = ld2(rk1 + 156);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c17;
      jc9 += 1.0 * dq;
      jc17 += 1.0 * dq;
      jc26 += -1.0 * dq;
      jc30 += -1.0 * dq;
    }
    {
      const double dq = scr[39 * SST] * c29 + 1.5 * scr[38 * SST];
      jc9 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = -scr[52 * SST] * c2*c27;
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc16 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 172), a1 = ld2(rk1 + 172);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c14;
      jc9 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc16 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 182), a1 = ld2(rk1 + 182);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c16 - (a0.y + (a1.y - a0.y) * Td) * c29;
      jc9 += 0.0 * dq;
      jc16 += -1.0 * dq;
      jc29 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 186), a1 = ld2(rk1 + 186);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c10;
      jc9 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc16 += -1.0 * dq;
      jc30 += -1.0 * dq;
    }
    {
      const double dq = 1.5 * scr[41 * SST];
      jc11 += 1.0 * dq;
      jc18 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 196), a1 = ld2(rk1 + 196);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c3;
      jc3 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 204), a1 = ld2(rk1 + 204);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M5 * c2;
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 206), a1 = ld2(rk1 + 206);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M10 * c2 + 1.5 * scr[45 * SST];
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 208), a1 = ld2(rk1 + 208);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c24;
      jc9 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc17 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    {
      const double dq = 1.5 * scr[46 * SST];
      jc4 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 230), a1 = ld2(rk1 + 230);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc9 += -1.0 * dq;
      jc15 += 1.0 * dq;
      jc26 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    {
      const double dq = scr[49 * SST];
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc20 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 266), a1 = ld2(rk1 + 266);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M11 * c18 + scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 364), a1 = ld2(rk1 + 364);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c13;
      jc0 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc13 += -1.0 * dq;
    }
    double acc = 0.0;
    { const double v = 71.099 * gj * jc0; sink(0, 9, v); acc += e0 * v; }
    sink(1, 9, 0.0);
    { const double v = 1.008 * gj * jc2; sink(2, 9, v); acc += e2 * v; }
    { const double v = 16.043 * gj * jc3; sink(3, 9, v); acc += e3 * v; }
    { const double v = 26.038 * gj * jc4; sink(4, 9, v); acc += e4 * v; }
    sink(5, 9, 0.0);
    { const double v = 27.046 * gj * jc6; sink(6, 9, v); acc += e6 * v; }
    { const double v = 15.999 * gj * jc7; sink(7, 9, v); acc += e7 * v; }
    { const double v = 34.014 * gj * jc8; sink(8, 9, v); acc += e8 * v; }
    { const double v = 28.01 * gj * jc9; sink(9, 9, v); acc += e9 * v; }
    { const double v = 30.026 * gj * jc10; sink(10, 9, v); acc += e10 * v; }
    { const double v = 30.07 * gj * jc11; sink(11, 9, v); acc += e11 * v; }
    { const double v = 42.037 * gj * jc12; sink(12, 9, v); acc += e12 * v; }
    { const double v = 43.089 * gj * jc13; sink(13, 9, v); acc += e13 * v; }
    { const double v = 18.015 * gj * jc14; sink(14, 9, v); acc += e14 * v; }
    { const double v = 31.998 * gj * jc15; sink(15, 9, v); acc += e15 * v; }
    { const double v = 14.027 * gj * jc16; sink(16, 9, v); acc += e16 * v; }
    { const double v = 29.018 * gj * jc17; sink(17, 9, v); acc += e17 * v; }
    { const double v = 15.035 * gj * jc18; sink(18, 9, v); acc += e18 * v; }
    { const double v = 2.016 * gj * jc19; sink(19, 9, v); acc += e19 * v; }
    { const double v = 43.045 * gj * jc20; sink(20, 9, v); acc += e20 * v; }
    sink(21, 9, 0.0);
    sink(22, 9, 0.0);
    sink(23, 9, 0.0);
    { const double v = 33.006 * gj * jc24; sink(24, 9, v); acc += e24 * v; }
    { const double v = 28.054 * gj * jc25; sink(25, 9, v); acc += e25 * v; }
    { const double v = 13.019 * gj * jc26; sink(26, 9, v); acc += e26 * v; }
    { const double v = 17.007 * gj * jc27; sink(27, 9, v); acc += e27 * v; }
    { const double v = 29.062 * gj * jc28; sink(28, 9, v); acc += e28 * v; }
    { const double v = 14.027 * gj * jc29; sink(29, 9, v); acc += e29 * v; }
    { const double v = 44.009 * gj * jc30; sink(30, 9, v); acc += e30 * v; }
    const double cvj = (ld2(rh0 + 18).x + (ld2(rh1 + 18).x - ld2(rh0 + 18).x) * Td) - 296.84077115315955;
    sink(NS, 9, -(acc + Tdot * cvj * (((negmask >> 9) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 10: d/drho_10 (CH2O)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 10) & 1u ? 0.0 : 0.033304469459801506;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc10 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc17 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc21 = 0.0;
    double jc24 = 0.0;
    double jc25 = 0.0;
    double jc26 = 0.0;
    double jc27 = 0.0;
    double jc28 = 0.0;
    double jc29 = 0.0;
    double jc30 = 0.0;
    {
      const double dq = scr[0 * SST];
      jc7 += -2.0 * dq;
      jc15 += 1.0 * dq;
    }
    {
      const double dq = scr[1 * SST];
      jc2 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = -scr[68 * SST] * c2;
      jc2 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc7 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    {
      const double dq = scr[75 * SST] * c7;
      jc7 += -1.0 * dq;
      jc10 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc27 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 36), a1 = ld2(rk1 + 36);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c18;
      jc7 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 44), a1 = ld2(rk1 + 44);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c15;
      jc10 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc24 += 1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc2 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    {
      const double dq = scr[11 * SST];
      jc2 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc2 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = -scr[19 * SST] * 1.0 + scr[17 * SST];
      jc2 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 80), a1 = ld2(rk1 + 80);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c2;
      jc2 += -1.0 * dq;
      jc10 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[20 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
    }
    {
      const double dq = scr[23 * SST];
      jc2 += -1.0 * dq;
      jc6 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = scr[29 * SST];
      jc2 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    {
      const double dq = -scr[34 * SST] * 1.0 + scr[32 * SST];
      jc9 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[35 * SST];
      jc8 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 104), a1 = ld2(rk1 + 104);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc27 += -1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 108), a1 = ld2(rk1 + 108);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc16 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 120), a1 = ld2(rk1 + 120);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc10 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc17 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 136), a1 = ld2(rk1 + 136);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c27;
      jc10 += 1.0 * dq;
      jc24 += -1.0 * dq;
      jc27 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 142), a1 = ld2(rk1 + 142);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c24;
      jc8 += 1.0 * dq;
      jc10 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc24 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 148), a1 = ld2(rk1 + 148);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc14 += -1.0 * dq;
      jc26 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 158), a1 = ld2(rk1 + 158);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c26;
      jc2 += 1.0 * dq;
      jc10 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc26 += -1.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc9 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 186), a1 = ld2(rk1 + 186);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c9;
      jc9 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc16 += -1.0 * dq;
      jc30 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 190), a1 = ld2(rk1 + 190);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c27;
      jc10 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[41 * SST];
      jc11 += 1.0 * dq;
      jc18 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 198), a1 = ld2(rk1 + 198);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc3 += 1.0 * dq;
      jc10 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[45 * SST];
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 210), a1 = ld2(rk1 + 210);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c17;
      jc6 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc17 += 1.0 * dq;
    }
    {
      const double dq = scr[46 * SST];
      jc4 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 214), a1 = ld2(rk1 + 214);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c19;
      jc10 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 242), a1 = ld2(rk1 + 242);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c17;
      jc10 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc17 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[49 * SST];
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc20 += -1.0 * dq;
    }
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 302), a1 = ld2(rk1 + 302);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc10 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc25 += 1.0 * dq;
      jc27 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 326), a1 = ld2(rk1 + 326);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c25;
      jc10 += -1.0 * dq;
      jc21 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    double acc = 0.0;
    sink(0, 10, 0.0);
    sink(1, 10, 0.0);
    { const double v = 1.008 * gj * jc2; sink(2, 10, v); acc += e2 * v; }
    { const double v = 16.043 * gj * jc3; sink(3, 10, v); acc += e3 * v; }
    { const double v = 26.038 * gj * jc4; sink(4, 10, v); acc += e4 * v; }
    sink(5, 10, 0.0);
    { const double v = 27.046 * gj * jc6; sink(6, 10, v); acc += e6 * v; }
    { const double v = 15.999 * gj * jc7; sink(7, 10, v); acc += e7 * v; }
    { const double v = 34.014 * gj * jc8; sink(8, 10, v); acc += e8 * v; }
    { const double v = 28.01 * gj * jc9; sink(9, 10, v); acc += e9 * v; }
    { const double v = 30.026 * gj * jc10; sink(10, 10, v); acc += e10 * v; }
    { const double v = 30.07 * gj * jc11; sink(11, 10, v); acc += e11 * v; }
    { const double v = 42.037 * gj * jc12; sink(12, 10, v); acc += e12 * v; }
    sink(13, 10, 0.0);
    { const double v = 18.015 * gj * jc14; sink(14, 10, v); acc += e14 * v; }
    { const double v = 31.998 * gj * jc15; sink(15, 10, v); acc += e15 * v; }
    { const double v = 14.027 * gj * jc16; sink(16, 10, v); acc += e16 * v; }
    { const double v = 29.018 * gj * jc17; sink(17, 10, v); acc += e17 * v; }
    { const double v = 15.035 * gj * jc18; sink(18, 10, v); acc += e18 * v; }
    { const double v = 2.016 * gj * jc19; sink(19, 10, v); acc += e19 * v; }
    { const double v = 43.045 * gj * jc20; sink(20, 10, v); acc += e20 * v; }
    { const double v = 58.08 * gj * jc21; sink(21, 10, v); acc += e21 * v; }
    sink(22, 10, 0.0);
    sink(23, 10, 0.0);
    { const double v = 33.006 * gj * jc24; sink(24, 10, v); acc += e24 * v; }
    { const double v = 28.054 * gj * jc25; sink(25, 10, v); acc += e25 * v; }
    { const double v = 13.019 * gj * jc26; sink(26, 10, v); acc += e26 * v; }
    { const double v = 17.007 * gj * jc27; sink(27, 10, v); acc += e27 * v; }
    { const double v = 29.062 * gj * jc28; sink(28, 10, v); acc += e28 * v; }
    { const double v = 14.027 * gj * jc29; sink(29, 10, v); acc += e29 * v; }
    { const double v = 44.009 * gj * jc30; sink(30, 10, v); acc += e30 * v; }
    const double cvj = (ld2(rh0 + 20).x + (ld2(rh1 + 20).x - ld2(rh0 + 20).x) * Td) - 276.9103443682142;
    sink(NS, 10, -(acc + Tdot * cvj * (((negmask >> 10) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 11: d/drho_11 (C2H6)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 11) & 1u ? 0.0 : 0.03325573661456601;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc10 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc17 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc24 = 0.0;
    double jc25 = 0.0;
    double jc26 = 0.0;
    double jc27 = 0.0;
    double jc28 = 0.0;
    double jc29 = 0.0;
    double jc30 = 0.0;
    {
      const double dq = 3.0 * scr[0 * SST];
      jc7 += -2.0 * dq;
      jc15 += 1.0 * dq;
    }
    {
      const double dq = 3.0 * scr[1 * SST];
      jc2 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = 3.0 * scr[2 * SST];
      jc7 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 38), a1 = ld2(rk1 + 38);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c7;
      jc7 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc27 += 1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = 1.5 * scr[3 * SST];
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    {
      const double dq = 3.0 * scr[6 * SST];
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = 3.0 * scr[10 * SST];
      jc2 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    {
      const double dq = 3.0 * scr[11 * SST];
      jc2 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = 3.0 * scr[14 * SST];
      jc2 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = 3.0 * scr[17 * SST];
      jc2 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = 3.0 * scr[20 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
    }
    {
      const double dq = 3.0 * scr[23 * SST];
      jc2 += -1.0 * dq;
      jc6 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    {
      const double dq = 3.0 * scr[26 * SST];
      jc2 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = -scr[31 * SST] * 1.0 + 3.0 * scr[29 * SST];
      jc2 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 88), a1 = ld2(rk1 + 88);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c2;
      jc2 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = 3.0 * scr[32 * SST];
      jc9 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = 3.0 * scr[35 * SST];
      jc8 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 130), a1 = ld2(rk1 + 130);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc11 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = 3.0 * scr[38 * SST];
      jc9 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 188), a1 = ld2(rk1 + 188);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c16;
      jc11 += -1.0 * dq;
      jc16 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = -scr[43 * SST] * 1.0 + 3.0 * scr[41 * SST];
      jc11 += 1.0 * dq;
      jc18 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 202), a1 = ld2(rk1 + 202);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc3 += 1.0 * dq;
      jc11 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = 3.0 * scr[45 * SST];
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = 3.0 * scr[46 * SST];
      jc4 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 244), a1 = ld2(rk1 + 244);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c15;
      jc11 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
      jc28 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 248), a1 = ld2(rk1 + 248);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c24;
      jc8 += 1.0 * dq;
      jc11 += -1.0 * dq;
      jc24 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 252), a1 = ld2(rk1 + 252);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c25;
      jc11 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 2.0 * dq;
    }
    {
      const double dq = scr[49 * SST];
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc20 += -1.0 * dq;
    }
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 292), a1 = ld2(rk1 + 292);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c26;
      jc11 += -1.0 * dq;
      jc26 += -1.0 * dq;
      jc28 += 1.0 * dq;
      jc29 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 296), a1 = ld2(rk1 + 296);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c29;
      jc11 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc28 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    double acc = 0.0;
    sink(0, 11, 0.0);
    sink(1, 11, 0.0);
    { const double v = 1.008 * gj * jc2; sink(2, 11, v); acc += e2 * v; }
    { const double v = 16.043 * gj * jc3; sink(3, 11, v); acc += e3 * v; }
    { const double v = 26.038 * gj * jc4; sink(4, 11, v); acc += e4 * v; }
    sink(5, 11, 0.0);
    { const double v = 27.046 * gj * jc6; sink(6, 11, v); acc += e6 * v; }
    { const double v = 15.999 * gj * jc7; sink(7, 11, v); acc += e7 * v; }
    { const double v = 34.014 * gj * jc8; sink(8, 11, v); acc += e8 * v; }
    { const double v = 28.01 * gj * jc9; sink(9, 11, v); acc += e9 * v; }
    { const double v = 30.026 * gj * jc10; sink(10, 11, v); acc += e10 * v; }
    { const double v = 30.07 * gj * jc11; sink(11, 11, v); acc += e11 * v; }
    { const double v = 42.037 * gj * jc12; sink(12, 11, v); acc += e12 * v; }
    sink(13, 11, 0.0);
    { const double v = 18.015 * gj * jc14; sink(14, 11, v); acc += e14 * v; }
    { const double v = 31.998 * gj * jc15; sink(15, 11, v); acc += e15 * v; }
    { const double v = 14.027 * gj * jc16; sink(16, 11, v); acc += e16 * v; }
    { const double v = 29.018 * gj * jc17; sink(17, 11, v); acc += e17 * v; }
    { const double v = 15.035 * gj * jc18; sink(18, 11, v); acc += e18 * v; }
    { const double v = 2.016 * gj * jc19; sink(19, 11, v); acc += e19 * v; }
    { const double v = 43.045 * gj * jc20; sink(20, 11, v); acc += e20 * v; }
    sink(21, 11, 0.0);
    sink(22, 11, 0.0);
    sink(23, 11, 0.0);
    { const double v = 33.006 * gj * jc24; sink(24, 11, v); acc += e24 * v; }
    { const double v = 28.054 * gj * jc25; sink(25, 11, v); acc += e25 * v; }
    { const double v = 13.019 * gj * jc26; sink(26, 11, v); acc += e26 * v; }
    { const double v = 17.007 * gj * jc27; sink(27, 11, v); acc += e27 * v; }
    { const double v = 29.062 * gj * jc28; sink(28, 11, v); acc += e28 * v; }
    { const double v = 14.027 * gj * jc29; sink(29, 11, v); acc += e29 * v; }
    { const double v = 44.009 * gj * jc30; sink(30, 11, v); acc += e30 * v; }
    const double cvj = (ld2(rh0 + 22).x + (ld2(rh1 + 22).x - ld2(rh0 + 22).x) * Td) - 276.50515463917526;
    sink(NS, 11, -(acc + Tdot * cvj * (((negmask >> 11) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 12: d/drho_12 (CH2CO)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 12) & 1u ? 0.0 : 0.023788567214596667;
    double jc1 = 0.0;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc5 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc10 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc17 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc24 = 0.0;
    double jc25 = 0.0;
    double jc26 = 0.0;
    double jc27 = 0.0;
    double jc28 = 0.0;
    double jc29 = 0.0;
    double jc30 = 0.0;
    {
      const double dq = scr[0 * SST];
      jc7 += -2.0 * dq;
      jc15 += 1.0 * dq;
    }
    {
      const double dq = scr[1 * SST];
      jc2 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc7 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 32), a1 = ld2(rk1 + 32);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc12 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 40), a1 = ld2(rk1 + 40);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c7;
      jc7 += -1.0 * dq;
      jc12 += -1.0 * dq;
      jc29 += 1.0 * dq;
      jc30 += 1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc2 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    {
      const double dq = scr[11 * SST];
      jc2 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc2 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[17 * SST];
      jc2 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[20 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
    }
    {
      const double dq = scr[23 * SST];
      jc2 += -1.0 * dq;
      jc6 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = scr[29 * SST];
      jc2 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 90), a1 = ld2(rk1 + 90);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c2;
      jc2 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc12 += -1.0 * dq;
      jc18 += 1.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc9 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[35 * SST];
      jc8 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 122), a1 = ld2(rk1 + 122);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc4 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 158), a1 = ld2(rk1 + 158);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc10 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc26 += -1.0 * dq;
    }
    {
      const double dq = -scr[40 * SST] * 1.0 + scr[38 * SST];
      jc9 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[41 * SST];
      jc11 += 1.0 * dq;
      jc18 += -2.0 * dq;
    }
    {
      const double dq = scr[45 * SST];
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[46 * SST];
      jc4 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 236), a1 = ld2(rk1 + 236);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c7;
      jc7 += -1.0 * dq;
      jc12 += -1.0 * dq;
      jc17 += 2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 240), a1 = ld2(rk1 + 240);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc10 += 1.0 * dq;
      jc12 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    {
      const double dq = scr[49 * SST];
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc20 += -1.0 * dq;
    }
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 282), a1 = ld2(rk1 + 282);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc7 += 1.0 * dq;
      jc12 += -1.0 * dq;
      jc20 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 286), a1 = ld2(rk1 + 286);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c3;
      jc3 += -1.0 * dq;
      jc12 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 340), a1 = ld2(rk1 + 340);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c17;
      jc5 += 1.0 * dq;
      jc12 += -1.0 * dq;
      jc15 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 352), a1 = ld2(rk1 + 352);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c2;
      jc1 += 1.0 * dq;
      jc2 += -1.0 * dq;
      jc12 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 356), a1 = ld2(rk1 + 356);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c24;
      jc1 += 1.0 * dq;
      jc12 += -1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 358), a1 = ld2(rk1 + 358);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc12 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc27 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    double acc = 0.0;
    sink(0, 12, 0.0);
    { const double v = 43.045 * gj * jc1; sink(1, 12, v); acc += e1 * v; }
    { const double v = 1.008 * gj * jc2; sink(2, 12, v); acc += e2 * v; }
    { const double v = 16.043 * gj * jc3; sink(3, 12, v); acc += e3 * v; }
    { const double v = 26.038 * gj * jc4; sink(4, 12, v); acc += e4 * v; }
    { const double v = 39.057 * gj * jc5; sink(5, 12, v); acc += e5 * v; }
    { const double v = 27.046 * gj * jc6; sink(6, 12, v); acc += e6 * v; }
    { const double v = 15.999 * gj * jc7; sink(7, 12, v); acc += e7 * v; }
    { const double v = 34.014 * gj * jc8; sink(8, 12, v); acc += e8 * v; }
    { const double v = 28.01 * gj * jc9; sink(9, 12, v); acc += e9 * v; }
    { const double v = 30.026 * gj * jc10; sink(10, 12, v); acc += e10 * v; }
    { const double v = 30.07 * gj * jc11; sink(11, 12, v); acc += e11 * v; }
    { const double v = 42.037 * gj * jc12; sink(12, 12, v); acc += e12 * v; }
    sink(13, 12, 0.0);
    { const double v = 18.015 * gj * jc14; sink(14, 12, v); acc += e14 * v; }
    { const double v = 31.998 * gj * jc15; sink(15, 12, v); acc += e15 * v; }
    sink(16, 12, 0.0);
    { const double v = 29.018 * gj * jc17; sink(17, 12, v); acc += e17 * v; }
    { const double v = 15.035 * gj * jc18; sink(18, 12, v); acc += e18 * v; }
    { const double v = 2.016 * gj * jc19; sink(19, 12, v); acc += e19 * v; }
    { const double v = 43.045 * gj * jc20; sink(20, 12, v); acc += e20 * v; }
    sink(21, 12, 0.0);
    sink(22, 12, 0.0);
    sink(23, 12, 0.0);
    { const double v = 33.006 * gj * jc24; sink(24, 12, v); acc += e24 * v; }
    { const double v = 28.054 * gj * jc25; sink(25, 12, v); acc += e25 * v; }
    { const double v = 13.019 * gj * jc26; sink(26, 12, v); acc += e26 * v; }
    { const double v = 17.007 * gj * jc27; sink(27, 12, v); acc += e27 * v; }
    { const double v = 29.062 * gj * jc28; sink(28, 12, v); acc += e28 * v; }
    { const double v = 14.027 * gj * jc29; sink(29, 12, v); acc += e29 * v; }
    { const double v = 44.009 * gj * jc30; sink(30, 12, v); acc += e30 * v; }
    const double cvj = (ld2(rh0 + 24).x + (ld2(rh1 + 24).x - ld2(rh0 + 24).x) * Td) - 197.79027999143614;
    sink(NS, 12, -(acc + Tdot * cvj * (((negmask >> 12) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 13: d/drho_13 (IC3H7)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 13) & 1u ? 0.0 : 0.023207779247603796;
    double jc0 = 0.0;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc10 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc13 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc17 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc24 = 0.0;
    double jc25 = 0.0;
    double jc27 = 0.0;
    double jc28 = 0.0;
    double jc29 = 0.0;
    double jc30 = 0.0;
    {
      const double dq = scr[0 * SST];
      jc7 += -2.0 * dq;
      jc15 += 1.0 * dq;
    }
    {
      const double dq = scr[1 * SST];
      jc2 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc7 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc2 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    {
      const double dq = scr[11 * SST];
      jc2 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc2 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[17 * SST];
      jc2 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[20 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
    }
    {
      const double dq = scr[23 * SST];
      jc2 += -1.0 * dq;
      jc6 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = scr[29 * SST];
      jc2 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc9 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[35 * SST];
      jc8 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc9 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[41 * SST];
      jc11 += 1.0 * dq;
      jc18 += -2.0 * dq;
    }
    {
      const double dq = scr[45 * SST];
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[46 * SST];
      jc4 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    {
      const double dq = scr[49 * SST];
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc20 += -1.0 * dq;
    }
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 276), a1 = ld2(rk1 + 276);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 1.0;
      jc13 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc25 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 364), a1 = ld2(rk1 + 364);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c9;
      jc0 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc13 += -1.0 * dq;
    }
    double acc = 0.0;
    { const double v = 71.099 * gj * jc0; sink(0, 13, v); acc += e0 * v; }
    sink(1, 13, 0.0);
    { const double v = 1.008 * gj * jc2; sink(2, 13, v); acc += e2 * v; }
    { const double v = 16.043 * gj * jc3; sink(3, 13, v); acc += e3 * v; }
    { const double v = 26.038 * gj * jc4; sink(4, 13, v); acc += e4 * v; }
    sink(5, 13, 0.0);
    { const double v = 27.046 * gj * jc6; sink(6, 13, v); acc += e6 * v; }
    { const double v = 15.999 * gj * jc7; sink(7, 13, v); acc += e7 * v; }
    { const double v = 34.014 * gj * jc8; sink(8, 13, v); acc += e8 * v; }
    { const double v = 28.01 * gj * jc9; sink(9, 13, v); acc += e9 * v; }
    { const double v = 30.026 * gj * jc10; sink(10, 13, v); acc += e10 * v; }
    { const double v = 30.07 * gj * jc11; sink(11, 13, v); acc += e11 * v; }
    { const double v = 42.037 * gj * jc12; sink(12, 13, v); acc += e12 * v; }
    { const double v = 43.089 * gj * jc13; sink(13, 13, v); acc += e13 * v; }
    { const double v = 18.015 * gj * jc14; sink(14, 13, v); acc += e14 * v; }
    { const double v = 31.998 * gj * jc15; sink(15, 13, v); acc += e15 * v; }
    sink(16, 13, 0.0);
    { const double v = 29.018 * gj * jc17; sink(17, 13, v); acc += e17 * v; }
    { const double v = 15.035 * gj * jc18; sink(18, 13, v); acc += e18 * v; }
    { const double v = 2.016 * gj * jc19; sink(19, 13, v); acc += e19 * v; }
    { const double v = 43.045 * gj * jc20; sink(20, 13, v); acc += e20 * v; }
    sink(21, 13, 0.0);
    sink(22, 13, 0.0);
    sink(23, 13, 0.0);
    { const double v = 33.006 * gj * jc24; sink(24, 13, v); acc += e24 * v; }
    { const double v = 28.054 * gj * jc25; sink(25, 13, v); acc += e25 * v; }
    sink(26, 13, 0.0);
    { const double v = 17.007 * gj * jc27; sink(27, 13, v); acc += e27 * v; }
    { const double v = 29.062 * gj * jc28; sink(28, 13, v); acc += e28 * v; }
    { const double v = 14.027 * gj * jc29; sink(29, 13, v); acc += e29 * v; }
    { const double v = 44.009 * gj * jc30; sink(30, 13, v); acc += e30 * v; }
    const double cvj = (ld2(rh0 + 26).x + (ld2(rh1 + 26).x - ld2(rh0 + 26).x) * Td) - 192.96131263199425;
    sink(NS, 13, -(acc + Tdot * cvj * (((negmask >> 13) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 14: d/drho_14 (H2O)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 14) & 1u ? 0.0 : 0.055509297807382736;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc10 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc17 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc22 = 0.0;
    double jc23 = 0.0;
    double jc24 = 0.0;
    double jc25 = 0.0;
    double jc26 = 0.0;
    double jc27 = 0.0;
    double jc28 = 0.0;
    double jc29 = 0.0;
    double jc30 = 0.0;
    {
      const double dq = 15.399999618530273 * scr[0 * SST];
      jc7 += -2.0 * dq;
      jc15 += 1.0 * dq;
    }
    {
      const double dq = 6.0 * scr[1 * SST];
      jc2 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = 6.0 * scr[2 * SST];
      jc7 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    {
      const double dq = scr[5 * SST];
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    {
      const double dq = scr[8 * SST];
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 62), a1 = ld2(rk1 + 62);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M9 * 1.0 + 3.6500000953674316 * scr[10 * SST];
      jc2 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 64), a1 = ld2(rk1 + 64);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c7;
      jc2 += -1.0 * dq;
      jc7 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc24 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 72), a1 = ld2(rk1 + 72);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c27;
      jc2 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = 6.0 * scr[11 * SST];
      jc2 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = 6.0 * scr[14 * SST];
      jc2 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = 6.0 * scr[17 * SST];
      jc2 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = 6.0 * scr[20 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
    }
    {
      const double dq = 6.0 * scr[23 * SST];
      jc2 += -1.0 * dq;
      jc6 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    {
      const double dq = 6.0 * scr[26 * SST];
      jc2 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = 6.0 * scr[29 * SST];
      jc2 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    {
      const double dq = 6.0 * scr[32 * SST];
      jc9 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 92), a1 = ld2(rk1 + 92);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc19 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    {
      const double dq = 6.0 * scr[35 * SST];
      jc8 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 94), a1 = ld2(rk1 + 94);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c7;
      jc7 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 96), a1 = ld2(rk1 + 96);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c15;
      jc14 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 98), a1 = ld2(rk1 + 98);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c24;
      jc8 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc24 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 100), a1 = ld2(rk1 + 100);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c24;
      jc8 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc24 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 106), a1 = ld2(rk1 + 106);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c26;
      jc14 += 1.0 * dq;
      jc26 += 1.0 * dq;
      jc27 += -1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 110), a1 = ld2(rk1 + 110);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c29;
      jc14 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc27 += -1.0 * dq;
      jc29 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 112), a1 = ld2(rk1 + 112);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c16;
      jc14 += 1.0 * dq;
      jc16 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 114), a1 = ld2(rk1 + 114);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c18;
      jc3 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 118), a1 = ld2(rk1 + 118);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c9;
      jc9 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc17 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 120), a1 = ld2(rk1 + 120);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c17;
      jc10 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc17 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 126), a1 = ld2(rk1 + 126);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c4;
      jc4 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 128), a1 = ld2(rk1 + 128);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c6;
      jc6 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc25 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 130), a1 = ld2(rk1 + 130);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c28;
      jc11 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 148), a1 = ld2(rk1 + 148);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c26;
      jc2 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc14 += -1.0 * dq;
      jc26 += -1.0 * dq;
    }
    {
      const double dq = 6.0 * scr[38 * SST];
      jc9 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 172), a1 = ld2(rk1 + 172);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c9;
      jc9 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc16 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 176), a1 = ld2(rk1 + 176);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c16 - (a0.y + (a1.y - a0.y) * Td) * c29;
      jc14 += 0.0 * dq;
      jc16 += -1.0 * dq;
      jc29 += 1.0 * dq;
    }
    {
      const double dq = 6.0 * scr[41 * SST];
      jc11 += 1.0 * dq;
      jc18 += -2.0 * dq;
    }
    {
      const double dq = scr[44 * SST];
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = 6.0 * scr[46 * SST];
      jc4 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    {
      const double dq = scr[49 * SST];
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc20 += -1.0 * dq;
    }
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 310), a1 = ld2(rk1 + 310);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c23;
      jc14 += -1.0 * dq;
      jc22 += 1.0 * dq;
      jc23 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    double acc = 0.0;
    sink(0, 14, 0.0);
    sink(1, 14, 0.0);
    { const double v = 1.008 * gj * jc2; sink(2, 14, v); acc += e2 * v; }
    { const double v = 16.043 * gj * jc3; sink(3, 14, v); acc += e3 * v; }
    { const double v = 26.038 * gj * jc4; sink(4, 14, v); acc += e4 * v; }
    sink(5, 14, 0.0);
    { const double v = 27.046 * gj * jc6; sink(6, 14, v); acc += e6 * v; }
    { const double v = 15.999 * gj * jc7; sink(7, 14, v); acc += e7 * v; }
    { const double v = 34.014 * gj * jc8; sink(8, 14, v); acc += e8 * v; }
    { const double v = 28.01 * gj * jc9; sink(9, 14, v); acc += e9 * v; }
    { const double v = 30.026 * gj * jc10; sink(10, 14, v); acc += e10 * v; }
    { const double v = 30.07 * gj * jc11; sink(11, 14, v); acc += e11 * v; }
    { const double v = 42.037 * gj * jc12; sink(12, 14, v); acc += e12 * v; }
    sink(13, 14, 0.0);
    { const double v = 18.015 * gj * jc14; sink(14, 14, v); acc += e14 * v; }
    { const double v = 31.998 * gj * jc15; sink(15, 14, v); acc += e15 * v; }
    { const double v = 14.027 * gj * jc16; sink(16, 14, v); acc += e16 * v; }
    { const double v = 29.018 * gj * jc17; sink(17, 14, v); acc += e17 * v; }
    { const double v = 15.035 * gj * jc18; sink(18, 14, v); acc += e18 * v; }
    { const double v = 2.016 * gj * jc19; sink(19, 14, v); acc += e19 * v; }
    { const double v = 43.045 * gj * jc20; sink(20, 14, v); acc += e20 * v; }
    sink(21, 14, 0.0);
    { const double v = 44.053 * gj * jc22; sink(22, 14, v); acc += e22 * v; }
    { const double v = 43.045 * gj * jc23; sink(23, 14, v); acc += e23 * v; }
    { const double v = 33.006 * gj * jc24; sink(24, 14, v); acc += e24 * v; }
    { const double v = 28.054 * gj * jc25; sink(25, 14, v); acc += e25 * v; }
    { const double v = 13.019 * gj * jc26; sink(26, 14, v); acc += e26 * v; }
    { const double v = 17.007 * gj * jc27; sink(27, 14, v); acc += e27 * v; }
    { const double v = 29.062 * gj * jc28; sink(28, 14, v); acc += e28 * v; }
    { const double v = 14.027 * gj * jc29; sink(29, 14, v); acc += e29 * v; }
    { const double v = 44.009 * gj * jc30; sink(30, 14, v); acc += e30 * v; }
    const double cvj = (ld2(rh0 + 28).x + (ld2(rh1 + 28).x - ld2(rh0 + 28).x) * Td) - 461.53261171246186;
    sink(NS, 14, -(acc + Tdot * cvj * (((negmask >> 14) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 15: d/drho_15 (O2)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 15) & 1u ? 0.0 : 0.03125195324707794;
    double jc1 = 0.0;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc5 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc10 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc17 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc24 = 0.0;
    double jc25 = 0.0;
    double jc26 = 0.0;
    double jc27 = 0.0;
    double jc28 = 0.0;
    double jc29 = 0.0;
    double jc30 = 0.0;
    { const double2 a0 = ld2(rk0 + 0), a1 = ld2(rk1 + 0);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M0 * 1.0 + scr[0 * SST];
      jc7 += -2.0 * dq;
      jc15 += 1.0 * dq;
    }
    {
      const double dq = scr[1 * SST];
      jc2 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = -scr[56 * SST] * c27;
      jc7 += -1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = 6.0 * scr[2 * SST];
      jc7 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 42), a1 = ld2(rk1 + 42);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c9;
      jc7 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 44), a1 = ld2(rk1 + 44);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c10;
      jc10 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc24 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 46), a1 = ld2(rk1 + 46);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M3 * c2;
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 48), a1 = ld2(rk1 + 48);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M4 * c2 + scr[4 * SST];
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 50), a1 = ld2(rk1 + 50);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M5 * c2;
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 52), a1 = ld2(rk1 + 52);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c2;
      jc2 += -1.0 * dq;
      jc7 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc2 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 66), a1 = ld2(rk1 + 66);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c19;
      jc2 += -1.0 * dq;
      jc15 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc24 += -1.0 * dq;
    }
    {
      const double dq = scr[11 * SST];
      jc2 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc2 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[17 * SST];
      jc2 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[20 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
    }
    {
      const double dq = scr[23 * SST];
      jc2 += -1.0 * dq;
      jc6 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = scr[29 * SST];
      jc2 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc9 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[35 * SST];
      jc8 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 96), a1 = ld2(rk1 + 96);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c14;
      jc14 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 132), a1 = ld2(rk1 + 132);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c8;
      jc8 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 134), a1 = ld2(rk1 + 134);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c8;
      jc8 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 138), a1 = ld2(rk1 + 138);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c3;
      jc3 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc24 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 144), a1 = ld2(rk1 + 144);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c26;
      jc7 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc26 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 160), a1 = ld2(rk1 + 160);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c29;
      jc15 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc27 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc9 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[51 * SST] * c16;
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc16 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 172), a1 = ld2(rk1 + 172);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c16;
      jc9 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc16 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 190), a1 = ld2(rk1 + 190);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc10 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[41 * SST];
      jc11 += 1.0 * dq;
      jc18 += -2.0 * dq;
    }
    {
      const double dq = scr[45 * SST];
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 208), a1 = ld2(rk1 + 208);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c17;
      jc9 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc17 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 210), a1 = ld2(rk1 + 210);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c6;
      jc6 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc17 += 1.0 * dq;
    }
    {
      const double dq = scr[46 * SST];
      jc4 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 212), a1 = ld2(rk1 + 212);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c28;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
      jc25 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 224), a1 = ld2(rk1 + 224);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c4;
      jc4 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc17 += 2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 228), a1 = ld2(rk1 + 228);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c26;
      jc9 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc26 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 232), a1 = ld2(rk1 + 232);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c6;
      jc4 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 244), a1 = ld2(rk1 + 244);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c11;
      jc11 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = scr[49 * SST];
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc20 += -1.0 * dq;
    }
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 268), a1 = ld2(rk1 + 268);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c25;
      jc6 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 338), a1 = ld2(rk1 + 338);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c5;
      jc5 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc17 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 354), a1 = ld2(rk1 + 354);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c1;
      jc1 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    double acc = 0.0;
    sink(0, 15, 0.0);
    { const double v = 43.045 * gj * jc1; sink(1, 15, v); acc += e1 * v; }
    { const double v = 1.008 * gj * jc2; sink(2, 15, v); acc += e2 * v; }
    { const double v = 16.043 * gj * jc3; sink(3, 15, v); acc += e3 * v; }
    { const double v = 26.038 * gj * jc4; sink(4, 15, v); acc += e4 * v; }
    { const double v = 39.057 * gj * jc5; sink(5, 15, v); acc += e5 * v; }
    { const double v = 27.046 * gj * jc6; sink(6, 15, v); acc += e6 * v; }
    { const double v = 15.999 * gj * jc7; sink(7, 15, v); acc += e7 * v; }
    { const double v = 34.014 * gj * jc8; sink(8, 15, v); acc += e8 * v; }
    { const double v = 28.01 * gj * jc9; sink(9, 15, v); acc += e9 * v; }
    { const double v = 30.026 * gj * jc10; sink(10, 15, v); acc += e10 * v; }
    { const double v = 30.07 * gj * jc11; sink(11, 15, v); acc += e11 * v; }
    { const double v = 42.037 * gj * jc12; sink(12, 15, v); acc += e12 * v; }
    sink(13, 15, 0.0);
    { const double v = 18.015 * gj * jc14; sink(14, 15, v); acc += e14 * v; }
    { const double v = 31.998 * gj * jc15; sink(15, 15, v); acc += e15 * v; }
    { const double v = 14.027 * gj * jc16; sink(16, 15, v); acc += e16 * v; }
    { const double v = 29.018 * gj * jc17; sink(17, 15, v); acc += e17 * v; }
    { const double v = 15.035 * gj * jc18; sink(18, 15, v); acc += e18 * v; }
    { const double v = 2.016 * gj * jc19; sink(19, 15, v); acc += e19 * v; }
    { const double v = 43.045 * gj * jc20; sink(20, 15, v); acc += e20 * v; }
    sink(21, 15, 0.0);
    sink(22, 15, 0.0);
    sink(23, 15, 0.0);
    { const double v = 33.006 * gj * jc24; sink(24, 15, v); acc += e24 * v; }
    { const double v = 28.054 * gj * jc25; sink(25, 15, v); acc += e25 * v; }
    { const double v = 13.019 * gj * jc26; sink(26, 15, v); acc += e26 * v; }
    { const double v = 17.007 * gj * jc27; sink(27, 15, v); acc += e27 * v; }
    { const double v = 29.062 * gj * jc28; sink(28, 15, v); acc += e28 * v; }
    { const double v = 14.027 * gj * jc29; sink(29, 15, v); acc += e29 * v; }
    { const double v = 44.009 * gj * jc30; sink(30, 15, v); acc += e30 * v; }
    const double cvj = (ld2(rh0 + 30).x + (ld2(rh1 + 30).x - ld2(rh0 + 30).x) * Td) - 259.844677792362;
    sink(NS, 15, -(acc + Tdot * cvj * (((negmask >> 15) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 16: d/drho_16 (CH2(S))
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 16) & 1u ? 0.0 : 0.07129108148570615;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc10 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc17 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc24 = 0.0;
    double jc25 = 0.0;
    double jc26 = 0.0;
    double jc27 = 0.0;
    double jc28 = 0.0;
    double jc29 = 0.0;
    double jc30 = 0.0;
    {
      const double dq = scr[0 * SST];
      jc7 += -2.0 * dq;
      jc15 += 1.0 * dq;
    }
    {
      const double dq = scr[1 * SST];
      jc2 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[63 * SST] * c7;
      jc7 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc16 += -1.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[65 * SST] * c7;
      jc2 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc16 += -1.0 * dq;
      jc17 += 1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc7 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc2 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    {
      const double dq = scr[11 * SST];
      jc2 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 74), a1 = ld2(rk1 + 74);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c2;
      jc2 += -1.0 * dq;
      jc16 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc26 += 1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc2 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[17 * SST];
      jc2 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[20 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
    }
    {
      const double dq = scr[23 * SST];
      jc2 += -1.0 * dq;
      jc6 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = scr[29 * SST];
      jc2 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc9 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[35 * SST];
      jc8 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 108), a1 = ld2(rk1 + 108);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc2 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc16 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 112), a1 = ld2(rk1 + 112);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c14;
      jc14 += 1.0 * dq;
      jc16 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc9 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[51 * SST] * c15;
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc16 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 172), a1 = ld2(rk1 + 172);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c15;
      jc9 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc16 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 174), a1 = ld2(rk1 + 174);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c19;
      jc2 += 1.0 * dq;
      jc16 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 176), a1 = ld2(rk1 + 176);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c14;
      jc14 += 0.0 * dq;
      jc16 += -1.0 * dq;
      jc29 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 178), a1 = ld2(rk1 + 178);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc2 += 1.0 * dq;
      jc16 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 180), a1 = ld2(rk1 + 180);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c3;
      jc3 += -1.0 * dq;
      jc16 += -1.0 * dq;
      jc18 += 2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 182), a1 = ld2(rk1 + 182);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c9;
      jc9 += 0.0 * dq;
      jc16 += -1.0 * dq;
      jc29 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 184), a1 = ld2(rk1 + 184);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c30;
      jc16 += -1.0 * dq;
      jc29 += 1.0 * dq;
      jc30 += 0.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 186), a1 = ld2(rk1 + 186);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c30;
      jc9 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc16 += -1.0 * dq;
      jc30 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 188), a1 = ld2(rk1 + 188);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c11;
      jc11 += -1.0 * dq;
      jc16 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = scr[41 * SST];
      jc11 += 1.0 * dq;
      jc18 += -2.0 * dq;
    }
    {
      const double dq = scr[45 * SST];
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[46 * SST];
      jc4 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    {
      const double dq = scr[49 * SST];
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc20 += -1.0 * dq;
    }
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    double acc = 0.0;
    sink(0, 16, 0.0);
    sink(1, 16, 0.0);
    { const double v = 1.008 * gj * jc2; sink(2, 16, v); acc += e2 * v; }
    { const double v = 16.043 * gj * jc3; sink(3, 16, v); acc += e3 * v; }
    { const double v = 26.038 * gj * jc4; sink(4, 16, v); acc += e4 * v; }
    sink(5, 16, 0.0);
    { const double v = 27.046 * gj * jc6; sink(6, 16, v); acc += e6 * v; }
    { const double v = 15.999 * gj * jc7; sink(7, 16, v); acc += e7 * v; }
    { const double v = 34.014 * gj * jc8; sink(8, 16, v); acc += e8 * v; }
    { const double v = 28.01 * gj * jc9; sink(9, 16, v); acc += e9 * v; }
    { const double v = 30.026 * gj * jc10; sink(10, 16, v); acc += e10 * v; }
    { const double v = 30.07 * gj * jc11; sink(11, 16, v); acc += e11 * v; }
    { const double v = 42.037 * gj * jc12; sink(12, 16, v); acc += e12 * v; }
    sink(13, 16, 0.0);
    { const double v = 18.015 * gj * jc14; sink(14, 16, v); acc += e14 * v; }
    { const double v = 31.998 * gj * jc15; sink(15, 16, v); acc += e15 * v; }
    { const double v = 14.027 * gj * jc16; sink(16, 16, v); acc += e16 * v; }
    { const double v = 29.018 * gj * jc17; sink(17, 16, v); acc += e17 * v; }
    { const double v = 15.035 * gj * jc18; sink(18, 16, v); acc += e18 * v; }
    { const double v = 2.016 * gj * jc19; sink(19, 16, v); acc += e19 * v; }
    { const double v = 43.045 * gj * jc20; sink(20, 16, v); acc += e20 * v; }
    sink(21, 16, 0.0);
    sink(22, 16, 0.0);
    sink(23, 16, 0.0);
    { const double v = 33.006 * gj * jc24; sink(24, 16, v); acc += e24 * v; }
    { const double v = 28.054 * gj * jc25; sink(25, 16, v); acc += e25 * v; }
    { const double v = 13.019 * gj * jc26; sink(26, 16, v); acc += e26 * v; }
    { const double v = 17.007 * gj * jc27; sink(27, 16, v); acc += e27 * v; }
    { const double v = 29.062 * gj * jc28; sink(28, 16, v); acc += e28 * v; }
    { const double v = 14.027 * gj * jc29; sink(29, 16, v); acc += e29 * v; }
    { const double v = 44.009 * gj * jc30; sink(30, 16, v); acc += e30 * v; }
    const double cvj = (ld2(rh0 + 32).x + (ld2(rh1 + 32).x - ld2(rh0 + 32).x) * Td) - 592.7504099237186;
    sink(NS, 16, -(acc + Tdot * cvj * (((negmask >> 16) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 17: d/drho_17 (HCO)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 17) & 1u ? 0.0 : 0.03446136880556896;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc5 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc10 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc17 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc22 = 0.0;
    double jc24 = 0.0;
    double jc25 = 0.0;
    double jc26 = 0.0;
    double jc27 = 0.0;
    double jc28 = 0.0;
    double jc29 = 0.0;
    double jc30 = 0.0;
    {
      const double dq = scr[0 * SST];
      jc7 += -2.0 * dq;
      jc15 += 1.0 * dq;
    }
    {
      const double dq = scr[1 * SST];
      jc2 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = -scr[62 * SST] * c2;
      jc2 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = -scr[66 * SST] * c2;
      jc2 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc16 += -1.0 * dq;
      jc17 += 1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc7 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    {
      const double dq = scr[71 * SST] * c7;
      jc7 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[73 * SST] * c7;
      jc2 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc17 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    {
      const double dq = -scr[76 * SST] * c27;
      jc7 += -1.0 * dq;
      jc10 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc27 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 34), a1 = ld2(rk1 + 34);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c18;
      jc7 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 44), a1 = ld2(rk1 + 44);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c24;
      jc10 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc24 += 1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc2 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    {
      const double dq = scr[11 * SST];
      jc2 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc2 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[18 * SST] * c2 + scr[17 * SST];
      jc2 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 78), a1 = ld2(rk1 + 78);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c2;
      jc2 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
      jc19 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 80), a1 = ld2(rk1 + 80);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c19;
      jc2 += -1.0 * dq;
      jc10 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[20 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
    }
    {
      const double dq = scr[23 * SST];
      jc2 += -1.0 * dq;
      jc6 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = scr[29 * SST];
      jc2 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc9 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[35 * SST];
      jc8 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 102), a1 = ld2(rk1 + 102);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc17 += 1.0 * dq;
      jc26 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 118), a1 = ld2(rk1 + 118);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc9 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc17 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 120), a1 = ld2(rk1 + 120);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c14;
      jc10 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc17 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 142), a1 = ld2(rk1 + 142);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c8;
      jc8 += 1.0 * dq;
      jc10 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc24 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 144), a1 = ld2(rk1 + 144);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c7;
      jc7 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc26 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 156), a1 = ld2(rk1 + 156);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c9;
      jc9 += 1.0 * dq;
      jc17 += 1.0 * dq;
      jc26 += -1.0 * dq;
      jc30 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 160), a1 = ld2(rk1 + 160);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c27;
      jc15 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc27 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc9 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[41 * SST];
      jc11 += 1.0 * dq;
      jc18 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 196), a1 = ld2(rk1 + 196);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc3 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 198), a1 = ld2(rk1 + 198);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c3;
      jc3 += 1.0 * dq;
      jc10 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 204), a1 = ld2(rk1 + 204);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M5 * 1.0;
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 206), a1 = ld2(rk1 + 206);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M10 * 1.0 + scr[45 * SST];
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 208), a1 = ld2(rk1 + 208);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c15;
      jc9 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc17 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 210), a1 = ld2(rk1 + 210);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c10;
      jc6 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc17 += 1.0 * dq;
    }
    {
      const double dq = scr[46 * SST];
      jc4 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 226), a1 = ld2(rk1 + 226);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 2.0*c17;
      jc4 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc17 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 238), a1 = ld2(rk1 + 238);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 2.0*c17;
      jc7 += 1.0 * dq;
      jc12 += 1.0 * dq;
      jc17 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 242), a1 = ld2(rk1 + 242);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c10;
      jc10 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc17 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[49 * SST];
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc20 += -1.0 * dq;
    }
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 306), a1 = ld2(rk1 + 306);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc17 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc22 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 340), a1 = ld2(rk1 + 340);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c12;
      jc5 += 1.0 * dq;
      jc12 += -1.0 * dq;
      jc15 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    double acc = 0.0;
    sink(0, 17, 0.0);
    sink(1, 17, 0.0);
    { const double v = 1.008 * gj * jc2; sink(2, 17, v); acc += e2 * v; }
    { const double v = 16.043 * gj * jc3; sink(3, 17, v); acc += e3 * v; }
    { const double v = 26.038 * gj * jc4; sink(4, 17, v); acc += e4 * v; }
    { const double v = 39.057 * gj * jc5; sink(5, 17, v); acc += e5 * v; }
    { const double v = 27.046 * gj * jc6; sink(6, 17, v); acc += e6 * v; }
    { const double v = 15.999 * gj * jc7; sink(7, 17, v); acc += e7 * v; }
    { const double v = 34.014 * gj * jc8; sink(8, 17, v); acc += e8 * v; }
    { const double v = 28.01 * gj * jc9; sink(9, 17, v); acc += e9 * v; }
    { const double v = 30.026 * gj * jc10; sink(10, 17, v); acc += e10 * v; }
    { const double v = 30.07 * gj * jc11; sink(11, 17, v); acc += e11 * v; }
    { const double v = 42.037 * gj * jc12; sink(12, 17, v); acc += e12 * v; }
    sink(13, 17, 0.0);
    { const double v = 18.015 * gj * jc14; sink(14, 17, v); acc += e14 * v; }
    { const double v = 31.998 * gj * jc15; sink(15, 17, v); acc += e15 * v; }
    { const double v = 14.027 * gj * jc16; sink(16, 17, v); acc += e16 * v; }
    { const double v = 29.018 * gj * jc17; sink(17, 17, v); acc += e17 * v; }
    { const double v = 15.035 * gj * jc18; sink(18, 17, v); acc += e18 * v; }
    { const double v = 2.016 * gj * jc19; sink(19, 17, v); acc += e19 * v; }
    { const double v = 43.045 * gj * jc20; sink(20, 17, v); acc += e20 * v; }
    sink(21, 17, 0.0);
    { const double v = 44.053 * gj * jc22; sink(22, 17, v); acc += e22 * v; }
    sink(23, 17, 0.0);
    { const double v = 33.006 * gj * jc24; sink(24, 17, v); acc += e24 * v; }
    { const double v = 28.054 * gj * jc25; sink(25, 17, v); acc += e25 * v; }
    { const double v = 13.019 * gj * jc26; sink(26, 17, v); acc += e26 * v; }
    { const double v = 17.007 * gj * jc27; sink(27, 17, v); acc += e27 * v; }
    { const double v = 29.062 * gj * jc28; sink(28, 17, v); acc += e28 * v; }
    { const double v = 14.027 * gj * jc29; sink(29, 17, v); acc += e29 * v; }
    { const double v = 44.009 * gj * jc30; sink(30, 17, v); acc += e30 * v; }
    const double cvj = (ld2(rh0 + 34).x + (ld2(rh1 + 34).x - ld2(rh0 + 34).x) * Td) - 286.52939554759115;
    sink(NS, 17, -(acc + Tdot * cvj * (((negmask >> 17) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 18: d/drho_18 (CH3)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 18) & 1u ? 0.0 : 0.06651147322913202;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc10 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc13 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc17 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc21 = 0.0;
    double jc22 = 0.0;
    double jc23 = 0.0;
    double jc24 = 0.0;
    double jc25 = 0.0;
    double jc26 = 0.0;
    double jc27 = 0.0;
    double jc28 = 0.0;
    double jc29 = 0.0;
    double jc30 = 0.0;
    {
      const double dq = scr[0 * SST];
      jc7 += -2.0 * dq;
      jc15 += 1.0 * dq;
    }
    {
      const double dq = scr[1 * SST];
      jc2 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[67 * SST] * c7;
      jc2 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = -scr[70 * SST] * c27;
      jc3 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc7 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 34), a1 = ld2(rk1 + 34);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c17;
      jc7 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 36), a1 = ld2(rk1 + 36);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c10;
      jc7 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc2 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    {
      const double dq = -scr[13 * SST] * 1.0 + scr[11 * SST];
      jc2 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[15 * SST] * c2 + scr[14 * SST];
      jc2 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 76), a1 = ld2(rk1 + 76);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c19;
      jc2 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[17 * SST];
      jc2 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[20 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
    }
    {
      const double dq = scr[23 * SST];
      jc2 += -1.0 * dq;
      jc6 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = scr[29 * SST];
      jc2 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 90), a1 = ld2(rk1 + 90);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c9;
      jc2 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc12 += -1.0 * dq;
      jc18 += 1.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc9 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[35 * SST];
      jc8 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 110), a1 = ld2(rk1 + 110);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc14 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc27 += -1.0 * dq;
      jc29 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 112), a1 = ld2(rk1 + 112);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc14 += 1.0 * dq;
      jc16 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 114), a1 = ld2(rk1 + 114);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c14;
      jc3 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 124), a1 = ld2(rk1 + 124);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c9;
      jc4 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 138), a1 = ld2(rk1 + 138);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c24;
      jc3 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc24 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 152), a1 = ld2(rk1 + 152);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c26;
      jc2 += 1.0 * dq;
      jc6 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc26 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 162), a1 = ld2(rk1 + 162);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc19 += -1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 166), a1 = ld2(rk1 + 166);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c29;
      jc2 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc25 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 168), a1 = ld2(rk1 + 168);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * 2.0*c18;
      jc3 += -1.0 * dq;
      jc18 += 2.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc9 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 174), a1 = ld2(rk1 + 174);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc16 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 178), a1 = ld2(rk1 + 178);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c16;
      jc2 += 1.0 * dq;
      jc16 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 180), a1 = ld2(rk1 + 180);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * 2.0*c18;
      jc3 += -1.0 * dq;
      jc16 += -1.0 * dq;
      jc18 += 2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 188), a1 = ld2(rk1 + 188);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c28;
      jc11 += -1.0 * dq;
      jc16 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc28 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 190), a1 = ld2(rk1 + 190);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c15;
      jc10 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 192), a1 = ld2(rk1 + 192);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c8;
      jc3 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    {
      const double dq = scr[42 * SST] * 2.0*c18 + scr[41 * SST];
      jc11 += 1.0 * dq;
      jc18 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 194), a1 = ld2(rk1 + 194);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 2.0*c18;
      jc2 += 1.0 * dq;
      jc18 += -2.0 * dq;
      jc28 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 196), a1 = ld2(rk1 + 196);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c17;
      jc3 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 198), a1 = ld2(rk1 + 198);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c10;
      jc3 += 1.0 * dq;
      jc10 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 200), a1 = ld2(rk1 + 200);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c25;
      jc3 += 1.0 * dq;
      jc6 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 202), a1 = ld2(rk1 + 202);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c11;
      jc3 += 1.0 * dq;
      jc11 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = scr[45 * SST];
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[46 * SST];
      jc4 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 214), a1 = ld2(rk1 + 214);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc10 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 256), a1 = ld2(rk1 + 256);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c6;
      jc3 += 1.0 * dq;
      jc4 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 260), a1 = ld2(rk1 + 260);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c28;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc25 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    {
      const double dq = scr[49 * SST];
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc20 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 266), a1 = ld2(rk1 + 266);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M11 * c9 + scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 278), a1 = ld2(rk1 + 278);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c25;
      jc13 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 284), a1 = ld2(rk1 + 284);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c20;
      jc3 += 1.0 * dq;
      jc12 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 298), a1 = ld2(rk1 + 298);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c28;
      jc11 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc28 += -1.0 * dq;
      jc29 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 302), a1 = ld2(rk1 + 302);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c10;
      jc10 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc25 += 1.0 * dq;
      jc27 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 306), a1 = ld2(rk1 + 306);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c17;
      jc17 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc22 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 320), a1 = ld2(rk1 + 320);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c22;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 336), a1 = ld2(rk1 + 336);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c21;
      jc3 += 1.0 * dq;
      jc6 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc21 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 360), a1 = ld2(rk1 + 360);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c30;
      jc12 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc27 += 1.0 * dq;
      jc30 += -1.0 * dq;
    }
    double acc = 0.0;
    sink(0, 18, 0.0);
    sink(1, 18, 0.0);
    { const double v = 1.008 * gj * jc2; sink(2, 18, v); acc += e2 * v; }
    { const double v = 16.043 * gj * jc3; sink(3, 18, v); acc += e3 * v; }
    { const double v = 26.038 * gj * jc4; sink(4, 18, v); acc += e4 * v; }
    sink(5, 18, 0.0);
    { const double v = 27.046 * gj * jc6; sink(6, 18, v); acc += e6 * v; }
    { const double v = 15.999 * gj * jc7; sink(7, 18, v); acc += e7 * v; }
    { const double v = 34.014 * gj * jc8; sink(8, 18, v); acc += e8 * v; }
    { const double v = 28.01 * gj * jc9; sink(9, 18, v); acc += e9 * v; }
    { const double v = 30.026 * gj * jc10; sink(10, 18, v); acc += e10 * v; }
    { const double v = 30.07 * gj * jc11; sink(11, 18, v); acc += e11 * v; }
    { const double v = 42.037 * gj * jc12; sink(12, 18, v); acc += e12 * v; }
    { const double v = 43.089 * gj * jc13; sink(13, 18, v); acc += e13 * v; }
    { const double v = 18.015 * gj * jc14; sink(14, 18, v); acc += e14 * v; }
    { const double v = 31.998 * gj * jc15; sink(15, 18, v); acc += e15 * v; }
    { const double v = 14.027 * gj * jc16; sink(16, 18, v); acc += e16 * v; }
    { const double v = 29.018 * gj * jc17; sink(17, 18, v); acc += e17 * v; }
    { const double v = 15.035 * gj * jc18; sink(18, 18, v); acc += e18 * v; }
    { const double v = 2.016 * gj * jc19; sink(19, 18, v); acc += e19 * v; }
    { const double v = 43.045 * gj * jc20; sink(20, 18, v); acc += e20 * v; }
    { const double v = 58.08 * gj * jc21; sink(21, 18, v); acc += e21 * v; }
    { const double v = 44.053 * gj * jc22; sink(22, 18, v); acc += e22 * v; }
    { const double v = 43.045 * gj * jc23; sink(23, 18, v); acc += e23 * v; }
    { const double v = 33.006 * gj * jc24; sink(24, 18, v); acc += e24 * v; }
    { const double v = 28.054 * gj * jc25; sink(25, 18, v); acc += e25 * v; }
    { const double v = 13.019 * gj * jc26; sink(26, 18, v); acc += e26 * v; }
    { const double v = 17.007 * gj * jc27; sink(27, 18, v); acc += e27 * v; }
    { const double v = 29.062 * gj * jc28; sink(28, 18, v); acc += e28 * v; }
    { const double v = 14.027 * gj * jc29; sink(29, 18, v); acc += e29 * v; }
    { const double v = 44.009 * gj * jc30; sink(30, 18, v); acc += e30 * v; }
    const double cvj = (ld2(rh0 + 36).x + (ld2(rh1 + 36).x - ld2(rh0 + 36).x) * Td) - 553.0103092783505;
    sink(NS, 18, -(acc + Tdot * cvj * (((negmask >> 18) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 19: d/drho_19 (H2)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 19) & 1u ? 0.0 : 0.49603174603174605;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc10 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc17 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc22 = 0.0;
    double jc23 = 0.0;
    double jc24 = 0.0;
    double jc25 = 0.0;
    double jc26 = 0.0;
    double jc27 = 0.0;
    double jc28 = 0.0;
    double jc29 = 0.0;
    double jc30 = 0.0;
    {
      const double dq = 2.4000000953674316 * scr[0 * SST];
      jc7 += -2.0 * dq;
      jc15 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[1 * SST];
      jc2 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[53 * SST] * c7;
      jc2 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc19 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = -scr[64 * SST] * c9;
      jc7 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc16 += -1.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[2 * SST];
      jc7 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 54), a1 = ld2(rk1 + 54);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M6 * 1.0;
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 56), a1 = ld2(rk1 + 56);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M7 * 1.0 + scr[7 * SST];
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 58), a1 = ld2(rk1 + 58);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M5 * 1.0;
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 60), a1 = ld2(rk1 + 60);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M8 * 1.0;
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = 0.7300000190734863 * scr[10 * SST];
      jc2 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 66), a1 = ld2(rk1 + 66);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c15;
      jc2 += -1.0 * dq;
      jc15 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc24 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 70), a1 = ld2(rk1 + 70);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c24;
      jc2 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc24 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[11 * SST];
      jc2 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 74), a1 = ld2(rk1 + 74);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c26;
      jc2 += -1.0 * dq;
      jc16 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc26 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[14 * SST];
      jc2 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 76), a1 = ld2(rk1 + 76);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c18;
      jc2 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[17 * SST];
      jc2 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 78), a1 = ld2(rk1 + 78);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c9;
      jc2 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
      jc19 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 80), a1 = ld2(rk1 + 80);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c17;
      jc2 += -1.0 * dq;
      jc10 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[20 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[23 * SST];
      jc2 += -1.0 * dq;
      jc6 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 82), a1 = ld2(rk1 + 82);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c4;
      jc2 += -1.0 * dq;
      jc4 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[26 * SST];
      jc2 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 84), a1 = ld2(rk1 + 84);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c6;
      jc2 += -1.0 * dq;
      jc6 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[29 * SST];
      jc2 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 86), a1 = ld2(rk1 + 86);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c25;
      jc2 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 88), a1 = ld2(rk1 + 88);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c28;
      jc2 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = scr[33 * SST] * c9 + 2.0 * scr[32 * SST];
      jc9 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 92), a1 = ld2(rk1 + 92);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc2 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc19 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[35 * SST];
      jc8 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 146), a1 = ld2(rk1 + 146);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c26;
      jc2 += 1.0 * dq;
      jc19 += -1.0 * dq;
      jc26 += -1.0 * dq;
      jc29 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 162), a1 = ld2(rk1 + 162);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c29;
      jc2 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc19 += -1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 164), a1 = ld2(rk1 + 164);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c4;
      jc4 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc29 += -2.0 * dq;
    }
    {
      const double dq = 2.0 * scr[38 * SST];
      jc9 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 174), a1 = ld2(rk1 + 174);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c16;
      jc2 += 1.0 * dq;
      jc16 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[41 * SST];
      jc11 += 1.0 * dq;
      jc18 += -2.0 * dq;
    }
    {
      const double dq = 2.0 * scr[45 * SST];
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = -scr[48 * SST] * c4 + 2.0 * scr[46 * SST];
      jc4 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 214), a1 = ld2(rk1 + 214);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c10;
      jc10 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    {
      const double dq = scr[49 * SST];
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc20 += -1.0 * dq;
    }
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 314), a1 = ld2(rk1 + 314);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c23;
      jc2 += 1.0 * dq;
      jc19 += -1.0 * dq;
      jc22 += 1.0 * dq;
      jc23 += -1.0 * dq;
    }
    double acc = 0.0;
    sink(0, 19, 0.0);
    sink(1, 19, 0.0);
    { const double v = 1.008 * gj * jc2; sink(2, 19, v); acc += e2 * v; }
    { const double v = 16.043 * gj * jc3; sink(3, 19, v); acc += e3 * v; }
    { const double v = 26.038 * gj * jc4; sink(4, 19, v); acc += e4 * v; }
    sink(5, 19, 0.0);
    { const double v = 27.046 * gj * jc6; sink(6, 19, v); acc += e6 * v; }
    { const double v = 15.999 * gj * jc7; sink(7, 19, v); acc += e7 * v; }
    { const double v = 34.014 * gj * jc8; sink(8, 19, v); acc += e8 * v; }
    { const double v = 28.01 * gj * jc9; sink(9, 19, v); acc += e9 * v; }
    { const double v = 30.026 * gj * jc10; sink(10, 19, v); acc += e10 * v; }
    { const double v = 30.07 * gj * jc11; sink(11, 19, v); acc += e11 * v; }
    { const double v = 42.037 * gj * jc12; sink(12, 19, v); acc += e12 * v; }
    sink(13, 19, 0.0);
    { const double v = 18.015 * gj * jc14; sink(14, 19, v); acc += e14 * v; }
    { const double v = 31.998 * gj * jc15; sink(15, 19, v); acc += e15 * v; }
    { const double v = 14.027 * gj * jc16; sink(16, 19, v); acc += e16 * v; }
    { const double v = 29.018 * gj * jc17; sink(17, 19, v); acc += e17 * v; }
    { const double v = 15.035 * gj * jc18; sink(18, 19, v); acc += e18 * v; }
    { const double v = 2.016 * gj * jc19; sink(19, 19, v); acc += e19 * v; }
    { const double v = 43.045 * gj * jc20; sink(20, 19, v); acc += e20 * v; }
    sink(21, 19, 0.0);
    { const double v = 44.053 * gj * jc22; sink(22, 19, v); acc += e22 * v; }
    { const double v = 43.045 * gj * jc23; sink(23, 19, v); acc += e23 * v; }
    { const double v = 33.006 * gj * jc24; sink(24, 19, v); acc += e24 * v; }
    { const double v = 28.054 * gj * jc25; sink(25, 19, v); acc += e25 * v; }
    { const double v = 13.019 * gj * jc26; sink(26, 19, v); acc += e26 * v; }
    { const double v = 17.007 * gj * jc27; sink(27, 19, v); acc += e27 * v; }
    { const double v = 29.062 * gj * jc28; sink(28, 19, v); acc += e28 * v; }
    { const double v = 14.027 * gj * jc29; sink(29, 19, v); acc += e29 * v; }
    { const double v = 44.009 * gj * jc30; sink(30, 19, v); acc += e30 * v; }
    const double cvj = (ld2(rh0 + 38).x + (ld2(rh1 + 38).x - ld2(rh0 + 38).x) * Td) - 4124.2609126984125;
    sink(NS, 19, -(acc + Tdot * cvj * (((negmask >> 19) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 20: d/drho_20 (CH3CO)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 20) & 1u ? 0.0 : 0.023231501916598906;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc10 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc17 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc23 = 0.0;
    double jc24 = 0.0;
    double jc25 = 0.0;
    double jc27 = 0.0;
    double jc28 = 0.0;
    double jc29 = 0.0;
    double jc30 = 0.0;
    {
      const double dq = scr[0 * SST];
      jc7 += -2.0 * dq;
      jc15 += 1.0 * dq;
    }
    {
      const double dq = scr[1 * SST];
      jc2 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc7 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc2 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    {
      const double dq = scr[11 * SST];
      jc2 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc2 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[17 * SST];
      jc2 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[20 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
    }
    {
      const double dq = scr[23 * SST];
      jc2 += -1.0 * dq;
      jc6 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = scr[29 * SST];
      jc2 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc9 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[35 * SST];
      jc8 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc9 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[41 * SST];
      jc11 += 1.0 * dq;
      jc18 += -2.0 * dq;
    }
    {
      const double dq = scr[45 * SST];
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[46 * SST];
      jc4 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 264), a1 = ld2(rk1 + 264);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M11 * 1.0 + scr[49 * SST];
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc20 += -1.0 * dq;
    }
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 280), a1 = ld2(rk1 + 280);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c7;
      jc7 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc20 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 284), a1 = ld2(rk1 + 284);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc3 += 1.0 * dq;
      jc12 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 344), a1 = ld2(rk1 + 344);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 1.0;
      jc20 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    double acc = 0.0;
    sink(0, 20, 0.0);
    sink(1, 20, 0.0);
    { const double v = 1.008 * gj * jc2; sink(2, 20, v); acc += e2 * v; }
    { const double v = 16.043 * gj * jc3; sink(3, 20, v); acc += e3 * v; }
    { const double v = 26.038 * gj * jc4; sink(4, 20, v); acc += e4 * v; }
    sink(5, 20, 0.0);
    { const double v = 27.046 * gj * jc6; sink(6, 20, v); acc += e6 * v; }
    { const double v = 15.999 * gj * jc7; sink(7, 20, v); acc += e7 * v; }
    { const double v = 34.014 * gj * jc8; sink(8, 20, v); acc += e8 * v; }
    { const double v = 28.01 * gj * jc9; sink(9, 20, v); acc += e9 * v; }
    { const double v = 30.026 * gj * jc10; sink(10, 20, v); acc += e10 * v; }
    { const double v = 30.07 * gj * jc11; sink(11, 20, v); acc += e11 * v; }
    { const double v = 42.037 * gj * jc12; sink(12, 20, v); acc += e12 * v; }
    sink(13, 20, 0.0);
    { const double v = 18.015 * gj * jc14; sink(14, 20, v); acc += e14 * v; }
    { const double v = 31.998 * gj * jc15; sink(15, 20, v); acc += e15 * v; }
    sink(16, 20, 0.0);
    { const double v = 29.018 * gj * jc17; sink(17, 20, v); acc += e17 * v; }
    { const double v = 15.035 * gj * jc18; sink(18, 20, v); acc += e18 * v; }
    { const double v = 2.016 * gj * jc19; sink(19, 20, v); acc += e19 * v; }
    { const double v = 43.045 * gj * jc20; sink(20, 20, v); acc += e20 * v; }
    sink(21, 20, 0.0);
    sink(22, 20, 0.0);
    { const double v = 43.045 * gj * jc23; sink(23, 20, v); acc += e23 * v; }
    { const double v = 33.006 * gj * jc24; sink(24, 20, v); acc += e24 * v; }
    { const double v = 28.054 * gj * jc25; sink(25, 20, v); acc += e25 * v; }
    sink(26, 20, 0.0);
    { const double v = 17.007 * gj * jc27; sink(27, 20, v); acc += e27 * v; }
    { const double v = 29.062 * gj * jc28; sink(28, 20, v); acc += e28 * v; }
    { const double v = 14.027 * gj * jc29; sink(29, 20, v); acc += e29 * v; }
    { const double v = 44.009 * gj * jc30; sink(30, 20, v); acc += e30 * v; }
    const double cvj = (ld2(rh0 + 40).x + (ld2(rh1 + 40).x - ld2(rh0 + 40).x) * Td) - 193.15855500058078;
    sink(NS, 20, -(acc + Tdot * cvj * (((negmask >> 20) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 21: d/drho_21 (C3H6O1-3)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 21) & 1u ? 0.0 : 0.01721763085399449;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc10 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc17 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc21 = 0.0;
    double jc24 = 0.0;
    double jc25 = 0.0;
    double jc27 = 0.0;
    double jc28 = 0.0;
    double jc29 = 0.0;
    double jc30 = 0.0;
    {
      const double dq = scr[0 * SST];
      jc7 += -2.0 * dq;
      jc15 += 1.0 * dq;
    }
    {
      const double dq = scr[1 * SST];
      jc2 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc7 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc2 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    {
      const double dq = scr[11 * SST];
      jc2 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc2 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[17 * SST];
      jc2 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[20 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
    }
    {
      const double dq = scr[23 * SST];
      jc2 += -1.0 * dq;
      jc6 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = scr[29 * SST];
      jc2 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc9 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[35 * SST];
      jc8 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc9 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[41 * SST];
      jc11 += 1.0 * dq;
      jc18 += -2.0 * dq;
    }
    {
      const double dq = scr[45 * SST];
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[46 * SST];
      jc4 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    {
      const double dq = scr[49 * SST];
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc20 += -1.0 * dq;
    }
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 324), a1 = ld2(rk1 + 324);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 1.0;
      jc10 += 1.0 * dq;
      jc21 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 328), a1 = ld2(rk1 + 328);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc6 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc21 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 330), a1 = ld2(rk1 + 330);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c7;
      jc6 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc21 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 332), a1 = ld2(rk1 + 332);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c2;
      jc2 += -1.0 * dq;
      jc6 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc21 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 334), a1 = ld2(rk1 + 334);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c24;
      jc6 += 1.0 * dq;
      jc8 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc21 += -1.0 * dq;
      jc24 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 336), a1 = ld2(rk1 + 336);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc3 += 1.0 * dq;
      jc6 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc21 += -1.0 * dq;
    }
    double acc = 0.0;
    sink(0, 21, 0.0);
    sink(1, 21, 0.0);
    { const double v = 1.008 * gj * jc2; sink(2, 21, v); acc += e2 * v; }
    { const double v = 16.043 * gj * jc3; sink(3, 21, v); acc += e3 * v; }
    { const double v = 26.038 * gj * jc4; sink(4, 21, v); acc += e4 * v; }
    sink(5, 21, 0.0);
    { const double v = 27.046 * gj * jc6; sink(6, 21, v); acc += e6 * v; }
    { const double v = 15.999 * gj * jc7; sink(7, 21, v); acc += e7 * v; }
    { const double v = 34.014 * gj * jc8; sink(8, 21, v); acc += e8 * v; }
    { const double v = 28.01 * gj * jc9; sink(9, 21, v); acc += e9 * v; }
    { const double v = 30.026 * gj * jc10; sink(10, 21, v); acc += e10 * v; }
    { const double v = 30.07 * gj * jc11; sink(11, 21, v); acc += e11 * v; }
    { const double v = 42.037 * gj * jc12; sink(12, 21, v); acc += e12 * v; }
    sink(13, 21, 0.0);
    { const double v = 18.015 * gj * jc14; sink(14, 21, v); acc += e14 * v; }
    { const double v = 31.998 * gj * jc15; sink(15, 21, v); acc += e15 * v; }
    sink(16, 21, 0.0);
    { const double v = 29.018 * gj * jc17; sink(17, 21, v); acc += e17 * v; }
    { const double v = 15.035 * gj * jc18; sink(18, 21, v); acc += e18 * v; }
    { const double v = 2.016 * gj * jc19; sink(19, 21, v); acc += e19 * v; }
    { const double v = 43.045 * gj * jc20; sink(20, 21, v); acc += e20 * v; }
    { const double v = 58.08 * gj * jc21; sink(21, 21, v); acc += e21 * v; }
    sink(22, 21, 0.0);
    sink(23, 21, 0.0);
    { const double v = 33.006 * gj * jc24; sink(24, 21, v); acc += e24 * v; }
    { const double v = 28.054 * gj * jc25; sink(25, 21, v); acc += e25 * v; }
    sink(26, 21, 0.0);
    { const double v = 17.007 * gj * jc27; sink(27, 21, v); acc += e27 * v; }
    { const double v = 29.062 * gj * jc28; sink(28, 21, v); acc += e28 * v; }
    { const double v = 14.027 * gj * jc29; sink(29, 21, v); acc += e29 * v; }
    { const double v = 44.009 * gj * jc30; sink(30, 21, v); acc += e30 * v; }
    const double cvj = (ld2(rh0 + 42).x + (ld2(rh1 + 42).x - ld2(rh0 + 42).x) * Td) - 143.15616391184574;
    sink(NS, 21, -(acc + Tdot * cvj * (((negmask >> 21) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 22: d/drho_22 (C2H4O1-2)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 22) & 1u ? 0.0 : 0.02269992963021815;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc10 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc17 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc22 = 0.0;
    double jc23 = 0.0;
    double jc24 = 0.0;
    double jc25 = 0.0;
    double jc27 = 0.0;
    double jc28 = 0.0;
    double jc29 = 0.0;
    double jc30 = 0.0;
    {
      const double dq = scr[0 * SST];
      jc7 += -2.0 * dq;
      jc15 += 1.0 * dq;
    }
    {
      const double dq = scr[1 * SST];
      jc2 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc7 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc2 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    {
      const double dq = scr[11 * SST];
      jc2 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc2 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[17 * SST];
      jc2 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[20 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
    }
    {
      const double dq = scr[23 * SST];
      jc2 += -1.0 * dq;
      jc6 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = scr[29 * SST];
      jc2 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc9 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[35 * SST];
      jc8 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc9 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[41 * SST];
      jc11 += 1.0 * dq;
      jc18 += -2.0 * dq;
    }
    {
      const double dq = scr[45 * SST];
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[46 * SST];
      jc4 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    {
      const double dq = scr[49 * SST];
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc20 += -1.0 * dq;
    }
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 304), a1 = ld2(rk1 + 304);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 1.0;
      jc17 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc22 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 308), a1 = ld2(rk1 + 308);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc14 += 1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 312), a1 = ld2(rk1 + 312);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c2;
      jc2 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 316), a1 = ld2(rk1 + 316);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c24;
      jc8 += 1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
      jc24 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 320), a1 = ld2(rk1 + 320);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 368), a1 = ld2(rk1 + 368);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc22 += -1.0 * dq;
      jc24 += 1.0 * dq;
      jc25 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    double acc = 0.0;
    sink(0, 22, 0.0);
    sink(1, 22, 0.0);
    { const double v = 1.008 * gj * jc2; sink(2, 22, v); acc += e2 * v; }
    { const double v = 16.043 * gj * jc3; sink(3, 22, v); acc += e3 * v; }
    { const double v = 26.038 * gj * jc4; sink(4, 22, v); acc += e4 * v; }
    sink(5, 22, 0.0);
    { const double v = 27.046 * gj * jc6; sink(6, 22, v); acc += e6 * v; }
    { const double v = 15.999 * gj * jc7; sink(7, 22, v); acc += e7 * v; }
    { const double v = 34.014 * gj * jc8; sink(8, 22, v); acc += e8 * v; }
    { const double v = 28.01 * gj * jc9; sink(9, 22, v); acc += e9 * v; }
    { const double v = 30.026 * gj * jc10; sink(10, 22, v); acc += e10 * v; }
    { const double v = 30.07 * gj * jc11; sink(11, 22, v); acc += e11 * v; }
    { const double v = 42.037 * gj * jc12; sink(12, 22, v); acc += e12 * v; }
    sink(13, 22, 0.0);
    { const double v = 18.015 * gj * jc14; sink(14, 22, v); acc += e14 * v; }
    { const double v = 31.998 * gj * jc15; sink(15, 22, v); acc += e15 * v; }
    sink(16, 22, 0.0);
    { const double v = 29.018 * gj * jc17; sink(17, 22, v); acc += e17 * v; }
    { const double v = 15.035 * gj * jc18; sink(18, 22, v); acc += e18 * v; }
    { const double v = 2.016 * gj * jc19; sink(19, 22, v); acc += e19 * v; }
    { const double v = 43.045 * gj * jc20; sink(20, 22, v); acc += e20 * v; }
    sink(21, 22, 0.0);
    { const double v = 44.053 * gj * jc22; sink(22, 22, v); acc += e22 * v; }
    { const double v = 43.045 * gj * jc23; sink(23, 22, v); acc += e23 * v; }
    { const double v = 33.006 * gj * jc24; sink(24, 22, v); acc += e24 * v; }
    { const double v = 28.054 * gj * jc25; sink(25, 22, v); acc += e25 * v; }
    sink(26, 22, 0.0);
    { const double v = 17.007 * gj * jc27; sink(27, 22, v); acc += e27 * v; }
    { const double v = 29.062 * gj * jc28; sink(28, 22, v); acc += e28 * v; }
    { const double v = 14.027 * gj * jc29; sink(29, 22, v); acc += e29 * v; }
    { const double v = 44.009 * gj * jc30; sink(30, 22, v); acc += e30 * v; }
    const double cvj = (ld2(rh0 + 44).x + (ld2(rh1 + 44).x - ld2(rh0 + 44).x) * Td) - 188.7387919097451;
    sink(NS, 22, -(acc + Tdot * cvj * (((negmask >> 22) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 23: d/drho_23 (C2H3O1-2)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 23) & 1u ? 0.0 : 0.023231501916598906;
    double jc1 = 0.0;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc10 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc17 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc22 = 0.0;
    double jc23 = 0.0;
    double jc24 = 0.0;
    double jc25 = 0.0;
    double jc27 = 0.0;
    double jc28 = 0.0;
    double jc29 = 0.0;
    double jc30 = 0.0;
    {
      const double dq = scr[0 * SST];
      jc7 += -2.0 * dq;
      jc15 += 1.0 * dq;
    }
    {
      const double dq = scr[1 * SST];
      jc2 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc7 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc2 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    {
      const double dq = scr[11 * SST];
      jc2 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc2 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[17 * SST];
      jc2 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[20 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
    }
    {
      const double dq = scr[23 * SST];
      jc2 += -1.0 * dq;
      jc6 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = scr[29 * SST];
      jc2 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc9 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[35 * SST];
      jc8 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc9 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[41 * SST];
      jc11 += 1.0 * dq;
      jc18 += -2.0 * dq;
    }
    {
      const double dq = scr[45 * SST];
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[46 * SST];
      jc4 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    {
      const double dq = scr[49 * SST];
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc20 += -1.0 * dq;
    }
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 310), a1 = ld2(rk1 + 310);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c14;
      jc14 += -1.0 * dq;
      jc22 += 1.0 * dq;
      jc23 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 314), a1 = ld2(rk1 + 314);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c19;
      jc2 += 1.0 * dq;
      jc19 += -1.0 * dq;
      jc22 += 1.0 * dq;
      jc23 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 318), a1 = ld2(rk1 + 318);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c8;
      jc8 += -1.0 * dq;
      jc22 += 1.0 * dq;
      jc23 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 322), a1 = ld2(rk1 + 322);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c3;
      jc3 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc22 += 1.0 * dq;
      jc23 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 342), a1 = ld2(rk1 + 342);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 1.0;
      jc20 += 1.0 * dq;
      jc23 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 346), a1 = ld2(rk1 + 346);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 1.0;
      jc1 += 1.0 * dq;
      jc23 += -1.0 * dq;
    }
    double acc = 0.0;
    sink(0, 23, 0.0);
    { const double v = 43.045 * gj * jc1; sink(1, 23, v); acc += e1 * v; }
    { const double v = 1.008 * gj * jc2; sink(2, 23, v); acc += e2 * v; }
    { const double v = 16.043 * gj * jc3; sink(3, 23, v); acc += e3 * v; }
    { const double v = 26.038 * gj * jc4; sink(4, 23, v); acc += e4 * v; }
    sink(5, 23, 0.0);
    { const double v = 27.046 * gj * jc6; sink(6, 23, v); acc += e6 * v; }
    { const double v = 15.999 * gj * jc7; sink(7, 23, v); acc += e7 * v; }
    { const double v = 34.014 * gj * jc8; sink(8, 23, v); acc += e8 * v; }
    { const double v = 28.01 * gj * jc9; sink(9, 23, v); acc += e9 * v; }
    { const double v = 30.026 * gj * jc10; sink(10, 23, v); acc += e10 * v; }
    { const double v = 30.07 * gj * jc11; sink(11, 23, v); acc += e11 * v; }
    { const double v = 42.037 * gj * jc12; sink(12, 23, v); acc += e12 * v; }
    sink(13, 23, 0.0);
    { const double v = 18.015 * gj * jc14; sink(14, 23, v); acc += e14 * v; }
    { const double v = 31.998 * gj * jc15; sink(15, 23, v); acc += e15 * v; }
    sink(16, 23, 0.0);
    { const double v = 29.018 * gj * jc17; sink(17, 23, v); acc += e17 * v; }
    { const double v = 15.035 * gj * jc18; sink(18, 23, v); acc += e18 * v; }
    { const double v = 2.016 * gj * jc19; sink(19, 23, v); acc += e19 * v; }
    { const double v = 43.045 * gj * jc20; sink(20, 23, v); acc += e20 * v; }
    sink(21, 23, 0.0);
    { const double v = 44.053 * gj * jc22; sink(22, 23, v); acc += e22 * v; }
    { const double v = 43.045 * gj * jc23; sink(23, 23, v); acc += e23 * v; }
    { const double v = 33.006 * gj * jc24; sink(24, 23, v); acc += e24 * v; }
    { const double v = 28.054 * gj * jc25; sink(25, 23, v); acc += e25 * v; }
    sink(26, 23, 0.0);
    { const double v = 17.007 * gj * jc27; sink(27, 23, v); acc += e27 * v; }
    { const double v = 29.062 * gj * jc28; sink(28, 23, v); acc += e28 * v; }
    { const double v = 14.027 * gj * jc29; sink(29, 23, v); acc += e29 * v; }
    { const double v = 44.009 * gj * jc30; sink(30, 23, v); acc += e30 * v; }
    const double cvj = (ld2(rh0 + 46).x + (ld2(rh1 + 46).x - ld2(rh0 + 46).x) * Td) - 193.15855500058078;
    sink(NS, 23, -(acc + Tdot * cvj * (((negmask >> 23) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 24: d/drho_24 (HO2)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 24) & 1u ? 0.0 : 0.030297521662727988;
    double jc1 = 0.0;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc10 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc17 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc21 = 0.0;
    double jc22 = 0.0;
    double jc23 = 0.0;
    double jc24 = 0.0;
    double jc25 = 0.0;
    double jc27 = 0.0;
    double jc28 = 0.0;
    double jc29 = 0.0;
    double jc30 = 0.0;
    {
      const double dq = scr[0 * SST];
      jc7 += -2.0 * dq;
      jc15 += 1.0 * dq;
    }
    {
      const double dq = scr[1 * SST];
      jc2 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[55 * SST] * c7;
      jc7 += -1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = -scr[58 * SST] * c27;
      jc7 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc24 += 1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc7 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 44), a1 = ld2(rk1 + 44);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c17;
      jc10 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc24 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 46), a1 = ld2(rk1 + 46);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M3 * 1.0 + scr[3 * SST];
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 48), a1 = ld2(rk1 + 48);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M4 * 1.0;
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 50), a1 = ld2(rk1 + 50);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M5 * 1.0;
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc2 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 64), a1 = ld2(rk1 + 64);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c2;
      jc2 += -1.0 * dq;
      jc7 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc24 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 66), a1 = ld2(rk1 + 66);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c2;
      jc2 += -1.0 * dq;
      jc15 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc24 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 68), a1 = ld2(rk1 + 68);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c2;
      jc2 += -1.0 * dq;
      jc24 += -1.0 * dq;
      jc27 += 2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 70), a1 = ld2(rk1 + 70);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c19;
      jc2 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc24 += 1.0 * dq;
    }
    {
      const double dq = scr[11 * SST];
      jc2 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc2 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[17 * SST];
      jc2 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[20 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
    }
    {
      const double dq = scr[23 * SST];
      jc2 += -1.0 * dq;
      jc6 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = scr[29 * SST];
      jc2 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc9 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[35 * SST];
      jc8 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 96), a1 = ld2(rk1 + 96);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc14 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 98), a1 = ld2(rk1 + 98);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c14;
      jc8 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc24 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 100), a1 = ld2(rk1 + 100);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c14;
      jc8 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc24 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 132), a1 = ld2(rk1 + 132);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 2.0*c24;
      jc8 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 134), a1 = ld2(rk1 + 134);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 2.0*c24;
      jc8 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 136), a1 = ld2(rk1 + 136);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c29;
      jc10 += 1.0 * dq;
      jc24 += -1.0 * dq;
      jc27 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 138), a1 = ld2(rk1 + 138);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc3 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc24 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 140), a1 = ld2(rk1 + 140);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c9;
      jc9 += -1.0 * dq;
      jc24 += -1.0 * dq;
      jc27 += 1.0 * dq;
      jc30 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 142), a1 = ld2(rk1 + 142);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c10;
      jc8 += 1.0 * dq;
      jc10 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc24 += -1.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc9 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 192), a1 = ld2(rk1 + 192);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c3;
      jc3 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    {
      const double dq = scr[41 * SST];
      jc11 += 1.0 * dq;
      jc18 += -2.0 * dq;
    }
    {
      const double dq = scr[45 * SST];
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 208), a1 = ld2(rk1 + 208);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c9;
      jc9 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc17 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    {
      const double dq = scr[46 * SST];
      jc4 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 212), a1 = ld2(rk1 + 212);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c25;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
      jc25 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 234), a1 = ld2(rk1 + 234);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c4;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 246), a1 = ld2(rk1 + 246);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c28;
      jc11 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -1.0 * dq;
      jc28 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 248), a1 = ld2(rk1 + 248);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c11;
      jc8 += 1.0 * dq;
      jc11 += -1.0 * dq;
      jc24 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = scr[49 * SST];
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc20 += -1.0 * dq;
    }
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 270), a1 = ld2(rk1 + 270);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c6;
      jc6 += -1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 272), a1 = ld2(rk1 + 272);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c25;
      jc6 += 1.0 * dq;
      jc8 += 1.0 * dq;
      jc24 += -1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 316), a1 = ld2(rk1 + 316);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c22;
      jc8 += 1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
      jc24 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 334), a1 = ld2(rk1 + 334);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c21;
      jc6 += 1.0 * dq;
      jc8 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc21 += -1.0 * dq;
      jc24 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 356), a1 = ld2(rk1 + 356);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c12;
      jc1 += 1.0 * dq;
      jc12 += -1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 366), a1 = ld2(rk1 + 366);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c25;
      jc22 += 1.0 * dq;
      jc24 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    double acc = 0.0;
    sink(0, 24, 0.0);
    { const double v = 43.045 * gj * jc1; sink(1, 24, v); acc += e1 * v; }
    { const double v = 1.008 * gj * jc2; sink(2, 24, v); acc += e2 * v; }
    { const double v = 16.043 * gj * jc3; sink(3, 24, v); acc += e3 * v; }
    { const double v = 26.038 * gj * jc4; sink(4, 24, v); acc += e4 * v; }
    sink(5, 24, 0.0);
    { const double v = 27.046 * gj * jc6; sink(6, 24, v); acc += e6 * v; }
    { const double v = 15.999 * gj * jc7; sink(7, 24, v); acc += e7 * v; }
    { const double v = 34.014 * gj * jc8; sink(8, 24, v); acc += e8 * v; }
    { const double v = 28.01 * gj * jc9; sink(9, 24, v); acc += e9 * v; }
    { const double v = 30.026 * gj * jc10; sink(10, 24, v); acc += e10 * v; }
    { const double v = 30.07 * gj * jc11; sink(11, 24, v); acc += e11 * v; }
    { const double v = 42.037 * gj * jc12; sink(12, 24, v); acc += e12 * v; }
    sink(13, 24, 0.0);
    { const double v = 18.015 * gj * jc14; sink(14, 24, v); acc += e14 * v; }
    { const double v = 31.998 * gj * jc15; sink(15, 24, v); acc += e15 * v; }
    sink(16, 24, 0.0);
    { const double v = 29.018 * gj * jc17; sink(17, 24, v); acc += e17 * v; }
    { const double v = 15.035 * gj * jc18; sink(18, 24, v); acc += e18 * v; }
    { const double v = 2.016 * gj * jc19; sink(19, 24, v); acc += e19 * v; }
    { const double v = 43.045 * gj * jc20; sink(20, 24, v); acc += e20 * v; }
    { const double v = 58.08 * gj * jc21; sink(21, 24, v); acc += e21 * v; }
    { const double v = 44.053 * gj * jc22; sink(22, 24, v); acc += e22 * v; }
    { const double v = 43.045 * gj * jc23; sink(23, 24, v); acc += e23 * v; }
    { const double v = 33.006 * gj * jc24; sink(24, 24, v); acc += e24 * v; }
    { const double v = 28.054 * gj * jc25; sink(25, 24, v); acc += e25 * v; }
    sink(26, 24, 0.0);
    { const double v = 17.007 * gj * jc27; sink(27, 24, v); acc += e27 * v; }
    { const double v = 29.062 * gj * jc28; sink(28, 24, v); acc += e28 * v; }
    { const double v = 14.027 * gj * jc29; sink(29, 24, v); acc += e29 * v; }
    { const double v = 44.009 * gj * jc30; sink(30, 24, v); acc += e30 * v; }
    const double cvj = (ld2(rh0 + 48).x + (ld2(rh1 + 48).x - ld2(rh0 + 48).x) * Td) - 251.90904683996848;
    sink(NS, 24, -(acc + Tdot * cvj * (((negmask >> 24) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 25: d/drho_25 (C2H4)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 25) & 1u ? 0.0 : 0.035645540742853074;
    double jc1 = 0.0;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc10 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc13 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc17 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc21 = 0.0;
    double jc22 = 0.0;
    double jc24 = 0.0;
    double jc25 = 0.0;
    double jc26 = 0.0;
    double jc27 = 0.0;
    double jc28 = 0.0;
    double jc29 = 0.0;
    double jc30 = 0.0;
    {
      const double dq = scr[0 * SST];
      jc7 += -2.0 * dq;
      jc15 += 1.0 * dq;
    }
    {
      const double dq = scr[1 * SST];
      jc2 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc7 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 34), a1 = ld2(rk1 + 34);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c7;
      jc7 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc2 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    {
      const double dq = scr[11 * SST];
      jc2 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc2 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[17 * SST];
      jc2 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[20 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
    }
    {
      const double dq = -scr[25 * SST] * 1.0 + scr[23 * SST];
      jc2 += -1.0 * dq;
      jc6 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    {
      const double dq = scr[27 * SST] * c2 + scr[26 * SST];
      jc2 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 84), a1 = ld2(rk1 + 84);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c2;
      jc2 += -1.0 * dq;
      jc6 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    {
      const double dq = scr[29 * SST];
      jc2 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 86), a1 = ld2(rk1 + 86);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c19;
      jc2 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc9 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[35 * SST];
      jc8 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 128), a1 = ld2(rk1 + 128);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc6 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc25 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 154), a1 = ld2(rk1 + 154);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc3 += -1.0 * dq;
      jc25 += 1.0 * dq;
      jc26 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 166), a1 = ld2(rk1 + 166);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc25 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc9 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 178), a1 = ld2(rk1 + 178);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc16 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    {
      const double dq = scr[41 * SST];
      jc11 += 1.0 * dq;
      jc18 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 200), a1 = ld2(rk1 + 200);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc3 += 1.0 * dq;
      jc6 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc25 += -1.0 * dq;
    }
    {
      const double dq = scr[45 * SST];
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[47 * SST] * 1.0 + scr[46 * SST];
      jc4 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 212), a1 = ld2(rk1 + 212);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c24;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
      jc25 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 218), a1 = ld2(rk1 + 218);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 2.0*c25;
      jc6 += 1.0 * dq;
      jc25 += -2.0 * dq;
      jc28 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 220), a1 = ld2(rk1 + 220);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c7;
      jc6 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 252), a1 = ld2(rk1 + 252);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c11;
      jc11 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 262), a1 = ld2(rk1 + 262);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c3;
      jc3 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = scr[49 * SST];
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc20 += -1.0 * dq;
    }
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 268), a1 = ld2(rk1 + 268);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c15;
      jc6 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 272), a1 = ld2(rk1 + 272);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c24;
      jc6 += 1.0 * dq;
      jc8 += 1.0 * dq;
      jc24 += -1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 278), a1 = ld2(rk1 + 278);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc13 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 288), a1 = ld2(rk1 + 288);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c7;
      jc1 += 1.0 * dq;
      jc2 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 300), a1 = ld2(rk1 + 300);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc10 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc25 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 326), a1 = ld2(rk1 + 326);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c10;
      jc10 += -1.0 * dq;
      jc21 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 366), a1 = ld2(rk1 + 366);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c24;
      jc22 += 1.0 * dq;
      jc24 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    double acc = 0.0;
    sink(0, 25, 0.0);
    { const double v = 43.045 * gj * jc1; sink(1, 25, v); acc += e1 * v; }
    { const double v = 1.008 * gj * jc2; sink(2, 25, v); acc += e2 * v; }
    { const double v = 16.043 * gj * jc3; sink(3, 25, v); acc += e3 * v; }
    { const double v = 26.038 * gj * jc4; sink(4, 25, v); acc += e4 * v; }
    sink(5, 25, 0.0);
    { const double v = 27.046 * gj * jc6; sink(6, 25, v); acc += e6 * v; }
    { const double v = 15.999 * gj * jc7; sink(7, 25, v); acc += e7 * v; }
    { const double v = 34.014 * gj * jc8; sink(8, 25, v); acc += e8 * v; }
    { const double v = 28.01 * gj * jc9; sink(9, 25, v); acc += e9 * v; }
    { const double v = 30.026 * gj * jc10; sink(10, 25, v); acc += e10 * v; }
    { const double v = 30.07 * gj * jc11; sink(11, 25, v); acc += e11 * v; }
    { const double v = 42.037 * gj * jc12; sink(12, 25, v); acc += e12 * v; }
    { const double v = 43.089 * gj * jc13; sink(13, 25, v); acc += e13 * v; }
    { const double v = 18.015 * gj * jc14; sink(14, 25, v); acc += e14 * v; }
    { const double v = 31.998 * gj * jc15; sink(15, 25, v); acc += e15 * v; }
    { const double v = 14.027 * gj * jc16; sink(16, 25, v); acc += e16 * v; }
    { const double v = 29.018 * gj * jc17; sink(17, 25, v); acc += e17 * v; }
    { const double v = 15.035 * gj * jc18; sink(18, 25, v); acc += e18 * v; }
    { const double v = 2.016 * gj * jc19; sink(19, 25, v); acc += e19 * v; }
    { const double v = 43.045 * gj * jc20; sink(20, 25, v); acc += e20 * v; }
    { const double v = 58.08 * gj * jc21; sink(21, 25, v); acc += e21 * v; }
    { const double v = 44.053 * gj * jc22; sink(22, 25, v); acc += e22 * v; }
    sink(23, 25, 0.0);
    { const double v = 33.006 * gj * jc24; sink(24, 25, v); acc += e24 * v; }
    { const double v = 28.054 * gj * jc25; sink(25, 25, v); acc += e25 * v; }
    { const double v = 13.019 * gj * jc26; sink(26, 25, v); acc += e26 * v; }
    { const double v = 17.007 * gj * jc27; sink(27, 25, v); acc += e27 * v; }
    { const double v = 29.062 * gj * jc28; sink(28, 25, v); acc += e28 * v; }
    { const double v = 14.027 * gj * jc29; sink(29, 25, v); acc += e29 * v; }
    { const double v = 44.009 * gj * jc30; sink(30, 25, v); acc += e30 * v; }
    const double cvj = (ld2(rh0 + 50).x + (ld2(rh1 + 50).x - ld2(rh0 + 50).x) * Td) - 296.3752049618593;
    sink(NS, 25, -(acc + Tdot * cvj * (((negmask >> 25) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 26: d/drho_26 (CH)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 26) & 1u ? 0.0 : 0.07681081496274675;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc10 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc17 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc24 = 0.0;
    double jc25 = 0.0;
    double jc26 = 0.0;
    double jc27 = 0.0;
    double jc28 = 0.0;
    double jc29 = 0.0;
    double jc30 = 0.0;
    {
      const double dq = scr[0 * SST];
      jc7 += -2.0 * dq;
      jc15 += 1.0 * dq;
    }
    {
      const double dq = scr[1 * SST];
      jc2 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[59 * SST] * c7;
      jc2 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc26 += -1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc7 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc2 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    {
      const double dq = scr[11 * SST];
      jc2 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 74), a1 = ld2(rk1 + 74);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c19;
      jc2 += -1.0 * dq;
      jc16 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc26 += 1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc2 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[17 * SST];
      jc2 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[20 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
    }
    {
      const double dq = scr[23 * SST];
      jc2 += -1.0 * dq;
      jc6 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = scr[29 * SST];
      jc2 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc9 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[35 * SST];
      jc8 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 102), a1 = ld2(rk1 + 102);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc2 += 1.0 * dq;
      jc17 += 1.0 * dq;
      jc26 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 106), a1 = ld2(rk1 + 106);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c14;
      jc14 += 1.0 * dq;
      jc26 += 1.0 * dq;
      jc27 += -1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 144), a1 = ld2(rk1 + 144);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c15;
      jc7 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc26 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 146), a1 = ld2(rk1 + 146);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c19;
      jc2 += 1.0 * dq;
      jc19 += -1.0 * dq;
      jc26 += -1.0 * dq;
      jc29 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 148), a1 = ld2(rk1 + 148);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c14;
      jc2 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc14 += -1.0 * dq;
      jc26 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 150), a1 = ld2(rk1 + 150);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c29;
      jc2 += 1.0 * dq;
      jc4 += 1.0 * dq;
      jc26 += -1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 152), a1 = ld2(rk1 + 152);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc2 += 1.0 * dq;
      jc6 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc26 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 154), a1 = ld2(rk1 + 154);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c3;
      jc2 += 1.0 * dq;
      jc3 += -1.0 * dq;
      jc25 += 1.0 * dq;
      jc26 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 156), a1 = ld2(rk1 + 156);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c30;
      jc9 += 1.0 * dq;
      jc17 += 1.0 * dq;
      jc26 += -1.0 * dq;
      jc30 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 158), a1 = ld2(rk1 + 158);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c10;
      jc2 += 1.0 * dq;
      jc10 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc26 += -1.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc9 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[41 * SST];
      jc11 += 1.0 * dq;
      jc18 += -2.0 * dq;
    }
    {
      const double dq = scr[45 * SST];
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[46 * SST];
      jc4 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 228), a1 = ld2(rk1 + 228);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c15;
      jc9 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc26 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[49 * SST];
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc20 += -1.0 * dq;
    }
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 292), a1 = ld2(rk1 + 292);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c11;
      jc11 += -1.0 * dq;
      jc26 += -1.0 * dq;
      jc28 += 1.0 * dq;
      jc29 += 1.0 * dq;
    }
    double acc = 0.0;
    sink(0, 26, 0.0);
    sink(1, 26, 0.0);
    { const double v = 1.008 * gj * jc2; sink(2, 26, v); acc += e2 * v; }
    { const double v = 16.043 * gj * jc3; sink(3, 26, v); acc += e3 * v; }
    { const double v = 26.038 * gj * jc4; sink(4, 26, v); acc += e4 * v; }
    sink(5, 26, 0.0);
    { const double v = 27.046 * gj * jc6; sink(6, 26, v); acc += e6 * v; }
    { const double v = 15.999 * gj * jc7; sink(7, 26, v); acc += e7 * v; }
    { const double v = 34.014 * gj * jc8; sink(8, 26, v); acc += e8 * v; }
    { const double v = 28.01 * gj * jc9; sink(9, 26, v); acc += e9 * v; }
    { const double v = 30.026 * gj * jc10; sink(10, 26, v); acc += e10 * v; }
    { const double v = 30.07 * gj * jc11; sink(11, 26, v); acc += e11 * v; }
    { const double v = 42.037 * gj * jc12; sink(12, 26, v); acc += e12 * v; }
    sink(13, 26, 0.0);
    { const double v = 18.015 * gj * jc14; sink(14, 26, v); acc += e14 * v; }
    { const double v = 31.998 * gj * jc15; sink(15, 26, v); acc += e15 * v; }
    { const double v = 14.027 * gj * jc16; sink(16, 26, v); acc += e16 * v; }
    { const double v = 29.018 * gj * jc17; sink(17, 26, v); acc += e17 * v; }
    { const double v = 15.035 * gj * jc18; sink(18, 26, v); acc += e18 * v; }
    { const double v = 2.016 * gj * jc19; sink(19, 26, v); acc += e19 * v; }
    { const double v = 43.045 * gj * jc20; sink(20, 26, v); acc += e20 * v; }
    sink(21, 26, 0.0);
    sink(22, 26, 0.0);
    sink(23, 26, 0.0);
    { const double v = 33.006 * gj * jc24; sink(24, 26, v); acc += e24 * v; }
    { const double v = 28.054 * gj * jc25; sink(25, 26, v); acc += e25 * v; }
    { const double v = 13.019 * gj * jc26; sink(26, 26, v); acc += e26 * v; }
    { const double v = 17.007 * gj * jc27; sink(27, 26, v); acc += e27 * v; }
    { const double v = 29.062 * gj * jc28; sink(28, 26, v); acc += e28 * v; }
    { const double v = 14.027 * gj * jc29; sink(29, 26, v); acc += e29 * v; }
    { const double v = 44.009 * gj * jc30; sink(30, 26, v); acc += e30 * v; }
    const double cvj = (ld2(rh0 + 52).x + (ld2(rh1 + 52).x - ld2(rh0 + 52).x) * Td) - 638.6442891159076;
    sink(NS, 26, -(acc + Tdot * cvj * (((negmask >> 26) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 27: d/drho_27 (OH)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 27) & 1u ? 0.0 : 0.05879931792791203;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc10 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc17 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc21 = 0.0;
    double jc22 = 0.0;
    double jc23 = 0.0;
    double jc24 = 0.0;
    double jc25 = 0.0;
    double jc26 = 0.0;
    double jc27 = 0.0;
    double jc28 = 0.0;
    double jc29 = 0.0;
    double jc30 = 0.0;
    {
      const double dq = scr[0 * SST];
      jc7 += -2.0 * dq;
      jc15 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 2), a1 = ld2(rk1 + 2);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M1 * 1.0 + scr[1 * SST];
      jc2 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = -scr[54 * SST] * c2;
      jc2 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc19 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = -scr[56 * SST] * c15;
      jc7 += -1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = -scr[58 * SST] * c24;
      jc7 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc24 += 1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = -scr[70 * SST] * c18;
      jc3 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc7 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    {
      const double dq = -scr[72 * SST] * c9;
      jc7 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = -scr[76 * SST] * c17;
      jc7 += -1.0 * dq;
      jc10 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc27 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 38), a1 = ld2(rk1 + 38);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c28;
      jc7 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc27 += 1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 52), a1 = ld2(rk1 + 52);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c7;
      jc2 += -1.0 * dq;
      jc7 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 62), a1 = ld2(rk1 + 62);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M9 * c2 + scr[10 * SST];
      jc2 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 68), a1 = ld2(rk1 + 68);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * 2.0*c27;
      jc2 += -1.0 * dq;
      jc24 += -1.0 * dq;
      jc27 += 2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 72), a1 = ld2(rk1 + 72);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c14;
      jc2 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[11 * SST];
      jc2 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc2 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[17 * SST];
      jc2 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[20 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
    }
    {
      const double dq = scr[23 * SST];
      jc2 += -1.0 * dq;
      jc6 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = scr[29 * SST];
      jc2 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc9 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 92), a1 = ld2(rk1 + 92);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c19;
      jc2 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc19 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    {
      const double dq = scr[36 * SST] * 2.0*c27 + scr[35 * SST];
      jc8 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 94), a1 = ld2(rk1 + 94);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 2.0*c27;
      jc7 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 96), a1 = ld2(rk1 + 96);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c24;
      jc14 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 98), a1 = ld2(rk1 + 98);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c8;
      jc8 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc24 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 100), a1 = ld2(rk1 + 100);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c8;
      jc8 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc24 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 102), a1 = ld2(rk1 + 102);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c26;
      jc2 += 1.0 * dq;
      jc17 += 1.0 * dq;
      jc26 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 104), a1 = ld2(rk1 + 104);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c29;
      jc2 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc27 += -1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 106), a1 = ld2(rk1 + 106);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c29;
      jc14 += 1.0 * dq;
      jc26 += 1.0 * dq;
      jc27 += -1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 108), a1 = ld2(rk1 + 108);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c16;
      jc2 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc16 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 110), a1 = ld2(rk1 + 110);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc14 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc27 += -1.0 * dq;
      jc29 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 112), a1 = ld2(rk1 + 112);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc14 += 1.0 * dq;
      jc16 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 114), a1 = ld2(rk1 + 114);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c3;
      jc3 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 116), a1 = ld2(rk1 + 116);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c9;
      jc2 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc27 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 118), a1 = ld2(rk1 + 118);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c17;
      jc9 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc17 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 120), a1 = ld2(rk1 + 120);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c10;
      jc10 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc17 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 122), a1 = ld2(rk1 + 122);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c4;
      jc2 += 1.0 * dq;
      jc4 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 124), a1 = ld2(rk1 + 124);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c4;
      jc4 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 126), a1 = ld2(rk1 + 126);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c6;
      jc4 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 128), a1 = ld2(rk1 + 128);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c25;
      jc6 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc25 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 130), a1 = ld2(rk1 + 130);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c11;
      jc11 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 136), a1 = ld2(rk1 + 136);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c10;
      jc10 += 1.0 * dq;
      jc24 += -1.0 * dq;
      jc27 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 140), a1 = ld2(rk1 + 140);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c30;
      jc9 += -1.0 * dq;
      jc24 += -1.0 * dq;
      jc27 += 1.0 * dq;
      jc30 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 160), a1 = ld2(rk1 + 160);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c17;
      jc15 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc27 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc9 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = -scr[52 * SST] * c9*c2;
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc16 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 190), a1 = ld2(rk1 + 190);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c10;
      jc10 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[41 * SST];
      jc11 += 1.0 * dq;
      jc18 += -2.0 * dq;
    }
    {
      const double dq = scr[45 * SST];
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[46 * SST];
      jc4 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 214), a1 = ld2(rk1 + 214);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc10 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 222), a1 = ld2(rk1 + 222);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c6;
      jc6 += -1.0 * dq;
      jc7 += 1.0 * dq;
      jc25 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 230), a1 = ld2(rk1 + 230);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c9;
      jc9 += -1.0 * dq;
      jc15 += 1.0 * dq;
      jc26 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 240), a1 = ld2(rk1 + 240);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c12;
      jc10 += 1.0 * dq;
      jc12 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    {
      const double dq = scr[49 * SST];
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc20 += -1.0 * dq;
    }
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 282), a1 = ld2(rk1 + 282);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c12;
      jc7 += 1.0 * dq;
      jc12 += -1.0 * dq;
      jc20 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 300), a1 = ld2(rk1 + 300);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c25;
      jc10 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc25 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 308), a1 = ld2(rk1 + 308);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c22;
      jc14 += 1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 328), a1 = ld2(rk1 + 328);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c21;
      jc6 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc21 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 358), a1 = ld2(rk1 + 358);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c12;
      jc12 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc27 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 368), a1 = ld2(rk1 + 368);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c22;
      jc22 += -1.0 * dq;
      jc24 += 1.0 * dq;
      jc25 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    double acc = 0.0;
    sink(0, 27, 0.0);
    sink(1, 27, 0.0);
    { const double v = 1.008 * gj * jc2; sink(2, 27, v); acc += e2 * v; }
    { const double v = 16.043 * gj * jc3; sink(3, 27, v); acc += e3 * v; }
    { const double v = 26.038 * gj * jc4; sink(4, 27, v); acc += e4 * v; }
    sink(5, 27, 0.0);
    { const double v = 27.046 * gj * jc6; sink(6, 27, v); acc += e6 * v; }
    { const double v = 15.999 * gj * jc7; sink(7, 27, v); acc += e7 * v; }
    { const double v = 34.014 * gj * jc8; sink(8, 27, v); acc += e8 * v; }
    { const double v = 28.01 * gj * jc9; sink(9, 27, v); acc += e9 * v; }
    { const double v = 30.026 * gj * jc10; sink(10, 27, v); acc += e10 * v; }
    { const double v = 30.07 * gj * jc11; sink(11, 27, v); acc += e11 * v; }
    { const double v = 42.037 * gj * jc12; sink(12, 27, v); acc += e12 * v; }
    sink(13, 27, 0.0);
    { const double v = 18.015 * gj * jc14; sink(14, 27, v); acc += e14 * v; }
    { const double v = 31.998 * gj * jc15; sink(15, 27, v); acc += e15 * v; }
    { const double v = 14.027 * gj * jc16; sink(16, 27, v); acc += e16 * v; }
    { const double v = 29.018 * gj * jc17; sink(17, 27, v); acc += e17 * v; }
    { const double v = 15.035 * gj * jc18; sink(18, 27, v); acc += e18 * v; }
    { const double v = 2.016 * gj * jc19; sink(19, 27, v); acc += e19 * v; }
    { const double v = 43.045 * gj * jc20; sink(20, 27, v); acc += e20 * v; }
    { const double v = 58.08 * gj * jc21; sink(21, 27, v); acc += e21 * v; }
    { const double v = 44.053 * gj * jc22; sink(22, 27, v); acc += e22 * v; }
    { const double v = 43.045 * gj * jc23; sink(23, 27, v); acc += e23 * v; }
    { const double v = 33.006 * gj * jc24; sink(24, 27, v); acc += e24 * v; }
    { const double v = 28.054 * gj * jc25; sink(25, 27, v); acc += e25 * v; }
    { const double v = 13.019 * gj * jc26; sink(26, 27, v); acc += e26 * v; }
    { const double v = 17.007 * gj * jc27; sink(27, 27, v); acc += e27 * v; }
    { const double v = 29.062 * gj * jc28; sink(28, 27, v); acc += e28 * v; }
    { const double v = 14.027 * gj * jc29; sink(29, 27, v); acc += e29 * v; }
    { const double v = 44.009 * gj * jc30; sink(30, 27, v); acc += e30 * v; }
    const double cvj = (ld2(rh0 + 54).x + (ld2(rh1 + 54).x - ld2(rh0 + 54).x) * Td) - 488.8875169048039;
    sink(NS, 27, -(acc + Tdot * cvj * (((negmask >> 27) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 28: d/drho_28 (C2H5)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 28) & 1u ? 0.0 : 0.034409194136673316;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc10 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc17 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc24 = 0.0;
    double jc25 = 0.0;
    double jc26 = 0.0;
    double jc27 = 0.0;
    double jc28 = 0.0;
    double jc29 = 0.0;
    double jc30 = 0.0;
    {
      const double dq = scr[0 * SST];
      jc7 += -2.0 * dq;
      jc15 += 1.0 * dq;
    }
    {
      const double dq = scr[1 * SST];
      jc2 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc7 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 36), a1 = ld2(rk1 + 36);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c7;
      jc7 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 38), a1 = ld2(rk1 + 38);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c27;
      jc7 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc27 += 1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc2 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    {
      const double dq = scr[11 * SST];
      jc2 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc2 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[17 * SST];
      jc2 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[20 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
    }
    {
      const double dq = scr[23 * SST];
      jc2 += -1.0 * dq;
      jc6 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    {
      const double dq = -scr[28 * SST] * 1.0 + scr[26 * SST];
      jc2 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = scr[30 * SST] * c2 + scr[29 * SST];
      jc2 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 86), a1 = ld2(rk1 + 86);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c2;
      jc2 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 88), a1 = ld2(rk1 + 88);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c19;
      jc2 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc9 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[35 * SST];
      jc8 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 130), a1 = ld2(rk1 + 130);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c14;
      jc11 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc9 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 188), a1 = ld2(rk1 + 188);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c18;
      jc11 += -1.0 * dq;
      jc16 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = scr[41 * SST];
      jc11 += 1.0 * dq;
      jc18 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 194), a1 = ld2(rk1 + 194);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc18 += -2.0 * dq;
      jc28 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 202), a1 = ld2(rk1 + 202);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c3;
      jc3 += 1.0 * dq;
      jc11 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = scr[45 * SST];
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[46 * SST];
      jc4 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 212), a1 = ld2(rk1 + 212);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c15;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
      jc25 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 216), a1 = ld2(rk1 + 216);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c6;
      jc6 += -1.0 * dq;
      jc25 += 2.0 * dq;
      jc28 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 246), a1 = ld2(rk1 + 246);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c24;
      jc11 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -1.0 * dq;
      jc28 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 250), a1 = ld2(rk1 + 250);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c8;
      jc8 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc24 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 254), a1 = ld2(rk1 + 254);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 2.0*c28;
      jc11 += 1.0 * dq;
      jc25 += 1.0 * dq;
      jc28 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 260), a1 = ld2(rk1 + 260);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc25 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    {
      const double dq = scr[49 * SST];
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc20 += -1.0 * dq;
    }
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 294), a1 = ld2(rk1 + 294);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c29;
      jc11 += 1.0 * dq;
      jc26 += 1.0 * dq;
      jc28 += -1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 298), a1 = ld2(rk1 + 298);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc11 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc28 += -1.0 * dq;
      jc29 += 1.0 * dq;
    }
    double acc = 0.0;
    sink(0, 28, 0.0);
    sink(1, 28, 0.0);
    { const double v = 1.008 * gj * jc2; sink(2, 28, v); acc += e2 * v; }
    { const double v = 16.043 * gj * jc3; sink(3, 28, v); acc += e3 * v; }
    { const double v = 26.038 * gj * jc4; sink(4, 28, v); acc += e4 * v; }
    sink(5, 28, 0.0);
    { const double v = 27.046 * gj * jc6; sink(6, 28, v); acc += e6 * v; }
    { const double v = 15.999 * gj * jc7; sink(7, 28, v); acc += e7 * v; }
    { const double v = 34.014 * gj * jc8; sink(8, 28, v); acc += e8 * v; }
    { const double v = 28.01 * gj * jc9; sink(9, 28, v); acc += e9 * v; }
    { const double v = 30.026 * gj * jc10; sink(10, 28, v); acc += e10 * v; }
    { const double v = 30.07 * gj * jc11; sink(11, 28, v); acc += e11 * v; }
    { const double v = 42.037 * gj * jc12; sink(12, 28, v); acc += e12 * v; }
    sink(13, 28, 0.0);
    { const double v = 18.015 * gj * jc14; sink(14, 28, v); acc += e14 * v; }
    { const double v = 31.998 * gj * jc15; sink(15, 28, v); acc += e15 * v; }
    { const double v = 14.027 * gj * jc16; sink(16, 28, v); acc += e16 * v; }
    { const double v = 29.018 * gj * jc17; sink(17, 28, v); acc += e17 * v; }
    { const double v = 15.035 * gj * jc18; sink(18, 28, v); acc += e18 * v; }
    { const double v = 2.016 * gj * jc19; sink(19, 28, v); acc += e19 * v; }
    { const double v = 43.045 * gj * jc20; sink(20, 28, v); acc += e20 * v; }
    sink(21, 28, 0.0);
    sink(22, 28, 0.0);
    sink(23, 28, 0.0);
    { const double v = 33.006 * gj * jc24; sink(24, 28, v); acc += e24 * v; }
    { const double v = 28.054 * gj * jc25; sink(25, 28, v); acc += e25 * v; }
    { const double v = 13.019 * gj * jc26; sink(26, 28, v); acc += e26 * v; }
    { const double v = 17.007 * gj * jc27; sink(27, 28, v); acc += e27 * v; }
    { const double v = 29.062 * gj * jc28; sink(28, 28, v); acc += e28 * v; }
    { const double v = 14.027 * gj * jc29; sink(29, 28, v); acc += e29 * v; }
    { const double v = 44.009 * gj * jc30; sink(30, 28, v); acc += e30 * v; }
    const double cvj = (ld2(rh0 + 56).x + (ld2(rh1 + 56).x - ld2(rh0 + 56).x) * Td) - 286.0955887413117;
    sink(NS, 28, -(acc + Tdot * cvj * (((negmask >> 28) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 29: d/drho_29 (CH2)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 29) & 1u ? 0.0 : 0.07129108148570615;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc10 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc17 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc24 = 0.0;
    double jc25 = 0.0;
    double jc26 = 0.0;
    double jc27 = 0.0;
    double jc28 = 0.0;
    double jc29 = 0.0;
    double jc30 = 0.0;
    {
      const double dq = scr[0 * SST];
      jc7 += -2.0 * dq;
      jc15 += 1.0 * dq;
    }
    {
      const double dq = scr[1 * SST];
      jc2 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    {
      const double dq = scr[61 * SST] * c7;
      jc2 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc7 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    {
      const double dq = -scr[78 * SST] * c9;
      jc4 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc29 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 40), a1 = ld2(rk1 + 40);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c30;
      jc7 += -1.0 * dq;
      jc12 += -1.0 * dq;
      jc29 += 1.0 * dq;
      jc30 += 1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc2 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    {
      const double dq = scr[12 * SST] * c2 + scr[11 * SST];
      jc2 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc2 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[17 * SST];
      jc2 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[20 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
    }
    {
      const double dq = scr[23 * SST];
      jc2 += -1.0 * dq;
      jc6 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = scr[29 * SST];
      jc2 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc9 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[35 * SST];
      jc8 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 104), a1 = ld2(rk1 + 104);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc2 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc27 += -1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 106), a1 = ld2(rk1 + 106);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc14 += 1.0 * dq;
      jc26 += 1.0 * dq;
      jc27 += -1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 110), a1 = ld2(rk1 + 110);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c14;
      jc14 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc27 += -1.0 * dq;
      jc29 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 136), a1 = ld2(rk1 + 136);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c24;
      jc10 += 1.0 * dq;
      jc24 += -1.0 * dq;
      jc27 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 146), a1 = ld2(rk1 + 146);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc19 += -1.0 * dq;
      jc26 += -1.0 * dq;
      jc29 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 150), a1 = ld2(rk1 + 150);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c26;
      jc2 += 1.0 * dq;
      jc4 += 1.0 * dq;
      jc26 += -1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 160), a1 = ld2(rk1 + 160);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c15;
      jc15 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc27 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 162), a1 = ld2(rk1 + 162);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c19;
      jc2 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc19 += -1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 164), a1 = ld2(rk1 + 164);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 2.0*c29;
      jc4 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc29 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 166), a1 = ld2(rk1 + 166);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc2 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc25 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 168), a1 = ld2(rk1 + 168);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c3;
      jc3 += -1.0 * dq;
      jc18 += 2.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = scr[39 * SST] * c9 + scr[38 * SST];
      jc9 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 176), a1 = ld2(rk1 + 176);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c14;
      jc14 += 0.0 * dq;
      jc16 += -1.0 * dq;
      jc29 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 182), a1 = ld2(rk1 + 182);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c9;
      jc9 += 0.0 * dq;
      jc16 += -1.0 * dq;
      jc29 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 184), a1 = ld2(rk1 + 184);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c30;
      jc16 += -1.0 * dq;
      jc29 += 1.0 * dq;
      jc30 += 0.0 * dq;
    }
    {
      const double dq = scr[41 * SST];
      jc11 += 1.0 * dq;
      jc18 += -2.0 * dq;
    }
    {
      const double dq = scr[45 * SST];
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[46 * SST];
      jc4 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    {
      const double dq = scr[49 * SST];
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc20 += -1.0 * dq;
    }
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 294), a1 = ld2(rk1 + 294);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c28;
      jc11 += 1.0 * dq;
      jc26 += 1.0 * dq;
      jc28 += -1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 296), a1 = ld2(rk1 + 296);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c11;
      jc11 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc28 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    double acc = 0.0;
    sink(0, 29, 0.0);
    sink(1, 29, 0.0);
    { const double v = 1.008 * gj * jc2; sink(2, 29, v); acc += e2 * v; }
    { const double v = 16.043 * gj * jc3; sink(3, 29, v); acc += e3 * v; }
    { const double v = 26.038 * gj * jc4; sink(4, 29, v); acc += e4 * v; }
    sink(5, 29, 0.0);
    { const double v = 27.046 * gj * jc6; sink(6, 29, v); acc += e6 * v; }
    { const double v = 15.999 * gj * jc7; sink(7, 29, v); acc += e7 * v; }
    { const double v = 34.014 * gj * jc8; sink(8, 29, v); acc += e8 * v; }
    { const double v = 28.01 * gj * jc9; sink(9, 29, v); acc += e9 * v; }
    { const double v = 30.026 * gj * jc10; sink(10, 29, v); acc += e10 * v; }
    { const double v = 30.07 * gj * jc11; sink(11, 29, v); acc += e11 * v; }
    { const double v = 42.037 * gj * jc12; sink(12, 29, v); acc += e12 * v; }
    sink(13, 29, 0.0);
    { const double v = 18.015 * gj * jc14; sink(14, 29, v); acc += e14 * v; }
    { const double v = 31.998 * gj * jc15; sink(15, 29, v); acc += e15 * v; }
    { const double v = 14.027 * gj * jc16; sink(16, 29, v); acc += e16 * v; }
    { const double v = 29.018 * gj * jc17; sink(17, 29, v); acc += e17 * v; }
    { const double v = 15.035 * gj * jc18; sink(18, 29, v); acc += e18 * v; }
    { const double v = 2.016 * gj * jc19; sink(19, 29, v); acc += e19 * v; }
    { const double v = 43.045 * gj * jc20; sink(20, 29, v); acc += e20 * v; }
    sink(21, 29, 0.0);
    sink(22, 29, 0.0);
    sink(23, 29, 0.0);
    { const double v = 33.006 * gj * jc24; sink(24, 29, v); acc += e24 * v; }
    { const double v = 28.054 * gj * jc25; sink(25, 29, v); acc += e25 * v; }
    { const double v = 13.019 * gj * jc26; sink(26, 29, v); acc += e26 * v; }
    { const double v = 17.007 * gj * jc27; sink(27, 29, v); acc += e27 * v; }
    { const double v = 29.062 * gj * jc28; sink(28, 29, v); acc += e28 * v; }
    { const double v = 14.027 * gj * jc29; sink(29, 29, v); acc += e29 * v; }
    { const double v = 44.009 * gj * jc30; sink(30, 29, v); acc += e30 * v; }
    const double cvj = (ld2(rh0 + 58).x + (ld2(rh1 + 58).x - ld2(rh0 + 58).x) * Td) - 592.7504099237186;
    sink(NS, 29, -(acc + Tdot * cvj * (((negmask >> 29) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 30: d/drho_30 (CO2)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 30) & 1u ? 0.0 : 0.022722624917630486;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc10 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc17 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc24 = 0.0;
    double jc25 = 0.0;
    double jc26 = 0.0;
    double jc27 = 0.0;
    double jc28 = 0.0;
    double jc29 = 0.0;
    double jc30 = 0.0;
    {
      const double dq = 3.5999999046325684 * scr[0 * SST];
      jc7 += -2.0 * dq;
      jc15 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[1 * SST];
      jc2 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 22), a1 = ld2(rk1 + 22);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M2 * 1.0 + 3.5 * scr[2 * SST];
      jc7 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    {
      const double dq = -scr[74 * SST] * c2;
      jc2 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc17 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 40), a1 = ld2(rk1 + 40);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c29;
      jc7 += -1.0 * dq;
      jc12 += -1.0 * dq;
      jc29 += 1.0 * dq;
      jc30 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 42), a1 = ld2(rk1 + 42);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c7;
      jc7 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    {
      const double dq = 1.5 * scr[3 * SST];
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    {
      const double dq = scr[9 * SST];
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc2 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[11 * SST];
      jc2 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[14 * SST];
      jc2 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[17 * SST];
      jc2 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[20 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[23 * SST];
      jc2 += -1.0 * dq;
      jc6 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[26 * SST];
      jc2 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[29 * SST];
      jc2 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[32 * SST];
      jc9 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[35 * SST];
      jc8 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 116), a1 = ld2(rk1 + 116);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc27 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 140), a1 = ld2(rk1 + 140);
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c27;
      jc9 += -1.0 * dq;
      jc24 += -1.0 * dq;
      jc27 += 1.0 * dq;
      jc30 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 156), a1 = ld2(rk1 + 156);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c26;
      jc9 += 1.0 * dq;
      jc17 += 1.0 * dq;
      jc26 += -1.0 * dq;
      jc30 += -1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[38 * SST];
      jc9 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 184), a1 = ld2(rk1 + 184);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c16 - (a0.y + (a1.y - a0.y) * Td) * c29;
      jc16 += -1.0 * dq;
      jc29 += 1.0 * dq;
      jc30 += 0.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 186), a1 = ld2(rk1 + 186);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c16;
      jc9 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc16 += -1.0 * dq;
      jc30 += -1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[41 * SST];
      jc11 += 1.0 * dq;
      jc18 += -2.0 * dq;
    }
    {
      const double dq = 2.0 * scr[45 * SST];
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[46 * SST];
      jc4 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    {
      const double dq = scr[49 * SST];
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc20 += -1.0 * dq;
    }
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = ld2(rk0 + 360), a1 = ld2(rk1 + 360);
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc12 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc27 += 1.0 * dq;
      jc30 += -1.0 * dq;
    }
    double acc = 0.0;
    sink(0, 30, 0.0);
    sink(1, 30, 0.0);
    { const double v = 1.008 * gj * jc2; sink(2, 30, v); acc += e2 * v; }
    { const double v = 16.043 * gj * jc3; sink(3, 30, v); acc += e3 * v; }
    { const double v = 26.038 * gj * jc4; sink(4, 30, v); acc += e4 * v; }
    sink(5, 30, 0.0);
    { const double v = 27.046 * gj * jc6; sink(6, 30, v); acc += e6 * v; }
    { const double v = 15.999 * gj * jc7; sink(7, 30, v); acc += e7 * v; }
    { const double v = 34.014 * gj * jc8; sink(8, 30, v); acc += e8 * v; }
    { const double v = 28.01 * gj * jc9; sink(9, 30, v); acc += e9 * v; }
    { const double v = 30.026 * gj * jc10; sink(10, 30, v); acc += e10 * v; }
    { const double v = 30.07 * gj * jc11; sink(11, 30, v); acc += e11 * v; }
    { const double v = 42.037 * gj * jc12; sink(12, 30, v); acc += e12 * v; }
    sink(13, 30, 0.0);
    { const double v = 18.015 * gj * jc14; sink(14, 30, v); acc += e14 * v; }
    { const double v = 31.998 * gj * jc15; sink(15, 30, v); acc += e15 * v; }
    { const double v = 14.027 * gj * jc16; sink(16, 30, v); acc += e16 * v; }
    { const double v = 29.018 * gj * jc17; sink(17, 30, v); acc += e17 * v; }
    { const double v = 15.035 * gj * jc18; sink(18, 30, v); acc += e18 * v; }
    { const double v = 2.016 * gj * jc19; sink(19, 30, v); acc += e19 * v; }
    { const double v = 43.045 * gj * jc20; sink(20, 30, v); acc += e20 * v; }
    sink(21, 30, 0.0);
    sink(22, 30, 0.0);
    sink(23, 30, 0.0);
    { const double v = 33.006 * gj * jc24; sink(24, 30, v); acc += e24 * v; }
    { const double v = 28.054 * gj * jc25; sink(25, 30, v); acc += e25 * v; }
    { const double v = 13.019 * gj * jc26; sink(26, 30, v); acc += e26 * v; }
    { const double v = 17.007 * gj * jc27; sink(27, 30, v); acc += e27 * v; }
    { const double v = 29.062 * gj * jc28; sink(28, 30, v); acc += e28 * v; }
    { const double v = 14.027 * gj * jc29; sink(29, 30, v); acc += e29 * v; }
    { const double v = 44.009 * gj * jc30; sink(30, 30, v); acc += e30 * v; }
    const double cvj = (ld2(rh0 + 60).x + (ld2(rh1 + 60).x - ld2(rh0 + 60).x) * Td) - 188.92749210388786;
    sink(NS, 30, -(acc + Tdot * cvj * (((negmask >> 30) & 1u) ? 0.0 : 1.0)) * invD);
  }
  return status;
}

// chemistry_source over a batch: rho[k*ld+i], T[i] -> wdot[k*ld+i] (no clipping, as the generated
// routines; Lib_ChemMech/*.f90).
__global__ __launch_bounds__(256) void wdot_kernel(long n, long ld, const double* __restrict__ rho,
    const double* __restrict__ T, double* __restrict__ wdot, int* __restrict__ status,
    const double* __restrict__ tabk, int Tk_min, int Tk_max) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double r[NS], F[NZ];
#pragma unroll
  for (int k = 0; k < NS; ++k) r[k] = stream_load(rho + k * ld + i);
  const int st = eval<false, false, 1>(tabk, Tk_min, Tk_max, nullptr, 0, 0, r, stream_load(T + i), F, nullptr,
      NullSink{}, NullOut{});
#pragma unroll
  for (int k = 0; k < NS; ++k) stream_store(wdot + k * ld + i, F[k]);
  if (status) status[i] = st;
}

constexpr int JAC_BLOCK = 128;
// rhs_native and its Jacobian over a batch: Z[k*ld+i] (rho_1..rho_NS, T) -> F[k*ld+i],
// J[(row*NZ+col)*ld+i] (J may be null: RHS only).
template <bool JAC>
__global__ __launch_bounds__(JAC_BLOCK) void rhs_jac_kernel(long n, long ld, const double* __restrict__ Z,
    double* __restrict__ Fo, double* __restrict__ J, int* __restrict__ status,
    const double* __restrict__ tabk, int Tk_min, int Tk_max, const double* __restrict__ tabh,
    int Th_min, int Th_max) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double r[NS], F[NZ];
#pragma unroll
  for (int k = 0; k < NS; ++k) r[k] = stream_load(Z + k * ld + i);
  const double T = stream_load(Z + NS * ld + i);
  double* __restrict__ Ji = J + i;
  __shared__ double smem[JAC ? NSLOT * JAC_BLOCK : 1];
  const int st = eval<JAC, true, JAC_BLOCK>(tabk, Tk_min, Tk_max, tabh, Th_min, Th_max, r, T, F, smem + threadIdx.x,
      [&](int row, int col, double v) { stream_store(Ji + (long)(row * NZ + col) * ld, v); },
      [&](const double (&Fv)[NZ]) {
#pragma unroll
        for (int k = 0; k < NZ; ++k) stream_store(Fo + k * ld + i, Fv[k]);
      });
  if (!JAC) {
#pragma unroll
    for (int k = 0; k < NZ; ++k) stream_store(Fo + k * ld + i, F[k]);
  }
  if (status) status[i] = st;
}
static const int p_rx_kind[] = {5, 5, 4, 4, 4, 4, 4, 4, 4, 4, 4, 5, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 5, 5, 5, 4, 5, 5, 5, 5, 5, 4, 4, 4, 4, 4, 6, 4, 6, 4, 6, 4, 4, 6, 6, 4, 6, 4, 6, 4, 4, 4, 6, 4, 6, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 6, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 6, 4, 4, 4, 4, 4, 5, 5, 4, 4, 6, 4, 4, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
static const int p_rx_col[] = {0, 2, 4, 6, 8, 10, 12, 14, 16, 18, 20, 22, 24, 26, 28, 30, 32, 34, 36, 38, 40, 42, 44, 46, 48, 50, 52, 54, 56, 58, 60, 62, 64, 66, 68, 70, 72, 370, 74, 374, 76, 378, 78, 80, 382, 386, 82, 390, 84, 394, 86, 88, 90, 398, 92, 402, 94, 96, 98, 100, 102, 104, 106, 108, 110, 112, 114, 116, 118, 120, 122, 124, 126, 128, 130, 132, 134, 136, 138, 140, 142, 144, 146, 148, 150, 152, 154, 156, 158, 160, 162, 164, 166, 168, 406, 170, 172, 174, 176, 178, 180, 182, 184, 186, 188, 190, 192, 410, 194, 196, 198, 200, 202, 204, 206, 208, 210, 414, 212, 214, 216, 218, 220, 222, 224, 226, 228, 230, 232, 234, 236, 238, 240, 242, 244, 246, 248, 250, 252, 254, 256, 258, 260, 262, 264, 266, 268, 270, 272, 274, 276, 278, 280, 282, 284, 286, 288, 290, 292, 294, 296, 298, 300, 302, 304, 306, 308, 310, 312, 314, 316, 318, 320, 322, 324, 326, 328, 330, 332, 334, 336, 338, 340, 342, 344, 346, 348, 350, 352, 354, 356, 358, 360, 362, 364, 366, 368};
static const int p_rx_mset[] = {0, 1, -1, -1, -1, -1, -1, -1, -1, -1, -1, 2, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, 3, 4, 5, -1, 6, 7, 5, 8, 9, -1, -1, -1, -1, -1, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1, 1, -1, 1, -1, -1, -1, 1, -1, 1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, 1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, 1, -1, -1, -1, -1, -1, 5, 10, -1, -1, 1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, 11, 11, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1};
static const int p_rx_fac[] = {519, -1, -1, -1, 271, -1, -1, -1, 258, 263, -1, -1, 283, -1, -1, -1, 275, 263, -1, -1, 258, 283, -1, -1, 280, 263, -1, -1, 271, 283, -1, -1, 264, 263, -1, -1, 280, 283, -1, -1, 282, 263, -1, -1, 265, 258, -1, -1, 285, 263, -1, -1, 258, 273, -1, -1, 272, 263, -1, -1, 265, 275, -1, -1, 272, 263, -1, -1, 258, 273, -1, -1, 274, 263, -1, -1, 266, 258, -1, -1, 259, 263, -1, -1, 274, 283, -1, -1, 265, 263, -1, -1, 286, -1, -1, -1, 273, 263, -1, -1, 265, 283, -1, -1, 273, 263, -1, -1, 286, 258, -1, -1, 266, 263, -1, -1, 273, 283, -1, -1, 260, 263, -1, -1, 285, 265, -1, -1, 262, 263, -1, -1, 268, 258, -1, -1, 281, 263, -1, -1, 274, 273, -1, -1, 284, 263, -1, -1, 266, 274, -1, -1, 267, 263, -1, -1, 284, 283, -1, -1, 268, 263, -1, -1, 285, 286, -1, -1, 265, 271, -1, -1, 286, 263, -1, -1, 266, 271, -1, -1, 273, 280, -1, -1, 258, 271, -1, -1, 280, -1, -1, -1, 258, 271, -1, -1, 280, -1, -1, -1, 258, 271, -1, -1, 280, -1, -1, -1, 258, 271, -1, -1, 263, 283, -1, -1, 514, -1, -1, -1, 275, -1, -1, -1, 514, -1, -1, -1, 275, -1, -1, -1, 514, -1, -1, -1, 275, -1, -1, -1, 514, -1, -1, -1, 275, -1, -1, -1, 258, 283, -1, -1, 270, -1, -1, -1, 258, 280, -1, -1, 270, 263, -1, -1, 258, 280, -1, -1, 275, 271, -1, -1, 258, 280, -1, -1, 539, -1, -1, -1, 258, 264, -1, -1, 275, 280, -1, -1, 258, 264, -1, -1, 270, 283, -1, -1, 285, 258, -1, -1, 274, -1, -1, -1, 272, 258, -1, -1, 282, 275, -1, -1, 274, 258, -1, -1, 259, -1, -1, -1, 259, 258, -1, -1, 274, 275, -1, -1, 258, 273, -1, -1, 266, -1, -1, -1, 258, 273, -1, -1, 265, 275, -1, -1, 266, 258, -1, -1, 275, 273, -1, -1, 260, 258, -1, -1, 262, -1, -1, -1, 262, 258, -1, -1, 281, -1, -1, -1, 262, 258, -1, -1, 260, 275, -1, -1, 281, 258, -1, -1, 284, -1, -1, -1, 281, 258, -1, -1, 262, 275, -1, -1, 284, 258, -1, -1, 267, -1, -1, -1, 284, 258, -1, -1, 281, 275, -1, -1, 267, 258, -1, -1, 284, 275, -1, -1, 268, 258, -1, -1, 274, 265, -1, -1, 265, 275, -1, -1, 266, -1, -1, -1, 275, 283, -1, -1, 258, 270, -1, -1, 539, -1, -1, -1, 264, -1, -1, -1, 539, -1, -1, -1, 270, 263, -1, -1, 280, 283, -1, -1, 270, 271, -1, -1, 264, 283, -1, -1, 270, 280, -1, -1, 264, 283, -1, -1, 270, 280, -1, -1, 282, 283, -1, -1, 258, 273, -1, -1, 285, 283, -1, -1, 266, 258, -1, -1, 285, 283, -1, -1, 282, 270, -1, -1, 272, 283, -1, -1, 266, 258, -1, -1, 274, 283, -1, -1, 285, 270, -1, -1, 274, 283, -1, -1, 272, 270, -1, -1, 259, 283, -1, -1, 274, 270, -1, -1, 265, 283, -1, -1, 286, 258, -1, -1, 273, 283, -1, -1, 265, 270, -1, -1, 266, 283, -1, -1, 270, 273, -1, -1, 260, 283, -1, -1, 268, 258, -1, -1, 260, 283, -1, -1, 274, 265, -1, -1, 262, 283, -1, -1, 260, 270, -1, -1, 281, 283, -1, -1, 262, 270, -1, -1, 267, 283, -1, -1, 284, 270, -1, -1, 536, -1, -1, -1, 264, 271, -1, -1, 536, -1, -1, -1, 264, 271, -1, -1, 285, 280, -1, -1, 266, 283, -1, -1, 274, 280, -1, -1, 259, 271, -1, -1, 265, 280, -1, -1, 286, 283, -1, -1, 266, 280, -1, -1, 264, 273, -1, -1, 282, 271, -1, -1, 273, 263, -1, -1, 282, 275, -1, -1, 285, 258, -1, -1, 282, 270, -1, -1, 266, 258, -1, -1, 282, 285, -1, -1, 260, 258, -1, -1, 282, 274, -1, -1, 262, 258, -1, -1, 282, 259, -1, -1, 281, 258, -1, -1, 282, 286, -1, -1, 265, 273, -1, -1, 282, 266, -1, -1, 268, 258, -1, -1, 285, 271, -1, -1, 273, 283, -1, -1, 285, 275, -1, -1, 274, 258, -1, -1, 541, -1, -1, -1, 260, 275, -1, -1, 285, 274, -1, -1, 281, 258, -1, -1, 285, 259, -1, -1, 530, -1, -1, -1, 285, 265, -1, -1, 268, -1, -1, -1, 272, 271, -1, -1, 265, 258, 283, -1, 272, 271, -1, -1, 265, 270, -1, -1, 272, 275, -1, -1, 274, 258, -1, -1, 272, 270, -1, -1, 285, 270, -1, -1, 272, 274, -1, -1, 281, 258, -1, -1, 272, 259, -1, -1, 530, -1, -1, -1, 272, 265, -1, -1, 285, 265, -1, -1, 272, 286, -1, -1, 285, 286, -1, -1, 272, 286, -1, -1, 266, 265, -1, -1, 267, 272, -1, -1, 284, 274, -1, -1, 274, 271, -1, -1, 266, 283, -1, -1, 274, 264, -1, -1, 259, 280, -1, -1, 530, -1, -1, -1, 267, -1, -1, -1, 530, -1, -1, -1, 284, 258, -1, -1, 274, 273, -1, -1, 259, 265, -1, -1, 266, 274, -1, -1, 259, 273, -1, -1, 281, 274, -1, -1, 262, 259, -1, -1, 267, 274, -1, -1, 284, 259, -1, -1, 273, -1, -1, -1, 265, 258, -1, -1, 273, -1, -1, -1, 265, 258, -1, -1, 273, 271, -1, -1, 265, 280, -1, -1, 262, 271, -1, -1, 266, 273, -1, -1, 281, -1, -1, -1, 260, 275, -1, -1, 284, 271, -1, -1, 281, 280, -1, -1, 274, 283, -1, -1, 266, 275, -1, -1, 262, 284, -1, -1, 537, -1, -1, -1, 537, -1, -1, -1, 262, 284, -1, -1, 281, 263, -1, -1, 262, 283, -1, -1, 262, 283, -1, -1, 281, 263, -1, -1, 260, 271, -1, -1, 529, -1, -1, -1, 529, -1, -1, -1, 260, 271, -1, -1, 282, 271, -1, -1, 265, 283, -1, -1, 265, 283, -1, -1, 282, 271, -1, -1, 262, 271, -1, -1, 260, 280, -1, -1, 260, 280, -1, -1, 262, 271, -1, -1, 268, 263, -1, -1, 529, -1, -1, -1, 529, -1, -1, -1, 268, 263, -1, -1, 268, 283, -1, -1, 266, 273, -1, -1, 266, 273, -1, -1, 268, 283, -1, -1, 267, 271, -1, -1, 284, 280, -1, -1, 284, 280, -1, -1, 267, 271, -1, -1, 267, 280, -1, -1, 284, 264, -1, -1, 284, 264, -1, -1, 267, 280, -1, -1, 281, 267, -1, -1, 540, -1, -1, -1, 540, -1, -1, -1, 281, 267, -1, -1, 262, 274, -1, -1, 260, 259, -1, -1, 260, 259, -1, -1, 262, 274, -1, -1, 284, 274, -1, -1, 281, 259, -1, -1, 281, 259, -1, -1, 284, 274, -1, -1, 276, -1, -1, -1, 274, 265, -1, -1, 274, 265, -1, -1, 276, -1, -1, -1, 281, 271, -1, -1, 262, 280, -1, -1, 262, 280, -1, -1, 281, 271, -1, -1, 281, 280, -1, -1, 262, 264, -1, -1, 262, 264, -1, -1, 281, 280, -1, -1, 269, -1, -1, -1, 281, 274, -1, -1, 281, 274, -1, -1, 269, -1, -1, -1, 276, 263, -1, -1, 268, 283, -1, -1, 268, 283, -1, -1, 276, 263, -1, -1, 274, 276, -1, -1, 268, 259, -1, -1, 268, 259, -1, -1, 274, 276, -1, -1, 281, 263, -1, -1, 257, 258, -1, -1, 257, 258, -1, -1, 281, 263, -1, -1, 267, 282, -1, -1, 284, 285, -1, -1, 284, 285, -1, -1, 267, 282, -1, -1, 267, 285, -1, -1, 284, 274, -1, -1, 284, 274, -1, -1, 267, 285, -1, -1, 281, 283, -1, -1, 266, 274, -1, -1, 266, 274, -1, -1, 281, 283, -1, -1, 278, -1, -1, -1, 274, 273, -1, -1, 274, 273, -1, -1, 278, -1, -1, -1, 278, 283, -1, -1, 279, 270, -1, -1, 279, 270, -1, -1, 278, 283, -1, -1, 278, 258, -1, -1, 279, 275, -1, -1, 279, 275, -1, -1, 278, 258, -1, -1, 278, 280, -1, -1, 279, 264, -1, -1, 279, 264, -1, -1, 278, 280, -1, -1, 278, 274, -1, -1, 279, 259, -1, -1, 279, 259, -1, -1, 278, 274, -1, -1, 277, -1, -1, -1, 281, 266, -1, -1, 281, 266, -1, -1, 277, -1, -1, -1, 277, 283, -1, -1, 262, 266, 270, -1, 277, 263, -1, -1, 262, 266, 283, -1, 277, 258, -1, -1, 262, 266, 275, -1, 277, 280, -1, -1, 262, 266, 264, -1, 277, 274, -1, -1, 262, 266, 259, -1, 261, 271, -1, -1, 268, 273, -1, -1, 268, 273, -1, -1, 261, 271, -1, -1, 279, -1, -1, -1, 276, -1, -1, -1, 276, -1, -1, -1, 279, -1, -1, -1, 279, -1, -1, -1, 257, -1, -1, -1, 257, -1, -1, -1, 279, -1, -1, -1, 257, -1, -1, -1, 268, 258, -1, -1, 268, 258, -1, -1, 257, -1, -1, -1, 257, 271, -1, -1, 268, 280, -1, -1, 268, 280, -1, -1, 257, 271, -1, -1, 268, 283, -1, -1, 274, 286, -1, -1, 274, 286, -1, -1, 268, 283, -1, -1, 256, -1, -1, -1, 265, 269, -1, -1, 265, 269, -1, -1, 256, -1, -1, -1, 281, 280, -1, -1, 278, 283, -1, -1, 278, 283, -1, -1, 281, 280, -1, -1};
static const int p_wd_ptr[] = {0, 2, 10, 75, 97, 112, 114, 141, 176, 193, 222, 252, 269, 291, 295, 318, 353, 368, 401, 452, 479, 487, 494, 506, 518, 557, 591, 607, 664, 689, 712, 723};
static const int p_wd_rx[] = {193, 194, 156, 157, 185, 186, 187, 188, 189, 190, 1, 2, 5, 6, 8, 9, 13, 16, 23, 24, 25, 26, 27, 28, 29, 30, 31, 32, 33, 34, 35, 36, 37, 38, 39, 40, 41, 42, 43, 44, 45, 46, 47, 48, 49, 50, 51, 52, 54, 60, 61, 63, 67, 70, 82, 83, 84, 85, 86, 88, 90, 92, 95, 97, 99, 108, 113, 114, 156, 157, 168, 169, 178, 187, 188, 10, 39, 40, 66, 78, 86, 93, 100, 106, 109, 110, 111, 112, 140, 141, 142, 143, 154, 155, 172, 173, 180, 15, 44, 46, 70, 71, 72, 84, 91, 117, 124, 125, 128, 129, 140, 141, 181, 182, 16, 44, 45, 46, 48, 72, 73, 85, 111, 116, 120, 121, 122, 123, 128, 129, 140, 141, 146, 147, 148, 149, 176, 177, 178, 179, 180, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 26, 32, 56, 81, 122, 123, 130, 131, 152, 153, 156, 157, 177, 4, 35, 36, 55, 58, 59, 75, 76, 80, 106, 136, 137, 148, 149, 170, 171, 179, 5, 7, 11, 12, 15, 21, 42, 52, 53, 67, 68, 71, 79, 87, 94, 95, 96, 101, 103, 109, 113, 114, 115, 126, 127, 144, 145, 193, 194, 9, 14, 18, 22, 41, 43, 53, 61, 63, 69, 77, 80, 83, 88, 103, 105, 110, 116, 119, 132, 133, 162, 163, 174, 175, 176, 177, 178, 179, 180, 19, 49, 51, 74, 104, 107, 112, 134, 135, 136, 137, 138, 139, 158, 159, 160, 161, 16, 20, 52, 70, 88, 94, 130, 131, 132, 133, 152, 153, 154, 155, 181, 182, 187, 188, 189, 190, 191, 192, 150, 151, 193, 194, 31, 32, 36, 54, 56, 57, 58, 59, 62, 64, 65, 66, 68, 69, 72, 73, 74, 83, 96, 98, 166, 167, 176, 0, 3, 21, 22, 23, 24, 25, 26, 33, 57, 75, 76, 78, 81, 89, 95, 96, 105, 115, 116, 118, 124, 125, 126, 127, 128, 129, 134, 135, 146, 147, 181, 182, 189, 190, 7, 8, 38, 63, 65, 95, 96, 97, 98, 99, 100, 101, 102, 103, 104, 6, 8, 12, 13, 14, 17, 22, 41, 42, 43, 60, 68, 69, 80, 81, 87, 89, 109, 110, 113, 114, 115, 116, 124, 125, 130, 131, 132, 133, 164, 165, 181, 182, 9, 10, 17, 18, 37, 39, 40, 52, 64, 65, 66, 71, 78, 85, 90, 92, 93, 97, 99, 100, 104, 105, 106, 107, 108, 109, 110, 111, 112, 119, 140, 141, 142, 143, 144, 145, 150, 151, 154, 155, 160, 161, 162, 163, 164, 165, 172, 173, 180, 191, 192, 2, 7, 27, 28, 29, 30, 33, 35, 38, 40, 42, 43, 46, 48, 50, 51, 53, 54, 82, 90, 91, 97, 117, 119, 168, 169, 178, 144, 145, 152, 153, 154, 155, 183, 184, 174, 175, 176, 177, 178, 179, 180, 164, 165, 166, 167, 168, 169, 170, 171, 172, 173, 195, 196, 166, 167, 168, 169, 170, 171, 172, 173, 183, 184, 185, 186, 3, 4, 22, 23, 24, 25, 32, 33, 34, 35, 57, 58, 59, 75, 76, 77, 78, 79, 80, 106, 115, 118, 128, 129, 134, 135, 136, 137, 146, 147, 148, 149, 170, 171, 179, 189, 190, 195, 196, 17, 45, 47, 48, 50, 73, 86, 92, 99, 111, 117, 118, 120, 121, 122, 123, 138, 139, 142, 143, 146, 147, 148, 149, 150, 151, 156, 157, 162, 163, 174, 175, 195, 196, 5, 38, 60, 62, 81, 82, 83, 84, 85, 86, 87, 88, 126, 127, 158, 159, 1, 2, 3, 4, 10, 12, 14, 19, 26, 31, 34, 36, 54, 55, 56, 57, 58, 59, 60, 61, 62, 63, 64, 65, 66, 67, 68, 69, 70, 71, 72, 73, 74, 77, 79, 89, 95, 105, 119, 122, 123, 126, 127, 132, 133, 152, 153, 162, 163, 166, 167, 176, 177, 191, 192, 195, 196, 18, 19, 47, 49, 50, 51, 74, 104, 108, 112, 118, 120, 121, 134, 135, 136, 137, 138, 139, 142, 143, 158, 159, 160, 161, 6, 15, 20, 37, 61, 62, 64, 77, 82, 84, 89, 90, 91, 92, 93, 94, 98, 101, 102, 158, 159, 160, 161, 11, 13, 20, 21, 67, 79, 87, 102, 103, 191, 192};
static const double p_wd_coef[] = {-1.0, 1.0, 1.0, -1.0, 1.0, -1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, -1.0, -1.0, -1.0, -1.0, -2.0, -2.0, -2.0, -2.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, -1.0, -1.0, 1.0, -1.0, 1.0, -1.0, -1.0, 1.0, -1.0, -1.0, 1.0, -1.0, -1.0, -1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, -1.0, 1.0, -1.0, -1.0, 1.0, 1.0, 1.0, 1.0, -1.0, 1.0, 1.0, -1.0, 1.0, -1.0, -1.0, 1.0, -1.0, 1.0, -1.0, -1.0, 1.0, -1.0, 1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0, -1.0, -1.0, 1.0, -1.0, 1.0, 1.0, -1.0, 1.0, -1.0, 1.0, 1.0, 1.0, 1.0, 1.0, -2.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, 1.0, 1.0, 1.0, 1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, -1.0, -1.0, -1.0, 1.0, -1.0, -1.0, 1.0, 1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, 1.0, 1.0, -1.0, 1.0, 1.0, -1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0, -1.0, 1.0, -1.0, 1.0, 1.0, 0.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, 1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, 1.0, -1.0, 1.0, 1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, 1.0, 1.0, 1.0, 1.0, -1.0, 1.0, -1.0, -1.0, -1.0, 1.0, -1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0, 1.0, -1.0, 1.0, -1.0, 1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, -1.0, 1.0, -1.0, 1.0, 1.0, -1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, -1.0, 1.0, 0.0, 1.0, -1.0, 1.0, 1.0, 1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, 1.0, 1.0, 1.0, 1.0, 1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, -1.0, -1.0, -1.0, 1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0, -1.0, 1.0, 1.0, 1.0, 1.0, 1.0, -1.0, 1.0, -1.0, -1.0, -1.0, 1.0, 2.0, -2.0, 2.0, -2.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, -1.0, 1.0, 1.0, 1.0, 1.0, -1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0, -1.0, -1.0, 1.0, -1.0, 2.0, 1.0, -1.0, 2.0, 1.0, -1.0, -1.0, -2.0, -2.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, 1.0, -1.0, 1.0, 1.0, -1.0, 1.0, -1.0, -1.0, 1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, -1.0, 1.0, -1.0, 1.0, -1.0, -1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, -1.0, -1.0, -1.0, -1.0, 1.0, -1.0, 1.0, 1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, 1.0, -1.0, -1.0, 1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, 1.0, 1.0, 1.0, 1.0, -1.0, -1.0, -1.0, 1.0, -1.0, 1.0, 1.0, -2.0, -2.0, -1.0, -1.0, -1.0, -1.0, 1.0, 1.0, 1.0, 1.0, -1.0, 1.0, -1.0, -1.0, 1.0, 1.0, -1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, -1.0, 1.0, -1.0, 1.0, -1.0, -1.0, 1.0, -1.0, 1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 2.0, -2.0, -1.0, 1.0, -1.0, 1.0, 1.0, -1.0, -1.0, 1.0, -1.0, 1.0, 1.0, -1.0, -1.0, 1.0, -1.0, 1.0, 1.0, -1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, 1.0, -1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, -1.0, 2.0, 1.0, -1.0, -2.0, -2.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, 1.0, 1.0, 1.0, 1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, -1.0, 1.0, 1.0, -1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0, 1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0, -1.0, 1.0, -1.0, 2.0, -2.0, -1.0, 1.0, 1.0, -1.0, 1.0, -1.0, -1.0, 1.0, 1.0, -1.0, -1.0, -1.0, 1.0, -1.0, 1.0, -1.0, -1.0, -1.0, -2.0, -1.0, -1.0, -1.0, 1.0, 1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, -1.0, 0.0, -1.0, 1.0, -1.0};
static const int p_ms_ptr[] = {0, 31, 62, 93, 122, 123, 124, 152, 153, 154, 185, 215, 246};
static const int p_ms_sp[] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 15, 14, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 15, 16, 17, 18, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 19, 30, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30};
static const double p_ms_eff[] = {1.0, 1.0, 1.0, 2.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.75, 1.0, 3.0, 1.0, 1.0, 15.399999618530273, 1.0, 1.0, 1.0, 1.0, 2.4000000953674316, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 3.5999999046325684, 1.0, 1.0, 1.0, 2.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.5, 1.0, 3.0, 1.0, 1.0, 6.0, 1.0, 1.0, 1.0, 1.0, 2.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 2.0, 1.0, 1.0, 1.0, 2.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.5, 1.0, 3.0, 1.0, 1.0, 6.0, 6.0, 1.0, 1.0, 1.0, 2.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 3.5, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 0.75, 1.0, 1.5, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.5, 1.0, 1.0, 1.0, 1.0, 1.0, 2.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 3.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 2.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 3.0, 1.0, 1.0, 3.6500000953674316, 1.0, 1.0, 1.0, 1.0, 0.7300000190734863, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 2.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.5, 1.0, 3.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 2.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 2.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0};
const flint_host_prog host_prog = {31, 197, 12, 723, 246, p_rx_kind, p_rx_col, p_rx_mset, p_rx_fac, p_wd_ptr, p_wd_rx, p_wd_coef, p_ms_ptr, p_ms_sp, p_ms_eff};
// element matrix for the equilibrium kernel: a(e, j) = atoms of element e in species j
struct Cea {
  static constexpr int NS = 31, NE = 3;
  __host__ __device__ static constexpr double a(int e, int j) {
    constexpr double t[NE][NS] = {{4.0, 2.0, 0.0, 1.0, 2.0, 3.0, 2.0, 0.0, 0.0, 1.0, 1.0, 2.0, 2.0, 3.0, 0.0, 0.0, 1.0, 1.0, 1.0, 0.0, 2.0, 3.0, 2.0, 2.0, 0.0, 2.0, 1.0, 0.0, 2.0, 1.0, 1.0}, {7.0, 3.0, 1.0, 4.0, 2.0, 3.0, 3.0, 0.0, 2.0, 0.0, 2.0, 6.0, 2.0, 7.0, 2.0, 0.0, 2.0, 1.0, 3.0, 2.0, 3.0, 6.0, 4.0, 3.0, 1.0, 4.0, 1.0, 1.0, 5.0, 2.0, 0.0}, {1.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 2.0, 1.0, 1.0, 0.0, 1.0, 0.0, 1.0, 2.0, 0.0, 1.0, 0.0, 0.0, 1.0, 1.0, 1.0, 1.0, 2.0, 0.0, 0.0, 1.0, 0.0, 0.0, 2.0}};
    return t[e][j];
  }
};
static const double cea_A[] = {4.0, 2.0, 0.0, 1.0, 2.0, 3.0, 2.0, 0.0, 0.0, 1.0, 1.0, 2.0, 2.0, 3.0, 0.0, 0.0, 1.0, 1.0, 1.0, 0.0, 2.0, 3.0, 2.0, 2.0, 0.0, 2.0, 1.0, 0.0, 2.0, 1.0, 1.0, 7.0, 3.0, 1.0, 4.0, 2.0, 3.0, 3.0, 0.0, 2.0, 0.0, 2.0, 6.0, 2.0, 7.0, 2.0, 0.0, 2.0, 1.0, 3.0, 2.0, 3.0, 6.0, 4.0, 3.0, 1.0, 4.0, 1.0, 1.0, 5.0, 2.0, 0.0, 1.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 2.0, 1.0, 1.0, 0.0, 1.0, 0.0, 1.0, 2.0, 0.0, 1.0, 0.0, 0.0, 1.0, 1.0, 1.0, 1.0, 2.0, 0.0, 0.0, 1.0, 0.0, 0.0, 2.0};
}  // namespace flint_gen_tsr_rich_31

#include "../flint_cea_kernel.h"
FLINT_REGISTER_MECH_CEA(tsr_rich_31, flint_gen_tsr_rich_31)
