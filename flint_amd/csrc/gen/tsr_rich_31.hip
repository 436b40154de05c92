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
  const double2 A0_1 = ld2(rk0 + 0), A1_1 = ld2(rk1 + 0);
  const double2 A0_2 = ld2(rk0 + 2), A1_2 = ld2(rk1 + 2);
  const double2 A0_3 = ld2(rk0 + 4), A1_3 = ld2(rk1 + 4);
  const double2 A0_4 = ld2(rk0 + 6), A1_4 = ld2(rk1 + 6);
  {  // reaction 1: 2 O + M <=> O2 + M
    const double2 a0 = A0_1, a1 = A1_1;
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
    const double2 a0 = A0_2, a1 = A1_2;
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
    const double2 a0 = A0_3, a1 = A1_3;
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
    const double2 a0 = A0_4, a1 = A1_4;
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
  const double2 A0_5 = ld2(rk0 + 8), A1_5 = ld2(rk1 + 8);
  const double2 A0_6 = ld2(rk0 + 10), A1_6 = ld2(rk1 + 10);
  const double2 A0_7 = ld2(rk0 + 12), A1_7 = ld2(rk1 + 12);
  const double2 A0_8 = ld2(rk0 + 14), A1_8 = ld2(rk1 + 14);
  {  // reaction 5: H2O2 + O <=> HO2 + OH
    const double2 a0 = A0_5, a1 = A1_5;
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
    const double2 a0 = A0_6, a1 = A1_6;
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
    const double2 a0 = A0_7, a1 = A1_7;
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
    const double2 a0 = A0_8, a1 = A1_8;
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
  const double2 A0_9 = ld2(rk0 + 16), A1_9 = ld2(rk1 + 16);
  const double2 A0_10 = ld2(rk0 + 18), A1_10 = ld2(rk1 + 18);
  const double2 A0_11 = ld2(rk0 + 20), A1_11 = ld2(rk1 + 20);
  const double2 A0_12 = ld2(rk0 + 22), A1_12 = ld2(rk1 + 22);
  {  // reaction 9: CH2(S) + O <=> H + HCO
    const double2 a0 = A0_9, a1 = A1_9;
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
    const double2 a0 = A0_10, a1 = A1_10;
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
    const double2 a0 = A0_11, a1 = A1_11;
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
    const double2 a0 = A0_12, a1 = A1_12;
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
  const double2 A0_13 = ld2(rk0 + 24), A1_13 = ld2(rk1 + 24);
  const double2 A0_14 = ld2(rk0 + 26), A1_14 = ld2(rk1 + 26);
  const double2 A0_15 = ld2(rk0 + 28), A1_15 = ld2(rk1 + 28);
  const double2 A0_16 = ld2(rk0 + 30), A1_16 = ld2(rk1 + 30);
  {  // reaction 13: HCO + O <=> CO + OH
    const double2 a0 = A0_13, a1 = A1_13;
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
    const double2 a0 = A0_14, a1 = A1_14;
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
    const double2 a0 = A0_15, a1 = A1_15;
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
    const double2 a0 = A0_16, a1 = A1_16;
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
  const double2 A0_17 = ld2(rk0 + 32), A1_17 = ld2(rk1 + 32);
  const double2 A0_18 = ld2(rk0 + 34), A1_18 = ld2(rk1 + 34);
  const double2 A0_19 = ld2(rk0 + 36), A1_19 = ld2(rk1 + 36);
  const double2 A0_20 = ld2(rk0 + 38), A1_20 = ld2(rk1 + 38);
  {  // reaction 17: C2H3 + O <=> CH2CO + H
    const double2 a0 = A0_17, a1 = A1_17;
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
    const double2 a0 = A0_18, a1 = A1_18;
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
    const double2 a0 = A0_19, a1 = A1_19;
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
    const double2 a0 = A0_20, a1 = A1_20;
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
  const double2 A0_21 = ld2(rk0 + 40), A1_21 = ld2(rk1 + 40);
  const double2 A0_22 = ld2(rk0 + 42), A1_22 = ld2(rk1 + 42);
  const double2 A0_23 = ld2(rk0 + 44), A1_23 = ld2(rk1 + 44);
  const double2 A0_24 = ld2(rk0 + 46), A1_24 = ld2(rk1 + 46);
  {  // reaction 21: CH2CO + O <=> CH2 + CO2
    const double2 a0 = A0_21, a1 = A1_21;
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
    const double2 a0 = A0_22, a1 = A1_22;
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
    const double2 a0 = A0_23, a1 = A1_23;
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
    const double2 a0 = A0_24, a1 = A1_24;
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
  const double2 A0_25 = ld2(rk0 + 48), A1_25 = ld2(rk1 + 48);
  const double2 A0_26 = ld2(rk0 + 50), A1_26 = ld2(rk1 + 50);
  const double2 A0_27 = ld2(rk0 + 52), A1_27 = ld2(rk1 + 52);
  const double2 A0_28 = ld2(rk0 + 54), A1_28 = ld2(rk1 + 54);
  {  // reaction 25: H + O2 + O2 <=> HO2 + O2
    const double2 a0 = A0_25, a1 = A1_25;
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
    const double2 a0 = A0_26, a1 = A1_26;
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
    const double2 a0 = A0_27, a1 = A1_27;
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
    const double2 a0 = A0_28, a1 = A1_28;
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
  const double2 A0_29 = ld2(rk0 + 56), A1_29 = ld2(rk1 + 56);
  const double2 A0_30 = ld2(rk0 + 58), A1_30 = ld2(rk1 + 58);
  const double2 A0_31 = ld2(rk0 + 60), A1_31 = ld2(rk1 + 60);
  const double2 A0_32 = ld2(rk0 + 62), A1_32 = ld2(rk1 + 62);
  {  // reaction 29: 2 H + H2 <=> H2 + H2
    const double2 a0 = A0_29, a1 = A1_29;
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
    const double2 a0 = A0_30, a1 = A1_30;
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
    const double2 a0 = A0_31, a1 = A1_31;
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
    const double2 a0 = A0_32, a1 = A1_32;
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
  const double2 A0_33 = ld2(rk0 + 64), A1_33 = ld2(rk1 + 64);
  const double2 A0_34 = ld2(rk0 + 66), A1_34 = ld2(rk1 + 66);
  const double2 A0_35 = ld2(rk0 + 68), A1_35 = ld2(rk1 + 68);
  const double2 A0_36 = ld2(rk0 + 70), A1_36 = ld2(rk1 + 70);
  {  // reaction 33: H + HO2 <=> H2O + O
    const double2 a0 = A0_33, a1 = A1_33;
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
    const double2 a0 = A0_34, a1 = A1_34;
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
    const double2 a0 = A0_35, a1 = A1_35;
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
    const double2 a0 = A0_36, a1 = A1_36;
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
  const double2 A0_37 = ld2(rk0 + 72), A1_37 = ld2(rk1 + 72);
  const double2 A0_38 = ld2(rk0 + 74), A1_38 = ld2(rk1 + 74);
  {  // reaction 37: H + H2O2 <=> H2O + OH
    const double2 a0 = A0_37, a1 = A1_37;
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
    const double2 a0 = A0_38, a1 = A1_38;
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
  const double2 A0_39 = ld2(rk0 + 76), A1_39 = ld2(rk1 + 76);
  const double2 A0_40 = ld2(rk0 + 78), A1_40 = ld2(rk1 + 78);
  const double2 A0_41 = ld2(rk0 + 80), A1_41 = ld2(rk1 + 80);
  {  // reaction 41: CH4 + H <=> CH3 + H2
    const double2 a0 = A0_39, a1 = A1_39;
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
    const double2 a0 = A0_40, a1 = A1_40;
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
    const double2 a0 = A0_41, a1 = A1_41;
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
  const double2 A0_42 = ld2(rk0 + 82), A1_42 = ld2(rk1 + 82);
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
    const double2 a0 = A0_42, a1 = A1_42;
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
  const double2 A0_43 = ld2(rk0 + 84), A1_43 = ld2(rk1 + 84);
  const double2 A0_44 = ld2(rk0 + 86), A1_44 = ld2(rk1 + 86);
  const double2 A0_45 = ld2(rk0 + 88), A1_45 = ld2(rk1 + 88);
  {  // reaction 49: C2H4 + H <=> C2H3 + H2
    const double2 a0 = A0_43, a1 = A1_43;
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
    const double2 a0 = A0_44, a1 = A1_44;
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
    const double2 a0 = A0_45, a1 = A1_45;
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
  const double2 A0_46 = ld2(rk0 + 90), A1_46 = ld2(rk1 + 90);
  const double2 A0_47 = ld2(rk0 + 92), A1_47 = ld2(rk1 + 92);
  {  // reaction 53: CH2CO + H <=> CH3 + CO
    const double2 a0 = A0_46, a1 = A1_46;
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
    const double2 a0 = A0_47, a1 = A1_47;
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
  const double2 A0_48 = ld2(rk0 + 94), A1_48 = ld2(rk1 + 94);
  const double2 A0_49 = ld2(rk0 + 96), A1_49 = ld2(rk1 + 96);
  const double2 A0_50 = ld2(rk0 + 98), A1_50 = ld2(rk1 + 98);
  const double2 A0_51 = ld2(rk0 + 100), A1_51 = ld2(rk1 + 100);
  {  // reaction 57: 2 OH <=> H2O + O
    const double2 a0 = A0_48, a1 = A1_48;
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
    const double2 a0 = A0_49, a1 = A1_49;
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
    const double2 a0 = A0_50, a1 = A1_50;
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
    const double2 a0 = A0_51, a1 = A1_51;
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
  const double2 A0_52 = ld2(rk0 + 102), A1_52 = ld2(rk1 + 102);
  const double2 A0_53 = ld2(rk0 + 104), A1_53 = ld2(rk1 + 104);
  const double2 A0_54 = ld2(rk0 + 106), A1_54 = ld2(rk1 + 106);
  const double2 A0_55 = ld2(rk0 + 108), A1_55 = ld2(rk1 + 108);
  {  // reaction 61: CH + OH <=> H + HCO
    const double2 a0 = A0_52, a1 = A1_52;
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
    const double2 a0 = A0_53, a1 = A1_53;
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
    const double2 a0 = A0_54, a1 = A1_54;
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
    const double2 a0 = A0_55, a1 = A1_55;
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
  const double2 A0_56 = ld2(rk0 + 110), A1_56 = ld2(rk1 + 110);
  const double2 A0_57 = ld2(rk0 + 112), A1_57 = ld2(rk1 + 112);
  const double2 A0_58 = ld2(rk0 + 114), A1_58 = ld2(rk1 + 114);
  const double2 A0_59 = ld2(rk0 + 116), A1_59 = ld2(rk1 + 116);
  {  // reaction 65: CH3 + OH <=> CH2 + H2O
    const double2 a0 = A0_56, a1 = A1_56;
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
    const double2 a0 = A0_57, a1 = A1_57;
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
    const double2 a0 = A0_58, a1 = A1_58;
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
    const double2 a0 = A0_59, a1 = A1_59;
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
  const double2 A0_60 = ld2(rk0 + 118), A1_60 = ld2(rk1 + 118);
  const double2 A0_61 = ld2(rk0 + 120), A1_61 = ld2(rk1 + 120);
  const double2 A0_62 = ld2(rk0 + 122), A1_62 = ld2(rk1 + 122);
  const double2 A0_63 = ld2(rk0 + 124), A1_63 = ld2(rk1 + 124);
  {  // reaction 69: HCO + OH <=> CO + H2O
    const double2 a0 = A0_60, a1 = A1_60;
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
    const double2 a0 = A0_61, a1 = A1_61;
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
    const double2 a0 = A0_62, a1 = A1_62;
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
    const double2 a0 = A0_63, a1 = A1_63;
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
  const double2 A0_64 = ld2(rk0 + 126), A1_64 = ld2(rk1 + 126);
  const double2 A0_65 = ld2(rk0 + 128), A1_65 = ld2(rk1 + 128);
  const double2 A0_66 = ld2(rk0 + 130), A1_66 = ld2(rk1 + 130);
  const double2 A0_67 = ld2(rk0 + 132), A1_67 = ld2(rk1 + 132);
  {  // reaction 73: C2H3 + OH <=> C2H2 + H2O
    const double2 a0 = A0_64, a1 = A1_64;
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
    const double2 a0 = A0_65, a1 = A1_65;
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
    const double2 a0 = A0_66, a1 = A1_66;
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
    const double2 a0 = A0_67, a1 = A1_67;
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
  const double2 A0_68 = ld2(rk0 + 134), A1_68 = ld2(rk1 + 134);
  const double2 A0_69 = ld2(rk0 + 136), A1_69 = ld2(rk1 + 136);
  const double2 A0_70 = ld2(rk0 + 138), A1_70 = ld2(rk1 + 138);
  const double2 A0_71 = ld2(rk0 + 140), A1_71 = ld2(rk1 + 140);
  {  // reaction 77: 2 HO2 <=> H2O2 + O2
    const double2 a0 = A0_68, a1 = A1_68;
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
    const double2 a0 = A0_69, a1 = A1_69;
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
    const double2 a0 = A0_70, a1 = A1_70;
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
    const double2 a0 = A0_71, a1 = A1_71;
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
  const double2 A0_72 = ld2(rk0 + 142), A1_72 = ld2(rk1 + 142);
  const double2 A0_73 = ld2(rk0 + 144), A1_73 = ld2(rk1 + 144);
  const double2 A0_74 = ld2(rk0 + 146), A1_74 = ld2(rk1 + 146);
  const double2 A0_75 = ld2(rk0 + 148), A1_75 = ld2(rk1 + 148);
  {  // reaction 81: CH2O + HO2 <=> H2O2 + HCO
    const double2 a0 = A0_72, a1 = A1_72;
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
    const double2 a0 = A0_73, a1 = A1_73;
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
    const double2 a0 = A0_74, a1 = A1_74;
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
    const double2 a0 = A0_75, a1 = A1_75;
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
  const double2 A0_76 = ld2(rk0 + 150), A1_76 = ld2(rk1 + 150);
  const double2 A0_77 = ld2(rk0 + 152), A1_77 = ld2(rk1 + 152);
  const double2 A0_78 = ld2(rk0 + 154), A1_78 = ld2(rk1 + 154);
  const double2 A0_79 = ld2(rk0 + 156), A1_79 = ld2(rk1 + 156);
  {  // reaction 85: CH + CH2 <=> C2H2 + H
    const double2 a0 = A0_76, a1 = A1_76;
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
    const double2 a0 = A0_77, a1 = A1_77;
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
    const double2 a0 = A0_78, a1 = A1_78;
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
    const double2 a0 = A0_79, a1 = A1_79;
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
  const double2 A0_80 = ld2(rk0 + 158), A1_80 = ld2(rk1 + 158);
  const double2 A0_81 = ld2(rk0 + 160), A1_81 = ld2(rk1 + 160);
  const double2 A0_82 = ld2(rk0 + 162), A1_82 = ld2(rk1 + 162);
  const double2 A0_83 = ld2(rk0 + 164), A1_83 = ld2(rk1 + 164);
  {  // reaction 89: CH + CH2O <=> CH2CO + H
    const double2 a0 = A0_80, a1 = A1_80;
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
    const double2 a0 = A0_81, a1 = A1_81;
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
    const double2 a0 = A0_82, a1 = A1_82;
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
    const double2 a0 = A0_83, a1 = A1_83;
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
  const double2 A0_84 = ld2(rk0 + 166), A1_84 = ld2(rk1 + 166);
  const double2 A0_85 = ld2(rk0 + 168), A1_85 = ld2(rk1 + 168);
  const double2 A0_86 = ld2(rk0 + 170), A1_86 = ld2(rk1 + 170);
  {  // reaction 93: CH2 + CH3 <=> C2H4 + H
    const double2 a0 = A0_84, a1 = A1_84;
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
    const double2 a0 = A0_85, a1 = A1_85;
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
    const double2 a0 = A0_86, a1 = A1_86;
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
  const double2 A0_87 = ld2(rk0 + 172), A1_87 = ld2(rk1 + 172);
  const double2 A0_88 = ld2(rk0 + 174), A1_88 = ld2(rk1 + 174);
  const double2 A0_89 = ld2(rk0 + 176), A1_89 = ld2(rk1 + 176);
  const double2 A0_90 = ld2(rk0 + 178), A1_90 = ld2(rk1 + 178);
  {  // reaction 97: CH2(S) + O2 <=> CO + H2O
    const double2 a0 = A0_87, a1 = A1_87;
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
    const double2 a0 = A0_88, a1 = A1_88;
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
    const double2 a0 = A0_89, a1 = A1_89;
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
    const double2 a0 = A0_90, a1 = A1_90;
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
  const double2 A0_91 = ld2(rk0 + 180), A1_91 = ld2(rk1 + 180);
  const double2 A0_92 = ld2(rk0 + 182), A1_92 = ld2(rk1 + 182);
  const double2 A0_93 = ld2(rk0 + 184), A1_93 = ld2(rk1 + 184);
  const double2 A0_94 = ld2(rk0 + 186), A1_94 = ld2(rk1 + 186);
  {  // reaction 101: CH2(S) + CH4 <=> 2 CH3
    const double2 a0 = A0_91, a1 = A1_91;
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
    const double2 a0 = A0_92, a1 = A1_92;
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
    const double2 a0 = A0_93, a1 = A1_93;
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
    const double2 a0 = A0_94, a1 = A1_94;
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
  const double2 A0_95 = ld2(rk0 + 188), A1_95 = ld2(rk1 + 188);
  const double2 A0_96 = ld2(rk0 + 190), A1_96 = ld2(rk1 + 190);
  const double2 A0_97 = ld2(rk0 + 192), A1_97 = ld2(rk1 + 192);
  {  // reaction 105: C2H6 + CH2(S) <=> C2H5 + CH3
    const double2 a0 = A0_95, a1 = A1_95;
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
    const double2 a0 = A0_96, a1 = A1_96;
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
    const double2 a0 = A0_97, a1 = A1_97;
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
  const double2 A0_98 = ld2(rk0 + 194), A1_98 = ld2(rk1 + 194);
  const double2 A0_99 = ld2(rk0 + 196), A1_99 = ld2(rk1 + 196);
  const double2 A0_100 = ld2(rk0 + 198), A1_100 = ld2(rk1 + 198);
  const double2 A0_101 = ld2(rk0 + 200), A1_101 = ld2(rk1 + 200);
  {  // reaction 109: 2 CH3 <=> C2H5 + H
    const double2 a0 = A0_98, a1 = A1_98;
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
    const double2 a0 = A0_99, a1 = A1_99;
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
    const double2 a0 = A0_100, a1 = A1_100;
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
    const double2 a0 = A0_101, a1 = A1_101;
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
  const double2 A0_102 = ld2(rk0 + 202), A1_102 = ld2(rk1 + 202);
  const double2 A0_103 = ld2(rk0 + 204), A1_103 = ld2(rk1 + 204);
  const double2 A0_104 = ld2(rk0 + 206), A1_104 = ld2(rk1 + 206);
  const double2 A0_105 = ld2(rk0 + 208), A1_105 = ld2(rk1 + 208);
  {  // reaction 113: C2H6 + CH3 <=> C2H5 + CH4
    const double2 a0 = A0_102, a1 = A1_102;
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
    const double2 a0 = A0_103, a1 = A1_103;
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
    const double2 a0 = A0_104, a1 = A1_104;
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
    const double2 a0 = A0_105, a1 = A1_105;
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
  const double2 A0_106 = ld2(rk0 + 210), A1_106 = ld2(rk1 + 210);
  const double2 A0_107 = ld2(rk0 + 212), A1_107 = ld2(rk1 + 212);
  const double2 A0_108 = ld2(rk0 + 214), A1_108 = ld2(rk1 + 214);
  {  // reaction 117: C2H3 + O2 <=> CH2O + HCO
    const double2 a0 = A0_106, a1 = A1_106;
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
    const double2 a0 = A0_107, a1 = A1_107;
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
    const double2 a0 = A0_108, a1 = A1_108;
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
  const double2 A0_109 = ld2(rk0 + 216), A1_109 = ld2(rk1 + 216);
  const double2 A0_110 = ld2(rk0 + 218), A1_110 = ld2(rk1 + 218);
  const double2 A0_111 = ld2(rk0 + 220), A1_111 = ld2(rk1 + 220);
  const double2 A0_112 = ld2(rk0 + 222), A1_112 = ld2(rk1 + 222);
  {  // reaction 121: C2H3 + C2H5 => 2 C2H4
    const double2 a0 = A0_109, a1 = A1_109;
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
    const double2 a0 = A0_110, a1 = A1_110;
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
    const double2 a0 = A0_111, a1 = A1_111;
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
    const double2 a0 = A0_112, a1 = A1_112;
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
  const double2 A0_113 = ld2(rk0 + 224), A1_113 = ld2(rk1 + 224);
  const double2 A0_114 = ld2(rk0 + 226), A1_114 = ld2(rk1 + 226);
  const double2 A0_115 = ld2(rk0 + 228), A1_115 = ld2(rk1 + 228);
  const double2 A0_116 = ld2(rk0 + 230), A1_116 = ld2(rk1 + 230);
  {  // reaction 125: C2H2 + O2 => 2 HCO
    const double2 a0 = A0_113, a1 = A1_113;
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
    const double2 a0 = A0_114, a1 = A1_114;
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
    const double2 a0 = A0_115, a1 = A1_115;
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
    const double2 a0 = A0_116, a1 = A1_116;
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
  const double2 A0_117 = ld2(rk0 + 232), A1_117 = ld2(rk1 + 232);
  const double2 A0_118 = ld2(rk0 + 234), A1_118 = ld2(rk1 + 234);
  const double2 A0_119 = ld2(rk0 + 236), A1_119 = ld2(rk1 + 236);
  const double2 A0_120 = ld2(rk0 + 238), A1_120 = ld2(rk1 + 238);
  {  // reaction 129: C2H3 + O2 => C2H2 + HO2
    const double2 a0 = A0_117, a1 = A1_117;
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
    const double2 a0 = A0_118, a1 = A1_118;
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
    const double2 a0 = A0_119, a1 = A1_119;
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
    const double2 a0 = A0_120, a1 = A1_120;
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
  const double2 A0_121 = ld2(rk0 + 240), A1_121 = ld2(rk1 + 240);
  const double2 A0_122 = ld2(rk0 + 242), A1_122 = ld2(rk1 + 242);
  const double2 A0_123 = ld2(rk0 + 244), A1_123 = ld2(rk1 + 244);
  const double2 A0_124 = ld2(rk0 + 246), A1_124 = ld2(rk1 + 246);
  {  // reaction 133: CH2CO + OH => CH2O + HCO
    const double2 a0 = A0_121, a1 = A1_121;
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
    const double2 a0 = A0_122, a1 = A1_122;
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
    const double2 a0 = A0_123, a1 = A1_123;
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
    const double2 a0 = A0_124, a1 = A1_124;
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
  const double2 A0_125 = ld2(rk0 + 248), A1_125 = ld2(rk1 + 248);
  const double2 A0_126 = ld2(rk0 + 250), A1_126 = ld2(rk1 + 250);
  const double2 A0_127 = ld2(rk0 + 252), A1_127 = ld2(rk1 + 252);
  const double2 A0_128 = ld2(rk0 + 254), A1_128 = ld2(rk1 + 254);
  {  // reaction 137: C2H6 + HO2 => C2H5 + H2O2
    const double2 a0 = A0_125, a1 = A1_125;
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
    const double2 a0 = A0_126, a1 = A1_126;
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
    const double2 a0 = A0_127, a1 = A1_127;
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
    const double2 a0 = A0_128, a1 = A1_128;
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
  const double2 A0_129 = ld2(rk0 + 256), A1_129 = ld2(rk1 + 256);
  const double2 A0_130 = ld2(rk0 + 258), A1_130 = ld2(rk1 + 258);
  const double2 A0_131 = ld2(rk0 + 260), A1_131 = ld2(rk1 + 260);
  const double2 A0_132 = ld2(rk0 + 262), A1_132 = ld2(rk1 + 262);
  {  // reaction 141: C2H3 + CH3 => C2H2 + CH4
    const double2 a0 = A0_129, a1 = A1_129;
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
    const double2 a0 = A0_130, a1 = A1_130;
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
    const double2 a0 = A0_131, a1 = A1_131;
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
    const double2 a0 = A0_132, a1 = A1_132;
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
  const double2 A0_133 = ld2(rk0 + 264), A1_133 = ld2(rk1 + 264);
  const double2 A0_134 = ld2(rk0 + 266), A1_134 = ld2(rk1 + 266);
  const double2 A0_135 = ld2(rk0 + 268), A1_135 = ld2(rk1 + 268);
  const double2 A0_136 = ld2(rk0 + 270), A1_136 = ld2(rk1 + 270);
  {  // reaction 145: CH3CO + M => CH3 + CO + M
    const double2 a0 = A0_133, a1 = A1_133;
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
    const double2 a0 = A0_134, a1 = A1_134;
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
    const double2 a0 = A0_135, a1 = A1_135;
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
    const double2 a0 = A0_136, a1 = A1_136;
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
  const double2 A0_137 = ld2(rk0 + 272), A1_137 = ld2(rk1 + 272);
  const double2 A0_138 = ld2(rk0 + 274), A1_138 = ld2(rk1 + 274);
  const double2 A0_139 = ld2(rk0 + 276), A1_139 = ld2(rk1 + 276);
  const double2 A0_140 = ld2(rk0 + 278), A1_140 = ld2(rk1 + 278);
  {  // reaction 149: C2H4 + HO2 => C2H3 + H2O2
    const double2 a0 = A0_137, a1 = A1_137;
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
    const double2 a0 = A0_138, a1 = A1_138;
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
    const double2 a0 = A0_139, a1 = A1_139;
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
    const double2 a0 = A0_140, a1 = A1_140;
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
  const double2 A0_141 = ld2(rk0 + 280), A1_141 = ld2(rk1 + 280);
  const double2 A0_142 = ld2(rk0 + 282), A1_142 = ld2(rk1 + 282);
  const double2 A0_143 = ld2(rk0 + 284), A1_143 = ld2(rk1 + 284);
  const double2 A0_144 = ld2(rk0 + 286), A1_144 = ld2(rk1 + 286);
  {  // reaction 153: CH3CO + O => CH2CO + OH
    const double2 a0 = A0_141, a1 = A1_141;
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
    const double2 a0 = A0_142, a1 = A1_142;
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
    const double2 a0 = A0_143, a1 = A1_143;
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
    const double2 a0 = A0_144, a1 = A1_144;
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
  const double2 A0_145 = ld2(rk0 + 288), A1_145 = ld2(rk1 + 288);
  const double2 A0_146 = ld2(rk0 + 290), A1_146 = ld2(rk1 + 290);
  const double2 A0_147 = ld2(rk0 + 292), A1_147 = ld2(rk1 + 292);
  const double2 A0_148 = ld2(rk0 + 294), A1_148 = ld2(rk1 + 294);
  {  // reaction 157: C2H4 + O => CH2CHO + H
    const double2 a0 = A0_145, a1 = A1_145;
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
    const double2 a0 = A0_146, a1 = A1_146;
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
    const double2 a0 = A0_147, a1 = A1_147;
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
    const double2 a0 = A0_148, a1 = A1_148;
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
  const double2 A0_149 = ld2(rk0 + 296), A1_149 = ld2(rk1 + 296);
  const double2 A0_150 = ld2(rk0 + 298), A1_150 = ld2(rk1 + 298);
  const double2 A0_151 = ld2(rk0 + 300), A1_151 = ld2(rk1 + 300);
  const double2 A0_152 = ld2(rk0 + 302), A1_152 = ld2(rk1 + 302);
  {  // reaction 161: C2H6 + CH2 => C2H5 + CH3
    const double2 a0 = A0_149, a1 = A1_149;
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
    const double2 a0 = A0_150, a1 = A1_150;
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
    const double2 a0 = A0_151, a1 = A1_151;
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
    const double2 a0 = A0_152, a1 = A1_152;
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
  const double2 A0_153 = ld2(rk0 + 304), A1_153 = ld2(rk1 + 304);
  const double2 A0_154 = ld2(rk0 + 306), A1_154 = ld2(rk1 + 306);
  const double2 A0_155 = ld2(rk0 + 308), A1_155 = ld2(rk1 + 308);
  const double2 A0_156 = ld2(rk0 + 310), A1_156 = ld2(rk1 + 310);
  {  // reaction 165: C2H4O1-2 => CH3 + HCO
    const double2 a0 = A0_153, a1 = A1_153;
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
    const double2 a0 = A0_154, a1 = A1_154;
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
    const double2 a0 = A0_155, a1 = A1_155;
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
    const double2 a0 = A0_156, a1 = A1_156;
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
  const double2 A0_157 = ld2(rk0 + 312), A1_157 = ld2(rk1 + 312);
  const double2 A0_158 = ld2(rk0 + 314), A1_158 = ld2(rk1 + 314);
  const double2 A0_159 = ld2(rk0 + 316), A1_159 = ld2(rk1 + 316);
  const double2 A0_160 = ld2(rk0 + 318), A1_160 = ld2(rk1 + 318);
  {  // reaction 169: C2H4O1-2 + H => C2H3O1-2 + H2
    const double2 a0 = A0_157, a1 = A1_157;
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
    const double2 a0 = A0_158, a1 = A1_158;
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
    const double2 a0 = A0_159, a1 = A1_159;
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
    const double2 a0 = A0_160, a1 = A1_160;
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
  const double2 A0_161 = ld2(rk0 + 320), A1_161 = ld2(rk1 + 320);
  const double2 A0_162 = ld2(rk0 + 322), A1_162 = ld2(rk1 + 322);
  const double2 A0_163 = ld2(rk0 + 324), A1_163 = ld2(rk1 + 324);
  const double2 A0_164 = ld2(rk0 + 326), A1_164 = ld2(rk1 + 326);
  {  // reaction 173: C2H4O1-2 + CH3 => C2H3O1-2 + CH4
    const double2 a0 = A0_161, a1 = A1_161;
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
    const double2 a0 = A0_162, a1 = A1_162;
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
    const double2 a0 = A0_163, a1 = A1_163;
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
    const double2 a0 = A0_164, a1 = A1_164;
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
  const double2 A0_165 = ld2(rk0 + 328), A1_165 = ld2(rk1 + 328);
  const double2 A0_166 = ld2(rk0 + 330), A1_166 = ld2(rk1 + 330);
  const double2 A0_167 = ld2(rk0 + 332), A1_167 = ld2(rk1 + 332);
  const double2 A0_168 = ld2(rk0 + 334), A1_168 = ld2(rk1 + 334);
  {  // reaction 177: C3H6O1-3 + OH => C2H3 + CH2O + H2O
    const double2 a0 = A0_165, a1 = A1_165;
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
    const double2 a0 = A0_166, a1 = A1_166;
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
    const double2 a0 = A0_167, a1 = A1_167;
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
    const double2 a0 = A0_168, a1 = A1_168;
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
  const double2 A0_169 = ld2(rk0 + 336), A1_169 = ld2(rk1 + 336);
  const double2 A0_170 = ld2(rk0 + 338), A1_170 = ld2(rk1 + 338);
  const double2 A0_171 = ld2(rk0 + 340), A1_171 = ld2(rk1 + 340);
  const double2 A0_172 = ld2(rk0 + 342), A1_172 = ld2(rk1 + 342);
  {  // reaction 181: C3H6O1-3 + CH3 => C2H3 + CH2O + CH4
    const double2 a0 = A0_169, a1 = A1_169;
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
    const double2 a0 = A0_170, a1 = A1_170;
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
    const double2 a0 = A0_171, a1 = A1_171;
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
    const double2 a0 = A0_172, a1 = A1_172;
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
  const double2 A0_173 = ld2(rk0 + 344), A1_173 = ld2(rk1 + 344);
  const double2 A0_174 = ld2(rk0 + 346), A1_174 = ld2(rk1 + 346);
  const double2 A0_175 = ld2(rk0 + 348), A1_175 = ld2(rk1 + 348);
  const double2 A0_176 = ld2(rk0 + 350), A1_176 = ld2(rk1 + 350);
  {  // reaction 185: CH3CO => C2H3O1-2
    const double2 a0 = A0_173, a1 = A1_173;
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
    const double2 a0 = A0_174, a1 = A1_174;
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
    const double2 a0 = A0_175, a1 = A1_175;
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
    const double2 a0 = A0_176, a1 = A1_176;
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
  const double2 A0_177 = ld2(rk0 + 352), A1_177 = ld2(rk1 + 352);
  const double2 A0_178 = ld2(rk0 + 354), A1_178 = ld2(rk1 + 354);
  const double2 A0_179 = ld2(rk0 + 356), A1_179 = ld2(rk1 + 356);
  const double2 A0_180 = ld2(rk0 + 358), A1_180 = ld2(rk1 + 358);
  {  // reaction 189: CH2CO + H => CH2CHO
    const double2 a0 = A0_177, a1 = A1_177;
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
    const double2 a0 = A0_178, a1 = A1_178;
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
    const double2 a0 = A0_179, a1 = A1_179;
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
    const double2 a0 = A0_180, a1 = A1_180;
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
  const double2 A0_181 = ld2(rk0 + 360), A1_181 = ld2(rk1 + 360);
  const double2 A0_182 = ld2(rk0 + 362), A1_182 = ld2(rk1 + 362);
  const double2 A0_183 = ld2(rk0 + 364), A1_183 = ld2(rk1 + 364);
  const double2 A0_184 = ld2(rk0 + 366), A1_184 = ld2(rk1 + 366);
  {  // reaction 193: CH3 + CO2 => CH2CO + OH
    const double2 a0 = A0_181, a1 = A1_181;
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
    const double2 a0 = A0_182, a1 = A1_182;
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
    const double2 a0 = A0_183, a1 = A1_183;
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
    const double2 a0 = A0_184, a1 = A1_184;
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
  const double2 A0_185 = ld2(rk0 + 368), A1_185 = ld2(rk1 + 368);
  {  // reaction 197: C2H4O1-2 + OH => C2H4 + HO2
    const double2 a0 = A0_185, a1 = A1_185;
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
    const double2 A0_186 = ld2(rk0 + 362), A1_186 = ld2(rk1 + 362);
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = A0_186, a1 = A1_186;
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
    const double2 A0_187 = ld2(rk0 + 290), A1_187 = ld2(rk1 + 290);
    const double2 A0_188 = ld2(rk0 + 348), A1_188 = ld2(rk1 + 348);
    const double2 A0_189 = ld2(rk0 + 350), A1_189 = ld2(rk1 + 350);
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = A0_187, a1 = A1_187;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c2;
      jc1 += -1.0 * dq;
      jc2 += -1.0 * dq;
      jc7 += 1.0 * dq;
      jc25 += 1.0 * dq;
    }
    { const double2 a0 = A0_188, a1 = A1_188;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 1.0;
      jc1 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    { const double2 a0 = A0_189, a1 = A1_189;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 1.0;
      jc1 += -1.0 * dq;
      jc2 += 1.0 * dq;
      jc12 += 1.0 * dq;
    }
    const double2 A0_190 = ld2(rk0 + 354), A1_190 = ld2(rk1 + 354);
    { const double2 a0 = A0_190, a1 = A1_190;
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
    const double2 A0_191 = ld2(rk0 + 2), A1_191 = ld2(rk1 + 2);
    {
      const double dq = scr[0 * SST];
      jc7 += -2.0 * dq;
      jc15 += 1.0 * dq;
    }
    { const double2 a0 = A0_191, a1 = A1_191;
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
    const double2 A0_192 = ld2(rk0 + 32), A1_192 = ld2(rk1 + 32);
    const double2 A0_193 = ld2(rk0 + 46), A1_193 = ld2(rk1 + 46);
    const double2 A0_194 = ld2(rk0 + 48), A1_194 = ld2(rk1 + 48);
    {
      const double dq = -scr[74 * SST] * c30;
      jc2 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc17 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    { const double2 a0 = A0_192, a1 = A1_192;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c12;
      jc2 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc12 += 1.0 * dq;
    }
    { const double2 a0 = A0_193, a1 = A1_193;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M3 * c15 + scr[3 * SST];
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    { const double2 a0 = A0_194, a1 = A1_194;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M4 * c15;
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    const double2 A0_195 = ld2(rk0 + 50), A1_195 = ld2(rk1 + 50);
    const double2 A0_196 = ld2(rk0 + 52), A1_196 = ld2(rk1 + 52);
    const double2 A0_197 = ld2(rk0 + 54), A1_197 = ld2(rk1 + 54);
    const double2 A0_198 = ld2(rk0 + 56), A1_198 = ld2(rk1 + 56);
    { const double2 a0 = A0_195, a1 = A1_195;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M5 * c15;
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    { const double2 a0 = A0_196, a1 = A1_196;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c15;
      jc2 += -1.0 * dq;
      jc7 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    { const double2 a0 = A0_197, a1 = A1_197;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M6 * 2.0*c2 + scr[6 * SST];
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    { const double2 a0 = A0_198, a1 = A1_198;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M7 * 2.0*c2;
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    const double2 A0_199 = ld2(rk0 + 58), A1_199 = ld2(rk1 + 58);
    const double2 A0_200 = ld2(rk0 + 60), A1_200 = ld2(rk1 + 60);
    const double2 A0_201 = ld2(rk0 + 62), A1_201 = ld2(rk1 + 62);
    const double2 A0_202 = ld2(rk0 + 64), A1_202 = ld2(rk1 + 64);
    { const double2 a0 = A0_199, a1 = A1_199;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M5 * 2.0*c2;
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    { const double2 a0 = A0_200, a1 = A1_200;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M8 * 2.0*c2;
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    { const double2 a0 = A0_201, a1 = A1_201;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M9 * c27 + scr[10 * SST];
      jc2 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_202, a1 = A1_202;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c24;
      jc2 += -1.0 * dq;
      jc7 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc24 += -1.0 * dq;
    }
    const double2 A0_203 = ld2(rk0 + 66), A1_203 = ld2(rk1 + 66);
    const double2 A0_204 = ld2(rk0 + 68), A1_204 = ld2(rk1 + 68);
    const double2 A0_205 = ld2(rk0 + 70), A1_205 = ld2(rk1 + 70);
    const double2 A0_206 = ld2(rk0 + 72), A1_206 = ld2(rk1 + 72);
    { const double2 a0 = A0_203, a1 = A1_203;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c24;
      jc2 += -1.0 * dq;
      jc15 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc24 += -1.0 * dq;
    }
    { const double2 a0 = A0_204, a1 = A1_204;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c24;
      jc2 += -1.0 * dq;
      jc24 += -1.0 * dq;
      jc27 += 2.0 * dq;
    }
    { const double2 a0 = A0_205, a1 = A1_205;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c8;
      jc2 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc24 += 1.0 * dq;
    }
    { const double2 a0 = A0_206, a1 = A1_206;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c8;
      jc2 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += 1.0 * dq;
    }
    const double2 A0_207 = ld2(rk0 + 74), A1_207 = ld2(rk1 + 74);
    const double2 A0_208 = ld2(rk0 + 76), A1_208 = ld2(rk1 + 76);
    {
      const double dq = scr[12 * SST] * c29 + scr[11 * SST];
      jc2 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = A0_207, a1 = A1_207;
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
    { const double2 a0 = A0_208, a1 = A1_208;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c3;
      jc2 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc19 += 1.0 * dq;
    }
    const double2 A0_209 = ld2(rk0 + 78), A1_209 = ld2(rk1 + 78);
    const double2 A0_210 = ld2(rk0 + 80), A1_210 = ld2(rk1 + 80);
    {
      const double dq = scr[18 * SST] * c17 + scr[17 * SST];
      jc2 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    { const double2 a0 = A0_209, a1 = A1_209;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c17;
      jc2 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
      jc19 += 1.0 * dq;
    }
    { const double2 a0 = A0_210, a1 = A1_210;
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
    const double2 A0_211 = ld2(rk0 + 82), A1_211 = ld2(rk1 + 82);
    const double2 A0_212 = ld2(rk0 + 84), A1_212 = ld2(rk1 + 84);
    {
      const double dq = scr[24 * SST] * c6 + scr[23 * SST];
      jc2 += -1.0 * dq;
      jc6 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    { const double2 a0 = A0_211, a1 = A1_211;
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
    { const double2 a0 = A0_212, a1 = A1_212;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c25;
      jc2 += -1.0 * dq;
      jc6 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    const double2 A0_213 = ld2(rk0 + 86), A1_213 = ld2(rk1 + 86);
    const double2 A0_214 = ld2(rk0 + 88), A1_214 = ld2(rk1 + 88);
    const double2 A0_215 = ld2(rk0 + 90), A1_215 = ld2(rk1 + 90);
    {
      const double dq = scr[30 * SST] * c28 + scr[29 * SST];
      jc2 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    { const double2 a0 = A0_213, a1 = A1_213;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c28;
      jc2 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    { const double2 a0 = A0_214, a1 = A1_214;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c11;
      jc2 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc28 += 1.0 * dq;
    }
    { const double2 a0 = A0_215, a1 = A1_215;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c12;
      jc2 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc12 += -1.0 * dq;
      jc18 += 1.0 * dq;
    }
    const double2 A0_216 = ld2(rk0 + 92), A1_216 = ld2(rk1 + 92);
    const double2 A0_217 = ld2(rk0 + 102), A1_217 = ld2(rk1 + 102);
    {
      const double dq = scr[32 * SST];
      jc9 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    { const double2 a0 = A0_216, a1 = A1_216;
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
    { const double2 a0 = A0_217, a1 = A1_217;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c17;
      jc2 += 1.0 * dq;
      jc17 += 1.0 * dq;
      jc26 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    const double2 A0_218 = ld2(rk0 + 104), A1_218 = ld2(rk1 + 104);
    const double2 A0_219 = ld2(rk0 + 108), A1_219 = ld2(rk1 + 108);
    const double2 A0_220 = ld2(rk0 + 116), A1_220 = ld2(rk1 + 116);
    const double2 A0_221 = ld2(rk0 + 122), A1_221 = ld2(rk1 + 122);
    { const double2 a0 = A0_218, a1 = A1_218;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c10;
      jc2 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc27 += -1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = A0_219, a1 = A1_219;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c10;
      jc2 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc16 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_220, a1 = A1_220;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c30;
      jc2 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc27 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    { const double2 a0 = A0_221, a1 = A1_221;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c12;
      jc2 += 1.0 * dq;
      jc4 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    const double2 A0_222 = ld2(rk0 + 146), A1_222 = ld2(rk1 + 146);
    const double2 A0_223 = ld2(rk0 + 148), A1_223 = ld2(rk1 + 148);
    const double2 A0_224 = ld2(rk0 + 150), A1_224 = ld2(rk1 + 150);
    const double2 A0_225 = ld2(rk0 + 152), A1_225 = ld2(rk1 + 152);
    { const double2 a0 = A0_222, a1 = A1_222;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c29;
      jc2 += 1.0 * dq;
      jc19 += -1.0 * dq;
      jc26 += -1.0 * dq;
      jc29 += 1.0 * dq;
    }
    { const double2 a0 = A0_223, a1 = A1_223;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c10;
      jc2 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc14 += -1.0 * dq;
      jc26 += -1.0 * dq;
    }
    { const double2 a0 = A0_224, a1 = A1_224;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c4;
      jc2 += 1.0 * dq;
      jc4 += 1.0 * dq;
      jc26 += -1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = A0_225, a1 = A1_225;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c6;
      jc2 += 1.0 * dq;
      jc6 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc26 += -1.0 * dq;
    }
    const double2 A0_226 = ld2(rk0 + 154), A1_226 = ld2(rk1 + 154);
    const double2 A0_227 = ld2(rk0 + 158), A1_227 = ld2(rk1 + 158);
    const double2 A0_228 = ld2(rk0 + 162), A1_228 = ld2(rk1 + 162);
    const double2 A0_229 = ld2(rk0 + 166), A1_229 = ld2(rk1 + 166);
    { const double2 a0 = A0_226, a1 = A1_226;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c25;
      jc2 += 1.0 * dq;
      jc3 += -1.0 * dq;
      jc25 += 1.0 * dq;
      jc26 += -1.0 * dq;
    }
    { const double2 a0 = A0_227, a1 = A1_227;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c12;
      jc2 += 1.0 * dq;
      jc10 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc26 += -1.0 * dq;
    }
    { const double2 a0 = A0_228, a1 = A1_228;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c18;
      jc2 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc19 += -1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = A0_229, a1 = A1_229;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c25;
      jc2 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc25 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    const double2 A0_230 = ld2(rk0 + 174), A1_230 = ld2(rk1 + 174);
    const double2 A0_231 = ld2(rk0 + 178), A1_231 = ld2(rk1 + 178);
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
    { const double2 a0 = A0_230, a1 = A1_230;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c18;
      jc2 += 1.0 * dq;
      jc16 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    { const double2 a0 = A0_231, a1 = A1_231;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c25;
      jc2 += 1.0 * dq;
      jc16 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    const double2 A0_232 = ld2(rk0 + 194), A1_232 = ld2(rk1 + 194);
    const double2 A0_233 = ld2(rk0 + 204), A1_233 = ld2(rk1 + 204);
    const double2 A0_234 = ld2(rk0 + 206), A1_234 = ld2(rk1 + 206);
    {
      const double dq = scr[41 * SST];
      jc11 += 1.0 * dq;
      jc18 += -2.0 * dq;
    }
    { const double2 a0 = A0_232, a1 = A1_232;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c28;
      jc2 += 1.0 * dq;
      jc18 += -2.0 * dq;
      jc28 += 1.0 * dq;
    }
    { const double2 a0 = A0_233, a1 = A1_233;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M5 * c9;
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    { const double2 a0 = A0_234, a1 = A1_234;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M10 * c9 + scr[45 * SST];
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    const double2 A0_235 = ld2(rk0 + 290), A1_235 = ld2(rk1 + 290);
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
    { const double2 a0 = A0_235, a1 = A1_235;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c1;
      jc1 += -1.0 * dq;
      jc2 += -1.0 * dq;
      jc7 += 1.0 * dq;
      jc25 += 1.0 * dq;
    }
    const double2 A0_236 = ld2(rk0 + 312), A1_236 = ld2(rk1 + 312);
    const double2 A0_237 = ld2(rk0 + 332), A1_237 = ld2(rk1 + 332);
    const double2 A0_238 = ld2(rk0 + 352), A1_238 = ld2(rk1 + 352);
    { const double2 a0 = A0_236, a1 = A1_236;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c22;
      jc2 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    { const double2 a0 = A0_237, a1 = A1_237;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c21;
      jc2 += -1.0 * dq;
      jc6 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc21 += -1.0 * dq;
    }
    { const double2 a0 = A0_238, a1 = A1_238;
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
    const double2 A0_239 = ld2(rk0 + 76), A1_239 = ld2(rk1 + 76);
    {
      const double dq = -scr[16 * SST] * 1.0 + 2.0 * scr[14 * SST];
      jc2 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_239, a1 = A1_239;
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
    const double2 A0_240 = ld2(rk0 + 114), A1_240 = ld2(rk1 + 114);
    const double2 A0_241 = ld2(rk0 + 138), A1_241 = ld2(rk1 + 138);
    const double2 A0_242 = ld2(rk0 + 154), A1_242 = ld2(rk1 + 154);
    {
      const double dq = 2.0 * scr[35 * SST];
      jc8 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    { const double2 a0 = A0_240, a1 = A1_240;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc3 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_241, a1 = A1_241;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c15;
      jc3 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc24 += -1.0 * dq;
    }
    { const double2 a0 = A0_242, a1 = A1_242;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c26;
      jc2 += 1.0 * dq;
      jc3 += -1.0 * dq;
      jc25 += 1.0 * dq;
      jc26 += -1.0 * dq;
    }
    const double2 A0_243 = ld2(rk0 + 168), A1_243 = ld2(rk1 + 168);
    const double2 A0_244 = ld2(rk0 + 180), A1_244 = ld2(rk1 + 180);
    const double2 A0_245 = ld2(rk0 + 192), A1_245 = ld2(rk1 + 192);
    { const double2 a0 = A0_243, a1 = A1_243;
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
    { const double2 a0 = A0_244, a1 = A1_244;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c16;
      jc3 += -1.0 * dq;
      jc16 += -1.0 * dq;
      jc18 += 2.0 * dq;
    }
    { const double2 a0 = A0_245, a1 = A1_245;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c24;
      jc3 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    const double2 A0_246 = ld2(rk0 + 196), A1_246 = ld2(rk1 + 196);
    const double2 A0_247 = ld2(rk0 + 198), A1_247 = ld2(rk1 + 198);
    const double2 A0_248 = ld2(rk0 + 200), A1_248 = ld2(rk1 + 200);
    {
      const double dq = 2.0 * scr[41 * SST];
      jc11 += 1.0 * dq;
      jc18 += -2.0 * dq;
    }
    { const double2 a0 = A0_246, a1 = A1_246;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c9;
      jc3 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_247, a1 = A1_247;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c17;
      jc3 += 1.0 * dq;
      jc10 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_248, a1 = A1_248;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c6;
      jc3 += 1.0 * dq;
      jc6 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc25 += -1.0 * dq;
    }
    const double2 A0_249 = ld2(rk0 + 202), A1_249 = ld2(rk1 + 202);
    const double2 A0_250 = ld2(rk0 + 258), A1_250 = ld2(rk1 + 258);
    { const double2 a0 = A0_249, a1 = A1_249;
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
    { const double2 a0 = A0_250, a1 = A1_250;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c4;
      jc3 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
      jc18 += 1.0 * dq;
    }
    const double2 A0_251 = ld2(rk0 + 262), A1_251 = ld2(rk1 + 262);
    const double2 A0_252 = ld2(rk0 + 286), A1_252 = ld2(rk1 + 286);
    { const double2 a0 = A0_251, a1 = A1_251;
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
    { const double2 a0 = A0_252, a1 = A1_252;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c12;
      jc3 += -1.0 * dq;
      jc12 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc20 += 1.0 * dq;
    }
    const double2 A0_253 = ld2(rk0 + 322), A1_253 = ld2(rk1 + 322);
    { const double2 a0 = A0_253, a1 = A1_253;
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
    const double2 A0_254 = ld2(rk0 + 82), A1_254 = ld2(rk1 + 82);
    { const double2 a0 = A0_254, a1 = A1_254;
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
    const double2 A0_255 = ld2(rk0 + 122), A1_255 = ld2(rk1 + 122);
    const double2 A0_256 = ld2(rk0 + 124), A1_256 = ld2(rk1 + 124);
    const double2 A0_257 = ld2(rk0 + 126), A1_257 = ld2(rk1 + 126);
    {
      const double dq = scr[35 * SST];
      jc8 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    { const double2 a0 = A0_255, a1 = A1_255;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc2 += 1.0 * dq;
      jc4 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_256, a1 = A1_256;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc4 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_257, a1 = A1_257;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c14;
      jc4 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    const double2 A0_258 = ld2(rk0 + 150), A1_258 = ld2(rk1 + 150);
    const double2 A0_259 = ld2(rk0 + 164), A1_259 = ld2(rk1 + 164);
    { const double2 a0 = A0_258, a1 = A1_258;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc4 += 1.0 * dq;
      jc26 += -1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = A0_259, a1 = A1_259;
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
    const double2 A0_260 = ld2(rk0 + 224), A1_260 = ld2(rk1 + 224);
    const double2 A0_261 = ld2(rk0 + 234), A1_261 = ld2(rk1 + 234);
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
    { const double2 a0 = A0_260, a1 = A1_260;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c15;
      jc4 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc17 += 2.0 * dq;
    }
    { const double2 a0 = A0_261, a1 = A1_261;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c24;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -1.0 * dq;
    }
    const double2 A0_262 = ld2(rk0 + 258), A1_262 = ld2(rk1 + 258);
    { const double2 a0 = A0_262, a1 = A1_262;
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
    const double2 A0_263 = ld2(rk0 + 338), A1_263 = ld2(rk1 + 338);
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = A0_263, a1 = A1_263;
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
    const double2 A0_264 = ld2(rk0 + 32), A1_264 = ld2(rk1 + 32);
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
    { const double2 a0 = A0_264, a1 = A1_264;
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
    const double2 A0_265 = ld2(rk0 + 82), A1_265 = ld2(rk1 + 82);
    const double2 A0_266 = ld2(rk0 + 84), A1_266 = ld2(rk1 + 84);
    { const double2 a0 = A0_265, a1 = A1_265;
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
    { const double2 a0 = A0_266, a1 = A1_266;
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
    const double2 A0_267 = ld2(rk0 + 126), A1_267 = ld2(rk1 + 126);
    const double2 A0_268 = ld2(rk0 + 128), A1_268 = ld2(rk1 + 128);
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
    { const double2 a0 = A0_267, a1 = A1_267;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc4 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_268, a1 = A1_268;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c14;
      jc6 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc25 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    const double2 A0_269 = ld2(rk0 + 152), A1_269 = ld2(rk1 + 152);
    const double2 A0_270 = ld2(rk0 + 200), A1_270 = ld2(rk1 + 200);
    { const double2 a0 = A0_269, a1 = A1_269;
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
    { const double2 a0 = A0_270, a1 = A1_270;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c3;
      jc3 += 1.0 * dq;
      jc6 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc25 += -1.0 * dq;
    }
    const double2 A0_271 = ld2(rk0 + 210), A1_271 = ld2(rk1 + 210);
    const double2 A0_272 = ld2(rk0 + 216), A1_272 = ld2(rk1 + 216);
    {
      const double dq = scr[45 * SST];
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    { const double2 a0 = A0_271, a1 = A1_271;
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
    { const double2 a0 = A0_272, a1 = A1_272;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c28;
      jc6 += -1.0 * dq;
      jc25 += 2.0 * dq;
      jc28 += -1.0 * dq;
    }
    const double2 A0_273 = ld2(rk0 + 222), A1_273 = ld2(rk1 + 222);
    const double2 A0_274 = ld2(rk0 + 232), A1_274 = ld2(rk1 + 232);
    const double2 A0_275 = ld2(rk0 + 256), A1_275 = ld2(rk1 + 256);
    { const double2 a0 = A0_273, a1 = A1_273;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc6 += -1.0 * dq;
      jc7 += 1.0 * dq;
      jc25 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_274, a1 = A1_274;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c15;
      jc4 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    { const double2 a0 = A0_275, a1 = A1_275;
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
    const double2 A0_276 = ld2(rk0 + 270), A1_276 = ld2(rk1 + 270);
    const double2 A0_277 = ld2(rk0 + 274), A1_277 = ld2(rk1 + 274);
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = A0_276, a1 = A1_276;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c24;
      jc6 += -1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    { const double2 a0 = A0_277, a1 = A1_277;
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
    const double2 A0_278 = ld2(rk0 + 0), A1_278 = ld2(rk1 + 0);
    const double2 A0_279 = ld2(rk0 + 2), A1_279 = ld2(rk1 + 2);
    { const double2 a0 = A0_278, a1 = A1_278;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M0 * 2.0*c7 + scr[0 * SST];
      jc7 += -2.0 * dq;
      jc15 += 1.0 * dq;
    }
    { const double2 a0 = A0_279, a1 = A1_279;
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
    const double2 A0_280 = ld2(rk0 + 22), A1_280 = ld2(rk1 + 22);
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
    { const double2 a0 = A0_280, a1 = A1_280;
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
    const double2 A0_281 = ld2(rk0 + 32), A1_281 = ld2(rk1 + 32);
    const double2 A0_282 = ld2(rk0 + 34), A1_282 = ld2(rk1 + 34);
    const double2 A0_283 = ld2(rk0 + 36), A1_283 = ld2(rk1 + 36);
    const double2 A0_284 = ld2(rk0 + 38), A1_284 = ld2(rk1 + 38);
    { const double2 a0 = A0_281, a1 = A1_281;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c6;
      jc2 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc12 += 1.0 * dq;
    }
    { const double2 a0 = A0_282, a1 = A1_282;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c25;
      jc7 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = A0_283, a1 = A1_283;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c28;
      jc7 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    { const double2 a0 = A0_284, a1 = A1_284;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c11;
      jc7 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc27 += 1.0 * dq;
      jc28 += 1.0 * dq;
    }
    const double2 A0_285 = ld2(rk0 + 40), A1_285 = ld2(rk1 + 40);
    const double2 A0_286 = ld2(rk0 + 42), A1_286 = ld2(rk1 + 42);
    const double2 A0_287 = ld2(rk0 + 52), A1_287 = ld2(rk1 + 52);
    { const double2 a0 = A0_285, a1 = A1_285;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c12;
      jc7 += -1.0 * dq;
      jc12 += -1.0 * dq;
      jc29 += 1.0 * dq;
      jc30 += 1.0 * dq;
    }
    { const double2 a0 = A0_286, a1 = A1_286;
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
    { const double2 a0 = A0_287, a1 = A1_287;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c27;
      jc2 += -1.0 * dq;
      jc7 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    const double2 A0_288 = ld2(rk0 + 64), A1_288 = ld2(rk1 + 64);
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
    { const double2 a0 = A0_288, a1 = A1_288;
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
    const double2 A0_289 = ld2(rk0 + 94), A1_289 = ld2(rk1 + 94);
    const double2 A0_290 = ld2(rk0 + 144), A1_290 = ld2(rk1 + 144);
    { const double2 a0 = A0_289, a1 = A1_289;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c14;
      jc7 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    { const double2 a0 = A0_290, a1 = A1_290;
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
    const double2 A0_291 = ld2(rk0 + 220), A1_291 = ld2(rk1 + 220);
    const double2 A0_292 = ld2(rk0 + 236), A1_292 = ld2(rk1 + 236);
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
    { const double2 a0 = A0_291, a1 = A1_291;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c25;
      jc6 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    { const double2 a0 = A0_292, a1 = A1_292;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c12;
      jc7 += -1.0 * dq;
      jc12 += -1.0 * dq;
      jc17 += 2.0 * dq;
    }
    const double2 A0_293 = ld2(rk0 + 280), A1_293 = ld2(rk1 + 280);
    const double2 A0_294 = ld2(rk0 + 288), A1_294 = ld2(rk1 + 288);
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
    { const double2 a0 = A0_293, a1 = A1_293;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c20;
      jc7 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc20 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    { const double2 a0 = A0_294, a1 = A1_294;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c25;
      jc1 += 1.0 * dq;
      jc2 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc25 += -1.0 * dq;
    }
    const double2 A0_295 = ld2(rk0 + 330), A1_295 = ld2(rk1 + 330);
    { const double2 a0 = A0_295, a1 = A1_295;
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
    const double2 A0_296 = ld2(rk0 + 70), A1_296 = ld2(rk1 + 70);
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
    { const double2 a0 = A0_296, a1 = A1_296;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c2;
      jc2 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc24 += 1.0 * dq;
    }
    const double2 A0_297 = ld2(rk0 + 72), A1_297 = ld2(rk1 + 72);
    { const double2 a0 = A0_297, a1 = A1_297;
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
    const double2 A0_298 = ld2(rk0 + 98), A1_298 = ld2(rk1 + 98);
    const double2 A0_299 = ld2(rk0 + 100), A1_299 = ld2(rk1 + 100);
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
    { const double2 a0 = A0_298, a1 = A1_298;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc8 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc24 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_299, a1 = A1_299;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc8 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc24 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    const double2 A0_300 = ld2(rk0 + 132), A1_300 = ld2(rk1 + 132);
    const double2 A0_301 = ld2(rk0 + 134), A1_301 = ld2(rk1 + 134);
    const double2 A0_302 = ld2(rk0 + 142), A1_302 = ld2(rk1 + 142);
    { const double2 a0 = A0_300, a1 = A1_300;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c15;
      jc8 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -2.0 * dq;
    }
    { const double2 a0 = A0_301, a1 = A1_301;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c15;
      jc8 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -2.0 * dq;
    }
    { const double2 a0 = A0_302, a1 = A1_302;
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
    const double2 A0_303 = ld2(rk0 + 192), A1_303 = ld2(rk1 + 192);
    { const double2 a0 = A0_303, a1 = A1_303;
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
    const double2 A0_304 = ld2(rk0 + 250), A1_304 = ld2(rk1 + 250);
    const double2 A0_305 = ld2(rk0 + 274), A1_305 = ld2(rk1 + 274);
    { const double2 a0 = A0_304, a1 = A1_304;
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
    { const double2 a0 = A0_305, a1 = A1_305;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c6;
      jc6 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc24 += 1.0 * dq;
      jc25 += 1.0 * dq;
    }
    const double2 A0_306 = ld2(rk0 + 318), A1_306 = ld2(rk1 + 318);
    { const double2 a0 = A0_306, a1 = A1_306;
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
    const double2 A0_307 = ld2(rk0 + 22), A1_307 = ld2(rk1 + 22);
    const double2 A0_308 = ld2(rk0 + 42), A1_308 = ld2(rk1 + 42);
    { const double2 a0 = A0_307, a1 = A1_307;
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
    { const double2 a0 = A0_308, a1 = A1_308;
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
    const double2 A0_309 = ld2(rk0 + 78), A1_309 = ld2(rk1 + 78);
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
    { const double2 a0 = A0_309, a1 = A1_309;
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
    const double2 A0_310 = ld2(rk0 + 90), A1_310 = ld2(rk1 + 90);
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
    { const double2 a0 = A0_310, a1 = A1_310;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c18;
      jc2 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc12 += -1.0 * dq;
      jc18 += 1.0 * dq;
    }
    const double2 A0_311 = ld2(rk0 + 116), A1_311 = ld2(rk1 + 116);
    const double2 A0_312 = ld2(rk0 + 118), A1_312 = ld2(rk1 + 118);
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
    { const double2 a0 = A0_311, a1 = A1_311;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc2 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc27 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    { const double2 a0 = A0_312, a1 = A1_312;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c14;
      jc9 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc17 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    const double2 A0_313 = ld2(rk0 + 124), A1_313 = ld2(rk1 + 124);
    const double2 A0_314 = ld2(rk0 + 140), A1_314 = ld2(rk1 + 140);
    const double2 A0_315 = ld2(rk0 + 156), A1_315 = ld2(rk1 + 156);
    { const double2 a0 = A0_313, a1 = A1_313;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c18;
      jc4 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_314, a1 = A1_314;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c24;
      jc9 += -1.0 * dq;
      jc24 += -1.0 * dq;
      jc27 += 1.0 * dq;
      jc30 += 1.0 * dq;
    }
    { const double2 a0 = A0_315, a1 = A1_315;
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
    const double2 A0_316 = ld2(rk0 + 172), A1_316 = ld2(rk1 + 172);
    const double2 A0_317 = ld2(rk0 + 182), A1_317 = ld2(rk1 + 182);
    const double2 A0_318 = ld2(rk0 + 186), A1_318 = ld2(rk1 + 186);
    {
      const double dq = -scr[52 * SST] * c2*c27;
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc16 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    { const double2 a0 = A0_316, a1 = A1_316;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c14;
      jc9 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc16 += -1.0 * dq;
    }
    { const double2 a0 = A0_317, a1 = A1_317;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c16 - (a0.y + (a1.y - a0.y) * Td) * c29;
      jc9 += 0.0 * dq;
      jc16 += -1.0 * dq;
      jc29 += 1.0 * dq;
    }
    { const double2 a0 = A0_318, a1 = A1_318;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c10;
      jc9 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc16 += -1.0 * dq;
      jc30 += -1.0 * dq;
    }
    const double2 A0_319 = ld2(rk0 + 196), A1_319 = ld2(rk1 + 196);
    const double2 A0_320 = ld2(rk0 + 204), A1_320 = ld2(rk1 + 204);
    const double2 A0_321 = ld2(rk0 + 206), A1_321 = ld2(rk1 + 206);
    {
      const double dq = 1.5 * scr[41 * SST];
      jc11 += 1.0 * dq;
      jc18 += -2.0 * dq;
    }
    { const double2 a0 = A0_319, a1 = A1_319;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c3;
      jc3 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_320, a1 = A1_320;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M5 * c2;
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    { const double2 a0 = A0_321, a1 = A1_321;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M10 * c2 + 1.5 * scr[45 * SST];
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    const double2 A0_322 = ld2(rk0 + 208), A1_322 = ld2(rk1 + 208);
    const double2 A0_323 = ld2(rk0 + 230), A1_323 = ld2(rk1 + 230);
    { const double2 a0 = A0_322, a1 = A1_322;
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
    { const double2 a0 = A0_323, a1 = A1_323;
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
    const double2 A0_324 = ld2(rk0 + 266), A1_324 = ld2(rk1 + 266);
    const double2 A0_325 = ld2(rk0 + 364), A1_325 = ld2(rk1 + 364);
    { const double2 a0 = A0_324, a1 = A1_324;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M11 * c18 + scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = A0_325, a1 = A1_325;
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
    const double2 A0_326 = ld2(rk0 + 36), A1_326 = ld2(rk1 + 36);
    const double2 A0_327 = ld2(rk0 + 44), A1_327 = ld2(rk1 + 44);
    {
      const double dq = scr[75 * SST] * c7;
      jc7 += -1.0 * dq;
      jc10 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc27 += 1.0 * dq;
    }
    { const double2 a0 = A0_326, a1 = A1_326;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c18;
      jc7 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    { const double2 a0 = A0_327, a1 = A1_327;
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
    const double2 A0_328 = ld2(rk0 + 80), A1_328 = ld2(rk1 + 80);
    {
      const double dq = -scr[19 * SST] * 1.0 + scr[17 * SST];
      jc2 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    { const double2 a0 = A0_328, a1 = A1_328;
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
    const double2 A0_329 = ld2(rk0 + 104), A1_329 = ld2(rk1 + 104);
    const double2 A0_330 = ld2(rk0 + 108), A1_330 = ld2(rk1 + 108);
    const double2 A0_331 = ld2(rk0 + 120), A1_331 = ld2(rk1 + 120);
    const double2 A0_332 = ld2(rk0 + 136), A1_332 = ld2(rk1 + 136);
    { const double2 a0 = A0_329, a1 = A1_329;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc27 += -1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = A0_330, a1 = A1_330;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc16 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_331, a1 = A1_331;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc10 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc17 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_332, a1 = A1_332;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c27;
      jc10 += 1.0 * dq;
      jc24 += -1.0 * dq;
      jc27 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    const double2 A0_333 = ld2(rk0 + 142), A1_333 = ld2(rk1 + 142);
    const double2 A0_334 = ld2(rk0 + 148), A1_334 = ld2(rk1 + 148);
    const double2 A0_335 = ld2(rk0 + 158), A1_335 = ld2(rk1 + 158);
    { const double2 a0 = A0_333, a1 = A1_333;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c24;
      jc8 += 1.0 * dq;
      jc10 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc24 += -1.0 * dq;
    }
    { const double2 a0 = A0_334, a1 = A1_334;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc14 += -1.0 * dq;
      jc26 += -1.0 * dq;
    }
    { const double2 a0 = A0_335, a1 = A1_335;
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
    const double2 A0_336 = ld2(rk0 + 186), A1_336 = ld2(rk1 + 186);
    const double2 A0_337 = ld2(rk0 + 190), A1_337 = ld2(rk1 + 190);
    const double2 A0_338 = ld2(rk0 + 198), A1_338 = ld2(rk1 + 198);
    { const double2 a0 = A0_336, a1 = A1_336;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c9;
      jc9 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc16 += -1.0 * dq;
      jc30 += -1.0 * dq;
    }
    { const double2 a0 = A0_337, a1 = A1_337;
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
    { const double2 a0 = A0_338, a1 = A1_338;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc3 += 1.0 * dq;
      jc10 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    const double2 A0_339 = ld2(rk0 + 210), A1_339 = ld2(rk1 + 210);
    const double2 A0_340 = ld2(rk0 + 214), A1_340 = ld2(rk1 + 214);
    {
      const double dq = scr[45 * SST];
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    { const double2 a0 = A0_339, a1 = A1_339;
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
    { const double2 a0 = A0_340, a1 = A1_340;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c19;
      jc10 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    const double2 A0_341 = ld2(rk0 + 242), A1_341 = ld2(rk1 + 242);
    const double2 A0_342 = ld2(rk0 + 302), A1_342 = ld2(rk1 + 302);
    { const double2 a0 = A0_341, a1 = A1_341;
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
    { const double2 a0 = A0_342, a1 = A1_342;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc10 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc25 += 1.0 * dq;
      jc27 += 1.0 * dq;
    }
    const double2 A0_343 = ld2(rk0 + 326), A1_343 = ld2(rk1 + 326);
    { const double2 a0 = A0_343, a1 = A1_343;
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
    const double2 A0_344 = ld2(rk0 + 38), A1_344 = ld2(rk1 + 38);
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
    { const double2 a0 = A0_344, a1 = A1_344;
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
    const double2 A0_345 = ld2(rk0 + 88), A1_345 = ld2(rk1 + 88);
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
    { const double2 a0 = A0_345, a1 = A1_345;
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
    const double2 A0_346 = ld2(rk0 + 130), A1_346 = ld2(rk1 + 130);
    const double2 A0_347 = ld2(rk0 + 188), A1_347 = ld2(rk1 + 188);
    {
      const double dq = 3.0 * scr[35 * SST];
      jc8 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    { const double2 a0 = A0_346, a1 = A1_346;
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
    { const double2 a0 = A0_347, a1 = A1_347;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c16;
      jc11 += -1.0 * dq;
      jc16 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc28 += 1.0 * dq;
    }
    const double2 A0_348 = ld2(rk0 + 202), A1_348 = ld2(rk1 + 202);
    {
      const double dq = -scr[43 * SST] * 1.0 + 3.0 * scr[41 * SST];
      jc11 += 1.0 * dq;
      jc18 += -2.0 * dq;
    }
    { const double2 a0 = A0_348, a1 = A1_348;
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
    const double2 A0_349 = ld2(rk0 + 244), A1_349 = ld2(rk1 + 244);
    const double2 A0_350 = ld2(rk0 + 248), A1_350 = ld2(rk1 + 248);
    const double2 A0_351 = ld2(rk0 + 252), A1_351 = ld2(rk1 + 252);
    { const double2 a0 = A0_349, a1 = A1_349;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c15;
      jc11 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
      jc28 += 1.0 * dq;
    }
    { const double2 a0 = A0_350, a1 = A1_350;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c24;
      jc8 += 1.0 * dq;
      jc11 += -1.0 * dq;
      jc24 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    { const double2 a0 = A0_351, a1 = A1_351;
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
    const double2 A0_352 = ld2(rk0 + 292), A1_352 = ld2(rk1 + 292);
    const double2 A0_353 = ld2(rk0 + 296), A1_353 = ld2(rk1 + 296);
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = A0_352, a1 = A1_352;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c26;
      jc11 += -1.0 * dq;
      jc26 += -1.0 * dq;
      jc28 += 1.0 * dq;
      jc29 += 1.0 * dq;
    }
    { const double2 a0 = A0_353, a1 = A1_353;
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
    const double2 A0_354 = ld2(rk0 + 32), A1_354 = ld2(rk1 + 32);
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
    { const double2 a0 = A0_354, a1 = A1_354;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc12 += 1.0 * dq;
    }
    const double2 A0_355 = ld2(rk0 + 40), A1_355 = ld2(rk1 + 40);
    { const double2 a0 = A0_355, a1 = A1_355;
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
    const double2 A0_356 = ld2(rk0 + 90), A1_356 = ld2(rk1 + 90);
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
    { const double2 a0 = A0_356, a1 = A1_356;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c2;
      jc2 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc12 += -1.0 * dq;
      jc18 += 1.0 * dq;
    }
    const double2 A0_357 = ld2(rk0 + 122), A1_357 = ld2(rk1 + 122);
    const double2 A0_358 = ld2(rk0 + 158), A1_358 = ld2(rk1 + 158);
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
    { const double2 a0 = A0_357, a1 = A1_357;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc4 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_358, a1 = A1_358;
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
    const double2 A0_359 = ld2(rk0 + 236), A1_359 = ld2(rk1 + 236);
    const double2 A0_360 = ld2(rk0 + 240), A1_360 = ld2(rk1 + 240);
    { const double2 a0 = A0_359, a1 = A1_359;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c7;
      jc7 += -1.0 * dq;
      jc12 += -1.0 * dq;
      jc17 += 2.0 * dq;
    }
    { const double2 a0 = A0_360, a1 = A1_360;
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
    const double2 A0_361 = ld2(rk0 + 282), A1_361 = ld2(rk1 + 282);
    const double2 A0_362 = ld2(rk0 + 286), A1_362 = ld2(rk1 + 286);
    const double2 A0_363 = ld2(rk0 + 340), A1_363 = ld2(rk1 + 340);
    const double2 A0_364 = ld2(rk0 + 352), A1_364 = ld2(rk1 + 352);
    { const double2 a0 = A0_361, a1 = A1_361;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc7 += 1.0 * dq;
      jc12 += -1.0 * dq;
      jc20 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_362, a1 = A1_362;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c3;
      jc3 += -1.0 * dq;
      jc12 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = A0_363, a1 = A1_363;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c17;
      jc5 += 1.0 * dq;
      jc12 += -1.0 * dq;
      jc15 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    { const double2 a0 = A0_364, a1 = A1_364;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c2;
      jc1 += 1.0 * dq;
      jc2 += -1.0 * dq;
      jc12 += -1.0 * dq;
    }
    const double2 A0_365 = ld2(rk0 + 356), A1_365 = ld2(rk1 + 356);
    const double2 A0_366 = ld2(rk0 + 358), A1_366 = ld2(rk1 + 358);
    { const double2 a0 = A0_365, a1 = A1_365;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c24;
      jc1 += 1.0 * dq;
      jc12 += -1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -1.0 * dq;
    }
    { const double2 a0 = A0_366, a1 = A1_366;
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
    const double2 A0_367 = ld2(rk0 + 276), A1_367 = ld2(rk1 + 276);
    const double2 A0_368 = ld2(rk0 + 364), A1_368 = ld2(rk1 + 364);
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = A0_367, a1 = A1_367;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 1.0;
      jc13 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc25 += 1.0 * dq;
    }
    { const double2 a0 = A0_368, a1 = A1_368;
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
    const double2 A0_369 = ld2(rk0 + 62), A1_369 = ld2(rk1 + 62);
    const double2 A0_370 = ld2(rk0 + 64), A1_370 = ld2(rk1 + 64);
    const double2 A0_371 = ld2(rk0 + 72), A1_371 = ld2(rk1 + 72);
    {
      const double dq = scr[8 * SST];
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    { const double2 a0 = A0_369, a1 = A1_369;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M9 * 1.0 + 3.6500000953674316 * scr[10 * SST];
      jc2 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_370, a1 = A1_370;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c7;
      jc2 += -1.0 * dq;
      jc7 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc24 += -1.0 * dq;
    }
    { const double2 a0 = A0_371, a1 = A1_371;
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
    const double2 A0_372 = ld2(rk0 + 92), A1_372 = ld2(rk1 + 92);
    const double2 A0_373 = ld2(rk0 + 94), A1_373 = ld2(rk1 + 94);
    const double2 A0_374 = ld2(rk0 + 96), A1_374 = ld2(rk1 + 96);
    { const double2 a0 = A0_372, a1 = A1_372;
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
    { const double2 a0 = A0_373, a1 = A1_373;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c7;
      jc7 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    { const double2 a0 = A0_374, a1 = A1_374;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c15;
      jc14 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    const double2 A0_375 = ld2(rk0 + 98), A1_375 = ld2(rk1 + 98);
    const double2 A0_376 = ld2(rk0 + 100), A1_376 = ld2(rk1 + 100);
    const double2 A0_377 = ld2(rk0 + 106), A1_377 = ld2(rk1 + 106);
    const double2 A0_378 = ld2(rk0 + 110), A1_378 = ld2(rk1 + 110);
    { const double2 a0 = A0_375, a1 = A1_375;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c24;
      jc8 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc24 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_376, a1 = A1_376;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c24;
      jc8 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc24 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_377, a1 = A1_377;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c26;
      jc14 += 1.0 * dq;
      jc26 += 1.0 * dq;
      jc27 += -1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = A0_378, a1 = A1_378;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c29;
      jc14 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc27 += -1.0 * dq;
      jc29 += 1.0 * dq;
    }
    const double2 A0_379 = ld2(rk0 + 112), A1_379 = ld2(rk1 + 112);
    const double2 A0_380 = ld2(rk0 + 114), A1_380 = ld2(rk1 + 114);
    const double2 A0_381 = ld2(rk0 + 118), A1_381 = ld2(rk1 + 118);
    const double2 A0_382 = ld2(rk0 + 120), A1_382 = ld2(rk1 + 120);
    { const double2 a0 = A0_379, a1 = A1_379;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c16;
      jc14 += 1.0 * dq;
      jc16 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_380, a1 = A1_380;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c18;
      jc3 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_381, a1 = A1_381;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c9;
      jc9 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc17 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_382, a1 = A1_382;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c17;
      jc10 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc17 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    const double2 A0_383 = ld2(rk0 + 126), A1_383 = ld2(rk1 + 126);
    const double2 A0_384 = ld2(rk0 + 128), A1_384 = ld2(rk1 + 128);
    const double2 A0_385 = ld2(rk0 + 130), A1_385 = ld2(rk1 + 130);
    const double2 A0_386 = ld2(rk0 + 148), A1_386 = ld2(rk1 + 148);
    { const double2 a0 = A0_383, a1 = A1_383;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c4;
      jc4 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_384, a1 = A1_384;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c6;
      jc6 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc25 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_385, a1 = A1_385;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c28;
      jc11 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    { const double2 a0 = A0_386, a1 = A1_386;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c26;
      jc2 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc14 += -1.0 * dq;
      jc26 += -1.0 * dq;
    }
    const double2 A0_387 = ld2(rk0 + 172), A1_387 = ld2(rk1 + 172);
    const double2 A0_388 = ld2(rk0 + 176), A1_388 = ld2(rk1 + 176);
    {
      const double dq = 6.0 * scr[38 * SST];
      jc9 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = A0_387, a1 = A1_387;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c9;
      jc9 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc16 += -1.0 * dq;
    }
    { const double2 a0 = A0_388, a1 = A1_388;
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
    const double2 A0_389 = ld2(rk0 + 310), A1_389 = ld2(rk1 + 310);
    { const double2 a0 = A0_389, a1 = A1_389;
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
    const double2 A0_390 = ld2(rk0 + 0), A1_390 = ld2(rk1 + 0);
    { const double2 a0 = A0_390, a1 = A1_390;
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
    const double2 A0_391 = ld2(rk0 + 42), A1_391 = ld2(rk1 + 42);
    const double2 A0_392 = ld2(rk0 + 44), A1_392 = ld2(rk1 + 44);
    const double2 A0_393 = ld2(rk0 + 46), A1_393 = ld2(rk1 + 46);
    const double2 A0_394 = ld2(rk0 + 48), A1_394 = ld2(rk1 + 48);
    { const double2 a0 = A0_391, a1 = A1_391;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c9;
      jc7 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    { const double2 a0 = A0_392, a1 = A1_392;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c10;
      jc10 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc24 += 1.0 * dq;
    }
    { const double2 a0 = A0_393, a1 = A1_393;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M3 * c2;
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    { const double2 a0 = A0_394, a1 = A1_394;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M4 * c2 + scr[4 * SST];
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    const double2 A0_395 = ld2(rk0 + 50), A1_395 = ld2(rk1 + 50);
    const double2 A0_396 = ld2(rk0 + 52), A1_396 = ld2(rk1 + 52);
    { const double2 a0 = A0_395, a1 = A1_395;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M5 * c2;
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    { const double2 a0 = A0_396, a1 = A1_396;
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
    const double2 A0_397 = ld2(rk0 + 66), A1_397 = ld2(rk1 + 66);
    { const double2 a0 = A0_397, a1 = A1_397;
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
    const double2 A0_398 = ld2(rk0 + 96), A1_398 = ld2(rk1 + 96);
    const double2 A0_399 = ld2(rk0 + 132), A1_399 = ld2(rk1 + 132);
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
    { const double2 a0 = A0_398, a1 = A1_398;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c14;
      jc14 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_399, a1 = A1_399;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c8;
      jc8 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -2.0 * dq;
    }
    const double2 A0_400 = ld2(rk0 + 134), A1_400 = ld2(rk1 + 134);
    const double2 A0_401 = ld2(rk0 + 138), A1_401 = ld2(rk1 + 138);
    const double2 A0_402 = ld2(rk0 + 144), A1_402 = ld2(rk1 + 144);
    const double2 A0_403 = ld2(rk0 + 160), A1_403 = ld2(rk1 + 160);
    { const double2 a0 = A0_400, a1 = A1_400;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c8;
      jc8 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -2.0 * dq;
    }
    { const double2 a0 = A0_401, a1 = A1_401;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c3;
      jc3 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc24 += -1.0 * dq;
    }
    { const double2 a0 = A0_402, a1 = A1_402;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c26;
      jc7 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc26 += -1.0 * dq;
    }
    { const double2 a0 = A0_403, a1 = A1_403;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c29;
      jc15 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc27 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    const double2 A0_404 = ld2(rk0 + 172), A1_404 = ld2(rk1 + 172);
    const double2 A0_405 = ld2(rk0 + 190), A1_405 = ld2(rk1 + 190);
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
    { const double2 a0 = A0_404, a1 = A1_404;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c16;
      jc9 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc16 += -1.0 * dq;
    }
    { const double2 a0 = A0_405, a1 = A1_405;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc10 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    const double2 A0_406 = ld2(rk0 + 208), A1_406 = ld2(rk1 + 208);
    const double2 A0_407 = ld2(rk0 + 210), A1_407 = ld2(rk1 + 210);
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
    { const double2 a0 = A0_406, a1 = A1_406;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c17;
      jc9 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc17 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    { const double2 a0 = A0_407, a1 = A1_407;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c6;
      jc6 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc17 += 1.0 * dq;
    }
    const double2 A0_408 = ld2(rk0 + 212), A1_408 = ld2(rk1 + 212);
    const double2 A0_409 = ld2(rk0 + 224), A1_409 = ld2(rk1 + 224);
    const double2 A0_410 = ld2(rk0 + 228), A1_410 = ld2(rk1 + 228);
    {
      const double dq = scr[46 * SST];
      jc4 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = A0_408, a1 = A1_408;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c28;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
      jc25 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    { const double2 a0 = A0_409, a1 = A1_409;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c4;
      jc4 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc17 += 2.0 * dq;
    }
    { const double2 a0 = A0_410, a1 = A1_410;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c26;
      jc9 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc26 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    const double2 A0_411 = ld2(rk0 + 232), A1_411 = ld2(rk1 + 232);
    const double2 A0_412 = ld2(rk0 + 244), A1_412 = ld2(rk1 + 244);
    { const double2 a0 = A0_411, a1 = A1_411;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c6;
      jc4 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    { const double2 a0 = A0_412, a1 = A1_412;
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
    const double2 A0_413 = ld2(rk0 + 268), A1_413 = ld2(rk1 + 268);
    const double2 A0_414 = ld2(rk0 + 338), A1_414 = ld2(rk1 + 338);
    const double2 A0_415 = ld2(rk0 + 354), A1_415 = ld2(rk1 + 354);
    { const double2 a0 = A0_413, a1 = A1_413;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c25;
      jc6 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = A0_414, a1 = A1_414;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c5;
      jc5 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc17 += 1.0 * dq;
    }
    { const double2 a0 = A0_415, a1 = A1_415;
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
    const double2 A0_416 = ld2(rk0 + 74), A1_416 = ld2(rk1 + 74);
    {
      const double dq = scr[11 * SST];
      jc2 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = A0_416, a1 = A1_416;
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
    const double2 A0_417 = ld2(rk0 + 108), A1_417 = ld2(rk1 + 108);
    const double2 A0_418 = ld2(rk0 + 112), A1_418 = ld2(rk1 + 112);
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
    { const double2 a0 = A0_417, a1 = A1_417;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc2 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc16 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_418, a1 = A1_418;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c14;
      jc14 += 1.0 * dq;
      jc16 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    const double2 A0_419 = ld2(rk0 + 172), A1_419 = ld2(rk1 + 172);
    const double2 A0_420 = ld2(rk0 + 174), A1_420 = ld2(rk1 + 174);
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
    { const double2 a0 = A0_419, a1 = A1_419;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c15;
      jc9 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc16 += -1.0 * dq;
    }
    { const double2 a0 = A0_420, a1 = A1_420;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c19;
      jc2 += 1.0 * dq;
      jc16 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    const double2 A0_421 = ld2(rk0 + 176), A1_421 = ld2(rk1 + 176);
    const double2 A0_422 = ld2(rk0 + 178), A1_422 = ld2(rk1 + 178);
    const double2 A0_423 = ld2(rk0 + 180), A1_423 = ld2(rk1 + 180);
    const double2 A0_424 = ld2(rk0 + 182), A1_424 = ld2(rk1 + 182);
    { const double2 a0 = A0_421, a1 = A1_421;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c14;
      jc14 += 0.0 * dq;
      jc16 += -1.0 * dq;
      jc29 += 1.0 * dq;
    }
    { const double2 a0 = A0_422, a1 = A1_422;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc2 += 1.0 * dq;
      jc16 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    { const double2 a0 = A0_423, a1 = A1_423;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c3;
      jc3 += -1.0 * dq;
      jc16 += -1.0 * dq;
      jc18 += 2.0 * dq;
    }
    { const double2 a0 = A0_424, a1 = A1_424;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c9;
      jc9 += 0.0 * dq;
      jc16 += -1.0 * dq;
      jc29 += 1.0 * dq;
    }
    const double2 A0_425 = ld2(rk0 + 184), A1_425 = ld2(rk1 + 184);
    const double2 A0_426 = ld2(rk0 + 186), A1_426 = ld2(rk1 + 186);
    const double2 A0_427 = ld2(rk0 + 188), A1_427 = ld2(rk1 + 188);
    { const double2 a0 = A0_425, a1 = A1_425;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c30;
      jc16 += -1.0 * dq;
      jc29 += 1.0 * dq;
      jc30 += 0.0 * dq;
    }
    { const double2 a0 = A0_426, a1 = A1_426;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c30;
      jc9 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc16 += -1.0 * dq;
      jc30 += -1.0 * dq;
    }
    { const double2 a0 = A0_427, a1 = A1_427;
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
    const double2 A0_428 = ld2(rk0 + 34), A1_428 = ld2(rk1 + 34);
    const double2 A0_429 = ld2(rk0 + 44), A1_429 = ld2(rk1 + 44);
    { const double2 a0 = A0_428, a1 = A1_428;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c18;
      jc7 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = A0_429, a1 = A1_429;
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
    const double2 A0_430 = ld2(rk0 + 78), A1_430 = ld2(rk1 + 78);
    const double2 A0_431 = ld2(rk0 + 80), A1_431 = ld2(rk1 + 80);
    { const double2 a0 = A0_430, a1 = A1_430;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c2;
      jc2 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
      jc19 += 1.0 * dq;
    }
    { const double2 a0 = A0_431, a1 = A1_431;
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
    const double2 A0_432 = ld2(rk0 + 102), A1_432 = ld2(rk1 + 102);
    const double2 A0_433 = ld2(rk0 + 118), A1_433 = ld2(rk1 + 118);
    const double2 A0_434 = ld2(rk0 + 120), A1_434 = ld2(rk1 + 120);
    const double2 A0_435 = ld2(rk0 + 142), A1_435 = ld2(rk1 + 142);
    { const double2 a0 = A0_432, a1 = A1_432;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc17 += 1.0 * dq;
      jc26 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_433, a1 = A1_433;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc9 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc17 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_434, a1 = A1_434;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c14;
      jc10 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc17 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_435, a1 = A1_435;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c8;
      jc8 += 1.0 * dq;
      jc10 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc24 += -1.0 * dq;
    }
    const double2 A0_436 = ld2(rk0 + 144), A1_436 = ld2(rk1 + 144);
    const double2 A0_437 = ld2(rk0 + 156), A1_437 = ld2(rk1 + 156);
    const double2 A0_438 = ld2(rk0 + 160), A1_438 = ld2(rk1 + 160);
    { const double2 a0 = A0_436, a1 = A1_436;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c7;
      jc7 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc26 += -1.0 * dq;
    }
    { const double2 a0 = A0_437, a1 = A1_437;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c9;
      jc9 += 1.0 * dq;
      jc17 += 1.0 * dq;
      jc26 += -1.0 * dq;
      jc30 += -1.0 * dq;
    }
    { const double2 a0 = A0_438, a1 = A1_438;
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
    const double2 A0_439 = ld2(rk0 + 196), A1_439 = ld2(rk1 + 196);
    const double2 A0_440 = ld2(rk0 + 198), A1_440 = ld2(rk1 + 198);
    const double2 A0_441 = ld2(rk0 + 204), A1_441 = ld2(rk1 + 204);
    {
      const double dq = scr[41 * SST];
      jc11 += 1.0 * dq;
      jc18 += -2.0 * dq;
    }
    { const double2 a0 = A0_439, a1 = A1_439;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc3 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_440, a1 = A1_440;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c3;
      jc3 += 1.0 * dq;
      jc10 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_441, a1 = A1_441;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M5 * 1.0;
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    const double2 A0_442 = ld2(rk0 + 206), A1_442 = ld2(rk1 + 206);
    const double2 A0_443 = ld2(rk0 + 208), A1_443 = ld2(rk1 + 208);
    const double2 A0_444 = ld2(rk0 + 210), A1_444 = ld2(rk1 + 210);
    { const double2 a0 = A0_442, a1 = A1_442;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M10 * 1.0 + scr[45 * SST];
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    { const double2 a0 = A0_443, a1 = A1_443;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c15;
      jc9 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc17 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    { const double2 a0 = A0_444, a1 = A1_444;
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
    const double2 A0_445 = ld2(rk0 + 226), A1_445 = ld2(rk1 + 226);
    const double2 A0_446 = ld2(rk0 + 238), A1_446 = ld2(rk1 + 238);
    const double2 A0_447 = ld2(rk0 + 242), A1_447 = ld2(rk1 + 242);
    { const double2 a0 = A0_445, a1 = A1_445;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 2.0*c17;
      jc4 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc17 += -2.0 * dq;
    }
    { const double2 a0 = A0_446, a1 = A1_446;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 2.0*c17;
      jc7 += 1.0 * dq;
      jc12 += 1.0 * dq;
      jc17 += -2.0 * dq;
    }
    { const double2 a0 = A0_447, a1 = A1_447;
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
    const double2 A0_448 = ld2(rk0 + 306), A1_448 = ld2(rk1 + 306);
    const double2 A0_449 = ld2(rk0 + 340), A1_449 = ld2(rk1 + 340);
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = A0_448, a1 = A1_448;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc17 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc22 += 1.0 * dq;
    }
    { const double2 a0 = A0_449, a1 = A1_449;
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
    const double2 A0_450 = ld2(rk0 + 34), A1_450 = ld2(rk1 + 34);
    const double2 A0_451 = ld2(rk0 + 36), A1_451 = ld2(rk1 + 36);
    {
      const double dq = scr[2 * SST];
      jc7 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    { const double2 a0 = A0_450, a1 = A1_450;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c17;
      jc7 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = A0_451, a1 = A1_451;
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
    const double2 A0_452 = ld2(rk0 + 76), A1_452 = ld2(rk1 + 76);
    { const double2 a0 = A0_452, a1 = A1_452;
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
    const double2 A0_453 = ld2(rk0 + 90), A1_453 = ld2(rk1 + 90);
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
    { const double2 a0 = A0_453, a1 = A1_453;
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
    const double2 A0_454 = ld2(rk0 + 110), A1_454 = ld2(rk1 + 110);
    const double2 A0_455 = ld2(rk0 + 112), A1_455 = ld2(rk1 + 112);
    const double2 A0_456 = ld2(rk0 + 114), A1_456 = ld2(rk1 + 114);
    {
      const double dq = scr[35 * SST];
      jc8 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    { const double2 a0 = A0_454, a1 = A1_454;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc14 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc27 += -1.0 * dq;
      jc29 += 1.0 * dq;
    }
    { const double2 a0 = A0_455, a1 = A1_455;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc14 += 1.0 * dq;
      jc16 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_456, a1 = A1_456;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c14;
      jc3 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    const double2 A0_457 = ld2(rk0 + 124), A1_457 = ld2(rk1 + 124);
    const double2 A0_458 = ld2(rk0 + 138), A1_458 = ld2(rk1 + 138);
    const double2 A0_459 = ld2(rk0 + 152), A1_459 = ld2(rk1 + 152);
    const double2 A0_460 = ld2(rk0 + 162), A1_460 = ld2(rk1 + 162);
    { const double2 a0 = A0_457, a1 = A1_457;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c9;
      jc4 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_458, a1 = A1_458;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c24;
      jc3 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc24 += -1.0 * dq;
    }
    { const double2 a0 = A0_459, a1 = A1_459;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c26;
      jc2 += 1.0 * dq;
      jc6 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc26 += -1.0 * dq;
    }
    { const double2 a0 = A0_460, a1 = A1_460;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc19 += -1.0 * dq;
      jc29 += -1.0 * dq;
    }
    const double2 A0_461 = ld2(rk0 + 166), A1_461 = ld2(rk1 + 166);
    const double2 A0_462 = ld2(rk0 + 168), A1_462 = ld2(rk1 + 168);
    const double2 A0_463 = ld2(rk0 + 174), A1_463 = ld2(rk1 + 174);
    { const double2 a0 = A0_461, a1 = A1_461;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c29;
      jc2 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc25 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = A0_462, a1 = A1_462;
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
    { const double2 a0 = A0_463, a1 = A1_463;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc16 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    const double2 A0_464 = ld2(rk0 + 178), A1_464 = ld2(rk1 + 178);
    const double2 A0_465 = ld2(rk0 + 180), A1_465 = ld2(rk1 + 180);
    const double2 A0_466 = ld2(rk0 + 188), A1_466 = ld2(rk1 + 188);
    const double2 A0_467 = ld2(rk0 + 190), A1_467 = ld2(rk1 + 190);
    { const double2 a0 = A0_464, a1 = A1_464;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c16;
      jc2 += 1.0 * dq;
      jc16 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    { const double2 a0 = A0_465, a1 = A1_465;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * 2.0*c18;
      jc3 += -1.0 * dq;
      jc16 += -1.0 * dq;
      jc18 += 2.0 * dq;
    }
    { const double2 a0 = A0_466, a1 = A1_466;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c28;
      jc11 += -1.0 * dq;
      jc16 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc28 += 1.0 * dq;
    }
    { const double2 a0 = A0_467, a1 = A1_467;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c15;
      jc10 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    const double2 A0_468 = ld2(rk0 + 192), A1_468 = ld2(rk1 + 192);
    const double2 A0_469 = ld2(rk0 + 194), A1_469 = ld2(rk1 + 194);
    const double2 A0_470 = ld2(rk0 + 196), A1_470 = ld2(rk1 + 196);
    { const double2 a0 = A0_468, a1 = A1_468;
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
    { const double2 a0 = A0_469, a1 = A1_469;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 2.0*c18;
      jc2 += 1.0 * dq;
      jc18 += -2.0 * dq;
      jc28 += 1.0 * dq;
    }
    { const double2 a0 = A0_470, a1 = A1_470;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c17;
      jc3 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    const double2 A0_471 = ld2(rk0 + 198), A1_471 = ld2(rk1 + 198);
    const double2 A0_472 = ld2(rk0 + 200), A1_472 = ld2(rk1 + 200);
    const double2 A0_473 = ld2(rk0 + 202), A1_473 = ld2(rk1 + 202);
    { const double2 a0 = A0_471, a1 = A1_471;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c10;
      jc3 += 1.0 * dq;
      jc10 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_472, a1 = A1_472;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c25;
      jc3 += 1.0 * dq;
      jc6 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = A0_473, a1 = A1_473;
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
    const double2 A0_474 = ld2(rk0 + 214), A1_474 = ld2(rk1 + 214);
    const double2 A0_475 = ld2(rk0 + 256), A1_475 = ld2(rk1 + 256);
    const double2 A0_476 = ld2(rk0 + 260), A1_476 = ld2(rk1 + 260);
    {
      const double dq = scr[46 * SST];
      jc4 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = A0_474, a1 = A1_474;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc10 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_475, a1 = A1_475;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c6;
      jc3 += 1.0 * dq;
      jc4 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_476, a1 = A1_476;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c28;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc25 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    const double2 A0_477 = ld2(rk0 + 266), A1_477 = ld2(rk1 + 266);
    const double2 A0_478 = ld2(rk0 + 278), A1_478 = ld2(rk1 + 278);
    const double2 A0_479 = ld2(rk0 + 284), A1_479 = ld2(rk1 + 284);
    {
      const double dq = scr[49 * SST];
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc20 += -1.0 * dq;
    }
    { const double2 a0 = A0_477, a1 = A1_477;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M11 * c9 + scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = A0_478, a1 = A1_478;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c25;
      jc13 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = A0_479, a1 = A1_479;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c20;
      jc3 += 1.0 * dq;
      jc12 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += -1.0 * dq;
    }
    const double2 A0_480 = ld2(rk0 + 298), A1_480 = ld2(rk1 + 298);
    const double2 A0_481 = ld2(rk0 + 302), A1_481 = ld2(rk1 + 302);
    const double2 A0_482 = ld2(rk0 + 306), A1_482 = ld2(rk1 + 306);
    const double2 A0_483 = ld2(rk0 + 320), A1_483 = ld2(rk1 + 320);
    { const double2 a0 = A0_480, a1 = A1_480;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c28;
      jc11 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc28 += -1.0 * dq;
      jc29 += 1.0 * dq;
    }
    { const double2 a0 = A0_481, a1 = A1_481;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c10;
      jc10 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc25 += 1.0 * dq;
      jc27 += 1.0 * dq;
    }
    { const double2 a0 = A0_482, a1 = A1_482;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c17;
      jc17 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc22 += 1.0 * dq;
    }
    { const double2 a0 = A0_483, a1 = A1_483;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c22;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    const double2 A0_484 = ld2(rk0 + 336), A1_484 = ld2(rk1 + 336);
    const double2 A0_485 = ld2(rk0 + 360), A1_485 = ld2(rk1 + 360);
    { const double2 a0 = A0_484, a1 = A1_484;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c21;
      jc3 += 1.0 * dq;
      jc6 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc21 += -1.0 * dq;
    }
    { const double2 a0 = A0_485, a1 = A1_485;
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
    const double2 A0_486 = ld2(rk0 + 54), A1_486 = ld2(rk1 + 54);
    const double2 A0_487 = ld2(rk0 + 56), A1_487 = ld2(rk1 + 56);
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
    { const double2 a0 = A0_486, a1 = A1_486;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M6 * 1.0;
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    { const double2 a0 = A0_487, a1 = A1_487;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M7 * 1.0 + scr[7 * SST];
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    const double2 A0_488 = ld2(rk0 + 58), A1_488 = ld2(rk1 + 58);
    const double2 A0_489 = ld2(rk0 + 60), A1_489 = ld2(rk1 + 60);
    const double2 A0_490 = ld2(rk0 + 66), A1_490 = ld2(rk1 + 66);
    { const double2 a0 = A0_488, a1 = A1_488;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M5 * 1.0;
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    { const double2 a0 = A0_489, a1 = A1_489;
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
    { const double2 a0 = A0_490, a1 = A1_490;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c15;
      jc2 += -1.0 * dq;
      jc15 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc24 += -1.0 * dq;
    }
    const double2 A0_491 = ld2(rk0 + 70), A1_491 = ld2(rk1 + 70);
    const double2 A0_492 = ld2(rk0 + 74), A1_492 = ld2(rk1 + 74);
    { const double2 a0 = A0_491, a1 = A1_491;
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
    { const double2 a0 = A0_492, a1 = A1_492;
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
    const double2 A0_493 = ld2(rk0 + 76), A1_493 = ld2(rk1 + 76);
    const double2 A0_494 = ld2(rk0 + 78), A1_494 = ld2(rk1 + 78);
    const double2 A0_495 = ld2(rk0 + 80), A1_495 = ld2(rk1 + 80);
    { const double2 a0 = A0_493, a1 = A1_493;
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
    { const double2 a0 = A0_494, a1 = A1_494;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c9;
      jc2 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
      jc19 += 1.0 * dq;
    }
    { const double2 a0 = A0_495, a1 = A1_495;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c17;
      jc2 += -1.0 * dq;
      jc10 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc19 += 1.0 * dq;
    }
    const double2 A0_496 = ld2(rk0 + 82), A1_496 = ld2(rk1 + 82);
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
    { const double2 a0 = A0_496, a1 = A1_496;
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
    const double2 A0_497 = ld2(rk0 + 84), A1_497 = ld2(rk1 + 84);
    const double2 A0_498 = ld2(rk0 + 86), A1_498 = ld2(rk1 + 86);
    const double2 A0_499 = ld2(rk0 + 88), A1_499 = ld2(rk1 + 88);
    { const double2 a0 = A0_497, a1 = A1_497;
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
    { const double2 a0 = A0_498, a1 = A1_498;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c25;
      jc2 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    { const double2 a0 = A0_499, a1 = A1_499;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c28;
      jc2 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc28 += 1.0 * dq;
    }
    const double2 A0_500 = ld2(rk0 + 92), A1_500 = ld2(rk1 + 92);
    const double2 A0_501 = ld2(rk0 + 146), A1_501 = ld2(rk1 + 146);
    {
      const double dq = scr[33 * SST] * c9 + 2.0 * scr[32 * SST];
      jc9 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    { const double2 a0 = A0_500, a1 = A1_500;
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
    { const double2 a0 = A0_501, a1 = A1_501;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c26;
      jc2 += 1.0 * dq;
      jc19 += -1.0 * dq;
      jc26 += -1.0 * dq;
      jc29 += 1.0 * dq;
    }
    const double2 A0_502 = ld2(rk0 + 162), A1_502 = ld2(rk1 + 162);
    const double2 A0_503 = ld2(rk0 + 164), A1_503 = ld2(rk1 + 164);
    const double2 A0_504 = ld2(rk0 + 174), A1_504 = ld2(rk1 + 174);
    { const double2 a0 = A0_502, a1 = A1_502;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c29;
      jc2 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc19 += -1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = A0_503, a1 = A1_503;
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
    { const double2 a0 = A0_504, a1 = A1_504;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c16;
      jc2 += 1.0 * dq;
      jc16 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    const double2 A0_505 = ld2(rk0 + 214), A1_505 = ld2(rk1 + 214);
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
    { const double2 a0 = A0_505, a1 = A1_505;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c10;
      jc10 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    const double2 A0_506 = ld2(rk0 + 314), A1_506 = ld2(rk1 + 314);
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
    { const double2 a0 = A0_506, a1 = A1_506;
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
    const double2 A0_507 = ld2(rk0 + 264), A1_507 = ld2(rk1 + 264);
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
    { const double2 a0 = A0_507, a1 = A1_507;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M11 * 1.0 + scr[49 * SST];
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc20 += -1.0 * dq;
    }
    const double2 A0_508 = ld2(rk0 + 280), A1_508 = ld2(rk1 + 280);
    const double2 A0_509 = ld2(rk0 + 284), A1_509 = ld2(rk1 + 284);
    const double2 A0_510 = ld2(rk0 + 344), A1_510 = ld2(rk1 + 344);
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = A0_508, a1 = A1_508;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c7;
      jc7 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc20 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    { const double2 a0 = A0_509, a1 = A1_509;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc3 += 1.0 * dq;
      jc12 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += -1.0 * dq;
    }
    { const double2 a0 = A0_510, a1 = A1_510;
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
    const double2 A0_511 = ld2(rk0 + 324), A1_511 = ld2(rk1 + 324);
    const double2 A0_512 = ld2(rk0 + 328), A1_512 = ld2(rk1 + 328);
    const double2 A0_513 = ld2(rk0 + 330), A1_513 = ld2(rk1 + 330);
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = A0_511, a1 = A1_511;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 1.0;
      jc10 += 1.0 * dq;
      jc21 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    { const double2 a0 = A0_512, a1 = A1_512;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc6 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc21 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_513, a1 = A1_513;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c7;
      jc6 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc21 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    const double2 A0_514 = ld2(rk0 + 332), A1_514 = ld2(rk1 + 332);
    const double2 A0_515 = ld2(rk0 + 334), A1_515 = ld2(rk1 + 334);
    const double2 A0_516 = ld2(rk0 + 336), A1_516 = ld2(rk1 + 336);
    { const double2 a0 = A0_514, a1 = A1_514;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c2;
      jc2 += -1.0 * dq;
      jc6 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc21 += -1.0 * dq;
    }
    { const double2 a0 = A0_515, a1 = A1_515;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c24;
      jc6 += 1.0 * dq;
      jc8 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc21 += -1.0 * dq;
      jc24 += -1.0 * dq;
    }
    { const double2 a0 = A0_516, a1 = A1_516;
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
    const double2 A0_517 = ld2(rk0 + 304), A1_517 = ld2(rk1 + 304);
    const double2 A0_518 = ld2(rk0 + 308), A1_518 = ld2(rk1 + 308);
    const double2 A0_519 = ld2(rk0 + 312), A1_519 = ld2(rk1 + 312);
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = A0_517, a1 = A1_517;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 1.0;
      jc17 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc22 += -1.0 * dq;
    }
    { const double2 a0 = A0_518, a1 = A1_518;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc14 += 1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_519, a1 = A1_519;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c2;
      jc2 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    const double2 A0_520 = ld2(rk0 + 316), A1_520 = ld2(rk1 + 316);
    const double2 A0_521 = ld2(rk0 + 320), A1_521 = ld2(rk1 + 320);
    const double2 A0_522 = ld2(rk0 + 368), A1_522 = ld2(rk1 + 368);
    { const double2 a0 = A0_520, a1 = A1_520;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c24;
      jc8 += 1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
      jc24 += -1.0 * dq;
    }
    { const double2 a0 = A0_521, a1 = A1_521;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc3 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    { const double2 a0 = A0_522, a1 = A1_522;
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
    const double2 A0_523 = ld2(rk0 + 310), A1_523 = ld2(rk1 + 310);
    const double2 A0_524 = ld2(rk0 + 314), A1_524 = ld2(rk1 + 314);
    const double2 A0_525 = ld2(rk0 + 318), A1_525 = ld2(rk1 + 318);
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = A0_523, a1 = A1_523;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c14;
      jc14 += -1.0 * dq;
      jc22 += 1.0 * dq;
      jc23 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    { const double2 a0 = A0_524, a1 = A1_524;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c19;
      jc2 += 1.0 * dq;
      jc19 += -1.0 * dq;
      jc22 += 1.0 * dq;
      jc23 += -1.0 * dq;
    }
    { const double2 a0 = A0_525, a1 = A1_525;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c8;
      jc8 += -1.0 * dq;
      jc22 += 1.0 * dq;
      jc23 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    const double2 A0_526 = ld2(rk0 + 322), A1_526 = ld2(rk1 + 322);
    const double2 A0_527 = ld2(rk0 + 342), A1_527 = ld2(rk1 + 342);
    const double2 A0_528 = ld2(rk0 + 346), A1_528 = ld2(rk1 + 346);
    { const double2 a0 = A0_526, a1 = A1_526;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c3;
      jc3 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc22 += 1.0 * dq;
      jc23 += -1.0 * dq;
    }
    { const double2 a0 = A0_527, a1 = A1_527;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 1.0;
      jc20 += 1.0 * dq;
      jc23 += -1.0 * dq;
    }
    { const double2 a0 = A0_528, a1 = A1_528;
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
    const double2 A0_529 = ld2(rk0 + 44), A1_529 = ld2(rk1 + 44);
    const double2 A0_530 = ld2(rk0 + 46), A1_530 = ld2(rk1 + 46);
    const double2 A0_531 = ld2(rk0 + 48), A1_531 = ld2(rk1 + 48);
    {
      const double dq = scr[2 * SST];
      jc7 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    { const double2 a0 = A0_529, a1 = A1_529;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c17;
      jc10 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc24 += 1.0 * dq;
    }
    { const double2 a0 = A0_530, a1 = A1_530;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M3 * 1.0 + scr[3 * SST];
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    { const double2 a0 = A0_531, a1 = A1_531;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M4 * 1.0;
      jc2 += -1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
    }
    const double2 A0_532 = ld2(rk0 + 50), A1_532 = ld2(rk1 + 50);
    const double2 A0_533 = ld2(rk0 + 64), A1_533 = ld2(rk1 + 64);
    { const double2 a0 = A0_532, a1 = A1_532;
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
    { const double2 a0 = A0_533, a1 = A1_533;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c2;
      jc2 += -1.0 * dq;
      jc7 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc24 += -1.0 * dq;
    }
    const double2 A0_534 = ld2(rk0 + 66), A1_534 = ld2(rk1 + 66);
    const double2 A0_535 = ld2(rk0 + 68), A1_535 = ld2(rk1 + 68);
    const double2 A0_536 = ld2(rk0 + 70), A1_536 = ld2(rk1 + 70);
    { const double2 a0 = A0_534, a1 = A1_534;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c2;
      jc2 += -1.0 * dq;
      jc15 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc24 += -1.0 * dq;
    }
    { const double2 a0 = A0_535, a1 = A1_535;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c2;
      jc2 += -1.0 * dq;
      jc24 += -1.0 * dq;
      jc27 += 2.0 * dq;
    }
    { const double2 a0 = A0_536, a1 = A1_536;
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
    const double2 A0_537 = ld2(rk0 + 96), A1_537 = ld2(rk1 + 96);
    const double2 A0_538 = ld2(rk0 + 98), A1_538 = ld2(rk1 + 98);
    const double2 A0_539 = ld2(rk0 + 100), A1_539 = ld2(rk1 + 100);
    const double2 A0_540 = ld2(rk0 + 132), A1_540 = ld2(rk1 + 132);
    { const double2 a0 = A0_537, a1 = A1_537;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc14 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_538, a1 = A1_538;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c14;
      jc8 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc24 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_539, a1 = A1_539;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c14;
      jc8 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc24 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_540, a1 = A1_540;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 2.0*c24;
      jc8 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -2.0 * dq;
    }
    const double2 A0_541 = ld2(rk0 + 134), A1_541 = ld2(rk1 + 134);
    const double2 A0_542 = ld2(rk0 + 136), A1_542 = ld2(rk1 + 136);
    const double2 A0_543 = ld2(rk0 + 138), A1_543 = ld2(rk1 + 138);
    const double2 A0_544 = ld2(rk0 + 140), A1_544 = ld2(rk1 + 140);
    { const double2 a0 = A0_541, a1 = A1_541;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 2.0*c24;
      jc8 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -2.0 * dq;
    }
    { const double2 a0 = A0_542, a1 = A1_542;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c29;
      jc10 += 1.0 * dq;
      jc24 += -1.0 * dq;
      jc27 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = A0_543, a1 = A1_543;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc3 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc24 += -1.0 * dq;
    }
    { const double2 a0 = A0_544, a1 = A1_544;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c9;
      jc9 += -1.0 * dq;
      jc24 += -1.0 * dq;
      jc27 += 1.0 * dq;
      jc30 += 1.0 * dq;
    }
    const double2 A0_545 = ld2(rk0 + 142), A1_545 = ld2(rk1 + 142);
    const double2 A0_546 = ld2(rk0 + 192), A1_546 = ld2(rk1 + 192);
    { const double2 a0 = A0_545, a1 = A1_545;
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
    { const double2 a0 = A0_546, a1 = A1_546;
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
    const double2 A0_547 = ld2(rk0 + 208), A1_547 = ld2(rk1 + 208);
    const double2 A0_548 = ld2(rk0 + 212), A1_548 = ld2(rk1 + 212);
    {
      const double dq = scr[45 * SST];
      jc2 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc17 += -1.0 * dq;
    }
    { const double2 a0 = A0_547, a1 = A1_547;
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
    { const double2 a0 = A0_548, a1 = A1_548;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c25;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
      jc25 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    const double2 A0_549 = ld2(rk0 + 234), A1_549 = ld2(rk1 + 234);
    const double2 A0_550 = ld2(rk0 + 246), A1_550 = ld2(rk1 + 246);
    const double2 A0_551 = ld2(rk0 + 248), A1_551 = ld2(rk1 + 248);
    { const double2 a0 = A0_549, a1 = A1_549;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c4;
      jc4 += -1.0 * dq;
      jc6 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -1.0 * dq;
    }
    { const double2 a0 = A0_550, a1 = A1_550;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c28;
      jc11 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -1.0 * dq;
      jc28 += -1.0 * dq;
    }
    { const double2 a0 = A0_551, a1 = A1_551;
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
    const double2 A0_552 = ld2(rk0 + 270), A1_552 = ld2(rk1 + 270);
    const double2 A0_553 = ld2(rk0 + 272), A1_553 = ld2(rk1 + 272);
    const double2 A0_554 = ld2(rk0 + 316), A1_554 = ld2(rk1 + 316);
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = A0_552, a1 = A1_552;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c6;
      jc6 += -1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -1.0 * dq;
      jc25 += 1.0 * dq;
    }
    { const double2 a0 = A0_553, a1 = A1_553;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c25;
      jc6 += 1.0 * dq;
      jc8 += 1.0 * dq;
      jc24 += -1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = A0_554, a1 = A1_554;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c22;
      jc8 += 1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
      jc24 += -1.0 * dq;
    }
    const double2 A0_555 = ld2(rk0 + 334), A1_555 = ld2(rk1 + 334);
    const double2 A0_556 = ld2(rk0 + 356), A1_556 = ld2(rk1 + 356);
    const double2 A0_557 = ld2(rk0 + 366), A1_557 = ld2(rk1 + 366);
    { const double2 a0 = A0_555, a1 = A1_555;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c21;
      jc6 += 1.0 * dq;
      jc8 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc21 += -1.0 * dq;
      jc24 += -1.0 * dq;
    }
    { const double2 a0 = A0_556, a1 = A1_556;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c12;
      jc1 += 1.0 * dq;
      jc12 += -1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -1.0 * dq;
    }
    { const double2 a0 = A0_557, a1 = A1_557;
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
    const double2 A0_558 = ld2(rk0 + 34), A1_558 = ld2(rk1 + 34);
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
    { const double2 a0 = A0_558, a1 = A1_558;
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
    const double2 A0_559 = ld2(rk0 + 84), A1_559 = ld2(rk1 + 84);
    const double2 A0_560 = ld2(rk0 + 86), A1_560 = ld2(rk1 + 86);
    {
      const double dq = scr[27 * SST] * c2 + scr[26 * SST];
      jc2 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    { const double2 a0 = A0_559, a1 = A1_559;
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
    { const double2 a0 = A0_560, a1 = A1_560;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c19;
      jc2 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    const double2 A0_561 = ld2(rk0 + 128), A1_561 = ld2(rk1 + 128);
    const double2 A0_562 = ld2(rk0 + 154), A1_562 = ld2(rk1 + 154);
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
    { const double2 a0 = A0_561, a1 = A1_561;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc6 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc25 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_562, a1 = A1_562;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc3 += -1.0 * dq;
      jc25 += 1.0 * dq;
      jc26 += -1.0 * dq;
    }
    const double2 A0_563 = ld2(rk0 + 166), A1_563 = ld2(rk1 + 166);
    const double2 A0_564 = ld2(rk0 + 178), A1_564 = ld2(rk1 + 178);
    { const double2 a0 = A0_563, a1 = A1_563;
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
    { const double2 a0 = A0_564, a1 = A1_564;
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
    const double2 A0_565 = ld2(rk0 + 200), A1_565 = ld2(rk1 + 200);
    const double2 A0_566 = ld2(rk0 + 212), A1_566 = ld2(rk1 + 212);
    { const double2 a0 = A0_565, a1 = A1_565;
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
    { const double2 a0 = A0_566, a1 = A1_566;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c24;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
      jc25 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    const double2 A0_567 = ld2(rk0 + 218), A1_567 = ld2(rk1 + 218);
    const double2 A0_568 = ld2(rk0 + 220), A1_568 = ld2(rk1 + 220);
    const double2 A0_569 = ld2(rk0 + 252), A1_569 = ld2(rk1 + 252);
    const double2 A0_570 = ld2(rk0 + 262), A1_570 = ld2(rk1 + 262);
    { const double2 a0 = A0_567, a1 = A1_567;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 2.0*c25;
      jc6 += 1.0 * dq;
      jc25 += -2.0 * dq;
      jc28 += 1.0 * dq;
    }
    { const double2 a0 = A0_568, a1 = A1_568;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c7;
      jc6 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    { const double2 a0 = A0_569, a1 = A1_569;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c11;
      jc11 += -1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 2.0 * dq;
    }
    { const double2 a0 = A0_570, a1 = A1_570;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c3;
      jc3 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc25 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    const double2 A0_571 = ld2(rk0 + 268), A1_571 = ld2(rk1 + 268);
    const double2 A0_572 = ld2(rk0 + 272), A1_572 = ld2(rk1 + 272);
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
    { const double2 a0 = A0_571, a1 = A1_571;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c15;
      jc6 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = A0_572, a1 = A1_572;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c24;
      jc6 += 1.0 * dq;
      jc8 += 1.0 * dq;
      jc24 += -1.0 * dq;
      jc25 += -1.0 * dq;
    }
    const double2 A0_573 = ld2(rk0 + 278), A1_573 = ld2(rk1 + 278);
    const double2 A0_574 = ld2(rk0 + 288), A1_574 = ld2(rk1 + 288);
    const double2 A0_575 = ld2(rk0 + 300), A1_575 = ld2(rk1 + 300);
    const double2 A0_576 = ld2(rk0 + 326), A1_576 = ld2(rk1 + 326);
    { const double2 a0 = A0_573, a1 = A1_573;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc13 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = A0_574, a1 = A1_574;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c7;
      jc1 += 1.0 * dq;
      jc2 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc25 += -1.0 * dq;
    }
    { const double2 a0 = A0_575, a1 = A1_575;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc10 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc25 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_576, a1 = A1_576;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c10;
      jc10 += -1.0 * dq;
      jc21 += 1.0 * dq;
      jc25 += -1.0 * dq;
    }
    const double2 A0_577 = ld2(rk0 + 366), A1_577 = ld2(rk1 + 366);
    { const double2 a0 = A0_577, a1 = A1_577;
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
    const double2 A0_578 = ld2(rk0 + 74), A1_578 = ld2(rk1 + 74);
    { const double2 a0 = A0_578, a1 = A1_578;
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
    const double2 A0_579 = ld2(rk0 + 102), A1_579 = ld2(rk1 + 102);
    const double2 A0_580 = ld2(rk0 + 106), A1_580 = ld2(rk1 + 106);
    const double2 A0_581 = ld2(rk0 + 144), A1_581 = ld2(rk1 + 144);
    {
      const double dq = scr[35 * SST];
      jc8 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    { const double2 a0 = A0_579, a1 = A1_579;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc2 += 1.0 * dq;
      jc17 += 1.0 * dq;
      jc26 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_580, a1 = A1_580;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c14;
      jc14 += 1.0 * dq;
      jc26 += 1.0 * dq;
      jc27 += -1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = A0_581, a1 = A1_581;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c15;
      jc7 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc26 += -1.0 * dq;
    }
    const double2 A0_582 = ld2(rk0 + 146), A1_582 = ld2(rk1 + 146);
    const double2 A0_583 = ld2(rk0 + 148), A1_583 = ld2(rk1 + 148);
    const double2 A0_584 = ld2(rk0 + 150), A1_584 = ld2(rk1 + 150);
    const double2 A0_585 = ld2(rk0 + 152), A1_585 = ld2(rk1 + 152);
    { const double2 a0 = A0_582, a1 = A1_582;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c19;
      jc2 += 1.0 * dq;
      jc19 += -1.0 * dq;
      jc26 += -1.0 * dq;
      jc29 += 1.0 * dq;
    }
    { const double2 a0 = A0_583, a1 = A1_583;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c14;
      jc2 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc14 += -1.0 * dq;
      jc26 += -1.0 * dq;
    }
    { const double2 a0 = A0_584, a1 = A1_584;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c29;
      jc2 += 1.0 * dq;
      jc4 += 1.0 * dq;
      jc26 += -1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = A0_585, a1 = A1_585;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc2 += 1.0 * dq;
      jc6 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc26 += -1.0 * dq;
    }
    const double2 A0_586 = ld2(rk0 + 154), A1_586 = ld2(rk1 + 154);
    const double2 A0_587 = ld2(rk0 + 156), A1_587 = ld2(rk1 + 156);
    const double2 A0_588 = ld2(rk0 + 158), A1_588 = ld2(rk1 + 158);
    { const double2 a0 = A0_586, a1 = A1_586;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c3;
      jc2 += 1.0 * dq;
      jc3 += -1.0 * dq;
      jc25 += 1.0 * dq;
      jc26 += -1.0 * dq;
    }
    { const double2 a0 = A0_587, a1 = A1_587;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c30;
      jc9 += 1.0 * dq;
      jc17 += 1.0 * dq;
      jc26 += -1.0 * dq;
      jc30 += -1.0 * dq;
    }
    { const double2 a0 = A0_588, a1 = A1_588;
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
    const double2 A0_589 = ld2(rk0 + 228), A1_589 = ld2(rk1 + 228);
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
    { const double2 a0 = A0_589, a1 = A1_589;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c15;
      jc9 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc26 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    const double2 A0_590 = ld2(rk0 + 292), A1_590 = ld2(rk1 + 292);
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
    { const double2 a0 = A0_590, a1 = A1_590;
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
    const double2 A0_591 = ld2(rk0 + 2), A1_591 = ld2(rk1 + 2);
    {
      const double dq = scr[0 * SST];
      jc7 += -2.0 * dq;
      jc15 += 1.0 * dq;
    }
    { const double2 a0 = A0_591, a1 = A1_591;
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
    const double2 A0_592 = ld2(rk0 + 38), A1_592 = ld2(rk1 + 38);
    const double2 A0_593 = ld2(rk0 + 52), A1_593 = ld2(rk1 + 52);
    {
      const double dq = -scr[76 * SST] * c17;
      jc7 += -1.0 * dq;
      jc10 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc27 += 1.0 * dq;
    }
    { const double2 a0 = A0_592, a1 = A1_592;
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
    { const double2 a0 = A0_593, a1 = A1_593;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c7;
      jc2 += -1.0 * dq;
      jc7 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc27 += 1.0 * dq;
    }
    const double2 A0_594 = ld2(rk0 + 62), A1_594 = ld2(rk1 + 62);
    const double2 A0_595 = ld2(rk0 + 68), A1_595 = ld2(rk1 + 68);
    const double2 A0_596 = ld2(rk0 + 72), A1_596 = ld2(rk1 + 72);
    {
      const double dq = scr[6 * SST];
      jc2 += -2.0 * dq;
      jc19 += 1.0 * dq;
    }
    { const double2 a0 = A0_594, a1 = A1_594;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M9 * c2 + scr[10 * SST];
      jc2 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_595, a1 = A1_595;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * 2.0*c27;
      jc2 += -1.0 * dq;
      jc24 += -1.0 * dq;
      jc27 += 2.0 * dq;
    }
    { const double2 a0 = A0_596, a1 = A1_596;
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
    const double2 A0_597 = ld2(rk0 + 92), A1_597 = ld2(rk1 + 92);
    const double2 A0_598 = ld2(rk0 + 94), A1_598 = ld2(rk1 + 94);
    const double2 A0_599 = ld2(rk0 + 96), A1_599 = ld2(rk1 + 96);
    { const double2 a0 = A0_597, a1 = A1_597;
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
    { const double2 a0 = A0_598, a1 = A1_598;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 2.0*c27;
      jc7 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -2.0 * dq;
    }
    { const double2 a0 = A0_599, a1 = A1_599;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c24;
      jc14 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    const double2 A0_600 = ld2(rk0 + 98), A1_600 = ld2(rk1 + 98);
    const double2 A0_601 = ld2(rk0 + 100), A1_601 = ld2(rk1 + 100);
    const double2 A0_602 = ld2(rk0 + 102), A1_602 = ld2(rk1 + 102);
    const double2 A0_603 = ld2(rk0 + 104), A1_603 = ld2(rk1 + 104);
    { const double2 a0 = A0_600, a1 = A1_600;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c8;
      jc8 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc24 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_601, a1 = A1_601;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c8;
      jc8 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc24 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_602, a1 = A1_602;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c26;
      jc2 += 1.0 * dq;
      jc17 += 1.0 * dq;
      jc26 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_603, a1 = A1_603;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c29;
      jc2 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc27 += -1.0 * dq;
      jc29 += -1.0 * dq;
    }
    const double2 A0_604 = ld2(rk0 + 106), A1_604 = ld2(rk1 + 106);
    const double2 A0_605 = ld2(rk0 + 108), A1_605 = ld2(rk1 + 108);
    const double2 A0_606 = ld2(rk0 + 110), A1_606 = ld2(rk1 + 110);
    const double2 A0_607 = ld2(rk0 + 112), A1_607 = ld2(rk1 + 112);
    { const double2 a0 = A0_604, a1 = A1_604;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c29;
      jc14 += 1.0 * dq;
      jc26 += 1.0 * dq;
      jc27 += -1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = A0_605, a1 = A1_605;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c16;
      jc2 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc16 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_606, a1 = A1_606;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc14 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc27 += -1.0 * dq;
      jc29 += 1.0 * dq;
    }
    { const double2 a0 = A0_607, a1 = A1_607;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc14 += 1.0 * dq;
      jc16 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    const double2 A0_608 = ld2(rk0 + 114), A1_608 = ld2(rk1 + 114);
    const double2 A0_609 = ld2(rk0 + 116), A1_609 = ld2(rk1 + 116);
    const double2 A0_610 = ld2(rk0 + 118), A1_610 = ld2(rk1 + 118);
    const double2 A0_611 = ld2(rk0 + 120), A1_611 = ld2(rk1 + 120);
    { const double2 a0 = A0_608, a1 = A1_608;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c3;
      jc3 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_609, a1 = A1_609;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c9;
      jc2 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc27 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    { const double2 a0 = A0_610, a1 = A1_610;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c17;
      jc9 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc17 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_611, a1 = A1_611;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c10;
      jc10 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc17 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    const double2 A0_612 = ld2(rk0 + 122), A1_612 = ld2(rk1 + 122);
    const double2 A0_613 = ld2(rk0 + 124), A1_613 = ld2(rk1 + 124);
    const double2 A0_614 = ld2(rk0 + 126), A1_614 = ld2(rk1 + 126);
    const double2 A0_615 = ld2(rk0 + 128), A1_615 = ld2(rk1 + 128);
    { const double2 a0 = A0_612, a1 = A1_612;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c4;
      jc2 += 1.0 * dq;
      jc4 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_613, a1 = A1_613;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c4;
      jc4 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_614, a1 = A1_614;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c6;
      jc4 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_615, a1 = A1_615;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c25;
      jc6 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc25 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    const double2 A0_616 = ld2(rk0 + 130), A1_616 = ld2(rk1 + 130);
    const double2 A0_617 = ld2(rk0 + 136), A1_617 = ld2(rk1 + 136);
    const double2 A0_618 = ld2(rk0 + 140), A1_618 = ld2(rk1 + 140);
    const double2 A0_619 = ld2(rk0 + 160), A1_619 = ld2(rk1 + 160);
    { const double2 a0 = A0_616, a1 = A1_616;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c11;
      jc11 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    { const double2 a0 = A0_617, a1 = A1_617;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c10;
      jc10 += 1.0 * dq;
      jc24 += -1.0 * dq;
      jc27 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = A0_618, a1 = A1_618;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c30;
      jc9 += -1.0 * dq;
      jc24 += -1.0 * dq;
      jc27 += 1.0 * dq;
      jc30 += 1.0 * dq;
    }
    { const double2 a0 = A0_619, a1 = A1_619;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c17;
      jc15 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc27 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    const double2 A0_620 = ld2(rk0 + 190), A1_620 = ld2(rk1 + 190);
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
    { const double2 a0 = A0_620, a1 = A1_620;
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
    const double2 A0_621 = ld2(rk0 + 214), A1_621 = ld2(rk1 + 214);
    const double2 A0_622 = ld2(rk0 + 222), A1_622 = ld2(rk1 + 222);
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
    { const double2 a0 = A0_621, a1 = A1_621;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc10 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_622, a1 = A1_622;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c6;
      jc6 += -1.0 * dq;
      jc7 += 1.0 * dq;
      jc25 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    const double2 A0_623 = ld2(rk0 + 230), A1_623 = ld2(rk1 + 230);
    const double2 A0_624 = ld2(rk0 + 240), A1_624 = ld2(rk1 + 240);
    { const double2 a0 = A0_623, a1 = A1_623;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c9;
      jc9 += -1.0 * dq;
      jc15 += 1.0 * dq;
      jc26 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_624, a1 = A1_624;
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
    const double2 A0_625 = ld2(rk0 + 282), A1_625 = ld2(rk1 + 282);
    const double2 A0_626 = ld2(rk0 + 300), A1_626 = ld2(rk1 + 300);
    const double2 A0_627 = ld2(rk0 + 308), A1_627 = ld2(rk1 + 308);
    const double2 A0_628 = ld2(rk0 + 328), A1_628 = ld2(rk1 + 328);
    { const double2 a0 = A0_625, a1 = A1_625;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c12;
      jc7 += 1.0 * dq;
      jc12 += -1.0 * dq;
      jc20 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_626, a1 = A1_626;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c25;
      jc10 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc25 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_627, a1 = A1_627;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c22;
      jc14 += 1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
      jc27 += -1.0 * dq;
    }
    { const double2 a0 = A0_628, a1 = A1_628;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c21;
      jc6 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc14 += 1.0 * dq;
      jc21 += -1.0 * dq;
      jc27 += -1.0 * dq;
    }
    const double2 A0_629 = ld2(rk0 + 358), A1_629 = ld2(rk1 + 358);
    const double2 A0_630 = ld2(rk0 + 368), A1_630 = ld2(rk1 + 368);
    { const double2 a0 = A0_629, a1 = A1_629;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c12;
      jc12 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc27 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    { const double2 a0 = A0_630, a1 = A1_630;
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
    const double2 A0_631 = ld2(rk0 + 36), A1_631 = ld2(rk1 + 36);
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
    { const double2 a0 = A0_631, a1 = A1_631;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c7;
      jc7 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    const double2 A0_632 = ld2(rk0 + 38), A1_632 = ld2(rk1 + 38);
    { const double2 a0 = A0_632, a1 = A1_632;
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
    const double2 A0_633 = ld2(rk0 + 86), A1_633 = ld2(rk1 + 86);
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
    { const double2 a0 = A0_633, a1 = A1_633;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c2;
      jc2 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc25 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    const double2 A0_634 = ld2(rk0 + 88), A1_634 = ld2(rk1 + 88);
    const double2 A0_635 = ld2(rk0 + 130), A1_635 = ld2(rk1 + 130);
    { const double2 a0 = A0_634, a1 = A1_634;
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
    { const double2 a0 = A0_635, a1 = A1_635;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c14;
      jc11 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc27 += -1.0 * dq;
      jc28 += 1.0 * dq;
    }
    const double2 A0_636 = ld2(rk0 + 188), A1_636 = ld2(rk1 + 188);
    const double2 A0_637 = ld2(rk0 + 194), A1_637 = ld2(rk1 + 194);
    {
      const double dq = scr[38 * SST];
      jc9 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = A0_636, a1 = A1_636;
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
    { const double2 a0 = A0_637, a1 = A1_637;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc18 += -2.0 * dq;
      jc28 += 1.0 * dq;
    }
    const double2 A0_638 = ld2(rk0 + 202), A1_638 = ld2(rk1 + 202);
    const double2 A0_639 = ld2(rk0 + 212), A1_639 = ld2(rk1 + 212);
    { const double2 a0 = A0_638, a1 = A1_638;
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
    { const double2 a0 = A0_639, a1 = A1_639;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c15;
      jc15 += -1.0 * dq;
      jc24 += 1.0 * dq;
      jc25 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    const double2 A0_640 = ld2(rk0 + 216), A1_640 = ld2(rk1 + 216);
    const double2 A0_641 = ld2(rk0 + 246), A1_641 = ld2(rk1 + 246);
    const double2 A0_642 = ld2(rk0 + 250), A1_642 = ld2(rk1 + 250);
    const double2 A0_643 = ld2(rk0 + 254), A1_643 = ld2(rk1 + 254);
    { const double2 a0 = A0_640, a1 = A1_640;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c6;
      jc6 += -1.0 * dq;
      jc25 += 2.0 * dq;
      jc28 += -1.0 * dq;
    }
    { const double2 a0 = A0_641, a1 = A1_641;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c24;
      jc11 += 1.0 * dq;
      jc15 += 1.0 * dq;
      jc24 += -1.0 * dq;
      jc28 += -1.0 * dq;
    }
    { const double2 a0 = A0_642, a1 = A1_642;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c8;
      jc8 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc24 += 1.0 * dq;
      jc28 += -1.0 * dq;
    }
    { const double2 a0 = A0_643, a1 = A1_643;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 2.0*c28;
      jc11 += 1.0 * dq;
      jc25 += 1.0 * dq;
      jc28 += -2.0 * dq;
    }
    const double2 A0_644 = ld2(rk0 + 260), A1_644 = ld2(rk1 + 260);
    const double2 A0_645 = ld2(rk0 + 294), A1_645 = ld2(rk1 + 294);
    { const double2 a0 = A0_644, a1 = A1_644;
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
    { const double2 a0 = A0_645, a1 = A1_645;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c29;
      jc11 += 1.0 * dq;
      jc26 += 1.0 * dq;
      jc28 += -1.0 * dq;
      jc29 += -1.0 * dq;
    }
    const double2 A0_646 = ld2(rk0 + 298), A1_646 = ld2(rk1 + 298);
    { const double2 a0 = A0_646, a1 = A1_646;
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
    const double2 A0_647 = ld2(rk0 + 40), A1_647 = ld2(rk1 + 40);
    {
      const double dq = -scr[78 * SST] * c9;
      jc4 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc29 += 1.0 * dq;
    }
    { const double2 a0 = A0_647, a1 = A1_647;
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
    const double2 A0_648 = ld2(rk0 + 104), A1_648 = ld2(rk1 + 104);
    const double2 A0_649 = ld2(rk0 + 106), A1_649 = ld2(rk1 + 106);
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
    { const double2 a0 = A0_648, a1 = A1_648;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc2 += 1.0 * dq;
      jc10 += 1.0 * dq;
      jc27 += -1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = A0_649, a1 = A1_649;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c27;
      jc14 += 1.0 * dq;
      jc26 += 1.0 * dq;
      jc27 += -1.0 * dq;
      jc29 += -1.0 * dq;
    }
    const double2 A0_650 = ld2(rk0 + 110), A1_650 = ld2(rk1 + 110);
    const double2 A0_651 = ld2(rk0 + 136), A1_651 = ld2(rk1 + 136);
    const double2 A0_652 = ld2(rk0 + 146), A1_652 = ld2(rk1 + 146);
    const double2 A0_653 = ld2(rk0 + 150), A1_653 = ld2(rk1 + 150);
    { const double2 a0 = A0_650, a1 = A1_650;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c14;
      jc14 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc27 += -1.0 * dq;
      jc29 += 1.0 * dq;
    }
    { const double2 a0 = A0_651, a1 = A1_651;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c24;
      jc10 += 1.0 * dq;
      jc24 += -1.0 * dq;
      jc27 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = A0_652, a1 = A1_652;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc19 += -1.0 * dq;
      jc26 += -1.0 * dq;
      jc29 += 1.0 * dq;
    }
    { const double2 a0 = A0_653, a1 = A1_653;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c26;
      jc2 += 1.0 * dq;
      jc4 += 1.0 * dq;
      jc26 += -1.0 * dq;
      jc29 += -1.0 * dq;
    }
    const double2 A0_654 = ld2(rk0 + 160), A1_654 = ld2(rk1 + 160);
    const double2 A0_655 = ld2(rk0 + 162), A1_655 = ld2(rk1 + 162);
    const double2 A0_656 = ld2(rk0 + 164), A1_656 = ld2(rk1 + 164);
    const double2 A0_657 = ld2(rk0 + 166), A1_657 = ld2(rk1 + 166);
    { const double2 a0 = A0_654, a1 = A1_654;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c15;
      jc15 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc27 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = A0_655, a1 = A1_655;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c19;
      jc2 += 1.0 * dq;
      jc18 += 1.0 * dq;
      jc19 += -1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = A0_656, a1 = A1_656;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 2.0*c29;
      jc4 += 1.0 * dq;
      jc19 += 1.0 * dq;
      jc29 += -2.0 * dq;
    }
    { const double2 a0 = A0_657, a1 = A1_657;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc2 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc25 += 1.0 * dq;
      jc29 += -1.0 * dq;
    }
    const double2 A0_658 = ld2(rk0 + 168), A1_658 = ld2(rk1 + 168);
    const double2 A0_659 = ld2(rk0 + 176), A1_659 = ld2(rk1 + 176);
    const double2 A0_660 = ld2(rk0 + 182), A1_660 = ld2(rk1 + 182);
    { const double2 a0 = A0_658, a1 = A1_658;
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
    { const double2 a0 = A0_659, a1 = A1_659;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c14;
      jc14 += 0.0 * dq;
      jc16 += -1.0 * dq;
      jc29 += 1.0 * dq;
    }
    { const double2 a0 = A0_660, a1 = A1_660;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c9;
      jc9 += 0.0 * dq;
      jc16 += -1.0 * dq;
      jc29 += 1.0 * dq;
    }
    const double2 A0_661 = ld2(rk0 + 184), A1_661 = ld2(rk1 + 184);
    { const double2 a0 = A0_661, a1 = A1_661;
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
    const double2 A0_662 = ld2(rk0 + 294), A1_662 = ld2(rk1 + 294);
    const double2 A0_663 = ld2(rk0 + 296), A1_663 = ld2(rk1 + 296);
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
    { const double2 a0 = A0_662, a1 = A1_662;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c28;
      jc11 += 1.0 * dq;
      jc26 += 1.0 * dq;
      jc28 += -1.0 * dq;
      jc29 += -1.0 * dq;
    }
    { const double2 a0 = A0_663, a1 = A1_663;
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
    const double2 A0_664 = ld2(rk0 + 22), A1_664 = ld2(rk1 + 22);
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
    { const double2 a0 = A0_664, a1 = A1_664;
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
    const double2 A0_665 = ld2(rk0 + 40), A1_665 = ld2(rk1 + 40);
    const double2 A0_666 = ld2(rk0 + 42), A1_666 = ld2(rk1 + 42);
    { const double2 a0 = A0_665, a1 = A1_665;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c29;
      jc7 += -1.0 * dq;
      jc12 += -1.0 * dq;
      jc29 += 1.0 * dq;
      jc30 += 1.0 * dq;
    }
    { const double2 a0 = A0_666, a1 = A1_666;
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
    const double2 A0_667 = ld2(rk0 + 116), A1_667 = ld2(rk1 + 116);
    const double2 A0_668 = ld2(rk0 + 140), A1_668 = ld2(rk1 + 140);
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
    { const double2 a0 = A0_667, a1 = A1_667;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc27 += -1.0 * dq;
      jc30 += 1.0 * dq;
    }
    { const double2 a0 = A0_668, a1 = A1_668;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c27;
      jc9 += -1.0 * dq;
      jc24 += -1.0 * dq;
      jc27 += 1.0 * dq;
      jc30 += 1.0 * dq;
    }
    const double2 A0_669 = ld2(rk0 + 156), A1_669 = ld2(rk1 + 156);
    const double2 A0_670 = ld2(rk0 + 184), A1_670 = ld2(rk1 + 184);
    const double2 A0_671 = ld2(rk0 + 186), A1_671 = ld2(rk1 + 186);
    { const double2 a0 = A0_669, a1 = A1_669;
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
    { const double2 a0 = A0_670, a1 = A1_670;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c16 - (a0.y + (a1.y - a0.y) * Td) * c29;
      jc16 += -1.0 * dq;
      jc29 += 1.0 * dq;
      jc30 += 0.0 * dq;
    }
    { const double2 a0 = A0_671, a1 = A1_671;
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
    const double2 A0_672 = ld2(rk0 + 360), A1_672 = ld2(rk1 + 360);
    {
      const double dq = scr[50 * SST];
      jc9 += -1.0 * dq;
      jc18 += -1.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = A0_672, a1 = A1_672;
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
