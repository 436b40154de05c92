// Generated by flint_amd/codegen.py from the TSR-GP-24 mechanism descriptor. Do not edit.
// 24 species, 110 reactions (102 Arrhenius/three-body, 8 Troe, 0 Lindemann).
#include "../flint_device.h"
namespace flint_gen_tsr_gp_24 {
constexpr int NS = 24, NZ = 25, WK = 236, WH = 48, NSLOT = 79;

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
  const double c0 = r0 / 33.006;
  const double c1 = r1 / 34.014;
  const double c2 = r2 / 28.01;
  const double c3 = r3 / 17.007;
  const double c4 = r4 / 2.016;
  const double c5 = r5 / 29.018;
  const double c6 = r6 / 31.034;
  const double c7 = r7 / 14.027;
  const double c8 = r8 / 1.008;
  const double c9 = r9 / 15.999;
  const double c10 = r10 / 38.049;
  const double c11 = r11 / 15.035;
  const double c12 = r12 / 26.038;
  const double c13 = r13 / 30.026;
  const double c14 = r14 / 39.057;
  const double c15 = r15 / 25.03;
  const double c16 = r16 / 18.015;
  const double c17 = r17 / 47.033;
  const double c18 = r18 / 31.998;
  const double c19 = r19 / 14.027;
  const double c20 = r20 / 30.07;
  const double c21 = r21 / 16.043;
  const double c22 = r22 / 45.017;
  const double c23 = r23 / 44.009;
  const int T_i = (int)T;
  const double Td = T - (double)T_i;
  const int rowk = min(max(T_i, Tk_min), Tk_max) - Tk_min;
  const double* __restrict__ rk0 = tabk + (long)rowk * WK;
  const double* __restrict__ rk1 = rk0 + WK;
  const double M0 = c0+c1+c2*1.75+c3+c4*2.4000000953674316+c5+c6+c7+c8+c9+c10+c11+c12+c13+c14+c15+c16*15.399999618530273+c17+c18+c19+c20*3.0+c21*2.0+c22+c23*3.5999999046325684;
  const double M1 = c0+c1+c2*1.5+c3+c4*2.0+c5+c6+c7+c8+c9+c10+c11+c12+c13+c14+c15+c16*6.0+c17+c18+c19+c20*3.0+c21*2.0+c22+c23*2.0;
  const double M2 = c0+c1+c2*1.5+c3+c4*2.0+c5+c6+c7+c8+c9+c10+c11+c12+c13+c14+c15+c16*6.0+c17+c18*6.0+c19+c20*3.0+c21*2.0+c22+c23*3.5;
  const double M3 = c0+c1+c2*0.75+c3+c4+c5+c6+c7+c8+c9+c10+c11+c12+c13+c14+c15+c17+c19+c20*1.5+c21+c22+c23*1.5;
  const double M4 = c18;
  const double M5 = c16;
  const double M6 = c0+c1+c2+c3+c5+c6+c7+c8+c9+c10+c11+c12+c13+c14+c15+c17+c18+c19+c20*3.0+c21*2.0+c22;
  const double M7 = c4;
  const double M8 = c23;
  const double M9 = c0+c1+c2+c3+c4*0.7300000190734863+c5+c6+c7+c8+c9+c10+c11+c12+c13+c14+c15+c16*3.6500000953674316+c17+c18+c19+c20*3.0+c21*2.0+c22+c23;
  const double M10 = c0+c1+c2*1.5+c3+c4*2.0+c5+c6+c7+c8+c9+c10+c11+c12+c13+c14+c15+c17+c18+c19+c20*3.0+c21*2.0+c22+c23*2.0;
  const double M11 = c0+c1+c2+c3+c4+c5+c6+c7+c8+c9+c10+c11+c12+c13+c14+c15+c16+c17+c18+c19+c20+c21+c22+c23;
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
  // JAC: dT-derivative accumulators
  [[maybe_unused]] double dT[NS] = {};
  const double2 A0_1 = ld2(rk0 + 0), A1_1 = ld2(rk1 + 0);
  const double2 A0_2 = ld2(rk0 + 2), A1_2 = ld2(rk1 + 2);
  const double2 A0_3 = ld2(rk0 + 4), A1_3 = ld2(rk1 + 4);
  const double2 A0_4 = ld2(rk0 + 6), A1_4 = ld2(rk1 + 6);
  const double2 A0_5 = ld2(rk0 + 8), A1_5 = ld2(rk1 + 8);
  const double2 A0_6 = ld2(rk0 + 10), A1_6 = ld2(rk1 + 10);
  const double2 A0_7 = ld2(rk0 + 12), A1_7 = ld2(rk1 + 12);
  const double2 A0_8 = ld2(rk0 + 14), A1_8 = ld2(rk1 + 14);
  const double2 A0_9 = ld2(rk0 + 16), A1_9 = ld2(rk1 + 16);
  const double2 A0_10 = ld2(rk0 + 18), A1_10 = ld2(rk1 + 18);
  const double2 A0_11 = ld2(rk0 + 20), A1_11 = ld2(rk1 + 20);
  const double2 A0_12 = ld2(rk0 + 22), A1_12 = ld2(rk1 + 22);
  const double2 A0_13 = ld2(rk0 + 24), A1_13 = ld2(rk1 + 24);
  const double2 A0_14 = ld2(rk0 + 26), A1_14 = ld2(rk1 + 26);
  const double2 A0_15 = ld2(rk0 + 28), A1_15 = ld2(rk1 + 28);
  const double2 A0_16 = ld2(rk0 + 30), A1_16 = ld2(rk1 + 30);
  {  // reaction 1: 2 O + M <=> O2 + M
    const double2 a0 = A0_1, a1 = A1_1;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c9*c9)*M0;
    const double pb = kb*(c18)*M0;
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c9*c9), Pb = (c18);
      const double dqT = ((a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb) * M0;
      scr[0 * SST] = kf * Pf - kb * Pb;
      dT[9] += -2.0 * dqT;
      dT[18] += 1.0 * dqT;
    }
    w9 = -2.0 * q;
    w18 = 1.0 * q;
  }
  {  // reaction 2: H + O + M <=> OH + M
    const double2 a0 = A0_2, a1 = A1_2;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c8)*(c9)*M1;
    const double pb = kb*(c3)*M1;
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c8)*(c9), Pb = (c3);
      const double dqT = ((a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb) * M1;
      scr[1 * SST] = kf * Pf - kb * Pb;
      dT[3] += 1.0 * dqT;
      dT[8] += -1.0 * dqT;
      dT[9] += -1.0 * dqT;
    }
    w3 = 1.0 * q;
    w8 = -1.0 * q;
    w9 = w9 + -1.0 * q;
  }
  {  // reaction 3: H2 + O <=> H + OH
    const double2 a0 = A0_3, a1 = A1_3;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    if (JAC) { scr[43 * SST] = kf; scr[44 * SST] = kb; }
    const double pf = kf*(c4)*(c9);
    const double pb = kb*(c8)*(c3);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c4)*(c9), Pb = (c8)*(c3);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[3] += 1.0 * dqT;
      dT[4] += -1.0 * dqT;
      dT[8] += 1.0 * dqT;
      dT[9] += -1.0 * dqT;
    }
    w3 = w3 + 1.0 * q;
    w4 = -1.0 * q;
    w8 = w8 + 1.0 * q;
    w9 = w9 + -1.0 * q;
  }
  {  // reaction 4: HO2 + O <=> O2 + OH
    const double2 a0 = A0_4, a1 = A1_4;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    if (JAC) { scr[45 * SST] = kf; scr[46 * SST] = kb; }
    const double pf = kf*(c0)*(c9);
    const double pb = kb*(c18)*(c3);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c0)*(c9), Pb = (c18)*(c3);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[0] += -1.0 * dqT;
      dT[3] += 1.0 * dqT;
      dT[9] += -1.0 * dqT;
      dT[18] += 1.0 * dqT;
    }
    w0 = -1.0 * q;
    w3 = w3 + 1.0 * q;
    w9 = w9 + -1.0 * q;
    w18 = w18 + 1.0 * q;
  }
  {  // reaction 5: H2O2 + O <=> HO2 + OH
    const double2 a0 = A0_5, a1 = A1_5;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    if (JAC) { scr[47 * SST] = kf; scr[48 * SST] = kb; }
    const double pf = kf*(c1)*(c9);
    const double pb = kb*(c0)*(c3);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c1)*(c9), Pb = (c0)*(c3);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[0] += 1.0 * dqT;
      dT[1] += -1.0 * dqT;
      dT[3] += 1.0 * dqT;
      dT[9] += -1.0 * dqT;
    }
    w0 = w0 + 1.0 * q;
    w1 = -1.0 * q;
    w3 = w3 + 1.0 * q;
    w9 = w9 + -1.0 * q;
  }
  {  // reaction 6: CH2 + O <=> H + HCO
    const double2 a0 = A0_6, a1 = A1_6;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    if (JAC) { scr[49 * SST] = kf; scr[50 * SST] = kb; }
    const double pf = kf*(c19)*(c9);
    const double pb = kb*(c8)*(c5);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c19)*(c9), Pb = (c8)*(c5);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[5] += 1.0 * dqT;
      dT[8] += 1.0 * dqT;
      dT[9] += -1.0 * dqT;
      dT[19] += -1.0 * dqT;
    }
    w5 = 1.0 * q;
    w8 = w8 + 1.0 * q;
    w9 = w9 + -1.0 * q;
    w19 = -1.0 * q;
  }
  {  // reaction 7: CH2(S) + O <=> CO + H2
    const double2 a0 = A0_7, a1 = A1_7;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    if (JAC) { scr[51 * SST] = kf; scr[52 * SST] = kb; }
    const double pf = kf*(c7)*(c9);
    const double pb = kb*(c2)*(c4);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c7)*(c9), Pb = (c2)*(c4);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += 1.0 * dqT;
      dT[4] += 1.0 * dqT;
      dT[7] += -1.0 * dqT;
      dT[9] += -1.0 * dqT;
    }
    w2 = 1.0 * q;
    w4 = w4 + 1.0 * q;
    w7 = -1.0 * q;
    w9 = w9 + -1.0 * q;
  }
  {  // reaction 8: CH2(S) + O <=> H + HCO
    const double2 a0 = A0_8, a1 = A1_8;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    if (JAC) { scr[53 * SST] = kf; scr[54 * SST] = kb; }
    const double pf = kf*(c7)*(c9);
    const double pb = kb*(c8)*(c5);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c7)*(c9), Pb = (c8)*(c5);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[5] += 1.0 * dqT;
      dT[7] += -1.0 * dqT;
      dT[8] += 1.0 * dqT;
      dT[9] += -1.0 * dqT;
    }
    w5 = w5 + 1.0 * q;
    w7 = w7 + -1.0 * q;
    w8 = w8 + 1.0 * q;
    w9 = w9 + -1.0 * q;
  }
  {  // reaction 9: CH3 + O <=> CH2O + H
    const double2 a0 = A0_9, a1 = A1_9;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    if (JAC) { scr[55 * SST] = kf; scr[56 * SST] = kb; }
    const double pf = kf*(c11)*(c9);
    const double pb = kb*(c13)*(c8);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c11)*(c9), Pb = (c13)*(c8);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[8] += 1.0 * dqT;
      dT[9] += -1.0 * dqT;
      dT[11] += -1.0 * dqT;
      dT[13] += 1.0 * dqT;
    }
    w8 = w8 + 1.0 * q;
    w9 = w9 + -1.0 * q;
    w11 = -1.0 * q;
    w13 = 1.0 * q;
  }
  {  // reaction 10: CH4 + O <=> CH3 + OH
    const double2 a0 = A0_10, a1 = A1_10;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    if (JAC) { scr[57 * SST] = kf; scr[58 * SST] = kb; }
    const double pf = kf*(c21)*(c9);
    const double pb = kb*(c11)*(c3);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c21)*(c9), Pb = (c11)*(c3);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[3] += 1.0 * dqT;
      dT[9] += -1.0 * dqT;
      dT[11] += 1.0 * dqT;
      dT[21] += -1.0 * dqT;
    }
    w3 = w3 + 1.0 * q;
    w9 = w9 + -1.0 * q;
    w11 = w11 + 1.0 * q;
    w21 = -1.0 * q;
  }
  {  // reaction 11: CO + O + M <=> CO2 + M
    const double2 a0 = A0_11, a1 = A1_11;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c2)*(c9)*M2;
    const double pb = kb*(c23)*M2;
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c2)*(c9), Pb = (c23);
      const double dqT = ((a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb) * M2;
      scr[2 * SST] = kf * Pf - kb * Pb;
      dT[2] += -1.0 * dqT;
      dT[9] += -1.0 * dqT;
      dT[23] += 1.0 * dqT;
    }
    w2 = w2 + -1.0 * q;
    w9 = w9 + -1.0 * q;
    w23 = 1.0 * q;
  }
  {  // reaction 12: HCO + O <=> CO + OH
    const double2 a0 = A0_12, a1 = A1_12;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    if (JAC) { scr[59 * SST] = kf; scr[60 * SST] = kb; }
    const double pf = kf*(c5)*(c9);
    const double pb = kb*(c2)*(c3);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c5)*(c9), Pb = (c2)*(c3);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += 1.0 * dqT;
      dT[3] += 1.0 * dqT;
      dT[5] += -1.0 * dqT;
      dT[9] += -1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w3 = w3 + 1.0 * q;
    w5 = w5 + -1.0 * q;
    w9 = w9 + -1.0 * q;
  }
  {  // reaction 13: HCO + O <=> CO2 + H
    const double2 a0 = A0_13, a1 = A1_13;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    if (JAC) { scr[61 * SST] = kf; scr[62 * SST] = kb; }
    const double pf = kf*(c5)*(c9);
    const double pb = kb*(c23)*(c8);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c5)*(c9), Pb = (c23)*(c8);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[5] += -1.0 * dqT;
      dT[8] += 1.0 * dqT;
      dT[9] += -1.0 * dqT;
      dT[23] += 1.0 * dqT;
    }
    w5 = w5 + -1.0 * q;
    w8 = w8 + 1.0 * q;
    w9 = w9 + -1.0 * q;
    w23 = w23 + 1.0 * q;
  }
  {  // reaction 14: CH2O + O <=> HCO + OH
    const double2 a0 = A0_14, a1 = A1_14;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    if (JAC) { scr[63 * SST] = kf; scr[64 * SST] = kb; }
    const double pf = kf*(c13)*(c9);
    const double pb = kb*(c5)*(c3);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c13)*(c9), Pb = (c5)*(c3);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[3] += 1.0 * dqT;
      dT[5] += 1.0 * dqT;
      dT[9] += -1.0 * dqT;
      dT[13] += -1.0 * dqT;
    }
    w3 = w3 + 1.0 * q;
    w5 = w5 + 1.0 * q;
    w9 = w9 + -1.0 * q;
    w13 = w13 + -1.0 * q;
  }
  {  // reaction 15: CH3O + O <=> CH2O + OH
    const double2 a0 = A0_15, a1 = A1_15;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    if (JAC) { scr[65 * SST] = kf; scr[66 * SST] = kb; }
    const double pf = kf*(c6)*(c9);
    const double pb = kb*(c13)*(c3);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c6)*(c9), Pb = (c13)*(c3);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[3] += 1.0 * dqT;
      dT[6] += -1.0 * dqT;
      dT[9] += -1.0 * dqT;
      dT[13] += 1.0 * dqT;
    }
    w3 = w3 + 1.0 * q;
    w6 = -1.0 * q;
    w9 = w9 + -1.0 * q;
    w13 = w13 + 1.0 * q;
  }
  {  // reaction 16: C2H2 + O <=> C2H + OH
    const double2 a0 = A0_16, a1 = A1_16;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    if (JAC) { scr[67 * SST] = kf; scr[68 * SST] = kb; }
    const double pf = kf*(c12)*(c9);
    const double pb = kb*(c15)*(c3);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c12)*(c9), Pb = (c15)*(c3);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[3] += 1.0 * dqT;
      dT[9] += -1.0 * dqT;
      dT[12] += -1.0 * dqT;
      dT[15] += 1.0 * dqT;
    }
    w3 = w3 + 1.0 * q;
    w9 = w9 + -1.0 * q;
    w12 = -1.0 * q;
    w15 = 1.0 * q;
  }
  const double2 A0_17 = ld2(rk0 + 32), A1_17 = ld2(rk1 + 32);
  const double2 A0_18 = ld2(rk0 + 34), A1_18 = ld2(rk1 + 34);
  const double2 A0_19 = ld2(rk0 + 36), A1_19 = ld2(rk1 + 36);
  const double2 A0_20 = ld2(rk0 + 38), A1_20 = ld2(rk1 + 38);
  const double2 A0_21 = ld2(rk0 + 40), A1_21 = ld2(rk1 + 40);
  const double2 A0_22 = ld2(rk0 + 42), A1_22 = ld2(rk1 + 42);
  const double2 A0_23 = ld2(rk0 + 44), A1_23 = ld2(rk1 + 44);
  const double2 A0_24 = ld2(rk0 + 46), A1_24 = ld2(rk1 + 46);
  const double2 A0_25 = ld2(rk0 + 48), A1_25 = ld2(rk1 + 48);
  const double2 A0_26 = ld2(rk0 + 50), A1_26 = ld2(rk1 + 50);
  const double2 A0_27 = ld2(rk0 + 52), A1_27 = ld2(rk1 + 52);
  const double2 A0_28 = ld2(rk0 + 54), A1_28 = ld2(rk1 + 54);
  const double2 A0_29 = ld2(rk0 + 56), A1_29 = ld2(rk1 + 56);
  const double2 A0_30 = ld2(rk0 + 58), A1_30 = ld2(rk1 + 58);
  const double2 A0_31 = ld2(rk0 + 60), A1_31 = ld2(rk1 + 60);
  const double2 A0_32 = ld2(rk0 + 62), A1_32 = ld2(rk1 + 62);
  {  // reaction 17: C2H2 + O <=> CH2 + CO
    const double2 a0 = A0_17, a1 = A1_17;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    if (JAC) { scr[69 * SST] = kf; scr[70 * SST] = kb; }
    const double pf = kf*(c12)*(c9);
    const double pb = kb*(c19)*(c2);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c12)*(c9), Pb = (c19)*(c2);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += 1.0 * dqT;
      dT[9] += -1.0 * dqT;
      dT[12] += -1.0 * dqT;
      dT[19] += 1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w9 = w9 + -1.0 * q;
    w12 = w12 + -1.0 * q;
    w19 = w19 + 1.0 * q;
  }
  {  // reaction 18: CO + O2 <=> CO2 + O
    const double2 a0 = A0_18, a1 = A1_18;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    if (JAC) { scr[71 * SST] = kf; scr[72 * SST] = kb; }
    const double pf = kf*(c2)*(c18);
    const double pb = kb*(c23)*(c9);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c2)*(c18), Pb = (c23)*(c9);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += -1.0 * dqT;
      dT[9] += 1.0 * dqT;
      dT[18] += -1.0 * dqT;
      dT[23] += 1.0 * dqT;
    }
    w2 = w2 + -1.0 * q;
    w9 = w9 + 1.0 * q;
    w18 = w18 + -1.0 * q;
    w23 = w23 + 1.0 * q;
  }
  {  // reaction 19: CH2O + O2 <=> HCO + HO2
    const double2 a0 = A0_19, a1 = A1_19;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    if (JAC) { scr[73 * SST] = kf; scr[74 * SST] = kb; }
    const double pf = kf*(c13)*(c18);
    const double pb = kb*(c5)*(c0);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c13)*(c18), Pb = (c5)*(c0);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[0] += 1.0 * dqT;
      dT[5] += 1.0 * dqT;
      dT[13] += -1.0 * dqT;
      dT[18] += -1.0 * dqT;
    }
    w0 = w0 + 1.0 * q;
    w5 = w5 + 1.0 * q;
    w13 = w13 + -1.0 * q;
    w18 = w18 + -1.0 * q;
  }
  {  // reaction 20: H + O2 + M <=> HO2 + M
    const double2 a0 = A0_20, a1 = A1_20;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c8)*(c18)*M3;
    const double pb = kb*(c0)*M3;
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c8)*(c18), Pb = (c0);
      const double dqT = ((a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb) * M3;
      scr[3 * SST] = kf * Pf - kb * Pb;
      dT[0] += 1.0 * dqT;
      dT[8] += -1.0 * dqT;
      dT[18] += -1.0 * dqT;
    }
    w0 = w0 + 1.0 * q;
    w8 = w8 + -1.0 * q;
    w18 = w18 + -1.0 * q;
  }
  {  // reaction 21: H + 2 O2 <=> HO2 + O2
    const double2 a0 = A0_21, a1 = A1_21;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c8)*(c18)*M4;
    const double pb = kb*(c0)*M4;
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c8)*(c18), Pb = (c0);
      const double dqT = ((a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb) * M4;
      scr[4 * SST] = kf * Pf - kb * Pb;
      dT[0] += 1.0 * dqT;
      dT[8] += -1.0 * dqT;
      dT[18] += -1.0 * dqT;
    }
    w0 = w0 + 1.0 * q;
    w8 = w8 + -1.0 * q;
    w18 = w18 + -1.0 * q;
  }
  {  // reaction 22: H + H2O + O2 <=> H2O + HO2
    const double2 a0 = A0_22, a1 = A1_22;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c8)*(c18)*M5;
    const double pb = kb*(c0)*M5;
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c8)*(c18), Pb = (c0);
      const double dqT = ((a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb) * M5;
      scr[5 * SST] = kf * Pf - kb * Pb;
      dT[0] += 1.0 * dqT;
      dT[8] += -1.0 * dqT;
      dT[18] += -1.0 * dqT;
    }
    w0 = w0 + 1.0 * q;
    w8 = w8 + -1.0 * q;
    w18 = w18 + -1.0 * q;
  }
  {  // reaction 23: H + O2 <=> O + OH
    const double2 a0 = A0_23, a1 = A1_23;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    if (JAC) { scr[75 * SST] = kf; scr[76 * SST] = kb; }
    const double pf = kf*(c8)*(c18);
    const double pb = kb*(c9)*(c3);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c8)*(c18), Pb = (c9)*(c3);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[3] += 1.0 * dqT;
      dT[8] += -1.0 * dqT;
      dT[9] += 1.0 * dqT;
      dT[18] += -1.0 * dqT;
    }
    w3 = w3 + 1.0 * q;
    w8 = w8 + -1.0 * q;
    w9 = w9 + 1.0 * q;
    w18 = w18 + -1.0 * q;
  }
  {  // reaction 24: 2 H + M <=> H2 + M
    const double2 a0 = A0_24, a1 = A1_24;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c8*c8)*M6;
    const double pb = kb*(c4)*M6;
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c8*c8), Pb = (c4);
      const double dqT = ((a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb) * M6;
      scr[6 * SST] = kf * Pf - kb * Pb;
      dT[4] += 1.0 * dqT;
      dT[8] += -2.0 * dqT;
    }
    w4 = w4 + 1.0 * q;
    w8 = w8 + -2.0 * q;
  }
  {  // reaction 25: 2 H + H2 <=> 2 H2
    const double2 a0 = A0_25, a1 = A1_25;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c8*c8)*M7;
    const double pb = kb*(c4)*M7;
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c8*c8), Pb = (c4);
      const double dqT = ((a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb) * M7;
      scr[7 * SST] = kf * Pf - kb * Pb;
      dT[4] += 1.0 * dqT;
      dT[8] += -2.0 * dqT;
    }
    w4 = w4 + 1.0 * q;
    w8 = w8 + -2.0 * q;
  }
  {  // reaction 26: 2 H + H2O <=> H2 + H2O
    const double2 a0 = A0_26, a1 = A1_26;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c8*c8)*M5;
    const double pb = kb*(c4)*M5;
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c8*c8), Pb = (c4);
      const double dqT = ((a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb) * M5;
      scr[8 * SST] = kf * Pf - kb * Pb;
      dT[4] += 1.0 * dqT;
      dT[8] += -2.0 * dqT;
    }
    w4 = w4 + 1.0 * q;
    w8 = w8 + -2.0 * q;
  }
  {  // reaction 27: CO2 + 2 H <=> CO2 + H2
    const double2 a0 = A0_27, a1 = A1_27;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c8*c8)*M8;
    const double pb = kb*(c4)*M8;
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c8*c8), Pb = (c4);
      const double dqT = ((a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb) * M8;
      scr[9 * SST] = kf * Pf - kb * Pb;
      dT[4] += 1.0 * dqT;
      dT[8] += -2.0 * dqT;
    }
    w4 = w4 + 1.0 * q;
    w8 = w8 + -2.0 * q;
  }
  {  // reaction 28: H + OH + M <=> H2O + M
    const double2 a0 = A0_28, a1 = A1_28;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c8)*(c3)*M9;
    const double pb = kb*(c16)*M9;
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c8)*(c3), Pb = (c16);
      const double dqT = ((a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb) * M9;
      scr[10 * SST] = kf * Pf - kb * Pb;
      dT[3] += -1.0 * dqT;
      dT[8] += -1.0 * dqT;
      dT[16] += 1.0 * dqT;
    }
    w3 = w3 + -1.0 * q;
    w8 = w8 + -1.0 * q;
    w16 = 1.0 * q;
  }
  {  // reaction 29: H + HO2 <=> H2O + O
    const double2 a0 = A0_29, a1 = A1_29;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    if (JAC) { scr[77 * SST] = kf; scr[78 * SST] = kb; }
    const double pf = kf*(c8)*(c0);
    const double pb = kb*(c16)*(c9);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c8)*(c0), Pb = (c16)*(c9);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[0] += -1.0 * dqT;
      dT[8] += -1.0 * dqT;
      dT[9] += 1.0 * dqT;
      dT[16] += 1.0 * dqT;
    }
    w0 = w0 + -1.0 * q;
    w8 = w8 + -1.0 * q;
    w9 = w9 + 1.0 * q;
    w16 = w16 + 1.0 * q;
  }
  {  // reaction 30: H + HO2 <=> H2 + O2
    const double2 a0 = A0_30, a1 = A1_30;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c8)*(c0);
    const double pb = kb*(c4)*(c18);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c8)*(c0), Pb = (c4)*(c18);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[0] += -1.0 * dqT;
      dT[4] += 1.0 * dqT;
      dT[8] += -1.0 * dqT;
      dT[18] += 1.0 * dqT;
    }
    w0 = w0 + -1.0 * q;
    w4 = w4 + 1.0 * q;
    w8 = w8 + -1.0 * q;
    w18 = w18 + 1.0 * q;
  }
  {  // reaction 31: H + HO2 <=> 2 OH
    const double2 a0 = A0_31, a1 = A1_31;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c8)*(c0);
    const double pb = kb*(c3*c3);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c8)*(c0), Pb = (c3*c3);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[0] += -1.0 * dqT;
      dT[3] += 2.0 * dqT;
      dT[8] += -1.0 * dqT;
    }
    w0 = w0 + -1.0 * q;
    w3 = w3 + 2.0 * q;
    w8 = w8 + -1.0 * q;
  }
  {  // reaction 32: H + H2O2 <=> H2 + HO2
    const double2 a0 = A0_32, a1 = A1_32;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c8)*(c1);
    const double pb = kb*(c4)*(c0);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c8)*(c1), Pb = (c4)*(c0);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[0] += 1.0 * dqT;
      dT[1] += -1.0 * dqT;
      dT[4] += 1.0 * dqT;
      dT[8] += -1.0 * dqT;
    }
    w0 = w0 + 1.0 * q;
    w1 = w1 + -1.0 * q;
    w4 = w4 + 1.0 * q;
    w8 = w8 + -1.0 * q;
  }
  const double2 A0_33 = ld2(rk0 + 64), A1_33 = ld2(rk1 + 64);
  const double2 A0_34 = ld2(rk0 + 66), A1_34 = ld2(rk1 + 66);
  const double2 A0_35 = ld2(rk0 + 68), A1_35 = ld2(rk1 + 68);
  const double2 A0_36 = ld2(rk0 + 70), A1_36 = ld2(rk1 + 70);
  const double2 A0_37 = ld2(rk0 + 72), A1_37 = ld2(rk1 + 72);
  const double2 A0_38 = ld2(rk0 + 74), A1_38 = ld2(rk1 + 74);
  const double2 A0_39 = ld2(rk0 + 76), A1_39 = ld2(rk1 + 76);
  const double2 A0_40 = ld2(rk0 + 78), A1_40 = ld2(rk1 + 78);
  const double2 A0_41 = ld2(rk0 + 80), A1_41 = ld2(rk1 + 80);
  {  // reaction 33: H + H2O2 <=> H2O + OH
    const double2 a0 = A0_33, a1 = A1_33;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c8)*(c1);
    const double pb = kb*(c16)*(c3);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c8)*(c1), Pb = (c16)*(c3);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[1] += -1.0 * dqT;
      dT[3] += 1.0 * dqT;
      dT[8] += -1.0 * dqT;
      dT[16] += 1.0 * dqT;
    }
    w1 = w1 + -1.0 * q;
    w3 = w3 + 1.0 * q;
    w8 = w8 + -1.0 * q;
    w16 = w16 + 1.0 * q;
  }
  {  // reaction 34: CH2 + H (+M) <=> CH3 (+M)
    const double2 x0 = ld2(rk0 + 204), y0 = ld2(rk0 + 206);
    const double2 x1 = ld2(rk1 + 204), y1 = ld2(rk1 + 206);
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
    const double pf = k1*(c19)*(c8);
    const double pb = k2*(c11);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c19)*(c8), Pb = (c11);
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
      dT[8] += -1.0 * dqT;
      dT[11] += 1.0 * dqT;
      dT[19] += -1.0 * dqT;
    }
    w8 = w8 + -1.0 * q;
    w11 = w11 + 1.0 * q;
    w19 = w19 + -1.0 * q;
  }
  {  // reaction 35: CH3 + H (+M) <=> CH4 (+M)
    const double2 x0 = ld2(rk0 + 208), y0 = ld2(rk0 + 210);
    const double2 x1 = ld2(rk1 + 208), y1 = ld2(rk1 + 210);
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
    const double pf = k1*(c11)*(c8);
    const double pb = k2*(c21);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c11)*(c8), Pb = (c21);
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
      dT[8] += -1.0 * dqT;
      dT[11] += -1.0 * dqT;
      dT[21] += 1.0 * dqT;
    }
    w8 = w8 + -1.0 * q;
    w11 = w11 + -1.0 * q;
    w21 = w21 + 1.0 * q;
  }
  {  // reaction 36: CH4 + H <=> CH3 + H2
    const double2 a0 = A0_34, a1 = A1_34;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c21)*(c8);
    const double pb = kb*(c11)*(c4);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c21)*(c8), Pb = (c11)*(c4);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[4] += 1.0 * dqT;
      dT[8] += -1.0 * dqT;
      dT[11] += 1.0 * dqT;
      dT[21] += -1.0 * dqT;
    }
    w4 = w4 + 1.0 * q;
    w8 = w8 + -1.0 * q;
    w11 = w11 + 1.0 * q;
    w21 = w21 + -1.0 * q;
  }
  {  // reaction 37: H + HCO (+M) <=> CH2O (+M)
    const double2 x0 = ld2(rk0 + 212), y0 = ld2(rk0 + 214);
    const double2 x1 = ld2(rk1 + 212), y1 = ld2(rk1 + 214);
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
    const double pf = k1*(c8)*(c5);
    const double pb = k2*(c13);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c8)*(c5), Pb = (c13);
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
      dT[5] += -1.0 * dqT;
      dT[8] += -1.0 * dqT;
      dT[13] += 1.0 * dqT;
    }
    w5 = w5 + -1.0 * q;
    w8 = w8 + -1.0 * q;
    w13 = w13 + 1.0 * q;
  }
  {  // reaction 38: H + HCO <=> CO + H2
    const double2 a0 = A0_35, a1 = A1_35;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c8)*(c5);
    const double pb = kb*(c2)*(c4);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c8)*(c5), Pb = (c2)*(c4);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += 1.0 * dqT;
      dT[4] += 1.0 * dqT;
      dT[5] += -1.0 * dqT;
      dT[8] += -1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w4 = w4 + 1.0 * q;
    w5 = w5 + -1.0 * q;
    w8 = w8 + -1.0 * q;
  }
  {  // reaction 39: CH2O + H (+M) <=> CH3O (+M)
    const double2 x0 = ld2(rk0 + 216), y0 = ld2(rk0 + 218);
    const double2 x1 = ld2(rk1 + 216), y1 = ld2(rk1 + 218);
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
    const double pf = k1*(c13)*(c8);
    const double pb = k2*(c6);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c13)*(c8), Pb = (c6);
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
      dT[6] += 1.0 * dqT;
      dT[8] += -1.0 * dqT;
      dT[13] += -1.0 * dqT;
    }
    w6 = w6 + 1.0 * q;
    w8 = w8 + -1.0 * q;
    w13 = w13 + -1.0 * q;
  }
  {  // reaction 40: CH2O + H <=> H2 + HCO
    const double2 a0 = A0_36, a1 = A1_36;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c13)*(c8);
    const double pb = kb*(c4)*(c5);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c13)*(c8), Pb = (c4)*(c5);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[4] += 1.0 * dqT;
      dT[5] += 1.0 * dqT;
      dT[8] += -1.0 * dqT;
      dT[13] += -1.0 * dqT;
    }
    w4 = w4 + 1.0 * q;
    w5 = w5 + 1.0 * q;
    w8 = w8 + -1.0 * q;
    w13 = w13 + -1.0 * q;
  }
  {  // reaction 41: CH3O + H <=> CH2O + H2
    const double2 a0 = A0_37, a1 = A1_37;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c6)*(c8);
    const double pb = kb*(c13)*(c4);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c6)*(c8), Pb = (c13)*(c4);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[4] += 1.0 * dqT;
      dT[6] += -1.0 * dqT;
      dT[8] += -1.0 * dqT;
      dT[13] += 1.0 * dqT;
    }
    w4 = w4 + 1.0 * q;
    w6 = w6 + -1.0 * q;
    w8 = w8 + -1.0 * q;
    w13 = w13 + 1.0 * q;
  }
  {  // reaction 42: CH3O + H <=> CH3 + OH
    const double2 a0 = A0_38, a1 = A1_38;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c6)*(c8);
    const double pb = kb*(c11)*(c3);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c6)*(c8), Pb = (c11)*(c3);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[3] += 1.0 * dqT;
      dT[6] += -1.0 * dqT;
      dT[8] += -1.0 * dqT;
      dT[11] += 1.0 * dqT;
    }
    w3 = w3 + 1.0 * q;
    w6 = w6 + -1.0 * q;
    w8 = w8 + -1.0 * q;
    w11 = w11 + 1.0 * q;
  }
  {  // reaction 43: CH3O + H <=> CH2(S) + H2O
    const double2 a0 = A0_39, a1 = A1_39;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c6)*(c8);
    const double pb = kb*(c7)*(c16);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c6)*(c8), Pb = (c7)*(c16);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[6] += -1.0 * dqT;
      dT[7] += 1.0 * dqT;
      dT[8] += -1.0 * dqT;
      dT[16] += 1.0 * dqT;
    }
    w6 = w6 + -1.0 * q;
    w7 = w7 + 1.0 * q;
    w8 = w8 + -1.0 * q;
    w16 = w16 + 1.0 * q;
  }
  {  // reaction 44: C2H + H (+M) <=> C2H2 (+M)
    const double2 x0 = ld2(rk0 + 220), y0 = ld2(rk0 + 222);
    const double2 x1 = ld2(rk1 + 220), y1 = ld2(rk1 + 222);
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
    const double pf = k1*(c15)*(c8);
    const double pb = k2*(c12);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c15)*(c8), Pb = (c12);
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
      dT[8] += -1.0 * dqT;
      dT[12] += 1.0 * dqT;
      dT[15] += -1.0 * dqT;
    }
    w8 = w8 + -1.0 * q;
    w12 = w12 + 1.0 * q;
    w15 = w15 + -1.0 * q;
  }
  {  // reaction 45: CO + H2 (+M) <=> CH2O (+M)
    const double2 x0 = ld2(rk0 + 224), y0 = ld2(rk0 + 226);
    const double2 x1 = ld2(rk1 + 224), y1 = ld2(rk1 + 226);
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
    const double pf = k1*(c2)*(c4);
    const double pb = k2*(c13);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c2)*(c4), Pb = (c13);
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
      dT[4] += -1.0 * dqT;
      dT[13] += 1.0 * dqT;
    }
    w2 = w2 + -1.0 * q;
    w4 = w4 + -1.0 * q;
    w13 = w13 + 1.0 * q;
  }
  {  // reaction 46: H2 + OH <=> H + H2O
    const double2 a0 = A0_40, a1 = A1_40;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c4)*(c3);
    const double pb = kb*(c8)*(c16);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c4)*(c3), Pb = (c8)*(c16);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[3] += -1.0 * dqT;
      dT[4] += -1.0 * dqT;
      dT[8] += 1.0 * dqT;
      dT[16] += 1.0 * dqT;
    }
    w3 = w3 + -1.0 * q;
    w4 = w4 + -1.0 * q;
    w8 = w8 + 1.0 * q;
    w16 = w16 + 1.0 * q;
  }
  {  // reaction 47: 2 OH (+M) <=> H2O2 (+M)
    const double2 x0 = ld2(rk0 + 228), y0 = ld2(rk0 + 230);
    const double2 x1 = ld2(rk1 + 228), y1 = ld2(rk1 + 230);
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
    const double pf = k1*(c3*c3);
    const double pb = k2*(c1);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c3*c3), Pb = (c1);
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
      dT[1] += 1.0 * dqT;
      dT[3] += -2.0 * dqT;
    }
    w1 = w1 + 1.0 * q;
    w3 = w3 + -2.0 * q;
  }
  {  // reaction 48: 2 OH <=> H2O + O
    const double2 a0 = A0_41, a1 = A1_41;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c3*c3);
    const double pb = kb*(c16)*(c9);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c3*c3), Pb = (c16)*(c9);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[3] += -2.0 * dqT;
      dT[9] += 1.0 * dqT;
      dT[16] += 1.0 * dqT;
    }
    w3 = w3 + -2.0 * q;
    w9 = w9 + 1.0 * q;
    w16 = w16 + 1.0 * q;
  }
  const double2 A0_42 = ld2(rk0 + 82), A1_42 = ld2(rk1 + 82);
  const double2 A0_43 = ld2(rk0 + 84), A1_43 = ld2(rk1 + 84);
  const double2 A0_44 = ld2(rk0 + 86), A1_44 = ld2(rk1 + 86);
  const double2 A0_45 = ld2(rk0 + 88), A1_45 = ld2(rk1 + 88);
  const double2 A0_46 = ld2(rk0 + 90), A1_46 = ld2(rk1 + 90);
  const double2 A0_47 = ld2(rk0 + 92), A1_47 = ld2(rk1 + 92);
  const double2 A0_48 = ld2(rk0 + 94), A1_48 = ld2(rk1 + 94);
  const double2 A0_49 = ld2(rk0 + 96), A1_49 = ld2(rk1 + 96);
  const double2 A0_50 = ld2(rk0 + 98), A1_50 = ld2(rk1 + 98);
  const double2 A0_51 = ld2(rk0 + 100), A1_51 = ld2(rk1 + 100);
  const double2 A0_52 = ld2(rk0 + 102), A1_52 = ld2(rk1 + 102);
  const double2 A0_53 = ld2(rk0 + 104), A1_53 = ld2(rk1 + 104);
  const double2 A0_54 = ld2(rk0 + 106), A1_54 = ld2(rk1 + 106);
  const double2 A0_55 = ld2(rk0 + 108), A1_55 = ld2(rk1 + 108);
  const double2 A0_56 = ld2(rk0 + 110), A1_56 = ld2(rk1 + 110);
  const double2 A0_57 = ld2(rk0 + 112), A1_57 = ld2(rk1 + 112);
  {  // reaction 49: HO2 + OH <=> H2O + O2
    const double2 a0 = A0_42, a1 = A1_42;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c0)*(c3);
    const double pb = kb*(c16)*(c18);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c0)*(c3), Pb = (c16)*(c18);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[0] += -1.0 * dqT;
      dT[3] += -1.0 * dqT;
      dT[16] += 1.0 * dqT;
      dT[18] += 1.0 * dqT;
    }
    w0 = w0 + -1.0 * q;
    w3 = w3 + -1.0 * q;
    w16 = w16 + 1.0 * q;
    w18 = w18 + 1.0 * q;
  }
  {  // reaction 50: H2O2 + OH <=> H2O + HO2
    const double2 a0 = A0_43, a1 = A1_43;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c1)*(c3);
    const double pb = kb*(c16)*(c0);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c1)*(c3), Pb = (c16)*(c0);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[0] += 1.0 * dqT;
      dT[1] += -1.0 * dqT;
      dT[3] += -1.0 * dqT;
      dT[16] += 1.0 * dqT;
    }
    w0 = w0 + 1.0 * q;
    w1 = w1 + -1.0 * q;
    w3 = w3 + -1.0 * q;
    w16 = w16 + 1.0 * q;
  }
  {  // reaction 51: H2O2 + OH <=> H2O + HO2
    const double2 a0 = A0_44, a1 = A1_44;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c1)*(c3);
    const double pb = kb*(c16)*(c0);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c1)*(c3), Pb = (c16)*(c0);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[0] += 1.0 * dqT;
      dT[1] += -1.0 * dqT;
      dT[3] += -1.0 * dqT;
      dT[16] += 1.0 * dqT;
    }
    w0 = w0 + 1.0 * q;
    w1 = w1 + -1.0 * q;
    w3 = w3 + -1.0 * q;
    w16 = w16 + 1.0 * q;
  }
  {  // reaction 52: CH2 + OH <=> CH2O + H
    const double2 a0 = A0_45, a1 = A1_45;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c19)*(c3);
    const double pb = kb*(c13)*(c8);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c19)*(c3), Pb = (c13)*(c8);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[3] += -1.0 * dqT;
      dT[8] += 1.0 * dqT;
      dT[13] += 1.0 * dqT;
      dT[19] += -1.0 * dqT;
    }
    w3 = w3 + -1.0 * q;
    w8 = w8 + 1.0 * q;
    w13 = w13 + 1.0 * q;
    w19 = w19 + -1.0 * q;
  }
  {  // reaction 53: CH2(S) + OH <=> CH2O + H
    const double2 a0 = A0_46, a1 = A1_46;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c7)*(c3);
    const double pb = kb*(c13)*(c8);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c7)*(c3), Pb = (c13)*(c8);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[3] += -1.0 * dqT;
      dT[7] += -1.0 * dqT;
      dT[8] += 1.0 * dqT;
      dT[13] += 1.0 * dqT;
    }
    w3 = w3 + -1.0 * q;
    w7 = w7 + -1.0 * q;
    w8 = w8 + 1.0 * q;
    w13 = w13 + 1.0 * q;
  }
  {  // reaction 54: CH3 + OH <=> CH2 + H2O
    const double2 a0 = A0_47, a1 = A1_47;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c11)*(c3);
    const double pb = kb*(c19)*(c16);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c11)*(c3), Pb = (c19)*(c16);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[3] += -1.0 * dqT;
      dT[11] += -1.0 * dqT;
      dT[16] += 1.0 * dqT;
      dT[19] += 1.0 * dqT;
    }
    w3 = w3 + -1.0 * q;
    w11 = w11 + -1.0 * q;
    w16 = w16 + 1.0 * q;
    w19 = w19 + 1.0 * q;
  }
  {  // reaction 55: CH3 + OH <=> CH2(S) + H2O
    const double2 a0 = A0_48, a1 = A1_48;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c11)*(c3);
    const double pb = kb*(c7)*(c16);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c11)*(c3), Pb = (c7)*(c16);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[3] += -1.0 * dqT;
      dT[7] += 1.0 * dqT;
      dT[11] += -1.0 * dqT;
      dT[16] += 1.0 * dqT;
    }
    w3 = w3 + -1.0 * q;
    w7 = w7 + 1.0 * q;
    w11 = w11 + -1.0 * q;
    w16 = w16 + 1.0 * q;
  }
  {  // reaction 56: CH4 + OH <=> CH3 + H2O
    const double2 a0 = A0_49, a1 = A1_49;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c21)*(c3);
    const double pb = kb*(c11)*(c16);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c21)*(c3), Pb = (c11)*(c16);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[3] += -1.0 * dqT;
      dT[11] += 1.0 * dqT;
      dT[16] += 1.0 * dqT;
      dT[21] += -1.0 * dqT;
    }
    w3 = w3 + -1.0 * q;
    w11 = w11 + 1.0 * q;
    w16 = w16 + 1.0 * q;
    w21 = w21 + -1.0 * q;
  }
  {  // reaction 57: CO + OH <=> CO2 + H
    const double2 a0 = A0_50, a1 = A1_50;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c2)*(c3);
    const double pb = kb*(c23)*(c8);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c2)*(c3), Pb = (c23)*(c8);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += -1.0 * dqT;
      dT[3] += -1.0 * dqT;
      dT[8] += 1.0 * dqT;
      dT[23] += 1.0 * dqT;
    }
    w2 = w2 + -1.0 * q;
    w3 = w3 + -1.0 * q;
    w8 = w8 + 1.0 * q;
    w23 = w23 + 1.0 * q;
  }
  {  // reaction 58: HCO + OH <=> CO + H2O
    const double2 a0 = A0_51, a1 = A1_51;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c5)*(c3);
    const double pb = kb*(c2)*(c16);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c5)*(c3), Pb = (c2)*(c16);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += 1.0 * dqT;
      dT[3] += -1.0 * dqT;
      dT[5] += -1.0 * dqT;
      dT[16] += 1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w3 = w3 + -1.0 * q;
    w5 = w5 + -1.0 * q;
    w16 = w16 + 1.0 * q;
  }
  {  // reaction 59: CH2O + OH <=> H2O + HCO
    const double2 a0 = A0_52, a1 = A1_52;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c13)*(c3);
    const double pb = kb*(c16)*(c5);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c13)*(c3), Pb = (c16)*(c5);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[3] += -1.0 * dqT;
      dT[5] += 1.0 * dqT;
      dT[13] += -1.0 * dqT;
      dT[16] += 1.0 * dqT;
    }
    w3 = w3 + -1.0 * q;
    w5 = w5 + 1.0 * q;
    w13 = w13 + -1.0 * q;
    w16 = w16 + 1.0 * q;
  }
  {  // reaction 60: CH3O + OH <=> CH2O + H2O
    const double2 a0 = A0_53, a1 = A1_53;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c6)*(c3);
    const double pb = kb*(c13)*(c16);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c6)*(c3), Pb = (c13)*(c16);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[3] += -1.0 * dqT;
      dT[6] += -1.0 * dqT;
      dT[13] += 1.0 * dqT;
      dT[16] += 1.0 * dqT;
    }
    w3 = w3 + -1.0 * q;
    w6 = w6 + -1.0 * q;
    w13 = w13 + 1.0 * q;
    w16 = w16 + 1.0 * q;
  }
  {  // reaction 61: C2H2 + OH <=> C2H + H2O
    const double2 a0 = A0_54, a1 = A1_54;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c12)*(c3);
    const double pb = kb*(c15)*(c16);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c12)*(c3), Pb = (c15)*(c16);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[3] += -1.0 * dqT;
      dT[12] += -1.0 * dqT;
      dT[15] += 1.0 * dqT;
      dT[16] += 1.0 * dqT;
    }
    w3 = w3 + -1.0 * q;
    w12 = w12 + -1.0 * q;
    w15 = w15 + 1.0 * q;
    w16 = w16 + 1.0 * q;
  }
  {  // reaction 62: C2H2 + OH <=> CH3 + CO
    const double2 a0 = A0_55, a1 = A1_55;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c12)*(c3);
    const double pb = kb*(c11)*(c2);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c12)*(c3), Pb = (c11)*(c2);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += 1.0 * dqT;
      dT[3] += -1.0 * dqT;
      dT[11] += 1.0 * dqT;
      dT[12] += -1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w3 = w3 + -1.0 * q;
    w11 = w11 + 1.0 * q;
    w12 = w12 + -1.0 * q;
  }
  {  // reaction 63: 2 HO2 <=> H2O2 + O2
    const double2 a0 = A0_56, a1 = A1_56;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c0*c0);
    const double pb = kb*(c1)*(c18);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c0*c0), Pb = (c1)*(c18);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[0] += -2.0 * dqT;
      dT[1] += 1.0 * dqT;
      dT[18] += 1.0 * dqT;
    }
    w0 = w0 + -2.0 * q;
    w1 = w1 + 1.0 * q;
    w18 = w18 + 1.0 * q;
  }
  {  // reaction 64: 2 HO2 <=> H2O2 + O2
    const double2 a0 = A0_57, a1 = A1_57;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c0*c0);
    const double pb = kb*(c1)*(c18);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c0*c0), Pb = (c1)*(c18);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[0] += -2.0 * dqT;
      dT[1] += 1.0 * dqT;
      dT[18] += 1.0 * dqT;
    }
    w0 = w0 + -2.0 * q;
    w1 = w1 + 1.0 * q;
    w18 = w18 + 1.0 * q;
  }
  const double2 A0_58 = ld2(rk0 + 114), A1_58 = ld2(rk1 + 114);
  const double2 A0_59 = ld2(rk0 + 116), A1_59 = ld2(rk1 + 116);
  const double2 A0_60 = ld2(rk0 + 118), A1_60 = ld2(rk1 + 118);
  const double2 A0_61 = ld2(rk0 + 120), A1_61 = ld2(rk1 + 120);
  const double2 A0_62 = ld2(rk0 + 122), A1_62 = ld2(rk1 + 122);
  const double2 A0_63 = ld2(rk0 + 124), A1_63 = ld2(rk1 + 124);
  const double2 A0_64 = ld2(rk0 + 126), A1_64 = ld2(rk1 + 126);
  const double2 A0_65 = ld2(rk0 + 128), A1_65 = ld2(rk1 + 128);
  const double2 A0_66 = ld2(rk0 + 130), A1_66 = ld2(rk1 + 130);
  const double2 A0_67 = ld2(rk0 + 132), A1_67 = ld2(rk1 + 132);
  const double2 A0_68 = ld2(rk0 + 134), A1_68 = ld2(rk1 + 134);
  const double2 A0_69 = ld2(rk0 + 136), A1_69 = ld2(rk1 + 136);
  const double2 A0_70 = ld2(rk0 + 138), A1_70 = ld2(rk1 + 138);
  const double2 A0_71 = ld2(rk0 + 140), A1_71 = ld2(rk1 + 140);
  const double2 A0_72 = ld2(rk0 + 142), A1_72 = ld2(rk1 + 142);
  const double2 A0_73 = ld2(rk0 + 144), A1_73 = ld2(rk1 + 144);
  {  // reaction 65: CH2 + HO2 <=> CH2O + OH
    const double2 a0 = A0_58, a1 = A1_58;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c19)*(c0);
    const double pb = kb*(c13)*(c3);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c19)*(c0), Pb = (c13)*(c3);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[0] += -1.0 * dqT;
      dT[3] += 1.0 * dqT;
      dT[13] += 1.0 * dqT;
      dT[19] += -1.0 * dqT;
    }
    w0 = w0 + -1.0 * q;
    w3 = w3 + 1.0 * q;
    w13 = w13 + 1.0 * q;
    w19 = w19 + -1.0 * q;
  }
  {  // reaction 66: CH3 + HO2 <=> CH4 + O2
    const double2 a0 = A0_59, a1 = A1_59;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c11)*(c0);
    const double pb = kb*(c21)*(c18);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c11)*(c0), Pb = (c21)*(c18);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[0] += -1.0 * dqT;
      dT[11] += -1.0 * dqT;
      dT[18] += 1.0 * dqT;
      dT[21] += 1.0 * dqT;
    }
    w0 = w0 + -1.0 * q;
    w11 = w11 + -1.0 * q;
    w18 = w18 + 1.0 * q;
    w21 = w21 + 1.0 * q;
  }
  {  // reaction 67: CH3 + HO2 <=> CH3O + OH
    const double2 a0 = A0_60, a1 = A1_60;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c11)*(c0);
    const double pb = kb*(c6)*(c3);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c11)*(c0), Pb = (c6)*(c3);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[0] += -1.0 * dqT;
      dT[3] += 1.0 * dqT;
      dT[6] += 1.0 * dqT;
      dT[11] += -1.0 * dqT;
    }
    w0 = w0 + -1.0 * q;
    w3 = w3 + 1.0 * q;
    w6 = w6 + 1.0 * q;
    w11 = w11 + -1.0 * q;
  }
  {  // reaction 68: CO + HO2 <=> CO2 + OH
    const double2 a0 = A0_61, a1 = A1_61;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c2)*(c0);
    const double pb = kb*(c23)*(c3);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c2)*(c0), Pb = (c23)*(c3);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[0] += -1.0 * dqT;
      dT[2] += -1.0 * dqT;
      dT[3] += 1.0 * dqT;
      dT[23] += 1.0 * dqT;
    }
    w0 = w0 + -1.0 * q;
    w2 = w2 + -1.0 * q;
    w3 = w3 + 1.0 * q;
    w23 = w23 + 1.0 * q;
  }
  {  // reaction 69: CH2O + HO2 <=> H2O2 + HCO
    const double2 a0 = A0_62, a1 = A1_62;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c13)*(c0);
    const double pb = kb*(c1)*(c5);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c13)*(c0), Pb = (c1)*(c5);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[0] += -1.0 * dqT;
      dT[1] += 1.0 * dqT;
      dT[5] += 1.0 * dqT;
      dT[13] += -1.0 * dqT;
    }
    w0 = w0 + -1.0 * q;
    w1 = w1 + 1.0 * q;
    w5 = w5 + 1.0 * q;
    w13 = w13 + -1.0 * q;
  }
  {  // reaction 70: CH2 + O2 <=> HCO + OH
    const double2 a0 = A0_63, a1 = A1_63;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c19)*(c18);
    const double pb = kb*(c5)*(c3);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c19)*(c18), Pb = (c5)*(c3);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[3] += 1.0 * dqT;
      dT[5] += 1.0 * dqT;
      dT[18] += -1.0 * dqT;
      dT[19] += -1.0 * dqT;
    }
    w3 = w3 + 1.0 * q;
    w5 = w5 + 1.0 * q;
    w18 = w18 + -1.0 * q;
    w19 = w19 + -1.0 * q;
  }
  {  // reaction 71: CH2 + H2 <=> CH3 + H
    const double2 a0 = A0_64, a1 = A1_64;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c19)*(c4);
    const double pb = kb*(c11)*(c8);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c19)*(c4), Pb = (c11)*(c8);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[4] += -1.0 * dqT;
      dT[8] += 1.0 * dqT;
      dT[11] += 1.0 * dqT;
      dT[19] += -1.0 * dqT;
    }
    w4 = w4 + -1.0 * q;
    w8 = w8 + 1.0 * q;
    w11 = w11 + 1.0 * q;
    w19 = w19 + -1.0 * q;
  }
  {  // reaction 72: 2 CH2 <=> C2H2 + H2
    const double2 a0 = A0_65, a1 = A1_65;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c19*c19);
    const double pb = kb*(c12)*(c4);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c19*c19), Pb = (c12)*(c4);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[4] += 1.0 * dqT;
      dT[12] += 1.0 * dqT;
      dT[19] += -2.0 * dqT;
    }
    w4 = w4 + 1.0 * q;
    w12 = w12 + 1.0 * q;
    w19 = w19 + -2.0 * q;
  }
  {  // reaction 73: CH2 + CH4 <=> 2 CH3
    const double2 a0 = A0_66, a1 = A1_66;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c19)*(c21);
    const double pb = kb*(c11*c11);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c19)*(c21), Pb = (c11*c11);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[11] += 2.0 * dqT;
      dT[19] += -1.0 * dqT;
      dT[21] += -1.0 * dqT;
    }
    w11 = w11 + 2.0 * q;
    w19 = w19 + -1.0 * q;
    w21 = w21 + -1.0 * q;
  }
  {  // reaction 74: CH2(S) + O2 <=> CO + H + OH
    const double2 a0 = A0_67, a1 = A1_67;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    if (JAC) { scr[41 * SST] = kf; scr[42 * SST] = kb; }
    const double pf = kf*(c7)*(c18);
    const double pb = kb*(c2)*(c8)*(c3);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c7)*(c18), Pb = (c2)*(c8)*(c3);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += 1.0 * dqT;
      dT[3] += 1.0 * dqT;
      dT[7] += -1.0 * dqT;
      dT[8] += 1.0 * dqT;
      dT[18] += -1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w3 = w3 + 1.0 * q;
    w7 = w7 + -1.0 * q;
    w8 = w8 + 1.0 * q;
    w18 = w18 + -1.0 * q;
  }
  {  // reaction 75: CH2(S) + O2 <=> CO + H2O
    const double2 a0 = A0_68, a1 = A1_68;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c7)*(c18);
    const double pb = kb*(c2)*(c16);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c7)*(c18), Pb = (c2)*(c16);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += 1.0 * dqT;
      dT[7] += -1.0 * dqT;
      dT[16] += 1.0 * dqT;
      dT[18] += -1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w7 = w7 + -1.0 * q;
    w16 = w16 + 1.0 * q;
    w18 = w18 + -1.0 * q;
  }
  {  // reaction 76: CH2(S) + H2 <=> CH3 + H
    const double2 a0 = A0_69, a1 = A1_69;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c7)*(c4);
    const double pb = kb*(c11)*(c8);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c7)*(c4), Pb = (c11)*(c8);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[4] += -1.0 * dqT;
      dT[7] += -1.0 * dqT;
      dT[8] += 1.0 * dqT;
      dT[11] += 1.0 * dqT;
    }
    w4 = w4 + -1.0 * q;
    w7 = w7 + -1.0 * q;
    w8 = w8 + 1.0 * q;
    w11 = w11 + 1.0 * q;
  }
  {  // reaction 77: CH2(S) + H2O <=> CH2 + H2O
    const double2 a0 = A0_70, a1 = A1_70;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c7)*(c16);
    const double pb = kb*(c19)*(c16);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c7)*(c16), Pb = (c19)*(c16);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[7] += -1.0 * dqT;
      dT[16] += 0.0 * dqT;
      dT[19] += 1.0 * dqT;
    }
    w7 = w7 + -1.0 * q;
    w16 = w16 + 0.0 * q;
    w19 = w19 + 1.0 * q;
  }
  {  // reaction 78: CH2(S) + CH4 <=> 2 CH3
    const double2 a0 = A0_71, a1 = A1_71;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c7)*(c21);
    const double pb = kb*(c11*c11);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c7)*(c21), Pb = (c11*c11);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[7] += -1.0 * dqT;
      dT[11] += 2.0 * dqT;
      dT[21] += -1.0 * dqT;
    }
    w7 = w7 + -1.0 * q;
    w11 = w11 + 2.0 * q;
    w21 = w21 + -1.0 * q;
  }
  {  // reaction 79: CH2(S) + CO <=> CH2 + CO
    const double2 a0 = A0_72, a1 = A1_72;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c7)*(c2);
    const double pb = kb*(c19)*(c2);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c7)*(c2), Pb = (c19)*(c2);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += 0.0 * dqT;
      dT[7] += -1.0 * dqT;
      dT[19] += 1.0 * dqT;
    }
    w2 = w2 + 0.0 * q;
    w7 = w7 + -1.0 * q;
    w19 = w19 + 1.0 * q;
  }
  {  // reaction 80: CH2(S) + CO2 <=> CH2 + CO2
    const double2 a0 = A0_73, a1 = A1_73;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c7)*(c23);
    const double pb = kb*(c19)*(c23);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c7)*(c23), Pb = (c19)*(c23);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[7] += -1.0 * dqT;
      dT[19] += 1.0 * dqT;
      dT[23] += 0.0 * dqT;
    }
    w7 = w7 + -1.0 * q;
    w19 = w19 + 1.0 * q;
    w23 = w23 + 0.0 * q;
  }
  const double2 A0_74 = ld2(rk0 + 146), A1_74 = ld2(rk1 + 146);
  const double2 A0_75 = ld2(rk0 + 148), A1_75 = ld2(rk1 + 148);
  const double2 A0_76 = ld2(rk0 + 150), A1_76 = ld2(rk1 + 150);
  const double2 A0_77 = ld2(rk0 + 152), A1_77 = ld2(rk1 + 152);
  const double2 A0_78 = ld2(rk0 + 154), A1_78 = ld2(rk1 + 154);
  const double2 A0_79 = ld2(rk0 + 156), A1_79 = ld2(rk1 + 156);
  const double2 A0_80 = ld2(rk0 + 158), A1_80 = ld2(rk1 + 158);
  const double2 A0_81 = ld2(rk0 + 160), A1_81 = ld2(rk1 + 160);
  const double2 A0_82 = ld2(rk0 + 162), A1_82 = ld2(rk1 + 162);
  const double2 A0_83 = ld2(rk0 + 164), A1_83 = ld2(rk1 + 164);
  const double2 A0_84 = ld2(rk0 + 166), A1_84 = ld2(rk1 + 166);
  const double2 A0_85 = ld2(rk0 + 168), A1_85 = ld2(rk1 + 168);
  const double2 A0_86 = ld2(rk0 + 170), A1_86 = ld2(rk1 + 170);
  const double2 A0_87 = ld2(rk0 + 172), A1_87 = ld2(rk1 + 172);
  const double2 A0_88 = ld2(rk0 + 174), A1_88 = ld2(rk1 + 174);
  {  // reaction 81: CH2(S) + CO2 <=> CH2O + CO
    const double2 a0 = A0_74, a1 = A1_74;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c7)*(c23);
    const double pb = kb*(c13)*(c2);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c7)*(c23), Pb = (c13)*(c2);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += 1.0 * dqT;
      dT[7] += -1.0 * dqT;
      dT[13] += 1.0 * dqT;
      dT[23] += -1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w7 = w7 + -1.0 * q;
    w13 = w13 + 1.0 * q;
    w23 = w23 + -1.0 * q;
  }
  {  // reaction 82: CH3 + O2 <=> CH3O + O
    const double2 a0 = A0_75, a1 = A1_75;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c11)*(c18);
    const double pb = kb*(c6)*(c9);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c11)*(c18), Pb = (c6)*(c9);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[6] += 1.0 * dqT;
      dT[9] += 1.0 * dqT;
      dT[11] += -1.0 * dqT;
      dT[18] += -1.0 * dqT;
    }
    w6 = w6 + 1.0 * q;
    w9 = w9 + 1.0 * q;
    w11 = w11 + -1.0 * q;
    w18 = w18 + -1.0 * q;
  }
  {  // reaction 83: CH3 + O2 <=> CH2O + OH
    const double2 a0 = A0_76, a1 = A1_76;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c11)*(c18);
    const double pb = kb*(c13)*(c3);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c11)*(c18), Pb = (c13)*(c3);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[3] += 1.0 * dqT;
      dT[11] += -1.0 * dqT;
      dT[13] += 1.0 * dqT;
      dT[18] += -1.0 * dqT;
    }
    w3 = w3 + 1.0 * q;
    w11 = w11 + -1.0 * q;
    w13 = w13 + 1.0 * q;
    w18 = w18 + -1.0 * q;
  }
  {  // reaction 84: CH3 + H2O2 <=> CH4 + HO2
    const double2 a0 = A0_77, a1 = A1_77;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c11)*(c1);
    const double pb = kb*(c21)*(c0);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c11)*(c1), Pb = (c21)*(c0);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[0] += 1.0 * dqT;
      dT[1] += -1.0 * dqT;
      dT[11] += -1.0 * dqT;
      dT[21] += 1.0 * dqT;
    }
    w0 = w0 + 1.0 * q;
    w1 = w1 + -1.0 * q;
    w11 = w11 + -1.0 * q;
    w21 = w21 + 1.0 * q;
  }
  {  // reaction 85: 2 CH3 (+M) <=> C2H6 (+M)
    const double2 x0 = ld2(rk0 + 232), y0 = ld2(rk0 + 234);
    const double2 x1 = ld2(rk1 + 232), y1 = ld2(rk1 + 234);
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
    const double pf = k1*(c11*c11);
    const double pb = k2*(c20);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c11*c11), Pb = (c20);
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
      dT[11] += -2.0 * dqT;
      dT[20] += 1.0 * dqT;
    }
    w11 = w11 + -2.0 * q;
    w20 = 1.0 * q;
  }
  {  // reaction 86: CH3 + HCO <=> CH4 + CO
    const double2 a0 = A0_78, a1 = A1_78;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c11)*(c5);
    const double pb = kb*(c21)*(c2);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c11)*(c5), Pb = (c21)*(c2);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += 1.0 * dqT;
      dT[5] += -1.0 * dqT;
      dT[11] += -1.0 * dqT;
      dT[21] += 1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w5 = w5 + -1.0 * q;
    w11 = w11 + -1.0 * q;
    w21 = w21 + 1.0 * q;
  }
  {  // reaction 87: CH2O + CH3 <=> CH4 + HCO
    const double2 a0 = A0_79, a1 = A1_79;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c13)*(c11);
    const double pb = kb*(c21)*(c5);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c13)*(c11), Pb = (c21)*(c5);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[5] += 1.0 * dqT;
      dT[11] += -1.0 * dqT;
      dT[13] += -1.0 * dqT;
      dT[21] += 1.0 * dqT;
    }
    w5 = w5 + 1.0 * q;
    w11 = w11 + -1.0 * q;
    w13 = w13 + -1.0 * q;
    w21 = w21 + 1.0 * q;
  }
  {  // reaction 88: H2O + HCO <=> CO + H + H2O
    const double2 a0 = A0_80, a1 = A1_80;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c5)*M5;
    const double pb = kb*(c2)*(c8)*M5;
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c5), Pb = (c2)*(c8);
      const double dqT = ((a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb) * M5;
      scr[35 * SST] = kf * Pf - kb * Pb;
      dT[2] += 1.0 * dqT;
      dT[5] += -1.0 * dqT;
      dT[8] += 1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w5 = w5 + -1.0 * q;
    w8 = w8 + 1.0 * q;
  }
  {  // reaction 89: HCO + M <=> CO + H + M
    const double2 a0 = A0_81, a1 = A1_81;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c5)*M10;
    const double pb = kb*(c2)*(c8)*M10;
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c5), Pb = (c2)*(c8);
      const double dqT = ((a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb) * M10;
      scr[36 * SST] = kf * Pf - kb * Pb;
      dT[2] += 1.0 * dqT;
      dT[5] += -1.0 * dqT;
      dT[8] += 1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w5 = w5 + -1.0 * q;
    w8 = w8 + 1.0 * q;
  }
  {  // reaction 90: HCO + O2 <=> CO + HO2
    const double2 a0 = A0_82, a1 = A1_82;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c5)*(c18);
    const double pb = kb*(c2)*(c0);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c5)*(c18), Pb = (c2)*(c0);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[0] += 1.0 * dqT;
      dT[2] += 1.0 * dqT;
      dT[5] += -1.0 * dqT;
      dT[18] += -1.0 * dqT;
    }
    w0 = w0 + 1.0 * q;
    w2 = w2 + 1.0 * q;
    w5 = w5 + -1.0 * q;
    w18 = w18 + -1.0 * q;
  }
  {  // reaction 91: CH3O + O2 <=> CH2O + HO2
    const double2 a0 = A0_83, a1 = A1_83;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c6)*(c18);
    const double pb = kb*(c13)*(c0);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c6)*(c18), Pb = (c13)*(c0);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[0] += 1.0 * dqT;
      dT[6] += -1.0 * dqT;
      dT[13] += 1.0 * dqT;
      dT[18] += -1.0 * dqT;
    }
    w0 = w0 + 1.0 * q;
    w6 = w6 + -1.0 * q;
    w13 = w13 + 1.0 * q;
    w18 = w18 + -1.0 * q;
  }
  {  // reaction 92: C2H + O2 <=> CO + HCO
    const double2 a0 = A0_84, a1 = A1_84;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c15)*(c18);
    const double pb = kb*(c2)*(c5);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c15)*(c18), Pb = (c2)*(c5);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[2] += 1.0 * dqT;
      dT[5] += 1.0 * dqT;
      dT[15] += -1.0 * dqT;
      dT[18] += -1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w5 = w5 + 1.0 * q;
    w15 = w15 + -1.0 * q;
    w18 = w18 + -1.0 * q;
  }
  {  // reaction 93: C2H + H2 <=> C2H2 + H
    const double2 a0 = A0_85, a1 = A1_85;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c15)*(c4);
    const double pb = kb*(c12)*(c8);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c15)*(c4), Pb = (c12)*(c8);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[4] += -1.0 * dqT;
      dT[8] += 1.0 * dqT;
      dT[12] += 1.0 * dqT;
      dT[15] += -1.0 * dqT;
    }
    w4 = w4 + -1.0 * q;
    w8 = w8 + 1.0 * q;
    w12 = w12 + 1.0 * q;
    w15 = w15 + -1.0 * q;
  }
  {  // reaction 94: CH3 + OH <=> CH2O + H2
    const double2 a0 = A0_86, a1 = A1_86;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c11)*(c3);
    const double pb = kb*(c13)*(c4);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c11)*(c3), Pb = (c13)*(c4);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[3] += -1.0 * dqT;
      dT[4] += 1.0 * dqT;
      dT[11] += -1.0 * dqT;
      dT[13] += 1.0 * dqT;
    }
    w3 = w3 + -1.0 * q;
    w4 = w4 + 1.0 * q;
    w11 = w11 + -1.0 * q;
    w13 = w13 + 1.0 * q;
  }
  {  // reaction 95: CH3 + O2 <=> CH3O2
    const double2 a0 = A0_87, a1 = A1_87;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c11)*(c18);
    const double pb = kb*(c17);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c11)*(c18), Pb = (c17);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[11] += -1.0 * dqT;
      dT[17] += 1.0 * dqT;
      dT[18] += -1.0 * dqT;
    }
    w11 = w11 + -1.0 * q;
    w17 = 1.0 * q;
    w18 = w18 + -1.0 * q;
  }
  {  // reaction 96: CH3O + HO2 <=> CH2O + H2O2
    const double2 a0 = A0_88, a1 = A1_88;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c6)*(c0);
    const double pb = kb*(c13)*(c1);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c6)*(c0), Pb = (c13)*(c1);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[0] += -1.0 * dqT;
      dT[1] += 1.0 * dqT;
      dT[6] += -1.0 * dqT;
      dT[13] += 1.0 * dqT;
    }
    w0 = w0 + -1.0 * q;
    w1 = w1 + 1.0 * q;
    w6 = w6 + -1.0 * q;
    w13 = w13 + 1.0 * q;
  }
  const double2 A0_89 = ld2(rk0 + 176), A1_89 = ld2(rk1 + 176);
  const double2 A0_90 = ld2(rk0 + 178), A1_90 = ld2(rk1 + 178);
  const double2 A0_91 = ld2(rk0 + 180), A1_91 = ld2(rk1 + 180);
  const double2 A0_92 = ld2(rk0 + 182), A1_92 = ld2(rk1 + 182);
  const double2 A0_93 = ld2(rk0 + 184), A1_93 = ld2(rk1 + 184);
  const double2 A0_94 = ld2(rk0 + 186), A1_94 = ld2(rk1 + 186);
  const double2 A0_95 = ld2(rk0 + 188), A1_95 = ld2(rk1 + 188);
  const double2 A0_96 = ld2(rk0 + 190), A1_96 = ld2(rk1 + 190);
  const double2 A0_97 = ld2(rk0 + 192), A1_97 = ld2(rk1 + 192);
  const double2 A0_98 = ld2(rk0 + 194), A1_98 = ld2(rk1 + 194);
  const double2 A0_99 = ld2(rk0 + 196), A1_99 = ld2(rk1 + 196);
  const double2 A0_100 = ld2(rk0 + 198), A1_100 = ld2(rk1 + 198);
  const double2 A0_101 = ld2(rk0 + 200), A1_101 = ld2(rk1 + 200);
  const double2 A0_102 = ld2(rk0 + 202), A1_102 = ld2(rk1 + 202);
  {  // reaction 97: CH3 + CH3O <=> CH2O + CH4
    const double2 a0 = A0_89, a1 = A1_89;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c11)*(c6);
    const double pb = kb*(c13)*(c21);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c11)*(c6), Pb = (c13)*(c21);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[6] += -1.0 * dqT;
      dT[11] += -1.0 * dqT;
      dT[13] += 1.0 * dqT;
      dT[21] += 1.0 * dqT;
    }
    w6 = w6 + -1.0 * q;
    w11 = w11 + -1.0 * q;
    w13 = w13 + 1.0 * q;
    w21 = w21 + 1.0 * q;
  }
  {  // reaction 98: CH3 + CH3O2 <=> 2 CH3O
    const double2 a0 = A0_90, a1 = A1_90;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c11)*(c17);
    const double pb = kb*(c6*c6);
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c11)*(c17), Pb = (c6*c6);
      const double dqT = (a1.x - a0.x) * Pf - (a1.y - a0.y) * Pb;
      dT[6] += 2.0 * dqT;
      dT[11] += -1.0 * dqT;
      dT[17] += -1.0 * dqT;
    }
    w6 = w6 + 2.0 * q;
    w11 = w11 + -1.0 * q;
    w17 = w17 + -1.0 * q;
  }
  {  // reaction 99: C2H2 + O2 => 2 HCO
    const double2 a0 = A0_91, a1 = A1_91;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c12)*(c18);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c12)*(c18), Pb = (c5*c5);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[5] += 2.0 * dqT;
      dT[12] += -1.0 * dqT;
      dT[18] += -1.0 * dqT;
    }
    w5 = w5 + 2.0 * q;
    w12 = w12 + -1.0 * q;
    w18 = w18 + -1.0 * q;
  }
  {  // reaction 100: 2 HCO => C2H2 + O2
    const double2 a0 = A0_92, a1 = A1_92;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c5*c5);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c5*c5), Pb = (c12)*(c18);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[5] += -2.0 * dqT;
      dT[12] += 1.0 * dqT;
      dT[18] += 1.0 * dqT;
    }
    w5 = w5 + -2.0 * q;
    w12 = w12 + 1.0 * q;
    w18 = w18 + 1.0 * q;
  }
  {  // reaction 101: CH3O2 + HO2 => CH3O + O2 + OH
    const double2 a0 = A0_93, a1 = A1_93;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c17)*(c0);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c17)*(c0), Pb = (c6)*(c18)*(c3);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[0] += -1.0 * dqT;
      dT[3] += 1.0 * dqT;
      dT[6] += 1.0 * dqT;
      dT[17] += -1.0 * dqT;
      dT[18] += 1.0 * dqT;
    }
    w0 = w0 + -1.0 * q;
    w3 = w3 + 1.0 * q;
    w6 = w6 + 1.0 * q;
    w17 = w17 + -1.0 * q;
    w18 = w18 + 1.0 * q;
  }
  {  // reaction 102: 2 CH3O2 => 2 CH3O + O2
    const double2 a0 = A0_94, a1 = A1_94;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c17*c17);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c17*c17), Pb = (c6*c6)*(c18);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[6] += 2.0 * dqT;
      dT[17] += -2.0 * dqT;
      dT[18] += 1.0 * dqT;
    }
    w6 = w6 + 2.0 * q;
    w17 = w17 + -2.0 * q;
    w18 = w18 + 1.0 * q;
  }
  {  // reaction 103: C3H3 + O => C2H + CH2O
    const double2 a0 = A0_95, a1 = A1_95;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c14)*(c9);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c14)*(c9), Pb = (c15)*(c13);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[9] += -1.0 * dqT;
      dT[13] += 1.0 * dqT;
      dT[14] += -1.0 * dqT;
      dT[15] += 1.0 * dqT;
    }
    w9 = w9 + -1.0 * q;
    w13 = w13 + 1.0 * q;
    w14 = -1.0 * q;
    w15 = w15 + 1.0 * q;
  }
  {  // reaction 104: C2H + CH2O => C3H3 + O
    const double2 a0 = A0_96, a1 = A1_96;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c15)*(c13);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c15)*(c13), Pb = (c14)*(c9);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[9] += 1.0 * dqT;
      dT[13] += -1.0 * dqT;
      dT[14] += 1.0 * dqT;
      dT[15] += -1.0 * dqT;
    }
    w9 = w9 + 1.0 * q;
    w13 = w13 + -1.0 * q;
    w14 = w14 + 1.0 * q;
    w15 = w15 + -1.0 * q;
  }
  {  // reaction 105: C3H3 + OH => C3H2 + H2O
    const double2 a0 = A0_97, a1 = A1_97;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c14)*(c3);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c14)*(c3), Pb = (c10)*(c16);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[3] += -1.0 * dqT;
      dT[10] += 1.0 * dqT;
      dT[14] += -1.0 * dqT;
      dT[16] += 1.0 * dqT;
    }
    w3 = w3 + -1.0 * q;
    w10 = 1.0 * q;
    w14 = w14 + -1.0 * q;
    w16 = w16 + 1.0 * q;
  }
  {  // reaction 106: C3H2 + H2O => C3H3 + OH
    const double2 a0 = A0_98, a1 = A1_98;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c10)*(c16);
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c10)*(c16), Pb = (c14)*(c3);
      const double dqT = (a1.x - a0.x) * Pf - 0.0;
      dT[3] += 1.0 * dqT;
      dT[10] += -1.0 * dqT;
      dT[14] += 1.0 * dqT;
      dT[16] += -1.0 * dqT;
    }
    w3 = w3 + 1.0 * q;
    w10 = w10 + -1.0 * q;
    w14 = w14 + 1.0 * q;
    w16 = w16 + -1.0 * q;
  }
  {  // reaction 107: HCO2 + M => CO2 + H + M
    const double2 a0 = A0_99, a1 = A1_99;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c22)*M11;
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c22), Pb = (c23)*(c8);
      const double dqT = ((a1.x - a0.x) * Pf - 0.0) * M11;
      scr[37 * SST] = kf * Pf - 0.0;
      dT[8] += 1.0 * dqT;
      dT[22] += -1.0 * dqT;
      dT[23] += 1.0 * dqT;
    }
    w8 = w8 + 1.0 * q;
    w22 = -1.0 * q;
    w23 = w23 + 1.0 * q;
  }
  {  // reaction 108: CO2 + H + M => HCO2 + M
    const double2 a0 = A0_100, a1 = A1_100;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c23)*(c8)*M11;
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c23)*(c8), Pb = (c22);
      const double dqT = ((a1.x - a0.x) * Pf - 0.0) * M11;
      scr[38 * SST] = kf * Pf - 0.0;
      dT[8] += -1.0 * dqT;
      dT[22] += 1.0 * dqT;
      dT[23] += -1.0 * dqT;
    }
    w8 = w8 + -1.0 * q;
    w22 = w22 + 1.0 * q;
    w23 = w23 + -1.0 * q;
  }
  {  // reaction 109: HCO2 + M => CO + OH + M
    const double2 a0 = A0_101, a1 = A1_101;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c22)*M11;
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c22), Pb = (c2)*(c3);
      const double dqT = ((a1.x - a0.x) * Pf - 0.0) * M11;
      scr[39 * SST] = kf * Pf - 0.0;
      dT[2] += 1.0 * dqT;
      dT[3] += 1.0 * dqT;
      dT[22] += -1.0 * dqT;
    }
    w2 = w2 + 1.0 * q;
    w3 = w3 + 1.0 * q;
    w22 = w22 + -1.0 * q;
  }
  {  // reaction 110: CO + OH + M => HCO2 + M
    const double2 a0 = A0_102, a1 = A1_102;
    const double kf = a0.x + (a1.x - a0.x) * Td, kb = a0.y + (a1.y - a0.y) * Td;
    const double pf = kf*(c2)*(c3)*M11;
    const double pb = 0.0;  // irreversible: kb table is identically 0
    const double q = pf - pb;
    if (JAC) {
      const double Pf = (c2)*(c3), Pb = (c22);
      const double dqT = ((a1.x - a0.x) * Pf - 0.0) * M11;
      scr[40 * SST] = kf * Pf - 0.0;
      dT[2] += -1.0 * dqT;
      dT[3] += -1.0 * dqT;
      dT[22] += 1.0 * dqT;
    }
    w2 = w2 + -1.0 * q;
    w3 = w3 + -1.0 * q;
    w22 = w22 + 1.0 * q;
  }
  F[0] = 33.006 * w0;
  F[1] = 34.014 * w1;
  F[2] = 28.01 * w2;
  F[3] = 17.007 * w3;
  F[4] = 2.016 * w4;
  F[5] = 29.018 * w5;
  F[6] = 31.034 * w6;
  F[7] = 14.027 * w7;
  F[8] = 1.008 * w8;
  F[9] = 15.999 * w9;
  F[10] = 38.049 * w10;
  F[11] = 15.035 * w11;
  F[12] = 26.038 * w12;
  F[13] = 30.026 * w13;
  F[14] = 39.057 * w14;
  F[15] = 25.03 * w15;
  F[16] = 18.015 * w16;
  F[17] = 47.033 * w17;
  F[18] = 31.998 * w18;
  F[19] = 14.027 * w19;
  F[20] = 30.07 * w20;
  F[21] = 16.043 * w21;
  F[22] = 45.017 * w22;
  F[23] = 44.009 * w23;
  if (!RHS) return status;
  const int rowh = T_i - Th_min;
  const double* __restrict__ rh0 = tabh + (long)rowh * WH;
  const double* __restrict__ rh1 = rh0 + WH;
  double eiroi = 0.0, rho_cv = 0.0;
  [[maybe_unused]] double s_de = 0.0, s_dcv = 0.0;
  [[maybe_unused]] double e0, e1, e2, e3, e4, e5, e6, e7, e8, e9, e10, e11, e12, e13, e14, e15, e16, e17, e18, e19, e20, e21, e22, e23;
  { const double2 t0 = ld2(rh0 + 0), t1 = ld2(rh1 + 0);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 251.90904683996848 * T;
    eiroi = eiroi + ek * F[0];
    rho_cv = rho_cv + r0 * (cpk - 251.90904683996848);
    if (JAC) { e0 = ek; s_de += ((t1.y - t0.y) - 251.90904683996848) * F[0]; s_dcv += r0 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 2), t1 = ld2(rh1 + 2);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 244.44375845240194 * T;
    eiroi = eiroi + ek * F[1];
    rho_cv = rho_cv + r1 * (cpk - 244.44375845240194);
    if (JAC) { e1 = ek; s_de += ((t1.y - t0.y) - 244.44375845240194) * F[1]; s_dcv += r1 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 4), t1 = ld2(rh1 + 4);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 296.84077115315955 * T;
    eiroi = eiroi + ek * F[2];
    rho_cv = rho_cv + r2 * (cpk - 296.84077115315955);
    if (JAC) { e2 = ek; s_de += ((t1.y - t0.y) - 296.84077115315955) * F[2]; s_dcv += r2 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 6), t1 = ld2(rh1 + 6);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 488.8875169048039 * T;
    eiroi = eiroi + ek * F[3];
    rho_cv = rho_cv + r3 * (cpk - 488.8875169048039);
    if (JAC) { e3 = ek; s_de += ((t1.y - t0.y) - 488.8875169048039) * F[3]; s_dcv += r3 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 8), t1 = ld2(rh1 + 8);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 4124.2609126984125 * T;
    eiroi = eiroi + ek * F[4];
    rho_cv = rho_cv + r4 * (cpk - 4124.2609126984125);
    if (JAC) { e4 = ek; s_de += ((t1.y - t0.y) - 4124.2609126984125) * F[4]; s_dcv += r4 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 10), t1 = ld2(rh1 + 10);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 286.52939554759115 * T;
    eiroi = eiroi + ek * F[5];
    rho_cv = rho_cv + r5 * (cpk - 286.52939554759115);
    if (JAC) { e5 = ek; s_de += ((t1.y - t0.y) - 286.52939554759115) * F[5]; s_dcv += r5 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 12), t1 = ld2(rh1 + 12);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 267.9161564735452 * T;
    eiroi = eiroi + ek * F[6];
    rho_cv = rho_cv + r6 * (cpk - 267.9161564735452);
    if (JAC) { e6 = ek; s_de += ((t1.y - t0.y) - 267.9161564735452) * F[6]; s_dcv += r6 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 14), t1 = ld2(rh1 + 14);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 592.7504099237186 * T;
    eiroi = eiroi + ek * F[7];
    rho_cv = rho_cv + r7 * (cpk - 592.7504099237186);
    if (JAC) { e7 = ek; s_de += ((t1.y - t0.y) - 592.7504099237186) * F[7]; s_dcv += r7 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 16), t1 = ld2(rh1 + 16);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 8248.521825396825 * T;
    eiroi = eiroi + ek * F[8];
    rho_cv = rho_cv + r8 * (cpk - 8248.521825396825);
    if (JAC) { e8 = ek; s_de += ((t1.y - t0.y) - 8248.521825396825) * F[8]; s_dcv += r8 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 18), t1 = ld2(rh1 + 18);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 519.689355584724 * T;
    eiroi = eiroi + ek * F[9];
    rho_cv = rho_cv + r9 * (cpk - 519.689355584724);
    if (JAC) { e9 = ek; s_de += ((t1.y - t0.y) - 519.689355584724) * F[9]; s_dcv += r9 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 20), t1 = ld2(rh1 + 20);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 218.52111750637337 * T;
    eiroi = eiroi + ek * F[10];
    rho_cv = rho_cv + r10 * (cpk - 218.52111750637337);
    if (JAC) { e10 = ek; s_de += ((t1.y - t0.y) - 218.52111750637337) * F[10]; s_dcv += r10 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 22), t1 = ld2(rh1 + 22);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 553.0103092783505 * T;
    eiroi = eiroi + ek * F[11];
    rho_cv = rho_cv + r11 * (cpk - 553.0103092783505);
    if (JAC) { e11 = ek; s_de += ((t1.y - t0.y) - 553.0103092783505) * F[11]; s_dcv += r11 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 24), t1 = ld2(rh1 + 24);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 319.3221445579538 * T;
    eiroi = eiroi + ek * F[12];
    rho_cv = rho_cv + r12 * (cpk - 319.3221445579538);
    if (JAC) { e12 = ek; s_de += ((t1.y - t0.y) - 319.3221445579538) * F[12]; s_dcv += r12 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 26), t1 = ld2(rh1 + 26);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 276.9103443682142 * T;
    eiroi = eiroi + ek * F[13];
    rho_cv = rho_cv + r13 * (cpk - 276.9103443682142);
    if (JAC) { e13 = ek; s_de += ((t1.y - t0.y) - 276.9103443682142) * F[13]; s_dcv += r13 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 28), t1 = ld2(rh1 + 28);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 212.8814297053025 * T;
    eiroi = eiroi + ek * F[14];
    rho_cv = rho_cv + r14 * (cpk - 212.8814297053025);
    if (JAC) { e14 = ek; s_de += ((t1.y - t0.y) - 212.8814297053025) * F[14]; s_dcv += r14 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 30), t1 = ld2(rh1 + 30);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 332.1817818617659 * T;
    eiroi = eiroi + ek * F[15];
    rho_cv = rho_cv + r15 * (cpk - 332.1817818617659);
    if (JAC) { e15 = ek; s_de += ((t1.y - t0.y) - 332.1817818617659) * F[15]; s_dcv += r15 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 32), t1 = ld2(rh1 + 32);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 461.53261171246186 * T;
    eiroi = eiroi + ek * F[16];
    rho_cv = rho_cv + r16 * (cpk - 461.53261171246186);
    if (JAC) { e16 = ek; s_de += ((t1.y - t0.y) - 461.53261171246186) * F[16]; s_dcv += r16 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 34), t1 = ld2(rh1 + 34);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 176.78034571471096 * T;
    eiroi = eiroi + ek * F[17];
    rho_cv = rho_cv + r17 * (cpk - 176.78034571471096);
    if (JAC) { e17 = ek; s_de += ((t1.y - t0.y) - 176.78034571471096) * F[17]; s_dcv += r17 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 36), t1 = ld2(rh1 + 36);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 259.844677792362 * T;
    eiroi = eiroi + ek * F[18];
    rho_cv = rho_cv + r18 * (cpk - 259.844677792362);
    if (JAC) { e18 = ek; s_de += ((t1.y - t0.y) - 259.844677792362) * F[18]; s_dcv += r18 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 38), t1 = ld2(rh1 + 38);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 592.7504099237186 * T;
    eiroi = eiroi + ek * F[19];
    rho_cv = rho_cv + r19 * (cpk - 592.7504099237186);
    if (JAC) { e19 = ek; s_de += ((t1.y - t0.y) - 592.7504099237186) * F[19]; s_dcv += r19 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 40), t1 = ld2(rh1 + 40);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 276.50515463917526 * T;
    eiroi = eiroi + ek * F[20];
    rho_cv = rho_cv + r20 * (cpk - 276.50515463917526);
    if (JAC) { e20 = ek; s_de += ((t1.y - t0.y) - 276.50515463917526) * F[20]; s_dcv += r20 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 42), t1 = ld2(rh1 + 42);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 518.264040391448 * T;
    eiroi = eiroi + ek * F[21];
    rho_cv = rho_cv + r21 * (cpk - 518.264040391448);
    if (JAC) { e21 = ek; s_de += ((t1.y - t0.y) - 518.264040391448) * F[21]; s_dcv += r21 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 44), t1 = ld2(rh1 + 44);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 184.69711442344 * T;
    eiroi = eiroi + ek * F[22];
    rho_cv = rho_cv + r22 * (cpk - 184.69711442344);
    if (JAC) { e22 = ek; s_de += ((t1.y - t0.y) - 184.69711442344) * F[22]; s_dcv += r22 * (t1.x - t0.x); } }
  { const double2 t0 = ld2(rh0 + 46), t1 = ld2(rh1 + 46);
    const double cpk = t0.x + (t1.x - t0.x) * Td, hk = t0.y + (t1.y - t0.y) * Td;
    const double ek = hk - 188.92749210388786 * T;
    eiroi = eiroi + ek * F[23];
    rho_cv = rho_cv + r23 * (cpk - 188.92749210388786);
    if (JAC) { e23 = ek; s_de += ((t1.y - t0.y) - 188.92749210388786) * F[23]; s_dcv += r23 * (t1.x - t0.x); } }
  F[NS] = -eiroi / rho_cv;
  if (JAC) fout(F);
  if (!JAC) return status;
  const double Tdot = F[NS], invD = 1.0 / rho_cv;
  const double* const rk0_ = rk0;
  {  // column NS: d/dT
    double acc = s_de;
    { const double v = 33.006 * dT[0]; sink(0, NS, v); acc += e0 * v; }
    { const double v = 34.014 * dT[1]; sink(1, NS, v); acc += e1 * v; }
    { const double v = 28.01 * dT[2]; sink(2, NS, v); acc += e2 * v; }
    { const double v = 17.007 * dT[3]; sink(3, NS, v); acc += e3 * v; }
    { const double v = 2.016 * dT[4]; sink(4, NS, v); acc += e4 * v; }
    { const double v = 29.018 * dT[5]; sink(5, NS, v); acc += e5 * v; }
    { const double v = 31.034 * dT[6]; sink(6, NS, v); acc += e6 * v; }
    { const double v = 14.027 * dT[7]; sink(7, NS, v); acc += e7 * v; }
    { const double v = 1.008 * dT[8]; sink(8, NS, v); acc += e8 * v; }
    { const double v = 15.999 * dT[9]; sink(9, NS, v); acc += e9 * v; }
    { const double v = 38.049 * dT[10]; sink(10, NS, v); acc += e10 * v; }
    { const double v = 15.035 * dT[11]; sink(11, NS, v); acc += e11 * v; }
    { const double v = 26.038 * dT[12]; sink(12, NS, v); acc += e12 * v; }
    { const double v = 30.026 * dT[13]; sink(13, NS, v); acc += e13 * v; }
    { const double v = 39.057 * dT[14]; sink(14, NS, v); acc += e14 * v; }
    { const double v = 25.03 * dT[15]; sink(15, NS, v); acc += e15 * v; }
    { const double v = 18.015 * dT[16]; sink(16, NS, v); acc += e16 * v; }
    { const double v = 47.033 * dT[17]; sink(17, NS, v); acc += e17 * v; }
    { const double v = 31.998 * dT[18]; sink(18, NS, v); acc += e18 * v; }
    { const double v = 14.027 * dT[19]; sink(19, NS, v); acc += e19 * v; }
    { const double v = 30.07 * dT[20]; sink(20, NS, v); acc += e20 * v; }
    { const double v = 16.043 * dT[21]; sink(21, NS, v); acc += e21 * v; }
    { const double v = 45.017 * dT[22]; sink(22, NS, v); acc += e22 * v; }
    { const double v = 44.009 * dT[23]; sink(23, NS, v); acc += e23 * v; }
    sink(NS, NS, -(acc + Tdot * s_dcv) * invD);
  }
  {  // column 0: d/drho_0 (HO2)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 0) & 1u ? 0.0 : 0.030297521662727988;
    double jc0 = 0.0;
    double jc1 = 0.0;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc5 = 0.0;
    double jc6 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc13 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc17 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc21 = 0.0;
    double jc22 = 0.0;
    double jc23 = 0.0;
    const double2 A0_103 = ld2(rk0 + 38), A1_103 = ld2(rk1 + 38);
    const double2 A0_104 = ld2(rk0 + 40), A1_104 = ld2(rk1 + 40);
    const double2 A0_105 = ld2(rk0 + 42), A1_105 = ld2(rk1 + 42);
    const double2 A0_106 = ld2(rk0 + 58), A1_106 = ld2(rk1 + 58);
    const double2 A0_107 = ld2(rk0 + 60), A1_107 = ld2(rk1 + 60);
    const double2 A0_108 = ld2(rk0 + 62), A1_108 = ld2(rk1 + 62);
    {
      const double dq = scr[0 * SST];
      jc9 += -2.0 * dq;
      jc18 += 1.0 * dq;
    }
    {
      const double dq = scr[1 * SST];
      jc3 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = scr[45 * SST] * c9;
      jc0 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc18 += 1.0 * dq;
    }
    {
      const double dq = -scr[48 * SST] * c3;
      jc0 += 1.0 * dq;
      jc1 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc2 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = -scr[74 * SST] * c5;
      jc0 += 1.0 * dq;
      jc5 += 1.0 * dq;
      jc13 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_103, a1 = A1_103;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M3 * 1.0 + scr[3 * SST];
      jc0 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_104, a1 = A1_104;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M4 * 1.0;
      jc0 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_105, a1 = A1_105;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M5 * 1.0;
      jc0 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc4 += 1.0 * dq;
      jc8 += -2.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc3 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    {
      const double dq = scr[77 * SST] * c8;
      jc0 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc16 += 1.0 * dq;
    }
    { const double2 a0 = A0_106, a1 = A1_106;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c8;
      jc0 += -1.0 * dq;
      jc4 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += 1.0 * dq;
    }
    { const double2 a0 = A0_107, a1 = A1_107;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c8;
      jc0 += -1.0 * dq;
      jc3 += 2.0 * dq;
      jc8 += -1.0 * dq;
    }
    { const double2 a0 = A0_108, a1 = A1_108;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c4;
      jc0 += 1.0 * dq;
      jc1 += -1.0 * dq;
      jc4 += 1.0 * dq;
      jc8 += -1.0 * dq;
    }
    {
      const double dq = scr[11 * SST];
      jc8 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    const double2 A0_109 = ld2(rk0 + 82), A1_109 = ld2(rk1 + 82);
    const double2 A0_110 = ld2(rk0 + 84), A1_110 = ld2(rk1 + 84);
    const double2 A0_111 = ld2(rk0 + 86), A1_111 = ld2(rk1 + 86);
    const double2 A0_112 = ld2(rk0 + 110), A1_112 = ld2(rk1 + 110);
    const double2 A0_113 = ld2(rk0 + 112), A1_113 = ld2(rk1 + 112);
    const double2 A0_114 = ld2(rk0 + 114), A1_114 = ld2(rk1 + 114);
    const double2 A0_115 = ld2(rk0 + 116), A1_115 = ld2(rk1 + 116);
    const double2 A0_116 = ld2(rk0 + 118), A1_116 = ld2(rk1 + 118);
    const double2 A0_117 = ld2(rk0 + 120), A1_117 = ld2(rk1 + 120);
    const double2 A0_118 = ld2(rk0 + 122), A1_118 = ld2(rk1 + 122);
    {
      const double dq = scr[14 * SST];
      jc8 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc21 += 1.0 * dq;
    }
    {
      const double dq = scr[17 * SST];
      jc5 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = scr[20 * SST];
      jc6 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += -1.0 * dq;
    }
    {
      const double dq = scr[23 * SST];
      jc8 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc15 += -1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = scr[29 * SST];
      jc1 += 1.0 * dq;
      jc3 += -2.0 * dq;
    }
    { const double2 a0 = A0_109, a1 = A1_109;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c3;
      jc0 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc16 += 1.0 * dq;
      jc18 += 1.0 * dq;
    }
    { const double2 a0 = A0_110, a1 = A1_110;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c16;
      jc0 += 1.0 * dq;
      jc1 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    { const double2 a0 = A0_111, a1 = A1_111;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c16;
      jc0 += 1.0 * dq;
      jc1 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    { const double2 a0 = A0_112, a1 = A1_112;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 2.0*c0;
      jc0 += -2.0 * dq;
      jc1 += 1.0 * dq;
      jc18 += 1.0 * dq;
    }
    { const double2 a0 = A0_113, a1 = A1_113;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 2.0*c0;
      jc0 += -2.0 * dq;
      jc1 += 1.0 * dq;
      jc18 += 1.0 * dq;
    }
    { const double2 a0 = A0_114, a1 = A1_114;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c19;
      jc0 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc13 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    { const double2 a0 = A0_115, a1 = A1_115;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c11;
      jc0 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc21 += 1.0 * dq;
    }
    { const double2 a0 = A0_116, a1 = A1_116;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c11;
      jc0 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc6 += 1.0 * dq;
      jc11 += -1.0 * dq;
    }
    { const double2 a0 = A0_117, a1 = A1_117;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c2;
      jc0 += -1.0 * dq;
      jc2 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc23 += 1.0 * dq;
    }
    { const double2 a0 = A0_118, a1 = A1_118;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c13;
      jc0 += -1.0 * dq;
      jc1 += 1.0 * dq;
      jc5 += 1.0 * dq;
      jc13 += -1.0 * dq;
    }
    const double2 A0_119 = ld2(rk0 + 152), A1_119 = ld2(rk1 + 152);
    const double2 A0_120 = ld2(rk0 + 162), A1_120 = ld2(rk1 + 162);
    const double2 A0_121 = ld2(rk0 + 164), A1_121 = ld2(rk1 + 164);
    const double2 A0_122 = ld2(rk0 + 174), A1_122 = ld2(rk1 + 174);
    const double2 A0_123 = ld2(rk0 + 184), A1_123 = ld2(rk1 + 184);
    { const double2 a0 = A0_119, a1 = A1_119;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c21;
      jc0 += 1.0 * dq;
      jc1 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc21 += 1.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc11 += -2.0 * dq;
      jc20 += 1.0 * dq;
    }
    {
      const double dq = scr[36 * SST];
      jc2 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc8 += 1.0 * dq;
    }
    { const double2 a0 = A0_120, a1 = A1_120;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc0 += 1.0 * dq;
      jc2 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_121, a1 = A1_121;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c13;
      jc0 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc13 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_122, a1 = A1_122;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c6;
      jc0 += -1.0 * dq;
      jc1 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    { const double2 a0 = A0_123, a1 = A1_123;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c17;
      jc0 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc6 += 1.0 * dq;
      jc17 += -1.0 * dq;
      jc18 += 1.0 * dq;
    }
    {
      const double dq = scr[37 * SST];
      jc8 += 1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc8 += -1.0 * dq;
      jc22 += 1.0 * dq;
      jc23 += -1.0 * dq;
    }
    {
      const double dq = scr[39 * SST];
      jc2 += 1.0 * dq;
      jc3 += 1.0 * dq;
      jc22 += -1.0 * dq;
    }
    {
      const double dq = scr[40 * SST];
      jc2 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc22 += 1.0 * dq;
    }
    double acc = 0.0;
    { const double v = 33.006 * gj * jc0; sink(0, 0, v); acc += e0 * v; }
    { const double v = 34.014 * gj * jc1; sink(1, 0, v); acc += e1 * v; }
    { const double v = 28.01 * gj * jc2; sink(2, 0, v); acc += e2 * v; }
    { const double v = 17.007 * gj * jc3; sink(3, 0, v); acc += e3 * v; }
    { const double v = 2.016 * gj * jc4; sink(4, 0, v); acc += e4 * v; }
    { const double v = 29.018 * gj * jc5; sink(5, 0, v); acc += e5 * v; }
    { const double v = 31.034 * gj * jc6; sink(6, 0, v); acc += e6 * v; }
    sink(7, 0, 0.0);
    { const double v = 1.008 * gj * jc8; sink(8, 0, v); acc += e8 * v; }
    { const double v = 15.999 * gj * jc9; sink(9, 0, v); acc += e9 * v; }
    sink(10, 0, 0.0);
    { const double v = 15.035 * gj * jc11; sink(11, 0, v); acc += e11 * v; }
    { const double v = 26.038 * gj * jc12; sink(12, 0, v); acc += e12 * v; }
    { const double v = 30.026 * gj * jc13; sink(13, 0, v); acc += e13 * v; }
    sink(14, 0, 0.0);
    { const double v = 25.03 * gj * jc15; sink(15, 0, v); acc += e15 * v; }
    { const double v = 18.015 * gj * jc16; sink(16, 0, v); acc += e16 * v; }
    { const double v = 47.033 * gj * jc17; sink(17, 0, v); acc += e17 * v; }
    { const double v = 31.998 * gj * jc18; sink(18, 0, v); acc += e18 * v; }
    { const double v = 14.027 * gj * jc19; sink(19, 0, v); acc += e19 * v; }
    { const double v = 30.07 * gj * jc20; sink(20, 0, v); acc += e20 * v; }
    { const double v = 16.043 * gj * jc21; sink(21, 0, v); acc += e21 * v; }
    { const double v = 45.017 * gj * jc22; sink(22, 0, v); acc += e22 * v; }
    { const double v = 44.009 * gj * jc23; sink(23, 0, v); acc += e23 * v; }
    const double cvj = (ld2(rh0 + 0).x + (ld2(rh1 + 0).x - ld2(rh0 + 0).x) * Td) - 251.90904683996848;
    sink(NS, 0, -(acc + Tdot * cvj * (((negmask >> 0) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 1: d/drho_1 (H2O2)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 1) & 1u ? 0.0 : 0.029399658963956014;
    double jc0 = 0.0;
    double jc1 = 0.0;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc5 = 0.0;
    double jc6 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc13 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc21 = 0.0;
    double jc22 = 0.0;
    double jc23 = 0.0;
    const double2 A0_124 = ld2(rk0 + 62), A1_124 = ld2(rk1 + 62);
    const double2 A0_125 = ld2(rk0 + 64), A1_125 = ld2(rk1 + 64);
    {
      const double dq = scr[0 * SST];
      jc9 += -2.0 * dq;
      jc18 += 1.0 * dq;
    }
    {
      const double dq = scr[1 * SST];
      jc3 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = scr[47 * SST] * c9;
      jc0 += 1.0 * dq;
      jc1 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc2 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc0 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc4 += 1.0 * dq;
      jc8 += -2.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc3 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    { const double2 a0 = A0_124, a1 = A1_124;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c8;
      jc0 += 1.0 * dq;
      jc1 += -1.0 * dq;
      jc4 += 1.0 * dq;
      jc8 += -1.0 * dq;
    }
    { const double2 a0 = A0_125, a1 = A1_125;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c8;
      jc1 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    {
      const double dq = scr[11 * SST];
      jc8 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc8 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc21 += 1.0 * dq;
    }
    {
      const double dq = scr[17 * SST];
      jc5 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = scr[20 * SST];
      jc6 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += -1.0 * dq;
    }
    {
      const double dq = scr[23 * SST];
      jc8 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc15 += -1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = -scr[31 * SST] * 1.0 + scr[29 * SST];
      jc1 += 1.0 * dq;
      jc3 += -2.0 * dq;
    }
    const double2 A0_126 = ld2(rk0 + 84), A1_126 = ld2(rk1 + 84);
    const double2 A0_127 = ld2(rk0 + 86), A1_127 = ld2(rk1 + 86);
    const double2 A0_128 = ld2(rk0 + 110), A1_128 = ld2(rk1 + 110);
    const double2 A0_129 = ld2(rk0 + 112), A1_129 = ld2(rk1 + 112);
    const double2 A0_130 = ld2(rk0 + 122), A1_130 = ld2(rk1 + 122);
    const double2 A0_131 = ld2(rk0 + 152), A1_131 = ld2(rk1 + 152);
    const double2 A0_132 = ld2(rk0 + 174), A1_132 = ld2(rk1 + 174);
    { const double2 a0 = A0_126, a1 = A1_126;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c3;
      jc0 += 1.0 * dq;
      jc1 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    { const double2 a0 = A0_127, a1 = A1_127;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c3;
      jc0 += 1.0 * dq;
      jc1 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    { const double2 a0 = A0_128, a1 = A1_128;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c18;
      jc0 += -2.0 * dq;
      jc1 += 1.0 * dq;
      jc18 += 1.0 * dq;
    }
    { const double2 a0 = A0_129, a1 = A1_129;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c18;
      jc0 += -2.0 * dq;
      jc1 += 1.0 * dq;
      jc18 += 1.0 * dq;
    }
    { const double2 a0 = A0_130, a1 = A1_130;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c5;
      jc0 += -1.0 * dq;
      jc1 += 1.0 * dq;
      jc5 += 1.0 * dq;
      jc13 += -1.0 * dq;
    }
    { const double2 a0 = A0_131, a1 = A1_131;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c11;
      jc0 += 1.0 * dq;
      jc1 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc21 += 1.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc11 += -2.0 * dq;
      jc20 += 1.0 * dq;
    }
    {
      const double dq = scr[36 * SST];
      jc2 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc8 += 1.0 * dq;
    }
    { const double2 a0 = A0_132, a1 = A1_132;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c13;
      jc0 += -1.0 * dq;
      jc1 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = scr[37 * SST];
      jc8 += 1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc8 += -1.0 * dq;
      jc22 += 1.0 * dq;
      jc23 += -1.0 * dq;
    }
    {
      const double dq = scr[39 * SST];
      jc2 += 1.0 * dq;
      jc3 += 1.0 * dq;
      jc22 += -1.0 * dq;
    }
    {
      const double dq = scr[40 * SST];
      jc2 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc22 += 1.0 * dq;
    }
    double acc = 0.0;
    { const double v = 33.006 * gj * jc0; sink(0, 1, v); acc += e0 * v; }
    { const double v = 34.014 * gj * jc1; sink(1, 1, v); acc += e1 * v; }
    { const double v = 28.01 * gj * jc2; sink(2, 1, v); acc += e2 * v; }
    { const double v = 17.007 * gj * jc3; sink(3, 1, v); acc += e3 * v; }
    { const double v = 2.016 * gj * jc4; sink(4, 1, v); acc += e4 * v; }
    { const double v = 29.018 * gj * jc5; sink(5, 1, v); acc += e5 * v; }
    { const double v = 31.034 * gj * jc6; sink(6, 1, v); acc += e6 * v; }
    sink(7, 1, 0.0);
    { const double v = 1.008 * gj * jc8; sink(8, 1, v); acc += e8 * v; }
    { const double v = 15.999 * gj * jc9; sink(9, 1, v); acc += e9 * v; }
    sink(10, 1, 0.0);
    { const double v = 15.035 * gj * jc11; sink(11, 1, v); acc += e11 * v; }
    { const double v = 26.038 * gj * jc12; sink(12, 1, v); acc += e12 * v; }
    { const double v = 30.026 * gj * jc13; sink(13, 1, v); acc += e13 * v; }
    sink(14, 1, 0.0);
    { const double v = 25.03 * gj * jc15; sink(15, 1, v); acc += e15 * v; }
    { const double v = 18.015 * gj * jc16; sink(16, 1, v); acc += e16 * v; }
    sink(17, 1, 0.0);
    { const double v = 31.998 * gj * jc18; sink(18, 1, v); acc += e18 * v; }
    { const double v = 14.027 * gj * jc19; sink(19, 1, v); acc += e19 * v; }
    { const double v = 30.07 * gj * jc20; sink(20, 1, v); acc += e20 * v; }
    { const double v = 16.043 * gj * jc21; sink(21, 1, v); acc += e21 * v; }
    { const double v = 45.017 * gj * jc22; sink(22, 1, v); acc += e22 * v; }
    { const double v = 44.009 * gj * jc23; sink(23, 1, v); acc += e23 * v; }
    const double cvj = (ld2(rh0 + 2).x + (ld2(rh1 + 2).x - ld2(rh0 + 2).x) * Td) - 244.44375845240194;
    sink(NS, 1, -(acc + Tdot * cvj * (((negmask >> 1) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 2: d/drho_2 (CO)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 2) & 1u ? 0.0 : 0.035701535166012134;
    double jc0 = 0.0;
    double jc1 = 0.0;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc5 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc13 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc21 = 0.0;
    double jc22 = 0.0;
    double jc23 = 0.0;
    const double2 A0_133 = ld2(rk0 + 20), A1_133 = ld2(rk1 + 20);
    const double2 A0_134 = ld2(rk0 + 68), A1_134 = ld2(rk1 + 68);
    {
      const double dq = 1.75 * scr[0 * SST];
      jc9 += -2.0 * dq;
      jc18 += 1.0 * dq;
    }
    {
      const double dq = 1.5 * scr[1 * SST];
      jc3 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = -scr[52 * SST] * c4;
      jc2 += 1.0 * dq;
      jc4 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc9 += -1.0 * dq;
    }
    { const double2 a0 = A0_133, a1 = A1_133;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M2 * c9 + 1.5 * scr[2 * SST];
      jc2 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = -scr[60 * SST] * c3;
      jc2 += 1.0 * dq;
      jc3 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = -scr[70 * SST] * c19;
      jc2 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc12 += -1.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[71 * SST] * c18;
      jc2 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = 0.75 * scr[3 * SST];
      jc0 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc4 += 1.0 * dq;
      jc8 += -2.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc3 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    {
      const double dq = 1.5 * scr[11 * SST];
      jc8 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = 1.5 * scr[14 * SST];
      jc8 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc21 += 1.0 * dq;
    }
    {
      const double dq = 1.5 * scr[17 * SST];
      jc5 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    { const double2 a0 = A0_134, a1 = A1_134;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c4;
      jc2 += 1.0 * dq;
      jc4 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc8 += -1.0 * dq;
    }
    {
      const double dq = 1.5 * scr[20 * SST];
      jc6 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += -1.0 * dq;
    }
    {
      const double dq = 1.5 * scr[23 * SST];
      jc8 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc15 += -1.0 * dq;
    }
    const double2 A0_135 = ld2(rk0 + 98), A1_135 = ld2(rk1 + 98);
    const double2 A0_136 = ld2(rk0 + 100), A1_136 = ld2(rk1 + 100);
    const double2 A0_137 = ld2(rk0 + 108), A1_137 = ld2(rk1 + 108);
    const double2 A0_138 = ld2(rk0 + 120), A1_138 = ld2(rk1 + 120);
    const double2 A0_139 = ld2(rk0 + 134), A1_139 = ld2(rk1 + 134);
    const double2 A0_140 = ld2(rk0 + 142), A1_140 = ld2(rk1 + 142);
    const double2 A0_141 = ld2(rk0 + 146), A1_141 = ld2(rk1 + 146);
    const double2 A0_142 = ld2(rk0 + 154), A1_142 = ld2(rk1 + 154);
    const double2 A0_143 = ld2(rk0 + 158), A1_143 = ld2(rk1 + 158);
    const double2 A0_144 = ld2(rk0 + 160), A1_144 = ld2(rk1 + 160);
    const double2 A0_145 = ld2(rk0 + 162), A1_145 = ld2(rk1 + 162);
    const double2 A0_146 = ld2(rk0 + 166), A1_146 = ld2(rk1 + 166);
    {
      const double dq = scr[27 * SST] * c4 + 1.5 * scr[26 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = 1.5 * scr[29 * SST];
      jc1 += 1.0 * dq;
      jc3 += -2.0 * dq;
    }
    { const double2 a0 = A0_135, a1 = A1_135;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c3;
      jc2 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc23 += 1.0 * dq;
    }
    { const double2 a0 = A0_136, a1 = A1_136;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c16;
      jc2 += 1.0 * dq;
      jc3 += -1.0 * dq;
      jc5 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    { const double2 a0 = A0_137, a1 = A1_137;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c11;
      jc2 += 1.0 * dq;
      jc3 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc12 += -1.0 * dq;
    }
    { const double2 a0 = A0_138, a1 = A1_138;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c0;
      jc0 += -1.0 * dq;
      jc2 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = -scr[42 * SST] * c8*c3;
      jc2 += 1.0 * dq;
      jc3 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_139, a1 = A1_139;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c16;
      jc2 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc16 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_140, a1 = A1_140;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c7 - (a0.y + (a1.y - a0.y) * Td) * c19;
      jc2 += 0.0 * dq;
      jc7 += -1.0 * dq;
      jc19 += 1.0 * dq;
    }
    { const double2 a0 = A0_141, a1 = A1_141;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c13;
      jc2 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc13 += 1.0 * dq;
      jc23 += -1.0 * dq;
    }
    {
      const double dq = 1.5 * scr[32 * SST];
      jc11 += -2.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = A0_142, a1 = A1_142;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c21;
      jc2 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc21 += 1.0 * dq;
    }
    { const double2 a0 = A0_143, a1 = A1_143;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M5 * c8;
      jc2 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc8 += 1.0 * dq;
    }
    { const double2 a0 = A0_144, a1 = A1_144;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M10 * c8 + 1.5 * scr[36 * SST];
      jc2 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc8 += 1.0 * dq;
    }
    { const double2 a0 = A0_145, a1 = A1_145;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c0;
      jc0 += 1.0 * dq;
      jc2 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_146, a1 = A1_146;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c5;
      jc2 += 1.0 * dq;
      jc5 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    const double2 A0_147 = ld2(rk0 + 202), A1_147 = ld2(rk1 + 202);
    {
      const double dq = scr[37 * SST];
      jc8 += 1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc8 += -1.0 * dq;
      jc22 += 1.0 * dq;
      jc23 += -1.0 * dq;
    }
    {
      const double dq = scr[39 * SST];
      jc2 += 1.0 * dq;
      jc3 += 1.0 * dq;
      jc22 += -1.0 * dq;
    }
    { const double2 a0 = A0_147, a1 = A1_147;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M11 * c3 + scr[40 * SST];
      jc2 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc22 += 1.0 * dq;
    }
    double acc = 0.0;
    { const double v = 33.006 * gj * jc0; sink(0, 2, v); acc += e0 * v; }
    { const double v = 34.014 * gj * jc1; sink(1, 2, v); acc += e1 * v; }
    { const double v = 28.01 * gj * jc2; sink(2, 2, v); acc += e2 * v; }
    { const double v = 17.007 * gj * jc3; sink(3, 2, v); acc += e3 * v; }
    { const double v = 2.016 * gj * jc4; sink(4, 2, v); acc += e4 * v; }
    { const double v = 29.018 * gj * jc5; sink(5, 2, v); acc += e5 * v; }
    { const double v = 31.034 * gj * jc6; sink(6, 2, v); acc += e6 * v; }
    { const double v = 14.027 * gj * jc7; sink(7, 2, v); acc += e7 * v; }
    { const double v = 1.008 * gj * jc8; sink(8, 2, v); acc += e8 * v; }
    { const double v = 15.999 * gj * jc9; sink(9, 2, v); acc += e9 * v; }
    sink(10, 2, 0.0);
    { const double v = 15.035 * gj * jc11; sink(11, 2, v); acc += e11 * v; }
    { const double v = 26.038 * gj * jc12; sink(12, 2, v); acc += e12 * v; }
    { const double v = 30.026 * gj * jc13; sink(13, 2, v); acc += e13 * v; }
    sink(14, 2, 0.0);
    { const double v = 25.03 * gj * jc15; sink(15, 2, v); acc += e15 * v; }
    { const double v = 18.015 * gj * jc16; sink(16, 2, v); acc += e16 * v; }
    sink(17, 2, 0.0);
    { const double v = 31.998 * gj * jc18; sink(18, 2, v); acc += e18 * v; }
    { const double v = 14.027 * gj * jc19; sink(19, 2, v); acc += e19 * v; }
    { const double v = 30.07 * gj * jc20; sink(20, 2, v); acc += e20 * v; }
    { const double v = 16.043 * gj * jc21; sink(21, 2, v); acc += e21 * v; }
    { const double v = 45.017 * gj * jc22; sink(22, 2, v); acc += e22 * v; }
    { const double v = 44.009 * gj * jc23; sink(23, 2, v); acc += e23 * v; }
    const double cvj = (ld2(rh0 + 4).x + (ld2(rh1 + 4).x - ld2(rh0 + 4).x) * Td) - 296.84077115315955;
    sink(NS, 2, -(acc + Tdot * cvj * (((negmask >> 2) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 3: d/drho_3 (OH)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 3) & 1u ? 0.0 : 0.05879931792791203;
    double jc0 = 0.0;
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
    double jc13 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc21 = 0.0;
    double jc22 = 0.0;
    double jc23 = 0.0;
    const double2 A0_148 = ld2(rk0 + 2), A1_148 = ld2(rk1 + 2);
    const double2 A0_149 = ld2(rk0 + 54), A1_149 = ld2(rk1 + 54);
    const double2 A0_150 = ld2(rk0 + 60), A1_150 = ld2(rk1 + 60);
    {
      const double dq = scr[0 * SST];
      jc9 += -2.0 * dq;
      jc18 += 1.0 * dq;
    }
    { const double2 a0 = A0_148, a1 = A1_148;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M1 * 1.0 + scr[1 * SST];
      jc3 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = -scr[44 * SST] * c8;
      jc3 += 1.0 * dq;
      jc4 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = -scr[46 * SST] * c18;
      jc0 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc18 += 1.0 * dq;
    }
    {
      const double dq = -scr[48 * SST] * c0;
      jc0 += 1.0 * dq;
      jc1 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = -scr[58 * SST] * c11;
      jc3 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc21 += -1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc2 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = -scr[60 * SST] * c2;
      jc2 += 1.0 * dq;
      jc3 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = -scr[64 * SST] * c5;
      jc3 += 1.0 * dq;
      jc5 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc13 += -1.0 * dq;
    }
    {
      const double dq = -scr[66 * SST] * c13;
      jc3 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = -scr[68 * SST] * c15;
      jc3 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc12 += -1.0 * dq;
      jc15 += 1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc0 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = -scr[76 * SST] * c9;
      jc3 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc4 += 1.0 * dq;
      jc8 += -2.0 * dq;
    }
    { const double2 a0 = A0_149, a1 = A1_149;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M9 * c8 + scr[10 * SST];
      jc3 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    { const double2 a0 = A0_150, a1 = A1_150;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * 2.0*c3;
      jc0 += -1.0 * dq;
      jc3 += 2.0 * dq;
      jc8 += -1.0 * dq;
    }
    const double2 A0_151 = ld2(rk0 + 64), A1_151 = ld2(rk1 + 64);
    const double2 A0_152 = ld2(rk0 + 74), A1_152 = ld2(rk1 + 74);
    const double2 A0_153 = ld2(rk0 + 78), A1_153 = ld2(rk1 + 78);
    const double2 A0_154 = ld2(rk0 + 80), A1_154 = ld2(rk1 + 80);
    const double2 A0_155 = ld2(rk0 + 82), A1_155 = ld2(rk1 + 82);
    const double2 A0_156 = ld2(rk0 + 84), A1_156 = ld2(rk1 + 84);
    const double2 A0_157 = ld2(rk0 + 86), A1_157 = ld2(rk1 + 86);
    const double2 A0_158 = ld2(rk0 + 88), A1_158 = ld2(rk1 + 88);
    const double2 A0_159 = ld2(rk0 + 90), A1_159 = ld2(rk1 + 90);
    { const double2 a0 = A0_151, a1 = A1_151;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c16;
      jc1 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    {
      const double dq = scr[11 * SST];
      jc8 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc8 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc21 += 1.0 * dq;
    }
    {
      const double dq = scr[17 * SST];
      jc5 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = scr[20 * SST];
      jc6 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += -1.0 * dq;
    }
    { const double2 a0 = A0_152, a1 = A1_152;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c11;
      jc3 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc11 += 1.0 * dq;
    }
    {
      const double dq = scr[23 * SST];
      jc8 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc15 += -1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    { const double2 a0 = A0_153, a1 = A1_153;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c4;
      jc3 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc16 += 1.0 * dq;
    }
    {
      const double dq = scr[30 * SST] * 2.0*c3 + scr[29 * SST];
      jc1 += 1.0 * dq;
      jc3 += -2.0 * dq;
    }
    { const double2 a0 = A0_154, a1 = A1_154;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 2.0*c3;
      jc3 += -2.0 * dq;
      jc9 += 1.0 * dq;
      jc16 += 1.0 * dq;
    }
    { const double2 a0 = A0_155, a1 = A1_155;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c0;
      jc0 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc16 += 1.0 * dq;
      jc18 += 1.0 * dq;
    }
    { const double2 a0 = A0_156, a1 = A1_156;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c1;
      jc0 += 1.0 * dq;
      jc1 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    { const double2 a0 = A0_157, a1 = A1_157;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c1;
      jc0 += 1.0 * dq;
      jc1 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    { const double2 a0 = A0_158, a1 = A1_158;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c19;
      jc3 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc13 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    { const double2 a0 = A0_159, a1 = A1_159;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c7;
      jc3 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc13 += 1.0 * dq;
    }
    const double2 A0_160 = ld2(rk0 + 92), A1_160 = ld2(rk1 + 92);
    const double2 A0_161 = ld2(rk0 + 94), A1_161 = ld2(rk1 + 94);
    const double2 A0_162 = ld2(rk0 + 96), A1_162 = ld2(rk1 + 96);
    const double2 A0_163 = ld2(rk0 + 98), A1_163 = ld2(rk1 + 98);
    const double2 A0_164 = ld2(rk0 + 100), A1_164 = ld2(rk1 + 100);
    const double2 A0_165 = ld2(rk0 + 102), A1_165 = ld2(rk1 + 102);
    const double2 A0_166 = ld2(rk0 + 104), A1_166 = ld2(rk1 + 104);
    const double2 A0_167 = ld2(rk0 + 106), A1_167 = ld2(rk1 + 106);
    const double2 A0_168 = ld2(rk0 + 108), A1_168 = ld2(rk1 + 108);
    const double2 A0_169 = ld2(rk0 + 114), A1_169 = ld2(rk1 + 114);
    const double2 A0_170 = ld2(rk0 + 118), A1_170 = ld2(rk1 + 118);
    const double2 A0_171 = ld2(rk0 + 120), A1_171 = ld2(rk1 + 120);
    const double2 A0_172 = ld2(rk0 + 124), A1_172 = ld2(rk1 + 124);
    const double2 A0_173 = ld2(rk0 + 150), A1_173 = ld2(rk1 + 150);
    { const double2 a0 = A0_160, a1 = A1_160;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c11;
      jc3 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc16 += 1.0 * dq;
      jc19 += 1.0 * dq;
    }
    { const double2 a0 = A0_161, a1 = A1_161;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c11;
      jc3 += -1.0 * dq;
      jc7 += 1.0 * dq;
      jc11 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    { const double2 a0 = A0_162, a1 = A1_162;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c21;
      jc3 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc16 += 1.0 * dq;
      jc21 += -1.0 * dq;
    }
    { const double2 a0 = A0_163, a1 = A1_163;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c2;
      jc2 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc23 += 1.0 * dq;
    }
    { const double2 a0 = A0_164, a1 = A1_164;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c5;
      jc2 += 1.0 * dq;
      jc3 += -1.0 * dq;
      jc5 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    { const double2 a0 = A0_165, a1 = A1_165;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c13;
      jc3 += -1.0 * dq;
      jc5 += 1.0 * dq;
      jc13 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    { const double2 a0 = A0_166, a1 = A1_166;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c6;
      jc3 += -1.0 * dq;
      jc6 += -1.0 * dq;
      jc13 += 1.0 * dq;
      jc16 += 1.0 * dq;
    }
    { const double2 a0 = A0_167, a1 = A1_167;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c12;
      jc3 += -1.0 * dq;
      jc12 += -1.0 * dq;
      jc15 += 1.0 * dq;
      jc16 += 1.0 * dq;
    }
    { const double2 a0 = A0_168, a1 = A1_168;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c12;
      jc2 += 1.0 * dq;
      jc3 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc12 += -1.0 * dq;
    }
    { const double2 a0 = A0_169, a1 = A1_169;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c13;
      jc0 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc13 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    { const double2 a0 = A0_170, a1 = A1_170;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c6;
      jc0 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc6 += 1.0 * dq;
      jc11 += -1.0 * dq;
    }
    { const double2 a0 = A0_171, a1 = A1_171;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c23;
      jc0 += -1.0 * dq;
      jc2 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc23 += 1.0 * dq;
    }
    { const double2 a0 = A0_172, a1 = A1_172;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c5;
      jc3 += 1.0 * dq;
      jc5 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = -scr[42 * SST] * c2*c8;
      jc2 += 1.0 * dq;
      jc3 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_173, a1 = A1_173;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c13;
      jc3 += 1.0 * dq;
      jc11 += -1.0 * dq;
      jc13 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc11 += -2.0 * dq;
      jc20 += 1.0 * dq;
    }
    const double2 A0_174 = ld2(rk0 + 170), A1_174 = ld2(rk1 + 170);
    const double2 A0_175 = ld2(rk0 + 192), A1_175 = ld2(rk1 + 192);
    const double2 A0_176 = ld2(rk0 + 202), A1_176 = ld2(rk1 + 202);
    {
      const double dq = scr[36 * SST];
      jc2 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc8 += 1.0 * dq;
    }
    { const double2 a0 = A0_174, a1 = A1_174;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c11;
      jc3 += -1.0 * dq;
      jc4 += 1.0 * dq;
      jc11 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    { const double2 a0 = A0_175, a1 = A1_175;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c14;
      jc3 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc14 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    {
      const double dq = scr[37 * SST];
      jc8 += 1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc8 += -1.0 * dq;
      jc22 += 1.0 * dq;
      jc23 += -1.0 * dq;
    }
    {
      const double dq = scr[39 * SST];
      jc2 += 1.0 * dq;
      jc3 += 1.0 * dq;
      jc22 += -1.0 * dq;
    }
    { const double2 a0 = A0_176, a1 = A1_176;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M11 * c2 + scr[40 * SST];
      jc2 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc22 += 1.0 * dq;
    }
    double acc = 0.0;
    { const double v = 33.006 * gj * jc0; sink(0, 3, v); acc += e0 * v; }
    { const double v = 34.014 * gj * jc1; sink(1, 3, v); acc += e1 * v; }
    { const double v = 28.01 * gj * jc2; sink(2, 3, v); acc += e2 * v; }
    { const double v = 17.007 * gj * jc3; sink(3, 3, v); acc += e3 * v; }
    { const double v = 2.016 * gj * jc4; sink(4, 3, v); acc += e4 * v; }
    { const double v = 29.018 * gj * jc5; sink(5, 3, v); acc += e5 * v; }
    { const double v = 31.034 * gj * jc6; sink(6, 3, v); acc += e6 * v; }
    { const double v = 14.027 * gj * jc7; sink(7, 3, v); acc += e7 * v; }
    { const double v = 1.008 * gj * jc8; sink(8, 3, v); acc += e8 * v; }
    { const double v = 15.999 * gj * jc9; sink(9, 3, v); acc += e9 * v; }
    { const double v = 38.049 * gj * jc10; sink(10, 3, v); acc += e10 * v; }
    { const double v = 15.035 * gj * jc11; sink(11, 3, v); acc += e11 * v; }
    { const double v = 26.038 * gj * jc12; sink(12, 3, v); acc += e12 * v; }
    { const double v = 30.026 * gj * jc13; sink(13, 3, v); acc += e13 * v; }
    { const double v = 39.057 * gj * jc14; sink(14, 3, v); acc += e14 * v; }
    { const double v = 25.03 * gj * jc15; sink(15, 3, v); acc += e15 * v; }
    { const double v = 18.015 * gj * jc16; sink(16, 3, v); acc += e16 * v; }
    sink(17, 3, 0.0);
    { const double v = 31.998 * gj * jc18; sink(18, 3, v); acc += e18 * v; }
    { const double v = 14.027 * gj * jc19; sink(19, 3, v); acc += e19 * v; }
    { const double v = 30.07 * gj * jc20; sink(20, 3, v); acc += e20 * v; }
    { const double v = 16.043 * gj * jc21; sink(21, 3, v); acc += e21 * v; }
    { const double v = 45.017 * gj * jc22; sink(22, 3, v); acc += e22 * v; }
    { const double v = 44.009 * gj * jc23; sink(23, 3, v); acc += e23 * v; }
    const double cvj = (ld2(rh0 + 6).x + (ld2(rh1 + 6).x - ld2(rh0 + 6).x) * Td) - 488.8875169048039;
    sink(NS, 3, -(acc + Tdot * cvj * (((negmask >> 3) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 4: d/drho_4 (H2)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 4) & 1u ? 0.0 : 0.49603174603174605;
    double jc0 = 0.0;
    double jc1 = 0.0;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc5 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc13 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc21 = 0.0;
    double jc22 = 0.0;
    double jc23 = 0.0;
    const double2 A0_177 = ld2(rk0 + 46), A1_177 = ld2(rk1 + 46);
    const double2 A0_178 = ld2(rk0 + 48), A1_178 = ld2(rk1 + 48);
    const double2 A0_179 = ld2(rk0 + 50), A1_179 = ld2(rk1 + 50);
    const double2 A0_180 = ld2(rk0 + 52), A1_180 = ld2(rk1 + 52);
    const double2 A0_181 = ld2(rk0 + 58), A1_181 = ld2(rk1 + 58);
    const double2 A0_182 = ld2(rk0 + 62), A1_182 = ld2(rk1 + 62);
    const double2 A0_183 = ld2(rk0 + 66), A1_183 = ld2(rk1 + 66);
    {
      const double dq = 2.4000000953674316 * scr[0 * SST];
      jc9 += -2.0 * dq;
      jc18 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[1 * SST];
      jc3 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = scr[43 * SST] * c9;
      jc3 += 1.0 * dq;
      jc4 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = -scr[52 * SST] * c2;
      jc2 += 1.0 * dq;
      jc4 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[2 * SST];
      jc2 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc0 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_177, a1 = A1_177;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M6 * 1.0;
      jc4 += 1.0 * dq;
      jc8 += -2.0 * dq;
    }
    { const double2 a0 = A0_178, a1 = A1_178;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M7 * 1.0 + scr[7 * SST];
      jc4 += 1.0 * dq;
      jc8 += -2.0 * dq;
    }
    { const double2 a0 = A0_179, a1 = A1_179;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M5 * 1.0;
      jc4 += 1.0 * dq;
      jc8 += -2.0 * dq;
    }
    { const double2 a0 = A0_180, a1 = A1_180;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M8 * 1.0;
      jc4 += 1.0 * dq;
      jc8 += -2.0 * dq;
    }
    {
      const double dq = 0.7300000190734863 * scr[10 * SST];
      jc3 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    { const double2 a0 = A0_181, a1 = A1_181;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c18;
      jc0 += -1.0 * dq;
      jc4 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += 1.0 * dq;
    }
    { const double2 a0 = A0_182, a1 = A1_182;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c0;
      jc0 += 1.0 * dq;
      jc1 += -1.0 * dq;
      jc4 += 1.0 * dq;
      jc8 += -1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[11 * SST];
      jc8 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[14 * SST];
      jc8 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc21 += 1.0 * dq;
    }
    { const double2 a0 = A0_183, a1 = A1_183;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c11;
      jc4 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc21 += -1.0 * dq;
    }
    const double2 A0_184 = ld2(rk0 + 68), A1_184 = ld2(rk1 + 68);
    const double2 A0_185 = ld2(rk0 + 70), A1_185 = ld2(rk1 + 70);
    const double2 A0_186 = ld2(rk0 + 72), A1_186 = ld2(rk1 + 72);
    const double2 A0_187 = ld2(rk0 + 78), A1_187 = ld2(rk1 + 78);
    const double2 A0_188 = ld2(rk0 + 126), A1_188 = ld2(rk1 + 126);
    const double2 A0_189 = ld2(rk0 + 128), A1_189 = ld2(rk1 + 128);
    const double2 A0_190 = ld2(rk0 + 136), A1_190 = ld2(rk1 + 136);
    const double2 A0_191 = ld2(rk0 + 168), A1_191 = ld2(rk1 + 168);
    const double2 A0_192 = ld2(rk0 + 170), A1_192 = ld2(rk1 + 170);
    {
      const double dq = 2.0 * scr[17 * SST];
      jc5 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    { const double2 a0 = A0_184, a1 = A1_184;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc4 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc8 += -1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[20 * SST];
      jc6 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += -1.0 * dq;
    }
    { const double2 a0 = A0_185, a1 = A1_185;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c5;
      jc4 += 1.0 * dq;
      jc5 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += -1.0 * dq;
    }
    { const double2 a0 = A0_186, a1 = A1_186;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c13;
      jc4 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[23 * SST];
      jc8 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc15 += -1.0 * dq;
    }
    {
      const double dq = scr[27 * SST] * c2 + 2.0 * scr[26 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    { const double2 a0 = A0_187, a1 = A1_187;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c3;
      jc3 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc16 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[29 * SST];
      jc1 += 1.0 * dq;
      jc3 += -2.0 * dq;
    }
    { const double2 a0 = A0_188, a1 = A1_188;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c19;
      jc4 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc11 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    { const double2 a0 = A0_189, a1 = A1_189;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c12;
      jc4 += 1.0 * dq;
      jc12 += 1.0 * dq;
      jc19 += -2.0 * dq;
    }
    { const double2 a0 = A0_190, a1 = A1_190;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c7;
      jc4 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc11 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[32 * SST];
      jc11 += -2.0 * dq;
      jc20 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[36 * SST];
      jc2 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc8 += 1.0 * dq;
    }
    { const double2 a0 = A0_191, a1 = A1_191;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c15;
      jc4 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc12 += 1.0 * dq;
      jc15 += -1.0 * dq;
    }
    { const double2 a0 = A0_192, a1 = A1_192;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c13;
      jc3 += -1.0 * dq;
      jc4 += 1.0 * dq;
      jc11 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = scr[37 * SST];
      jc8 += 1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc8 += -1.0 * dq;
      jc22 += 1.0 * dq;
      jc23 += -1.0 * dq;
    }
    {
      const double dq = scr[39 * SST];
      jc2 += 1.0 * dq;
      jc3 += 1.0 * dq;
      jc22 += -1.0 * dq;
    }
    {
      const double dq = scr[40 * SST];
      jc2 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc22 += 1.0 * dq;
    }
    double acc = 0.0;
    { const double v = 33.006 * gj * jc0; sink(0, 4, v); acc += e0 * v; }
    { const double v = 34.014 * gj * jc1; sink(1, 4, v); acc += e1 * v; }
    { const double v = 28.01 * gj * jc2; sink(2, 4, v); acc += e2 * v; }
    { const double v = 17.007 * gj * jc3; sink(3, 4, v); acc += e3 * v; }
    { const double v = 2.016 * gj * jc4; sink(4, 4, v); acc += e4 * v; }
    { const double v = 29.018 * gj * jc5; sink(5, 4, v); acc += e5 * v; }
    { const double v = 31.034 * gj * jc6; sink(6, 4, v); acc += e6 * v; }
    { const double v = 14.027 * gj * jc7; sink(7, 4, v); acc += e7 * v; }
    { const double v = 1.008 * gj * jc8; sink(8, 4, v); acc += e8 * v; }
    { const double v = 15.999 * gj * jc9; sink(9, 4, v); acc += e9 * v; }
    sink(10, 4, 0.0);
    { const double v = 15.035 * gj * jc11; sink(11, 4, v); acc += e11 * v; }
    { const double v = 26.038 * gj * jc12; sink(12, 4, v); acc += e12 * v; }
    { const double v = 30.026 * gj * jc13; sink(13, 4, v); acc += e13 * v; }
    sink(14, 4, 0.0);
    { const double v = 25.03 * gj * jc15; sink(15, 4, v); acc += e15 * v; }
    { const double v = 18.015 * gj * jc16; sink(16, 4, v); acc += e16 * v; }
    sink(17, 4, 0.0);
    { const double v = 31.998 * gj * jc18; sink(18, 4, v); acc += e18 * v; }
    { const double v = 14.027 * gj * jc19; sink(19, 4, v); acc += e19 * v; }
    { const double v = 30.07 * gj * jc20; sink(20, 4, v); acc += e20 * v; }
    { const double v = 16.043 * gj * jc21; sink(21, 4, v); acc += e21 * v; }
    { const double v = 45.017 * gj * jc22; sink(22, 4, v); acc += e22 * v; }
    { const double v = 44.009 * gj * jc23; sink(23, 4, v); acc += e23 * v; }
    const double cvj = (ld2(rh0 + 8).x + (ld2(rh1 + 8).x - ld2(rh0 + 8).x) * Td) - 4124.2609126984125;
    sink(NS, 4, -(acc + Tdot * cvj * (((negmask >> 4) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 5: d/drho_5 (HCO)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 5) & 1u ? 0.0 : 0.03446136880556896;
    double jc0 = 0.0;
    double jc1 = 0.0;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc5 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc13 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc21 = 0.0;
    double jc22 = 0.0;
    double jc23 = 0.0;
    const double2 A0_193 = ld2(rk0 + 68), A1_193 = ld2(rk1 + 68);
    {
      const double dq = scr[0 * SST];
      jc9 += -2.0 * dq;
      jc18 += 1.0 * dq;
    }
    {
      const double dq = scr[1 * SST];
      jc3 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = -scr[50 * SST] * c8;
      jc5 += 1.0 * dq;
      jc8 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = -scr[54 * SST] * c8;
      jc5 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc2 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[59 * SST] * c9;
      jc2 += 1.0 * dq;
      jc3 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = scr[61 * SST] * c9;
      jc5 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = -scr[64 * SST] * c3;
      jc3 += 1.0 * dq;
      jc5 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc13 += -1.0 * dq;
    }
    {
      const double dq = -scr[74 * SST] * c0;
      jc0 += 1.0 * dq;
      jc5 += 1.0 * dq;
      jc13 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc0 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc4 += 1.0 * dq;
      jc8 += -2.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc3 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    {
      const double dq = scr[11 * SST];
      jc8 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc8 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc21 += 1.0 * dq;
    }
    {
      const double dq = scr[18 * SST] * c8 + scr[17 * SST];
      jc5 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    { const double2 a0 = A0_193, a1 = A1_193;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c8;
      jc2 += 1.0 * dq;
      jc4 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc8 += -1.0 * dq;
    }
    const double2 A0_194 = ld2(rk0 + 70), A1_194 = ld2(rk1 + 70);
    const double2 A0_195 = ld2(rk0 + 100), A1_195 = ld2(rk1 + 100);
    const double2 A0_196 = ld2(rk0 + 102), A1_196 = ld2(rk1 + 102);
    const double2 A0_197 = ld2(rk0 + 122), A1_197 = ld2(rk1 + 122);
    const double2 A0_198 = ld2(rk0 + 124), A1_198 = ld2(rk1 + 124);
    const double2 A0_199 = ld2(rk0 + 154), A1_199 = ld2(rk1 + 154);
    const double2 A0_200 = ld2(rk0 + 156), A1_200 = ld2(rk1 + 156);
    const double2 A0_201 = ld2(rk0 + 158), A1_201 = ld2(rk1 + 158);
    const double2 A0_202 = ld2(rk0 + 160), A1_202 = ld2(rk1 + 160);
    const double2 A0_203 = ld2(rk0 + 162), A1_203 = ld2(rk1 + 162);
    const double2 A0_204 = ld2(rk0 + 166), A1_204 = ld2(rk1 + 166);
    {
      const double dq = scr[20 * SST];
      jc6 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += -1.0 * dq;
    }
    { const double2 a0 = A0_194, a1 = A1_194;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c4;
      jc4 += 1.0 * dq;
      jc5 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += -1.0 * dq;
    }
    {
      const double dq = scr[23 * SST];
      jc8 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc15 += -1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = scr[29 * SST];
      jc1 += 1.0 * dq;
      jc3 += -2.0 * dq;
    }
    { const double2 a0 = A0_195, a1 = A1_195;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c3;
      jc2 += 1.0 * dq;
      jc3 += -1.0 * dq;
      jc5 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    { const double2 a0 = A0_196, a1 = A1_196;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c16;
      jc3 += -1.0 * dq;
      jc5 += 1.0 * dq;
      jc13 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    { const double2 a0 = A0_197, a1 = A1_197;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c1;
      jc0 += -1.0 * dq;
      jc1 += 1.0 * dq;
      jc5 += 1.0 * dq;
      jc13 += -1.0 * dq;
    }
    { const double2 a0 = A0_198, a1 = A1_198;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c3;
      jc3 += 1.0 * dq;
      jc5 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc11 += -2.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = A0_199, a1 = A1_199;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c11;
      jc2 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc21 += 1.0 * dq;
    }
    { const double2 a0 = A0_200, a1 = A1_200;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c21;
      jc5 += 1.0 * dq;
      jc11 += -1.0 * dq;
      jc13 += -1.0 * dq;
      jc21 += 1.0 * dq;
    }
    { const double2 a0 = A0_201, a1 = A1_201;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M5 * 1.0;
      jc2 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc8 += 1.0 * dq;
    }
    { const double2 a0 = A0_202, a1 = A1_202;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M10 * 1.0 + scr[36 * SST];
      jc2 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc8 += 1.0 * dq;
    }
    { const double2 a0 = A0_203, a1 = A1_203;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc0 += 1.0 * dq;
      jc2 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_204, a1 = A1_204;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc5 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    const double2 A0_205 = ld2(rk0 + 182), A1_205 = ld2(rk1 + 182);
    { const double2 a0 = A0_205, a1 = A1_205;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 2.0*c5;
      jc5 += -2.0 * dq;
      jc12 += 1.0 * dq;
      jc18 += 1.0 * dq;
    }
    {
      const double dq = scr[37 * SST];
      jc8 += 1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc8 += -1.0 * dq;
      jc22 += 1.0 * dq;
      jc23 += -1.0 * dq;
    }
    {
      const double dq = scr[39 * SST];
      jc2 += 1.0 * dq;
      jc3 += 1.0 * dq;
      jc22 += -1.0 * dq;
    }
    {
      const double dq = scr[40 * SST];
      jc2 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc22 += 1.0 * dq;
    }
    double acc = 0.0;
    { const double v = 33.006 * gj * jc0; sink(0, 5, v); acc += e0 * v; }
    { const double v = 34.014 * gj * jc1; sink(1, 5, v); acc += e1 * v; }
    { const double v = 28.01 * gj * jc2; sink(2, 5, v); acc += e2 * v; }
    { const double v = 17.007 * gj * jc3; sink(3, 5, v); acc += e3 * v; }
    { const double v = 2.016 * gj * jc4; sink(4, 5, v); acc += e4 * v; }
    { const double v = 29.018 * gj * jc5; sink(5, 5, v); acc += e5 * v; }
    { const double v = 31.034 * gj * jc6; sink(6, 5, v); acc += e6 * v; }
    { const double v = 14.027 * gj * jc7; sink(7, 5, v); acc += e7 * v; }
    { const double v = 1.008 * gj * jc8; sink(8, 5, v); acc += e8 * v; }
    { const double v = 15.999 * gj * jc9; sink(9, 5, v); acc += e9 * v; }
    sink(10, 5, 0.0);
    { const double v = 15.035 * gj * jc11; sink(11, 5, v); acc += e11 * v; }
    { const double v = 26.038 * gj * jc12; sink(12, 5, v); acc += e12 * v; }
    { const double v = 30.026 * gj * jc13; sink(13, 5, v); acc += e13 * v; }
    sink(14, 5, 0.0);
    { const double v = 25.03 * gj * jc15; sink(15, 5, v); acc += e15 * v; }
    { const double v = 18.015 * gj * jc16; sink(16, 5, v); acc += e16 * v; }
    sink(17, 5, 0.0);
    { const double v = 31.998 * gj * jc18; sink(18, 5, v); acc += e18 * v; }
    { const double v = 14.027 * gj * jc19; sink(19, 5, v); acc += e19 * v; }
    { const double v = 30.07 * gj * jc20; sink(20, 5, v); acc += e20 * v; }
    { const double v = 16.043 * gj * jc21; sink(21, 5, v); acc += e21 * v; }
    { const double v = 45.017 * gj * jc22; sink(22, 5, v); acc += e22 * v; }
    { const double v = 44.009 * gj * jc23; sink(23, 5, v); acc += e23 * v; }
    const double cvj = (ld2(rh0 + 10).x + (ld2(rh1 + 10).x - ld2(rh0 + 10).x) * Td) - 286.52939554759115;
    sink(NS, 5, -(acc + Tdot * cvj * (((negmask >> 5) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 6: d/drho_6 (CH3O)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 6) & 1u ? 0.0 : 0.032222723464587225;
    double jc0 = 0.0;
    double jc1 = 0.0;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc5 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc13 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc17 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc21 = 0.0;
    double jc22 = 0.0;
    double jc23 = 0.0;
    const double2 A0_206 = ld2(rk0 + 72), A1_206 = ld2(rk1 + 72);
    const double2 A0_207 = ld2(rk0 + 74), A1_207 = ld2(rk1 + 74);
    const double2 A0_208 = ld2(rk0 + 76), A1_208 = ld2(rk1 + 76);
    {
      const double dq = scr[0 * SST];
      jc9 += -2.0 * dq;
      jc18 += 1.0 * dq;
    }
    {
      const double dq = scr[1 * SST];
      jc3 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc2 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[65 * SST] * c9;
      jc3 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc0 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc4 += 1.0 * dq;
      jc8 += -2.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc3 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    {
      const double dq = scr[11 * SST];
      jc8 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc8 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc21 += 1.0 * dq;
    }
    {
      const double dq = scr[17 * SST];
      jc5 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = -scr[22 * SST] * 1.0 + scr[20 * SST];
      jc6 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += -1.0 * dq;
    }
    { const double2 a0 = A0_206, a1 = A1_206;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c8;
      jc4 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    { const double2 a0 = A0_207, a1 = A1_207;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c8;
      jc3 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc11 += 1.0 * dq;
    }
    { const double2 a0 = A0_208, a1 = A1_208;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c8;
      jc6 += -1.0 * dq;
      jc7 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    {
      const double dq = scr[23 * SST];
      jc8 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc15 += -1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    const double2 A0_209 = ld2(rk0 + 104), A1_209 = ld2(rk1 + 104);
    const double2 A0_210 = ld2(rk0 + 118), A1_210 = ld2(rk1 + 118);
    const double2 A0_211 = ld2(rk0 + 148), A1_211 = ld2(rk1 + 148);
    const double2 A0_212 = ld2(rk0 + 164), A1_212 = ld2(rk1 + 164);
    const double2 A0_213 = ld2(rk0 + 174), A1_213 = ld2(rk1 + 174);
    const double2 A0_214 = ld2(rk0 + 176), A1_214 = ld2(rk1 + 176);
    const double2 A0_215 = ld2(rk0 + 178), A1_215 = ld2(rk1 + 178);
    {
      const double dq = scr[29 * SST];
      jc1 += 1.0 * dq;
      jc3 += -2.0 * dq;
    }
    { const double2 a0 = A0_209, a1 = A1_209;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c3;
      jc3 += -1.0 * dq;
      jc6 += -1.0 * dq;
      jc13 += 1.0 * dq;
      jc16 += 1.0 * dq;
    }
    { const double2 a0 = A0_210, a1 = A1_210;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c3;
      jc0 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc6 += 1.0 * dq;
      jc11 += -1.0 * dq;
    }
    { const double2 a0 = A0_211, a1 = A1_211;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c9;
      jc6 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc11 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc11 += -2.0 * dq;
      jc20 += 1.0 * dq;
    }
    {
      const double dq = scr[36 * SST];
      jc2 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc8 += 1.0 * dq;
    }
    { const double2 a0 = A0_212, a1 = A1_212;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc0 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc13 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_213, a1 = A1_213;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c0;
      jc0 += -1.0 * dq;
      jc1 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    { const double2 a0 = A0_214, a1 = A1_214;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c11;
      jc6 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc13 += 1.0 * dq;
      jc21 += 1.0 * dq;
    }
    { const double2 a0 = A0_215, a1 = A1_215;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * 2.0*c6;
      jc6 += 2.0 * dq;
      jc11 += -1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[37 * SST];
      jc8 += 1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc8 += -1.0 * dq;
      jc22 += 1.0 * dq;
      jc23 += -1.0 * dq;
    }
    {
      const double dq = scr[39 * SST];
      jc2 += 1.0 * dq;
      jc3 += 1.0 * dq;
      jc22 += -1.0 * dq;
    }
    {
      const double dq = scr[40 * SST];
      jc2 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc22 += 1.0 * dq;
    }
    double acc = 0.0;
    { const double v = 33.006 * gj * jc0; sink(0, 6, v); acc += e0 * v; }
    { const double v = 34.014 * gj * jc1; sink(1, 6, v); acc += e1 * v; }
    { const double v = 28.01 * gj * jc2; sink(2, 6, v); acc += e2 * v; }
    { const double v = 17.007 * gj * jc3; sink(3, 6, v); acc += e3 * v; }
    { const double v = 2.016 * gj * jc4; sink(4, 6, v); acc += e4 * v; }
    { const double v = 29.018 * gj * jc5; sink(5, 6, v); acc += e5 * v; }
    { const double v = 31.034 * gj * jc6; sink(6, 6, v); acc += e6 * v; }
    { const double v = 14.027 * gj * jc7; sink(7, 6, v); acc += e7 * v; }
    { const double v = 1.008 * gj * jc8; sink(8, 6, v); acc += e8 * v; }
    { const double v = 15.999 * gj * jc9; sink(9, 6, v); acc += e9 * v; }
    sink(10, 6, 0.0);
    { const double v = 15.035 * gj * jc11; sink(11, 6, v); acc += e11 * v; }
    { const double v = 26.038 * gj * jc12; sink(12, 6, v); acc += e12 * v; }
    { const double v = 30.026 * gj * jc13; sink(13, 6, v); acc += e13 * v; }
    sink(14, 6, 0.0);
    { const double v = 25.03 * gj * jc15; sink(15, 6, v); acc += e15 * v; }
    { const double v = 18.015 * gj * jc16; sink(16, 6, v); acc += e16 * v; }
    { const double v = 47.033 * gj * jc17; sink(17, 6, v); acc += e17 * v; }
    { const double v = 31.998 * gj * jc18; sink(18, 6, v); acc += e18 * v; }
    { const double v = 14.027 * gj * jc19; sink(19, 6, v); acc += e19 * v; }
    { const double v = 30.07 * gj * jc20; sink(20, 6, v); acc += e20 * v; }
    { const double v = 16.043 * gj * jc21; sink(21, 6, v); acc += e21 * v; }
    { const double v = 45.017 * gj * jc22; sink(22, 6, v); acc += e22 * v; }
    { const double v = 44.009 * gj * jc23; sink(23, 6, v); acc += e23 * v; }
    const double cvj = (ld2(rh0 + 12).x + (ld2(rh1 + 12).x - ld2(rh0 + 12).x) * Td) - 267.9161564735452;
    sink(NS, 6, -(acc + Tdot * cvj * (((negmask >> 6) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 7: d/drho_7 (CH2(S))
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 7) & 1u ? 0.0 : 0.07129108148570615;
    double jc0 = 0.0;
    double jc1 = 0.0;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc5 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc13 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc21 = 0.0;
    double jc22 = 0.0;
    double jc23 = 0.0;
    const double2 A0_216 = ld2(rk0 + 76), A1_216 = ld2(rk1 + 76);
    {
      const double dq = scr[0 * SST];
      jc9 += -2.0 * dq;
      jc18 += 1.0 * dq;
    }
    {
      const double dq = scr[1 * SST];
      jc3 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = scr[51 * SST] * c9;
      jc2 += 1.0 * dq;
      jc4 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = scr[53 * SST] * c9;
      jc5 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc2 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc0 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc4 += 1.0 * dq;
      jc8 += -2.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc3 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    {
      const double dq = scr[11 * SST];
      jc8 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc8 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc21 += 1.0 * dq;
    }
    {
      const double dq = scr[17 * SST];
      jc5 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = scr[20 * SST];
      jc6 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += -1.0 * dq;
    }
    { const double2 a0 = A0_216, a1 = A1_216;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c16;
      jc6 += -1.0 * dq;
      jc7 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    {
      const double dq = scr[23 * SST];
      jc8 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc15 += -1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = scr[29 * SST];
      jc1 += 1.0 * dq;
      jc3 += -2.0 * dq;
    }
    const double2 A0_217 = ld2(rk0 + 90), A1_217 = ld2(rk1 + 90);
    const double2 A0_218 = ld2(rk0 + 94), A1_218 = ld2(rk1 + 94);
    const double2 A0_219 = ld2(rk0 + 134), A1_219 = ld2(rk1 + 134);
    const double2 A0_220 = ld2(rk0 + 136), A1_220 = ld2(rk1 + 136);
    const double2 A0_221 = ld2(rk0 + 138), A1_221 = ld2(rk1 + 138);
    const double2 A0_222 = ld2(rk0 + 140), A1_222 = ld2(rk1 + 140);
    const double2 A0_223 = ld2(rk0 + 142), A1_223 = ld2(rk1 + 142);
    const double2 A0_224 = ld2(rk0 + 144), A1_224 = ld2(rk1 + 144);
    const double2 A0_225 = ld2(rk0 + 146), A1_225 = ld2(rk1 + 146);
    { const double2 a0 = A0_217, a1 = A1_217;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c3;
      jc3 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc13 += 1.0 * dq;
    }
    { const double2 a0 = A0_218, a1 = A1_218;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c16;
      jc3 += -1.0 * dq;
      jc7 += 1.0 * dq;
      jc11 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    {
      const double dq = scr[41 * SST] * c18;
      jc2 += 1.0 * dq;
      jc3 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_219, a1 = A1_219;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc2 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc16 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_220, a1 = A1_220;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c4;
      jc4 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc11 += 1.0 * dq;
    }
    { const double2 a0 = A0_221, a1 = A1_221;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c16;
      jc7 += -1.0 * dq;
      jc16 += 0.0 * dq;
      jc19 += 1.0 * dq;
    }
    { const double2 a0 = A0_222, a1 = A1_222;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c21;
      jc7 += -1.0 * dq;
      jc11 += 2.0 * dq;
      jc21 += -1.0 * dq;
    }
    { const double2 a0 = A0_223, a1 = A1_223;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c2;
      jc2 += 0.0 * dq;
      jc7 += -1.0 * dq;
      jc19 += 1.0 * dq;
    }
    { const double2 a0 = A0_224, a1 = A1_224;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c23;
      jc7 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc23 += 0.0 * dq;
    }
    { const double2 a0 = A0_225, a1 = A1_225;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c23;
      jc2 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc13 += 1.0 * dq;
      jc23 += -1.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc11 += -2.0 * dq;
      jc20 += 1.0 * dq;
    }
    {
      const double dq = scr[36 * SST];
      jc2 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc8 += 1.0 * dq;
    }
    {
      const double dq = scr[37 * SST];
      jc8 += 1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc8 += -1.0 * dq;
      jc22 += 1.0 * dq;
      jc23 += -1.0 * dq;
    }
    {
      const double dq = scr[39 * SST];
      jc2 += 1.0 * dq;
      jc3 += 1.0 * dq;
      jc22 += -1.0 * dq;
    }
    {
      const double dq = scr[40 * SST];
      jc2 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc22 += 1.0 * dq;
    }
    double acc = 0.0;
    { const double v = 33.006 * gj * jc0; sink(0, 7, v); acc += e0 * v; }
    { const double v = 34.014 * gj * jc1; sink(1, 7, v); acc += e1 * v; }
    { const double v = 28.01 * gj * jc2; sink(2, 7, v); acc += e2 * v; }
    { const double v = 17.007 * gj * jc3; sink(3, 7, v); acc += e3 * v; }
    { const double v = 2.016 * gj * jc4; sink(4, 7, v); acc += e4 * v; }
    { const double v = 29.018 * gj * jc5; sink(5, 7, v); acc += e5 * v; }
    { const double v = 31.034 * gj * jc6; sink(6, 7, v); acc += e6 * v; }
    { const double v = 14.027 * gj * jc7; sink(7, 7, v); acc += e7 * v; }
    { const double v = 1.008 * gj * jc8; sink(8, 7, v); acc += e8 * v; }
    { const double v = 15.999 * gj * jc9; sink(9, 7, v); acc += e9 * v; }
    sink(10, 7, 0.0);
    { const double v = 15.035 * gj * jc11; sink(11, 7, v); acc += e11 * v; }
    { const double v = 26.038 * gj * jc12; sink(12, 7, v); acc += e12 * v; }
    { const double v = 30.026 * gj * jc13; sink(13, 7, v); acc += e13 * v; }
    sink(14, 7, 0.0);
    { const double v = 25.03 * gj * jc15; sink(15, 7, v); acc += e15 * v; }
    { const double v = 18.015 * gj * jc16; sink(16, 7, v); acc += e16 * v; }
    sink(17, 7, 0.0);
    { const double v = 31.998 * gj * jc18; sink(18, 7, v); acc += e18 * v; }
    { const double v = 14.027 * gj * jc19; sink(19, 7, v); acc += e19 * v; }
    { const double v = 30.07 * gj * jc20; sink(20, 7, v); acc += e20 * v; }
    { const double v = 16.043 * gj * jc21; sink(21, 7, v); acc += e21 * v; }
    { const double v = 45.017 * gj * jc22; sink(22, 7, v); acc += e22 * v; }
    { const double v = 44.009 * gj * jc23; sink(23, 7, v); acc += e23 * v; }
    const double cvj = (ld2(rh0 + 14).x + (ld2(rh1 + 14).x - ld2(rh0 + 14).x) * Td) - 592.7504099237186;
    sink(NS, 7, -(acc + Tdot * cvj * (((negmask >> 7) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 8: d/drho_8 (H)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 8) & 1u ? 0.0 : 0.9920634920634921;
    double jc0 = 0.0;
    double jc1 = 0.0;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc5 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc13 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc21 = 0.0;
    double jc22 = 0.0;
    double jc23 = 0.0;
    const double2 A0_226 = ld2(rk0 + 2), A1_226 = ld2(rk1 + 2);
    const double2 A0_227 = ld2(rk0 + 38), A1_227 = ld2(rk1 + 38);
    const double2 A0_228 = ld2(rk0 + 40), A1_228 = ld2(rk1 + 40);
    const double2 A0_229 = ld2(rk0 + 42), A1_229 = ld2(rk1 + 42);
    const double2 A0_230 = ld2(rk0 + 46), A1_230 = ld2(rk1 + 46);
    const double2 A0_231 = ld2(rk0 + 48), A1_231 = ld2(rk1 + 48);
    const double2 A0_232 = ld2(rk0 + 50), A1_232 = ld2(rk1 + 50);
    const double2 A0_233 = ld2(rk0 + 52), A1_233 = ld2(rk1 + 52);
    {
      const double dq = scr[0 * SST];
      jc9 += -2.0 * dq;
      jc18 += 1.0 * dq;
    }
    { const double2 a0 = A0_226, a1 = A1_226;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M1 * c9 + scr[1 * SST];
      jc3 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = -scr[44 * SST] * c3;
      jc3 += 1.0 * dq;
      jc4 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = -scr[50 * SST] * c5;
      jc5 += 1.0 * dq;
      jc8 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = -scr[54 * SST] * c5;
      jc5 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = -scr[56 * SST] * c13;
      jc8 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc2 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = -scr[62 * SST] * c23;
      jc5 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    { const double2 a0 = A0_227, a1 = A1_227;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M3 * c18 + scr[3 * SST];
      jc0 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_228, a1 = A1_228;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M4 * c18;
      jc0 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_229, a1 = A1_229;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M5 * c18;
      jc0 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[75 * SST] * c18;
      jc3 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_230, a1 = A1_230;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M6 * 2.0*c8 + scr[6 * SST];
      jc4 += 1.0 * dq;
      jc8 += -2.0 * dq;
    }
    { const double2 a0 = A0_231, a1 = A1_231;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M7 * 2.0*c8;
      jc4 += 1.0 * dq;
      jc8 += -2.0 * dq;
    }
    { const double2 a0 = A0_232, a1 = A1_232;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M5 * 2.0*c8;
      jc4 += 1.0 * dq;
      jc8 += -2.0 * dq;
    }
    { const double2 a0 = A0_233, a1 = A1_233;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M8 * 2.0*c8;
      jc4 += 1.0 * dq;
      jc8 += -2.0 * dq;
    }
    const double2 A0_234 = ld2(rk0 + 54), A1_234 = ld2(rk1 + 54);
    const double2 A0_235 = ld2(rk0 + 58), A1_235 = ld2(rk1 + 58);
    const double2 A0_236 = ld2(rk0 + 60), A1_236 = ld2(rk1 + 60);
    const double2 A0_237 = ld2(rk0 + 62), A1_237 = ld2(rk1 + 62);
    const double2 A0_238 = ld2(rk0 + 64), A1_238 = ld2(rk1 + 64);
    const double2 A0_239 = ld2(rk0 + 66), A1_239 = ld2(rk1 + 66);
    const double2 A0_240 = ld2(rk0 + 68), A1_240 = ld2(rk1 + 68);
    const double2 A0_241 = ld2(rk0 + 70), A1_241 = ld2(rk1 + 70);
    const double2 A0_242 = ld2(rk0 + 72), A1_242 = ld2(rk1 + 72);
    const double2 A0_243 = ld2(rk0 + 74), A1_243 = ld2(rk1 + 74);
    const double2 A0_244 = ld2(rk0 + 76), A1_244 = ld2(rk1 + 76);
    { const double2 a0 = A0_234, a1 = A1_234;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M9 * c3 + scr[10 * SST];
      jc3 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    {
      const double dq = scr[77 * SST] * c0;
      jc0 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc16 += 1.0 * dq;
    }
    { const double2 a0 = A0_235, a1 = A1_235;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c0;
      jc0 += -1.0 * dq;
      jc4 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += 1.0 * dq;
    }
    { const double2 a0 = A0_236, a1 = A1_236;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c0;
      jc0 += -1.0 * dq;
      jc3 += 2.0 * dq;
      jc8 += -1.0 * dq;
    }
    { const double2 a0 = A0_237, a1 = A1_237;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c1;
      jc0 += 1.0 * dq;
      jc1 += -1.0 * dq;
      jc4 += 1.0 * dq;
      jc8 += -1.0 * dq;
    }
    { const double2 a0 = A0_238, a1 = A1_238;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c1;
      jc1 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    {
      const double dq = scr[12 * SST] * c19 + scr[11 * SST];
      jc8 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[15 * SST] * c11 + scr[14 * SST];
      jc8 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc21 += 1.0 * dq;
    }
    { const double2 a0 = A0_239, a1 = A1_239;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c21;
      jc4 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc21 += -1.0 * dq;
    }
    {
      const double dq = scr[18 * SST] * c5 + scr[17 * SST];
      jc5 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    { const double2 a0 = A0_240, a1 = A1_240;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c5;
      jc2 += 1.0 * dq;
      jc4 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc8 += -1.0 * dq;
    }
    {
      const double dq = scr[21 * SST] * c13 + scr[20 * SST];
      jc6 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += -1.0 * dq;
    }
    { const double2 a0 = A0_241, a1 = A1_241;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c13;
      jc4 += 1.0 * dq;
      jc5 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += -1.0 * dq;
    }
    { const double2 a0 = A0_242, a1 = A1_242;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c6;
      jc4 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    { const double2 a0 = A0_243, a1 = A1_243;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c6;
      jc3 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc11 += 1.0 * dq;
    }
    { const double2 a0 = A0_244, a1 = A1_244;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c6;
      jc6 += -1.0 * dq;
      jc7 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    const double2 A0_245 = ld2(rk0 + 78), A1_245 = ld2(rk1 + 78);
    const double2 A0_246 = ld2(rk0 + 88), A1_246 = ld2(rk1 + 88);
    const double2 A0_247 = ld2(rk0 + 90), A1_247 = ld2(rk1 + 90);
    const double2 A0_248 = ld2(rk0 + 98), A1_248 = ld2(rk1 + 98);
    const double2 A0_249 = ld2(rk0 + 126), A1_249 = ld2(rk1 + 126);
    const double2 A0_250 = ld2(rk0 + 136), A1_250 = ld2(rk1 + 136);
    const double2 A0_251 = ld2(rk0 + 158), A1_251 = ld2(rk1 + 158);
    const double2 A0_252 = ld2(rk0 + 160), A1_252 = ld2(rk1 + 160);
    const double2 A0_253 = ld2(rk0 + 168), A1_253 = ld2(rk1 + 168);
    const double2 A0_254 = ld2(rk0 + 198), A1_254 = ld2(rk1 + 198);
    {
      const double dq = scr[24 * SST] * c15 + scr[23 * SST];
      jc8 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc15 += -1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    { const double2 a0 = A0_245, a1 = A1_245;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c16;
      jc3 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc16 += 1.0 * dq;
    }
    {
      const double dq = scr[29 * SST];
      jc1 += 1.0 * dq;
      jc3 += -2.0 * dq;
    }
    { const double2 a0 = A0_246, a1 = A1_246;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c13;
      jc3 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc13 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    { const double2 a0 = A0_247, a1 = A1_247;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c13;
      jc3 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc13 += 1.0 * dq;
    }
    { const double2 a0 = A0_248, a1 = A1_248;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c23;
      jc2 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc23 += 1.0 * dq;
    }
    { const double2 a0 = A0_249, a1 = A1_249;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c11;
      jc4 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc11 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = -scr[42 * SST] * c2*c3;
      jc2 += 1.0 * dq;
      jc3 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_250, a1 = A1_250;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c11;
      jc4 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc11 += 1.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc11 += -2.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = A0_251, a1 = A1_251;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M5 * c2;
      jc2 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc8 += 1.0 * dq;
    }
    { const double2 a0 = A0_252, a1 = A1_252;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M10 * c2 + scr[36 * SST];
      jc2 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc8 += 1.0 * dq;
    }
    { const double2 a0 = A0_253, a1 = A1_253;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c12;
      jc4 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc12 += 1.0 * dq;
      jc15 += -1.0 * dq;
    }
    {
      const double dq = scr[37 * SST];
      jc8 += 1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    { const double2 a0 = A0_254, a1 = A1_254;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M11 * c23 + scr[38 * SST];
      jc8 += -1.0 * dq;
      jc22 += 1.0 * dq;
      jc23 += -1.0 * dq;
    }
    {
      const double dq = scr[39 * SST];
      jc2 += 1.0 * dq;
      jc3 += 1.0 * dq;
      jc22 += -1.0 * dq;
    }
    {
      const double dq = scr[40 * SST];
      jc2 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc22 += 1.0 * dq;
    }
    double acc = 0.0;
    { const double v = 33.006 * gj * jc0; sink(0, 8, v); acc += e0 * v; }
    { const double v = 34.014 * gj * jc1; sink(1, 8, v); acc += e1 * v; }
    { const double v = 28.01 * gj * jc2; sink(2, 8, v); acc += e2 * v; }
    { const double v = 17.007 * gj * jc3; sink(3, 8, v); acc += e3 * v; }
    { const double v = 2.016 * gj * jc4; sink(4, 8, v); acc += e4 * v; }
    { const double v = 29.018 * gj * jc5; sink(5, 8, v); acc += e5 * v; }
    { const double v = 31.034 * gj * jc6; sink(6, 8, v); acc += e6 * v; }
    { const double v = 14.027 * gj * jc7; sink(7, 8, v); acc += e7 * v; }
    { const double v = 1.008 * gj * jc8; sink(8, 8, v); acc += e8 * v; }
    { const double v = 15.999 * gj * jc9; sink(9, 8, v); acc += e9 * v; }
    sink(10, 8, 0.0);
    { const double v = 15.035 * gj * jc11; sink(11, 8, v); acc += e11 * v; }
    { const double v = 26.038 * gj * jc12; sink(12, 8, v); acc += e12 * v; }
    { const double v = 30.026 * gj * jc13; sink(13, 8, v); acc += e13 * v; }
    sink(14, 8, 0.0);
    { const double v = 25.03 * gj * jc15; sink(15, 8, v); acc += e15 * v; }
    { const double v = 18.015 * gj * jc16; sink(16, 8, v); acc += e16 * v; }
    sink(17, 8, 0.0);
    { const double v = 31.998 * gj * jc18; sink(18, 8, v); acc += e18 * v; }
    { const double v = 14.027 * gj * jc19; sink(19, 8, v); acc += e19 * v; }
    { const double v = 30.07 * gj * jc20; sink(20, 8, v); acc += e20 * v; }
    { const double v = 16.043 * gj * jc21; sink(21, 8, v); acc += e21 * v; }
    { const double v = 45.017 * gj * jc22; sink(22, 8, v); acc += e22 * v; }
    { const double v = 44.009 * gj * jc23; sink(23, 8, v); acc += e23 * v; }
    const double cvj = (ld2(rh0 + 16).x + (ld2(rh1 + 16).x - ld2(rh0 + 16).x) * Td) - 8248.521825396825;
    sink(NS, 8, -(acc + Tdot * cvj * (((negmask >> 8) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 9: d/drho_9 (O)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 9) & 1u ? 0.0 : 0.06250390649415588;
    double jc0 = 0.0;
    double jc1 = 0.0;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc5 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc13 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc21 = 0.0;
    double jc22 = 0.0;
    double jc23 = 0.0;
    const double2 A0_255 = ld2(rk0 + 0), A1_255 = ld2(rk1 + 0);
    const double2 A0_256 = ld2(rk0 + 2), A1_256 = ld2(rk1 + 2);
    const double2 A0_257 = ld2(rk0 + 20), A1_257 = ld2(rk1 + 20);
    { const double2 a0 = A0_255, a1 = A1_255;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M0 * 2.0*c9 + scr[0 * SST];
      jc9 += -2.0 * dq;
      jc18 += 1.0 * dq;
    }
    { const double2 a0 = A0_256, a1 = A1_256;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M1 * c8 + scr[1 * SST];
      jc3 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = scr[43 * SST] * c4;
      jc3 += 1.0 * dq;
      jc4 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = scr[45 * SST] * c0;
      jc0 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc18 += 1.0 * dq;
    }
    {
      const double dq = scr[47 * SST] * c1;
      jc0 += 1.0 * dq;
      jc1 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = scr[49 * SST] * c19;
      jc5 += 1.0 * dq;
      jc8 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[51 * SST] * c7;
      jc2 += 1.0 * dq;
      jc4 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = scr[53 * SST] * c7;
      jc5 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = scr[55 * SST] * c11;
      jc8 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = scr[57 * SST] * c21;
      jc3 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc21 += -1.0 * dq;
    }
    { const double2 a0 = A0_257, a1 = A1_257;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M2 * c2 + scr[2 * SST];
      jc2 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[59 * SST] * c5;
      jc2 += 1.0 * dq;
      jc3 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = scr[61 * SST] * c5;
      jc5 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[63 * SST] * c13;
      jc3 += 1.0 * dq;
      jc5 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc13 += -1.0 * dq;
    }
    {
      const double dq = scr[65 * SST] * c6;
      jc3 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = scr[67 * SST] * c12;
      jc3 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc12 += -1.0 * dq;
      jc15 += 1.0 * dq;
    }
    const double2 A0_258 = ld2(rk0 + 80), A1_258 = ld2(rk1 + 80);
    const double2 A0_259 = ld2(rk0 + 148), A1_259 = ld2(rk1 + 148);
    {
      const double dq = scr[69 * SST] * c12;
      jc2 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc12 += -1.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = -scr[72 * SST] * c23;
      jc2 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc0 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = -scr[76 * SST] * c3;
      jc3 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc4 += 1.0 * dq;
      jc8 += -2.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc3 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    {
      const double dq = -scr[78 * SST] * c16;
      jc0 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc16 += 1.0 * dq;
    }
    {
      const double dq = scr[11 * SST];
      jc8 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc8 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc21 += 1.0 * dq;
    }
    {
      const double dq = scr[17 * SST];
      jc5 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = scr[20 * SST];
      jc6 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += -1.0 * dq;
    }
    {
      const double dq = scr[23 * SST];
      jc8 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc15 += -1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = scr[29 * SST];
      jc1 += 1.0 * dq;
      jc3 += -2.0 * dq;
    }
    { const double2 a0 = A0_258, a1 = A1_258;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c16;
      jc3 += -2.0 * dq;
      jc9 += 1.0 * dq;
      jc16 += 1.0 * dq;
    }
    { const double2 a0 = A0_259, a1 = A1_259;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c6;
      jc6 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc11 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    const double2 A0_260 = ld2(rk0 + 188), A1_260 = ld2(rk1 + 188);
    {
      const double dq = scr[32 * SST];
      jc11 += -2.0 * dq;
      jc20 += 1.0 * dq;
    }
    {
      const double dq = scr[36 * SST];
      jc2 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc8 += 1.0 * dq;
    }
    { const double2 a0 = A0_260, a1 = A1_260;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c14;
      jc9 += -1.0 * dq;
      jc13 += 1.0 * dq;
      jc14 += -1.0 * dq;
      jc15 += 1.0 * dq;
    }
    {
      const double dq = scr[37 * SST];
      jc8 += 1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc8 += -1.0 * dq;
      jc22 += 1.0 * dq;
      jc23 += -1.0 * dq;
    }
    {
      const double dq = scr[39 * SST];
      jc2 += 1.0 * dq;
      jc3 += 1.0 * dq;
      jc22 += -1.0 * dq;
    }
    {
      const double dq = scr[40 * SST];
      jc2 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc22 += 1.0 * dq;
    }
    double acc = 0.0;
    { const double v = 33.006 * gj * jc0; sink(0, 9, v); acc += e0 * v; }
    { const double v = 34.014 * gj * jc1; sink(1, 9, v); acc += e1 * v; }
    { const double v = 28.01 * gj * jc2; sink(2, 9, v); acc += e2 * v; }
    { const double v = 17.007 * gj * jc3; sink(3, 9, v); acc += e3 * v; }
    { const double v = 2.016 * gj * jc4; sink(4, 9, v); acc += e4 * v; }
    { const double v = 29.018 * gj * jc5; sink(5, 9, v); acc += e5 * v; }
    { const double v = 31.034 * gj * jc6; sink(6, 9, v); acc += e6 * v; }
    { const double v = 14.027 * gj * jc7; sink(7, 9, v); acc += e7 * v; }
    { const double v = 1.008 * gj * jc8; sink(8, 9, v); acc += e8 * v; }
    { const double v = 15.999 * gj * jc9; sink(9, 9, v); acc += e9 * v; }
    sink(10, 9, 0.0);
    { const double v = 15.035 * gj * jc11; sink(11, 9, v); acc += e11 * v; }
    { const double v = 26.038 * gj * jc12; sink(12, 9, v); acc += e12 * v; }
    { const double v = 30.026 * gj * jc13; sink(13, 9, v); acc += e13 * v; }
    { const double v = 39.057 * gj * jc14; sink(14, 9, v); acc += e14 * v; }
    { const double v = 25.03 * gj * jc15; sink(15, 9, v); acc += e15 * v; }
    { const double v = 18.015 * gj * jc16; sink(16, 9, v); acc += e16 * v; }
    sink(17, 9, 0.0);
    { const double v = 31.998 * gj * jc18; sink(18, 9, v); acc += e18 * v; }
    { const double v = 14.027 * gj * jc19; sink(19, 9, v); acc += e19 * v; }
    { const double v = 30.07 * gj * jc20; sink(20, 9, v); acc += e20 * v; }
    { const double v = 16.043 * gj * jc21; sink(21, 9, v); acc += e21 * v; }
    { const double v = 45.017 * gj * jc22; sink(22, 9, v); acc += e22 * v; }
    { const double v = 44.009 * gj * jc23; sink(23, 9, v); acc += e23 * v; }
    const double cvj = (ld2(rh0 + 18).x + (ld2(rh1 + 18).x - ld2(rh0 + 18).x) * Td) - 519.689355584724;
    sink(NS, 9, -(acc + Tdot * cvj * (((negmask >> 9) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 10: d/drho_10 (C3H2)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 10) & 1u ? 0.0 : 0.026281899655707115;
    double jc0 = 0.0;
    double jc1 = 0.0;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc5 = 0.0;
    double jc6 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc10 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc13 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc21 = 0.0;
    double jc22 = 0.0;
    double jc23 = 0.0;
    const double2 A0_261 = ld2(rk0 + 194), A1_261 = ld2(rk1 + 194);
    {
      const double dq = scr[0 * SST];
      jc9 += -2.0 * dq;
      jc18 += 1.0 * dq;
    }
    {
      const double dq = scr[1 * SST];
      jc3 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc2 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc0 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc4 += 1.0 * dq;
      jc8 += -2.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc3 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    {
      const double dq = scr[11 * SST];
      jc8 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc8 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc21 += 1.0 * dq;
    }
    {
      const double dq = scr[17 * SST];
      jc5 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = scr[20 * SST];
      jc6 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += -1.0 * dq;
    }
    {
      const double dq = scr[23 * SST];
      jc8 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc15 += -1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = scr[29 * SST];
      jc1 += 1.0 * dq;
      jc3 += -2.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc11 += -2.0 * dq;
      jc20 += 1.0 * dq;
    }
    {
      const double dq = scr[36 * SST];
      jc2 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc8 += 1.0 * dq;
    }
    { const double2 a0 = A0_261, a1 = A1_261;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c16;
      jc3 += 1.0 * dq;
      jc10 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc16 += -1.0 * dq;
    }
    {
      const double dq = scr[37 * SST];
      jc8 += 1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc8 += -1.0 * dq;
      jc22 += 1.0 * dq;
      jc23 += -1.0 * dq;
    }
    {
      const double dq = scr[39 * SST];
      jc2 += 1.0 * dq;
      jc3 += 1.0 * dq;
      jc22 += -1.0 * dq;
    }
    {
      const double dq = scr[40 * SST];
      jc2 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc22 += 1.0 * dq;
    }
    double acc = 0.0;
    { const double v = 33.006 * gj * jc0; sink(0, 10, v); acc += e0 * v; }
    { const double v = 34.014 * gj * jc1; sink(1, 10, v); acc += e1 * v; }
    { const double v = 28.01 * gj * jc2; sink(2, 10, v); acc += e2 * v; }
    { const double v = 17.007 * gj * jc3; sink(3, 10, v); acc += e3 * v; }
    { const double v = 2.016 * gj * jc4; sink(4, 10, v); acc += e4 * v; }
    { const double v = 29.018 * gj * jc5; sink(5, 10, v); acc += e5 * v; }
    { const double v = 31.034 * gj * jc6; sink(6, 10, v); acc += e6 * v; }
    sink(7, 10, 0.0);
    { const double v = 1.008 * gj * jc8; sink(8, 10, v); acc += e8 * v; }
    { const double v = 15.999 * gj * jc9; sink(9, 10, v); acc += e9 * v; }
    { const double v = 38.049 * gj * jc10; sink(10, 10, v); acc += e10 * v; }
    { const double v = 15.035 * gj * jc11; sink(11, 10, v); acc += e11 * v; }
    { const double v = 26.038 * gj * jc12; sink(12, 10, v); acc += e12 * v; }
    { const double v = 30.026 * gj * jc13; sink(13, 10, v); acc += e13 * v; }
    { const double v = 39.057 * gj * jc14; sink(14, 10, v); acc += e14 * v; }
    { const double v = 25.03 * gj * jc15; sink(15, 10, v); acc += e15 * v; }
    { const double v = 18.015 * gj * jc16; sink(16, 10, v); acc += e16 * v; }
    sink(17, 10, 0.0);
    { const double v = 31.998 * gj * jc18; sink(18, 10, v); acc += e18 * v; }
    { const double v = 14.027 * gj * jc19; sink(19, 10, v); acc += e19 * v; }
    { const double v = 30.07 * gj * jc20; sink(20, 10, v); acc += e20 * v; }
    { const double v = 16.043 * gj * jc21; sink(21, 10, v); acc += e21 * v; }
    { const double v = 45.017 * gj * jc22; sink(22, 10, v); acc += e22 * v; }
    { const double v = 44.009 * gj * jc23; sink(23, 10, v); acc += e23 * v; }
    const double cvj = (ld2(rh0 + 20).x + (ld2(rh1 + 20).x - ld2(rh0 + 20).x) * Td) - 218.52111750637337;
    sink(NS, 10, -(acc + Tdot * cvj * (((negmask >> 10) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 11: d/drho_11 (CH3)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 11) & 1u ? 0.0 : 0.06651147322913202;
    double jc0 = 0.0;
    double jc1 = 0.0;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc5 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc13 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc17 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc21 = 0.0;
    double jc22 = 0.0;
    double jc23 = 0.0;
    const double2 A0_262 = ld2(rk0 + 66), A1_262 = ld2(rk1 + 66);
    const double2 A0_263 = ld2(rk0 + 74), A1_263 = ld2(rk1 + 74);
    {
      const double dq = scr[0 * SST];
      jc9 += -2.0 * dq;
      jc18 += 1.0 * dq;
    }
    {
      const double dq = scr[1 * SST];
      jc3 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = scr[55 * SST] * c9;
      jc8 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = -scr[58 * SST] * c3;
      jc3 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc21 += -1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc2 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc0 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc4 += 1.0 * dq;
      jc8 += -2.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc3 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    {
      const double dq = -scr[13 * SST] * 1.0 + scr[11 * SST];
      jc8 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[15 * SST] * c8 + scr[14 * SST];
      jc8 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc21 += 1.0 * dq;
    }
    { const double2 a0 = A0_262, a1 = A1_262;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c4;
      jc4 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc21 += -1.0 * dq;
    }
    {
      const double dq = scr[17 * SST];
      jc5 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = scr[20 * SST];
      jc6 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += -1.0 * dq;
    }
    { const double2 a0 = A0_263, a1 = A1_263;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c3;
      jc3 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc11 += 1.0 * dq;
    }
    {
      const double dq = scr[23 * SST];
      jc8 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc15 += -1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    const double2 A0_264 = ld2(rk0 + 92), A1_264 = ld2(rk1 + 92);
    const double2 A0_265 = ld2(rk0 + 94), A1_265 = ld2(rk1 + 94);
    const double2 A0_266 = ld2(rk0 + 96), A1_266 = ld2(rk1 + 96);
    const double2 A0_267 = ld2(rk0 + 108), A1_267 = ld2(rk1 + 108);
    const double2 A0_268 = ld2(rk0 + 116), A1_268 = ld2(rk1 + 116);
    const double2 A0_269 = ld2(rk0 + 118), A1_269 = ld2(rk1 + 118);
    const double2 A0_270 = ld2(rk0 + 126), A1_270 = ld2(rk1 + 126);
    const double2 A0_271 = ld2(rk0 + 130), A1_271 = ld2(rk1 + 130);
    const double2 A0_272 = ld2(rk0 + 136), A1_272 = ld2(rk1 + 136);
    const double2 A0_273 = ld2(rk0 + 140), A1_273 = ld2(rk1 + 140);
    const double2 A0_274 = ld2(rk0 + 148), A1_274 = ld2(rk1 + 148);
    const double2 A0_275 = ld2(rk0 + 150), A1_275 = ld2(rk1 + 150);
    const double2 A0_276 = ld2(rk0 + 152), A1_276 = ld2(rk1 + 152);
    const double2 A0_277 = ld2(rk0 + 154), A1_277 = ld2(rk1 + 154);
    {
      const double dq = scr[29 * SST];
      jc1 += 1.0 * dq;
      jc3 += -2.0 * dq;
    }
    { const double2 a0 = A0_264, a1 = A1_264;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c3;
      jc3 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc16 += 1.0 * dq;
      jc19 += 1.0 * dq;
    }
    { const double2 a0 = A0_265, a1 = A1_265;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c3;
      jc3 += -1.0 * dq;
      jc7 += 1.0 * dq;
      jc11 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    { const double2 a0 = A0_266, a1 = A1_266;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c16;
      jc3 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc16 += 1.0 * dq;
      jc21 += -1.0 * dq;
    }
    { const double2 a0 = A0_267, a1 = A1_267;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc3 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc12 += -1.0 * dq;
    }
    { const double2 a0 = A0_268, a1 = A1_268;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c0;
      jc0 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc21 += 1.0 * dq;
    }
    { const double2 a0 = A0_269, a1 = A1_269;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c0;
      jc0 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc6 += 1.0 * dq;
      jc11 += -1.0 * dq;
    }
    { const double2 a0 = A0_270, a1 = A1_270;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c8;
      jc4 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc11 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    { const double2 a0 = A0_271, a1 = A1_271;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * 2.0*c11;
      jc11 += 2.0 * dq;
      jc19 += -1.0 * dq;
      jc21 += -1.0 * dq;
    }
    { const double2 a0 = A0_272, a1 = A1_272;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c8;
      jc4 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc11 += 1.0 * dq;
    }
    { const double2 a0 = A0_273, a1 = A1_273;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * 2.0*c11;
      jc7 += -1.0 * dq;
      jc11 += 2.0 * dq;
      jc21 += -1.0 * dq;
    }
    { const double2 a0 = A0_274, a1 = A1_274;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc6 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc11 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_275, a1 = A1_275;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc3 += 1.0 * dq;
      jc11 += -1.0 * dq;
      jc13 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_276, a1 = A1_276;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c1;
      jc0 += 1.0 * dq;
      jc1 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc21 += 1.0 * dq;
    }
    {
      const double dq = scr[33 * SST] * 2.0*c11 + scr[32 * SST];
      jc11 += -2.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = A0_277, a1 = A1_277;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c5;
      jc2 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc21 += 1.0 * dq;
    }
    const double2 A0_278 = ld2(rk0 + 156), A1_278 = ld2(rk1 + 156);
    const double2 A0_279 = ld2(rk0 + 170), A1_279 = ld2(rk1 + 170);
    const double2 A0_280 = ld2(rk0 + 172), A1_280 = ld2(rk1 + 172);
    const double2 A0_281 = ld2(rk0 + 176), A1_281 = ld2(rk1 + 176);
    const double2 A0_282 = ld2(rk0 + 178), A1_282 = ld2(rk1 + 178);
    { const double2 a0 = A0_278, a1 = A1_278;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c13;
      jc5 += 1.0 * dq;
      jc11 += -1.0 * dq;
      jc13 += -1.0 * dq;
      jc21 += 1.0 * dq;
    }
    {
      const double dq = scr[36 * SST];
      jc2 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc8 += 1.0 * dq;
    }
    { const double2 a0 = A0_279, a1 = A1_279;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c3;
      jc3 += -1.0 * dq;
      jc4 += 1.0 * dq;
      jc11 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    { const double2 a0 = A0_280, a1 = A1_280;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc11 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_281, a1 = A1_281;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c6;
      jc6 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc13 += 1.0 * dq;
      jc21 += 1.0 * dq;
    }
    { const double2 a0 = A0_282, a1 = A1_282;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c17;
      jc6 += 2.0 * dq;
      jc11 += -1.0 * dq;
      jc17 += -1.0 * dq;
    }
    {
      const double dq = scr[37 * SST];
      jc8 += 1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc8 += -1.0 * dq;
      jc22 += 1.0 * dq;
      jc23 += -1.0 * dq;
    }
    {
      const double dq = scr[39 * SST];
      jc2 += 1.0 * dq;
      jc3 += 1.0 * dq;
      jc22 += -1.0 * dq;
    }
    {
      const double dq = scr[40 * SST];
      jc2 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc22 += 1.0 * dq;
    }
    double acc = 0.0;
    { const double v = 33.006 * gj * jc0; sink(0, 11, v); acc += e0 * v; }
    { const double v = 34.014 * gj * jc1; sink(1, 11, v); acc += e1 * v; }
    { const double v = 28.01 * gj * jc2; sink(2, 11, v); acc += e2 * v; }
    { const double v = 17.007 * gj * jc3; sink(3, 11, v); acc += e3 * v; }
    { const double v = 2.016 * gj * jc4; sink(4, 11, v); acc += e4 * v; }
    { const double v = 29.018 * gj * jc5; sink(5, 11, v); acc += e5 * v; }
    { const double v = 31.034 * gj * jc6; sink(6, 11, v); acc += e6 * v; }
    { const double v = 14.027 * gj * jc7; sink(7, 11, v); acc += e7 * v; }
    { const double v = 1.008 * gj * jc8; sink(8, 11, v); acc += e8 * v; }
    { const double v = 15.999 * gj * jc9; sink(9, 11, v); acc += e9 * v; }
    sink(10, 11, 0.0);
    { const double v = 15.035 * gj * jc11; sink(11, 11, v); acc += e11 * v; }
    { const double v = 26.038 * gj * jc12; sink(12, 11, v); acc += e12 * v; }
    { const double v = 30.026 * gj * jc13; sink(13, 11, v); acc += e13 * v; }
    sink(14, 11, 0.0);
    { const double v = 25.03 * gj * jc15; sink(15, 11, v); acc += e15 * v; }
    { const double v = 18.015 * gj * jc16; sink(16, 11, v); acc += e16 * v; }
    { const double v = 47.033 * gj * jc17; sink(17, 11, v); acc += e17 * v; }
    { const double v = 31.998 * gj * jc18; sink(18, 11, v); acc += e18 * v; }
    { const double v = 14.027 * gj * jc19; sink(19, 11, v); acc += e19 * v; }
    { const double v = 30.07 * gj * jc20; sink(20, 11, v); acc += e20 * v; }
    { const double v = 16.043 * gj * jc21; sink(21, 11, v); acc += e21 * v; }
    { const double v = 45.017 * gj * jc22; sink(22, 11, v); acc += e22 * v; }
    { const double v = 44.009 * gj * jc23; sink(23, 11, v); acc += e23 * v; }
    const double cvj = (ld2(rh0 + 22).x + (ld2(rh1 + 22).x - ld2(rh0 + 22).x) * Td) - 553.0103092783505;
    sink(NS, 11, -(acc + Tdot * cvj * (((negmask >> 11) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 12: d/drho_12 (C2H2)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 12) & 1u ? 0.0 : 0.038405407481373374;
    double jc0 = 0.0;
    double jc1 = 0.0;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc5 = 0.0;
    double jc6 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc13 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc21 = 0.0;
    double jc22 = 0.0;
    double jc23 = 0.0;
    const double2 A0_283 = ld2(rk0 + 106), A1_283 = ld2(rk1 + 106);
    {
      const double dq = scr[0 * SST];
      jc9 += -2.0 * dq;
      jc18 += 1.0 * dq;
    }
    {
      const double dq = scr[1 * SST];
      jc3 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc2 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[67 * SST] * c9;
      jc3 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc12 += -1.0 * dq;
      jc15 += 1.0 * dq;
    }
    {
      const double dq = scr[69 * SST] * c9;
      jc2 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc12 += -1.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc0 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc4 += 1.0 * dq;
      jc8 += -2.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc3 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    {
      const double dq = scr[11 * SST];
      jc8 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc8 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc21 += 1.0 * dq;
    }
    {
      const double dq = scr[17 * SST];
      jc5 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = scr[20 * SST];
      jc6 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += -1.0 * dq;
    }
    {
      const double dq = -scr[25 * SST] * 1.0 + scr[23 * SST];
      jc8 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc15 += -1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = scr[29 * SST];
      jc1 += 1.0 * dq;
      jc3 += -2.0 * dq;
    }
    { const double2 a0 = A0_283, a1 = A1_283;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c3;
      jc3 += -1.0 * dq;
      jc12 += -1.0 * dq;
      jc15 += 1.0 * dq;
      jc16 += 1.0 * dq;
    }
    const double2 A0_284 = ld2(rk0 + 108), A1_284 = ld2(rk1 + 108);
    const double2 A0_285 = ld2(rk0 + 128), A1_285 = ld2(rk1 + 128);
    const double2 A0_286 = ld2(rk0 + 168), A1_286 = ld2(rk1 + 168);
    const double2 A0_287 = ld2(rk0 + 180), A1_287 = ld2(rk1 + 180);
    { const double2 a0 = A0_284, a1 = A1_284;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c3;
      jc2 += 1.0 * dq;
      jc3 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc12 += -1.0 * dq;
    }
    { const double2 a0 = A0_285, a1 = A1_285;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c4;
      jc4 += 1.0 * dq;
      jc12 += 1.0 * dq;
      jc19 += -2.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc11 += -2.0 * dq;
      jc20 += 1.0 * dq;
    }
    {
      const double dq = scr[36 * SST];
      jc2 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc8 += 1.0 * dq;
    }
    { const double2 a0 = A0_286, a1 = A1_286;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c8;
      jc4 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc12 += 1.0 * dq;
      jc15 += -1.0 * dq;
    }
    { const double2 a0 = A0_287, a1 = A1_287;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc5 += 2.0 * dq;
      jc12 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[37 * SST];
      jc8 += 1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc8 += -1.0 * dq;
      jc22 += 1.0 * dq;
      jc23 += -1.0 * dq;
    }
    {
      const double dq = scr[39 * SST];
      jc2 += 1.0 * dq;
      jc3 += 1.0 * dq;
      jc22 += -1.0 * dq;
    }
    {
      const double dq = scr[40 * SST];
      jc2 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc22 += 1.0 * dq;
    }
    double acc = 0.0;
    { const double v = 33.006 * gj * jc0; sink(0, 12, v); acc += e0 * v; }
    { const double v = 34.014 * gj * jc1; sink(1, 12, v); acc += e1 * v; }
    { const double v = 28.01 * gj * jc2; sink(2, 12, v); acc += e2 * v; }
    { const double v = 17.007 * gj * jc3; sink(3, 12, v); acc += e3 * v; }
    { const double v = 2.016 * gj * jc4; sink(4, 12, v); acc += e4 * v; }
    { const double v = 29.018 * gj * jc5; sink(5, 12, v); acc += e5 * v; }
    { const double v = 31.034 * gj * jc6; sink(6, 12, v); acc += e6 * v; }
    sink(7, 12, 0.0);
    { const double v = 1.008 * gj * jc8; sink(8, 12, v); acc += e8 * v; }
    { const double v = 15.999 * gj * jc9; sink(9, 12, v); acc += e9 * v; }
    sink(10, 12, 0.0);
    { const double v = 15.035 * gj * jc11; sink(11, 12, v); acc += e11 * v; }
    { const double v = 26.038 * gj * jc12; sink(12, 12, v); acc += e12 * v; }
    { const double v = 30.026 * gj * jc13; sink(13, 12, v); acc += e13 * v; }
    sink(14, 12, 0.0);
    { const double v = 25.03 * gj * jc15; sink(15, 12, v); acc += e15 * v; }
    { const double v = 18.015 * gj * jc16; sink(16, 12, v); acc += e16 * v; }
    sink(17, 12, 0.0);
    { const double v = 31.998 * gj * jc18; sink(18, 12, v); acc += e18 * v; }
    { const double v = 14.027 * gj * jc19; sink(19, 12, v); acc += e19 * v; }
    { const double v = 30.07 * gj * jc20; sink(20, 12, v); acc += e20 * v; }
    { const double v = 16.043 * gj * jc21; sink(21, 12, v); acc += e21 * v; }
    { const double v = 45.017 * gj * jc22; sink(22, 12, v); acc += e22 * v; }
    { const double v = 44.009 * gj * jc23; sink(23, 12, v); acc += e23 * v; }
    const double cvj = (ld2(rh0 + 24).x + (ld2(rh1 + 24).x - ld2(rh0 + 24).x) * Td) - 319.3221445579538;
    sink(NS, 12, -(acc + Tdot * cvj * (((negmask >> 12) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 13: d/drho_13 (CH2O)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 13) & 1u ? 0.0 : 0.033304469459801506;
    double jc0 = 0.0;
    double jc1 = 0.0;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc5 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc13 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc21 = 0.0;
    double jc22 = 0.0;
    double jc23 = 0.0;
    const double2 A0_288 = ld2(rk0 + 70), A1_288 = ld2(rk1 + 70);
    const double2 A0_289 = ld2(rk0 + 72), A1_289 = ld2(rk1 + 72);
    {
      const double dq = scr[0 * SST];
      jc9 += -2.0 * dq;
      jc18 += 1.0 * dq;
    }
    {
      const double dq = scr[1 * SST];
      jc3 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = -scr[56 * SST] * c8;
      jc8 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc2 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[63 * SST] * c9;
      jc3 += 1.0 * dq;
      jc5 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc13 += -1.0 * dq;
    }
    {
      const double dq = -scr[66 * SST] * c3;
      jc3 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = scr[73 * SST] * c18;
      jc0 += 1.0 * dq;
      jc5 += 1.0 * dq;
      jc13 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc0 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc4 += 1.0 * dq;
      jc8 += -2.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc3 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    {
      const double dq = scr[11 * SST];
      jc8 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc8 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc21 += 1.0 * dq;
    }
    {
      const double dq = -scr[19 * SST] * 1.0 + scr[17 * SST];
      jc5 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = scr[21 * SST] * c8 + scr[20 * SST];
      jc6 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += -1.0 * dq;
    }
    { const double2 a0 = A0_288, a1 = A1_288;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c8;
      jc4 += 1.0 * dq;
      jc5 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += -1.0 * dq;
    }
    { const double2 a0 = A0_289, a1 = A1_289;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c4;
      jc4 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    const double2 A0_290 = ld2(rk0 + 88), A1_290 = ld2(rk1 + 88);
    const double2 A0_291 = ld2(rk0 + 90), A1_291 = ld2(rk1 + 90);
    const double2 A0_292 = ld2(rk0 + 102), A1_292 = ld2(rk1 + 102);
    const double2 A0_293 = ld2(rk0 + 104), A1_293 = ld2(rk1 + 104);
    const double2 A0_294 = ld2(rk0 + 114), A1_294 = ld2(rk1 + 114);
    const double2 A0_295 = ld2(rk0 + 122), A1_295 = ld2(rk1 + 122);
    const double2 A0_296 = ld2(rk0 + 146), A1_296 = ld2(rk1 + 146);
    const double2 A0_297 = ld2(rk0 + 150), A1_297 = ld2(rk1 + 150);
    const double2 A0_298 = ld2(rk0 + 156), A1_298 = ld2(rk1 + 156);
    const double2 A0_299 = ld2(rk0 + 164), A1_299 = ld2(rk1 + 164);
    const double2 A0_300 = ld2(rk0 + 170), A1_300 = ld2(rk1 + 170);
    {
      const double dq = scr[23 * SST];
      jc8 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc15 += -1.0 * dq;
    }
    {
      const double dq = -scr[28 * SST] * 1.0 + scr[26 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = scr[29 * SST];
      jc1 += 1.0 * dq;
      jc3 += -2.0 * dq;
    }
    { const double2 a0 = A0_290, a1 = A1_290;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c8;
      jc3 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc13 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    { const double2 a0 = A0_291, a1 = A1_291;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c8;
      jc3 += -1.0 * dq;
      jc7 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc13 += 1.0 * dq;
    }
    { const double2 a0 = A0_292, a1 = A1_292;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c3;
      jc3 += -1.0 * dq;
      jc5 += 1.0 * dq;
      jc13 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    { const double2 a0 = A0_293, a1 = A1_293;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c16;
      jc3 += -1.0 * dq;
      jc6 += -1.0 * dq;
      jc13 += 1.0 * dq;
      jc16 += 1.0 * dq;
    }
    { const double2 a0 = A0_294, a1 = A1_294;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c3;
      jc0 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc13 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    { const double2 a0 = A0_295, a1 = A1_295;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c0;
      jc0 += -1.0 * dq;
      jc1 += 1.0 * dq;
      jc5 += 1.0 * dq;
      jc13 += -1.0 * dq;
    }
    { const double2 a0 = A0_296, a1 = A1_296;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc13 += 1.0 * dq;
      jc23 += -1.0 * dq;
    }
    { const double2 a0 = A0_297, a1 = A1_297;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c3;
      jc3 += 1.0 * dq;
      jc11 += -1.0 * dq;
      jc13 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc11 += -2.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = A0_298, a1 = A1_298;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c11;
      jc5 += 1.0 * dq;
      jc11 += -1.0 * dq;
      jc13 += -1.0 * dq;
      jc21 += 1.0 * dq;
    }
    {
      const double dq = scr[36 * SST];
      jc2 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc8 += 1.0 * dq;
    }
    { const double2 a0 = A0_299, a1 = A1_299;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c0;
      jc0 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc13 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_300, a1 = A1_300;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c4;
      jc3 += -1.0 * dq;
      jc4 += 1.0 * dq;
      jc11 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    const double2 A0_301 = ld2(rk0 + 174), A1_301 = ld2(rk1 + 174);
    const double2 A0_302 = ld2(rk0 + 176), A1_302 = ld2(rk1 + 176);
    const double2 A0_303 = ld2(rk0 + 190), A1_303 = ld2(rk1 + 190);
    { const double2 a0 = A0_301, a1 = A1_301;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c1;
      jc0 += -1.0 * dq;
      jc1 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    { const double2 a0 = A0_302, a1 = A1_302;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c21;
      jc6 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc13 += 1.0 * dq;
      jc21 += 1.0 * dq;
    }
    { const double2 a0 = A0_303, a1 = A1_303;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c15;
      jc9 += 1.0 * dq;
      jc13 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc15 += -1.0 * dq;
    }
    {
      const double dq = scr[37 * SST];
      jc8 += 1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc8 += -1.0 * dq;
      jc22 += 1.0 * dq;
      jc23 += -1.0 * dq;
    }
    {
      const double dq = scr[39 * SST];
      jc2 += 1.0 * dq;
      jc3 += 1.0 * dq;
      jc22 += -1.0 * dq;
    }
    {
      const double dq = scr[40 * SST];
      jc2 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc22 += 1.0 * dq;
    }
    double acc = 0.0;
    { const double v = 33.006 * gj * jc0; sink(0, 13, v); acc += e0 * v; }
    { const double v = 34.014 * gj * jc1; sink(1, 13, v); acc += e1 * v; }
    { const double v = 28.01 * gj * jc2; sink(2, 13, v); acc += e2 * v; }
    { const double v = 17.007 * gj * jc3; sink(3, 13, v); acc += e3 * v; }
    { const double v = 2.016 * gj * jc4; sink(4, 13, v); acc += e4 * v; }
    { const double v = 29.018 * gj * jc5; sink(5, 13, v); acc += e5 * v; }
    { const double v = 31.034 * gj * jc6; sink(6, 13, v); acc += e6 * v; }
    { const double v = 14.027 * gj * jc7; sink(7, 13, v); acc += e7 * v; }
    { const double v = 1.008 * gj * jc8; sink(8, 13, v); acc += e8 * v; }
    { const double v = 15.999 * gj * jc9; sink(9, 13, v); acc += e9 * v; }
    sink(10, 13, 0.0);
    { const double v = 15.035 * gj * jc11; sink(11, 13, v); acc += e11 * v; }
    { const double v = 26.038 * gj * jc12; sink(12, 13, v); acc += e12 * v; }
    { const double v = 30.026 * gj * jc13; sink(13, 13, v); acc += e13 * v; }
    { const double v = 39.057 * gj * jc14; sink(14, 13, v); acc += e14 * v; }
    { const double v = 25.03 * gj * jc15; sink(15, 13, v); acc += e15 * v; }
    { const double v = 18.015 * gj * jc16; sink(16, 13, v); acc += e16 * v; }
    sink(17, 13, 0.0);
    { const double v = 31.998 * gj * jc18; sink(18, 13, v); acc += e18 * v; }
    { const double v = 14.027 * gj * jc19; sink(19, 13, v); acc += e19 * v; }
    { const double v = 30.07 * gj * jc20; sink(20, 13, v); acc += e20 * v; }
    { const double v = 16.043 * gj * jc21; sink(21, 13, v); acc += e21 * v; }
    { const double v = 45.017 * gj * jc22; sink(22, 13, v); acc += e22 * v; }
    { const double v = 44.009 * gj * jc23; sink(23, 13, v); acc += e23 * v; }
    const double cvj = (ld2(rh0 + 26).x + (ld2(rh1 + 26).x - ld2(rh0 + 26).x) * Td) - 276.9103443682142;
    sink(NS, 13, -(acc + Tdot * cvj * (((negmask >> 13) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 14: d/drho_14 (C3H3)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 14) & 1u ? 0.0 : 0.02560360498758225;
    double jc0 = 0.0;
    double jc1 = 0.0;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc5 = 0.0;
    double jc6 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc10 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc13 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc21 = 0.0;
    double jc22 = 0.0;
    double jc23 = 0.0;
    const double2 A0_304 = ld2(rk0 + 188), A1_304 = ld2(rk1 + 188);
    {
      const double dq = scr[0 * SST];
      jc9 += -2.0 * dq;
      jc18 += 1.0 * dq;
    }
    {
      const double dq = scr[1 * SST];
      jc3 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc2 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc0 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc4 += 1.0 * dq;
      jc8 += -2.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc3 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    {
      const double dq = scr[11 * SST];
      jc8 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc8 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc21 += 1.0 * dq;
    }
    {
      const double dq = scr[17 * SST];
      jc5 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = scr[20 * SST];
      jc6 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += -1.0 * dq;
    }
    {
      const double dq = scr[23 * SST];
      jc8 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc15 += -1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = scr[29 * SST];
      jc1 += 1.0 * dq;
      jc3 += -2.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc11 += -2.0 * dq;
      jc20 += 1.0 * dq;
    }
    {
      const double dq = scr[36 * SST];
      jc2 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc8 += 1.0 * dq;
    }
    { const double2 a0 = A0_304, a1 = A1_304;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c9;
      jc9 += -1.0 * dq;
      jc13 += 1.0 * dq;
      jc14 += -1.0 * dq;
      jc15 += 1.0 * dq;
    }
    const double2 A0_305 = ld2(rk0 + 192), A1_305 = ld2(rk1 + 192);
    { const double2 a0 = A0_305, a1 = A1_305;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c3;
      jc3 += -1.0 * dq;
      jc10 += 1.0 * dq;
      jc14 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    {
      const double dq = scr[37 * SST];
      jc8 += 1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc8 += -1.0 * dq;
      jc22 += 1.0 * dq;
      jc23 += -1.0 * dq;
    }
    {
      const double dq = scr[39 * SST];
      jc2 += 1.0 * dq;
      jc3 += 1.0 * dq;
      jc22 += -1.0 * dq;
    }
    {
      const double dq = scr[40 * SST];
      jc2 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc22 += 1.0 * dq;
    }
    double acc = 0.0;
    { const double v = 33.006 * gj * jc0; sink(0, 14, v); acc += e0 * v; }
    { const double v = 34.014 * gj * jc1; sink(1, 14, v); acc += e1 * v; }
    { const double v = 28.01 * gj * jc2; sink(2, 14, v); acc += e2 * v; }
    { const double v = 17.007 * gj * jc3; sink(3, 14, v); acc += e3 * v; }
    { const double v = 2.016 * gj * jc4; sink(4, 14, v); acc += e4 * v; }
    { const double v = 29.018 * gj * jc5; sink(5, 14, v); acc += e5 * v; }
    { const double v = 31.034 * gj * jc6; sink(6, 14, v); acc += e6 * v; }
    sink(7, 14, 0.0);
    { const double v = 1.008 * gj * jc8; sink(8, 14, v); acc += e8 * v; }
    { const double v = 15.999 * gj * jc9; sink(9, 14, v); acc += e9 * v; }
    { const double v = 38.049 * gj * jc10; sink(10, 14, v); acc += e10 * v; }
    { const double v = 15.035 * gj * jc11; sink(11, 14, v); acc += e11 * v; }
    { const double v = 26.038 * gj * jc12; sink(12, 14, v); acc += e12 * v; }
    { const double v = 30.026 * gj * jc13; sink(13, 14, v); acc += e13 * v; }
    { const double v = 39.057 * gj * jc14; sink(14, 14, v); acc += e14 * v; }
    { const double v = 25.03 * gj * jc15; sink(15, 14, v); acc += e15 * v; }
    { const double v = 18.015 * gj * jc16; sink(16, 14, v); acc += e16 * v; }
    sink(17, 14, 0.0);
    { const double v = 31.998 * gj * jc18; sink(18, 14, v); acc += e18 * v; }
    { const double v = 14.027 * gj * jc19; sink(19, 14, v); acc += e19 * v; }
    { const double v = 30.07 * gj * jc20; sink(20, 14, v); acc += e20 * v; }
    { const double v = 16.043 * gj * jc21; sink(21, 14, v); acc += e21 * v; }
    { const double v = 45.017 * gj * jc22; sink(22, 14, v); acc += e22 * v; }
    { const double v = 44.009 * gj * jc23; sink(23, 14, v); acc += e23 * v; }
    const double cvj = (ld2(rh0 + 28).x + (ld2(rh1 + 28).x - ld2(rh0 + 28).x) * Td) - 212.8814297053025;
    sink(NS, 14, -(acc + Tdot * cvj * (((negmask >> 14) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 15: d/drho_15 (C2H)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 15) & 1u ? 0.0 : 0.03995205753096284;
    double jc0 = 0.0;
    double jc1 = 0.0;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc5 = 0.0;
    double jc6 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc13 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc21 = 0.0;
    double jc22 = 0.0;
    double jc23 = 0.0;
    const double2 A0_306 = ld2(rk0 + 106), A1_306 = ld2(rk1 + 106);
    {
      const double dq = scr[0 * SST];
      jc9 += -2.0 * dq;
      jc18 += 1.0 * dq;
    }
    {
      const double dq = scr[1 * SST];
      jc3 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc2 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = -scr[68 * SST] * c3;
      jc3 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc12 += -1.0 * dq;
      jc15 += 1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc0 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc4 += 1.0 * dq;
      jc8 += -2.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc3 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    {
      const double dq = scr[11 * SST];
      jc8 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc8 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc21 += 1.0 * dq;
    }
    {
      const double dq = scr[17 * SST];
      jc5 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = scr[20 * SST];
      jc6 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += -1.0 * dq;
    }
    {
      const double dq = scr[24 * SST] * c8 + scr[23 * SST];
      jc8 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc15 += -1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = scr[29 * SST];
      jc1 += 1.0 * dq;
      jc3 += -2.0 * dq;
    }
    { const double2 a0 = A0_306, a1 = A1_306;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c16;
      jc3 += -1.0 * dq;
      jc12 += -1.0 * dq;
      jc15 += 1.0 * dq;
      jc16 += 1.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc11 += -2.0 * dq;
      jc20 += 1.0 * dq;
    }
    const double2 A0_307 = ld2(rk0 + 166), A1_307 = ld2(rk1 + 166);
    const double2 A0_308 = ld2(rk0 + 168), A1_308 = ld2(rk1 + 168);
    const double2 A0_309 = ld2(rk0 + 190), A1_309 = ld2(rk1 + 190);
    {
      const double dq = scr[36 * SST];
      jc2 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc8 += 1.0 * dq;
    }
    { const double2 a0 = A0_307, a1 = A1_307;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc2 += 1.0 * dq;
      jc5 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_308, a1 = A1_308;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c4;
      jc4 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc12 += 1.0 * dq;
      jc15 += -1.0 * dq;
    }
    { const double2 a0 = A0_309, a1 = A1_309;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c13;
      jc9 += 1.0 * dq;
      jc13 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc15 += -1.0 * dq;
    }
    {
      const double dq = scr[37 * SST];
      jc8 += 1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc8 += -1.0 * dq;
      jc22 += 1.0 * dq;
      jc23 += -1.0 * dq;
    }
    {
      const double dq = scr[39 * SST];
      jc2 += 1.0 * dq;
      jc3 += 1.0 * dq;
      jc22 += -1.0 * dq;
    }
    {
      const double dq = scr[40 * SST];
      jc2 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc22 += 1.0 * dq;
    }
    double acc = 0.0;
    { const double v = 33.006 * gj * jc0; sink(0, 15, v); acc += e0 * v; }
    { const double v = 34.014 * gj * jc1; sink(1, 15, v); acc += e1 * v; }
    { const double v = 28.01 * gj * jc2; sink(2, 15, v); acc += e2 * v; }
    { const double v = 17.007 * gj * jc3; sink(3, 15, v); acc += e3 * v; }
    { const double v = 2.016 * gj * jc4; sink(4, 15, v); acc += e4 * v; }
    { const double v = 29.018 * gj * jc5; sink(5, 15, v); acc += e5 * v; }
    { const double v = 31.034 * gj * jc6; sink(6, 15, v); acc += e6 * v; }
    sink(7, 15, 0.0);
    { const double v = 1.008 * gj * jc8; sink(8, 15, v); acc += e8 * v; }
    { const double v = 15.999 * gj * jc9; sink(9, 15, v); acc += e9 * v; }
    sink(10, 15, 0.0);
    { const double v = 15.035 * gj * jc11; sink(11, 15, v); acc += e11 * v; }
    { const double v = 26.038 * gj * jc12; sink(12, 15, v); acc += e12 * v; }
    { const double v = 30.026 * gj * jc13; sink(13, 15, v); acc += e13 * v; }
    { const double v = 39.057 * gj * jc14; sink(14, 15, v); acc += e14 * v; }
    { const double v = 25.03 * gj * jc15; sink(15, 15, v); acc += e15 * v; }
    { const double v = 18.015 * gj * jc16; sink(16, 15, v); acc += e16 * v; }
    sink(17, 15, 0.0);
    { const double v = 31.998 * gj * jc18; sink(18, 15, v); acc += e18 * v; }
    { const double v = 14.027 * gj * jc19; sink(19, 15, v); acc += e19 * v; }
    { const double v = 30.07 * gj * jc20; sink(20, 15, v); acc += e20 * v; }
    { const double v = 16.043 * gj * jc21; sink(21, 15, v); acc += e21 * v; }
    { const double v = 45.017 * gj * jc22; sink(22, 15, v); acc += e22 * v; }
    { const double v = 44.009 * gj * jc23; sink(23, 15, v); acc += e23 * v; }
    const double cvj = (ld2(rh0 + 30).x + (ld2(rh1 + 30).x - ld2(rh0 + 30).x) * Td) - 332.1817818617659;
    sink(NS, 15, -(acc + Tdot * cvj * (((negmask >> 15) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 16: d/drho_16 (H2O)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 16) & 1u ? 0.0 : 0.055509297807382736;
    double jc0 = 0.0;
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
    double jc13 = 0.0;
    double jc14 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc21 = 0.0;
    double jc22 = 0.0;
    double jc23 = 0.0;
    const double2 A0_310 = ld2(rk0 + 54), A1_310 = ld2(rk1 + 54);
    const double2 A0_311 = ld2(rk0 + 64), A1_311 = ld2(rk1 + 64);
    const double2 A0_312 = ld2(rk0 + 76), A1_312 = ld2(rk1 + 76);
    const double2 A0_313 = ld2(rk0 + 78), A1_313 = ld2(rk1 + 78);
    {
      const double dq = 15.399999618530273 * scr[0 * SST];
      jc9 += -2.0 * dq;
      jc18 += 1.0 * dq;
    }
    {
      const double dq = 6.0 * scr[1 * SST];
      jc3 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = 6.0 * scr[2 * SST];
      jc2 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[5 * SST];
      jc0 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[8 * SST];
      jc4 += 1.0 * dq;
      jc8 += -2.0 * dq;
    }
    { const double2 a0 = A0_310, a1 = A1_310;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M9 * 1.0 + 3.6500000953674316 * scr[10 * SST];
      jc3 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    {
      const double dq = -scr[78 * SST] * c9;
      jc0 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc16 += 1.0 * dq;
    }
    { const double2 a0 = A0_311, a1 = A1_311;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c3;
      jc1 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    {
      const double dq = 6.0 * scr[11 * SST];
      jc8 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = 6.0 * scr[14 * SST];
      jc8 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc21 += 1.0 * dq;
    }
    {
      const double dq = 6.0 * scr[17 * SST];
      jc5 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = 6.0 * scr[20 * SST];
      jc6 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += -1.0 * dq;
    }
    { const double2 a0 = A0_312, a1 = A1_312;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c7;
      jc6 += -1.0 * dq;
      jc7 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    {
      const double dq = 6.0 * scr[23 * SST];
      jc8 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc15 += -1.0 * dq;
    }
    {
      const double dq = 6.0 * scr[26 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    { const double2 a0 = A0_313, a1 = A1_313;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c8;
      jc3 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc16 += 1.0 * dq;
    }
    const double2 A0_314 = ld2(rk0 + 80), A1_314 = ld2(rk1 + 80);
    const double2 A0_315 = ld2(rk0 + 82), A1_315 = ld2(rk1 + 82);
    const double2 A0_316 = ld2(rk0 + 84), A1_316 = ld2(rk1 + 84);
    const double2 A0_317 = ld2(rk0 + 86), A1_317 = ld2(rk1 + 86);
    const double2 A0_318 = ld2(rk0 + 92), A1_318 = ld2(rk1 + 92);
    const double2 A0_319 = ld2(rk0 + 94), A1_319 = ld2(rk1 + 94);
    const double2 A0_320 = ld2(rk0 + 96), A1_320 = ld2(rk1 + 96);
    const double2 A0_321 = ld2(rk0 + 100), A1_321 = ld2(rk1 + 100);
    const double2 A0_322 = ld2(rk0 + 102), A1_322 = ld2(rk1 + 102);
    const double2 A0_323 = ld2(rk0 + 104), A1_323 = ld2(rk1 + 104);
    const double2 A0_324 = ld2(rk0 + 106), A1_324 = ld2(rk1 + 106);
    const double2 A0_325 = ld2(rk0 + 134), A1_325 = ld2(rk1 + 134);
    const double2 A0_326 = ld2(rk0 + 138), A1_326 = ld2(rk1 + 138);
    {
      const double dq = 6.0 * scr[29 * SST];
      jc1 += 1.0 * dq;
      jc3 += -2.0 * dq;
    }
    { const double2 a0 = A0_314, a1 = A1_314;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c9;
      jc3 += -2.0 * dq;
      jc9 += 1.0 * dq;
      jc16 += 1.0 * dq;
    }
    { const double2 a0 = A0_315, a1 = A1_315;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c18;
      jc0 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc16 += 1.0 * dq;
      jc18 += 1.0 * dq;
    }
    { const double2 a0 = A0_316, a1 = A1_316;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c0;
      jc0 += 1.0 * dq;
      jc1 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    { const double2 a0 = A0_317, a1 = A1_317;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c0;
      jc0 += 1.0 * dq;
      jc1 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    { const double2 a0 = A0_318, a1 = A1_318;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c19;
      jc3 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc16 += 1.0 * dq;
      jc19 += 1.0 * dq;
    }
    { const double2 a0 = A0_319, a1 = A1_319;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c7;
      jc3 += -1.0 * dq;
      jc7 += 1.0 * dq;
      jc11 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    { const double2 a0 = A0_320, a1 = A1_320;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c11;
      jc3 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc16 += 1.0 * dq;
      jc21 += -1.0 * dq;
    }
    { const double2 a0 = A0_321, a1 = A1_321;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc3 += -1.0 * dq;
      jc5 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    { const double2 a0 = A0_322, a1 = A1_322;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c5;
      jc3 += -1.0 * dq;
      jc5 += 1.0 * dq;
      jc13 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    { const double2 a0 = A0_323, a1 = A1_323;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c13;
      jc3 += -1.0 * dq;
      jc6 += -1.0 * dq;
      jc13 += 1.0 * dq;
      jc16 += 1.0 * dq;
    }
    { const double2 a0 = A0_324, a1 = A1_324;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c15;
      jc3 += -1.0 * dq;
      jc12 += -1.0 * dq;
      jc15 += 1.0 * dq;
      jc16 += 1.0 * dq;
    }
    { const double2 a0 = A0_325, a1 = A1_325;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc16 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_326, a1 = A1_326;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c7 - (a0.y + (a1.y - a0.y) * Td) * c19;
      jc7 += -1.0 * dq;
      jc16 += 0.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = 6.0 * scr[32 * SST];
      jc11 += -2.0 * dq;
      jc20 += 1.0 * dq;
    }
    {
      const double dq = scr[35 * SST];
      jc2 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc8 += 1.0 * dq;
    }
    const double2 A0_327 = ld2(rk0 + 194), A1_327 = ld2(rk1 + 194);
    { const double2 a0 = A0_327, a1 = A1_327;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c10;
      jc3 += 1.0 * dq;
      jc10 += -1.0 * dq;
      jc14 += 1.0 * dq;
      jc16 += -1.0 * dq;
    }
    {
      const double dq = scr[37 * SST];
      jc8 += 1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc8 += -1.0 * dq;
      jc22 += 1.0 * dq;
      jc23 += -1.0 * dq;
    }
    {
      const double dq = scr[39 * SST];
      jc2 += 1.0 * dq;
      jc3 += 1.0 * dq;
      jc22 += -1.0 * dq;
    }
    {
      const double dq = scr[40 * SST];
      jc2 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc22 += 1.0 * dq;
    }
    double acc = 0.0;
    { const double v = 33.006 * gj * jc0; sink(0, 16, v); acc += e0 * v; }
    { const double v = 34.014 * gj * jc1; sink(1, 16, v); acc += e1 * v; }
    { const double v = 28.01 * gj * jc2; sink(2, 16, v); acc += e2 * v; }
    { const double v = 17.007 * gj * jc3; sink(3, 16, v); acc += e3 * v; }
    { const double v = 2.016 * gj * jc4; sink(4, 16, v); acc += e4 * v; }
    { const double v = 29.018 * gj * jc5; sink(5, 16, v); acc += e5 * v; }
    { const double v = 31.034 * gj * jc6; sink(6, 16, v); acc += e6 * v; }
    { const double v = 14.027 * gj * jc7; sink(7, 16, v); acc += e7 * v; }
    { const double v = 1.008 * gj * jc8; sink(8, 16, v); acc += e8 * v; }
    { const double v = 15.999 * gj * jc9; sink(9, 16, v); acc += e9 * v; }
    { const double v = 38.049 * gj * jc10; sink(10, 16, v); acc += e10 * v; }
    { const double v = 15.035 * gj * jc11; sink(11, 16, v); acc += e11 * v; }
    { const double v = 26.038 * gj * jc12; sink(12, 16, v); acc += e12 * v; }
    { const double v = 30.026 * gj * jc13; sink(13, 16, v); acc += e13 * v; }
    { const double v = 39.057 * gj * jc14; sink(14, 16, v); acc += e14 * v; }
    { const double v = 25.03 * gj * jc15; sink(15, 16, v); acc += e15 * v; }
    { const double v = 18.015 * gj * jc16; sink(16, 16, v); acc += e16 * v; }
    sink(17, 16, 0.0);
    { const double v = 31.998 * gj * jc18; sink(18, 16, v); acc += e18 * v; }
    { const double v = 14.027 * gj * jc19; sink(19, 16, v); acc += e19 * v; }
    { const double v = 30.07 * gj * jc20; sink(20, 16, v); acc += e20 * v; }
    { const double v = 16.043 * gj * jc21; sink(21, 16, v); acc += e21 * v; }
    { const double v = 45.017 * gj * jc22; sink(22, 16, v); acc += e22 * v; }
    { const double v = 44.009 * gj * jc23; sink(23, 16, v); acc += e23 * v; }
    const double cvj = (ld2(rh0 + 32).x + (ld2(rh1 + 32).x - ld2(rh0 + 32).x) * Td) - 461.53261171246186;
    sink(NS, 16, -(acc + Tdot * cvj * (((negmask >> 16) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 17: d/drho_17 (CH3O2)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 17) & 1u ? 0.0 : 0.0212616673399528;
    double jc0 = 0.0;
    double jc1 = 0.0;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc5 = 0.0;
    double jc6 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc13 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc17 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc21 = 0.0;
    double jc22 = 0.0;
    double jc23 = 0.0;
    const double2 A0_328 = ld2(rk0 + 172), A1_328 = ld2(rk1 + 172);
    {
      const double dq = scr[0 * SST];
      jc9 += -2.0 * dq;
      jc18 += 1.0 * dq;
    }
    {
      const double dq = scr[1 * SST];
      jc3 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc2 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc0 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc4 += 1.0 * dq;
      jc8 += -2.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc3 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    {
      const double dq = scr[11 * SST];
      jc8 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc8 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc21 += 1.0 * dq;
    }
    {
      const double dq = scr[17 * SST];
      jc5 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = scr[20 * SST];
      jc6 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += -1.0 * dq;
    }
    {
      const double dq = scr[23 * SST];
      jc8 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc15 += -1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = scr[29 * SST];
      jc1 += 1.0 * dq;
      jc3 += -2.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc11 += -2.0 * dq;
      jc20 += 1.0 * dq;
    }
    {
      const double dq = scr[36 * SST];
      jc2 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc8 += 1.0 * dq;
    }
    { const double2 a0 = A0_328, a1 = A1_328;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * 1.0;
      jc11 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    const double2 A0_329 = ld2(rk0 + 178), A1_329 = ld2(rk1 + 178);
    const double2 A0_330 = ld2(rk0 + 184), A1_330 = ld2(rk1 + 184);
    const double2 A0_331 = ld2(rk0 + 186), A1_331 = ld2(rk1 + 186);
    { const double2 a0 = A0_329, a1 = A1_329;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c11;
      jc6 += 2.0 * dq;
      jc11 += -1.0 * dq;
      jc17 += -1.0 * dq;
    }
    { const double2 a0 = A0_330, a1 = A1_330;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c0;
      jc0 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc6 += 1.0 * dq;
      jc17 += -1.0 * dq;
      jc18 += 1.0 * dq;
    }
    { const double2 a0 = A0_331, a1 = A1_331;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 2.0*c17;
      jc6 += 2.0 * dq;
      jc17 += -2.0 * dq;
      jc18 += 1.0 * dq;
    }
    {
      const double dq = scr[37 * SST];
      jc8 += 1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc8 += -1.0 * dq;
      jc22 += 1.0 * dq;
      jc23 += -1.0 * dq;
    }
    {
      const double dq = scr[39 * SST];
      jc2 += 1.0 * dq;
      jc3 += 1.0 * dq;
      jc22 += -1.0 * dq;
    }
    {
      const double dq = scr[40 * SST];
      jc2 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc22 += 1.0 * dq;
    }
    double acc = 0.0;
    { const double v = 33.006 * gj * jc0; sink(0, 17, v); acc += e0 * v; }
    { const double v = 34.014 * gj * jc1; sink(1, 17, v); acc += e1 * v; }
    { const double v = 28.01 * gj * jc2; sink(2, 17, v); acc += e2 * v; }
    { const double v = 17.007 * gj * jc3; sink(3, 17, v); acc += e3 * v; }
    { const double v = 2.016 * gj * jc4; sink(4, 17, v); acc += e4 * v; }
    { const double v = 29.018 * gj * jc5; sink(5, 17, v); acc += e5 * v; }
    { const double v = 31.034 * gj * jc6; sink(6, 17, v); acc += e6 * v; }
    sink(7, 17, 0.0);
    { const double v = 1.008 * gj * jc8; sink(8, 17, v); acc += e8 * v; }
    { const double v = 15.999 * gj * jc9; sink(9, 17, v); acc += e9 * v; }
    sink(10, 17, 0.0);
    { const double v = 15.035 * gj * jc11; sink(11, 17, v); acc += e11 * v; }
    { const double v = 26.038 * gj * jc12; sink(12, 17, v); acc += e12 * v; }
    { const double v = 30.026 * gj * jc13; sink(13, 17, v); acc += e13 * v; }
    sink(14, 17, 0.0);
    { const double v = 25.03 * gj * jc15; sink(15, 17, v); acc += e15 * v; }
    { const double v = 18.015 * gj * jc16; sink(16, 17, v); acc += e16 * v; }
    { const double v = 47.033 * gj * jc17; sink(17, 17, v); acc += e17 * v; }
    { const double v = 31.998 * gj * jc18; sink(18, 17, v); acc += e18 * v; }
    { const double v = 14.027 * gj * jc19; sink(19, 17, v); acc += e19 * v; }
    { const double v = 30.07 * gj * jc20; sink(20, 17, v); acc += e20 * v; }
    { const double v = 16.043 * gj * jc21; sink(21, 17, v); acc += e21 * v; }
    { const double v = 45.017 * gj * jc22; sink(22, 17, v); acc += e22 * v; }
    { const double v = 44.009 * gj * jc23; sink(23, 17, v); acc += e23 * v; }
    const double cvj = (ld2(rh0 + 34).x + (ld2(rh1 + 34).x - ld2(rh0 + 34).x) * Td) - 176.78034571471096;
    sink(NS, 17, -(acc + Tdot * cvj * (((negmask >> 17) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 18: d/drho_18 (O2)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 18) & 1u ? 0.0 : 0.03125195324707794;
    double jc0 = 0.0;
    double jc1 = 0.0;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc5 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc13 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc17 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc21 = 0.0;
    double jc22 = 0.0;
    double jc23 = 0.0;
    const double2 A0_332 = ld2(rk0 + 0), A1_332 = ld2(rk1 + 0);
    const double2 A0_333 = ld2(rk0 + 38), A1_333 = ld2(rk1 + 38);
    const double2 A0_334 = ld2(rk0 + 40), A1_334 = ld2(rk1 + 40);
    const double2 A0_335 = ld2(rk0 + 42), A1_335 = ld2(rk1 + 42);
    const double2 A0_336 = ld2(rk0 + 58), A1_336 = ld2(rk1 + 58);
    { const double2 a0 = A0_332, a1 = A1_332;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M0 * 1.0 + scr[0 * SST];
      jc9 += -2.0 * dq;
      jc18 += 1.0 * dq;
    }
    {
      const double dq = scr[1 * SST];
      jc3 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = -scr[46 * SST] * c3;
      jc0 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc18 += 1.0 * dq;
    }
    {
      const double dq = 6.0 * scr[2 * SST];
      jc2 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[71 * SST] * c2;
      jc2 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[73 * SST] * c13;
      jc0 += 1.0 * dq;
      jc5 += 1.0 * dq;
      jc13 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_333, a1 = A1_333;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M3 * c8;
      jc0 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_334, a1 = A1_334;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M4 * c8 + scr[4 * SST];
      jc0 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_335, a1 = A1_335;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M5 * c8;
      jc0 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[75 * SST] * c8;
      jc3 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc4 += 1.0 * dq;
      jc8 += -2.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc3 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    { const double2 a0 = A0_336, a1 = A1_336;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c4;
      jc0 += -1.0 * dq;
      jc4 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += 1.0 * dq;
    }
    {
      const double dq = scr[11 * SST];
      jc8 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc8 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc21 += 1.0 * dq;
    }
    {
      const double dq = scr[17 * SST];
      jc5 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    const double2 A0_337 = ld2(rk0 + 82), A1_337 = ld2(rk1 + 82);
    const double2 A0_338 = ld2(rk0 + 110), A1_338 = ld2(rk1 + 110);
    const double2 A0_339 = ld2(rk0 + 112), A1_339 = ld2(rk1 + 112);
    const double2 A0_340 = ld2(rk0 + 116), A1_340 = ld2(rk1 + 116);
    const double2 A0_341 = ld2(rk0 + 124), A1_341 = ld2(rk1 + 124);
    const double2 A0_342 = ld2(rk0 + 134), A1_342 = ld2(rk1 + 134);
    const double2 A0_343 = ld2(rk0 + 148), A1_343 = ld2(rk1 + 148);
    const double2 A0_344 = ld2(rk0 + 150), A1_344 = ld2(rk1 + 150);
    const double2 A0_345 = ld2(rk0 + 162), A1_345 = ld2(rk1 + 162);
    {
      const double dq = scr[20 * SST];
      jc6 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += -1.0 * dq;
    }
    {
      const double dq = scr[23 * SST];
      jc8 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc15 += -1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = scr[29 * SST];
      jc1 += 1.0 * dq;
      jc3 += -2.0 * dq;
    }
    { const double2 a0 = A0_337, a1 = A1_337;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c16;
      jc0 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc16 += 1.0 * dq;
      jc18 += 1.0 * dq;
    }
    { const double2 a0 = A0_338, a1 = A1_338;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c1;
      jc0 += -2.0 * dq;
      jc1 += 1.0 * dq;
      jc18 += 1.0 * dq;
    }
    { const double2 a0 = A0_339, a1 = A1_339;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c1;
      jc0 += -2.0 * dq;
      jc1 += 1.0 * dq;
      jc18 += 1.0 * dq;
    }
    { const double2 a0 = A0_340, a1 = A1_340;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c21;
      jc0 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc21 += 1.0 * dq;
    }
    { const double2 a0 = A0_341, a1 = A1_341;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c19;
      jc3 += 1.0 * dq;
      jc5 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[41 * SST] * c7;
      jc2 += 1.0 * dq;
      jc3 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_342, a1 = A1_342;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c7;
      jc2 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc16 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_343, a1 = A1_343;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c11;
      jc6 += 1.0 * dq;
      jc9 += 1.0 * dq;
      jc11 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_344, a1 = A1_344;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c11;
      jc3 += 1.0 * dq;
      jc11 += -1.0 * dq;
      jc13 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc11 += -2.0 * dq;
      jc20 += 1.0 * dq;
    }
    {
      const double dq = scr[36 * SST];
      jc2 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc8 += 1.0 * dq;
    }
    { const double2 a0 = A0_345, a1 = A1_345;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c5;
      jc0 += 1.0 * dq;
      jc2 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    const double2 A0_346 = ld2(rk0 + 164), A1_346 = ld2(rk1 + 164);
    const double2 A0_347 = ld2(rk0 + 166), A1_347 = ld2(rk1 + 166);
    const double2 A0_348 = ld2(rk0 + 172), A1_348 = ld2(rk1 + 172);
    const double2 A0_349 = ld2(rk0 + 180), A1_349 = ld2(rk1 + 180);
    { const double2 a0 = A0_346, a1 = A1_346;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c6;
      jc0 += 1.0 * dq;
      jc6 += -1.0 * dq;
      jc13 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_347, a1 = A1_347;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c15;
      jc2 += 1.0 * dq;
      jc5 += 1.0 * dq;
      jc15 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_348, a1 = A1_348;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c11;
      jc11 += -1.0 * dq;
      jc17 += 1.0 * dq;
      jc18 += -1.0 * dq;
    }
    { const double2 a0 = A0_349, a1 = A1_349;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c12;
      jc5 += 2.0 * dq;
      jc12 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[37 * SST];
      jc8 += 1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc8 += -1.0 * dq;
      jc22 += 1.0 * dq;
      jc23 += -1.0 * dq;
    }
    {
      const double dq = scr[39 * SST];
      jc2 += 1.0 * dq;
      jc3 += 1.0 * dq;
      jc22 += -1.0 * dq;
    }
    {
      const double dq = scr[40 * SST];
      jc2 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc22 += 1.0 * dq;
    }
    double acc = 0.0;
    { const double v = 33.006 * gj * jc0; sink(0, 18, v); acc += e0 * v; }
    { const double v = 34.014 * gj * jc1; sink(1, 18, v); acc += e1 * v; }
    { const double v = 28.01 * gj * jc2; sink(2, 18, v); acc += e2 * v; }
    { const double v = 17.007 * gj * jc3; sink(3, 18, v); acc += e3 * v; }
    { const double v = 2.016 * gj * jc4; sink(4, 18, v); acc += e4 * v; }
    { const double v = 29.018 * gj * jc5; sink(5, 18, v); acc += e5 * v; }
    { const double v = 31.034 * gj * jc6; sink(6, 18, v); acc += e6 * v; }
    { const double v = 14.027 * gj * jc7; sink(7, 18, v); acc += e7 * v; }
    { const double v = 1.008 * gj * jc8; sink(8, 18, v); acc += e8 * v; }
    { const double v = 15.999 * gj * jc9; sink(9, 18, v); acc += e9 * v; }
    sink(10, 18, 0.0);
    { const double v = 15.035 * gj * jc11; sink(11, 18, v); acc += e11 * v; }
    { const double v = 26.038 * gj * jc12; sink(12, 18, v); acc += e12 * v; }
    { const double v = 30.026 * gj * jc13; sink(13, 18, v); acc += e13 * v; }
    sink(14, 18, 0.0);
    { const double v = 25.03 * gj * jc15; sink(15, 18, v); acc += e15 * v; }
    { const double v = 18.015 * gj * jc16; sink(16, 18, v); acc += e16 * v; }
    { const double v = 47.033 * gj * jc17; sink(17, 18, v); acc += e17 * v; }
    { const double v = 31.998 * gj * jc18; sink(18, 18, v); acc += e18 * v; }
    { const double v = 14.027 * gj * jc19; sink(19, 18, v); acc += e19 * v; }
    { const double v = 30.07 * gj * jc20; sink(20, 18, v); acc += e20 * v; }
    { const double v = 16.043 * gj * jc21; sink(21, 18, v); acc += e21 * v; }
    { const double v = 45.017 * gj * jc22; sink(22, 18, v); acc += e22 * v; }
    { const double v = 44.009 * gj * jc23; sink(23, 18, v); acc += e23 * v; }
    const double cvj = (ld2(rh0 + 36).x + (ld2(rh1 + 36).x - ld2(rh0 + 36).x) * Td) - 259.844677792362;
    sink(NS, 18, -(acc + Tdot * cvj * (((negmask >> 18) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 19: d/drho_19 (CH2)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 19) & 1u ? 0.0 : 0.07129108148570615;
    double jc0 = 0.0;
    double jc1 = 0.0;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc5 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc13 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc21 = 0.0;
    double jc22 = 0.0;
    double jc23 = 0.0;
    const double2 A0_350 = ld2(rk0 + 88), A1_350 = ld2(rk1 + 88);
    {
      const double dq = scr[0 * SST];
      jc9 += -2.0 * dq;
      jc18 += 1.0 * dq;
    }
    {
      const double dq = scr[1 * SST];
      jc3 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = scr[49 * SST] * c9;
      jc5 += 1.0 * dq;
      jc8 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc2 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = -scr[70 * SST] * c2;
      jc2 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc12 += -1.0 * dq;
      jc19 += 1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc0 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc4 += 1.0 * dq;
      jc8 += -2.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc3 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    {
      const double dq = scr[12 * SST] * c8 + scr[11 * SST];
      jc8 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc8 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc21 += 1.0 * dq;
    }
    {
      const double dq = scr[17 * SST];
      jc5 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = scr[20 * SST];
      jc6 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += -1.0 * dq;
    }
    {
      const double dq = scr[23 * SST];
      jc8 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc15 += -1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = scr[29 * SST];
      jc1 += 1.0 * dq;
      jc3 += -2.0 * dq;
    }
    { const double2 a0 = A0_350, a1 = A1_350;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c3;
      jc3 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc13 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    const double2 A0_351 = ld2(rk0 + 92), A1_351 = ld2(rk1 + 92);
    const double2 A0_352 = ld2(rk0 + 114), A1_352 = ld2(rk1 + 114);
    const double2 A0_353 = ld2(rk0 + 124), A1_353 = ld2(rk1 + 124);
    const double2 A0_354 = ld2(rk0 + 126), A1_354 = ld2(rk1 + 126);
    const double2 A0_355 = ld2(rk0 + 128), A1_355 = ld2(rk1 + 128);
    const double2 A0_356 = ld2(rk0 + 130), A1_356 = ld2(rk1 + 130);
    const double2 A0_357 = ld2(rk0 + 138), A1_357 = ld2(rk1 + 138);
    const double2 A0_358 = ld2(rk0 + 142), A1_358 = ld2(rk1 + 142);
    const double2 A0_359 = ld2(rk0 + 144), A1_359 = ld2(rk1 + 144);
    { const double2 a0 = A0_351, a1 = A1_351;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c16;
      jc3 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc16 += 1.0 * dq;
      jc19 += 1.0 * dq;
    }
    { const double2 a0 = A0_352, a1 = A1_352;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c0;
      jc0 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc13 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    { const double2 a0 = A0_353, a1 = A1_353;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c18;
      jc3 += 1.0 * dq;
      jc5 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc19 += -1.0 * dq;
    }
    { const double2 a0 = A0_354, a1 = A1_354;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c4;
      jc4 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc11 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    { const double2 a0 = A0_355, a1 = A1_355;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * 2.0*c19;
      jc4 += 1.0 * dq;
      jc12 += 1.0 * dq;
      jc19 += -2.0 * dq;
    }
    { const double2 a0 = A0_356, a1 = A1_356;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c21;
      jc11 += 2.0 * dq;
      jc19 += -1.0 * dq;
      jc21 += -1.0 * dq;
    }
    { const double2 a0 = A0_357, a1 = A1_357;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c16;
      jc7 += -1.0 * dq;
      jc16 += 0.0 * dq;
      jc19 += 1.0 * dq;
    }
    { const double2 a0 = A0_358, a1 = A1_358;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 0.0 * dq;
      jc7 += -1.0 * dq;
      jc19 += 1.0 * dq;
    }
    { const double2 a0 = A0_359, a1 = A1_359;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c23;
      jc7 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc23 += 0.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc11 += -2.0 * dq;
      jc20 += 1.0 * dq;
    }
    {
      const double dq = scr[36 * SST];
      jc2 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc8 += 1.0 * dq;
    }
    {
      const double dq = scr[37 * SST];
      jc8 += 1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc8 += -1.0 * dq;
      jc22 += 1.0 * dq;
      jc23 += -1.0 * dq;
    }
    {
      const double dq = scr[39 * SST];
      jc2 += 1.0 * dq;
      jc3 += 1.0 * dq;
      jc22 += -1.0 * dq;
    }
    {
      const double dq = scr[40 * SST];
      jc2 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc22 += 1.0 * dq;
    }
    double acc = 0.0;
    { const double v = 33.006 * gj * jc0; sink(0, 19, v); acc += e0 * v; }
    { const double v = 34.014 * gj * jc1; sink(1, 19, v); acc += e1 * v; }
    { const double v = 28.01 * gj * jc2; sink(2, 19, v); acc += e2 * v; }
    { const double v = 17.007 * gj * jc3; sink(3, 19, v); acc += e3 * v; }
    { const double v = 2.016 * gj * jc4; sink(4, 19, v); acc += e4 * v; }
    { const double v = 29.018 * gj * jc5; sink(5, 19, v); acc += e5 * v; }
    { const double v = 31.034 * gj * jc6; sink(6, 19, v); acc += e6 * v; }
    { const double v = 14.027 * gj * jc7; sink(7, 19, v); acc += e7 * v; }
    { const double v = 1.008 * gj * jc8; sink(8, 19, v); acc += e8 * v; }
    { const double v = 15.999 * gj * jc9; sink(9, 19, v); acc += e9 * v; }
    sink(10, 19, 0.0);
    { const double v = 15.035 * gj * jc11; sink(11, 19, v); acc += e11 * v; }
    { const double v = 26.038 * gj * jc12; sink(12, 19, v); acc += e12 * v; }
    { const double v = 30.026 * gj * jc13; sink(13, 19, v); acc += e13 * v; }
    sink(14, 19, 0.0);
    { const double v = 25.03 * gj * jc15; sink(15, 19, v); acc += e15 * v; }
    { const double v = 18.015 * gj * jc16; sink(16, 19, v); acc += e16 * v; }
    sink(17, 19, 0.0);
    { const double v = 31.998 * gj * jc18; sink(18, 19, v); acc += e18 * v; }
    { const double v = 14.027 * gj * jc19; sink(19, 19, v); acc += e19 * v; }
    { const double v = 30.07 * gj * jc20; sink(20, 19, v); acc += e20 * v; }
    { const double v = 16.043 * gj * jc21; sink(21, 19, v); acc += e21 * v; }
    { const double v = 45.017 * gj * jc22; sink(22, 19, v); acc += e22 * v; }
    { const double v = 44.009 * gj * jc23; sink(23, 19, v); acc += e23 * v; }
    const double cvj = (ld2(rh0 + 38).x + (ld2(rh1 + 38).x - ld2(rh0 + 38).x) * Td) - 592.7504099237186;
    sink(NS, 19, -(acc + Tdot * cvj * (((negmask >> 19) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 20: d/drho_20 (C2H6)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 20) & 1u ? 0.0 : 0.03325573661456601;
    double jc0 = 0.0;
    double jc1 = 0.0;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc5 = 0.0;
    double jc6 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc13 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc21 = 0.0;
    double jc22 = 0.0;
    double jc23 = 0.0;
    {
      const double dq = 3.0 * scr[0 * SST];
      jc9 += -2.0 * dq;
      jc18 += 1.0 * dq;
    }
    {
      const double dq = 3.0 * scr[1 * SST];
      jc3 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = 3.0 * scr[2 * SST];
      jc2 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = 1.5 * scr[3 * SST];
      jc0 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = 3.0 * scr[6 * SST];
      jc4 += 1.0 * dq;
      jc8 += -2.0 * dq;
    }
    {
      const double dq = 3.0 * scr[10 * SST];
      jc3 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    {
      const double dq = 3.0 * scr[11 * SST];
      jc8 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = 3.0 * scr[14 * SST];
      jc8 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc21 += 1.0 * dq;
    }
    {
      const double dq = 3.0 * scr[17 * SST];
      jc5 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = 3.0 * scr[20 * SST];
      jc6 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += -1.0 * dq;
    }
    {
      const double dq = 3.0 * scr[23 * SST];
      jc8 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc15 += -1.0 * dq;
    }
    {
      const double dq = 3.0 * scr[26 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = 3.0 * scr[29 * SST];
      jc1 += 1.0 * dq;
      jc3 += -2.0 * dq;
    }
    {
      const double dq = -scr[34 * SST] * 1.0 + 3.0 * scr[32 * SST];
      jc11 += -2.0 * dq;
      jc20 += 1.0 * dq;
    }
    {
      const double dq = 3.0 * scr[36 * SST];
      jc2 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc8 += 1.0 * dq;
    }
    {
      const double dq = scr[37 * SST];
      jc8 += 1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc8 += -1.0 * dq;
      jc22 += 1.0 * dq;
      jc23 += -1.0 * dq;
    }
    {
      const double dq = scr[39 * SST];
      jc2 += 1.0 * dq;
      jc3 += 1.0 * dq;
      jc22 += -1.0 * dq;
    }
    {
      const double dq = scr[40 * SST];
      jc2 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc22 += 1.0 * dq;
    }
    double acc = 0.0;
    { const double v = 33.006 * gj * jc0; sink(0, 20, v); acc += e0 * v; }
    { const double v = 34.014 * gj * jc1; sink(1, 20, v); acc += e1 * v; }
    { const double v = 28.01 * gj * jc2; sink(2, 20, v); acc += e2 * v; }
    { const double v = 17.007 * gj * jc3; sink(3, 20, v); acc += e3 * v; }
    { const double v = 2.016 * gj * jc4; sink(4, 20, v); acc += e4 * v; }
    { const double v = 29.018 * gj * jc5; sink(5, 20, v); acc += e5 * v; }
    { const double v = 31.034 * gj * jc6; sink(6, 20, v); acc += e6 * v; }
    sink(7, 20, 0.0);
    { const double v = 1.008 * gj * jc8; sink(8, 20, v); acc += e8 * v; }
    { const double v = 15.999 * gj * jc9; sink(9, 20, v); acc += e9 * v; }
    sink(10, 20, 0.0);
    { const double v = 15.035 * gj * jc11; sink(11, 20, v); acc += e11 * v; }
    { const double v = 26.038 * gj * jc12; sink(12, 20, v); acc += e12 * v; }
    { const double v = 30.026 * gj * jc13; sink(13, 20, v); acc += e13 * v; }
    sink(14, 20, 0.0);
    { const double v = 25.03 * gj * jc15; sink(15, 20, v); acc += e15 * v; }
    { const double v = 18.015 * gj * jc16; sink(16, 20, v); acc += e16 * v; }
    sink(17, 20, 0.0);
    { const double v = 31.998 * gj * jc18; sink(18, 20, v); acc += e18 * v; }
    { const double v = 14.027 * gj * jc19; sink(19, 20, v); acc += e19 * v; }
    { const double v = 30.07 * gj * jc20; sink(20, 20, v); acc += e20 * v; }
    { const double v = 16.043 * gj * jc21; sink(21, 20, v); acc += e21 * v; }
    { const double v = 45.017 * gj * jc22; sink(22, 20, v); acc += e22 * v; }
    { const double v = 44.009 * gj * jc23; sink(23, 20, v); acc += e23 * v; }
    const double cvj = (ld2(rh0 + 40).x + (ld2(rh1 + 40).x - ld2(rh0 + 40).x) * Td) - 276.50515463917526;
    sink(NS, 20, -(acc + Tdot * cvj * (((negmask >> 20) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 21: d/drho_21 (CH4)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 21) & 1u ? 0.0 : 0.06233248145608677;
    double jc0 = 0.0;
    double jc1 = 0.0;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc5 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc13 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc21 = 0.0;
    double jc22 = 0.0;
    double jc23 = 0.0;
    const double2 A0_360 = ld2(rk0 + 66), A1_360 = ld2(rk1 + 66);
    const double2 A0_361 = ld2(rk0 + 96), A1_361 = ld2(rk1 + 96);
    {
      const double dq = 2.0 * scr[0 * SST];
      jc9 += -2.0 * dq;
      jc18 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[1 * SST];
      jc3 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = scr[57 * SST] * c9;
      jc3 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc21 += -1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[2 * SST];
      jc2 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc0 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[6 * SST];
      jc4 += 1.0 * dq;
      jc8 += -2.0 * dq;
    }
    {
      const double dq = 2.0 * scr[10 * SST];
      jc3 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[11 * SST];
      jc8 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = -scr[16 * SST] * 1.0 + 2.0 * scr[14 * SST];
      jc8 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc21 += 1.0 * dq;
    }
    { const double2 a0 = A0_360, a1 = A1_360;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c8;
      jc4 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc21 += -1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[17 * SST];
      jc5 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[20 * SST];
      jc6 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += -1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[23 * SST];
      jc8 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc15 += -1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[26 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[29 * SST];
      jc1 += 1.0 * dq;
      jc3 += -2.0 * dq;
    }
    { const double2 a0 = A0_361, a1 = A1_361;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c3;
      jc3 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc16 += 1.0 * dq;
      jc21 += -1.0 * dq;
    }
    const double2 A0_362 = ld2(rk0 + 116), A1_362 = ld2(rk1 + 116);
    const double2 A0_363 = ld2(rk0 + 130), A1_363 = ld2(rk1 + 130);
    const double2 A0_364 = ld2(rk0 + 140), A1_364 = ld2(rk1 + 140);
    const double2 A0_365 = ld2(rk0 + 152), A1_365 = ld2(rk1 + 152);
    const double2 A0_366 = ld2(rk0 + 154), A1_366 = ld2(rk1 + 154);
    const double2 A0_367 = ld2(rk0 + 156), A1_367 = ld2(rk1 + 156);
    const double2 A0_368 = ld2(rk0 + 176), A1_368 = ld2(rk1 + 176);
    { const double2 a0 = A0_362, a1 = A1_362;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c18;
      jc0 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc18 += 1.0 * dq;
      jc21 += 1.0 * dq;
    }
    { const double2 a0 = A0_363, a1 = A1_363;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c19;
      jc11 += 2.0 * dq;
      jc19 += -1.0 * dq;
      jc21 += -1.0 * dq;
    }
    { const double2 a0 = A0_364, a1 = A1_364;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c7;
      jc7 += -1.0 * dq;
      jc11 += 2.0 * dq;
      jc21 += -1.0 * dq;
    }
    { const double2 a0 = A0_365, a1 = A1_365;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c0;
      jc0 += 1.0 * dq;
      jc1 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc21 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[32 * SST];
      jc11 += -2.0 * dq;
      jc20 += 1.0 * dq;
    }
    { const double2 a0 = A0_366, a1 = A1_366;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c2;
      jc2 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc21 += 1.0 * dq;
    }
    { const double2 a0 = A0_367, a1 = A1_367;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c5;
      jc5 += 1.0 * dq;
      jc11 += -1.0 * dq;
      jc13 += -1.0 * dq;
      jc21 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[36 * SST];
      jc2 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc8 += 1.0 * dq;
    }
    { const double2 a0 = A0_368, a1 = A1_368;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c13;
      jc6 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc13 += 1.0 * dq;
      jc21 += 1.0 * dq;
    }
    {
      const double dq = scr[37 * SST];
      jc8 += 1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[38 * SST];
      jc8 += -1.0 * dq;
      jc22 += 1.0 * dq;
      jc23 += -1.0 * dq;
    }
    {
      const double dq = scr[39 * SST];
      jc2 += 1.0 * dq;
      jc3 += 1.0 * dq;
      jc22 += -1.0 * dq;
    }
    {
      const double dq = scr[40 * SST];
      jc2 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc22 += 1.0 * dq;
    }
    double acc = 0.0;
    { const double v = 33.006 * gj * jc0; sink(0, 21, v); acc += e0 * v; }
    { const double v = 34.014 * gj * jc1; sink(1, 21, v); acc += e1 * v; }
    { const double v = 28.01 * gj * jc2; sink(2, 21, v); acc += e2 * v; }
    { const double v = 17.007 * gj * jc3; sink(3, 21, v); acc += e3 * v; }
    { const double v = 2.016 * gj * jc4; sink(4, 21, v); acc += e4 * v; }
    { const double v = 29.018 * gj * jc5; sink(5, 21, v); acc += e5 * v; }
    { const double v = 31.034 * gj * jc6; sink(6, 21, v); acc += e6 * v; }
    { const double v = 14.027 * gj * jc7; sink(7, 21, v); acc += e7 * v; }
    { const double v = 1.008 * gj * jc8; sink(8, 21, v); acc += e8 * v; }
    { const double v = 15.999 * gj * jc9; sink(9, 21, v); acc += e9 * v; }
    sink(10, 21, 0.0);
    { const double v = 15.035 * gj * jc11; sink(11, 21, v); acc += e11 * v; }
    { const double v = 26.038 * gj * jc12; sink(12, 21, v); acc += e12 * v; }
    { const double v = 30.026 * gj * jc13; sink(13, 21, v); acc += e13 * v; }
    sink(14, 21, 0.0);
    { const double v = 25.03 * gj * jc15; sink(15, 21, v); acc += e15 * v; }
    { const double v = 18.015 * gj * jc16; sink(16, 21, v); acc += e16 * v; }
    sink(17, 21, 0.0);
    { const double v = 31.998 * gj * jc18; sink(18, 21, v); acc += e18 * v; }
    { const double v = 14.027 * gj * jc19; sink(19, 21, v); acc += e19 * v; }
    { const double v = 30.07 * gj * jc20; sink(20, 21, v); acc += e20 * v; }
    { const double v = 16.043 * gj * jc21; sink(21, 21, v); acc += e21 * v; }
    { const double v = 45.017 * gj * jc22; sink(22, 21, v); acc += e22 * v; }
    { const double v = 44.009 * gj * jc23; sink(23, 21, v); acc += e23 * v; }
    const double cvj = (ld2(rh0 + 42).x + (ld2(rh1 + 42).x - ld2(rh0 + 42).x) * Td) - 518.264040391448;
    sink(NS, 21, -(acc + Tdot * cvj * (((negmask >> 21) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 22: d/drho_22 (HCO2)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 22) & 1u ? 0.0 : 0.02221383033076393;
    double jc0 = 0.0;
    double jc1 = 0.0;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc5 = 0.0;
    double jc6 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc13 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc21 = 0.0;
    double jc22 = 0.0;
    double jc23 = 0.0;
    const double2 A0_369 = ld2(rk0 + 196), A1_369 = ld2(rk1 + 196);
    {
      const double dq = scr[0 * SST];
      jc9 += -2.0 * dq;
      jc18 += 1.0 * dq;
    }
    {
      const double dq = scr[1 * SST];
      jc3 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc9 += -1.0 * dq;
    }
    {
      const double dq = scr[2 * SST];
      jc2 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = scr[3 * SST];
      jc0 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[6 * SST];
      jc4 += 1.0 * dq;
      jc8 += -2.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc3 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    {
      const double dq = scr[11 * SST];
      jc8 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = scr[14 * SST];
      jc8 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc21 += 1.0 * dq;
    }
    {
      const double dq = scr[17 * SST];
      jc5 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = scr[20 * SST];
      jc6 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += -1.0 * dq;
    }
    {
      const double dq = scr[23 * SST];
      jc8 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc15 += -1.0 * dq;
    }
    {
      const double dq = scr[26 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = scr[29 * SST];
      jc1 += 1.0 * dq;
      jc3 += -2.0 * dq;
    }
    {
      const double dq = scr[32 * SST];
      jc11 += -2.0 * dq;
      jc20 += 1.0 * dq;
    }
    {
      const double dq = scr[36 * SST];
      jc2 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc8 += 1.0 * dq;
    }
    { const double2 a0 = A0_369, a1 = A1_369;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M11 * 1.0 + scr[37 * SST];
      jc8 += 1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    const double2 A0_370 = ld2(rk0 + 200), A1_370 = ld2(rk1 + 200);
    {
      const double dq = scr[38 * SST];
      jc8 += -1.0 * dq;
      jc22 += 1.0 * dq;
      jc23 += -1.0 * dq;
    }
    { const double2 a0 = A0_370, a1 = A1_370;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M11 * 1.0 + scr[39 * SST];
      jc2 += 1.0 * dq;
      jc3 += 1.0 * dq;
      jc22 += -1.0 * dq;
    }
    {
      const double dq = scr[40 * SST];
      jc2 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc22 += 1.0 * dq;
    }
    double acc = 0.0;
    { const double v = 33.006 * gj * jc0; sink(0, 22, v); acc += e0 * v; }
    { const double v = 34.014 * gj * jc1; sink(1, 22, v); acc += e1 * v; }
    { const double v = 28.01 * gj * jc2; sink(2, 22, v); acc += e2 * v; }
    { const double v = 17.007 * gj * jc3; sink(3, 22, v); acc += e3 * v; }
    { const double v = 2.016 * gj * jc4; sink(4, 22, v); acc += e4 * v; }
    { const double v = 29.018 * gj * jc5; sink(5, 22, v); acc += e5 * v; }
    { const double v = 31.034 * gj * jc6; sink(6, 22, v); acc += e6 * v; }
    sink(7, 22, 0.0);
    { const double v = 1.008 * gj * jc8; sink(8, 22, v); acc += e8 * v; }
    { const double v = 15.999 * gj * jc9; sink(9, 22, v); acc += e9 * v; }
    sink(10, 22, 0.0);
    { const double v = 15.035 * gj * jc11; sink(11, 22, v); acc += e11 * v; }
    { const double v = 26.038 * gj * jc12; sink(12, 22, v); acc += e12 * v; }
    { const double v = 30.026 * gj * jc13; sink(13, 22, v); acc += e13 * v; }
    sink(14, 22, 0.0);
    { const double v = 25.03 * gj * jc15; sink(15, 22, v); acc += e15 * v; }
    { const double v = 18.015 * gj * jc16; sink(16, 22, v); acc += e16 * v; }
    sink(17, 22, 0.0);
    { const double v = 31.998 * gj * jc18; sink(18, 22, v); acc += e18 * v; }
    { const double v = 14.027 * gj * jc19; sink(19, 22, v); acc += e19 * v; }
    { const double v = 30.07 * gj * jc20; sink(20, 22, v); acc += e20 * v; }
    { const double v = 16.043 * gj * jc21; sink(21, 22, v); acc += e21 * v; }
    { const double v = 45.017 * gj * jc22; sink(22, 22, v); acc += e22 * v; }
    { const double v = 44.009 * gj * jc23; sink(23, 22, v); acc += e23 * v; }
    const double cvj = (ld2(rh0 + 44).x + (ld2(rh1 + 44).x - ld2(rh0 + 44).x) * Td) - 184.69711442344;
    sink(NS, 22, -(acc + Tdot * cvj * (((negmask >> 22) & 1u) ? 0.0 : 1.0)) * invD);
  }
  {  // column 23: d/drho_23 (CO2)
    const double* __restrict__ rk0 = opaque(rk0_);
    const double* __restrict__ rk1 = rk0 + WK;
    const double gj = (negmask >> 23) & 1u ? 0.0 : 0.022722624917630486;
    double jc0 = 0.0;
    double jc1 = 0.0;
    double jc2 = 0.0;
    double jc3 = 0.0;
    double jc4 = 0.0;
    double jc5 = 0.0;
    double jc6 = 0.0;
    double jc7 = 0.0;
    double jc8 = 0.0;
    double jc9 = 0.0;
    double jc11 = 0.0;
    double jc12 = 0.0;
    double jc13 = 0.0;
    double jc15 = 0.0;
    double jc16 = 0.0;
    double jc18 = 0.0;
    double jc19 = 0.0;
    double jc20 = 0.0;
    double jc21 = 0.0;
    double jc22 = 0.0;
    double jc23 = 0.0;
    const double2 A0_371 = ld2(rk0 + 20), A1_371 = ld2(rk1 + 20);
    const double2 A0_372 = ld2(rk0 + 98), A1_372 = ld2(rk1 + 98);
    {
      const double dq = 3.5999999046325684 * scr[0 * SST];
      jc9 += -2.0 * dq;
      jc18 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[1 * SST];
      jc3 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc9 += -1.0 * dq;
    }
    { const double2 a0 = A0_371, a1 = A1_371;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * M2 * 1.0 + 3.5 * scr[2 * SST];
      jc2 += -1.0 * dq;
      jc9 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = -scr[62 * SST] * c8;
      jc5 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc9 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = -scr[72 * SST] * c9;
      jc2 += -1.0 * dq;
      jc9 += 1.0 * dq;
      jc18 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    {
      const double dq = 1.5 * scr[3 * SST];
      jc0 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc18 += -1.0 * dq;
    }
    {
      const double dq = scr[9 * SST];
      jc4 += 1.0 * dq;
      jc8 += -2.0 * dq;
    }
    {
      const double dq = scr[10 * SST];
      jc3 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc16 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[11 * SST];
      jc8 += -1.0 * dq;
      jc11 += 1.0 * dq;
      jc19 += -1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[14 * SST];
      jc8 += -1.0 * dq;
      jc11 += -1.0 * dq;
      jc21 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[17 * SST];
      jc5 += -1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[20 * SST];
      jc6 += 1.0 * dq;
      jc8 += -1.0 * dq;
      jc13 += -1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[23 * SST];
      jc8 += -1.0 * dq;
      jc12 += 1.0 * dq;
      jc15 += -1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[26 * SST];
      jc2 += -1.0 * dq;
      jc4 += -1.0 * dq;
      jc13 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[29 * SST];
      jc1 += 1.0 * dq;
      jc3 += -2.0 * dq;
    }
    { const double2 a0 = A0_372, a1 = A1_372;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c8;
      jc2 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc8 += 1.0 * dq;
      jc23 += 1.0 * dq;
    }
    const double2 A0_373 = ld2(rk0 + 120), A1_373 = ld2(rk1 + 120);
    const double2 A0_374 = ld2(rk0 + 144), A1_374 = ld2(rk1 + 144);
    const double2 A0_375 = ld2(rk0 + 146), A1_375 = ld2(rk1 + 146);
    const double2 A0_376 = ld2(rk0 + 198), A1_376 = ld2(rk1 + 198);
    { const double2 a0 = A0_373, a1 = A1_373;
      const double dq = -(a0.y + (a1.y - a0.y) * Td) * c3;
      jc0 += -1.0 * dq;
      jc2 += -1.0 * dq;
      jc3 += 1.0 * dq;
      jc23 += 1.0 * dq;
    }
    { const double2 a0 = A0_374, a1 = A1_374;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c7 - (a0.y + (a1.y - a0.y) * Td) * c19;
      jc7 += -1.0 * dq;
      jc19 += 1.0 * dq;
      jc23 += 0.0 * dq;
    }
    { const double2 a0 = A0_375, a1 = A1_375;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * c7;
      jc2 += 1.0 * dq;
      jc7 += -1.0 * dq;
      jc13 += 1.0 * dq;
      jc23 += -1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[32 * SST];
      jc11 += -2.0 * dq;
      jc20 += 1.0 * dq;
    }
    {
      const double dq = 2.0 * scr[36 * SST];
      jc2 += 1.0 * dq;
      jc5 += -1.0 * dq;
      jc8 += 1.0 * dq;
    }
    {
      const double dq = scr[37 * SST];
      jc8 += 1.0 * dq;
      jc22 += -1.0 * dq;
      jc23 += 1.0 * dq;
    }
    { const double2 a0 = A0_376, a1 = A1_376;
      const double dq = (a0.x + (a1.x - a0.x) * Td) * M11 * c8 + scr[38 * SST];
      jc8 += -1.0 * dq;
      jc22 += 1.0 * dq;
      jc23 += -1.0 * dq;
    }
    {
      const double dq = scr[39 * SST];
      jc2 += 1.0 * dq;
      jc3 += 1.0 * dq;
      jc22 += -1.0 * dq;
    }
    {
      const double dq = scr[40 * SST];
      jc2 += -1.0 * dq;
      jc3 += -1.0 * dq;
      jc22 += 1.0 * dq;
    }
    double acc = 0.0;
    { const double v = 33.006 * gj * jc0; sink(0, 23, v); acc += e0 * v; }
    { const double v = 34.014 * gj * jc1; sink(1, 23, v); acc += e1 * v; }
    { const double v = 28.01 * gj * jc2; sink(2, 23, v); acc += e2 * v; }
    { const double v = 17.007 * gj * jc3; sink(3, 23, v); acc += e3 * v; }
    { const double v = 2.016 * gj * jc4; sink(4, 23, v); acc += e4 * v; }
    { const double v = 29.018 * gj * jc5; sink(5, 23, v); acc += e5 * v; }
    { const double v = 31.034 * gj * jc6; sink(6, 23, v); acc += e6 * v; }
    { const double v = 14.027 * gj * jc7; sink(7, 23, v); acc += e7 * v; }
    { const double v = 1.008 * gj * jc8; sink(8, 23, v); acc += e8 * v; }
    { const double v = 15.999 * gj * jc9; sink(9, 23, v); acc += e9 * v; }
    sink(10, 23, 0.0);
    { const double v = 15.035 * gj * jc11; sink(11, 23, v); acc += e11 * v; }
    { const double v = 26.038 * gj * jc12; sink(12, 23, v); acc += e12 * v; }
    { const double v = 30.026 * gj * jc13; sink(13, 23, v); acc += e13 * v; }
    sink(14, 23, 0.0);
    { const double v = 25.03 * gj * jc15; sink(15, 23, v); acc += e15 * v; }
    { const double v = 18.015 * gj * jc16; sink(16, 23, v); acc += e16 * v; }
    sink(17, 23, 0.0);
    { const double v = 31.998 * gj * jc18; sink(18, 23, v); acc += e18 * v; }
    { const double v = 14.027 * gj * jc19; sink(19, 23, v); acc += e19 * v; }
    { const double v = 30.07 * gj * jc20; sink(20, 23, v); acc += e20 * v; }
    { const double v = 16.043 * gj * jc21; sink(21, 23, v); acc += e21 * v; }
    { const double v = 45.017 * gj * jc22; sink(22, 23, v); acc += e22 * v; }
    { const double v = 44.009 * gj * jc23; sink(23, 23, v); acc += e23 * v; }
    const double cvj = (ld2(rh0 + 46).x + (ld2(rh1 + 46).x - ld2(rh0 + 46).x) * Td) - 188.92749210388786;
    sink(NS, 23, -(acc + Tdot * cvj * (((negmask >> 23) & 1u) ? 0.0 : 1.0)) * invD);
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
static const int p_rx_kind[] = {5, 5, 4, 4, 4, 4, 4, 4, 4, 4, 5, 4, 4, 4, 4, 4, 4, 4, 4, 5, 5, 5, 4, 5, 5, 5, 5, 5, 4, 4, 4, 4, 4, 6, 6, 4, 6, 4, 6, 4, 4, 4, 4, 6, 6, 4, 6, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 6, 4, 4, 5, 5, 4, 4, 4, 4, 4, 4, 4, 4, 4, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1};
static const int p_rx_col[] = {0, 2, 4, 6, 8, 10, 12, 14, 16, 18, 20, 22, 24, 26, 28, 30, 32, 34, 36, 38, 40, 42, 44, 46, 48, 50, 52, 54, 56, 58, 60, 62, 64, 204, 208, 66, 212, 68, 216, 70, 72, 74, 76, 220, 224, 78, 228, 80, 82, 84, 86, 88, 90, 92, 94, 96, 98, 100, 102, 104, 106, 108, 110, 112, 114, 116, 118, 120, 122, 124, 126, 128, 130, 132, 134, 136, 138, 140, 142, 144, 146, 148, 150, 152, 232, 154, 156, 158, 160, 162, 164, 166, 168, 170, 172, 174, 176, 178, 180, 182, 184, 186, 188, 190, 192, 194, 196, 198, 200, 202};
static const int p_rx_mset[] = {0, 1, -1, -1, -1, -1, -1, -1, -1, -1, 2, -1, -1, -1, -1, -1, -1, -1, -1, 3, 4, 5, -1, 6, 7, 5, 8, 9, -1, -1, -1, -1, -1, 1, 1, -1, 1, -1, 1, -1, -1, -1, -1, 1, 1, -1, 1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, 1, -1, -1, 5, 10, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, 11, 11, 11, 11};
static const int p_rx_fac[] = {521, -1, -1, -1, 274, -1, -1, -1, 264, 265, -1, -1, 259, -1, -1, -1, 260, 265, -1, -1, 264, 259, -1, -1, 256, 265, -1, -1, 274, 259, -1, -1, 257, 265, -1, -1, 256, 259, -1, -1, 275, 265, -1, -1, 264, 261, -1, -1, 263, 265, -1, -1, 258, 260, -1, -1, 263, 265, -1, -1, 264, 261, -1, -1, 267, 265, -1, -1, 269, 264, -1, -1, 277, 265, -1, -1, 267, 259, -1, -1, 258, 265, -1, -1, 279, -1, -1, -1, 261, 265, -1, -1, 258, 259, -1, -1, 261, 265, -1, -1, 279, 264, -1, -1, 269, 265, -1, -1, 261, 259, -1, -1, 262, 265, -1, -1, 269, 259, -1, -1, 268, 265, -1, -1, 271, 259, -1, -1, 268, 265, -1, -1, 275, 258, -1, -1, 258, 274, -1, -1, 279, 265, -1, -1, 269, 274, -1, -1, 261, 256, -1, -1, 264, 274, -1, -1, 256, -1, -1, -1, 264, 274, -1, -1, 256, -1, -1, -1, 264, 274, -1, -1, 256, -1, -1, -1, 264, 274, -1, -1, 265, 259, -1, -1, 520, -1, -1, -1, 260, -1, -1, -1, 520, -1, -1, -1, 260, -1, -1, -1, 520, -1, -1, -1, 260, -1, -1, -1, 520, -1, -1, -1, 260, -1, -1, -1, 264, 259, -1, -1, 272, -1, -1, -1, 264, 256, -1, -1, 272, 265, -1, -1, 264, 256, -1, -1, 260, 274, -1, -1, 264, 256, -1, -1, 515, -1, -1, -1, 264, 257, -1, -1, 260, 256, -1, -1, 264, 257, -1, -1, 272, 259, -1, -1, 275, 264, -1, -1, 267, -1, -1, -1, 267, 264, -1, -1, 277, -1, -1, -1, 277, 264, -1, -1, 267, 260, -1, -1, 264, 261, -1, -1, 269, -1, -1, -1, 264, 261, -1, -1, 258, 260, -1, -1, 269, 264, -1, -1, 262, -1, -1, -1, 269, 264, -1, -1, 260, 261, -1, -1, 262, 264, -1, -1, 269, 260, -1, -1, 262, 264, -1, -1, 267, 259, -1, -1, 262, 264, -1, -1, 263, 272, -1, -1, 271, 264, -1, -1, 268, -1, -1, -1, 258, 260, -1, -1, 269, -1, -1, -1, 260, 259, -1, -1, 264, 272, -1, -1, 515, -1, -1, -1, 257, -1, -1, -1, 515, -1, -1, -1, 272, 265, -1, -1, 256, 259, -1, -1, 272, 274, -1, -1, 257, 259, -1, -1, 272, 256, -1, -1, 257, 259, -1, -1, 272, 256, -1, -1, 275, 259, -1, -1, 269, 264, -1, -1, 263, 259, -1, -1, 269, 264, -1, -1, 267, 259, -1, -1, 275, 272, -1, -1, 267, 259, -1, -1, 263, 272, -1, -1, 277, 259, -1, -1, 267, 272, -1, -1, 258, 259, -1, -1, 279, 264, -1, -1, 261, 259, -1, -1, 258, 272, -1, -1, 269, 259, -1, -1, 272, 261, -1, -1, 262, 259, -1, -1, 269, 272, -1, -1, 268, 259, -1, -1, 271, 272, -1, -1, 268, 259, -1, -1, 267, 258, -1, -1, 512, -1, -1, -1, 257, 274, -1, -1, 512, -1, -1, -1, 257, 274, -1, -1, 275, 256, -1, -1, 269, 259, -1, -1, 267, 256, -1, -1, 277, 274, -1, -1, 267, 256, -1, -1, 262, 259, -1, -1, 258, 256, -1, -1, 279, 259, -1, -1, 269, 256, -1, -1, 257, 261, -1, -1, 275, 274, -1, -1, 261, 259, -1, -1, 275, 260, -1, -1, 267, 264, -1, -1, 531, -1, -1, -1, 268, 260, -1, -1, 275, 277, -1, -1, 523, -1, -1, -1, 263, 274, -1, -1, 258, 264, 259, -1, 263, 274, -1, -1, 258, 272, -1, -1, 263, 260, -1, -1, 267, 264, -1, -1, 263, 272, -1, -1, 275, 272, -1, -1, 263, 277, -1, -1, 523, -1, -1, -1, 263, 258, -1, -1, 275, 258, -1, -1, 263, 279, -1, -1, 275, 279, -1, -1, 263, 279, -1, -1, 269, 258, -1, -1, 267, 274, -1, -1, 262, 265, -1, -1, 267, 274, -1, -1, 269, 259, -1, -1, 267, 257, -1, -1, 277, 256, -1, -1, 523, -1, -1, -1, 276, -1, -1, -1, 267, 261, -1, -1, 277, 258, -1, -1, 269, 267, -1, -1, 277, 261, -1, -1, 261, -1, -1, -1, 258, 264, -1, -1, 261, -1, -1, -1, 258, 264, -1, -1, 261, 274, -1, -1, 258, 256, -1, -1, 262, 274, -1, -1, 269, 256, -1, -1, 271, 274, -1, -1, 258, 261, -1, -1, 271, 260, -1, -1, 268, 264, -1, -1, 267, 259, -1, -1, 269, 260, -1, -1, 267, 274, -1, -1, 273, -1, -1, -1, 262, 256, -1, -1, 269, 257, -1, -1, 267, 262, -1, -1, 269, 277, -1, -1, 267, 273, -1, -1, 518, -1, -1, -1, 268, 274, -1, -1, 517, -1, -1, -1, 517, -1, -1, -1, 268, 274, -1, -1, 273, 256, -1, -1, 262, 274, 259, -1, 529, -1, -1, -1, 518, 274, -1, -1, 270, 265, -1, -1, 271, 269, -1, -1, 271, 269, -1, -1, 270, 265, -1, -1, 270, 259, -1, -1, 266, 272, -1, -1, 266, 272, -1, -1, 270, 259, -1, -1, 278, -1, -1, -1, 279, 264, -1, -1, 279, 264, -1, -1, 278, -1, -1, -1, 278, -1, -1, -1, 258, 259, -1, -1, 258, 259, -1, -1, 278, -1, -1, -1};
static const int p_wd_ptr[] = {0, 25, 36, 58, 101, 120, 141, 155, 168, 211, 235, 237, 263, 272, 296, 300, 307, 327, 331, 357, 370, 371, 382, 386, 395};
static const int p_wd_rx[] = {3, 4, 18, 19, 20, 21, 28, 29, 30, 31, 48, 49, 50, 62, 63, 64, 65, 66, 67, 68, 83, 89, 90, 95, 100, 4, 31, 32, 46, 49, 50, 62, 63, 68, 83, 95, 6, 10, 11, 16, 17, 37, 44, 56, 57, 61, 67, 73, 74, 78, 80, 85, 87, 88, 89, 91, 108, 109, 1, 2, 3, 4, 9, 11, 13, 14, 15, 22, 27, 30, 32, 41, 45, 46, 47, 48, 49, 50, 51, 52, 53, 54, 55, 56, 57, 58, 59, 60, 61, 64, 66, 67, 69, 73, 82, 93, 100, 104, 105, 108, 109, 2, 6, 23, 24, 25, 26, 29, 31, 35, 37, 39, 40, 44, 45, 70, 71, 75, 92, 93, 5, 7, 11, 12, 13, 18, 36, 37, 39, 57, 58, 68, 69, 85, 86, 87, 88, 89, 91, 98, 99, 14, 38, 40, 41, 42, 59, 66, 81, 90, 95, 96, 97, 100, 101, 6, 7, 42, 52, 54, 73, 74, 75, 76, 77, 78, 79, 80, 1, 2, 5, 7, 8, 12, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 31, 32, 33, 34, 35, 36, 37, 38, 39, 40, 41, 42, 43, 45, 51, 52, 56, 70, 73, 75, 87, 88, 92, 106, 107, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 22, 28, 47, 81, 102, 103, 104, 105, 8, 9, 33, 34, 35, 41, 53, 54, 55, 61, 65, 66, 70, 72, 75, 77, 81, 82, 83, 84, 85, 86, 93, 94, 96, 97, 15, 16, 43, 60, 61, 71, 92, 98, 99, 8, 13, 14, 18, 36, 38, 39, 40, 44, 51, 52, 58, 59, 64, 68, 80, 82, 86, 90, 93, 95, 96, 102, 103, 102, 103, 104, 105, 15, 43, 60, 91, 92, 102, 103, 27, 28, 32, 42, 45, 47, 48, 49, 50, 53, 54, 55, 57, 58, 59, 60, 74, 76, 104, 105, 94, 97, 100, 101, 0, 3, 17, 18, 19, 20, 21, 22, 29, 48, 62, 63, 65, 69, 73, 74, 81, 82, 89, 90, 91, 94, 98, 99, 100, 101, 5, 16, 33, 51, 53, 64, 69, 70, 71, 72, 76, 78, 79, 84, 9, 34, 35, 55, 65, 72, 77, 83, 85, 86, 96, 106, 107, 108, 109, 10, 12, 17, 56, 67, 79, 80, 106, 107};
static const double p_wd_coef[] = {-1.0, 1.0, 1.0, 1.0, 1.0, 1.0, -1.0, -1.0, -1.0, 1.0, -1.0, 1.0, 1.0, -2.0, -2.0, -1.0, -1.0, -1.0, -1.0, -1.0, 1.0, 1.0, 1.0, -1.0, -1.0, -1.0, -1.0, -1.0, 1.0, -1.0, -1.0, 1.0, 1.0, 1.0, -1.0, 1.0, 1.0, -1.0, 1.0, 1.0, -1.0, 1.0, -1.0, -1.0, 1.0, 1.0, -1.0, 1.0, 1.0, 0.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, -1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, -1.0, 2.0, 1.0, 1.0, -1.0, -2.0, -2.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, -1.0, 1.0, -1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, -1.0, -1.0, -1.0, 1.0, -1.0, -1.0, 1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0, -1.0, -1.0, 1.0, -1.0, 1.0, 1.0, 1.0, -1.0, 1.0, -1.0, -1.0, -1.0, 1.0, 2.0, -2.0, -1.0, 1.0, -1.0, -1.0, -1.0, -1.0, 1.0, 1.0, -1.0, -1.0, -1.0, 2.0, 1.0, 2.0, -1.0, -1.0, 1.0, -1.0, 1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, 1.0, 1.0, 1.0, 1.0, 1.0, -1.0, -1.0, -1.0, -1.0, -2.0, -2.0, -2.0, -2.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, -1.0, -2.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, 1.0, 1.0, 1.0, 1.0, 1.0, -1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0, -1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 2.0, 1.0, 2.0, -1.0, -1.0, -1.0, -2.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, 1.0, -1.0, -1.0, 1.0, 1.0, -1.0, 1.0, 1.0, -1.0, 1.0, -1.0, 1.0, -1.0, -1.0, 1.0, 1.0, 1.0, 1.0, -1.0, 1.0, 1.0, -1.0, 1.0, 1.0, -1.0, 1.0, 1.0, 1.0, 1.0, 1.0, -1.0, -1.0, 1.0, -1.0, 1.0, 1.0, -1.0, 1.0, -1.0, -1.0, 1.0, -1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 0.0, 1.0, -1.0, 1.0, -1.0, -1.0, -2.0, 1.0, 1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, 1.0, 1.0, 1.0, 1.0, 1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, 1.0, 1.0, 1.0, -1.0, 1.0, -1.0, -1.0, 1.0, -1.0, -1.0, -1.0, -2.0, -1.0, 1.0, 1.0, 1.0, 1.0, -1.0, 1.0, -1.0, -1.0, 1.0, -1.0, -1.0, 1.0, 1.0, 1.0, 1.0, -1.0, 1.0, -1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 0.0, -1.0, 1.0, -1.0};
static const int p_ms_ptr[] = {0, 24, 48, 72, 94, 95, 96, 117, 118, 119, 143, 166, 190};
static const int p_ms_sp[] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 17, 19, 20, 21, 22, 23, 18, 16, 0, 1, 2, 3, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 17, 18, 19, 20, 21, 22, 4, 23, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 17, 18, 19, 20, 21, 22, 23, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23};
static const double p_ms_eff[] = {1.0, 1.0, 1.75, 1.0, 2.4000000953674316, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 15.399999618530273, 1.0, 1.0, 1.0, 3.0, 2.0, 1.0, 3.5999999046325684, 1.0, 1.0, 1.5, 1.0, 2.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 6.0, 1.0, 1.0, 1.0, 3.0, 2.0, 1.0, 2.0, 1.0, 1.0, 1.5, 1.0, 2.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 6.0, 1.0, 6.0, 1.0, 3.0, 2.0, 1.0, 3.5, 1.0, 1.0, 0.75, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.5, 1.0, 1.0, 1.5, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 3.0, 2.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 0.7300000190734863, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 3.6500000953674316, 1.0, 1.0, 1.0, 3.0, 2.0, 1.0, 1.0, 1.0, 1.0, 1.5, 1.0, 2.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 3.0, 2.0, 1.0, 2.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0};
const flint_host_prog host_prog = {24, 110, 12, 395, 190, p_rx_kind, p_rx_col, p_rx_mset, p_rx_fac, p_wd_ptr, p_wd_rx, p_wd_coef, p_ms_ptr, p_ms_sp, p_ms_eff};
// element matrix for the equilibrium kernel: a(e, j) = atoms of element e in species j
struct Cea {
  static constexpr int NS = 24, NE = 3;
  __host__ __device__ static constexpr double a(int e, int j) {
    constexpr double t[NE][NS] = {{0.0, 0.0, 1.0, 0.0, 0.0, 1.0, 1.0, 1.0, 0.0, 0.0, 3.0, 1.0, 2.0, 1.0, 3.0, 2.0, 0.0, 1.0, 0.0, 1.0, 2.0, 1.0, 1.0, 1.0}, {1.0, 2.0, 0.0, 1.0, 2.0, 1.0, 3.0, 2.0, 1.0, 0.0, 2.0, 3.0, 2.0, 2.0, 3.0, 1.0, 2.0, 3.0, 0.0, 2.0, 6.0, 4.0, 1.0, 0.0}, {2.0, 2.0, 1.0, 1.0, 0.0, 1.0, 1.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 1.0, 2.0, 2.0, 0.0, 0.0, 0.0, 2.0, 2.0}};
    return t[e][j];
  }
};
static const double cea_A[] = {0.0, 0.0, 1.0, 0.0, 0.0, 1.0, 1.0, 1.0, 0.0, 0.0, 3.0, 1.0, 2.0, 1.0, 3.0, 2.0, 0.0, 1.0, 0.0, 1.0, 2.0, 1.0, 1.0, 1.0, 1.0, 2.0, 0.0, 1.0, 2.0, 1.0, 3.0, 2.0, 1.0, 0.0, 2.0, 3.0, 2.0, 2.0, 3.0, 1.0, 2.0, 3.0, 0.0, 2.0, 6.0, 4.0, 1.0, 0.0, 2.0, 2.0, 1.0, 1.0, 0.0, 1.0, 1.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 1.0, 2.0, 2.0, 0.0, 0.0, 0.0, 2.0, 2.0};
}  // namespace flint_gen_tsr_gp_24

#include "../flint_cea_kernel.h"
FLINT_REGISTER_MECH_CEA(tsr_gp_24, flint_gen_tsr_gp_24)
