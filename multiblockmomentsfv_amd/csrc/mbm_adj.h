/*
 * mbm_adj.h -- hand-written reverse-mode derivatives (vector-Jacobian
 * products) of the cell and face functions of mbm_phys.h, for the discrete
 * adjoint of the Heun step.
 *
 * Convention: for y = f(x), f_vjp(x, gy, gx) ADDS gy^T df/dx to gx. The
 * limiter, the HLL speed clipping and |.| are differentiated almost
 * everywhere (the derivative of the branch that is taken). The time step is
 * a frozen input: the CFL reductions are not differentiated.
 */
#ifndef MBM_ADJ_H
#define MBM_ADJ_H

#include "mbm_phys.h"

namespace mbm {

/* ---------------- primitive (U -> P) ---------------- */
MBM_DEV void primitive_vjp(const double *U, const double *gP, double *gU)
{
  const double ir = rcp(U[0]);
  const double v1 = U[1] * ir, v2 = U[2] * ir;
  const double t11 = U[4] * ir - v1 * v1;
  const double t12 = U[5] * ir - v1 * v2;
  const double t22 = U[6] * ir - v2 * v2;
  double g_t11 = gP[4], g_t12 = gP[5], g_t22 = gP[6], g_v1 = gP[1], g_v2 = gP[2];
  double g_ir = 0.0, gU0 = gP[0], gU3 = 0.0;
  /* q1 = U7 - U3 v1 - U0 (t11 v1 + t12 v2) */
  const double q1g = gP[7];
  gU[7] += q1g;
  gU3 -= v1 * q1g;
  g_v1 -= U[3] * q1g + U[0] * t11 * q1g;
  gU0 -= (t11 * v1 + t12 * v2) * q1g;
  g_t11 -= U[0] * v1 * q1g;
  g_t12 -= U[0] * v2 * q1g;
  g_v2 -= U[0] * t12 * q1g;
  /* q2 = U8 - U3 v2 - U0 (t12 v1 + t22 v2) */
  const double q2g = gP[8];
  gU[8] += q2g;
  gU3 -= v2 * q2g;
  g_v2 -= U[3] * q2g + U[0] * t22 * q2g;
  gU0 -= (t12 * v1 + t22 * v2) * q2g;
  g_t12 -= U[0] * v1 * q2g;
  g_v1 -= U[0] * t12 * q2g;
  g_t22 -= U[0] * v2 * q2g;
  /* theta = (2 U3 ir - vs)/3 */
  g_ir += (2.0 / 3.0) * U[3] * gP[3];
  gU3 += (2.0 / 3.0) * ir * gP[3];
  const double g_vs = -(1.0 / 3.0) * gP[3];
  /* T_ij = U_k ir - v_i v_j */
  gU[4] += ir * g_t11;
  g_ir += U[4] * g_t11;
  g_v1 -= 2.0 * v1 * g_t11;
  gU[5] += ir * g_t12;
  g_ir += U[5] * g_t12;
  g_v1 -= v2 * g_t12;
  g_v2 -= v1 * g_t12;
  gU[6] += ir * g_t22;
  g_ir += U[6] * g_t22;
  g_v2 -= 2.0 * v2 * g_t22;
  g_v1 += 2.0 * v1 * g_vs;
  g_v2 += 2.0 * v2 * g_vs;
  gU[1] += ir * g_v1;
  gU[2] += ir * g_v2;
  g_ir += U[1] * g_v1 + U[2] * g_v2;
  gU[0] += gU0 - ir * ir * g_ir;
  gU[3] += gU3;
}

/* ---------------- conservative (P -> U) ---------------- */
MBM_DEV void conservative_vjp(const double *P, const double *gU, double *gP)
{
  const double rho = P[0], v1 = P[1], v2 = P[2], th = P[3];
  const double t11 = P[4], t12 = P[5], t22 = P[6];
  const double vs = v1 * v1 + v2 * v2;
  const double E = rho * (0.5 * vs + 1.5 * th);
  double grho = gU[0] + v1 * gU[1] + v2 * gU[2] + (t11 + v1 * v1) * gU[4]
              + (t12 + v1 * v2) * gU[5] + (t22 + v2 * v2) * gU[6]
              + (t11 * v1 + t12 * v2) * gU[7] + (t12 * v1 + t22 * v2) * gU[8];
  double gv1 = rho * gU[1] + 2.0 * rho * v1 * gU[4] + rho * v2 * gU[5]
             + (E + rho * t11) * gU[7] + rho * t12 * gU[8];
  double gv2 = rho * gU[2] + rho * v1 * gU[5] + 2.0 * rho * v2 * gU[6]
             + rho * t12 * gU[7] + (E + rho * t22) * gU[8];
  const double gE = gU[3] + v1 * gU[7] + v2 * gU[8];
  grho += (0.5 * vs + 1.5 * th) * gE;
  gv1 += rho * v1 * gE;
  gv2 += rho * v2 * gE;
  gP[0] += grho;
  gP[1] += gv1;
  gP[2] += gv2;
  gP[3] += 1.5 * rho * gE;
  gP[4] += rho * gU[4] + rho * v1 * gU[7];
  gP[5] += rho * gU[5] + rho * v2 * gU[7] + rho * v1 * gU[8];
  gP[6] += rho * gU[6] + rho * v2 * gU[8];
  gP[7] += gU[7];
  gP[8] += gU[8];
}

/* ---------------- relaxation of T and q ---------------- */
MBM_DEV void relax_vjp(const double *P, Damp d, const double *gR, double *gP, double &gd1,
                       double &gd2)
{
  gP[0] += gR[0];
  gP[1] += gR[1];
  gP[2] += gR[2];
  gP[3] += gR[3] + (1.0 - d.d1) * (gR[4] + gR[6]);
  gP[4] += d.d1 * gR[4];
  gP[5] += d.d1 * gR[5];
  gP[6] += d.d1 * gR[6];
  gP[7] += d.d2 * gR[7];
  gP[8] += d.d2 * gR[8];
  gd1 += (P[4] - P[3]) * gR[4] + P[5] * gR[5] + (P[6] - P[3]) * gR[6];
  gd2 += P[7] * gR[7] + P[8] * gR[8];
}

/* derivatives of the damping factors with respect to z (2z/3 in d2) */
MBM_DEV Damp damp0_dz(double z, double i1, double i2)
{
  /* f(y) = (1 + y/2 - y^2/4)/(1+y/2)^2, f'(y) = ((1/2 - y/2)(1+y/2) - (1+y/2-y^2/4))/(1+y/2)^3 */
  const double y = (2.0 / 3.0) * z;
  Damp d;
  d.d1 = ((0.5 - 0.5 * z) * (1.0 + 0.5 * z) - (1.0 + 0.5 * z - 0.25 * z * z)) * (i1 * i1 * i1);
  d.d2 = (2.0 / 3.0) * ((0.5 - 0.5 * y) * (1.0 + 0.5 * y) - (1.0 + 0.5 * y - 0.25 * y * y)) * (i2 * i2 * i2);
  return d;
}

MBM_DEV Damp damp1_dz(double z, double i1, double i2)
{
  /* f(y) = (1+y)/(1+y/2)^2, f'(y) = -(y/2)/(1+y/2)^3 */
  const double y = (2.0 / 3.0) * z;
  Damp d;
  d.d1 = -0.5 * z * (i1 * i1 * i1);
  d.d2 = (2.0 / 3.0) * (-0.5 * y) * (i2 * i2 * i2);
  return d;
}

MBM_DEV Damp damp3_dz(double i1, double i2)
{
  /* f(y) = 1/(1+y/2), f'(y) = -1/2 /(1+y/2)^2 ; i2 = 1/(1+z/3) */
  Damp d;
  d.d1 = -0.5 * i1 * i1;
  d.d2 = -(1.0 / 3.0) * i2 * i2;
  return d;
}

/* z = rho sqrt(theta) dt/Kn: adds to gQ and to gkn (d/dKn = -z/Kn) */
MBM_DEV void z_vjp(const double *Q, double dtkn, double kn, double gz, double *gQ, double &gkn)
{
  const double sq = sqrtp(Q[3]);
  gQ[0] += sq * dtkn * gz;
  gQ[3] += Q[0] * dtkn * 0.5 / sq * gz;
  gkn -= Q[0] * sq * dtkn / kn * gz;
}

/* ---------------- slope limiters ---------------- */
template <int LIM>
MBM_DEV void limit_vjp(double d1, double d2, double gs, double &gd1, double &gd2)
{
  if constexpr (LIM == 0) {
  } else if constexpr (LIM == 1) {
    gd1 += 0.5 * gs;
    gd2 += 0.5 * gs;
  } else if constexpr (LIM == 2) {
    const bool same = (d1 > 0.0 && d2 > 0.0) || (d1 < 0.0 && d2 < 0.0);
    if (same) {
      if (fabs(d1) <= fabs(d2)) gd1 += gs;
      else gd2 += gs;
    }
  } else if constexpr (LIM == 3) {
    const bool same = (d1 > 0.0 && d2 > 0.0) || (d1 < 0.0 && d2 < 0.0);
    if (same) {
      const double a1 = fabs(d1), a2 = fabs(d2);
      if (fmin(2.0 * a2, a1) >= fmin(a2, 2.0 * a1)) {
        if (2.0 * a2 <= a1) gd2 += 2.0 * gs;
        else gd1 += gs;
      } else {
        if (a2 <= 2.0 * a1) gd2 += gs;
        else gd1 += 2.0 * gs;
      }
    }
  } else {
    const double w1 = fabs(d1) + 1.0e-10, w2 = fabs(d2) + 1.0e-10;
    const double W = w1 + w2, iW = rcp(W), iW2 = iW * iW;
    const double s1 = d1 > 0.0 ? 1.0 : (d1 < 0.0 ? -1.0 : 0.0);
    const double s2 = d2 > 0.0 ? 1.0 : (d2 < 0.0 ? -1.0 : 0.0);
    gd1 += w2 * (W + s1 * (d2 - d1)) * iW2 * gs;
    gd2 += w1 * (W + s2 * (d1 - d2)) * iW2 * gs;
  }
}

/* three limiter VJPs at once; for the harmonic limiter the three
 * denominators share one reciprocal (as limit3 in the forward) */
template <int LIM>
MBM_DEV void limit_vjp3(const double *d1, const double *d2, const double *gs, double *gd1,
                        double *gd2)
{
  if constexpr (LIM == 4) {
    double w1[3], w2[3], W[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      w1[k] = fabs(d1[k]) + 1.0e-10;
      w2[k] = fabs(d2[k]) + 1.0e-10;
      W[k] = w1[k] + w2[k];
    }
    const double p01 = W[0] * W[1];
    const double inv = rcp(p01 * W[2]);
    const double i01 = inv * W[2];
    const double r[3] = { i01 * W[1], i01 * W[0], inv * p01 };
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const double iW2g = r[k] * r[k] * gs[k];
      const double s1 = d1[k] > 0.0 ? 1.0 : (d1[k] < 0.0 ? -1.0 : 0.0);
      const double s2 = d2[k] > 0.0 ? 1.0 : (d2[k] < 0.0 ? -1.0 : 0.0);
      gd1[k] += w2[k] * (W[k] + s1 * (d2[k] - d1[k])) * iW2g;
      gd2[k] += w1[k] * (W[k] + s2 * (d1[k] - d2[k])) * iW2g;
    }
  } else {
#pragma unroll
    for (int k = 0; k < 3; k++) limit_vjp<LIM>(d1[k], d2[k], gs[k], gd1[k], gd2[k]);
  }
}

/* ---------------- (U, F) of one face state ---------------- */
/* gW += gu^T dU/dW + gf^T dF/dW for the frame-local state W */
MBM_DEV void uf_vjp(const double *W, const double *gu, const double *gf, double *gW)
{
  const double rho = W[0], un = W[1], ut = W[2], th = W[3];
  const double tnn = W[4], tnt = W[5], ttt = W[6], qn = W[7], qt = W[8];
  const double un2 = un * un, ut2 = ut * ut, vs = un2 + ut2, unut = un * ut;
  const double E = rho * (0.5 * vs + 1.5 * th);
  const double An = tnn * un + tnt * ut, At = tnt * un + ttt * ut;
  /* combined adjoints of the shared expressions: F0=U1, F1=U4, F2=U5, F3=U7 */
  const double G1 = gu[1] + gf[0], G4 = gu[4] + gf[1], G5 = gu[5] + gf[2], G7 = gu[7] + gf[3];
  const double G2 = gu[2], G6 = gu[6], G8 = gu[8];
  double grho = gu[0], gun = 0.0, gut = 0.0, gth = 0.0, gtnn = 0.0, gtnt = 0.0, gttt = 0.0;
  double gqn = 0.0, gqt = 0.0, gE = gu[3], gAn = 0.0, gAt = 0.0, gvs = 0.0, gunut = 0.0;
  double gun2 = 0.0;
  /* rho un, rho ut */
  grho += un * G1 + ut * G2;
  gun += rho * G1;
  gut += rho * G2;
  /* rho (tnn + un2), rho (tnt + unut), rho (ttt + ut2) */
  grho += (tnn + un2) * G4 + (tnt + unut) * G5 + (ttt + ut2) * G6;
  gtnn += rho * G4;
  gun += 2.0 * rho * un * G4 + rho * ut * G5;
  gtnt += rho * G5;
  gut += rho * un * G5 + 2.0 * rho * ut * G6;
  gttt += rho * G6;
  /* qn + E un + rho An, qt + E ut + rho At */
  gqn += G7;
  gE += un * G7 + ut * G8;
  gun += E * G7;
  grho += An * G7 + At * G8;
  gAn += rho * G7;
  gqt += G8;
  gut += E * G8;
  gAt += rho * G8;
  /* f4 = 1.2 qn + rho un (3 tnn + un2) */
  const double g4 = gf[4];
  gqn += 1.2 * g4;
  grho += un * (3.0 * tnn + un2) * g4;
  gun += rho * (3.0 * tnn + 3.0 * un2) * g4;
  gtnn += 3.0 * rho * un * g4;
  /* f5 = 0.4 qt + rho (tnn ut + 2 tnt un + un2 ut) */
  const double g5 = gf[5];
  gqt += 0.4 * g5;
  grho += (tnn * ut + 2.0 * tnt * un + un2 * ut) * g5;
  gtnn += rho * ut * g5;
  gut += rho * (tnn + un2) * g5;
  gtnt += 2.0 * rho * un * g5;
  gun += rho * (2.0 * tnt + 2.0 * unut) * g5;
  /* f6 = 0.4 qn + rho (ttt un + 2 tnt ut + ut2 un) */
  const double g6 = gf[6];
  gqn += 0.4 * g6;
  grho += (ttt * un + 2.0 * tnt * ut + ut2 * un) * g6;
  gttt += rho * un * g6;
  gun += rho * (ttt + ut2) * g6;
  gtnt += 2.0 * rho * ut * g6;
  gut += rho * (2.0 * tnt + 2.0 * unut) * g6;
  /* f7 = 2(qn + rho An) un + E un2 + 1.2 qn un + 0.4 qt ut
   *      + rho/2 tnn vs + 3.5 rho th tnn - rho th^2 */
  const double g7 = gf[7];
  gqn += 3.2 * un * g7;
  gAn += 2.0 * rho * un * g7;
  grho += (2.0 * An * un + 0.5 * tnn * vs + 3.5 * th * tnn - th * th) * g7;
  gun += (2.0 * (qn + rho * An) + 1.2 * qn) * g7;
  gE += un2 * g7;
  gun2 += E * g7;
  gqt += 0.4 * ut * g7;
  gut += 0.4 * qt * g7;
  gtnn += (0.5 * rho * vs + 3.5 * rho * th) * g7;
  gvs += 0.5 * rho * tnn * g7;
  gth += (3.5 * rho * tnn - 2.0 * rho * th) * g7;
  /* f8 = 1.4 qt un + 1.4 qn ut + rho unut (tnn + ttt) + 1.5 rho tnt vs + E unut
   *      + 3.5 rho th tnt */
  const double g8 = gf[8];
  gqt += 1.4 * un * g8;
  gun += 1.4 * qt * g8;
  gqn += 1.4 * ut * g8;
  gut += 1.4 * qn * g8;
  grho += (unut * (tnn + ttt) + 1.5 * tnt * vs + 3.5 * th * tnt) * g8;
  gunut += (rho * (tnn + ttt) + E) * g8;
  gtnn += rho * unut * g8;
  gttt += rho * unut * g8;
  gtnt += (1.5 * rho * vs + 3.5 * rho * th) * g8;
  gvs += 1.5 * rho * tnt * g8;
  gE += unut * g8;
  gth += 3.5 * rho * tnt * g8;
  /* An, At */
  gtnn += un * gAn;
  gun += tnn * gAn + tnt * gAt;
  gtnt += ut * gAn + un * gAt;
  gut += tnt * gAn + ttt * gAt;
  gttt += ut * gAt;
  /* E = rho (vs/2 + 1.5 th) */
  grho += (0.5 * vs + 1.5 * th) * gE;
  gvs += 0.5 * rho * gE;
  gth += 1.5 * rho * gE;
  /* vs = un2 + ut2, unut */
  gun += 2.0 * un * (gvs + gun2) + ut * gunut;
  gut += 2.0 * ut * gvs + un * gunut;
  gW[0] += grho;
  gW[1] += gun;
  gW[2] += gut;
  gW[3] += gth;
  gW[4] += gtnn;
  gW[5] += gtnt;
  gW[6] += gttt;
  gW[7] += gqn;
  gW[8] += gqt;
}

/* ---------------- HLL face (frame-local) ---------------- */
/* g . U(W) */
MBM_DEV double dot_u(const double *W, const double *g)
{
  const Side sd = side_of(W);
  double u[9], f;
  uf<0>(W, sd, u[0], f); uf<1>(W, sd, u[1], f); uf<2>(W, sd, u[2], f);
  uf<3>(W, sd, u[3], f); uf<4>(W, sd, u[4], f); uf<5>(W, sd, u[5], f);
  uf<6>(W, sd, u[6], f); uf<7>(W, sd, u[7], f); uf<8>(W, sd, u[8], f);
  double acc = 0.0;
  for (int n = 0; n < 9; n++) acc += g[n] * u[n];
  return acc;
}

/* Fs = ca F(WL) + cb F(WR) + cc (U(WR) - U(WL)) with cc = lamL lamR s, so the
 * WL part is ca (F(WL) - pL U(WL)) and the WR part cb (F(WR) - pR U(WR))
 * (pL = lamL, pR = lamR; both 0 on the alph <= 0 branch, where cc = 0). The
 * speeds need g.F(W) and g.U(W): U and F are linear in (rho, qn, qt) at fixed
 * velocity, theta and T, so F(W) = J_F(W) e and U(W) = J_U(W) e with
 * e = (rho, 0, ..., 0, qn, qt), and n = J_U^T(-p g) + J_F^T g gives
 * n . e = g.F(W) - p g.U(W) without evaluating F. */
MBM_DEV void hll_vjp(const double *WL, const double *WR, double dtdx, const double *gF,
                     double *gWL, double *gWR)
{
  const double k = 20.0 / 3.0;
  const double cL = sqrtp(k * WL[3]), cR = sqrtp(k * WR[3]);
  const double lamL = fmin(WL[1] - cL, 0.0);
  const double lamR = fmax(WR[1] + cR, 0.0);
  const double alph = lamR - lamL;
  double ca, cb, s = 0.0, pL = 0.0, pR = 0.0;
  if (alph > 0.0) {
    s = dtdx * rcp(alph);
    ca = lamR * s;
    cb = -lamL * s;
    pL = lamL;
    pR = lamR;
  } else {
    ca = 0.5 * dtdx;
    cb = 0.5 * dtdx;
  }
  double mL[9], mR[9], nL[9] = { 0 }, nR[9] = { 0 };
  for (int n = 0; n < 9; n++) {
    mL[n] = -pL * gF[n];
    mR[n] = -pR * gF[n];
  }
  uf_vjp(WL, mL, gF, nL);
  uf_vjp(WR, mR, gF, nR);
  if (alph > 0.0) {
    const double guL = dot_u(WL, gF), guR = dot_u(WR, gF);
    const double gca = nL[0] * WL[0] + nL[7] * WL[7] + nL[8] * WL[8] + pL * guL;  /* g.F(WL) */
    const double gcb = nR[0] * WR[0] + nR[7] * WR[7] + nR[8] * WR[8] + pR * guR;  /* g.F(WR) */
    const double gcc = guR - guL;
    double glamR = s * gca + lamL * s * gcc;
    double glamL = -s * gcb + lamR * s * gcc;
    const double gs = lamR * gca - lamL * gcb + lamL * lamR * gcc;
    const double galph = -s * rcp(alph) * gs;
    glamR += galph;
    glamL -= galph;
    /* dc/dtheta = k / (2c) */
    if (WL[1] - cL < 0.0) {
      gWL[1] += glamL;
      gWL[3] -= glamL * k * 0.5 * rcp(cL);
    }
    if (WR[1] + cR > 0.0) {
      gWR[1] += glamR;
      gWR[3] += glamR * k * 0.5 * rcp(cR);
    }
  }
  for (int n = 0; n < 9; n++) {
    gWL[n] += ca * nL[n];
    gWR[n] += cb * nR[n];
  }
}

template <int DIR>
MBM_DEV void face_vjp(const double *L, const double *R, double dtdx, const double *gFs,
                      double *gL, double *gR)
{
  double WL[9], WR[9], gF[9], gWL[9] = {0}, gWR[9] = {0};
  for (int n = 0; n < 9; n++) {
    WL[n] = L[perm<DIR>(n)];
    WR[n] = R[perm<DIR>(n)];
    gF[n] = gFs[perm<DIR>(n)];
  }
  hll_vjp(WL, WR, dtdx, gF, gWL, gWR);
  for (int n = 0; n < 9; n++) {
    gL[perm<DIR>(n)] += gWL[n];
    gR[perm<DIR>(n)] += gWR[n];
  }
}

} // namespace mbm

#endif
