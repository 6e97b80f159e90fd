/*
 * mbm_adj_stage.h -- per-cell bodies of the stage adjoint (host+device, so
 * the same code runs in the GPU kernels of mbm_adjoint.hip and in CPU tests).
 * BLK is any type with off, plane, lx, ly, pitch, dx, dy.
 *
 * adj_prim_cell : stage primitives P of cell (a,b) (Main.c:314-336);
 * adj_face_cell : VJP of the x-face a+1/2 and y-face b+1/2 (reconstruction +
 *                 HLL flux) w.r.t. their two face states, for the faces that
 *                 update an interior cell; zeros elsewhere;
 * adj_cell_cell : gather of the face-state adjoints of cell (a,b), VJP of the
 *                 slopes that use P(a,b), then of the projection and the
 *                 initial value; returns the cell's dJ/dKn term.
 */
#ifndef MBM_ADJ_STAGE_H
#define MBM_ADJ_STAGE_H

#include "mbm_adj.h"

namespace mbm {

MBM_DEV bool adj_interior(int a, int b, int lx, int ly)
{
  return a >= 3 && a <= lx - 4 && b >= 3 && b <= ly - 4;
}

MBM_DEV void adj_ld(const double *base, long long plane, long long off, double *v)
{
  for (int m = 0; m < 9; m++) v[m] = base[off + m * plane];
}

MBM_DEV void adj_st(double *base, long long plane, long long off, const double *v)
{
  for (int m = 0; m < 9; m++) base[off + m * plane] = v[m];
}

template <int STAGE, class BLK>
MBM_DEV void adj_prim_cell(const double *in, double *P, const BLK &B, int a, int b, double dt,
                           double dtkn, double force)
{
  const long long off = B.off + (long long)b * B.pitch + a;
  double U[9], Q[9], I[9];
  double cp;
  adj_ld(in, B.plane, off, U);
  const bool ring3 = a == 0 || a == B.lx - 1 || b == 0 || b == B.ly - 1;
  project<STAGE, false>(U, U, ring3, dt, dtkn, force, Q, I, cp);
  adj_st(P, B.plane, off, Q);
}

template <int LIM, class BLK>
MBM_DEV void adj_face_cell(const double *P, const double *gout, double *FxL, double *FxR,
                           double *FyL, double *FyR, const BLK &B, int a, int b, double dt,
                           double scale)
{
  const int lx = B.lx, ly = B.ly;
  const long long pl = B.plane;
  const long long off = B.off + (long long)b * B.pitch + a;
  double gL[9], gR[9];
  for (int d = 0; d < 2; d++) {
    for (int n = 0; n < 9; n++) { gL[n] = 0.0; gR[n] = 0.0; }
    const bool in_set = d == 0 ? (a >= 2 && a <= lx - 4 && b >= 3 && b <= ly - 1 - 3)
                               : (b >= 2 && b <= ly - 4 && a >= 3 && a <= lx - 1 - 3);
    if (in_set) {
      const long long step = d == 0 ? 1 : B.pitch;
      double Pm[9], P0[9], P1[9], P2[9], WL[9], WR[9], gF[9];
      adj_ld(P, pl, off - step, Pm);
      adj_ld(P, pl, off, P0);
      adj_ld(P, pl, off + step, P1);
      adj_ld(P, pl, off + 2 * step, P2);
      for (int n = 0; n < 9; n++) {
        const double s0 = limit<LIM>(P0[n] - Pm[n], P1[n] - P0[n]);
        const double s1 = limit<LIM>(P1[n] - P0[n], P2[n] - P1[n]);
        WL[n] = P0[n] + 0.5 * s0;
        WR[n] = P1[n] - 0.5 * s1;
      }
      const int a1 = d == 0 ? a + 1 : a, b1 = d == 0 ? b : b + 1;
      const double w0 = adj_interior(a, b, lx, ly) ? scale : 0.0;
      const double w1 = adj_interior(a1, b1, lx, ly) ? scale : 0.0;
      double g0[9], g1[9];
      adj_ld(gout, pl, off, g0);
      adj_ld(gout, pl, off + step, g1);
      for (int n = 0; n < 9; n++) gF[n] = w1 * g1[n] - w0 * g0[n];
      if (d == 0) face_vjp<0>(WL, WR, dt / B.dx, gF, gL, gR);
      else face_vjp<1>(WL, WR, dt / B.dy, gF, gL, gR);
    }
    if (d == 0) { adj_st(FxL, pl, off, gL); adj_st(FxR, pl, off, gR); }
    else { adj_st(FyL, pl, off, gL); adj_st(FyR, pl, off, gR); }
  }
}

/* VJP of the projection of one cell (Main.c:314-336), given Q = prim(U), z and
 * damp_inv(z): gP is the adjoint of its stage primitives, gI (may be null)
 * that of its initial value. Adds to
 * gU (stage input) and, in stage 2 with DATA0, gU0 (DATA); returns the dJ/dKn
 * term. Stage 2's initial value is cons(relax(prim(U), damp3)) + the DATA
 * part cons(relax(prim(U0), damp0)); DATA0 = false leaves out the VJP of the
 * DATA part (its adjoint is gI itself; see adj_x). */
template <int STAGE, bool DATA0 = true>
MBM_DEV double proj_vjp_q(const double *U, const double *Q, double z, double i1, double i2,
                          const double *U0, const double *gP, const double *gI, double dt,
                          double kn, double force, double *gU, double *gU0)
{
  double gkn = 0.0;
  const double dtkn = dt / kn;
  double gQ[9] = { 0 }, gz = 0.0, g1 = 0.0, g2 = 0.0;
  if (STAGE == 1) {
    relax_vjp(Q, damp1(z, i1, i2), gP, gQ, g1, g2);
    Damp dz = damp1_dz(z, i1, i2);
    gz += dz.d1 * g1 + dz.d2 * g2;
    if (gI) {
      const Damp d0 = damp0(z, i1, i2);
      double R[9], gR[9] = { 0 };
      relax(Q, d0, R);
      R[1] += dt * force;
      conservative_vjp(R, gI, gR);
      g1 = 0.0;
      g2 = 0.0;
      relax_vjp(Q, d0, gR, gQ, g1, g2);
      dz = damp0_dz(z, i1, i2);
      gz += dz.d1 * g1 + dz.d2 * g2;
    }
  } else {
    Damp d3;
    d3.d1 = i1;
    d3.d2 = i2;
    relax_vjp(Q, d3, gP, gQ, g1, g2);
    const Damp dz = damp3_dz(i1, i2);
    gz += dz.d1 * g1 + dz.d2 * g2;
    if (gI) {
      double Qf[9];
      for (int n = 0; n < 9; n++) Qf[n] = Q[n];
      Qf[1] += dt * force;
      conservative_vjp(Qf, gI, gQ);
    }
    if (gI && DATA0) {
      /* Data0 part: cons(relax(prim(U0), damp0(z0))) */
      double Q0[9], R0[9], gR0[9] = { 0 }, gQ0[9] = { 0 };
      primitive(U0, Q0);
      const double z0 = Q0[0] * sqrtp(Q0[3]) * dtkn;
      double j1, j2;
      damp_inv(z0, j1, j2);
      const Damp e0 = damp0(z0, j1, j2);
      relax(Q0, e0, R0);
      conservative_vjp(R0, gI, gR0);
      double h1 = 0.0, h2 = 0.0;
      relax_vjp(Q0, e0, gR0, gQ0, h1, h2);
      const Damp ez = damp0_dz(z0, j1, j2);
      z_vjp(Q0, dtkn, kn, ez.d1 * h1 + ez.d2 * h2, gQ0, gkn);
      primitive_vjp(U0, gQ0, gU0);
    }
  }
  z_vjp(Q, dtkn, kn, gz, gQ, gkn);
  primitive_vjp(U, gQ, gU);
  return gkn;
}

/* the same from U alone: Q = prim(U), z = rho sqrt(theta) dt/Kn, (i1, i2) =
 * damp_inv(z), as the projection computes them */
template <int STAGE, bool DATA0 = true>
MBM_DEV double proj_vjp(const double *U, const double *U0, const double *gP, const double *gI,
                        double dt, double kn, double force, double *gU, double *gU0)
{
  double Q[9];
  primitive(U, Q);
  const double z = Q[0] * sqrtp(Q[3]) * (dt / kn);
  double i1, i2;
  damp_inv(z, i1, i2);
  return proj_vjp_q<STAGE, DATA0>(U, Q, z, i1, i2, U0, gP, gI, dt, kn, force, gU, gU0);
}

template <int STAGE, int LIM, class BLK>
MBM_DEV double adj_cell_cell(const double *in, const double *in0, const double *P,
                             const double *FxL, const double *FxR, const double *FyL,
                             const double *FyR, const double *gout, double *gin, double *gin0,
                             const BLK &B, int a, int b, double dt, double kn, double force,
                             double scale)
{
  double gkn = 0.0;
  {
    const int lx = B.lx, ly = B.ly;
    const long long pl = B.plane;
    const long long off = B.off + (long long)b * B.pitch + a;
    double gU[9] = { 0 }, gU0[9] = { 0 };
    if (a >= 1 && a <= lx - 2 && b >= 1 && b <= ly - 2) {
      double gP[9] = { 0 };
      for (int d = 0; d < 2; d++) {
        const long long step = d == 0 ? 1 : B.pitch;
        const int pos = d == 0 ? a : b, len = d == 0 ? lx : ly;
        const double *FL = d == 0 ? FxL : FyL;
        const double *FR = d == 0 ? FxR : FyR;
        double v[9], w[9];
        /* direct: left state of face pos+1/2, right state of face pos-1/2 */
        adj_ld(FL, pl, off, v);
        adj_ld(FR, pl, off - step, w);
        for (int n = 0; n < 9; n++) gP[n] += v[n] + w[n];
        /* slopes at e = pos-1, pos, pos+1 use P(pos) */
        for (int k = -1; k <= 1; k++) {
          const int e = pos + k;
          if (e < 2 || e > len - 3) continue;   /* slopes used by faces in the set */
          const long long oe = off + k * step;
          double fl[9], fr[9], pm[9], p0[9], pp[9];
          adj_ld(FL, pl, oe, fl);
          adj_ld(FR, pl, oe - step, fr);
          adj_ld(P, pl, oe - step, pm);
          adj_ld(P, pl, oe, p0);
          adj_ld(P, pl, oe + step, pp);
          for (int n = 0; n < 9; n++) {
            const double gs = 0.5 * fl[n] - 0.5 * fr[n];
            double gd1 = 0.0, gd2 = 0.0;
            limit_vjp<LIM>(p0[n] - pm[n], pp[n] - p0[n], gs, gd1, gd2);
            /* d1(e) = P(e) - P(e-1), d2(e) = P(e+1) - P(e) */
            if (k == 0) gP[n] += gd1 - gd2;
            else if (k == -1) gP[n] += gd2;
            else gP[n] -= gd1;
          }
        }
      }
      double gI[9], U[9], U0[9];
      adj_ld(gout, pl, off, gI);
      const double wI = adj_interior(a, b, lx, ly) ? scale : 0.0;
      for (int n = 0; n < 9; n++) gI[n] *= wI;
      adj_ld(in, pl, off, U);
      if (STAGE == 2) adj_ld(in0, pl, off, U0);
      gkn += proj_vjp<STAGE>(U, U0, gP, gI, dt, kn, force, gU, gU0);
    }
    if (STAGE == 1) {
      for (int m = 0; m < 9; m++) gin[off + m * pl] += gU[m];
    } else {
      for (int m = 0; m < 9; m++) {
        gin[off + m * pl] = gU[m];
        gin0[off + m * pl] += gU0[m];
      }
    }
  }
  return gkn;
}

} // namespace mbm

#endif
