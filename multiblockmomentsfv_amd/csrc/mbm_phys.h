/*
 * mbm_phys.h -- device-side cell and face physics of the explicit update,
 * FP64, written for throughput on gfx950's FP64 VALU.
 *
 * Same mathematics as the reference (Flux.c, Limiter.c, Scenario.c), in an
 * algebraically equivalent form with fewer divides:
 *  - one reciprocal of rho per primitive recovery (Flux.c:30-56);
 *  - the Grad-13 flux is evaluated from primitives without the
 *    Conservative -> Primitive round trip of Flux.c:79-119 (rho*M_ijk = c*q,
 *    so the flux needs no division at all);
 *  - the relaxation factors R(z), R(2z/3) of Flux.c:213-223 share their
 *    denominators (1+z/2)^2 and (1+z/3)^2, so both come from one reciprocal;
 *  - the HLL combine (Flux.c:364-376) uses one reciprocal per face and
 *    folds dt/dx into it;
 *  - the slope division by dx (Main.c:346-349) and the 0.5*dx factor of the
 *    reconstruction (Main.c:360-374) cancel and are dropped.
 * Results agree with the reference to rounding (<= 1e-12 relative L-inf per
 * field on the checked decks).
 *
 * Non-finite data: these fast forms give the reference's NaN pattern only on
 * finite data with a nonzero face density. Where a NaN is present, the
 * reference's exact expression structure decides which components turn NaN
 * (its limiters compare against NaN, sgn(NaN) = 1 and min(NaN, b) = b, its
 * flux divides by the density after Conservative() and rotates components
 * into the face frame, where NaN * 0 is NaN). The *_ref functions at the end
 * restate those expressions; the stage kernel switches a row to them when a
 * primitive in the row's stencil is not finite. A face whose reference
 * density is exactly zero has an all-NaN flux (0/0 in Flux.c:85-103); the
 * kernel decides that exact zero in the reference's arithmetic.
 */
#ifndef MBM_PHYS_H
#define MBM_PHYS_H

#include <math.h>

/* The functions are host+device so the same code can be unit-tested on the
 * CPU (the host build uses IEEE division for rcp). */
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MBM_DEV __host__ __device__ __forceinline__
#else
#define MBM_DEV static inline
#endif

namespace mbm {

/* reciprocal: hardware estimate r (5e-8 relative) and one cubic correction,
 * 1/x = r / (1 - e) = r (1 + e + e^2 + ...) with e = 1 - x r, whose
 * truncation (e^3, about 1e-22) is far below an ulp: correctly rounded on
 * 2^24 arguments (log-uniform in [1e-8, 1e8], and mantissas next to 1 and
 * 2), as two Newton steps are, in three FMAs instead of four; 0 and inf give
 * NaN */
MBM_DEV double rcp(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
  const double r = __builtin_amdgcn_rcp(x);
  const double e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(__builtin_fma(e, e, e), r, r);
#else
  return 1.0 / x;
#endif
}

/* square root of a positive number: hardware rsq estimate y (5e-8
 * relative), one Goldschmidt step s = x y (1 + r) (36 ulp at most) and one
 * Newton correction with h = y / 2 (the correction is about 36 ulp, so h's
 * own error is negligible in it): correctly rounded on the same 2^24
 * arguments (the compiler's own sequence refines h and corrects twice, and
 * rescales subnormal and huge arguments, which theta never reaches);
 * 0 -> 0, negative -> NaN */
MBM_DEV double sqrtp(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
  const double y = __builtin_amdgcn_rsq(x);
  const double s0 = x * y, h = 0.5 * y;
  const double s = __builtin_fma(s0, __builtin_fma(-h, s0, 0.5), s0);
  const double d = __builtin_fma(-s, s, x);
  return x > 0.0 ? __builtin_fma(d, h, s) : (x == 0.0 ? 0.0 : __builtin_nan(""));
#else
  return sqrt(x);
#endif
}

/* true when the relaxed components 4..8 of a projection are finite (their
 * sum is; an overflowing sum of finite values only errs towards the slower
 * reference-form path). A non-finite conservative component reaches them:
 * U[0..3] through theta and the damping factors, U[4..8] directly; so does
 * a NaN damping factor (negative temperature). */
MBM_DEV bool finite_relaxed(const double *P)
{
  const double s = (P[4] + P[5]) + (P[6] + P[7]) + P[8];
  return __builtin_isfinite(s);
}

/* primitive variables {rho, v1, v2, theta, T11, T12, T22, q1, q2} */
MBM_DEV void primitive(const double *U, double *P)
{
  const double ir = rcp(U[0]);
  const double v1 = U[1] * ir, v2 = U[2] * ir;
  const double vs = v1 * v1 + v2 * v2;
  const double t11 = U[4] * ir - v1 * v1;
  const double t12 = U[5] * ir - v1 * v2;
  const double t22 = U[6] * ir - v2 * v2;
  P[0] = U[0];
  P[1] = v1;
  P[2] = v2;
  P[3] = (2.0 * U[3] * ir - vs) * (1.0 / 3.0);
  P[4] = t11;
  P[5] = t12;
  P[6] = t22;
  P[7] = U[7] - U[3] * v1 - U[0] * (t11 * v1 + t12 * v2);
  P[8] = U[8] - U[3] * v2 - U[0] * (t12 * v1 + t22 * v2);
}

/* conservative variables from primitives, E recomputed (primDef.h) */
MBM_DEV void conservative(const double *P, double *U)
{
  const double rho = P[0], v1 = P[1], v2 = P[2];
  const double E = rho * (0.5 * (v1 * v1 + v2 * v2) + 1.5 * P[3]);
  U[0] = rho;
  U[1] = rho * v1;
  U[2] = rho * v2;
  U[3] = E;
  U[4] = rho * (P[4] + v1 * v1);
  U[5] = rho * (P[5] + v1 * v2);
  U[6] = rho * (P[6] + v2 * v2);
  U[7] = P[7] + E * v1 + rho * (P[4] * v1 + P[5] * v2);
  U[8] = P[8] + E * v2 + rho * (P[5] * v1 + P[6] * v2);
}

/* damping factors of the relaxation for z = nu*dt.
 * flag 0/-1: (1+z/2-z^2/4)/(1+z/2)^2, flag 1: (1+z)/(1+z/2)^2,
 * flag 3: 1/(1+z/2); the second factor uses 2z/3 (Flux.c:247-248). */
struct Damp { double d1, d2; };

/* returns i1 = 1/(1+z/2), i2 = 1/(1+z/3) with one reciprocal */
MBM_DEV void damp_inv(double z, double &i1, double &i2)
{
  const double a = 1.0 + 0.5 * z, b = 1.0 + (1.0 / 3.0) * z;
  const double r = rcp(a * b);
  i1 = b * r;
  i2 = a * r;
}

MBM_DEV Damp damp0(double z, double i1, double i2)
{
  const double z3 = (2.0 / 3.0) * z;
  Damp d;
  d.d1 = (1.0 + 0.5 * z - 0.25 * z * z) * (i1 * i1);
  d.d2 = (1.0 + 0.5 * z3 - 0.25 * z3 * z3) * (i2 * i2);
  return d;
}

MBM_DEV Damp damp1(double z, double i1, double i2)
{
  Damp d;
  d.d1 = (1.0 + z) * (i1 * i1);
  d.d2 = (1.0 + (2.0 / 3.0) * z) * (i2 * i2);
  return d;
}

/* relaxed primitives (Flux.c:249-253): T_ij -> theta + d1 (T_ij - theta),
 * T12 -> d1 T12, q -> d2 q */
MBM_DEV void relax(const double *P, Damp d, double *R)
{
  R[0] = P[0];
  R[1] = P[1];
  R[2] = P[2];
  R[3] = P[3];
  R[4] = __builtin_fma(d.d1, P[4] - P[3], P[3]);
  R[5] = d.d1 * P[5];
  R[6] = __builtin_fma(d.d1, P[6] - P[3], P[3]);
  R[7] = d.d2 * P[7];
  R[8] = d.d2 * P[8];
}

/* Conservative(relax(P, d)) with P = Primitive(U), taken from U itself
 * (relaxation without a body force): relaxing removes rho (1-d1)
 * (T_ij - theta delta_ij) from the second moments, and (1-d2) q plus the
 * work of that removed stress against v from the heat fluxes; density,
 * momentum and energy are unchanged. */
MBM_DEV void relax_cons(const double *U, const double *P, Damp d, double *V)
{
  const double rom = P[0] * (1.0 - d.d1), om2 = 1.0 - d.d2;
  const double a11 = P[4] - P[3], a22 = P[6] - P[3];
  V[0] = U[0];
  V[1] = U[1];
  V[2] = U[2];
  V[3] = U[3];
  V[4] = U[4] - rom * a11;
  V[5] = U[5] - rom * P[5];
  V[6] = U[6] - rom * a22;
  V[7] = U[7] - om2 * P[7] - rom * (a11 * P[1] + P[5] * P[2]);
  V[8] = U[8] - om2 * P[8] - rom * (P[5] * P[1] + a22 * P[2]);
}

/* ---------------- slope limiters (Limiter.c:26-59) ---------------- */
template <int LIM>
MBM_DEV double limit(double d1, double d2)
{
  if constexpr (LIM == 0) {
    return 0.0;
  } else if constexpr (LIM == 1) {
    return 0.5 * (d1 + d2);
  } else if constexpr (LIM == 2) {
    const double m = fmin(fabs(d1), fabs(d2));
    return (d1 > 0.0 && d2 > 0.0) ? m : ((d1 < 0.0 && d2 < 0.0) ? -m : 0.0);
  } else if constexpr (LIM == 3) {
    /* d1 * phi(d2/d1), phi = max(0, min(2r,1), min(r,2)) */
    const double a1 = fabs(d1), a2 = fabs(d2);
    const double m = fmax(fmin(2.0 * a2, a1), fmin(a2, 2.0 * a1));
    return (d1 > 0.0 && d2 > 0.0) ? m : ((d1 < 0.0 && d2 < 0.0) ? -m : 0.0);
  } else {
    const double w1 = fabs(d1) + 1.0e-10, w2 = fabs(d2) + 1.0e-10;
    return (w1 * d2 + w2 * d1) * rcp(w1 + w2);
  }
}

/* Three limited slopes at once. For the harmonic limiter the three
 * denominators share one reciprocal (r_k = prod_{j!=k} W_j / prod W_j). */
template <int LIM>
MBM_DEV void limit3(const double *d1, const double *d2, double *s)
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
    for (int k = 0; k < 3; k++) s[k] = (w1[k] * d2[k] + w2[k] * d1[k]) * r[k];
  } else {
#pragma unroll
    for (int k = 0; k < 3; k++) s[k] = limit<LIM>(d1[k], d2[k]);
  }
}

/* ---------------- Grad-13 flux (Flux.c:79-119) ----------------
 * W: primitives in the face frame {rho, un, ut, theta, Tnn, Tnt, Ttt, qn, qt}.
 * Side holds the products shared by the 9 components of the conservative
 * state U and the normal flux F, so the components can be produced one at a
 * time (keeps the register footprint of a face small). */
struct Side {
  double un2, ut2, vs, unut, E, An, At;
};

MBM_DEV Side side_of(const double *W)
{
  Side s;
  s.un2 = W[1] * W[1];
  s.ut2 = W[2] * W[2];
  s.vs = s.un2 + s.ut2;
  s.unut = W[1] * W[2];
  s.E = W[0] * (0.5 * s.vs + 1.5 * W[3]);
  s.An = W[4] * W[1] + W[5] * W[2];     /* (T.v)_n */
  s.At = W[5] * W[1] + W[6] * W[2];     /* (T.v)_t */
  return s;
}

/* component n of U and of the normal flux F */
template <int N>
MBM_DEV void uf(const double *W, const Side &s, double &u, double &f)
{
  const double rho = W[0], un = W[1], ut = W[2], th = W[3];
  const double tnn = W[4], tnt = W[5], ttt = W[6], qn = W[7], qt = W[8];
  if constexpr (N == 0) { u = rho; f = rho * un; }
  else if constexpr (N == 1) { u = rho * un; f = rho * (tnn + s.un2); }
  else if constexpr (N == 2) { u = rho * ut; f = rho * (tnt + s.unut); }
  else if constexpr (N == 3) { u = s.E; f = qn + s.E * un + rho * s.An; }
  else if constexpr (N == 4) { u = rho * (tnn + s.un2); f = 1.2 * qn + rho * un * (3.0 * tnn + s.un2); }
  else if constexpr (N == 5) { u = rho * (tnt + s.unut); f = 0.4 * qt + rho * (tnn * ut + 2.0 * tnt * un + s.un2 * ut); }
  else if constexpr (N == 6) { u = rho * (ttt + s.ut2); f = 0.4 * qn + rho * (ttt * un + 2.0 * tnt * ut + s.ut2 * un); }
  else if constexpr (N == 7) {
    u = qn + s.E * un + rho * s.An;
    /* R11 = 5 th^2 + 7 th (Tnn - th) */
    const double r11 = th * (5.0 * th + 7.0 * (tnn - th));
    f = 2.0 * (qn + rho * s.An) * un + s.E * s.un2 + (1.2 * qn * un + 0.4 * qt * ut)
      + 0.5 * rho * (tnn * s.vs + r11);
  } else {
    u = qt + s.E * ut + rho * s.At;
    /* R12 = 7 th Tnt */
    f = qt * un + qn * ut + rho * (s.unut * (tnn + ttt) + 1.5 * tnt * s.vs) + s.E * s.unut
      + 0.4 * (qt * un + qn * ut) + 3.5 * rho * th * tnt;
  }
}

/* full U and F (used where all components are needed at once) */
MBM_DEV void cons_flux(const double *W, double *U, double *F)
{
  const Side s = side_of(W);
  uf<0>(W, s, U[0], F[0]); uf<1>(W, s, U[1], F[1]); uf<2>(W, s, U[2], F[2]);
  uf<3>(W, s, U[3], F[3]); uf<4>(W, s, U[4], F[4]); uf<5>(W, s, U[5], F[5]);
  uf<6>(W, s, U[6], F[6]); uf<7>(W, s, U[7], F[7]); uf<8>(W, s, U[8], F[8]);
}

/* component n of the shifted flux g = F - lam U = w U + p of one side, with
 * w = un - lam and p = F - un U, the part of the Grad-13 flux that is not
 * carried by the normal velocity (uf above, regrouped):
 *   p = (0, rho Tnn, rho Tnt, qn + rho (T.v)_n, 1.2 qn + 2 un rho Tnn,
 *        0.4 qt + rho (Tnn ut + Tnt un), 0.4 qn + 2 ut rho Tnt,
 *        un (2.2 qn + rho (T.v)_n) + 0.4 qt ut + rho (Tnn |v|^2 + R11) / 2,
 *        1.4 qn ut + 0.4 qt un + rho (un ut Tnn + Tnt (un^2/2 + 3 ut^2/2 + 7 th/2)))
 * with R11 = th (7 Tnn - 2 th). */
template <int N>
MBM_DEV double gshift(const double *W, const Side &s, double w)
{
  const double rho = W[0], un = W[1], ut = W[2], th = W[3];
  const double tnn = W[4], tnt = W[5], ttt = W[6], qn = W[7], qt = W[8];
  if constexpr (N == 0) {
    return w * rho;
  } else if constexpr (N == 1) {
    return w * (rho * un) + rho * tnn;
  } else if constexpr (N == 2) {
    return w * (rho * ut) + rho * tnt;
  } else if constexpr (N == 3) {
    return w * s.E + (qn + rho * s.An);
  } else if constexpr (N == 4) {
    const double rt = rho * tnn;
    return w * (rt + rho * s.un2) + (1.2 * qn + 2.0 * un * rt);
  } else if constexpr (N == 5) {
    return w * (rho * (tnt + s.unut)) + (0.4 * qt + rho * (tnn * ut + tnt * un));
  } else if constexpr (N == 6) {
    return w * (rho * (ttt + s.ut2)) + (0.4 * qn + 2.0 * ut * (rho * tnt));
  } else if constexpr (N == 7) {
    const double u = qn + s.E * un + rho * s.An;
    const double p = un * (2.2 * qn + rho * s.An) + 0.4 * qt * ut
                   + 0.5 * rho * (tnn * s.vs + th * (7.0 * tnn - 2.0 * th));
    return w * u + p;
  } else {
    const double u = qt + s.E * ut + rho * s.At;
    const double p = 1.4 * qn * ut + 0.4 * qt * un
                   + rho * (s.unut * tnn + tnt * (0.5 * s.un2 + 1.5 * s.ut2 + 3.5 * th));
    return w * u + p;
  }
}

/* HLL face flux (Flux.c:332-379), frame-local. Fs = (dt/dx) * flux.
 * smax = |(un_L+un_R)/2| + max(c_L, c_R), c = 2 sqrt(5 theta / 3).
 * With a = lamR - lamL > 0 the reference's
 *   (lamR F_L - lamL F_R + lamL lamR (U_R - U_L)) / a
 * equals (lamR g_L - lamL g_R) / a for the shifted fluxes g = F - lam U of
 * each side (gshift): two products per component instead of four. With
 * a <= 0 (lamL = lamR = 0) both sides' g are F and the flux is their mean. */
MBM_DEV void hll(const double *WL, const double *WR, double dtdx, double *Fs,
                 double &smax)
{
  const double cL = sqrtp((20.0 / 3.0) * WL[3]);
  const double cR = sqrtp((20.0 / 3.0) * WR[3]);
  smax = fabs(0.5 * (WL[1] + WR[1])) + (cR > cL ? cR : cL);   /* max(cR,cL), tools.h:3 */
  const double lamL = fmin(WL[1] - cL, 0.0);
  const double lamR = fmax(WR[1] + cR, 0.0);
  const double alph = lamR - lamL;
  double ca, cb;
  if (alph > 0.0) {
    const double s = dtdx * rcp(alph);
    ca = lamR * s;
    cb = -lamL * s;
  } else {
    ca = 0.5 * dtdx;
    cb = 0.5 * dtdx;
  }
  const double wL = WL[1] - lamL, wR = WR[1] - lamR;
  const Side sL = side_of(WL), sR = side_of(WR);
  Fs[0] = ca * gshift<0>(WL, sL, wL) + cb * gshift<0>(WR, sR, wR);
  Fs[1] = ca * gshift<1>(WL, sL, wL) + cb * gshift<1>(WR, sR, wR);
  Fs[2] = ca * gshift<2>(WL, sL, wL) + cb * gshift<2>(WR, sR, wR);
  Fs[3] = ca * gshift<3>(WL, sL, wL) + cb * gshift<3>(WR, sR, wR);
  Fs[4] = ca * gshift<4>(WL, sL, wL) + cb * gshift<4>(WR, sR, wR);
  Fs[5] = ca * gshift<5>(WL, sL, wL) + cb * gshift<5>(WR, sR, wR);
  Fs[6] = ca * gshift<6>(WL, sL, wL) + cb * gshift<6>(WR, sR, wR);
  Fs[7] = ca * gshift<7>(WL, sL, wL) + cb * gshift<7>(WR, sR, wR);
  Fs[8] = ca * gshift<8>(WL, sL, wL) + cb * gshift<8>(WR, sR, wR);
}

/* x-face: frame = Cartesian. y-face: swap (1,2), (4,6), (7,8) in and out
 * (the rotation of Flux.c:126-172 with n=(0,1) is exactly this permutation). */
template <int DIR>
MBM_DEV constexpr int perm(int n)
{
  if constexpr (DIR == 0) {
    return n;
  } else {
    return n == 1 ? 2 : n == 2 ? 1 : n == 4 ? 6 : n == 6 ? 4 : n == 7 ? 8 : n == 8 ? 7 : n;
  }
}

template <int DIR>
MBM_DEV void face(const double *L, const double *R, double dtdx, double *Fs,
                  double &smax)
{
  double WL[9], WR[9], Ff[9];
#pragma unroll
  for (int n = 0; n < 9; n++) {
    WL[n] = L[perm<DIR>(n)];
    WR[n] = R[perm<DIR>(n)];
  }
  hll(WL, WR, dtdx, Ff, smax);
#pragma unroll
  for (int n = 0; n < 9; n++) Fs[perm<DIR>(n)] = Ff[n];
}

/* a face whose reference density is exactly zero: every component of the
 * reference's flux is 0/0 (Flux.c:85-103 divide by rho after Conservative) */
MBM_DEV void nan_face(double *Fs)
{
#pragma unroll
  for (int n = 0; n < 9; n++) Fs[n] = __builtin_nan("");
}

/* Projection of one loaded cell (Main.c:314-336):
 *   P  = primitives used for slopes and faces: relaxed with flag 1 (stage 1)
 *        or 3 (stage 2) on rings <= 2, unrelaxed on ring 3;
 *   I  = initial value of the new state: relaxation flag 0 (stage 1) or
 *        flag 2 of the stage input plus flag -1 of DATA (stage 2);
 *   cp = production CFL 0.5*(1-R0(2z/3)) (stage 1 only). */
/* bad is set when the relaxed primitives are not all finite (a non-finite
 * input, or a NaN damping factor from a negative temperature) */
template <int STAGE, bool NEED_INIT>
MBM_DEV void project(const double *U, const double *U0, bool ring3, double dt,
                     double dtkn, double force, double *P, double *I, double &cp, bool &bad)
{
  double Q[9];
  primitive(U, Q);
  const double z = Q[0] * sqrtp(Q[3]) * dtkn;
  double i1, i2;
  damp_inv(z, i1, i2);
  if constexpr (STAGE == 1) {
    relax(Q, damp1(z, i1, i2), P);
    bad = !finite_relaxed(P);
    if constexpr (NEED_INIT) {
      const Damp d0 = damp0(z, i1, i2);
      cp = 0.5 * (1.0 - d0.d2);
      if (force == 0.0 && __builtin_expect(!bad, 1)) {
        relax_cons(U, Q, d0, I);
      } else {
        /* a body force (Flux.c:239-253 adds dt*force to v1 before the
         * round trip), or non-finite data, whose NaNs follow the
         * reference's round trip */
        double R[9];
        relax(Q, d0, R);
        R[1] += dt * force;
        conservative(R, I);
      }
    }
  } else {
    Damp d3;
    d3.d1 = i1;
    d3.d2 = i2;
    relax(Q, d3, P);
    bad = !finite_relaxed(P);
    if constexpr (NEED_INIT) {
      double R[9], Q0[9], V[9];
      if (force != 0.0) {
        /* flag 2 (Flux.c:247-253): force, then the identity relaxation */
#pragma unroll
        for (int n = 0; n < 9; n++) R[n] = Q[n];
        R[1] += dt * force;
        conservative(R, I);
      } else if (__builtin_expect(!bad, 1)) {
        /* without force flag 2 is the identity: Conservative(Primitive(U)) = U */
#pragma unroll
        for (int n = 0; n < 9; n++) I[n] = U[n];
      } else {
        /* non-finite data: the round trip of Flux.c:230-267 with the factor
         * R2(z) = (1+z/2)/(1+z/2) (Flux.c:220), which is NaN when z is */
        const double a = 1.0 + 0.5 * z, two = 1.0 + 0.5 * ((2.0 / 3.0) * z);
        Damp dI;
        dI.d1 = a / a;
        dI.d2 = two / two;
        relax(Q, dI, R);
        conservative(R, I);
      }
      primitive(U0, Q0);
      const double z0 = Q0[0] * sqrtp(Q0[3]) * dtkn;
      double j1, j2;
      damp_inv(z0, j1, j2);
      relax(Q0, damp0(z0, j1, j2), R);
      conservative(R, V);
#pragma unroll
      for (int n = 0; n < 9; n++) I[n] += V[n];
      /* production CFL of flag 2 (Flux.c:264): 0.5*(1 - R2(2z/3)), zero unless
       * that factor is 0/0 or inf/inf */
      const double two = 1.0 + 0.5 * ((2.0 / 3.0) * z);
      cp = (__builtin_isfinite(two) && two != 0.0) ? 0.0 : __builtin_nan("");
    }
  }
  if (ring3) {
#pragma unroll
    for (int n = 0; n < 9; n++) P[n] = Q[n];
  }
}

template <int STAGE, bool NEED_INIT>
MBM_DEV void project(const double *U, const double *U0, bool ring3, double dt,
                     double dtkn, double force, double *P, double *I, double &cp)
{
  bool bad;
  project<STAGE, NEED_INIT>(U, U0, ring3, dt, dtkn, force, P, I, cp, bad);
}

/* ---------------- reference-form paths for non-finite data ----------------
 * Restatements of the reference expressions whose NaN behaviour the fast
 * forms do not share; used only for rows whose stencil holds a non-finite
 * primitive. Values agree with the fast forms to rounding. */

MBM_DEV double ref_max(double a, double b) { return a > b ? a : b; }          /* tools.h:3-6 */
MBM_DEV double ref_min(double a, double b) { return a < b ? a : b; }          /* tools.h:8-11 */
MBM_DEV double ref_sgn(double a) { return a == 0 ? 0.0 : (a < 0 ? -1.0 : 1.0); }   /* tools.h:13-17 */

/* Limit1 (Limiter.c:26-59) without the /dx that the reconstruction cancels */
template <int LIM>
MBM_DEV double limit_ref(double d1, double d2)
{
#pragma clang fp contract(off)
  if constexpr (LIM == 0) {
    return 0.0;
  } else if constexpr (LIM == 1) {
    return 0.5 * (d1 + d2);
  } else if constexpr (LIM == 2) {
    return 0.5 * (ref_sgn(d1) + ref_sgn(d2)) * ref_min(fabs(d1), fabs(d2));
  } else if constexpr (LIM == 3) {
    const double r = d2 / d1;
    double phi = 2.0;
    if (r < 0.0) phi = 0.0;
    else if (r < 0.5) phi = 2.0 * r;
    else if (r < 1.0) phi = 1.0;
    else if (r < 2.0) phi = r;
    return d1 * phi;
  } else {
    const double w1 = fabs(d1) + 1.0e-10, w2 = fabs(d2) + 1.0e-10;
    return (w1 * d2 + w2 * d1) / (w1 + w2);
  }
}

/* Face states of one cell as Main.c:346-374 forms them, from the one-sided
 * differences d1 = p - p_minus, d2 = p_plus - p: slope Limit1(d1, d2)/dx,
 * then p -/+ (0.5*dx)*slope for the lower/upper face, each operation
 * rounded on its own as in the reference build */
template <int LIM>
MBM_DEV void reconstruct_ref(double p, double d1, double d2, double dx, double &lo, double &hi)
{
#pragma clang fp contract(off)
  const double sl = limit_ref<LIM>(d1, d2) / dx;
  const double h = 0.5 * dx;
  lo = p - h * sl;
  hi = p + h * sl;
}

/* The reference's face densities of a cell may be exactly zero only where
 * the fast ones, flo = p - s/2 and fhi = p + s/2 (s the fast limited slope),
 * are within 1e-13 |p| of zero: both forms agree to a few ulps of p. */
MBM_DEV bool maybe_zero_density(double p, double flo, double fhi)
{
  const double tiny = 1e-13 * fabs(p);
  return fabs(flo) <= tiny || fabs(fhi) <= tiny;
}

/* Whether the reference's face densities of a cell are exactly zero, decided
 * in the reference's arithmetic where maybe_zero_density holds (the density
 * is never transformed by the projection, so p and the differences are the
 * reference's own values). */
template <int LIM>
MBM_DEV void zero_density_ref(double p, double d1, double d2, double dx, double flo, double fhi,
                              bool &zlo, bool &zhi)
{
  zlo = zhi = false;
  if (__builtin_expect(maybe_zero_density(p, flo, fhi), 0)) {
    double lo, hi;
    reconstruct_ref<LIM>(p, d1, d2, dx, lo, hi);
    zlo = lo == 0.0;
    zhi = hi == 0.0;
  }
}

/* Conservative with E from primDef.h (Flux.c:61-74) */
MBM_DEV void conservative_ref(const double *P, double *U)
{
#pragma clang fp contract(off)
  const double rho = P[0], v1 = P[1], v2 = P[2], th = P[3];
  const double vs = v1 * v1 + v2 * v2;
  const double E = 0.5 * rho * vs + 1.5 * rho * th;
  U[0] = rho;
  U[1] = rho * v1;
  U[2] = rho * v2;
  U[3] = E;
  U[4] = rho * (P[4] + v1 * v1);
  U[5] = rho * (P[5] + v1 * v2);
  U[6] = rho * (P[6] + v2 * v2);
  U[7] = P[7] + E * v1 + rho * (P[4] * v1 + P[5] * v2);
  U[8] = P[8] + E * v2 + rho * (P[5] * v1 + P[6] * v2);
}

/* component n of conservative_ref, with the same expressions (so the same
 * rounding) */
MBM_DEV double conservative_ref_comp(const double *P, int n)
{
#pragma clang fp contract(off)
  const double rho = P[0], v1 = P[1], v2 = P[2], th = P[3];
  const double vs = v1 * v1 + v2 * v2;
  const double E = 0.5 * rho * vs + 1.5 * rho * th;
  switch (n) {
    case 0: return rho;
    case 1: return rho * v1;
    case 2: return rho * v2;
    case 3: return E;
    case 4: return rho * (P[4] + v1 * v1);
    case 5: return rho * (P[5] + v1 * v2);
    case 6: return rho * (P[6] + v2 * v2);
    case 7: return P[7] + E * v1 + rho * (P[4] * v1 + P[5] * v2);
    default: return P[8] + E * v2 + rho * (P[5] * v1 + P[6] * v2);
  }
}

/* Grad13Flux1D (Flux.c:79-119): primitives recovered by division */
MBM_DEV void flux1d_ref(const double *U, double *F)
{
#pragma clang fp contract(off)
  const double rho = U[0], v1 = U[1] / rho, v2 = U[2] / rho;
  const double vs = v1 * v1 + v2 * v2;
  const double th = (2 * U[3] / rho - vs) / 3, E = U[3];
  const double T11 = U[4] / rho - v1 * v1, T12 = U[5] / rho - v1 * v2, T22 = U[6] / rho - v2 * v2;
  const double q1 = U[7] - E * v1 - rho * (T11 * v1 + T12 * v2);
  const double q2 = U[8] - E * v2 - rho * (T12 * v1 + T22 * v2);
  const double M111 = 1.2 * q1 / rho, M121 = 0.4 * q2 / rho, M112 = M121;
  const double M221 = 0.4 * q1 / rho, M122 = M221;
  const double R11 = 5 * th * th + 7 * th * (T11 - th), R12 = 7 * th * T12;
  F[0] = rho * v1;
  F[1] = rho * (T11 + v1 * v1);
  F[2] = rho * (T12 + v1 * v2);
  F[3] = E * v1 + rho * (T11 * v1 + T12 * v2) + q1;
  F[4] = rho * (M111 + 3 * T11 * v1 + v1 * v1 * v1);
  F[5] = rho * (M121 + T11 * v2 + 2 * T12 * v1 + v1 * v2 * v1);
  F[6] = rho * (M221 + T22 * v1 + 2 * T12 * v2 + v2 * v2 * v1);
  F[7] = 2 * (q1 + rho * (T11 * v1 + T12 * v2)) * v1 + E * v1 * v1 + rho * (M111 * v1 + M112 * v2)
       + 0.5 * rho * (T11 * vs + R11);
  F[8] = q2 * v1 + q1 * v2 + rho * (v1 * v2 * (T11 + T22) + 1.5 * T12 * vs) + E * v1 * v2
       + rho * (M121 * v1 + M122 * v2) + 0.5 * rho * R12;
}

/* Grad13Flux2D (Flux.c:126-172) for n = (nx, ny), t = (ny, -nx); the
 * products with 0 and 1 are kept: they carry NaN across components */
MBM_DEV void flux2d_ref(const double *U, double *F, double nx, double ny)
{
#pragma clang fp contract(off)
  const double tx = ny, ty = -nx;
  double UR[9], FR[9], s1, s2;
  UR[0] = U[0];
  UR[1] = U[1] * nx + U[2] * ny;
  UR[2] = U[1] * tx + U[2] * ty;
  UR[3] = U[3];
  s1 = U[4] * nx + U[5] * ny;
  s2 = U[5] * nx + U[6] * ny;
  UR[4] = s1 * nx + s2 * ny;
  UR[5] = s1 * tx + s2 * ty;
  s1 = U[4] * tx + U[5] * ty;
  s2 = U[5] * tx + U[6] * ty;
  UR[6] = s1 * tx + s2 * ty;
  UR[7] = U[7] * nx + U[8] * ny;
  UR[8] = U[7] * tx + U[8] * ty;
  flux1d_ref(UR, FR);
  F[0] = FR[0];
  F[1] = FR[1] * nx + FR[2] * tx;
  F[2] = FR[1] * ny + FR[2] * ty;
  F[3] = FR[3];
  s1 = FR[4] * nx + FR[5] * tx;
  s2 = FR[5] * nx + FR[6] * tx;
  F[4] = s1 * nx + s2 * tx;
  F[5] = s1 * ny + s2 * ty;
  s1 = FR[4] * ny + FR[5] * ty;
  s2 = FR[5] * ny + FR[6] * ty;
  F[6] = s1 * ny + s2 * ty;
  F[7] = FR[7] * nx + FR[8] * tx;
  F[8] = FR[7] * ny + FR[8] * ty;
}

/* FnumHLLE (Flux.c:332-379) on Cartesian primitives L, R, x-face (DIR 0) or
 * y-face (DIR 1); Fs = (dt/dx) * flux, smax = cMax (lamDiff is 1 there).
 * The combine keeps the reference's order, lamR fL - lamL fR + lamL lamR
 * (UR - UL), so that values agree to rounding also where UR - UL cancels;
 * UL and UR are formed per component when they are needed, which keeps few
 * temporaries live (the stage kernel runs this inside its march). */
template <int DIR>
MBM_DEV void face_ref(const double *L, const double *R, double dtdx, double *Fs, double &smax)
{
#pragma clang fp contract(off)
  const double nx = DIR == 0 ? 1.0 : 0.0, ny = DIR == 0 ? 0.0 : 1.0;
  const double cR = 2.0 * sqrt(5.0 / 3 * R[3]), cL = 2.0 * sqrt(5.0 / 3 * L[3]);
  const double vL = L[1 + DIR], vR = R[1 + DIR];
  double fL[9], fR[9];
  {
    double U[9];
    conservative_ref(L, U);
    flux2d_ref(U, fL, nx, ny);
  }
  {
    double U[9];
    conservative_ref(R, U);
    flux2d_ref(U, fR, nx, ny);
  }
  smax = fabs(0.5 * (vL + vR)) + ref_max(cR, cL);
  const double lamL = ref_min(vL - cL, 0.0), lamR = ref_max(vR + cR, 0.0);
  const double alph = lamR - lamL;
  if (alph > 0.0) {
#pragma unroll
    for (int n = 0; n < 9; n++) {
      const double uL = conservative_ref_comp(L, n), uR = conservative_ref_comp(R, n);
      Fs[n] = dtdx * ((lamR * fL[n] - lamL * fR[n] + lamL * lamR * (uR - uL)) / alph);
    }
  } else {
#pragma unroll
    for (int n = 0; n < 9; n++) Fs[n] = dtdx * (0.5 * (fL[n] + fR[n]));
  }
}

} // namespace mbm

#endif
