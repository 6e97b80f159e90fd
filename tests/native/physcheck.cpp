// Host build of the engine's cell/face functions and their hand-written
// VJPs, for finite-difference checks in tests/test_adjoint_local.py.
#include "../../multiblockmomentsfv_amd/csrc/mbm_adj_stage.h"
#include <vector>

using namespace mbm;

extern "C" {
void pc_primitive(const double *U, double *P) { primitive(U, P); }
void pc_primitive_vjp(const double *U, const double *g, double *gU) { primitive_vjp(U, g, gU); }
void pc_conservative(const double *P, double *U) { conservative(P, U); }
void pc_conservative_vjp(const double *P, const double *g, double *gP) { conservative_vjp(P, g, gP); }
void pc_face(int dir, const double *L, const double *R, double dtdx, double *F)
{
  double s;
  if (dir == 0) face<0>(L, R, dtdx, F, s);
  else face<1>(L, R, dtdx, F, s);
}
void pc_face_vjp(int dir, const double *L, const double *R, double dtdx, const double *g,
                 double *gL, double *gR)
{
  if (dir == 0) face_vjp<0>(L, R, dtdx, g, gL, gR);
  else face_vjp<1>(L, R, dtdx, g, gL, gR);
}
double pc_limit(int lim, double d1, double d2)
{
  switch (lim) {
    case 0: return limit<0>(d1, d2);
    case 1: return limit<1>(d1, d2);
    case 2: return limit<2>(d1, d2);
    case 3: return limit<3>(d1, d2);
    default: return limit<4>(d1, d2);
  }
}
void pc_limit_vjp(int lim, double d1, double d2, double g, double *gd)
{
  switch (lim) {
    case 0: limit_vjp<0>(d1, d2, g, gd[0], gd[1]); break;
    case 1: limit_vjp<1>(d1, d2, g, gd[0], gd[1]); break;
    case 2: limit_vjp<2>(d1, d2, g, gd[0], gd[1]); break;
    case 3: limit_vjp<3>(d1, d2, g, gd[0], gd[1]); break;
    default: limit_vjp<4>(d1, d2, g, gd[0], gd[1]); break;
  }
}
void pc_relax(const double *P, double d1, double d2, double *R) { Damp d{d1, d2}; relax(P, d, R); }
void pc_relax_vjp(const double *P, double d1, double d2, const double *g, double *gP, double *gd)
{
  Damp d{d1, d2};
  relax_vjp(P, d, g, gP, gd[0], gd[1]);
}
void pc_damp(int flag, double z, double *d)
{
  double i1, i2;
  damp_inv(z, i1, i2);
  Damp r = flag == 0 ? damp0(z, i1, i2) : flag == 1 ? damp1(z, i1, i2) : Damp{i1, i2};
  d[0] = r.d1;
  d[1] = r.d2;
}
void pc_damp_dz(int flag, double z, double *d)
{
  double i1, i2;
  damp_inv(z, i1, i2);
  Damp r = flag == 0 ? damp0_dz(z, i1, i2) : flag == 1 ? damp1_dz(z, i1, i2) : damp3_dz(i1, i2);
  d[0] = r.d1;
  d[1] = r.d2;
}
}

/* Stage adjoint of one block (SoA: component m of cell (a,b) at
 * m*lx*ly + b*lx + a), built from the same per-cell functions as the GPU
 * kernels. stage 1: gin += S1^T gout; stage 2: gin = S2^T gout (w.r.t. the
 * stage input), gin0 += (w.r.t. DATA). Returns dJ/dKn. */
struct HostBlk {
  long long off, plane;
  int lx, ly, pitch;
  double dx, dy;
};

template <int STAGE, int LIM>
static double stage_vjp(const HostBlk &B, const double *in, const double *in0, const double *gout,
                        double *gin, double *gin0, double dt, double kn, double force)
{
  const size_t n = (size_t)B.plane * 9;
  std::vector<double> P(n), FxL(n), FxR(n), FyL(n), FyR(n);
  const double scale = STAGE == 2 ? 0.5 : 1.0;
  for (int b = 0; b < B.ly; b++)
    for (int a = 0; a < B.lx; a++) adj_prim_cell<STAGE>(in, P.data(), B, a, b, dt, dt / kn, force);
  for (int b = 0; b < B.ly; b++)
    for (int a = 0; a < B.lx; a++)
      adj_face_cell<LIM>(P.data(), gout, FxL.data(), FxR.data(), FyL.data(), FyR.data(), B, a, b,
                         dt, scale);
  double gkn = 0.0;
  for (int b = 0; b < B.ly; b++)
    for (int a = 0; a < B.lx; a++)
      gkn += adj_cell_cell<STAGE, LIM>(in, in0, P.data(), FxL.data(), FxR.data(), FyL.data(),
                                       FyR.data(), gout, gin, gin0, B, a, b, dt, kn, force, scale);
  return gkn;
}

extern "C" double pc_stage_vjp(int stage, int lim, int nx, int ny, double dx, double dy, const double *in,
                    const double *in0, const double *gout, double *gin, double *gin0, double dt,
                    double kn, double force)
{
  HostBlk B;
  B.lx = nx + 6;
  B.ly = ny + 6;
  B.pitch = B.lx;
  B.plane = (long long)B.lx * B.ly;
  B.off = 0;
  B.dx = dx;
  B.dy = dy;
#define CASE(S, L) if (stage == S && lim == L) return stage_vjp<S, L>(B, in, in0, gout, gin, gin0, dt, kn, force)
  CASE(1, 0); CASE(1, 1); CASE(1, 2); CASE(1, 3); CASE(1, 4);
  CASE(2, 0); CASE(2, 1); CASE(2, 2); CASE(2, 3); CASE(2, 4);
#undef CASE
  return 0.0;
}

/* Host emulation of one stage of the fused GPU stage kernel on one block:
 * the same per-cell functions (project, limit3 in groups of three, face)
 * and the same accumulation order per cell (stage 1, upward march:
 * init + G(j-1/2) + (F(i-1/2) - F(i+1/2)) - G(j+1/2); stage 2, downward:
 * init - G(j+1/2) + (F(i-1/2) - F(i+1/2)) + G(j-1/2), then halved). Arrays
 * are AoS (ny+6, nx+6, 9), x fastest, as the oracle's layers. Writes the
 * interior cells of out; returns dt*cx/dx + dt*cy/dy + cprod with each
 * maximum taken NaN-ignoring over the kernel's index sets. */
/* the three maxima of the last emulated stage: x-face speed, y-face speed,
 * production CFL */
extern "C" { double pc_last_maxima[3]; }

template <int STAGE, int LIM>
static double emu_stage(int nx, int ny, double dx, double dy, const double *in, const double *in0,
                        double *out, double dt, double kn, double force)
{
  const int lx = nx + 6, ly = ny + 6;
  const size_t nc = (size_t)lx * ly;
  std::vector<double> P(nc * 9), I(nc * 9), CP(nc);
  std::vector<char> BAD(nc);
  const double dtkn = dt / kn;
  for (int b = 0; b < ly; b++)
    for (int a = 0; a < lx; a++) {
      const size_t c = (size_t)b * lx + a;
      const bool ring3 = a == 0 || a == lx - 1 || b == 0 || b == ly - 1;
      double cp = 0.0;
      bool bad = false;
      project<STAGE, true>(in + 9 * c, in0 + 9 * c, ring3, dt, dtkn, force, &P[9 * c], &I[9 * c], cp, bad);
      CP[c] = cp;
      BAD[c] = bad;
    }
  auto p = [&](int a, int b) { return &P[9 * ((size_t)b * lx + a)]; };
  auto bad_at = [&](int a, int b) { return BAD[(size_t)b * lx + a] != 0; };
  /* limited slopes of cell (a,b) in direction dir (0: x, 1: y), groups of
   * three; reference form (ref) for stencils with non-finite data */
  /* face states of cell (a,b) in direction dir: lo/hi; zlo/zhi: the
   * reference's density there is exactly zero */
  auto states = [&](int a, int b, int dir, double *lo, double *hi, bool ref, bool &zlo, bool &zhi) {
    const double *m = dir ? p(a, b - 1) : p(a - 1, b), *c = p(a, b), *q = dir ? p(a, b + 1) : p(a + 1, b);
    const double h = dir ? dy : dx;
    if (ref) {
      for (int n = 0; n < 9; n++) reconstruct_ref<LIM>(c[n], c[n] - m[n], q[n] - c[n], h, lo[n], hi[n]);
      zlo = lo[0] == 0.0;
      zhi = hi[0] == 0.0;
      return;
    }
    double s[9];
    for (int g = 0; g < 9; g += 3) {
      double e1[3], e2[3];
      for (int k = 0; k < 3; k++) { e1[k] = c[g + k] - m[g + k]; e2[k] = q[g + k] - c[g + k]; }
      limit3<LIM>(e1, e2, s + g);
    }
    for (int n = 0; n < 9; n++) { lo[n] = c[n] - 0.5 * s[n]; hi[n] = c[n] + 0.5 * s[n]; }
    zero_density_ref<LIM>(c[0], c[0] - m[0], q[0] - c[0], h, lo[0], hi[0], zlo, zhi);
  };
  const double dtdx = dt / dx, dtdy = dt / dy;
  std::vector<double> FX(nc * 9, 0.0), FY(nc * 9, 0.0);   /* face i+1/2 / j+1/2 of cell (a,b) */
  double mxx = 0.0, mxy = 0.0, mxp = 0.0;
  for (int b = 1; b <= ly - 2; b++)
    for (int a = 1; a <= lx - 3; a++) {
      double lo0[9], L[9], R[9], hi1[9], s;
      bool z0, z1, z2, z3;
      const bool ref = bad_at(a - 1, b) || bad_at(a, b) || bad_at(a + 1, b) || (a + 2 < lx && bad_at(a + 2, b));
      states(a, b, 0, lo0, L, ref, z0, z1);
      states(a + 1, b, 0, R, hi1, ref, z2, z3);
      double *F = &FX[9 * ((size_t)b * lx + a)];
      if (ref) face_ref<0>(L, R, dtdx, F, s);
      else {
        face<0>(L, R, dtdx, F, s);
        if (z1 || z2) nan_face(F);
      }
      if (b <= ly - 3) mxx = s > mxx ? s : mxx;
    }
  for (int b = 1; b <= ly - 3; b++)
    for (int a = 1; a <= lx - 2; a++) {
      double lo0[9], L[9], R[9], hi1[9], s;
      bool z0, z1, z2, z3;
      const bool ref = bad_at(a, b - 1) || bad_at(a, b) || bad_at(a, b + 1) || (b + 2 < ly && bad_at(a, b + 2));
      states(a, b, 1, lo0, L, ref, z0, z1);
      states(a, b + 1, 1, R, hi1, ref, z2, z3);
      double *F = &FY[9 * ((size_t)b * lx + a)];
      if (ref) face_ref<1>(L, R, dtdy, F, s);
      else {
        face<1>(L, R, dtdy, F, s);
        if (z1 || z2) nan_face(F);
      }
      if (a <= lx - 3) mxy = s > mxy ? s : mxy;
    }
  /* production CFL as the kernel and controller form it: the NaN-ignoring
   * maximum, and 1 + the last position holding a NaN in the reference's
   * scan order (i outer, j inner over rings >= 1); with a NaN, the maximum
   * over the positions after it (NaN if there are none) */
  unsigned long long pnan = 0;
  for (int b = 1; b <= ly - 2; b++)
    for (int a = 1; a <= lx - 2; a++) {
      const double v = CP[(size_t)b * lx + a];
      mxp = v > mxp ? v : mxp;
      if (v != v) {
        const unsigned long long pos = (unsigned long long)(a - 1) * (ly - 2) + (b - 1) + 1;
        pnan = pos > pnan ? pos : pnan;
      }
    }
  if (pnan) {
    const unsigned long long ncell = (unsigned long long)(lx - 2) * (ly - 2);
    double m = -HUGE_VAL;
    for (unsigned long long q = pnan; q < ncell; q++) {
      const int a = 1 + (int)(q / (ly - 2)), b = 1 + (int)(q % (ly - 2));
      const double v = CP[(size_t)b * lx + a];
      m = v > m ? v : m;
    }
    mxp = pnan >= ncell ? __builtin_nan("") : m;
  }
  for (int b = 3; b < ly - 3; b++)
    for (int a = 3; a < lx - 3; a++) {
      const size_t c = (size_t)b * lx + a;
      for (int n = 0; n < 9; n++) {
        const double fx = FX[9 * (c - 1) + n] - FX[9 * c + n];
        const double gm = FY[9 * (c - lx) + n], gp = FY[9 * c + n];
        double v;
        if (STAGE == 1) v = ((I[9 * c + n] + gm) + fx) - gp;
        else v = 0.5 * (((I[9 * c + n] - gp) + fx) + gm);
        out[9 * c + n] = v;
      }
    }
  pc_last_maxima[0] = mxx;
  pc_last_maxima[1] = mxy;
  pc_last_maxima[2] = mxp;
  return dt * mxx / dx + dt * mxy / dy + mxp;
}

extern "C" double pc_emu_stage(int stage, int lim, int nx, int ny, double dx, double dy,
                               const double *in, const double *in0, double *out, double dt,
                               double kn, double force)
{
#define CASE(S, L) if (stage == S && lim == L) return emu_stage<S, L>(nx, ny, dx, dy, in, in0, out, dt, kn, force)
  CASE(1, 0); CASE(1, 1); CASE(1, 2); CASE(1, 3); CASE(1, 4);
  CASE(2, 0); CASE(2, 1); CASE(2, 2); CASE(2, 3); CASE(2, 4);
#undef CASE
  return 0.0;
}

/* the device projection of one cell (stage 1/2, with the initial value) */
extern "C" double pc_project(int stage, const double *U, const double *U0, int ring3, double dt,
                             double kn, double force, double *P, double *I)
{
  double cp = 0.0;
  if (stage == 1) project<1, true>(U, U0, ring3 != 0, dt, dt / kn, force, P, I, cp);
  else project<2, true>(U, U0, ring3 != 0, dt, dt / kn, force, P, I, cp);
  return cp;
}

extern "C" void pc_limit3(int lim, const double *d1, const double *d2, double *s)
{
  switch (lim) {
    case 0: limit3<0>(d1, d2, s); break;
    case 1: limit3<1>(d1, d2, s); break;
    case 2: limit3<2>(d1, d2, s); break;
    case 3: limit3<3>(d1, d2, s); break;
    default: limit3<4>(d1, d2, s); break;
  }
}

/* the reference-form limiter and face used for rows with non-finite data */
extern "C" double pc_limit_ref(int lim, double d1, double d2)
{
  switch (lim) {
    case 0: return limit_ref<0>(d1, d2);
    case 1: return limit_ref<1>(d1, d2);
    case 2: return limit_ref<2>(d1, d2);
    case 3: return limit_ref<3>(d1, d2);
    default: return limit_ref<4>(d1, d2);
  }
}
extern "C" double pc_face_ref(int dir, const double *L, const double *R, double dtdx, double *F)
{
  double s;
  if (dir == 0) face_ref<0>(L, R, dtdx, F, s);
  else face_ref<1>(L, R, dtdx, F, s);
  return s;
}
