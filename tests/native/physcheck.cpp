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
