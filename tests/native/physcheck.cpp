// Host build of the engine's cell/face functions and their hand-written
// VJPs, for finite-difference checks in tests/test_adjoint_local.py.
#include "../../multiblockmomentsfv_amd/csrc/mbm_adj.h"

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
