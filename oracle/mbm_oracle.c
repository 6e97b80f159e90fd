/*
 * mbm_oracle.c -- CPU oracle (test infrastructure only; see mbm_oracle.h).
 *
 * Every expression below keeps the evaluation order of the reference source
 * it cites, because the goal is bit-identity with the reference build:
 * operands are grouped exactly as C parses the reference's expressions.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <assert.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "mbm_oracle.h"

#define NV ORC_NV
#define MBC ORC_MBC

typedef struct {
  int nx, ny;      /* interior cells */
  int lx, ly;      /* nx+6, ny+6: line lengths including 3 ghost rings */
  double bx, by;   /* GetPoint origin: xmin-2.5dx (Mesh.c:169-172) */
  double dx, dy;
  int bc[4];       /* left, right, bottom, top */
} oblk;

struct orc_ctx {
  int nb;
  oblk b[ORC_MAXB];
  double kn, force;
  int limiter, scenario;
  double tj[5];
  double *lay[3][ORC_MAXB];
  double *prim[ORC_MAXB], *sx[ORC_MAXB], *sy[ORC_MAXB];
};

static inline double mx(double a, double b) { return a > b ? a : b; } /* tools.h:3 */
static inline double mn(double a, double b) { return a < b ? a : b; } /* tools.h:8 */
static inline double sg(double a) { if (a == 0) return 0; return a < 0 ? -1 : 1; }

void orc_set_threads(int n)
{
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ------------------------------------------------------------------ */
/* cell physics                                                        */
/* ------------------------------------------------------------------ */

/* Flux.c:30-56 */
void orc_primitive(const double *U, double *P)
{
  double rho = U[0];
  double v1 = U[1] / rho;
  double v2 = U[2] / rho;
  double vs = v1 * v1 + v2 * v2;
  double th = (2 * U[3] / rho - vs) / 3;
  double E = U[3];
  double t11 = U[4] / rho - v1 * v1;
  double t12 = U[5] / rho - v1 * v2;
  double t22 = U[6] / rho - v2 * v2;
  P[0] = rho; P[1] = v1; P[2] = v2; P[3] = th;
  P[4] = t11; P[5] = t12; P[6] = t22;
  P[7] = U[7] - E * v1 - rho * (t11 * v1 + t12 * v2);
  P[8] = U[8] - E * v2 - rho * (t12 * v1 + t22 * v2);
}

/* Flux.c:61-74 with primDef.h:8-20 (E recomputed from the primitives) */
void orc_conservative(const double *P, double *U)
{
  double rho = P[0], v1 = P[1], v2 = P[2], th = P[3];
  double t11 = P[4], t12 = P[5], t22 = P[6], q1 = P[7], q2 = P[8];
  double vs = v1 * v1 + v2 * v2;
  double E = 0.5 * rho * vs + 1.5 * rho * th;
  U[0] = rho;
  U[1] = rho * v1;
  U[2] = rho * v2;
  U[3] = E;
  U[4] = rho * (t11 + v1 * v1);
  U[5] = rho * (t12 + v1 * v2);
  U[6] = rho * (t22 + v2 * v2);
  U[7] = q1 + E * v1 + rho * (t11 * v1 + t12 * v2);
  U[8] = q2 + E * v2 + rho * (t12 * v1 + t22 * v2);
}

/* Grad-13 flux in the local normal direction (Flux.c:79-119). The primitive
 * recovery inside it is the same arithmetic as Flux.c:30-56. */
static void g13_flux_normal(const double *U, double *F)
{
  double P[NV];
  orc_primitive(U, P);
  double rho = P[0], v1 = P[1], v2 = P[2], th = P[3];
  double t11 = P[4], t12 = P[5], t22 = P[6], q1 = P[7], q2 = P[8];
  double vs = v1 * v1 + v2 * v2;
  double E = U[3];
  double m111 = 1.2 * q1 / rho;
  double m121 = 0.4 * q2 / rho, m112 = m121;
  double m221 = 0.4 * q1 / rho, m122 = m221;
  double r11 = 5 * th * th + 7 * th * (t11 - th);
  double r12 = 7 * th * t12;
  F[0] = rho * v1;
  F[1] = rho * (t11 + v1 * v1);
  F[2] = rho * (t12 + v1 * v2);
  F[3] = E * v1 + rho * (t11 * v1 + t12 * v2) + q1;
  F[4] = rho * (m111 + 3 * t11 * v1 + v1 * v1 * v1);
  F[5] = rho * (m121 + t11 * v2 + 2 * t12 * v1 + v1 * v2 * v1);
  F[6] = rho * (m221 + t22 * v1 + 2 * t12 * v2 + v2 * v2 * v1);
  F[7] = 2 * (q1 + rho * (t11 * v1 + t12 * v2)) * v1 + E * v1 * v1
       + rho * (m111 * v1 + m112 * v2) + 0.5 * rho * (t11 * vs + r11);
  F[8] = q2 * v1 + q1 * v2 + rho * (v1 * v2 * (t11 + t22) + 1.5 * t12 * vs)
       + E * v1 * v2 + rho * (m121 * v1 + m122 * v2) + 0.5 * rho * r12;
}

/* rotation into (n,t) with t=(ny,-nx), flux, rotation back (Flux.c:126-172).
 * The multiplications by 0 and +-1 are kept: they fix the signs of zeros. */
void orc_flux_dir(const double *U, double *F, double nx, double ny)
{
  double R[NV], FR[NV], s1, s2;
  double tx = ny, ty = -nx;
  R[0] = U[0];
  R[1] = U[1] * nx + U[2] * ny;
  R[2] = U[1] * tx + U[2] * ty;
  R[3] = U[3];
  s1 = U[4] * nx + U[5] * ny;
  s2 = U[5] * nx + U[6] * ny;
  R[4] = s1 * nx + s2 * ny;
  R[5] = s1 * tx + s2 * ty;
  s1 = U[4] * tx + U[5] * ty;
  s2 = U[5] * tx + U[6] * ty;
  R[6] = s1 * tx + s2 * ty;
  R[7] = U[7] * nx + U[8] * ny;
  R[8] = U[7] * tx + U[8] * ty;
  g13_flux_normal(R, FR);
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

/* damping factor of the relaxation (Flux.c:213-223) */
static inline double rdamp(double z, int flag)
{
  switch (flag) {
    case -1:
    case 0: return (1 + 0.5 * z - 0.25 * z * z) / (1 + z + 0.25 * z * z);
    case 1: return (1 + z) / (1 + z + 0.25 * z * z);
    case 2: return (1 + 0.5 * z) / (1 + 0.5 * z);
    case 3: return 1.0 / (1 + 0.5 * z);
  }
  return 0.0;
}

/* ProductionsDirect (Flux.c:230-267) with CollFreq and ExternalForce
 * (Scenario.c:28-42). Returns the production CFL 0.5*(1-R(2z/3)). */
double orc_relax(const orc_ctx *c, const double *U, double *out, double dt,
                 int flag)
{
  double P[NV];
  orc_primitive(U, P);
  double rho = P[0], v1 = P[1], v2 = P[2], th = P[3];
  double t11 = P[4], t12 = P[5], t22 = P[6], q1 = P[7], q2 = P[8];
  double vs = v1 * v1 + v2 * v2;
  double E = 0.5 * rho * vs + 1.5 * rho * th;
  double nu0 = 1.0;
  double coll = (nu0 * rho * sqrt(th) / c->kn) * dt;
  if (flag == 0 || flag == 2) {
    double g0 = c->force, g1 = 0.0;
    v1 += dt * g0;
    v2 += dt * g1;
    vs = v1 * v1 + v2 * v2;
    E = 0.5 * rho * vs + 1.5 * rho * th;
  }
  double d1 = rdamp(coll, flag);
  double d2 = rdamp(2.0 / 3 * coll, flag);
  t11 = d1 * (t11 - th) + th;
  t12 = d1 * t12;
  t22 = d1 * (t22 - th) + th;
  q1 = d2 * q1;
  q2 = d2 * q2;
  out[0] = rho;
  out[1] = rho * v1;
  out[2] = rho * v2;
  out[3] = E;
  out[4] = rho * (t11 + v1 * v1);
  out[5] = rho * (t12 + v1 * v2);
  out[6] = rho * (t22 + v2 * v2);
  out[7] = q1 + E * v1 + rho * (t11 * v1 + t12 * v2);
  out[8] = q2 + E * v2 + rho * (t12 * v1 + t22 * v2);
  return 0.5 * (1 - d2);
}

/* Limit1 (Limiter.c:26-59) */
static inline double superbee_phi(double r)
{
  if (r < 0.0) return 0.0;
  if (r < 0.5) return 2 * r;
  if (r < 1.0) return 1.0;
  if (r < 2.0) return r;
  return 2.0;
}

double orc_limit(int limiter, double d1, double d2)
{
  const double eps = 1.0e-10;
  double w1, w2;
  switch (limiter) {
    case ORC_LIM_NONE: return 0.0;
    case ORC_LIM_AVERAGE: return 0.5 * (d1 + d2);
    case ORC_LIM_MINMOD: return 0.5 * (sg(d1) + sg(d2)) * mn(fabs(d1), fabs(d2));
    case ORC_LIM_SUPERBEE: return d1 * superbee_phi(d2 / d1);
    default:
      w1 = fabs(d1) + eps;
      w2 = fabs(d2) + eps;
      return (w1 * d2 + w2 * d1) / (w1 + w2);
  }
}

/* lamDiff (Flux.c:318-325) */
static inline double lam_diff(double k)
{
  if (k > 4.1422533434191076) return (10 + 4 * (k + sqrt(6 + 4 * k))) / ((1 + 2 * k) * (1 + 2 * k));
  if (k > 1) return 1.15 - 0.15 * k;
  return 1.0;
}

/* FnumHLLE (Flux.c:332-379); dir 1 = x, 2 = y. Returns cMax. */
double orc_hll(const double *PL, const double *PR, double dx, int dir, double *flx)
{
  double fL[NV], fR[NV], UL[NV], UR[NV], Ut[NV];
  double cR = 2.0 * sqrt(5.0 / 3 * PR[3]);
  double cL = 2.0 * sqrt(5.0 / 3 * PL[3]);
  double vL, vR, sL = 0.0, sR = 0.0;
  if (dir == 1) {
    vR = PR[1]; vL = PL[1];
    orc_conservative(PL, Ut); orc_flux_dir(Ut, fL, 1.0, 0.0);
    orc_conservative(PR, Ut); orc_flux_dir(Ut, fR, 1.0, 0.0);
  } else {
    vR = PR[2]; vL = PL[2];
    orc_conservative(PL, Ut); orc_flux_dir(Ut, fL, 0.0, 1.0);
    orc_conservative(PR, Ut); orc_flux_dir(Ut, fR, 0.0, 1.0);
  }
  double cmax = fabs(0.5 * (vL + vR)) + mx(cR, cL);
  cmax = cmax / lam_diff(mx(sL, sR) / (cmax * dx));
  double lamL = mn(vL - cL, 0.0);
  double lamR = mx(vR + cR, 0.0);
  double alph = lamR - lamL;
  orc_conservative(PL, UL);
  orc_conservative(PR, UR);
  if (alph > 0.0)
    for (int n = 0; n < NV; n++)
      flx[n] = (lamR * fL[n] - lamL * fR[n] + lamL * lamR * (UR[n] - UL[n])) / alph;
  else
    for (int n = 0; n < NV; n++) flx[n] = 0.5 * (fL[n] + fR[n]);
  return cmax;
}

/* ------------------------------------------------------------------ */
/* scenarios (Scenario.c:72-153)                                      */
/* ------------------------------------------------------------------ */
void orc_initial_state(const orc_ctx *c, double x, double y, double t, double *U)
{
  (void)t;
  double rho, v1, v2, th, t11, t12, t22, q1, q2, vs, E;
  v1 = 0.0; v2 = 0.0; t12 = 0.0; q1 = 0.0; q2 = 0.0; th = 1.0;
  if (c->scenario == ORC_SCN_FFS) {
    rho = 1.0;
    v1 = 2.0 * sqrt(5.0 / 3.0);
  } else if (c->scenario == ORC_SCN_TEMPJUMP) {
    /* build-defined: smoothed temperature jump at uniform density */
    double r = sqrt(x * x + y * y);
    double s = 0.5 * (1.0 - tanh((r - c->tj[2]) / c->tj[3]));
    rho = c->tj[4];
    th = c->tj[1] + (c->tj[0] - c->tj[1]) * s;
  } else {
    rho = (sqrt(x * x + y * y) < 0.4) ? 5.0 : 1.0;
  }
  t11 = th;
  t22 = th;
  vs = v1 * v1 + v2 * v2;
  E = 0.5 * rho * vs + 1.5 * rho * th;
  U[0] = rho;
  U[1] = rho * v1;
  U[2] = rho * v2;
  U[3] = E;
  U[4] = rho * (t11 + v1 * v1);
  U[5] = rho * (t12 + v1 * v2);
  U[6] = rho * (t22 + v2 * v2);
  U[7] = q1 + E * v1 + rho * (t11 * v1 + t12 * v2);
  U[8] = q2 + E * v2 + rho * (t12 * v1 + t22 * v2);
}

/* ------------------------------------------------------------------ */
/* context                                                             */
/* ------------------------------------------------------------------ */
orc_ctx *orc_create(int nblocks, const int *nx, const int *ny,
                    const double *xmin, const double *xmax,
                    const double *ymin, const double *ymax,
                    const int *bc4, double kn, double force, int limiter,
                    int scenario)
{
  if (nblocks < 1 || nblocks > ORC_MAXB) return NULL;
  orc_ctx *c = calloc(1, sizeof *c);
  c->nb = nblocks;
  c->kn = kn;
  c->force = force;
  c->limiter = limiter;
  c->scenario = scenario;
  c->tj[0] = 5.0; c->tj[1] = 1.0; c->tj[2] = 0.4; c->tj[3] = 0.05; c->tj[4] = 1.0;
  for (int k = 0; k < nblocks; k++) {
    oblk *B = &c->b[k];
    int nst = 2 * MBC;
    B->nx = nx[k]; B->ny = ny[k];
    B->lx = nx[k] + nst; B->ly = ny[k] + nst;
    B->dx = (xmax[k] - xmin[k]) / nx[k];
    B->dy = (ymax[k] - ymin[k]) / ny[k];
    B->bx = xmin[k] - (nst / 2 - 0.5) * B->dx;
    B->by = ymin[k] - (nst / 2 - 0.5) * B->dy;
    for (int s = 0; s < 4; s++) B->bc[s] = bc4[4 * k + s];
    size_t n = (size_t)B->lx * B->ly * NV;
    for (int l = 0; l < 3; l++) c->lay[l][k] = calloc(n, sizeof(double));
    c->prim[k] = calloc(n, sizeof(double));
    c->sx[k] = calloc(n, sizeof(double));
    c->sy[k] = calloc(n, sizeof(double));
  }
  return c;
}

void orc_set_tempjump(orc_ctx *c, double th_in, double th_out, double radius,
                      double width, double rho)
{
  c->tj[0] = th_in; c->tj[1] = th_out; c->tj[2] = radius; c->tj[3] = width; c->tj[4] = rho;
}

void orc_destroy(orc_ctx *c)
{
  if (!c) return;
  for (int k = 0; k < c->nb; k++) {
    for (int l = 0; l < 3; l++) free(c->lay[l][k]);
    free(c->prim[k]); free(c->sx[k]); free(c->sy[k]);
  }
  free(c);
}

int orc_block_cells(const orc_ctx *c, int blk) { return c->b[blk].lx * c->b[blk].ly; }

void orc_get_layer(const orc_ctx *c, int layer, int blk, double *out)
{
  memcpy(out, c->lay[layer][blk], sizeof(double) * NV * orc_block_cells(c, blk));
}

void orc_set_layer(orc_ctx *c, int layer, int blk, const double *in)
{
  memcpy(c->lay[layer][blk], in, sizeof(double) * NV * orc_block_cells(c, blk));
}

/* ------------------------------------------------------------------ */
/* ghost cells (Boundary.c:88-193)                                     */
/* ------------------------------------------------------------------ */
enum { K_SYNC, K_OPEN, K_WALL, K_DIR };

static const double wall_sign[4][NV] = {
  { +1, -1, +1, +1, +1, -1, +1, -1, +1 },
  { +1, -1, +1, +1, +1, -1, +1, -1, +1 },
  { +1, +1, -1, +1, +1, -1, +1, +1, -1 },
  { +1, +1, -1, +1, +1, -1, +1, +1, -1 } };

static inline int kind_of(int code)
{
  if (code > 0) return K_SYNC;
  if (code == 0) return K_OPEN;
  if (code == -1) return K_WALL;
  return K_DIR;
}

/* local 0-based (a,b) of ghost layer n at tangential index ti */
static inline void ghost_ab(const oblk *B, int side, int ti, int n, int *a, int *b)
{
  switch (side) {
    case 0: *a = MBC - n; *b = ti; break;
    case 1: *a = B->lx - 1 - MBC + n; *b = ti; break;
    case 2: *a = ti; *b = MBC - n; break;
    default: *a = ti; *b = B->ly - 1 - MBC + n; break;
  }
}

/* local (a,b) of the s-th interior layer counted inward from 'side' */
static inline void inner_ab(const oblk *B, int side, int ti, int s, int *a, int *b)
{
  switch (side) {
    case 0: *a = MBC - 1 + s; *b = ti; break;
    case 1: *a = B->lx - MBC - s; *b = ti; break;
    case 2: *a = ti; *b = MBC - 1 + s; break;
    default: *a = ti; *b = B->ly - MBC - s; break;
  }
}

static void fill_edge(orc_ctx *c, int layer, int d, int s, int side, int kind, double t)
{
  static const int opp[4] = { 1, 0, 3, 2 };
  oblk *D = &c->b[d], *S = &c->b[s];
  double *dst = c->lay[layer][d];
  const double *src = c->lay[layer][s];
  int t1 = (side < 2) ? D->ly : D->lx;
  if (kind == K_SYNC)
    assert((side < 2) ? (D->ly == S->ly) : (D->lx == S->lx));
  for (int ti = 0; ti < t1; ti++)
    for (int n = 1; n <= MBC; n++) {
      int ga, gb, sa, sb;
      ghost_ab(D, side, ti, n, &ga, &gb);
      double *g = dst + NV * ((size_t)gb * D->lx + ga);
      switch (kind) {
        case K_SYNC: {
          inner_ab(S, opp[side], ti, n, &sa, &sb);
          const double *q = src + NV * ((size_t)sb * S->lx + sa);
          for (int m = 0; m < NV; m++) g[m] = q[m];
        } break;
        case K_OPEN: {
          inner_ab(D, side, ti, 1, &sa, &sb);
          const double *q = dst + NV * ((size_t)sb * D->lx + sa);
          for (int m = 0; m < NV; m++) g[m] = q[m];
        } break;
        case K_WALL: {
          inner_ab(D, side, ti, n, &sa, &sb);
          const double *q = dst + NV * ((size_t)sb * D->lx + sa);
          for (int m = 0; m < NV; m++) g[m] = wall_sign[side][m] * q[m];
        } break;
        default:
          orc_initial_state(c, D->bx + ga * D->dx, D->by + gb * D->dy, t, g);
          break;
      }
    }
}

void orc_set_boundary(orc_ctx *c, int layer, double t)
{
  for (int k = 0; k < c->nb; k++)
    for (int s = 0; s < 4; s++) {
      int code = c->b[k].bc[s];
      int kind = kind_of(code);
      fill_edge(c, layer, k, kind == K_SYNC ? code - 1 : k, s, kind, t);
    }
  for (int k = 0; k < c->nb; k++)
    for (int s = 0; s < 4; s++) {
      int code = c->b[k].bc[s];
      if (code > 0) fill_edge(c, layer, k, code - 1, s, K_SYNC, t);
    }
}

/* ------------------------------------------------------------------ */
/* one stage of one block (Main.c:297-396)                             */
/* ------------------------------------------------------------------ */
static double update_block(orc_ctx *c, int k, double *New, const double *Data,
                           const double *Data0, double dt, int heun)
{
  const oblk *B = &c->b[k];
  int lx = B->lx, ly = B->ly;
  double *P = c->prim[k], *SX = c->sx[k], *SY = c->sy[k];
  double U0[NV], F[NV], G[NV], pl[NV], pr[NV];
  double cflx = 0, cfly = 0, cflp = 0, cf;
  double dtdx = dt / B->dx, dtdy = dt / B->dy;

  for (int a = 0; a < lx; a++)
    for (int b = 0; b < ly; b++) {
      size_t q = (size_t)b * lx + a;
      orc_primitive(Data + NV * q, P + NV * q);
    }

  for (int a = 1; a <= lx - 2; a++)
    for (int b = 1; b <= ly - 2; b++) {
      size_t q = (size_t)b * lx + a;
      orc_relax(c, Data + NV * q, U0, dt, 2 * heun + 1);
      orc_primitive(U0, P + NV * q);
      cflp = mx(cflp, orc_relax(c, Data + NV * q, New + NV * q, dt, 2 * heun));
      if (heun == 1) {
        orc_relax(c, Data0 + NV * q, U0, dt, -1);
        for (int n = 0; n < NV; n++) New[NV * q + n] += U0[n];
      }
      for (int n = 0; n < NV; n++) { SX[NV * q + n] = 0.0; SY[NV * q + n] = 0.0; }
    }

  for (int a = 1; a <= lx - 2; a++)
    for (int b = 1; b <= ly - 2; b++) {
      size_t q = (size_t)b * lx + a;
      const double *p0 = P + NV * q;
      const double *pxm = p0 - NV, *pxp = p0 + NV;
      const double *pym = p0 - NV * lx, *pyp = p0 + NV * lx;
      for (int n = 0; n < NV; n++) {
        SX[NV * q + n] = orc_limit(c->limiter, p0[n] - pxm[n], pxp[n] - p0[n]) / B->dx;
        SY[NV * q + n] = orc_limit(c->limiter, p0[n] - pym[n], pyp[n] - p0[n]) / B->dy;
      }
    }

  for (int a = 1; a <= lx - 3; a++)
    for (int b = 1; b <= ly - 3; b++) {
      size_t q = (size_t)b * lx + a;
      size_t qx = q + 1, qy = q + lx;
      for (int n = 0; n < NV; n++) {
        pl[n] = P[NV * q + n] + 0.5 * B->dy * SY[NV * q + n];
        pr[n] = P[NV * qy + n] - 0.5 * B->dy * SY[NV * qy + n];
      }
      cf = orc_hll(pl, pr, B->dy, 2, G);
      cfly = mx(cf, cfly);
      for (int n = 0; n < NV; n++) {
        New[NV * q + n] -= dtdy * G[n];
        New[NV * qy + n] += dtdy * G[n];
      }
      for (int n = 0; n < NV; n++) {
        pl[n] = P[NV * q + n] + 0.5 * B->dx * SX[NV * q + n];
        pr[n] = P[NV * qx + n] - 0.5 * B->dx * SX[NV * qx + n];
      }
      cf = orc_hll(pl, pr, B->dx, 1, F);
      cflx = mx(cf, cflx);
      for (int n = 0; n < NV; n++) {
        New[NV * q + n] -= dtdx * F[n];
        New[NV * qx + n] += dtdx * F[n];
      }
    }

  if (heun == 1) {
    size_t n = (size_t)lx * ly * NV;
    for (size_t i = 0; i < n; i++) New[i] = 0.5 * New[i];
  }
  return dt * cflx / B->dx + dt * cfly / B->dy + cflp;
}

double orc_update_block(orc_ctx *c, int blk, double dt, double t, int heun)
{
  (void)t;
  if (heun == 0)
    return update_block(c, blk, c->lay[ORC_L_TMP][blk], c->lay[ORC_L_DATA][blk],
                        c->lay[ORC_L_DATA][blk], dt, 0);
  return update_block(c, blk, c->lay[ORC_L_NEW][blk], c->lay[ORC_L_TMP][blk],
                      c->lay[ORC_L_DATA][blk], dt, 1);
}

/* TimeUpdate (Main.c:253-284) */
double orc_time_update(orc_ctx *c, double dt, double t)
{
  double cfl1 = 0.0, cfl2 = 0.0, cflmax = 0.0;
  int k;
#pragma omp parallel for schedule(dynamic) reduction(max:cfl1)
  for (k = 0; k < c->nb; k++) {
    double v = update_block(c, k, c->lay[ORC_L_TMP][k], c->lay[ORC_L_DATA][k],
                            c->lay[ORC_L_DATA][k], dt, 0);
    if (v > cfl1) cfl1 = v;
  }
  orc_set_boundary(c, ORC_L_TMP, t + dt);
#pragma omp parallel for schedule(dynamic) reduction(max:cfl2)
  for (k = 0; k < c->nb; k++) {
    double v = update_block(c, k, c->lay[ORC_L_NEW][k], c->lay[ORC_L_TMP][k],
                            c->lay[ORC_L_DATA][k], dt, 1);
    if (v > cfl2) cfl2 = v;
  }
  orc_set_boundary(c, ORC_L_NEW, t + dt);
  return mx(cfl1, mx(cfl2, cflmax));
}

void orc_commit(orc_ctx *c)
{
  int k;
#pragma omp parallel for schedule(dynamic)
  for (k = 0; k < c->nb; k++)
    memcpy(c->lay[ORC_L_DATA][k], c->lay[ORC_L_NEW][k],
           sizeof(double) * NV * orc_block_cells(c, k));
}

/* Steps (Main.c:189-244); DT_INITIAL = 1e-4 (Main.c:28) on every call */
int orc_steps(orc_ctx *c, double *t, double tend, double cflmax,
              double cflreg, int maxsteps, double *log, int logcap,
              int *nlog, int *nrej)
{
  int nstep = 0, rej = 0, nl = 0;
  double cfl, dt = 0.0001;
  while (*t < tend) {
    if (maxsteps > 0 && nstep >= maxsteps) break;
    nstep++;
    do {
      double dt_used = dt;
      cfl = orc_time_update(c, dt, *t);
      int acc = !(cfl > cflmax);
      if (!acc) {
        dt = dt * cflreg / cfl;
        rej++;
      } else {
        *t += dt;
        dt = dt * cflreg / cfl;
        if (*t + dt > tend) dt = tend - (*t);
      }
      if (log && nl < logcap) {
        log[4 * nl + 0] = cfl;
        log[4 * nl + 1] = dt_used;
        log[4 * nl + 2] = *t;
        log[4 * nl + 3] = acc;
      }
      nl++;
    } while (cfl > cflmax);
    orc_commit(c);
  }
  if (nlog) *nlog = nl;
  if (nrej) *nrej = rej;
  return nstep;
}

/* Main.c:126-138: interior cells from the scenario, then the ghost fill */
void orc_init(orc_ctx *c)
{
  for (int k = 0; k < c->nb; k++) {
    oblk *B = &c->b[k];
    double *D = c->lay[ORC_L_DATA][k];
    memset(D, 0, sizeof(double) * NV * orc_block_cells(c, k));
    for (int a = MBC; a < B->lx - MBC; a++)
      for (int b = MBC; b < B->ly - MBC; b++)
        orc_initial_state(c, B->bx + a * B->dx, B->by + b * B->dy, 0.0,
                          D + NV * ((size_t)b * B->lx + a));
  }
  orc_set_boundary(c, ORC_L_DATA, 0.0);
}

/* Output.c:217-243 */
void orc_output_values(const orc_ctx *c, int blk, double *vals)
{
  const oblk *B = &c->b[blk];
  const double *D = c->lay[ORC_L_DATA][blk];
  double P[NV];
  size_t p = 0;
  for (int a = MBC - 1; a <= B->lx - MBC; a++)
    for (int b = MBC - 1; b <= B->ly - MBC; b++) {
      orc_primitive(D + NV * ((size_t)b * B->lx + a), P);
      double *v = vals + 10 * p++;
      v[0] = P[0]; v[1] = P[1]; v[2] = P[2]; v[3] = P[3];
      v[4] = P[0] * P[3];
      v[5] = P[0] * P[4]; v[6] = P[0] * P[5]; v[7] = P[0] * P[6];
      v[8] = P[7]; v[9] = P[8];
    }
}
