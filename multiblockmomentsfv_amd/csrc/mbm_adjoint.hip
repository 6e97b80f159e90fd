/*
 * mbm_adjoint.hip -- discrete adjoint of the Heun step (no reference
 * counterpart): reverse-mode derivative of TimeUpdate with a frozen dt
 * schedule, with respect to the state and to Kn.
 *
 * Forward:  T = fill(S1(D)),  N = fill(S2(T, D)).
 * Reverse of one step, given gN (adjoint of N, ghosts included):
 *   fill^T gN; gT, gD <- S2^T gN; fill^T gT; gD += S1^T gT.
 * fill^T folds each ghost's adjoint into its source cell (with the wall
 * signs) and clears the ghosts; it is a gather over the inverse ghost map,
 * so the sums are deterministic.
 *
 * S^T is split by direction:
 *   adj_y : every lane marches up one column over a run of rows (64 columns
 *           of one run per wavefront, or a block's last columns with several
 *           runs packed into one wavefront), with the primitives of four rows,
 *           the face state, its adjoint and gout of the current row in
 *           lane-private LDS and the partial adjoints of two rows in
 *           registers; per row: VJP of the y-face and of the y-limiter. The
 *           y-part of the adjoint of the stage primitives, gP_y, is written
 *           out;
 *   adj_x : row-local; one wavefront per 58 owned columns (3 halo lanes per
 *           side), lane shifts for neighbours. Per row: primitives, x-slopes,
 *           VJP of the x-faces and of the x-limiter, then the VJP of the
 *           projection of gP_x + gP_y and of the initial value, written to gin
 *           (and gin0 in stage 2). Every cell is finalised by exactly one lane.
 * Both sweeps load row r+1 while row r is computed. Without a body force the
 * initial value of stage 1 and the DATA part of stage 2's are the same
 * function of DATA; their VJPs are taken once, in stage 1.
 * dJ/dKn: one partial sum per adj_x wavefront, summed on the host in a fixed
 * order. Checkpointing: DATA is stored every 'stride' steps during the
 * forward run; the reverse sweep recomputes one segment at a time and keeps
 * its DATA and stage states (memory: N/stride + 2 x (2*stride - 1) layers,
 * kept between runs until mbm_adj_release). Segment s-1 is recomputed on a
 * second stream while segment s is reversed.
 * Resident recording (stride 0 with a given schedule, when the memory
 * budget holds N + 2 layers): the forward run writes the DATA state of every
 * step, and the stage state of the last steps as far as the budget goes,
 * into layers of their own (no copies; a step is the same launches as in
 * the checkpointed run, so the results are bitwise the same). The reverse
 * sweep then recomputes only stage 1 of the steps whose stage state was not
 * kept, on the second stream, one step ahead. At 4 x 512^2 (77 MB a layer)
 * a 2199-step run keeps its 2199 DATA states and about 1200 stage states in
 * 288 GB.
 */
#include "mbm_internal.h"
#include "mbm_dev.h"
#include "mbm_adj.h"
#include "mbm_adj_stage.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <tuple>
#include <type_traits>

using namespace mbmi;
using namespace mbm;

/* inverse ghost map for fill^T: element offsets of component 0 in a layer */
struct FoldGhost {
  long long off;
  int plane, sign;      /* sign: bit m set, component m negated */
};
struct FoldSrc {
  long long off;
  int plane, j0, j1;    /* ghosts j0..j1-1 copy this cell; j0 is inlined below */
  int pad;
  FoldGhost first;
};

/* one adj_y wavefront: lane l marches column a0 + l % ncols over the rows
 * r0 + (l / ncols) * R .. + R - 1 (R = AdjArgs::yrows), clipped to the
 * block's rows 1..ly-2, for l / ncols < ngroups */
struct YTileDev {
  int blk, a0, ncols, r0, ngroups, pad;
};

struct mbm_adj_state {
  int stride = 0;
  std::vector<double> dts;        /* accepted dt of every step */
  /* layers are allocated on first need and kept until mbm_adj_release */
  std::vector<double *> ckpt;     /* ckpt[i], i < nckpt: DATA after i*stride steps */
  int nckpt = 0;
  std::vector<double *> seg;      /* segment states during the reverse sweep */
  /* resident recording: ckpt[k] is the DATA state of step k (stride 1), and
   * step k >= kept_from has its stage state in tkeep[k - kept_from] */
  bool resident = false;
  int kept_from = 0;
  std::vector<double *> tkeep;
  std::vector<double *> spare;    /* held layers no pool uses at present */
  double *g[3] = { nullptr, nullptr, nullptr };   /* gN, gD, gT */
  double *gpy = nullptr;          /* y-part of gP of the current stage */
  TileDev *xtiles = nullptr;
  YTileDev *ytiles = nullptr;
  int nxt = 0, nyt = 0, yrows = 0;
  unsigned long long *slots = nullptr;
  double *knpart = nullptr;
  int nknpart = 0;
  /* the recomputation of the next segment runs on its own stream while the
   * current one is reversed on the context's stream */
  hipStream_t rstream = nullptr;
  hipEvent_t ev_start = nullptr, ev_ready[2] = { nullptr, nullptr }, ev_done[2] = { nullptr, nullptr };
  FoldSrc *fold_src = nullptr;     /* cells that ghosts copy, with their ghost ranges */
  FoldGhost *fold_ghost = nullptr; /* ghosts grouped by source cell */
  FoldGhost *fold_dir = nullptr;   /* Dirichlet ghosts (no source) */
  int nsrc = 0, ndir = 0;
};

namespace {
constexpr int ADJ_THREADS = 256;
}

/* arguments of the direction-split stage adjoint */
struct AdjArgs {
  const double *in, *in0, *gout;
  double *gin, *gin0;
  const double *gprev;                /* stage 1 without body force: stage 2's gout */
  double *gpy;                        /* y-part of the adjoint of the stage primitives */
  const BlockDev *blocks;
  const TileDev *tiles;               /* adj_x */
  const YTileDev *ytiles;             /* adj_y */
  int ntiles;
  int yrows;                          /* adj_y: rows per lane */
  double dt, kn, force, scale;
  double *knpart;                     /* one slot per tile, accumulated over launches */
};

constexpr int ADJX_STRIP = 58;        /* owned columns of an adj_x wavefront */

MBM_DEV bool cell_interior(int a, int b, int lx, int ly)
{
  return a >= 3 && a <= lx - 4 && b >= 3 && b <= ly - 4;
}

/* x-direction part of S^T plus the VJP of the initial value, row by row.
 * The state and adjoint of row r+1 are loaded while row r is computed (two
 * register sets, alternating), so the loop waits on memory only when a
 * row's compute is shorter than the load latency. */
template <int STAGE, int LIM, bool MERGE>
__global__ void __launch_bounds__(64) adj_x(const AdjArgs A)
{
  const int lane = threadIdx.x & 63;
  const int tix = __builtin_amdgcn_readfirstlane((int)blockIdx.x);
  if (tix >= A.ntiles) return;
  const TileDev T = A.tiles[tix];
  if (T.blk < 0) return;
  const BlockDev B = A.blocks[T.blk];
  const int lx = B.lx, ly = B.ly;
  const int a = T.a0 + lane;
  const int ac = a < 0 ? 0 : (a > lx - 1 ? lx - 1 : a);
  const bool own = lane >= 3 && lane <= 60 && a >= 0 && a <= lx - 1;
  const long long nel = (long long)B.npos * NV * B.pitch;
  const __amdgpu_buffer_rsrc_t rin = block_rsrc(A.in + B.off, nel);
  const __amdgpu_buffer_rsrc_t rin0 = block_rsrc((STAGE == 2 ? A.in0 : A.in) + B.off, nel);
  const __amdgpu_buffer_rsrc_t rg = block_rsrc(A.gout + B.off, nel);
  const __amdgpu_buffer_rsrc_t rgi = block_rsrc(A.gin + B.off, nel);
  const __amdgpu_buffer_rsrc_t rgi0 = block_rsrc((STAGE == 2 ? A.gin0 : A.gin) + B.off, nel);
  const __amdgpu_buffer_rsrc_t rgo = block_rsrc((STAGE == 1 && MERGE ? A.gprev : A.gin) + B.off,
                                                nel);
  const __amdgpu_buffer_rsrc_t rgy = block_rsrc(A.gpy + B.off, nel);
  const int pb = (int)(8 * B.plane);
  const unsigned col = 8u * (unsigned)ac;
  /* byte offset of row x (its position in the block's row order) */
  auto rowb = [&](int x) { return 8u * (unsigned)row_elem(B, x); };
  const double dtdx = A.dt / B.dx, dtkn = A.dt / A.kn;
  double gkn = 0.0;

  auto fetch = [&](int r, double *U, double *U0, double *g) {
    const unsigned boff = col + rowb(r);
    load9b(rin, pb, boff, U);
    if constexpr (STAGE == 2 && !MERGE) load9b(rin0, pb, boff, U0);   /* DATA part's VJP */
    load9b(rg, pb, boff, g);
  };
  /* row r from (U, U0, gr); loads row rn into (Un, U0n, gn) */
  auto row = [&](int r, const double *U, const double *U0, const double *gr, int rn, double *Un,
                 double *U0n, double *gn) {
    const unsigned boff = col + rowb(r);
    double gy[NV], o[NV];
    load9b(rgy, pb, boff, gy);
    if constexpr (STAGE == 1) load9b(rgo, pb, boff, o);
    fetch(rn, Un, U0n, gn);
    const bool inner = a >= 1 && a <= lx - 2 && r >= 1 && r <= ly - 2;
    const bool ring3 = a <= 0 || a >= lx - 1 || r == 0 || r == ly - 1;
    const double w = cell_interior(a, r, lx, ly) ? A.scale : 0.0;
    double g[NV], P[NV], Q[NV], z, i1, i2;
#pragma unroll
    for (int n = 0; n < NV; n++) g[n] = gr[n] * w;
    /* the projection (as project<STAGE, false>), keeping Q, z and the damping
     * reciprocals for its VJP */
    primitive(U, Q);
    z = Q[0] * sqrtp(Q[3]) * dtkn;
    damp_inv(z, i1, i2);
    if constexpr (STAGE == 1) {
      relax(Q, damp1(z, i1, i2), P);
    } else {
      Damp d3;
      d3.d1 = i1;
      d3.d2 = i2;
      relax(Q, d3, P);
    }
    if (ring3) {
#pragma unroll
      for (int n = 0; n < NV; n++) P[n] = Q[n];
    }

    /* x-slopes, face states; the right state of face a+1/2 comes from lane a+1 */
    double d1[NV], d2[NV], L[NV], Rn[NV], gF[NV];
#pragma unroll
    for (int gq = 0; gq < NV; gq += 3) {
      double sx[3];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        d1[gq + k] = P[gq + k] - from_left(P[gq + k]);
        d2[gq + k] = from_right(P[gq + k]) - P[gq + k];
      }
      limit3<LIM>(d1 + gq, d2 + gq, sx);
#pragma unroll
      for (int k = 0; k < 3; k++) {
        L[gq + k] = P[gq + k] + 0.5 * sx[k];
        Rn[gq + k] = from_right(P[gq + k] - 0.5 * sx[k]);
      }
    }
#pragma unroll
    for (int n = 0; n < NV; n++) gF[n] = from_right(g[n]) - g[n];
    double gL[NV] = { 0 }, gRn[NV] = { 0 };
    face_vjp<0>(L, Rn, dtdx, gF, gL, gRn);

    /* gather: own left state, right state (from face a-1/2), slope VJPs of
     * the slopes at a-1, a, a+1 */
    double gP[NV];
#pragma unroll
    for (int gq = 0; gq < NV; gq += 3) {
      double gR[3], gs[3], gd1[3] = { 0, 0, 0 }, gd2[3] = { 0, 0, 0 };
#pragma unroll
      for (int j = 0; j < 3; j++) {
        gR[j] = from_left(gRn[gq + j]);
        gs[j] = 0.5 * gL[gq + j] - 0.5 * gR[j];
      }
      limit_vjp3<LIM>(d1 + gq, d2 + gq, gs, gd1, gd2);
#pragma unroll
      for (int j = 0; j < 3; j++) {
        const int n = gq + j;
        gP[n] = gL[n] + gR[j] + gd1[j] - gd2[j] - from_right(gd1[j]) + from_left(gd2[j]);
      }
    }
    /* plus the y-part computed by adj_y */
#pragma unroll
    for (int n = 0; n < NV; n++) gP[n] = inner ? gP[n] + gy[n] : 0.0;
    /* MERGE (no body force): stage 1's initial value and the DATA part of
     * stage 2's are the same function of DATA, so their VJPs merge: stage 2
     * skips its DATA part and stage 1 adds that part's cotangent, stage 2's
     * weighted gout (1/2 gprev on interior cells), to its own. */
    if (own) {
      double gU[NV] = { 0 }, gU0[NV] = { 0 };
      if constexpr (STAGE == 1) {
        if constexpr (MERGE) {
          const double w2 = cell_interior(a, r, lx, ly) ? 0.5 : 0.0;
#pragma unroll
          for (int n = 0; n < NV; n++) g[n] += o[n] * w2;
        }
        if (inner) gkn += proj_vjp_q<1>(U, Q, z, i1, i2, U0, gP, g, A.dt, A.kn, A.force, gU, gU0);
#pragma unroll
        for (int n = 0; n < NV; n++) store1b(rgi, n * pb, boff, MERGE ? gU[n] : o[n] + gU[n]);
      } else {
        if (inner)
          gkn += proj_vjp_q<2, !MERGE>(U, Q, z, i1, i2, U0, gP, g, A.dt, A.kn, A.force, gU, gU0);
#pragma unroll
        for (int n = 0; n < NV; n++) {
          store1b(rgi, n * pb, boff, gU[n]);
          if constexpr (!MERGE) store1b(rgi0, n * pb, boff, gU0[n]);
        }
      }
    }
  };

  double Ua[NV], U0a[NV], ga[NV], Ub[NV], U0b[NV], gb[NV];
  fetch(T.r0, Ua, U0a, ga);
  for (int r = T.r0; r <= T.r1; r += 2) {
    /* the last row's look-ahead reloads that row (in bounds, unused) */
    row(r, Ua, U0a, ga, min(r + 1, T.r1), Ub, U0b, gb);
    if (r + 1 <= T.r1) row(r + 1, Ub, U0b, gb, min(r + 2, T.r1), Ua, U0a, ga);
  }
  gkn = wave_sum(gkn);
  if (lane == 0) A.knpart[tix] += gkn;
}

/* y-direction part of S^T: every lane marches up one column over a run of
 * rows; the lanes are independent, so a wavefront may hold 64 columns of one
 * run or fewer columns of several runs (a block's last columns) */
template <int STAGE, int LIM>
__global__ void __launch_bounds__(64) adj_y(const AdjArgs A)
{
  const int lane = threadIdx.x & 63;
  const int tix = __builtin_amdgcn_readfirstlane((int)blockIdx.x);
  if (tix >= A.ntiles) return;
  const YTileDev T = A.ytiles[tix];
  const BlockDev B = A.blocks[T.blk];
  const int lx = B.lx, ly = B.ly;
  const int R = A.yrows;
  const int grp = lane / T.ncols;
  const int a = T.a0 + (lane - grp * T.ncols);
  const int rs = T.r0 + grp * R;                 /* this lane's rows rs..re */
  const int re = min(rs + R - 1, ly - 2);
  const bool act = grp < T.ngroups && a <= lx - 2 && rs <= ly - 2;
  const int ac = a > lx - 1 ? lx - 1 : a;
  const bool colin = act && a >= 1;
  const bool colr3 = a == 0 || a >= lx - 1;
  const long long nel = (long long)B.npos * NV * B.pitch;
  const __amdgpu_buffer_rsrc_t rin = block_rsrc(A.in + B.off, nel);
  const __amdgpu_buffer_rsrc_t rg = block_rsrc(A.gout + B.off, nel);
  const __amdgpu_buffer_rsrc_t rgy = block_rsrc(A.gpy + B.off, nel);
  const int pb = (int)(8 * B.plane);
  const unsigned col = 8u * (unsigned)ac;
  /* byte offset of row x (its position in the block's row order) */
  auto rowb = [&](int x) { return 8u * (unsigned)row_elem(B, x); };
  const double dtdy = A.dt / B.dy, dtkn = A.dt / A.kn;
  auto rowc = [&](int r) { return r < 0 ? 0 : (r > ly - 1 ? ly - 1 : r); };
  auto prims = [&](int r, double *P) {
    const int rr = rowc(r);
    double U[NV], dummy[NV], cp;
    load9b(rin, pb, col + rowb(rr), U);
    project<STAGE, false>(U, U, colr3 || rr == 0 || rr == ly - 1, A.dt, dtkn, A.force, P, dummy,
                          cp);
  };
  auto gweighted = [&](int r, double *g) {
    const int rr = rowc(r);
    load9b(rg, pb, col + rowb(rr), g);
    const double w = cell_interior(a, r, lx, ly) ? A.scale : 0.0;
#pragma unroll
    for (int n = 0; n < NV; n++) g[n] *= w;
  };

  /* lane-private LDS rows: the stage primitives of rows k-1..k+2 (a ring
   * indexed by the iteration), the upper face state of row k, the adjoint of
   * its lower face state and its weighted gout. Every slot is read before it
   * is rewritten, by the same lane. Iteration it handles row k = rs - 2 + it. */
  __shared__ double sm[7][NV][64];
  auto Prow = [&](int j) -> double(&)[NV][64] { return sm[(j + 8) & 3]; };   /* row rs-2+j */
  double(&Tk)[NV][64] = sm[4];
  double(&gBk)[NV][64] = sm[5];
  double(&gk)[NV][64] = sm[6];
  /* partial gP of rows k-1 and k */
  double gPm[NV], gP0[NV];
  const int rb = rs - 2;

  /* state at k = rs-2 */
  {
    double Pm[NV], Pk[NV], Pk1[NV];
    prims(rb - 1, Pm);
    prims(rb, Pk);
    prims(rb + 1, Pk1);
#pragma unroll
    for (int gq = 0; gq < NV; gq += 3) {
      double d1[3], d2[3], s3[3];
#pragma unroll
      for (int j = 0; j < 3; j++) {
        d1[j] = Pk[gq + j] - Pm[gq + j];
        d2[j] = Pk1[gq + j] - Pk[gq + j];
      }
      limit3<LIM>(d1, d2, s3);
#pragma unroll
      for (int j = 0; j < 3; j++) Tk[gq + j][lane] = Pk[gq + j] + 0.5 * s3[j];
    }
    double g0[NV];
    gweighted(rb, g0);
#pragma unroll
    for (int n = 0; n < NV; n++) {
      Prow(-1)[n][lane] = Pm[n];
      Prow(0)[n][lane] = Pk[n];
      Prow(1)[n][lane] = Pk1[n];
      gk[n][lane] = g0[n];
      gBk[n][lane] = 0.0;
      gPm[n] = gP0[n] = 0.0;
    }
  }
  /* inputs of iteration it (the state of row k+2, gout of row k+1) are
   * loaded one iteration ahead, into two alternating register sets */
  auto fetch = [&](int it, double *U, double *g) {
    load9b(rin, pb, col + rowb(rowc(rb + it + 2)), U);
    load9b(rg, pb, col + rowb(rowc(rb + it + 1)), g);
  };
  auto step = [&](int it, const double *Uc, const double *gc, double *Un, double *gn) {
    const int k = rb + it;
    double(&Pk)[NV][64] = Prow(it);
    double(&Pk1)[NV][64] = Prow(it + 1);
    double(&P2)[NV][64] = Prow(it + 2);   /* slot of row k-2, no longer read */
    /* face k+1/2: the upper state of row k (replaced below by that of row
     * k+1) and the gout difference */
    double Tl[NV], gG[NV];
    {
      const int rr = rowc(k + 2);
      double P[NV], dummy[NV], cp;
      project<STAGE, false>(Uc, Uc, colr3 || rr == 0 || rr == ly - 1, A.dt, dtkn, A.force, P,
                            dummy, cp);
      const double w = cell_interior(a, k + 1, lx, ly) ? A.scale : 0.0;
#pragma unroll
      for (int n = 0; n < NV; n++) {
        P2[n][lane] = P[n];
        Tl[n] = Tk[n][lane];
        const double g1 = gc[n] * w;
        gG[n] = g1 - gk[n][lane];
        gk[n][lane] = g1;
      }
    }
    /* slope of row k+1 and its face states */
    double B1[NV];
#pragma unroll
    for (int gq = 0; gq < NV; gq += 3) {
      double e1[3], e2[3], s3[3];
#pragma unroll
      for (int j = 0; j < 3; j++) {
        e1[j] = Pk1[gq + j][lane] - Pk[gq + j][lane];
        e2[j] = P2[gq + j][lane] - Pk1[gq + j][lane];
      }
      limit3<LIM>(e1, e2, s3);
#pragma unroll
      for (int j = 0; j < 3; j++) {
        Tk[gq + j][lane] = Pk1[gq + j][lane] + 0.5 * s3[j];
        B1[gq + j] = Pk1[gq + j][lane] - 0.5 * s3[j];
      }
    }
    double gT[NV] = { 0 }, gB1[NV] = { 0 };
    face_vjp<1>(Tl, B1, dtdy, gG, gT, gB1);
    fetch(it + 1, Un, gn);
    /* slope of row k: T_k = P_k + s/2, B_k = P_k - s/2 */
    double(&Pm)[NV][64] = Prow(it - 1);
    double gP1[NV];
#pragma unroll
    for (int gq = 0; gq < NV; gq += 3) {
      double d1[3], d2[3], gs[3], gd1[3] = { 0, 0, 0 }, gd2[3] = { 0, 0, 0 };
#pragma unroll
      for (int j = 0; j < 3; j++) {
        d1[j] = Pk[gq + j][lane] - Pm[gq + j][lane];
        d2[j] = Pk1[gq + j][lane] - Pk[gq + j][lane];
        gs[j] = 0.5 * gT[gq + j] - 0.5 * gBk[gq + j][lane];
      }
      limit_vjp3<LIM>(d1, d2, gs, gd1, gd2);
#pragma unroll
      for (int j = 0; j < 3; j++) {
        const int n = gq + j;
        gP0[n] += gT[n] + gBk[n][lane] + gd1[j] - gd2[j];
        gPm[n] -= gd1[j];
        gP1[n] = gd2[j];
      }
    }
    /* row k-1 is complete: its y-part goes to gpy for adj_x */
    const int rf = k - 1;
    if (colin && rf >= rs && rf <= re) {
      const unsigned boff = col + rowb(rf);
#pragma unroll
      for (int n = 0; n < NV; n++) store1b(rgy, n * pb, boff, gPm[n]);
    }
#pragma unroll
    for (int n = 0; n < NV; n++) {
      gBk[n][lane] = gB1[n];
      gPm[n] = gP0[n];
      gP0[n] = gP1[n];
    }
  };
  double Ua[NV], ga[NV], Ub[NV], gb[NV];
  const int nit = R + 3;
  fetch(0, Ua, ga);
  for (int it = 0; it < nit; it += 2) {
    step(it, Ua, ga, Ub, gb);
    if (it + 1 < nit) step(it + 1, Ub, gb, Ua, ga);
  }
}


/* fill^T in one pass, one thread per (source cell, component m =
 * blockIdx.y): the source gathers sign * g(ghost) over the ghost cells that
 * copy it, in a fixed order, and clears them (every ghost copies exactly one
 * interior cell, so the reads and writes of different threads never meet);
 * the threads past nsrc clear the Dirichlet ghosts, which copy nothing. */
__global__ void __launch_bounds__(ADJ_THREADS) adj_fill(double *g, const FoldSrc *src, int nsrc,
                                                        const FoldGhost *ghost,
                                                        const FoldGhost *dir, int ndir)
{
  const int m = blockIdx.y;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsrc) {
    if (s - nsrc >= ndir) return;
    const FoldGhost d = dir[s - nsrc];
    g[d.off + (long long)m * d.plane] = 0.0;
    return;
  }
  const FoldSrc S = src[s];
  /* the source and its first ghost load together (ghosts are never sources) */
  double *ps = g + S.off + (long long)m * S.plane;
  double *p = g + S.first.off + (long long)m * S.first.plane;
  const double vs = *ps, v = *p;
  double acc = ((S.first.sign >> m) & 1) ? -v : v;
  *p = 0.0;
  for (int j = S.j0 + 1; j < S.j1; j++) {
    const FoldGhost q = ghost[j];
    double *pq = g + q.off + (long long)m * q.plane;
    const double w = *pq;
    acc += ((q.sign >> m) & 1) ? -w : w;
    *pq = 0.0;
  }
  *ps = vs + acc;
}

/* ------------------------------------------------------------------ */
namespace {

int alloc_layer(mbm_ctx *c, double **p)
{
  if (*p) return MBM_OK;
  if (hipMalloc(p, c->layer_elems * sizeof(double)) != hipSuccess) {
    *p = nullptr;
    return set_err(MBM_E_NOMEM, "adjoint: device allocation of %zu bytes failed",
                   c->layer_elems * sizeof(double));
  }
  HIPCHK(hipMemsetAsync(*p, 0, c->layer_elems * sizeof(double), c->stream));
  return MBM_OK;
}

int adj_setup(mbm_ctx *c)
{
  if (c->adj) return MBM_OK;
  mbm_adj_state *A = new mbm_adj_state();
  c->adj = A;
  int rc;
  for (int k = 0; k < 3; k++) if ((rc = alloc_layer(c, &A->g[k]))) return rc;
  if ((rc = alloc_layer(c, &A->gpy))) return rc;
  HIPCHK(hipMalloc(&A->slots, (size_t)2 * SLOTS_PER_BLOCK * c->nb * sizeof(unsigned long long)));
  HIPCHK(hipMemsetAsync(A->slots, 0, (size_t)2 * SLOTS_PER_BLOCK * c->nb * sizeof(unsigned long long), c->stream));
  /* adj_x: 58 owned columns per wavefront (lanes 3..60), every row of the
   * block. adj_y: one column per lane over R rows of 1..ly-2, 64 columns per
   * wavefront, a block's remaining columns packed several row runs to a
   * wavefront. Both run one wavefront per SIMD; the rows per wavefront
   * minimise rounds x (rows + warm-up) over the device's wave slots (adj_y
   * spends three rows of warm-up per run). */
  const int slots = c->cus * 4;
  auto xcount = [&](int R) {
    long long n = 0;
    for (int k = 0; k < c->nb; k++) {
      const BlockDev &B = c->hblk[k];
      n += (long long)((B.lx + ADJX_STRIP - 1) / ADJX_STRIP) * ((B.ly + R - 1) / R);
    }
    return n;
  };
  auto ycount = [&](int R) {
    long long n = 0;
    for (int k = 0; k < c->nb; k++) {
      const BlockDev &B = c->hblk[k];
      const int cols = B.lx - 2, runs = (B.ly - 2 + R - 1) / R, rem = cols % 64;
      n += (long long)(cols / 64) * runs;
      if (rem) n += (runs + 64 / rem - 1) / (64 / rem);
    }
    return n;
  };
  auto pick_rows = [&](auto count, int maxrows, int warm) {
    int best = 8;
    double bestc = 1e300;
    for (int R = 4; R <= 256; R++) {
      const double cost = (double)((count(R) + slots - 1) / slots) * (std::min(R, maxrows) + warm);
      if (cost < bestc - 1e-9) {
        bestc = cost;
        best = R;
      }
    }
    return best;
  };
  int maxly = 0;
  for (int k = 0; k < c->nb; k++) maxly = std::max(maxly, c->hblk[k].ly);
  const int rx = pick_rows(xcount, maxly, 0), ry = pick_rows(ycount, maxly - 2, 3);
  std::vector<TileDev> xt;
  std::vector<YTileDev> yt;
  for (int k = 0; k < c->nb; k++) {
    const BlockDev &B = c->hblk[k];
    for (int r0 = 0; r0 <= B.ly - 1; r0 += rx)
      for (int a0 = -3; a0 + 3 <= B.lx - 1; a0 += ADJX_STRIP)
        xt.push_back(TileDev{ k, a0, r0, std::min(r0 + rx - 1, B.ly - 1) });
    const int cols = B.lx - 2, runs = (B.ly - 2 + ry - 1) / ry, rem = cols % 64;
    for (int j = 0; j < runs; j++)
      for (int a0 = 1; a0 + 63 <= B.lx - 2; a0 += 64)
        yt.push_back(YTileDev{ k, a0, 64, 1 + j * ry, 1, 0 });
    if (rem) {
      const int g = 64 / rem;
      for (int j = 0; j < runs; j += g)
        yt.push_back(YTileDev{ k, 1 + 64 * (cols / 64), rem, 1 + j * ry, std::min(g, runs - j), 0 });
    }
  }
  A->yrows = ry;
  A->nxt = (int)xt.size();
  A->nyt = (int)yt.size();
  HIPCHK(hipMalloc(&A->xtiles, xt.size() * sizeof(TileDev)));
  HIPCHK(hipMalloc(&A->ytiles, yt.size() * sizeof(YTileDev)));
  HIPCHK(hipMemcpy(A->xtiles, xt.data(), xt.size() * sizeof(TileDev), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(A->ytiles, yt.data(), yt.size() * sizeof(YTileDev), hipMemcpyHostToDevice));
  A->nknpart = A->nxt;
  HIPCHK(hipMalloc(&A->knpart, A->nknpart * sizeof(double)));
  /* inverse ghost map: CSR over source cells */
  std::vector<GhostDev> gm = build_ghost_map(c->desc, c->hblk);
  /* source cells in (block, row, column) order */
  std::map<std::tuple<int, int, int>, std::vector<int>> inv;
  std::vector<int> dir;
  for (int i = 0; i < (int)gm.size(); i++) {
    if (gm[i].sblk >= 0) inv[{ gm[i].sblk, gm[i].sb, gm[i].sa }].push_back(i);
    else dir.push_back(i);
  }
  auto ghost_at = [&](int i) {
    const GhostDev &q = gm[i];
    const BlockDev &D = c->hblk[q.dblk];
    return FoldGhost{ cell_elem(D, q.db, q.da), (int)D.plane, q.sign };
  };
  std::vector<FoldGhost> fd, fg;
  std::vector<FoldSrc> fs;
  for (int i : dir) fd.push_back(ghost_at(i));
  for (auto &kv : inv) {
    const BlockDev &S = c->hblk[std::get<0>(kv.first)];
    FoldSrc src{ cell_elem(S, std::get<1>(kv.first), std::get<2>(kv.first)), (int)S.plane,
                 (int)fg.size(), 0, 0, ghost_at(kv.second[0]) };
    for (int i : kv.second) fg.push_back(ghost_at(i));
    src.j1 = (int)fg.size();
    fs.push_back(src);
  }
  A->ndir = (int)fd.size();
  A->nsrc = (int)fs.size();
  auto upload = [&](auto **dst, const auto &v) -> int {
    using T = typename std::decay_t<decltype(v)>::value_type;
    HIPCHK(hipMalloc(dst, std::max<size_t>(1, v.size()) * sizeof(T)));
    if (!v.empty())
      HIPCHK(hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return MBM_OK;
  };
  if ((rc = upload(&A->fold_src, fs)) || (rc = upload(&A->fold_ghost, fg)) ||
      (rc = upload(&A->fold_dir, fd)))
    return rc;
  HIPCHK(hipStreamCreateWithFlags(&A->rstream, hipStreamNonBlocking));
  HIPCHK(hipEventCreateWithFlags(&A->ev_start, hipEventDisableTiming));
  for (int q = 0; q < 2; q++) {
    HIPCHK(hipEventCreateWithFlags(&A->ev_ready[q], hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&A->ev_done[q], hipEventDisableTiming));
  }
  return MBM_OK;
}

int fill_adjoint(mbm_ctx *c, double *g)
{
  mbm_adj_state *A = c->adj;
  const int n = A->nsrc + A->ndir;
  if (n > 0)
    hipLaunchKernelGGL(adj_fill, dim3((n + ADJ_THREADS - 1) / ADJ_THREADS, NV), dim3(ADJ_THREADS),
                       0, c->stream, g, A->fold_src, A->nsrc, A->fold_ghost, A->fold_dir, A->ndir);
  HIPCHK(hipGetLastError());
  return MBM_OK;
}

template <int STAGE, int LIM>
int stage_adjoint_lim(mbm_ctx *c, const double *in, const double *in0, const double *gout,
                      double *gin, double *gin0, const double *gprev, double dt)
{
  mbm_adj_state *A = c->adj;
  AdjArgs X{};
  X.in = in;
  X.in0 = in0;
  X.gout = gout;
  X.gin = gin;
  X.gin0 = gin0;
  X.gprev = gprev;
  X.blocks = c->dblk;
  X.dt = dt;
  X.kn = c->par.kn;
  X.force = c->par.force;
  X.scale = STAGE == 2 ? 0.5 : 1.0;
  X.knpart = A->knpart;
  X.gpy = A->gpy;
  AdjArgs Y = X;
  Y.ytiles = A->ytiles;
  Y.ntiles = A->nyt;
  Y.yrows = A->yrows;
  hipLaunchKernelGGL((adj_y<STAGE, LIM>), dim3(A->nyt), dim3(64), 0, c->stream, Y);
  X.tiles = A->xtiles;
  X.ntiles = A->nxt;
  if (X.force == 0.0)
    hipLaunchKernelGGL((adj_x<STAGE, LIM, true>), dim3(A->nxt), dim3(64), 0, c->stream, X);
  else
    hipLaunchKernelGGL((adj_x<STAGE, LIM, false>), dim3(A->nxt), dim3(64), 0, c->stream, X);
  HIPCHK(hipGetLastError());
  return MBM_OK;
}

template <int STAGE>
int stage_adjoint(mbm_ctx *c, const double *in, const double *in0, const double *gout,
                  double *gin, double *gin0, const double *gprev, double dt)
{
  switch (c->par.limiter) {
    case 0: return stage_adjoint_lim<STAGE, 0>(c, in, in0, gout, gin, gin0, gprev, dt);
    case 1: return stage_adjoint_lim<STAGE, 1>(c, in, in0, gout, gin, gin0, gprev, dt);
    case 2: return stage_adjoint_lim<STAGE, 2>(c, in, in0, gout, gin, gin0, gprev, dt);
    case 3: return stage_adjoint_lim<STAGE, 3>(c, in, in0, gout, gin, gin0, gprev, dt);
    default: return stage_adjoint_lim<STAGE, 4>(c, in, in0, gout, gin, gin0, gprev, dt);
  }
}

/* stage 1 of a step with its ghost fill: D -> T */
int stage1_buffers(mbm_ctx *c, const double *D, double *T, double dt)
{
  mbm_adj_state *A = c->adj;
  int rc;
  if ((rc = enqueue_stage(c, 1, D, D, T, dt, A->slots))) return rc;
  return fill_after_stage(c, T);
}

/* forward step between explicit buffers (frozen dt, no controller); the
 * intermediate stage state is left in T */
int step_buffers(mbm_ctx *c, const double *D, double *T, double *N, double dt)
{
  mbm_adj_state *A = c->adj;
  int rc;
  if ((rc = stage1_buffers(c, D, T, dt))) return rc;
  if ((rc = enqueue_stage(c, 2, T, D, N, dt, A->slots + SLOTS_PER_BLOCK * c->nb))) return rc;
  return fill_after_stage(c, N);
}

/* reverse of one step at DATA = D with stage state T: gN (in, destroyed)
 * -> gD (out) */
int step_adjoint(mbm_ctx *c, const double *D, const double *T, double dt, double *gN, double *gD)
{
  mbm_adj_state *A = c->adj;
  int rc;
  if ((rc = fill_adjoint(c, gN))) return rc;
  /* stage 2 writes every cell of gT and gD, stage 1 then adds to gD; without
   * a body force stage 2 leaves gD alone and stage 1 writes every cell of it,
   * taking the DATA part of stage 2's initial value from gN */
  if ((rc = stage_adjoint<2>(c, T, D, gN, A->g[2], gD, nullptr, dt))) return rc;
  if ((rc = fill_adjoint(c, A->g[2]))) return rc;
  return stage_adjoint<1>(c, D, D, A->g[2], gD, nullptr, gN, dt);
}

/* layer i of a pool: on first use taken from the spare layers, else
 * allocated (a failed allocation leaves no error state in the runtime) */
int pool_layer(mbm_ctx *c, std::vector<double *> &pool, int i, double **p)
{
  mbm_adj_state *A = c->adj;
  while ((int)pool.size() <= i) {
    double *q = nullptr;
    if (!A->spare.empty()) {
      q = A->spare.back();
      A->spare.pop_back();
    } else if (hipMalloc(&q, c->layer_elems * sizeof(double)) != hipSuccess) {
      (void)hipGetLastError();
      return set_err(MBM_E_NOMEM, "adjoint: allocation of checkpoint layer %d failed", i);
    }
    pool.push_back(q);
  }
  *p = pool[i];
  return MBM_OK;
}

/* a pool keeps its first n layers; the rest become spare */
void trim_pool(mbm_adj_state *A, std::vector<double *> &pool, size_t n)
{
  while (pool.size() > n) {
    A->spare.push_back(pool.back());
    pool.pop_back();
  }
}

/* the layers of a resident recording of N steps keeping the stage states of
 * the last *nkeep: N DATA states, two layers for recomputed stage states
 * (unless all are kept), then the kept stage states as far as the
 * allocations go (*nkeep lowered if they stop early). false when the DATA
 * states or the two layers cannot be had. */
bool reserve_resident(mbm_ctx *c, int N, int *nkeep)
{
  mbm_adj_state *A = c->adj;
  trim_pool(A, A->tkeep, (size_t)*nkeep);
  trim_pool(A, A->ckpt, (size_t)N);
  trim_pool(A, A->seg, *nkeep < N ? 2 : 0);
  double *p;
  for (int i = 0; i < N; i++)
    if (pool_layer(c, A->ckpt, i, &p)) return false;
  for (int i = 0; i < 2 && *nkeep < N; i++)
    if (pool_layer(c, A->seg, i, &p)) return false;
  for (int i = 0; i < *nkeep; i++) {
    if (pool_layer(c, A->tkeep, i, &p)) {
      *nkeep = *nkeep < N ? i : std::max(0, i - 2);
      trim_pool(A, A->tkeep, (size_t)*nkeep);
      for (int j = 0; j < 2; j++)
        if (pool_layer(c, A->seg, j, &p)) return false;
      break;
    }
  }
  return true;
}

int push_checkpoint(mbm_ctx *c, const double *src)
{
  mbm_adj_state *A = c->adj;
  double *p;
  int rc;
  if ((rc = pool_layer(c, A->ckpt, A->nckpt, &p))) return rc;
  A->nckpt++;
  HIPCHK(hipMemcpyAsync(p, src, c->layer_elems * sizeof(double), hipMemcpyDeviceToDevice,
                        c->stream));
  return MBM_OK;
}

void free_checkpoints(mbm_adj_state *A)
{
  for (double *p : A->ckpt) (void)hipFree(p);
  A->ckpt.clear();
  A->nckpt = 0;
  for (double *p : A->seg) (void)hipFree(p);
  A->seg.clear();
  for (double *p : A->tkeep) (void)hipFree(p);
  A->tkeep.clear();
  for (double *p : A->spare) (void)hipFree(p);
  A->spare.clear();
}

/* layers the recording may hold: the set budget, or the free device memory
 * plus the layers already held less a reserve of max(8 GiB, 5 %) */
long long budget_layers(mbm_ctx *c)
{
  mbm_adj_state *A = c->adj;
  const long long lb = (long long)(c->layer_elems * sizeof(double));
  if (c->adj_budget > 0) return c->adj_budget / lb;
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) return 0;
  const long long held =
      (long long)(A->ckpt.size() + A->seg.size() + A->tkeep.size() + A->spare.size()) * lb;
  const long long reserve = std::max<long long>(8ll << 30, (long long)(tot / 20));
  return std::max<long long>(0, ((long long)fr + held - reserve) / lb);
}

} // namespace

extern "C" long long mbm_layer_elems(const mbm_ctx *c) { return c ? (long long)c->layer_elems : 0; }

namespace {
int check_layer_count(const mbm_ctx *c, long long n)
{
  if (n != (long long)c->layer_elems)
    return set_err(MBM_E_ARG, "layer buffer holds %lld doubles, the layout needs %zu", n,
                   c->layer_elems);
  return MBM_OK;
}
} // namespace

extern "C" int mbm_layer_get(mbm_ctx *c, int layer, void *dst, long long n)
{
  if (!c || !dst || layer < 0 || layer > 2) return set_err(MBM_E_ARG, "invalid argument");
  if (int rc = check_layer_count(c, n)) return rc;
  HIPCHK(hipMemcpyAsync(dst, c->layer[c->role[layer]], c->layer_elems * sizeof(double),
                        hipMemcpyDefault, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MBM_OK;
}

extern "C" int mbm_layer_set(mbm_ctx *c, int layer, const void *src, long long n)
{
  if (!c || !src || layer < 0 || layer > 2) return set_err(MBM_E_ARG, "invalid argument");
  if (int rc = check_layer_count(c, n)) return rc;
  HIPCHK(hipMemcpyAsync(c->layer[c->role[layer]], src, c->layer_elems * sizeof(double),
                        hipMemcpyDefault, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MBM_OK;
}

static int adj_forward_run(mbm_ctx *c, double t0, double tend, double cflmax, double cflreg,
                           int maxsteps, const double *dts, int ndt, int stride, int *nsteps,
                           double *t_end)
{
  if (!c || stride < 0 || (dts && ndt < 0)) return set_err(MBM_E_ARG, "invalid argument");
  int rc = adj_setup(c);
  if (rc) return rc;
  mbm_adj_state *A = c->adj;
  A->nckpt = 0;
  A->dts.clear();
  A->resident = false;
  A->kept_from = 0;
  if (stride == 0) {
    const long long avail = dts && ndt > 0 ? budget_layers(c) : 0;
    if (avail >= (long long)ndt + 2) {
      /* every DATA state; the stage states of the last steps, all of them
       * when 2N layers fit, else as many as leave two layers for the stage
       * states the reverse sweep recomputes */
      const long long extra = avail - ndt;
      int nkeep = extra >= ndt ? ndt : (int)(extra - 2);
      if (reserve_resident(c, ndt, &nkeep)) {
        A->resident = true;
        A->kept_from = ndt - nkeep;
        stride = 1;
      }
    }
    if (!A->resident) {
      const int n = dts ? ndt : (maxsteps > 0 ? maxsteps : 1000);
      stride = std::max(1, (int)std::ceil(std::sqrt((double)std::max(n, 1))));
    }
  }
  if (!A->resident) {
    /* the checkpointed run's layers: N/stride checkpoints (N known for a
     * replay), 2 x (2 stride - 1) segment layers; the rest become spare */
    trim_pool(A, A->tkeep, 0);
    if (dts) trim_pool(A, A->ckpt, (size_t)std::max(1, (ndt + stride - 1) / stride));
    trim_pool(A, A->seg, (size_t)(2 * (2 * stride - 1)));
  }
  A->stride = stride;
  double t = t0;
  int k = 0;
  if (A->resident) {
    /* step k reads ckpt[k], writes its stage state to tkeep[k - kept_from]
     * or TMP and its result to ckpt[k + 1], the last one to DATA */
    HIPCHK(hipMemcpyAsync(A->ckpt[0], c->layer[c->role[MBM_L_DATA]],
                          c->layer_elems * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    A->nckpt = ndt;
    for (k = 0; k < ndt; k++) {
      double *T = k >= A->kept_from ? A->tkeep[k - A->kept_from] : c->layer[c->role[MBM_L_TMP]];
      double *N = k + 1 < ndt ? A->ckpt[k + 1] : c->layer[c->role[MBM_L_DATA]];
      if ((rc = step_buffers(c, A->ckpt[k], T, N, dts[k]))) return rc;
      t += dts[k];
      A->dts.push_back(dts[k]);
    }
  } else if ((rc = push_checkpoint(c, c->layer[c->role[MBM_L_DATA]]))) {
    return rc;
  } else if (dts) {
    /* replay: no CFL, the steps queue without host round trips */
    for (k = 0; k < ndt; k++) {
      if ((rc = step_buffers(c, c->layer[c->role[MBM_L_DATA]], c->layer[c->role[MBM_L_TMP]],
                             c->layer[c->role[MBM_L_NEW]], dts[k])))
        return rc;
      mbm_commit(c);
      t += dts[k];
      A->dts.push_back(dts[k]);
      if ((k + 1) % stride == 0 && k + 1 < ndt)
        if ((rc = push_checkpoint(c, c->layer[c->role[MBM_L_DATA]]))) return rc;
    }
  } else {
    double dt = 0.0001, cfl;
    while (t < tend && (maxsteps <= 0 || k < maxsteps)) {
      double used;
      do {
        used = dt;
        if ((rc = mbm_time_update(c, dt, t, &cfl))) return rc;
        if (cfl > cflmax) dt = dt * cflreg / cfl;
        else {
          t += dt;
          dt = dt * cflreg / cfl;
          if (t + dt > tend) dt = tend - t;
        }
      } while (cfl > cflmax);
      mbm_commit(c);
      A->dts.push_back(used);
      k++;
      if (k % stride == 0 && t < tend && (maxsteps <= 0 || k < maxsteps))
        if ((rc = push_checkpoint(c, c->layer[c->role[MBM_L_DATA]]))) return rc;
    }
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  if (nsteps) *nsteps = (int)A->dts.size();
  if (t_end) *t_end = t;
  return MBM_OK;
}

/* a forward run that fails leaves no recording behind (a backward pass is
 * then refused instead of reversing a partial run) */
extern "C" int mbm_adj_forward(mbm_ctx *c, double t0, double tend, double cflmax, double cflreg,
                               int maxsteps, const double *dts, int ndt, int stride,
                               int *nsteps, double *t_end)
{
  const int rc = adj_forward_run(c, t0, tend, cflmax, cflreg, maxsteps, dts, ndt, stride, nsteps,
                                 t_end);
  if (rc && c && c->adj) {
    c->adj->nckpt = 0;
    c->adj->dts.clear();
  }
  return rc;
}

extern "C" int mbm_adj_schedule(mbm_ctx *c, double *dts, int cap, int *n)
{
  if (!c || !c->adj) return set_err(MBM_E_STATE, "no recorded forward run");
  const int m = (int)c->adj->dts.size();
  for (int i = 0; i < m && i < cap; i++) dts[i] = c->adj->dts[i];
  if (n) *n = m;
  return MBM_OK;
}

extern "C" int mbm_adj_backward(mbm_ctx *c, const void *g_final, void *g_initial, long long n,
                                double *g_kn)
{
  if (!c || !g_final || !g_initial) return set_err(MBM_E_ARG, "NULL argument");
  if (int rc = check_layer_count(c, n)) return rc;
  if (!c->adj || c->adj->nckpt == 0) return set_err(MBM_E_STATE, "no recorded forward run");
  mbm_adj_state *A = c->adj;
  const size_t bytes = c->layer_elems * sizeof(double);
  const int N = (int)A->dts.size(), S = A->stride;
  int rc;
  HIPCHK(hipMemcpyAsync(A->g[0], g_final, bytes, hipMemcpyDefault, c->stream));
  HIPCHK(hipMemsetAsync(A->knpart, 0, A->nknpart * sizeof(double), c->stream));
  /* per segment: the DATA state and the stage state of every step (the
   * segment's first DATA state is its checkpoint), in two buffer sets: the
   * recomputation of segment s-1 (recompute stream, set (s-1)&1) runs while
   * segment s is reversed (context stream, set s&1) */
  const int L = std::min(S, std::max(N, 1));     /* steps in the longest segment */
  /* resident recording: segments of one step; steps k >= kept_from have
   * their stage state kept and need no recomputation */
  const int kept = A->resident ? A->kept_from : N;
  std::vector<double *> sD[2], sT[2];
  for (int q = 0; q < 2; q++) {
    sD[q].assign(L, nullptr);
    sT[q].assign(L, nullptr);
    if (kept == 0) continue;
    const int base = q * (2 * L - 1);
    for (int i = 0; i < L; i++) {
      if (i > 0 && (rc = pool_layer(c, A->seg, base + i - 1, &sD[q][i]))) return rc;
      if ((rc = pool_layer(c, A->seg, base + L - 1 + i, &sT[q][i]))) return rc;
    }
  }
  double *gN = A->g[0], *gD = A->g[1];
  const int nseg = (N + S - 1) / S;
  /* the recompute stream starts after everything queued so far (the
   * checkpoints of the forward run) */
  HIPCHK(hipEventRecord(A->ev_start, c->stream));
  HIPCHK(hipStreamWaitEvent(A->rstream, A->ev_start, 0));
  auto recompute = [&](int s) -> int {
    const int q = s & 1, k0 = s * S, k1 = std::min(N, k0 + S);
    sD[q][0] = A->ckpt[s];
    if (k0 >= kept) return MBM_OK;
    /* the stage and fill launchers enqueue on c->stream */
    hipStream_t keep = c->stream;
    c->stream = A->rstream;
    int r = MBM_OK;
    for (int k = k0; k < k1 - 1 && !r; k++)
      r = step_buffers(c, sD[q][k - k0], sT[q][k - k0], sD[q][k - k0 + 1], A->dts[k]);
    if (!r) r = stage1_buffers(c, sD[q][k1 - 1 - k0], sT[q][k1 - 1 - k0], A->dts[k1 - 1]);
    c->stream = keep;
    if (r) return r;
    HIPCHK(hipEventRecord(A->ev_ready[q], A->rstream));
    return MBM_OK;
  };
  if (nseg > 0 && (rc = recompute(nseg - 1))) return rc;
  for (int s = nseg - 1; s >= 0; s--) {
    const int q = s & 1, k0 = s * S, k1 = std::min(N, k0 + S);
    if (s > 0) {
      /* set (s-1)&1 was last read by the reversal of segment s+1 */
      if (s + 1 < nseg) HIPCHK(hipStreamWaitEvent(A->rstream, A->ev_done[(s + 1) & 1], 0));
      if ((rc = recompute(s - 1))) return rc;
    }
    if (k0 < kept) HIPCHK(hipStreamWaitEvent(c->stream, A->ev_ready[q], 0));
    for (int k = k1 - 1; k >= k0; k--) {
      const double *T = k >= kept ? A->tkeep[k - kept] : sT[q][k - k0];
      if ((rc = step_adjoint(c, sD[q][k - k0], T, A->dts[k], gN, gD))) return rc;
      std::swap(gN, gD);
    }
    HIPCHK(hipEventRecord(A->ev_done[q], c->stream));
  }
  if ((rc = fill_adjoint(c, gN))) return rc;
  HIPCHK(hipMemcpyAsync(g_initial, gN, bytes, hipMemcpyDefault, c->stream));
  std::vector<double> part(A->nknpart);
  HIPCHK(hipMemcpyAsync(part.data(), A->knpart, A->nknpart * sizeof(double), hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  double s = 0.0;
  for (double v : part) s += v;
  if (g_kn) *g_kn = s;
  return MBM_OK;
}

extern "C" int mbm_adj_set_budget(mbm_ctx *c, long long bytes)
{
  if (!c || bytes < 0) return set_err(MBM_E_ARG, "invalid argument");
  c->adj_budget = bytes;
  return MBM_OK;
}

extern "C" int mbm_adj_memory(mbm_ctx *c, long long *bytes)
{
  if (!c || !bytes) return set_err(MBM_E_ARG, "NULL argument");
  const mbm_adj_state *A = c->adj;
  *bytes = A ? (long long)(A->ckpt.size() + A->seg.size() + A->tkeep.size() + A->spare.size()) *
                   (long long)(c->layer_elems * sizeof(double))
             : 0;
  return MBM_OK;
}

extern "C" int mbm_adj_release(mbm_ctx *c)
{
  if (!c || !c->adj) return MBM_OK;
  mbm_adj_state *A = c->adj;
  (void)hipStreamSynchronize(c->stream);
  if (A->rstream) (void)hipStreamSynchronize(A->rstream);
  free_checkpoints(A);
  for (double *p : { A->g[0], A->g[1], A->g[2], A->gpy })
    if (p) (void)hipFree(p);
  for (hipEvent_t e : { A->ev_start, A->ev_ready[0], A->ev_ready[1], A->ev_done[0], A->ev_done[1] })
    if (e) (void)hipEventDestroy(e);
  if (A->rstream) (void)hipStreamDestroy(A->rstream);
  if (A->xtiles) (void)hipFree(A->xtiles);
  if (A->ytiles) (void)hipFree(A->ytiles);
  if (A->slots) (void)hipFree(A->slots);
  if (A->knpart) (void)hipFree(A->knpart);
  for (void *p : { (void *)A->fold_src, (void *)A->fold_ghost, (void *)A->fold_dir })
    if (p) (void)hipFree(p);
  delete A;
  c->adj = nullptr;
  return MBM_OK;
}
