/*
 * mbm_engine.hip -- gfx950 engine for the explicit update of
 * MultiBlockMomentsFV: stage kernel, ghost-fill kernel, CFL finalisation,
 * initial state and output conversion, behind the C ABI of mbmfv.h.
 *
 * Stage kernel (one Heun stage, Main.c:297-396, fused):
 *   one wavefront owns a strip of 60 columns (64 lanes: 2 halo lanes on each
 *   side) and a run of rows of one block. It marches row by row (stage 1
 *   upward, stage 2 downward), keeping the primitives of rows r, r+-1 and the
 *   partial update of row r in registers, and the reconstructed state of row
 *   r on the face ahead and the initial values of the next two rows in
 *   lane-private LDS slots; the input of row r+-2 is loaded while the x-faces
 *   of row r are computed. x-neighbours are
 *   exchanged with cross-lane moves, so there is no barrier; every wavefront
 *   is independent. Each x-face and each y-face is evaluated once. Face
 *   speeds and the production CFL are
 *   reduced per wavefront and folded into per-block maxima with 64-bit
 *   atomics (non-negative doubles order like their bit patterns).
 */
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cstdarg>
#include <vector>
#include <string>
#include <algorithm>
#include <mutex>
#include <type_traits>

#include "mbmfv.h"
#include "mbmfv_host.h"
#include "mbm_phys.h"
#include "mbm_scenario.h"

#include "mbm_internal.h"
#include "mbm_dev.h"

namespace mbmi {
thread_local char g_err[512];

int set_err(int code, const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}
} // namespace mbmi

using namespace mbmi;

using namespace mbm;

/* Staging buffer of mbm_output_values_noalloc: one copy per GPU, held for
 * the life of the process. On the host the name `scratch` is a stand-in, not
 * a device address: host code uses only the per-device address that
 * hipGetSymbolAddress returns (resolved once per device, with that device
 * current, when the first context is created on it). Every call on one GPU
 * shares the buffer, so calls that overlap on one GPU race. */
static __device__ __align__(256) unsigned char scratch[MBM_OUTPUT_SCRATCH_BYTES];

namespace mbmi {
constexpr int kMaxScratchDevices = 64;
void *g_scratch[kMaxScratchDevices];
std::mutex g_scratch_mu;

/* device address of `scratch` on the current device dev */
int resolve_scratch(int dev)
{
  if (dev < 0 || dev >= kMaxScratchDevices)
    return set_err(MBM_E_NODEV, "device %d outside [0,%d)", dev, kMaxScratchDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[dev]) return MBM_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return set_err(MBM_E_HIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", dev,
                   hipGetErrorString(e));
  g_scratch[dev] = p;
  return MBM_OK;
}
} // namespace mbmi

/* waves per SIMD the stage kernel is compiled for: 2 (256 registers per
 * wave, 220 used) lets one wave's memory waits overlap the other's FP64 work */
#ifndef MBM_STAGE_WAVES
#define MBM_STAGE_WAVES 2
#endif

/* stage 2 marches its tiles downward (MBM_STAGE2_DOWN=1): stage 1 ends at
 * the top rows of every tile, whose input and output are then the most
 * recently touched lines in the memory-side cache, and stage 2 starts there */
#ifndef MBM_STAGE2_DOWN
#define MBM_STAGE2_DOWN 1
#endif

/* the stage launches copy the interior-sourced ghosts at the end of the tile
 * that stores their sources (MBM_PUSH=1) instead of leaving them to the
 * fill kernels */
#ifndef MBM_PUSH
#define MBM_PUSH 1
#endif

#ifndef MBM_STAGE2_REVERSE
#define MBM_STAGE2_REVERSE 1
#endif

/* stage 2 reads DATA for the rows the tile stores only (MBM_DATA_OWN_ROWS=1):
 * the march's last projections, two rows past the tile, need no initial
 * value */
#ifndef MBM_DATA_OWN_ROWS
#define MBM_DATA_OWN_ROWS 1
#endif

/* the CFL parts of one tile: face-speed and production maxima (NaN
 * ignored), and 1 + the scan position of its last NaN production CFL */
struct TileOut {
  double mxx, mxy, mxp;
  unsigned long long pnan;
  int blk;
  double *out;        /* the layer the tile wrote */
};

/* One pass of a wavefront over its tile: one Heun stage of Update
 * (Main.c:297-396) on a strip of 60 columns and a run of rows. sm: the
 * wavefront's lane-private LDS slots (the reconstructed state of the current
 * row on the face ahead of the march -- the upper state when marching up,
 * the lower one when marching down -- and the initial values of the next two
 * rows, by row parity), which keep that state out of registers.
 *
 * Returns 0 for a tile with nothing to do (padding, or its deck is done).
 * SLOWOK false: the fast forms only; returns 2 if some row's stencil held
 * non-finite data or some face density came within rounding of zero, else 1.
 * SLOWOK true: rows with non-finite data use the reference's forms
 * (reconstruct_ref, face_ref), and a face whose density is exactly zero in
 * the reference's arithmetic is 0/0. A tile whose fast pass returned true is
 * run again with SLOWOK true; that pass rewrites every stored cell and the
 * CFL parts. The fast pass has no branches for these cases: its flags are
 * lane masks accumulated in scalar registers. */
template <int STAGE, int LIM, bool DOWN, bool SLOWOK>
__device__ __forceinline__ int tile_pass(const StageArgs &A, int tix, double (*sm)[NV][64], TileOut &o)
{
  /* tile, block and deck data (wave-uniform, scalar registers); each pass
   * loads them itself, so that nothing of the first pass is kept for the
   * second */
  const TileDev T = A.tiles[tix];
  if (T.blk < 0) return 0;               /* padding of the XCD ordering */
  const BlockDev B = A.blocks[T.blk];
  const double *in = A.in, *in0 = A.in0;
  double *out = A.out;
  double dt = A.dt, dtkn = A.dtkn, force = A.force;
  if (A.decks) {                       /* device-controlled: per-deck dt and buffers */
    const DeckState &S = A.decks[B.deck];
    if (S.done) return 0;
    dt = S.dt;
    dtkn = S.dt / S.kn;
    force = S.force;
    const double *data = S.parity ? A.buf1 : A.buf0;
    double *newb = const_cast<double *>(S.parity ? A.buf0 : A.buf1);
    in = STAGE == 1 ? data : A.tmp;
    in0 = data;
    out = STAGE == 1 ? A.tmp : newb;
  }
  const int lane = threadIdx.x & 63;
  const int lx = B.lx, ly = B.ly;
  const int a = T.a0 + lane;
  const int ac = a < lx ? a : lx - 1;
  /* block bases (uniform) and per-lane byte offsets within a block; the host
   * guarantees that a block spans less than 2^31 bytes */
  const long long nel = (long long)B.npos * NV * B.pitch;
  const __amdgpu_buffer_rsrc_t rin = block_rsrc(in + B.off, nel);
  const __amdgpu_buffer_rsrc_t rin0 = block_rsrc(in0 + B.off, nel);
  const __amdgpu_buffer_rsrc_t rout = block_rsrc(out + B.off, nel);
  const int pb = (int)(8 * B.plane);
  /* byte offset of row x (its position in the block's row order) */
  auto rowb = [&](int x) { return 8u * (unsigned)row_elem(B, x); };
  const unsigned cin = 8u * (unsigned)ac, cout = 8u * (unsigned)a;
  const double dtdx = dt / B.dx, dtdy = dt / B.dy;

  const bool col_ring3 = (a == 0 || a >= lx - 1);
  const bool col_prod = (a >= 1 && a <= lx - 2);
  const bool col_face = (a >= 1 && a <= lx - 3);
  const bool xface_ok = col_face && lane >= 1 && lane <= 61;
  const bool st_col = (lane >= 2 && lane <= 61) && a >= MBC && a <= lx - 1 - MBC;
#define TUP(n) sm[0][n][lane]
#define INIT(row, n) sm[1 + ((row) & 1)][n][lane]

  bool anyslow = false;
  double mxx = 0.0, mxy = 0.0, mxp = 0.0;
  unsigned long long pnan = 0ull;
  double U[NV], U0[NV], dummy[NV];
  double cp = 0.0;

  auto row_of = [&](int r) { return r < 0 ? 0 : (r > ly - 1 ? ly - 1 : r); };

  /* Non-finite data (wave-uniform bookkeeping, scalar registers only):
   * nfrows has bit 0 set when the last projected row holds a non-finite
   * relaxed primitive in some lane, bit 1 for the row before, ...; a face
   * whose stencil rows have a bit set is computed in the reference's form
   * (reconstruct_ref, face_ref). zcarry marks the lanes whose carried face
   * state has a density that is exactly zero in the reference's arithmetic. */
  unsigned nfrows = 0;
  unsigned long long zcarry = 0ull;
  /* lanes whose fast face density came within rounding of zero: the exact
   * decision (zero_density_ref) is left to the second pass */
  unsigned long long mzero = 0ull;
  auto note_maybe_zero = [&](bool c) { mzero |= __builtin_amdgcn_ballot_w64(c); };
  /* the form a face takes: reference's only in the second pass */
  auto slow_rows = [&](bool nonfinite) {
    anyslow = anyslow || nonfinite;
    return SLOWOK && nonfinite;
  };
  auto note_row = [&](bool bad) {
    nfrows = (nfrows << 1) | (__builtin_amdgcn_ballot_w64(bad) != 0ull ? 1u : 0u);
  };
  /* a NaN production CFL of cell (a, row) at ring >= 1: 1 + its position in
   * the reference's scan order (i outer, j inner) for the controller */
  auto note_prod = [&](int row) {
    if (__builtin_expect(cp != cp, 0) && col_prod && row >= 1 && row <= ly - 2) {
      const unsigned long long q = (unsigned long long)(a - 1) * (ly - 2) + row;
      pnan = q > pnan ? q : pnan;
    }
  };

  double Pa[NV], Pb[NV], Pc[NV], Pn[NV], acc[NV];

  /* x-faces of the current row (primitives in Pc): acc += F(i-1/2) - F(i+1/2) */
  auto xfaces = [&](bool slow) {
    double L[NV], Rr[NV], F[NV], s;
    if (__builtin_expect(slow, 0)) {
#pragma unroll
      for (int n = 0; n < NV; n++) {
        double lo, hi;
        reconstruct_ref<LIM>(Pc[n], Pc[n] - from_left(Pc[n]), from_right(Pc[n]) - Pc[n], B.dx, lo, hi);
        L[n] = hi;
        Rr[n] = from_right(lo);
      }
      face_ref<0>(L, Rr, dtdx, F, s);
    } else {
      double d10 = 0.0, d20 = 0.0, sx0 = 0.0;
#pragma unroll
      for (int g = 0; g < NV; g += 3) {
        double e1[3], e2[3], sx[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
          e1[k] = Pc[g + k] - from_left(Pc[g + k]);
          e2[k] = from_right(Pc[g + k]) - Pc[g + k];
        }
        limit3<LIM>(e1, e2, sx);
        if (g == 0) {
          d10 = e1[0];
          d20 = e2[0];
          sx0 = sx[0];
        }
#pragma unroll
        for (int k = 0; k < 3; k++) {
          L[g + k] = Pc[g + k] + 0.5 * sx[k];
          Rr[g + k] = from_right(Pc[g + k] - 0.5 * sx[k]);
        }
      }
      face<0>(L, Rr, dtdx, F, s);
      if constexpr (SLOWOK) {
        bool zlo, zhi;
        zero_density_ref<LIM>(Pc[0], d10, d20, B.dx, Pc[0] - 0.5 * sx0, L[0], zlo, zhi);
        /* face a+1/2: upper state of lane a, lower state of lane a+1 */
        const unsigned long long zf = __builtin_amdgcn_ballot_w64(zhi) | (__builtin_amdgcn_ballot_w64(zlo) >> 1);
        if (__builtin_expect(zf != 0ull, 0))
          if ((zf >> lane) & 1ull) nan_face(F);
      } else {
        note_maybe_zero(maybe_zero_density(Pc[0], Pc[0] - 0.5 * sx0, L[0]));
      }
    }
    if (xface_ok) fold_max(mxx, s);
#pragma unroll
    for (int n = 0; n < NV; n++) acc[n] += from_left(F[n]) - F[n];
  };
  /* input rows, clamped to the block (row 0 and ly-1 are ghost ring 3) */
  /* the stage input of a row, and in stage 2 its DATA, which only the
   * initial values of the tile's own rows use (MBM_DATA_OWN_ROWS) */
  auto fetch = [&](int row, double *Ud, double *U0d) {
    const unsigned roff = cin + rowb(row_of(row));
    load9b(rin, pb, roff, Ud);
    if constexpr (STAGE == 2)
      if (!MBM_DATA_OWN_ROWS || (row >= T.r0 && row <= T.r1)) load9b(rin0, pb, roff, U0d);
  };
  /* one row of the block: the fast face states, or the reference's
   * (slow), of the row with primitives pc between pm (below) and pp; zlo,
   * zhi: the reference's density of the lower/upper state is exactly 0.
   * CARRY 1: the upper state goes to the lane's LDS slot instead of hi,
   * CARRY 2: the lower state does (it is carried to the next row's face) */
  auto ystates = [&](auto carry, const double *pm, const double *pc, const double *pp, bool slow,
                     double *lo, double *hi, bool &zlo, bool &zhi) {
    constexpr int CARRY = decltype(carry)::value;
    auto put = [&](int n, double l, double h) {
      if constexpr (CARRY == 1) { TUP(n) = h; lo[n] = l; }
      else if constexpr (CARRY == 2) { TUP(n) = l; hi[n] = h; }
      else { lo[n] = l; hi[n] = h; }
    };
    if (__builtin_expect(slow, 0)) {
#pragma unroll
      for (int n = 0; n < NV; n++) {
        double l, h;
        reconstruct_ref<LIM>(pc[n], pc[n] - pm[n], pp[n] - pc[n], B.dy, l, h);
        if (n == 0) {
          zlo = l == 0.0;
          zhi = h == 0.0;
        }
        put(n, l, h);
      }
      return;
    }
    double d10 = 0.0, d20 = 0.0, lo0 = 0.0, hi0 = 0.0;
#pragma unroll
    for (int g = 0; g < NV; g += 3) {
      double e1[3], e2[3], sy[3];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        e1[k] = pc[g + k] - pm[g + k];
        e2[k] = pp[g + k] - pc[g + k];
      }
      limit3<LIM>(e1, e2, sy);
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const double l = pc[g + k] - 0.5 * sy[k], h = pc[g + k] + 0.5 * sy[k];
        if (g + k == 0) {
          d10 = e1[0];
          d20 = e2[0];
          lo0 = l;
          hi0 = h;
        }
        put(g + k, l, h);
      }
    }
    if constexpr (SLOWOK) {
      zero_density_ref<LIM>(pc[0], d10, d20, B.dy, lo0, hi0, zlo, zhi);
    } else {
      zlo = zhi = false;
      note_maybe_zero(maybe_zero_density(pc[0], lo0, hi0));
    }
  };
  using both_ = std::integral_constant<int, 0>;
  using up_ = std::integral_constant<int, 1>;
  using down_ = std::integral_constant<int, 2>;
  /* the y-face between lower state Lo (from below) and upper state Up; znan:
   * lanes whose face is 0/0 in the reference */
  auto yface = [&](const double *Lo, const double *Up, bool slow, unsigned long long znan, double *G,
                   double &s) {
    if (__builtin_expect(slow, 0)) face_ref<1>(Lo, Up, dtdy, G, s);
    else face<1>(Lo, Up, dtdy, G, s);
    if constexpr (SLOWOK) {
      if (__builtin_expect(znan != 0ull, 0))
        if ((znan >> lane) & 1ull) nan_face(G);
    }
  };

  if constexpr (!DOWN) {
    /* ---- warm-up: rows r0-2 .. r0+1 ---- */
    const int r0 = T.r0;
    {
      double Ir[NV];
      bool bad;
      int rr = row_of(r0 - 2);
      load9b(rin, pb, cin + rowb(rr), U);
      project<STAGE, false>(U, U0, col_ring3 || rr == 0 || rr == ly - 1, dt, dtkn, force, Pa, dummy, cp, bad);
      note_row(bad);
      rr = row_of(r0 - 1);
      load9b(rin, pb, cin + rowb(rr), U);
      project<STAGE, false>(U, U0, col_ring3 || rr == 0 || rr == ly - 1, dt, dtkn, force, Pb, dummy, cp, bad);
      note_row(bad);
      rr = row_of(r0);
      load9b(rin, pb, cin + rowb(rr), U);
      if constexpr (STAGE == 2) load9b(rin0, pb, cin + rowb(rr), U0);
      project<STAGE, true>(U, U0, col_ring3 || rr == 0 || rr == ly - 1, dt, dtkn, force, Pc, acc, cp, bad);
      note_row(bad);
      note_prod(rr);
      if (STAGE == 1 && col_prod && rr >= 1 && rr <= ly - 2) fold_max(mxp, cp);
      rr = row_of(r0 + 1);
      load9b(rin, pb, cin + rowb(rr), U);
      if constexpr (STAGE == 2) load9b(rin0, pb, cin + rowb(rr), U0);
      project<STAGE, true>(U, U0, col_ring3 || rr == 0 || rr == ly - 1, dt, dtkn, force, Pn, Ir, cp, bad);
      note_row(bad);
      note_prod(rr);
      if (STAGE == 1 && col_prod && rr >= 1 && rr <= ly - 2) fold_max(mxp, cp);
#pragma unroll
      for (int n = 0; n < NV; n++) INIT(r0 + 1, n) = Ir[n];

      /* face r0-1/2 from the states of rows r0-1 and r0 (same rounding as
       * the march computes them) */
      const bool slow = slow_rows((nfrows & 0xfu) != 0u);
      double lo[NV], Tprev[NV], Bot[NV], G[NV], s;
      bool z0, zp, zb, zt;
      ystates(both_{}, Pa, Pb, Pc, slow, lo, Tprev, z0, zp);
      ystates(up_{}, Pb, Pc, Pn, slow, Bot, nullptr, zb, zt);
      zcarry = __builtin_amdgcn_ballot_w64(zt);
      yface(Tprev, Bot, slow, __builtin_amdgcn_ballot_w64(zp) | __builtin_amdgcn_ballot_w64(zb), G, s);
      if (col_face && r0 - 1 >= 1 && r0 - 1 <= ly - 3) fold_max(mxy, s);
#pragma unroll
      for (int n = 0; n < NV; n++) acc[n] += G[n];
    }

    /* ---- march: row r, with the input of row r+2 in (Uc, U0c) ---- */
    auto march_row = [&](int r, const double *Uc, const double *U0c) {
      xfaces(slow_rows(((nfrows >> 1) & 1u) != 0u));
      if constexpr (SLOWOK) fetch(r + 2, U, U0);
      __builtin_amdgcn_sched_barrier(0);

      /* row r+2: projection */
      double P2[NV];
      const int rr = r + 2;
      {
        double I2[NV];
        bool bad;
        project<STAGE, true>(Uc, U0c, col_ring3 || rr == ly - 1, dt, dtkn, force, P2, I2, cp, bad);
        note_row(bad);
#pragma unroll
        for (int n = 0; n < NV; n++) INIT(rr, n) = I2[n];
      }
      note_prod(rr);
      if (STAGE == 1 && col_prod && rr <= ly - 2) fold_max(mxp, cp);
      __builtin_amdgcn_sched_barrier(0);

      /* y-face r+1/2: upper state of row r (carried), lower state of row r+1 */
      {
        const bool slow = slow_rows((nfrows & 0xfu) != 0u);
        double Tu[NV], Bn[NV], G[NV], s;
        bool zb, zt;
#pragma unroll
        for (int n = 0; n < NV; n++) Tu[n] = TUP(n);
        ystates(up_{}, Pc, Pn, P2, slow, Bn, nullptr, zb, zt);
        yface(Tu, Bn, slow, zcarry | __builtin_amdgcn_ballot_w64(zb), G, s);
        zcarry = __builtin_amdgcn_ballot_w64(zt);
        if (col_face) fold_max(mxy, s);
        if (st_col && r >= MBC && r <= ly - 1 - MBC) {
          const unsigned ob = cout + rowb(r);
#pragma unroll
          for (int n = 0; n < NV; n++) {
            const double v = acc[n] - G[n];
            store1b(rout, n * pb, ob, (STAGE == 2) ? 0.5 * v : v);
          }
        }
#pragma unroll
        for (int n = 0; n < NV; n++) {
          acc[n] = INIT(r + 1, n) + G[n];
          Pc[n] = Pn[n];
          Pn[n] = P2[n];
        }
      }
    };
    for (int r = r0; r <= T.r1; r++) {
      /* row r+2 is loaded while the x-faces of row r are computed (in the
       * second pass after them: fewer registers live) */
      if constexpr (!SLOWOK) fetch(r + 2, U, U0);
      march_row(r, U, U0);
    }
  } else {
    /* ---- warm-up: rows r1+2 .. r1-1 (the mirror image of the above) ---- */
    const int r1 = T.r1;
    {
      double Ir[NV];
      bool bad;
      int rr = row_of(r1 + 2);
      load9b(rin, pb, cin + rowb(rr), U);
      project<STAGE, false>(U, U0, col_ring3 || rr == 0 || rr == ly - 1, dt, dtkn, force, Pa, dummy, cp, bad);
      note_row(bad);
      rr = row_of(r1 + 1);
      load9b(rin, pb, cin + rowb(rr), U);
      project<STAGE, false>(U, U0, col_ring3 || rr == 0 || rr == ly - 1, dt, dtkn, force, Pb, dummy, cp, bad);
      note_row(bad);
      rr = row_of(r1);
      load9b(rin, pb, cin + rowb(rr), U);
      load9b(rin0, pb, cin + rowb(rr), U0);
      project<STAGE, true>(U, U0, col_ring3 || rr == 0 || rr == ly - 1, dt, dtkn, force, Pc, acc, cp, bad);
      note_row(bad);
      note_prod(rr);
      rr = row_of(r1 - 1);
      load9b(rin, pb, cin + rowb(rr), U);
      load9b(rin0, pb, cin + rowb(rr), U0);
      project<STAGE, true>(U, U0, col_ring3 || rr == 0 || rr == ly - 1, dt, dtkn, force, Pn, Ir, cp, bad);
      note_row(bad);
      note_prod(rr);
#pragma unroll
      for (int n = 0; n < NV; n++) INIT(r1 - 1, n) = Ir[n];

      /* face r1+1/2 from the states of rows r1+1 and r1, as the march
       * computes them, so that it equals the face the tile above computes */
      const bool slow = slow_rows((nfrows & 0xfu) != 0u);
      double Bu[NV], hi[NV], Tl[NV], G[NV], s;
      bool zb, z0, zlc, zt;
      ystates(both_{}, Pc, Pb, Pa, slow, Bu, hi, zb, z0);
      ystates(down_{}, Pn, Pc, Pb, slow, nullptr, Tl, zlc, zt);
      zcarry = __builtin_amdgcn_ballot_w64(zlc);
      yface(Tl, Bu, slow, __builtin_amdgcn_ballot_w64(zt) | __builtin_amdgcn_ballot_w64(zb), G, s);
      if (col_face) fold_max(mxy, s);
#pragma unroll
      for (int n = 0; n < NV; n++) acc[n] -= G[n];
    }

    /* ---- march down: row r, with the input of row r-2 in (Uc, U0c); a cell
     * accumulates init - G(j+1/2) + (F(i-1/2) - F(i+1/2)) + G(j-1/2) ---- */
    auto march_row = [&](int r, const double *Uc, const double *U0c) {
      xfaces(slow_rows(((nfrows >> 1) & 1u) != 0u));
      if constexpr (SLOWOK) fetch(r - 2, U, U0);
      __builtin_amdgcn_sched_barrier(0);

      /* row r-2 (clamped to ring 3 below r = 1): projection */
      double P2[NV];
      {
        const int rr = row_of(r - 2);
        double I2[NV];
        bool bad;
        project<STAGE, true>(Uc, U0c, col_ring3 || rr == 0, dt, dtkn, force, P2, I2, cp, bad);
        note_row(bad);
        note_prod(r - 2);
#pragma unroll
        for (int n = 0; n < NV; n++) INIT(r - 2, n) = I2[n];
      }
      __builtin_amdgcn_sched_barrier(0);

      /* y-face r-1/2: upper state of row r-1, lower state of row r (carried) */
      {
        const bool slow = slow_rows((nfrows & 0xfu) != 0u);
        double Bu[NV], Tl[NV], G[NV], s;
        bool zb, zt;
#pragma unroll
        for (int n = 0; n < NV; n++) Bu[n] = TUP(n);
        ystates(down_{}, P2, Pn, Pc, slow, nullptr, Tl, zb, zt);
        yface(Tl, Bu, slow, zcarry | __builtin_amdgcn_ballot_w64(zt), G, s);
        zcarry = __builtin_amdgcn_ballot_w64(zb);
        if (col_face && r - 1 >= 1) fold_max(mxy, s);
        if (st_col && r >= MBC && r <= ly - 1 - MBC) {
          const unsigned ob = cout + rowb(r);
#pragma unroll
          for (int n = 0; n < NV; n++) {
            const double v = acc[n] + G[n];
            store1b(rout, n * pb, ob, (STAGE == 2) ? 0.5 * v : v);
          }
        }
#pragma unroll
        for (int n = 0; n < NV; n++) {
          acc[n] = INIT(r - 1, n) - G[n];
          Pc[n] = Pn[n];
          Pn[n] = P2[n];
        }
      }
    };
    for (int r = r1; r >= T.r0; r--) {
      /* row r-2 is loaded while the x-faces of row r are computed (in the
       * second pass after them) */
      if constexpr (!SLOWOK) fetch(r - 2, U, U0);
      march_row(r, U, U0);
    }
  }
#undef TUP
#undef INIT
  o.mxx = mxx;
  o.mxy = mxy;
  o.mxp = mxp;
  o.pnan = pnan;
  o.blk = T.blk;
  o.out = out;
  return (anyslow || mzero != 0ull) ? 2 : 1;
}

/* bytes of dynamic LDS a stage workgroup of w waves needs (per wave: the
 * carried face state and the initial values of two rows, 3 x 9 x 64 doubles) */
constexpr size_t stage_lds_bytes(int w) { return (size_t)w * 3 * NV * 64 * sizeof(double); }

/* One workgroup = blockDim.x / 64 waves, each running consecutive tiles of
 * the launch order (the context's wave group, 1 to MAX_WAVES_PER_WG). */
template <int STAGE, int LIM, bool DOWN>
__global__ void __launch_bounds__(64 * MAX_WAVES_PER_WG)
__attribute__((amdgpu_waves_per_eu(MBM_STAGE_WAVES, MBM_STAGE_WAVES)))
stage_kernel(const StageArgs A)
{
  /* the production CFL (stage 1) is folded by the upward march only */
  static_assert(!DOWN || STAGE == 2, "downward march is stage 2 only");
  const int lane = threadIdx.x & 63;
  const int raw = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
  if (raw >= A.ntiles) return;
  /* stage 2 takes the launch list from its end (MBM_STAGE2_REVERSE): a launch
   * of several rounds of waves then starts on the tiles stage 1 finished
   * last, whose lines the memory-side cache still holds; the list is padded
   * to whole rounds of 8 XCDs, so every XCD keeps a contiguous patch */
  const int tix = (MBM_STAGE2_REVERSE && STAGE == 2) ? A.ntiles - 1 - raw : raw;
  extern __shared__ double stage_lds[];
  double (*sm)[3][NV][64] = reinterpret_cast<double (*)[3][NV][64]>(stage_lds);
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  TileOut o;
  const int pass = tile_pass<STAGE, LIM, DOWN, false>(A, tix, sm[wv], o);
  if (pass == 0) return;
  if (__builtin_expect(pass == 2, 0)) tile_pass<STAGE, LIM, DOWN, true>(A, tix, sm[wv], o);
  double mxx = o.mxx, mxy = o.mxy, mxp = o.mxp;
  const unsigned long long pnan = o.pnan;

  /* ---- CFL maxima of this wavefront ---- */
  mxx = wave_max(mxx);
  mxy = wave_max(mxy);
  if (STAGE == 1) mxp = wave_max(mxp);
  if (lane == 0) {
    unsigned long long *s = A.slots + SLOTS_PER_BLOCK * o.blk;
    atomicMax(s + 0, (unsigned long long)__double_as_longlong(mxx));
    atomicMax(s + 1, (unsigned long long)__double_as_longlong(mxy));
    if (STAGE == 1) atomicMax(s + 2, (unsigned long long)__double_as_longlong(mxp));
  }
  /* the last NaN production CFL of this wavefront (rare) */
  if (__builtin_expect(__builtin_amdgcn_ballot_w64(pnan != 0ull) != 0ull, 0))
    if (pnan) atomicMax(A.slots + SLOTS_PER_BLOCK * o.blk + 3, pnan);

  /* ---- ghost copies whose sources this tile stored (SetBoundary's
   * interior-sourced ghosts, Boundary.c:51-132): the source lines were just
   * written by this wavefront and are read back from its L2; the ghost
   * cells of the output layer are read by no tile of this launch ---- */
  if (A.prange) {
    const int2 pr = A.prange[tix];
    if (pr.x < pr.y) {
      /* this wavefront's stores complete before its loads of the same lines */
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
      double *out = o.out;
      for (int i = pr.x + lane; i < pr.y; i += 64) {
        const FillRec g = A.push[i];
        const double *src = out + g.src;
        double v[NV];
#pragma unroll
        for (int m = 0; m < NV; m++) v[m] = src[(long long)m * g.splane];
        double *dst = out + g.dst;
#pragma unroll
        for (int m = 0; m < NV; m++) dst[(long long)m * g.dplane] = ((g.sign >> m) & 1) ? -v[m] : v[m];
      }
    }
  }
}

/* SetBoundary as a gather over the precomputed ghost map (Boundary.c:173-193).
 * decks != nullptr: layer chosen per deck (which 1: TMP, 2: NEW by parity),
 * decks that are done are skipped. */
__global__ void __launch_bounds__(256) fill_kernel(double *layer, const FillRec *rec,
                                                   const double2 *xy, int n, int scenario,
                                                   const double *tj, const DeckState *decks,
                                                   double *buf0, double *buf1, double *tmp,
                                                   int which)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const FillRec g = rec[i];
  if (decks) {
    const DeckState &S = decks[g.deck];
    if (S.done) return;
    layer = which == 1 ? tmp : (S.parity ? buf0 : buf1);
  }
  double v[NV];
  if (g.src >= 0) {
    const double *src = layer + g.src;
#pragma unroll
    for (int m = 0; m < NV; m++) v[m] = src[(long long)m * g.splane];
  } else {
    double tjl[5];
#pragma unroll
    for (int m = 0; m < 5; m++) tjl[m] = tj[m];
    const double2 p = xy[i];
    mbm_scn_eval(scenario, tjl, p.x, p.y, v);
  }
  double *dst = layer + g.dst;
#pragma unroll
  for (int m = 0; m < NV; m++) dst[(long long)m * g.dplane] = ((g.sign >> m) & 1) ? -v[m] : v[m];
}

/* Production CFL of block B (stage STAGE) in the reference's fold when a NaN
 * occurred: Main.c:325-326 folds CFL_prod = max(CFL_prod, new), which keeps
 * the running value only while new is a number, so the result is the
 * maximum over the cells after the last NaN in the i-outer, j-inner scan
 * (pnan = 1 + its position), NaN when that NaN is the last cell. Recomputed
 * from the stage input by the calling wavefront (64 lanes, all lanes call). */
template <int STAGE>
__device__ double prod_after_nan(const double *layer, const BlockDev &B, unsigned long long pnan,
                                 double dt, double dtkn, double force)
{
  const unsigned long long ny2 = (unsigned long long)(B.ly - 2);
  const unsigned long long ncell = (unsigned long long)(B.lx - 2) * ny2;
  if (pnan >= ncell) return __builtin_nan("");
  double m = -HUGE_VAL;
  for (unsigned long long q = pnan + (threadIdx.x & 63); q < ncell; q += 64) {
    const int a = 1 + (int)(q / ny2), b = 1 + (int)(q % ny2);
    const double *p = layer + cell_elem(B, b, a);
    double U[NV], P[NV], I[NV], cp;
    bool bad;
    for (int n = 0; n < NV; n++) U[n] = p[n * B.plane];
    project<STAGE, true>(U, U, false, dt, dtkn, force, P, I, cp, bad);
    if (cp > m) m = cp;
  }
  return wave_max(m);
}

/* CFL of blocks [k0, k1) in one stage from their slots (slots of block k at
 * SLOTS_PER_BLOCK * k): per block dt*cx/dx + dt*cy/dy + cprod (Main.c:395),
 * the production part refolded after a NaN; over the blocks the running
 * maximum of Main.c:268-282, which starts at 0 and skips a NaN block, i.e.
 * the largest non-NaN block CFL or 0 (the order does not matter). One lane
 * per block reads and clears its slots; a block with a NaN production CFL
 * is refolded by the whole wavefront. All 64 lanes call; the result is
 * uniform. */
template <int STAGE>
__device__ double stage_cfl(unsigned long long *slots, const BlockDev *blocks, int k0, int k1,
                            const double *in, double dt, double dtkn, double force)
{
  const int lane = threadIdx.x & 63;
  double c = 0.0;
  for (int base = k0; base < k1; base += 64) {
    const int k = base + lane;
    const bool have = k < k1;
    unsigned long long s0 = 0ull, s1 = 0ull, s2 = 0ull, s3 = 0ull;
    if (have) {
      unsigned long long *sk = slots + (size_t)SLOTS_PER_BLOCK * k;
      s0 = sk[0];
      s1 = sk[1];
      s2 = sk[2];
      s3 = sk[3];
      sk[0] = 0ull;
      sk[1] = 0ull;
      sk[2] = 0ull;
      sk[3] = 0ull;
    }
    double cprod = __longlong_as_double((long long)s2);
    unsigned long long nan_lanes = __builtin_amdgcn_ballot_w64(have && s3 != 0ull);
    while (nan_lanes) {
      const int l = __builtin_ctzll(nan_lanes);
      nan_lanes &= nan_lanes - 1ull;
      const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(s3 & 0xffffffffull), l);
      const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(s3 >> 32), l);
      const double r = prod_after_nan<STAGE>(in, blocks[base + l],
                                             ((unsigned long long)hi << 32) | lo, dt, dtkn, force);
      if (lane == l) cprod = r;
    }
    if (have) {
      const BlockDev &B = blocks[k];
      const double cx = __longlong_as_double((long long)s0);
      const double cy = __longlong_as_double((long long)s1);
      const double v = dt * cx / B.dx + dt * cy / B.dy + cprod;
      if (v > c) c = v;
    }
  }
  return wave_max(c);
}

/* Steps controller of every deck (Main.c:202-216) after one TimeUpdate:
 * combined CFL (Main.c:395, :268-282; a block whose CFL is NaN is skipped,
 * as the reference's `cfl > CFL1` skips it), accept/reject, dt update, clamp
 * to tend, parity flip instead of the New->Data copy; clears the slots.
 * One 64-lane workgroup per deck. */
__global__ void __launch_bounds__(64) control_kernel(DeckState *decks, unsigned long long *slots,
                                                     const BlockDev *blocks, const int *first, int nb,
                                                     RecordDev *rec, const double *buf0,
                                                     const double *buf1, const double *tmp)
{
  const int d = blockIdx.x;
  DeckState S = decks[d];
  if (S.done) return;
  const double dtkn = S.dt / S.kn;
  const double *data = S.parity ? buf1 : buf0;
  const double c1 = stage_cfl<1>(slots, blocks, first[d], first[d + 1], data, S.dt, dtkn, S.force);
  const double c2 = stage_cfl<2>(slots + (size_t)SLOTS_PER_BLOCK * nb, blocks, first[d], first[d + 1],
                                 tmp, S.dt, dtkn, S.force);
  if (threadIdx.x != 0) return;
  const double m2 = c2 > 0.0 ? c2 : 0.0;
  const double cfl = c1 > m2 ? c1 : m2;
  const double used = S.dt;
  const int acc = !(cfl > S.cflmax);
  if (!acc) {
    S.dt = S.dt * S.cflreg / cfl;
    S.nrej++;
  } else {
    S.t += S.dt;
    S.dt = S.dt * S.cflreg / cfl;
    if (S.t + S.dt > S.tend) S.dt = S.tend - S.t;
    S.nstep++;
    S.parity ^= 1;
    if (!(S.t < S.tend) || (S.maxsteps > 0 && S.nstep >= S.maxsteps)) S.done = 1;
  }
  S.cfl = cfl;
  if (rec && d == 0 && S.nrec < S.reccap) {
    RecordDev r;
    r.cfl = cfl;
    r.dt = used;
    r.t = S.t;
    r.accepted = acc;
    r.step = S.nstep + (acc ? 0 : 1);
    rec[S.nrec] = r;
  }
  S.nrec++;
  decks[d] = S;
}

/* Combined CFL (Main.c:395 per block, Main.c:268-282 over blocks and stages),
 * evaluated in the reference's order; clears the slots for the next step.
 * One 64-lane workgroup; in1/in2: the inputs of stages 1 and 2. */
__global__ void __launch_bounds__(64) finalize_kernel(unsigned long long *slots, const BlockDev *blocks,
                                                      int nb, double dt, double kn, double force,
                                                      const double *in1, const double *in2, double *out)
{
  const double c1 = stage_cfl<1>(slots, blocks, 0, nb, in1, dt, dt / kn, force);
  const double c2 = stage_cfl<2>(slots + (size_t)SLOTS_PER_BLOCK * nb, blocks, 0, nb, in2, dt, dt / kn, force);
  if (threadIdx.x != 0) return;
  const double m2 = c2 > 0.0 ? c2 : 0.0;
  out[0] = c1 > m2 ? c1 : m2;
}

__global__ void zero_kernel(unsigned long long *p, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

/* copies every cell of one block between two row orders (n = ly*NV*lx) */
__global__ void relayout_kernel(double *dst, const double *src, const BlockDev to, const BlockDev from,
                                int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int a = i % from.lx, t = i / from.lx, m = t % NV, b = t / NV;
  dst[cell_elem(to, b, a) + m * to.plane] = src[cell_elem(from, b, a) + m * from.plane];
}

/* interior initial state (Main.c:126-136) */
__global__ void init_kernel(double *layer, const BlockDev *blocks, int blk, int nx, int ny,
                            int scenario, const double *tj)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nx * ny) return;
  const BlockDev B = blocks[blk];
  const int a = MBC + i % nx, b = MBC + i / nx;
  double tjl[5], U[NV];
  for (int m = 0; m < 5; m++) tjl[m] = tj[m];
  /* GetPoint (tools.c:97-104) without contraction */
  const double x = __dadd_rn(B.bx, __dmul_rn((double)a, B.dx));
  const double y = __dadd_rn(B.by, __dmul_rn((double)b, B.dy));
  mbm_scn_eval(scenario, tjl, x, y, U);
  double *d = layer + cell_elem(B, b, a);
  for (int m = 0; m < NV; m++) d[m * B.plane] = U[m];
}

/* Output.c:217-243: interior + 1 ring, point p = (a-2)*(ny+2) + (b-2) */
template <typename OUT>
__global__ void output_kernel(const double *layer, const BlockDev *blocks, int blk, int nx,
                              int ny, OUT *out)
{
#pragma clang fp contract(off)
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int ni = ny + 2;
  if (p >= (nx + 2) * ni) return;
  const BlockDev B = blocks[blk];
  const int a = MBC - 1 + p / ni, b = MBC - 1 + p % ni;
  const double *s = layer + cell_elem(B, b, a);
  double U[NV];
  for (int m = 0; m < NV; m++) U[m] = s[m * B.plane];
  /* Primitive with true divisions (Flux.c:30-56), output only */
  const double rho = U[0], v1 = U[1] / rho, v2 = U[2] / rho;
  const double vs = v1 * v1 + v2 * v2;
  const double th = (2 * U[3] / rho - vs) / 3;
  const double t11 = U[4] / rho - v1 * v1, t12 = U[5] / rho - v1 * v2, t22 = U[6] / rho - v2 * v2;
  const double q1 = U[7] - U[3] * v1 - rho * (t11 * v1 + t12 * v2);
  const double q2 = U[8] - U[3] * v2 - rho * (t12 * v1 + t22 * v2);
  OUT *o = out + (size_t)10 * p;
  o[0] = (OUT)rho; o[1] = (OUT)v1; o[2] = (OUT)v2; o[3] = (OUT)th;
  o[4] = (OUT)(rho * th); o[5] = (OUT)(rho * t11); o[6] = (OUT)(rho * t12);
  o[7] = (OUT)(rho * t22); o[8] = (OUT)q1; o[9] = (OUT)q2;
}

/* ------------------------------------------------------------------ */
/* host side                                                           */
/* ------------------------------------------------------------------ */

extern "C" const char *mbm_last_error(void) { return g_err; }

namespace mbmi {

/* ghost cell resolution: returns the final value source of ghost (a,b) of
 * block k after both sweeps of SetBoundary (Boundary.c:145-193). */
struct Src {
  int blk, a, b;   /* source cell, blk < 0: scenario at ghost (a,b) of block -blk-1 */
  int sign;
};

inline int kind_of(int code) { return code > 0 ? 0 : code == 0 ? 1 : code == -1 ? 2 : 3; }

const int kWallSign[4] = {
  (1 << 1) | (1 << 5) | (1 << 7),     /* left: flip U1,U5,U7 (Scenario.c:57-63) */
  (1 << 1) | (1 << 5) | (1 << 7),
  (1 << 2) | (1 << 5) | (1 << 8),     /* bottom: flip U2,U5,U8 */
  (1 << 2) | (1 << 5) | (1 << 8) };

void inner_ab(const mbm_block_desc &B, int side, int ti, int s, int &a, int &b)
{
  const int lx = B.nx + 2 * MBC, ly = B.ny + 2 * MBC;
  switch (side) {
    case 0: a = MBC - 1 + s; b = ti; break;
    case 1: a = lx - MBC - s; b = ti; break;
    case 2: a = ti; b = MBC - 1 + s; break;
    default: a = ti; b = ly - MBC - s; break;
  }
}

Src resolve(const std::vector<mbm_block_desc> &D, int k, int a, int b, int depth);

/* non-corner ghost on 'side' at tangential ti, layer n (FillEdge, Boundary.c:88-132) */
Src resolve_edge(const std::vector<mbm_block_desc> &D, int k, int side, int ti, int n, int ga,
                 int gb, int depth)
{
  static const int opp[4] = { 1, 0, 3, 2 };
  const int code = D[k].bc[side];
  Src s{ k, 0, 0, 0 };
  switch (kind_of(code)) {
    case 0: {
      const int o = code - 1;
      inner_ab(D[o], opp[side], ti, n, s.a, s.b);
      s.blk = o;
      const int lxo = D[o].nx + 2 * MBC, lyo = D[o].ny + 2 * MBC;
      const bool inner = s.a >= MBC && s.a < lxo - MBC && s.b >= MBC && s.b < lyo - MBC;
      if (!inner) return resolve(D, o, s.a, s.b, depth + 1);
      return s;
    }
    case 1: inner_ab(D[k], side, ti, 1, s.a, s.b); return s;
    case 2: inner_ab(D[k], side, ti, n, s.a, s.b); s.sign = kWallSign[side]; return s;
    default: s.blk = -k - 1; s.a = ga; s.b = gb; return s;
  }
}

Src resolve(const std::vector<mbm_block_desc> &D, int k, int a, int b, int depth)
{
  static const int opp[4] = { 1, 0, 3, 2 };
  const int lx = D[k].nx + 2 * MBC, ly = D[k].ny + 2 * MBC;
  if (depth > 4) return Src{ k, a, b, 0 };  /* unreachable for nx,ny >= 3 */
  const int xs = a < MBC ? 0 : (a >= lx - MBC ? 1 : -1);
  const int ys = b < MBC ? 2 : (b >= ly - MBC ? 3 : -1);
  const int nxg = xs == 0 ? MBC - a : a - (lx - 1 - MBC);
  const int nyg = ys == 2 ? MBC - b : b - (ly - 1 - MBC);
  if (xs < 0 && ys < 0) return Src{ k, a, b, 0 };
  if (ys < 0) return resolve_edge(D, k, xs, b, nxg, a, b, depth);
  if (xs < 0) return resolve_edge(D, k, ys, a, nyg, a, b, depth);
  /* corner: last writer of the two sweeps */
  const int cy = D[k].bc[ys], cx = D[k].bc[xs];
  int sa, sb;
  if (kind_of(cy) == 0) {
    const int o = cy - 1;
    inner_ab(D[o], opp[ys], a, nyg, sa, sb);
    return resolve(D, o, sa, sb, depth + 1);
  }
  if (kind_of(cx) == 0) {
    const int o = cx - 1;
    inner_ab(D[o], opp[xs], b, nxg, sa, sb);
    return resolve(D, o, sa, sb, depth + 1);
  }
  switch (kind_of(cy)) {
    case 1:
      inner_ab(D[k], ys, a, 1, sa, sb);
      return resolve(D, k, sa, sb, depth + 1);
    case 2: {
      inner_ab(D[k], ys, a, nyg, sa, sb);
      Src s = resolve(D, k, sa, sb, depth + 1);
      s.sign ^= kWallSign[ys];
      return s;
    }
    default:
      return Src{ -k - 1, a, b, 0 };
  }
}

int validate(const mbm_block_desc *blocks, int nb, const mbm_params *p)
{
  if (nb < 1 || nb > MBM_MAX_BLOCKS) return set_err(MBM_E_ARG, "nblocks %d outside [1,%d]", nb, MBM_MAX_BLOCKS);
  if (!p) return set_err(MBM_E_ARG, "params is NULL");
  if (p->limiter < 0 || p->limiter > 4) return set_err(MBM_E_ARG, "limiter %d invalid", p->limiter);
  if (p->scenario < 0 || p->scenario > 2) return set_err(MBM_E_ARG, "scenario %d invalid", p->scenario);
  for (int k = 0; k < nb; k++) {
    const mbm_block_desc &B = blocks[k];
    if (B.nx < MBC || B.ny < MBC)
      return set_err(MBM_E_ARG, "block %d: grid %dx%d smaller than 3x3", k + 1, B.nx, B.ny);
    /* the kernels address a block with 32-bit byte offsets: 72 bytes per
     * cell of the padded rows < 2^31 (the row order's padding rows are
     * checked when the layers are laid out) */
    if ((long long)(B.nx + 6 + 7) * (B.ny + 6) * 8 * NV + 256 >= (1ll << 31))
      return set_err(MBM_E_ARG, "block %d too large (%dx%d)", k + 1, B.nx, B.ny);
    if (!(B.xmax > B.xmin) || !(B.ymax > B.ymin))
      return set_err(MBM_E_ARG, "block %d has an empty extent", k + 1);
    for (int s = 0; s < 4; s++) {
      const int c = B.bc[s];
      if (c > nb) return set_err(MBM_E_ARG, "block %d side %d syncs with missing block %d", k + 1, s, c);
      if (c > 0) {
        const mbm_block_desc &O = blocks[c - 1];
        const bool conf = s < 2 ? O.ny == B.ny : O.nx == B.nx;
        if (!conf)
          return set_err(MBM_E_ARG, "non-conforming sync interface: block %d side %d vs block %d", k + 1, s, c);
      }
    }
  }
  return MBM_OK;
}

/* geometry exactly as SetBlocks (Mesh.c:169-172) */
void geometry(const mbm_block_desc &B, double &bx, double &by, double &dx, double &dy)
{
  const int nst = 2 * MBC;
  dx = (B.xmax - B.xmin) / B.nx;
  dy = (B.ymax - B.ymin) / B.ny;
  bx = B.xmin - (nst / 2 - 0.5) * dx;
  by = B.ymin - (nst / 2 - 0.5) * dy;
}

std::vector<GhostDev> build_ghost_map(const std::vector<mbm_block_desc> &D,
                                      const std::vector<BlockDev> &hb)
{
  std::vector<GhostDev> m;
  for (int k = 0; k < (int)D.size(); k++) {
    const int lx = D[k].nx + 2 * MBC, ly = D[k].ny + 2 * MBC;
    for (int b = 0; b < ly; b++)
      for (int a = 0; a < lx; a++) {
        const bool ghost = a < MBC || a >= lx - MBC || b < MBC || b >= ly - MBC;
        if (!ghost) continue;
        const Src s = resolve(D, k, a, b, 0);
        GhostDev g{};
        g.dblk = k;
        g.da = a;
        g.db = b;
        g.sign = s.sign;
        if (s.blk >= 0) {
          g.sblk = s.blk;
          g.sa = s.a;
          g.sb = s.b;
        } else {
          const int kk = -s.blk - 1;
          g.sblk = -1;
          g.x = hb[kk].bx + s.a * hb[kk].dx;
          g.y = hb[kk].by + s.b * hb[kk].dy;
        }
        m.push_back(g);
      }
  }
  return m;
}

/* Tiles of every block in the order block -> row range -> strip, so that
 * consecutive tiles are horizontally adjacent strips of the same rows. Each
 * block splits its rows evenly (tile_band): no short last tile. */
std::vector<TileDev> build_tiles(const std::vector<mbm_block_desc> &D, int rows)
{
  std::vector<TileDev> t;
  for (int k = 0; k < (int)D.size(); k++) {
    const int lx = D[k].nx + 2 * MBC, ly = D[k].ny + 2 * MBC;
    const int band = tile_band(ly, rows);
    for (int r0 = 1; r0 <= ly - 3; r0 += band) {
      /* strip j: lanes [60j, 60j+63], x-faces [60j+1, 60j+61]; need 60j+61 >= lx-3 */
      for (int a0 = 0;; a0 += STRIP) {
        TileDev d;
        d.blk = k;
        d.a0 = a0;
        d.r0 = r0;
        d.r1 = std::min(r0 + band - 1, ly - 3);
        t.push_back(d);
        if (a0 + 61 >= lx - 3) break;
      }
    }
  }
  return t;
}

/* Launch order of the tiles. A workgroup runs w (the wave group) consecutive tiles
 * (adjacent strips: they share the cache lines at strip edges and the halo
 * columns), and workgroup w goes to XCD w mod 8. The groups are dealt so
 * that each XCD receives one contiguous run of groups -- a patch of
 * neighbouring strips and row ranges -- keeping halo and edge lines in that
 * XCD's L2. (Ordering each XCD's run tallest tile first was measured slower
 * on the 256-deck sweep: 9.91 against 10.03 G cell-updates/s, the patches
 * lose their locality.) Padding entries (blk = -1) exit at once. */
std::vector<TileDev> xcd_order(const std::vector<TileDev> &t, int w)
{
  constexpr int NXCD = 8;
  const int G = (int)((t.size() + w - 1) / w);
  const int per = (G + NXCD - 1) / NXCD;
  std::vector<TileDev> o((size_t)per * NXCD * w);
  for (auto &d : o) {
    d.blk = -1;
    d.a0 = d.r0 = d.r1 = 0;
  }
  for (int g = 0; g < G; g++) {
    const int x = g / per, j = g % per, wg = j * NXCD + x;
    for (int k = 0; k < w; k++) {
      const size_t src = (size_t)g * w + k;
      if (src < t.size()) o[(size_t)wg * w + k] = t[src];
    }
  }
  return o;
}

/* Tile height for a deck: every tile costs rows + 4 (warm-up), and a SIMD
 * runs w = ceil(tiles / SIMDs) waves over the launch. With two waves resident
 * per SIMD a row costs 0.87 of its one-wave time (the other wave covers the
 * waits; measured on 4x512^2: 78.5 us at 10 rows and two waves per SIMD,
 * 74.1 us at 19 rows and one). Minimises w * (tallest tile + 4) * that
 * factor, with each block's rows split evenly (tile_band). */
int auto_tile_rows(const std::vector<mbm_block_desc> &D, int simds)
{
  int best = 32;
  double bestc = 1e300;
  int maxrows = 0;
  for (const auto &b : D) maxrows = std::max(maxrows, b.ny + 3);
  for (int R = 4; R <= 512; R++) {
    long long nt = 0;
    int tallest = 0;
    for (const auto &b : D) {
      const int lx = b.nx + 2 * MBC, band = tile_band(b.ny + 2 * MBC, R);
      int strips = 0;
      for (int a0 = 0;; a0 += STRIP) {
        strips++;
        if (a0 + 61 >= lx - 3) break;
      }
      nt += (long long)strips * ((b.ny + 3 + band - 1) / band);
      tallest = std::max(tallest, band);
    }
    const long long w = (nt + simds - 1) / simds;
    const double cost = (double)w * (tallest + 4) *
                        (w >= 2 && MBM_STAGE_WAVES >= 2 ? 0.87 : 1.0);
    if (cost < bestc - 1e-9) {
      bestc = cost;
      best = R;
    }
    if (R >= maxrows) break;
  }
  return best;
}

int relayout(mbm_ctx *c, int rows);
int build_push(mbm_ctx *c, const std::vector<TileDev> &t);

int upload_tiles(mbm_ctx *c)
{
  const int rows = c->tile_rows > 0 ? c->tile_rows
                                    : auto_tile_rows(c->desc, c->cus * 4);
  std::vector<TileDev> t;
  int grp[3] = { 0, 0, 0 };
  if (c->ngroups == 2) {
    /* tiles come in block order: group 0's blocks first */
    const std::vector<TileDev> all = build_tiles(c->desc, rows);
    const int b1 = c->deck_first[c->grp_deck[1]];
    std::vector<TileDev> t0, t1;
    for (const TileDev &d : all) (d.blk < b1 ? t0 : t1).push_back(d);
    t = xcd_order(t0, c->wpg);
    grp[1] = (int)t.size();
    const std::vector<TileDev> o1 = xcd_order(t1, c->wpg);
    t.insert(t.end(), o1.begin(), o1.end());
    grp[2] = (int)t.size();
  } else {
    t = xcd_order(build_tiles(c->desc, rows), c->wpg);
  }
  /* the new tile list is on the device before anything of the context
   * changes: a failure leaves tiles, layout and tile height as they were */
  TileDev *nt = nullptr;
  if (hipMalloc(&nt, t.size() * sizeof(TileDev)) != hipSuccess)
    return set_err(MBM_E_NOMEM, "device allocation of the tile list failed");
  if (hipMemcpy(nt, t.data(), t.size() * sizeof(TileDev), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(nt);
    return set_err(MBM_E_HIP, "upload of the tile list failed");
  }
  /* the layers follow the tile height, so that the rows marched together
   * stay adjacent */
  bool moved = false;
  for (const BlockDev &B : c->hblk) moved = moved || B.band != tile_band(B.ly, rows);
  if (moved) {
    const int rc = relayout(c, rows);
    if (rc) {
      (void)hipFree(nt);
      return rc;
    }
  }
  if (c->dtiles) (void)hipFree(c->dtiles);
  c->dtiles = nt;
  c->ntiles = (int)t.size();
  c->rows_used = rows;
  if (c->ngroups == 2)
    for (int g = 0; g < 3; g++) c->grp_tile[g] = grp[g];
  return build_push(c, t);
}

/* raises the dynamic LDS limit of every stage kernel instance to the largest
 * wave group (on the current device; called when a context is created) */
template <int STAGE, int LIM>
hipError_t raise_lds_limit()
{
  constexpr bool down = STAGE == 2 && MBM_STAGE2_DOWN;
  return hipFuncSetAttribute(reinterpret_cast<const void *>(&stage_kernel<STAGE, LIM, down>),
                             hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)stage_lds_bytes(MAX_WAVES_PER_WG));
}

int raise_stage_lds_limits()
{
  const hipError_t e[] = { raise_lds_limit<1, 0>(), raise_lds_limit<1, 1>(), raise_lds_limit<1, 2>(),
                           raise_lds_limit<1, 3>(), raise_lds_limit<1, 4>(), raise_lds_limit<2, 0>(),
                           raise_lds_limit<2, 1>(), raise_lds_limit<2, 2>(), raise_lds_limit<2, 3>(),
                           raise_lds_limit<2, 4>() };
  for (hipError_t x : e)
    if (x != hipSuccess)
      return set_err(MBM_E_HIP, "raising the LDS limit of the stage kernels failed: %s", hipGetErrorString(x));
  return MBM_OK;
}

/* launches one stage kernel instance with the context's wave group */
template <int STAGE, int LIM>
int launch_stage_lim(mbm_ctx *c, const StageArgs &A)
{
  constexpr bool down = STAGE == 2 && MBM_STAGE2_DOWN;
  const int w = c->wpg;
  const dim3 grid((A.ntiles + w - 1) / w), block(64 * w);
  hipLaunchKernelGGL((stage_kernel<STAGE, LIM, down>), grid, block, stage_lds_bytes(w), c->stream, A);
  HIPCHK(hipGetLastError());
  return MBM_OK;
}

template <int STAGE>
int launch_stage(mbm_ctx *c, const StageArgs &A)
{
  switch (c->par.limiter) {
    case 0: return launch_stage_lim<STAGE, 0>(c, A);
    case 1: return launch_stage_lim<STAGE, 1>(c, A);
    case 2: return launch_stage_lim<STAGE, 2>(c, A);
    case 3: return launch_stage_lim<STAGE, 3>(c, A);
    default: return launch_stage_lim<STAGE, 4>(c, A);
  }
}

int fill_buffer(mbm_ctx *c, double *layer)
{
  if (c->nghost == 0) return MBM_OK;
  const int nblk = (c->nghost + 255) / 256;
  hipLaunchKernelGGL(fill_kernel, dim3(nblk), dim3(256), 0, c->stream, layer, c->dfill,
                     c->dfillxy, c->nghost, c->par.scenario, c->dtj, (const DeckState *)nullptr,
                     (double *)nullptr, (double *)nullptr, (double *)nullptr, 0);
  HIPCHK(hipGetLastError());
  return MBM_OK;
}

int do_fill(mbm_ctx *c, int layer) { return fill_buffer(c, c->layer[c->role[layer]]); }

/* the ghost fill that follows a stage launch: with the stage kernel's ghost
 * copies only the scenario ghosts remain */
int fill_after_stage(mbm_ctx *c, double *layer)
{
  if (!c->push_on) return fill_buffer(c, layer);
  if (c->nresid == 0) return MBM_OK;
  hipLaunchKernelGGL(fill_kernel, dim3((c->nresid + 255) / 256), dim3(256), 0, c->stream, layer,
                     c->dresid, c->dresidxy, c->nresid, c->par.scenario, c->dtj,
                     (const DeckState *)nullptr, (double *)nullptr, (double *)nullptr,
                     (double *)nullptr, 0);
  HIPCHK(hipGetLastError());
  return MBM_OK;
}

/* the stage kernel's ghost copies over the whole tile list, when on */
void set_push(const mbm_ctx *c, StageArgs &A)
{
  if (c->push_on) {
    A.prange = c->dprange;
    A.push = c->dpush;
  }
}

int enqueue_stage(mbm_ctx *c, int stage, const double *in, const double *in0, double *out,
                  double dt, unsigned long long *slots)
{
  StageArgs A{};
  set_push(c, A);
  A.blocks = c->dblk;
  A.tiles = c->dtiles;
  A.ntiles = c->ntiles;
  A.dt = dt;
  A.dtkn = dt / c->par.kn;
  A.force = c->par.force;
  A.in = in;
  A.in0 = in0;
  A.out = out;
  A.slots = slots;
  return stage == 1 ? launch_stage<1>(c, A) : launch_stage<2>(c, A);
}

int upload_params(mbm_ctx *c)
{
  HIPCHK(hipMemcpyAsync(c->dtj, c->par.scenario_params, 5 * sizeof(double), hipMemcpyHostToDevice,
                        c->stream));
  return MBM_OK;
}

} // namespace mbmi

namespace mbmi {

/* Row order and offsets of the blocks of a layer for tiles of height rows
 * (each block's own tile_band; blocks 256-byte aligned); off returns the
 * layer's size in elements. */
int lay_out(std::vector<BlockDev> &hb, int rows, size_t &off)
{
  off = 0;
  for (size_t k = 0; k < hb.size(); k++) {
    BlockDev &b = hb[k];
    const int band = tile_band(b.ly, rows);
    if ((long long)(b.ly - 1) * band >= (1ll << 32))
      return set_err(MBM_E_ARG, "block %zu: %d rows too many for row bands of %d", k + 1, b.ly, band);
    set_block_layout(b, band);
    b.off = (long long)off;
    /* the kernels address a block with 32-bit byte offsets */
    if (8 * block_elems(b) + 256 >= (1ll << 31))
      return set_err(MBM_E_ARG, "block %zu too large (%dx%d)", k + 1, b.lx - 2 * MBC, b.ly - 2 * MBC);
    off += (size_t)((block_elems(b) + 31) & ~31ll);
  }
  return MBM_OK;
}

/* Ghost fill records of the context's blocks, in block order: element offsets
 * of component 0 of every ghost and its source. */
void fill_records(const mbm_ctx *c, std::vector<FillRec> &fr, std::vector<double2> &fxy)
{
  const std::vector<GhostDev> gm = build_ghost_map(c->desc, c->hblk);
  fr.resize(gm.size());
  fxy.resize(gm.size());
  for (size_t i = 0; i < gm.size(); i++) {
    const GhostDev &g = gm[i];
    const BlockDev &D = c->hblk[g.dblk];
    fr[i].dst = cell_elem(D, g.db, g.da);
    fr[i].dplane = (int)D.plane;
    fr[i].src = g.sblk >= 0 ? cell_elem(c->hblk[g.sblk], g.sb, g.sa) : -1;
    fr[i].splane = g.sblk >= 0 ? (int)c->hblk[g.sblk].plane : 0;
    fr[i].deck = D.deck;
    fr[i].sign = g.sign;
    fxy[i] = make_double2(g.x, g.y);
  }
}

/* The ghost copies of the stage kernel for the launch list t: each ghost
 * whose source is an interior cell goes to the tile of t that stores that
 * cell (strip (a - 2) / STRIP, row band (b - 1) / band, as build_tiles lays
 * them out), ghosts grouped by tile; scenario ghosts form the residual list
 * of the fill kernel. Without any ghost, or if the lists cannot be
 * allocated, the copies stay off and the fill kernels make them. */
int build_push(mbm_ctx *c, const std::vector<TileDev> &t)
{
  auto drop = [&]() {
    for (void *p : { (void *)c->dpush, (void *)c->dresid, (void *)c->dresidxy, (void *)c->dprange })
      if (p) (void)hipFree(p);
    c->dpush = c->dresid = nullptr;
    c->dresidxy = nullptr;
    c->dprange = nullptr;
    c->push_on = false;
  };
  drop();
#if MBM_PUSH
  if (c->nghost == 0) return MBM_OK;
  /* tile position by (block, row band, strip) */
  std::vector<int> base(c->nb + 1, 0), nstrip(c->nb);
  for (int k = 0; k < c->nb; k++) {
    const BlockDev &B = c->hblk[k];
    int ns = 1;
    while (STRIP * (ns - 1) + 61 < B.lx - 3) ns++;
    nstrip[k] = ns;
    base[k + 1] = base[k] + ns * ((B.ly - 3 + B.band - 1) / B.band);
  }
  std::vector<int> pos(base[c->nb], -1);
  for (size_t i = 0; i < t.size(); i++) {
    const TileDev &d = t[i];
    if (d.blk < 0) continue;
    const int q = base[d.blk] + ((d.r0 - 1) / c->hblk[d.blk].band) * nstrip[d.blk] + d.a0 / STRIP;
    if (q < base[d.blk + 1]) pos[q] = (int)i;
  }
  const std::vector<GhostDev> gm = build_ghost_map(c->desc, c->hblk);
  std::vector<int> owner(gm.size(), -1), cnt(t.size() + 1, 0);
  std::vector<FillRec> resid;
  std::vector<double2> residxy;
  int b1 = c->nb;
  if (c->ngroups == 2) b1 = c->deck_first[c->grp_deck[1]];
  c->grp_resid[0] = 0;
  c->grp_resid[1] = 0;
  auto rec = [&](const GhostDev &g) {
    const BlockDev &D = c->hblk[g.dblk];
    FillRec r;
    r.dst = cell_elem(D, g.db, g.da);
    r.dplane = (int)D.plane;
    r.src = g.sblk >= 0 ? cell_elem(c->hblk[g.sblk], g.sb, g.sa) : -1;
    r.splane = g.sblk >= 0 ? (int)c->hblk[g.sblk].plane : 0;
    r.deck = D.deck;
    r.sign = g.sign;
    return r;
  };
  for (size_t i = 0; i < gm.size(); i++) {
    const GhostDev &g = gm[i];
    if (g.sblk < 0) {
      resid.push_back(rec(g));
      residxy.push_back(make_double2(g.x, g.y));
      if (g.dblk < b1) c->grp_resid[1] = (int)resid.size();
      continue;
    }
    /* every source is an interior cell stored by exactly one tile; if one
     * were not, the fill kernel keeps every ghost */
    const BlockDev &S = c->hblk[g.sblk];
    if (g.sa < MBC || g.sa > S.lx - 1 - MBC || g.sb < MBC || g.sb > S.ly - 1 - MBC) return MBM_OK;
    const int p = pos[base[g.sblk] + ((g.sb - 1) / S.band) * nstrip[g.sblk] + (g.sa - 2) / STRIP];
    if (p < 0 || t[p].blk != g.sblk || g.sa < t[p].a0 + 2 || g.sa > t[p].a0 + 61 || g.sb < t[p].r0 ||
        g.sb > t[p].r1)
      return MBM_OK;
    owner[i] = p;
    cnt[p + 1]++;
  }
  c->grp_resid[2] = (int)resid.size();
  if (c->ngroups != 2) c->grp_resid[1] = c->grp_resid[2];
  for (size_t i = 0; i < t.size(); i++) cnt[i + 1] += cnt[i];
  std::vector<int2> range(t.size());
  for (size_t i = 0; i < t.size(); i++) range[i] = make_int2(cnt[i], cnt[i + 1]);
  std::vector<FillRec> push((size_t)cnt[t.size()]);
  for (size_t i = 0; i < gm.size(); i++)
    if (owner[i] >= 0) push[cnt[owner[i]]++] = rec(gm[i]);
  c->nresid = (int)resid.size();
  auto up = [&](void **d, const void *h, size_t n) {
    return hipMalloc(d, std::max<size_t>(n, 1)) == hipSuccess &&
           (n == 0 || hipMemcpy(*d, h, n, hipMemcpyHostToDevice) == hipSuccess);
  };
  if (!up((void **)&c->dpush, push.data(), push.size() * sizeof(FillRec)) ||
      !up((void **)&c->dprange, range.data(), range.size() * sizeof(int2)) ||
      !up((void **)&c->dresid, resid.data(), resid.size() * sizeof(FillRec)) ||
      !up((void **)&c->dresidxy, residxy.data(), residxy.size() * sizeof(double2))) {
    /* the fill kernels do the copies instead; the tile list and layout
     * stay as they are */
    (void)hipGetLastError();
    drop();
    return MBM_OK;
  }
  c->push_on = true;
#endif
  return MBM_OK;
}

/* Moves the three layers to the row order of tile height rows; the ghost
 * records follow, captured graphs and adjoint buffers are dropped. */
int relayout(mbm_ctx *c, int rows)
{
  std::vector<BlockDev> nh = c->hblk;
  size_t off = 0;
  int rc = lay_out(nh, rows, off);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  if (c->stream2) HIPCHK(hipStreamSynchronize(c->stream2));
  /* all three new layers first: on failure the context keeps its layout */
  double *nl[3] = { nullptr, nullptr, nullptr };
  for (int l = 0; l < 3; l++)
    if (hipMalloc(&nl[l], off * sizeof(double)) != hipSuccess) {
      for (int q = 0; q < l; q++) (void)hipFree(nl[q]);
      return set_err(MBM_E_NOMEM, "device allocation of %zu bytes failed", off * sizeof(double));
    }
  auto drop_new = [&](int code) {
    (void)hipStreamSynchronize(c->stream);
    for (int q = 0; q < 3; q++) (void)hipFree(nl[q]);
    return code;
  };
  if (c->adj && (rc = mbm_adj_release(c))) return drop_new(rc);
  release_graphs(c);
  for (int l = 0; l < 3; l++) {
    hipError_t e = hipMemsetAsync(nl[l], 0, off * sizeof(double), c->stream);
    for (int k = 0; k < c->nb && e == hipSuccess; k++) {
      const BlockDev &B = c->hblk[k];
      const int n = B.ly * NV * B.lx;
      hipLaunchKernelGGL(relayout_kernel, dim3((n + 255) / 256), dim3(256), 0, c->stream, nl[l],
                         c->layer[l], nh[k], B, n);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess)
      return drop_new(set_err(MBM_E_HIP, "relayout failed: %s", hipGetErrorString(e)));
  }
  for (int l = 0; l < 3; l++) {
    (void)hipFree(c->layer[l]);
    c->layer[l] = nl[l];
  }
  c->hblk = nh;
  c->layer_elems = off;
  HIPCHK(hipMemcpy(c->dblk, c->hblk.data(), c->nb * sizeof(BlockDev), hipMemcpyHostToDevice));
  std::vector<FillRec> fr;
  std::vector<double2> fxy;
  fill_records(c, fr, fxy);
  if (!fr.empty())
    HIPCHK(hipMemcpy(c->dfill, fr.data(), fr.size() * sizeof(FillRec), hipMemcpyHostToDevice));
  return MBM_OK;
}

/* Device state of one or more decks whose blocks are listed back to back;
 * deck d owns blocks [deck_first[d], deck_first[d+1]) and its sync codes are
 * already global (1-based over all blocks). */
mbm_ctx *create_ctx(const mbm_block_desc *blocks, int nblocks, const int *deck_first, int ndecks,
                    const mbm_params *deck_par, void *stream, int *rcp)
{
  int &rc = *rcp;
  rc = MBM_OK;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    rc = set_err(MBM_E_NODEV, "no HIP device available");
    return nullptr;
  }
  mbm_ctx *c = new mbm_ctx();
  auto fail = [&](int code) { rc = code; mbm_destroy(c); return (mbm_ctx *)nullptr; };
  auto hipok = [&](hipError_t e, const char *what) {
    if (e != hipSuccess) rc = set_err(MBM_E_HIP, "%s failed: %s", what, hipGetErrorString(e));
    return e == hipSuccess;
  };
  c->nb = nblocks;
  c->desc.assign(blocks, blocks + nblocks);
  c->par = deck_par[0];
  c->tile_rows = 0;                      /* automatic */
  c->wpg = WAVES_PER_WG;
  c->cus = 256;
  {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
      c->cus = cus;
  }
  c->ndecks = ndecks;
  c->deck_first.assign(deck_first, deck_first + ndecks + 1);
  c->deck_par.assign(deck_par, deck_par + ndecks);
  if (!hipok(hipGetDevice(&c->device), "hipGetDevice")) return fail(rc);
  if ((rc = resolve_scratch(c->device))) return fail(rc);
  if ((rc = raise_stage_lds_limits())) return fail(rc);
  for (int d = 0; d < ndecks; d++)
    for (int k = deck_first[d]; k < deck_first[d + 1]; k++) {
      BlockDev b{};
      b.lx = blocks[k].nx + 2 * MBC;
      b.ly = blocks[k].ny + 2 * MBC;
      b.deck = d;
      geometry(blocks[k], b.bx, b.by, b.dx, b.dy);
      c->hblk.push_back(b);
    }
  /* rows ordered for the automatic tile height */
  size_t off = 0;
  if ((rc = lay_out(c->hblk, auto_tile_rows(c->desc, c->cus * 4), off))) return fail(rc);
  c->layer_elems = off;
  if (stream) {
    c->stream = (hipStream_t)stream;
    c->own_stream = false;
  } else {
    if (!hipok(hipStreamCreate(&c->stream), "hipStreamCreate")) return fail(rc);
    c->own_stream = true;
  }
  for (int l = 0; l < 3; l++) {
    if (hipMalloc(&c->layer[l], off * sizeof(double)) != hipSuccess) {
      c->layer[l] = nullptr;
      return fail(set_err(MBM_E_NOMEM, "device allocation of %zu bytes failed", off * sizeof(double)));
    }
    if (!hipok(hipMemsetAsync(c->layer[l], 0, off * sizeof(double), c->stream), "hipMemsetAsync"))
      return fail(rc);
    c->role[l] = l;
  }
  std::vector<FillRec> fr;
  std::vector<double2> fxy;
  fill_records(c, fr, fxy);
  c->nghost = (int)fr.size();
  const bool ok =
      hipok(hipMalloc(&c->dblk, nblocks * sizeof(BlockDev)), "hipMalloc") &&
      hipok(hipMemcpy(c->dblk, c->hblk.data(), nblocks * sizeof(BlockDev), hipMemcpyHostToDevice), "hipMemcpy") &&
      hipok(hipMalloc(&c->dfill, std::max<size_t>(1, fr.size()) * sizeof(FillRec)), "hipMalloc") &&
      (fr.empty() || hipok(hipMemcpy(c->dfill, fr.data(), fr.size() * sizeof(FillRec), hipMemcpyHostToDevice), "hipMemcpy")) &&
      hipok(hipMalloc(&c->dfillxy, std::max<size_t>(1, fxy.size()) * sizeof(double2)), "hipMalloc") &&
      (fxy.empty() || hipok(hipMemcpy(c->dfillxy, fxy.data(), fxy.size() * sizeof(double2), hipMemcpyHostToDevice), "hipMemcpy")) &&
      hipok(hipMalloc(&c->dslots, (size_t)2 * SLOTS_PER_BLOCK * nblocks * sizeof(unsigned long long)), "hipMalloc") &&
      hipok(hipMemsetAsync(c->dslots, 0, (size_t)2 * SLOTS_PER_BLOCK * nblocks * sizeof(unsigned long long), c->stream), "hipMemsetAsync") &&
      hipok(hipMalloc(&c->dcfl, sizeof(double)), "hipMalloc") &&
      hipok(hipMalloc(&c->dtj, 5 * sizeof(double)), "hipMalloc") &&
      hipok(hipHostMalloc(&c->hcfl, sizeof(double)), "hipHostMalloc") &&
      hipok(hipMalloc(&c->ddeck_first, (ndecks + 1) * sizeof(int)), "hipMalloc") &&
      hipok(hipMemcpy(c->ddeck_first, deck_first, (ndecks + 1) * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy") &&
      hipok(hipMalloc(&c->ddecks, ndecks * sizeof(DeckState)), "hipMalloc");
  if (!ok) return fail(rc);
  for (int i = 0; i < 4; i++)
    if (!hipok(hipEventCreate(&c->ev[i]), "hipEventCreate")) return fail(rc);
  if ((rc = upload_params(c)) || (rc = upload_tiles(c))) return fail(rc);
  if (!hipok(hipStreamSynchronize(c->stream), "hipStreamSynchronize")) return fail(rc);
  return c;
}

void release_graphs(mbm_ctx *c)
{
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      if (c->graph_ok[i][j]) {
        (void)hipGraphExecDestroy(c->graph[i][j]);
        c->graph_ok[i][j] = false;
      }
      if (c->graphk_ok[i][j]) {
        (void)hipGraphExecDestroy(c->graphk[i][j]);
        c->graphk_ok[i][j] = false;
      }
    }
  if (c->graph2_ok) {
    (void)hipGraphExecDestroy(c->graph2);
    c->graph2_ok = false;
  }
  if (c->graph2k_ok) {
    (void)hipGraphExecDestroy(c->graph2k);
    c->graph2k_ok = false;
  }
}

namespace {

/* Pieces of one device-controlled attempt of group g (g < 0: every deck):
 * stage s with the fill that follows it, and the controller. */
int enqueue_group_stage(mbm_ctx *c, int g, int stage, double *buf0, double *buf1, double *tmp)
{
  StageArgs A{};
  const int t0 = g < 0 ? 0 : c->grp_tile[g], t1 = g < 0 ? c->ntiles : c->grp_tile[g + 1];
  const int h0 = g < 0 ? 0 : c->grp_ghost[g], h1 = g < 0 ? c->nghost : c->grp_ghost[g + 1];
  A.blocks = c->dblk;
  A.tiles = c->dtiles + t0;
  A.ntiles = t1 - t0;
  A.decks = c->ddecks;
  A.buf0 = buf0;
  A.buf1 = buf1;
  A.tmp = tmp;
  A.slots = c->dslots + (stage == 1 ? 0 : SLOTS_PER_BLOCK * c->nb);
  if (c->push_on) {
    A.prange = c->dprange + t0;
    A.push = c->dpush;
  }
  int rc;
  if ((rc = stage == 1 ? launch_stage<1>(c, A) : launch_stage<2>(c, A))) return rc;
  if (c->push_on) {
    /* the stage kernel copied the interior-sourced ghosts; the scenario
     * ghosts remain */
    const int r0 = g < 0 ? 0 : c->grp_resid[g], r1 = g < 0 ? c->nresid : c->grp_resid[g + 1];
    if (r1 > r0)
      hipLaunchKernelGGL(fill_kernel, dim3((r1 - r0 + 255) / 256), dim3(256), 0, c->stream,
                         (double *)nullptr, c->dresid + r0, c->dresidxy + r0, r1 - r0, c->par.scenario,
                         c->dtj, (const DeckState *)c->ddecks, buf0, buf1, tmp, stage);
  } else if (h1 > h0) {
    hipLaunchKernelGGL(fill_kernel, dim3((h1 - h0 + 255) / 256), dim3(256), 0, c->stream,
                       (double *)nullptr, c->dfill + h0, c->dfillxy + h0, h1 - h0, c->par.scenario, c->dtj,
                       (const DeckState *)c->ddecks, buf0, buf1, tmp, stage);
  }
  HIPCHK(hipGetLastError());
  return MBM_OK;
}

int enqueue_group_control(mbm_ctx *c, int g, const double *buf0, const double *buf1, const double *tmp)
{
  const int d0 = g < 0 ? 0 : c->grp_deck[g], d1 = g < 0 ? c->ndecks : c->grp_deck[g + 1];
  hipLaunchKernelGGL(control_kernel, dim3(d1 - d0), dim3(64), 0, c->stream,
                     c->ddecks + d0, c->dslots, c->dblk, c->ddeck_first + d0, c->nb,
                     g > 0 ? (RecordDev *)nullptr : c->drec, buf0, buf1, tmp);
  HIPCHK(hipGetLastError());
  return MBM_OK;
}

/* one device-controlled attempt of every deck: 2 stages, 2 fills, controller */
int enqueue_device_step(mbm_ctx *c, double *buf0, double *buf1, double *tmp)
{
  int rc;
  if ((rc = enqueue_group_stage(c, -1, 1, buf0, buf1, tmp))) return rc;
  if ((rc = enqueue_group_stage(c, -1, 2, buf0, buf1, tmp))) return rc;
  return enqueue_group_control(c, -1, buf0, buf1, tmp);
}

} // namespace

/* stages and fills of every deck at its current dt, without the controller
 * (writes TMP and NEW only; used to time tile shapes). The fills belong to
 * the timed work, and without them stage 2 would read stale TMP ghosts. A
 * two-group batch runs its groups on their two streams, as its attempts do;
 * group 1's stage 2 then reads the TMP of its stage 1 in the previous call. */
int enqueue_device_stages(mbm_ctx *c)
{
  double *buf0 = c->layer[c->role[MBM_L_DATA]], *buf1 = c->layer[c->role[MBM_L_NEW]];
  double *tmp = c->layer[c->role[MBM_L_TMP]];
  int rc;
  if (c->ngroups == 2) {
    /* the two groups on their two streams, as a batched attempt runs them
     * (group 1 half a step behind: its stage 2, then its stage 1); the
     * context's stream waits for both */
    HIPCHK(hipEventRecord(c->grp_ev, c->stream));
    HIPCHK(hipStreamWaitEvent(c->stream2, c->grp_ev, 0));
    if ((rc = enqueue_group_stage(c, 0, 1, buf0, buf1, tmp))) return rc;
    hipStream_t keep = c->stream;
    c->stream = c->stream2;
    rc = enqueue_group_stage(c, 1, 2, buf0, buf1, tmp);
    if (!rc) rc = enqueue_group_stage(c, 1, 1, buf0, buf1, tmp);
    c->stream = keep;
    if (rc) return rc;
    if ((rc = enqueue_group_stage(c, 0, 2, buf0, buf1, tmp))) return rc;
    HIPCHK(hipEventRecord(c->grp_ev, c->stream2));
    HIPCHK(hipStreamWaitEvent(c->stream, c->grp_ev, 0));
    return MBM_OK;
  }
  if ((rc = enqueue_group_stage(c, -1, 1, buf0, buf1, tmp))) return rc;
  return enqueue_group_stage(c, -1, 2, buf0, buf1, tmp);
}

int clear_slots(mbm_ctx *c)
{
  const int n = 2 * SLOTS_PER_BLOCK * c->nb;
  hipLaunchKernelGGL(zero_kernel, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->dslots, n);
  HIPCHK(hipGetLastError());
  return MBM_OK;
}

/* Runs the controllers of all decks on the device until every deck is done:
 * the step (2 stages, 2 fills, controller) is captured once in a hipGraph
 * and replayed in chunks; the host reads the deck states between chunks.
 * The current DATA/NEW roles are buf0/buf1 at parity 0; on return the roles
 * of the physical buffers are updated from deck 0's parity (a plain context
 * has one deck; batches keep per-deck parity in ddecks). */
/* capture what 'enq' enqueues on stream s into an executable graph */
template <class F>
int capture(mbm_ctx *c, hipStream_t s, hipGraphExec_t *out, F enq)
{
  hipGraph_t g;
  hipStream_t keep = c->stream;
  c->stream = s;                  /* the enqueue helpers launch on c->stream */
  HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  int rc = enq();
  hipError_t e = hipStreamEndCapture(s, &g);
  c->stream = keep;
  if (rc) return rc;
  if (e != hipSuccess) return set_err(MBM_E_HIP, "graph capture failed: %s", hipGetErrorString(e));
  e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return set_err(MBM_E_HIP, "graph instantiation failed: %s", hipGetErrorString(e));
  /* move the executable graph to the device now, not at its first launch */
  e = hipGraphUpload(*out, s);
  if (e != hipSuccess) return set_err(MBM_E_HIP, "graph upload failed: %s", hipGetErrorString(e));
  return MBM_OK;
}

/* the captured attempts (1, or GRAPH_ATTEMPTS with long) for DATA in buffer
 * iD and NEW in buffer iN; the buffers of each attempt follow the deck
 * states on the device, so one graph serves any number of attempts */
int step_graph(mbm_ctx *c, int iD, int iN, int iT, bool lng)
{
  bool &ok = lng ? c->graphk_ok[iD][iN] : c->graph_ok[iD][iN];
  if (ok) return MBM_OK;
  hipGraphExec_t &ge = lng ? c->graphk[iD][iN] : c->graph[iD][iN];
  double *buf0 = c->layer[iD], *buf1 = c->layer[iN], *tmp = c->layer[iT];
  const int n = lng ? GRAPH_ATTEMPTS : 1;
  int rc;
  if (c->ngroups == 2)   /* group 0: stage 1, stage 2, controller */
    rc = capture(c, c->stream, &ge, [&] {
      int r = 0;
      for (int i = 0; i < n && !r; i++) {
        r = enqueue_group_stage(c, 0, 1, buf0, buf1, tmp);
        if (!r) r = enqueue_group_stage(c, 0, 2, buf0, buf1, tmp);
        if (!r) r = enqueue_group_control(c, 0, buf0, buf1, tmp);
      }
      return r;
    });
  else
    rc = capture(c, c->stream, &ge, [&] {
      int r = 0;
      for (int i = 0; i < n && !r; i++) r = enqueue_device_step(c, buf0, buf1, tmp);
      return r;
    });
  if (rc) return rc;
  ok = true;
  return MBM_OK;
}

/* group 1 of a two-group batch, half an attempt later: stage 2,
 * controller, next stage 1 (1 or GRAPH_ATTEMPTS times) */
int group1_graph(mbm_ctx *c, double *buf0, double *buf1, double *tmp, bool lng)
{
  bool &ok = lng ? c->graph2k_ok : c->graph2_ok;
  if (ok) return MBM_OK;
  const int n = lng ? GRAPH_ATTEMPTS : 1;
  int rc = capture(c, c->stream2, lng ? &c->graph2k : &c->graph2, [&] {
    int r = 0;
    for (int i = 0; i < n && !r; i++) {
      r = enqueue_group_stage(c, 1, 2, buf0, buf1, tmp);
      if (!r) r = enqueue_group_control(c, 1, buf0, buf1, tmp);
      if (!r) r = enqueue_group_stage(c, 1, 1, buf0, buf1, tmp);
    }
    return r;
  });
  if (rc) return rc;
  ok = true;
  return MBM_OK;
}

int device_steps(mbm_ctx *c, int maxchunk, const DeckState *known, DeckState *last)
{
  const int iD = c->role[MBM_L_DATA], iN = c->role[MBM_L_NEW], iT = c->role[MBM_L_TMP];
  double *buf0 = c->layer[iD], *buf1 = c->layer[iN], *tmp = c->layer[iT];
  const bool two = c->ngroups == 2;
  int rc;
  for (int lng = 0; lng < 2; lng++) {
    if ((rc = step_graph(c, iD, iN, iT, lng))) return rc;
    /* a plain context swaps the DATA and NEW roles after an odd number of
     * accepted steps: the graphs of the swapped assignment are captured now,
     * so that no later call captures one between its launches */
    if (!c->batch && (rc = step_graph(c, iN, iD, iT, lng))) return rc;
    if (two && (rc = group1_graph(c, buf0, buf1, tmp, lng))) return rc;
  }
  std::vector<DeckState> st(c->ndecks);
  if (known) {
    /* the caller has just written these states to the device */
    st.assign(known, known + c->ndecks);
  } else {
    HIPCHK(hipMemcpyAsync(st.data(), c->ddecks, c->ndecks * sizeof(DeckState), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  if (two) HIPCHK(hipStreamSynchronize(c->stream));
  if (two) {
    /* group 1's first stage 1 (the stream is idle: the host has synchronised
     * with everything queued on c->stream) */
    hipStream_t keep = c->stream;
    c->stream = c->stream2;
    rc = enqueue_group_stage(c, 1, 1, buf0, buf1, tmp);
    c->stream = keep;
    if (rc) return rc;
  }
  for (;;) {
    int remaining = 0;
    for (const DeckState &S : st)
      if (!S.done) {
        /* attempts left: bounded by maxsteps, estimated from (tend-t)/dt */
        double est = S.dt > 0 ? (S.tend - S.t) / S.dt : 1.0;
        int r = est > 1e6 ? 1000000 : (int)est + 1;
        if (S.maxsteps > 0) r = std::min(r, S.maxsteps - S.nstep);
        remaining = std::max(remaining, std::max(r, 1));
      }
    if (remaining == 0) break;
    /* whole runs of GRAPH_ATTEMPTS, then single attempts: no attempt is
     * launched beyond the estimate */
    const int chunk = std::min(remaining, maxchunk);
    for (int i = 0; i < chunk; ) {
      const bool lng = chunk - i >= GRAPH_ATTEMPTS;
      HIPCHK(hipGraphLaunch(lng ? c->graphk[iD][iN] : c->graph[iD][iN], c->stream));
      if (two) HIPCHK(hipGraphLaunch(lng ? c->graph2k : c->graph2, c->stream2));
      i += lng ? GRAPH_ATTEMPTS : 1;
    }
    if (two) {
      HIPCHK(hipEventRecord(c->grp_ev, c->stream2));
      HIPCHK(hipStreamWaitEvent(c->stream, c->grp_ev, 0));
    }
    HIPCHK(hipMemcpyAsync(st.data(), c->ddecks, c->ndecks * sizeof(DeckState), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  if (!c->batch && st[0].parity) std::swap(c->role[MBM_L_DATA], c->role[MBM_L_NEW]);
  if (last) std::copy(st.begin(), st.end(), last);
  return MBM_OK;
}

} // namespace mbmi

extern "C" int mbm_create(const mbm_block_desc *blocks, int nblocks, const mbm_params *p,
                          void *stream, mbm_ctx **out)
{
  if (!out || !blocks) return set_err(MBM_E_ARG, "NULL argument");
  *out = nullptr;
  int rc = validate(blocks, nblocks, p);
  if (rc) return rc;
  const int first[2] = { 0, nblocks };
  mbm_ctx *c = create_ctx(blocks, nblocks, first, 1, p, stream, &rc);
  if (!c) return rc;
  *out = c;
  return MBM_OK;
}

extern "C" int mbm_destroy(mbm_ctx *c)
{
  if (!c) return MBM_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->stream2) (void)hipStreamSynchronize(c->stream2);
  if (c->adj) mbm_adj_release(c);
  release_graphs(c);
  for (int l = 0; l < 3; l++) if (c->layer[l]) (void)hipFree(c->layer[l]);
  if (c->dblk) (void)hipFree(c->dblk);
  if (c->dtiles) (void)hipFree(c->dtiles);
  if (c->dfill) (void)hipFree(c->dfill);
  if (c->dfillxy) (void)hipFree(c->dfillxy);
  for (void *p : { (void *)c->dpush, (void *)c->dresid, (void *)c->dresidxy, (void *)c->dprange })
    if (p) (void)hipFree(p);
  if (c->dslots) (void)hipFree(c->dslots);
  if (c->dcfl) (void)hipFree(c->dcfl);
  if (c->dtj) (void)hipFree(c->dtj);
  if (c->hcfl) (void)hipHostFree(c->hcfl);
  if (c->ddeck_first) (void)hipFree(c->ddeck_first);
  if (c->ddecks) (void)hipFree(c->ddecks);
  if (c->drec) (void)hipFree(c->drec);
  for (int i = 0; i < 4; i++) if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
  if (c->stream2) (void)hipStreamDestroy(c->stream2);
  if (c->grp_ev) (void)hipEventDestroy(c->grp_ev);
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return MBM_OK;
}

extern "C" int mbm_set_stream(mbm_ctx *c, void *stream)
{
  if (!c) return set_err(MBM_E_ARG, "NULL context");
  HIPCHK(hipStreamSynchronize(c->stream));
  release_graphs(c);
  if (c->own_stream) HIPCHK(hipStreamDestroy(c->stream));
  if (stream) { c->stream = (hipStream_t)stream; c->own_stream = false; }
  else { HIPCHK(hipStreamCreate(&c->stream)); c->own_stream = true; }
  return MBM_OK;
}

extern "C" int mbm_set_params(mbm_ctx *c, const mbm_params *p)
{
  if (!c) return set_err(MBM_E_ARG, "NULL context");
  int rc = validate(c->desc.data(), c->nb, p);
  if (rc) return rc;
  if (p->limiter != c->par.limiter) release_graphs(c);
  c->par = *p;
  for (auto &dp : c->deck_par) dp = *p;
  return upload_params(c);
}

extern "C" int mbm_tile_rows(const mbm_ctx *c) { return c ? c->rows_used : 0; }

extern "C" int mbm_set_tile_rows(mbm_ctx *c, int rows)
{
  if (!c) return set_err(MBM_E_ARG, "NULL context");
  if (rows < 0 || rows > 4096) return set_err(MBM_E_ARG, "tile rows %d outside [0,4096]", rows);
  const int keep = c->tile_rows;
  c->tile_rows = rows;
  HIPCHK(hipStreamSynchronize(c->stream));
  release_graphs(c);
  const int rc = upload_tiles(c);
  if (rc) c->tile_rows = keep;   /* the layers are unchanged when the move failed */
  return rc;
}

extern "C" int mbm_set_wave_group(mbm_ctx *c, int waves)
{
  if (!c) return set_err(MBM_E_ARG, "NULL context");
  if (waves != 1 && waves != 2 && waves != 4 && waves != MAX_WAVES_PER_WG)
    return set_err(MBM_E_ARG, "wave group %d not one of 1, 2, 4, %d", waves, MAX_WAVES_PER_WG);
  if (waves == c->wpg) return MBM_OK;
  const int keep = c->wpg;
  HIPCHK(hipStreamSynchronize(c->stream));
  if (c->stream2) HIPCHK(hipStreamSynchronize(c->stream2));
  release_graphs(c);
  c->wpg = waves;
  const int rc = upload_tiles(c);
  if (rc) c->wpg = keep;         /* the tile list is unchanged when the upload failed */
  return rc;
}

extern "C" int mbm_wave_group(const mbm_ctx *c) { return c ? c->wpg : 0; }

extern "C" int mbm_init(mbm_ctx *c)
{
  if (!c) return set_err(MBM_E_ARG, "NULL context");
  c->role[0] = 0; c->role[1] = 1; c->role[2] = 2;
  std::fill(c->deck_parity.begin(), c->deck_parity.end(), 0);
  for (int l = 0; l < 3; l++)
    HIPCHK(hipMemsetAsync(c->layer[l], 0, c->layer_elems * sizeof(double), c->stream));
  for (int k = 0; k < c->nb; k++) {
    const int n = c->desc[k].nx * c->desc[k].ny;
    hipLaunchKernelGGL(init_kernel, dim3((n + 255) / 256), dim3(256), 0, c->stream,
                       c->layer[c->role[MBM_L_DATA]], c->dblk, k, c->desc[k].nx, c->desc[k].ny,
                       c->par.scenario, c->dtj);
    HIPCHK(hipGetLastError());
  }
  int rc = do_fill(c, MBM_L_DATA);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  return MBM_OK;
}

namespace {
int check_lb(mbm_ctx *c, int layer, int block)
{
  if (!c) return set_err(MBM_E_ARG, "NULL context");
  if (layer < 0 || layer > 2) return set_err(MBM_E_ARG, "layer %d invalid", layer);
  if (block < 0 || block >= c->nb) return set_err(MBM_E_ARG, "block %d invalid", block);
  return MBM_OK;
}
} // namespace

extern "C" int mbm_upload(mbm_ctx *c, int layer, int block, const double *host)
{
  int rc = check_lb(c, layer, block);
  if (rc) return rc;
  const BlockDev &B = c->hblk[block];
  std::vector<double> soa((size_t)block_elems(B), 0.0);
  for (int b = 0; b < B.ly; b++)
    for (int a = 0; a < B.lx; a++)
      for (int m = 0; m < NV; m++)
        soa[(size_t)(row_elem(B, b) + m * B.plane + a)] = host[(size_t)NV * ((size_t)b * B.lx + a) + m];
  HIPCHK(hipMemcpyAsync(c->layer[c->role[layer]] + B.off, soa.data(), soa.size() * sizeof(double),
                        hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MBM_OK;
}

extern "C" int mbm_download(mbm_ctx *c, int layer, int block, double *host)
{
  int rc = check_lb(c, layer, block);
  if (rc) return rc;
  const BlockDev &B = c->hblk[block];
  std::vector<double> soa((size_t)block_elems(B));
  HIPCHK(hipMemcpyAsync(soa.data(), c->layer[c->role[layer]] + B.off, soa.size() * sizeof(double),
                        hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int b = 0; b < B.ly; b++)
    for (int a = 0; a < B.lx; a++)
      for (int m = 0; m < NV; m++)
        host[(size_t)NV * ((size_t)b * B.lx + a) + m] = soa[(size_t)(row_elem(B, b) + m * B.plane + a)];
  return MBM_OK;
}

extern "C" int mbm_layer_ptr(mbm_ctx *c, int layer, int block, void **dptr, long long *plane,
                             int *pitch)
{
  int rc = check_lb(c, layer, block);
  if (rc) return rc;
  const BlockDev &B = c->hblk[block];
  if (dptr) *dptr = c->layer[c->role[layer]] + B.off;
  if (plane) *plane = B.plane;
  if (pitch) *pitch = NV * B.pitch;
  return MBM_OK;
}

extern "C" int mbm_layer_rows(mbm_ctx *c, int block, int *pos, int n)
{
  int rc = check_lb(c, MBM_L_DATA, block);
  if (rc) return rc;
  if (!pos || n < 0) return set_err(MBM_E_ARG, "invalid row buffer");
  const BlockDev &B = c->hblk[block];
  for (int b = 0; b < n && b < B.ly; b++) pos[b] = row_pos(B, b);
  return MBM_OK;
}

extern "C" int mbm_row_order(int ny, int band, int *pos)
{
  if (ny < 1 || band < 0 || band > 4096 || !pos) return set_err(MBM_E_ARG, "invalid argument");
  BlockDev B{};
  B.lx = 1;
  B.ly = ny + 2 * MBC;
  set_block_layout(B, std::max(band, 2));
  for (int b = 0; b < B.ly; b++) pos[b] = row_pos(B, b);
  return B.npos;
}

extern "C" int mbm_set_boundary(mbm_ctx *c, int layer, double t)
{
  int rc = check_lb(c, layer, 0);
  if (rc) return rc;
  (void)t; /* the scenarios are steady: Dirichlet values equal the initial state */
  rc = do_fill(c, layer);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  return MBM_OK;
}

namespace mbmi {

/* enqueues one TimeUpdate; the combined CFL lands in c->dcfl */
int enqueue_time_update(mbm_ctx *c, double dt)
{
  StageArgs A{};
  set_push(c, A);
  A.blocks = c->dblk;
  A.tiles = c->dtiles;
  A.ntiles = c->ntiles;
  A.dt = dt;
  A.dtkn = dt / c->par.kn;
  A.force = c->par.force;
  A.in = c->layer[c->role[MBM_L_DATA]];
  A.in0 = A.in;
  A.out = c->layer[c->role[MBM_L_TMP]];
  A.slots = c->dslots;
  int rc;
  if (c->timing) HIPCHK(hipEventRecord(c->ev[0], c->stream));
  if ((rc = launch_stage<1>(c, A))) return rc;
  if (c->timing) HIPCHK(hipEventRecord(c->ev[1], c->stream));
  if ((rc = fill_after_stage(c, c->layer[c->role[MBM_L_TMP]]))) return rc;
  A.in = c->layer[c->role[MBM_L_TMP]];
  A.in0 = c->layer[c->role[MBM_L_DATA]];
  A.out = c->layer[c->role[MBM_L_NEW]];
  A.slots = c->dslots + SLOTS_PER_BLOCK * c->nb;
  if (c->timing) HIPCHK(hipEventRecord(c->ev[2], c->stream));
  if ((rc = launch_stage<2>(c, A))) return rc;
  if (c->timing) HIPCHK(hipEventRecord(c->ev[3], c->stream));
  if ((rc = fill_after_stage(c, c->layer[c->role[MBM_L_NEW]]))) return rc;
  hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(64), 0, c->stream, c->dslots, c->dblk, c->nb,
                     dt, c->par.kn, c->par.force, c->layer[c->role[MBM_L_DATA]],
                     c->layer[c->role[MBM_L_TMP]], c->dcfl);
  HIPCHK(hipGetLastError());
  return MBM_OK;
}

int read_cfl(mbm_ctx *c, double *cfl)
{
  HIPCHK(hipMemcpyAsync(c->hcfl, c->dcfl, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (c->timing) {
    float ms1 = 0.f, ms2 = 0.f;
    HIPCHK(hipEventElapsedTime(&ms1, c->ev[0], c->ev[1]));
    HIPCHK(hipEventElapsedTime(&ms2, c->ev[2], c->ev[3]));
    float msf = 0.f;
    HIPCHK(hipEventElapsedTime(&msf, c->ev[1], c->ev[2]));
    c->stage_ms += ms1 + ms2;
    c->nstage += 2;
    c->fill_ms += msf;            /* the TMP fill between the two stages */
    c->nfill += 1;
  }
  *cfl = *c->hcfl;
  return MBM_OK;
}

} // namespace mbmi

extern "C" int mbm_time_update(mbm_ctx *c, double dt, double t, double *cfl)
{
  if (!c || !cfl) return set_err(MBM_E_ARG, "NULL argument");
  (void)t;
  int rc = enqueue_time_update(c, dt);
  if (rc) return rc;
  return read_cfl(c, cfl);
}

extern "C" int mbm_commit(mbm_ctx *c)
{
  if (!c) return set_err(MBM_E_ARG, "NULL context");
  std::swap(c->role[MBM_L_DATA], c->role[MBM_L_NEW]);
  return MBM_OK;
}

extern "C" int mbm_steps(mbm_ctx *c, double *t, double tend, double cflmax, double cflreg,
                         int maxsteps, double *dt_io, mbm_step_record *rec, int reccap, int *nrec,
                         int *nsteps, int *nrej)
{
  if (!c || !t) return set_err(MBM_E_ARG, "NULL argument");
  if (c->batch) return set_err(MBM_E_STATE, "mbm_steps needs a plain (single-deck) context");
  if (!(cflreg < cflmax)) return set_err(MBM_E_ARG, "CFL target %g >= CFLmax %g never accepts", cflreg, cflmax);
  if (rec && reccap > c->reccap) {
    if (c->drec) HIPCHK(hipFree(c->drec));
    c->drec = nullptr;
    HIPCHK(hipMalloc(&c->drec, (size_t)reccap * sizeof(RecordDev)));
    c->reccap = reccap;
    release_graphs(c);   /* the controller's record pointer is baked into the graph */
  }
  if (!c->drec) {
    HIPCHK(hipMalloc(&c->drec, sizeof(RecordDev)));
    c->reccap = 1;
  }
  DeckState S{};
  S.t = *t;
  S.dt = (dt_io && *dt_io > 0.0) ? *dt_io : 0.0001;
  S.tend = tend;
  S.cflmax = cflmax;
  S.cflreg = cflreg;
  S.kn = c->par.kn;
  S.force = c->par.force;
  S.maxsteps = maxsteps;
  S.done = !(*t < tend);
  S.reccap = rec ? reccap : 0;
  HIPCHK(hipMemcpyAsync(c->ddecks, &S, sizeof S, hipMemcpyHostToDevice, c->stream));
  /* S: the deck state the loop read back last (after the last attempt) */
  int rc = device_steps(c, 64, &S, &S);
  if (rc) return rc;
  if (rec && S.nrec > 0) {
    std::vector<RecordDev> r(std::min(S.nrec, reccap));
    HIPCHK(hipMemcpy(r.data(), c->drec, r.size() * sizeof(RecordDev), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < r.size(); i++) {
      rec[i].cfl = r[i].cfl;
      rec[i].dt = r[i].dt;
      rec[i].t = r[i].t;
      rec[i].accepted = r[i].accepted;
      rec[i].step = r[i].step;
    }
  }
  *t = S.t;
  if (nrec) *nrec = S.nrec;
  if (nsteps) *nsteps = S.nstep;
  if (nrej) *nrej = S.nrej;
  if (dt_io) *dt_io = S.dt;
  return MBM_OK;
}

namespace {
/* bytes of the Output.c conversion of one block: (nx+2)(ny+2) points of 10 values */
size_t output_bytes(const mbm_ctx *c, int block, int float32)
{
  const size_t npt = (size_t)(c->desc[block].nx + 2) * (c->desc[block].ny + 2);
  return npt * 10 * (float32 ? sizeof(float) : sizeof(double));
}

/* converts DATA of one block into the device buffer d and copies it to host */
hipError_t output_through(mbm_ctx *c, int block, int float32, void *d, void *host)
{
  const int nx = c->desc[block].nx, ny = c->desc[block].ny;
  const size_t npt = (size_t)(nx + 2) * (ny + 2);
  const int nblk = (int)((npt + 255) / 256);
  if (float32)
    hipLaunchKernelGGL(output_kernel<float>, dim3(nblk), dim3(256), 0, c->stream,
                       c->layer[c->role[MBM_L_DATA]], c->dblk, block, nx, ny, (float *)d);
  else
    hipLaunchKernelGGL(output_kernel<double>, dim3(nblk), dim3(256), 0, c->stream,
                       c->layer[c->role[MBM_L_DATA]], c->dblk, block, nx, ny, (double *)d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(host, d, output_bytes(c, block, float32), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  return e;
}
} // namespace

extern "C" int mbm_output_values(mbm_ctx *c, int block, int float32, void *host)
{
  int rc = check_lb(c, 0, block);
  if (rc) return rc;
  if (!host) return set_err(MBM_E_ARG, "NULL host buffer");
  void *d = nullptr;
  HIPCHK(hipMalloc(&d, output_bytes(c, block, float32)));
  const hipError_t e = output_through(c, block, float32, d, host);
  (void)hipFree(d);
  if (e != hipSuccess) return set_err(MBM_E_HIP, "output conversion failed: %s", hipGetErrorString(e));
  return MBM_OK;
}

extern "C" int mbm_output_values_noalloc(mbm_ctx *c, int block, int float32, void *host)
{
  int rc = check_lb(c, 0, block);
  if (rc) return rc;
  if (!host) return set_err(MBM_E_ARG, "NULL host buffer");
  const size_t bytes = output_bytes(c, block, float32);
  if (bytes > (size_t)MBM_OUTPUT_SCRATCH_BYTES)
    return set_err(MBM_E_ARG,
                   "block %d: its output (%zu bytes) exceeds the %d-byte static staging buffer of "
                   "mbm_output_values_noalloc; use mbm_output_values", block, bytes,
                   MBM_OUTPUT_SCRATCH_BYTES);
  void *d = c->device >= 0 && c->device < kMaxScratchDevices ? g_scratch[c->device] : nullptr;
  if (!d) return set_err(MBM_E_STATE, "static staging buffer not resolved on device %d", c->device);
  const hipError_t e = output_through(c, block, float32, d, host);
  if (e != hipSuccess) return set_err(MBM_E_HIP, "output conversion failed: %s", hipGetErrorString(e));
  return MBM_OK;
}

extern "C" long long mbm_interior_cells(const mbm_ctx *c)
{
  if (!c) return 0;
  long long s = 0;
  for (int k = 0; k < c->nb; k++) s += (long long)c->desc[k].nx * c->desc[k].ny;
  return s;
}

extern "C" int mbm_block_dims(const mbm_ctx *c, int block, int *nx, int *ny)
{
  if (!c || block < 0 || block >= c->nb) return set_err(MBM_E_ARG, "invalid block");
  if (nx) *nx = c->desc[block].nx;
  if (ny) *ny = c->desc[block].ny;
  return MBM_OK;
}

extern "C" int mbm_device(const mbm_ctx *c) { return c ? c->device : -1; }

extern "C" int mbm_enable_timing(mbm_ctx *c, int on)
{
  if (!c) return set_err(MBM_E_ARG, "NULL context");
  c->timing = on != 0;
  c->stage_ms = 0.0;
  c->nstage = 0;
  c->fill_ms = 0.0;
  c->nfill = 0;
  return MBM_OK;
}

extern "C" int mbm_kernel_time(mbm_ctx *c, double *stage_ms, int *nstage, double *fill_ms,
                               int *nfill)
{
  if (!c) return set_err(MBM_E_ARG, "NULL context");
  if (stage_ms) *stage_ms = c->stage_ms;
  if (nstage) *nstage = c->nstage;
  if (fill_ms) *fill_ms = c->fill_ms;
  if (nfill) *nfill = c->nfill;
  return MBM_OK;
}

/* CPU application of the ghost map, for checking the map against the
 * reference semantics without a GPU. layers[k]: AoS block k, (ny+6)*(nx+6)
 * cells x fastest. */
extern "C" int mbm_host_set_boundary(const mbm_block_desc *blocks, int nblocks,
                                     const mbm_params *p, double *const *layers, double t)
{
  (void)t;
  int rc = validate(blocks, nblocks, p);
  if (rc) return rc;
  std::vector<mbm_block_desc> D(blocks, blocks + nblocks);
  std::vector<BlockDev> hb;
  for (int k = 0; k < nblocks; k++) {
    BlockDev b{};
    b.lx = blocks[k].nx + 2 * MBC;
    b.ly = blocks[k].ny + 2 * MBC;
    geometry(blocks[k], b.bx, b.by, b.dx, b.dy);
    hb.push_back(b);
  }
  std::vector<GhostDev> gm = build_ghost_map(D, hb);
  std::vector<double> vals(gm.size() * NV);
  for (size_t i = 0; i < gm.size(); i++) {
    const GhostDev &g = gm[i];
    double v[NV];
    if (g.sblk >= 0) {
      const size_t sc = (size_t)g.sb * hb[g.sblk].lx + g.sa;
      for (int m = 0; m < NV; m++) v[m] = layers[g.sblk][NV * sc + m];
    } else {
      mbm_scn_eval(p->scenario, p->scenario_params, g.x, g.y, v);
    }
    for (int m = 0; m < NV; m++) vals[i * NV + m] = ((g.sign >> m) & 1) ? -v[m] : v[m];
  }
  for (size_t i = 0; i < gm.size(); i++) {
    const size_t dc = (size_t)gm[i].db * hb[gm[i].dblk].lx + gm[i].da;
    for (int m = 0; m < NV; m++) layers[gm[i].dblk][NV * dc + m] = vals[i * NV + m];
  }
  return MBM_OK;
}
