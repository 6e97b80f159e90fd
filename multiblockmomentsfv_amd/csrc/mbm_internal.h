/*
 * mbm_internal.h -- structures shared by the engine's translation units
 * (stage/fill kernels in mbm_engine.hip, adjoint in mbm_adjoint.hip).
 */
#ifndef MBM_INTERNAL_H
#define MBM_INTERNAL_H

#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdarg>
#include <vector>
#include <algorithm>

#include "mbmfv.h"
#include "mbmfv_host.h"

struct mbm_adj_state;

namespace mbmi {

int set_err(int code, const char *fmt, ...);

#define HIPCHK(expr)                                                           \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return set_err(MBM_E_HIP, "%s failed: %s (%s:%d)", #expr,              \
                     hipGetErrorString(e_), __FILE__, __LINE__);               \
  } while (0)

constexpr int NV = 9;
constexpr int MBC = 3;
constexpr int STRIP = 60;          /* owned columns per wavefront */
/* CFL slots per block and stage: max x-face speed, max y-face speed, max
 * production CFL (bit patterns of non-negative doubles, NaN ignored), and 1 +
 * the scan position of the last NaN production CFL (0: none) */
constexpr int SLOTS_PER_BLOCK = 4;
/* waves (tiles) per workgroup of the stage kernels: the default, and the
 * largest a context may choose (8 waves hold 108 KiB of LDS, one workgroup
 * per CU) */
constexpr int WAVES_PER_WG = 4;
constexpr int MAX_WAVES_PER_WG = 8;

/* A block in a layer buffer. Rows are stored at row positions: a position
 * holds the 9 components of one row back to back (component m at m*pitch
 * elements, pitch >= lx), and the positions are ordered so that the rows
 * which tiles of height `band` march through at the same time are adjacent:
 * row b >= 1 lies at position ((b-1) % band) * nbands + (b-1) / band, row 0
 * at band * nbands. */
struct BlockDev {
  long long off;      /* element offset of block in a layer buffer */
  long long plane;    /* elements between components of a cell (= pitch) */
  int lx, ly, pitch, deck;   /* deck: index of the deck the block belongs to */
  double dx, dy, bx, by;
  int band, nbands;   /* row order, see above */
  unsigned bmul;      /* ceil(2^32 / band): (y * bmul) >> 32 == y / band while y * band < 2^32 */
  int npos;           /* row positions: band * nbands + 1 */
};

__host__ __device__ inline int row_pos(const BlockDev &B, int b)
{
  if (b <= 0) return B.band * B.nbands;
  const unsigned y = (unsigned)(b - 1);
  const unsigned q = (unsigned)(((unsigned long long)y * B.bmul) >> 32);
  return (int)(y - q * (unsigned)B.band) * B.nbands + (int)q;
}

/* element offset of row b within the block, and of cell (a, b), component 0,
 * within the layer */
__host__ __device__ inline long long row_elem(const BlockDev &B, int b)
{
  return (long long)row_pos(B, b) * (NV * (long long)B.pitch);
}
__host__ __device__ inline long long cell_elem(const BlockDev &B, int b, int a)
{
  return B.off + row_elem(B, b) + a;
}
inline long long block_elems(const BlockDev &B) { return (long long)B.npos * NV * B.pitch; }

/* sets pitch, plane and the row order of a block with lx, ly set (band >= 2,
 * so that bmul fits 32 bits) */
inline void set_block_layout(BlockDev &B, int band)
{
  B.pitch = (B.lx + 7) & ~7;
  B.plane = B.pitch;
  B.band = band;
  B.nbands = (B.ly - 1 + band - 1) / band;
  B.bmul = (unsigned)((0x100000000ull + (unsigned long long)band - 1) / (unsigned long long)band);
  B.npos = band * B.nbands + 1;
}

/* Tile height of a block with ly rows (ghosts included) for a requested
 * height: its iteration rows 1..ly-3 split into k = ceil((ly-3)/rows) tiles
 * of equal height (the last one at most k-1 rows shorter), so that no block
 * ends in a short tile that costs a full warm-up; at least 2, the smallest
 * band the row order allows. The block's row order uses the same height. */
inline int tile_band(int ly, int rows)
{
  const int n = ly - 3;
  const int k = (n + std::max(rows, 1) - 1) / std::max(rows, 1);
  return std::max((n + k - 1) / k, 2);
}

/* Device-resident state of the Steps controller of one deck (Main.c:189-244). */
struct DeckState {
  double t, dt, tend, cflmax, cflreg, kn, force;
  double cfl;         /* combined CFL of the last attempt */
  int nstep, nrej, maxsteps, done;
  int parity;         /* 0: DATA in buf0, NEW in buf1; 1: swapped */
  int nrec, reccap, pad;
};

/* one attempt of the controller (mirrors mbm_step_record) */
struct RecordDev {
  double cfl, dt, t;
  int accepted, step;
};

struct TileDev {
  int blk;
  int a0;             /* column of lane 0 */
  int r0, r1;         /* first/last iteration row */
};

/* one ghost cell: value = sign * source, source is a cell of the same layer
 * or the scenario state at (x,y) */
struct GhostDev {
  int dblk, da, db;   /* destination block, column, row */
  int sblk, sa, sb;   /* source block, column, row; sblk < 0: scenario */
  int sign;           /* bit m set: component m negated */
  int pad;
  double x, y;
};

/* device form of one ghost for the fill: element offsets of component 0 in
 * a layer (src < 0: the scenario state at the ghost's (x, y)) */
struct FillRec {
  long long dst, src;
  int dplane, splane;
  int deck, sign;     /* sign: bit m set, component m negated */
};

struct StageArgs {
  const double *in;   /* stage 1: DATA, stage 2: TMP */
  const double *in0;  /* stage 2: DATA */
  double *out;        /* stage 1: TMP, stage 2: NEW */
  const BlockDev *blocks;
  const TileDev *tiles;
  unsigned long long *slots;   /* SLOTS_PER_BLOCK per block */
  int ntiles;
  double dt, dtkn, force;
  /* device-controlled mode (decks != nullptr): dt, Kn and F come from the
   * block's deck state, DATA/NEW from its parity; in/in0/out are unused. */
  const DeckState *decks;
  const double *buf0, *buf1;
  double *tmp;
  /* ghost copies at the end of a tile (null: the fill kernel does them):
   * tile i copies the ghosts push[prange[i].x .. prange[i].y - 1], whose
   * sources it stores */
  const int2 *prange;
  const FillRec *push;
};


} // namespace mbmi

using mbmi::BlockDev;
using mbmi::TileDev;
using mbmi::GhostDev;
using mbmi::DeckState;
using mbmi::RecordDev;
using mbmi::FillRec;

struct mbm_ctx {
  int nb;
  std::vector<mbm_block_desc> desc;
  std::vector<BlockDev> hblk;
  mbm_params par;
  size_t layer_elems;
  double *layer[3];           /* physical buffers */
  int role[3];                /* role[L] = physical buffer holding layer L */
  BlockDev *dblk;
  TileDev *dtiles;
  int ntiles;
  int tile_rows;             /* requested tile height, 0 = automatic */
  int rows_used;             /* tile height of the current tile list */
  int cus;                   /* compute units of the device */
  int wpg;                   /* waves per workgroup of the stage kernels (1, 2, 4, 8) */
  FillRec *dfill;             /* ghost fill records, in block order */
  double2 *dfillxy;           /* ghost cell centres (scenario ghosts) */
  int nghost;
  unsigned long long *dslots; /* 2 stages x nb x SLOTS_PER_BLOCK */
  double *dcfl;               /* combined CFL */
  double *dtj;                /* scenario params */
  double *hcfl;               /* pinned */
  hipStream_t stream;
  bool own_stream;
  bool timing;
  hipEvent_t ev[4];
  double stage_ms, fill_ms;
  int nstage, nfill;
  int device;
  mbm_adj_state *adj;         /* adjoint buffers, created on demand */
  long long adj_budget;       /* bytes for recorded adjoint states, 0 = automatic */
  /* decks (one for a plain context): blocks [first[d], first[d+1]) */
  int ndecks;
  std::vector<int> deck_first;
  std::vector<mbm_params> deck_par;
  int *ddeck_first;
  DeckState *ddecks;
  RecordDev *drec;
  int reccap;
  bool batch;                     /* created by mbm_batch_create: roles never rotate */
  std::vector<int> deck_parity;   /* batches: DATA of deck d is buffer role[DATA] or role[NEW] */
  /* captured Heun step + controller, per (DATA, NEW) buffer assignment:
   * one attempt (graph) and GRAPH_ATTEMPTS attempts in a row (graphk) */
  hipGraphExec_t graph[3][3], graphk[3][3];
  bool graph_ok[3][3], graphk_ok[3][3];
  /* batches of two or more decks: decks [grp_deck[g], grp_deck[g+1]) form
   * group g (g = 0, 1), with tiles [grp_tile[g], grp_tile[g+1]) and ghosts
   * [grp_ghost[g], grp_ghost[g+1]). Group 1 runs on stream2, its attempt
   * shifted by half (stage 2 of one group overlaps stage 1 of the other). */
  int ngroups;                    /* 0: one group (plain contexts) */
  int grp_deck[3], grp_tile[3], grp_ghost[3];
  /* ghost copies by the stage kernel (push_on): per tile of the launch
   * list its range of dpush (ghosts with an interior source, grouped by the
   * tile that stores the source); the scenario ghosts dresid, ranges
   * grp_resid per group, stay with the fill kernel */
  bool push_on;
  FillRec *dpush, *dresid;
  double2 *dresidxy;
  int2 *dprange;
  int nresid, grp_resid[3];
  hipStream_t stream2;
  hipEvent_t grp_ev;
  hipGraphExec_t graph2, graph2k;
  bool graph2_ok, graph2k_ok;
};

/* attempts per launch of the long step graphs: one graph launch costs the
 * device a few microseconds more than a kernel boundary inside a graph */
constexpr int GRAPH_ATTEMPTS = 8;

namespace mbmi {
std::vector<GhostDev> build_ghost_map(const std::vector<mbm_block_desc> &D,
                                      const std::vector<BlockDev> &hb);
int do_fill(mbm_ctx *c, int layer);
int fill_buffer(mbm_ctx *c, double *layer);
int fill_after_stage(mbm_ctx *c, double *layer);
int enqueue_stage(mbm_ctx *c, int stage, const double *in, const double *in0, double *out,
                  double dt, unsigned long long *slots);
mbm_ctx *create_ctx(const mbm_block_desc *blocks, int nblocks, const int *deck_first, int ndecks,
                    const mbm_params *deck_par, void *stream, int *rc);
/* known: the deck states just written to the device (NULL: read them back);
 * last (may be NULL, may be known): receives the deck states after the run */
int device_steps(mbm_ctx *c, int maxchunk, const DeckState *known, DeckState *last);
int enqueue_time_update(mbm_ctx *c, double dt);   /* both stages and fills, CFL to device */
int upload_tiles(mbm_ctx *c);
int validate(const mbm_block_desc *blocks, int nb, const mbm_params *p);
int enqueue_device_stages(mbm_ctx *c);
int clear_slots(mbm_ctx *c);
void release_graphs(mbm_ctx *c);
} // namespace mbmi

#endif
