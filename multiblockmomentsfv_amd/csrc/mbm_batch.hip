/*
 * mbm_batch.hip -- many independent decks advanced together (no reference
 * counterpart: the reference runs one deck per process).
 *
 * All blocks of all decks share the three layer buffers, the tile list and
 * the ghost map; every block knows its deck. Each deck has its own Steps
 * controller on the device (dt, rejection, t, DATA/NEW parity), so one Heun
 * attempt of a set of decks is five launches (stage 1, fill, stage 2, fill,
 * controller), captured once in a hipGraph and replayed without host round
 * trips. The decks form two groups of about equal size on two streams, the
 * second half an attempt behind the first: stage 1 is bound by FP64 issue
 * and stage 2 by memory, and one group's stage 1 runs beside the other's
 * stage 2. Tile height is chosen by timing candidate heights on the batch.
 */
#include "mbm_internal.h"

#include <algorithm>

using namespace mbmi;

extern "C" int mbm_batch_create(const mbm_block_desc *blocks, const int *nblocks_per_deck,
                                int ndecks, const mbm_params *params, void *stream,
                                mbm_ctx **out)
{
  if (!out || !blocks || !nblocks_per_deck || !params || ndecks < 1)
    return set_err(MBM_E_ARG, "invalid argument");
  *out = nullptr;
  std::vector<int> first(ndecks + 1, 0);
  for (int d = 0; d < ndecks; d++) {
    if (nblocks_per_deck[d] < 1 || nblocks_per_deck[d] > MBM_MAX_BLOCKS)
      return set_err(MBM_E_ARG, "deck %d: %d blocks outside [1,%d]", d, nblocks_per_deck[d], MBM_MAX_BLOCKS);
    if (params[d].limiter != params[0].limiter || params[d].scenario != params[0].scenario)
      return set_err(MBM_E_ARG, "deck %d: the limiter and scenario must be the same in a batch", d);
    first[d + 1] = first[d] + nblocks_per_deck[d];
  }
  const int nb = first[ndecks];
  std::vector<mbm_block_desc> glob(blocks, blocks + nb);
  for (int d = 0; d < ndecks; d++) {
    int rc = validate(blocks + first[d], nblocks_per_deck[d], &params[d]);
    if (rc) return set_err(rc, "deck %d: %s", d, mbm_last_error());
    for (int k = first[d]; k < first[d + 1]; k++)
      for (int s = 0; s < 4; s++)
        if (glob[k].bc[s] > 0) glob[k].bc[s] += first[d];
  }
  int rc;
  mbm_ctx *c = create_ctx(glob.data(), nb, first.data(), ndecks, params, stream, &rc);
  if (!c) return rc;
  c->batch = true;
  c->deck_parity.assign(ndecks, 0);
  if (ndecks >= 2) {
    /* two groups of about equal cell count, on two streams */
    std::vector<double> cells(ndecks + 1, 0.0);
    for (int d = 0; d < ndecks; d++) {
      cells[d + 1] = cells[d];
      for (int k = first[d]; k < first[d + 1]; k++) cells[d + 1] += (double)blocks[k].nx * blocks[k].ny;
    }
    int split = 1;
    while (split < ndecks - 1 && cells[split] < 0.5 * cells[ndecks]) split++;
    const std::vector<GhostDev> gm = build_ghost_map(c->desc, c->hblk);
    int h1 = 0;
    while (h1 < (int)gm.size() && gm[h1].dblk < first[split]) h1++;
    c->ngroups = 2;
    c->grp_deck[0] = 0;
    c->grp_deck[1] = split;
    c->grp_deck[2] = ndecks;
    c->grp_ghost[0] = 0;
    c->grp_ghost[1] = h1;
    c->grp_ghost[2] = (int)gm.size();
    if (hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->grp_ev, hipEventDisableTiming) != hipSuccess ||
        (rc = upload_tiles(c))) {
      mbm_destroy(c);
      return rc ? rc : set_err(MBM_E_HIP, "batch: second stream");
    }
  }
  *out = c;
  return MBM_OK;
}

extern "C" int mbm_batch_run(mbm_ctx *c, const double *tend, const double *cflreg, double cflmax,
                             int maxsteps, double *t, int *nsteps, int *nrej)
{
  if (!c || !tend || !cflreg || !t) return set_err(MBM_E_ARG, "NULL argument");
  if (c->deck_parity.size() != (size_t)c->ndecks) c->deck_parity.assign(c->ndecks, 0);
  std::vector<DeckState> st(c->ndecks);
  for (int d = 0; d < c->ndecks; d++) {
    if (!(cflreg[d] < cflmax)) return set_err(MBM_E_ARG, "deck %d: CFL target >= CFLmax", d);
    DeckState &S = st[d];
    S = DeckState{};
    S.t = t[d];
    S.dt = 0.0001;
    S.tend = tend[d];
    S.cflmax = cflmax;
    S.cflreg = cflreg[d];
    S.kn = c->deck_par[d].kn;
    S.force = c->deck_par[d].force;
    S.maxsteps = maxsteps;
    S.done = !(t[d] < tend[d]);
    S.parity = c->deck_parity[d];
  }
  HIPCHK(hipMemcpyAsync(c->ddecks, st.data(), st.size() * sizeof(DeckState), hipMemcpyHostToDevice,
                        c->stream));
  /* roles stay fixed in a batch: each deck's parity says where its DATA is */
  const int ndecks = c->ndecks;
  int rc = device_steps(c, 64, st.data(), st.data());
  if (rc) return rc;
  for (int d = 0; d < ndecks; d++) {
    t[d] = st[d].t;
    if (nsteps) nsteps[d] = st[d].nstep;
    if (nrej) nrej[d] = st[d].nrej;
    c->deck_parity[d] = st[d].parity;
  }
  return MBM_OK;
}

extern "C" int mbm_batch_download(mbm_ctx *c, int deck, int block, double *host)
{
  if (!c || deck < 0 || deck >= c->ndecks) return set_err(MBM_E_ARG, "invalid deck");
  const int k = c->deck_first[deck] + block;
  if (block < 0 || k >= c->deck_first[deck + 1]) return set_err(MBM_E_ARG, "invalid block");
  const int par = c->deck_parity.empty() ? 0 : c->deck_parity[deck];
  const int phys = c->role[par ? MBM_L_NEW : MBM_L_DATA];
  const BlockDev &B = c->hblk[k];
  std::vector<double> soa((size_t)block_elems(B));
  HIPCHK(hipMemcpyAsync(soa.data(), c->layer[phys] + B.off, soa.size() * sizeof(double),
                        hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int b = 0; b < B.ly; b++)
    for (int a = 0; a < B.lx; a++)
      for (int m = 0; m < NV; m++)
        host[(size_t)NV * ((size_t)b * B.lx + a) + m] = soa[(size_t)(row_elem(B, b) + m * B.plane + a)];
  return MBM_OK;
}

extern "C" int mbm_batch_ndecks(const mbm_ctx *c) { return c ? c->ndecks : 0; }

/* Times 'reps' stage pairs (with their fills) of every deck, at the decks'
 * current dt, for each pair of a candidate wave group and tile height, and
 * keeps the fastest pair. Two passes over the candidates; each candidate
 * counts with the faster of its two timings, so that a change of the GPU's
 * clock during one pass does not decide the choice. The state is not
 * changed: the stages write TMP and the NEW side only. */
extern "C" int mbm_batch_autotune_shape(mbm_ctx *c, const int *rows, int nrows, const int *waves,
                                        int nwaves, int reps, int *best_rows, int *best_waves,
                                        double *best_ms)
{
  if (!c || !rows || nrows < 1 || !waves || nwaves < 1 || reps < 1)
    return set_err(MBM_E_ARG, "invalid argument");
  if (c->deck_parity.size() != (size_t)c->ndecks) c->deck_parity.assign(c->ndecks, 0);
  std::vector<DeckState> st(c->ndecks);
  for (int d = 0; d < c->ndecks; d++) {
    DeckState &S = st[d];
    S = DeckState{};
    S.dt = 1e-5;
    S.tend = 1.0;
    S.cflmax = 1.0;
    S.cflreg = 0.9;
    S.kn = c->deck_par[d].kn;
    S.force = c->deck_par[d].force;
    S.parity = c->deck_parity[d];
  }
  HIPCHK(hipMemcpyAsync(c->ddecks, st.data(), st.size() * sizeof(DeckState), hipMemcpyHostToDevice,
                        c->stream));
  int bestr = c->tile_rows, bestw = c->wpg, rc;
  double bestt = 1e300;
  std::vector<double> t((size_t)nwaves * nrows, 1e300);
  for (int pass = 0; pass < 2; pass++)
    for (int k = 0; k < nwaves; k++) {
      if ((rc = mbm_set_wave_group(c, waves[k]))) return rc;
      for (int i = 0; i < nrows; i++) {
        if ((rc = mbm_set_tile_rows(c, rows[i]))) return rc;
        if ((rc = enqueue_device_stages(c))) return rc;    /* warm-up */
        HIPCHK(hipEventRecord(c->ev[0], c->stream));
        for (int r = 0; r < reps; r++)
          if ((rc = enqueue_device_stages(c))) return rc;
        HIPCHK(hipEventRecord(c->ev[1], c->stream));
        HIPCHK(hipEventSynchronize(c->ev[1]));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
        double &tk = t[(size_t)k * nrows + i];
        tk = std::min(tk, (double)ms / reps);
      }
    }
  for (int k = 0; k < nwaves; k++)
    for (int i = 0; i < nrows; i++)
      if (t[(size_t)k * nrows + i] < bestt) {
        bestt = t[(size_t)k * nrows + i];
        bestr = rows[i];
        bestw = waves[k];
      }
  if ((rc = clear_slots(c))) return rc;
  if ((rc = mbm_set_wave_group(c, bestw))) return rc;
  if ((rc = mbm_set_tile_rows(c, bestr))) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  if (best_rows) *best_rows = bestr;
  if (best_waves) *best_waves = bestw;
  if (best_ms) *best_ms = bestt;
  return MBM_OK;
}

extern "C" int mbm_batch_autotune(mbm_ctx *c, const int *rows, int nrows, int reps, int *best,
                                  double *best_ms)
{
  if (!c) return set_err(MBM_E_ARG, "invalid argument");
  const int w = c->wpg;
  return mbm_batch_autotune_shape(c, rows, nrows, &w, 1, reps, best, nullptr, best_ms);
}
