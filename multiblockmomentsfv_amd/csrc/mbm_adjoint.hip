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
 * so the sums are deterministic. S^T runs as three kernels:
 *   adj_prim : recompute the stage primitives P of every cell;
 *   adj_face : per x-face and y-face, the VJP of the reconstruction + HLL
 *              flux with respect to its two face states;
 *   adj_cell : per cell, gather the face-state adjoints, apply the limiter
 *              VJPs of the slopes that use the cell, then the VJP of the
 *              projection and of the initial value (relaxation, primitive
 *              recovery), and dJ/dKn partial sums per workgroup.
 * Checkpointing: DATA is stored every 'stride' steps during the forward
 * run; the reverse sweep recomputes one segment at a time (sqrt(N) memory).
 */
#include "mbm_internal.h"
#include "mbm_adj.h"
#include "mbm_adj_stage.h"

#include <algorithm>
#include <map>

using namespace mbmi;
using namespace mbm;

struct mbm_adj_state {
  int stride = 0;
  std::vector<double> dts;        /* accepted dt of every step */
  std::vector<double *> ckpt;     /* ckpt[i]: DATA after i*stride steps */
  std::vector<double *> seg;      /* segment states during the reverse sweep */
  double *g[3] = { nullptr, nullptr, nullptr };   /* gN, gD, gT */
  double *P = nullptr, *FxL = nullptr, *FxR = nullptr, *FyL = nullptr, *FyR = nullptr;
  double *tmp = nullptr;          /* primal TMP recompute */
  unsigned long long *slots = nullptr;
  double *knpart = nullptr;
  int nknpart = 0;
  int *inv_off = nullptr, *inv_list = nullptr, *inv_src_blk = nullptr, *inv_src_cell = nullptr;
  int *ghost_blk = nullptr, *ghost_cell = nullptr;
  int nsrc = 0, nghost = 0;
  int maxcells = 0;
};

namespace {
constexpr int ADJ_THREADS = 256;
}

template <int STAGE>
__global__ void __launch_bounds__(ADJ_THREADS) adj_prim(const double *in, double *P,
                                                        const BlockDev *blocks, double dt,
                                                        double dtkn, double force)
{
  const BlockDev B = blocks[blockIdx.y];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B.lx * B.ly) return;
  adj_prim_cell<STAGE>(in, P, B, i % B.lx, i / B.lx, dt, dtkn, force);
}

template <int LIM>
__global__ void __launch_bounds__(ADJ_THREADS) adj_face(const double *P, const double *gout,
                                                        double *FxL, double *FxR, double *FyL,
                                                        double *FyR, const BlockDev *blocks,
                                                        double dt, double scale)
{
  const BlockDev B = blocks[blockIdx.y];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B.lx * B.ly) return;
  adj_face_cell<LIM>(P, gout, FxL, FxR, FyL, FyR, B, i % B.lx, i / B.lx, dt, scale);
}

template <int STAGE, int LIM>
__global__ void __launch_bounds__(ADJ_THREADS) adj_cell(
    const double *in, const double *in0, const double *P, const double *FxL, const double *FxR,
    const double *FyL, const double *FyR, const double *gout, double *gin, double *gin0,
    const BlockDev *blocks, double dt, double kn, double force, double scale, double *knpart)
{
  const BlockDev B = blocks[blockIdx.y];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double gkn = 0.0;
  if (i < B.lx * B.ly)
    gkn = adj_cell_cell<STAGE, LIM>(in, in0, P, FxL, FxR, FyL, FyR, gout, gin, gin0, B, i % B.lx,
                                    i / B.lx, dt, kn, force, scale);
  /* dJ/dKn: one partial sum per workgroup, fixed order */
  __shared__ double red[ADJ_THREADS];
  red[threadIdx.x] = gkn;
  __syncthreads();
  for (int s = ADJ_THREADS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) knpart[blockIdx.y * gridDim.x + blockIdx.x] += red[0];
}

/* fill^T: g(src) += sum over its ghosts of sign * g(ghost) */
__global__ void __launch_bounds__(ADJ_THREADS) adj_fill_gather(
    double *g, const BlockDev *blocks, const GhostDev *map, const int *off, const int *list,
    const int *src_blk, const int *src_cell, int nsrc)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsrc) return;
  const BlockDev S = blocks[src_blk[s]];
  double acc[NV] = { 0 };
  for (int j = off[s]; j < off[s + 1]; j++) {
    const GhostDev gh = map[list[j]];
    const BlockDev D = blocks[gh.dblk];
    for (int m = 0; m < NV; m++) {
      const double v = g[D.off + gh.dcell + m * D.plane];
      acc[m] += ((gh.sign >> m) & 1) ? -v : v;
    }
  }
  for (int m = 0; m < NV; m++) g[S.off + src_cell[s] + m * S.plane] += acc[m];
}

__global__ void __launch_bounds__(ADJ_THREADS) adj_fill_clear(double *g, const BlockDev *blocks,
                                                              const GhostDev *map, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const GhostDev gh = map[i];
  const BlockDev D = blocks[gh.dblk];
  for (int m = 0; m < NV; m++) g[D.off + gh.dcell + m * D.plane] = 0.0;
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
  if ((rc = alloc_layer(c, &A->P))) return rc;
  if ((rc = alloc_layer(c, &A->FxL))) return rc;
  if ((rc = alloc_layer(c, &A->FxR))) return rc;
  if ((rc = alloc_layer(c, &A->FyL))) return rc;
  if ((rc = alloc_layer(c, &A->FyR))) return rc;
  if ((rc = alloc_layer(c, &A->tmp))) return rc;
  HIPCHK(hipMalloc(&A->slots, (size_t)6 * c->nb * sizeof(unsigned long long)));
  HIPCHK(hipMemsetAsync(A->slots, 0, (size_t)6 * c->nb * sizeof(unsigned long long), c->stream));
  for (int k = 0; k < c->nb; k++)
    A->maxcells = std::max(A->maxcells, c->hblk[k].lx * c->hblk[k].ly);
  A->nknpart = ((A->maxcells + ADJ_THREADS - 1) / ADJ_THREADS) * c->nb;
  HIPCHK(hipMalloc(&A->knpart, A->nknpart * sizeof(double)));
  /* inverse ghost map: CSR over source cells */
  std::vector<GhostDev> gm = build_ghost_map(c->desc, c->hblk);
  std::map<std::pair<int, int>, std::vector<int>> inv;
  for (int i = 0; i < (int)gm.size(); i++)
    if (gm[i].sblk >= 0) inv[{ gm[i].sblk, gm[i].scell }].push_back(i);
  std::vector<int> off(1, 0), list, sb, sc;
  for (auto &kv : inv) {
    sb.push_back(kv.first.first);
    sc.push_back(kv.first.second);
    for (int g : kv.second) list.push_back(g);
    off.push_back((int)list.size());
  }
  A->nsrc = (int)sb.size();
  A->nghost = (int)gm.size();
  HIPCHK(hipMalloc(&A->inv_off, off.size() * sizeof(int)));
  HIPCHK(hipMalloc(&A->inv_list, std::max<size_t>(1, list.size()) * sizeof(int)));
  HIPCHK(hipMalloc(&A->inv_src_blk, std::max<size_t>(1, sb.size()) * sizeof(int)));
  HIPCHK(hipMalloc(&A->inv_src_cell, std::max<size_t>(1, sc.size()) * sizeof(int)));
  HIPCHK(hipMemcpy(A->inv_off, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice));
  if (!list.empty())
    HIPCHK(hipMemcpy(A->inv_list, list.data(), list.size() * sizeof(int), hipMemcpyHostToDevice));
  if (!sb.empty()) {
    HIPCHK(hipMemcpy(A->inv_src_blk, sb.data(), sb.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(A->inv_src_cell, sc.data(), sc.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  return MBM_OK;
}

int fill_adjoint(mbm_ctx *c, double *g)
{
  mbm_adj_state *A = c->adj;
  if (A->nsrc > 0)
    hipLaunchKernelGGL(adj_fill_gather, dim3((A->nsrc + ADJ_THREADS - 1) / ADJ_THREADS),
                       dim3(ADJ_THREADS), 0, c->stream, g, c->dblk, c->dghost, A->inv_off,
                       A->inv_list, A->inv_src_blk, A->inv_src_cell, A->nsrc);
  if (A->nghost > 0)
    hipLaunchKernelGGL(adj_fill_clear, dim3((A->nghost + ADJ_THREADS - 1) / ADJ_THREADS),
                       dim3(ADJ_THREADS), 0, c->stream, g, c->dblk, c->dghost, A->nghost);
  HIPCHK(hipGetLastError());
  return MBM_OK;
}

template <int STAGE, int LIM>
int stage_adjoint_lim(mbm_ctx *c, const double *in, const double *in0, const double *gout,
                      double *gin, double *gin0, double dt)
{
  mbm_adj_state *A = c->adj;
  const dim3 grid((A->maxcells + ADJ_THREADS - 1) / ADJ_THREADS, c->nb), block(ADJ_THREADS);
  const double scale = STAGE == 2 ? 0.5 : 1.0;
  hipLaunchKernelGGL(adj_prim<STAGE>, grid, block, 0, c->stream, in, A->P, c->dblk, dt,
                     dt / c->par.kn, c->par.force);
  hipLaunchKernelGGL(adj_face<LIM>, grid, block, 0, c->stream, A->P, gout, A->FxL, A->FxR,
                     A->FyL, A->FyR, c->dblk, dt, scale);
  hipLaunchKernelGGL((adj_cell<STAGE, LIM>), grid, block, 0, c->stream, in, in0, A->P, A->FxL,
                     A->FxR, A->FyL, A->FyR, gout, gin, gin0, c->dblk, dt, c->par.kn,
                     c->par.force, scale, A->knpart);
  HIPCHK(hipGetLastError());
  return MBM_OK;
}

template <int STAGE>
int stage_adjoint(mbm_ctx *c, const double *in, const double *in0, const double *gout,
                  double *gin, double *gin0, double dt)
{
  switch (c->par.limiter) {
    case 0: return stage_adjoint_lim<STAGE, 0>(c, in, in0, gout, gin, gin0, dt);
    case 1: return stage_adjoint_lim<STAGE, 1>(c, in, in0, gout, gin, gin0, dt);
    case 2: return stage_adjoint_lim<STAGE, 2>(c, in, in0, gout, gin, gin0, dt);
    case 3: return stage_adjoint_lim<STAGE, 3>(c, in, in0, gout, gin, gin0, dt);
    default: return stage_adjoint_lim<STAGE, 4>(c, in, in0, gout, gin, gin0, dt);
  }
}

/* forward step between explicit buffers (frozen dt, no controller) */
int step_buffers(mbm_ctx *c, const double *D, double *N, double dt)
{
  mbm_adj_state *A = c->adj;
  int rc;
  if ((rc = enqueue_stage(c, 1, D, D, A->tmp, dt, A->slots))) return rc;
  if ((rc = fill_buffer(c, A->tmp))) return rc;
  if ((rc = enqueue_stage(c, 2, A->tmp, D, N, dt, A->slots + 3 * c->nb))) return rc;
  return fill_buffer(c, N);
}

/* reverse of one step: gN (in, destroyed) -> gD (out) */
int step_adjoint(mbm_ctx *c, const double *D, double dt, double *gN, double *gD)
{
  mbm_adj_state *A = c->adj;
  int rc;
  if ((rc = enqueue_stage(c, 1, D, D, A->tmp, dt, A->slots))) return rc;
  if ((rc = fill_buffer(c, A->tmp))) return rc;
  if ((rc = fill_adjoint(c, gN))) return rc;
  HIPCHK(hipMemsetAsync(gD, 0, c->layer_elems * sizeof(double), c->stream));
  if ((rc = stage_adjoint<2>(c, A->tmp, D, gN, A->g[2], gD, dt))) return rc;
  if ((rc = fill_adjoint(c, A->g[2]))) return rc;
  return stage_adjoint<1>(c, D, D, A->g[2], gD, nullptr, dt);
}

int push_checkpoint(mbm_ctx *c, const double *src)
{
  double *p = nullptr;
  if (hipMalloc(&p, c->layer_elems * sizeof(double)) != hipSuccess)
    return set_err(MBM_E_NOMEM, "adjoint: checkpoint allocation failed");
  c->adj->ckpt.push_back(p);
  HIPCHK(hipMemcpyAsync(p, src, c->layer_elems * sizeof(double), hipMemcpyDeviceToDevice,
                        c->stream));
  return MBM_OK;
}

void free_checkpoints(mbm_adj_state *A)
{
  for (double *p : A->ckpt) (void)hipFree(p);
  A->ckpt.clear();
  for (double *p : A->seg) (void)hipFree(p);
  A->seg.clear();
}

} // namespace

extern "C" long long mbm_layer_elems(const mbm_ctx *c) { return c ? (long long)c->layer_elems : 0; }

extern "C" int mbm_layer_get(mbm_ctx *c, int layer, void *dst)
{
  if (!c || !dst || layer < 0 || layer > 2) return set_err(MBM_E_ARG, "invalid argument");
  HIPCHK(hipMemcpyAsync(dst, c->layer[c->role[layer]], c->layer_elems * sizeof(double),
                        hipMemcpyDefault, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MBM_OK;
}

extern "C" int mbm_layer_set(mbm_ctx *c, int layer, const void *src)
{
  if (!c || !src || layer < 0 || layer > 2) return set_err(MBM_E_ARG, "invalid argument");
  HIPCHK(hipMemcpyAsync(c->layer[c->role[layer]], src, c->layer_elems * sizeof(double),
                        hipMemcpyDefault, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return MBM_OK;
}

extern "C" int mbm_adj_forward(mbm_ctx *c, double t0, double tend, double cflmax, double cflreg,
                               int maxsteps, const double *dts, int ndt, int stride,
                               int *nsteps, double *t_end)
{
  if (!c || stride < 1) return set_err(MBM_E_ARG, "invalid argument");
  int rc = adj_setup(c);
  if (rc) return rc;
  mbm_adj_state *A = c->adj;
  free_checkpoints(A);
  A->dts.clear();
  A->stride = stride;
  if ((rc = push_checkpoint(c, c->layer[c->role[MBM_L_DATA]]))) return rc;
  double t = t0;
  int k = 0;
  if (dts) {
    for (k = 0; k < ndt; k++) {
      double cfl;
      if ((rc = mbm_time_update(c, dts[k], t, &cfl))) return rc;
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

extern "C" int mbm_adj_schedule(mbm_ctx *c, double *dts, int cap, int *n)
{
  if (!c || !c->adj) return set_err(MBM_E_STATE, "no recorded forward run");
  const int m = (int)c->adj->dts.size();
  for (int i = 0; i < m && i < cap; i++) dts[i] = c->adj->dts[i];
  if (n) *n = m;
  return MBM_OK;
}

extern "C" int mbm_adj_backward(mbm_ctx *c, const void *g_final, void *g_initial, double *g_kn)
{
  if (!c || !c->adj || c->adj->ckpt.empty()) return set_err(MBM_E_STATE, "no recorded forward run");
  mbm_adj_state *A = c->adj;
  const size_t bytes = c->layer_elems * sizeof(double);
  const int N = (int)A->dts.size(), S = A->stride;
  int rc;
  HIPCHK(hipMemcpyAsync(A->g[0], g_final, bytes, hipMemcpyDefault, c->stream));
  HIPCHK(hipMemsetAsync(A->knpart, 0, A->nknpart * sizeof(double), c->stream));
  while ((int)A->seg.size() < S) {
    double *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return set_err(MBM_E_NOMEM, "adjoint: segment allocation failed");
    A->seg.push_back(p);
  }
  double *gN = A->g[0], *gD = A->g[1];
  const int nseg = (N + S - 1) / S;
  for (int s = nseg - 1; s >= 0; s--) {
    const int k0 = s * S, k1 = std::min(N, k0 + S);
    HIPCHK(hipMemcpyAsync(A->seg[0], A->ckpt[s], bytes, hipMemcpyDeviceToDevice, c->stream));
    for (int k = k0; k < k1 - 1; k++)
      if ((rc = step_buffers(c, A->seg[k - k0], A->seg[k - k0 + 1], A->dts[k]))) return rc;
    for (int k = k1 - 1; k >= k0; k--) {
      if ((rc = step_adjoint(c, A->seg[k - k0], A->dts[k], gN, gD))) return rc;
      std::swap(gN, gD);
    }
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

extern "C" int mbm_adj_release(mbm_ctx *c)
{
  if (!c || !c->adj) return MBM_OK;
  mbm_adj_state *A = c->adj;
  (void)hipStreamSynchronize(c->stream);
  free_checkpoints(A);
  for (double *p : { A->g[0], A->g[1], A->g[2], A->P, A->FxL, A->FxR, A->FyL, A->FyR, A->tmp })
    if (p) (void)hipFree(p);
  if (A->slots) (void)hipFree(A->slots);
  if (A->knpart) (void)hipFree(A->knpart);
  for (int *p : { A->inv_off, A->inv_list, A->inv_src_blk, A->inv_src_cell })
    if (p) (void)hipFree(p);
  delete A;
  c->adj = nullptr;
  return MBM_OK;
}
