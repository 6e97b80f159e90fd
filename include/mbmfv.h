/*
 * mbmfv.h -- C ABI of the MI355X (gfx950) engine for the explicit update of
 * MultiBlockMomentsFV.
 *
 * The reference has no plugin or FFI layer: its hot path is a set of plain C
 * calls inside one process. Each entry point below replaces one of them; the
 * comment names the reference interface it stands for. Plain C types only:
 * pointers, sizes and doubles; device pointers are void*. Every function
 * returns MBM_OK (0) or a negative MBM_E_* code; mbm_last_error() gives the
 * message of the last failure on the calling thread. A context is used from
 * one host thread at a time; work is ordered on the context's HIP stream.
 *
 * Device layout (per layer = DATA, TMP, NEW): one allocation holding every
 * block. Block k stores its ly = ny+6 rows at row positions; a position holds
 * the 9 components of one row back to back, each row padded to a length
 * plane[k] >= lx = nx+6. Cell (a,b) of component m is at
 *   base + off[k] + m*plane[k] + pos_k[b]*pitch[k] + a      (pitch = 9*plane),
 * a,b counted from the outermost ghost ring (0-based). The row order pos_k
 * (mbm_layer_rows) keeps adjacent the rows that the stage kernel's tiles
 * march through at the same time; it follows the tile height, and a change
 * of tile height moves the layers to the new order. FP64 throughout.
 */
#ifndef MBMFV_H
#define MBMFV_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MBM_OK 0
#define MBM_E_ARG -1       /* invalid argument or deck */
#define MBM_E_NOMEM -2     /* host or device allocation failed */
#define MBM_E_HIP -3       /* HIP runtime error */
#define MBM_E_IO -4        /* file I/O */
#define MBM_E_FORMAT -5    /* malformed block file / output file */
#define MBM_E_STATE -6     /* call out of order */
#define MBM_E_NODEV -7     /* no usable gfx950 device */

/* layers (defs.h:21-24) */
#define MBM_L_DATA 0
#define MBM_L_TMP 1
#define MBM_L_NEW 2

typedef struct mbm_ctx mbm_ctx;

/* One block of the block file (Mesh.c:148-179): interior cells, extent and
 * edge codes left,right,bottom,top (>0 sync with block id, 0 open, -1 wall,
 * -2 dirichlet). nx, ny >= 3. */
typedef struct {
  int nx, ny;
  double xmin, xmax, ymin, ymax;
  int bc[4];
} mbm_block_desc;

/* Physics/numerics parameters: Kn and F (Scenario.c:22-23), the limiter
 * (Limiter.c:20), the scenario used for the initial state and Dirichlet
 * ghosts (Scenario.c:72), TempJump parameters {theta_in, theta_out, radius,
 * width, rho}. */
typedef struct {
  double kn;
  double force;
  int limiter;
  int scenario;
  double scenario_params[5];
} mbm_params;

/* Per-step record written by mbm_steps (Main.c:199-230 prints these). */
typedef struct {
  double cfl;       /* combined CFL of the attempt */
  double dt;        /* dt used by the attempt */
  double t;         /* time after the attempt */
  int accepted;     /* 0 = rejected (CFL > CFLmax) */
  int step;         /* 1-based step number within this mbm_steps call */
} mbm_step_record;

const char *mbm_last_error(void);

/* Creates the device state for a deck: three FP64 layers, the ghost map and
 * the tile list. stream: hipStream_t to order work on, or NULL for the
 * context's own stream. Replaces main's per-block malloc (Main.c:113-123). */
int mbm_create(const mbm_block_desc *blocks, int nblocks, const mbm_params *p,
               void *stream, mbm_ctx **out);
int mbm_destroy(mbm_ctx *c);                                  /* Main.c:158-161 */
int mbm_set_stream(mbm_ctx *c, void *stream);
int mbm_set_params(mbm_ctx *c, const mbm_params *p);          /* CLI overrides, Main.c:97-106 */
/* Tile height (rows per wavefront) of the stage kernel; 0 (the default)
 * chooses it from the deck and the device's compute-unit count. */
int mbm_set_tile_rows(mbm_ctx *c, int rows);
int mbm_tile_rows(const mbm_ctx *c);                          /* height in use */
/* Waves (tiles) per workgroup of the stage kernels: 1, 2, 4 (the default) or
 * 8. No counterpart in the reference (a launch-shape choice); results do not
 * depend on it. Drops captured step graphs. */
int mbm_set_wave_group(mbm_ctx *c, int waves);
int mbm_wave_group(const mbm_ctx *c);

/* Initial state on interior cells then the ghost fill (Main.c:124-138). */
int mbm_init(mbm_ctx *c);

/* Host <-> device copies of one block of one layer, AoS (9 doubles per
 * cell, (ny+6)*(nx+6) cells, x fastest) as the reference's CVector arrays. */
int mbm_upload(mbm_ctx *c, int layer, int block, const double *host);
int mbm_download(mbm_ctx *c, int layer, int block, double *host);

/* Device pointer and geometry of one block of one layer (for callers that
 * keep the state on the device, e.g. a PyTorch wrapper): component stride
 * 'plane' and row-position stride 'pitch' in doubles. */
int mbm_layer_ptr(mbm_ctx *c, int layer, int block, void **dptr,
                  long long *plane, int *pitch);

/* Row order of one block: pos[b] = row position of row b, b < min(n, ny+6).
 * Valid until the next change of tile height. No reference counterpart: the
 * reference's layers are host arrays in natural row order (Mesh.c:161-164). */
int mbm_layer_rows(mbm_ctx *c, int block, int *pos, int n);

/* The row order of a block of ny interior rows for tiles of 'band' rows, as
 * mbm_layer_rows reports it (pos holds ny+6 entries); needs no device and
 * has no reference counterpart. Returns the number of row positions, or a
 * negative MBM_E_* code. */
int mbm_row_order(int ny, int band, int *pos);

/* SetBoundary (Boundary.c:173-193): fills the three ghost rings of every
 * block of 'layer' at time t, with the reference's corner semantics. */
int mbm_set_boundary(mbm_ctx *c, int layer, double t);

/* TimeUpdate (Main.c:253-284): stage 1 -> fill(TMP) -> stage 2 -> fill(NEW).
 * Writes TMP and NEW, leaves DATA untouched, returns the combined CFL
 * max over blocks and stages of dt*cx/dx + dt*cy/dy + cprod (Main.c:395). */
int mbm_time_update(mbm_ctx *c, double dt, double t, double *cfl);

/* The New->Data copy (Main.c:236-240) as a buffer rotation: NEW becomes DATA. */
int mbm_commit(mbm_ctx *c);

/* Steps (Main.c:189-244): adaptive dt from DT_INITIAL=1e-4, rejection while
 * CFL > cflmax, clamp to tend. Advances *t, returns accepted steps in
 * *nsteps and rejections in *nrej. If rec != NULL, up to reccap attempt
 * records are stored and *nrec receives the attempt count. maxsteps > 0
 * stops after that many accepted steps. dt_io: NULL or *dt_io <= 0 starts
 * from DT_INITIAL as the reference does on every call; otherwise the run
 * continues with *dt_io, and *dt_io receives the next dt on return. */
int mbm_steps(mbm_ctx *c, double *t, double tend, double cflmax, double cflreg,
              int maxsteps, double *dt_io, mbm_step_record *rec, int reccap,
              int *nrec, int *nsteps, int *nrej);

/* Output.c:217-243 conversion of DATA for one block, computed on the
 * device: interior plus one ghost ring, i (x) outer / j (y) inner, values
 * rho vx vy T p Pxx Pxy Pyy qx qy per point, written to host as float
 * (float32 != 0) or double. */
int mbm_output_values(mbm_ctx *c, int block, int float32, void *host);

/* The conversion of mbm_output_values without any device allocation. No
 * reference counterpart: it is for callers that cannot allocate device
 * memory (tight output loops, code that forbids hipMalloc/hipFree). The
 * device-side values are staged in a static per-GPU buffer of
 * MBM_OUTPUT_SCRATCH_BYTES, held until the process ends; a block whose output
 * ((nx+2)(ny+2) points x 10 values x 4 or 8 bytes) exceeds it is rejected
 * with MBM_E_ARG. Calls on one GPU share that buffer: calls that overlap in
 * time on one GPU (other threads, other streams) race. */
#define MBM_OUTPUT_SCRATCH_BYTES (64 * 1024 * 1024)
int mbm_output_values_noalloc(mbm_ctx *c, int block, int float32, void *host);

/* Number of interior cells (tools.c:87-95) and of cells incl. ghosts. */
long long mbm_interior_cells(const mbm_ctx *c);
int mbm_block_dims(const mbm_ctx *c, int block, int *nx, int *ny);
/* HIP device index the context was created on (no reference counterpart:
 * lets a caller check that its device buffers live on the same GPU). */
int mbm_device(const mbm_ctx *c);

/* Kernel timing of mbm_time_update calls made while timing is enabled: HIP
 * events on the context's stream bracket each stage launch and the TMP ghost
 * fill; mbm_kernel_time returns the summed ms and the launch counts since
 * mbm_enable_timing. (mbm_steps replays a captured graph and is not timed.) */
int mbm_enable_timing(mbm_ctx *c, int on);
int mbm_kernel_time(mbm_ctx *c, double *stage_ms, int *nstage, double *fill_ms,
                    int *nfill);

/* Whole-layer copies in the device layout above (all blocks, layer_elems
 * doubles); src/dst may be host or device pointers holding n doubles, and n
 * must equal mbm_layer_elems (it changes with the tile height). */
long long mbm_layer_elems(const mbm_ctx *c);
int mbm_layer_get(mbm_ctx *c, int layer, void *dst, long long n);
int mbm_layer_set(mbm_ctx *c, int layer, const void *src, long long n);

/* ---- discrete adjoint of the Heun step --------------------------------
 * No counterpart in the reference (it has no derivatives). Reverse-mode
 * derivative of a run of TimeUpdate steps with the dt schedule held fixed,
 * with respect to the state and Kn; limiter, |.| and HLL speed clipping are
 * differentiated almost everywhere. Layers are allocated on first use and
 * kept (for the next forward/backward pair) until mbm_adj_release or
 * mbm_destroy. A backward pass leaves the recorded run intact, so it may be
 * repeated.
 *
 * mbm_adj_forward: advances DATA from its current value. dts == NULL: the
 * Steps controller of Main.c:189-244 (from DT_INITIAL, rejection, clamp to
 * tend, at most maxsteps accepted steps) and its accepted dt are recorded;
 * otherwise the ndt given steps are replayed. stride >= 1: DATA is
 * checkpointed every 'stride' steps and the backward pass recomputes each
 * segment (N/stride + 2*(2*stride - 1) layers). stride == 0: a replay whose
 * N + 2 layers fit the budget keeps every step's DATA state and, as far as
 * the budget goes, the stage states of the last steps, so the backward pass
 * recomputes at most stage 1 of a step; otherwise stride ceil(sqrt(N))
 * (N = ndt, or maxsteps, or 1000 without either). The gradients do not
 * depend on the mode. */
int mbm_adj_forward(mbm_ctx *c, double t0, double tend, double cflmax, double cflreg,
                    int maxsteps, const double *dts, int ndt, int stride,
                    int *nsteps, double *t_end);
/* the dt schedule of the last mbm_adj_forward */
int mbm_adj_schedule(mbm_ctx *c, double *dts, int cap, int *n);
/* g_final: dJ/d(final DATA), layer layout (ghost entries are folded through
 * the last ghost fill). g_initial receives dJ/d(initial DATA) on interior
 * cells (ghost entries 0: the initial ghosts are a function of the interior
 * through SetBoundary); *g_kn receives dJ/dKn. Host or device pointers of n
 * doubles each, n == mbm_layer_elems. */
int mbm_adj_backward(mbm_ctx *c, const void *g_final, void *g_initial, long long n,
                     double *g_kn);
int mbm_adj_release(mbm_ctx *c);
/* device memory that a stride-0 recording may hold for its states, in
 * bytes; 0 (the default) = the free device memory plus the layers the
 * recording already holds, less max(8 GiB, 5 % of the device) */
int mbm_adj_set_budget(mbm_ctx *c, long long bytes);
/* device memory held by the recorded states (checkpoints, kept stage
 * states, the reverse sweep's segment layers, and layers kept for reuse by
 * the next recording) */
int mbm_adj_memory(mbm_ctx *c, long long *bytes);

/* ---- batches of independent decks ---------------------------------------
 * No counterpart in the reference (one deck per process there). Decks are
 * listed back to back in 'blocks'; each deck's sync codes are 1-based within
 * the deck. Per deck: Kn, F (params[d]) and its own Steps controller (dt,
 * rejection, tend); the limiter and scenario must be common. The returned
 * context works with mbm_init, mbm_set_tile_rows, mbm_layer_* and
 * mbm_destroy; mbm_time_update/mbm_steps are for single decks. */
int mbm_batch_create(const mbm_block_desc *blocks, const int *nblocks_per_deck, int ndecks,
                     const mbm_params *params, void *stream, mbm_ctx **out);
int mbm_batch_ndecks(const mbm_ctx *c);
/* Advances every deck d from t[d] (dt from DT_INITIAL) with its controller
 * until t[d] >= tend[d] or maxsteps accepted steps; one hipGraph-captured
 * step of all decks per launch sequence, no host round trip per step. */
int mbm_batch_run(mbm_ctx *c, const double *tend, const double *cflreg, double cflmax,
                  int maxsteps, double *t, int *nsteps, int *nrej);
/* DATA of one block of one deck, AoS as mbm_download */
int mbm_batch_download(mbm_ctx *c, int deck, int block, double *host);
/* Times the candidate tile heights on the batch and keeps the fastest. Each
 * candidate moves the layers to its row order (the state is kept); it
 * writes TMP and NEW, and drops a recorded adjoint run. */
int mbm_batch_autotune(mbm_ctx *c, const int *rows, int nrows, int reps, int *best,
                       double *best_ms);
/* As mbm_batch_autotune over every pair of a candidate tile height and a
 * candidate wave group (1, 2, 4 or 8 waves per workgroup); keeps the fastest
 * pair. */
int mbm_batch_autotune_shape(mbm_ctx *c, const int *rows, int nrows, const int *waves, int nwaves,
                             int reps, int *best_rows, int *best_waves, double *best_ms);

#ifdef __cplusplus
}
#endif
#endif
