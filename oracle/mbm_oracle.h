/*
 * mbm_oracle.h -- CPU oracle for the explicit update of MultiBlockMomentsFV.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's
 * FP64 explicit update (Heun stages, relaxation, MUSCL + limiter, HLL fluxes,
 * ghost fill, adaptive-CFL controller).  It exists to check the HIP engine and
 * to time a CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product never links it.
 *
 * Arithmetic is written in the reference's evaluation order and built with
 * gcc -O3 -ffp-contract=off, so the results are bit-identical to the
 * reference's own build (pinned by tests/test_oracle_*.py against the Float32
 * goldens and FP64 fixtures written by the compiled reference).
 *
 * Pinning status per function (see the tests named):
 *   cell/face physics, limiters, Update, TimeUpdate, SetBoundary, Steps,
 *   SetInitial(Riemann2D), output conversion  -> pinned bitwise against the
 *       reference goldens and FP64 fixtures (test_oracle_golden.py,
 *       test_oracle_fixtures.py).
 *   SetInitial(ForwardFacingStep)             -> pinned bitwise against FP64
 *       fixtures of the unmodified reference build (ffs decks).
 *   SetInitial(TempJump)                       -> build-defined scenario with
 *       no reference counterpart; pinned by closed-form checks of the initial
 *       state (test_oracle_math.py).
 *   The oracle has no derivatives: the engine's adjoint is checked against
 *   central differences of orc_time_update / orc_update_block
 *   (test_adjoint_local.py, test_adjoint_step_ref.py).
 */
#ifndef MBM_ORACLE_H
#define MBM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NV 9
#define ORC_MBC 3
#define ORC_MAXB 32

enum { ORC_LIM_NONE = 0, ORC_LIM_AVERAGE = 1, ORC_LIM_MINMOD = 2,
       ORC_LIM_SUPERBEE = 3, ORC_LIM_HARMONIC = 4 };

enum { ORC_SCN_RIEMANN2D = 0, ORC_SCN_FFS = 1, ORC_SCN_TEMPJUMP = 2 };

enum { ORC_L_DATA = 0, ORC_L_TMP = 1, ORC_L_NEW = 2 };

typedef struct orc_ctx orc_ctx;

/* Block table as read from a block file (Mesh.c:148-179): per block the
 * interior cell counts, the x/y ranges and the four edge codes
 * (left,right,bottom,top; >0 sync with block id, 0 open, -1 wall, -2 dirichlet). */
orc_ctx *orc_create(int nblocks, const int *nx, const int *ny,
                    const double *xmin, const double *xmax,
                    const double *ymin, const double *ymax,
                    const int *bc4, double kn, double force, int limiter,
                    int scenario);
void orc_destroy(orc_ctx *c);
void orc_set_threads(int n);

/* Scenario parameters for TempJump (build-defined): theta_in, theta_out,
 * radius, width, rho. Ignored by the other scenarios. */
void orc_set_tempjump(orc_ctx *c, double th_in, double th_out, double radius,
                      double width, double rho);

/* interior initial state (Main.c:126-136) followed by SetBoundary(DATA,0). */
void orc_init(orc_ctx *c);
void orc_set_boundary(orc_ctx *c, int layer, double t);
/* one Heun step (Main.c:253-284); returns the combined CFL. NEW holds the
 * candidate state, DATA is untouched. */
double orc_time_update(orc_ctx *c, double dt, double t);
/* NEW -> DATA copy of all cells incl. ghosts (Main.c:236-240). */
void orc_commit(orc_ctx *c);
/* one stage on one block (Main.c:297-396); heun=0 writes TMP from DATA,
 * heun=1 writes NEW from TMP and DATA. Returns the block CFL. */
double orc_update_block(orc_ctx *c, int blk, double dt, double t, int heun);

/* Steps (Main.c:189-244). Advances *t to tend; returns the accepted step
 * count. If log!=NULL, for each TimeUpdate call appends 4 doubles
 * {cfl, dt_used, t_after, accepted} up to logcap entries; *nlog receives the
 * number of attempts. maxsteps>0 stops after that many accepted steps. */
int orc_steps(orc_ctx *c, double *t, double tend, double cflmax,
              double cflreg, int maxsteps, double *log, int logcap,
              int *nlog, int *nrej);

int orc_block_cells(const orc_ctx *c, int blk);   /* (nx+6)*(ny+6) */
/* AoS copy of one layer of one block: cell-major, 9 components per cell,
 * a (x) fastest, ghosts included. */
void orc_get_layer(const orc_ctx *c, int layer, int blk, double *out);
void orc_set_layer(orc_ctx *c, int layer, int blk, const double *in);
/* Output.c:217-243 conversion of DATA for one block: interior plus one ghost
 * ring, i (x) outer / j (y) inner, 10 values per point. */
void orc_output_values(const orc_ctx *c, int blk, double *vals);

/* single-cell physics, exposed for closed-form tests */
void orc_primitive(const double *U, double *P);
void orc_conservative(const double *P, double *U);
void orc_flux_dir(const double *U, double *F, double nx, double ny);
double orc_relax(const orc_ctx *c, const double *U, double *out, double dt,
                 int flag);
double orc_limit(int limiter, double d1, double d2);
double orc_hll(const double *PL, const double *PR, double dx, int dir,
               double *flx);
void orc_initial_state(const orc_ctx *c, double x, double y, double t,
                       double *U);

#ifdef __cplusplus
}
#endif
#endif
