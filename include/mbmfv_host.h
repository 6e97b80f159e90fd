/*
 * mbmfv_host.h -- host-side C interface of the MI355X MultiBlockMomentsFV
 * build: block-file reader, output writer, restart reader and the scenario
 * initial states. These keep the reference's file formats and CLI behaviour
 * unchanged; the device engine is declared in mbmfv.h.
 *
 * All functions are plain C, single-threaded, and return 0 on success or a
 * negative MBM_E* code (mbmfv.h) with a message available from
 * mbm_host_last_error().
 */
#ifndef MBMFV_HOST_H
#define MBMFV_HOST_H

#ifdef __cplusplus
extern "C" {
#endif

#define MBM_NV 9              /* conservative components per cell (defs.h:11) */
#define MBM_MBC 3             /* ghost rings (defs.h:14) */
#define MBM_MAX_BLOCKS 32     /* NB (defs.h:13) */
#define MBM_MAX_OUT 50        /* NOUT (defs.h:15) */
#define MBM_NVARS_OUT 10      /* rho vx vy T p Pxx Pxy Pyy qx qy (Output.c:30) */

/* slope limiters, same codes as Limiter.h:10-16 */
enum { MBM_LIM_NONE = 0, MBM_LIM_AVERAGE = 1, MBM_LIM_MINMOD = 2,
       MBM_LIM_SUPERBEE = 3, MBM_LIM_HARMONIC = 4 };

/* initial/Dirichlet states. Riemann2D and ForwardFacingStep are the two
 * problems of Scenario.c:72-153 (the reference picks one by editing that
 * file); TempJump is a smoothed temperature jump used by the adjoint. */
enum { MBM_SCN_RIEMANN2D = 0, MBM_SCN_FFS = 1, MBM_SCN_TEMPJUMP = 2 };

/* boundary codes as in the block file: >0 sync with that block id,
 * 0 open, -1 wall, -2 dirichlet (defs.h:26-31) */

/* Parsed block file: Mesh.c:60-184. */
typedef struct {
  int nblocks;                          /* number of block lines */
  int nx[MBM_MAX_BLOCKS], ny[MBM_MAX_BLOCKS];   /* interior cells, refine applied */
  double xmin[MBM_MAX_BLOCKS], xmax[MBM_MAX_BLOCKS];
  double ymin[MBM_MAX_BLOCKS], ymax[MBM_MAX_BLOCKS];
  int bc[MBM_MAX_BLOCKS][4];            /* left, right, bottom, top */
  double kn, r13, force, cfl;           /* Kn, R13, F, CFL (Scenario.c:22-24, Mesh.c:24) */
  int limiter;                          /* MBM_LIM_* */
  int refine;
  int nout;                             /* post-IC output times */
  double tout[MBM_MAX_OUT + 1];         /* 0-terminated (Mesh.c:26) */
} mbm_deck;

/* Geometry of block k as built by SetBlocks (Mesh.c:161-172): centre of the
 * local cell (0,0) (ghosts included) and the spacing. */
typedef struct { double bx, by, dx, dy; } mbm_geom;

const char *mbm_host_last_error(void);

/* Defaults of Scenario.c:22-24, Mesh.c:24-32, Limiter.c:20. */
void mbm_deck_defaults(mbm_deck *d);

/* Reads a block file (ParseHeader Mesh.c:60-122, SetBlocks Mesh.c:130-184).
 * refine_override > 0 replaces the file's 'refine' (the -refine flag,
 * Main.c:87-91). verbose != 0 prints the reference's progress lines. */
int mbm_deck_read(const char *path, mbm_deck *d, int refine_override, int verbose);

/* Checks what SetBoundary asserts (Boundary.c:101-103): every sync edge
 * names an existing block with the same tangential extent. */
int mbm_deck_check(const mbm_deck *d);

mbm_geom mbm_deck_geom(const mbm_deck *d, int k);
int mbm_full_grid_size(const mbm_deck *d);           /* tools.c:87-95 */
int mbm_limiter_code(const char *name);              /* Limiter.c:66-75 */
const char *mbm_limiter_name(int code);              /* Limiter.c:77-86 */
const char *mbm_scenario_type(int scenario);         /* 'Type' string */
char *mbm_convert(double a, char *s);                /* Output.c:60-71 */

/* Conservative initial/Dirichlet state at (x,y,t) (Scenario.c:72-153).
 * params: TempJump {theta_in, theta_out, radius, width, rho}; may be NULL. */
void mbm_initial_state(int scenario, const double *params, double x, double y,
                       double t, double U[MBM_NV]);

/* Output writer (Output.c:124-267). Writes/appends
 * <outdir>/<Grad13|Reg13><type>_nst<N>_Kn<kn>_t<All|time>.{hdr,dat}.
 * vals[k] holds block k's 10 values per point for the interior plus one
 * ghost ring, i (x) outer / j (y) inner, as doubles; float32 selects the
 * payload precision. add: -1 new tAll files (STRANDID 1), >0 append with
 * STRANDID add+1, 0 single-time files. */
int mbm_output_write(const char *outdir, const char *type, const mbm_deck *d,
                     int float32, double t, int add,
                     const double *const *vals);

/* Restart reader (LoadData Mesh.c:198-283), bug-compatible: it places the
 * stored p and P_ij where the reference does (into the theta and T_ij slots).
 * cons[k] receives block k's conservative state, AoS (9 per cell), cells
 * (nx+6)*(ny+6) with x fastest; only the interior plus one ring is written.
 * type_out (>=200 bytes) receives the manifest's Type. */
int mbm_restart_read(const char *path, const mbm_deck *d, double *const *cons,
                     char *type_out);

#ifdef __cplusplus
}
#endif
#endif
