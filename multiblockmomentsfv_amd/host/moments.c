/*
 * moments.c -- drop-in replacement for the reference's ./build/moments
 * (Main.c:61-244): same flags, same block file, same stdout log lines and
 * the same .hdr/.dat output; the time loop runs on the gfx950 engine.
 *
 *   moments -bfile <file> -out <dir> [-Kn v] [-R13 v] [-CFL v]
 *           [-limiter name] [-sfile <restart>] [-double] [-refine n]
 *           [-scenario riemann2d|ffs|tempjump]
 *
 * -scenario is an addition: the reference selects its initial state by
 * editing Scenario.c; the default here is Riemann2D, the problem of the
 * reference's goldens.
 */
#define _GNU_SOURCE   /* program_invocation_short_name, as glibc's assert message uses */
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/time.h>

#include "mbmfv.h"
#include "mbmfv_host.h"

#define DT_INITIAL 0.0001   /* Main.c:28 */

static int arg_index(int argc, const char **argv, const char *key)
{
  for (int i = 1; i < argc; i++)
    if (strcmp(argv[i], key) == 0) return (i + 1 < argc) ? i + 1 : -1;
  return -1;
}

static int has_arg(int argc, const char **argv, const char *key)
{
  for (int i = 1; i < argc; i++)
    if (strcmp(argv[i], key) == 0) return 1;
  return 0;
}

/* the reference parses numeric flags into a float (tools.c:38-46) */
static float float_arg(int argc, const char **argv, const char *key)
{
  int v = arg_index(argc, argv, key);
  if (v < 0) {
    printf("Error: get_float_argument(%s)\n", key);
    return (float)-999999999;
  }
  return atof(argv[v]);
}

static int string_arg(int argc, const char **argv, const char *key, char *out, size_t n)
{
  int v = arg_index(argc, argv, key);
  if (v < 0) {
    fprintf(stderr, "Error: get_string_argument(%s)\n", key);
    return -1;
  }
  snprintf(out, n, "%s", argv[v]);
  return 0;
}

static void make_dirs(const char *path)
{
  char buf[200];
  snprintf(buf, sizeof buf, "%s", path);
  for (int i = 1; buf[i]; i++)
    if (buf[i] == '/') { buf[i] = '\0'; mkdir(buf, 0755); buf[i] = '/'; }
  mkdir(buf, 0755);
}

/* every sync interface joins edges of the same cell count (the condition
 * the reference asserts in FillEdge, Boundary.c:101-103) */
static int conforming(const mbm_deck *d)
{
  for (int k = 0; k < d->nblocks; k++)
    for (int s = 0; s < 4; s++) {
      const int o = d->bc[k][s] - 1;
      if (o >= 0 && o < d->nblocks && ((s < 2) ? d->ny[o] != d->ny[k] : d->nx[o] != d->nx[k]))
        return 0;
    }
  return 1;
}

static int die(const char *what, int rc)
{
  fprintf(stderr, "moments: %s failed (%d): %s\n", what, rc, mbm_last_error());
  return 1;
}

static int write_output(mbm_ctx *ctx, const mbm_deck *d, const char *outdir, const char *type,
                        int float32, double t, int add)
{
  double *vals[MBM_MAX_BLOCKS] = { 0 };
  int rc = 0;
  for (int k = 0; k < d->nblocks && !rc; k++) {
    size_t n = (size_t)(d->nx[k] + 2) * (d->ny[k] + 2) * MBM_NVARS_OUT;
    vals[k] = malloc(n * sizeof(double));
    if (!vals[k]) { rc = MBM_E_NOMEM; break; }
    rc = mbm_output_values(ctx, k, 0, vals[k]);
  }
  if (!rc) rc = mbm_output_write(outdir, type, d, float32, t, add, (const double *const *)vals);
  for (int k = 0; k < d->nblocks; k++) free(vals[k]);
  return rc;
}

static double now_ms(const struct timeval *t0)
{
  struct timeval t;
  gettimeofday(&t, NULL);
  return 1.0e3 * (t.tv_sec - t0->tv_sec) + 1.0e-3 * (t.tv_usec - t0->tv_usec);
}

/* Steps (Main.c:189-244). The whole output interval runs on the device
 * (mbm_steps: controller, rejection and clamp to tend in a replayed graph);
 * the lines the reference prints every 20 steps are written afterwards from
 * the attempt records, in the same format. The perf figure is the interval's
 * mean time per cell-update, as the reference's running average. */
#define STEP_CHUNK 1000
#define REC_CAP 8192

static int steps(mbm_ctx *ctx, int gridsize, double *t, double tend, double cflmax,
                 double cflreg, int *nout)
{
  const int show = 20;
  int nstep = 0, nrej = 0, minutes = 0, seconds = 0, rc = 0;
  double dt = DT_INITIAL, perf = 0.0, diff = 0.0;
  mbm_step_record *rec = malloc(REC_CAP * sizeof *rec);
  if (!rec) return MBM_E_NOMEM;
  struct timeval t0;
  gettimeofday(&t0, NULL);
  while (*t < tend) {
    int nrec = 0, ns = 0, nr = 0;
    const int base = nstep;
    if ((rc = mbm_steps(ctx, t, tend, cflmax, cflreg, STEP_CHUNK, &dt, rec, REC_CAP, &nrec, &ns,
                        &nr)))
      break;
    const double ms = now_ms(&t0);
    const int shown = nrec < REC_CAP ? nrec : REC_CAP;
    for (int i = 0; i < shown; i++) {
      const mbm_step_record *r = &rec[i];
      const int s = base + r->step;
      const int first = i == 0 || rec[i - 1].step != r->step || rec[i - 1].accepted;
      const double next_dt = i + 1 < shown ? rec[i + 1].dt : dt;
      if (s % show) continue;
      if (first) printf(" step: %3d...", s);
      if (!r->accepted) {
        printf("\n   step rejected: CFL = %f (next dt = %f)\n   repeat ...", r->cfl, next_dt);
        continue;
      }
      if (next_dt > 0) {
        diff = 0.0;
        perf = ms / (1.0e-3 * gridsize * (double)(base + ns + nrej + nr));
        minutes = (int)(1.0e-6 * perf * gridsize * (tend - r->t) / (60 * next_dt));
        seconds = (int)(1.0e-6 * perf * gridsize * (tend - r->t) / next_dt) % 60;
      }
      printf(" done.  CFL: %lf, time: %lf, next dt: %lf, diff: %lf, perf = %.2f (ca. %dm%ds)\n",
             r->cfl, r->t, next_dt, diff, perf, minutes, seconds);
    }
    fflush(stdout);
    nstep += ns;
    nrej += nr;
  }
  free(rec);
  *nout = nstep;
  return rc;
}

int main(int argc, const char **argv)
{
  char bfile[200] = "blockfile.txt", sfile[200] = "", outdir[200] = "", name[64];
  char type[200];
  int float32 = 1, refine = 0, scenario = MBM_SCN_RIEMANN2D, rc;
  mbm_deck d;

  if (has_arg(argc, argv, "-sfile")) string_arg(argc, argv, "-sfile", sfile, sizeof sfile);
  if (has_arg(argc, argv, "-bfile")) string_arg(argc, argv, "-bfile", bfile, sizeof bfile);
  if (has_arg(argc, argv, "-out")) {
    string_arg(argc, argv, "-out", outdir, sizeof outdir);
    make_dirs(outdir);
  }
  if (has_arg(argc, argv, "-double")) float32 = 0;
  if (has_arg(argc, argv, "-refine")) {
    refine = (int)float_arg(argc, argv, "-refine");
    if (refine < 1) refine = 1;
  }
  if (has_arg(argc, argv, "-scenario")) {
    string_arg(argc, argv, "-scenario", name, sizeof name);
    if (!strcmp(name, "ffs")) scenario = MBM_SCN_FFS;
    else if (!strcmp(name, "tempjump")) scenario = MBM_SCN_TEMPJUMP;
    else if (strcmp(name, "riemann2d")) {
      fprintf(stderr, "moments: unknown scenario '%s'\n", name);
      return 2;
    }
  }

  if ((rc = mbm_deck_read(bfile, &d, refine, 1))) {
    fprintf(stderr, "moments: %s\n", mbm_host_last_error());
    return 1;
  }
  if (has_arg(argc, argv, "-CFL")) d.cfl = (double)float_arg(argc, argv, "-CFL");
  if (has_arg(argc, argv, "-Kn")) d.kn = (double)float_arg(argc, argv, "-Kn");
  if (has_arg(argc, argv, "-R13")) d.r13 = (double)float_arg(argc, argv, "-R13");
  if (has_arg(argc, argv, "-limiter")) {
    string_arg(argc, argv, "-limiter", name, sizeof name);
    d.limiter = mbm_limiter_code(name);
  }
  const int conf = conforming(&d);
  if (conf && (rc = mbm_deck_check(&d))) {
    fprintf(stderr, "moments: %s\n", mbm_host_last_error());
    return 1;
  }

  printf("\n Running: Kn = %.4g, R13 = %g, CFL = %.2f, limiter = %s, BlockFile = \"%s\"\n",
         d.kn, d.r13, d.cfl, mbm_limiter_name(d.limiter), bfile);
  printf("   Output directory: \"%s\"\n", outdir[0] ? outdir : ".");
  printf("   Initialize: ");
  for (int k = 0; k < d.nblocks; k++) printf("Block %d (Size %d), ", k + 1, d.nx[k] * d.ny[k]);
  if (!conf) {
    /* the reference stops at its conformity assert in the first SetBoundary
     * (Boundary.c:101-103), after the lines above (still buffered) */
    fprintf(stderr, "%s: src/Boundary.c:102: FillEdge: Assertion `(side<2) ? "
            "(Rd.rbt.j-Rd.lfb.j)==(Rs.rbt.j-Rs.lfb.j) : (Rd.rbt.i-Rd.lfb.i)==(Rs.rbt.i-Rs.lfb.i)' "
            "failed.\n", program_invocation_short_name);
    abort();
  }

  mbm_block_desc blk[MBM_MAX_BLOCKS];
  for (int k = 0; k < d.nblocks; k++) {
    blk[k].nx = d.nx[k]; blk[k].ny = d.ny[k];
    blk[k].xmin = d.xmin[k]; blk[k].xmax = d.xmax[k];
    blk[k].ymin = d.ymin[k]; blk[k].ymax = d.ymax[k];
    for (int s = 0; s < 4; s++) blk[k].bc[s] = d.bc[k][s];
  }
  mbm_params par = { d.kn, d.force, d.limiter, scenario, { 5.0, 1.0, 0.4, 0.05, 1.0 } };
  mbm_ctx *ctx = NULL;
  if ((rc = mbm_create(blk, d.nblocks, &par, NULL, &ctx))) return die("mbm_create", rc);

  snprintf(type, sizeof type, "%s", mbm_scenario_type(scenario));
  if (strlen(sfile) > 0) {
    double *cons[MBM_MAX_BLOCKS] = { 0 };
    for (int k = 0; k < d.nblocks; k++)
      cons[k] = calloc((size_t)(d.nx[k] + 6) * (d.ny[k] + 6) * MBM_NV, sizeof(double));
    rc = mbm_restart_read(sfile, &d, cons, type);
    for (int k = 0; k < d.nblocks && !rc; k++) rc = mbm_upload(ctx, MBM_L_DATA, k, cons[k]);
    for (int k = 0; k < d.nblocks; k++) free(cons[k]);
    if (!rc) rc = mbm_set_boundary(ctx, MBM_L_DATA, 0.0);
    if (rc) { mbm_destroy(ctx); return die("restart", rc); }
  } else if ((rc = mbm_init(ctx))) {
    mbm_destroy(ctx);
    return die("mbm_init", rc);
  }
  printf("    done\n");

  if ((rc = write_output(ctx, &d, outdir, type, float32, 0.0, -1))) {
    mbm_destroy(ctx);
    return die("output", rc);
  }
  double t = 0.0, cflmax = 1.0;
  int ntot = 0, gridsize = mbm_full_grid_size(&d);
  for (int n = 0; d.tout[n] > 0; n++) {
    int ns = 0;
    if ((rc = steps(ctx, gridsize, &t, d.tout[n], cflmax, d.cfl, &ns))) {
      mbm_destroy(ctx);
      return die("time step", rc);
    }
    ntot += ns;
    if ((rc = write_output(ctx, &d, outdir, type, float32, t, ntot))) {
      mbm_destroy(ctx);
      return die("output", rc);
    }
  }
  printf("\n Done after %d steps.\n\n", ntot);
  mbm_destroy(ctx);
  return 0;
}
