/*
 * mbm_host.c -- block-file reader, output writer and restart reader of the
 * MI355X build. File formats and printed lines follow the reference
 * (Mesh.c, Output.c) so the build is a drop-in for ./build/moments.
 */
#define _GNU_SOURCE
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>
#include <stdarg.h>
#include <math.h>
#include <time.h>

#include "mbmfv_host.h"
#include "mbmfv.h"
#include "../csrc/mbm_scenario.h"

#ifndef MBM_GIT_COMMIT
#define MBM_GIT_COMMIT "unknown"
#endif

static char g_err[512];

static int fail(int code, const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

const char *mbm_host_last_error(void) { return g_err; }

void mbm_deck_defaults(mbm_deck *d)
{
  memset(d, 0, sizeof *d);
  d->kn = 0.001;
  d->r13 = 0.0;
  d->force = 0.0;
  d->cfl = 0.95;
  d->limiter = MBM_LIM_HARMONIC;
  d->refine = 1;
}

int mbm_limiter_code(const char *name)
{
  static const char *names[] = { "none", "average", "minmod", "superbee", "harmonic" };
  for (int k = 0; k < 5; k++)
    if (strcmp(name, names[k]) == 0) return k;
  fprintf(stderr, "   WARNING: unknown limiter '%s', using 'harmonic'.\n", name);
  return MBM_LIM_HARMONIC;
}

const char *mbm_limiter_name(int code)
{
  switch (code) {
    case MBM_LIM_NONE: return "none";
    case MBM_LIM_AVERAGE: return "average";
    case MBM_LIM_MINMOD: return "minmod";
    case MBM_LIM_SUPERBEE: return "superbee";
    default: return "harmonic";
  }
}

const char *mbm_scenario_type(int scenario)
{
  switch (scenario) {
    case MBM_SCN_FFS: return "ForwardFacingStep";
    case MBM_SCN_TEMPJUMP: return "TempJump";
    default: return "Riemann2D";
  }
}

void mbm_initial_state(int scenario, const double *params, double x, double y,
                       double t, double U[MBM_NV])
{
  static const double tj_default[5] = { 5.0, 1.0, 0.4, 0.05, 1.0 };
  (void)t;
  mbm_scn_eval(scenario, params ? params : tj_default, x, y, U);
}

/* key = value header (Mesh.c:60-122) */
static int read_header(FILE *fp, mbm_deck *d, int refine_locked, int verbose)
{
  char line[300], key[64], val[200], *p;
  double dt_out = 0.0, tl[MBM_MAX_OUT + 1] = { 0.0 };
  int nlist = 0, n;

  if (verbose) printf("\n   Reading parameters...\n");
  while (fgets(line, sizeof line, fp) != NULL) {
    p = line;
    while (*p == ' ' || *p == '\t') p++;
    if (*p == '\0' || *p == '\n' || *p == '#') continue;
    if (p[0] == '/' && p[1] == '/') continue;
    if (strncasecmp(p, "Block data", 10) == 0) break;
    if (sscanf(p, " %63[^= \t] = %199[^\n]", key, val) != 2) continue;
    if ((p = strchr(val, '#')) != NULL) *p = '\0';

    if (!strcasecmp(key, "Kn")) d->kn = atof(val);
    else if (!strcasecmp(key, "R13")) d->r13 = atof(val);
    else if (!strcasecmp(key, "F")) d->force = atof(val);
    else if (!strcasecmp(key, "CFL")) d->cfl = atof(val);
    else if (!strcasecmp(key, "dt_out")) dt_out = atof(val);
    else if (!strcasecmp(key, "t_list")) {
      char *q = val, *e;
      double v;
      nlist = 0;
      while (nlist < MBM_MAX_OUT && (v = strtod(q, &e)) != 0.0 && e != q) { tl[nlist++] = v; q = e; }
      tl[nlist] = 0.0;
    } else if (!strcasecmp(key, "refine")) {
      if (!refine_locked) d->refine = atoi(val);
    } else if (!strcasecmp(key, "limiter")) {
      char name[64] = "";
      sscanf(val, " %63s", name);
      d->limiter = mbm_limiter_code(name);
    } else
      fprintf(stderr, "   WARNING: unknown parameter '%s' ignored.\n", key);
  }

  if (dt_out > 0.0) {
    double tfin = tl[0];
    n = (int)((tfin + 1e-10) / dt_out);
    if (n > MBM_MAX_OUT - 1) n = MBM_MAX_OUT - 1;
    for (int i = 0; i < n; i++) d->tout[i] = (i + 1) * dt_out;
    if (n * dt_out < tfin - 1e-10) d->tout[n++] = tfin;
    d->tout[n] = 0.0;
    d->nout = n;
    if (verbose)
      printf("   Output every dt_out = %g until t_end = %g (%d outputs)\n", dt_out, tfin, n);
  } else {
    for (int i = 0; i < nlist; i++) d->tout[i] = tl[i];
    d->tout[nlist] = 0.0;
    d->nout = nlist;
    if (verbose) {
      printf("   Output at t = ");
      for (int i = 0; i < nlist; i++) printf("%g%s", d->tout[i], i < nlist - 1 ? ", " : "\n");
    }
  }
  if (d->refine < 1) d->refine = 1;
  if (verbose)
    printf("   Kn = %.4g, R13 = %g, F = %g, CFL = %.2f, limiter = %s, refine = %d\n",
           d->kn, d->r13, d->force, d->cfl, mbm_limiter_name(d->limiter), d->refine);
  return 0;
}

int mbm_deck_read(const char *path, mbm_deck *d, int refine_override, int verbose)
{
  char line[300];
  FILE *fp = fopen(path, "r");
  int id, nx, ny, b0, b1, b2, b3, count = 0, seen[MBM_MAX_BLOCKS] = { 0 };
  double x0, x1, y0, y1;

  mbm_deck_defaults(d);
  if (refine_override > 0) d->refine = refine_override;
  if (!fp) { perror(path); return fail(MBM_E_IO, "cannot open block file '%s'", path); }
  read_header(fp, d, refine_override > 0, verbose);
  if (fgets(line, sizeof line, fp) == NULL) {
    fclose(fp);
    fprintf(stderr, "Unexpected EOF before block definitions\n");
    return fail(MBM_E_FORMAT, "unexpected EOF before block definitions in '%s'", path);
  }
  if (verbose) printf("\n   Reading block definitions...\n");
  while (fscanf(fp, "%d/%lf,%lf/%lf,%lf/%d,%d/%d,%d,%d,%d/%*[^\n]",
                &id, &x0, &x1, &y0, &y1, &nx, &ny, &b0, &b1, &b2, &b3) != EOF) {
    if (id < 1 || id > MBM_MAX_BLOCKS) {
      fprintf(stderr, "   ERROR: block number %d out of range [1,%d=NB]. "
              "Fix the block file (this build allows at most %d blocks).\n", id, MBM_MAX_BLOCKS, MBM_MAX_BLOCKS);
      fclose(fp);
      return fail(MBM_E_FORMAT, "block number %d out of range [1,%d]", id, MBM_MAX_BLOCKS);
    }
    nx *= d->refine;
    ny *= d->refine;
    if (verbose)
      printf("    Block %d, x-range: (%+7.2lf,%+7.2lf), y-range: (%+7.2lf,%+7.2lf), grid %dx%d\n",
             id, x0, x1, y0, y1, nx, ny);
    int k = id - 1;
    d->nx[k] = nx; d->ny[k] = ny;
    d->xmin[k] = x0; d->xmax[k] = x1; d->ymin[k] = y0; d->ymax[k] = y1;
    d->bc[k][0] = b0; d->bc[k][1] = b1; d->bc[k][2] = b2; d->bc[k][3] = b3;
    seen[k] = 1;
    count++;
  }
  if (verbose) printf("   ... done.\n\n");
  fclose(fp);
  d->nblocks = count;
  for (int k = 0; k < count; k++)
    if (!seen[k])
      return fail(MBM_E_FORMAT, "block ids must be 1..%d without gaps (missing %d)", count, k + 1);
  return 0;
}

int mbm_deck_check(const mbm_deck *d)
{
  if (d->nblocks < 1) return fail(MBM_E_ARG, "deck has no blocks");
  for (int k = 0; k < d->nblocks; k++) {
    if (d->nx[k] < 1 || d->ny[k] < 1)
      return fail(MBM_E_ARG, "block %d has an empty grid %dx%d", k + 1, d->nx[k], d->ny[k]);
    for (int s = 0; s < 4; s++) {
      int code = d->bc[k][s];
      if (code > 0) {
        int o = code - 1;
        if (o >= d->nblocks)
          return fail(MBM_E_ARG, "block %d side %d syncs with missing block %d", k + 1, s, code);
        int same = (s < 2) ? d->ny[o] == d->ny[k] : d->nx[o] == d->nx[k];
        if (!same)
          return fail(MBM_E_ARG, "non-conforming sync interface: block %d side %d vs block %d",
                      k + 1, s, code);
      }
    }
  }
  return 0;
}

mbm_geom mbm_deck_geom(const mbm_deck *d, int k)
{
  mbm_geom g;
  int nst = 2 * MBM_MBC;
  g.dx = (d->xmax[k] - d->xmin[k]) / d->nx[k];
  g.dy = (d->ymax[k] - d->ymin[k]) / d->ny[k];
  g.bx = d->xmin[k] - (nst / 2 - 0.5) * g.dx;
  g.by = d->ymin[k] - (nst / 2 - 0.5) * g.dy;
  return g;
}

int mbm_full_grid_size(const mbm_deck *d)
{
  int s = 0;
  for (int k = 0; k < d->nblocks; k++) s += d->nx[k] * d->ny[k];
  return s;
}

char *mbm_convert(double a, char *s)
{
  int i = fabs(a + 1e-7);
  int j = fabs(10 * (a - i + 1e-7));
  int k = fabs(100 * (a - i - 0.1 * j + 1e-7));
  int l = fabs(1000 * (a - i - 0.1 * j - 0.01 * k + 1e-7));
  char t[64];   /* room for any four ints; the caller's buffer holds 20 */
  if (l != 0) snprintf(t, sizeof t, "%dp%d%d%d", i, j, k, l);
  else if (k != 0) snprintf(t, sizeof t, "%dp%d%d", i, j, k);
  else snprintf(t, sizeof t, "%dp%d", i, j);
  memcpy(s, t, 19);
  s[19] = '\0';
  return s;
}

static void write_manifest(FILE *fh, const char *datbase, const char *solv,
                           const char *type, const mbm_deck *d, int float32)
{
  char tbuf[64];
  time_t now = time(NULL);
  strftime(tbuf, sizeof tbuf, "%Y-%m-%dT%H:%M:%SZ", gmtime(&now));
  fprintf(fh, "# MultiBlockMomentsFV binary output header\n");
  fprintf(fh, "# FORMAT_VERSION = %d\n", 1);
  fprintf(fh, "# CREATED_UTC = %s\n", tbuf);
  fprintf(fh, "# GIT_COMMIT = %s\n\n", MBM_GIT_COMMIT);
  fprintf(fh, "[RUN]\n");
  fprintf(fh, "solver = %s\n", solv);
  fprintf(fh, "Kn = %.10g\n", d->kn);
  fprintf(fh, "R13 = %.10g\n", d->r13);
  fprintf(fh, "Type = %s\n", type);
  fprintf(fh, "FullGridSize = %d\n", mbm_full_grid_size(d));
  fprintf(fh, "refine = %d\n", d->refine);
  fprintf(fh, "t_output = 0.0");
  for (int k = 0; d->tout[k] > 0; k++) fprintf(fh, " %.10g", d->tout[k]);
  fprintf(fh, "\n\n");
  fprintf(fh, "[BINARY]\n");
  fprintf(fh, "BINARY_FILE = %s\n", datbase);
  fprintf(fh, "NVARS = %d\n", MBM_NVARS_OUT);
  fprintf(fh, "ORDER = variable-fastest\n");
  fprintf(fh, "PRECISION = %s\n", float32 ? "Float32" : "Float64");
  fprintf(fh, "ENDIAN = little\n\n");
  fprintf(fh, "[VARIABLES]\n");
  fprintf(fh, "rho vx vy T p Pxx Pxy Pyy qx qy\n\n");
  fprintf(fh, "[ZONES]\n");
  fprintf(fh, "# block I J NVARS BX BY DX DY LFB_I LFB_J ISTART JSTART NGHOST SOLUTIONTIME STRANDID OFFSET NBYTES\n");
}

int mbm_output_write(const char *outdir, const char *type, const mbm_deck *d,
                     int float32, double t, int add, const double *const *vals)
{
  char s1[20], s2[20], solv[1024], prefix[224];
  char *base = NULL, *hdr = NULL, *dat = NULL;
  size_t esz = float32 ? sizeof(float) : sizeof(double);
  FILE *fh, *fd;
  int rc = 0;

  if (outdir && outdir[0]) snprintf(prefix, sizeof prefix, "%s/", outdir);
  else prefix[0] = '\0';
  snprintf(solv, sizeof solv, "%s%s", d->r13 < 1.0 ? "Grad13" : "Reg13", type);
  if (!add) mbm_convert(t, s2);
  else snprintf(s2, sizeof s2, "All");
  if (asprintf(&base, "%s%s_nst%d_Kn%s_t%s", prefix, solv, mbm_full_grid_size(d),
               mbm_convert(d->kn, s1), s2) < 0)
    return fail(MBM_E_NOMEM, "asprintf failed");
  if (asprintf(&hdr, "%s.hdr", base) < 0) hdr = NULL;
  if (asprintf(&dat, "%s.dat", base) < 0) dat = NULL;
  if (!hdr || !dat) { free(base); free(hdr); free(dat); return fail(MBM_E_NOMEM, "asprintf failed"); }
  const char *datbase = strrchr(dat, '/');
  datbase = datbase ? datbase + 1 : dat;

  if (add <= 0) {
    add = 0;
    fh = fopen(hdr, "w");
    fd = fopen(dat, "wb");
  } else {
    fh = fopen(hdr, "a");
    fd = fopen(dat, "ab");
  }
  if (!fh || !fd) {
    perror(!fh ? hdr : dat);
    rc = fail(MBM_E_IO, "cannot open '%s'", !fh ? hdr : dat);
    goto done;
  }
  if (add == 0) write_manifest(fh, datbase, solv, type, d, float32);
  if (add == 0) printf("\n  Create File: %s(.hdr/.dat)...", base);
  else printf("\n  Append (%d) File: %s(.hdr/.dat)...", add, base);
  fflush(stdout);

  for (int k = 0; k < d->nblocks; k++) {
    int ni = d->ny[k] + 2, nj = d->nx[k] + 2;
    size_t npt = (size_t)ni * nj, nval = npt * MBM_NVARS_OUT;
    long long nbytes = (long long)nval * (long long)esz;
    long offset = ftell(fd);
    if (float32) {
      float *buf = malloc(nval * sizeof(float));
      if (!buf) { rc = fail(MBM_E_NOMEM, "out of host memory"); goto done; }
      for (size_t m = 0; m < nval; m++) buf[m] = (float)vals[k][m];
      fwrite(buf, sizeof(float), nval, fd);
      free(buf);
    } else {
      fwrite(vals[k], sizeof(double), nval, fd);
    }
    mbm_geom g = mbm_deck_geom(d, k);
    fprintf(fh, "ZONE block=%d I=%d J=%d NVARS=%d "
            "BX=%.16g BY=%.16g DX=%.16g DY=%.16g LFB_I=%d LFB_J=%d "
            "ISTART=%d JSTART=%d NGHOST=%d "
            "SOLUTIONTIME=%.16g STRANDID=%d OFFSET=%ld NBYTES=%lld\n",
            k + 1, ni, nj, MBM_NVARS_OUT, g.bx, g.by, g.dx, g.dy, 1, 1,
            1 + MBM_MBC - 1, 1 + MBM_MBC - 1, 1, t, add + 1, offset, nbytes);
  }
  printf(" done\n\n");
done:
  if (fh) fclose(fh);
  if (fd) fclose(fd);
  free(base); free(hdr); free(dat);
  return rc;
}

int mbm_restart_read(const char *path, const mbm_deck *d, double *const *cons,
                     char *type_out)
{
  char base[500], hdrname[520], datname[520], line[600];
  int k = 0, single = 0;
  size_t len;
  FILE *fh, *fd;

  snprintf(type_out, 200, "FromFile(%s)", path);
  strncpy(base, path, sizeof base - 1);
  base[sizeof base - 1] = '\0';
  len = strlen(base);
  if (len > 4 && (!strcmp(base + len - 4, ".hdr") || !strcmp(base + len - 4, ".dat")))
    base[len - 4] = '\0';
  snprintf(hdrname, sizeof hdrname, "%s.hdr", base);
  snprintf(datname, sizeof datname, "%s.dat", base);
  printf("\n   Read file: %s(.hdr/.dat)...", base);
  fh = fopen(hdrname, "r");
  fd = fopen(datname, "rb");
  if (!fh || !fd) {
    perror(!fh ? hdrname : datname);
    if (fh) fclose(fh);
    if (fd) fclose(fd);
    return fail(MBM_E_IO, "cannot open restart file '%s'", !fh ? hdrname : datname);
  }
  while (fgets(line, sizeof line, fh) != NULL && k < d->nblocks) {
    char *q;
    if ((q = strstr(line, "PRECISION")) != NULL) { single = strstr(q, "Float32") != NULL; continue; }
    if (strncmp(line, "Type", 4) == 0 && (q = strchr(line, '=')) != NULL) {
      sscanf(q + 1, " %199[^\n]", type_out);
      continue;
    }
    if (strncmp(line, "ZONE", 4) != 0) continue;
    long off = 0;
    if ((q = strstr(line, "OFFSET=")) != NULL) off = atol(q + 7);
    fseek(fd, off, SEEK_SET);
    int lx = d->nx[k] + 2 * MBM_MBC;
    for (int a = MBM_MBC - 1; a <= d->nx[k] + MBM_MBC; a++)
      for (int b = MBM_MBC - 1; b <= d->ny[k] + MBM_MBC; b++) {
        double v[10], P[MBM_NV];
        if (single) {
          float f[10];
          if (fread(f, sizeof(float), 10, fd) != 10) goto eof;
          for (int m = 0; m < 10; m++) v[m] = (double)f[m];
        } else if (fread(v, sizeof(double), 10, fd) != 10) goto eof;
        /* stored: rho vx vy T p Pxx Pxy Pyy qx qy; placed as Mesh.c:263-271 */
        P[0] = v[0]; P[1] = v[1]; P[2] = v[2]; P[3] = v[4];
        P[4] = v[5]; P[5] = v[6]; P[6] = v[7]; P[7] = v[8]; P[8] = v[9];
        double *U = cons[k] + (size_t)MBM_NV * ((size_t)b * lx + a);
        {
          double rho = P[0], v1 = P[1], v2 = P[2], th = P[3];
          double vs = v1 * v1 + v2 * v2;
          double E = 0.5 * rho * vs + 1.5 * rho * th;
          U[0] = rho;
          U[1] = rho * v1;
          U[2] = rho * v2;
          U[3] = E;
          U[4] = rho * (P[4] + v1 * v1);
          U[5] = rho * (P[5] + v1 * v2);
          U[6] = rho * (P[6] + v2 * v2);
          U[7] = P[7] + E * v1 + rho * (P[4] * v1 + P[5] * v2);
          U[8] = P[8] + E * v2 + rho * (P[5] * v1 + P[6] * v2);
        }
      }
    k++;
  }
  fclose(fh);
  fclose(fd);
  printf(" done\n");
  return 0;
eof:
  fclose(fh);
  fclose(fd);
  fprintf(stderr, "Unexpected EOF in %s\n", datname);
  return fail(MBM_E_IO, "unexpected EOF in %s", datname);
}
