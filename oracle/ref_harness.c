/*
 * ref_harness.c -- harness linked against the reference's own objects
 * (oracle/Makefile builds them from the reference sources into oracle/_ref).
 * TEST INFRASTRUCTURE ONLY.
 *
 *   ref_harness <moments CLI args>          runs the reference main unchanged
 *   ref_harness steplog -bfile F -nsteps N -dump OUT [-threads T] [-time]
 *       initialises like main (Main.c:113-138), then runs the Steps controller
 *       (Main.c:196-241) towards the first output time for N accepted steps,
 *       calling the reference TimeUpdate directly. Writes to OUT:
 *         OUT.log : one line per TimeUpdate call: cfl dt_used t_after accepted
 *                   (each value printed with %a, exact)
 *         OUT.bin : DATA layer of every block after the last step, float64,
 *                   cell-major AoS (9 per cell), ghosts included, block order.
 *       With -time only the step loop is timed and printed (no files).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "database_range.h"
#include "defs.h"
#include "tools.h"
#include "Flux.h"
#include "Scenario.h"
#include "Limiter.h"
#include "Mesh.h"
#include "Output.h"

int ref_cli_main(int argc, const char **argv);
double TimeUpdate(DB_Block *B, int bMax, double dt, double t);
void SetBoundary(DB_Block *B, int bMax, int L, double t);

static const char *arg(int argc, const char **argv, const char *k)
{
  for (int i = 1; i + 1 < argc; i++)
    if (!strcmp(argv[i], k)) return argv[i + 1];
  return NULL;
}

static int steplog(int argc, const char **argv)
{
  static DB_Block B[NB];
  const char *bfile = arg(argc, argv, "-bfile");
  const char *ns = arg(argc, argv, "-nsteps");
  const char *out = arg(argc, argv, "-dump");
  const char *th = arg(argc, argv, "-threads");
  int timing = 0;
  for (int i = 1; i < argc; i++) if (!strcmp(argv[i], "-time")) timing = 1;
  if (!bfile || !ns || (!out && !timing)) {
    fprintf(stderr, "usage: steplog -bfile F -nsteps N (-dump OUT | -time) [-threads T]\n");
    return 2;
  }
#ifdef _OPENMP
  if (th) omp_set_num_threads(atoi(th));
#endif
  int nsteps = atoi(ns);
  int bMax = SetBlocks(B, (char *)bfile);
  if (bMax <= 0) return 3;
  for (int nb = 0; nb < bMax; nb++) {
    DB_Range R = B[nb].range;
    for (int l = 0; l < N_LAYER; l++) B[nb].layer[l] = calloc(DBR_Size(R), sizeof(CVector));
    B[nb].Aux = calloc(DBR_Size(R), sizeof(AuxData));
  }
  double U[N];
  for (int nb = 0; nb < bMax; nb++) {
    DB_Range R = B[nb].range;
    int line = R.rbt.i - R.lfb.i + 1;
    for (int i = R.lfb.i + MBC; i < R.rbt.i - 2; i++)
      for (int j = R.lfb.j + MBC; j < R.rbt.j - 2; j++) {
        int idx = (j - R.lfb.j) * line + (i - R.lfb.i);
        SetInitial(U, GetPoint(i, j, 1, B[nb]), 0.0);
        for (int m = 0; m < N; m++) B[nb].layer[L_DATA][idx].U[m] = U[m];
      }
  }
  SetBoundary(B, bMax, L_DATA, 0.0);

  FILE *lg = NULL;
  if (!timing) {
    char name[1024];
    snprintf(name, sizeof name, "%s.log", out);
    lg = fopen(name, "w");
    if (!lg) { perror(name); return 4; }
  }
  double t = 0.0, tend = TOut[0], dt = 0.0001, cfl, cflmax = 1.0;
  int nstep = 0, nrej = 0;
  struct timeval t0, t1;
  gettimeofday(&t0, NULL);
  while (t < tend && nstep < nsteps) {
    nstep++;
    do {
      double dt_used = dt;
      cfl = TimeUpdate(B, bMax, dt, t);
      int acc = !(cfl > cflmax);
      if (!acc) { dt = dt * CFLreg / cfl; nrej++; }
      else { t += dt; dt = dt * CFLreg / cfl; if (t + dt > tend) dt = tend - t; }
      if (lg) fprintf(lg, "%a %a %a %d\n", cfl, dt_used, t, acc);
    } while (cfl > cflmax);
    for (int nb = 0; nb < bMax; nb++)
      memcpy(B[nb].layer[L_DATA], B[nb].layer[L_NEW], DBR_Size(B[nb].range) * sizeof(CVector));
  }
  gettimeofday(&t1, NULL);
  double sec = (t1.tv_sec - t0.tv_sec) + 1e-6 * (t1.tv_usec - t0.tv_usec);
  if (timing) {
    int threads = 1;
#ifdef _OPENMP
    threads = omp_get_max_threads();
#endif
    printf("{\"steps\": %d, \"rejected\": %d, \"seconds\": %.6f, \"cells\": %d, \"threads\": %d}\n",
           nstep, nrej, sec, FullGridSize(B, bMax), threads);
    return 0;
  }
  fclose(lg);
  char name[1024];
  snprintf(name, sizeof name, "%s.bin", out);
  FILE *fb = fopen(name, "wb");
  if (!fb) { perror(name); return 4; }
  for (int nb = 0; nb < bMax; nb++)
    fwrite(B[nb].layer[L_DATA], sizeof(CVector), DBR_Size(B[nb].range), fb);
  fclose(fb);
  return 0;
}

int main(int argc, const char **argv)
{
  if (argc > 1 && !strcmp(argv[1], "steplog")) return steplog(argc - 1, argv + 1);
  return ref_cli_main(argc, argv);
}
