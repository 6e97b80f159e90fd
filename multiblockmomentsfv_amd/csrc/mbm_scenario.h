/*
 * mbm_scenario.h -- initial / Dirichlet states, shared by the host C code
 * (gcc) and the device kernels (hipcc) so both evaluate the same arithmetic.
 *
 * Riemann2D and ForwardFacingStep restate Scenario.c:72-153 (the reference
 * selects one by editing that file; here it is a run-time choice).
 * TempJump is a smoothed temperature jump with no reference counterpart:
 *   rho = params[4], v = 0, T_ij = theta*delta_ij, q = 0,
 *   theta = th_out + (th_in - th_out) * 0.5 * (1 - tanh((r - radius)/width)).
 */
#ifndef MBM_SCENARIO_H
#define MBM_SCENARIO_H

#include <math.h>

#ifdef __HIPCC__
#define MBM_HD __host__ __device__
#else
#define MBM_HD
#endif

#define MBM_SCN_RIEMANN2D_ 0
#define MBM_SCN_FFS_ 1
#define MBM_SCN_TEMPJUMP_ 2

/* conservative vector of a Maxwellian-like state (Scenario.c:139-152) */
MBM_HD static inline void mbm_scn_state(double rho, double v1, double v2,
                                        double th, double *U)
{
#ifdef __clang__
#pragma clang fp contract(off)
#endif
  double t11 = th, t22 = th, t12 = 0.0, q1 = 0.0, q2 = 0.0;
  double vs = v1 * v1 + v2 * v2;
  double E = 0.5 * rho * vs + 1.5 * rho * th;
  U[0] = rho;
  U[1] = rho * v1;
  U[2] = rho * v2;
  U[3] = E;
  U[4] = rho * (t11 + v1 * v1);
  U[5] = rho * (t12 + v1 * v2);
  U[6] = rho * (t22 + v2 * v2);
  U[7] = q1 + E * v1 + rho * (t11 * v1 + t12 * v2);
  U[8] = q2 + E * v2 + rho * (t12 * v1 + t22 * v2);
}

MBM_HD static inline void mbm_scn_eval(int scenario, const double *tj,
                                       double x, double y, double *U)
{
#ifdef __clang__
#pragma clang fp contract(off)
#endif
  if (scenario == MBM_SCN_FFS_) {
    mbm_scn_state(1.0, 2.0 * sqrt(5.0 / 3.0), 0.0, 1.0, U);
  } else if (scenario == MBM_SCN_TEMPJUMP_) {
    double r = sqrt(x * x + y * y);
    double s = 0.5 * (1.0 - tanh((r - tj[2]) / tj[3]));
    mbm_scn_state(tj[4], 0.0, 0.0, tj[1] + (tj[0] - tj[1]) * s, U);
  } else {
    double rho = (sqrt(x * x + y * y) < 0.4) ? 5.0 : 1.0;
    mbm_scn_state(rho, 0.0, 0.0, 1.0, U);
  }
}

#endif
