/* CPU restatement of the block-time-step Hermite pieces of the reference's
 * sample integrator — TEST INFRASTRUCTURE ONLY (included by pg_oracle.c and
 * cpu_batch.c).  FP64, in the reference's operation order; compile without
 * FMA contraction.  The jerk is kept in double (the reference's struct
 * Particle stores it as float, PG6/banana/particle.h:17-19).
 *
 * Pinning: the predictor and corrector are the closed-form Hermite
 * polynomials (tests check them against the Taylor series of a known orbit);
 * calc_modulus against its definition on the block-time grid; the step
 * criteria against direct evaluation of their formulas. */
#ifndef HERMITE_REF_H
#define HERMITE_REF_H
#include <math.h>

#define HR_SECOND 5e-1                 /* hermite.c:16 */
#define HR_THIRD 1.666666666666667e-1  /* hermite.c:17 */
#define HR_FOURTH 4.166666666666667e-2 /* hermite.c:18 */
#define HR_FIFTH 8.333333333333333e-3  /* hermite.c:19 */

/* calc_modulus, timestep.c:176-190 */
static inline double hr_calc_modulus(double num, double mod, double modmax) {
  double tmpmod = modmax;
  while (num > 0.0) {
    while (tmpmod > num) tmpmod *= 0.5;
    if (tmpmod < mod) break;
    num -= tmpmod;
  }
  return num;
}

/* dt_startup, timestep.c:24-48 (0 instead of exit when no step fits) */
static inline double hr_dt_startup(double eta_s, double dtmax, double dtmin, double t_gl,
                                   const double *a, const double *j) {
  double acc = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
  double jrk = j[0] * j[0] + j[1] * j[1] + j[2] * j[2];
  double dtcrit = eta_s * sqrt(acc / jrk);
  double dtblock = dtmax;
  while (dtblock > dtcrit && dtblock > dtmin) dtblock = dtblock * 0.5;
  while (hr_calc_modulus(t_gl, dtblock, dtmax) != 0.0) {
    dtblock = dtblock * 0.5;
    if (dtblock < dtmin) return 0.0;
  }
  return dtblock;
}

/* predict_exmotion, hermite.c:39-62, one particle */
static inline void hr_predict(double t_gl, double t, const double *x, const double *v,
                              const double *a, const double *j, double *xp, double *vp) {
  double dt = t_gl - t, dt2 = dt * dt, dt3 = dt * dt2;
  for (int d = 0; d < 3; ++d) {
    xp[d] = x[d] + v[d] * dt + a[d] * dt2 * HR_SECOND + j[d] * dt3 * HR_THIRD;
    vp[d] = v[d] + a[d] * dt + j[d] * dt2 * HR_SECOND;
  }
}

/* correct (hermite.c:66-110) then dt_criterion (timestep.c:50-79) for one
 * particle stepping from t over dt; (xp, vp) predicted, (a1, j1) the new
 * force.  x1, v1 get the corrected state; returns the new step. */
static inline double hr_correct(double eta, double dtmax, double dtmin, double t, double dts,
                                const double *a0, const double *j0, const double *xp,
                                const double *vp, const double *a1, const double *j1,
                                double *x1, double *v1) {
  double dtsinv = 1.0 / dts, s0[3], c0[3], s1[3];
  for (int k = 0; k < 3; ++k) {
    x1[k] = xp[k];
    v1[k] = vp[k];
    s0[k] = 2.0 * (-3.0 * (a0[k] - a1[k]) - (2.0 * j0[k] + j1[k]) * dts) * dtsinv * dtsinv;
    c0[k] = 6.0 * (2.0 * (a0[k] - a1[k]) + (j0[k] + j1[k]) * dts) * dtsinv * dtsinv * dtsinv;
    x1[k] += s0[k] * dts * dts * dts * dts * HR_FOURTH +
             c0[k] * dts * dts * dts * dts * dts * HR_FIFTH;
    v1[k] += s0[k] * dts * dts * dts * HR_THIRD + c0[k] * dts * dts * dts * dts * HR_FOURTH;
  }
  double ts = t + dts;
  for (int k = 0; k < 3; ++k) s1[k] = s0[k] + c0[k] * dts;
  double acc = a1[0] * a1[0] + a1[1] * a1[1] + a1[2] * a1[2];
  double jrk = j1[0] * j1[0] + j1[1] * j1[1] + j1[2] * j1[2];
  double snp = s1[0] * s1[0] + s1[1] * s1[1] + s1[2] * s1[2];
  double crc = c0[0] * c0[0] + c0[1] * c0[1] + c0[2] * c0[2];
  acc = sqrt(acc);
  jrk = sqrt(jrk);
  snp = sqrt(snp);
  crc = sqrt(crc);
  double dti = eta * sqrt((acc * snp + jrk * jrk) / (jrk * crc + snp * snp));
  if (dti < dts) {
    while (dts > dti && dts > dtmin) dts = dts * 0.5;
  } else if (dti > 2.0 * dts) {
    if (hr_calc_modulus(ts, 2.0 * dts, dtmax) == 0.0) dts = 2.0 * dts;
  }
  return dts;
}

#endif
