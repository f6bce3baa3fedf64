/* CPU oracle — TEST INFRASTRUCTURE ONLY (see pg_oracle.h for the pinning of
 * each function).  Plain C, FP64, compiled with FMA contraction off so the
 * table routines reproduce the reference's IEEE operation sequence exactly.
 * OpenMP over i only changes speed, never results (each i is summed serially
 * in j order). */
#include "pg_oracle.h"

#include <math.h>
#include <stdint.h>
#include <string.h>
#include <xmmintrin.h>
#include "hermite_ref.h"

/* ---------------------------------------------------------------- GRAPE-5 */
/* PG5/newton/libpg5/rsqrt.c:42-81, the loop in its order, on given values. */
double or_rsqrt_bias_of(const float *y0v, const float *y1v) {
  union {
    int i;
    float f;
  } b0, b1;
  double sw = 0.0, se = 0.0;
  for (int i = 0; i < (1 << 23); i++) {
    b0.i = (127 << 23) + i;
    b1.i = (128 << 23) + i;
    const double x0 = (double)b0.f, x1 = (double)b1.f;
    const double y0 = (double)y0v[i], y1 = (double)y1v[i];
    const double e0 = y0 * y0 * x0 - 1., e1 = y1 * y1 * x1 - 1.;
    const double w = y0 * y0;
    sw += w;
    se += w * (e0 + e1);
  }
  sw *= 2.0;
  return 0.5 * se / sw;
}

/* The reference's rsqrt() (rsqrt.c:12-22): the SSE rsqrtss estimate. */
void or_rsqrtss_table(float *y0, float *y1) {
  union {
    int i;
    float f;
  } b;
  for (int i = 0; i < (1 << 23); i++) {
    float r;
    b.i = (127 << 23) + i;
    __m128 v = _mm_rsqrt_ss(_mm_load_ss(&b.f));
    _mm_store_ss(&r, v);
    y0[i] = r;
    b.i = (128 << 23) + i;
    v = _mm_rsqrt_ss(_mm_load_ss(&b.f));
    _mm_store_ss(&r, v);
    y1[i] = r;
  }
}

/* PG5/newton/libpg5/gravity_avx2.c:95-154: dx = xj - xi, r2 = |dx|^2 + eps2,
 * phi accumulates m/r (output sign made positive by phantom_g5.c:52,258),
 * a accumulates m dx / r^3.  Symmetric softening adds eps_j^2
 * (gravity_avx2.c:178-250, direct.c:96-131). */
void or_pg5_force(int ni, const double *xi, const double *eps2i, int nj,
                  const double *xj, const double *mj, const double *eps2j,
                  double *acc, double *pot) {
  int i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < ni; i++) {
    double ax = 0, ay = 0, az = 0, p = 0;
    for (int j = 0; j < nj; j++) {
      double dx = xj[3 * j] - xi[3 * i];
      double dy = xj[3 * j + 1] - xi[3 * i + 1];
      double dz = xj[3 * j + 2] - xi[3 * i + 2];
      double r2 = dx * dx + dy * dy + dz * dz + eps2i[i];
      if (eps2j) r2 += eps2j[j];
      double rinv = 1.0 / sqrt(r2);
      double mr = mj[j] * rinv;
      double mr3 = mr * rinv * rinv;
      p += mr;
      ax += mr3 * dx;
      ay += mr3 * dy;
      az += mr3 * dz;
    }
    acc[3 * i] = ax;
    acc[3 * i + 1] = ay;
    acc[3 * i + 2] = az;
    pot[i] = p;
  }
}

/* ---------------------------------------------------------------- GRAPE-6 */
/* PG6/libpg6/gravity.c:19-20 macro predict(dt,x,v,a2,j6) and the velocity
 * form of :171-176 (2 a2 and 3 j6 with a zero cubic term), left to right. */
void or_pg6_predict(int nj, double ti, const double *x, const double *v,
                    const double *a2, const double *j6, const double *tj,
                    double *xp, double *vp) {
  for (int j = 0; j < nj; j++) {
    double dt = ti - tj[j];
    for (int k = 0; k < 3; k++) {
      int q = 3 * j + k;
      xp[q] = x[q] + v[q] * dt + a2[q] * dt * dt + j6[q] * dt * dt * dt;
      vp[q] = v[q] + (2.0 * a2[q]) * dt + (3.0 * j6[q]) * dt * dt;
    }
  }
}

/* Scalar FP64 recipe of gravity.c:195-258 (force, jerk, pot, self skipped by
 * id) with the nearest-neighbour rule of gravity.c:554-564: the smallest
 * r^2 + eps^2 wins, the first j in address order on ties. */
void or_pg6_force(int ni, const double *xi, const double *vi, const int *idi,
                  double eps2, int nj, const double *xj, const double *vj,
                  const double *mj, const int *idj, double *acc, double *jrk,
                  double *pot, int *nnb, double *rnnb) {
  int i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < ni; i++) {
    double ax = 0, ay = 0, az = 0, jx = 0, jy = 0, jz = 0, p = 0;
    double r2min = INFINITY;
    int nn = -1;
    for (int j = 0; j < nj; j++) {
      if (idj[j] == idi[i]) continue;
      double dx = xj[3 * j] - xi[3 * i];
      double dy = xj[3 * j + 1] - xi[3 * i + 1];
      double dz = xj[3 * j + 2] - xi[3 * i + 2];
      double dvx = vj[3 * j] - vi[3 * i];
      double dvy = vj[3 * j + 1] - vi[3 * i + 1];
      double dvz = vj[3 * j + 2] - vi[3 * i + 2];
      double r2 = dx * dx + dy * dy + dz * dz + eps2;
      double rv = dx * dvx + dy * dvy + dz * dvz;
      double rinv2 = 1.0 / r2;
      double rinv = sqrt(rinv2);
      double mr = mj[j] * rinv;
      double mr3 = mr * rinv2;
      double alpha = 3.0 * rv * rinv2;
      p -= mr;
      ax += mr3 * dx;
      ay += mr3 * dy;
      az += mr3 * dz;
      jx += mr3 * (dvx - alpha * dx);
      jy += mr3 * (dvy - alpha * dy);
      jz += mr3 * (dvz - alpha * dz);
      if (r2 < r2min) {
        r2min = r2;
        nn = j;
      }
    }
    acc[3 * i] = ax;
    acc[3 * i + 1] = ay;
    acc[3 * i + 2] = az;
    jrk[3 * i] = jx;
    jrk[3 * i + 1] = jy;
    jrk[3 * i + 2] = jz;
    pot[i] = p;
    if (nnb) nnb[i] = nn >= 0 ? idj[nn] : -1;
    if (rnnb) rnnb[i] = nn >= 0 ? sqrt(r2min) : 0.0;
  }
}

/* gravity.c:788-793 flags -2/r <= -2/h, i.e. sqrt(|dx|^2+eps^2) <= h, with
 * the self pair masked; :892-909 compacts in j order.  h2 == 0 means none
 * (gravity.c:689-696). */
void or_pg6_nblist(int ni, const double *xi, const int *idi, const double *h2,
                   double eps2, int nj, const double *xj, const int *idj,
                   int maxlen, int *nbl, int *nblen) {
  int i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < ni; i++) {
    int n = 0;
    if (h2[i] > 0.0) {
      for (int j = 0; j < nj; j++) {
        if (idj[j] == idi[i]) continue;
        double dx = xj[3 * j] - xi[3 * i];
        double dy = xj[3 * j + 1] - xi[3 * i + 1];
        double dz = xj[3 * j + 2] - xi[3 * i + 2];
        double r2 = dx * dx + dy * dy + dz * dz + eps2;
        if (r2 <= h2[i]) {
          if (n < maxlen) nbl[(size_t)i * maxlen + n] = idj[j];
          n++;
        }
      }
    }
    nblen[i] = n;
  }
}

/* ------------------------------------------------------------ TreePM table */
/* S2 force shape in r^-3 units (pg5_table.c:8-22).  The polynomial is kept in
 * the reference's Horner order so that table entries are bit-identical. */
double or_s2_reff(double sft, double r) {
  double eta = 2.0 * r / sft;
  if (eta < 1.0) {
    return (224. + eta * eta * (-224. + eta * (70. + eta * (48. - eta * 21.)))) * 2.e0 /
           (35.0 * pow(sft, 3));
  } else if (eta < 2.0) {
    return (12. / (eta * eta) - 224. +
            eta * (896. + eta * (-840. + eta * (224. + eta * (70. + eta * (-48. + eta * 7.)))))) *
           2.e0 / (35.0 * eta * pow(sft, 3));
  }
  return 1.0 / pow(r, 3);
}

/* pg5_table.c:27-29 */
double or_s2_force(double eps_pp, double eps_pm, double r) {
  return or_s2_reff(eps_pp, r) - or_s2_reff(eps_pm, r);
}

typedef union {
  float f;
  uint32_t u;
} f32bits;

/* pg5_table.c:43-113 with force_func = S2 and rcut = eps_pm
 * (pg5_gen_s2_force_table, pg5_table.c:31-35).  Mixed float/double
 * expressions are written with the same C types as the reference so the
 * roundings are the same. */
void or_table_build(double eps_pp, double eps_pm, int exp_bit, int frc_bit,
                    float *tab, float *r2scale_out, double *xscale_out) {
  const int tbl = 1 << (exp_bit + frc_bit);
  const uint32_t tick = 1u << (23 - frc_bit);
  const double rcut = eps_pm;
  const float r2max = (float)(rcut * rcut);
  const float fmax = (float)((double)(1 << (1 << exp_bit)) * (2.0 - 1.0 / (1 << frc_bit)));
  const float r2scale = (fmax - 2.0f) / r2max;
  f32bits m;
  int i;
  m.f = 2.0f;
  for (i = 0; i < tbl; i++, m.u += tick) {
    float f = m.f;
    float r2 = (float)(((double)f - 2.0) / (double)r2scale);
    float r = sqrtf(r2);
    tab[2 * i] = (float)or_s2_force(eps_pp, eps_pm, (double)r);
  }
  m.f = 2.0f;
  for (i = 0; i < tbl - 1; i++) {
    float x0 = m.f;
    m.u += tick;
    float x1 = m.f;
    float y0 = tab[2 * i];
    float y1 = tab[2 * (i + 1)];
    tab[2 * i + 1] = (y1 - y0) / (x1 - x0);
  }
  tab[2 * (tbl - 1) + 1] = 0.0f;
  *r2scale_out = r2scale;
  *xscale_out = (double)sqrtf(r2scale);
}

/* gravity_kernel_avx2.c:69-107: s clamped to the last entry's bit pattern
 * (phantom_g5.c:35-40), index from the high bits, linear term from the
 * remaining low bits, value + frac * slope as one fused multiply-add. */
float or_table_lookup(const float *tab, int exp_bit, int frc_bit, float s) {
  const int shift = 23 - frc_bit;
  const int tbl = 1 << (exp_bit + frc_bit);
  f32bits cut, x, t;
  cut.u = (1u << 30) | ((uint32_t)(tbl - 1) << shift);
  x.f = s < cut.f ? s : cut.f;
  uint32_t hi = x.u >> shift;
  int idx = (int)hi - (1 << (30 - shift));
  t.u = hi << shift;
  float frac = x.f - t.f;
  return fmaf(frac, tab[2 * idx + 1], tab[2 * idx]);
}

void or_table_force_emul(int ni, const double *xi, int nj, const double *xj,
                         const double *mj, const float *tab, int exp_bit,
                         int frc_bit, double xscale, double *acc) {
  /* set_xi: (float)x * Xscale in double, stored as float
   * (PG5/table/libpg5/phantom_g5.c:102-104); set_xmj: float(x) * float(xscale)
   * in float, mass unscaled (:116-127); kernel arithmetic as in
   * gravity_kernel_avx2.c:56-118; final scale by float(1/xscale) (:129-131). */
  const float xs = (float)xscale;
  int i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < ni; i++) {
    float xis = (float)((double)(float)xi[3 * i] * xscale);
    float yis = (float)((double)(float)xi[3 * i + 1] * xscale);
    float zis = (float)((double)(float)xi[3 * i + 2] * xscale);
    double ax = 0, ay = 0, az = 0;
    for (int j = 0; j < nj; j++) {
      float xjs = (float)xj[3 * j] * xs;
      float yjs = (float)xj[3 * j + 1] * xs;
      float zjs = (float)xj[3 * j + 2] * xs;
      float m = (float)mj[j];
      float dx = xjs - xis, dy = yjs - yis, dz = zjs - zis;
      float s = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, 2.0f)));
      float f = or_table_lookup(tab, exp_bit, frc_bit, s);
      ax += (double)f * (double)(m * dx);
      ay += (double)f * (double)(m * dy);
      az += (double)f * (double)(m * dz);
    }
    double sc = (double)(float)(1.0 / xscale);
    acc[3 * i] = ax * sc;
    acc[3 * i + 1] = ay * sc;
    acc[3 * i + 2] = az * sc;
  }
}

void or_table_force_f64(int ni, const double *xi, int nj, const double *xj,
                        const double *mj, double eps_pp, double eps_pm,
                        double *acc) {
  int i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < ni; i++) {
    double ax = 0, ay = 0, az = 0;
    for (int j = 0; j < nj; j++) {
      double dx = xj[3 * j] - xi[3 * i];
      double dy = xj[3 * j + 1] - xi[3 * i + 1];
      double dz = xj[3 * j + 2] - xi[3 * i + 2];
      double r = sqrt(dx * dx + dy * dy + dz * dz);
      if (r >= eps_pm) continue;
      double f = or_s2_force(eps_pp, eps_pm, r);
      ax += mj[j] * f * dx;
      ay += mj[j] * f * dy;
      az += mj[j] * f * dz;
    }
    acc[3 * i] = ax;
    acc[3 * i + 1] = ay;
    acc[3 * i + 2] = az;
  }
}

/* ------------------------------------------------------------- quadrupole */
/* quad_avx2.c:35-149 algebra: d = x_cell - x_i, r2 = |d|^2 + eps2,
 * phi_p = M/r, qd = Q d, phi_q = (d.qd)/(2 r^5), pot += phi_p + phi_q,
 * a += (phi_p + 5 phi_q) d / r^2 - qd / r^5. */
void or_quad_force(int ni, const double *xi, double eps2, int nc,
                   const double *xc, const double *mc, const double *qc,
                   double *acc, double *pot) {
  int i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < ni; i++) {
    double ax = 0, ay = 0, az = 0, p = 0;
    for (int c = 0; c < nc; c++) {
      const double *q = qc + 6 * c;
      double dx = xc[3 * c] - xi[3 * i];
      double dy = xc[3 * c + 1] - xi[3 * i + 1];
      double dz = xc[3 * c + 2] - xi[3 * i + 2];
      double r2 = dx * dx + dy * dy + dz * dz + eps2;
      double rinv = 1.0 / sqrt(r2);
      double rinv2 = rinv * rinv;
      double rinv5 = rinv2 * rinv2 * rinv;
      double phip = mc[c] * rinv;
      double qx = q[0] * dx + q[1] * dy + q[2] * dz;
      double qy = q[1] * dx + q[3] * dy + q[4] * dz;
      double qz = q[2] * dx + q[4] * dy + q[5] * dz;
      double phiq = 0.5 * (dx * qx + dy * qy + dz * qz) * rinv5;
      double f = (phip + 5.0 * phiq) * rinv2;
      p += phip + phiq;
      ax += f * dx - rinv5 * qx;
      ay += f * dy - rinv5 * qy;
      az += f * dz - rinv5 * qz;
    }
    acc[3 * i] = ax;
    acc[3 * i + 1] = ay;
    acc[3 * i + 2] = az;
    pot[i] = p;
  }
}

/* source.cpp:215-248: Q_jk = sum m (3 d_j d_k - |d|^2 delta_jk) about the
 * centre of mass, stored {xx, xy, xz, yy, yz, zz} with zz = -(xx + yy). */
void or_quad_moments(int n, const double *x, const double *m, double *com,
                     double *mass, double *q6) {
  double M = 0, c[3] = {0, 0, 0};
  for (int i = 0; i < n; i++) {
    M += m[i];
    for (int k = 0; k < 3; k++) c[k] += m[i] * x[3 * i + k];
  }
  for (int k = 0; k < 3; k++) c[k] /= M;
  double q[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < n; i++) {
    double d[3] = {x[3 * i] - c[0], x[3 * i + 1] - c[1], x[3 * i + 2] - c[2]};
    double dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    int kk = 0;
    for (int a = 0; a < 2; a++)
      for (int b = a; b < 3; b++) {
        double v = 3.0 * d[a] * d[b];
        if (a == b) v -= dd;
        q[kk++] += m[i] * v;
      }
  }
  q[5] = -(q[0] + q[3]);
  memcpy(com, c, sizeof(c));
  *mass = M;
  memcpy(q6, q, sizeof(q));
}


/* ------------------------------------------------------------------ Hermite
 * Vectorised entry points over hermite_ref.h (PG6/banana/hermite.c:39-110,
 * timestep.c:24-79,176-190). */
double or_calc_modulus(double num, double mod, double modmax) {
  return hr_calc_modulus(num, mod, modmax);
}

void or_dt_startup(int n, double eta_s, double dtmax, double t_gl, const double *a,
                   const double *j, double *dt) {
  const double dtmin = dtmax / pow(2.0, 51.0);
  for (int i = 0; i < n; ++i) dt[i] = hr_dt_startup(eta_s, dtmax, dtmin, t_gl, a + 3 * i, j + 3 * i);
}

void or_hermite_predict(int n, double t_gl, const double *t, const double *x, const double *v,
                        const double *a, const double *j, double *xp, double *vp) {
  for (int i = 0; i < n; ++i)
    hr_predict(t_gl, t[i], x + 3 * i, v + 3 * i, a + 3 * i, j + 3 * i, xp + 3 * i, vp + 3 * i);
}

void or_hermite_correct(int n, double eta, double dtmax, const double *t, const double *dt,
                        const double *a0, const double *j0, const double *xp, const double *vp,
                        const double *a1, const double *j1, double *x1, double *v1,
                        double *dtn) {
  const double dtmin = dtmax / pow(2.0, 51.0);
  for (int i = 0; i < n; ++i)
    dtn[i] = hr_correct(eta, dtmax, dtmin, t[i], dt[i], a0 + 3 * i, j0 + 3 * i, xp + 3 * i,
                        vp + 3 * i, a1 + 3 * i, j1 + 3 * i, x1 + 3 * i, v1 + 3 * i);
}
