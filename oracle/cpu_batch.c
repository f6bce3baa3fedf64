/* CPU baseline drivers for the batched workloads — TEST INFRASTRUCTURE ONLY.
 *
 * They drive the reference libraries (compiled from their own sources into
 * oracle/_ref/ by oracle/ref.mk, loaded here with dlopen) the way the
 * reference's own tree code drives them, so bench.py can time the CPU side of
 * the TreePM, Barnes-Hut and many-cluster configurations:
 *
 *   cb_treepm   per i-group g5_set_xmjMC + g5_set_nMC + g5_calculate_force_on_xMC
 *               inside an OpenMP loop over groups with devid = thread number
 *               (the pattern of PG5/newton_quad/source.cpp:813-860)
 *   cb_bh       the same for the particle list, then g5c_set_xmjMC +
 *               g5c_set_nMC + g5c_calculate_force_on_xMC for the cell list
 *               (source.cpp:845-867)
 *   cb_clusters one GRAPE-6 session per cluster: g6_set_j_particle for every
 *               member, g6_set_ti, then g6calc_firsthalf + g6calc_lasthalf2
 *               over chunks of g6_npipes() i-particles (PG6/banana/hermite.c:
 *               221-277); the library parallelises each call internally.
 *   cb_hermite  the block-time-step Hermite integration of the reference's
 *               sample program PG6/banana (banana01.c:57-166) on the
 *               reference library, with the integrator restated in
 *               hermite_ref.h.
 *
 * Only bench.py's cpu_baseline leg and the tests load this library.
 */
#include <dlfcn.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hermite_ref.h"

typedef void (*f_open)(void);
typedef void (*f_eps)(double);
typedef void (*f_setn)(int, int);
typedef void (*f_xmj)(int, int, int, double (*)[3], double *);
typedef void (*f_force)(int, double (*)[3], double (*)[3], double *, int);
typedef void (*f_cxmj)(int, int, int, double (*)[3], double *, double (*)[6]);
typedef void (*f_s2)(double, double);

static void *need(void *h, const char *name) {
  void *p = dlsym(h, name);
  if (!p) fprintf(stderr, "cpu_batch: %s missing: %s\n", name, dlerror());
  return p;
}

static double now(void) { return omp_get_wtime(); }

/* TreePM: pos[n][3] are the physical positions (ghosts included), groups as
 * CSR lists.  The table is the reference's S2 table for (eps_pp, eps_pm).
 * acc[n][3] is written for every listed i.  Returns seconds, or -1. */
double cb_treepm(const char *lib, int ngroups, const int *ioff, const int *icnt,
                 const int *ilist, const int *joff, const int *jcnt, const int *jlist,
                 const double *pos, const double *mass, double eps_pp, double eps_pm,
                 int nthreads, double *acc) {
  void *h = dlopen(lib, RTLD_NOW | RTLD_LOCAL);
  if (!h) {
    fprintf(stderr, "cpu_batch: %s\n", dlerror());
    return -1.0;
  }
  f_open g5_open = (f_open)need(h, "g5_open");
  f_s2 gen = (f_s2)need(h, "pg5_gen_s2_force_table");
  f_setn set_n = (f_setn)need(h, "g5_set_nMC");
  f_xmj set_xmj = (f_xmj)need(h, "g5_set_xmjMC");
  f_force force = (f_force)need(h, "g5_calculate_force_on_xMC");
  if (!g5_open || !gen || !set_n || !set_xmj || !force) return -1.0;
  g5_open();
  gen(eps_pp, eps_pm);
  if (nthreads > 24) nthreads = 24; /* MAXDEV of the reference (phantom_g5.c:12) */
  int maxi = 0, maxj = 0;
  for (int g = 0; g < ngroups; ++g) {
    if (icnt[g] > maxi) maxi = icnt[g];
    if (jcnt[g] > maxj) maxj = jcnt[g];
  }
  const double t0 = now();
#pragma omp parallel num_threads(nthreads)
  {
    const int dev = omp_get_thread_num();
    double(*xi)[3] = malloc(sizeof(double[3]) * (size_t)(maxi + 4));
    double(*ai)[3] = malloc(sizeof(double[3]) * (size_t)(maxi + 4));
    double *pi = malloc(sizeof(double) * (size_t)(maxi + 4));
    double(*xj)[3] = malloc(sizeof(double[3]) * (size_t)(maxj + 8));
    double *mj = malloc(sizeof(double) * (size_t)(maxj + 8));
#pragma omp for schedule(dynamic, 4)
    for (int g = 0; g < ngroups; ++g) {
      const int ni = icnt[g], nj = jcnt[g];
      for (int k = 0; k < nj; ++k) {
        const int j = jlist[joff[g] + k];
        memcpy(xj[k], pos + 3 * (size_t)j, sizeof(double[3]));
        mj[k] = mass[j];
      }
      /* the AVX2 table kernel reads one j beyond nj (gravity_kernel_avx2.c:71-72) */
      memset(xj[nj], 0, sizeof(double[3]));
      mj[nj] = 0.0;
      for (int k = 0; k < ni; ++k)
        memcpy(xi[k], pos + 3 * (size_t)ilist[ioff[g] + k], sizeof(double[3]));
      set_xmj(dev, 0, nj + (nj & 1), xj, mj);
      set_n(dev, nj + (nj & 1));
      force(dev, xi, ai, pi, ni);
      for (int k = 0; k < ni; ++k)
        memcpy(acc + 3 * (size_t)ilist[ioff[g] + k], ai[k], sizeof(double[3]));
    }
    free(xi);
    free(ai);
    free(pi);
    free(xj);
    free(mj);
  }
  return now() - t0;
}

/* Barnes-Hut: particle lists (jlist into pos/mass) and cell lists (clist into
 * cells[nc][10] = {x, y, z, M, Qxx, Qxy, Qxz, Qyy, Qyz, Qzz}).  acc and pot
 * get the library's outputs, particle part plus cell part, with the library's
 * sign convention.  Returns seconds, or -1. */
double cb_bh(const char *lib, int ngroups, const int *ioff, const int *icnt, const int *ilist,
             const int *joff, const int *jcnt, const int *jlist, const int *coff,
             const int *ccnt, const int *clist, const double *pos, const double *mass,
             const double *cells, double eps, int nthreads, double *acc, double *pot) {
  void *h = dlopen(lib, RTLD_NOW | RTLD_LOCAL);
  if (!h) {
    fprintf(stderr, "cpu_batch: %s\n", dlerror());
    return -1.0;
  }
  f_open g5_open = (f_open)need(h, "g5_open");
  f_eps set_eps = (f_eps)need(h, "g5_set_eps_to_all");
  f_setn set_n = (f_setn)need(h, "g5_set_nMC");
  f_xmj set_xmj = (f_xmj)need(h, "g5_set_xmjMC");
  f_force force = (f_force)need(h, "g5_calculate_force_on_xMC");
  f_setn cset_n = (f_setn)need(h, "g5c_set_nMC");
  f_cxmj cset_xmj = (f_cxmj)need(h, "g5c_set_xmjMC");
  f_force cforce = (f_force)need(h, "g5c_calculate_force_on_xMC");
  if (!g5_open || !set_eps || !set_n || !set_xmj || !force || !cset_n || !cset_xmj || !cforce)
    return -1.0;
  g5_open();
  set_eps(eps);
  if (nthreads > 24) nthreads = 24;
  int maxi = 0, maxj = 0, maxc = 0;
  for (int g = 0; g < ngroups; ++g) {
    if (icnt[g] > maxi) maxi = icnt[g];
    if (jcnt[g] > maxj) maxj = jcnt[g];
    if (ccnt[g] > maxc) maxc = ccnt[g];
  }
  const double t0 = now();
#pragma omp parallel num_threads(nthreads)
  {
    const int dev = omp_get_thread_num();
    double(*xi)[3] = malloc(sizeof(double[3]) * (size_t)(maxi + 4));
    double(*ai)[3] = malloc(sizeof(double[3]) * (size_t)(maxi + 4));
    double *pi = malloc(sizeof(double) * (size_t)(maxi + 4));
    double(*ci)[3] = malloc(sizeof(double[3]) * (size_t)(maxi + 4));
    double *qi = malloc(sizeof(double) * (size_t)(maxi + 4));
    double(*xj)[3] = malloc(sizeof(double[3]) * (size_t)(maxj + 8));
    double *mj = malloc(sizeof(double) * (size_t)(maxj + 8));
    double(*xc)[3] = malloc(sizeof(double[3]) * (size_t)(maxc + 2));
    double *mc = malloc(sizeof(double) * (size_t)(maxc + 2));
    double(*qc)[6] = malloc(sizeof(double[6]) * (size_t)(maxc + 2));
#pragma omp for schedule(dynamic, 4)
    for (int g = 0; g < ngroups; ++g) {
      const int ni = icnt[g], nj = jcnt[g], nc = ccnt[g];
      for (int k = 0; k < ni; ++k)
        memcpy(xi[k], pos + 3 * (size_t)ilist[ioff[g] + k], sizeof(double[3]));
      for (int k = 0; k < nj; ++k) {
        const int j = jlist[joff[g] + k];
        memcpy(xj[k], pos + 3 * (size_t)j, sizeof(double[3]));
        mj[k] = mass[j];
      }
      set_xmj(dev, 0, nj, xj, mj);
      set_n(dev, nj);
      force(dev, xi, ai, pi, ni);
      /* cells are stored in pairs and xj[nc] is read for odd nc (quad_g5c.c:198-283) */
      for (int k = 0; k < nc; ++k) {
        const double *c = cells + 10 * (size_t)clist[coff[g] + k];
        memcpy(xc[k], c, sizeof(double[3]));
        mc[k] = c[3];
        memcpy(qc[k], c + 4, sizeof(double[6]));
      }
      memset(xc[nc], 0, sizeof(double[3]));
      mc[nc] = 0.0;
      memset(qc[nc], 0, sizeof(double[6]));
      for (int k = 0; k < ni; ++k) {
        ci[k][0] = ci[k][1] = ci[k][2] = 0.0;
        qi[k] = 0.0;
      }
      if (nc > 0) {
        cset_xmj(dev, 0, nc + (nc & 1), xc, mc, qc);
        cset_n(dev, nc + (nc & 1));
        cforce(dev, xi, ci, qi, ni);
      }
      for (int k = 0; k < ni; ++k) {
        const size_t i = (size_t)ilist[ioff[g] + k];
        for (int d = 0; d < 3; ++d) acc[3 * i + d] = ai[k][d] + ci[k][d];
        pot[i] = pi[k] + qi[k];
      }
    }
    free(xi);
    free(ai);
    free(pi);
    free(ci);
    free(qi);
    free(xj);
    free(mj);
    free(xc);
    free(mc);
    free(qc);
  }
  return now() - t0;
}

typedef void (*f_g6open)(int);
typedef void (*f_g6setj)(int, int, int, double, double, double, double *, double *, double *,
                         double *, double *);
typedef void (*f_g6ti)(int, double);
typedef void (*f_g6first)(int, int, int, int *, double (*)[3], double (*)[3], double (*)[3],
                          double (*)[3], double *, double, double *);
typedef int (*f_g6last2)(int, int, int, int *, double (*)[3], double (*)[3], double, double *,
                         double (*)[3], double (*)[3], double *, int *);

/* Many clusters: clusters `sel[0..nsel)` of the CSR set coff; members have
 * pos, vel (FP64) and mass at t = 0 with zero acc/jerk (prediction is the
 * identity), eps2[c] per cluster.  acc, jrk, pot are written for the members
 * of the selected clusters.  Returns seconds, or -1. */
double cb_clusters(const char *lib, int nsel, const int *sel, const int *coff,
                   const double *pos, const double *vel, const double *mass,
                   const double *eps2, double *acc, double *jrk, double *pot) {
  void *h = dlopen(lib, RTLD_NOW | RTLD_LOCAL);
  if (!h) {
    fprintf(stderr, "cpu_batch: %s\n", dlerror());
    return -1.0;
  }
  f_g6open g6_open = (f_g6open)need(h, "g6_open");
  f_g6setj setj = (f_g6setj)need(h, "g6_set_j_particle");
  f_g6ti set_ti = (f_g6ti)need(h, "g6_set_ti");
  f_g6first first = (f_g6first)need(h, "g6calc_firsthalf");
  f_g6last2 last2 = (f_g6last2)need(h, "g6calc_lasthalf2");
  if (!g6_open || !setj || !set_ti || !first || !last2) return -1.0;
  g6_open(0);
  const int npipes = 48; /* g6_npipes() of the reference (avx_type.h:1) */
  double z[3] = {0.0, 0.0, 0.0};
  double h2[48];
  int nnb[48];
  double pbuf[48];
  memset(h2, 0, sizeof h2);
  const double t0 = now();
  for (int s = 0; s < nsel; ++s) {
    const int c = sel[s];
    const int b = coff[c], n = coff[c + 1] - coff[c];
    for (int k = 0; k < n; ++k) {
      const size_t i = (size_t)(b + k);
      setj(0, k, k, 0.0, 0.0, mass[i], z, z, z, (double *)(vel + 3 * i), (double *)(pos + 3 * i));
    }
    set_ti(0, 0.0);
    for (int o = 0; o < n; o += npipes) {
      const int ni = n - o < npipes ? n - o : npipes;
      int idx[48];
      for (int k = 0; k < ni; ++k) idx[k] = o + k;
      double(*xi)[3] = (double(*)[3])(pos + 3 * (size_t)(b + o));
      double(*vi)[3] = (double(*)[3])(vel + 3 * (size_t)(b + o));
      double(*ai)[3] = (double(*)[3])(acc + 3 * (size_t)(b + o));
      double(*ji)[3] = (double(*)[3])(jrk + 3 * (size_t)(b + o));
      first(0, n, ni, idx, xi, vi, ai, ji, pbuf, eps2[c], h2);
      last2(0, n, ni, idx, xi, vi, eps2[c], h2, ai, ji, pot + b + o, nnb);
    }
  }
  return now() - t0;
}

typedef void (*f_g6last)(int, int, int, int *, double (*)[3], double (*)[3], double, double *,
                         double (*)[3], double (*)[3], double *);

/* Integrates n particles (x, v, m at t = 0) to t_end as banana01.c does:
 * set every particle on the board, start-up forces and steps, then block
 * steps (global_time, predict_exmotion, firsthalf + lasthalf over chunks of
 * 48, correct + dt_criterion, set_j_particle) with the time reset at every
 * dtmax.  x, v are overwritten with the final state; info[3] = {particle
 * steps, block steps, windows}; energy[2 * w] = {ek, ep} at the end of
 * window w (up to maxw).  Returns seconds, or -1. */
double cb_hermite(const char *lib, int n, double *x, double *v, const double *m, double eps2,
                  double eta_s, double eta, double dtmax, double t_end, long long *info,
                  double *energy, int maxw) {
  void *h = dlopen(lib, RTLD_NOW | RTLD_LOCAL);
  if (!h) {
    fprintf(stderr, "cpu_batch: %s\n", dlerror());
    return -1.0;
  }
  f_g6open g6_open = (f_g6open)need(h, "g6_open");
  f_g6setj setj = (f_g6setj)need(h, "g6_set_j_particle");
  f_g6ti set_ti = (f_g6ti)need(h, "g6_set_ti");
  f_g6first first = (f_g6first)need(h, "g6calc_firsthalf");
  f_g6last last = (f_g6last)need(h, "g6calc_lasthalf");
  if (!g6_open || !setj || !set_ti || !first || !last) return -1.0;
  g6_open(0);
  const int np = 48;
  const double dtmin = dtmax / pow(2.0, 51.0);
  double *a = calloc(3 * (size_t)n, sizeof(double)), *j = calloc(3 * (size_t)n, sizeof(double));
  double *t = calloc((size_t)n, sizeof(double)), *dt = calloc((size_t)n, sizeof(double));
  double *pot = calloc((size_t)n, sizeof(double));
  double *xp = malloc(sizeof(double) * 3 * (size_t)n), *vp = malloc(sizeof(double) * 3 * (size_t)n);
  double *a1 = malloc(sizeof(double) * 3 * (size_t)n), *j1 = malloc(sizeof(double) * 3 * (size_t)n);
  double *p1 = malloc(sizeof(double) * (size_t)n);
  int *act = malloc(sizeof(int) * (size_t)n);
  double z[3] = {0, 0, 0}, h2[48];
  memset(h2, 0, sizeof h2);
  const double over2 = 0.5, over6 = 1.0 / 6.0;
#define SETJ(i)                                                                        \
  do {                                                                                 \
    double aj[3], jj[3];                                                               \
    for (int d = 0; d < 3; ++d) {                                                      \
      aj[d] = a[3 * (i) + d] * over2;                                                  \
      jj[d] = j[3 * (i) + d] * over6;                                                  \
    }                                                                                  \
    setj(0, (i), (i), t[i], dt[i], m[i], z, jj, aj, v + 3 * (size_t)(i), x + 3 * (size_t)(i)); \
  } while (0)
  const double t0 = now();
  for (int i = 0; i < n; ++i) {
    dt[i] = dtmax;
    SETJ(i);
  }
  double t_gl = 0.0;
  long long t_mod = 0, nsteps = 0, nblocks = 0;
  int nact = n;
  for (int i = 0; i < n; ++i) act[i] = i;
  int startup = 1;
  int w = 0;
  for (;;) {
    set_ti(0, t_gl);
    for (int k = 0; k < nact; ++k) {
      const int i = act[k];
      hr_predict(t_gl, t[i], x + 3 * i, v + 3 * i, a + 3 * i, j + 3 * i, xp + 3 * k, vp + 3 * k);
    }
    for (int o = 0; o < nact; o += np) {
      const int ni = nact - o < np ? nact - o : np;
      first(0, n, ni, act + o, (double(*)[3])(xp + 3 * o), (double(*)[3])(vp + 3 * o),
            (double(*)[3])(a1 + 3 * o), (double(*)[3])(j1 + 3 * o), p1 + o, eps2, h2);
      last(0, n, ni, act + o, (double(*)[3])(xp + 3 * o), (double(*)[3])(vp + 3 * o), eps2, h2,
           (double(*)[3])(a1 + 3 * o), (double(*)[3])(j1 + 3 * o), p1 + o);
    }
    for (int k = 0; k < nact; ++k) {
      const int i = act[k];
      if (startup) {
        memcpy(a + 3 * i, a1 + 3 * k, sizeof(double) * 3);
        memcpy(j + 3 * i, j1 + 3 * k, sizeof(double) * 3);
        dt[i] = hr_dt_startup(eta_s, dtmax, dtmin, t[i], a + 3 * i, j + 3 * i);
      } else {
        double x1[3], v1[3];
        const double dtn = hr_correct(eta, dtmax, dtmin, t[i], dt[i], a + 3 * i, j + 3 * i,
                                      xp + 3 * k, vp + 3 * k, a1 + 3 * k, j1 + 3 * k, x1, v1);
        memcpy(x + 3 * i, x1, sizeof x1);
        memcpy(v + 3 * i, v1, sizeof v1);
        memcpy(a + 3 * i, a1 + 3 * k, sizeof(double) * 3);
        memcpy(j + 3 * i, j1 + 3 * k, sizeof(double) * 3);
        t[i] = t[i] + dt[i];
        dt[i] = dtn;
      }
      pot[i] = p1[k];
      SETJ(i);
    }
    nsteps += nact;
    nblocks += 1;
    startup = 0;
    if (t_gl == dtmax) {  /* end of a window: energy, reset_time, set_particle_on_grape */
      if (w < maxw && energy) {
        double ek = 0.0, ew = 0.0;
        for (int i = 0; i < n; ++i) {
          double v2 = v[3 * i] * v[3 * i];
          v2 += v[3 * i + 1] * v[3 * i + 1];
          v2 += v[3 * i + 2] * v[3 * i + 2];
          ek += m[i] * v2;
          ew += m[i] * pot[i];
        }
        energy[2 * w] = 0.5 * ek;
        energy[2 * w + 1] = 0.5 * ew;
      }
      ++w;
      t_gl = 0.0;
      ++t_mod;
      for (int i = 0; i < n; ++i) {
        t[i] = 0.0;
        SETJ(i);
      }
    }
    if ((double)t_mod * dtmax + t_gl >= t_end) break;
    /* global_time: the particles with the smallest t + dt, in index order */
    double tn = 1e300;
    for (int i = 0; i < n; ++i) tn = fmin(tn, t[i] + dt[i]);
    nact = 0;
    for (int i = 0; i < n; ++i)
      if (t[i] + dt[i] == tn) act[nact++] = i;
    t_gl = tn;
  }
#undef SETJ
  const double sec = now() - t0;
  if (info) {
    info[0] = nsteps;
    info[1] = nblocks;
    info[2] = w;
  }
  free(a); free(j); free(t); free(dt); free(pot); free(xp); free(vp); free(a1); free(j1);
  free(p1); free(act);
  return sec;
}
