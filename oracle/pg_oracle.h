/* CPU oracle for the MI355X Phantom-GRAPE library — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C FP64 restatement of the reference algorithms on the hot path.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker.  The product libraries never link it.
 *
 * Pinning (what checks each function against something other than itself):
 *   or_pg5_force     - compiled reference libpg5 (oracle/_ref), two-body
 *                      closed form, momentum conservation, a = -grad(phi)
 *                      by finite differences (tests/test_oracle.py)
 *   or_pg6_predict   - closed-form polynomial; compiled reference predictor
 *                      through g6calc on a one-body j-set
 *   or_pg6_force     - compiled reference libpg6, two-body closed form,
 *                      jerk = d(acc)/dt by finite differences
 *   or_pg6_nblist    - brute force (the definition itself) and the
 *                      compiled reference g6calc_lasthalfn lists
 *   or_s2_reff/force - piecewise polynomial continuity at eta = 1, 2 and
 *                      the 1/r^3 tail; compiled reference table values
 *   or_table_build   - bit-exact against the compiled reference
 *                      Force_table (extern global of libref_pg5table.so)
 *   or_table_lookup  - bit-exact against the compiled reference through
 *                      single-pair force evaluations
 *   or_quad_force    - FP64 direct sum over the particles of a cell at large
 *                      distance (multipole residual O(theta^3)), traceless
 *                      Q, compiled reference libpg5quad
 *   or_quad_moments  - definition (source.cpp:215-248), tracelessness
 */
#ifndef PG_ORACLE_H
#define PG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* GRAPE-5 softened force, FP64 (PG5/newton/libpg5/gravity_avx2.c:28-173,
 * symmetric :178-250).  eps2j may be NULL.  pot = + sum m / r including j=i.
 * Positions are arrays of [n][3]. */
void or_pg5_force(int ni, const double *xi, const double *eps2i, int nj,
                  const double *xj, const double *mj, const double *eps2j,
                  double *acc, double *pot);

/* Hermite predictor (PG6/libpg6/gravity.c:19-20,149-193) evaluated with the
 * same left-to-right operation order; vel stays FP64 here. */
void or_pg6_predict(int nj, double ti, const double *x, const double *v,
                    const double *a2, const double *j6, const double *tj,
                    double *xp, double *vp);

/* GRAPE-6 force/jerk/pot/nearest neighbour, FP64 (scalar recipe
 * PG6/libpg6/gravity.c:195-258; NN gravity.c:554-564,652-673).
 * Self excluded by id.  pot = - sum m / r.  nnb = id of nearest j
 * (ties -> lowest j address), rnnb = sqrt(r^2 + eps^2) to it; -1/0 if none. */
void or_pg6_force(int ni, const double *xi, const double *vi, const int *idi,
                  double eps2, int nj, const double *xj, const double *vj,
                  const double *mj, const int *idj, double *acc, double *jrk,
                  double *pot, int *nnb, double *rnnb);

/* Neighbour lists (gravity.c:788-793,892-909): ids of j with
 * |dx|^2 + eps^2 <= h2_i and id != id_i, in j order, at most maxlen each;
 * nblen gets the full count. */
void or_pg6_nblist(int ni, const double *xi, const int *idi, const double *h2,
                   double eps2, int nj, const double *xj, const int *idj,
                   int maxlen, int *nbl, int *nblen);

/* S2 (Hockney-Eastwood) force shape, r^-3 units (PG5/table/libpg5/
 * pg5_table.c:8-29). */
/* rsqrt bias of PG5/newton/libpg5/rsqrt.c:42-81 for a given reciprocal
 * square root: y0[i], y1[i] = rsqrt of the floats with bits (127 << 23) + i
 * and (128 << 23) + i, i < 2^23 (x in [1, 2) and [2, 4)).  b = sum w (e0 + e1)
 * / (2 * 2 sum w) with e = y^2 x - 1, w = y0^2.  Pinned: with the SSE
 * rsqrtss values (or_rsqrtss_table, the reference's rsqrt(), rsqrt.c:12-22)
 * it equals the compiled reference's rsqrt_bias() bit for bit; with the
 * correctly rounded FP32 1/sqrt(x) it is ~0. */
double or_rsqrt_bias_of(const float *y0, const float *y1);
void or_rsqrtss_table(float *y0, float *y1);
double or_s2_reff(double sft, double r);
double or_s2_force(double eps_pp, double eps_pm, double r);

/* Table generation, bit-exact with pg5_gen_force_table for the S2 shape
 * (pg5_table.c:43-113): tab[2*k] = value, tab[2*k+1] = slope,
 * k < 2^(exp_bit+frc_bit).  Returns r2scale (float) and xscale. */
void or_table_build(double eps_pp, double eps_pm, int exp_bit, int frc_bit,
                    float *tab, float *r2scale, double *xscale);

/* Table lookup on s = 2 + r^2 r2scale, bit-exact with the AVX2 kernel
 * (gravity_kernel_avx2.c:69-107, phantom_g5.c:35-40). */
float or_table_lookup(const float *tab, int exp_bit, int frc_bit, float s);

/* Table-path force, per-pair arithmetic of the AVX2 kernel in FP32 (scaled
 * positions, FMA order of gravity_kernel_avx2.c:56-118), summed in FP64 and
 * scaled by 1/xscale.  acc in [ni][3]. */
void or_table_force_emul(int ni, const double *xi, int nj, const double *xj,
                         const double *mj, const float *tab, int exp_bit,
                         int frc_bit, double xscale, double *acc);

/* Closed-form S2 short-range force in FP64: sum m f(r) dx, f = 0 for
 * r >= eps_pm (the cut-off). */
void or_table_force_f64(int ni, const double *xi, int nj, const double *xj,
                        const double *mj, double eps_pp, double eps_pm,
                        double *acc);

/* Monopole + quadrupole cell force, FP64 (quad_avx2.c:35-149 algebra,
 * source.cpp:1163-1184 scalar form).  q = [nc][6] {xx,xy,xz,yy,yz,zz}.
 * pot = + sum (M/r + dx.Q.dx/(2 r^5)). */
void or_quad_force(int ni, const double *xi, double eps2, int nc,
                   const double *xc, const double *mc, const double *qc,
                   double *acc, double *pot);

/* Mass, centre of mass and traceless quadrupole of a particle set
 * (source.cpp:215-248). */
void or_quad_moments(int n, const double *x, const double *m, double *com,
                     double *mass, double *q6);

/* Block-time-step Hermite pieces of the reference's sample integrator
 * (PG6/banana/hermite.c:39-110, timestep.c:24-79,176-190), restated in
 * hermite_ref.h; arrays [n][3] or [n]. */
double or_calc_modulus(double num, double mod, double modmax);
void or_dt_startup(int n, double eta_s, double dtmax, double t_gl, const double *a,
                   const double *j, double *dt);
void or_hermite_predict(int n, double t_gl, const double *t, const double *x, const double *v,
                        const double *a, const double *j, double *xp, double *vp);
void or_hermite_correct(int n, double eta, double dtmax, const double *t, const double *dt,
                        const double *a0, const double *j0, const double *xp, const double *vp,
                        const double *a1, const double *j1, double *x1, double *v1,
                        double *dtn);

#ifdef __cplusplus
}
#endif
#endif
