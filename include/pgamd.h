/* Entry points of the MI355X libraries that go beyond the reference API:
 * device-pointer variants of the force calls (inputs already in HBM, results
 * left in HBM, asynchronous on a caller-chosen HIP stream), launch-shape
 * control for the autotuner, and error reporting.  None of these has a
 * counterpart in the reference (a CPU library whose data always lives in
 * host memory); each names the reference call whose work it performs.
 * Stream arguments are hipStream_t passed as void*; nullptr = the board's
 * own stream.  All pointers named d_* or documented "device" are device
 * pointers.  Returns: 0 on success, -1 after an API error.
 */
#ifndef PGAMD_H
#define PGAMD_H

#if defined(__cplusplus)
extern "C" {
#endif

/* ------------------------------------------------------------------ every library
 * pgamd-library: pg5 pg5_sym pg5_table pg6 */
/* 1 (default): API misuse prints a message and aborts, like the reference's
 * assert()s; 0: the call returns without effect and records the error. */
void pgamd_set_abort_on_error(int flag);
/* Returns and clears the last recorded error code (0 = none); copies its message. */
int pgamd_last_error(char *buf, int buflen);

/* ------------------------------------------------------------------ libpg5.so
 * pgamd-library: pg5 pg5_sym */
/* Work of g5_calculate_force_on_xMC (PG5/newton/libpg5/phantom_g5.c:276):
 * xi[ni] float4 {x,y,z,eps_i^2}, xj[nj] float4 {x,y,z,m}, epsj2[nj] (symmetric
 * build only), acc[ni][3] and pot[ni] FP64 with pot = + sum m/r. */
int pg5_dev_force(int devid, int ni, const void *xi, int nj, const void *xj,
                  const float *epsj2, double *acc, double *pot, void *stream);
/* Work of g5c_calculate_force_on_xMC (PG5/newton_quad/quad/quad_g5c.c:285):
 * cells[nc] = 12 floats {x,y,z,M, Qxx,Qxy,Qxz,Qyy, Qyz,Qzz,0,0};
 * pot = + sum (M/r + d.Q.d/(2 r^5)). */
int pg5_dev_quad(int devid, int ni, const void *xi, int nc, const void *cells,
                 double *acc, double *pot, void *stream);
void *pg5_dev_stream(int devid);
/* Allocation-free form of pg5_dev_force, for callers that cannot allocate
 * device memory (HIP graph capture, serving loops): no counterpart in the
 * reference.  The j-range partial sums live in a static device array of
 * 16 MiB per GPU (the split is reduced until they fit: ni * nsplit * 32 bytes
 * <= 16 MiB); no board, stream or host buffer is created, `stream` (null =
 * the default stream) carries the work.  Every call on one GPU shares that
 * array: calls that overlap in time on one GPU race.  Results equal
 * pg5_dev_force's whenever its split fits. */
int pg5_dev_force_noalloc(int ni, const void *xi, int nj, const void *xj, const float *epsj2,
                          double *acc, double *pot, void *stream);
/* rsqrt-bias correction of the g5_* outputs, the reference's rule
 * (PG5/newton/libpg5/rsqrt.c:42-81 measures the bias b of the rsqrt the
 * kernel uses; phantom_g5.c:45-60,244-269 scales acc by float(1 - 3b) and pot
 * by float(1 - b), the cell calls included): mode 0 (default) none, 1 with b
 * measured by that rule for the device's v_rsq_f32 (once per process). */
int pg5_set_rsqrt_bias_mode(int mode);
/* b of the device's rsqrt by that rule (measured on the first call). */
double pg5_rsqrt_bias(void);
/* bi in {0 = automatic, 64, 128, 256} i-particles per block; blocks_per_cu and
 * min_jper (> 0) set the j-range split (defaults: particles 64 i per block up
 * to 32768 i and 256 beyond, 32 blocks per CU, 256 j per range at least;
 * cells 64 / 128 / 256 i by ni, 8 blocks per CU, 128 j). */
int pg5_dev_set_launch(int bi, int blocks_per_cu, int min_jper);
/* Particle force kernel: 0 = j-tiles staged in LDS (default), 1 = j
 * broadcast through scalar registers.  Same pair arithmetic; speed only. */
int pg5_dev_set_kernel(int variant);
/* Batched tree-list force (the work of the per-group g5_calculate_force_on_xMC
 * + g5c_calculate_force_on_xMC pair of PG5/newton_quad/source.cpp:845-867 for
 * every group) in one persistent launch.  Device arrays: pos[np] float4
 * {x,y,z,m}; cells[nc] (12 floats as pg5_dev_quad); per group g the i-list
 * ilist[ioff[g]..+icnt[g]), particle list plist[poff[g]..+pcnt[g]), cell list
 * clist[coff[g]..+ccnt[g]); order (nullable) = processing order.  Writes
 * acc[i][3] and pot[i] = + sum m/r + sum (phi_p + phi_q) for the listed i. */
int pg5_bh_batch(int devid, int ngroups, const int *ioff, const int *icnt, const int *ilist,
                 const int *poff, const int *pcnt, const int *plist, const int *coff,
                 const int *ccnt, const int *clist, const int *order, const void *pos,
                 const void *cells, double eps2, double *acc, double *pot, void *stream);
/* The same launch captured in a HIP graph; replay / destroy. */
void *pg5_bh_graph(int devid, int ngroups, const int *ioff, const int *icnt, const int *ilist,
                   const int *poff, const int *pcnt, const int *plist, const int *coff,
                   const int *ccnt, const int *clist, const int *order, const void *pos,
                   const void *cells, double eps2, double *acc, double *pot);
int pg5_bh_graph_launch(void *graph, void *stream);
int pg5_bh_graph_destroy(void *graph);
/* Autotuner over blocks of {256, 512, 1024} threads (an LDS tile holds that
 * many particles or half as many cells) x persistent blocks per CU
 * {1,2,3,4,6,8,12,16}; keeps the fastest pair.  *best = its blocks per CU,
 * ms[8] (nullable) = the timings at the block kept. */
int pg5_bh_autotune(int devid, int ngroups, const int *ioff, const int *icnt, const int *ilist,
                    const int *poff, const int *pcnt, const int *plist, const int *coff,
                    const int *ccnt, const int *clist, const int *order, const void *pos,
                    const void *cells, double eps2, double *acc, double *pot, int reps,
                    int *best, float *ms);
/* Launch shape of the tree-list batch: block in {0 = default 1024, 256, 512,
 * 1024}, blocks_per_cu (0 = default 2); a negative value keeps that part.
 * pg5_bh_launch_shape: out2 = {block, blocks_per_cu} in use. */
int pg5_bh_set_launch(int block, int blocks_per_cu);
int pg5_bh_launch_shape(int *out2);

/* ------------------------------------------------------------------ libpg5_table.so
 * pgamd-library: pg5_table */
/* Table resolution 2^frc_bit entries per octave of s (reference FRC_BIT = 5,
 * compile-time, PG5/table/libpg5/pg5_table.h:7-15); rebuilds the table. */
int pg5_set_table_frc_bit(int frc_bit);
/* Interpolation order of the table in the batched calls (pg5t_batch_*):
 * 1 = the reference's value + slope per float-bit tick (default), 2 = a
 * quadratic through each interval's ends and midpoint from the shape in FP64
 * (16 bytes per entry; frc_bit <= 8).  The drop-in calls keep the
 * reference's table. */
int pg5_set_table_order(int order);
/* Copies the current table (2*ntab floats) if out != NULL; returns ntab. */
int pg5_get_table(float *out, double *xscale, float *r2scale);
/* Work of g5_calculate_force_on_xMC (PG5/table/libpg5/phantom_g5.c:160) on
 * scaled device positions: xi[ni] float4 {x*xs,y*xs,z*xs,0}, xj[nj] float4
 * {x*xs,y*xs,z*xs,m}; acc[ni][3] FP64 (unscaled). */
int pg5t_dev_force(int devid, int ni, const void *xi, int nj, const void *xj,
                   double *acc, void *stream);
/* Table value f(s) at n device values s (gravity_kernel_avx2.c:69-107). */
int pg5t_dev_eval(int n, const float *s, float *out);
/* As pg5_dev_set_launch (defaults: 64 i per block up to 32768 i and 256
 * beyond, 16 blocks per CU up to 16384 i and 32 beyond, 256 j per range at
 * least). */
int pg5t_dev_set_launch(int bi, int blocks_per_cu, int min_jper);
void *pg5t_dev_stream(int devid);
/* Learned cut-off: a 1-16-16-1 tanh FP32 network, trained on the GPU on
 * samples of the current cut-off shape, replacing the table lookup of
 * gravity_kernel_avx2.c:74-100.  For the S2 shape of
 * pg5_gen_s2_force_table it learns H(u) = R(u, 1) of pg5_table.c:8-22 on
 * u in [0, 1] (form 1: rc = eps_PM, r0 = eps_PP); for other shapes
 * Y(v) = f(r) (r^3 + r0^3), v = sqrt(r / rc) (form 0).  Parameters p321 in
 * the tanh form, FP64 (layout in mlp_kernels.h).
 * mode 0 = table (default), 1 = network, for the drop-in calls. */
int pg5_set_cutoff_mode(int mode);
int pg5_get_cutoff_mode(void);
int pg5_mlp_set_params(const double *p321, int form, double rc, double r0);
/* Returns 1 and copies the loaded network (each output nullable), 0 if none. */
int pg5_mlp_get_params(double *p321, int *form, double *rc, double *r0);
/* Levenberg-Marquardt on the GPU: `starts` seeded starts of `iters` trial
 * steps each (FP64 Jacobians and normal equations from the training kernels,
 * the damped 321 x 321 solve on the host), then `refine` more trial steps for
 * the best of them, which is kept.  *loss = weighted mean squared error of
 * the kept network. */
int pg5_mlp_train(int iters, int starts, int refine, unsigned seed, double *loss);
/* Normal equations of the weighted fit at p321 to n host samples, from the
 * FP64 training kernels: jtj[321*321] (nullable), jtr[321], *rtr, with
 * residuals sqrt(wt) (y(v) - t).  Synchronous; for checking the kernels. */
int pg5_mlp_normal_eq(int n, const double *v, const double *t, const double *wt,
                      const double *p321, double *jtj, double *jtr, double *rtr);
/* Learned f(r) at n physical radii (device arrays), as the force kernels decode it. */
int pg5_mlp_eval_dev(int n, const float *r, float *f);
/* Batched TreePM short-range force of every group in one persistent launch
 * (the per-group g5_set_xmjMC + g5_calculate_force_on_xMC loop of a TreePM
 * caller).  Device arrays: pos[np] float4 {x*xs,y*xs,z*xs,m}; pos64[np][3]
 * (nullable) the unscaled FP64 positions, from which separations are formed
 * relative to each group's first i-particle (the reference's ingest
 * conversions applied to x - o, so close pairs keep their digits); per
 * group g ilist[ioff[g]..+icnt[g]), jlist[joff[g]..+jcnt[g]); order
 * nullable; acc[i][3] FP64 written for the listed i; mode 0 table,
 * 1 network. */
int pg5t_batch_force(int devid, int ngroups, const int *ioff, const int *icnt, const int *ilist,
                     const int *joff, const int *jcnt, const int *jlist, const int *order,
                     const void *pos, const double *pos64, double *acc, int mode, void *stream);
void *pg5t_batch_graph(int devid, int ngroups, const int *ioff, const int *icnt,
                       const int *ilist, const int *joff, const int *jcnt, const int *jlist,
                       const int *order, const void *pos, const double *pos64, double *acc,
                       int mode);
int pg5t_batch_graph_launch(void *graph, void *stream);
int pg5t_batch_graph_destroy(void *graph);
/* Autotuner over the table kernel's block of {256, 512, 1024} threads (each
 * also its LDS j-tile; the network kernel runs 256) x persistent blocks per
 * CU {1,2,3,4,6,8,12,16}; keeps the fastest pair for this mode.  *best = its
 * blocks per CU, ms[8] (nullable) = the timings at the block kept. */
int pg5t_batch_autotune(int devid, int ngroups, const int *ioff, const int *icnt,
                        const int *ilist, const int *joff, const int *jcnt, const int *jlist,
                        const int *order, const void *pos, const double *pos64, double *acc,
                        int mode, int reps, int *best, float *ms);
int pg5t_batch_set_blocks_per_cu(int mode, int bpc);
/* Table-mode block in {0 = default 1024, 256, 512, 1024}. */
int pg5t_batch_set_block(int block);
/* out2 = {threads per block, blocks per CU} the batched calls of this mode
 * use now. */
int pg5t_batch_launch_shape(int mode, int *out2);
/* Long-range (particle-mesh) side of the TreePM force, which the reference
 * does not ship (PG5/table/libpg5/pppm.h:4-6 names its mesh): the periodic
 * force of the S2 shape R(r, eps_pm) that the table subtracts, by TSC
 * assignment on a mesh^3 grid, FFT, Green's function -4 pi S(k a)^2 /
 * (k^2 W(k)^2), spectral gradient and TSC interpolation.  pos[n][3] in the
 * unit box [0, 1), mass[n], acc[n][3]: device FP64; acc is written.  Table
 * force + this = the S2-softened periodic Newtonian force. */
int pg5t_pm_force(int n, const double *pos, const double *mass, int mesh, double eps_pm,
                  double *acc, void *stream);

/* ------------------------------------------------------------------ libpg6.so
 * pgamd-library: pg6 */
/* Bulk form of nj g6_set_j_particle calls (PG6/libpg6/phantom_g6.c:118) at
 * addresses 0..nj-1: pos, vel, acc2 (= a/2), jrk6 (= j/6) as [nj][3]. */
int pg6_dev_set_j(int nj, const double *pos, const double *vel, const double *acc2,
                  const double *jrk6, const double *mass, const double *tj, const int *idx);
/* Work of g6_set_ti + g6calc_firsthalf (phantom_g6.c:138,153): predict
 * j-particles 0..nj-1 to ti.  The prediction runs with the first force that
 * needs it; a force on one i-block predicts inside its own launch without
 * storing the predicted set (pg6_dev_read_predicted, the neighbour-list
 * calls and j-records written before the next ti store it first), so the
 * results are those of a stored prediction, bit for bit.
 * Ordering: every call that reads or writes the board's j-memory or its
 * partial sums (pg6_dev_set_j, _predict, _force, the g6calc_* calls) runs on
 * the device after the previous such call, whichever streams they were
 * queued on (the library makes the later stream wait for the earlier one). */
int pg6_dev_predict(int nj, double ti, void *stream);
/* Work of g6calc_lasthalf2p (phantom_g6.c:270) on device arrays: pos, vel
 * [ni][3] FP64, id[ni]; acc, jrk [ni][3], pot[ni] FP64; nnb/rnnb may be NULL
 * (no nearest-neighbour search, the work of g6calc_lasthalf). */
int pg6_dev_force(int nj, int ni, const double *pos, const double *vel, const int *id,
                  double eps2, double *acc, double *jrk, double *pot, int *nnb, float *rnnb,
                  void *stream);
int pg6_dev_sync(void);
void *pg6_dev_stream(void);
/* Launch-shape overrides: bi in {0 = auto, 64, 128, 256}; blocks_per_cu
 * (0 = auto) sets how many j-ranges the j-set is split into; min_jper the
 * smallest j-range (0 = 128). */
int pg6_dev_set_launch(int bi, int blocks_per_cu, int min_jper);
/* Force kernel: 0 = j-tiles staged in LDS, 1 = j broadcast through scalar
 * registers, 2 = as 1 with dx formed in FP32 from two-float positions and
 * the per-pair self test skipped on j-chunks whose id range cannot hold the
 * self pair, 3 = prediction, force and j-range reduction in one launch (up
 * to 1024 i), 4 = j in Morton order as FP32 offsets from the FP64 origin of
 * their 64-particle chunk (dx error grows with the chunk extent: chosen
 * automatically only from 16384 j), 5 = as 4 with the j-range of each
 * 64-i block split over its wavefronts, -1 = automatic (autotuned or default
 * per i-count).  0 and 1: same arithmetic, results equal up to summation
 * order; 2 and 3: dx within about one FP32 ulp of theirs. */
int pg6_dev_set_kernel(int variant);
/* Launch shape for (ni, nj): out5 = {variant, bi, blocks_per_cu, nsplit,
 * jper}. */
int pg6_dev_plan(int ni, int nj, int *out5);
/* Autotuner: times every (kernel, bi, blocks_per_cu) candidate for
 * ni = 64, 128, ..., min(nj, 65536) on the first nj predicted j-particles and
 * keeps the fastest per i-count bucket.  out[4k..4k+3] = {variant, bi,
 * blocks_per_cu, microseconds} for bucket k (may be NULL).  Returns the
 * number of buckets tuned, or -1.  No counterpart in the reference. */
int pg6_dev_autotune(int nj, double eps2, int reps, int *out);
int pg6_dev_autotune_reset(void);
/* Block-time-step Hermite integration on the GPU: the loop of the reference's
 * sample integrator PG6/banana (banana01.c:57-166, hermite.c:39-208,
 * timestep.c:24-92) with its state in HBM.  It owns the GRAPE-6 board (the
 * particles are the j-memory, address = id = index).  init: host x[n][3],
 * v[n][3], m[n]; start-up forces and steps (eta_s); eta = corrector step
 * accuracy; dtmax = largest step and time window.  step: one block step;
 * *t = time after it, *nact = particles advanced.  run: block steps until
 * t >= t_end (finite), queued without a host round trip per step; returns
 * their number.  state: host copies (each may be null),
 * t inside the current window; info[4] = {windows done, particle steps, block
 * steps, n}.  energy: 0.5 sum m v^2, 0.5 sum m pot. */
int pg6_hermite_init(int n, const double *x, const double *v, const double *m, double eps2,
                     double eta_s, double eta, double dtmax);
int pg6_hermite_step(double *t, int *nact);
long long pg6_hermite_run(double t_end);
int pg6_hermite_state(double *x, double *v, double *a, double *j, double *pot, double *t,
                      double *dt, long long *info);
int pg6_hermite_energy(double *ek, double *ep);
int pg6_hermite_end(void);
/* Read-back of the predicted j-particles (what the force kernel reads):
 * pos[nj][3] FP64, vel[nj][3] FP32, mass[nj] FP32 (may be NULL).  The
 * reference keeps them in its static Pred_Mem (PG6/libpg6/gravity.c:42-47). */
int pg6_dev_read_predicted(int nj, double *pos, float *vel, float *mass);
/* Force-kernel timing: events recorded on the launch stream around each of
 * the next max_launches pg6_force kernel launches; read returns the summed
 * kernel time (ms) and the number of launches timed, and stops timing. */
int pg6_dev_timing_start(int max_launches);
int pg6_dev_timing_read(double *ms, int *count);
/* Many independent clusters (one GRAPE-6 session per cluster, each
 * g6calc_firsthalf/lasthalf over its own members) in one persistent launch.
 * pg6_cluster_items builds the work items (cluster, first i) from host
 * offsets coff[ncl+1], largest first; returns their number. */
int pg6_cluster_items(int ncl, const int *coff, int *items, int maxitems);
/* Device arrays: items[2*nitems], coff[ncl+1], pos/vel [n][3], mass[n] FP64,
 * eps2[ncl] FP32; outputs acc/jrk [n][3], pot[n] FP64. */
int pg6_cluster_force(int nitems, const int *items, const int *coff, const double *pos,
                      const double *vel, const double *mass, const float *eps2, double *acc,
                      double *jrk, double *pot, void *stream);
void *pg6_cluster_graph(int nitems, const int *items, const int *coff, const double *pos,
                        const double *vel, const double *mass, const float *eps2, double *acc,
                        double *jrk, double *pot);
int pg6_cluster_graph_launch(void *graph, void *stream);
int pg6_cluster_graph_destroy(void *graph);
/* Autotuner over j-tiles of {128, 256, 512} particles x persistent blocks
 * per CU {1,2,3,4,6,8,12,16}; keeps the fastest pair.  *best = its blocks
 * per CU, ms[8] (nullable) = the timings at the tile kept. */
int pg6_cluster_autotune(int nitems, const int *items, const int *coff, const double *pos,
                         const double *vel, const double *mass, const float *eps2, double *acc,
                         double *jrk, double *pot, int reps, int *best, float *ms);
/* Launch shape of the cluster batch: tile in {0 = default 256, 128, 256,
 * 512} j-particles per LDS tile, blocks_per_cu (0 = default 8); a negative
 * value keeps that part.  pg6_cluster_launch_shape: out2 = {tile,
 * blocks_per_cu} in use. */
int pg6_cluster_set_launch(int tile, int blocks_per_cu);
int pg6_cluster_launch_shape(int *out2);

#if defined(__cplusplus)
}
#endif

#endif
