// Block-time-step 4th-order Hermite integration on the GPU: the host loop of
// the reference's sample integrator (PG6/banana/banana01.c:120-166,
// hermite.c:39-208, timestep.c:24-92) as device kernels around the GRAPE-6
// force.  A block step needs no host round trip: the step's time, active
// set and launch plan live in device memory (Ctl), so the host only queues
// kernels and reads Ctl back when it wants to know where the run stands.
//
// State (device, FP64, particle order = j address = id): x[n][3], v[n][3],
// a[n][3], j[n][3], pot[n], t[n], dt[n], m[n].  One block step, with
// t_next = min_i (t_i + dt_i) known from the step before:
//   1. predict every j to t_next; active = { i : t_i + dt_i == t_next },
//      appended to the active list with their predicted i data
//      (predict_exmotion); the minimum of t + dt over the others
//                                                          (pg6h_select)
//   2. force, jerk, pot of the active i against all j      (pg6h_force)
//   3. sum of the j-range partials, corrector + new time step, j-memory
//      update, reset of every time at the end of a dtmax window; the
//      minimum of the new t + dt folded in, which makes the next t_next;
//      step bookkeeping                                    (pg6h_finish)
// The first t_next comes from pg6h_min_time after the start-up.
// Every FP64 expression is written in the reference's operation order with
// contraction off, so given the same forces the corrected state and the time
// steps equal those of the reference's code bit for bit.  One deviation: the
// reference keeps the jerk in float (PG6/banana/particle.h:17-19); here it
// stays FP64.  The active list is filled by atomics, so its order varies
// from run to run; nothing computed per particle depends on it (each i's
// pair sums run over the j addresses in a fixed order).
#pragma once
#include <hip/hip_runtime.h>
#include "pg6_kernels.h"

namespace pg6h {

// Taylor coefficients as the reference writes them (hermite.c:16-19).
constexpr double kSecond = 5e-1;
constexpr double kThird = 1.666666666666667e-1;
constexpr double kFourth = 4.166666666666667e-2;
constexpr double kFifth = 8.333333333333333e-3;

struct State {
  double *x, *v, *a, *j, *pot, *t, *dt, *m;
};

struct Params {
  double eta_s, eta, dtmax, dtmin;
};

// calc_modulus (timestep.c:176-190): what remains of `num` after removing
// the powers-of-two multiples of `mod` (up to modmax) it is built from; 0 when
// `num` is a multiple of `mod` on the block-time grid.
__host__ __device__ inline double calc_modulus(double num, double mod, double modmax) {
  double tmpmod = modmax;
  while (num > 0.0) {
    while (tmpmod > num) tmpmod *= 0.5;
    if (tmpmod < mod) break;
    num -= tmpmod;
  }
  return num;
}

// dt_startup (timestep.c:24-48), without the reference's exit on failure:
// returns 0 when no admissible step exists (the caller reports it).
__device__ inline double dt_startup(const Params& P, double t_gl, const double* a,
                                    const double* j) {
#pragma clang fp contract(off)
  const double acc = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
  const double jrk = j[0] * j[0] + j[1] * j[1] + j[2] * j[2];
  const double dtcrit = P.eta_s * sqrt(acc / jrk);
  double dtblock = P.dtmax;
  while (dtblock > dtcrit && dtblock > P.dtmin) dtblock = dtblock * 0.5;
  while (calc_modulus(t_gl, dtblock, P.dtmax) != 0.0) {
    dtblock = dtblock * 0.5;
    if (dtblock < P.dtmin) return 0.0;
  }
  return dtblock;
}

// ---------------------------------------------------------------- control
// Step state in device memory.  Written by pg6h_select (nact), pg6h_finish
// (everything else, by its last block) and the host between runs (t_end).
struct Ctl {
  unsigned long long tmin;       // bits of t_next = min(t + dt) for the next step
  unsigned long long tmin_next;  // bits of min(t + dt) being folded for the step after
  double t_gl;              // time inside the current window [0, dtmax)
  double t_end;             // a step is taken only while t_mod * dtmax + t_gl < t_end
  long long t_mod;          // completed windows (t = t_mod * dtmax + t_gl)
  long long nsteps;         // particle steps taken
  long long nblocks;        // block steps taken
  int nact;                 // length of the active list being formed
  int last_nact;            // active particles of the last step taken
  int done;                 // pg6h_finish blocks finished with the current step
  int bad;                  // particles whose step fell below dtmin
};

// +infinity as FP64 bits: the neutral element of the min over t + dt (all
// t + dt are positive, so their bit patterns order as the values do).
constexpr unsigned long long kInfBits = 0x7FF0000000000000ull;

// Whether the step being queued is past the end of the run.  Every kernel of
// a step evaluates it on the same Ctl fields, which only pg6h_finish's last
// block changes, so all of them take the same decision.
__device__ inline bool pg6h_skip(const Ctl* c, double dtmax) {
#pragma clang fp contract(off)
  return (double)c->t_mod * dtmax + c->t_gl >= c->t_end;
}

// Split of the force work of `nact` active particles over all nj
// j-particles: tiles of 64 i times ns j-ranges of jper addresses, one block
// of kForceWaves wavefronts per (tile, range) item; wavefront w of the block
// takes the w-th of kForceWaves equal pieces of the range (a multiple of 16
// addresses, at least min_jper).  About `target` wavefronts of work in all.
// Partials of range s for active slot k sit at s * stride + k.
constexpr int kForceWaves = 8;
// pg6h_force sums the seven components with one wavefront each
static_assert(kForceWaves >= 7, "a block needs a wavefront per force component");

struct Plan {
  int ntiles, ns, jper, stride;
};

__host__ __device__ inline Plan pg6h_plan(int nact, int nj, int target, int min_jper) {
  Plan p;
  p.ntiles = (nact + 63) / 64;
  const int blocks = target / kForceWaves;
  int ns = p.ntiles > 0 ? (blocks + p.ntiles - 1) / p.ntiles : 1;
  if (ns > nj / (kForceWaves * min_jper)) ns = nj / (kForceWaves * min_jper);
  if (ns < 1) ns = 1;
  const int q = 16 * kForceWaves;
  p.jper = ((nj + ns - 1) / ns + q - 1) / q * q;
  p.ns = (nj + p.jper - 1) / p.jper;
  p.stride = p.ntiles * 64;
  return p;
}

// ---------------------------------------------------------------- next time
// Folds the wavefront minimum of v (lanes without a value pass 1e300) into
// *dst.  The atomic's result is consumed so that it has completed before the
// wavefront goes on.
__device__ inline void fold_min(double v, unsigned long long* dst) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
  if ((threadIdx.x & 63) == 0 && v < 1e300) {
    const unsigned long long old = atomicMin(dst, (unsigned long long)__double_as_longlong(v));
    asm volatile("" ::"v"(old));
  }
}

// The same for the whole block, with one atomic per block (all the block's
// atomics on one address would queue behind each other).  Every thread of
// the block calls it; thread 0 waits for the atomic's result (pg6h_finish's
// last block relies on that).  Blocks of at most 1024 threads.
__device__ inline void fold_min_block(double v, unsigned long long* dst) {
  __shared__ double s_min[16];
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
  if ((threadIdx.x & 63) == 0) s_min[threadIdx.x / 64] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x + 63) / 64; ++w) v = fmin(v, s_min[w]);
    if (v < 1e300) {
      const unsigned long long old = atomicMin(dst, (unsigned long long)__double_as_longlong(v));
      asm volatile("" ::"v"(old));
    }
  }
}

// t_next = min_i (t_i + dt_i) into Ctl::tmin (after the start-up).
__global__ __launch_bounds__(256) void pg6h_min_time(int n, State S, Ctl* ctl) {
#pragma clang fp contract(off)
  double v = 1e300;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    v = fmin(v, S.t[i] + S.dt[i]);
  fold_min(v, &ctl->tmin);
}

// ---------------------------------------------------------------- selection
// Every j predicted to t_next (the board's predictor, pg6_predict_j); the
// active particles (t + dt == t_next, global_time of banana01.c /
// timestep.c:81-92) appended to the active list with their i data as
// predict_exmotion (hermite.c:39-62) computes it: FP64 position and
// velocity at t_next, id = address.  Block = 256 threads, one particle each.
__global__ __launch_bounds__(256) void pg6h_select(int n, State S, Params P, Ctl* ctl,
                                                   pg6::JState js, pg6::JPred jp,
                                                   int* __restrict__ act,
                                                   double* __restrict__ ipos,
                                                   double* __restrict__ ivel,
                                                   int* __restrict__ iid) {
#pragma clang fp contract(off)
  if (pg6h_skip(ctl, P.dtmax)) return;
  const double tnext = __longlong_as_double((long long)ctl->tmin);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < n;
  pg6::pg6_predict_j(i, valid, tnext, js, jp);
  const double tn = valid ? S.t[i] + S.dt[i] : 1e300;
  const bool a = valid && tn == tnext;
  fold_min_block(a ? 1e300 : tn, &ctl->tmin_next);  // the particles that do not move
  const unsigned long long mask = __ballot(a);
  const int lane = threadIdx.x & 63;
  int base = 0;
  if (lane == 0 && mask != 0ull) base = atomicAdd(&ctl->nact, __popcll(mask));
  base = __shfl(base, 0);
  if (!a) return;
  const int k = base + __popcll(mask & ((1ull << lane) - 1ull));
  act[k] = i;
  const double dt = tnext - S.t[i];
  const double dt2 = dt * dt;
  const double dt3 = dt * dt2;
  for (int d = 0; d < 3; ++d) {
    const double x = S.x[3 * i + d], v = S.v[3 * i + d];
    const double ac = S.a[3 * i + d], j = S.j[3 * i + d];
    ipos[3 * k + d] = x + v * dt + ac * dt2 * kSecond + j * dt3 * kThird;
    ivel[3 * k + d] = v + ac * dt + j * dt2 * kSecond;
  }
  iid[k] = i;
}

// ---------------------------------------------------------------- force
// Acc, jerk and pot of the active list against all n j-particles with the
// two-float pair loop (pg6::pg6_range_h).  Block = kForceWaves wavefronts;
// blocks walk the (i-tile, j-range) items of pg6h_plan.  The wavefronts'
// sums over their pieces of a range are added in LDS in wavefront order and
// written as one FP64 partial per range.
__global__ __launch_bounds__(64 * kForceWaves) void pg6h_force(
    int n, float eps2, Params P, const Ctl* ctl, int target, int min_jper,
    const double* __restrict__ ipos, const double* __restrict__ ivel, const int* __restrict__ iid,
    pg6::JPred jp, pg6::Partials part) {
  if (pg6h_skip(ctl, P.dtmax)) return;
  const int nact = ctl->nact;
  const Plan pl = pg6h_plan(nact, n, target, min_jper);
  const int items = pl.ntiles * pl.ns;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
  const int piece = pl.jper / kForceWaves;
  double* dst[7] = {part.ax, part.ay, part.az, part.pot, part.jx, part.jy, part.jz};
  __shared__ double red[7][kForceWaves][64];
  for (int it = blockIdx.x; it < items; it += gridDim.x) {
    const int tile = it % pl.ntiles, sp = it / pl.ntiles;
    const int i = tile * 64 + lane;
    const int ic = i < nact ? i : nact - 1;  // padding lanes repeat the last slot
    const int j0 = min(n, sp * pl.jper + w * piece);
    const pg6::RangeSums r = pg6::pg6_range_h<false>(
        ipos[3 * ic], ipos[3 * ic + 1], ipos[3 * ic + 2], (float)ivel[3 * ic],
        (float)ivel[3 * ic + 1], (float)ivel[3 * ic + 2], iid[ic], eps2, j0, min(n, j0 + piece),
        jp);
    red[0][w][lane] = r.ax;
    red[1][w][lane] = r.ay;
    red[2][w][lane] = r.az;
    red[3][w][lane] = r.pot;
    red[4][w][lane] = r.jx;
    red[5][w][lane] = r.jy;
    red[6][w][lane] = r.jz;
    __syncthreads();
    if (threadIdx.x < 7 * 64) {
      const int c = threadIdx.x / 64;
      double sum = 0.0;
      for (int q = 0; q < kForceWaves; ++q) sum += red[c][q][lane];
      const int k = tile * 64 + lane;
      if (k < nact) dst[c][(size_t)sp * pl.stride + k] = sum;
    }
    __syncthreads();
  }
}

// j-memory record of particle i as set_particle_on_grape / the end of
// integrate_exmotion write it (banana01.c:206-236, hermite.c:244-274):
// acc/2, jerk/6, snap 0.
__device__ inline void store_j(const State& S, int i, pg6::JState js) {
#pragma clang fp contract(off)
  const double over2 = 0.5, over6 = 1.0 / 6.0;
  js.x[i] = S.x[3 * i];
  js.y[i] = S.x[3 * i + 1];
  js.z[i] = S.x[3 * i + 2];
  js.vx[i] = S.v[3 * i];
  js.vy[i] = S.v[3 * i + 1];
  js.vz[i] = S.v[3 * i + 2];
  js.ax[i] = S.a[3 * i] * over2;
  js.ay[i] = S.a[3 * i + 1] * over2;
  js.az[i] = S.a[3 * i + 2] * over2;
  js.kx[i] = S.j[3 * i] * over6;
  js.ky[i] = S.j[3 * i + 1] * over6;
  js.kz[i] = S.j[3 * i + 2] * over6;
  js.m[i] = S.m[i];
  js.t[i] = S.t[i];
  js.idx[i] = i;
}

// ---------------------------------------------------------------- corrector
// correct (hermite.c:66-110) and dt_criterion (timestep.c:50-79) for active
// slot k = particle i, given its forces f = {ax, ay, az, pot, jx, jy, jz} at
// the predicted position and velocity (ipos, ivel); then the j-memory
// update.  A step that ends the dtmax window stores t = 0 (reset_time,
// timestep.c:166-174: every particle is active then).  Returns the new
// t + dt.
__device__ inline double correct_one(int k, int i, const State& S, const Params& P,
                                   const double* __restrict__ ipos,
                                   const double* __restrict__ ivel, const double f[7],
                                   const pg6::JState& js, int* bad) {
#pragma clang fp contract(off)
  const double dts = S.dt[i];
  const double ts = S.t[i] + dts;  // p0.t (hermite.c:154)
  const double dtsinv = 1.0 / dts;
  const double a1[3] = {f[0], f[1], f[2]};
  const double j1[3] = {f[4], f[5], f[6]};
  double s0[3], c0[3], x1[3], v1[3];
  for (int d = 0; d < 3; ++d) {
    const double a0 = S.a[3 * i + d], j0 = S.j[3 * i + d];
    x1[d] = ipos[3 * k + d];
    v1[d] = ivel[3 * k + d];
    s0[d] = 2.0 * (-3.0 * (a0 - a1[d]) - (2.0 * j0 + j1[d]) * dts) * dtsinv * dtsinv;
    c0[d] = 6.0 * (2.0 * (a0 - a1[d]) + (j0 + j1[d]) * dts) * dtsinv * dtsinv * dtsinv;
    x1[d] += s0[d] * dts * dts * dts * dts * kFourth + c0[d] * dts * dts * dts * dts * dts * kFifth;
    v1[d] += s0[d] * dts * dts * dts * kThird + c0[d] * dts * dts * dts * dts * kFourth;
  }
  // dt_criterion
  double s1[3];
  for (int d = 0; d < 3; ++d) s1[d] = s0[d] + c0[d] * dts;
  double acc = a1[0] * a1[0] + a1[1] * a1[1] + a1[2] * a1[2];
  double jrk = j1[0] * j1[0] + j1[1] * j1[1] + j1[2] * j1[2];
  double snp = s1[0] * s1[0] + s1[1] * s1[1] + s1[2] * s1[2];
  double crc = c0[0] * c0[0] + c0[1] * c0[1] + c0[2] * c0[2];
  acc = sqrt(acc);
  jrk = sqrt(jrk);
  snp = sqrt(snp);
  crc = sqrt(crc);
  const double dti = P.eta * sqrt((acc * snp + jrk * jrk) / (jrk * crc + snp * snp));
  double dtn = dts;
  if (dti < dtn) {
    while (dtn > dti && dtn > P.dtmin) dtn = dtn * 0.5;
  } else if (dti > 2.0 * dtn) {
    if (calc_modulus(ts, 2.0 * dtn, P.dtmax) == 0.0) dtn = 2.0 * dtn;
  }
  if (!(dtn >= P.dtmin)) atomicAdd(bad, 1);
  for (int d = 0; d < 3; ++d) {
    S.x[3 * i + d] = x1[d];
    S.v[3 * i + d] = v1[d];
    S.a[3 * i + d] = a1[d];
    S.j[3 * i + d] = j1[d];
  }
  S.pot[i] = f[3];
  const double tnew = ts == P.dtmax ? 0.0 : ts;
  S.t[i] = tnew;
  S.dt[i] = dtn;
  store_j(S, i, js);
  return tnew + dtn;
}

// Sums each active particle's j-range partials in a fixed order (one thread
// per particle for up to 16 ranges, one wavefront per particle with a fixed
// butterfly for more), applies correct_one, folds the new t + dt into
// Ctl::tmin_next, and lets the last block to finish close the step in Ctl.
// Block = 256 threads.
__global__ __launch_bounds__(256) void pg6h_finish(int n, Params P, Ctl* ctl, int target,
                                                   int min_jper, const int* __restrict__ act,
                                                   State S, const double* __restrict__ ipos,
                                                   const double* __restrict__ ivel,
                                                   pg6::Partials part, pg6::JState js) {
#pragma clang fp contract(off)
  if (pg6h_skip(ctl, P.dtmax)) return;
  const int nact = ctl->nact;
  const double tnext = __longlong_as_double((long long)ctl->tmin);
  const Plan pl = pg6h_plan(nact, n, target, min_jper);
  const double* src[7] = {part.ax, part.ay, part.az, part.pot, part.jx, part.jy, part.jz};
  double tmin = 1e300;  // smallest new t + dt seen by this thread
  if (pl.ns <= 16) {
    const int k0 = blockIdx.x * blockDim.x + threadIdx.x;
    for (int k = k0; k < nact; k += gridDim.x * blockDim.x) {
      double f[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      for (int sp = 0; sp < pl.ns; ++sp)
        for (int c = 0; c < 7; ++c) f[c] += src[c][(size_t)sp * pl.stride + k];
      tmin = fmin(tmin, correct_one(k, act[k], S, P, ipos, ivel, f, js, &ctl->bad));
    }
  } else {
    const int lane = threadIdx.x & 63;
    const int wpb = blockDim.x / 64;
    for (int k = blockIdx.x * wpb + threadIdx.x / 64; k < nact; k += gridDim.x * wpb) {
      double f[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      for (int sp = lane; sp < pl.ns; sp += 64)
        for (int c = 0; c < 7; ++c) f[c] += src[c][(size_t)sp * pl.stride + k];
      // every lane ends with the same sums (a + b == b + a)
      for (int o = 32; o > 0; o >>= 1)
        for (int c = 0; c < 7; ++c) f[c] += __shfl_xor(f[c], o);
      if (lane == 0)
        tmin = fmin(tmin, correct_one(k, act[k], S, P, ipos, ivel, f, js, &ctl->bad));
    }
  }
  fold_min_block(tmin, &ctl->tmin_next);
  // The last block to get here closes the step.  It reads nothing that other
  // blocks wrote in this kernel (only nact and t_next, read by every block
  // before it counted itself done), so no memory fence is needed: a
  // device-scope fence per block would write back the L2 of its XCD.
  __shared__ int last;
  __syncthreads();
  if (threadIdx.x == 0) last = atomicAdd(&ctl->done, 1) == (int)gridDim.x - 1;
  __syncthreads();
  if (last && threadIdx.x == 0) {
    ctl->last_nact = nact;
    ctl->nsteps += nact;
    ctl->nblocks += 1;
    if (tnext == P.dtmax) {  // end of a window: every time restarts at 0
      ctl->t_gl = 0.0;
      ctl->t_mod += 1;
    } else {
      ctl->t_gl = tnext;
    }
    ctl->nact = 0;
    // every block's fold into tmin_next has completed (fold_min_block waits
    // for it)
    ctl->tmin = atomicExch(&ctl->tmin_next, kInfBits);
    ctl->done = 0;
  }
}

// startup_sgmotion (hermite.c:114-142) for every particle: forces at t = 0
// (acc1 etc. in particle order), first time steps, j-memory.
__global__ void pg6h_startup(int n, State S, Params P, const double* __restrict__ acc1,
                             const double* __restrict__ jrk1, const double* __restrict__ pot1,
                             pg6::JState js, int* __restrict__ bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int d = 0; d < 3; ++d) {
    S.a[3 * i + d] = acc1[3 * i + d];
    S.j[3 * i + d] = jrk1[3 * i + d];
  }
  S.pot[i] = pot1[i];
  const double dt = dt_startup(P, S.t[i], S.a + 3 * i, S.j + 3 * i);
  if (!(dt > 0.0)) atomicAdd(bad, 1);
  S.dt[i] = dt > 0.0 ? dt : P.dtmin;
  store_j(S, i, js);
}

// j-memory record of every particle (set_particle_on_grape).
__global__ void pg6h_store_all(int n, State S, pg6::JState js) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) store_j(S, i, js);
}

// Per-block partial sums of m v^2 and m pot, in FP64 (calc_energy.c:8-33).
__global__ void pg6h_energy_part(int n, State S, double* __restrict__ part) {
#pragma clang fp contract(off)
  __shared__ double s[2][256];
  double ek = 0.0, ew = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double v2 = S.v[3 * i] * S.v[3 * i];
    v2 += S.v[3 * i + 1] * S.v[3 * i + 1];
    v2 += S.v[3 * i + 2] * S.v[3 * i + 2];
    ek += S.m[i] * v2;
    ew += S.m[i] * S.pot[i];
  }
  s[0][threadIdx.x] = ek;
  s[1][threadIdx.x] = ew;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      s[0][threadIdx.x] += s[0][threadIdx.x + h];
      s[1][threadIdx.x] += s[1][threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s[0][0];
    part[2 * blockIdx.x + 1] = s[1][0];
  }
}

}  // namespace pg6h
