// Block-time-step 4th-order Hermite integration on the GPU: the host loop of
// the reference's sample integrator (PG6/banana/banana01.c:120-166,
// hermite.c:39-208, timestep.c:24-92) as device kernels around the GRAPE-6
// force, so a run needs no host round trip per block step beyond two scalars.
//
// State (device, FP64, particle order = j address = id): x[n][3], v[n][3],
// a[n][3], j[n][3], pot[n], t[n], dt[n], m[n].  One block step:
//   1. t_next = min_i (t_i + dt_i)                        (pg6h_next_time)
//   2. active = { i : t_i + dt_i == t_next }, index order  (pg6h_count/scan/compact)
//   3. predict every j to t_next                           (pg6_predict)
//   4. predict the active i, as predict_exmotion does      (pg6h_predict_i)
//   5. force, jerk, pot of the active i against all j      (pg6_force*)
//   6. corrector + new time step, j-memory update          (pg6h_correct)
// Every FP64 expression is written in the reference's operation order with
// contraction off, so given the same forces the corrected state and the time
// steps equal those of the reference's code bit for bit.  One deviation: the
// reference keeps the jerk in float (PG6/banana/particle.h:17-19); here it
// stays FP64.
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

// ---------------------------------------------------------------- next time
// Per-block minima of t + dt, then one block over them.
__global__ void pg6h_next_time_part(int n, State S, double* __restrict__ part) {
#pragma clang fp contract(off)
  __shared__ double s[256];
  double v = 1e300;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    v = fmin(v, S.t[i] + S.dt[i]);
  s[threadIdx.x] = v;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) s[threadIdx.x] = fmin(s[threadIdx.x], s[threadIdx.x + h]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ void pg6h_next_time_final(int nb, const double* __restrict__ part,
                                     double* __restrict__ tnext) {
  __shared__ double s[256];
  double v = 1e300;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) v = fmin(v, part[b]);
  s[threadIdx.x] = v;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) s[threadIdx.x] = fmin(s[threadIdx.x], s[threadIdx.x + h]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *tnext = s[0];
}

// ---------------------------------------------------------------- active set
// Block-wise stream compaction in index order: count per block, scan of the
// block counts (one block), then each block writes its actives.
__device__ inline bool is_active(const State& S, int i, double tnext) {
#pragma clang fp contract(off)
  return S.t[i] + S.dt[i] == tnext;
}

__global__ void pg6h_count(int n, State S, const double* __restrict__ tnext,
                           int* __restrict__ bcount) {
  __shared__ int s[256];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  s[threadIdx.x] = i < n && is_active(S, i, *tnext);
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) bcount[blockIdx.x] = s[0];
}

// boff[b] = sum of bcount[0..b); boff[nb] = total (also copied to *total).
__global__ void pg6h_scan(int nb, const int* __restrict__ bcount, int* __restrict__ boff,
                          int* __restrict__ total) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int acc = 0;
  for (int b = 0; b < nb; ++b) {
    boff[b] = acc;
    acc += bcount[b];
  }
  boff[nb] = acc;
  *total = acc;
}

__global__ void pg6h_compact(int n, State S, const double* __restrict__ tnext,
                             const int* __restrict__ boff, int* __restrict__ act) {
  __shared__ int s[256];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int f = i < n && is_active(S, i, *tnext);
  s[threadIdx.x] = f;
  __syncthreads();
  // inclusive scan in the block (Hillis-Steele), fixed order
  for (int h = 1; h < 256; h <<= 1) {
    const int v = (int)threadIdx.x >= h ? s[threadIdx.x - h] : 0;
    __syncthreads();
    s[threadIdx.x] += v;
    __syncthreads();
  }
  if (f) act[boff[blockIdx.x] + s[threadIdx.x] - 1] = i;
}

// ---------------------------------------------------------------- i prediction
// predict_exmotion (hermite.c:39-62) for the active particles: FP64 position
// and velocity at t_next, id = address.
__global__ void pg6h_predict_i(int nact, const int* __restrict__ act, State S,
                               const double* __restrict__ tnext, double* __restrict__ ipos,
                               double* __restrict__ ivel, int* __restrict__ iid) {
#pragma clang fp contract(off)
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nact) return;
  const int i = act[k];
  const double dt = *tnext - S.t[i];
  const double dt2 = dt * dt;
  const double dt3 = dt * dt2;
  for (int d = 0; d < 3; ++d) {
    const double x = S.x[3 * i + d], v = S.v[3 * i + d];
    const double a = S.a[3 * i + d], j = S.j[3 * i + d];
    ipos[3 * k + d] = x + v * dt + a * dt2 * kSecond + j * dt3 * kThird;
    ivel[3 * k + d] = v + a * dt + j * dt2 * kSecond;
  }
  iid[k] = i;
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
// correct (hermite.c:66-110) and dt_criterion (timestep.c:50-79) for the
// active particles; then the j-memory update.  acc1/jrk1/pot1 are the forces
// at the predicted positions (ipos, ivel), in active order.  *bad counts
// particles whose step fell below dtmin.
__global__ void pg6h_correct(int nact, const int* __restrict__ act, State S, Params P,
                             const double* __restrict__ ipos, const double* __restrict__ ivel,
                             const double* __restrict__ acc1, const double* __restrict__ jrk1,
                             const double* __restrict__ pot1, pg6::JState js,
                             int* __restrict__ bad) {
#pragma clang fp contract(off)
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nact) return;
  const int i = act[k];
  const double dts = S.dt[i];
  const double ts = S.t[i] + dts;  // p0.t (hermite.c:154)
  const double dtsinv = 1.0 / dts;
  double s0[3], c0[3], x1[3], v1[3];
  for (int d = 0; d < 3; ++d) {
    const double a0 = S.a[3 * i + d], j0 = S.j[3 * i + d];
    const double a1 = acc1[3 * k + d], j1 = jrk1[3 * k + d];
    x1[d] = ipos[3 * k + d];
    v1[d] = ivel[3 * k + d];
    s0[d] = 2.0 * (-3.0 * (a0 - a1) - (2.0 * j0 + j1) * dts) * dtsinv * dtsinv;
    c0[d] = 6.0 * (2.0 * (a0 - a1) + (j0 + j1) * dts) * dtsinv * dtsinv * dtsinv;
    x1[d] += s0[d] * dts * dts * dts * dts * kFourth + c0[d] * dts * dts * dts * dts * dts * kFifth;
    v1[d] += s0[d] * dts * dts * dts * kThird + c0[d] * dts * dts * dts * dts * kFourth;
  }
  // dt_criterion
  const double* a1 = acc1 + 3 * k;
  const double* j1 = jrk1 + 3 * k;
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
  S.pot[i] = pot1[k];
  S.t[i] = ts;
  S.dt[i] = dtn;
  store_j(S, i, js);
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

// reset_time (timestep.c:166-174) with set_particle_on_grape: every t = 0.
__global__ void pg6h_reset(int n, State S, pg6::JState js) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  S.t[i] = 0.0;
  js.t[i] = 0.0;
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
