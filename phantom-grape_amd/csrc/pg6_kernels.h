// GRAPE-6 (Hermite) device kernels for gfx950.
//
//  * pg6_scatter_j   : applies the j-particle records queued by
//                      g6_set_j_particle to the device j-state (SoA, FP64).
//  * pg6_predict     : 3rd-order Taylor predictor of every j-particle to the
//                      block time, FP64 position, FP32 velocity
//                      (PG6/libpg6/gravity.c:19-20,149-193).
//  * pg6_force<BI,NN>: acc, jerk, pot (+ nearest neighbour) of BI i-particles
//                      per block against one j-range
//                      (PG6/libpg6/gravity.c:260-453, 455-678).
//  * pg6_reduce      : fixed-order sum of the per-j-range partials.
//
// Arithmetic (matches the reference's precision split, gravity.c:313-421):
// Δx is formed in FP64 and rounded to FP32; pair math is FP32 with the
// hardware v_rsq_f32; acc and pot are summed in FP32 over one LDS tile and
// folded into FP64 accumulators per tile (jerk too, where the reference keeps
// an FP32 sum: a strict accuracy gain at no measurable cost).
// The self pair is excluded by particle id (gravity.c:348-351).
#pragma once
#include <hip/hip_runtime.h>
#include <cfloat>
#include <climits>

namespace pg6 {

constexpr int kThreads = 256;  // threads per block: 4 wave64
constexpr int kJTile = 256;    // j-particles staged in LDS per tile

// One queued j-particle record, as uploaded by g6calc_firsthalf.  The host
// passes acc/2 and jerk/6 (PG6/banana/banana01.c:199-212), stored as given.
struct JRecord {
  double pos[3], vel[3], acc2[3], jrk6[3];
  double mass, time;
  int idx, addr;
};
static_assert(sizeof(JRecord) == 120, "JRecord layout");

// Device j-state, SoA, capacity `cap`.
struct JState {
  double *x, *y, *z, *vx, *vy, *vz, *ax, *ay, *az, *kx, *ky, *kz, *m, *t;
  int* idx;
};

// j-particles per chunk of the self-pair filter (pg6_force_h).
constexpr int kChunk = 64;

// Predicted j-particles, SoA.
struct JPred {
  double *x, *y, *z;
  float4* vm;  // predicted velocity (FP32) and mass (FP32)
  int* id;
  // One 48-byte record per j for pg6_force_h, three float4:
  //   {x_hi, y_hi, z_hi, m}, {x_lo, y_lo, z_lo, vx}, {vy, vz, id (int bits), m},
  // the predicted position as an unevaluated sum of two floats, x = hi + lo
  // (hi = x rounded to FP32, lo = the remainder rounded to FP32).
  float4* rec;
  // Smallest and largest particle id of each chunk of kChunk consecutive
  // j addresses (chunk c = addresses [c*kChunk, (c+1)*kChunk) below nj).
  int2* crange;
};

// Per-j-range partial results, SoA over [split][i].
struct Partials {
  double *ax, *ay, *az, *pot, *jx, *jy, *jz;
  float* rmax;
  int* nn;  // j address of the nearest neighbour, -1 if none
  // pg6_force with in-kernel prediction (which leaves the JPred arrays
  // unwritten): that neighbour's id and distance
  int* nid;
  float* rdist;
};

// Final outputs in the API's layout: acc[i][3], jrk[i][3], pot[i], nnb[i], rnnb[i].
struct Outputs {
  double* acc;
  double* jrk;
  double* pot;
  int* nnb;
  float* rnnb;
};

__global__ void pg6_scatter_j(int n, const JRecord* __restrict__ rec, JState js) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const JRecord r = rec[k];
  const int a = r.addr;
  js.x[a] = r.pos[0];  js.y[a] = r.pos[1];  js.z[a] = r.pos[2];
  js.vx[a] = r.vel[0]; js.vy[a] = r.vel[1]; js.vz[a] = r.vel[2];
  js.ax[a] = r.acc2[0]; js.ay[a] = r.acc2[1]; js.az[a] = r.acc2[2];
  js.kx[a] = r.jrk6[0]; js.ky[a] = r.jrk6[1]; js.kz[a] = r.jrk6[2];
  js.m[a] = r.mass;
  js.t[a] = r.time;
  js.idx[a] = r.idx;
}

// Predictor, evaluated left to right as the reference's macro
// predict(dt,x,v,a2,j6) = x + v*dt + a2*dt*dt + j6*dt*dt*dt (gravity.c:19-20),
// velocity = v + (2 a2) dt + (3 j6) dt dt (gravity.c:171-176), with FMA
// contraction disabled so the result is the exactly rounded IEEE sequence.
__device__ inline double pg6_pred_pos(double dt, double x, double v, double a2, double j6) {
#pragma clang fp contract(off)
  return x + v * dt + a2 * dt * dt + j6 * dt * dt * dt;
}
__device__ inline double pg6_pred_vel(double dt, double v, double a2, double j6) {
#pragma clang fp contract(off)
  return v + (2.0 * a2) * dt + (3.0 * j6) * dt * dt;
}

// Predicts j-particle j (valid = j < nj) to time ti into every JPred array,
// and returns its two-float record in rec[3] (left unset if !valid).
// Called by whole wavefronts whose lane l holds j = 64 w + l, so that the
// wavefront fills in the id range of j-chunk w.
__device__ inline void pg6_predict_j(int j, bool valid, double ti, const JState& js,
                                     const JPred& jp, float4* rec = nullptr) {
  int id = 0;
  if (valid) {
    const double dt = ti - js.t[j];
    const double x = pg6_pred_pos(dt, js.x[j], js.vx[j], js.ax[j], js.kx[j]);
    const double y = pg6_pred_pos(dt, js.y[j], js.vy[j], js.ay[j], js.ky[j]);
    const double z = pg6_pred_pos(dt, js.z[j], js.vz[j], js.az[j], js.kz[j]);
    jp.x[j] = x;
    jp.y[j] = y;
    jp.z[j] = z;
    float4 vm;
    vm.x = (float)pg6_pred_vel(dt, js.vx[j], js.ax[j], js.kx[j]);
    vm.y = (float)pg6_pred_vel(dt, js.vy[j], js.ay[j], js.ky[j]);
    vm.z = (float)pg6_pred_vel(dt, js.vz[j], js.az[j], js.kz[j]);
    vm.w = (float)js.m[j];
    jp.vm[j] = vm;
    id = js.idx[j];
    jp.id[j] = id;
    const float hx = (float)x, hy = (float)y, hz = (float)z;
    float4* r = jp.rec + 3 * (size_t)j;
    const float4 r0 = make_float4(hx, hy, hz, vm.w);
    // x - (double)hx is exact in FP64
    const float4 r1 = make_float4((float)(x - (double)hx), (float)(y - (double)hy),
                                  (float)(z - (double)hz), vm.x);
    const float4 r2 = make_float4(vm.y, vm.z, __int_as_float(id), vm.w);
    r[0] = r0;
    r[1] = r1;
    r[2] = r2;
    if (rec) {
      rec[0] = r0;
      rec[1] = r1;
      rec[2] = r2;
    }
  }
  int lo = valid ? id : INT_MAX, hi = valid ? id : INT_MIN;
  for (int o = 32; o > 0; o >>= 1) {
    lo = min(lo, __shfl_xor(lo, o));
    hi = max(hi, __shfl_xor(hi, o));
  }
  if ((j % kChunk) == 0 && valid) jp.crange[j / kChunk] = make_int2(lo, hi);
}

// Block = 256 threads, one j per thread.
// Sorted j layout for pg6_force_o: the predicted j-set in the order of a
// Morton sort of their positions (taken when the layout was built), in
// chunks of kChunk.  Rewritten after every prediction by pg6_sort_records:
// chunk c carries an FP64 origin org[3c..3c+2] (the centre of its members'
// current box), a flag cflag[c] != 0 when some member's id differs from its
// address (then pg6_force_o tests the self pair by id on that chunk), and
// member k (sorted position) the record rec[2k] = {x - o, y - o, z - o
// (FP32), m}, rec[2k+1] = {vx, vy, vz, id bits}.  perm[k] = j address of
// sorted position k, inv[j] = k.
struct JSort {
  int* perm;
  int* inv;
  float4* rec;
  double* org;
  int* cflag;
};

// Block = 256 threads, one j per thread.
__global__ __launch_bounds__(256) void pg6_predict(int nj, double ti, JState js, JPred jp) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  pg6_predict_j(j, j < nj, ti, js, jp);
}

// ---------------------------------------------------------- sorted layout
// Box of the predicted positions: per block min and max of x, y, z in
// part[6 b .. 6 b + 5] (a host pass finishes it).
__global__ __launch_bounds__(256) void pg6_box_part(int nj, JPred jp, double* __restrict__ part) {
  __shared__ double s[6][256];
  double v[6] = {1e300, 1e300, 1e300, -1e300, -1e300, -1e300};
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < nj; j += gridDim.x * blockDim.x) {
    const double p[3] = {jp.x[j], jp.y[j], jp.z[j]};
    for (int d = 0; d < 3; ++d) {
      v[d] = fmin(v[d], p[d]);
      v[3 + d] = fmax(v[3 + d], p[d]);
    }
  }
  for (int d = 0; d < 6; ++d) s[d][threadIdx.x] = v[d];
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
      for (int d = 0; d < 6; ++d)
        s[d][threadIdx.x] = d < 3 ? fmin(s[d][threadIdx.x], s[d][threadIdx.x + h])
                                  : fmax(s[d][threadIdx.x], s[d][threadIdx.x + h]);
    __syncthreads();
  }
  if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

__device__ inline unsigned long long morton_spread(unsigned long long a) {
  a &= 0x1fffffull;
  a = (a | (a << 32)) & 0x1f00000000ffffull;
  a = (a | (a << 16)) & 0x1f0000ff0000ffull;
  a = (a | (a << 8)) & 0x100f00f00f00f00full;
  a = (a | (a << 4)) & 0x10c30c30c30c30c3ull;
  a = (a | (a << 2)) & 0x1249249249249249ull;
  return a;
}

// The box from nbox block partials: box[0..2] = lower corner, box[3] =
// 2^21 / (largest side), the Morton grid scale.  One block of 256 threads.
__global__ __launch_bounds__(256) void pg6_box_final(int nbox, const double* __restrict__ part,
                                                     double* __restrict__ box) {
  __shared__ double s[6][256];
  double v[6] = {1e300, 1e300, 1e300, -1e300, -1e300, -1e300};
  for (int k = threadIdx.x; k < nbox; k += 256)
    for (int d = 0; d < 6; ++d)
      v[d] = d < 3 ? fmin(v[d], part[6 * k + d]) : fmax(v[d], part[6 * k + d]);
  for (int d = 0; d < 6; ++d) s[d][threadIdx.x] = v[d];
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
      for (int d = 0; d < 6; ++d)
        s[d][threadIdx.x] = d < 3 ? fmin(s[d][threadIdx.x], s[d][threadIdx.x + h])
                                  : fmax(s[d][threadIdx.x], s[d][threadIdx.x + h]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double L = fmax(fmax(s[3][0] - s[0][0], s[4][0] - s[1][0]), fmax(s[5][0] - s[2][0], 1e-300));
    box[0] = s[0][0];
    box[1] = s[1][0];
    box[2] = s[2][0];
    box[3] = 2097152.0 / (L * (1.0 + 1e-9));
  }
}

// 63-bit Morton key of each predicted position in the box of pg6_box_final.
__global__ void pg6_sort_keys(int nj, JPred jp, const double* __restrict__ box,
                              unsigned long long* __restrict__ keys, int* __restrict__ vals) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nj) return;
  const double scale = box[3];
  const double p[3] = {jp.x[j] - box[0], jp.y[j] - box[1], jp.z[j] - box[2]};
  unsigned long long k = 0;
  for (int d = 0; d < 3; ++d) {
    double u = p[d] * scale;
    u = u < 0.0 ? 0.0 : (u > 2097151.0 ? 2097151.0 : u);
    k |= morton_spread((unsigned long long)u) << (2 - d);
  }
  keys[j] = k;
  vals[j] = j;
}

// Chunk origins, id flags and records of the sorted layout from the current
// prediction (after the sort, also the inverse map: with_inv).  One thread
// per sorted position k, one wavefront per chunk; block = 256 threads.
__global__ __launch_bounds__(256) void pg6_sort_records(int nj, JPred jp, JSort so,
                                                        bool with_inv) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const bool v = k < nj;
  const int j = v ? so.perm[k] : 0;
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  double p[3] = {0.0, 0.0, 0.0};
  int id = 0;
  if (v) {
    if (with_inv) so.inv[j] = k;
    p[0] = jp.x[j];
    p[1] = jp.y[j];
    p[2] = jp.z[j];
    id = jp.id[j];
    for (int d = 0; d < 3; ++d) lo[d] = hi[d] = p[d];
  }
  for (int o = 32; o > 0; o >>= 1)
    for (int d = 0; d < 3; ++d) {
      lo[d] = fmin(lo[d], __shfl_xor(lo[d], o));
      hi[d] = fmax(hi[d], __shfl_xor(hi[d], o));
    }
  // blocks of 256 threads hold four whole chunks of kChunk = 64; after the
  // butterfly every lane has its chunk's box
  const double o[3] = {0.5 * (lo[0] + hi[0]), 0.5 * (lo[1] + hi[1]), 0.5 * (lo[2] + hi[2])};
  const bool mixed = __builtin_amdgcn_ballot_w64(v && id != j) != 0ull;
  if ((k % kChunk) == 0 && v) {
    for (int d = 0; d < 3; ++d) so.org[3 * (size_t)(k / kChunk) + d] = o[d];
    so.cflag[k / kChunk] = mixed ? 1 : 0;
  }
  if (v) {
    const float4 vm = jp.vm[j];
    so.rec[2 * (size_t)k] = make_float4((float)(p[0] - o[0]), (float)(p[1] - o[1]),
                                        (float)(p[2] - o[2]), vm.w);
    so.rec[2 * (size_t)k + 1] = make_float4(vm.x, vm.y, vm.z, __int_as_float(id));
  }
}


// Nearest-neighbour key of one pair: the top 24 bits of 1/r (a larger key is
// a nearer j) with the tile slot 255 - jj in the low byte, so that max()
// keeps the lowest slot among equal 1/r.  One v_perm_b32: bytes 3..1 from
// 1/r, byte 0 from the (wave-uniform) slot.
__device__ inline unsigned nn_key(float rinv, unsigned, unsigned slot) {
  return __builtin_amdgcn_perm(__float_as_uint(rinv), slot, 0x07060500u);
}

// sqrt(|dx|^2 + eps^2) to j-particle j, with dx formed as in the force loop
// (FP64 difference rounded to FP32): the reference's rnnb
// (PG6/libpg6/gravity.c:652-673).
__device__ inline float nn_distance(const JPred& jp, int j, double xi, double yi, double zi,
                                    float eps2) {
  const float dx = (float)(jp.x[j] - xi);
  const float dy = (float)(jp.y[j] - yi);
  const float dz = (float)(jp.z[j] - zi);
  return sqrtf(fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2))));
}

// Force + jerk + potential (+ nearest neighbour) kernel.
//
// Block = NT threads (256; kOneBlockThreads for a launch with one i-block of
// 64 or 128).  BI i-particles per block (64, 128 or 256); the block's G = NT/BI wave
// groups split each LDS tile's j-particles between them (group g takes
// j = g, g+G, ...), and are combined in fixed order at the end.
// blockIdx.y selects the j-range [y*jper, min(nj,(y+1)*jper)).
// With nsplit == 1 the kernel writes final outputs, otherwise partials.
// PRED: the block predicts the j of each tile itself (the arithmetic of
// pg6_predict_j at time ti) straight into LDS instead of reading the
// predicted set, and writes no JPred array; for launches with one i-block,
// where every j is in exactly one block's range, so that the separate
// predictor launch and the predicted set's round trip through memory go
// away.  The nearest neighbour's id and distance then come from the tile in
// LDS (Partials::nid / rdist, or the outputs directly).
// A launch with one i-block of 64 or 128 runs blocks of kOneBlockThreads = 8
// wavefronts (two per SIMD at one block per CU): against 65536 j in 256
// j-ranges, 64 i take 18.9 us per predict + force call against 19.8 with 4
// wavefronts, 128 i 24.4 against 27.3 (no change at 16384 j); it is used with
// and without PRED alike, so the two sum in the same order.
constexpr int kOneBlockThreads = 512;

template <int BI, bool NN, bool PRED, int NT = kThreads>
__global__ __launch_bounds__(NT) void pg6_force(
    int ni, const double* __restrict__ ipos, const double* __restrict__ ivel,
    const int* __restrict__ iid, float eps2, int nj, int jper, JPred jp,
    Partials part, Outputs out, int nsplit, double ti, JState js) {
  static_assert(NT >= kJTile && NT % BI == 0, "one thread stages at most one j of a tile");
  constexpr int G = NT / BI;
  __shared__ double s_x[kJTile], s_y[kJTile], s_z[kJTile];
  __shared__ float4 s_vm[kJTile];
  __shared__ int s_id[kJTile];

  const int tid = threadIdx.x;
  const int il = tid % BI;
  const int g = __builtin_amdgcn_readfirstlane(tid / BI);  // uniform: BI is a multiple of 64
  const int i = blockIdx.x * BI + il;
  const bool ivalid = i < ni;
  const int ic = ivalid ? i : ni - 1;

  const double xi = ipos[3 * ic + 0];
  const double yi = ipos[3 * ic + 1];
  const double zi = ipos[3 * ic + 2];
  const float vxi = (float)ivel[3 * ic + 0];
  const float vyi = (float)ivel[3 * ic + 1];
  const float vzi = (float)ivel[3 * ic + 2];
  const int idi = iid[ic];

  const int j0 = blockIdx.y * jper;
  const int j1 = min(nj, j0 + jper);

  double ax = 0.0, ay = 0.0, az = 0.0, pot = 0.0;
  double jx = 0.0, jy = 0.0, jz = 0.0;
  // Nearest neighbour = largest 1/r.  Per pair only a packed key is kept:
  // the bits of 1/r (non-negative, so unsigned order = float order) with the
  // low 8 mantissa bits replaced by 255 - (tile slot), reduced with an
  // unsigned max; the tile's winner is decoded once per tile.  1/r is thereby
  // compared at 2^-15 relative resolution, and among candidates that close
  // the lowest j address wins.  The self pair has 1/r = 0 and never wins.
  float rmax = 0.f;  // quantised 1/r of the nearest neighbour so far
  int nn = -1;
  int nid = -1;       // PRED: its id
  float rdist = 0.f;  // PRED: its distance
  const unsigned kmask = 0xFFFFFF00u;

  for (int jt = j0; jt < j1; jt += kJTile) {
    const int nt = min(kJTile, j1 - jt);
    __syncthreads();
    if (tid < nt) {
      const int j = jt + tid;
      if (PRED) {  // pg6_predict_j's arithmetic, into LDS only
        const double dt = ti - js.t[j];
        s_x[tid] = pg6_pred_pos(dt, js.x[j], js.vx[j], js.ax[j], js.kx[j]);
        s_y[tid] = pg6_pred_pos(dt, js.y[j], js.vy[j], js.ay[j], js.ky[j]);
        s_z[tid] = pg6_pred_pos(dt, js.z[j], js.vz[j], js.az[j], js.kz[j]);
        float4 vm;
        vm.x = (float)pg6_pred_vel(dt, js.vx[j], js.ax[j], js.kx[j]);
        vm.y = (float)pg6_pred_vel(dt, js.vy[j], js.ay[j], js.ky[j]);
        vm.z = (float)pg6_pred_vel(dt, js.vz[j], js.az[j], js.kz[j]);
        vm.w = (float)js.m[j];
        s_vm[tid] = vm;
        s_id[tid] = js.idx[j];
      } else {
        s_x[tid] = jp.x[j];
        s_y[tid] = jp.y[j];
        s_z[tid] = jp.z[j];
        s_vm[tid] = jp.vm[j];
        s_id[tid] = jp.id[j];
      }
    }
    __syncthreads();

    float fax = 0.f, fay = 0.f, faz = 0.f, fpot = 0.f;
    float fjx = 0.f, fjy = 0.f, fjz = 0.f;
    unsigned tkey = 0u;
#pragma unroll 4
    for (int jj = g; jj < nt; jj += G) {
      const float dx = (float)(s_x[jj] - xi);
      const float dy = (float)(s_y[jj] - yi);
      const float dz = (float)(s_z[jj] - zi);
      const float4 vm = s_vm[jj];
      const float dvx = vm.x - vxi;
      const float dvy = vm.y - vyi;
      const float dvz = vm.z - vzi;
      const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2)));
      const float rv = fmaf(dz, dvz, fmaf(dy, dvy, dx * dvx));
      const bool self = s_id[jj] == idi;
      const float rinv = self ? 0.f : __builtin_amdgcn_rsqf(r2);
      const float rinv2 = rinv * rinv;
      const float mr = vm.w * rinv;
      const float mr3 = mr * rinv2;
      const float beta = -3.f * rv * rinv2;
      fpot -= mr;
      fax = fmaf(mr3, dx, fax);
      fay = fmaf(mr3, dy, fay);
      faz = fmaf(mr3, dz, faz);
      fjx = fmaf(mr3, fmaf(beta, dx, dvx), fjx);
      fjy = fmaf(mr3, fmaf(beta, dy, dvy), fjy);
      fjz = fmaf(mr3, fmaf(beta, dz, dvz), fjz);
      if (NN) {
        tkey = max(tkey, nn_key(rinv, kmask, (unsigned)(255 - jj)));
      }
    }
    if (NN) {
      const float rq = __uint_as_float(tkey & kmask);
      if (rq > rmax) {  // strict: an earlier tile keeps a tie
        rmax = rq;
        const int slot = 255 - (int)(tkey & 0xFFu);
        nn = jt + slot;
        if (PRED) {  // nn_distance on the tile's copy of the predicted position
          nid = s_id[slot];
          const float dx = (float)(s_x[slot] - xi);
          const float dy = (float)(s_y[slot] - yi);
          const float dz = (float)(s_z[slot] - zi);
          rdist = sqrtf(fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2))));
        }
      }
    }
    ax += (double)fax;
    ay += (double)fay;
    az += (double)faz;
    pot += (double)fpot;
    jx += (double)fjx;
    jy += (double)fjy;
    jz += (double)fjz;
  }

  // Combine the G wave groups in fixed order g = 0..G-1.
  if (G > 1) {
    __shared__ double r_d[7][NT];
    __shared__ float r_f[NT];
    __shared__ int r_n[NT];
    __shared__ int r_id[PRED ? NT : 1];
    __shared__ float r_dist[PRED ? NT : 1];
    __syncthreads();
    r_d[0][tid] = ax; r_d[1][tid] = ay; r_d[2][tid] = az; r_d[3][tid] = pot;
    r_d[4][tid] = jx; r_d[5][tid] = jy; r_d[6][tid] = jz;
    r_f[tid] = rmax;
    r_n[tid] = nn;
    if (PRED) {
      r_id[tid] = nid;
      r_dist[tid] = rdist;
    }
    __syncthreads();
    if (g == 0) {
      for (int q = 1; q < G; ++q) {
        const int t = q * BI + il;
        ax += r_d[0][t]; ay += r_d[1][t]; az += r_d[2][t]; pot += r_d[3][t];
        jx += r_d[4][t]; jy += r_d[5][t]; jz += r_d[6][t];
        if (NN) {
          const float rq = r_f[t];
          const int nq = r_n[t];
          // Larger 1/r wins; on a tie the lower j address wins.
          if (rq > rmax || (rq == rmax && nq >= 0 && (nn < 0 || nq < nn))) {
            rmax = rq;
            nn = nq;
            if (PRED) {
              nid = r_id[t];
              rdist = r_dist[t];
            }
          }
        }
      }
    }
  }
  if (g != 0 || !ivalid) return;

  if (nsplit == 1) {
    out.acc[3 * i + 0] = ax;
    out.acc[3 * i + 1] = ay;
    out.acc[3 * i + 2] = az;
    out.jrk[3 * i + 0] = jx;
    out.jrk[3 * i + 1] = jy;
    out.jrk[3 * i + 2] = jz;
    out.pot[i] = pot;
    if (NN) {
      if (PRED) {
        out.nnb[i] = nn >= 0 ? nid : -1;
        out.rnnb[i] = nn >= 0 ? rdist : 0.f;
      } else {
        out.nnb[i] = nn >= 0 ? jp.id[nn] : -1;
        out.rnnb[i] = nn >= 0 ? nn_distance(jp, nn, xi, yi, zi, eps2) : 0.f;
      }
    }
  } else {
    const size_t o = (size_t)blockIdx.y * ni + i;
    part.ax[o] = ax;
    part.ay[o] = ay;
    part.az[o] = az;
    part.pot[o] = pot;
    part.jx[o] = jx;
    part.jy[o] = jy;
    part.jz[o] = jz;
    if (NN) {
      part.rmax[o] = rmax;
      part.nn[o] = nn;
      if (PRED) {
        part.nid[o] = nid;
        part.rdist[o] = rdist;
      }
    }
  }
}

// Variant of pg6_force that broadcasts the j-particles through SGPRs instead
// of LDS: every wave walks the same j sequence, so j data is wave-uniform;
// loaded through the constant address space it arrives by scalar loads and
// the VALU instructions take it as a scalar operand (no LDS traffic, no
// per-lane address arithmetic).  Block = 64, 128 or 256 threads, one
// i-particle per lane; every wave walks the block's whole j-range.  Same
// arithmetic, nearest-neighbour rule and outputs as pg6_force.
typedef const __attribute__((address_space(4))) double cdouble;
typedef const __attribute__((address_space(4))) float cfloat_s;
typedef const __attribute__((address_space(4))) int cint;

template <bool NN>
__global__ __launch_bounds__(kThreads) void pg6_force_s(
    int ni, const double* __restrict__ ipos, const double* __restrict__ ivel,
    const int* __restrict__ iid, float eps2, int nj, int jper, JPred jp, Partials part,
    Outputs out, int nsplit) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool ivalid = i < ni;
  const int ic = ivalid ? i : ni - 1;
  const double xi = ipos[3 * ic + 0];
  const double yi = ipos[3 * ic + 1];
  const double zi = ipos[3 * ic + 2];
  const float vxi = (float)ivel[3 * ic + 0];
  const float vyi = (float)ivel[3 * ic + 1];
  const float vzi = (float)ivel[3 * ic + 2];
  const int idi = iid[ic];
  cdouble* px = (cdouble*)jp.x;
  cdouble* py = (cdouble*)jp.y;
  cdouble* pz = (cdouble*)jp.z;
  cfloat_s* pvm = (cfloat_s*)jp.vm;  // float4 records read as 4 floats
  cint* pid = (cint*)jp.id;
  const int j0 = blockIdx.y * jper;
  const int j1 = min(nj, j0 + jper);
  double ax = 0.0, ay = 0.0, az = 0.0, pot = 0.0, jx = 0.0, jy = 0.0, jz = 0.0;
  float rmax = 0.f;
  int nn = -1;
  const unsigned kmask = 0xFFFFFF00u;
  for (int jt = j0; jt < j1; jt += kJTile) {
    const int nt = min(kJTile, j1 - jt);
    float fax = 0.f, fay = 0.f, faz = 0.f, fpot = 0.f, fjx = 0.f, fjy = 0.f, fjz = 0.f;
    unsigned tkey = 0u;
#pragma unroll 4
    for (int jj = 0; jj < nt; ++jj) {
      const int j = jt + jj;
      const float dx = (float)(px[j] - xi);
      const float dy = (float)(py[j] - yi);
      const float dz = (float)(pz[j] - zi);
      const float dvx = pvm[4 * j + 0] - vxi;
      const float dvy = pvm[4 * j + 1] - vyi;
      const float dvz = pvm[4 * j + 2] - vzi;
      const float mj = pvm[4 * j + 3];
      const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2)));
      const float rv = fmaf(dz, dvz, fmaf(dy, dvy, dx * dvx));
      const bool self = pid[j] == idi;
      const float rinv = self ? 0.f : __builtin_amdgcn_rsqf(r2);
      const float rinv2 = rinv * rinv;
      const float mr = mj * rinv;
      const float mr3 = mr * rinv2;
      const float beta = -3.f * rv * rinv2;
      fpot -= mr;
      fax = fmaf(mr3, dx, fax);
      fay = fmaf(mr3, dy, fay);
      faz = fmaf(mr3, dz, faz);
      fjx = fmaf(mr3, fmaf(beta, dx, dvx), fjx);
      fjy = fmaf(mr3, fmaf(beta, dy, dvy), fjy);
      fjz = fmaf(mr3, fmaf(beta, dz, dvz), fjz);
      if (NN) {
        tkey = max(tkey, nn_key(rinv, kmask, (unsigned)(255 - jj)));
      }
    }
    if (NN) {
      const float rq = __uint_as_float(tkey & kmask);
      if (rq > rmax) {
        rmax = rq;
        nn = jt + 255 - (int)(tkey & 0xFFu);
      }
    }
    ax += (double)fax;
    ay += (double)fay;
    az += (double)faz;
    pot += (double)fpot;
    jx += (double)fjx;
    jy += (double)fjy;
    jz += (double)fjz;
  }
  if (!ivalid) return;
  if (nsplit == 1) {
    out.acc[3 * i + 0] = ax;
    out.acc[3 * i + 1] = ay;
    out.acc[3 * i + 2] = az;
    out.jrk[3 * i + 0] = jx;
    out.jrk[3 * i + 1] = jy;
    out.jrk[3 * i + 2] = jz;
    out.pot[i] = pot;
    if (NN) {
      out.nnb[i] = nn >= 0 ? jp.id[nn] : -1;
      out.rnnb[i] = nn >= 0 ? nn_distance(jp, nn, xi, yi, zi, eps2) : 0.f;
    }
  } else {
    const size_t o = (size_t)blockIdx.y * ni + i;
    part.ax[o] = ax;
    part.ay[o] = ay;
    part.az[o] = az;
    part.pot[o] = pot;
    part.jx[o] = jx;
    part.jy[o] = jy;
    part.jz[o] = jz;
    if (NN) {
      part.rmax[o] = rmax;
      part.nn[o] = nn;
    }
  }
}

// Accumulators of one i-particle over one j-tile (FP32), and its pair step
// for pg6_force_h.  SELF = whether the pair may be the self pair and must
// be tested by id.
struct TileAcc {
  float ax, ay, az, pot, jx, jy, jz;
  unsigned key;
};

template <bool SELF, bool NN>
__device__ __forceinline__ void pg6_pair_h(TileAcc& t, float dx, float dy, float dz, float dvx,
                                           float dvy, float dvz, float mj, float eps2,
                                           bool same_id, unsigned slot) {
  const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2)));
  const float rv = fmaf(dz, dvz, fmaf(dy, dvy, dx * dvx));
  const float rinv = (SELF && same_id) ? 0.f : __builtin_amdgcn_rsqf(r2);
  const float rinv2 = rinv * rinv;
  const float mr = mj * rinv;
  const float mr3 = mr * rinv2;
  const float beta = -3.f * rv * rinv2;
  t.pot -= mr;
  t.ax = fmaf(mr3, dx, t.ax);
  t.ay = fmaf(mr3, dy, t.ay);
  t.az = fmaf(mr3, dz, t.az);
  t.jx = fmaf(mr3, fmaf(beta, dx, dvx), t.jx);
  t.jy = fmaf(mr3, fmaf(beta, dy, dvy), t.jy);
  t.jz = fmaf(mr3, fmaf(beta, dz, dvz), t.jz);
  if (NN) t.key = max(t.key, nn_key(rinv, 0u, slot));
}

// Sums of one i-particle over one j-range (FP64 over FP32 tile sums), and the
// quantised 1/r and j address of its nearest neighbour (NN only).
struct RangeSums {
  double ax, ay, az, pot, jx, jy, jz;
  float rmax;
  int nn;
};

// Pair loop of pg6_force_h for the i-particle of this lane against the
// j-addresses [j0, j1):
//  * dx = (x_hi,j - x_hi,i) + (x_lo,j - x_lo,i), three FP32 operations on the
//    two-float positions (JPred::rec), in place of an FP64 subtraction and an
//    FP64->FP32 conversion.  It differs from the FP32-rounded FP64 difference
//    by at most about one FP32 ulp of dx (the hi difference is exact for
//    nearby pairs), which is the size of that rounding itself;
//  * the j-range is walked in pieces that stay inside one chunk of kChunk
//    addresses; a piece whose chunk's id range (JPred::crange) is disjoint
//    from the id range of the wavefront's i-particles cannot hold a self
//    pair, and runs without the per-pair id test.  The test is exact, so the
//    result does not depend on it.
// FP32 sums run over tiles of kJTile addresses from j0 and are folded into
// FP64 after each tile.  Every lane of the wavefront calls it with the same
// j0, j1 (the j data are wave-uniform and arrive through scalar loads).
template <bool NN>
__device__ __forceinline__ RangeSums pg6_range_h(double xi, double yi, double zi, float vxi,
                                                 float vyi, float vzi, int idi, float eps2,
                                                 int j0, int j1, const JPred& jp) {
  const float xih = (float)xi, yih = (float)yi, zih = (float)zi;
  const float xil = (float)(xi - (double)xih);
  const float yil = (float)(yi - (double)yih);
  const float zil = (float)(zi - (double)zih);
  // id range of this wavefront's i-particles
  int wlo = idi, whi = idi;
  for (int o = 32; o > 0; o >>= 1) {
    wlo = min(wlo, __shfl_xor(wlo, o));
    whi = max(whi, __shfl_xor(whi, o));
  }
  wlo = __builtin_amdgcn_readfirstlane(wlo);
  whi = __builtin_amdgcn_readfirstlane(whi);
  cint* pcr = (cint*)jp.crange;
  RangeSums r{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.f, -1};
  const unsigned kmask = 0xFFFFFF00u;
  for (int jt = j0; jt < j1; jt += kJTile) {
    const int nt = min(kJTile, j1 - jt);
    TileAcc t{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0u};
    for (int c = 0, ce; c < nt; c = ce) {
      const int ch = (jt + c) / kChunk;
      ce = min(nt, (ch + 1) * kChunk - jt);  // to the end of the chunk or the tile
      const bool clear = pcr[2 * ch + 1] < wlo || pcr[2 * ch] > whi;
      cfloat_s* q = (cfloat_s*)(jp.rec + 3 * (size_t)(jt + c));
#define PG6_PAIR_H(SELF)                                                                  \
  {                                                                                       \
    const float dx = (q[0] - xih) + (q[4] - xil);                                         \
    const float dy = (q[1] - yih) + (q[5] - yil);                                         \
    const float dz = (q[2] - zih) + (q[6] - zil);                                         \
    pg6_pair_h<SELF, NN>(t, dx, dy, dz, q[7] - vxi, q[8] - vyi, q[9] - vzi, q[3], eps2,   \
                         SELF && __float_as_int(q[10]) == idi, (unsigned)(255 - jj));     \
  }
      if (clear) {
#pragma unroll 4
        for (int jj = c; jj < ce; ++jj, q += 12) PG6_PAIR_H(false)
      } else {
#pragma unroll 4
        for (int jj = c; jj < ce; ++jj, q += 12) PG6_PAIR_H(true)
      }
#undef PG6_PAIR_H
    }
    if (NN) {
      const float rq = __uint_as_float(t.key & kmask);
      if (rq > r.rmax) {
        r.rmax = rq;
        r.nn = jt + 255 - (int)(t.key & 0xFFu);
      }
    }
    r.ax += (double)t.ax;
    r.ay += (double)t.ay;
    r.az += (double)t.az;
    r.pot += (double)t.pot;
    r.jx += (double)t.jx;
    r.jy += (double)t.jy;
    r.jz += (double)t.jz;
  }
  return r;
}

// Variant of pg6_force_s with no FP64 arithmetic and no id test in most of
// the pair loop (pg6_range_h).  Block = 64, 128 or 256 threads, one
// i-particle per lane, as pg6_force_s.
// Same outputs and nearest-neighbour rule (1/r from the two-float dx).
template <bool NN>
__global__ __launch_bounds__(kThreads) void pg6_force_h(
    int ni, const double* __restrict__ ipos, const double* __restrict__ ivel,
    const int* __restrict__ iid, float eps2, int nj, int jper, JPred jp, Partials part,
    Outputs out, int nsplit) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool ivalid = i < ni;
  const int ic = ivalid ? i : ni - 1;  // padding lanes repeat i = ni-1
  const double xi = ipos[3 * ic + 0];
  const double yi = ipos[3 * ic + 1];
  const double zi = ipos[3 * ic + 2];
  const int j0 = blockIdx.y * jper;
  const RangeSums r = pg6_range_h<NN>(xi, yi, zi, (float)ivel[3 * ic + 0],
                                      (float)ivel[3 * ic + 1], (float)ivel[3 * ic + 2], iid[ic],
                                      eps2, j0, min(nj, j0 + jper), jp);
  if (!ivalid) return;
  if (nsplit == 1) {
    out.acc[3 * i + 0] = r.ax;
    out.acc[3 * i + 1] = r.ay;
    out.acc[3 * i + 2] = r.az;
    out.jrk[3 * i + 0] = r.jx;
    out.jrk[3 * i + 1] = r.jy;
    out.jrk[3 * i + 2] = r.jz;
    out.pot[i] = r.pot;
    if (NN) {
      out.nnb[i] = r.nn >= 0 ? jp.id[r.nn] : -1;
      out.rnnb[i] = r.nn >= 0 ? nn_distance(jp, r.nn, xi, yi, zi, eps2) : 0.f;
    }
  } else {
    const size_t o = (size_t)blockIdx.y * ni + i;
    part.ax[o] = r.ax;
    part.ay[o] = r.ay;
    part.az[o] = r.az;
    part.pot[o] = r.pot;
    part.jx[o] = r.jx;
    part.jy[o] = r.jy;
    part.jz[o] = r.jz;
    if (NN) {
      part.rmax[o] = r.rmax;
      part.nn[o] = r.nn;
    }
  }
}

// Pair loop of pg6_force_o for the i-particle of this lane against the
// sorted positions [k0, k1) of the sorted layout (JSort):
//  * dx = u_j - d_i, one FP32 operation: u_j = x_j - o_c (FP32, stored) and
//    d_i = FP32(x_i - o_c), formed in FP64 once per chunk and lane.  The two
//    roundings are at most 2^-24 |x_j - o_c| and 2^-24 |x_i - o_c|, with
//    |x_i - o_c| <= |dx| + |x_j - o_c|: an FP32 rounding of dx itself plus
//    ulps of the chunk's size, so a close pair keeps its digits when the
//    chunk is small (the Morton order keeps it so);
//  * the self pair is tested by id only on chunks that hold the sorted
//    position of some lane's address idi (self_k = inv[idi], -1 when idi is
//    no address) or a member whose id differs from its address (cflag); the
//    j whose id is idi sits in one of them.
// FP32 sums over each chunk (at most kChunk positions), folded into FP64.
// rmax/nn as pg6_range_h per tile of kJTile positions, nn converted to the j
// address.  Every lane of the wavefront calls it with the same k0, k1 (the
// records are read by scalar loads).
template <bool NN>
__device__ __forceinline__ RangeSums pg6_range_o(double xi, double yi, double zi, float vxi,
                                                 float vyi, float vzi, int idi, int self_k,
                                                 float eps2, int k0, int k1, const JSort& so) {
  cdouble* porg = (cdouble*)so.org;
  RangeSums r{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.f, -1};
  const unsigned kmask = 0xFFFFFF00u;
  const int self_c = self_k >= 0 ? self_k / kChunk : -1;
  for (int kt = k0; kt < k1; kt += kJTile) {
    const int nt = min(kJTile, k1 - kt);
    TileAcc t{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0u};
    for (int c = 0, ce; c < nt; c = ce) {
      const int ch = (kt + c) / kChunk;
      ce = min(nt, (ch + 1) * kChunk - kt);  // to the end of the chunk or the tile
      const double ox = porg[3 * ch], oy = porg[3 * ch + 1], oz = porg[3 * ch + 2];
      const float dxi = (float)(xi - ox), dyi = (float)(yi - oy), dzi = (float)(zi - oz);
      const bool self_here =
          __builtin_amdgcn_ballot_w64(self_c == ch) != 0ull || so.cflag[ch] != 0;
      cfloat_s* q = (cfloat_s*)(so.rec + 2 * (size_t)(kt + c));
#define PG6_PAIR_O(SELF)                                                                   \
  {                                                                                        \
    const float dx = q[0] - dxi;                                                           \
    const float dy = q[1] - dyi;                                                           \
    const float dz = q[2] - dzi;                                                           \
    pg6_pair_h<SELF, NN>(t, dx, dy, dz, q[4] - vxi, q[5] - vyi, q[6] - vzi, q[3], eps2,    \
                         SELF && __float_as_int(q[7]) == idi, (unsigned)(255 - jj));       \
  }
      // two pairs per iteration: 2-3 % faster than four or eight at seven
      // wavefronts per SIMD (same registers, less to wait for per batch)
      if (self_here) {
#pragma unroll 2
        for (int jj = c; jj < ce; ++jj, q += 8) PG6_PAIR_O(true)
      } else {
#pragma unroll 2
        for (int jj = c; jj < ce; ++jj, q += 8) PG6_PAIR_O(false)
      }
#undef PG6_PAIR_O
      // fold the chunk's FP32 sums: neighbouring chunks hold close j on
      // opposite sides of nearby i, whose large terms partly cancel
      r.ax += (double)t.ax;
      r.ay += (double)t.ay;
      r.az += (double)t.az;
      r.pot += (double)t.pot;
      r.jx += (double)t.jx;
      r.jy += (double)t.jy;
      r.jz += (double)t.jz;
      t.ax = t.ay = t.az = t.pot = t.jx = t.jy = t.jz = 0.f;
    }
    if (NN) {
      const float rq = __uint_as_float(t.key & kmask);
      if (rq > r.rmax) {
        r.rmax = rq;
        r.nn = kt + 255 - (int)(t.key & 0xFFu);
      }
    }
  }
  if (NN && r.nn >= 0) r.nn = so.perm[r.nn];
  return r;
}

// pg6_force_h on the sorted layout (pg6_range_o): one i-particle per lane,
// blocks of BI threads, blockIdx.y = range of sorted positions.  Same outputs
// and partials as pg6_force_h (nearest neighbours by j address).
// At least eight wavefronts per SIMD (64 VGPRs; left alone the compiler
// takes 85 and five): more wavefronts to cover the scalar-load latency of
// the j records.
template <bool NN>
__global__ __launch_bounds__(kThreads, 8) void pg6_force_o(
    int ni, const double* __restrict__ ipos, const double* __restrict__ ivel,
    const int* __restrict__ iid, float eps2, int nj, int jper, JSort so, JPred jp, Partials part,
    Outputs out, int nsplit) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool ivalid = i < ni;
  const int ic = ivalid ? i : ni - 1;  // padding lanes repeat i = ni-1
  const double xi = ipos[3 * ic + 0];
  const double yi = ipos[3 * ic + 1];
  const double zi = ipos[3 * ic + 2];
  const int idi = iid[ic];
  const int self_k = (idi >= 0 && idi < nj) ? so.inv[idi] : -1;
  const int j0 = blockIdx.y * jper;
  const RangeSums r = pg6_range_o<NN>(xi, yi, zi, (float)ivel[3 * ic + 0],
                                      (float)ivel[3 * ic + 1], (float)ivel[3 * ic + 2], idi,
                                      self_k, eps2, j0, min(nj, j0 + jper), so);
  if (!ivalid) return;
  if (nsplit == 1) {
    out.acc[3 * i + 0] = r.ax;
    out.acc[3 * i + 1] = r.ay;
    out.acc[3 * i + 2] = r.az;
    out.jrk[3 * i + 0] = r.jx;
    out.jrk[3 * i + 1] = r.jy;
    out.jrk[3 * i + 2] = r.jz;
    out.pot[i] = r.pot;
    if (NN) {
      out.nnb[i] = r.nn >= 0 ? jp.id[r.nn] : -1;
      out.rnnb[i] = r.nn >= 0 ? nn_distance(jp, r.nn, xi, yi, zi, eps2) : 0.f;
    }
  } else {
    const size_t o = (size_t)blockIdx.y * ni + i;
    part.ax[o] = r.ax;
    part.ay[o] = r.ay;
    part.az[o] = r.az;
    part.pot[o] = r.pot;
    part.jx[o] = r.jx;
    part.jy[o] = r.jy;
    part.jz[o] = r.jz;
    if (NN) {
      part.rmax[o] = r.rmax;
      part.nn[o] = r.nn;
    }
  }
}

// Nearest-neighbour candidate order: larger (quantised) 1/r wins, equal
// values go to the lower j address; (0, -1) is "none".  The order is total,
// so any reduction tree gives the same winner.
__device__ inline bool nn_better(float r, int n, float rbest, int nbest) {
  return r > rbest || (r == rbest && n >= 0 && (nbest < 0 || n < nbest));
}

// pg6_force_o with the j-range of a block split over its W wavefronts, all on
// the same 64 i-particles (blockIdx.x = i-block of 64, blockIdx.y = range of
// sorted positions, jper a multiple of W * kChunk): wave w walks the w-th
// piece, and wave 0 adds the W sums in wave order (nearest neighbour by
// nn_better).  W times fewer partial sums than pg6_force_o for the same
// number of wavefronts.  Block = 64 W threads.
// Seven wavefronts per SIMD (72 VGPRs; at eight the four-wave version
// spills, left alone it takes 86 VGPRs and five): 2 % faster at N = 65536.
template <bool NN, int W>
__global__ __launch_bounds__(64 * W, 7) void pg6_force_ow(
    int ni, const double* __restrict__ ipos, const double* __restrict__ ivel,
    const int* __restrict__ iid, float eps2, int nj, int jper, JSort so, JPred jp, Partials part,
    Outputs out, int nsplit) {
  __shared__ double s_v[7][W][64];
  __shared__ float s_r[W][64];
  __shared__ int s_n[W][64];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i = blockIdx.x * 64 + lane;
  const bool ivalid = i < ni;
  const int ic = ivalid ? i : ni - 1;  // padding lanes repeat i = ni-1
  const double xi = ipos[3 * ic + 0];
  const double yi = ipos[3 * ic + 1];
  const double zi = ipos[3 * ic + 2];
  const int idi = iid[ic];
  const int self_k = (idi >= 0 && idi < nj) ? so.inv[idi] : -1;
  const int piece = jper / W;
  const int k0 = min(nj, blockIdx.y * jper + w * piece);
  const RangeSums r = pg6_range_o<NN>(xi, yi, zi, (float)ivel[3 * ic + 0],
                                      (float)ivel[3 * ic + 1], (float)ivel[3 * ic + 2], idi,
                                      self_k, eps2, k0, min(nj, k0 + piece), so);
  s_v[0][w][lane] = r.ax;
  s_v[1][w][lane] = r.ay;
  s_v[2][w][lane] = r.az;
  s_v[3][w][lane] = r.pot;
  s_v[4][w][lane] = r.jx;
  s_v[5][w][lane] = r.jy;
  s_v[6][w][lane] = r.jz;
  s_r[w][lane] = r.rmax;
  s_n[w][lane] = r.nn;
  __syncthreads();
  if (w != 0 || !ivalid) return;
  double v[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) v[k] = s_v[k][0][lane];
  float rmax = s_r[0][lane];
  int nn = s_n[0][lane];
#pragma unroll
  for (int g = 1; g < W; ++g) {
#pragma unroll
    for (int k = 0; k < 7; ++k) v[k] += s_v[k][g][lane];
    if (NN && nn_better(s_r[g][lane], s_n[g][lane], rmax, nn)) {
      rmax = s_r[g][lane];
      nn = s_n[g][lane];
    }
  }
  if (nsplit == 1) {
    out.acc[3 * i + 0] = v[0];
    out.acc[3 * i + 1] = v[1];
    out.acc[3 * i + 2] = v[2];
    out.jrk[3 * i + 0] = v[4];
    out.jrk[3 * i + 1] = v[5];
    out.jrk[3 * i + 2] = v[6];
    out.pot[i] = v[3];
    if (NN) {
      out.nnb[i] = nn >= 0 ? jp.id[nn] : -1;
      out.rnnb[i] = nn >= 0 ? nn_distance(jp, nn, xi, yi, zi, eps2) : 0.f;
    }
  } else {
    const size_t o = (size_t)blockIdx.y * ni + i;
    part.ax[o] = v[0];
    part.ay[o] = v[1];
    part.az[o] = v[2];
    part.pot[o] = v[3];
    part.jx[o] = v[4];
    part.jy[o] = v[5];
    part.jz[o] = v[6];
    if (NN) {
      part.rmax[o] = rmax;
      part.nn[o] = nn;
    }
  }
}

// Sums the nsplit partials of each i with a block of 256 threads per i:
// thread t takes splits t, t+256, ... in order (all its loads issued
// together), a fixed xor-butterfly combines each wavefront's lanes, and the
// four wavefront sums are added in wave order.  Deterministic.  For many
// splits and few i, where one thread per i would walk nsplit dependent loads
// and one wavefront per i still waited for several rounds of them.  NNP: the
// partials carry the neighbour's id and distance (pg6_force with in-kernel
// prediction), which then need no read of the predicted set.
template <bool NN, bool NNP = false>
__global__ __launch_bounds__(256) void pg6_reduce_w(int ni, int nsplit, Partials part, JPred jp,
                                                    const double* __restrict__ ipos, float eps2,
                                                    Outputs out) {
  __shared__ double s_v[7][4];
  __shared__ float s_r[4];
  __shared__ int s_n[4];
  __shared__ int s_id[4];
  __shared__ float s_d[4];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int i = blockIdx.x;
  double v[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  float rmax = 0.f;
  int nn = -1;
  int nid = -1;       // NNP: the neighbour's id and distance ride with it
  float rdist = 0.f;
#pragma unroll 4
  for (int s = threadIdx.x; s < nsplit; s += 256) {
    const size_t o = (size_t)s * ni + i;
    v[0] += part.ax[o];
    v[1] += part.ay[o];
    v[2] += part.az[o];
    v[3] += part.pot[o];
    v[4] += part.jx[o];
    v[5] += part.jy[o];
    v[6] += part.jz[o];
    if (NN) {
      const float r = part.rmax[o];
      const int n = part.nn[o];
      if (nn_better(r, n, rmax, nn)) {
        rmax = r;
        nn = n;
        if (NNP) {
          nid = part.nid[o];
          rdist = part.rdist[o];
        }
      }
    }
  }
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
#pragma unroll
    for (int k = 0; k < 7; ++k) v[k] += __shfl_xor(v[k], off, 64);
    if (NN) {
      const float r = __shfl_xor(rmax, off, 64);
      const int n = __shfl_xor(nn, off, 64);
      const int d = NNP ? __shfl_xor(nid, off, 64) : 0;
      const float e = NNP ? __shfl_xor(rdist, off, 64) : 0.f;
      if (nn_better(r, n, rmax, nn)) {
        rmax = r;
        nn = n;
        nid = d;
        rdist = e;
      }
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < 7; ++k) s_v[k][w] = v[k];
    s_r[w] = rmax;
    s_n[w] = nn;
    s_id[w] = nid;
    s_d[w] = rdist;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
#pragma unroll
  for (int k = 0; k < 7; ++k) v[k] = s_v[k][0];
  rmax = s_r[0];
  nn = s_n[0];
  nid = s_id[0];
  rdist = s_d[0];
  for (int g = 1; g < 4; ++g) {
#pragma unroll
    for (int k = 0; k < 7; ++k) v[k] += s_v[k][g];
    if (NN && nn_better(s_r[g], s_n[g], rmax, nn)) {
      rmax = s_r[g];
      nn = s_n[g];
      nid = s_id[g];
      rdist = s_d[g];
    }
  }
  out.acc[3 * i + 0] = v[0];
  out.acc[3 * i + 1] = v[1];
  out.acc[3 * i + 2] = v[2];
  out.jrk[3 * i + 0] = v[4];
  out.jrk[3 * i + 1] = v[5];
  out.jrk[3 * i + 2] = v[6];
  out.pot[i] = v[3];
  if (NN) {
    if (NNP) {
      out.nnb[i] = nn >= 0 ? nid : -1;
      out.rnnb[i] = nn >= 0 ? rdist : 0.f;
    } else {
      out.nnb[i] = nn >= 0 ? jp.id[nn] : -1;
      out.rnnb[i] = nn >= 0 ? nn_distance(jp, nn, ipos[3 * i], ipos[3 * i + 1], ipos[3 * i + 2],
                                          eps2)
                            : 0.f;
    }
  }
}

// Sums the nsplit partials of each i with four threads per i: thread q of i
// takes splits q, q+4, ... in order (consecutive threads hold consecutive i,
// so every load is coalesced), then thread 0 of each i adds the four sums in
// q order.  Deterministic.  For many i and many splits.  Block = 256 threads
// = 64 i.
template <bool NN>
__global__ __launch_bounds__(256) void pg6_reduce_q4(int ni, int nsplit, Partials part,
                                                     JPred jp, const double* __restrict__ ipos,
                                                     float eps2, Outputs out) {
  __shared__ double s_v[7][4][64];
  __shared__ float s_r[4][64];
  __shared__ int s_n[4][64];
  const int li = threadIdx.x & 63;
  const int q = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + li;
  double v[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  float rmax = 0.f;
  int nn = -1;
  if (i < ni) {
    for (int s = q; s < nsplit; s += 4) {
      const size_t o = (size_t)s * ni + i;
      v[0] += part.ax[o];
      v[1] += part.ay[o];
      v[2] += part.az[o];
      v[3] += part.pot[o];
      v[4] += part.jx[o];
      v[5] += part.jy[o];
      v[6] += part.jz[o];
      if (NN) {
        const float r = part.rmax[o];
        const int n = part.nn[o];
        if (nn_better(r, n, rmax, nn)) {
          rmax = r;
          nn = n;
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 7; ++k) s_v[k][q][li] = v[k];
  s_r[q][li] = rmax;
  s_n[q][li] = nn;
  __syncthreads();
  if (q != 0 || i >= ni) return;
  for (int t = 1; t < 4; ++t) {
#pragma unroll
    for (int k = 0; k < 7; ++k) v[k] += s_v[k][t][li];
    if (NN && nn_better(s_r[t][li], s_n[t][li], rmax, nn)) {
      rmax = s_r[t][li];
      nn = s_n[t][li];
    }
  }
  out.acc[3 * i + 0] = v[0];
  out.acc[3 * i + 1] = v[1];
  out.acc[3 * i + 2] = v[2];
  out.jrk[3 * i + 0] = v[4];
  out.jrk[3 * i + 1] = v[5];
  out.jrk[3 * i + 2] = v[6];
  out.pot[i] = v[3];
  if (NN) {
    out.nnb[i] = nn >= 0 ? jp.id[nn] : -1;
    out.rnnb[i] = nn >= 0 ? nn_distance(jp, nn, ipos[3 * i], ipos[3 * i + 1], ipos[3 * i + 2],
                                        eps2)
                          : 0.f;
  }
}

// Sums the nsplit partials of each i in split order (deterministic).
template <bool NN>
__global__ void pg6_reduce(int ni, int nsplit, Partials part, JPred jp,
                           const double* __restrict__ ipos, float eps2, Outputs out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ni) return;
  double ax = 0.0, ay = 0.0, az = 0.0, pot = 0.0, jx = 0.0, jy = 0.0, jz = 0.0;
  float rmax = 0.f;
  int nn = -1;
  for (int s = 0; s < nsplit; ++s) {
    const size_t o = (size_t)s * ni + i;
    ax += part.ax[o];
    ay += part.ay[o];
    az += part.az[o];
    pot += part.pot[o];
    jx += part.jx[o];
    jy += part.jy[o];
    jz += part.jz[o];
    if (NN) {
      // Splits cover increasing j addresses: strict '>' keeps the lowest address on ties.
      const float r = part.rmax[o];
      if (r > rmax) {
        rmax = r;
        nn = part.nn[o];
      }
    }
  }
  out.acc[3 * i + 0] = ax;
  out.acc[3 * i + 1] = ay;
  out.acc[3 * i + 2] = az;
  out.jrk[3 * i + 0] = jx;
  out.jrk[3 * i + 1] = jy;
  out.jrk[3 * i + 2] = jz;
  out.pot[i] = pot;
  if (NN) {
    out.nnb[i] = nn >= 0 ? jp.id[nn] : -1;
    out.rnnb[i] = nn >= 0 ? nn_distance(jp, nn, ipos[3 * i], ipos[3 * i + 1], ipos[3 * i + 2],
                                        eps2)
                          : 0.f;
  }
}

// ------------------------------------------------------------- small i-sets
// One launch per call for small i-sets (the block-step sizes of a Hermite
// code, PG6/banana/hermite.c:214-277): the predictor runs inside the force
// kernel on the block's own j-range, and the partial sums of the j-ranges
// are added in the same launch by the last block of each i-block to finish.
//
// Block = kFThreads = 8 wavefronts, 64 i-particles (one per lane), blockIdx.x
// = i-block, blockIdx.y = j-range [y jper, min(nj, (y + 1) jper)).  The range
// is staged in LDS tiles of kFTile two-float records (JPred::rec layout):
// with PRED, predicted from the j-state to ti (every JPred array is written
// too, so later calls at the same ti find the prediction); without, copied
// from JPred::rec.  Wave w takes tile slots w, w + 8, ... (FP32 sums per tile
// folded into FP64, as pg6_range_h), and the eight wave sums are added in
// wave order through LDS.  With nsplit > 1 each block writes its partials
// ([split][i], as pg6_force_h), releases them (device-scope fence) and counts
// itself done on counter[blockIdx.x]; the last block of the i-block acquires
// them and adds the nsplit partials of each i in a fixed order (eight threads
// per i take splits q, q + 8, ..., then the eight sums in q order, and the
// nearest neighbour by nn_better, a total order), writes the outputs and
// resets the counter to zero for the next launch.  Deterministic; the self
// pair is excluded by id on every pair.
constexpr int kFWaves = 8;
constexpr int kFThreads = 64 * kFWaves;
constexpr int kFTile = 256;

template <bool NN, bool PRED>
__global__ __launch_bounds__(kFThreads) void pg6_force_f(
    int ni, const double* __restrict__ ipos, const double* __restrict__ ivel,
    const int* __restrict__ iid, float eps2, int nj, int jper, double ti, JState js, JPred jp,
    Partials part, Outputs out, int nsplit, int* __restrict__ counter) {
  __shared__ float4 s_rec[kFTile][3];
  __shared__ double s_v[7][kFWaves][64];
  __shared__ float s_r[kFWaves][64];
  __shared__ int s_n[kFWaves][64];
  __shared__ int s_last;
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + lane;
  const bool ivalid = i < ni;
  const int ic = ivalid ? i : ni - 1;  // padding lanes repeat i = ni-1
  const double xi = ipos[3 * ic + 0], yi = ipos[3 * ic + 1], zi = ipos[3 * ic + 2];
  const float vxi = (float)ivel[3 * ic + 0], vyi = (float)ivel[3 * ic + 1],
              vzi = (float)ivel[3 * ic + 2];
  const int idi = iid[ic];
  const float xih = (float)xi, yih = (float)yi, zih = (float)zi;
  const float xil = (float)(xi - (double)xih);
  const float yil = (float)(yi - (double)yih);
  const float zil = (float)(zi - (double)zih);
  const int j0 = blockIdx.y * jper;
  const int j1 = min(nj, j0 + jper);
  RangeSums r{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.f, -1};
  const unsigned kmask = 0xFFFFFF00u;
  for (int jt = j0; jt < j1; jt += kFTile) {
    const int nt = min(kFTile, j1 - jt);
    __syncthreads();
    if (threadIdx.x < kFTile) {  // waves 0..3: one j each, 64-aligned chunks
      const int t = threadIdx.x;
      const bool v = t < nt;
      if (PRED) {
        float4 q[3];
        pg6_predict_j(jt + t, v, ti, js, jp, q);
        if (v) {
          s_rec[t][0] = q[0];
          s_rec[t][1] = q[1];
          s_rec[t][2] = q[2];
        }
      } else if (v) {
        const float4* q = jp.rec + 3 * (size_t)(jt + t);
        s_rec[t][0] = q[0];
        s_rec[t][1] = q[1];
        s_rec[t][2] = q[2];
      }
    }
    __syncthreads();
    TileAcc ta{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0u};
#pragma unroll 2
    for (int jj = w; jj < nt; jj += kFWaves) {
      const float4 a = s_rec[jj][0], b = s_rec[jj][1], c = s_rec[jj][2];
      const float dx = (a.x - xih) + (b.x - xil);
      const float dy = (a.y - yih) + (b.y - yil);
      const float dz = (a.z - zih) + (b.z - zil);
      pg6_pair_h<true, NN>(ta, dx, dy, dz, b.w - vxi, c.x - vyi, c.y - vzi, a.w, eps2,
                           __float_as_int(c.z) == idi, (unsigned)(255 - jj));
    }
    if (NN) {
      const float rq = __uint_as_float(ta.key & kmask);
      if (rq > r.rmax) {
        r.rmax = rq;
        r.nn = jt + 255 - (int)(ta.key & 0xFFu);
      }
    }
    r.ax += (double)ta.ax;
    r.ay += (double)ta.ay;
    r.az += (double)ta.az;
    r.pot += (double)ta.pot;
    r.jx += (double)ta.jx;
    r.jy += (double)ta.jy;
    r.jz += (double)ta.jz;
  }
  // the eight wave sums of each i, in wave order
  s_v[0][w][lane] = r.ax;
  s_v[1][w][lane] = r.ay;
  s_v[2][w][lane] = r.az;
  s_v[3][w][lane] = r.pot;
  s_v[4][w][lane] = r.jx;
  s_v[5][w][lane] = r.jy;
  s_v[6][w][lane] = r.jz;
  s_r[w][lane] = r.rmax;
  s_n[w][lane] = r.nn;
  __syncthreads();
  if (w == 0) {
    double v[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) v[k] = s_v[k][0][lane];
    float rmax = s_r[0][lane];
    int nn = s_n[0][lane];
    for (int g = 1; g < kFWaves; ++g) {
#pragma unroll
      for (int k = 0; k < 7; ++k) v[k] += s_v[k][g][lane];
      if (NN && nn_better(s_r[g][lane], s_n[g][lane], rmax, nn)) {
        rmax = s_r[g][lane];
        nn = s_n[g][lane];
      }
    }
    if (ivalid) {
      if (nsplit == 1) {
        out.acc[3 * i + 0] = v[0];
        out.acc[3 * i + 1] = v[1];
        out.acc[3 * i + 2] = v[2];
        out.jrk[3 * i + 0] = v[4];
        out.jrk[3 * i + 1] = v[5];
        out.jrk[3 * i + 2] = v[6];
        out.pot[i] = v[3];
        if (NN) {
          out.nnb[i] = nn >= 0 ? jp.id[nn] : -1;
          out.rnnb[i] = nn >= 0 ? nn_distance(jp, nn, xi, yi, zi, eps2) : 0.f;
        }
      } else {
        const size_t o = (size_t)blockIdx.y * ni + i;
        part.ax[o] = v[0];
        part.ay[o] = v[1];
        part.az[o] = v[2];
        part.pot[o] = v[3];
        part.jx[o] = v[4];
        part.jy[o] = v[5];
        part.jz[o] = v[6];
        if (NN) {
          part.rmax[o] = rmax;
          part.nn[o] = nn;
        }
      }
    }
  }
  if (nsplit == 1) return;
  // Release this block's partials (and, with PRED, its predicted j) at agent
  // scope once the block is done with them, then count it done; the last
  // block of the i-block acquires everything.  One release per block and one
  // acquire in the last block only: on this multi-L2 part an agent-scope
  // release writes back the L2 and an acquire invalidates it, so a full fence
  // in every block costs far more than the pair work of a small call.
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    s_last = atomicAdd(&counter[blockIdx.x], 1) == nsplit - 1;
  }
  __syncthreads();
  if (!s_last) return;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  const int q = w;  // thread (q, lane): i = lane of this i-block, splits q, q + 8, ...
  double v[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  float rmax = 0.f;
  int nn = -1;
  if (ivalid) {
    for (int sp = q; sp < nsplit; sp += kFWaves) {
      const size_t o = (size_t)sp * ni + i;
      v[0] += part.ax[o];
      v[1] += part.ay[o];
      v[2] += part.az[o];
      v[3] += part.pot[o];
      v[4] += part.jx[o];
      v[5] += part.jy[o];
      v[6] += part.jz[o];
      if (NN && nn_better(part.rmax[o], part.nn[o], rmax, nn)) {
        rmax = part.rmax[o];
        nn = part.nn[o];
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 7; ++k) s_v[k][q][lane] = v[k];
  s_r[q][lane] = rmax;
  s_n[q][lane] = nn;
  __syncthreads();
  if (q == 0) {
    for (int g = 1; g < kFWaves; ++g) {
#pragma unroll
      for (int k = 0; k < 7; ++k) v[k] += s_v[k][g][lane];
      if (NN && nn_better(s_r[g][lane], s_n[g][lane], rmax, nn)) {
        rmax = s_r[g][lane];
        nn = s_n[g][lane];
      }
    }
    if (ivalid) {
      out.acc[3 * i + 0] = v[0];
      out.acc[3 * i + 1] = v[1];
      out.acc[3 * i + 2] = v[2];
      out.jrk[3 * i + 0] = v[4];
      out.jrk[3 * i + 1] = v[5];
      out.jrk[3 * i + 2] = v[6];
      out.pot[i] = v[3];
      if (NN) {
        out.nnb[i] = nn >= 0 ? jp.id[nn] : -1;
        out.rnnb[i] = nn >= 0 ? nn_distance(jp, nn, xi, yi, zi, eps2) : 0.f;
      }
    }
    if (lane == 0) counter[blockIdx.x] = 0;
  }
}

}  // namespace pg6
