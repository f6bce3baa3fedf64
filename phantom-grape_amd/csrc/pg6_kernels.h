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

// Predicts j-particle j (valid = j < nj) to time ti into every JPred array.
// Called by whole wavefronts whose lane l holds j = 64 w + l, so that the
// wavefront fills in the id range of j-chunk w.
__device__ inline void pg6_predict_j(int j, bool valid, double ti, const JState& js,
                                     const JPred& jp) {
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
    r[0] = make_float4(hx, hy, hz, vm.w);
    // x - (double)hx is exact in FP64
    r[1] = make_float4((float)(x - (double)hx), (float)(y - (double)hy), (float)(z - (double)hz),
                       vm.x);
    r[2] = make_float4(vm.y, vm.z, __int_as_float(id), vm.w);
  }
  int lo = valid ? id : INT_MAX, hi = valid ? id : INT_MIN;
  for (int o = 32; o > 0; o >>= 1) {
    lo = min(lo, __shfl_xor(lo, o));
    hi = max(hi, __shfl_xor(hi, o));
  }
  if ((j % kChunk) == 0 && valid) jp.crange[j / kChunk] = make_int2(lo, hi);
}

// Block = 256 threads, one j per thread.
__global__ __launch_bounds__(256) void pg6_predict(int nj, double ti, JState js, JPred jp) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  pg6_predict_j(j, j < nj, ti, js, jp);
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
// Block = 256 threads.  BI i-particles per block (64, 128 or 256); the block's
// G = 256/BI wave groups split each LDS tile's j-particles between them
// (group g takes j = g, g+G, ...), and are combined in fixed order at the end.
// blockIdx.y selects the j-range [y*jper, min(nj,(y+1)*jper)).
// With nsplit == 1 the kernel writes final outputs, otherwise partials.
template <int BI, bool NN>
__global__ __launch_bounds__(kThreads) void pg6_force(
    int ni, const double* __restrict__ ipos, const double* __restrict__ ivel,
    const int* __restrict__ iid, float eps2, int nj, int jper, JPred jp,
    Partials part, Outputs out, int nsplit) {
  constexpr int G = kThreads / BI;
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
  const unsigned kmask = 0xFFFFFF00u;

  for (int jt = j0; jt < j1; jt += kJTile) {
    const int nt = min(kJTile, j1 - jt);
    __syncthreads();
    if (tid < nt) {
      const int j = jt + tid;
      s_x[tid] = jp.x[j];
      s_y[tid] = jp.y[j];
      s_z[tid] = jp.z[j];
      s_vm[tid] = jp.vm[j];
      s_id[tid] = jp.id[j];
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

  // Combine the G wave groups in fixed order g = 0..G-1.
  if (G > 1) {
    __shared__ double r_d[7][kThreads];
    __shared__ float r_f[kThreads];
    __shared__ int r_n[kThreads];
    __syncthreads();
    r_d[0][tid] = ax; r_d[1][tid] = ay; r_d[2][tid] = az; r_d[3][tid] = pot;
    r_d[4][tid] = jx; r_d[5][tid] = jy; r_d[6][tid] = jz;
    r_f[tid] = rmax;
    r_n[tid] = nn;
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

// Nearest-neighbour candidate order: larger (quantised) 1/r wins, equal
// values go to the lower j address; (0, -1) is "none".  The order is total,
// so any reduction tree gives the same winner.
__device__ inline bool nn_better(float r, int n, float rbest, int nbest) {
  return r > rbest || (r == rbest && n >= 0 && (nbest < 0 || n < nbest));
}

// Sums the nsplit partials of each i with one wavefront per i: lane l takes
// splits l, l+64, ... in order, then a fixed xor-butterfly combines the
// lanes (every lane ends with the same sums, so the result does not depend
// on which lane writes).  For many splits and few i, where one thread per i
// would walk nsplit dependent loads.  Block = 256 threads = 4 i.
template <bool NN>
__global__ __launch_bounds__(256) void pg6_reduce_w(int ni, int nsplit, Partials part, JPred jp,
                                                    const double* __restrict__ ipos, float eps2,
                                                    Outputs out) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= ni) return;  // uniform per wavefront
  double v[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  float rmax = 0.f;
  int nn = -1;
  for (int s = lane; s < nsplit; s += 64) {
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
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
#pragma unroll
    for (int k = 0; k < 7; ++k) v[k] += __shfl_xor(v[k], off, 64);
    if (NN) {
      const float r = __shfl_xor(rmax, off, 64);
      const int n = __shfl_xor(nn, off, 64);
      if (nn_better(r, n, rmax, nn)) {
        rmax = r;
        nn = n;
      }
    }
  }
  if (lane != 0) return;
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

}  // namespace pg6
