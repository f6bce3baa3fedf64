// GRAPE-5 device kernels for gfx950: softened Newtonian force + potential
// (PG5/newton/libpg5/gravity_avx2.c:28-173, symmetric softening :178-250) and
// the monopole + traceless-quadrupole cell force
// (PG5/newton_quad/quad/quad_avx2.c:8-163).
//
// Layout in HBM: j-particles as float4 {x, y, z, m} (16 B, one coalesced load
// per lane), optional eps_j^2 as a separate float array, cells as three
// float4 {x, y, z, M}, {Qxx, Qxy, Qxz, Qyy}, {Qyz, Qzz, 0, 0} (48 B);
// i-particles as float4 {x, y, z, eps_i^2}.
//
// Arithmetic: FP32 pair math as in the reference (which converts positions to
// FP32 in g5_set_xmj / g5_set_xi, phantom_g5.c:100-128,188-205), with the
// hardware v_rsq_f32 (about 1 ulp) where the reference uses the 12-bit
// rsqrtps estimate; per-tile FP32 partial sums folded into FP64.
// Outputs before the API's sign convention: acc = sum m dx / r^3 and
// pot = + sum m / r (self pair included), or for cells acc = sum
// (phi_p + 5 phi_q) dx / r^2 - Q.dx / r^5 and pot = + sum (phi_p + phi_q).
#pragma once
#include <hip/hip_runtime.h>

namespace pg5 {

constexpr int kThreads = 256;
constexpr int kJTile = 256;

struct Partials5 {
  double *ax, *ay, *az, *pot;
};
struct Outputs5 {
  double* acc;  // [ni][3]
  double* pot;  // [ni]
};

struct Cell {
  float4 xm;  // COM x, y, z and mass
  float4 qa;  // Qxx, Qxy, Qxz, Qyy
  float4 qb;  // Qyz, Qzz, 0, 0
};

__device__ inline void pg5_store(int i, int nsplit, int ni, double ax, double ay, double az,
                                 double pot, const Partials5& part, const Outputs5& out) {
  if (nsplit == 1) {
    out.acc[3 * i + 0] = ax;
    out.acc[3 * i + 1] = ay;
    out.acc[3 * i + 2] = az;
    out.pot[i] = pot;
  } else {
    const size_t o = (size_t)blockIdx.y * ni + i;
    part.ax[o] = ax;
    part.ay[o] = ay;
    part.az[o] = az;
    part.pot[o] = pot;
  }
}

// Combine the G wave groups of a block in fixed order (BI < 256 only).
template <int BI>
__device__ inline void pg5_group_reduce(double& ax, double& ay, double& az, double& pot) {
  constexpr int G = kThreads / BI;
  if (G == 1) return;
  __shared__ double r[4][kThreads];
  const int tid = threadIdx.x;
  __syncthreads();
  r[0][tid] = ax; r[1][tid] = ay; r[2][tid] = az; r[3][tid] = pot;
  __syncthreads();
  if (tid < BI) {
    for (int q = 1; q < G; ++q) {
      const int t = q * BI + tid;
      ax += r[0][t]; ay += r[1][t]; az += r[2][t]; pot += r[3][t];
    }
  }
}

// Particle force + potential.  Block: 256 threads, BI i-particles, G = 256/BI
// wave groups sharing each LDS j-tile (group g takes j = g, g+G, ...);
// blockIdx.y selects the j-range [y*jper, min(nj, (y+1)*jper)).
template <int BI, bool SYM>
__global__ __launch_bounds__(kThreads) void pg5_force(
    int ni, const float4* __restrict__ xi, int nj, int jper, const float4* __restrict__ xj,
    const float* __restrict__ ej2, Partials5 part, Outputs5 out, int nsplit) {
  constexpr int G = kThreads / BI;
  __shared__ float4 s_j[kJTile];
  __shared__ float s_e[SYM ? kJTile : 1];
  const int tid = threadIdx.x;
  const int il = tid % BI;
  const int g = __builtin_amdgcn_readfirstlane(tid / BI);  // uniform: BI is a multiple of 64
  const int i = blockIdx.x * BI + il;
  const bool ivalid = i < ni;
  const float4 ip = xi[ivalid ? i : ni - 1];
  const int j0 = blockIdx.y * jper;
  const int j1 = min(nj, j0 + jper);
  double ax = 0.0, ay = 0.0, az = 0.0, pot = 0.0;
  for (int jt = j0; jt < j1; jt += kJTile) {
    const int nt = min(kJTile, j1 - jt);
    __syncthreads();
    if (tid < nt) {
      s_j[tid] = xj[jt + tid];
      if (SYM) s_e[tid] = ej2[jt + tid];
    }
    __syncthreads();
    float fax = 0.f, fay = 0.f, faz = 0.f, fpot = 0.f;
#pragma unroll 8
    for (int jj = g; jj < nt; jj += G) {
      const float4 q = s_j[jj];
      const float dx = q.x - ip.x;
      const float dy = q.y - ip.y;
      const float dz = q.z - ip.z;
      const float e2 = SYM ? ip.w + s_e[jj] : ip.w;
      const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, e2)));
      const float rinv = __builtin_amdgcn_rsqf(r2);
      const float mr = q.w * rinv;
      const float mr3 = mr * rinv * rinv;
      fpot += mr;
      fax = fmaf(mr3, dx, fax);
      fay = fmaf(mr3, dy, fay);
      faz = fmaf(mr3, dz, faz);
    }
    ax += (double)fax;
    ay += (double)fay;
    az += (double)faz;
    pot += (double)fpot;
  }
  pg5_group_reduce<BI>(ax, ay, az, pot);
  if (g != 0 || !ivalid) return;
  pg5_store(i, nsplit, ni, ax, ay, az, pot, part, out);
}

// Variant of pg5_force that broadcasts the j-particles through scalar
// registers instead of LDS: every wave of a block walks the same j sequence,
// so each j is wave-uniform; read through the constant address space it
// arrives by scalar loads and the VALU instructions take it as an operand.
// With ~14 VALU instructions per pair the LDS variant issues one LDS read
// per pair per wave, enough to make LDS issue the limit; this one has none.
// Block = 64, 128 or 256 threads, one i-particle per lane.  Same pair
// arithmetic as pg5_force; FP32 sums per 256 j folded into FP64.
typedef const __attribute__((address_space(4))) float cfloat5;

template <bool SYM>
__global__ __launch_bounds__(kThreads) void pg5_force_s(
    int ni, const float4* __restrict__ xi, int nj, int jper, const float4* __restrict__ xj,
    const float* __restrict__ ej2, Partials5 part, Outputs5 out, int nsplit) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool ivalid = i < ni;
  const float4 ip = xi[ivalid ? i : ni - 1];
  cfloat5* pj = (cfloat5*)xj;
  cfloat5* pe = (cfloat5*)ej2;
  const int j0 = blockIdx.y * jper;
  const int j1 = min(nj, j0 + jper);
  double ax = 0.0, ay = 0.0, az = 0.0, pot = 0.0;
  for (int jt = j0; jt < j1; jt += kJTile) {
    const int nt = min(kJTile, j1 - jt);
    float fax = 0.f, fay = 0.f, faz = 0.f, fpot = 0.f;
#pragma unroll 8
    for (int jj = 0; jj < nt; ++jj) {
      const int j = jt + jj;
      const float dx = pj[4 * j + 0] - ip.x;
      const float dy = pj[4 * j + 1] - ip.y;
      const float dz = pj[4 * j + 2] - ip.z;
      const float mj = pj[4 * j + 3];
      const float e2 = SYM ? ip.w + pe[j] : ip.w;
      const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, e2)));
      const float rinv = __builtin_amdgcn_rsqf(r2);
      const float mr = mj * rinv;
      const float mr3 = mr * rinv * rinv;
      fpot += mr;
      fax = fmaf(mr3, dx, fax);
      fay = fmaf(mr3, dy, fay);
      faz = fmaf(mr3, dz, faz);
    }
    ax += (double)fax;
    ay += (double)fay;
    az += (double)faz;
    pot += (double)fpot;
  }
  if (!ivalid) return;
  pg5_store(i, nsplit, ni, ax, ay, az, pot, part, out);
}

// Monopole + quadrupole cell force (quad_avx2.c:35-149):
//   phi_p = M / r, phi_q = (dx.Q.dx) / (2 r^5),
//   acc  += (phi_p + 5 phi_q) dx / r^2 - Q.dx / r^5,  pot += phi_p + phi_q.
template <int BI>
__global__ __launch_bounds__(kThreads) void pg5_quad(
    int ni, const float4* __restrict__ xi, int nc, int jper, const Cell* __restrict__ cells,
    Partials5 part, Outputs5 out, int nsplit) {
  constexpr int G = kThreads / BI;
  __shared__ Cell s_c[kJTile];
  const int tid = threadIdx.x;
  const int il = tid % BI;
  const int g = __builtin_amdgcn_readfirstlane(tid / BI);  // uniform: BI is a multiple of 64
  const int i = blockIdx.x * BI + il;
  const bool ivalid = i < ni;
  const float4 ip = xi[ivalid ? i : ni - 1];
  const int j0 = blockIdx.y * jper;
  const int j1 = min(nc, j0 + jper);
  double ax = 0.0, ay = 0.0, az = 0.0, pot = 0.0;
  for (int jt = j0; jt < j1; jt += kJTile) {
    const int nt = min(kJTile, j1 - jt);
    __syncthreads();
    if (tid < nt) s_c[tid] = cells[jt + tid];
    __syncthreads();
    float fax = 0.f, fay = 0.f, faz = 0.f, fpot = 0.f;
#pragma unroll 4
    for (int jj = g; jj < nt; jj += G) {
      const Cell c = s_c[jj];
      const float dx = c.xm.x - ip.x;
      const float dy = c.xm.y - ip.y;
      const float dz = c.xm.z - ip.z;
      const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, ip.w)));
      const float rinv = __builtin_amdgcn_rsqf(r2);
      const float rinv2 = rinv * rinv;
      const float phip = c.xm.w * rinv;
      const float qx = fmaf(c.qa.z, dz, fmaf(c.qa.y, dy, c.qa.x * dx));
      const float qy = fmaf(c.qb.x, dz, fmaf(c.qa.w, dy, c.qa.y * dx));
      const float qz = fmaf(c.qb.y, dz, fmaf(c.qb.x, dy, c.qa.z * dx));
      const float dqd = fmaf(dz, qz, fmaf(dy, qy, dx * qx));
      const float rinv5 = rinv2 * rinv2 * rinv;
      const float phiq = 0.5f * dqd * rinv5;
      fpot += phip + phiq;
      const float f = fmaf(5.f, phiq, phip) * rinv2;
      fax = fmaf(f, dx, fmaf(-rinv5, qx, fax));
      fay = fmaf(f, dy, fmaf(-rinv5, qy, fay));
      faz = fmaf(f, dz, fmaf(-rinv5, qz, faz));
    }
    ax += (double)fax;
    ay += (double)fay;
    az += (double)faz;
    pot += (double)fpot;
  }
  pg5_group_reduce<BI>(ax, ay, az, pot);
  if (g != 0 || !ivalid) return;
  pg5_store(i, nsplit, ni, ax, ay, az, pot, part, out);
}

// Sums the nsplit partials of each i: four threads per i, thread q taking
// splits q, q + 4, ... in order (a wavefront holds 64 consecutive i of one q,
// so every load is coalesced), then the four sums added in q order.
// Deterministic.  Block = 256 threads = 64 i.
__global__ __launch_bounds__(256) void pg5_reduce(int ni, int nsplit, Partials5 part,
                                                  Outputs5 out) {
  __shared__ double s[4][4][64];
  const int il = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + il;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  if (i < ni)
    for (int sp = q; sp < nsplit; sp += 4) {
      const size_t o = (size_t)sp * ni + i;
      v[0] += part.ax[o];
      v[1] += part.ay[o];
      v[2] += part.az[o];
      v[3] += part.pot[o];
    }
  for (int c = 0; c < 4; ++c) s[c][q][il] = v[c];
  __syncthreads();
  if (q != 0 || i >= ni) return;
  double t[4];
  for (int c = 0; c < 4; ++c) t[c] = ((s[c][0][il] + s[c][1][il]) + s[c][2][il]) + s[c][3][il];
  out.acc[3 * i + 0] = t[0];
  out.acc[3 * i + 1] = t[1];
  out.acc[3 * i + 2] = t[2];
  out.pot[i] = t[3];
}

// Bias of the reciprocal square root the force kernels use (v_rsq_f32),
// measured as PG5/newton/libpg5/rsqrt.c:42-81 measures the reference's
// rsqrtss: over x0 = the 2^23 floats in [1, 2) (bits (127 << 23) + i) and
// x1 = 2 x0, e = y^2 x - 1, w = y0^2; each block sums sw = sum w and
// se = sum w (e0 + e1) over its i in a fixed order (FP64, contraction off),
// part[2 b] = sw, part[2 b + 1] = se.  Grid: kBiasBlocks blocks of 256.
constexpr int kBiasBlocks = 2048;
__global__ __launch_bounds__(256) void rsqrt_bias_sums(double* __restrict__ part) {
#pragma clang fp contract(off)
  __shared__ double s_w[256], s_e[256];
  constexpr int per = (1 << 23) / (kBiasBlocks * 256);
  double sw = 0.0, se = 0.0;
  for (int k = 0; k < per; ++k) {
    const int i = (blockIdx.x * per + k) * 256 + threadIdx.x;
    const float x0f = __int_as_float((127 << 23) + i);
    const float x1f = __int_as_float((128 << 23) + i);
    const double x0 = (double)x0f, x1 = (double)x1f;
    const double y0 = (double)__builtin_amdgcn_rsqf(x0f);
    const double y1 = (double)__builtin_amdgcn_rsqf(x1f);
    const double e0 = y0 * y0 * x0 - 1.0, e1 = y1 * y1 * x1 - 1.0;
    const double w = y0 * y0;
    sw += w;
    se += w * (e0 + e1);
  }
  s_w[threadIdx.x] = sw;
  s_e[threadIdx.x] = se;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      s_w[threadIdx.x] += s_w[threadIdx.x + h];
      s_e[threadIdx.x] += s_e[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s_w[0];
    part[2 * blockIdx.x + 1] = s_e[0];
  }
}

}  // namespace pg5
