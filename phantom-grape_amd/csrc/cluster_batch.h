// Batched many-cluster Hermite force: self-gravity acc, jerk and pot of many
// independent clusters in one persistent launch.  No counterpart in the
// reference; it replaces one GRAPE-6 board session per cluster (g6_set_j_
// particle for its members, then g6calc_firsthalf/lasthalf over chunks of
// g6_npipes() i-particles, PG6/libpg6/phantom_g6.c:118-211).
//
// Inputs (device): predicted positions pos[n][3] and velocities vel[n][3]
// (FP64), mass[n] (FP64), cluster c = particles [coff[c], coff[c+1]),
// eps2[c] per cluster, and work items (cluster, first i within it): a block
// takes an item from the counter and computes up to 256 i of that cluster
// against all of its members, staged TILE j at a time in LDS (TILE = 128,
// 256 or 512, chosen by the autotuner).  Arithmetic as pg6_force: FP64 dx
// rounded to FP32, FP32 pair math with v_rsq_f32, FP32 sums over at most 256
// j folded into FP64 (tiles of 256 and 512 give identical results); self
// pair excluded by index.  Output (FP64):
// acc[i][3], jrk[i][3], pot[i] = - sum m/r.
#pragma once
#include <hip/hip_runtime.h>

namespace pgc {

constexpr int kThreads = 256;  // block = i-particles per work item
constexpr int kTile = 256;     // default j-tile (128 and 512 are also built)

__device__ inline int take_item(int* counter, int* s_item) {
  __syncthreads();
  if (threadIdx.x == 0) *s_item = atomicAdd(counter, 1);
  __syncthreads();
  return *s_item;
}

template <int TILE>
__global__ __launch_bounds__(kThreads) void cluster_force(
    int nitems, const int2* __restrict__ items, const int* __restrict__ coff,
    const double* __restrict__ pos, const double* __restrict__ vel,
    const double* __restrict__ mass, const float* __restrict__ eps2,
    double* __restrict__ acc, double* __restrict__ jrk, double* __restrict__ pot, int* counter) {
  constexpr int kFold = TILE < 256 ? TILE : 256;
  __shared__ double s_x[TILE], s_y[TILE], s_z[TILE];
  __shared__ float4 s_vm[TILE];
  __shared__ double s_red[7][kThreads];
  __shared__ int s_item;
  const int tid = threadIdx.x;
  for (;;) {
    const int it = take_item(counter, &s_item);
    if (it >= nitems) break;
    const int2 item = items[it];
    const int c = item.x;
    const int start = coff[c];
    const int n = coff[c + 1] - start;
    const int ib = item.y;
    const int nib = min(kThreads, n - ib);
    if (nib <= 0) continue;  // uniform: every thread read the same item
    const float e2 = eps2[c];
    const int bi = nib <= 64 ? 64 : (nib <= 128 ? 128 : 256);
    const int G = kThreads / bi;
    const int il = tid % bi;
    const int g = __builtin_amdgcn_readfirstlane(tid / bi);  // uniform per wave
    const bool iv = il < nib;
    const int i = start + ib + (iv ? il : 0);
    const double xi = pos[3 * (size_t)i], yi = pos[3 * (size_t)i + 1], zi = pos[3 * (size_t)i + 2];
    const float vxi = (float)vel[3 * (size_t)i];
    const float vyi = (float)vel[3 * (size_t)i + 1];
    const float vzi = (float)vel[3 * (size_t)i + 2];
    double ax = 0.0, ay = 0.0, az = 0.0, pp = 0.0, jx = 0.0, jy = 0.0, jz = 0.0;
    for (int jt = 0; jt < n; jt += TILE) {
      const int nt = min(TILE, n - jt);
      __syncthreads();
      for (int k = tid; k < nt; k += kThreads) {
        const size_t j = (size_t)start + jt + k;
        s_x[k] = pos[3 * j];
        s_y[k] = pos[3 * j + 1];
        s_z[k] = pos[3 * j + 2];
        s_vm[k] = make_float4((float)vel[3 * j], (float)vel[3 * j + 1], (float)vel[3 * j + 2],
                              (float)mass[j]);
      }
      __syncthreads();
      const int self = i - start - jt;  // tile slot of the i-particle itself
      // FP32 sums over at most kFold j, then folded into FP64: a tile of 512
      // sums exactly as two tiles of 256
      for (int sub = 0; sub < nt; sub += kFold) {
        const int ne = min(sub + kFold, nt);
        float fax = 0.f, fay = 0.f, faz = 0.f, fp = 0.f, fjx = 0.f, fjy = 0.f, fjz = 0.f;
#pragma unroll 4
        for (int jj = sub + g; jj < ne; jj += G) {
          const float dx = (float)(s_x[jj] - xi);
          const float dy = (float)(s_y[jj] - yi);
          const float dz = (float)(s_z[jj] - zi);
          const float4 vm = s_vm[jj];
          const float dvx = vm.x - vxi;
          const float dvy = vm.y - vyi;
          const float dvz = vm.z - vzi;
          const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, e2)));
          const float rv = fmaf(dz, dvz, fmaf(dy, dvy, dx * dvx));
          const float rinv = jj == self ? 0.f : __builtin_amdgcn_rsqf(r2);
          const float rinv2 = rinv * rinv;
          const float mr = vm.w * rinv;
          const float mr3 = mr * rinv2;
          const float beta = -3.f * rv * rinv2;
          fp -= mr;
          fax = fmaf(mr3, dx, fax);
          fay = fmaf(mr3, dy, fay);
          faz = fmaf(mr3, dz, faz);
          fjx = fmaf(mr3, fmaf(beta, dx, dvx), fjx);
          fjy = fmaf(mr3, fmaf(beta, dy, dvy), fjy);
          fjz = fmaf(mr3, fmaf(beta, dz, dvz), fjz);
        }
        ax += (double)fax;
        ay += (double)fay;
        az += (double)faz;
        pp += (double)fp;
        jx += (double)fjx;
        jy += (double)fjy;
        jz += (double)fjz;
      }
    }
    if (G > 1) {
      __syncthreads();
      s_red[0][tid] = ax; s_red[1][tid] = ay; s_red[2][tid] = az; s_red[3][tid] = pp;
      s_red[4][tid] = jx; s_red[5][tid] = jy; s_red[6][tid] = jz;
      __syncthreads();
      if (g == 0)
        for (int q = 1; q < G; ++q) {
          const int t = q * bi + il;
          ax += s_red[0][t]; ay += s_red[1][t]; az += s_red[2][t]; pp += s_red[3][t];
          jx += s_red[4][t]; jy += s_red[5][t]; jz += s_red[6][t];
        }
    }
    if (g == 0 && iv) {
      acc[3 * (size_t)i + 0] = ax;
      acc[3 * (size_t)i + 1] = ay;
      acc[3 * (size_t)i + 2] = az;
      jrk[3 * (size_t)i + 0] = jx;
      jrk[3 * (size_t)i + 1] = jy;
      jrk[3 * (size_t)i + 2] = jz;
      pot[i] = pp;
    }
  }
}

__global__ void zero_counter(int* counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *counter = 0;
}

}  // namespace pgc
