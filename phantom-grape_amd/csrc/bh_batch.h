// Batched Barnes-Hut tree-list force: every i-group with its particle j-list
// (monopole, the GRAPE-5 kernel) and its cell list (monopole + traceless
// quadrupole, the quad_tree kernel) in one persistent launch.  No counterpart
// in the reference, whose caller makes two calls per group inside an OpenMP
// loop: g5_set_xmjMC + g5_calculate_force_on_xMC for the particles and
// g5c_set_xmjMC + g5c_calculate_force_on_xMC for the cells
// (PG5/newton_quad/source.cpp:813-867).
//
// Inputs (device): pos[np] float4 {x, y, z, m}; cells[nc] (pg5::Cell, 48 B);
// per group g: i-particles ilist[ioff[g] .. +icnt[g]), particle j-list
// plist[poff[g] .. +pcnt[g]), cell list clist[coff[g] .. +ccnt[g]); optional
// processing order; softening eps^2; a work counter that is 0 at launch.
// Output for every listed i: acc[i][3] = particle + cell acceleration and
// pot[i] = + sum m/r (particles, self included) + sum (phi_p + phi_q)
// (cells); the caller's potential is -pot + m_i/eps (source.cpp:850-867).
#pragma once
#include <hip/hip_runtime.h>

#include "pg5_kernels.h"

namespace pgbh {

constexpr int kThreads = 1024;  // default block (256 and 512 are also built)

__device__ inline int take_item(int* counter, int* s_item) {
  __syncthreads();
  if (threadIdx.x == 0) *s_item = atomicAdd(counter, 1);
  __syncthreads();
  return *s_item;
}

// Lanes are packed by phases: a group of n <= BLOCK i-particles takes
// p = BLOCK / n copies of the i-set; lane t holds i-particle t % n and
// walks the particles and cells of each LDS tile with jj = t / n, t / n + p,
// ...; the p phase sums of each i are added in phase order at the end.
// Groups larger than BLOCK run in passes.  An LDS tile holds BLOCK particles
// or BLOCK / 2 cells.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void bh_batch(
    int ngroups, const int* __restrict__ order, const int* __restrict__ ioff,
    const int* __restrict__ icnt, const int* __restrict__ ilist, const int* __restrict__ poff,
    const int* __restrict__ pcnt, const int* __restrict__ plist, const int* __restrict__ coff,
    const int* __restrict__ ccnt, const int* __restrict__ clist, const float4* __restrict__ pos,
    const pg5::Cell* __restrict__ cells, float eps2, double* __restrict__ acc,
    double* __restrict__ pot, int* counter) {
  constexpr int kPTile = BLOCK;      // particles staged per LDS tile
  constexpr int kCTile = BLOCK / 2;  // cells staged per LDS tile
  // one LDS buffer holds either a particle tile or a cell tile
  constexpr size_t kStage = sizeof(float4) * kPTile > sizeof(pg5::Cell) * kCTile
                                ? sizeof(float4) * kPTile
                                : sizeof(pg5::Cell) * kCTile;
  __shared__ __align__(16) char s_stage[kStage];
  float4* s_p = reinterpret_cast<float4*>(s_stage);
  pg5::Cell* s_c = reinterpret_cast<pg5::Cell*>(s_stage);
  __shared__ double s_red[4][BLOCK];
  __shared__ int s_item;
  const int tid = threadIdx.x;
  for (;;) {
    const int it = take_item(counter, &s_item);
    if (it >= ngroups) break;
    const int grp = order ? order[it] : it;
    const int i0 = ioff[grp], ni = icnt[grp];
    const int p0 = poff[grp], np = pcnt[grp];
    const int c0 = coff[grp], nc = ccnt[grp];
    for (int ib = 0; ib < ni; ib += BLOCK) {
      const int nib = min(BLOCK, ni - ib);
      const int nph = BLOCK / nib;  // phases
      const int il = tid % nib;
      const int phase = tid / nib;
      const bool live = phase < nph;
      const int ii = ilist[i0 + ib + il];
      const float4 ip = pos[ii];
      double ax = 0.0, ay = 0.0, az = 0.0, pp = 0.0;
      // particles: the GRAPE-5 pair force (gravity_avx2.c:95-154)
      for (int jt = 0; jt < np; jt += kPTile) {
        const int nt = min(kPTile, np - jt);
        __syncthreads();
        if (tid < nt) s_p[tid] = pos[plist[p0 + jt + tid]];
        __syncthreads();
        float fax = 0.f, fay = 0.f, faz = 0.f, fp = 0.f;
        if (live) {
#pragma unroll 4
          for (int jj = phase; jj < nt; jj += nph) {
            const float4 q = s_p[jj];
            const float dx = q.x - ip.x;
            const float dy = q.y - ip.y;
            const float dz = q.z - ip.z;
            const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2)));
            const float rinv = __builtin_amdgcn_rsqf(r2);
            const float mr = q.w * rinv;
            const float mr3 = mr * rinv * rinv;
            fp += mr;
            fax = fmaf(mr3, dx, fax);
            fay = fmaf(mr3, dy, fay);
            faz = fmaf(mr3, dz, faz);
          }
        }
        ax += (double)fax;
        ay += (double)fay;
        az += (double)faz;
        pp += (double)fp;
      }
      // cells: monopole + quadrupole (quad_avx2.c:35-149)
      for (int ct = 0; ct < nc; ct += kCTile) {
        const int nt = min(kCTile, nc - ct);
        __syncthreads();
        if (tid < nt) s_c[tid] = cells[clist[c0 + ct + tid]];
        __syncthreads();
        float fax = 0.f, fay = 0.f, faz = 0.f, fp = 0.f;
        if (live) {
#pragma unroll 2
          for (int jj = phase; jj < nt; jj += nph) {
            const pg5::Cell c = s_c[jj];
            const float dx = c.xm.x - ip.x;
            const float dy = c.xm.y - ip.y;
            const float dz = c.xm.z - ip.z;
            const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2)));
            const float rinv = __builtin_amdgcn_rsqf(r2);
            const float rinv2 = rinv * rinv;
            const float phip = c.xm.w * rinv;
            const float qx = fmaf(c.qa.z, dz, fmaf(c.qa.y, dy, c.qa.x * dx));
            const float qy = fmaf(c.qb.x, dz, fmaf(c.qa.w, dy, c.qa.y * dx));
            const float qz = fmaf(c.qb.y, dz, fmaf(c.qb.x, dy, c.qa.z * dx));
            const float dqd = fmaf(dz, qz, fmaf(dy, qy, dx * qx));
            const float rinv5 = rinv2 * rinv2 * rinv;
            const float phiq = 0.5f * dqd * rinv5;
            fp += phip + phiq;
            const float f = fmaf(5.f, phiq, phip) * rinv2;
            fax = fmaf(f, dx, fmaf(-rinv5, qx, fax));
            fay = fmaf(f, dy, fmaf(-rinv5, qy, fay));
            faz = fmaf(f, dz, fmaf(-rinv5, qz, faz));
          }
        }
        ax += (double)fax;
        ay += (double)fay;
        az += (double)faz;
        pp += (double)fp;
      }
      __syncthreads();
      s_red[0][tid] = ax;
      s_red[1][tid] = ay;
      s_red[2][tid] = az;
      s_red[3][tid] = pp;
      __syncthreads();
      if (tid < nib) {
        double sx = 0.0, sy = 0.0, sz = 0.0, sp = 0.0;
        for (int q = 0; q < nph; ++q) {
          sx += s_red[0][q * nib + tid];
          sy += s_red[1][q * nib + tid];
          sz += s_red[2][q * nib + tid];
          sp += s_red[3][q * nib + tid];
        }
        acc[3 * (size_t)ii + 0] = sx;
        acc[3 * (size_t)ii + 1] = sy;
        acc[3 * (size_t)ii + 2] = sz;
        pot[ii] = sp;
      }
    }
  }
}

__global__ void zero_counter(int* counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *counter = 0;
}

}  // namespace pgbh
