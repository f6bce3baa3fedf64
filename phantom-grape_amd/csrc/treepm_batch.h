// Batched TreePM short-range force: every tree i-group with its j-list in one
// persistent launch.  No counterpart in the reference, whose caller loops over
// groups with g5_set_xmjMC + g5_calculate_force_on_xMC inside an OpenMP loop
// (the pattern of PG5/newton_quad/source.cpp:813-860).
//
// Inputs (device): pos[np] float4 {x*xs, y*xs, z*xs, m} (positions scaled as
// in PG5/table/libpg5/phantom_g5.c:112-128; periodic images are ghost copies
// appended by the caller), per group g an i-list ilist[ioff[g] .. +icnt[g])
// and a j-list jlist[joff[g] .. +jcnt[g]) (groups may share list ranges), an
// optional processing order (groups by decreasing cost), and a work counter
// that must be 0 at launch.
// Each block repeatedly takes the next group from the counter and stages the
// group's j-list in LDS tiles (gathered through jlist).  The cut-off shape
// comes from the evaluator: the LDS table (treepm_batch_table, bit-identical
// per pair with the drop-in kernel) or the learned MLP (treepm_batch_mlp).
// Results: acc[i][3] FP64 (written, not accumulated) for every listed i.
#pragma once
#include <hip/hip_runtime.h>

#include "mlp_kernels.h"
#include "table_kernels.h"

namespace pgb {

constexpr int kThreads = 256;    // treepm_batch_mlp block
constexpr int kJTile = 256;
constexpr int kTableBlock = 1024;  // treepm_batch_table block

// Table evaluator: s = 2 + |dx_s|^2, f from the LDS table.
struct TableCut {
  const float2* gtab;
  int ntab, frc_bit;
};

// MLP evaluator, scaled units: r_s = |dx_s|, v = sqrt(r_s / rc_s),
// f = xs^3 Y(v) / (r_s^3 + r0_s^3) for 0 < r_s < rc_s.
struct MlpCut {
  const float* w;
  float inv_rc_s;  // 1 / (r_c xs)
  float rc2_s;     // (r_c xs)^2
  float r03_s;     // (r_0 xs)^3
  float xs3;       // xs^3
};

__device__ inline int take_item(int* counter, int* s_item) {
  __syncthreads();
  if (threadIdx.x == 0) *s_item = atomicAdd(counter, 1);
  __syncthreads();
  return *s_item;
}

// Learned cut-off: a group's i-particles (up to 256 at a time) over 64, 128
// or 256 lanes, the block's G = 256 / that wave groups splitting each j-tile
// (the network's skip test wants every lane of a wavefront on the same j).
__global__ __launch_bounds__(kThreads) void treepm_batch_mlp(
    int ngroups, const int* __restrict__ order, const int* __restrict__ ioff,
    const int* __restrict__ icnt, const int* __restrict__ ilist, const int* __restrict__ joff,
    const int* __restrict__ jcnt, const int* __restrict__ jlist,
    const float4* __restrict__ pos, MlpCut mc, double ascale, double* __restrict__ acc,
    int* counter) {
  __shared__ float4 s_j[kJTile];
  __shared__ double s_red[3][kThreads];
  __shared__ int s_item;
  const int tid = threadIdx.x;
  for (;;) {
    const int it = take_item(counter, &s_item);
    if (it >= ngroups) break;
    const int grp = order ? order[it] : it;
    const int i0 = ioff[grp], ni = icnt[grp];
    const int j0 = joff[grp], nj = jcnt[grp];
    for (int ib = 0; ib < ni; ib += kThreads) {
      const int nib = min(kThreads, ni - ib);
      const int bi = nib <= 64 ? 64 : (nib <= 128 ? 128 : 256);
      const int G = kThreads / bi;
      const int il = tid % bi;
      const int g = __builtin_amdgcn_readfirstlane(tid / bi);  // uniform per wave
      const bool iv = il < nib;
      const int ii = ilist[i0 + ib + (iv ? il : 0)];
      const float4 ip = pos[ii];
      double ax = 0.0, ay = 0.0, az = 0.0;
      for (int jt = 0; jt < nj; jt += kJTile) {
        const int nt = min(kJTile, nj - jt);
        __syncthreads();
        if (tid < nt) s_j[tid] = pos[jlist[j0 + jt + tid]];
        __syncthreads();
        float fax = 0.f, fay = 0.f, faz = 0.f;
        {
          // This wave's j are jj = g + G t.  Pass 1 marks (64 t at a time,
          // bit t of a wave-uniform mask) the j that some lane has inside
          // the cut-off; pass 2 evaluates the network only for those, four
          // j per parameter pass.  Pairs outside contribute exactly 0, as
          // in the table path, so skipping them changes no sum.
          const int nsub = nt > g ? (nt - g + G - 1) / G : 0;
          for (int c0 = 0; c0 < nsub; c0 += 64) {
            const int cn = min(64, nsub - c0);
            unsigned long long mask = 0ull;
            for (int t = 0; t < cn; ++t) {
              const float4 q = s_j[g + G * (c0 + t)];
              const float dx = q.x - ip.x;
              const float dy = q.y - ip.y;
              const float dz = q.z - ip.z;
              const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
              const bool in = iv && r2 > 0.f && r2 < mc.rc2_s;
              if (__builtin_amdgcn_ballot_w64(in)) mask |= 1ull << t;
            }
            while (mask) {
              float v[4], fac[4], ddx[4], ddy[4], ddz[4];
              bool inb[4];
#pragma unroll
              for (int u = 0; u < 4; ++u) {
                const bool ok = mask != 0ull;
                const int t = ok ? __builtin_ctzll(mask) : 0;
                if (ok) mask &= mask - 1ull;
                const float4 q = s_j[g + G * (c0 + t)];
                const float dx = q.x - ip.x;
                const float dy = q.y - ip.y;
                const float dz = q.z - ip.z;
                const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                const float rinv = __builtin_amdgcn_rsqf(r2);
                const float r = r2 * rinv;
                const float uu = fminf(r * mc.inv_rc_s, 1.f);
                v[u] = uu * __builtin_amdgcn_rsqf(uu);
                inb[u] = ok && r2 > 0.f && r2 < mc.rc2_s;
                fac[u] = mc.xs3 * __builtin_amdgcn_rcpf(fmaf(r2, r, mc.r03_s));
                ddx[u] = q.w * dx;
                ddy[u] = q.w * dy;
                ddz[u] = q.w * dz;
              }
              float y[4];
              mlp::eval_k<4>(mlp::fresh(mlp::as_const(mc.w)), v, y);
#pragma unroll
              for (int u = 0; u < 4; ++u) {
                const float f = inb[u] ? y[u] * fac[u] : 0.f;
                fax = fmaf(f, ddx[u], fax);
                fay = fmaf(f, ddy[u], fay);
                faz = fmaf(f, ddz[u], faz);
              }
            }
          }
        }
        ax += (double)fax;
        ay += (double)fay;
        az += (double)faz;
      }
      if (G > 1) {
        __syncthreads();
        s_red[0][tid] = ax;
        s_red[1][tid] = ay;
        s_red[2][tid] = az;
        __syncthreads();
        if (g == 0)
          for (int q = 1; q < G; ++q) {
            ax += s_red[0][q * bi + il];
            ay += s_red[1][q * bi + il];
            az += s_red[2][q * bi + il];
          }
      }
      if (g == 0 && iv) {
        acc[3 * (size_t)ii + 0] = ax * ascale;
        acc[3 * (size_t)ii + 1] = ay * ascale;
        acc[3 * (size_t)ii + 2] = az * ascale;
      }
    }
  }
}

// Table cut-off with lanes packed by phases.  Block = BLOCK threads.  A
// group of n <= BLOCK i-particles takes p = BLOCK / n copies of the i-set:
// lane t holds i-particle t % n and walks the j of each LDS tile with
// jj = t / n, t / n + p, ...; the p phase sums of each i are added in phase
// order at the end.  Lanes in use: n p of BLOCK (at least half, usually far
// more), where whole-wavefront i-blocks of 64, 128 or 256 lanes leave the
// rest of a wavefront idle for every group that is not a multiple of 64.
// Groups larger than BLOCK run in passes of BLOCK.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void treepm_batch_table(
    int ngroups, const int* __restrict__ order, const int* __restrict__ ioff,
    const int* __restrict__ icnt, const int* __restrict__ ilist, const int* __restrict__ joff,
    const int* __restrict__ jcnt, const int* __restrict__ jlist,
    const float4* __restrict__ pos, TableCut tc, double ascale, double* __restrict__ acc,
    int* counter) {
  extern __shared__ float2 s_tab[];
  __shared__ float4 s_j[BLOCK];
  __shared__ double s_red[3][BLOCK];
  __shared__ int s_item;
  const int tid = threadIdx.x;
  for (int k = tid; k < tc.ntab; k += BLOCK) s_tab[k] = tc.gtab[k];
  const int shift = 23 - tc.frc_bit;
  const int base = 1 << (30 - shift);
  const float scut = __uint_as_float((1u << 30) | ((unsigned)(tc.ntab - 1) << shift));
  for (;;) {
    const int it = take_item(counter, &s_item);
    if (it >= ngroups) break;
    const int grp = order ? order[it] : it;
    const int i0 = ioff[grp], ni = icnt[grp];
    const int j0 = joff[grp], nj = jcnt[grp];
    for (int ib = 0; ib < ni; ib += BLOCK) {
      const int nib = min(BLOCK, ni - ib);
      const int p = BLOCK / nib;  // phases
      const int il = tid % nib;
      const int phase = tid / nib;
      const bool live = phase < p;
      const int ii = ilist[i0 + ib + il];
      const float4 ip = pos[ii];
      double ax = 0.0, ay = 0.0, az = 0.0;
      for (int jt = 0; jt < nj; jt += BLOCK) {
        const int nt = min(BLOCK, nj - jt);
        __syncthreads();
        if (tid < nt) s_j[tid] = pos[jlist[j0 + jt + tid]];
        __syncthreads();
        float fax = 0.f, fay = 0.f, faz = 0.f;
        if (live) {
#pragma unroll 4
          for (int jj = phase; jj < nt; jj += p) {
            const float4 q = s_j[jj];
            const float dx = q.x - ip.x;
            const float dy = q.y - ip.y;
            const float dz = q.z - ip.z;
            const float s = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, 2.0f)));
            const float f = pgt::table_f(s_tab, s, scut, shift, base);
            fax = fmaf(f, q.w * dx, fax);
            fay = fmaf(f, q.w * dy, fay);
            faz = fmaf(f, q.w * dz, faz);
          }
        }
        ax += (double)fax;
        ay += (double)fay;
        az += (double)faz;
      }
      __syncthreads();
      s_red[0][tid] = ax;
      s_red[1][tid] = ay;
      s_red[2][tid] = az;
      __syncthreads();
      if (tid < nib) {
        double sx = 0.0, sy = 0.0, sz = 0.0;
        for (int q = 0; q < p; ++q) {
          sx += s_red[0][q * nib + tid];
          sy += s_red[1][q * nib + tid];
          sz += s_red[2][q * nib + tid];
        }
        acc[3 * (size_t)ii + 0] = sx * ascale;
        acc[3 * (size_t)ii + 1] = sy * ascale;
        acc[3 * (size_t)ii + 2] = sz * ascale;
      }
    }
  }
}

// Clears the work counter (a kernel, not a memset, so it can sit in a graph).
__global__ void zero_counter(int* counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *counter = 0;
}

}  // namespace pgb
