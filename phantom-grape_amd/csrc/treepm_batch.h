// Batched TreePM short-range force: every tree i-group with its j-list in one
// persistent launch.  No counterpart in the reference, whose caller loops over
// groups with g5_set_xmjMC + g5_calculate_force_on_xMC inside an OpenMP loop
// (the pattern of PG5/newton_quad/source.cpp:813-860).
//
// Inputs (device): pos[np] float4 {x*xs, y*xs, z*xs, m} (positions scaled as
// in PG5/table/libpg5/phantom_g5.c:112-128; periodic images are ghost copies
// appended by the caller); optionally pos64[np][3], the same positions in
// FP64 (unscaled), from which each group forms its separations relative to a
// group origin o (its first i-particle): a position enters the pair loop
// through the reference's ingest conversions applied to x - o instead of x
// (i: float(double(float(x - o)) * xs), PG5/table/libpg5/phantom_g5.c:102-104;
// j: float(x - o) * float(xs) in FP32, :116-127), so separations keep their
// digits where |x| >> |dx| (masses still come from pos[].w); per group g an
// i-list
// ilist[ioff[g] .. +icnt[g])
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
constexpr int kTableBlock = 1024;  // treepm_batch_table default block (256, 512 also built)

// Table evaluator: s = 2 + |dx_s|^2, f from the LDS table (linear: the
// reference's {value, slope}; quadratic: {c0, c1, c2, 0}).
struct TableCut {
  const float2* gtab;
  const float4* gquad;
  int ntab, frc_bit;
};

// Group-origin positions (see above).  o: the group's origin in FP64.
struct Origin {
  const double* pos64;  // null: take the scaled float4 positions as they are
  double xs;
  float xsf;  // float(xs)
};

template <bool I>
__device__ inline float4 load_pos(const float4* __restrict__ pos, const Origin& og, int k,
                                  double ox, double oy, double oz) {
  float4 q = pos[k];
  if (og.pos64) {
    const double* p = og.pos64 + 3 * (size_t)k;
    if (I) {
      q.x = (float)((double)(float)(p[0] - ox) * og.xs);
      q.y = (float)((double)(float)(p[1] - oy) * og.xs);
      q.z = (float)((double)(float)(p[2] - oz) * og.xs);
    } else {
      q.x = (float)(p[0] - ox) * og.xsf;
      q.y = (float)(p[1] - oy) * og.xsf;
      q.z = (float)(p[2] - oz) * og.xsf;
    }
  }
  return q;
}

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
    const float4* __restrict__ pos, Origin og, mlp::Cut mc, double ascale,
    double* __restrict__ acc, int* counter) {
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
    double ox = 0.0, oy = 0.0, oz = 0.0;
    if (og.pos64 && ni > 0) {
      const double* o = og.pos64 + 3 * (size_t)ilist[i0];
      ox = o[0];
      oy = o[1];
      oz = o[2];
    }
    for (int ib = 0; ib < ni; ib += kThreads) {
      const int nib = min(kThreads, ni - ib);
      const int bi = nib <= 64 ? 64 : (nib <= 128 ? 128 : 256);
      const int G = kThreads / bi;
      const int il = tid % bi;
      const int g = __builtin_amdgcn_readfirstlane(tid / bi);  // uniform per wave
      const bool iv = il < nib;
      const int ii = ilist[i0 + ib + (iv ? il : 0)];
      const float4 ip = load_pos<true>(pos, og, ii, ox, oy, oz);
      double ax = 0.0, ay = 0.0, az = 0.0;
      for (int jt = 0; jt < nj; jt += kJTile) {
        const int nt = min(kJTile, nj - jt);
        __syncthreads();
        if (tid < nt) s_j[tid] = load_pos<false>(pos, og, jlist[j0 + jt + tid], ox, oy, oz);
        __syncthreads();
        float fax = 0.f, fay = 0.f, faz = 0.f;
        {
          // This wave's j are jj = g + G t.  Pass 1 marks (64 t at a time,
          // bit t of a wave-uniform mask) the j that some lane has inside
          // the cut-off, and those inside the S2 core; pass 2 evaluates the
          // network only for the first, four j per parameter pass, pass 3
          // the core term for the second.  Pairs outside contribute exactly
          // 0, as in the table path, so skipping them changes no sum.
          const int nsub = nt > g ? (nt - g + G - 1) / G : 0;
          for (int c0 = 0; c0 < nsub; c0 += 64) {
            const int cn = min(64, nsub - c0);
            unsigned long long mask = 0ull, cmask = 0ull;
            for (int t = 0; t < cn; ++t) {
              const float4 q = s_j[g + G * (c0 + t)];
              const float dx = q.x - ip.x;
              const float dy = q.y - ip.y;
              const float dz = q.z - ip.z;
              const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
              const bool in = iv && r2 > 0.f && r2 < mc.rc2;
              if (__builtin_amdgcn_ballot_w64(in)) mask |= 1ull << t;
              if (__builtin_amdgcn_ballot_w64(in && r2 < mc.pp2)) cmask |= 1ull << t;
            }
            for (int pass = 0; pass < 2; ++pass) {
              unsigned long long mk = pass == 0 ? mask : cmask;
              while (mk) {
                float v[4], r2s[4], rs[4], rinvs[4], ddx[4], ddy[4], ddz[4];
                bool inb[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                  const bool ok = mk != 0ull;
                  const int t = ok ? __builtin_ctzll(mk) : 0;
                  if (ok) mk &= mk - 1ull;
                  const float4 q = s_j[g + G * (c0 + t)];
                  const float dx = q.x - ip.x;
                  const float dy = q.y - ip.y;
                  const float dz = q.z - ip.z;
                  const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                  const float rinv = __builtin_amdgcn_rsqf(r2);
                  const float r = r2 * rinv;
                  v[u] = pass == 0 ? mlp::cut_input(mc, r) : r * mc.inv_pp;
                  inb[u] = ok && r2 > 0.f && r2 < (pass == 0 ? mc.rc2 : mc.pp2);
                  r2s[u] = r2;
                  rs[u] = r;
                  rinvs[u] = rinv;
                  ddx[u] = q.w * dx;
                  ddy[u] = q.w * dy;
                  ddz[u] = q.w * dz;
                }
                float y[4];
                mlp::eval_k<4>(mlp::fresh(mlp::as_const(mc.w)), v, y);
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                  const float fm = pass == 0 ? mlp::cut_main(mc, y[u], rs[u], r2s[u], rinvs[u])
                                             : y[u] * mc.cpp;
                  const float f = inb[u] ? fm : 0.f;
                  fax = fmaf(f, ddx[u], fax);
                  fay = fmaf(f, ddy[u], fay);
                  faz = fmaf(f, ddz[u], faz);
                }
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
// Groups larger than BLOCK run in passes of BLOCK.  QUAD: the quadratic
// table (one more FMA and a 16-byte LDS read per pair).
template <int BLOCK, bool QUAD>
__global__ __launch_bounds__(BLOCK) void treepm_batch_table(
    int ngroups, const int* __restrict__ order, const int* __restrict__ ioff,
    const int* __restrict__ icnt, const int* __restrict__ ilist, const int* __restrict__ joff,
    const int* __restrict__ jcnt, const int* __restrict__ jlist,
    const float4* __restrict__ pos, Origin og, TableCut tc, double ascale,
    double* __restrict__ acc, int* counter) {
  // linear: {value, slope} per entry; quadratic: {c0, c1} per entry, then c2
  extern __shared__ float4 s_tab4[];
  float2* s_tab = reinterpret_cast<float2*>(s_tab4);
  float* s_c2 = reinterpret_cast<float*>(s_tab + tc.ntab);
  __shared__ float4 s_j[BLOCK];
  __shared__ double s_red[3][BLOCK];
  __shared__ int s_item;
  const int tid = threadIdx.x;
  for (int k = tid; k < tc.ntab; k += BLOCK) {
    if (QUAD) {
      const float4 q = tc.gquad[k];
      s_tab[k] = make_float2(q.x, q.y);
      s_c2[k] = q.z;
    } else {
      s_tab[k] = tc.gtab[k];
    }
  }
  const int shift = 23 - tc.frc_bit;
  const int base = 1 << (30 - shift);
  const float scut = __uint_as_float((1u << 30) | ((unsigned)(tc.ntab - 1) << shift));
  for (;;) {
    const int it = take_item(counter, &s_item);
    if (it >= ngroups) break;
    const int grp = order ? order[it] : it;
    const int i0 = ioff[grp], ni = icnt[grp];
    const int j0 = joff[grp], nj = jcnt[grp];
    double ox = 0.0, oy = 0.0, oz = 0.0;
    if (og.pos64 && ni > 0) {
      const double* o = og.pos64 + 3 * (size_t)ilist[i0];
      ox = o[0];
      oy = o[1];
      oz = o[2];
    }
    for (int ib = 0; ib < ni; ib += BLOCK) {
      const int nib = min(BLOCK, ni - ib);
      const int p = BLOCK / nib;  // phases
      const int il = tid % nib;
      const int phase = tid / nib;
      const bool live = phase < p;
      const int ii = ilist[i0 + ib + il];
      const float4 ip = load_pos<true>(pos, og, ii, ox, oy, oz);
      double ax = 0.0, ay = 0.0, az = 0.0;
      for (int jt = 0; jt < nj; jt += BLOCK) {
        const int nt = min(BLOCK, nj - jt);
        __syncthreads();
        if (tid < nt) s_j[tid] = load_pos<false>(pos, og, jlist[j0 + jt + tid], ox, oy, oz);
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
            const float f = QUAD ? pgt::table_f_quad(s_tab, s_c2, s, scut, shift, base)
                                 : pgt::table_f(s_tab, s, scut, shift, base);
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
