// Tabulated-cutoff (TreePM short-range) force kernel for gfx950
// (PG5/table/libpg5/gravity_kernel_avx2.c:22-137).
//
// Positions arrive pre-scaled by xscale = sqrt(r2scale), so that
// s = 2 + |dx_s|^2 runs over [2, fmax) for r in [0, r_cut); the force shape f
// (r^-3 units) is read from a table indexed directly by the float bits of s:
// idx = (bits(s) >> (23 - frc)) - 2^(7 + frc), with linear interpolation in
// the low bits, value + frac * slope as one FMA.  The per-pair arithmetic is
// the reference's, operation for operation, so each pair's contribution is
// bit-identical; only the summation order differs (FP32 per LDS tile, folded
// into FP64).  The table lives in LDS (512 or 1024 {value, slope} pairs).
#pragma once
#include <hip/hip_runtime.h>

namespace pgt {

constexpr int kThreads = 256;
constexpr int kJTile = 256;
constexpr int kMaxTable = 8192;  // entries (64 KB of LDS at most)

struct Part3 {
  double *ax, *ay, *az;
};

__device__ inline float table_f(const float2* __restrict__ tab, float s, float scut,
                                int shift, int base) {
  const float sc = fminf(s, scut);
  const unsigned bits = __float_as_uint(sc);
  const unsigned hi = bits >> shift;
  const float2 t = tab[(int)hi - base];
  const float frac = sc - __uint_as_float(hi << shift);
  return fmaf(frac, t.y, t.x);
}

// The same index with the quadratic entries {c0, c1} (q01) and c2 (q2),
// kept apart in LDS: a gather of 8 + 4 bytes per lane instead of one of 16
// for {c0, c1, c2, 0} (the quadratic kernel's time is spent in its LDS
// gathers).
__device__ inline float table_f_quad(const float2* __restrict__ q01,
                                     const float* __restrict__ q2, float s, float scut,
                                     int shift, int base) {
  const float sc = fminf(s, scut);
  const unsigned bits = __float_as_uint(sc);
  const unsigned hi = bits >> shift;
  const int k = (int)hi - base;
  const float2 t = q01[k];
  const float c2 = q2[k];
  const float d = sc - __uint_as_float(hi << shift);
  return fmaf(d, fmaf(d, c2, t.y), t.x);
}

// Block: 256 threads, BI i-particles, G = 256/BI groups share each j-tile.
template <int BI>
__global__ __launch_bounds__(kThreads) void table_force(
    int ni, const float4* __restrict__ xi, int nj, int jper, const float4* __restrict__ xj,
    const float2* __restrict__ gtab, int ntab, int frc_bit, double ascale, Part3 part,
    double* __restrict__ acc, int nsplit) {
  constexpr int G = kThreads / BI;
  extern __shared__ float2 s_tab[];  // ntab entries, sized at launch
  __shared__ float4 s_j[kJTile];
  const int tid = threadIdx.x;
  for (int k = tid; k < ntab; k += kThreads) s_tab[k] = gtab[k];
  const int shift = 23 - frc_bit;
  const int base = 1 << (30 - shift);
  const float scut = __uint_as_float((1u << 30) | ((unsigned)(ntab - 1) << shift));
  const int il = tid % BI;
  const int g = __builtin_amdgcn_readfirstlane(tid / BI);  // uniform: BI is a multiple of 64
  const int i = blockIdx.x * BI + il;
  const bool ivalid = i < ni;
  const float4 ip = xi[ivalid ? i : ni - 1];
  const int j0 = blockIdx.y * jper;
  const int j1 = min(nj, j0 + jper);
  double ax = 0.0, ay = 0.0, az = 0.0;
  for (int jt = j0; jt < j1; jt += kJTile) {
    const int nt = min(kJTile, j1 - jt);
    __syncthreads();
    if (tid < nt) s_j[tid] = xj[jt + tid];
    __syncthreads();
    float fax = 0.f, fay = 0.f, faz = 0.f;
#pragma unroll 4
    for (int jj = g; jj < nt; jj += G) {
      const float4 q = s_j[jj];
      const float dx = q.x - ip.x;
      const float dy = q.y - ip.y;
      const float dz = q.z - ip.z;
      const float s = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, 2.0f)));
      const float f = table_f(s_tab, s, scut, shift, base);
      fax = fmaf(f, q.w * dx, fax);
      fay = fmaf(f, q.w * dy, fay);
      faz = fmaf(f, q.w * dz, faz);
    }
    ax += (double)fax;
    ay += (double)fay;
    az += (double)faz;
  }
  if (G > 1) {
    __shared__ double r[3][kThreads];
    __syncthreads();
    r[0][tid] = ax; r[1][tid] = ay; r[2][tid] = az;
    __syncthreads();
    if (g == 0)
      for (int q = 1; q < G; ++q) {
        const int t = q * BI + il;
        ax += r[0][t]; ay += r[1][t]; az += r[2][t];
      }
  }
  if (g != 0 || !ivalid) return;
  if (nsplit == 1) {
    acc[3 * i + 0] = ax * ascale;
    acc[3 * i + 1] = ay * ascale;
    acc[3 * i + 2] = az * ascale;
  } else {
    const size_t o = (size_t)blockIdx.y * ni + i;
    part.ax[o] = ax;
    part.ay[o] = ay;
    part.az[o] = az;
  }
}

__global__ void table_reduce(int ni, int nsplit, Part3 part, double ascale, double* acc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ni) return;
  double ax = 0.0, ay = 0.0, az = 0.0;
  for (int s = 0; s < nsplit; ++s) {
    const size_t o = (size_t)s * ni + i;
    ax += part.ax[o];
    ay += part.ay[o];
    az += part.az[o];
  }
  acc[3 * i + 0] = ax * ascale;
  acc[3 * i + 1] = ay * ascale;
  acc[3 * i + 2] = az * ascale;
}

// Evaluates the table at n values of s (for parity tests of the lookup).
__global__ void table_eval(int n, const float* __restrict__ s, const float2* __restrict__ tab,
                           int ntab, int frc_bit, float* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int shift = 23 - frc_bit;
  const int base = 1 << (30 - shift);
  const float scut = __uint_as_float((1u << 30) | ((unsigned)(ntab - 1) << shift));
  out[k] = table_f(tab, s[k], scut, shift, base);
}

}  // namespace pgt
