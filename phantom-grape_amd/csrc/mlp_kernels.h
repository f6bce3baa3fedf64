// Learned cut-off function: a 1-16-16-1 tanh MLP in FP32 that replaces the
// table lookup of the tabulated-cutoff force (PG5/table/libpg5/
// gravity_kernel_avx2.c:74-100).  No counterpart in the reference.
//
// What is learned depends on the shape the table holds.
//  * S2 shape (pg5_gen_s2_force_table; f = R(r, eps_PP) - R(r, eps_PM) with
//    R the force of two S2 spheres, PG5/table/libpg5/pg5_table.c:8-34).  With
//    H(u) = a^3 R(u a, a), the universal S2 shape on u in [0, 1] (H(0) = 12.8,
//    H(1) = 1, R = 1/r^3 beyond u = 1), the network learns H on u in [0, 1]:
//      f(r) = [r >= eps_PP] / r^3 + [r < eps_PP] H(r/eps_PP) / eps_PP^3
//             - H(r/eps_PM) / eps_PM^3                     for r < eps_PM,
//    and f = 0 beyond eps_PM.  H is bounded and smooth, so the network's
//    absolute error is at most the same relative error of a pair's Newtonian
//    force.  Pairs inside eps_PP (rare) take a second evaluation.
//  * Any other shape f (pg5_gen_force_table): with r_0 = r_c / 30 the network
//    maps v = sqrt(r / r_c) to Y(v) = f(r) (r^3 + r_0^3), bounded on [0, 1],
//    and f = Y / (r^3 + r_0^3) for r < r_c.
// Both are decoded by Cut below:  f = cn [not core] / r^3 + y c1 / (c3 r^3 + c2)
// (+ y_core cpp in the S2 core).
//
// Parameters (321 floats): w1[16] @0, b1[16] @16, W2[16][16] @32 (row k =
// second-layer unit, column j = first-layer unit), b2[16] @288, w3[16] @304,
// b3 @320.  Input scaling x = 2 v - 1.
//
// Training: Levenberg-Marquardt in FP64.  mlp_fwd_jac64 writes every
// sample's residual and Jacobian row (FP64 forward and backward pass);
// mlp_jtj64 forms J^T J, J^T r and r^T r in a fixed order (deterministic); the
// host solves the damped 321 x 321 system.
//
// Inference form.  Since tanh(z) = 2 s(c z) - 1 with s(u) = 1 / (1 + 2^-u)
// and c = 2 log2(e), the same network is evaluated with the cheaper
// activation s (one exp2, one add, one reciprocal) after an exact affine
// change of parameters, done once on the host in FP64 (to_inference_form):
//   layer 1: u1 = (2 c w1) v + c (b1 - w1)     (input v, not x = 2 v - 1)
//   layer 2: u2 = (2 c W2) s1 + c (b2 - sum_j W2[k][j])
//   output:  y  = (2 w3) . s2 + (b3 - sum_k w3[k])
// stored at the same offsets as the tanh-form parameters.
#pragma once
#include <hip/hip_runtime.h>

namespace mlp {

constexpr int kH = 16;
constexpr int kParams = kH + kH + kH * kH + kH + kH + 1;  // 321
constexpr int kW1 = 0, kB1 = 16, kW2 = 32, kB2 = 288, kW3 = 304, kB3 = 320;

// Parameters are read through the constant address space: the loads are
// uniform across the wavefront and become scalar loads into SGPRs, which the
// FMAs take as operands.  The parameters are never written while a kernel
// that reads them runs.
typedef const __attribute__((address_space(4))) float cfloat;

__device__ inline cfloat* as_const(const float* p) { return (cfloat*)p; }

// Hides the pointer's value from the optimiser so that the 321 parameter
// loads are issued where they are used instead of being hoisted out of the
// pair loop (which would need 321 registers).
__device__ inline cfloat* fresh(cfloat* p) {
  asm volatile("" : "+s"(p));
  return p;
}

// Logistic activation of the inference form: 1 / (1 + 2^-u).
__device__ inline float sig2(float u) {
  return __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(-u));
}

// Inference-form forward pass for K inputs at once: every parameter is loaded
// once (scalar loads, uniform across the wavefront) and used for K pairs.
template <int K>
__device__ inline void eval_k(cfloat* w, const float (&v)[K], float (&y)[K]) {
  float s1[K][kH];
#pragma unroll
  for (int j = 0; j < kH; ++j) {
    const float a = w[kW1 + j], b = w[kB1 + j];
#pragma unroll
    for (int u = 0; u < K; ++u) s1[u][j] = sig2(fmaf(a, v[u], b));
  }
#pragma unroll
  for (int u = 0; u < K; ++u) y[u] = w[kB3];
  // one second-layer unit per iteration: its 18 parameters fit the scalar
  // registers (unrolled, all 288 would be live at once and spill)
#pragma unroll 1
  for (int k = 0; k < kH; ++k) {
    float z[K];
#pragma unroll
    for (int u = 0; u < K; ++u) z[u] = w[kB2 + k];
#pragma unroll
    for (int j = 0; j < kH; ++j) {
      const float wkj = w[kW2 + kH * k + j];
#pragma unroll
      for (int u = 0; u < K; ++u) z[u] = fmaf(wkj, s1[u][j], z[u]);
    }
    const float w3 = w[kW3 + k];
#pragma unroll
    for (int u = 0; u < K; ++u) y[u] = fmaf(w3, sig2(z[u]), y[u]);
  }
}

// ---------------------------------------------------------------- FP64 fit
// The network is fitted by Levenberg-Marquardt on the weighted least-squares
// loss sum_s wt_s (y(v_s) - t_s)^2.  Each iteration needs the residuals and
// the Jacobian of every sample, and their products J^T J, J^T r, r^T r; all in
// FP64 on the GPU (the 321 x 321 damped solve runs on the host).

constexpr int kExt = kParams + 1;  // Jacobian columns + the residual column
constexpr int kTile = 16;          // J^T J output tile
constexpr int kTiles = (kExt + kTile - 1) / kTile;  // 21
constexpr int kChunk = 64;         // samples staged per step of mlp_jtj64

// Per-sample forward pass and Jacobian row, FP64, transposed output:
// jac[p * n + s] = sqrt(wt_s) dy_s/dp for p < kParams and
// jac[kParams * n + s] = sqrt(wt_s) (y_s - t_s) (the residual).  p64: the
// tanh-form parameters in FP64 (uniform loads).
__global__ void mlp_fwd_jac64(int n, const double* __restrict__ v, const double* __restrict__ t,
                              const double* __restrict__ sw, const double* __restrict__ p64,
                              double* __restrict__ jac) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const double x = 2.0 * v[s] - 1.0;
  double h1[kH], h2[kH];
#pragma unroll
  for (int j = 0; j < kH; ++j) h1[j] = tanh(fma(p64[kW1 + j], x, p64[kB1 + j]));
  double y = p64[kB3];
#pragma unroll
  for (int k = 0; k < kH; ++k) {
    double a = p64[kB2 + k];
#pragma unroll
    for (int j = 0; j < kH; ++j) a = fma(p64[kW2 + kH * k + j], h1[j], a);
    h2[k] = tanh(a);
    y = fma(p64[kW3 + k], h2[k], y);
  }
  const double w = sw[s];
  const size_t N = (size_t)n;
  jac[(size_t)kParams * N + s] = w * (y - t[s]);
  jac[(size_t)kB3 * N + s] = w;
  double dh1[kH];
#pragma unroll
  for (int j = 0; j < kH; ++j) dh1[j] = 0.0;
#pragma unroll
  for (int k = 0; k < kH; ++k) {
    jac[(size_t)(kW3 + k) * N + s] = w * h2[k];
    const double da = w * p64[kW3 + k] * (1.0 - h2[k] * h2[k]);
    jac[(size_t)(kB2 + k) * N + s] = da;
#pragma unroll
    for (int j = 0; j < kH; ++j) {
      jac[(size_t)(kW2 + kH * k + j) * N + s] = da * h1[j];
      dh1[j] = fma(da, p64[kW2 + kH * k + j], dh1[j]);
    }
  }
#pragma unroll
  for (int j = 0; j < kH; ++j) {
    const double da = dh1[j] * (1.0 - h1[j] * h1[j]);
    jac[(size_t)(kB1 + j) * N + s] = da;
    jac[(size_t)(kW1 + j) * N + s] = da * x;
  }
}

// M = E^T E for the n x kExt matrix E = [J | r] held transposed in jac
// (jac[c * n + s]): one 16 x 16 output tile per block, blocks with tile row
// <= tile column only (M is symmetric), samples summed in a fixed order (64
// at a time through LDS, each thread's partial sum in sample order), so the
// result is the same on every run.  out[a * kExt + b] for a <= b.
// M holds J^T J (a, b < kParams), J^T r (b = kParams) and r^T r.
__global__ __launch_bounds__(256) void mlp_jtj64(int n, const double* __restrict__ jac,
                                                 double* __restrict__ out) {
  const int ta = blockIdx.y, tb = blockIdx.x;
  if (ta > tb) return;
  __shared__ double sa[kTile][kChunk + 1];
  __shared__ double sb[kTile][kChunk + 1];
  const int tid = threadIdx.x;
  const int ia = tid / kTile, ib = tid % kTile;
  const int a0 = ta * kTile, b0 = tb * kTile;
  double acc = 0.0;
  for (int s0 = 0; s0 < n; s0 += kChunk) {
    const int m = min(kChunk, n - s0);
    __syncthreads();
    for (int k = tid; k < kTile * kChunk; k += 256) {
      const int r = k / kChunk, c = k % kChunk;
      const int ca = a0 + r, cb = b0 + r;
      sa[r][c] = (c < m && ca < kExt) ? jac[(size_t)ca * n + s0 + c] : 0.0;
      sb[r][c] = (c < m && cb < kExt) ? jac[(size_t)cb * n + s0 + c] : 0.0;
    }
    __syncthreads();
#pragma unroll 8
    for (int c = 0; c < kChunk; ++c) acc = fma(sa[ia][c], sb[ib][c], acc);
  }
  const int a = a0 + ia, b = b0 + ib;
  if (a < kExt && b < kExt && a <= b) out[(size_t)a * kExt + b] = acc;
}

// Decoding of the network output into the pair shape f (physical r^-3
// units) at the scaled radius r = |dx_s| (see the top of this file).
struct Cut {
  const float* w;  // inference-form parameters (device)
  float inv_rc;    // 1 / r_c (scaled units)
  float rc2;       // r_c^2
  int sqrt_in;     // network input: 1: sqrt(r / r_c), 0: r / r_c
  float cn, c1, c2, c3;
  float inv_pp, pp2, cpp;  // S2 core: r^2 < pp2 adds y(r / eps_PP) cpp (pp2 = 0: none)
};

__device__ inline float cut_input(const Cut& c, float r) {
  const float u = fminf(r * c.inv_rc, 1.f);
  return c.sqrt_in ? u * __builtin_amdgcn_rsqf(u) : u;
}

// f from the main evaluation y at radius r (rinv = 1/r, r2 = r^2).
__device__ inline float cut_main(const Cut& c, float y, float r, float r2, float rinv) {
  const float rinv3 = rinv * rinv * rinv;
  const float newton = r2 < c.pp2 ? 0.f : c.cn * rinv3;
  return fmaf(y * c.c1, __builtin_amdgcn_rcpf(fmaf(c.c3, r2 * r, c.c2)), newton);
}

// Learned f at n physical radii (for tests and error measurement), decoded as
// the force kernels do: 0 for r >= r_c and for r = 0.
__global__ void mlp_eval_f(int n, const float* __restrict__ r, Cut c, float xs,
                           float* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float rr = r[k] * xs;
  const float r2 = rr * rr;
  float v[1] = {cut_input(c, rr)};
  float y[1];
  eval_k<1>(fresh(as_const(c.w)), v, y);
  float f = cut_main(c, y[0], rr, r2, 1.f / rr);
  if (r2 < c.pp2) {
    float vc[1] = {rr * c.inv_pp};
    eval_k<1>(fresh(as_const(c.w)), vc, y);
    f = fmaf(y[0], c.cpp, f);
  }
  out[k] = (r2 > 0.f && r2 < c.rc2) ? f : 0.f;
}

// Host: tanh-form parameters (FP64) -> inference form (see above), rounded
// once to FP32.
inline void to_inference_form(const double* p, float* q) {
  const double c = 2.0 / 0.69314718055994530942;  // 2 log2(e)
  for (int j = 0; j < kH; ++j) {
    q[kW1 + j] = (float)(2.0 * c * p[kW1 + j]);
    q[kB1 + j] = (float)(c * (p[kB1 + j] - p[kW1 + j]));
  }
  double sw3 = 0.0;
  for (int k = 0; k < kH; ++k) {
    double sw = 0.0;
    for (int j = 0; j < kH; ++j) {
      q[kW2 + kH * k + j] = (float)(2.0 * c * p[kW2 + kH * k + j]);
      sw += p[kW2 + kH * k + j];
    }
    q[kB2 + k] = (float)(c * (p[kB2 + k] - sw));
    q[kW3 + k] = (float)(2.0 * p[kW3 + k]);
    sw3 += p[kW3 + k];
  }
  q[kB3] = (float)(p[kB3] - sw3);
}

}  // namespace mlp
