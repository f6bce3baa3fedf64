// Learned cut-off function: a 1-16-16-1 tanh MLP in FP32 that replaces the
// table lookup of the tabulated-cutoff force (PG5/table/libpg5/
// gravity_kernel_avx2.c:74-100).  No counterpart in the reference.
//
// What is learned.  With r_c the cut-off radius and r_0 a core length (the PP
// softening for the S2 shape), the network maps v = sqrt(r / r_c) in [0, 1]
// to Y(v) = f(r) (r^3 + r_0^3), where f is the tabulated shape in r^-3 units.
// Y is bounded and smooth on [0, 1] (it tends to f(0) r_0^3 at r = 0 and to
// r^3 f(r) <= 1 outside the core), so an absolute error of the network is a
// relative error of the pair force wherever Y is O(1).  The kernel recovers
// f = Y / (r^3 + r_0^3) and uses f = 0 for r >= r_c, as the table does.
//
// Parameters (321 floats): w1[16] @0, b1[16] @16, W2[16][16] @32 (row k =
// second-layer unit, column j = first-layer unit), b2[16] @288, w3[16] @304,
// b3 @320.  Input scaling x = 2 v - 1.
//
// Training kernels: mlp_sample_grad writes each sample's weighted-MSE
// gradient (transposed, [param][sample]); mlp_reduce_adam sums them in a
// fixed order (deterministic) and applies one Adam step.
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

// tanh(x) = 1 - 2 / (1 + 2^(2 x log2 e)) with the hardware exp2 and
// reciprocal (each about 1 ulp); saturates cleanly to +-1.
__device__ inline float tanh_fast(float x) {
  const float e = __builtin_amdgcn_exp2f(x * 2.8853900817779268f);
  return 1.f - 2.f * __builtin_amdgcn_rcpf(1.f + e);
}

// Forward pass.  `w` is uniform across the wavefront (scalar loads).
__device__ inline float eval(cfloat* w, float v) {
  const float x = fmaf(2.f, v, -1.f);
  float h1[kH];
#pragma unroll
  for (int j = 0; j < kH; ++j) h1[j] = tanh_fast(fmaf(w[kW1 + j], x, w[kB1 + j]));
  float y = w[kB3];
#pragma unroll
  for (int k = 0; k < kH; ++k) {
    float a = w[kB2 + k];
#pragma unroll
    for (int j = 0; j < kH; ++j) a = fmaf(w[kW2 + kH * k + j], h1[j], a);
    y = fmaf(w[kW3 + k], tanh_fast(a), y);
  }
  return y;
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
#pragma unroll
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

// Per-sample gradient of wt * (y - t)^2 (divided by n), transposed output
// g[p * n + s]; loss[s] = wt * (y - t)^2 / n.
__global__ void mlp_sample_grad(int n, const float* __restrict__ v, const float* __restrict__ t,
                                const float* __restrict__ wt, const float* __restrict__ wg,
                                float* __restrict__ g, float* __restrict__ loss) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  cfloat* w = fresh(as_const(wg));
  const float x = fmaf(2.f, v[s], -1.f);
  float h1[kH], h2[kH];
#pragma unroll
  for (int j = 0; j < kH; ++j) h1[j] = tanh_fast(fmaf(w[kW1 + j], x, w[kB1 + j]));
  float y = w[kB3];
#pragma unroll
  for (int k = 0; k < kH; ++k) {
    float a = w[kB2 + k];
#pragma unroll
    for (int j = 0; j < kH; ++j) a = fmaf(w[kW2 + kH * k + j], h1[j], a);
    h2[k] = tanh_fast(a);
    y = fmaf(w[kW3 + k], h2[k], y);
  }
  w = fresh(w);  // reload for the backward pass
  const float inv_n = 1.f / (float)n;
  const float r = y - t[s];
  loss[s] = wt[s] * r * r * inv_n;
  const float dy = 2.f * wt[s] * r * inv_n;
  const size_t N = (size_t)n;
  g[(size_t)kB3 * N + s] = dy;
  float dh1[kH];
#pragma unroll
  for (int j = 0; j < kH; ++j) dh1[j] = 0.f;
#pragma unroll
  for (int k = 0; k < kH; ++k) {
    g[(size_t)(kW3 + k) * N + s] = dy * h2[k];
    const float da = dy * w[kW3 + k] * (1.f - h2[k] * h2[k]);
    g[(size_t)(kB2 + k) * N + s] = da;
#pragma unroll
    for (int j = 0; j < kH; ++j) {
      g[(size_t)(kW2 + kH * k + j) * N + s] = da * h1[j];
      dh1[j] = fmaf(da, w[kW2 + kH * k + j], dh1[j]);
    }
  }
#pragma unroll
  for (int j = 0; j < kH; ++j) {
    const float da = dh1[j] * (1.f - h1[j] * h1[j]);
    g[(size_t)(kB1 + j) * N + s] = da;
    g[(size_t)(kW1 + j) * N + s] = da * x;
  }
}

// One block per parameter (blockIdx.x < kParams) sums its n per-sample
// gradients in a fixed tree order; block kParams sums the loss.  With
// adam != 0, thread 0 then applies an Adam step to w[p] (m, u: moments,
// step = 1-based Adam step).  Sums are written to gsum[p] (gsum[kParams] =
// loss).
__global__ __launch_bounds__(256) void mlp_reduce_adam(int n, const float* __restrict__ g,
                                                       const float* __restrict__ loss,
                                                       float* __restrict__ w, float* m, float* u,
                                                       float* gsum, int adam, int step, float lr,
                                                       float b1, float b2, float eps) {
  __shared__ double part[256];
  const int p = blockIdx.x;
  const float* src = p < kParams ? g + (size_t)p * n : loss;
  double acc = 0.0;
  for (int s = threadIdx.x; s < n; s += 256) acc += (double)src[s];
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  const float gs = (float)part[0];
  gsum[p] = gs;
  if (!adam || p >= kParams) return;
  const float mp = b1 * m[p] + (1.f - b1) * gs;
  const float up = b2 * u[p] + (1.f - b2) * gs * gs;
  m[p] = mp;
  u[p] = up;
  const float mh = mp / (1.f - powf(b1, (float)step));
  const float uh = up / (1.f - powf(b2, (float)step));
  w[p] -= lr * mh / (sqrtf(uh) + eps);
}

// Evaluates the learned f(r) at n physical radii (for tests and error
// measurement) with the inference-form parameters `wi`, as the force kernel
// does: f = Y(sqrt(r/rc)) / (r^3 + r0^3), 0 for r >= rc.
__global__ void mlp_eval_f(int n, const float* __restrict__ r, const float* __restrict__ wi,
                           float rc, float r03, float* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float rr = r[k];
  float v[1] = {sqrtf(fminf(rr / rc, 1.f))};
  float y[1];
  eval_k<1>(fresh(as_const(wi)), v, y);
  out[k] = rr < rc ? y[0] / (rr * rr * rr + r03) : 0.f;
}

// Host: tanh-form parameters -> inference form (see above), in FP64.
inline void to_inference_form(const float* p, float* q) {
  const double c = 2.0 / 0.69314718055994530942;  // 2 log2(e)
  for (int j = 0; j < kH; ++j) {
    q[kW1 + j] = (float)(2.0 * c * p[kW1 + j]);
    q[kB1 + j] = (float)(c * ((double)p[kB1 + j] - p[kW1 + j]));
  }
  double sw3 = 0.0;
  for (int k = 0; k < kH; ++k) {
    double sw = 0.0;
    for (int j = 0; j < kH; ++j) {
      q[kW2 + kH * k + j] = (float)(2.0 * c * p[kW2 + kH * k + j]);
      sw += p[kW2 + kH * k + j];
    }
    q[kB2 + k] = (float)(c * ((double)p[kB2 + k] - sw));
    q[kW3 + k] = (float)(2.0 * p[kW3 + k]);
    sw3 += p[kW3 + k];
  }
  q[kB3] = (float)((double)p[kB3] - sw3);
}

}  // namespace mlp
