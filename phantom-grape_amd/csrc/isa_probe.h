// Issue-rate probe: the cost of the VALU operations the pair kernels are made
// of, measured on the device they run on.  Each lane runs eight independent
// chains of one operation, so the rate is set by issue, not by latency.
//   MODE 0: v_fma_f32          a = a * b + c
//   MODE 1: v_rsq_f32 + v_add  a = rsq(a) + c
//   MODE 2: v_rsq_f32          a = rsq(a)
// The result is stored so that no chain is optimised away.
#pragma once
#include <hip/hip_runtime.h>

namespace pgprobe {

constexpr int kChains = 8;

template <int MODE>
__global__ __launch_bounds__(256) void chains(int iters, float seed, float* __restrict__ out) {
  float a[kChains];
  for (int u = 0; u < kChains; ++u) a[u] = seed + (float)threadIdx.x * 1e-3f + (float)u;
  const float b = 1.0000001f, c = 1e-7f;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int u = 0; u < kChains; ++u) {
      if (MODE == 0) a[u] = fmaf(a[u], b, c);
      if (MODE == 1) a[u] = __builtin_amdgcn_rsqf(a[u]) + c;
      if (MODE == 2) a[u] = __builtin_amdgcn_rsqf(a[u]);
    }
  }
  float s = 0.f;
  for (int u = 0; u < kChains; ++u) s += a[u];
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace pgprobe
