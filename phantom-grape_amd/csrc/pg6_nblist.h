// Neighbour-list kernel for the g6calc_lasthalf*n entry points
// (PG6/libpg6/gravity.c:680-914, 916-1182).
//
// One block per i-particle scans the predicted j-particles in address order,
// flags r <= h_i with r^2 = |dx|^2 + eps^2 (dx formed in FP64 and rounded to
// FP32, as in the force kernel), excludes the particle itself by id, and
// compacts the flagged ids in j order with a block-wide prefix count.
// Lists longer than `maxlen` keep their first `maxlen` entries and report the
// full length, which is how the host detects overflow (gravity.c:905-906).
#pragma once
#include "pg6_kernels.h"

namespace pg6 {

__global__ __launch_bounds__(kThreads) void pg6_nblist(
    int ni, const double* __restrict__ ipos, const int* __restrict__ iid,
    const float* __restrict__ h2, float eps2, int nj, JPred jp, int maxlen,
    int* __restrict__ nbl, int* __restrict__ nblen) {
  __shared__ int s_wave_count[kThreads / 64];
  __shared__ int s_base;
  const int i = blockIdx.x;
  if (i >= ni) return;  // uniform per block
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const double xi = ipos[3 * i + 0], yi = ipos[3 * i + 1], zi = ipos[3 * i + 2];
  const int idi = iid[i];
  const float hh = h2[i];
  if (tid == 0) s_base = 0;
  __syncthreads();
  for (int jb = 0; jb < nj; jb += kThreads) {
    const int j = jb + tid;
    bool flag = false;
    if (j < nj && hh > 0.f) {
      const float dx = (float)(jp.x[j] - xi);
      const float dy = (float)(jp.y[j] - yi);
      const float dz = (float)(jp.z[j] - zi);
      const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2)));
      flag = (r2 <= hh) && (jp.id[j] != idi);
    }
    const unsigned long long ball = __ballot(flag);
    const int before = __popcll(ball & ((1ull << lane) - 1ull));
    if (lane == 0) s_wave_count[wave] = __popcll(ball);
    __syncthreads();
    int off = s_base;
    for (int w = 0; w < wave; ++w) off += s_wave_count[w];
    const int pos = off + before;
    if (flag && pos < maxlen) nbl[(size_t)i * maxlen + pos] = jp.id[j];
    __syncthreads();
    if (tid == 0) {
      int tot = 0;
      for (int w = 0; w < kThreads / 64; ++w) tot += s_wave_count[w];
      s_base += tot;
    }
    __syncthreads();
  }
  if (tid == 0) nblen[i] = s_base;
}

}  // namespace pg6
