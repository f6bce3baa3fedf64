// Particle-mesh long-range force of the TreePM split on the GPU.
//
// The table force of PG5/table (f = R(r, eps_PP) - R(r, eps_PM),
// PG5/table/libpg5/pg5_table.c:8-34) is the short-range side; this is the
// long-range side, the periodic force of R(r, eps_PM), i.e. of two S2 spheres
// of diameter a = eps_PM (Hockney & Eastwood sec. 8-3), so that PP + PM is
// the S2-softened Newtonian force of the periodic unit box.  Mesh method:
//   1. TSC assignment of the particles' density on an M^3 mesh (mesh points
//      at integer multiples of h = 1/M): the particles are sorted by the
//      Morton code of their nearest mesh point, each block sums a run of
//      kAssignRun of them into a box of the mesh held in LDS and adds the
//      box to the mesh (a run whose box does not fit adds straight to the
//      mesh).  The sums are 64-bit fixed point (pm_fix_scale): integer
//      addition is associative, so the mesh does not depend on the order in
//      which the atomics land and the force is the same bit for bit on
//      every run;
//   2. real-to-complex FFT (hipFFT, FP64);
//   3. Green's function -4 pi S(k a)^2 / (k^2 W(k)^2), with S the S2 form
//      factor and W the TSC window, times -i k for each axis (spectral
//      gradient), normalised by 1/M^3;
//   4. three complex-to-real FFTs;
//   5. TSC interpolation of the mesh accelerations back to the particles.
// G = 1, uniform neutralising background (k = 0 dropped), as in Ewald
// summation.
#pragma once
#include <hip/hip_runtime.h>
#include <climits>

namespace pm {

// Nearest mesh point and TSC weights along one axis, u in mesh units.
__device__ inline int tsc(double u, double w[3]) {
  const double g = floor(u + 0.5);
  const double d = u - g;
  w[0] = 0.5 * (0.5 - d) * (0.5 - d);
  w[1] = 0.75 - d * d;
  w[2] = 0.5 * (0.5 + d) * (0.5 + d);
  return (int)g;
}

__device__ inline int wrap(int i, int m) {
  i %= m;
  return i < 0 ? i + m : i;
}

// Nearest mesh point of u (mesh units), wrapped into [0, M).
__device__ inline int nearest(double u, int M) { return wrap((int)floor(u + 0.5), M); }

// Interleaves the low 10 bits of x, y, z (x in the highest position).
__device__ inline unsigned morton3(unsigned x, unsigned y, unsigned z) {
  auto spread = [](unsigned v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
  };
  return (spread(x) << 2) | (spread(y) << 1) | spread(z);
}

// Sort keys: Morton code of each particle's nearest mesh point, and its index.
__global__ void pm_keys(int n, const double* __restrict__ pos, int M, unsigned* __restrict__ key,
                        int* __restrict__ idx) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  key[p] = morton3(nearest(pos[3 * p] * M, M), nearest(pos[3 * p + 1] * M, M),
                   nearest(pos[3 * p + 2] * M, M));
  idx[p] = p;
}

// Run length and box size measured on a clumpy 2^20-particle box with a 128^3
// mesh: runs of 128 keep nearly every box (uniform background included)
// within 6144 points, so that few runs fall back to direct atomics.
constexpr int kAssignThreads = 128;
constexpr int kAssignPer = 1;                               // particles per thread
constexpr int kAssignRun = kAssignThreads * kAssignPer;     // particles per block
constexpr int kBoxCells = 6144;                             // LDS box: 48 KB of FP64

// Fixed-point scale of the mass assignment.  Every mesh point receives at
// most sum |m| M^3 (TSC weights of a particle sum to 1), so with
// 2^e <= 2^62 / (sum |m| M^3) no point and no partial sum can overflow a
// signed 64-bit integer; one unit is 2^-e, a relative resolution of
// M^3 / 2^62 of the mean density (4.5e-13 at M = 128).
constexpr int kMassBlocks = 256;

// Per-block sums of |m| (FP64, fixed order): part[kMassBlocks].  Block = 256.
__global__ __launch_bounds__(256) void pm_mass_part(int n, const double* __restrict__ mass,
                                                    double* __restrict__ part) {
  __shared__ double s[256];
  double v = 0.0;
  for (int p = blockIdx.x * 256 + threadIdx.x; p < n; p += kMassBlocks * 256) v += fabs(mass[p]);
  s[threadIdx.x] = v;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// scale[0] = 2^e from the block sums (one thread, in block order).
__global__ void pm_fix_scale(const double* __restrict__ part, int M, double* __restrict__ scale) {
  if (threadIdx.x != 0) return;
  double sum = 0.0;
  for (int b = 0; b < kMassBlocks; ++b) sum += part[b];
  const double bound = sum * (double)M * (double)M * (double)M;
  int e = 0;
  if (bound > 0.0) {
    frexp(bound, &e);  // bound < 2^e
    e = 62 - e;
  }
  scale[0] = ldexp(1.0, e);
}

// TSC assignment of the sorted run [blockIdx.x * kAssignRun, ...) of
// particles (order[] = particle indices) as fixed-point integers (units of
// 1 / scale[0]) into rho (M^3 signed 64-bit integers, two's complement sums
// through unsigned atomics).  The run's stencils span the mesh box [lo, hi]
// (unwrapped nearest points +-1); if it holds at most kBoxCells points the
// run is summed in LDS and the box added to rho point by point, else every
// stencil point goes straight to rho.
__global__ __launch_bounds__(kAssignThreads) void pm_assign_sorted(
    int n, const int* __restrict__ order, const double* __restrict__ pos,
    const double* __restrict__ mass, int M, const double* __restrict__ scale,
    unsigned long long* __restrict__ rho) {
  __shared__ unsigned long long box[kBoxCells];
  __shared__ int s_lo[3], s_hi[3];
  const int tid = threadIdx.x;
  const int k0 = blockIdx.x * kAssignRun;
  if (tid < 3) {
    s_lo[tid] = INT_MAX;
    s_hi[tid] = INT_MIN;
  }
  __syncthreads();
  int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
  for (int q = 0; q < kAssignPer; ++q) {
    const int k = k0 + q * kAssignThreads + tid;
    if (k >= n) break;
    const int p = order[k];
    for (int d = 0; d < 3; ++d) {
      const int g = (int)floor(pos[3 * p + d] * M + 0.5);
      lo[d] = min(lo[d], g - 1);
      hi[d] = max(hi[d], g + 1);
    }
  }
  for (int d = 0; d < 3; ++d) {
    atomicMin(&s_lo[d], lo[d]);
    atomicMax(&s_hi[d], hi[d]);
  }
  __syncthreads();
  const int bx = s_hi[0] - s_lo[0] + 1, by = s_hi[1] - s_lo[1] + 1, bz = s_hi[2] - s_lo[2] + 1;
  const bool in_lds = (long long)bx * by * bz <= kBoxCells;  // uniform over the block
  const int ox = s_lo[0], oy = s_lo[1], oz = s_lo[2];
  if (in_lds) {
    for (int t = tid; t < bx * by * bz; t += kAssignThreads) box[t] = 0ull;
    __syncthreads();
  }
  const double m3 = (double)M * (double)M * (double)M * scale[0];
  for (int q = 0; q < kAssignPer; ++q) {
    const int k = k0 + q * kAssignThreads + tid;
    if (k >= n) break;
    const int p = order[k];
    double wx[3], wy[3], wz[3];
    const int gx = tsc(pos[3 * p] * M, wx);
    const int gy = tsc(pos[3 * p + 1] * M, wy);
    const int gz = tsc(pos[3 * p + 2] * M, wz);
    const double qm = mass[p] * m3;  // mass / h^3 in fixed-point units
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        const double wab = qm * wx[a] * wy[b];
        for (int c = 0; c < 3; ++c) {
          const unsigned long long v = (unsigned long long)llrint(wab * wz[c]);
          if (in_lds)
            atomicAdd(&box[((gx + a - 1 - ox) * by + (gy + b - 1 - oy)) * bz + (gz + c - 1 - oz)],
                      v);
          else
            atomicAdd(&rho[((size_t)wrap(gx + a - 1, M) * M + wrap(gy + b - 1, M)) * M +
                           wrap(gz + c - 1, M)],
                      v);
        }
      }
  }
  if (!in_lds) return;
  __syncthreads();
  for (int t = tid; t < bx * by * bz; t += kAssignThreads) {
    const unsigned long long v = box[t];
    if (v == 0ull) continue;
    const int ix = t / (by * bz), iy = (t / bz) % by, iz = t % bz;
    atomicAdd(&rho[((size_t)wrap(ox + ix, M) * M + wrap(oy + iy, M)) * M + wrap(oz + iz, M)], v);
  }
}

// The fixed-point mesh back to FP64 in place: rho[i] = int64(bits) / scale.
__global__ void pm_fix_to_double(size_t nr, const double* __restrict__ scale, double* rho) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nr) return;
  const long long v = reinterpret_cast<const long long*>(rho)[i];
  rho[i] = (double)v / scale[0];
}

// S2 form factor of a sphere of diameter a at wave number k.
__device__ inline double s2_form(double k, double a) {
  const double u = k * a * 0.5;
  if (u < 1e-3) return 1.0 - u * u / 15.0;
  const double u2 = u * u;
  return 12.0 / (u2 * u2) * (2.0 - 2.0 * cos(u) - u * sin(u));
}

__device__ inline double sinc(double x) { return fabs(x) < 1e-12 ? 1.0 : sin(x) / x; }

// rhok: M x M x (M/2+1) complex (double2); out: 3 such arrays, the
// transformed accelerations of each axis, already scaled by 1/M^3.
__global__ void pm_green(int M, double a, const double2* __restrict__ rhok,
                         double2* __restrict__ fk) {
  const int Mz = M / 2 + 1;
  const size_t nk = (size_t)M * M * Mz;
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nk) return;
  const int iz = (int)(k % Mz);
  const int iy = (int)((k / Mz) % M);
  const int ix = (int)(k / ((size_t)Mz * M));
  const double tp = 2.0 * 3.14159265358979323846;
  const double kx = tp * (ix <= M / 2 - 1 || M == 1 ? ix : ix - M);
  const double ky = tp * (iy <= M / 2 - 1 || M == 1 ? iy : iy - M);
  const double kz = tp * iz;
  const double k2 = kx * kx + ky * ky + kz * kz;
  double2 f[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
  // the gradient of a Nyquist mode has no real counterpart: dropped
  const bool nyq = ix == M / 2 || iy == M / 2 || iz == M / 2;
  if (k2 > 0.0 && !nyq) {
    const double h2 = 0.5 / M;
    const double w = sinc(kx * h2) * sinc(ky * h2) * sinc(kz * h2);
    const double w3 = w * w * w;
    const double s = s2_form(sqrt(k2), a);
    const double g = -4.0 * 3.14159265358979323846 * s * s / (k2 * w3 * w3) /
                     ((double)M * (double)M * (double)M);
    const double2 r = rhok[k];
    const double pr = r.x * g, pi = r.y * g;  // phi_k
    const double kv[3] = {kx, ky, kz};
    for (int d = 0; d < 3; ++d) {  // -i k phi = (k pi, -k pr)
      f[d].x = kv[d] * pi;
      f[d].y = -kv[d] * pr;
    }
  }
  for (int d = 0; d < 3; ++d) fk[d * nk + k] = f[d];
}

// TSC interpolation, particles taken in the sorted order (neighbouring
// threads read neighbouring mesh points); order = null takes them as stored.
__global__ void pm_interp(int n, const int* __restrict__ order, const double* __restrict__ pos,
                          int M, const double* __restrict__ fx, const double* __restrict__ fy,
                          const double* __restrict__ fz, double* __restrict__ acc) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int p = order ? order[k] : k;
  double wx[3], wy[3], wz[3];
  const int gx = tsc(pos[3 * p] * M, wx);
  const int gy = tsc(pos[3 * p + 1] * M, wy);
  const int gz = tsc(pos[3 * p + 2] * M, wz);
  double ax = 0.0, ay = 0.0, az = 0.0;
  for (int a = 0; a < 3; ++a) {
    const size_t ix = (size_t)wrap(gx + a - 1, M) * M;
    for (int b = 0; b < 3; ++b) {
      const size_t iy = (ix + wrap(gy + b - 1, M)) * M;
      const double wab = wx[a] * wy[b];
      for (int c = 0; c < 3; ++c) {
        const size_t o = iy + wrap(gz + c - 1, M);
        const double w = wab * wz[c];
        ax += w * fx[o];
        ay += w * fy[o];
        az += w * fz[o];
      }
    }
  }
  acc[3 * p] = ax;
  acc[3 * p + 1] = ay;
  acc[3 * p + 2] = az;
}

}  // namespace pm
