// Particle-mesh long-range force of the TreePM split on the GPU.
//
// The table force of PG5/table (f = R(r, eps_PP) - R(r, eps_PM),
// PG5/table/libpg5/pg5_table.c:8-34) is the short-range side; this is the
// long-range side, the periodic force of R(r, eps_PM), i.e. of two S2 spheres
// of diameter a = eps_PM (Hockney & Eastwood sec. 8-3), so that PP + PM is
// the S2-softened Newtonian force of the periodic unit box.  Mesh method:
//   1. TSC assignment of the particles' density on an M^3 mesh (mesh points
//      at integer multiples of h = 1/M), FP64 atomics;
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

__global__ void pm_assign(int n, const double* __restrict__ pos, const double* __restrict__ mass,
                          int M, double* __restrict__ rho) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double wx[3], wy[3], wz[3];
  const int gx = tsc(pos[3 * p] * M, wx);
  const int gy = tsc(pos[3 * p + 1] * M, wy);
  const int gz = tsc(pos[3 * p + 2] * M, wz);
  const double q = mass[p] * (double)M * (double)M * (double)M;  // mass / h^3
  for (int a = 0; a < 3; ++a) {
    const size_t ix = (size_t)wrap(gx + a - 1, M) * M;
    for (int b = 0; b < 3; ++b) {
      const size_t iy = (ix + wrap(gy + b - 1, M)) * M;
      const double wab = q * wx[a] * wy[b];
      for (int c = 0; c < 3; ++c) atomicAdd(&rho[iy + wrap(gz + c - 1, M)], wab * wz[c]);
    }
  }
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

__global__ void pm_interp(int n, const double* __restrict__ pos, int M,
                          const double* __restrict__ fx, const double* __restrict__ fy,
                          const double* __restrict__ fz, double* __restrict__ acc) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
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
