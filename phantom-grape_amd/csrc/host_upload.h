// Uploads of small host buffers for the drop-in APIs: a kernel reads pinned
// host memory through its device mapping and writes device memory.  A DMA
// copy (hipMemcpyAsync) costs several microseconds of latency per call, and
// its first use in a process several milliseconds more; the drop-in calls pay
// that on every force evaluation.  Include from HIP sources only.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "pgamd_common.h"

namespace pgamd {
namespace {  // kernels with internal linkage: no exported launch stubs

__global__ void copy_in16(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    dst[k] = src[k];
}

__global__ void copy_in4(const unsigned* __restrict__ src, unsigned* __restrict__ dst, size_t n) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    dst[k] = src[k];
}

struct UploadSegs {
  const uint4* src[4];
  uint4* dst[4];
  unsigned long long n[4];
  int k;
};

__global__ void copy_in_segs(UploadSegs s) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (int q = 0; q < s.k; ++q)
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < s.n[q]; i += stride)
      s.dst[q][i] = s.src[q][i];
}

}  // namespace

// dst (device) = bytes [off, off + bytes) of the pinned buffer h; bytes is a
// multiple of 4.  Asynchronous on st: h must not change until it has run.
inline void upload(void* dst, const HostBuf& h, size_t off, size_t bytes, hipStream_t st) {
  if (bytes == 0) return;
  char* base = nullptr;
  PGAMD_HIP_CHECK(hipHostGetDevicePointer((void**)&base, h.ptr, 0));
  const char* src = base + off;
  const bool wide = (((uintptr_t)src | (uintptr_t)dst | bytes) & 15) == 0;
  const size_t n = wide ? bytes / 16 : bytes / 4;
  const int blocks = (int)std::min<size_t>(1024, (n + 255) / 256);
  if (wide)
    hipLaunchKernelGGL(copy_in16, dim3(blocks), dim3(256), 0, st,
                       reinterpret_cast<const uint4*>(src), static_cast<uint4*>(dst), n);
  else
    hipLaunchKernelGGL(copy_in4, dim3(blocks), dim3(256), 0, st,
                       reinterpret_cast<const unsigned*>(src), static_cast<unsigned*>(dst), n);
  PGAMD_HIP_CHECK(hipGetLastError());
}

// Several uploads in one launch (each launch costs microseconds of host and
// queue latency on a per-call path).  add() queues 16-byte-aligned segments
// (others are copied at once by upload()); flush() launches the copies.
class UploadBatch {
 public:
  void add(void* dst, const HostBuf& h, size_t off, size_t bytes, hipStream_t st) {
    if (bytes == 0) return;
    char* base = nullptr;
    PGAMD_HIP_CHECK(hipHostGetDevicePointer((void**)&base, h.ptr, 0));
    const char* src = base + off;
    if ((((uintptr_t)src | (uintptr_t)dst | bytes) & 15) != 0 || s_.k == 4) {
      upload(dst, h, off, bytes, st);
      return;
    }
    s_.src[s_.k] = reinterpret_cast<const uint4*>(src);
    s_.dst[s_.k] = static_cast<uint4*>(dst);
    s_.n[s_.k] = bytes / 16;
    total_ += bytes / 16;
    ++s_.k;
  }
  void flush(hipStream_t st) {
    if (s_.k == 0) return;
    const int blocks = (int)std::min<size_t>(1024, (total_ + 255) / 256);
    hipLaunchKernelGGL(copy_in_segs, dim3(blocks), dim3(256), 0, st, s_);
    PGAMD_HIP_CHECK(hipGetLastError());
    s_.k = 0;
    total_ = 0;
  }

 private:
  UploadSegs s_{{}, {}, {}, 0};
  size_t total_ = 0;
};

// Device address of a pinned host buffer (kernels write results through it).
template <class T>
T* mapped(const HostBuf& h) {
  void* d = nullptr;
  PGAMD_HIP_CHECK(hipHostGetDevicePointer(&d, h.ptr, 0));
  return static_cast<T*>(d);
}

}  // namespace pgamd
