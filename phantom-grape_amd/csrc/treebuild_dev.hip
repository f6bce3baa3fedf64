// libpgtree_dev.so: the interaction lists of the batched TreePM and
// Barnes-Hut calls, built on the GPU from positions already in HBM.
//
// Same algorithm and same output as the host builder libpgtree.so
// (treebuild.cpp), array for array: Morton keys (21 bits per axis), a stable
// radix sort, the octree split level by level, bottom-up moments, i-groups of
// <= ncrit particles in Morton order, TreePM ghosts and grid-based j-lists
// in grid-cell order, the Barnes-Hut walk with the opening test of
// PG5/newton_quad/source.cpp:660-707, and groups ordered by decreasing cost.
// Every FP64 operation that decides or produces a value is evaluated in the
// host builder's order with contraction off, so the two agree bit for bit;
// only the numbering of the Barnes-Hut cells differs (here: the order of the
// octree levels), so cell lists match as lists of cell contents.
//
// The build is synchronous on the caller's stream: list lengths are read back
// between phases to size the next allocation.
#include "pgamd_common.h"
#include "pgamd_api.h"

#include <hipcub/device/device_radix_sort.hpp>
#include <hipcub/device/device_scan.hpp>

#include <algorithm>
#include <climits>
#include <cmath>
#include <map>
#include <unordered_map>
#include <vector>

using namespace pgamd;

namespace {

using u64 = unsigned long long;
constexpr int kLevels = 21;  // bits per axis of the Morton key

// ------------------------------------------------------------ allocations
// Device memory of the builds comes from a cache: blocks rounded up to powers
// of two and kept for reuse once released, so that repeated builds make no
// hipMalloc / hipFree calls (hipFree also waits for the whole device).  A
// build releases its temporaries after its stream has drained, lists release
// theirs in pgt_dev_free after a device synchronisation; both run under the
// library's mutex.  The cache holds its peak size until the process ends.
struct Cache {
  std::map<size_t, std::vector<void*>> free;
  std::unordered_map<void*, size_t> size;
  void* get(size_t bytes) {
    size_t b = 256;
    while (b < bytes) b <<= 1;
    std::vector<void*>& v = free[b];
    if (!v.empty()) {
      void* p = v.back();
      v.pop_back();
      return p;
    }
    void* p = nullptr;
    PGAMD_HIP_CHECK(hipMalloc(&p, b));
    size[p] = b;
    return p;
  }
  void put(void* p) { free[size.at(p)].push_back(p); }
};
Cache g_cache;

struct Mem {
  std::vector<void*> ptrs;
  template <class T>
  T* get(size_t n) {
    void* p = g_cache.get(std::max<size_t>(n, 1) * sizeof(T));
    ptrs.push_back(p);
    return static_cast<T*>(p);
  }
  void free_all() {
    for (void* p : ptrs) g_cache.put(p);
    ptrs.clear();
  }
};

// out[n] = out[n - 1] + in[n - 1] (0 for n = 0): the total of an exclusive
// scan, also written to *host (mapped pinned memory).
__global__ void k_scan_total(const int* __restrict__ in, int* __restrict__ out, int n,
                             int* __restrict__ host) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const int t = n > 0 ? out[n - 1] + in[n - 1] : 0;
    out[n] = t;
    *host = t;
  }
}

// Exclusive scan of n <= kSmallScan ints in one block: thread q sums its
// consecutive chunk, the block scans the chunk sums, thread q writes its
// chunk's prefixes.  out[n] = total, also written to *host.
constexpr int kSmallScanThreads = 1024;
constexpr int kSmallScan = 4 * kSmallScanThreads;
__global__ __launch_bounds__(kSmallScanThreads) void k_scan_small(const int* __restrict__ in,
                                                                  int* __restrict__ out, int n,
                                                                  int* __restrict__ host) {
  __shared__ int s[kSmallScanThreads];
  const int q = threadIdx.x;
  const int per = (n + kSmallScanThreads - 1) / kSmallScanThreads;
  const int a = min(n, q * per), b = min(n, a + per);
  int sum = 0;
  for (int i = a; i < b; ++i) sum += in[i];
  s[q] = sum;
  __syncthreads();
  for (int d = 1; d < kSmallScanThreads; d <<= 1) {  // inclusive Hillis-Steele scan
    const int v = q >= d ? s[q - d] : 0;
    __syncthreads();
    s[q] += v;
    __syncthreads();
  }
  int run = s[q] - sum;  // exclusive prefix of this chunk
  for (int i = a; i < b; ++i) {
    out[i] = run;
    run += in[i];
  }
  if (q == kSmallScanThreads - 1) {
    out[n] = s[q];
    *host = s[q];
  }
}

// Result of one build: device arrays in the host builder's layout.
struct Lists {
  Mem mem;
  long long np = 0, ng = 0, ni = 0, nj = 0, nc = 0, ncells = 0, work_pp = 0, work_pc = 0;
  double* pos = nullptr;  // [np][3]
  int* src = nullptr;
  int *ioff = nullptr, *icnt = nullptr, *ilist = nullptr;
  int *joff = nullptr, *jcnt = nullptr, *jlist = nullptr;
  int *coff = nullptr, *ccnt = nullptr, *clist = nullptr;
  int* order = nullptr;
  double* cells = nullptr;  // [ncells][10]
};

// One pinned int per process for scan totals; builds run one at a time under
// the library's mutex.
int* scan_total_host() {
  static int* p = nullptr;
  if (!p) PGAMD_HIP_CHECK(hipHostMalloc((void**)&p, sizeof(int), hipHostMallocMapped));
  return p;
}

// Temporary storage for one build, released at its end.
struct Build {
  hipStream_t st;
  Mem tmp;
  void* cub = nullptr;
  size_t cub_bytes = 0;
  explicit Build(hipStream_t s) : st(s) {}
  ~Build() {
    PGAMD_HIP_CHECK(hipStreamSynchronize(st));
    if (cub) g_cache.put(cub);
    tmp.free_all();
  }
  void* cub_tmp(size_t bytes) {
    if (bytes > cub_bytes) {
      PGAMD_HIP_CHECK(hipStreamSynchronize(st));  // the old block may still be in use
      if (cub) g_cache.put(cub);
      cub = g_cache.get(bytes);
      cub_bytes = bytes;
    }
    return cub;
  }
  template <class T>
  T read(const T* d) {
    T v;
    PGAMD_HIP_CHECK(hipMemcpyAsync(&v, d, sizeof(T), hipMemcpyDeviceToHost, st));
    PGAMD_HIP_CHECK(hipStreamSynchronize(st));
    return v;
  }
  // out[k] = sum in[0..k), out[n] = total (out has n + 1 entries); returns
  // total.  The total comes back through mapped pinned memory: one kernel and
  // a synchronisation for small n, the hipCUB scan for large n.
  int scan(const int* in, int* out, int n) {
    int* host = scan_total_host();
    int* dev = nullptr;
    PGAMD_HIP_CHECK(hipHostGetDevicePointer((void**)&dev, host, 0));
    if (n <= kSmallScan) {
      hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(kSmallScanThreads), 0, st, in, out, n, dev);
    } else {
      size_t b = 0;
      PGAMD_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, b, in, out, n, st));
      PGAMD_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(cub_tmp(b), b, in, out, n, st));
      hipLaunchKernelGGL(k_scan_total, dim3(1), dim3(64), 0, st, in, out, n, dev);
    }
    PGAMD_HIP_CHECK(hipGetLastError());
    PGAMD_HIP_CHECK(hipStreamSynchronize(st));
    return *static_cast<volatile int*>(host);
  }
  // Stable radix sort of (key, value) pairs; int keys must be >= 0.
  void sort_pairs(const int* kin, int* kout, const int* vin, int* vout, int n, int end_bit) {
    sort_pairs(reinterpret_cast<const unsigned*>(kin), reinterpret_cast<unsigned*>(kout), vin,
               vout, n, end_bit);
  }
  template <class K>
  void sort_pairs(const K* kin, K* kout, const int* vin, int* vout, int n, int end_bit) {
    if (n <= 0) return;
    size_t b = 0;
    PGAMD_HIP_CHECK(
        hipcub::DeviceRadixSort::SortPairs(nullptr, b, kin, kout, vin, vout, n, 0, end_bit, st));
    PGAMD_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(cub_tmp(b), b, kin, kout, vin, vout, n, 0,
                                                       end_bit, st));
  }
};

// Bits needed to hold every value in [0, v].
inline int bits_for(long long v) {
  int b = 1;
  while ((1LL << b) <= v) ++b;
  return b;
}

inline dim3 grid_for(long long n, int threads = 256) {
  return dim3((unsigned)std::max<long long>(1, (n + threads - 1) / threads));
}

// ------------------------------------------------------------ Morton keys
__device__ inline u64 spread3(u64 v) {
  v &= 0x1fffffULL;
  v = (v | v << 32) & 0x1f00000000ffffULL;
  v = (v | v << 16) & 0x1f0000ff0000ffULL;
  v = (v | v << 8) & 0x100f00f00f00f00fULL;
  v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
  v = (v | v << 2) & 0x1249249249249249ULL;
  return v;
}

__global__ void k_keys(int n, const double* __restrict__ x, double lo0, double lo1, double lo2,
                       double s, u64* __restrict__ key, int* __restrict__ idx) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double lo[3] = {lo0, lo1, lo2};
  u64 k = 0;
  for (int d = 0; d < 3; ++d) {
    const double u = (x[3 * i + d] - lo[d]) * s;
    long long q = (long long)floor(u);
    q = q < 0 ? 0 : (q > (1 << kLevels) - 1 ? (1 << kLevels) - 1 : q);
    k |= spread3((u64)q) << (2 - d);
  }
  key[i] = k;
  idx[i] = i;
}

// Bounding box: per-block min/max of each axis, then one block over those.
__global__ void k_bbox_part(int n, const double* __restrict__ x, double* __restrict__ part) {
  __shared__ double s[6][256];
  double v[6] = {1e300, 1e300, 1e300, -1e300, -1e300, -1e300};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    for (int d = 0; d < 3; ++d) {
      v[d] = fmin(v[d], x[3 * i + d]);
      v[3 + d] = fmax(v[3 + d], x[3 * i + d]);
    }
  for (int k = 0; k < 6; ++k) s[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
      for (int k = 0; k < 6; ++k)
        s[k][threadIdx.x] = k < 3 ? fmin(s[k][threadIdx.x], s[k][threadIdx.x + h])
                                  : fmax(s[k][threadIdx.x], s[k][threadIdx.x + h]);
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 6; ++k) part[6 * blockIdx.x + k] = s[k][0];
}

// One wavefront over the per-block results (minima and maxima do not depend
// on the order).
__global__ void k_bbox_final(int nb, const double* __restrict__ part, double* __restrict__ out) {
  const int lane = threadIdx.x;
  if (blockIdx.x != 0 || lane >= 64) return;
  double v[6] = {1e300, 1e300, 1e300, -1e300, -1e300, -1e300};
  for (int b = lane; b < nb; b += 64)
    for (int k = 0; k < 6; ++k)
      v[k] = k < 3 ? fmin(v[k], part[6 * b + k]) : fmax(v[k], part[6 * b + k]);
  for (int off = 32; off > 0; off >>= 1)
    for (int k = 0; k < 6; ++k) {
      const double o = __shfl_xor(v[k], off);
      v[k] = k < 3 ? fmin(v[k], o) : fmax(v[k], o);
    }
  if (lane == 0)
    for (int k = 0; k < 6; ++k) out[k] = v[k];
}

// Particles in Morton order: pos[p] = x[perm[p]], mass likewise (optional).
__global__ void k_gather(int n, const int* __restrict__ perm, const double* __restrict__ x,
                         const double* __restrict__ m, double* __restrict__ pos,
                         double* __restrict__ mm, int* __restrict__ src) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int i = perm[p];
  for (int d = 0; d < 3; ++d) pos[3 * p + d] = x[3 * i + d];
  if (m) mm[p] = m[i];
  src[p] = i;
}

// ------------------------------------------------------------ octree
// Nodes in level order.  Geometry: centre c (FP64) and side L / 2^level.
struct Nodes {
  int *lo, *hi, *par, *lev, *child;  // child[8 * k + octant], -1 if empty
  double* c;                         // [k][3]
  double *com, *m, *q;               // [k][3], [k], [k][6]
  int* cell;                         // cell index, -1 if none
};

__device__ inline int digit(u64 key, int shift) { return (int)((key >> shift) & 7ULL); }

// First index in [a, b) whose octant digit at `shift` is >= o (digits are
// non-decreasing inside a node).
__device__ inline int lower_digit(const u64* key, int a, int b, int shift, int o) {
  while (a < b) {
    const int mid = a + (b - a) / 2;
    if (digit(key[mid], shift) < o)
      a = mid + 1;
    else
      b = mid;
  }
  return a;
}

// Octant ranges of the nodes [a, b) of one level: eight threads per node,
// thread o finding where octant o ends (the first index of the node's range
// whose octant digit exceeds o), so the eight binary searches run side by
// side.  end[8 * (k - a) + o] keeps the ends for k_split_fill; nchild[k - a]
// = the number of non-empty octants (0 for a node that is not split).
__global__ void k_split_count(int a, int b, Nodes nd, const u64* __restrict__ key, int split,
                              int* __restrict__ end, int* __restrict__ nchild) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int k = a + (int)(t >> 3), o = (int)(t & 7);
  int lo = 0, e = 0;
  bool splits = false;
  if (k < b) {
    const int hi = nd.hi[k], lev = nd.lev[k];
    lo = nd.lo[k];
    e = lo;
    splits = hi - lo > split && lev < kLevels;
    if (splits) e = o == 7 ? hi : lower_digit(key, lo, hi, 3 * (kLevels - 1 - lev), o + 1);
    end[8 * (size_t)(k - a) + o] = e;
  }
  // a node's eight threads are one aligned group of eight lanes: octant o
  // starts where octant o - 1 ends
  const int lane = threadIdx.x & 63;
  const int prev = __shfl(e, (lane + 63) & 63);
  const unsigned long long m = __ballot(splits && e > (o == 0 ? lo : prev));
  if (k < b && o == 0) nchild[k - a] = __popcll((m >> lane) & 0xffULL);
}

// Children of the nodes [a, b) from the octant ends of k_split_count: eight
// threads per node, thread o writing child o (numbered after the level, in
// node order, then octant order).
__global__ void k_split_fill(int a, int b, Nodes nd, const int* __restrict__ end,
                             const int* __restrict__ off, double L) {
#pragma clang fp contract(off)
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int k = a + (int)(t >> 3), o = (int)(t & 7);
  if (k >= b) return;
  const int* e = end + 8 * (size_t)(k - a);
  const int lo = nd.lo[k], lev = nd.lev[k];
  const int s0 = o == 0 ? lo : e[o - 1], s1 = e[o];
  if (s1 <= s0) return;
  int rank = 0;  // non-empty octants before o
  for (int q = 0; q < o; ++q) rank += e[q] > (q == 0 ? lo : e[q - 1]);
  const int id = b + off[k - a] + rank;
  const double l = ldexp(L, -lev);  // L halved lev times, exactly
  nd.lo[id] = s0;
  nd.hi[id] = s1;
  nd.par[id] = k;
  nd.lev[id] = lev + 1;
  for (int q = 0; q < 8; ++q) nd.child[8 * id + q] = -1;
  // octant bits: x is bit 2, y bit 1, z bit 0
  nd.c[3 * id + 0] = nd.c[3 * k + 0] + ((o & 4) ? 0.25 : -0.25) * l;
  nd.c[3 * id + 1] = nd.c[3 * k + 1] + ((o & 2) ? 0.25 : -0.25) * l;
  nd.c[3 * id + 2] = nd.c[3 * k + 2] + ((o & 1) ? 0.25 : -0.25) * l;
  nd.child[8 * k + o] = id;
}

// Moments of the nodes [a, b) of one level, children already done.
__global__ void k_moments(int a, int b, Nodes nd, const double* __restrict__ pos,
                          const double* __restrict__ mm) {
#pragma clang fp contract(off)
  const int k = a + blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= b) return;
  bool leaf = true;
  for (int o = 0; o < 8; ++o) leaf = leaf && nd.child[8 * k + o] < 0;
  double M = 0, c[3] = {0, 0, 0}, q[6] = {0, 0, 0, 0, 0, 0}, com[3];
  if (leaf) {
    for (int p = nd.lo[k]; p < nd.hi[k]; ++p) {
      const double mi = mm ? mm[p] : 1.0;
      M += mi;
      for (int d = 0; d < 3; ++d) c[d] += mi * pos[3 * p + d];
    }
    for (int d = 0; d < 3; ++d) com[d] = M > 0 ? c[d] / M : nd.c[3 * k + d];
    for (int p = nd.lo[k]; p < nd.hi[k]; ++p) {
      const double mi = mm ? mm[p] : 1.0;
      double dd[3];
      for (int d = 0; d < 3; ++d) dd[d] = pos[3 * p + d] - com[d];
      const double r2 = dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2];
      int kk = 0;
      for (int x = 0; x < 2; ++x)
        for (int y = x; y < 3; ++y) {
          double v = 3.0 * dd[x] * dd[y];
          if (x == y) v -= r2;
          q[kk++] += mi * v;
        }
    }
  } else {
    for (int o = 0; o < 8; ++o) {
      const int ch = nd.child[8 * k + o];
      if (ch < 0) continue;
      M += nd.m[ch];
      for (int d = 0; d < 3; ++d) c[d] += nd.m[ch] * nd.com[3 * ch + d];
    }
    for (int d = 0; d < 3; ++d) com[d] = M > 0 ? c[d] / M : nd.c[3 * k + d];
    for (int o = 0; o < 8; ++o) {
      const int ch = nd.child[8 * k + o];
      if (ch < 0) continue;
      double dd[3];
      for (int d = 0; d < 3; ++d) dd[d] = nd.com[3 * ch + d] - com[d];
      const double r2 = dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2];
      int kk = 0;
      for (int x = 0; x < 2; ++x)
        for (int y = x; y < 3; ++y) {
          double v = 3.0 * dd[x] * dd[y];
          if (x == y) v -= r2;
          q[kk] += nd.q[6 * ch + kk] + nd.m[ch] * v;
          ++kk;
        }
    }
  }
  nd.m[k] = M;
  for (int d = 0; d < 3; ++d) nd.com[3 * k + d] = com[d];
  q[5] = -(q[0] + q[3]);
  for (int t = 0; t < 6; ++t) nd.q[6 * k + t] = q[t];
}

// i-groups: nodes of <= ncrit particles whose parent holds more, and nodes
// the key cannot split (level kLevels: more than ncrit particles inside one
// cell of side L / 2^21), which form one group of all their particles.
__global__ void k_group_flag(int nn, Nodes nd, int ncrit, int* __restrict__ flag) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nn) return;
  const int cnt = nd.hi[k] - nd.lo[k];
  const int par = nd.par[k];
  flag[k] = (cnt <= ncrit || nd.lev[k] >= kLevels) &&
            (par < 0 || nd.hi[par] - nd.lo[par] > ncrit);
}

__global__ void k_compact(int n, const int* __restrict__ flag, const int* __restrict__ off,
                          int* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n && flag[k]) out[off[k]] = k;
}

__global__ void k_group_lo(int ng, const int* __restrict__ gnode, Nodes nd, int* __restrict__ lo) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g < ng) lo[g] = nd.lo[gnode[g]];
}

__global__ void k_group_icnt(int ng, const int* __restrict__ gnode, Nodes nd,
                             int* __restrict__ icnt) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g < ng) icnt[g] = nd.hi[gnode[g]] - nd.lo[gnode[g]];
}

__global__ void k_group_ilist(int ng, const int* __restrict__ gnode, Nodes nd,
                              const int* __restrict__ ioff, int* __restrict__ ilist) {
  const int g = blockIdx.x;
  const int lo = nd.lo[gnode[g]], hi = nd.hi[gnode[g]];
  for (int p = lo + threadIdx.x; p < hi; p += blockDim.x) ilist[ioff[g] + p - lo] = p;
}

// ------------------------------------------------------------ TreePM
__global__ void k_ghost_count(int n, const double* __restrict__ pos, double rcut,
                              int* __restrict__ cnt) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int c = 1;
  for (int d = 0; d < 3; ++d) {
    const double v = pos[3 * p + d];
    c *= 1 + (v >= 1.0 - rcut ? 1 : 0) + (v < rcut ? 1 : 0);
  }
  cnt[p] = c - 1;
}

__global__ void k_ghost_fill(int n, const int* __restrict__ goff, double rcut,
                             double* __restrict__ pos, int* __restrict__ src) {
#pragma clang fp contract(off)
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double x0 = pos[3 * p], x1 = pos[3 * p + 1], x2 = pos[3 * p + 2];
  int lo_s[3], hi_s[3];
  const double xp[3] = {x0, x1, x2};
  for (int d = 0; d < 3; ++d) {
    lo_s[d] = xp[d] >= 1.0 - rcut ? -1 : 0;
    hi_s[d] = xp[d] < rcut ? 1 : 0;
  }
  int o = n + goff[p];
  for (int sx = lo_s[0]; sx <= hi_s[0]; ++sx)
    for (int sy = lo_s[1]; sy <= hi_s[1]; ++sy)
      for (int sz = lo_s[2]; sz <= hi_s[2]; ++sz) {
        if (!sx && !sy && !sz) continue;
        pos[3 * o + 0] = x0 + sx;
        pos[3 * o + 1] = x1 + sy;
        pos[3 * o + 2] = x2 + sz;
        src[o] = src[p];
        ++o;
      }
}

struct Grid {
  double glo, h;
  int nc;
};

__device__ inline int cidx(const Grid& g, double v) {
#pragma clang fp contract(off)
  int c = (int)floor((v - g.glo) / g.h);
  return c < 0 ? 0 : (c > g.nc - 1 ? g.nc - 1 : c);
}

__global__ void k_cell_of(int np, const double* __restrict__ pos, Grid g, int* __restrict__ cell,
                          int* __restrict__ idx) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  cell[p] = (cidx(g, pos[3 * p]) * g.nc + cidx(g, pos[3 * p + 1])) * g.nc +
            cidx(g, pos[3 * p + 2]);
  idx[p] = p;
}

// Positions in cell order: cpos[k] = pos[cpart[k]], so the list walk reads
// each row of cells as one contiguous range.
__global__ void k_cell_pos(int np, const int* __restrict__ cpart, const double* __restrict__ pos,
                           double* __restrict__ cpos) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  const int p = cpart[k];
  for (int d = 0; d < 3; ++d) cpos[3 * k + d] = pos[3 * p + d];
}

// cstart[c] = first position of cell c in the sorted cell array (c <= ncell).
__global__ void k_cell_start(int ncell, int np, const int* __restrict__ scell,
                             int* __restrict__ cstart) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c > ncell) return;
  int a = 0, b = np;
  while (a < b) {
    const int mid = a + (b - a) / 2;
    if (scell[mid] < c)
      a = mid + 1;
    else
      b = mid;
  }
  cstart[c] = a;
}

// Group bounding boxes: [g][6] = min xyz, max xyz of the group's particles;
// one wavefront per group (minima and maxima do not depend on the order).
__global__ __launch_bounds__(256) void k_group_box(int ng, const int* __restrict__ gnode, Nodes nd,
                                                   const double* __restrict__ pos,
                                                   double* __restrict__ box) {
  const int lane = threadIdx.x & 63;
  const int g = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (g >= ng) return;  // uniform per wavefront
  const int k = gnode[g], lo = nd.lo[k], hi = nd.hi[k];
  double bmin[3] = {1e30, 1e30, 1e30}, bmax[3] = {-1e30, -1e30, -1e30};
  for (int p = lo + lane; p < hi; p += 64)
    for (int d = 0; d < 3; ++d) {
      bmin[d] = fmin(bmin[d], pos[3 * p + d]);
      bmax[d] = fmax(bmax[d], pos[3 * p + d]);
    }
  for (int off = 32; off > 0; off >>= 1)
    for (int d = 0; d < 3; ++d) {
      bmin[d] = fmin(bmin[d], __shfl_xor(bmin[d], off));
      bmax[d] = fmax(bmax[d], __shfl_xor(bmax[d], off));
    }
  if (lane < 3) {
    box[6 * g + lane] = lane == 0 ? bmin[0] : (lane == 1 ? bmin[1] : bmin[2]);
    box[6 * g + 3 + lane] = lane == 0 ? bmax[0] : (lane == 1 ? bmax[1] : bmax[2]);
  }
}

__device__ inline bool in_reach(const double* __restrict__ box, const double* __restrict__ v,
                                double rc2) {
#pragma clang fp contract(off)
  double d2 = 0;
  for (int d = 0; d < 3; ++d) {
    const double e = v[d] < box[d] ? box[d] - v[d] : (v[d] > box[3 + d] ? v[d] - box[3 + d] : 0.0);
    d2 += e * e;
  }
  return d2 < rc2;
}

// One wavefront per group walks the grid cells around the group box and
// writes the particles in reach to jlist[joff[g]...] and their number to
// jcnt[g], in cell order (cells by x, y, z index; a cell's particles by
// index).  The cells
// of one (x, y) row are consecutive, so a row is one range of the cell-sorted
// particles (cpos: positions in that order).
__global__ __launch_bounds__(256) void k_treepm_jlist(int ng, const double* __restrict__ box,
                                                      const double* __restrict__ cpos, Grid gr,
                                                      const int* __restrict__ cstart,
                                                      const int* __restrict__ cpart, double rcut,
                                                      int* __restrict__ jcnt,
                                                      const int* __restrict__ joff,
                                                      int* __restrict__ jlist) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63;
  const int g = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (g >= ng) return;  // uniform per wavefront
  const double* bx = box + 6 * g;
  const double rc2 = rcut * rcut;
  int c0[3], c1[3];
  for (int d = 0; d < 3; ++d) {
    c0[d] = cidx(gr, bx[d] - rcut);
    c1[d] = cidx(gr, bx[3 + d] + rcut);
  }
  int count = 0;
  const int base = joff[g];
  for (int a = c0[0]; a <= c1[0]; ++a)
    for (int b = c0[1]; b <= c1[1]; ++b) {
      const int row = (a * gr.nc + b) * gr.nc;
      const int k0 = cstart[row + c0[2]], k1 = cstart[row + c1[2] + 1];
      for (int k = k0; k < k1; k += 64) {
        const int kk = k + lane;
        const bool in = kk < k1 && in_reach(bx, cpos + 3 * (size_t)kk, rc2);
        const u64 mask = __ballot(in);
        if (in) {
          const int below = __popcll(mask & ((1ULL << lane) - 1ULL));
          jlist[base + count + below] = cpart[kk];
        }
        count += __popcll(mask);
      }
    }
  if (lane == 0) jcnt[g] = count;
}

// Candidates of each group: the particles in the rows of grid cells its walk
// visits (an upper bound of its list length).  One wavefront per group.
__global__ __launch_bounds__(256) void k_treepm_cand(int ng, const double* __restrict__ box,
                                                     Grid gr, const int* __restrict__ cstart,
                                                     double rcut, int* __restrict__ ccnt) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63;
  const int g = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (g >= ng) return;  // uniform per wavefront
  const double* bx = box + 6 * g;
  int c0[3], c1[3];
  for (int d = 0; d < 3; ++d) {
    c0[d] = cidx(gr, bx[d] - rcut);
    c1[d] = cidx(gr, bx[3 + d] + rcut);
  }
  const int ny = c1[1] - c0[1] + 1, nrows = (c1[0] - c0[0] + 1) * ny;
  int sum = 0;
  for (int r = lane; r < nrows; r += 64) {
    const int row = ((c0[0] + r / ny) * gr.nc + c0[1] + r % ny) * gr.nc;
    sum += cstart[row + c1[2] + 1] - cstart[row + c0[2]];
  }
  for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
  if (lane == 0) ccnt[g] = sum;
}

// jlist[joff[g] + k] = scratch[soff[g] + k] for k < jcnt[g]: the lists
// written at their candidate offsets, packed.  One wavefront per group.
__global__ __launch_bounds__(256) void k_pack_lists(int ng, const int* __restrict__ soff,
                                                    const int* __restrict__ jcnt,
                                                    const int* __restrict__ joff,
                                                    const int* __restrict__ scratch,
                                                    int* __restrict__ jlist) {
  const int lane = threadIdx.x & 63;
  const int g = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (g >= ng) return;
  const int* src = scratch + soff[g];
  int* dst = jlist + joff[g];
  for (int k = lane; k < jcnt[g]; k += 64) dst[k] = src[k];
}

// ------------------------------------------------------------ Barnes-Hut
__global__ void k_cell_flag(int nn, Nodes nd, int* __restrict__ flag) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nn) flag[k] = nd.hi[k] - nd.lo[k] > 1;
}

__global__ void k_cell_fill(int nn, Nodes nd, const int* __restrict__ flag,
                            const int* __restrict__ off, double* __restrict__ cells) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nn) return;
  if (!flag[k]) {
    nd.cell[k] = -1;
    return;
  }
  const int c = off[k];
  nd.cell[k] = c;
  double* o = cells + 10 * (size_t)c;
  o[0] = nd.com[3 * k];
  o[1] = nd.com[3 * k + 1];
  o[2] = nd.com[3 * k + 2];
  o[3] = nd.m[k];
  for (int t = 0; t < 6; ++t) o[4 + t] = nd.q[6 * k + t];
}

__device__ inline double bh_box_d2(const double* __restrict__ gc, double li,
                                   const double* __restrict__ c) {
#pragma clang fp contract(off)
  double d2 = 0;
  for (int d = 0; d < 3; ++d) {
    const double e = fmax(fabs(c[d] - gc[d]) - li, 0.0);
    d2 += e * e;
  }
  return d2;
}

// Node data the Barnes-Hut walk reads for one child, in one 64-byte record:
// centre of mass, particle range, cell index (nodes of more than one
// particle) and the children as the first child's index plus an octant mask
// (the children of a node are numbered consecutively in octant order).
struct WalkNode {
  int lo, hi, cell, first, mask, pad[3];
  double com[3], pad2;
};
static_assert(sizeof(WalkNode) == 64, "one record per 64 bytes");

__global__ void k_walk_nodes(int nn, Nodes nd, WalkNode* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nn) return;
  WalkNode w;
  for (int d = 0; d < 3; ++d) w.com[d] = nd.com[3 * k + d];
  w.lo = nd.lo[k];
  w.hi = nd.hi[k];
  w.cell = w.hi - w.lo > 1 ? nd.cell[k] : -1;
  w.first = -1;
  w.mask = 0;
  for (int o = 7; o >= 0; --o) {
    const int ch = nd.child[8 * k + o];
    if (ch >= 0) {
      w.first = ch;
      w.mask |= 1 << o;
    }
  }
  w.pad[0] = w.pad[1] = w.pad[2] = 0;
  w.pad2 = 0.0;
  out[k] = w;
}

// Depth of the walk stack: the depth-first walk keeps at most 7 pending
// siblings per level below the root, plus the 8 children of the node being
// expanded (internal nodes sit above level kLevels).
constexpr int kWalkDepth = 7 * (kLevels - 1) + 8;
constexpr int kWalkWaves = 4;  // groups (one wavefront each) per block

// Block b of a launch of nb blocks runs on XCD b % 8; renumber the blocks so
// that each XCD takes one contiguous range of groups, whose walks visit
// nearby nodes through that XCD's L2.
__device__ inline int xcd_block(int b, int nb) {
  const int x = b % 8, i = b / 8, q = nb / 8, r = nb % 8;
  return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// One wavefront per group: the depth-first walk of the host builder
// (source.cpp:832-867 order: the children of an opened node are examined in
// octant order 7..0, emitting singletons and accepted cells in that order and
// pushing the rest, so octant 0 is expanded first), counting (FILL == false)
// or writing the particle and cell lists.  Each step examines the children
// of the top 8 stack entries at once, one lane per (entry, octant), and keeps
// the work of the entries the serial walk would finish before the first one
// that pushes children: a run of entries whose children are all listed,
// ending at one that opens some, so the lists and the stack come out exactly
// as in the serial walk.  Ballots rank the emissions and pushes in the serial
// order.  A pushed entry holds what its expansion needs (the node's children
// and level, or the particle range of a leaf), so a step costs one round of
// global loads: the children's records.
template <bool FILL>
__global__ __launch_bounds__(64 * kWalkWaves) void k_bh_walk(
    int ng, const int* __restrict__ gnode, Nodes nd, const WalkNode* __restrict__ wn, double L,
    double th2, int nleaf, int* __restrict__ pcnt, int* __restrict__ ccnt,
    const int* __restrict__ poff, const int* __restrict__ coff, int* __restrict__ plist,
    int* __restrict__ clist) {
#pragma clang fp contract(off)
  // entry: {first child, octant mask, level} of a node to expand, or
  // {lo, hi, -1} of a leaf whose particles are listed
  __shared__ int3 s_stack[kWalkWaves][kWalkDepth];
  const int lane = threadIdx.x & 63;
  const int g = xcd_block(blockIdx.x, gridDim.x) * kWalkWaves + (threadIdx.x >> 6);
  if (g >= ng) return;  // uniform per wavefront
  int3* stk = s_stack[threadIdx.x >> 6];
  const int gk = gnode[g];
  const double gc[3] = {nd.c[3 * gk], nd.c[3 * gk + 1], nd.c[3 * gk + 2]};
  const double li = 0.5 * ldexp(L, -nd.lev[gk]);
  int np = 0, nc = 0;
  int* pl = FILL ? plist + poff[g] : nullptr;
  int* cl = FILL ? clist + coff[g] : nullptr;
  // the whole tree as one cell (source.cpp:832-840)
  if (bh_box_d2(gc, li, wn[0].com) * th2 >= L * L) {
    if (FILL && lane == 0) cl[0] = wn[0].cell;
    nc = 1;
  } else {
    if (lane == 0)
      stk[0] = wn[0].hi - wn[0].lo <= nleaf || wn[0].mask == 0
                   ? make_int3(wn[0].lo, wn[0].hi, -1)
                   : make_int3(wn[0].first, wn[0].mask, 0);
    int sp = 1;
    while (sp > 0) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      // lanes 8e..8e+7 take entry e from the top (e = 0 is the top), octant
      // o = lane % 8 of it
      const int npop = sp < 8 ? sp : 8;
      const int e = lane >> 3, o = lane & 7;
      int3 ent = make_int3(0, 0, 0);
      if (e < npop) ent = stk[sp - 1 - e];
      const bool leaf = e < npop && ent.z < 0;
      int kind = 0;  // 1 particle, 2 cell, 3 push
      int lo = 0, hi = 0, cell = 0, first = 0, mask = 0;
      if (e < npop && !leaf && ((ent.y >> o) & 1)) {
        const double l = ldexp(L, -ent.z);
        const double s2 = 0.25 * l * l;  // child side squared
        const WalkNode* w = wn + ent.x + __popc(ent.y & ((1 << o) - 1));
        const int4 r1 = *reinterpret_cast<const int4*>(w);
        lo = r1.x;
        hi = r1.y;
        cell = r1.z;
        first = r1.w;
        mask = w->mask;
        if (hi - lo == 1)
          kind = 1;
        else if (bh_box_d2(gc, li, w->com) * th2 > s2)
          kind = 2;
        else
          kind = 3;
      }
      // The serial walk would expand the entries from the top until one
      // pushes children; those entries (ncommit of them) are done here, the
      // rest stay on the stack.
      const u64 ms = __ballot(kind == 3);
      const int ncommit = ms ? (__builtin_ctzll(ms) >> 3) + 1 : npop;
      const bool commit = e < ncommit;
      const u64 mp = __ballot(commit && kind == 1), mc = __ballot(commit && kind == 2);
      const u64 mine = 0xffULL << (8 * e);           // this entry's lanes
      const u64 before = (1ULL << (8 * e)) - 1ULL;   // entries above this one
      const u64 above = ~((2ULL << lane) - 1ULL);    // lanes > this one
      // emission order: entries from the top; octants 7..0 inside an entry;
      // a leaf entry lists its particles lo..hi-1
      const int leafcnt = commit && leaf && o == 0 ? ent.y - ent.x : 0;
      int leafpre = 0, leaftot = 0;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int c = __builtin_amdgcn_readlane(leafcnt, 8 * q);
        leafpre += q < e ? c : 0;
        leaftot += c;
      }
      const int pbase = np + leafpre + __popcll(mp & before);
      if (FILL && commit && kind == 1) pl[pbase + __popcll(mp & mine & above)] = lo;
      if (FILL && commit && leaf)
        for (int p = ent.x + o; p < ent.y; p += 8) pl[pbase + (p - ent.x)] = p;
      if (FILL && commit && kind == 2) cl[nc + __popcll(mc & (before | (mine & above)))] = cell;
      const u64 mpush = ms & (0xffULL << (8 * (ncommit - 1)));  // the last committed entry
      sp -= ncommit;
      if (kind == 3 && commit)
        stk[sp + __popcll(mpush & above)] = hi - lo <= nleaf || mask == 0
                                                ? make_int3(lo, hi, -1)
                                                : make_int3(first, mask, ent.z + 1);
      np += __popcll(mp) + leaftot;
      nc += __popcll(mc);
      sp += __popcll(mpush);
    }
  }
  if (!FILL && lane == 0) {
    pcnt[g] = np;
    ccnt[g] = nc;
  }
}

// ------------------------------------------------------------ common tail
// cost key: groups by decreasing icnt * (jcnt + 2 ccnt), ties by index.
__global__ void k_cost(int ng, const int* __restrict__ icnt, const int* __restrict__ jcnt,
                       const int* __restrict__ ccnt, u64* __restrict__ key, int* __restrict__ idx,
                       unsigned long long* __restrict__ work) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ng) return;
  const long long c = (long long)icnt[g] * (jcnt[g] + (ccnt ? 2 * ccnt[g] : 0));
  key[g] = (u64)(LLONG_MAX - c);
  idx[g] = g;
  atomicAdd(&work[0], (unsigned long long)((long long)icnt[g] * jcnt[g]));
  if (ccnt) atomicAdd(&work[1], (unsigned long long)((long long)icnt[g] * ccnt[g]));
}

__global__ void k_pack(int np, const double* __restrict__ pos, const int* __restrict__ src,
                       const double* __restrict__ mass, float scale, float4* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  float4 v;
  v.x = (float)pos[3 * p] * scale;
  v.y = (float)pos[3 * p + 1] * scale;
  v.z = (float)pos[3 * p + 2] * scale;
  v.w = (float)mass[src[p]];
  out[p] = v;
}

__global__ void k_pack_cells(int nc, const double* __restrict__ cells, float* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  for (int t = 0; t < 10; ++t) out[12 * (size_t)c + t] = (float)cells[10 * (size_t)c + t];
  out[12 * (size_t)c + 10] = 0.f;
  out[12 * (size_t)c + 11] = 0.f;
}

// ------------------------------------------------------------ host drivers
struct Octree {
  Nodes nd{};
  int nn = 0;
  std::vector<int> level_start;  // level l occupies [level_start[l], level_start[l+1])
};

// Sorted keys, particles in Morton order and the octree.  m may be null.
void build_octree(Build& B, Lists& R, int n, const double* x, const double* m,
                  const double lo[3], double L, int split, Octree& T, u64** skey_out,
                  double** mm_out) {
  hipStream_t st = B.st;
  u64* key = B.tmp.get<u64>(n);
  u64* skey = B.tmp.get<u64>(n);
  int* idx = B.tmp.get<int>(n);
  int* perm = B.tmp.get<int>(n);
  const double s = (double)(1 << kLevels) / L;
  hipLaunchKernelGGL(k_keys, grid_for(n), dim3(256), 0, st, n, x, lo[0], lo[1], lo[2], s, key, idx);
  PGAMD_HIP_CHECK(hipGetLastError());
  B.sort_pairs(key, skey, idx, perm, n, 3 * kLevels);
  double* mm = m ? B.tmp.get<double>(n) : nullptr;
  hipLaunchKernelGGL(k_gather, grid_for(n), dim3(256), 0, st, n, perm, x, m, R.pos, mm, R.src);
  PGAMD_HIP_CHECK(hipGetLastError());
  // node storage: grown per level (a level holds at most 8x the previous)
  int cap = std::max(64, 2 * (n / std::max(1, split) + 1));
  auto alloc_nodes = [&](int c) {
    Nodes a;
    a.lo = B.tmp.get<int>(c);
    a.hi = B.tmp.get<int>(c);
    a.par = B.tmp.get<int>(c);
    a.lev = B.tmp.get<int>(c);
    a.child = B.tmp.get<int>(8 * (size_t)c);
    a.c = B.tmp.get<double>(3 * (size_t)c);
    a.com = B.tmp.get<double>(3 * (size_t)c);
    a.m = B.tmp.get<double>(c);
    a.q = B.tmp.get<double>(6 * (size_t)c);
    a.cell = B.tmp.get<int>(c);
    return a;
  };
  T.nd = alloc_nodes(cap);
  {
    const int z = 0, minus1 = -1;
    PGAMD_HIP_CHECK(hipMemcpyAsync(T.nd.lo, &z, sizeof(int), hipMemcpyHostToDevice, st));
    PGAMD_HIP_CHECK(hipMemcpyAsync(T.nd.hi, &n, sizeof(int), hipMemcpyHostToDevice, st));
    PGAMD_HIP_CHECK(hipMemcpyAsync(T.nd.par, &minus1, sizeof(int), hipMemcpyHostToDevice, st));
    PGAMD_HIP_CHECK(hipMemcpyAsync(T.nd.lev, &z, sizeof(int), hipMemcpyHostToDevice, st));
    std::vector<int> ch(8, -1);
    PGAMD_HIP_CHECK(
        hipMemcpyAsync(T.nd.child, ch.data(), 8 * sizeof(int), hipMemcpyHostToDevice, st));
    const double c0[3] = {lo[0] + 0.5 * L, lo[1] + 0.5 * L, lo[2] + 0.5 * L};
    PGAMD_HIP_CHECK(
        hipMemcpyAsync(T.nd.c, c0, 3 * sizeof(double), hipMemcpyHostToDevice, st));
    PGAMD_HIP_CHECK(hipStreamSynchronize(st));
  }
  T.nn = 1;
  T.level_start = {0, 1};
  for (int lev = 0; lev < kLevels; ++lev) {
    const int a = T.level_start[lev], b = T.level_start[lev + 1];
    if (b == a) break;
    int* nch = B.tmp.get<int>(b - a);
    int* off = B.tmp.get<int>(b - a + 1);
    int* end = B.tmp.get<int>(8 * (size_t)(b - a));
    hipLaunchKernelGGL(k_split_count, grid_for(8LL * (b - a)), dim3(256), 0, st, a, b, T.nd, skey,
                       split, end, nch);
    PGAMD_HIP_CHECK(hipGetLastError());
    const int add = B.scan(nch, off, b - a);
    if (add == 0) {
      T.level_start.push_back(b);
      break;
    }
    if (b + add > cap) {  // grow: copy the nodes so far into larger arrays
      int ncap = std::max(2 * cap, b + add);
      Nodes g = alloc_nodes(ncap);
      auto cp = [&](void* d, const void* s2, size_t bytes) {
        PGAMD_HIP_CHECK(hipMemcpyAsync(d, s2, bytes, hipMemcpyDeviceToDevice, st));
      };
      cp(g.lo, T.nd.lo, b * sizeof(int));
      cp(g.hi, T.nd.hi, b * sizeof(int));
      cp(g.par, T.nd.par, b * sizeof(int));
      cp(g.lev, T.nd.lev, b * sizeof(int));
      cp(g.child, T.nd.child, 8 * (size_t)b * sizeof(int));
      cp(g.c, T.nd.c, 3 * (size_t)b * sizeof(double));
      T.nd = g;
      cap = ncap;
    }
    hipLaunchKernelGGL(k_split_fill, grid_for(8LL * (b - a)), dim3(256), 0, st, a, b, T.nd, end,
                       off, L);
    PGAMD_HIP_CHECK(hipGetLastError());
    T.level_start.push_back(b + add);
    T.nn = b + add;
  }
  *skey_out = skey;
  *mm_out = mm;
}

// Groups of the octree sorted by first particle; fills ioff/icnt/ilist.
int build_groups(Build& B, Lists& R, const Octree& T, int ncrit, int** gnode_out) {
  hipStream_t st = B.st;
  const int nn = T.nn;
  int* flag = B.tmp.get<int>(nn);
  int* off = B.tmp.get<int>(nn + 1);
  hipLaunchKernelGGL(k_group_flag, grid_for(nn), dim3(256), 0, st, nn, T.nd, ncrit, flag);
  PGAMD_HIP_CHECK(hipGetLastError());
  const int ng = B.scan(flag, off, nn);
  int* gn = B.tmp.get<int>(ng);
  int* gn_sorted = B.tmp.get<int>(ng);
  int* lo = B.tmp.get<int>(ng);
  int* lo_sorted = B.tmp.get<int>(ng);
  hipLaunchKernelGGL(k_compact, grid_for(nn), dim3(256), 0, st, nn, flag, off, gn);
  hipLaunchKernelGGL(k_group_lo, grid_for(ng), dim3(256), 0, st, ng, gn, T.nd, lo);
  PGAMD_HIP_CHECK(hipGetLastError());
  B.sort_pairs(lo, lo_sorted, gn, gn_sorted, ng, 31);
  R.ng = ng;
  R.icnt = R.mem.get<int>(ng);
  R.ioff = R.mem.get<int>(ng + 1);
  hipLaunchKernelGGL(k_group_icnt, grid_for(ng), dim3(256), 0, st, ng, gn_sorted, T.nd, R.icnt);
  PGAMD_HIP_CHECK(hipGetLastError());
  R.ni = B.scan(R.icnt, R.ioff, ng);
  R.ilist = R.mem.get<int>(R.ni);
  if (ng > 0)
    hipLaunchKernelGGL(k_group_ilist, dim3(ng), dim3(256), 0, st, ng, gn_sorted, T.nd, R.ioff,
                       R.ilist);
  PGAMD_HIP_CHECK(hipGetLastError());
  *gnode_out = gn_sorted;
  return ng;
}

void finish(Build& B, Lists& R, bool cells) {
  hipStream_t st = B.st;
  const int ng = (int)R.ng;
  u64* key = B.tmp.get<u64>(ng);
  u64* skey = B.tmp.get<u64>(ng);
  int* idx = B.tmp.get<int>(ng);
  unsigned long long* work = B.tmp.get<unsigned long long>(2);
  PGAMD_HIP_CHECK(hipMemsetAsync(work, 0, 2 * sizeof(unsigned long long), st));
  R.order = R.mem.get<int>(ng);
  hipLaunchKernelGGL(k_cost, grid_for(ng), dim3(256), 0, st, ng, R.icnt, R.jcnt,
                     cells ? R.ccnt : nullptr, key, idx, work);
  PGAMD_HIP_CHECK(hipGetLastError());
  B.sort_pairs(key, skey, idx, R.order, ng, 64);
  unsigned long long w[2];
  PGAMD_HIP_CHECK(hipMemcpyAsync(w, work, sizeof(w), hipMemcpyDeviceToHost, st));
  PGAMD_HIP_CHECK(hipStreamSynchronize(st));
  R.work_pp = (long long)w[0];
  R.work_pc = cells ? (long long)w[1] : 0;
}

Lists* treepm_build(int n, const double* x, double rcut, int ncrit, hipStream_t st) {
  Lists* R = new Lists();
  Build B(st);
  // positions: originals then ghosts (at most 7 per particle)
  double* pos_all = R->mem.get<double>(3 * (size_t)n * 8);
  int* src_all = R->mem.get<int>((size_t)n * 8);
  R->pos = pos_all;
  R->src = src_all;
  Octree T;
  u64* skey = nullptr;
  double* mm = nullptr;
  const double lo[3] = {0, 0, 0};
  build_octree(B, *R, n, x, nullptr, lo, 1.0, std::max(1, ncrit), T, &skey, &mm);
  int* gnode = nullptr;
  const int ng = build_groups(B, *R, T, ncrit, &gnode);
  // ghosts
  int* gcnt = B.tmp.get<int>(n);
  int* goff = B.tmp.get<int>(n + 1);
  hipLaunchKernelGGL(k_ghost_count, grid_for(n), dim3(256), 0, st, n, R->pos, rcut, gcnt);
  PGAMD_HIP_CHECK(hipGetLastError());
  const int nghost = B.scan(gcnt, goff, n);
  hipLaunchKernelGGL(k_ghost_fill, grid_for(n), dim3(256), 0, st, n, goff, rcut, R->pos, R->src);
  PGAMD_HIP_CHECK(hipGetLastError());
  const int np = n + nghost;
  R->np = np;
  // grid over [-rcut, 1 + rcut) with cells of side >= rcut
  Grid gr;
  gr.glo = -rcut;
  const double gL = 1.0 + 2 * rcut;
  gr.nc = std::max(1, (int)std::floor(gL / rcut));
  gr.h = gL / gr.nc;
  const int ncell = gr.nc * gr.nc * gr.nc;
  int* cell = B.tmp.get<int>(np);
  int* scell = B.tmp.get<int>(np);
  int* idx = B.tmp.get<int>(np);
  int* cpart = B.tmp.get<int>(np);
  int* cstart = B.tmp.get<int>(ncell + 1);
  hipLaunchKernelGGL(k_cell_of, grid_for(np), dim3(256), 0, st, np, R->pos, gr, cell, idx);
  PGAMD_HIP_CHECK(hipGetLastError());
  B.sort_pairs(cell, scell, idx, cpart, np, bits_for(ncell));
  hipLaunchKernelGGL(k_cell_start, grid_for(ncell + 1), dim3(256), 0, st, ncell, np, scell,
                     cstart);
  PGAMD_HIP_CHECK(hipGetLastError());
  // j-lists: count, offsets, fill (cell order)
  double* box = B.tmp.get<double>(6 * (size_t)std::max(ng, 1));
  hipLaunchKernelGGL(k_group_box, grid_for(ng, 4), dim3(256), 0, st, ng, gnode, T.nd, R->pos, box);
  R->jcnt = R->mem.get<int>(ng);
  R->joff = R->mem.get<int>(ng + 1);
  double* cpos = B.tmp.get<double>(3 * (size_t)np);
  hipLaunchKernelGGL(k_cell_pos, grid_for(np), dim3(256), 0, st, np, cpart, R->pos, cpos);
  // One walk: each group writes its list at its candidate offset (the
  // candidate count bounds the list length), then the lists are packed.
  int* ccand = B.tmp.get<int>(ng);
  int* soff = B.tmp.get<int>(ng + 1);
  hipLaunchKernelGGL(k_treepm_cand, grid_for(ng, 4), dim3(256), 0, st, ng, box, gr, cstart, rcut,
                     ccand);
  PGAMD_HIP_CHECK(hipGetLastError());
  const int ncand = B.scan(ccand, soff, ng);
  int* scratch = B.tmp.get<int>(std::max(ncand, 1));
  hipLaunchKernelGGL(k_treepm_jlist, grid_for(ng, 4), dim3(256), 0, st, ng, box, cpos,
                     gr, cstart, cpart, rcut, R->jcnt, soff, scratch);
  PGAMD_HIP_CHECK(hipGetLastError());
  R->nj = B.scan(R->jcnt, R->joff, ng);
  R->jlist = R->mem.get<int>(R->nj);
  hipLaunchKernelGGL(k_pack_lists, grid_for(ng, 4), dim3(256), 0, st, ng, soff, R->jcnt, R->joff,
                     scratch, R->jlist);
  PGAMD_HIP_CHECK(hipGetLastError());
  R->coff = R->mem.get<int>(ng + 1);
  R->ccnt = R->mem.get<int>(ng);
  PGAMD_HIP_CHECK(hipMemsetAsync(R->coff, 0, sizeof(int) * (size_t)(ng + 1), st));
  PGAMD_HIP_CHECK(hipMemsetAsync(R->ccnt, 0, sizeof(int) * (size_t)std::max(ng, 1), st));
  R->clist = R->mem.get<int>(1);
  R->cells = R->mem.get<double>(10);
  finish(B, *R, false);
  return R;
}

Lists* bh_build(int n, const double* x, const double* m, double theta, int ncrit, int nleaf,
                hipStream_t st) {
  Lists* R = new Lists();
  Build B(st);
  // bounding box -> cube side, as the host builder
  const int nb = 256;
  double* part = B.tmp.get<double>(6 * nb);
  double* bb = B.tmp.get<double>(6);
  hipLaunchKernelGGL(k_bbox_part, dim3(nb), dim3(256), 0, st, n, x, part);
  hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(64), 0, st, nb, part, bb);
  PGAMD_HIP_CHECK(hipGetLastError());
  double h[6];
  PGAMD_HIP_CHECK(hipMemcpyAsync(h, bb, sizeof(h), hipMemcpyDeviceToHost, st));
  PGAMD_HIP_CHECK(hipStreamSynchronize(st));
  const double lo[3] = {h[0], h[1], h[2]};
  double L = 0;
  for (int d = 0; d < 3; ++d) L = std::max(L, h[3 + d] - h[d]);
  L *= 1.0 + 1e-9;
  R->pos = R->mem.get<double>(3 * (size_t)n);
  R->src = R->mem.get<int>(n);
  R->np = n;
  Octree T;
  u64* skey = nullptr;
  double* mm = nullptr;
  build_octree(B, *R, n, x, m, lo, L, std::max(1, std::min(ncrit, nleaf)), T, &skey, &mm);
  // moments, deepest level first
  for (int lev = (int)T.level_start.size() - 2; lev >= 0; --lev) {
    const int a = T.level_start[lev], b = T.level_start[lev + 1];
    if (b > a)
      hipLaunchKernelGGL(k_moments, grid_for(b - a, 64), dim3(64), 0, st, a, b, T.nd, R->pos,
                         mm);
    PGAMD_HIP_CHECK(hipGetLastError());
  }
  // cells: nodes with more than one particle, in node order
  int* flag = B.tmp.get<int>(T.nn);
  int* off = B.tmp.get<int>(T.nn + 1);
  hipLaunchKernelGGL(k_cell_flag, grid_for(T.nn), dim3(256), 0, st, T.nn, T.nd, flag);
  PGAMD_HIP_CHECK(hipGetLastError());
  R->ncells = B.scan(flag, off, T.nn);
  R->cells = R->mem.get<double>(10 * (size_t)R->ncells);
  hipLaunchKernelGGL(k_cell_fill, grid_for(T.nn), dim3(256), 0, st, T.nn, T.nd, flag, off,
                     R->cells);
  PGAMD_HIP_CHECK(hipGetLastError());
  int* gnode = nullptr;
  const int ng = build_groups(B, *R, T, ncrit, &gnode);
  const double th2 = theta * theta;
  R->jcnt = R->mem.get<int>(ng);
  R->joff = R->mem.get<int>(ng + 1);
  R->ccnt = R->mem.get<int>(ng);
  R->coff = R->mem.get<int>(ng + 1);
  WalkNode* wn = B.tmp.get<WalkNode>(T.nn);
  hipLaunchKernelGGL(k_walk_nodes, grid_for(T.nn), dim3(256), 0, st, T.nn, T.nd, wn);
  PGAMD_HIP_CHECK(hipGetLastError());
  const dim3 wgrid((ng + kWalkWaves - 1) / kWalkWaves), wblock(64 * kWalkWaves);
  hipLaunchKernelGGL((k_bh_walk<false>), wgrid, wblock, 0, st, ng, gnode, T.nd, wn, L, th2, nleaf,
                     R->jcnt, R->ccnt, nullptr, nullptr, nullptr, nullptr);
  PGAMD_HIP_CHECK(hipGetLastError());
  R->nj = B.scan(R->jcnt, R->joff, ng);
  R->nc = B.scan(R->ccnt, R->coff, ng);
  R->jlist = R->mem.get<int>(R->nj);
  R->clist = R->mem.get<int>(R->nc);
  hipLaunchKernelGGL((k_bh_walk<true>), wgrid, wblock, 0, st, ng, gnode, T.nd, wn, L, th2, nleaf,
                     nullptr, nullptr, R->joff, R->coff, R->jlist, R->clist);
  PGAMD_HIP_CHECK(hipGetLastError());
  finish(B, *R, true);
  return R;
}

}  // namespace

// ------------------------------------------------------------ C ABI
namespace {
hipStream_t g_stream = nullptr;
std::mutex g_mutex;

hipStream_t pick_stream(void* s) {
  if (s) return (hipStream_t)s;
  if (!g_stream) PGAMD_HIP_CHECK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  return g_stream;
}
}  // namespace

// TreePM lists of pgt_treepm_build (libpgtree.so), built on the GPU.
// x: device [n][3] FP64 positions in the periodic unit box [0, 1).  Returns
// a handle (null after an API error).  Synchronous on `stream` (null = the
// library's stream).
PGAMD_API void* pgt_dev_treepm_build(int n, const double* x, double rcut, int ncrit, void* stream) {
  std::lock_guard<std::mutex> lk(g_mutex);
  ensure_device();
  if (n <= 0 || !x || !(rcut > 0.0 && rcut < 0.5) || ncrit < 1) {
    api_error(3, "pgt_dev_treepm_build: n=%d rcut=%g ncrit=%d (need n > 0, 0 < rcut < 0.5, "
                 "ncrit >= 1, x non-null)", n, rcut, ncrit);
    return nullptr;
  }
  return treepm_build(n, x, rcut, ncrit, pick_stream(stream));
}

// Barnes-Hut lists of pgt_bh_build (libpgtree.so), built on the GPU.
// x: device [n][3], m: device [n], FP64.
PGAMD_API void* pgt_dev_bh_build(int n, const double* x, const double* m, double theta, int ncrit,
                                 int nleaf, void* stream) {
  std::lock_guard<std::mutex> lk(g_mutex);
  ensure_device();
  if (n <= 1 || !x || !m || !(theta > 0.0) || ncrit < 1 || nleaf < 1) {
    api_error(3, "pgt_dev_bh_build: n=%d theta=%g ncrit=%d nleaf=%d (need n > 1, theta > 0, "
                 "ncrit, nleaf >= 1, x and m non-null)", n, theta, ncrit, nleaf);
    return nullptr;
  }
  return bh_build(n, x, m, theta, ncrit, nleaf, pick_stream(stream));
}

// info[8] = {np, ngroups, len(ilist), len(jlist), len(clist), ncells,
//            sum ni*nj, sum ni*ncells}, as pgt_info.
PGAMD_API int pgt_dev_info(void* h, long long* info) {
  if (!h || !info) {
    api_error(3, "pgt_dev_info: null argument");
    return -1;
  }
  const Lists* r = static_cast<Lists*>(h);
  info[0] = r->np;
  info[1] = r->ng;
  info[2] = r->ni;
  info[3] = r->nj;
  info[4] = r->nc;
  info[5] = r->ncells;
  info[6] = r->work_pp;
  info[7] = r->work_pc;
  return 0;
}

namespace {
// Device pointer and byte size of array `what` (codes of pgt_get).
const void* array_of(const Lists* r, int what, size_t* bytes) {
  const size_t ng = (size_t)r->ng;
  switch (what) {
    case 0: *bytes = 3 * sizeof(double) * (size_t)r->np; return r->pos;
    case 1: *bytes = sizeof(int) * (size_t)r->np; return r->src;
    case 2: *bytes = sizeof(int) * ng; return r->ioff;
    case 3: *bytes = sizeof(int) * ng; return r->icnt;
    case 4: *bytes = sizeof(int) * (size_t)r->ni; return r->ilist;
    case 5: *bytes = sizeof(int) * ng; return r->joff;
    case 6: *bytes = sizeof(int) * ng; return r->jcnt;
    case 7: *bytes = sizeof(int) * (size_t)r->nj; return r->jlist;
    case 8: *bytes = sizeof(int) * ng; return r->coff;
    case 9: *bytes = sizeof(int) * ng; return r->ccnt;
    case 10: *bytes = sizeof(int) * (size_t)r->nc; return r->clist;
    case 11: *bytes = sizeof(int) * ng; return r->order;
    case 12: *bytes = 10 * sizeof(double) * (size_t)r->ncells; return r->cells;
    default: *bytes = 0; return nullptr;
  }
}
}  // namespace

// Device pointer of array `what` (0 pos, 1 src, 2 ioff, 3 icnt, 4 ilist,
// 5 joff, 6 jcnt, 7 jlist, 8 coff, 9 ccnt, 10 clist, 11 order, 12 cells),
// valid until pgt_dev_free.
PGAMD_API const void* pgt_dev_ptr(void* h, int what) {
  size_t b = 0;
  const void* p = h ? array_of(static_cast<Lists*>(h), what, &b) : nullptr;
  if (!p) api_error(3, "pgt_dev_ptr: bad handle or array code %d", what);
  return p;
}

// Copies array `what` to host memory `out` (sized as pgt_get's).  Synchronous.
PGAMD_API int pgt_dev_get(void* h, int what, void* out) {
  size_t b = 0;
  const void* p = h ? array_of(static_cast<Lists*>(h), what, &b) : nullptr;
  if (!p || !out) {
    api_error(3, "pgt_dev_get: bad handle, array code %d or output", what);
    return -1;
  }
  if (b) PGAMD_HIP_CHECK(hipMemcpy(out, p, b, hipMemcpyDeviceToHost));
  return 0;
}

// Packs the batch-call inputs on the device: pos4[np] = float4 {float(x) *
// float(scale), ..., float(mass[src])} and, for Barnes-Hut lists, cells12
// [ncells] = 12 floats {x, y, z, M, Qxx, Qxy, Qxz, Qyy, Qyz, Qzz, 0, 0}
// (either output may be null).  mass: device [n] FP64 of the original
// particles.  Asynchronous on `stream`.
PGAMD_API int pgt_dev_pack(void* h, const double* mass, double scale, void* pos4, void* cells12,
                           void* stream) {
  if (!h || (pos4 && !mass)) {
    api_error(3, "pgt_dev_pack: null handle or mass");
    return -1;
  }
  const Lists* r = static_cast<Lists*>(h);
  hipStream_t st = pick_stream(stream);
  if (pos4 && r->np > 0)
    hipLaunchKernelGGL(k_pack, grid_for(r->np), dim3(256), 0, st, (int)r->np, r->pos, r->src,
                       mass, (float)scale, (float4*)pos4);
  if (cells12 && r->ncells > 0)
    hipLaunchKernelGGL(k_pack_cells, grid_for(r->ncells), dim3(256), 0, st, (int)r->ncells,
                       r->cells, (float*)cells12);
  PGAMD_HIP_CHECK(hipGetLastError());
  return 0;
}

PGAMD_API void pgt_dev_free(void* h) {
  if (!h) return;
  std::lock_guard<std::mutex> lk(g_mutex);
  Lists* r = static_cast<Lists*>(h);
  PGAMD_HIP_CHECK(hipDeviceSynchronize());
  r->mem.free_all();
  delete r;
}
