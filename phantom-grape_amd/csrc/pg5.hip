// libpg5.so: GRAPE-5 API (PG5/newton/libpg5/gp5util.h, phantom_g5.c:35-383,
// pg5_fortran.c) and the quad_tree cell API (PG5/newton_quad/quad/
// quad_util.h, quad_g5c.c) on MI355X, plus device-pointer entry points
// (pg5_dev_*).
//
// Board model.  As in the reference each `devid` (0..kMaxDev-1) is an
// independent board with its own j-memory; the non-MC calls use devid 0
// (phantom_g5.c:13-28).  Here a board is a HIP stream plus HBM buffers; the
// j-memory is mirrored in pinned host memory, written by g5_set_xmj and
// uploaded (only the range written since the last upload) when a force is
// requested.  g5_calculate_force_on_x* sends all ni i-particles to the GPU in
// one launch instead of the reference's chunks of g5_get_number_of_pipelines().
// Different devids may be driven from different host threads at once.
//
// Softening: g5_set_eps_to_all sets a value shared by every board, as the
// reference's static Eps (phantom_g5.c:17,67-70).  Building with
// -DPG5_SYMMETRIC gives the reference's SYMMETRIC interface, where
// g5_set_xmj takes eps_j^2 and r^2 = |dx|^2 + eps_i^2 + eps_j^2.
//
// rsqrt bias: the reference scales outputs by (1-3b) and -(1-b), where b is
// the measured bias of its 12-bit rsqrtps (rsqrt.c:42-81).  The GPU uses the
// ~1 ulp v_rsq_f32, whose bias is negligible, so by default b = 0 and the
// output is acc = sum m dx / r^3, pot = + sum m / r including the self pair.
// pg5_set_rsqrt_bias_mode(1) applies the reference's rule with b measured the
// same way for v_rsq_f32 (pg5::rsqrt_bias_sums), once per process.
#include "pgamd_common.h"
#include "pgamd_api.h"
#include "pg5_kernels.h"
#include "host_upload.h"
#include "bh_batch.h"

#include <algorithm>
#include <cmath>

using namespace pgamd;
using pg5::Cell;
using pg5::Outputs5;
using pg5::Partials5;

#ifdef PG5_SYMMETRIC
constexpr bool kSym = true;
#else
constexpr bool kSym = false;
#endif

namespace {

constexpr int kMaxDev = 64;             // reference MAXDEV is 24 (phantom_g5.c:13-15)
constexpr int kDefaultPipes = 4;        // reference NUM_PIPE (phantom_g5.c:11)
constexpr int kJMemSize = 1 << 24;      // j-memory limit per board
constexpr int kMaxPipes = 1 << 20;

struct JMem {
  // Host mirror (pinned) and device copy of one j-memory of float4 records.
  HostBuf h, he;  // he: eps_j^2 (symmetric build)
  DevBuf d, de;
  int hcap = 0;   // records held by h
  int hmax = 0;   // one past the highest address ever written
  int dvalid = 0; // the device copy holds [0, dvalid) (and any range uploaded since)
  int lo = 0, hi = 0;  // dirty range [lo, hi)
  void write_range(int a0, int a1) {
    hmax = std::max(hmax, a1);
    if (lo >= hi) {
      lo = a0;
      hi = a1;
    } else {
      lo = std::min(lo, a0);
      hi = std::max(hi, a1);
    }
  }
};

struct Board {
  bool init = false;
  hipStream_t stream = nullptr;
  JMem jm;             // particle j-memory
  JMem cm[2];          // cell memories: [0] g5c_*MC, [1] g5c_*MC0
  int n = 0;           // g5_set_n
  int nc[2] = {0, 0};  // g5c_set_n
  // pipeline registers (g5_set_xi / g5_run / g5_get_force)
  HostBuf ipipe;       // float4 x, y, z, eps2
  int ni_pipe = 0;
  HostBuf opipe;       // double acc[3], pot per pipeline
  // staging of the calculate_force calls
  HostBuf h_in, h_out;
  DevBuf d_in, d_part, d_counter;
};

Board g_boards[kMaxDev];
std::mutex g_init_mutex;
double g_eps = 0.0;
int g_npipes = 0;
// rsqrt bias correction (phantom_g5.c:45-60,244-269): mode 0 = none, 1 = the
// reference's output scaling with the bias b of the device's rsqrt
int g_bias_mode = 0;
bool g_bias_ok = false;
double g_bias = 0.0;
std::mutex g_bias_mutex;

// launch tuning (pg5_dev_set_launch; 0 = the default of plan())
int g_force_bi = 0;
int g_blocks_per_cu = 0;
int g_min_jper = 0;

int npipes() {
  if (g_npipes == 0) {
    int v = kDefaultPipes;
    if (const char* s = std::getenv("PGAMD_G5_NPIPES")) {
      int e = std::atoi(s);
      if (e > 0 && e <= kMaxPipes) v = e;
    }
    g_npipes = v;
  }
  return g_npipes;
}

Board* board(int devid, const char* fn) {
  if (devid < 0 || devid >= kMaxDev) {
    api_error(4, "%s: devid %d outside [0, %d)", fn, devid, kMaxDev);
    return nullptr;
  }
  Board& b = g_boards[devid];
  if (!b.init) {
    std::lock_guard<std::mutex> lk(g_init_mutex);
    if (!b.init) {
      ensure_device();
      PGAMD_HIP_CHECK(hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking));
      b.ipipe.reserve(sizeof(float4) * (size_t)npipes());
      b.opipe.reserve(sizeof(double) * 4 * (size_t)npipes());
      b.init = true;
    }
  }
  return &b;
}

// Grows the host mirror to hold `need` records.  Records never written read
// as zero (x = y = z = m = 0: no force), as in the reference's zero-initialised
// static j-memory (PG5/newton/libpg5/phantom_g5.c:25-28).
void jmem_reserve_host(JMem& m, int need, size_t rec_bytes, bool eps) {
  if (need <= m.hcap) return;
  int cap = m.hcap > 0 ? m.hcap : 4096;
  while (cap < need) cap *= 2;
  m.h.reserve(rec_bytes * (size_t)cap, true);
  std::memset(static_cast<char*>(m.h.ptr) + rec_bytes * (size_t)m.hcap, 0,
              rec_bytes * (size_t)(cap - m.hcap));
  if (eps) {
    m.he.reserve(sizeof(float) * (size_t)cap, true);
    std::memset(m.he.as<float>() + m.hcap, 0, sizeof(float) * (size_t)(cap - m.hcap));
  }
  m.hcap = cap;
}

// Uploads, before a kernel reads [0, nread), the dirty range of a j-memory and
// every address below nread whose device copy was never written (a g5_set_n
// beyond the highest address set reads zero records, not stale memory).
void jmem_sync(JMem& m, size_t rec_bytes, bool eps, int nread, UploadBatch& ub,
               hipStream_t st) {
  jmem_reserve_host(m, nread, rec_bytes, eps);
  const size_t need = rec_bytes * (size_t)m.hcap;
  if (need > m.d.bytes) {
    // A grown device buffer starts empty: upload everything written so far.
    m.d.reserve(need);
    if (eps) m.de.reserve(sizeof(float) * (size_t)m.hcap);
    m.lo = 0;
    m.hi = m.hmax;
    m.dvalid = 0;
  }
  int lo = m.lo, hi = m.hi;
  if (nread > m.dvalid) {
    lo = lo < hi ? std::min(lo, m.dvalid) : m.dvalid;
    hi = std::max(hi, nread);
  }
  if (lo < hi) {
    const size_t off = rec_bytes * (size_t)lo;
    ub.add(static_cast<char*>(m.d.ptr) + off, m.h, off, rec_bytes * (size_t)(hi - lo), st);
    if (eps)
      ub.add(m.de.as<float>() + lo, m.he, sizeof(float) * (size_t)lo,
             sizeof(float) * (size_t)(hi - lo), st);
    if (lo <= m.dvalid) m.dvalid = std::max(m.dvalid, hi);
    m.lo = m.hi = 0;
  }
}

struct Plan {
  int bi, nsplit, jper, variant;
};

// Particle kernel: 0 = j-tiles in LDS (pg5_force), 1 = j broadcast through
// scalar registers (pg5_force_s).  Cells always use the LDS kernel.
int g_variant = 0;

// Default particle shape, measured on MI355X (LDS kernel, Plummer N x N,
// % of the FP32 peak against the earlier bi 256 / 128 / 64 by ni, 8 blocks
// per CU, 128 j per range): 64 i per block up to 32768 i, 256 beyond, 32
// blocks per CU, at least 256 j per range.  N = 4096: 40 % (30 %); 16384:
// 75 % (69 %); 32768: 81 % (78 %); 65536: 83 % (80 %); 131072: 84 % (81 %).
// Cells keep the earlier shape.
Plan plan(int ni, int nj, bool cells = false) {
  Plan p;
  if (g_force_bi > 0)
    p.bi = g_force_bi;
  else if (cells)
    p.bi = ni >= 16384 ? 256 : (ni >= 2048 ? 128 : 64);
  else
    p.bi = ni > 32768 ? 256 : 64;
  const int bpc = g_blocks_per_cu > 0 ? g_blocks_per_cu : (cells ? 8 : 32);
  const int min_jper = g_min_jper > 0 ? g_min_jper : (cells ? 128 : 256);
  p.variant = cells ? 0 : g_variant;
  const int nbi = ceil_div(ni, p.bi);
  // the scalar-broadcast kernel runs bi threads per block, the LDS one 256
  const int target = device_cu_count() * bpc * (p.variant == 1 ? 256 / p.bi : 1);
  int ns = std::max(1, ceil_div(target, nbi));
  ns = std::min(ns, std::max(1, nj / min_jper));
  int jper = ceil_div(std::max(nj, 1), ns);
  jper = (jper + 63) / 64 * 64;
  p.nsplit = std::max(1, ceil_div(std::max(nj, 1), jper));
  p.jper = jper;
  return p;
}

// Partial sums of nsplit j-ranges for ni i-particles in d (4 arrays of
// nsplit * ni doubles), or none when nsplit == 1.
Partials5 partials_at(double* d, int nsplit, int ni) {
  Partials5 part{};
  if (nsplit > 1) {
    const size_t n = (size_t)nsplit * ni;
    part.ax = d;
    part.ay = d + n;
    part.az = d + 2 * n;
    part.pot = d + 3 * n;
  }
  return part;
}

Partials5 partials(DevBuf& buf, int nsplit, int ni) {
  if (nsplit > 1) buf.reserve((size_t)nsplit * ni * 4 * sizeof(double));
  return partials_at(buf.as<double>(), nsplit, ni);
}

template <int BI>
void launch_particles_t(const Plan& p, int ni, const float4* xi, int nj, const float4* xj,
                        const float* ej2, const Partials5& part, const Outputs5& out,
                        hipStream_t st) {
  if (p.variant == 1)
    hipLaunchKernelGGL((pg5::pg5_force_s<kSym>), dim3(ceil_div(ni, BI), p.nsplit), dim3(BI), 0,
                       st, ni, xi, nj, p.jper, xj, ej2, part, out, p.nsplit);
  else
    hipLaunchKernelGGL((pg5::pg5_force<BI, kSym>), dim3(ceil_div(ni, BI), p.nsplit),
                       dim3(pg5::kThreads), 0, st, ni, xi, nj, p.jper, xj, ej2, part, out,
                       p.nsplit);
  PGAMD_HIP_CHECK(hipGetLastError());
}

template <int BI>
void launch_cells_t(const Plan& p, int ni, const float4* xi, int nc, const Cell* cells,
                    const Partials5& part, const Outputs5& out, hipStream_t st) {
  hipLaunchKernelGGL((pg5::pg5_quad<BI>), dim3(ceil_div(ni, BI), p.nsplit), dim3(pg5::kThreads),
                     0, st, ni, xi, nc, p.jper, cells, part, out, p.nsplit);
  PGAMD_HIP_CHECK(hipGetLastError());
}

// Launches the kernels of one force evaluation with plan p and partial sums
// in part (asynchronous on st).
void launch_force(const Plan& p, int ni, const float4* xi, int nj, const float4* xj,
                  const float* ej2, const Cell* cells, const Partials5& part,
                  const Outputs5& out, hipStream_t st);

// Force of nj particles (cells == nullptr) or nc cells on ni i-particles; all
// device pointers; asynchronous on st.  Partial sums live in `work`.
void run_force(DevBuf& work, int ni, const float4* xi, int nj, const float4* xj,
               const float* ej2, const Cell* cells, const Outputs5& out, hipStream_t st) {
  if (ni <= 0) return;
  if (nj <= 0) {
    PGAMD_HIP_CHECK(hipMemsetAsync(out.acc, 0, sizeof(double) * 3 * (size_t)ni, st));
    PGAMD_HIP_CHECK(hipMemsetAsync(out.pot, 0, sizeof(double) * (size_t)ni, st));
    return;
  }
  const Plan p = plan(ni, nj, cells != nullptr);
  launch_force(p, ni, xi, nj, xj, ej2, cells, partials(work, p.nsplit, ni), out, st);
}

void launch_force(const Plan& p, int ni, const float4* xi, int nj, const float4* xj,
                  const float* ej2, const Cell* cells, const Partials5& part,
                  const Outputs5& out, hipStream_t st) {
  if (cells) {
    switch (p.bi) {
      case 256: launch_cells_t<256>(p, ni, xi, nj, cells, part, out, st); break;
      case 128: launch_cells_t<128>(p, ni, xi, nj, cells, part, out, st); break;
      default: launch_cells_t<64>(p, ni, xi, nj, cells, part, out, st); break;
    }
  } else {
    switch (p.bi) {
      case 256: launch_particles_t<256>(p, ni, xi, nj, xj, ej2, part, out, st); break;
      case 128: launch_particles_t<128>(p, ni, xi, nj, xj, ej2, part, out, st); break;
      default: launch_particles_t<64>(p, ni, xi, nj, xj, ej2, part, out, st); break;
    }
  }
  if (p.nsplit > 1) {
    hipLaunchKernelGGL(pg5::pg5_reduce, dim3(ceil_div(ni, 64)), dim3(256), 0, st, ni,
                       p.nsplit, part, out);
    PGAMD_HIP_CHECK(hipGetLastError());
  }
}

// The bias b of the device's rsqrt by the rule of rsqrt.c:42-81 (measured on
// the first call, then kept): the block sums of pg5::rsqrt_bias_sums added in
// block order, b = (sum se) / (2 * 2 sum sw).
double rsqrt_bias(hipStream_t st) {
  std::lock_guard<std::mutex> lk(g_bias_mutex);
  if (!g_bias_ok) {
    DevBuf d;
    d.reserve(sizeof(double) * 2 * pg5::kBiasBlocks);
    hipLaunchKernelGGL(pg5::rsqrt_bias_sums, dim3(pg5::kBiasBlocks), dim3(256), 0, st,
                       d.as<double>());
    PGAMD_HIP_CHECK(hipGetLastError());
    std::vector<double> h(2 * pg5::kBiasBlocks);
    PGAMD_HIP_CHECK(hipMemcpyAsync(h.data(), d.ptr, sizeof(double) * h.size(),
                                   hipMemcpyDeviceToHost, st));
    PGAMD_HIP_CHECK(hipStreamSynchronize(st));
    d.release();
    double sw = 0.0, se = 0.0;
    for (int k = 0; k < pg5::kBiasBlocks; ++k) {
      sw += h[2 * k];
      se += h[2 * k + 1];
    }
    sw *= 2.0;
    g_bias = 0.5 * se / sw;
    g_bias_ok = true;
  }
  return g_bias;
}

// Host-side force evaluation for a board.  i-particles: x[ni][3] and either a
// per-i eps2 array (eps2 != nullptr, already squared) or the global Eps.
// cellmem: -1 = particles, 0/1 = cell memory.  Writes a[ni][3] and p[ni] with
// the sign convention of the API: +pot for particles, -pot for cells.
void board_calc(Board& b, const double (*x)[3], const double* eps2, const float4* pre, int ni,
                double (*a)[3], double* p, int cellmem) {
  if (ni <= 0) return;
  const int nread = cellmem < 0 ? b.n : b.nc[cellmem];
  if (nread <= 0) {  // nothing to sum over
    std::memset(a, 0, sizeof(double) * 3 * (size_t)ni);
    std::memset(p, 0, sizeof(double) * (size_t)ni);
    return;
  }
  const size_t in_bytes = sizeof(float4) * (size_t)ni;
  b.h_in.reserve(in_bytes);
  b.d_in.reserve(in_bytes);
  float4* hin = b.h_in.as<float4>();
  if (pre) {
    std::memcpy(hin, pre, in_bytes);
  } else {
    const float e2all = (float)(g_eps * g_eps);
    for (int i = 0; i < ni; ++i) {
      hin[i].x = (float)x[i][0];
      hin[i].y = (float)x[i][1];
      hin[i].z = (float)x[i][2];
      hin[i].w = eps2 ? (float)eps2[i] : e2all;
    }
  }
  UploadBatch ub;
  ub.add(b.d_in.ptr, b.h_in, 0, in_bytes, b.stream);
  // The kernels write the results straight into pinned host memory through
  // its device mapping; the synchronisation below makes them visible.
  const size_t out_bytes = sizeof(double) * 4 * (size_t)ni;
  b.h_out.reserve(out_bytes);
  double* dout = nullptr;
  PGAMD_HIP_CHECK(hipHostGetDevicePointer((void**)&dout, b.h_out.ptr, 0));
  Outputs5 o{dout, dout + 3 * (size_t)ni};
  if (cellmem < 0) {
    jmem_sync(b.jm, sizeof(float4), kSym, nread, ub, b.stream);
    ub.flush(b.stream);
    run_force(b.d_part, ni, b.d_in.as<float4>(), nread, b.jm.d.as<float4>(),
              kSym ? b.jm.de.as<float>() : nullptr, nullptr, o, b.stream);
  } else {
    jmem_sync(b.cm[cellmem], sizeof(Cell), false, nread, ub, b.stream);
    ub.flush(b.stream);
    run_force(b.d_part, ni, b.d_in.as<float4>(), nread, nullptr, nullptr,
              b.cm[cellmem].d.as<Cell>(), o, b.stream);
  }
  PGAMD_HIP_CHECK(hipStreamSynchronize(b.stream));
  const double* ha = b.h_out.as<double>();
  const double* hp = ha + 3 * (size_t)ni;
  const double psign = cellmem < 0 ? 1.0 : -1.0;
  std::memcpy(a, ha, sizeof(double) * 3 * (size_t)ni);
  for (int i = 0; i < ni; ++i) p[i] = psign * hp[i];
  if (g_bias_mode == 1) {
    // the reference's factors, FP32 as there (phantom_g5.c:49-50,250-257)
    const double bias = rsqrt_bias(b.stream);
    const double ca = (double)(float)(1.0 - 3.0 * bias), cp = (double)(float)(1.0 - bias);
    for (int i = 0; i < ni; ++i) {
      a[i][0] *= ca;
      a[i][1] *= ca;
      a[i][2] *= ca;
      p[i] *= cp;
    }
  }
}

void set_xmj_impl(int devid, int adr, int nj, const double (*xj)[3], const double* mj,
                  const double* epsj2) {
  Board* b = board(devid, "g5_set_xmj");
  if (!b) return;
  if (adr < 0 || nj < 0 || (long long)adr + nj > kJMemSize) {
    api_error(2, "g5_set_xmj: addresses [%d, %d) outside the j-memory of %d", adr, adr + nj,
              kJMemSize);
    return;
  }
  if (nj == 0) return;
  // The reference reads xj[j] for j in [adr, adr+nj) (phantom_g5.c:193-197).
  jmem_reserve_host(b->jm, adr + nj, sizeof(float4), kSym);
  float4* h = b->jm.h.as<float4>();
  for (int j = adr; j < adr + nj; ++j) {
    h[j].x = (float)xj[j][0];
    h[j].y = (float)xj[j][1];
    h[j].z = (float)xj[j][2];
    h[j].w = (float)mj[j];
  }
  if (kSym)
    for (int j = adr; j < adr + nj; ++j) b->jm.he.as<float>()[j] = (float)epsj2[j];
  b->jm.write_range(adr, adr + nj);
}

void set_xi_impl(int devid, int ni, const double (*xi)[3], const double* eps2) {
  Board* b = board(devid, "g5_set_xi");
  if (!b) return;
  if (ni < 0 || ni > npipes()) {
    api_error(3, "g5_set_xi: ni=%d exceeds the %d pipelines", ni, npipes());
    return;
  }
  float4* ip = b->ipipe.as<float4>();
  const float e2all = (float)(g_eps * g_eps);
  for (int i = 0; i < ni; ++i) {
    ip[i].x = (float)xi[i][0];
    ip[i].y = (float)xi[i][1];
    ip[i].z = (float)xi[i][2];
    ip[i].w = eps2 ? (float)eps2[i] : e2all;
  }
  b->ni_pipe = ni;
}

}  // namespace

// ------------------------------------------------------------------ GRAPE-5 API

PGAMD_API int g5_get_number_of_pipelines(void) { return npipes(); }
PGAMD_API int g5_get_jmemsize(void) { return kJMemSize; }

// Opening the board readies the device as the hardware's open did: the
// default board's stream and the library's kernels loaded (one empty
// launch), so the first force call pays none of it.
namespace {
__global__ void open_warm(int* flag) {
  if (flag) *flag = 0;
}
}  // namespace

PGAMD_API void g5_open(void) {
  ensure_device();
  Board* b = board(0, "g5_open");
  hipLaunchKernelGGL(open_warm, dim3(1), dim3(64), 0, b->stream, nullptr);
  PGAMD_HIP_CHECK(hipGetLastError());
  PGAMD_HIP_CHECK(hipStreamSynchronize(b->stream));
}
PGAMD_API void g5_close(void) {
  for (int d = 0; d < kMaxDev; ++d)
    if (g_boards[d].init) PGAMD_HIP_CHECK(hipStreamSynchronize(g_boards[d].stream));
}

PGAMD_API void g5_set_eps_to_all(double eps) { g_eps = eps; }

PGAMD_API void g5_set_epsMC(int devid, int ni, double* eps) {
  Board* b = board(devid, "g5_set_eps");
  if (!b) return;
  if (ni < 0 || ni > npipes()) {
    api_error(3, "g5_set_eps: ni=%d exceeds the %d pipelines", ni, npipes());
    return;
  }
  float4* ip = b->ipipe.as<float4>();
  for (int i = 0; i < ni; ++i) ip[i].w = (float)(eps[i] * eps[i]);
}
PGAMD_API void g5_set_eps(int ni, double* eps) { g5_set_epsMC(0, ni, eps); }

PGAMD_API void g5_set_nMC(int devid, int n) {
  Board* b = board(devid, "g5_set_n");
  if (!b) return;
  if (n < 0 || n > kJMemSize) {
    api_error(2, "g5_set_n: n=%d outside the j-memory of %d", n, kJMemSize);
    return;
  }
  b->n = n;
}
PGAMD_API void g5_set_n(int n) { g5_set_nMC(0, n); }

PGAMD_API void g5_set_xiMC(int devid, int ni, double (*xi)[3]) {
  set_xi_impl(devid, ni, xi, nullptr);
}
PGAMD_API void g5_set_xi(int ni, double (*xi)[3]) { g5_set_xiMC(0, ni, xi); }
PGAMD_API void g5_set_xeiMC(int devid, int ni, double (*xi)[3], double* eps2) {
  set_xi_impl(devid, ni, xi, eps2);
}
PGAMD_API void g5_set_xei(int ni, double (*xi)[3], double* eps2) {
  g5_set_xeiMC(0, ni, xi, eps2);
}

#ifdef PG5_SYMMETRIC
PGAMD_API void g5_set_xmjMC(int devid, int adr, int nj, double (*xj)[3], double* mj,
                            double* epsj2) {
  set_xmj_impl(devid, adr, nj, xj, mj, epsj2);
}
PGAMD_API void g5_set_xmj(int adr, int nj, double (*xj)[3], double* mj, double* epsj2) {
  set_xmj_impl(0, adr, nj, xj, mj, epsj2);
}
#else
PGAMD_API void g5_set_xmjMC(int devid, int adr, int nj, double (*xj)[3], double* mj) {
  set_xmj_impl(devid, adr, nj, xj, mj, nullptr);
}
PGAMD_API void g5_set_xmj(int adr, int nj, double (*xj)[3], double* mj) {
  set_xmj_impl(0, adr, nj, xj, mj, nullptr);
}
#endif

PGAMD_API void g5_runMC(int devid) {
  Board* b = board(devid, "g5_run");
  if (!b) return;
  const int ni = b->ni_pipe;
  if (ni <= 0) return;
  double(*a)[3] = reinterpret_cast<double(*)[3]>(b->opipe.as<double>());
  double* p = b->opipe.as<double>() + 3 * (size_t)ni;
  board_calc(*b, nullptr, nullptr, b->ipipe.as<float4>(), ni, a, p, -1);
}
PGAMD_API void g5_run(void) { g5_runMC(0); }

PGAMD_API void g5_get_forceMC(int devid, int ni, double (*a)[3], double* p) {
  Board* b = board(devid, "g5_get_force");
  if (!b) return;
  if (ni < 0 || ni > b->ni_pipe) {
    api_error(3, "g5_get_force: ni=%d but %d i-particles were run", ni, b->ni_pipe);
    return;
  }
  const int n = b->ni_pipe;
  const double* oa = b->opipe.as<double>();
  const double* op = oa + 3 * (size_t)n;
  std::memcpy(a, oa, sizeof(double) * 3 * (size_t)ni);
  std::memcpy(p, op, sizeof(double) * (size_t)ni);
}
PGAMD_API void g5_get_force(int ni, double (*a)[3], double* p) { g5_get_forceMC(0, ni, a, p); }

PGAMD_API void g5_calculate_force_on_xMC(int devid, double (*x)[3], double (*a)[3], double* p,
                                         int ni) {
  Board* b = board(devid, "g5_calculate_force_on_x");
  if (!b) return;
  board_calc(*b, x, nullptr, nullptr, ni, a, p, -1);
}
PGAMD_API void g5_calculate_force_on_x(double (*x)[3], double (*a)[3], double* p, int ni) {
  g5_calculate_force_on_xMC(0, x, a, p, ni);
}
PGAMD_API void g5_calculate_force_on_xeMC(int devid, double (*x)[3], double* eps2,
                                          double (*a)[3], double* p, int ni) {
  Board* b = board(devid, "g5_calculate_force_on_xe");
  if (!b) return;
  board_calc(*b, x, eps2, nullptr, ni, a, p, -1);
}
PGAMD_API void g5_calculate_force_on_xe(double (*x)[3], double* eps2, double (*a)[3], double* p,
                                        int ni) {
  g5_calculate_force_on_xeMC(0, x, eps2, a, p, ni);
}

// ------------------------------------------------- Fortran entry points
// (PG5/newton/libpg5/pg5_fortran.c:1-119): every argument by reference.

PGAMD_API int g5_get_number_of_pipelines_(void) { return g5_get_number_of_pipelines(); }
PGAMD_API int g5_get_jmemsize_(void) { return g5_get_jmemsize(); }
PGAMD_API void g5_open_(void) { g5_open(); }
PGAMD_API void g5_close_(void) { g5_close(); }
PGAMD_API void g5_set_eps_to_all_(double* eps) { g5_set_eps_to_all(*eps); }
PGAMD_API void g5_set_n_(int* n) { g5_set_n(*n); }
PGAMD_API void g5_set_nMC_(int* devid, int* n) { g5_set_nMC(*devid, *n); }
PGAMD_API void g5_set_xi_(int* ni, double (*xi)[3]) { g5_set_xi(*ni, xi); }
PGAMD_API void g5_set_xei_(int* ni, double (*xi)[3], double* eps2) { g5_set_xei(*ni, xi, eps2); }
PGAMD_API void g5_set_xiMC_(int* devid, int* ni, double (*xi)[3]) {
  g5_set_xiMC(*devid, *ni, xi);
}
PGAMD_API void g5_set_xeiMC_(int* devid, int* ni, double (*xi)[3], double* eps2) {
  g5_set_xeiMC(*devid, *ni, xi, eps2);
}
#ifdef PG5_SYMMETRIC
PGAMD_API void g5_set_xmj_(int* adr, int* nj, double (*xj)[3], double* mj, double* eps2) {
  g5_set_xmj(*adr, *nj, xj, mj, eps2);
}
PGAMD_API void g5_set_xmjMC_(int* devid, int* adr, int* nj, double (*xj)[3], double* mj,
                             double* eps2) {
  g5_set_xmjMC(*devid, *adr, *nj, xj, mj, eps2);
}
#else
PGAMD_API void g5_set_xmj_(int* adr, int* nj, double (*xj)[3], double* mj) {
  g5_set_xmj(*adr, *nj, xj, mj);
}
PGAMD_API void g5_set_xmjMC_(int* devid, int* adr, int* nj, double (*xj)[3], double* mj) {
  g5_set_xmjMC(*devid, *adr, *nj, xj, mj);
}
#endif
PGAMD_API void g5_run_(void) { g5_run(); }
PGAMD_API void g5_runMC_(int* devid) { g5_runMC(*devid); }
PGAMD_API void g5_get_force_(int* ni, double (*a)[3], double* p) { g5_get_force(*ni, a, p); }
PGAMD_API void g5_get_forceMC_(int* devid, int* ni, double (*a)[3], double* p) {
  g5_get_forceMC(*devid, *ni, a, p);
}
PGAMD_API void g5_calculate_force_on_x_(double (*x)[3], double (*a)[3], double* p, int* ni) {
  g5_calculate_force_on_x(x, a, p, *ni);
}
PGAMD_API void g5_calculate_force_on_xe_(double (*x)[3], double* eps2, double (*a)[3],
                                         double* p, int* ni) {
  g5_calculate_force_on_xe(x, eps2, a, p, *ni);
}
PGAMD_API void g5_calculate_force_on_xMC_(int* devid, double (*x)[3], double (*a)[3], double* p,
                                          int* ni) {
  g5_calculate_force_on_xMC(*devid, x, a, p, *ni);
}
PGAMD_API void g5_calculate_force_on_xeMC_(int* devid, double (*x)[3], double* eps2,
                                           double (*a)[3], double* p, int* ni) {
  g5_calculate_force_on_xeMC(*devid, x, eps2, a, p, *ni);
}

// ------------------------------------------------------------- quad_tree API
// Cells: COM xj[.][3], mass mj[.], traceless quadrupole qj[.][6] =
// {xx, xy, xz, yy, yz, zz} of sum m (3 d d - |d|^2 I) (source.cpp:215-248).
// Potential is returned negative: p = -sum (M/r + dx.Q.dx / (2 r^5))
// (quad_g5c.c:285-295 through g5_get_forceMC's -(1-b) scaling).

namespace {
void g5c_set_xmj_impl(int devid, int mem, int adr, int nj, const double (*xj)[3],
                      const double* mj, const double (*qj)[6]) {
  Board* b = board(devid, "g5c_set_xmj");
  if (!b) return;
  if (adr < 0 || nj < 0 || (long long)adr + nj > kJMemSize) {
    api_error(2, "g5c_set_xmj: addresses [%d, %d) outside the cell memory of %d", adr,
              adr + nj, kJMemSize);
    return;
  }
  if (nj == 0) return;
  JMem& m = b->cm[mem];
  jmem_reserve_host(m, adr + nj, sizeof(Cell), false);
  Cell* h = m.h.as<Cell>();
  for (int j = adr; j < adr + nj; ++j) {
    Cell c;
    c.xm = make_float4((float)xj[j][0], (float)xj[j][1], (float)xj[j][2], (float)mj[j]);
    c.qa = make_float4((float)qj[j][0], (float)qj[j][1], (float)qj[j][2], (float)qj[j][3]);
    c.qb = make_float4((float)qj[j][4], (float)qj[j][5], 0.f, 0.f);
    h[j] = c;
  }
  m.write_range(adr, adr + nj);
}
void g5c_set_n_impl(int devid, int mem, int n) {
  Board* b = board(devid, "g5c_set_n");
  if (!b) return;
  if (n < 0 || n > kJMemSize) {
    api_error(2, "g5c_set_n: n=%d outside the cell memory of %d", n, kJMemSize);
    return;
  }
  b->nc[mem] = n;
}
}  // namespace

PGAMD_API void g5c_set_nMC(int devid, int n) {
  g5c_set_n_impl(devid, 0, n);
  g5c_set_n_impl(devid, 1, n);
}
PGAMD_API void g5c_set_xmjMC(int devid, int adr, int nj, double (*xj)[3], double* mj,
                             double (*qj)[6]) {
  g5c_set_xmj_impl(devid, 0, adr, nj, xj, mj, qj);
}
PGAMD_API void g5c_set_xmjMC0(int devid, int adr, int nj, double (*xj)[3], double* mj,
                              double (*qj)[6]) {
  g5c_set_xmj_impl(devid, 1, adr, nj, xj, mj, qj);
}
PGAMD_API void g5c_runMC(int devid) {
  Board* b = board(devid, "g5c_run");
  if (!b) return;
  const int ni = b->ni_pipe;
  if (ni <= 0) return;
  double(*a)[3] = reinterpret_cast<double(*)[3]>(b->opipe.as<double>());
  double* p = b->opipe.as<double>() + 3 * (size_t)ni;
  board_calc(*b, nullptr, nullptr, b->ipipe.as<float4>(), ni, a, p, 0);
}
PGAMD_API void g5c_runMC0(int devid) {
  Board* b = board(devid, "g5c_run");
  if (!b) return;
  const int ni = b->ni_pipe;
  if (ni <= 0) return;
  double(*a)[3] = reinterpret_cast<double(*)[3]>(b->opipe.as<double>());
  double* p = b->opipe.as<double>() + 3 * (size_t)ni;
  board_calc(*b, nullptr, nullptr, b->ipipe.as<float4>(), ni, a, p, 1);
}
PGAMD_API void g5c_calculate_force_on_xMC(int devid, double (*x)[3], double (*a)[3], double* p,
                                          int ni) {
  Board* b = board(devid, "g5c_calculate_force_on_x");
  if (!b) return;
  board_calc(*b, x, nullptr, nullptr, ni, a, p, 0);
}
PGAMD_API void g5c_calculate_force_on_xMC0(int devid, double (*x)[3], double (*a)[3], double* p,
                                           int ni) {
  Board* b = board(devid, "g5c_calculate_force_on_x");
  if (!b) return;
  board_calc(*b, x, nullptr, nullptr, ni, a, p, 1);
}

// ------------------------------------------------- device-pointer entry points

// Force of nj j-particles on ni i-particles, all in device memory:
//   xi[ni] float4 {x, y, z, eps_i^2}, xj[nj] float4 {x, y, z, m},
//   epsj2[nj] (symmetric build only, else ignored), acc[ni][3], pot[ni] FP64
//   with pot = + sum m / r (self pair included, no sign change).
// Asynchronous on `stream` (nullptr = the board's stream); uses board devid's
// partial-sum scratch, so concurrent calls on one devid race.
PGAMD_API int pg5_dev_force(int devid, int ni, const void* xi, int nj, const void* xj,
                            const float* epsj2, double* acc, double* pot, void* stream) {
  Board* b = board(devid, "pg5_dev_force");
  if (!b) return -1;
  if (ni < 0 || nj < 0) {
    api_error(3, "pg5_dev_force: negative ni (%d) or nj (%d)", ni, nj);
    return -1;
  }
  Outputs5 o{acc, pot};
  run_force(b->d_part, ni, static_cast<const float4*>(xi), nj, static_cast<const float4*>(xj),
            epsj2, nullptr, o, stream ? (hipStream_t)stream : b->stream);
  return 0;
}

// ------------------------------------------------ allocation-free variant
// pg5_dev_force_noalloc keeps its j-range partial sums in this file-scope
// device array instead of memory the library allocates: every GPU the library
// runs on holds one copy (kScratchBytes) until the process ends.  Host code
// never passes `scratch` itself to a kernel (in host code the name is only a
// stand-in); it uses the device address looked up per device with
// hipGetSymbolAddress on the first call on that device.
constexpr size_t kScratchBytes = (size_t)16 << 20;
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

namespace {

constexpr int kMaxGpus = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxGpus] = {};

// Device address of this device's scratch array (current device), or null
// after reporting an error.
double* scratch_ptr(const char* fn) {
  int dev = 0;
  PGAMD_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxGpus) {
    api_error(4, "%s: device %d outside [0, %d)", fn, dev, kMaxGpus);
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      api_error(7, "%s: no device address for the scratch array on device %d (%s)", fn, dev,
                hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  return g_scratch_ptr[dev];
}

__global__ void zero_outputs(int ni, double* acc, double* pot) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ni) return;
  acc[3 * i + 0] = 0.0;
  acc[3 * i + 1] = 0.0;
  acc[3 * i + 2] = 0.0;
  pot[i] = 0.0;
}

}  // namespace

// Work of pg5_dev_force with no memory allocated by the call: the partial
// sums live in a static device array of 16 MiB per GPU, shared by every call
// on that GPU (calls that overlap in time on one GPU race on it).  The j-range
// split is the one pg5_dev_force plans, reduced until the partials fit
// (ni * nsplit * 32 bytes <= 16 MiB; one range needs none), so the results
// equal pg5_dev_force's whenever its plan fits.  No board, stream or host
// buffer is created: `stream` (null = the default stream) carries the work.
// For callers that cannot allocate (HIP graph capture, serving loops).
PGAMD_API int pg5_dev_force_noalloc(int ni, const void* xi, int nj, const void* xj,
                                    const float* epsj2, double* acc, double* pot,
                                    void* stream) {
  if (ni < 0 || nj < 0 || (ni > 0 && (!xi || !acc || !pot)) || (nj > 0 && !xj) ||
      (kSym && nj > 0 && !epsj2)) {
    api_error(3, "pg5_dev_force_noalloc: ni=%d nj=%d (need ni, nj >= 0 and non-null arrays)",
              ni, nj);
    return -1;
  }
  if (ni == 0) return 0;
  ensure_device();
  double* sp = scratch_ptr("pg5_dev_force_noalloc");
  if (!sp) return -1;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const Outputs5 o{acc, pot};
  if (nj == 0) {
    hipLaunchKernelGGL(zero_outputs, dim3(ceil_div(ni, 256)), dim3(256), 0, st, ni, acc, pot);
    PGAMD_HIP_CHECK(hipGetLastError());
    return 0;
  }
  Plan p = plan(ni, nj);
  const size_t cap = kScratchBytes / (4 * sizeof(double) * (size_t)ni);  // splits that fit
  if ((size_t)p.nsplit > cap) {
    const int ns = (int)std::max<size_t>(cap, 1);
    p.jper = (ceil_div(nj, ns) + 63) / 64 * 64;
    p.nsplit = ceil_div(nj, p.jper);
  }
  launch_force(p, ni, static_cast<const float4*>(xi), nj, static_cast<const float4*>(xj), epsj2,
               nullptr, partials_at(sp, p.nsplit, ni), o, st);
  return 0;
}

// Cell force of nc cells (device array of 12 floats each: x, y, z, M, Qxx,
// Qxy, Qxz, Qyy, Qyz, Qzz, 0, 0) on ni i-particles; pot = + sum (phi_p+phi_q).
PGAMD_API int pg5_dev_quad(int devid, int ni, const void* xi, int nc, const void* cells,
                           double* acc, double* pot, void* stream) {
  Board* b = board(devid, "pg5_dev_quad");
  if (!b) return -1;
  if (ni < 0 || nc < 0) {
    api_error(3, "pg5_dev_quad: negative ni (%d) or nc (%d)", ni, nc);
    return -1;
  }
  Outputs5 o{acc, pot};
  run_force(b->d_part, ni, static_cast<const float4*>(xi), nc, nullptr, nullptr,
            static_cast<const Cell*>(cells), o, stream ? (hipStream_t)stream : b->stream);
  return 0;
}

// rsqrt bias correction of the g5 outputs (the reference's A = acc (1 - 3b),
// pot (1 - b), PG5/newton/libpg5/phantom_g5.c:45-60,244-269, also for the
// cell calls, quad_g5c.c:285-295): mode 0 (default) none, 1 with b measured
// for the device's rsqrt by the rule of rsqrt.c:42-81.
PGAMD_API int pg5_set_rsqrt_bias_mode(int mode) {
  if (mode != 0 && mode != 1) {
    api_error(3, "pg5_set_rsqrt_bias_mode: mode %d not in {0, 1}", mode);
    return -1;
  }
  g_bias_mode = mode;
  return 0;
}

// b of the device's rsqrt (measured on the first call), whatever the mode.
PGAMD_API double pg5_rsqrt_bias(void) {
  Board* b = board(0, "pg5_rsqrt_bias");
  if (!b) return 0.0;
  return rsqrt_bias(b->stream);
}

PGAMD_API void* pg5_dev_stream(int devid) {
  Board* b = board(devid, "pg5_dev_stream");
  return b ? (void*)b->stream : nullptr;
}

PGAMD_API int pg5_dev_set_launch(int bi, int blocks_per_cu, int min_jper) {
  if (bi != 0 && bi != 64 && bi != 128 && bi != 256) {
    api_error(3, "pg5_dev_set_launch: bi=%d not in {0,64,128,256}", bi);
    return -1;
  }
  g_force_bi = bi;
  if (blocks_per_cu > 0) g_blocks_per_cu = blocks_per_cu;
  if (min_jper > 0) g_min_jper = min_jper;
  return 0;
}

// --------------------------------------------------- batched tree-list force

namespace {
int g_bh_bpc = 0;    // persistent blocks per CU (0 = default 2)
int g_bh_block = 0;  // threads per block (0 = default 1024)
constexpr int kBhBlocks[3] = {256, 512, 1024};

struct BhArgs {
  int ngroups;
  const int *ioff, *icnt, *ilist, *poff, *pcnt, *plist, *coff, *ccnt, *clist, *order;
  const float4* pos;
  const Cell* cells;
  float eps2;
  double *acc, *pot;
};

template <int BLOCK>
void launch_bh_t(const BhArgs& a, int* counter, int blocks, hipStream_t st) {
  hipLaunchKernelGGL(pgbh::bh_batch<BLOCK>, dim3(blocks), dim3(BLOCK), 0, st, a.ngroups, a.order,
                     a.ioff, a.icnt, a.ilist, a.poff, a.pcnt, a.plist, a.coff, a.ccnt, a.clist,
                     a.pos, a.cells, a.eps2, a.acc, a.pot, counter);
}

// Counter reset + persistent kernel on st; block in kBhBlocks.
void launch_bh(const BhArgs& a, int* counter, int bpc, int block, hipStream_t st) {
  hipLaunchKernelGGL(pgbh::zero_counter, dim3(1), dim3(64), 0, st, counter);
  PGAMD_HIP_CHECK(hipGetLastError());
  if (a.ngroups <= 0) return;
  const int blocks = std::max(1, std::min(a.ngroups, device_cu_count() * bpc));
  switch (block) {
    case 256: launch_bh_t<256>(a, counter, blocks, st); break;
    case 512: launch_bh_t<512>(a, counter, blocks, st); break;
    case 1024: launch_bh_t<1024>(a, counter, blocks, st); break;
    default:
      api_error(3, "tree-list batch: block %d not in {256, 512, 1024}", block);
      return;
  }
  PGAMD_HIP_CHECK(hipGetLastError());
}

int bh_bpc() { return g_bh_bpc > 0 ? g_bh_bpc : 2; }
int bh_block() { return g_bh_block > 0 ? g_bh_block : pgbh::kThreads; }

struct BhGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  DevBuf counter;
};

BhArgs bh_args(int ngroups, const int* ioff, const int* icnt, const int* ilist, const int* poff,
               const int* pcnt, const int* plist, const int* coff, const int* ccnt,
               const int* clist, const int* order, const void* pos, const void* cells,
               double eps2, double* acc, double* pot) {
  return BhArgs{ngroups, ioff, icnt, ilist, poff, pcnt, plist, coff, ccnt, clist, order,
                static_cast<const float4*>(pos), static_cast<const Cell*>(cells), (float)eps2,
                acc, pot};
}
}  // namespace

// Tree-list force of every i-group in one persistent launch (bh_batch.h).
// Device arrays: pos[np] float4 {x,y,z,m}; cells[nc] 12 floats {x,y,z,M,
// Qxx,Qxy,Qxz,Qyy,Qyz,Qzz,0,0}; per group g the i-list ilist[ioff[g]..
// +icnt[g]), particle list plist[poff[g]..+pcnt[g]) and cell list
// clist[coff[g]..+ccnt[g]) (indices into pos / cells); order nullable.
// Writes acc[i][3] and pot[i] = + sum m/r + sum (phi_p + phi_q) for every
// listed i.  Asynchronous on stream; calls on one devid share its counter.
PGAMD_API int pg5_bh_batch(int devid, int ngroups, const int* ioff, const int* icnt,
                           const int* ilist, const int* poff, const int* pcnt, const int* plist,
                           const int* coff, const int* ccnt, const int* clist, const int* order,
                           const void* pos, const void* cells, double eps2, double* acc,
                           double* pot, void* stream) {
  Board* b = board(devid, "pg5_bh_batch");
  if (!b) return -1;
  if (ngroups < 0) {
    api_error(3, "pg5_bh_batch: ngroups=%d", ngroups);
    return -1;
  }
  b->d_counter.reserve(256);
  launch_bh(bh_args(ngroups, ioff, icnt, ilist, poff, pcnt, plist, coff, ccnt, clist, order, pos,
                    cells, eps2, acc, pot),
            b->d_counter.as<int>(), bh_bpc(), bh_block(), stream ? (hipStream_t)stream : b->stream);
  return 0;
}

// The same launch captured into a HIP graph (own counter); replay with
// pg5_bh_graph_launch, free with pg5_bh_graph_destroy.
PGAMD_API void* pg5_bh_graph(int devid, int ngroups, const int* ioff, const int* icnt,
                             const int* ilist, const int* poff, const int* pcnt,
                             const int* plist, const int* coff, const int* ccnt,
                             const int* clist, const int* order, const void* pos,
                             const void* cells, double eps2, double* acc, double* pot) {
  Board* b = board(devid, "pg5_bh_graph");
  if (!b) return nullptr;
  BhGraph* g = new BhGraph();
  g->counter.reserve(256);
  hipStream_t cs;
  PGAMD_HIP_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  PGAMD_HIP_CHECK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
  launch_bh(bh_args(ngroups, ioff, icnt, ilist, poff, pcnt, plist, coff, ccnt, clist, order, pos,
                    cells, eps2, acc, pot),
            g->counter.as<int>(), bh_bpc(), bh_block(), cs);
  PGAMD_HIP_CHECK(hipStreamEndCapture(cs, &g->graph));
  PGAMD_HIP_CHECK(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
  PGAMD_HIP_CHECK(hipStreamDestroy(cs));
  return g;
}

PGAMD_API int pg5_bh_graph_launch(void* h, void* stream) {
  if (!h) {
    api_error(3, "pg5_bh_graph_launch: null graph");
    return -1;
  }
  Board* b = board(0, "pg5_bh_graph_launch");
  PGAMD_HIP_CHECK(hipGraphLaunch(static_cast<BhGraph*>(h)->exec,
                                 stream ? (hipStream_t)stream : b->stream));
  return 0;
}

PGAMD_API int pg5_bh_graph_destroy(void* h) {
  if (!h) return 0;
  BhGraph* g = static_cast<BhGraph*>(h);
  PGAMD_HIP_CHECK(hipDeviceSynchronize());
  PGAMD_HIP_CHECK(hipGraphExecDestroy(g->exec));
  PGAMD_HIP_CHECK(hipGraphDestroy(g->graph));
  g->counter.release();
  delete g;
  return 0;
}

// Autotuner: times the launch on these inputs for blocks of 256, 512 and
// 1024 threads (an LDS tile holds that many particles, or half as many
// cells) and {1, 2, 3, 4, 6, 8, 12, 16} persistent blocks per CU (one
// warm-up, mean of reps), keeps the fastest pair and returns its blocks per
// CU through *best; ms[8] (nullable) receives the timings at the block size
// kept, which pg5_bh_launch_shape reports.  Results depend on the block size
// only through the order of the FP32 partial sums.
PGAMD_API int pg5_bh_autotune(int devid, int ngroups, const int* ioff, const int* icnt,
                              const int* ilist, const int* poff, const int* pcnt,
                              const int* plist, const int* coff, const int* ccnt,
                              const int* clist, const int* order, const void* pos,
                              const void* cells, double eps2, double* acc, double* pot, int reps,
                              int* best, float* ms) {
  Board* b = board(devid, "pg5_bh_autotune");
  if (!b) return -1;
  if (ngroups < 0) {
    api_error(3, "pg5_bh_autotune: ngroups=%d", ngroups);
    return -1;
  }
  const BhArgs a = bh_args(ngroups, ioff, icnt, ilist, poff, pcnt, plist, coff, ccnt, clist,
                           order, pos, cells, eps2, acc, pot);
  b->d_counter.reserve(256);
  const int cand[8] = {1, 2, 3, 4, 6, 8, 12, 16};
  hipEvent_t e0, e1;
  PGAMD_HIP_CHECK(hipEventCreate(&e0));
  PGAMD_HIP_CHECK(hipEventCreate(&e1));
  float bestt = 1e30f, t_all[3][8];
  int bestc = bh_bpc(), bestkb = 2;
  while (bestkb > 0 && kBhBlocks[bestkb] != bh_block()) --bestkb;
  reps = std::max(1, reps);
  for (int kb = 0; kb < 3; ++kb)
    for (int k = 0; k < 8; ++k) {
      launch_bh(a, b->d_counter.as<int>(), cand[k], kBhBlocks[kb], b->stream);
      PGAMD_HIP_CHECK(hipEventRecord(e0, b->stream));
      for (int r = 0; r < reps; ++r)
        launch_bh(a, b->d_counter.as<int>(), cand[k], kBhBlocks[kb], b->stream);
      PGAMD_HIP_CHECK(hipEventRecord(e1, b->stream));
      PGAMD_HIP_CHECK(hipEventSynchronize(e1));
      float t = 0.f;
      PGAMD_HIP_CHECK(hipEventElapsedTime(&t, e0, e1));
      t /= reps;
      t_all[kb][k] = t;
      if (t < bestt) {
        bestt = t;
        bestc = cand[k];
        bestkb = kb;
      }
    }
  PGAMD_HIP_CHECK(hipEventDestroy(e0));
  PGAMD_HIP_CHECK(hipEventDestroy(e1));
  g_bh_bpc = bestc;
  g_bh_block = kBhBlocks[bestkb];
  if (ms)
    for (int k = 0; k < 8; ++k) ms[k] = t_all[bestkb][k];
  if (best) *best = bestc;
  return 0;
}

// Launch shape of the tree-list batch: block = threads per block (256, 512
// or 1024; 0 = default 1024), blocks_per_cu (0 = default 2); a negative value
// leaves that part as it is.
PGAMD_API int pg5_bh_set_launch(int block, int blocks_per_cu) {
  if (block > 0 && block != 256 && block != 512 && block != 1024) {
    api_error(3, "pg5_bh_set_launch: block %d not in {0, 256, 512, 1024}", block);
    return -1;
  }
  if (blocks_per_cu > 64) {
    api_error(3, "pg5_bh_set_launch: blocks_per_cu %d > 64", blocks_per_cu);
    return -1;
  }
  if (block >= 0) g_bh_block = block;
  if (blocks_per_cu >= 0) g_bh_bpc = blocks_per_cu;
  return 0;
}

// Launch shape the tree-list batch uses now: out2 = {threads per block,
// persistent blocks per CU}.
PGAMD_API int pg5_bh_launch_shape(int* out2) {
  if (!out2) {
    api_error(3, "pg5_bh_launch_shape: null output");
    return -1;
  }
  out2[0] = bh_block();
  out2[1] = bh_bpc();
  return 0;
}

// Particle force kernel: 0 = j-tiles staged in LDS (default, measured the
// faster on MI355X), 1 = j broadcast through scalar registers.  Same pair
// arithmetic; speed only.
PGAMD_API int pg5_dev_set_kernel(int variant) {
  if (variant != 0 && variant != 1) {
    api_error(3, "pg5_dev_set_kernel: variant %d not in {0, 1}", variant);
    return -1;
  }
  g_variant = variant;
  return 0;
}
