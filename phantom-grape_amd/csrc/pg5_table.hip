// libpg5_table.so: GRAPE-5 API with an arbitrary cut-off force shape by table
// interpolation (PG5/table/libpg5/gp5util.h, phantom_g5.c:43-200,
// pg5_table.h, pg5_table.c:31-113), on MI355X.  Default shape: the S2
// (Hockney-Eastwood) short-range TreePM force, eps_PP = 0.1/64,
// eps_PM = r_cut = 3/64 (pppm.h:4-6).
//
// Board model as in libpg5.so: per-devid boards (stream + HBM j-memory,
// pinned host mirror uploaded by dirty range).  Positions are stored scaled by
// xscale (phantom_g5.c:96-128); the kernel indexes the table by the float bits
// of s = 2 + |dx_s|^2.  The potential is not computed (returned as 0,
// phantom_g5.c:145-154).  Softening and eta are accepted and ignored, as in
// the reference (the softening is inside the tabulated shape).
#include "pgamd_common.h"
#include "host_upload.h"
#include "pgamd_api.h"
#include "table_kernels.h"
#include "treepm_batch.h"
#include "pm_kernels.h"

#include <hipcub/device/device_radix_sort.hpp>
#include <hipfft/hipfft.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>
#include <vector>

using namespace pgamd;

// Table of the default resolution, exported like the reference's global
// (pg5_table.h:31): Force_table[k][0] = value, [k][1] = slope in s.
extern "C" {
__attribute__((visibility("default"))) float Force_table[512][2];
}

namespace {

constexpr int kExpBit = 4;              // pg5_table.h:8
constexpr int kDefaultFrcBit = 5;       // pg5_table.h:9
constexpr int kMaxDev = 64;
constexpr int kDefaultPipes = 4;        // phantom_g5.c:9
constexpr int kJMemSize = 1 << 24;
constexpr double kMesh = 64.0;          // pppm.h:4
constexpr double kSftPM = 3.0 / kMesh;  // pppm.h:5
constexpr double kSftPP = 0.1 / kMesh;  // pppm.h:6

struct Table {
  int frc_bit = kDefaultFrcBit;
  int ntab = 1 << (kExpBit + kDefaultFrcBit);
  std::vector<float> host;  // 2*ntab
  DevBuf dev;
  bool dirty = true;
  double xscale = 0.0;
  float r2scale = 0.f;
  double (*func)(double) = nullptr;  // last shape given to pg5_gen_force_table
  double rcut = 0.0;
  double sft_pp = kSftPP, sft_pm = kSftPM;
  bool is_s2 = true;
};

struct Board {
  bool init = false;
  hipStream_t stream = nullptr;
  HostBuf hj;
  int hcap = 0;
  int hmax = 0;  // one past the highest address ever written
  int dvalid = 0;  // the device copy holds [0, dvalid) (and any range uploaded since)
  int lo = 0, hi = 0;
  DevBuf dj;
  int n = 0;
  HostBuf ipipe, opipe;
  int ni_pipe = 0;
  HostBuf h_in, h_out;
  DevBuf d_in, d_out, d_part, d_counter;
};

// Learned cut-off (mlp_kernels.h): parameters, the radii it was trained for.
struct Mlp {
  std::vector<float> params;  // mlp::kParams
  bool valid = false;
  double rc = 0.0, r0 = 0.0;
  DevBuf dev;
  bool dirty = true;
};

Table g_tab;
Mlp g_mlp;
int g_cut_mode = 0;  // 0: table, 1: learned MLP
int g_batch_bpc[2] = {0, 0};  // persistent blocks per CU per mode (0 = default)
Board g_boards[kMaxDev];
std::mutex g_mutex;  // guards the table and board creation
int g_force_bi = 0;
int g_blocks_per_cu = 8;
int g_min_jper = 128;

int npipes() {
  static int v = [] {
    int d = kDefaultPipes;
    if (const char* s = std::getenv("PGAMD_G5_NPIPES")) {
      int e = std::atoi(s);
      if (e > 0) d = e;
    }
    return d;
  }();
  return v;
}

// S2 force shape in r^-3 units, the reference's piecewise polynomial in
// eta = 2 r / eps (pg5_table.c:8-22) with the same Horner order, so that the
// table entries are the same floats.
double s2_reff(double sft, double r) {
  const double eta = 2.0 * r / sft;
  if (eta < 1.0)
    return (224. + eta * eta * (-224. + eta * (70. + eta * (48. - eta * 21.)))) * 2.e0 /
           (35.0 * std::pow(sft, 3));
  if (eta < 2.0)
    return (12. / (eta * eta) - 224. +
            eta * (896. + eta * (-840. + eta * (224. + eta * (70. + eta * (-48. + eta * 7.)))))) *
           2.e0 / (35.0 * eta * std::pow(sft, 3));
  return 1.0 / std::pow(r, 3);
}

double s2_force(double r) { return s2_reff(g_tab.sft_pp, r) - s2_reff(g_tab.sft_pm, r); }

union Pack32 {
  float f;
  uint32_t u;
};

// pg5_gen_force_table (pg5_table.c:43-113): entries at float-bit ticks of s
// from 2, r2 = (s - 2) / r2scale, value = func(sqrtf(r2)), forward-difference
// slope, last slope 0.  Expressions keep the reference's C types.
void build_table(double (*func)(double), double rcut, int frc_bit) {
  Table& t = g_tab;
  t.frc_bit = frc_bit;
  t.ntab = 1 << (kExpBit + frc_bit);
  t.host.assign(2 * (size_t)t.ntab, 0.f);
  const uint32_t tick = 1u << (23 - frc_bit);
  const float r2max = (float)(rcut * rcut);
  const float fmax = (float)((double)(1 << (1 << kExpBit)) * (2.0 - 1.0 / (1 << frc_bit)));
  const float r2scale = (fmax - 2.0f) / r2max;
  t.r2scale = r2scale;
  t.xscale = (double)sqrtf(r2scale);
  Pack32 m;
  m.f = 2.0f;
  for (int i = 0; i < t.ntab; i++, m.u += tick) {
    const float f = m.f;
    const float r2 = (float)(((double)f - 2.0) / (double)r2scale);
    const float r = sqrtf(r2);
    t.host[2 * i] = (float)func((double)r);
  }
  m.f = 2.0f;
  for (int i = 0; i < t.ntab - 1; i++) {
    const float x0 = m.f;
    m.u += tick;
    const float x1 = m.f;
    t.host[2 * i + 1] = (t.host[2 * (i + 1)] - t.host[2 * i]) / (x1 - x0);
  }
  t.host[2 * (t.ntab - 1) + 1] = 0.0f;
  t.func = func;
  t.rcut = rcut;
  t.dirty = true;
  if (t.ntab <= 512) std::memcpy(Force_table, t.host.data(), sizeof(float) * 2 * t.ntab);
}

void ensure_table() {
  if (g_tab.host.empty()) {
    g_tab.is_s2 = true;
    build_table(s2_force, g_tab.sft_pm, g_tab.frc_bit);
  }
}

// Upload the table if it changed (callers hold g_mutex).
const float2* table_dev(hipStream_t st) {
  Table& t = g_tab;
  if (t.dirty) {
    t.dev.reserve(sizeof(float) * t.host.size());
    PGAMD_HIP_CHECK(hipMemcpyAsync(t.dev.ptr, t.host.data(), sizeof(float) * t.host.size(),
                                   hipMemcpyHostToDevice, st));
    PGAMD_HIP_CHECK(hipStreamSynchronize(st));
    t.dirty = false;
  }
  return t.dev.as<float2>();
}

Board* board(int devid, const char* fn) {
  if (devid < 0 || devid >= kMaxDev) {
    api_error(4, "%s: devid %d outside [0, %d)", fn, devid, kMaxDev);
    return nullptr;
  }
  Board& b = g_boards[devid];
  if (!b.init) {
    std::lock_guard<std::mutex> lk(g_mutex);
    if (!b.init) {
      ensure_device();
      PGAMD_HIP_CHECK(hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking));
      b.ipipe.reserve(sizeof(float4) * (size_t)npipes());
      b.opipe.reserve(sizeof(double) * 3 * (size_t)npipes());
      b.init = true;
    }
  }
  return &b;
}

struct Plan {
  int bi, nsplit, jper;
};

Plan plan(int ni, int nj) {
  Plan p;
  p.bi = g_force_bi > 0 ? g_force_bi : (ni >= 16384 ? 256 : (ni >= 2048 ? 128 : 64));
  const int nbi = ceil_div(ni, p.bi);
  const int target = device_cu_count() * g_blocks_per_cu;
  int ns = std::max(1, ceil_div(target, nbi));
  ns = std::min(ns, std::max(1, nj / std::max(1, g_min_jper)));
  int jper = ceil_div(std::max(nj, 1), ns);
  jper = (jper + 63) / 64 * 64;
  p.nsplit = std::max(1, ceil_div(std::max(nj, 1), jper));
  p.jper = jper;
  return p;
}

template <int BI>
void launch_t(const Plan& p, int ni, const float4* xi, int nj, const float4* xj,
              const float2* tab, double ascale, const pgt::Part3& part, double* acc,
              hipStream_t st) {
  const size_t lds = sizeof(float2) * (size_t)g_tab.ntab;
  hipLaunchKernelGGL((pgt::table_force<BI>), dim3(ceil_div(ni, BI), p.nsplit),
                     dim3(pgt::kThreads), lds, st, ni, xi, nj, p.jper, xj, tab, g_tab.ntab,
                     g_tab.frc_bit, ascale, part, acc, p.nsplit);
  PGAMD_HIP_CHECK(hipGetLastError());
}

void run_mlp_dropin(DevBuf& scratch, int ni, const float4* xi, int nj, const float4* xj,
                    double* acc, hipStream_t st);

// Scaled i (float4) and j (float4) device arrays -> acc[ni][3] (device).
void run_table(DevBuf& scratch, int ni, const float4* xi, int nj, const float4* xj, double* acc,
               hipStream_t st) {
  if (ni <= 0) return;
  if (nj <= 0) {
    PGAMD_HIP_CHECK(hipMemsetAsync(acc, 0, sizeof(double) * 3 * (size_t)ni, st));
    return;
  }
  if (g_cut_mode == 1) {
    run_mlp_dropin(scratch, ni, xi, nj, xj, acc, st);
    return;
  }
  const float2* tab = table_dev(st);
  // Reference scales the FP32 sum by float(1/xscale) (phantom_g5.c:46-47).
  const double ascale = (double)(float)(1.0 / g_tab.xscale);
  const Plan p = plan(ni, nj);
  pgt::Part3 part{};
  if (p.nsplit > 1) {
    const size_t n = (size_t)p.nsplit * ni;
    scratch.reserve(n * 3 * sizeof(double));
    part.ax = scratch.as<double>();
    part.ay = part.ax + n;
    part.az = part.ay + n;
  }
  switch (p.bi) {
    case 256: launch_t<256>(p, ni, xi, nj, xj, tab, ascale, part, acc, st); break;
    case 128: launch_t<128>(p, ni, xi, nj, xj, tab, ascale, part, acc, st); break;
    default: launch_t<64>(p, ni, xi, nj, xj, tab, ascale, part, acc, st); break;
  }
  if (p.nsplit > 1) {
    hipLaunchKernelGGL(pgt::table_reduce, dim3(ceil_div(ni, 256)), dim3(256), 0, st, ni,
                       p.nsplit, part, ascale, acc);
    PGAMD_HIP_CHECK(hipGetLastError());
  }
}

inline float4 scale_i(const double* x, double xs) {
  // (float)x * Xscale evaluated in double, stored as float (phantom_g5.c:102-104)
  return make_float4((float)((double)(float)x[0] * xs), (float)((double)(float)x[1] * xs),
                     (float)((double)(float)x[2] * xs), 0.f);
}

void jmem_reserve(Board& b, int need) {
  if (need <= b.hcap) return;
  int cap = b.hcap > 0 ? b.hcap : 4096;
  while (cap < need) cap *= 2;
  b.hj.reserve(sizeof(float4) * (size_t)cap, true);
  std::memset(b.hj.as<float4>() + b.hcap, 0, sizeof(float4) * (size_t)(cap - b.hcap));
  b.hcap = cap;
}

void board_calc(Board& b, const double (*x)[3], const float4* pre, int ni, double (*a)[3],
                double* p) {
  if (ni <= 0) return;
  std::lock_guard<std::mutex> lk(g_mutex);
  ensure_table();
  for (int i = 0; i < ni; ++i) p[i] = 0.0;
  const int nj = b.n;
  if (nj <= 0) {  // nothing to sum over
    std::memset(a, 0, sizeof(double) * 3 * (size_t)ni);
    return;
  }
  const size_t in_bytes = sizeof(float4) * (size_t)ni;
  b.h_in.reserve(in_bytes);
  b.d_in.reserve(in_bytes);
  float4* hin = b.h_in.as<float4>();
  if (pre)
    std::memcpy(hin, pre, in_bytes);
  else
    for (int i = 0; i < ni; ++i) hin[i] = scale_i(x[i], g_tab.xscale);
  upload(b.d_in.ptr, b.h_in, 0, in_bytes, b.stream);
  // Addresses below nj never written read as zero records (no force), as in
  // the reference's zero-initialised static j-memory.
  jmem_reserve(b, nj);
  const size_t need = sizeof(float4) * (size_t)b.hcap;
  if (need > b.dj.bytes) {
    // A grown device buffer starts empty: upload everything written so far.
    b.dj.reserve(need);
    b.lo = 0;
    b.hi = b.hmax;
    b.dvalid = 0;
  }
  int lo = b.lo, hi = b.hi;
  if (nj > b.dvalid) {
    lo = lo < hi ? std::min(lo, b.dvalid) : b.dvalid;
    hi = std::max(hi, nj);
  }
  if (lo < hi) {
    upload(b.dj.as<float4>() + lo, b.hj, sizeof(float4) * (size_t)lo,
           sizeof(float4) * (size_t)(hi - lo), b.stream);
    if (lo <= b.dvalid) b.dvalid = std::max(b.dvalid, hi);
    b.lo = b.hi = 0;
  }
  // The table kernels write the results straight into pinned host memory
  // through its device mapping; the learned cut-off's path copies them from
  // a device buffer.
  const size_t out_bytes = sizeof(double) * 3 * (size_t)ni;
  b.h_out.reserve(out_bytes);
  double* out = mapped<double>(b.h_out);
  if (g_cut_mode == 1) {
    b.d_out.reserve(out_bytes);
    out = b.d_out.as<double>();
  }
  run_table(b.d_part, ni, b.d_in.as<float4>(), nj, b.dj.as<float4>(), out, b.stream);
  if (g_cut_mode == 1)
    PGAMD_HIP_CHECK(hipMemcpyAsync(b.h_out.ptr, b.d_out.ptr, out_bytes, hipMemcpyDeviceToHost,
                                   b.stream));
  PGAMD_HIP_CHECK(hipStreamSynchronize(b.stream));
  std::memcpy(a, b.h_out.ptr, out_bytes);
}

void mark(Board& b, int a0, int a1) {
  b.hmax = std::max(b.hmax, a1);
  if (b.lo >= b.hi) {
    b.lo = a0;
    b.hi = a1;
  } else {
    b.lo = std::min(b.lo, a0);
    b.hi = std::max(b.hi, a1);
  }
}

bool check_range(const char* fn, int adr, int nj) {
  if (adr < 0 || nj < 0 || (long long)adr + nj > kJMemSize) {
    api_error(2, "%s: addresses [%d, %d) outside the j-memory of %d", fn, adr, adr + nj,
              kJMemSize);
    return false;
  }
  return true;
}


// Device copy of the parameters in inference form (mlp_kernels.h), which the
// force and evaluation kernels read.
const float* mlp_dev(hipStream_t st) {
  Mlp& m = g_mlp;
  if (m.dirty) {
    float q[mlp::kParams];
    mlp::to_inference_form(m.params.data(), q);
    m.dev.reserve(sizeof(float) * mlp::kParams);
    PGAMD_HIP_CHECK(hipMemcpyAsync(m.dev.ptr, q, sizeof(float) * mlp::kParams,
                                   hipMemcpyHostToDevice, st));
    PGAMD_HIP_CHECK(hipStreamSynchronize(st));
    m.dirty = false;
  }
  return m.dev.as<float>();
}

// Core length of the learned target Y = f (r^3 + r0^3): the PP softening for
// the S2 shape, r_c/30 for a user shape.
double mlp_r0() { return g_tab.is_s2 ? g_tab.sft_pp : g_tab.rcut / 30.0; }

pgb::MlpCut mlp_cut(hipStream_t st) {
  pgb::MlpCut mc;
  const double xs = g_tab.xscale;
  mc.w = mlp_dev(st);
  mc.inv_rc_s = (float)(1.0 / (g_mlp.rc * xs));
  mc.rc2_s = (float)((g_mlp.rc * xs) * (g_mlp.rc * xs));
  mc.r03_s = (float)std::pow(g_mlp.r0 * xs, 3);
  mc.xs3 = (float)(xs * xs * xs);
  return mc;
}

// Launches the persistent batch kernel (counter reset + kernel) on st.
void launch_batch(int mode, int ngroups, const int* order, const int* ioff, const int* icnt,
                  const int* ilist, const int* joff, const int* jcnt, const int* jlist,
                  const float4* pos, double* acc, int* counter, int bpc, hipStream_t st) {
  hipLaunchKernelGGL(pgb::zero_counter, dim3(1), dim3(64), 0, st, counter);
  PGAMD_HIP_CHECK(hipGetLastError());
  if (ngroups <= 0) return;
  const int blocks = std::max(1, std::min(ngroups, device_cu_count() * bpc));
  const double ascale = (double)(float)(1.0 / g_tab.xscale);
  pgb::TableCut tc{table_dev(st), g_tab.ntab, g_tab.frc_bit};
  if (mode == 1) {
    pgb::MlpCut mc = mlp_cut(st);
    hipLaunchKernelGGL(pgb::treepm_batch_mlp, dim3(blocks), dim3(pgb::kThreads), 0, st, ngroups,
                       order, ioff, icnt, ilist, joff, jcnt, jlist, pos, mc, ascale, acc, counter);
  } else {
    const size_t lds = sizeof(float2) * (size_t)g_tab.ntab;
    hipLaunchKernelGGL((pgb::treepm_batch_table<pgb::kTableBlock>), dim3(blocks),
                       dim3(pgb::kTableBlock), lds, st, ngroups, order, ioff, icnt, ilist, joff,
                       jcnt, jlist, pos, tc, ascale, acc, counter);
  }
  PGAMD_HIP_CHECK(hipGetLastError());
}

int default_bpc(int mode) {
  if (g_batch_bpc[mode] > 0) return g_batch_bpc[mode];
  return mode == 1 ? 4 : 2;
}

// Drop-in force through the batch kernel (used in MLP mode): device i and j
// float4 arrays are placed in one array [j..., i...] with index lists.
void run_mlp_dropin(DevBuf& scratch, int ni, const float4* xi, int nj, const float4* xj,
                    double* acc, hipStream_t st) {
  const int ng = ceil_div(ni, pgb::kThreads);
  const size_t pos_b = sizeof(float4) * (size_t)(ni + nj);
  const size_t acc_b = sizeof(double) * 3 * (size_t)(ni + nj);
  const size_t idx_n = (size_t)ni + nj + 4 * (size_t)ng + 1;
  scratch.reserve(pos_b + acc_b + sizeof(int) * idx_n + 256);
  char* base = static_cast<char*>(scratch.ptr);
  float4* pos = reinterpret_cast<float4*>(base);
  double* accw = reinterpret_cast<double*>(base + pos_b);
  int* ilist = reinterpret_cast<int*>(base + pos_b + acc_b);
  int* jlist = ilist + ni;
  int* ioff = jlist + nj;
  int* icnt = ioff + ng;
  int* joff = icnt + ng;
  int* jcnt = joff + ng;
  int* counter = jcnt + ng;
  PGAMD_HIP_CHECK(hipMemcpyAsync(pos, xj, sizeof(float4) * (size_t)nj, hipMemcpyDeviceToDevice, st));
  PGAMD_HIP_CHECK(hipMemcpyAsync(pos + nj, xi, sizeof(float4) * (size_t)ni, hipMemcpyDeviceToDevice, st));
  std::vector<int> h((size_t)ni + nj + 4 * (size_t)ng);
  for (int i = 0; i < ni; ++i) h[i] = nj + i;
  for (int j = 0; j < nj; ++j) h[(size_t)ni + j] = j;
  for (int g = 0; g < ng; ++g) {
    h[(size_t)ni + nj + g] = g * pgb::kThreads;
    h[(size_t)ni + nj + ng + g] = std::min(pgb::kThreads, ni - g * pgb::kThreads);
    h[(size_t)ni + nj + 2 * ng + g] = 0;
    h[(size_t)ni + nj + 3 * ng + g] = nj;
  }
  PGAMD_HIP_CHECK(hipMemcpyAsync(ilist, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice, st));
  launch_batch(1, ng, nullptr, ioff, icnt, ilist, joff, jcnt, jlist, pos, accw, counter,
               default_bpc(1), st);
  PGAMD_HIP_CHECK(hipMemcpyAsync(acc, accw + 3 * (size_t)nj, sizeof(double) * 3 * (size_t)ni,
                                 hipMemcpyDeviceToDevice, st));
  PGAMD_HIP_CHECK(hipStreamSynchronize(st));
}

}  // namespace

// ------------------------------------------------------------ table control

PGAMD_API void pg5_set_xscale(double xscale) {
  std::lock_guard<std::mutex> lk(g_mutex);
  g_tab.xscale = xscale;
}

PGAMD_API void pg5_gen_force_table(double (*force_func)(double), double rcut) {
  std::lock_guard<std::mutex> lk(g_mutex);
  g_tab.is_s2 = false;
  build_table(force_func, rcut, g_tab.frc_bit);
}

PGAMD_API void pg5_gen_s2_force_table(double sft_for_PP, double sft_for_PM) {
  std::lock_guard<std::mutex> lk(g_mutex);
  g_tab.sft_pp = sft_for_PP;
  g_tab.sft_pm = sft_for_PM;
  g_tab.is_s2 = true;
  build_table(s2_force, sft_for_PM, g_tab.frc_bit);
}

// Table resolution: 2^frc_bit entries per octave of s (reference: 5, i.e. 512
// entries; 6 gives 1024).  Rebuilds the current shape.  No counterpart in the
// reference, where FRC_BIT is a compile-time constant (pg5_table.h:7-15).
PGAMD_API int pg5_set_table_frc_bit(int frc_bit) {
  if (frc_bit < 2 || frc_bit > 9) {
    api_error(3, "pg5_set_table_frc_bit: %d not in [2, 9]", frc_bit);
    return -1;
  }
  std::lock_guard<std::mutex> lk(g_mutex);
  double (*f)(double) = g_tab.func ? g_tab.func : s2_force;
  const double rc = g_tab.func ? g_tab.rcut : g_tab.sft_pm;
  build_table(f, rc, frc_bit);
  return 0;
}

// Copies the current table (2*ntab floats) into out if non-null; returns ntab.
PGAMD_API int pg5_get_table(float* out, double* xscale, float* r2scale) {
  std::lock_guard<std::mutex> lk(g_mutex);
  ensure_table();
  if (out) std::memcpy(out, g_tab.host.data(), sizeof(float) * g_tab.host.size());
  if (xscale) *xscale = g_tab.xscale;
  if (r2scale) *r2scale = g_tab.r2scale;
  return g_tab.ntab;
}

// ------------------------------------------------------------------ GRAPE-5 API

PGAMD_API int g5_get_number_of_pipelines(void) { return npipes(); }
PGAMD_API int g5_get_jmemsize(void) { return kJMemSize; }

// Opening the board readies the device as the hardware's open did: the
// default board's stream, the table in HBM, and the library's kernels loaded
// (one empty launch), so the first force call pays none of it.
namespace {
__global__ void open_warm(int* flag) {
  if (flag) *flag = 0;
}
}  // namespace

PGAMD_API void g5_open(void) {
  ensure_device();
  Board* b = board(0, "g5_open");
  std::lock_guard<std::mutex> lk(g_mutex);
  g_tab.sft_pp = kSftPP;
  g_tab.sft_pm = kSftPM;
  g_tab.is_s2 = true;
  build_table(s2_force, kSftPM, g_tab.frc_bit);
  table_dev(b->stream);
  hipLaunchKernelGGL(open_warm, dim3(1), dim3(64), 0, b->stream, nullptr);
  PGAMD_HIP_CHECK(hipGetLastError());
  PGAMD_HIP_CHECK(hipStreamSynchronize(b->stream));
}
PGAMD_API void g5_close(void) {
  for (int d = 0; d < kMaxDev; ++d)
    if (g_boards[d].init) PGAMD_HIP_CHECK(hipStreamSynchronize(g_boards[d].stream));
}

// Accepted and ignored, as in the reference (phantom_g5.c:69-84).
PGAMD_API void g5_set_eta(double) {}
PGAMD_API void g5_set_eps_to_all(double) {}
PGAMD_API void g5_set_range(double, double, double) {}
PGAMD_API void g5_set_cutoff_table(double (*)(double), double, double, double (*)(double),
                                   double, double) {}

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
  Board* b = board(devid, "g5_set_xi");
  if (!b) return;
  if (ni < 0 || ni > npipes()) {
    api_error(3, "g5_set_xi: ni=%d exceeds the %d pipelines", ni, npipes());
    return;
  }
  std::lock_guard<std::mutex> lk(g_mutex);
  ensure_table();
  for (int i = 0; i < ni; ++i) b->ipipe.as<float4>()[i] = scale_i(xi[i], g_tab.xscale);
  b->ni_pipe = ni;
}
PGAMD_API void g5_set_xi(int ni, double (*xi)[3]) { g5_set_xiMC(0, ni, xi); }

// j-memory: {x*xs, y*xs, z*xs, m} in float (phantom_g5.c:112-128).  Like the
// reference, xj and mj are indexed by the absolute address adr..adr+nj-1.
PGAMD_API void g5_set_xmjMC(int devid, int adr, int nj, double (*xj)[3], double* mj) {
  Board* b = board(devid, "g5_set_xmj");
  if (!b || !check_range("g5_set_xmj", adr, nj) || nj == 0) return;
  std::lock_guard<std::mutex> lk(g_mutex);
  ensure_table();
  jmem_reserve(*b, adr + nj);
  const float xs = (float)g_tab.xscale;
  float4* h = b->hj.as<float4>();
  for (int j = adr; j < adr + nj; ++j)
    h[j] = make_float4((float)xj[j][0] * xs, (float)xj[j][1] * xs, (float)xj[j][2] * xs,
                       (float)mj[j]);
  mark(*b, adr, adr + nj);
}
PGAMD_API void g5_set_xmj(int adr, int nj, double (*xj)[3], double* mj) {
  g5_set_xmjMC(0, adr, nj, xj, mj);
}

// Declared but never defined by the reference (gp5util.h:72-73); provided
// here with the obvious meaning: update positions or masses only.
PGAMD_API void g5_set_xj(int adr, int nj, double (*xj)[3]) {
  Board* b = board(0, "g5_set_xj");
  if (!b || !check_range("g5_set_xj", adr, nj) || nj == 0) return;
  std::lock_guard<std::mutex> lk(g_mutex);
  ensure_table();
  jmem_reserve(*b, adr + nj);
  const float xs = (float)g_tab.xscale;
  float4* h = b->hj.as<float4>();
  for (int j = adr; j < adr + nj; ++j) {
    h[j].x = (float)xj[j][0] * xs;
    h[j].y = (float)xj[j][1] * xs;
    h[j].z = (float)xj[j][2] * xs;
  }
  mark(*b, adr, adr + nj);
}
PGAMD_API void g5_set_mj(int adr, int nj, double* mj) {
  Board* b = board(0, "g5_set_mj");
  if (!b || !check_range("g5_set_mj", adr, nj) || nj == 0) return;
  std::lock_guard<std::mutex> lk(g_mutex);
  jmem_reserve(*b, adr + nj);
  float4* h = b->hj.as<float4>();
  for (int j = adr; j < adr + nj; ++j) h[j].w = (float)mj[j];
  mark(*b, adr, adr + nj);
}

PGAMD_API void g5_runMC(int devid) {
  Board* b = board(devid, "g5_run");
  if (!b) return;
  const int ni = b->ni_pipe;
  if (ni <= 0) return;
  double(*a)[3] = reinterpret_cast<double(*)[3]>(b->opipe.as<double>());
  std::vector<double> p((size_t)ni);
  board_calc(*b, nullptr, b->ipipe.as<float4>(), ni, a, p.data());
}
PGAMD_API void g5_run(void) { g5_runMC(0); }

PGAMD_API void g5_get_forceMC(int devid, int ni, double (*a)[3], double* pot) {
  Board* b = board(devid, "g5_get_force");
  if (!b) return;
  if (ni < 0 || ni > b->ni_pipe) {
    api_error(3, "g5_get_force: ni=%d but %d i-particles were run", ni, b->ni_pipe);
    return;
  }
  std::memcpy(a, b->opipe.ptr, sizeof(double) * 3 * (size_t)ni);
  for (int i = 0; i < ni; ++i) pot[i] = 0.0;
}
PGAMD_API void g5_get_force(int ni, double (*a)[3], double* p) { g5_get_forceMC(0, ni, a, p); }

PGAMD_API void g5_calculate_force_on_xMC(int devid, double (*x)[3], double (*a)[3], double* p,
                                         int ni) {
  Board* b = board(devid, "g5_calculate_force_on_x");
  if (!b) return;
  board_calc(*b, x, nullptr, ni, a, p);
}
PGAMD_API void g5_calculate_force_on_x(double (*x)[3], double (*a)[3], double* p, int ni) {
  g5_calculate_force_on_xMC(0, x, a, p, ni);
}

// ------------------------------------------------- device-pointer entry points

// Scaled positions (device float4 {x*xs, y*xs, z*xs, m} for j, {.., 0} for i)
// -> acc[ni][3] (device, FP64, unscaled).  Async on stream (nullptr = board's).
PGAMD_API int pg5t_dev_force(int devid, int ni, const void* xi, int nj, const void* xj,
                             double* acc, void* stream) {
  Board* b = board(devid, "pg5t_dev_force");
  if (!b) return -1;
  if (ni < 0 || nj < 0) {
    api_error(3, "pg5t_dev_force: negative ni (%d) or nj (%d)", ni, nj);
    return -1;
  }
  std::lock_guard<std::mutex> lk(g_mutex);
  ensure_table();
  run_table(b->d_part, ni, static_cast<const float4*>(xi), nj, static_cast<const float4*>(xj),
            acc, stream ? (hipStream_t)stream : b->stream);
  return 0;
}

// Evaluates the current table at n values of s (device arrays), synchronously.
PGAMD_API int pg5t_dev_eval(int n, const float* s, float* out) {
  Board* b = board(0, "pg5t_dev_eval");
  if (!b) return -1;
  std::lock_guard<std::mutex> lk(g_mutex);
  ensure_table();
  const float2* tab = table_dev(b->stream);
  if (n > 0) {
    hipLaunchKernelGGL(pgt::table_eval, dim3(ceil_div(n, 256)), dim3(256), 0, b->stream, n, s,
                       tab, g_tab.ntab, g_tab.frc_bit, out);
    PGAMD_HIP_CHECK(hipGetLastError());
  }
  PGAMD_HIP_CHECK(hipStreamSynchronize(b->stream));
  return 0;
}

PGAMD_API int pg5t_dev_set_launch(int bi, int blocks_per_cu, int min_jper) {
  if (bi != 0 && bi != 64 && bi != 128 && bi != 256) {
    api_error(3, "pg5t_dev_set_launch: bi=%d not in {0,64,128,256}", bi);
    return -1;
  }
  g_force_bi = bi;
  if (blocks_per_cu > 0) g_blocks_per_cu = blocks_per_cu;
  if (min_jper > 0) g_min_jper = min_jper;
  return 0;
}

PGAMD_API void* pg5t_dev_stream(int devid) {
  Board* b = board(devid, "pg5t_dev_stream");
  return b ? (void*)b->stream : nullptr;
}

// ------------------------------------------------------- learned cut-off

// 0 = table (default, the reference's method), 1 = learned MLP (requires a
// trained or loaded network).  Applies to g5_calculate_force_on_x* and
// pg5t_dev_force; the batch calls take the mode as an argument.
PGAMD_API int pg5_set_cutoff_mode(int mode) {
  std::lock_guard<std::mutex> lk(g_mutex);
  if (mode != 0 && mode != 1) {
    api_error(3, "pg5_set_cutoff_mode: mode %d not in {0, 1}", mode);
    return -1;
  }
  if (mode == 1 && !g_mlp.valid) {
    api_error(5, "pg5_set_cutoff_mode: no learned cut-off (train or load one first)");
    return -1;
  }
  g_cut_mode = mode;
  return 0;
}

PGAMD_API int pg5_get_cutoff_mode(void) { return g_cut_mode; }

// Loads a network: p[321] (layout in mlp_kernels.h), trained for cut-off
// radius rc and core length r0 (physical units).
PGAMD_API int pg5_mlp_set_params(const float* p, double rc, double r0) {
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!(rc > 0.0) || !(r0 >= 0.0)) {
    api_error(3, "pg5_mlp_set_params: rc=%g r0=%g", rc, r0);
    return -1;
  }
  g_mlp.params.assign(p, p + mlp::kParams);
  g_mlp.rc = rc;
  g_mlp.r0 = r0;
  g_mlp.valid = true;
  g_mlp.dirty = true;
  return 0;
}

// Copies the current network (p may be null); returns 1 if one is loaded.
PGAMD_API int pg5_mlp_get_params(float* p, double* rc, double* r0) {
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!g_mlp.valid) return 0;
  if (p) std::memcpy(p, g_mlp.params.data(), sizeof(float) * mlp::kParams);
  if (rc) *rc = g_mlp.rc;
  if (r0) *r0 = g_mlp.r0;
  return 1;
}

namespace {
struct TrainBufs {
  DevBuf v, t, wt, w, m, u, g, loss, gsum;
};

// One gradient evaluation (+ optional Adam step) on n samples at device
// offset `off` of the training set; returns the summed loss through gsum.
void mlp_step(TrainBufs& tb, int n, int off, int adam, int step, float lr, hipStream_t st) {
  hipLaunchKernelGGL(mlp::mlp_sample_grad, dim3(ceil_div(n, 256)), dim3(256), 0, st, n,
                     tb.v.as<float>() + off, tb.t.as<float>() + off, tb.wt.as<float>() + off,
                     tb.w.as<float>(), tb.g.as<float>(), tb.loss.as<float>());
  PGAMD_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(mlp::mlp_reduce_adam, dim3(mlp::kParams + 1), dim3(256), 0, st, n,
                     tb.g.as<float>(), tb.loss.as<float>(), tb.w.as<float>(), tb.m.as<float>(),
                     tb.u.as<float>(), tb.gsum.as<float>(), adam, step, lr, 0.9f, 0.999f, 1e-8f);
  PGAMD_HIP_CHECK(hipGetLastError());
}

void tb_alloc(TrainBufs& tb, int nset, int batch) {
  tb.v.reserve(sizeof(float) * nset);
  tb.t.reserve(sizeof(float) * nset);
  tb.wt.reserve(sizeof(float) * nset);
  tb.w.reserve(sizeof(float) * mlp::kParams);
  tb.m.reserve(sizeof(float) * mlp::kParams);
  tb.u.reserve(sizeof(float) * mlp::kParams);
  tb.g.reserve(sizeof(float) * (size_t)mlp::kParams * batch);
  tb.loss.reserve(sizeof(float) * batch);
  tb.gsum.reserve(sizeof(float) * (mlp::kParams + 1));
}

void tb_release(TrainBufs& tb) {
  for (DevBuf* d : {&tb.v, &tb.t, &tb.wt, &tb.w, &tb.m, &tb.u, &tb.g, &tb.loss, &tb.gsum})
    d->release();
}

// Full-batch loss and gradient at parameters x (host, FP64) over the nset
// training samples on the device: the per-sample kernels in FP32, the sums in
// FP64 in a fixed order.
double full_loss_grad(TrainBufs& tb, int nset, const std::vector<double>& x,
                      std::vector<double>& g, hipStream_t st) {
  std::vector<float> xf(x.begin(), x.end());
  PGAMD_HIP_CHECK(hipMemcpyAsync(tb.w.ptr, xf.data(), sizeof(float) * mlp::kParams,
                                 hipMemcpyHostToDevice, st));
  mlp_step(tb, nset, 0, 0, 1, 0.f, st);
  std::vector<float> gs(mlp::kParams + 1);
  PGAMD_HIP_CHECK(hipMemcpyAsync(gs.data(), tb.gsum.ptr, sizeof(float) * gs.size(),
                                 hipMemcpyDeviceToHost, st));
  PGAMD_HIP_CHECK(hipStreamSynchronize(st));
  g.assign(gs.begin(), gs.begin() + mlp::kParams);
  return gs[mlp::kParams];
}

// Limited-memory BFGS (two-loop recursion, 20 pairs) with Armijo
// backtracking, on full-batch GPU gradients.  Returns the final loss.
double lbfgs(TrainBufs& tb, int nset, std::vector<double>& x, int iters, hipStream_t st) {
  const int M = 20, P = mlp::kParams;
  std::vector<std::vector<double>> S, Y;
  std::vector<double> rho;
  std::vector<double> g, gn, d(P), xn(P);
  double f = full_loss_grad(tb, nset, x, g, st);
  auto dot = [](const std::vector<double>& a, const std::vector<double>& b) {
    double s = 0;
    for (size_t k = 0; k < a.size(); ++k) s += a[k] * b[k];
    return s;
  };
  for (int it = 0; it < iters; ++it) {
    // d = -H g
    std::vector<double> q = g, al(S.size());
    for (int k = (int)S.size() - 1; k >= 0; --k) {
      al[k] = rho[k] * dot(S[k], q);
      for (int p = 0; p < P; ++p) q[p] -= al[k] * Y[k][p];
    }
    const double gamma = S.empty() ? 1e-3 / std::max(1e-30, std::sqrt(dot(g, g)))
                                   : dot(S.back(), Y.back()) / dot(Y.back(), Y.back());
    for (int p = 0; p < P; ++p) q[p] *= gamma;
    for (size_t k = 0; k < S.size(); ++k) {
      const double b = rho[k] * dot(Y[k], q);
      for (int p = 0; p < P; ++p) q[p] += S[k][p] * (al[k] - b);
    }
    for (int p = 0; p < P; ++p) d[p] = -q[p];
    double gd = dot(g, d);
    if (!(gd < 0)) {  // not a descent direction: restart from steepest descent
      S.clear();
      Y.clear();
      rho.clear();
      for (int p = 0; p < P; ++p) d[p] = -g[p] * 1e-3 / std::max(1e-30, std::sqrt(dot(g, g)));
      gd = dot(g, d);
    }
    double step = 1.0, fn = 0.0;
    bool ok = false;
    for (int ls = 0; ls < 30; ++ls) {
      for (int p = 0; p < P; ++p) xn[p] = x[p] + step * d[p];
      fn = full_loss_grad(tb, nset, xn, gn, st);
      if (std::isfinite(fn) && fn <= f + 1e-4 * step * gd) {
        ok = true;
        break;
      }
      step *= 0.5;
    }
    if (!ok) break;
    std::vector<double> s(P), y(P);
    for (int p = 0; p < P; ++p) {
      s[p] = xn[p] - x[p];
      y[p] = gn[p] - g[p];
    }
    const double sy = dot(s, y);
    if (sy > 1e-20) {
      if ((int)S.size() == M) {
        S.erase(S.begin());
        Y.erase(Y.begin());
        rho.erase(rho.begin());
      }
      S.push_back(s);
      Y.push_back(y);
      rho.push_back(1.0 / sy);
    }
    x = xn;
    g = gn;
    const double rel = (f - fn) / std::max(f, 1e-300);
    f = fn;
    if (rel < 1e-12 && it > 50) break;
  }
  return f;
}
}  // namespace

// Summed gradient and loss of the weighted MSE sum_s wt_s (Y(v_s) - t_s)^2 / n
// for host arrays v, t, wt (n samples) at parameters p[321], computed by the
// training kernels on the GPU.  For checking them against a CPU reference.
PGAMD_API int pg5_mlp_grad(int n, const float* v, const float* t, const float* wt,
                           const float* p, float* grad, double* loss) {
  Board* b = board(0, "pg5_mlp_grad");
  if (!b) return -1;
  if (n <= 0) {
    api_error(3, "pg5_mlp_grad: n=%d", n);
    return -1;
  }
  TrainBufs tb;
  tb_alloc(tb, n, n);
  hipStream_t st = b->stream;
  PGAMD_HIP_CHECK(hipMemcpyAsync(tb.v.ptr, v, sizeof(float) * n, hipMemcpyHostToDevice, st));
  PGAMD_HIP_CHECK(hipMemcpyAsync(tb.t.ptr, t, sizeof(float) * n, hipMemcpyHostToDevice, st));
  PGAMD_HIP_CHECK(hipMemcpyAsync(tb.wt.ptr, wt, sizeof(float) * n, hipMemcpyHostToDevice, st));
  PGAMD_HIP_CHECK(hipMemcpyAsync(tb.w.ptr, p, sizeof(float) * mlp::kParams,
                                 hipMemcpyHostToDevice, st));
  mlp_step(tb, n, 0, 0, 1, 0.f, st);
  std::vector<float> gs(mlp::kParams + 1);
  PGAMD_HIP_CHECK(hipMemcpyAsync(gs.data(), tb.gsum.ptr, sizeof(float) * gs.size(),
                                 hipMemcpyDeviceToHost, st));
  PGAMD_HIP_CHECK(hipStreamSynchronize(st));
  std::memcpy(grad, gs.data(), sizeof(float) * mlp::kParams);
  if (loss) *loss = gs[mlp::kParams];
  tb_release(tb);
  return 0;
}

// Trains the network on the GPU for the current cut-off shape (the function
// the table was generated from, pg5_gen_force_table / pg5_gen_s2_force_table):
// 65536 seeded samples of r in [0, r_c) (half uniform in v = sqrt(r/r_c),
// half uniform in volume), target Y = f(r) (r^3 + r0^3) evaluated in FP64,
// weight 1 / (|Y| + 0.05)^2 (relative accuracy where Y is O(1)).  Optimiser:
// min(steps, 1000) Adam steps on mini-batches of `batch` at rate lr, then up
// to `steps` L-BFGS iterations on the full set (loss and gradient from the GPU
// kernels, the two-loop recursion on the host).  Deterministic for a given
// seed.  *loss_out = weighted MSE on the full set.
PGAMD_API int pg5_mlp_train(int steps, int batch, double lr, unsigned seed, double* loss_out) {
  Board* b = board(0, "pg5_mlp_train");
  if (!b) return -1;
  if (steps <= 0 || batch <= 0 || batch > (1 << 16) || !(lr > 0.0)) {
    api_error(3, "pg5_mlp_train: steps=%d batch=%d lr=%g", steps, batch, lr);
    return -1;
  }
  std::lock_guard<std::mutex> lk(g_mutex);
  ensure_table();
  double (*func)(double) = g_tab.func ? g_tab.func : s2_force;
  const double rc = g_tab.rcut;
  const double r0 = mlp_r0();
  const double r03 = r0 * r0 * r0;
  const int nset = 1 << 16;
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  std::vector<float> hv(nset), ht(nset), hw(nset);
  for (int k = 0; k < nset; ++k) {
    double r, v;
    if (k & 1) {
      r = rc * std::cbrt(U(rng));
      v = std::sqrt(r / rc);
    } else {
      v = U(rng);
      r = rc * v * v;
    }
    const double y = func(r) * (r * r * r + r03);
    hv[k] = (float)v;
    ht[k] = (float)y;
    hw[k] = (float)(1.0 / ((std::fabs(y) + 0.05) * (std::fabs(y) + 0.05)));
  }
  // initial parameters
  std::vector<float> p(mlp::kParams, 0.f);
  std::uniform_real_distribution<float> Uf(-1.f, 1.f);
  for (int j = 0; j < mlp::kH; ++j) {
    p[mlp::kW1 + j] = 4.f * Uf(rng);
    p[mlp::kB1 + j] = 2.f * Uf(rng);
  }
  const float lim = std::sqrt(6.f / (2 * mlp::kH));
  for (int k = 0; k < mlp::kH * mlp::kH; ++k) p[mlp::kW2 + k] = lim * Uf(rng);
  for (int k = 0; k < mlp::kH; ++k) p[mlp::kW3 + k] = lim * Uf(rng);
  double mean = 0.0;
  for (float y : ht) mean += y;
  p[mlp::kB3] = (float)(mean / nset);

  TrainBufs tb;
  tb_alloc(tb, nset, nset);
  hipStream_t st = b->stream;
  PGAMD_HIP_CHECK(hipMemcpyAsync(tb.v.ptr, hv.data(), sizeof(float) * nset, hipMemcpyHostToDevice, st));
  PGAMD_HIP_CHECK(hipMemcpyAsync(tb.t.ptr, ht.data(), sizeof(float) * nset, hipMemcpyHostToDevice, st));
  PGAMD_HIP_CHECK(hipMemcpyAsync(tb.wt.ptr, hw.data(), sizeof(float) * nset, hipMemcpyHostToDevice, st));
  PGAMD_HIP_CHECK(hipMemcpyAsync(tb.w.ptr, p.data(), sizeof(float) * mlp::kParams, hipMemcpyHostToDevice, st));
  std::vector<float> zero(mlp::kParams, 0.f);
  PGAMD_HIP_CHECK(hipMemcpyAsync(tb.m.ptr, zero.data(), sizeof(float) * mlp::kParams, hipMemcpyHostToDevice, st));
  PGAMD_HIP_CHECK(hipMemcpyAsync(tb.u.ptr, zero.data(), sizeof(float) * mlp::kParams, hipMemcpyHostToDevice, st));
  // Adam warm-up on mini-batches, then L-BFGS on the full set
  const int warm = std::min(steps, 1000);
  const int span = nset - batch + 1;
  for (int step = 1; step <= warm; ++step) {
    const int off = (int)(((long long)(step - 1) * batch) % span);
    mlp_step(tb, batch, off, 1, step, (float)lr, st);
  }
  PGAMD_HIP_CHECK(hipMemcpyAsync(p.data(), tb.w.ptr, sizeof(float) * mlp::kParams,
                                 hipMemcpyDeviceToHost, st));
  PGAMD_HIP_CHECK(hipStreamSynchronize(st));
  std::vector<double> x(p.begin(), p.end());
  const double tot = lbfgs(tb, nset, x, steps, st);
  for (int k = 0; k < mlp::kParams; ++k) p[k] = (float)x[k];
  tb_release(tb);
  g_mlp.params = p;
  g_mlp.rc = rc;
  g_mlp.r0 = r0;
  g_mlp.valid = true;
  g_mlp.dirty = true;
  if (loss_out) *loss_out = tot;
  return 0;
}

// Learned f(r) at n physical radii (device arrays), synchronously.
PGAMD_API int pg5_mlp_eval_dev(int n, const float* r, float* f) {
  Board* b = board(0, "pg5_mlp_eval_dev");
  if (!b) return -1;
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!g_mlp.valid) {
    api_error(5, "pg5_mlp_eval_dev: no learned cut-off");
    return -1;
  }
  const float* w = mlp_dev(b->stream);
  if (n > 0) {
    hipLaunchKernelGGL(mlp::mlp_eval_f, dim3(ceil_div(n, 256)), dim3(256), 0, b->stream, n, r, w,
                       (float)g_mlp.rc, (float)(g_mlp.r0 * g_mlp.r0 * g_mlp.r0), f);
    PGAMD_HIP_CHECK(hipGetLastError());
  }
  PGAMD_HIP_CHECK(hipStreamSynchronize(b->stream));
  return 0;
}

// ------------------------------------------------------------ batched TreePM

namespace {
bool batch_check(const char* fn, int mode, int ngroups) {
  if (mode != 0 && mode != 1) {
    api_error(3, "%s: mode %d not in {0, 1}", fn, mode);
    return false;
  }
  if (mode == 1 && !g_mlp.valid) {
    api_error(5, "%s: no learned cut-off (train or load one first)", fn);
    return false;
  }
  if (ngroups < 0) {
    api_error(3, "%s: ngroups=%d", fn, ngroups);
    return false;
  }
  return true;
}

struct BatchGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  DevBuf counter;
};
}  // namespace

// Short-range force of every group in one persistent launch (treepm_batch.h):
// pos[np] float4 scaled {x*xs, y*xs, z*xs, m}; group g: i-particles
// ilist[ioff[g] .. +icnt[g]), j-particles jlist[joff[g] .. +jcnt[g]); order
// (nullable) = processing order of the groups; acc[np][3] FP64 written for
// the listed i.  mode 0 = table, 1 = learned MLP.  All device pointers;
// asynchronous on stream.  Calls on one devid share its work counter.
PGAMD_API int pg5t_batch_force(int devid, int ngroups, const int* ioff, const int* icnt,
                               const int* ilist, const int* joff, const int* jcnt,
                               const int* jlist, const int* order, const void* pos, double* acc,
                               int mode, void* stream) {
  Board* b = board(devid, "pg5t_batch_force");
  if (!b) return -1;
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!batch_check("pg5t_batch_force", mode, ngroups)) return -1;
  ensure_table();
  hipStream_t st = stream ? (hipStream_t)stream : b->stream;
  b->d_counter.reserve(256);
  launch_batch(mode, ngroups, order, ioff, icnt, ilist, joff, jcnt, jlist,
               static_cast<const float4*>(pos), acc, b->d_counter.as<int>(), default_bpc(mode),
               st);
  return 0;
}

// Captures the same launch (counter reset + persistent kernel) into a HIP
// graph with its own work counter; replay with pg5t_batch_graph_launch.
// Inputs are read at replay time from the same device addresses.
PGAMD_API void* pg5t_batch_graph(int devid, int ngroups, const int* ioff, const int* icnt,
                                 const int* ilist, const int* joff, const int* jcnt,
                                 const int* jlist, const int* order, const void* pos,
                                 double* acc, int mode) {
  Board* b = board(devid, "pg5t_batch_graph");
  if (!b) return nullptr;
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!batch_check("pg5t_batch_graph", mode, ngroups)) return nullptr;
  ensure_table();
  // uploads happen before capture
  table_dev(b->stream);
  if (mode == 1) mlp_dev(b->stream);
  BatchGraph* bg = new BatchGraph();
  bg->counter.reserve(256);
  hipStream_t cs;
  PGAMD_HIP_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  PGAMD_HIP_CHECK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
  launch_batch(mode, ngroups, order, ioff, icnt, ilist, joff, jcnt, jlist,
               static_cast<const float4*>(pos), acc, bg->counter.as<int>(), default_bpc(mode), cs);
  PGAMD_HIP_CHECK(hipStreamEndCapture(cs, &bg->graph));
  PGAMD_HIP_CHECK(hipGraphInstantiate(&bg->exec, bg->graph, nullptr, nullptr, 0));
  PGAMD_HIP_CHECK(hipStreamDestroy(cs));
  return bg;
}

PGAMD_API int pg5t_batch_graph_launch(void* h, void* stream) {
  if (!h) {
    api_error(3, "pg5t_batch_graph_launch: null graph");
    return -1;
  }
  BatchGraph* bg = static_cast<BatchGraph*>(h);
  hipStream_t st = stream ? (hipStream_t)stream : board(0, "pg5t_batch_graph_launch")->stream;
  PGAMD_HIP_CHECK(hipGraphLaunch(bg->exec, st));
  return 0;
}

PGAMD_API int pg5t_batch_graph_destroy(void* h) {
  if (!h) return 0;
  BatchGraph* bg = static_cast<BatchGraph*>(h);
  PGAMD_HIP_CHECK(hipDeviceSynchronize());
  PGAMD_HIP_CHECK(hipGraphExecDestroy(bg->exec));
  PGAMD_HIP_CHECK(hipGraphDestroy(bg->graph));
  bg->counter.release();
  delete bg;
  return 0;
}

// Autotuner: times the batch launch for persistent-grid sizes of
// {1, 2, 3, 4, 6, 8, 12, 16} blocks per CU on these inputs (one warm-up, then
// the mean of `reps`), keeps the fastest for this mode and returns it through
// *best; ms[8] (nullable) receives the timings.
PGAMD_API int pg5t_batch_autotune(int devid, int ngroups, const int* ioff, const int* icnt,
                                  const int* ilist, const int* joff, const int* jcnt,
                                  const int* jlist, const int* order, const void* pos,
                                  double* acc, int mode, int reps, int* best, float* ms) {
  Board* b = board(devid, "pg5t_batch_autotune");
  if (!b) return -1;
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!batch_check("pg5t_batch_autotune", mode, ngroups)) return -1;
  ensure_table();
  const int cand[8] = {1, 2, 3, 4, 6, 8, 12, 16};
  hipStream_t st = b->stream;
  b->d_counter.reserve(256);
  hipEvent_t e0, e1;
  PGAMD_HIP_CHECK(hipEventCreate(&e0));
  PGAMD_HIP_CHECK(hipEventCreate(&e1));
  float bestt = 1e30f;
  int bestc = default_bpc(mode);
  reps = std::max(1, reps);
  for (int k = 0; k < 8; ++k) {
    auto go = [&] {
      launch_batch(mode, ngroups, order, ioff, icnt, ilist, joff, jcnt, jlist,
                   static_cast<const float4*>(pos), acc, b->d_counter.as<int>(), cand[k], st);
    };
    go();
    PGAMD_HIP_CHECK(hipEventRecord(e0, st));
    for (int r = 0; r < reps; ++r) go();
    PGAMD_HIP_CHECK(hipEventRecord(e1, st));
    PGAMD_HIP_CHECK(hipEventSynchronize(e1));
    float t = 0.f;
    PGAMD_HIP_CHECK(hipEventElapsedTime(&t, e0, e1));
    t /= reps;
    if (ms) ms[k] = t;
    if (t < bestt) {
      bestt = t;
      bestc = cand[k];
    }
  }
  PGAMD_HIP_CHECK(hipEventDestroy(e0));
  PGAMD_HIP_CHECK(hipEventDestroy(e1));
  g_batch_bpc[mode] = bestc;
  if (best) *best = bestc;
  return 0;
}

PGAMD_API int pg5t_batch_set_blocks_per_cu(int mode, int bpc) {
  if ((mode != 0 && mode != 1) || bpc < 0 || bpc > 64) {
    api_error(3, "pg5t_batch_set_blocks_per_cu: mode=%d bpc=%d", mode, bpc);
    return -1;
  }
  g_batch_bpc[mode] = bpc;
  return 0;
}

// ======================================================================
// Long-range (particle-mesh) side of the TreePM force: the periodic force of
// the S2 shape R(r, eps_PM) that the table subtracts (pm_kernels.h).  No
// counterpart in the reference, which ships only the short-range side
// (PG5/table/libpg5/pppm.h:4-6 names its mesh size).
namespace {

#define PGAMD_FFT_CHECK(expr)                                                        \
  do {                                                                              \
    hipfftResult _r = (expr);                                                       \
    if (_r != HIPFFT_SUCCESS) {                                                     \
      std::fprintf(stderr, "pgamd: hipFFT error %d at %s:%d in %s\n", (int)_r,      \
                   __FILE__, __LINE__, #expr);                                      \
      std::abort();                                                                 \
    }                                                                               \
  } while (0)

struct PmCache {
  int M = 0;
  hipfftHandle fwd{}, inv{};
  DevBuf rho, rhok, fk, f;
  DevBuf keys, order, sort_tmp;  // particle sort by mesh cell (pm_keys)
};
PmCache g_pm;
std::mutex g_pm_mutex;

void pm_setup(PmCache& c, int M) {
  if (c.M == M) return;
  if (c.M) {
    PGAMD_FFT_CHECK(hipfftDestroy(c.fwd));
    PGAMD_FFT_CHECK(hipfftDestroy(c.inv));
  }
  const size_t nr = (size_t)M * M * M, nk = (size_t)M * M * (M / 2 + 1);
  c.rho.reserve(sizeof(double) * nr);
  c.rhok.reserve(sizeof(double2) * nk);
  c.fk.reserve(sizeof(double2) * nk * 3);
  c.f.reserve(sizeof(double) * nr * 3);
  PGAMD_FFT_CHECK(hipfftPlan3d(&c.fwd, M, M, M, HIPFFT_D2Z));
  PGAMD_FFT_CHECK(hipfftPlan3d(&c.inv, M, M, M, HIPFFT_Z2D));
  c.M = M;
}

}  // namespace

// Long-range TreePM accelerations of n particles in the periodic unit box:
// pos[n][3] in [0, 1) and mass[n] (device, FP64) on an M^3 mesh (M even,
// 8..1024), with the S2 split at eps_pm (the table's eps_PM, which is also
// its cut-off radius).  acc[n][3] (device, FP64) is written.  Add the
// short-range accelerations of the table (pg5t_batch_force or the g5 calls)
// for the whole TreePM force.  Asynchronous on `stream` (null = the library's).
PGAMD_API int pg5t_pm_force(int n, const double* pos, const double* mass, int mesh, double eps_pm,
                            double* acc, void* stream) {
  Board* b = board(0, "pg5t_pm_force");
  if (!b) return -1;
  if (n < 0 || mesh < 8 || mesh > 1024 || (mesh & 1) || !(eps_pm > 0.0) ||
      (n > 0 && (!pos || !mass || !acc))) {
    api_error(3, "pg5t_pm_force: n=%d mesh=%d eps_pm=%g (need an even mesh in [8, 1024], "
                 "eps_pm > 0, non-null arrays)", n, mesh, eps_pm);
    return -1;
  }
  std::lock_guard<std::mutex> lk(g_pm_mutex);
  PmCache& c = g_pm;
  pm_setup(c, mesh);
  hipStream_t st = stream ? (hipStream_t)stream : b->stream;
  PGAMD_FFT_CHECK(hipfftSetStream(c.fwd, st));
  PGAMD_FFT_CHECK(hipfftSetStream(c.inv, st));
  const size_t nr = (size_t)mesh * mesh * mesh, nk = (size_t)mesh * mesh * (mesh / 2 + 1);
  double* rho = c.rho.as<double>();
  double2* rhok = c.rhok.as<double2>();
  double2* fk = c.fk.as<double2>();
  double* f = c.f.as<double>();
  PGAMD_HIP_CHECK(hipMemsetAsync(rho, 0, sizeof(double) * nr, st));
  // particles sorted by the Morton code of their nearest mesh point
  const int* order = nullptr;
  if (n > 0) {
    c.keys.reserve(2 * sizeof(unsigned) * (size_t)n);
    c.order.reserve(2 * sizeof(int) * (size_t)n);
    unsigned* kin = c.keys.as<unsigned>();
    int* vin = c.order.as<int>();
    hipLaunchKernelGGL(pm::pm_keys, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, pos, mesh, kin,
                       vin);
    PGAMD_HIP_CHECK(hipGetLastError());
    // the keys interleave three coordinates of ceil(log2 mesh) bits each
    int b = 1;
    while ((1 << b) < mesh) ++b;
    const int end_bit = 3 * b;
    size_t tmp = 0;
    PGAMD_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, kin, kin + n, vin, vin + n,
                                                       n, 0, end_bit, st));
    c.sort_tmp.reserve(tmp);
    PGAMD_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(c.sort_tmp.ptr, tmp, kin, kin + n, vin,
                                                       vin + n, n, 0, end_bit, st));
    order = vin + n;
    hipLaunchKernelGGL(pm::pm_assign_sorted, dim3(ceil_div(n, pm::kAssignRun)),
                       dim3(pm::kAssignThreads), 0, st, n, order, pos, mass, mesh, rho);
    PGAMD_HIP_CHECK(hipGetLastError());
  }
  PGAMD_FFT_CHECK(hipfftExecD2Z(c.fwd, rho, reinterpret_cast<hipfftDoubleComplex*>(rhok)));
  hipLaunchKernelGGL(pm::pm_green, dim3((unsigned)((nk + 255) / 256)), dim3(256), 0, st, mesh,
                     eps_pm, rhok, fk);
  PGAMD_HIP_CHECK(hipGetLastError());
  for (int d = 0; d < 3; ++d)
    PGAMD_FFT_CHECK(hipfftExecZ2D(c.inv, reinterpret_cast<hipfftDoubleComplex*>(fk + d * nk),
                                  f + d * nr));
  if (n > 0)
    hipLaunchKernelGGL(pm::pm_interp, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, order, pos,
                       mesh, f, f + nr, f + 2 * nr, acc);
  PGAMD_HIP_CHECK(hipGetLastError());
  return 0;
}
