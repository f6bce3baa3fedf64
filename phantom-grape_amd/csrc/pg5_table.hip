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
  // Quadratic form of the same table for the batched calls (order 2): per
  // interval {c0, c1, c2, 0} with f = c0 + d (c1 + d c2), d = s - s_k.
  int order = 1;
  std::vector<float> quad;  // 4*ntab
  DevBuf dev_q;
  bool dirty_q = true;
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

// Learned cut-off (mlp_kernels.h): tanh-form parameters in FP64, the form of
// the target and the radii it was trained for.  form 1 (S2 shape): rc =
// eps_PM, r0 = eps_PP; form 0 (other shapes): rc = r_c, r0 = core length.
struct Mlp {
  std::vector<double> params;  // mlp::kParams
  bool valid = false;
  int form = 1;
  double rc = 0.0, r0 = 0.0;
  DevBuf dev;
  bool dirty = true;
};

Table g_tab;
Mlp g_mlp;
int g_cut_mode = 0;  // 0: table, 1: learned MLP
int g_batch_bpc[2] = {0, 0};  // persistent blocks per CU per mode (0 = default)
int g_batch_block = 0;        // table-mode block (= j-tile) size (0 = default 1024)
Board g_boards[kMaxDev];
std::mutex g_mutex;  // guards the table and board creation
int g_force_bi = 0;       // pg5t_dev_set_launch; 0 = the default of plan()
int g_blocks_per_cu = 0;
int g_min_jper = 0;

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
  // Quadratic through each interval's ends and midpoint, from the shape in
  // FP64 at the same s ticks (the last entry is constant, as its slope is 0).
  t.quad.assign(4 * (size_t)t.ntab, 0.f);
  auto fs = [&](double s) { return func(std::sqrt(std::max(0.0, (s - 2.0) / (double)r2scale))); };
  m.f = 2.0f;
  for (int i = 0; i < t.ntab; i++) {
    const double s0 = m.f;
    m.u += tick;
    const double s1 = m.f;
    const double f0 = fs(s0);
    if (i == t.ntab - 1) {
      t.quad[4 * (size_t)i] = (float)f0;
      break;
    }
    const double h = s1 - s0, fm = fs(s0 + 0.5 * h), f1 = fs(s1);
    const double c2 = 2.0 * (f1 - 2.0 * fm + f0) / (h * h);
    t.quad[4 * (size_t)i] = (float)f0;
    t.quad[4 * (size_t)i + 1] = (float)((f1 - f0) / h - c2 * h);
    t.quad[4 * (size_t)i + 2] = (float)c2;
  }
  t.dirty_q = true;
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
const float4* table_quad_dev(hipStream_t st) {
  Table& t = g_tab;
  if (t.dirty_q) {
    t.dev_q.reserve(sizeof(float) * t.quad.size());
    PGAMD_HIP_CHECK(hipMemcpyAsync(t.dev_q.ptr, t.quad.data(), sizeof(float) * t.quad.size(),
                                   hipMemcpyHostToDevice, st));
    PGAMD_HIP_CHECK(hipStreamSynchronize(st));
    t.dirty_q = false;
  }
  return t.dev_q.as<float4>();
}

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

// Default shape, measured on MI355X (clumpy box, all N x N pairs listed, %
// of the 38-flop FP32 peak against the earlier 256 / 128 / 64 i per block by
// ni, 8 blocks per CU, 128 j per range): 64 i per block up to 32768 i, 256
// beyond; 16 blocks per CU up to 16384 i, 32 beyond; at least 256 j per
// range.  N = 4096: 30 % (19 %); 16384: 75 % (73 %); 32768: 85 % (80 %);
// 65536: 88 % (82 %).
Plan plan(int ni, int nj) {
  Plan p;
  p.bi = g_force_bi > 0 ? g_force_bi : (ni > 32768 ? 256 : 64);
  const int bpc = g_blocks_per_cu > 0 ? g_blocks_per_cu : (ni > 16384 ? 32 : 16);
  const int min_jper = g_min_jper > 0 ? g_min_jper : 256;
  const int nbi = ceil_div(ni, p.bi);
  const int target = device_cu_count() * bpc;
  int ns = std::max(1, ceil_div(target, nbi));
  ns = std::min(ns, std::max(1, nj / min_jper));
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
    // the network is read while a force kernel runs on any stream
    PGAMD_HIP_CHECK(hipDeviceSynchronize());
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

// Decoding constants of the network for the current scale (mlp_kernels.h):
// S2 form f = [not core] 1/r^3 - H(r/rc)/rc^3 (+ H(r/r0)/r0^3 in the core),
// other shapes f = Y(sqrt(r/rc)) / (r^3 + r0^3); scaled units, physical f.
mlp::Cut mlp_cut(hipStream_t st) {
  mlp::Cut c;
  const double xs = g_tab.xscale, xs3 = xs * xs * xs;
  const double rc = g_mlp.rc * xs, r0 = g_mlp.r0 * xs;
  c.w = mlp_dev(st);
  c.inv_rc = (float)(1.0 / rc);
  c.rc2 = (float)(rc * rc);
  if (g_mlp.form == 1) {
    c.sqrt_in = 0;
    c.cn = (float)xs3;
    c.c1 = (float)(-xs3 / (rc * rc * rc));
    c.c2 = 1.f;
    c.c3 = 0.f;
    c.inv_pp = (float)(1.0 / r0);
    c.pp2 = (float)(r0 * r0);
    c.cpp = (float)(xs3 / (r0 * r0 * r0));
  } else {
    c.sqrt_in = 1;
    c.cn = 0.f;
    c.c1 = (float)xs3;
    c.c2 = (float)(r0 * r0 * r0);
    c.c3 = 1.f;
    c.inv_pp = 0.f;
    c.pp2 = 0.f;
    c.cpp = 0.f;
  }
  return c;
}

template <int BLOCK, bool QUAD>
void launch_table_batch(int blocks, size_t lds, hipStream_t st, int ngroups, const int* order,
                        const int* ioff, const int* icnt, const int* ilist, const int* joff,
                        const int* jcnt, const int* jlist, const float4* pos,
                        const pgb::Origin& og, const pgb::TableCut& tc, double ascale,
                        double* acc, int* counter) {
  hipLaunchKernelGGL((pgb::treepm_batch_table<BLOCK, QUAD>), dim3(blocks), dim3(BLOCK), lds, st,
                     ngroups, order, ioff, icnt, ilist, joff, jcnt, jlist, pos, og, tc, ascale,
                     acc, counter);
}

// Table-mode block sizes (each also the j-tile staged in LDS per step).
constexpr int kTableBlocks[3] = {256, 512, 1024};

bool valid_table_block(int block) {
  for (int b : kTableBlocks)
    if (b == block) return true;
  return false;
}

// Launches the persistent batch kernel (counter reset + kernel) on st.
// pos64 (nullable): FP64 positions for group-origin separations.  block:
// threads per block of the table kernel (256, 512 or 1024); the network
// kernel always runs 256.
void launch_batch(int mode, int ngroups, const int* order, const int* ioff, const int* icnt,
                  const int* ilist, const int* joff, const int* jcnt, const int* jlist,
                  const float4* pos, const double* pos64, double* acc, int* counter, int bpc,
                  int block, hipStream_t st) {
  hipLaunchKernelGGL(pgb::zero_counter, dim3(1), dim3(64), 0, st, counter);
  PGAMD_HIP_CHECK(hipGetLastError());
  if (ngroups <= 0) return;
  const int blocks = std::max(1, std::min(ngroups, device_cu_count() * bpc));
  const double ascale = (double)(float)(1.0 / g_tab.xscale);
  const pgb::Origin og{pos64, g_tab.xscale, (float)g_tab.xscale};
  pgb::TableCut tc{table_dev(st), table_quad_dev(st), g_tab.ntab, g_tab.frc_bit};
  if (mode == 1) {
    const mlp::Cut mc = mlp_cut(st);
    hipLaunchKernelGGL(pgb::treepm_batch_mlp, dim3(blocks), dim3(pgb::kThreads), 0, st, ngroups,
                       order, ioff, icnt, ilist, joff, jcnt, jlist, pos, og, mc, ascale, acc,
                       counter);
  } else {
    const bool quad = g_tab.order == 2;
    const size_t lds = (quad ? sizeof(float2) + sizeof(float) : sizeof(float2)) * (size_t)g_tab.ntab;
#define PGB_TABLE_LAUNCH(B, Q)                                                                 \
  launch_table_batch<B, Q>(blocks, lds, st, ngroups, order, ioff, icnt, ilist, joff, jcnt, jlist, \
                           pos, og, tc, ascale, acc, counter)
    switch (block) {
      case 256:
        if (quad) PGB_TABLE_LAUNCH(256, true); else PGB_TABLE_LAUNCH(256, false);
        break;
      case 512:
        if (quad) PGB_TABLE_LAUNCH(512, true); else PGB_TABLE_LAUNCH(512, false);
        break;
      case 1024:
        if (quad) PGB_TABLE_LAUNCH(1024, true); else PGB_TABLE_LAUNCH(1024, false);
        break;
      default:
        api_error(3, "treepm batch: block %d not in {256, 512, 1024}", block);
        return;
    }
#undef PGB_TABLE_LAUNCH
  }
  PGAMD_HIP_CHECK(hipGetLastError());
}

int default_bpc(int mode) {
  if (g_batch_bpc[mode] > 0) return g_batch_bpc[mode];
  return mode == 1 ? 4 : 2;
}

int default_block(int mode) {
  if (mode == 1) return pgb::kThreads;
  return g_batch_block > 0 ? g_batch_block : pgb::kTableBlock;
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
  launch_batch(1, ng, nullptr, ioff, icnt, ilist, joff, jcnt, jlist, pos, nullptr, accw, counter,
               default_bpc(1), default_block(1), st);
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
  if (g_tab.order == 2 && frc_bit > 8) {
    api_error(3, "pg5_set_table_frc_bit: %d with the quadratic table (at most 8)", frc_bit);
    return -1;
  }
  double (*f)(double) = g_tab.func ? g_tab.func : s2_force;
  const double rc = g_tab.func ? g_tab.rcut : g_tab.sft_pm;
  build_table(f, rc, frc_bit);
  return 0;
}

// Interpolation order of the table in the batched calls (pg5t_batch_*):
// 1 = the reference's value + slope per float-bit tick (default), 2 = a
// quadratic through each interval's ends and midpoint, evaluated from the
// shape in FP64 (16 bytes per entry in LDS: frc_bit <= 8).  The drop-in calls
// keep the reference's table.  No counterpart in the reference.
PGAMD_API int pg5_set_table_order(int order) {
  std::lock_guard<std::mutex> lk(g_mutex);
  if (order != 1 && order != 2) {
    api_error(3, "pg5_set_table_order: %d not in {1, 2}", order);
    return -1;
  }
  if (order == 2 && g_tab.frc_bit > 8) {
    api_error(3, "pg5_set_table_order: the quadratic table needs frc_bit <= 8 (now %d)",
              g_tab.frc_bit);
    return -1;
  }
  g_tab.order = order;
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

// Loads a network: tanh-form parameters p[321] in FP64 (layout in
// mlp_kernels.h) of the given form (1: S2 shape H(u) with rc = eps_PM and
// r0 = eps_PP; 0: Y(v) = f (r^3 + r0^3), v = sqrt(r / rc)); physical units.
PGAMD_API int pg5_mlp_set_params(const double* p, int form, double rc, double r0) {
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!p || (form != 0 && form != 1) || !(rc > 0.0) || !(r0 >= 0.0) ||
      (form == 1 && !(r0 > 0.0 && r0 < rc))) {
    api_error(3, "pg5_mlp_set_params: p=%p form=%d rc=%g r0=%g", (const void*)p, form, rc, r0);
    return -1;
  }
  g_mlp.params.assign(p, p + mlp::kParams);
  g_mlp.form = form;
  g_mlp.rc = rc;
  g_mlp.r0 = r0;
  g_mlp.valid = true;
  g_mlp.dirty = true;
  return 0;
}

// Copies the current network (each output may be null); returns 1 if one is
// loaded, 0 if not.
PGAMD_API int pg5_mlp_get_params(double* p, int* form, double* rc, double* r0) {
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!g_mlp.valid) return 0;
  if (p) std::memcpy(p, g_mlp.params.data(), sizeof(double) * mlp::kParams);
  if (form) *form = g_mlp.form;
  if (rc) *rc = g_mlp.rc;
  if (r0) *r0 = g_mlp.r0;
  return 1;
}

namespace {

// Training set and work arrays of the FP64 fit (device).
struct Fit {
  int n = 0;
  DevBuf v, t, sw, p, jac, m;
  void alloc(int ns) {
    n = ns;
    v.reserve(sizeof(double) * ns);
    t.reserve(sizeof(double) * ns);
    sw.reserve(sizeof(double) * ns);
    p.reserve(sizeof(double) * mlp::kParams);
    jac.reserve(sizeof(double) * (size_t)mlp::kExt * ns);
    m.reserve(sizeof(double) * mlp::kExt * mlp::kExt);
  }
  void upload(const double* hv, const double* ht, const double* hsw, hipStream_t st) {
    PGAMD_HIP_CHECK(hipMemcpyAsync(v.ptr, hv, sizeof(double) * n, hipMemcpyHostToDevice, st));
    PGAMD_HIP_CHECK(hipMemcpyAsync(t.ptr, ht, sizeof(double) * n, hipMemcpyHostToDevice, st));
    PGAMD_HIP_CHECK(hipMemcpyAsync(sw.ptr, hsw, sizeof(double) * n, hipMemcpyHostToDevice, st));
  }
  void release() {
    for (DevBuf* d : {&v, &t, &sw, &p, &jac, &m}) d->release();
  }
};

// M = [J | r]^T [J | r] at parameters x (full symmetric, kExt x kExt, host):
// J^T J, J^T r and r^T r of the weighted residuals sqrt(wt) (y - t).
void fit_normal(Fit& f, const double* x, std::vector<double>& M, hipStream_t st) {
  constexpr int E = mlp::kExt;
  PGAMD_HIP_CHECK(hipMemcpyAsync(f.p.ptr, x, sizeof(double) * mlp::kParams,
                                 hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(mlp::mlp_fwd_jac64, dim3(ceil_div(f.n, 64)), dim3(64), 0, st, f.n,
                     f.v.as<double>(), f.t.as<double>(), f.sw.as<double>(), f.p.as<double>(),
                     f.jac.as<double>());
  PGAMD_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(mlp::mlp_jtj64, dim3(mlp::kTiles, mlp::kTiles), dim3(256), 0, st, f.n,
                     f.jac.as<double>(), f.m.as<double>());
  PGAMD_HIP_CHECK(hipGetLastError());
  M.resize((size_t)E * E);
  PGAMD_HIP_CHECK(hipMemcpyAsync(M.data(), f.m.ptr, sizeof(double) * E * E,
                                 hipMemcpyDeviceToHost, st));
  PGAMD_HIP_CHECK(hipStreamSynchronize(st));
  for (int a = 0; a < E; ++a)
    for (int b = 0; b < a; ++b) M[(size_t)a * E + b] = M[(size_t)b * E + a];
}

// Solves A x = b in place (b <- x) by Cholesky; A (P x P, row-major) is
// overwritten.  Returns false if A is not positive definite.
bool chol_solve(std::vector<double>& A, std::vector<double>& b, int P) {
  for (int j = 0; j < P; ++j) {
    double d = A[(size_t)j * P + j];
    for (int k = 0; k < j; ++k) d -= A[(size_t)j * P + k] * A[(size_t)j * P + k];
    if (!(d > 0.0)) return false;
    const double l = std::sqrt(d);
    A[(size_t)j * P + j] = l;
    for (int i = j + 1; i < P; ++i) {
      double s = A[(size_t)i * P + j];
      for (int k = 0; k < j; ++k) s -= A[(size_t)i * P + k] * A[(size_t)j * P + k];
      A[(size_t)i * P + j] = s / l;
    }
  }
  for (int i = 0; i < P; ++i) {
    double s = b[i];
    for (int k = 0; k < i; ++k) s -= A[(size_t)i * P + k] * b[k];
    b[i] = s / A[(size_t)i * P + i];
  }
  for (int i = P - 1; i >= 0; --i) {
    double s = b[i];
    for (int k = i + 1; k < P; ++k) s -= A[(size_t)k * P + i] * b[k];
    b[i] = s / A[(size_t)i * P + i];
  }
  return true;
}

// Levenberg-Marquardt state (Marquardt's diagonal scaling): parameters x,
// their normal equations M (kExt x kExt, the last column J^T r, the corner
// r^T r), the loss F = r^T r and the damping lam.
struct Lm {
  std::vector<double> x, M;
  double F = HUGE_VAL;
  double lam = 1e-3;
};

void lm_start(Fit& f, Lm& s, hipStream_t st) {
  fit_normal(f, s.x.data(), s.M, st);
  s.F = s.M[(size_t)mlp::kParams * mlp::kExt + mlp::kParams];
  s.lam = 1e-3;
}

// At most `iters` trial steps from s (fewer once the damping exceeds 1e12).
void lm_steps(Fit& f, Lm& s, int iters, hipStream_t st) {
  constexpr int P = mlp::kParams, E = mlp::kExt;
  std::vector<double> Mn, A((size_t)P * P), d(P), xn(P);
  for (int it = 0; it < iters && s.lam < 1e12; ++it) {
    for (int a = 0; a < P; ++a) {
      for (int b = 0; b < P; ++b) A[(size_t)a * P + b] = s.M[(size_t)a * E + b];
      A[(size_t)a * P + a] += s.lam * s.M[(size_t)a * E + a] + 1e-300;
      d[a] = -s.M[(size_t)a * E + P];
    }
    if (!chol_solve(A, d, P)) {
      s.lam *= 4.0;
      continue;
    }
    for (int a = 0; a < P; ++a) xn[a] = s.x[a] + d[a];
    fit_normal(f, xn.data(), Mn, st);
    const double Fn = Mn[(size_t)P * E + P];
    if (std::isfinite(Fn) && Fn < s.F) {
      s.x.swap(xn);
      s.M.swap(Mn);
      s.F = Fn;
      s.lam = std::max(s.lam / 3.0, 1e-15);
    } else {
      s.lam *= 4.0;
    }
  }
}

// Seeded initial parameters (tanh form): spread first-layer units, Glorot
// second layer and output, output bias = mean target.
void mlp_init(std::mt19937_64& rng, double tmean, std::vector<double>& p) {
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  p.assign(mlp::kParams, 0.0);
  for (int j = 0; j < mlp::kH; ++j) {
    p[mlp::kW1 + j] = 4.0 * U(rng);
    p[mlp::kB1 + j] = 2.0 * U(rng);
  }
  const double lim = std::sqrt(6.0 / (2 * mlp::kH));
  for (int k = 0; k < mlp::kH * mlp::kH; ++k) p[mlp::kW2 + k] = lim * U(rng);
  for (int k = 0; k < mlp::kH; ++k) p[mlp::kW3 + k] = lim * U(rng);
  p[mlp::kB3] = tmean;
}

// The S2 shape in units of a^-3 at u = r / a (pg5_table.c:8-22 with a = 1).
double s2_unit(double u) { return s2_reff(1.0, u); }

}  // namespace

// Normal equations of the weighted least-squares fit of the network at
// parameters p[321] (tanh form, FP64) to n host samples (v, t, wt), computed
// by the FP64 training kernels: jtj[321*321] (nullable) = J^T J, jtr[321] =
// J^T r and *rtr = r^T r, with r_s = sqrt(wt_s) (y(v_s) - t_s) and J = dr/dp.
// The gradient of the loss sum_s wt_s (y_s - t_s)^2 / n is 2 jtr / n.
// Synchronous.  For checking the kernels against a CPU restatement.
PGAMD_API int pg5_mlp_normal_eq(int n, const double* v, const double* t, const double* wt,
                                const double* p, double* jtj, double* jtr, double* rtr) {
  Board* b = board(0, "pg5_mlp_normal_eq");
  if (!b) return -1;
  if (n <= 0 || !v || !t || !wt || !p || !jtr || !rtr) {
    api_error(3, "pg5_mlp_normal_eq: n=%d (need n > 0 and non-null arrays)", n);
    return -1;
  }
  constexpr int P = mlp::kParams, E = mlp::kExt;
  std::vector<double> sw(n);
  for (int s = 0; s < n; ++s) sw[s] = std::sqrt(wt[s]);
  Fit f;
  f.alloc(n);
  f.upload(v, t, sw.data(), b->stream);
  std::vector<double> M;
  fit_normal(f, p, M, b->stream);
  f.release();
  for (int a = 0; a < P; ++a) {
    jtr[a] = M[(size_t)a * E + P];
    if (jtj)
      for (int c = 0; c < P; ++c) jtj[(size_t)a * P + c] = M[(size_t)a * E + c];
  }
  *rtr = M[(size_t)P * E + P];
  return 0;
}

// Trains the network on the GPU for the current cut-off shape (the function
// the table was generated from).  S2 shape (pg5_gen_s2_force_table): the
// network learns H(u) = a^3 R(u a, a) of pg5_table.c:8-22 on 8192 stratified
// samples of u in [0, 1] plus both ends, residuals weighted by u^3 + 0.2.  Other shapes
// (pg5_gen_force_table): Y(v) = f(r) (r^3 + r0^3), r0 = r_c / 30, on 8192
// seeded samples (half uniform in v = sqrt(r / r_c), half uniform in volume),
// weights 1 / (|Y| + 0.05)^2.  Targets are evaluated in FP64.  `starts`
// seeded initialisations each run `iters` Levenberg-Marquardt trial steps
// (FP64 residuals, Jacobians and normal equations from the GPU kernels, the
// damped 321 x 321 solve on the host); the one with the smallest loss then
// runs `refine` more (many starts stall on a plateau for hundreds of steps;
// the best one after a short run is the one that converges).  Deterministic
// for a given seed.  *loss_out = sum wt (y - t)^2 / n.
PGAMD_API int pg5_mlp_train(int iters, int starts, int refine, unsigned seed,
                            double* loss_out) {
  Board* b = board(0, "pg5_mlp_train");
  if (!b) return -1;
  if (iters <= 0 || starts <= 0 || starts > 64 || refine < 0) {
    api_error(3,
              "pg5_mlp_train: iters=%d starts=%d refine=%d (need iters > 0, 0 < starts <= 64, "
              "refine >= 0)",
              iters, starts, refine);
    return -1;
  }
  std::lock_guard<std::mutex> lk(g_mutex);
  ensure_table();
  const bool s2 = g_tab.is_s2;
  const int n = 8192 + 2;
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  std::vector<double> hv(n), ht(n), hsw(n);
  double rc, r0;
  if (s2) {
    rc = g_tab.sft_pm;
    r0 = g_tab.sft_pp;
    // A pair outside the core carries the network's error times u^3 relative
    // to its Newtonian force; inside the core (rare pairs) times 1/H:
    // weight the residuals by u^3 + 0.2.
    for (int k = 0; k < n - 2; ++k) {
      hv[k] = (k + U(rng)) / (n - 2);
      ht[k] = s2_unit(hv[k]);
      hsw[k] = hv[k] * hv[k] * hv[k] + 0.2;
    }
    for (int e = 0; e < 2; ++e) {
      hv[n - 2 + e] = e;
      ht[n - 2 + e] = s2_unit(e);
      hsw[n - 2 + e] = e + 0.2;
    }
  } else {
    double (*func)(double) = g_tab.func;
    rc = g_tab.rcut;
    r0 = g_tab.rcut / 30.0;
    const double r03 = r0 * r0 * r0;
    for (int k = 0; k < n; ++k) {
      double r, v;
      if (k & 1) {
        r = rc * std::cbrt(U(rng));
        v = std::sqrt(r / rc);
      } else {
        v = U(rng);
        r = rc * v * v;
      }
      const double y = func(r) * (r * r * r + r03);
      hv[k] = v;
      ht[k] = y;
      hsw[k] = 1.0 / (std::fabs(y) + 0.05);
    }
  }
  double tmean = 0.0;
  for (double y : ht) tmean += y;
  tmean /= n;
  Fit f;
  f.alloc(n);
  f.upload(hv.data(), ht.data(), hsw.data(), b->stream);
  Lm best;
  for (int s = 0; s < starts; ++s) {
    Lm c;
    mlp_init(rng, tmean, c.x);
    lm_start(f, c, b->stream);
    lm_steps(f, c, iters, b->stream);
    if (c.F < best.F) best = std::move(c);
  }
  if (best.x.empty()) {  // every start's loss was not finite
    f.release();
    api_error(6, "pg5_mlp_train: no start reached a finite loss");
    return -1;
  }
  lm_steps(f, best, refine, b->stream);
  f.release();
  g_mlp.params = best.x;
  g_mlp.form = s2 ? 1 : 0;
  g_mlp.rc = rc;
  g_mlp.r0 = r0;
  g_mlp.valid = true;
  g_mlp.dirty = true;
  if (loss_out) *loss_out = best.F / n;
  return 0;
}

// Learned f(r) at n physical radii (device arrays), decoded as the force
// kernels do (0 at r = 0 and beyond the cut-off).  Synchronous.
PGAMD_API int pg5_mlp_eval_dev(int n, const float* r, float* f) {
  Board* b = board(0, "pg5_mlp_eval_dev");
  if (!b) return -1;
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!g_mlp.valid) {
    api_error(5, "pg5_mlp_eval_dev: no learned cut-off");
    return -1;
  }
  const mlp::Cut c = mlp_cut(b->stream);
  if (n > 0) {
    hipLaunchKernelGGL(mlp::mlp_eval_f, dim3(ceil_div(n, 256)), dim3(256), 0, b->stream, n, r,
                       c, (float)g_tab.xscale, f);
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
// pos[np] float4 scaled {x*xs, y*xs, z*xs, m}; pos64[np][3] (nullable) the
// same positions unscaled in FP64, for separations formed relative to each
// group's origin; group g: i-particles ilist[ioff[g] .. +icnt[g]),
// j-particles jlist[joff[g] .. +jcnt[g]); order (nullable) = processing
// order of the groups; acc[np][3] FP64 written for the listed i.  mode 0 =
// table, 1 = learned MLP.  All device pointers; asynchronous on stream.
// Calls on one devid share its work counter.
PGAMD_API int pg5t_batch_force(int devid, int ngroups, const int* ioff, const int* icnt,
                               const int* ilist, const int* joff, const int* jcnt,
                               const int* jlist, const int* order, const void* pos,
                               const double* pos64, double* acc, int mode, void* stream) {
  Board* b = board(devid, "pg5t_batch_force");
  if (!b) return -1;
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!batch_check("pg5t_batch_force", mode, ngroups)) return -1;
  ensure_table();
  hipStream_t st = stream ? (hipStream_t)stream : b->stream;
  b->d_counter.reserve(256);
  launch_batch(mode, ngroups, order, ioff, icnt, ilist, joff, jcnt, jlist,
               static_cast<const float4*>(pos), pos64, acc, b->d_counter.as<int>(),
               default_bpc(mode), default_block(mode), st);
  return 0;
}

// Captures the same launch (counter reset + persistent kernel) into a HIP
// graph with its own work counter; replay with pg5t_batch_graph_launch.
// Inputs are read at replay time from the same device addresses.
PGAMD_API void* pg5t_batch_graph(int devid, int ngroups, const int* ioff, const int* icnt,
                                 const int* ilist, const int* joff, const int* jcnt,
                                 const int* jlist, const int* order, const void* pos,
                                 const double* pos64, double* acc, int mode) {
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
               static_cast<const float4*>(pos), pos64, acc, bg->counter.as<int>(),
               default_bpc(mode), default_block(mode), cs);
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

// Autotuner: times the batch launch on these inputs for every block size
// (table mode: 256, 512 and 1024 threads, each also the j-tile staged in
// LDS; network mode: 256) and persistent-grid size of {1, 2, 3, 4, 6, 8, 12,
// 16} blocks per CU (one warm-up, then the mean of `reps`), keeps the fastest
// pair for this mode and returns its blocks per CU through *best; ms[8]
// (nullable) receives the timings at the block size kept, which
// pg5t_batch_launch_shape reports.  Results depend on the block size only
// through the order of the FP32 partial sums.
PGAMD_API int pg5t_batch_autotune(int devid, int ngroups, const int* ioff, const int* icnt,
                                  const int* ilist, const int* joff, const int* jcnt,
                                  const int* jlist, const int* order, const void* pos,
                                  const double* pos64, double* acc, int mode, int reps,
                                  int* best, float* ms) {
  Board* b = board(devid, "pg5t_batch_autotune");
  if (!b) return -1;
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!batch_check("pg5t_batch_autotune", mode, ngroups)) return -1;
  ensure_table();
  const int cand[8] = {1, 2, 3, 4, 6, 8, 12, 16};
  const int nblk = mode == 1 ? 1 : 3;
  hipStream_t st = b->stream;
  b->d_counter.reserve(256);
  hipEvent_t e0, e1;
  PGAMD_HIP_CHECK(hipEventCreate(&e0));
  PGAMD_HIP_CHECK(hipEventCreate(&e1));
  float bestt = 1e30f, t_all[3][8];
  int bestc = default_bpc(mode), bestb = default_block(mode);
  reps = std::max(1, reps);
  for (int kb = 0; kb < nblk; ++kb) {
    const int blk = mode == 1 ? pgb::kThreads : kTableBlocks[kb];
    for (int k = 0; k < 8; ++k) {
      auto go = [&] {
        launch_batch(mode, ngroups, order, ioff, icnt, ilist, joff, jcnt, jlist,
                     static_cast<const float4*>(pos), pos64, acc, b->d_counter.as<int>(), cand[k],
                     blk, st);
      };
      go();
      PGAMD_HIP_CHECK(hipEventRecord(e0, st));
      for (int r = 0; r < reps; ++r) go();
      PGAMD_HIP_CHECK(hipEventRecord(e1, st));
      PGAMD_HIP_CHECK(hipEventSynchronize(e1));
      float t = 0.f;
      PGAMD_HIP_CHECK(hipEventElapsedTime(&t, e0, e1));
      t /= reps;
      t_all[kb][k] = t;
      if (t < bestt) {
        bestt = t;
        bestc = cand[k];
        bestb = blk;
      }
    }
  }
  PGAMD_HIP_CHECK(hipEventDestroy(e0));
  PGAMD_HIP_CHECK(hipEventDestroy(e1));
  g_batch_bpc[mode] = bestc;
  if (mode == 0) g_batch_block = bestb;
  if (ms) {
    int kb = 0;
    while (mode == 0 && kTableBlocks[kb] != bestb) ++kb;
    for (int k = 0; k < 8; ++k) ms[k] = t_all[kb][k];
  }
  if (best) *best = bestc;
  return 0;
}

// Table-mode block size for the batched calls: 256, 512 or 1024 threads
// (each also the j-tile staged in LDS per step), 0 = default (1024).
PGAMD_API int pg5t_batch_set_block(int block) {
  if (block != 0 && !valid_table_block(block)) {
    api_error(3, "pg5t_batch_set_block: block %d not in {0, 256, 512, 1024}", block);
    return -1;
  }
  g_batch_block = block;
  return 0;
}

// Launch shape the batched calls of this mode use now: out2 = {threads per
// block, persistent blocks per CU}.
PGAMD_API int pg5t_batch_launch_shape(int mode, int* out2) {
  if ((mode != 0 && mode != 1) || !out2) {
    api_error(3, "pg5t_batch_launch_shape: mode=%d out=%p", mode, (void*)out2);
    return -1;
  }
  out2[0] = default_block(mode);
  out2[1] = default_bpc(mode);
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
  DevBuf fix;  // pm_mass_part's block sums, then the fixed-point scale
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
    c.fix.reserve(sizeof(double) * (pm::kMassBlocks + 1));
    double* part = c.fix.as<double>();
    double* scale = part + pm::kMassBlocks;
    hipLaunchKernelGGL(pm::pm_mass_part, dim3(pm::kMassBlocks), dim3(256), 0, st, n, mass, part);
    hipLaunchKernelGGL(pm::pm_fix_scale, dim3(1), dim3(64), 0, st, part, mesh, scale);
    hipLaunchKernelGGL(pm::pm_assign_sorted, dim3(ceil_div(n, pm::kAssignRun)),
                       dim3(pm::kAssignThreads), 0, st, n, order, pos, mass, mesh, scale,
                       reinterpret_cast<unsigned long long*>(rho));
    hipLaunchKernelGGL(pm::pm_fix_to_double, dim3((unsigned)((nr + 255) / 256)), dim3(256), 0,
                       st, nr, scale, rho);
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
