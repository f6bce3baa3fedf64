// libpg6.so: GRAPE-6 API (PG6/libpg6/gp6util.h, phantom_g6.c:18-566) on
// MI355X, plus device-pointer entry points (pg6_dev_*) that work on the same
// board context without host round trips.
//
// Board model.  The reference keeps one global context (PG6/libpg6/gravity.c:
// 26-47): a j-particle memory written by g6_set_j_particle, a predicted copy
// rebuilt by g6calc_firsthalf when g6_set_ti has been called, and i-particles
// handed to g6calc_lasthalf*.  Here the j-memory lives in HBM as SoA arrays;
// g6_set_j_particle queues 120-byte records in pinned host memory (a later
// write to the same address replaces the earlier one), firsthalf uploads the
// queue in one copy and scatters it with a kernel, then predicts.  lasthalf*
// uploads the i-set in one copy, runs the force kernel (split over j-ranges
// when ni is small so every CU has work) and returns the results in one copy.
// The gpid argument is ignored, as in the reference.
#include "pgamd_common.h"
#include "pgamd_api.h"
#include "pg6_kernels.h"
#include "host_upload.h"
#include "pg6_nblist.h"
#include "cluster_batch.h"
#include "hermite_kernels.h"

#include <hipcub/device/device_radix_sort.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

using namespace pgamd;
using pg6::JPred;
using pg6::JRecord;
using pg6::JState;
using pg6::Outputs;
using pg6::Partials;

namespace {

constexpr int kDefaultPipes = 48;        // g6_npipes() of the reference (avx_type.h:1)
constexpr int kMaxLen = 1024;            // neighbour list length (avx_type.h:2)
constexpr int kJMax = 1 << 26;           // j-memory limit
constexpr int kInitialJCap = 1 << 18;    // reference JMEMSIZE (gravity.c:14)

struct LaunchPlan {
  int bi;       // i-particles per block
  int nsplit;   // j-ranges
  int jper;     // j-particles per range
  int variant;  // 0: j-tiles in LDS (pg6_force), 1: j broadcast in SGPRs (pg6_force_s),
                // 2: as 1 with two-float positions and the self filter (pg6_force_h),
                // 3: one launch with the predictor and the reduction inside (pg6_force_f),
                // 4: the sorted layout with chunk origins (pg6_force_o)
  int bpc;      // target blocks (of 256 threads) per CU
};

// Launch choice for one i-count bucket; bucket k holds ni <= 64 << k (the
// last bucket also every larger ni).
struct LaunchChoice {
  int variant, bi, bpc;
};
constexpr int kBuckets = 11;
constexpr int kFusedMaxNi = 1024;  // pg6_force_f (variant 3) up to this many i
// pg6_force_o (variant 4) is chosen automatically only from this many j up:
// its positions are FP32 offsets from the origin of their 64-particle chunk,
// so its dx error grows with the chunk extent, which is large where the
// j-set is sparse (p99 relative force error about 1e-6 on a 4096-particle
// Plummer sphere, 1.5e-7 at 65536).
constexpr int kSortedMinNj = 16384;

inline int ni_bucket(int ni) {
  int k = 0;
  while (k < kBuckets - 1 && ni > (64 << k)) ++k;
  return k;
}

// Default choice, from the autotuner's picks at nj = 65536 on MI355X: the
// LDS-tile kernel for small and medium i-sets, the sorted chunk-origin kernel
// for large ones (the two-float kernel below kSortedMinNj j; for the largest,
// four wavefronts per 64-i block sharing each j-range, 4x fewer partial sums
// than one wavefront per range), and many j-ranges so that the last wave of
// blocks is a small fraction of the launch.  Up to 128 i, one block per CU
// (one i-block of 64 or 128): the call is latency bound, and more j-ranges
// only lengthen the sum over them (64 i: 19.9 us per predict + force call
// against 24.1 us with twice the ranges).
inline LaunchChoice default_choice(int ni) {
  if (ni <= 64) return {0, 64, 1};
  if (ni <= 128) return {0, 128, 1};
  if (ni <= 2048) return {0, 64, 4};
  if (ni <= 8192) return {4, 256, 16};
  if (ni <= 32768) return {4, 64, 32};
  return {5, 256, 64};
}

struct Board {
  bool open = false;
  hipStream_t stream = nullptr;
  int npipes = kDefaultPipes;
  int cap = 0;  // j capacity
  DevBuf js_buf, jp_buf;
  JState js{};
  JPred jp{};
  // queued j records
  HostBuf pend;
  int npend = 0;
  std::vector<int> slot;  // address -> queue slot, -1 if none
  // recorded after the last flush's scatter: the queue is written again
  // only once it is done
  hipEvent_t pend_ev = nullptr;
  bool pend_busy = false;
  // Cross-stream order of the board's work (board_use_stream): the stream of
  // the last call that touched the j-memory or the partial-sum workspace, and
  // an event that marks the end of that call's work.
  hipStream_t last_st = nullptr;
  hipEvent_t order_ev = nullptr;
  double ti = 0.0;
  // predict_flag: the predicted j-set is stale (g6_set_ti since the last
  // prediction); the prediction runs with the first force that needs it (in
  // the force kernel itself for pg6_force_f).  pred_armed: a firsthalf (or
  // pg6_dev_predict) asked for it at pred_nj j-particles, so j-records flushed
  // before that force must not enter it (the reference predicts in firsthalf,
  // PG6/libpg6/phantom_g6.c:153-164): the flush predicts first.
  int predict_flag = 0;
  bool pred_armed = false;
  int pred_nj = 0;
  // pred_virtual: the last force predicted the first virt_nj j-particles to
  // ti inside its own launch and wrote no predicted set (predict_flag stays
  // set, so later calls at ti predict the same j-state again).  j-records
  // flushed before the next g6_set_ti must not enter this prediction, so the
  // flush writes it out first (board_predict); a new ti drops it.
  bool pred_virtual = false;
  int virt_nj = 0;
  int nj_pred = -1;  // j count of the last prediction
  // i-set and outputs of the host API
  HostBuf h_in, h_out;
  DevBuf d_in, d_part;
  DevBuf d_fcount;  // pg6_force_f's per-i-block completion counters (kept at zero)
  // Sorted j layout of pg6_force_o (pg6::JSort): built from the predicted
  // positions when missing or stale (kSortAge), its origins, flags and
  // records rewritten by every prediction made while it is in use (srec_ok:
  // they hold the current prediction).
  DevBuf d_sort, d_sort_tmp;
  pg6::JSort so{};
  int sort_nj = 0, sort_age = 0;
  long long sort_writes = 0;  // j-records written since the build
  bool srec_ok = false;
  // neighbour lists of the last lasthalf*n call
  DevBuf d_nb;
  std::vector<int> nbl, nblen;
  int nb_ni = 0;
  int nblerror = 0;
  // launch tuning: explicit overrides (0 / -1 = automatic), then the
  // autotuned choice of the ni bucket, then the default
  int force_bi = 0;
  int force_bpc = 0;
  int force_variant = -1;
  int min_jper = 128;
  LaunchChoice tuned[kBuckets] = {};
  bool tuned_ok[kBuckets] = {};
};

Board g_board;
std::mutex g_mutex;

// Optional per-launch timing of the force kernel: a HIP event pair recorded on
// the launch stream immediately around each pg6_force launch.
struct KernelTimer {
  bool on = false;
  std::vector<hipEvent_t> ev;  // 2 per launch
  int used = 0;
  void start(int pairs) {
    for (hipEvent_t e : ev) PGAMD_HIP_CHECK(hipEventDestroy(e));
    ev.assign(2 * (size_t)pairs, nullptr);
    for (hipEvent_t& e : ev) PGAMD_HIP_CHECK(hipEventCreate(&e));
    used = 0;
    on = true;
  }
  void before(hipStream_t st) {
    if (on && 2 * used + 1 < (int)ev.size()) PGAMD_HIP_CHECK(hipEventRecord(ev[2 * used], st));
  }
  void after(hipStream_t st) {
    if (on && 2 * used + 1 < (int)ev.size()) {
      PGAMD_HIP_CHECK(hipEventRecord(ev[2 * used + 1], st));
      ++used;
    }
  }
};
KernelTimer g_timer;

// Every call that reads or writes the j-memory or the partial-sum workspace
// runs on the device after the previous such call, whichever streams the two
// were queued on: board_use_stream(st) before queueing work on st makes st
// wait for the last call's work if that went to another stream, and
// board_done_on(st) after it marks its end.  Work on the board's own stream
// is marked lazily, when another stream comes next.
void board_use_stream(Board& b, hipStream_t st) {
  if (st == b.last_st) return;
  if (b.last_st == b.stream) PGAMD_HIP_CHECK(hipEventRecord(b.order_ev, b.stream));
  PGAMD_HIP_CHECK(hipStreamWaitEvent(st, b.order_ev, 0));
  b.last_st = st;
}

void board_done_on(Board& b, hipStream_t st) {
  if (st != b.stream) PGAMD_HIP_CHECK(hipEventRecord(b.order_ev, st));
}

void board_alloc_j(Board& b, int cap) {
  // Grow j-state and prediction arrays, preserving the j-state contents.
  if (b.cap > 0) board_use_stream(b, b.stream);
  DevBuf njs, njp;
  const size_t jsb = (size_t)cap * (14 * sizeof(double) + sizeof(int));
  const size_t jpb = (size_t)cap * (3 * sizeof(double) + 4 * sizeof(float4) + sizeof(int)) +
                     (size_t)(cap / pg6::kChunk + 1) * sizeof(int2);
  njs.reserve(jsb);
  njp.reserve(jpb);
  PGAMD_HIP_CHECK(hipMemsetAsync(njs.ptr, 0, jsb, b.stream));
  PGAMD_HIP_CHECK(hipMemsetAsync(njp.ptr, 0, jpb, b.stream));
  JState js;
  double* d = njs.as<double>();
  double** f[14] = {&js.x, &js.y, &js.z, &js.vx, &js.vy, &js.vz, &js.ax,
                    &js.ay, &js.az, &js.kx, &js.ky, &js.kz, &js.m, &js.t};
  for (int k = 0; k < 14; ++k) *f[k] = d + (size_t)k * cap;
  js.idx = reinterpret_cast<int*>(d + (size_t)14 * cap);
  JPred jp;
  double* e = njp.as<double>();
  jp.x = e;
  jp.y = e + cap;
  jp.z = e + 2 * (size_t)cap;
  jp.vm = reinterpret_cast<float4*>(e + 3 * (size_t)cap);
  jp.rec = jp.vm + cap;
  jp.crange = reinterpret_cast<int2*>(jp.rec + 3 * (size_t)cap);
  jp.id = reinterpret_cast<int*>(jp.crange + cap / pg6::kChunk + 1);
  if (b.cap > 0) {
    double** o[14] = {&b.js.x, &b.js.y, &b.js.z, &b.js.vx, &b.js.vy, &b.js.vz, &b.js.ax,
                      &b.js.ay, &b.js.az, &b.js.kx, &b.js.ky, &b.js.kz, &b.js.m, &b.js.t};
    for (int k = 0; k < 14; ++k)
      PGAMD_HIP_CHECK(hipMemcpyAsync(*f[k], *o[k], (size_t)b.cap * sizeof(double),
                                     hipMemcpyDeviceToDevice, b.stream));
    PGAMD_HIP_CHECK(hipMemcpyAsync(js.idx, b.js.idx, (size_t)b.cap * sizeof(int),
                                   hipMemcpyDeviceToDevice, b.stream));
    PGAMD_HIP_CHECK(hipStreamSynchronize(b.stream));
  }
  b.js_buf.release();
  b.jp_buf.release();
  b.js_buf = njs;
  b.jp_buf = njp;
  b.js = js;
  b.jp = jp;
  b.cap = cap;
  b.slot.assign((size_t)cap, -1);
  b.npend = 0;
  b.nj_pred = -1;
}

// Requires b.open.  Returns false (after reporting) if addr is out of range.
bool board_reserve(Board& b, int need) {
  if (need <= b.cap) return true;
  if (need > kJMax) {
    api_error(2, "j-particle address %d exceeds the j-memory limit %d", need - 1, kJMax);
    return false;
  }
  // Flush queued records before the arrays move.
  int cap = b.cap > 0 ? b.cap : kInitialJCap;
  while (cap < need) cap *= 2;
  if (cap > kJMax) cap = kJMax;
  std::vector<JRecord> keep;
  if (b.npend > 0) keep.assign(b.pend.as<JRecord>(), b.pend.as<JRecord>() + b.npend);
  board_alloc_j(b, cap);
  for (const JRecord& r : keep) {
    b.slot[r.addr] = b.npend;
    b.pend.as<JRecord>()[b.npend++] = r;
  }
  return true;
}

__global__ void open_warm(int* flag) {
  if (flag) *flag = 0;
}

void board_open(Board& b) {
  if (b.open) return;
  ensure_device();
  PGAMD_HIP_CHECK(hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking));
  if (const char* s = std::getenv("PGAMD_G6_NPIPES")) {
    int v = std::atoi(s);
    if (v > 0) b.npipes = v;
  }
  PGAMD_HIP_CHECK(hipEventCreateWithFlags(&b.pend_ev, hipEventDisableTiming));
  PGAMD_HIP_CHECK(hipEventCreateWithFlags(&b.order_ev, hipEventDisableTiming));
  b.last_st = b.stream;
  b.open = true;
  board_alloc_j(b, kInitialJCap);
  b.pend.reserve(sizeof(JRecord) * 1024);
  // one empty launch loads the library's kernels now, not in the first call
  hipLaunchKernelGGL(open_warm, dim3(1), dim3(64), 0, b.stream, nullptr);
  PGAMD_HIP_CHECK(hipGetLastError());
  PGAMD_HIP_CHECK(hipStreamSynchronize(b.stream));
}

void board_close(Board& b) {
  if (!b.open) return;
  board_use_stream(b, b.stream);
  PGAMD_HIP_CHECK(hipStreamSynchronize(b.stream));
  PGAMD_HIP_CHECK(hipEventDestroy(b.pend_ev));
  PGAMD_HIP_CHECK(hipEventDestroy(b.order_ev));
  b.js_buf.release();
  b.jp_buf.release();
  b.d_in.release();
  b.d_part.release();
  b.d_fcount.release();
  b.d_sort.release();
  b.d_sort_tmp.release();
  b.d_nb.release();
  b.pend.release();
  b.h_in.release();
  b.h_out.release();
  PGAMD_HIP_CHECK(hipStreamDestroy(b.stream));
  b = Board();
}

bool board_check_open(const char* fn) {
  if (!g_board.open) {
    api_error(1, "%s called before g6_open", fn);
    return false;
  }
  return true;
}

void board_queue_j(Board& b, int padr, int pidx, double t, double m, const double* acc2,
                   const double* jrk6, const double* vel, const double* pos) {
  if (b.pend_busy) {  // the previous flush may still be reading the queue
    PGAMD_HIP_CHECK(hipEventSynchronize(b.pend_ev));
    b.pend_busy = false;
  }
  int s = b.slot[padr];
  if (s < 0) {
    b.pend.reserve(sizeof(JRecord) * (size_t)(b.npend + 1), true);
    s = b.npend++;
    b.slot[padr] = s;
  }
  JRecord& r = b.pend.as<JRecord>()[s];
  for (int k = 0; k < 3; ++k) {
    r.pos[k] = pos[k];
    r.vel[k] = vel[k];
    r.acc2[k] = acc2[k];
    r.jrk6[k] = jrk6[k];
  }
  r.mass = m;
  r.time = t;
  r.idx = pidx;
  r.addr = padr;
}

void board_predict(Board& b, int nj, hipStream_t st, bool sorted = false);

void board_flush_j(Board& b) {
  if (b.npend == 0) return;
  if (b.predict_flag && b.pred_armed)
    board_predict(b, b.pred_nj, b.stream);
  else if (b.pred_virtual)
    board_predict(b, b.virt_nj, b.stream);
  board_use_stream(b, b.stream);
  // the scatter kernel reads the pinned queue through its device mapping
  hipLaunchKernelGGL(pg6::pg6_scatter_j, dim3(ceil_div(b.npend, 256)), dim3(256), 0, b.stream,
                     b.npend, mapped<JRecord>(b.pend), b.js);
  PGAMD_HIP_CHECK(hipGetLastError());
  b.sort_writes += b.npend;
  // No wait here: the next set_j_particle waits for this event before it
  // writes the pinned queue again (by then lasthalf has usually synchronised).
  PGAMD_HIP_CHECK(hipEventRecord(b.pend_ev, b.stream));
  b.pend_busy = true;
  const JRecord* r = b.pend.as<JRecord>();
  for (int k = 0; k < b.npend; ++k) b.slot[r[k].addr] = -1;
  b.npend = 0;
}

void board_predict(Board& b, int nj, hipStream_t st, bool sorted) {
  b.predict_flag = 0;
  b.pred_armed = false;
  b.pred_virtual = false;
  if (nj <= 0) return;
  board_use_stream(b, st);
  hipLaunchKernelGGL(pg6::pg6_predict, dim3(ceil_div(nj, 256)), dim3(256), 0, st, nj, b.ti,
                     b.js, b.jp);
  PGAMD_HIP_CHECK(hipGetLastError());
  // the sorted records are rewritten only while the layout fits this nj
  const bool srt = sorted && b.sort_nj == nj;
  if (srt) {
    hipLaunchKernelGGL(pg6::pg6_sort_records, dim3(ceil_div(nj, 256)), dim3(256), 0, st, nj,
                       b.jp, b.so, false);
    PGAMD_HIP_CHECK(hipGetLastError());
    ++b.sort_age;
  }
  board_done_on(b, st);
  b.nj_pred = nj;
  b.srec_ok = srt;
}

// The sorted layout is rebuilt (new Morton order) after kSortAge predictions
// or once as many j-records as it holds have been written since its build:
// between rebuilds each prediction recenters the chunk origins, and what ages
// is how compact the chunks stay as their members move apart.
constexpr int kSortAge = 64;

// (Re)builds the sorted layout of the first nj predicted j-particles on st
// (asynchronously): box of the positions, Morton keys, radix sort, then the
// inverse map, chunk origins, flags and records (pg6_sort_records).
void board_build_sort(Board& b, int nj, hipStream_t st) {
  const int nc = ceil_div(nj, pg6::kChunk);
  const int nbox = std::min(256, ceil_div(nj, 256));
  // keys (2 nj u64), vals (2 nj), perm, inv (nj each), rec (2 nj float4),
  // org (3 nc), box partials (6 nbox), box (4), cflag (nc)
  const size_t bytes = sizeof(unsigned long long) * 2 * (size_t)nj + sizeof(int) * 4 * (size_t)nj +
                       sizeof(float4) * 2 * (size_t)nj +
                       sizeof(double) * (3 * (size_t)nc + 6 * (size_t)nbox + 4) +
                       sizeof(int) * (size_t)nc + 9 * 256;
  b.d_sort.reserve(bytes);
  char* base = static_cast<char*>(b.d_sort.ptr);
  auto take = [&](size_t n) {
    char* p = base;
    base += (n + 255) / 256 * 256;
    return p;
  };
  unsigned long long* keys = reinterpret_cast<unsigned long long*>(take(sizeof(unsigned long long) * 2 * (size_t)nj));
  int* vals = reinterpret_cast<int*>(take(sizeof(int) * 2 * (size_t)nj));
  b.so.perm = reinterpret_cast<int*>(take(sizeof(int) * (size_t)nj));
  b.so.inv = reinterpret_cast<int*>(take(sizeof(int) * (size_t)nj));
  b.so.rec = reinterpret_cast<float4*>(take(sizeof(float4) * 2 * (size_t)nj));
  b.so.org = reinterpret_cast<double*>(take(sizeof(double) * 3 * (size_t)nc));
  double* boxp = reinterpret_cast<double*>(take(sizeof(double) * 6 * (size_t)nbox));
  double* box = reinterpret_cast<double*>(take(sizeof(double) * 4));
  b.so.cflag = reinterpret_cast<int*>(take(sizeof(int) * (size_t)nc));
  hipLaunchKernelGGL(pg6::pg6_box_part, dim3(nbox), dim3(256), 0, st, nj, b.jp, boxp);
  hipLaunchKernelGGL(pg6::pg6_box_final, dim3(1), dim3(256), 0, st, nbox, boxp, box);
  hipLaunchKernelGGL(pg6::pg6_sort_keys, dim3(ceil_div(nj, 256)), dim3(256), 0, st, nj, b.jp,
                     box, keys, vals);
  PGAMD_HIP_CHECK(hipGetLastError());
  size_t tmp = 0;
  PGAMD_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, keys, keys + nj, vals,
                                                     b.so.perm, nj, 0, 63, st));
  b.d_sort_tmp.reserve(tmp);
  PGAMD_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(b.d_sort_tmp.ptr, tmp, keys, keys + nj,
                                                     vals, b.so.perm, nj, 0, 63, st));
  hipLaunchKernelGGL(pg6::pg6_sort_records, dim3(ceil_div(nj, 256)), dim3(256), 0, st, nj,
                     b.jp, b.so, true);
  PGAMD_HIP_CHECK(hipGetLastError());
  b.sort_nj = nj;
  b.sort_age = 0;
  b.sort_writes = 0;
  b.srec_ok = true;
}

// The predicted j-set is needed for the first nj j-particles.
inline bool board_needs_pred(const Board& b, int nj) {
  return nj > 0 && (b.predict_flag || b.nj_pred < nj);
}

LaunchPlan plan_launch(const Board& b, int ni, int nj) {
  LaunchPlan p;
  const int k = ni_bucket(ni);
  const LaunchChoice c = b.tuned_ok[k] ? b.tuned[k] : default_choice(ni);
  p.bi = b.force_bi > 0 ? b.force_bi : c.bi;
  p.variant = b.force_variant >= 0 ? b.force_variant : c.variant;
  p.bpc = b.force_bpc > 0 ? b.force_bpc : c.bpc;
  if (p.variant == 3 && ni > kFusedMaxNi) p.variant = 2;
  if ((p.variant == 4 || p.variant == 5) && b.force_variant != p.variant && nj < kSortedMinNj)
    p.variant = 2;
  if (p.variant == 3) p.bi = 64;
  // variant 5: blocks of bi threads on 64 i each, the j-range split over the
  // block's bi / 64 wavefronts
  const int nbi = ceil_div(ni, p.variant == 5 ? 64 : p.bi);
  // the SGPR-broadcast kernels run bi threads per block, the LDS one 256, the
  // one-launch kernel 512 (bpc counts quarters of a block per CU there)
  const int target = p.variant == 3 ? std::max(1, device_cu_count() * p.bpc / 4)
                                    : device_cu_count() * p.bpc * (p.variant >= 1 ? 256 / p.bi : 1);
  int ns = std::max(1, ceil_div(target, nbi));
  const int maxs = std::max(1, nj / std::max(1, b.min_jper));
  ns = std::min(ns, maxs);
  int jper = ceil_div(nj, ns);
  static_assert(pg6::kChunk == 64, "j-ranges are aligned to the self-filter chunks");
  const int align = p.variant == 5 ? p.bi : 64;  // whole chunks per wavefront
  jper = (jper + align - 1) / align * align;
  ns = std::max(1, ceil_div(nj, jper));
  p.nsplit = ns;
  p.jper = jper;
  return p;
}

// pred_in: the LDS kernel predicts the j-set itself (variant 0 with one
// i-block only, board_force decides), at time ti from js, without writing
// the predicted set.
template <int BI, bool NN>
void launch_force_t(const LaunchPlan& p, int ni, const double* ipos, const double* ivel,
                    const int* iid, float eps2, int nj, const JPred& jp, const pg6::JSort& so,
                    const Partials& part, const Outputs& out, hipStream_t st, bool pred_in,
                    double ti, const JState& js) {
  dim3 grid(ceil_div(ni, BI), p.nsplit);
  // LDS kernel with one i-block: 8 wavefronts for up to 128 i, else 4
  constexpr int kNT1 = BI <= 128 ? pg6::kOneBlockThreads : pg6::kThreads;
  g_timer.before(st);
  if (p.variant == 5)  // 64 i per block, jper a multiple of BI sorted positions
    hipLaunchKernelGGL((pg6::pg6_force_ow<NN, BI / 64>), dim3(ceil_div(ni, 64), p.nsplit),
                       dim3(BI), 0, st, ni, ipos, ivel, iid, eps2, nj, p.jper, so, jp, part, out,
                       p.nsplit);
  else if (p.variant == 4)  // ranges of sorted positions, multiples of pg6::kChunk
    hipLaunchKernelGGL((pg6::pg6_force_o<NN>), grid, dim3(BI), 0, st, ni, ipos, ivel, iid, eps2,
                       nj, p.jper, so, jp, part, out, p.nsplit);
  else if (p.variant == 2)  // plan_launch keeps jper a multiple of pg6::kChunk
    hipLaunchKernelGGL((pg6::pg6_force_h<NN>), grid, dim3(BI), 0, st, ni, ipos, ivel, iid, eps2,
                       nj, p.jper, jp, part, out, p.nsplit);
  else if (p.variant == 1)
    hipLaunchKernelGGL((pg6::pg6_force_s<NN>), grid, dim3(BI), 0, st, ni, ipos, ivel, iid, eps2,
                       nj, p.jper, jp, part, out, p.nsplit);
  else if (grid.x == 1 && pred_in)  // (pred_in implies one i-block)
    hipLaunchKernelGGL((pg6::pg6_force<BI, NN, true, kNT1>), grid, dim3(kNT1), 0, st, ni, ipos,
                       ivel, iid, eps2, nj, p.jper, jp, part, out, p.nsplit, ti, js);
  else if (grid.x == 1)
    hipLaunchKernelGGL((pg6::pg6_force<BI, NN, false, kNT1>), grid, dim3(kNT1), 0, st, ni, ipos,
                       ivel, iid, eps2, nj, p.jper, jp, part, out, p.nsplit, ti, js);
  else
    hipLaunchKernelGGL((pg6::pg6_force<BI, NN, false>), grid, dim3(pg6::kThreads), 0, st, ni,
                       ipos, ivel, iid, eps2, nj, p.jper, jp, part, out, p.nsplit, ti, js);
  PGAMD_HIP_CHECK(hipGetLastError());
  g_timer.after(st);
  if (p.nsplit > 1) {
    // one thread per i walks the splits; with many splits a block of 256
    // threads per i (up to 1024 i: four-thread sums would leave most CUs
    // idle) or four threads per i (more i, coalesced loads; 2-5 % faster
    // calls at 2048 and 4096 i against N = 16384).  After an in-kernel
    // prediction (one i-block) the neighbour's id and distance come with the
    // partials: always the block-per-i sum.
    if (pred_in)
      hipLaunchKernelGGL((pg6::pg6_reduce_w<NN, true>), dim3(ni), dim3(256), 0, st, ni,
                         p.nsplit, part, jp, ipos, eps2, out);
    else if (p.nsplit > 16 && ni <= 1024)
      hipLaunchKernelGGL((pg6::pg6_reduce_w<NN>), dim3(ni), dim3(256), 0, st, ni,
                         p.nsplit, part, jp, ipos, eps2, out);
    else if (p.nsplit > 8)
      hipLaunchKernelGGL((pg6::pg6_reduce_q4<NN>), dim3(ceil_div(ni, 64)), dim3(256), 0, st, ni,
                         p.nsplit, part, jp, ipos, eps2, out);
    else
      hipLaunchKernelGGL((pg6::pg6_reduce<NN>), dim3(ceil_div(ni, 256)), dim3(256), 0, st, ni,
                         p.nsplit, part, jp, ipos, eps2, out);
    PGAMD_HIP_CHECK(hipGetLastError());
  }
}

// Runs the force kernels for ni i-particles (device arrays) against the first
// nj predicted j-particles.  out.nnb == nullptr selects the kernel without
// nearest-neighbour tracking.
void board_force(Board& b, int ni, int nj, const double* ipos, const double* ivel,
                 const int* iid, double eps2, const Outputs& out, hipStream_t st) {
  if (ni <= 0) return;
  board_use_stream(b, st);
  const LaunchPlan p = plan_launch(b, ni, nj > 0 ? nj : 1);
  const bool pred = board_needs_pred(b, nj);
  if ((p.variant == 4 || p.variant == 5) && nj > 0) {
    // the sorted layout of the current prediction
    const bool fresh = b.sort_nj == nj && b.sort_age < kSortAge && b.sort_writes < nj;
    // (a layout of another j count orders other particles: rebuild it)
    if (pred || !b.srec_ok || b.sort_nj != nj) {
      if (fresh) {
        board_predict(b, nj, st, true);
      } else {
        if (pred) board_predict(b, nj, st);
        board_build_sort(b, nj, st);
      }
    }
  }
  // one i-block of the LDS kernel: every j sits in one block's range, so the
  // blocks predict their own j (no predictor launch)
  const bool pred_in = pred && p.variant == 0 && ceil_div(ni, p.bi) == 1;
  if (pred && !pred_in && p.variant != 3 && p.variant != 4 && p.variant != 5)
    board_predict(b, nj, st);
  Partials part{};
  if (p.nsplit > 1) {
    const size_t n = (size_t)p.nsplit * ni;
    b.d_part.reserve(n * (7 * sizeof(double) + 2 * sizeof(float) + 2 * sizeof(int)));
    double* d = b.d_part.as<double>();
    part.ax = d;
    part.ay = d + n;
    part.az = d + 2 * n;
    part.pot = d + 3 * n;
    part.jx = d + 4 * n;
    part.jy = d + 5 * n;
    part.jz = d + 6 * n;
    part.rmax = reinterpret_cast<float*>(d + 7 * n);
    part.nn = reinterpret_cast<int*>(part.rmax + n);
    part.nid = part.nn + n;
    part.rdist = reinterpret_cast<float*>(part.nid + n);
  }
  const bool nn = out.nnb != nullptr;
  const float e2 = (float)eps2;
  const int njr = nj > 0 ? nj : 0;
  if (p.variant == 3) {
    // one launch: prediction (if due) and j-range reduction inside
    const int nib = ceil_div(ni, 64);
    if (b.d_fcount.bytes < sizeof(int) * (size_t)nib) {
      b.d_fcount.reserve(sizeof(int) * (size_t)std::max(nib, 1024));
      PGAMD_HIP_CHECK(hipMemsetAsync(b.d_fcount.ptr, 0, b.d_fcount.bytes, st));
    }
    int* cnt = b.d_fcount.as<int>();
    dim3 grid(nib, p.nsplit);
    g_timer.before(st);
#define PG6_FUSED(NN, PRED)                                                                  \
    hipLaunchKernelGGL((pg6::pg6_force_f<NN, PRED>), grid, dim3(pg6::kFThreads), 0, st, ni,  \
                       ipos, ivel, iid, e2, njr, p.jper, b.ti, b.js, b.jp, part, out,        \
                       p.nsplit, cnt)
    if (nn && pred) PG6_FUSED(true, true);
    else if (nn) PG6_FUSED(true, false);
    else if (pred) PG6_FUSED(false, true);
    else PG6_FUSED(false, false);
#undef PG6_FUSED
    PGAMD_HIP_CHECK(hipGetLastError());
    g_timer.after(st);
    if (pred) {
      b.nj_pred = nj;
      b.predict_flag = 0;
      b.pred_armed = false;
      b.pred_virtual = false;
      b.srec_ok = false;  // the sorted records hold an older prediction
    }
    board_done_on(b, st);
    return;
  }
#define PG6_LAUNCH(BI)                                                                     \
  do {                                                                                     \
    if (nn)                                                                                \
      launch_force_t<BI, true>(p, ni, ipos, ivel, iid, e2, njr, b.jp, b.so, part, out, st,  \
                               pred_in, b.ti, b.js);                                       \
    else                                                                                   \
      launch_force_t<BI, false>(p, ni, ipos, ivel, iid, e2, njr, b.jp, b.so, part, out, st, \
                                pred_in, b.ti, b.js);                                      \
  } while (0)
  switch (p.bi) {
    case 256: PG6_LAUNCH(256); break;
    case 128: PG6_LAUNCH(128); break;
    default: PG6_LAUNCH(64); break;
  }
#undef PG6_LAUNCH
  if (pred_in) {
    // the prediction lived in the kernel's LDS only: calls at this ti predict
    // again (same j-state, same bits) until a flush of new j-records writes
    // it out first (board_flush_j)
    b.pred_armed = false;
    b.pred_virtual = true;
    b.virt_nj = nj;
  }
  board_done_on(b, st);
}

// Host-API lasthalf: flags select nearest neighbour, rnnb and neighbour lists.
int lasthalf_impl(const char* fn, int nj, int ni, const int* pidx, const double (*pos)[3],
                  const double (*vel)[3], double eps2, const double* h2, double (*acc)[3],
                  double (*jrk)[3], double* pot, int* nnb, float* rnnb, bool want_nn,
                  bool want_nblist) {
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!board_check_open(fn)) return -1;
  Board& b = g_board;
  if (ni < 0 || nj < 0) {
    api_error(3, "%s: negative ni (%d) or nj (%d)", fn, ni, nj);
    return -1;
  }
  if (nj > b.cap) {
    api_error(2, "%s: nj=%d exceeds the j-particles set (capacity %d)", fn, nj, b.cap);
    return -1;
  }
  if (ni == 0) return 0;
  // Pending j records (a firsthalf's prediction, if armed, runs first); a
  // missing prediction runs with the force.
  board_flush_j(b);
  // i-set: pos[3] double, vel[3] double, id int, h2 float
  const size_t in_bytes = (size_t)ni * (6 * sizeof(double) + sizeof(int) + sizeof(float));
  b.h_in.reserve(in_bytes);
  b.d_in.reserve(in_bytes);
  double* hp = b.h_in.as<double>();
  std::memcpy(hp, pos, sizeof(double) * 3 * (size_t)ni);
  std::memcpy(hp + 3 * (size_t)ni, vel, sizeof(double) * 3 * (size_t)ni);
  int* hid = reinterpret_cast<int*>(hp + 6 * (size_t)ni);
  std::memcpy(hid, pidx, sizeof(int) * (size_t)ni);
  float* hh2 = reinterpret_cast<float*>(hid + ni);
  for (int i = 0; i < ni; ++i) hh2[i] = want_nblist && h2 ? (float)h2[i] : 0.f;
  upload(b.d_in.ptr, b.h_in, 0, in_bytes, b.stream);
  const double* dpos = b.d_in.as<double>();
  const double* dvel = dpos + 3 * (size_t)ni;
  const int* did = reinterpret_cast<const int*>(dvel + 3 * (size_t)ni);
  const float* dh2 = reinterpret_cast<const float*>(did + ni);

  // The kernels write the results straight into pinned host memory (no copy
  // back to queue behind them); the synchronisation below makes them visible.
  const size_t out_bytes = (size_t)ni * (7 * sizeof(double) + sizeof(int) + sizeof(float));
  b.h_out.reserve(out_bytes);
  void* dout = nullptr;
  PGAMD_HIP_CHECK(hipHostGetDevicePointer(&dout, b.h_out.ptr, 0));
  Outputs o;
  o.acc = static_cast<double*>(dout);
  o.jrk = o.acc + 3 * (size_t)ni;
  o.pot = o.jrk + 3 * (size_t)ni;
  o.nnb = want_nn ? reinterpret_cast<int*>(o.pot + ni) : nullptr;
  o.rnnb = reinterpret_cast<float*>(reinterpret_cast<int*>(o.pot + ni) + ni);
  board_force(b, ni, nj, dpos, dvel, did, eps2, o, b.stream);

  if (want_nblist) {
    // the list kernel reads the predicted set
    if (board_needs_pred(b, nj)) board_predict(b, nj, b.stream);
    b.d_nb.reserve((size_t)ni * (kMaxLen + 1) * sizeof(int));
    int* dnbl = b.d_nb.as<int>();
    int* dlen = dnbl + (size_t)ni * kMaxLen;
    hipLaunchKernelGGL(pg6::pg6_nblist, dim3(ni), dim3(pg6::kThreads), 0, b.stream, ni, dpos,
                       did, dh2, (float)eps2, nj, b.jp, kMaxLen, dnbl, dlen);
    PGAMD_HIP_CHECK(hipGetLastError());
    b.nbl.resize((size_t)ni * kMaxLen);
    b.nblen.resize((size_t)ni);
    PGAMD_HIP_CHECK(hipMemcpyAsync(b.nblen.data(), dlen, sizeof(int) * (size_t)ni,
                                   hipMemcpyDeviceToHost, b.stream));
    PGAMD_HIP_CHECK(hipMemcpyAsync(b.nbl.data(), dnbl, sizeof(int) * (size_t)ni * kMaxLen,
                                   hipMemcpyDeviceToHost, b.stream));
  }
  PGAMD_HIP_CHECK(hipStreamSynchronize(b.stream));
  const double* ha = b.h_out.as<double>();
  const double* hj = ha + 3 * (size_t)ni;
  const double* hpo = hj + 3 * (size_t)ni;
  const int* hn = reinterpret_cast<const int*>(hpo + ni);
  const float* hr = reinterpret_cast<const float*>(hn + ni);
  std::memcpy(acc, ha, sizeof(double) * 3 * (size_t)ni);
  std::memcpy(jrk, hj, sizeof(double) * 3 * (size_t)ni);
  std::memcpy(pot, hpo, sizeof(double) * (size_t)ni);
  if (want_nn && nnb) std::memcpy(nnb, hn, sizeof(int) * (size_t)ni);
  if (want_nn && rnnb) std::memcpy(rnnb, hr, sizeof(float) * (size_t)ni);
  if (want_nblist) {
    b.nb_ni = ni;
    for (int i = 0; i < ni; ++i)
      if (b.nblen[i] > kMaxLen) b.nblerror = 1;
  }
  return 0;
}

}  // namespace

// ---------------------------------------------------------------- GRAPE-6 API

PGAMD_API void g6_open(int gpid) {
  (void)gpid;
  std::lock_guard<std::mutex> lk(g_mutex);
  board_open(g_board);
}
PGAMD_API void g6_open_(int* gpid) { g6_open(*gpid); }

PGAMD_API void g6_close(int gpid) {
  (void)gpid;
  std::lock_guard<std::mutex> lk(g_mutex);
  board_close(g_board);
}
PGAMD_API void g6_close_(int* gpid) { g6_close(*gpid); }

// Hardware-specific settings of the GRAPE-6 board: no-ops, as in the reference
// (PG6/libpg6/phantom_g6.c:36-88,468-478).
PGAMD_API void g6_set_tunit(int) {}
PGAMD_API void g6_set_tunit_(int*) {}
PGAMD_API void g6_set_xunit(int) {}
PGAMD_API void g6_set_xunit_(int*) {}
PGAMD_API void g6_reset(int) {}
PGAMD_API void g6_reset_(int*) {}
PGAMD_API void g6_reset_fofpga(int) {}
PGAMD_API void g6_reset_fofpga_(int*) {}
PGAMD_API void g6_setup_njdata(int, int) {}
PGAMD_API void g6_setup_njdata_(int*, int*) {}

PGAMD_API int g6_npipes(void) {
  if (const char* s = std::getenv("PGAMD_G6_NPIPES")) {
    int v = std::atoi(s);
    if (v > 0) return v;
  }
  return g_board.open ? g_board.npipes : kDefaultPipes;
}
PGAMD_API int g6_npipes_(void) { return g6_npipes(); }

PGAMD_API void g6_set_j_particle(int gpid, int padr, int pidx, double t, double dt, double mss,
                                 double* snp, double* jrk, double* acc, double* vel,
                                 double* pos) {
  (void)gpid;
  (void)dt;
  (void)snp;
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!board_check_open("g6_set_j_particle")) return;
  if (padr < 0) {
    api_error(2, "g6_set_j_particle: negative address %d", padr);
    return;
  }
  if (!board_reserve(g_board, padr + 1)) return;
  board_queue_j(g_board, padr, pidx, t, mss, acc, jrk, vel, pos);
}
PGAMD_API void g6_set_j_particle_(int* gpid, int* padr, int* pidx, double* t, double* dt,
                                  double* mss, double* snp, double* jrk, double* acc,
                                  double* vel, double* pos) {
  g6_set_j_particle(*gpid, *padr, *pidx, *t, *dt, *mss, snp, jrk, acc, vel, pos);
}

PGAMD_API void g6_set_ti(int gpid, double ti) {
  (void)gpid;
  std::lock_guard<std::mutex> lk(g_mutex);
  g_board.ti = ti;
  g_board.predict_flag = 1;
  g_board.pred_virtual = false;
}
PGAMD_API void g6_set_ti_(int* gpid, double* ti) { g6_set_ti(*gpid, *ti); }

PGAMD_API void g6calc_firsthalf(int gpid, int nj, int ni, int* pidx, double (*pos)[3],
                                double (*vel)[3], double (*acc)[3], double (*jrk)[3],
                                double* pot, double eps2, double* h2) {
  (void)gpid; (void)ni; (void)pidx; (void)pos; (void)vel; (void)acc; (void)jrk; (void)pot;
  (void)eps2; (void)h2;
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!board_check_open("g6calc_firsthalf")) return;
  Board& b = g_board;
  if (nj < 0 || nj > b.cap) {
    api_error(2, "g6calc_firsthalf: nj=%d outside the j-memory (capacity %d)", nj, b.cap);
    return;
  }
  board_flush_j(b);
  if (b.predict_flag) {  // predicted with the first force that needs it
    b.pred_armed = true;
    b.pred_nj = nj;
  }
  b.nblerror = 0;
}
PGAMD_API void g6calc_firsthalf_(int* gpid, int* nj, int* ni, int* pidx, double (*pos)[3],
                                 double (*vel)[3], double (*acc)[3], double (*jrk)[3],
                                 double* pot, double* eps2, double* h2) {
  g6calc_firsthalf(*gpid, *nj, *ni, pidx, pos, vel, acc, jrk, pot, *eps2, h2);
}

PGAMD_API int g6calc_lasthalf(int gpid, int nj, int ni, int* pidx, double (*pos)[3],
                              double (*vel)[3], double eps2, double* h2, double (*acc)[3],
                              double (*jrk)[3], double* pot) {
  (void)gpid;
  int rc = lasthalf_impl("g6calc_lasthalf", nj, ni, pidx, pos, vel, eps2, h2, acc, jrk, pot,
                         nullptr, nullptr, false, false);
  return rc < 0 ? rc : 0;
}
PGAMD_API int g6calc_lasthalf_(int* gpid, int* nj, int* ni, int* pidx, double (*pos)[3],
                               double (*vel)[3], double* eps2, double* h2, double (*acc)[3],
                               double (*jrk)[3], double* pot) {
  return g6calc_lasthalf(*gpid, *nj, *ni, pidx, pos, vel, *eps2, h2, acc, jrk, pot);
}

PGAMD_API int g6calc_lasthalf2(int gpid, int nj, int ni, int* pidx, double (*pos)[3],
                               double (*vel)[3], double eps2, double* h2, double (*acc)[3],
                               double (*jrk)[3], double* pot, int* nnb) {
  (void)gpid;
  return lasthalf_impl("g6calc_lasthalf2", nj, ni, pidx, pos, vel, eps2, h2, acc, jrk, pot,
                       nnb, nullptr, true, false);
}
PGAMD_API int g6calc_lasthalf2_(int* gpid, int* nj, int* ni, int* pidx, double (*pos)[3],
                                double (*vel)[3], double* eps2, double* h2, double (*acc)[3],
                                double (*jrk)[3], double* pot, int* nnb) {
  return g6calc_lasthalf2(*gpid, *nj, *ni, pidx, pos, vel, *eps2, h2, acc, jrk, pot, nnb);
}

PGAMD_API int g6calc_lasthalf2p(int gpid, int nj, int ni, int* pidx, double (*pos)[3],
                                double (*vel)[3], double eps2, double* h2, double (*acc)[3],
                                double (*jrk)[3], double* pot, int* nnb, float* rnnb) {
  (void)gpid;
  return lasthalf_impl("g6calc_lasthalf2p", nj, ni, pidx, pos, vel, eps2, h2, acc, jrk, pot,
                       nnb, rnnb, true, false);
}
PGAMD_API int g6calc_lasthalf2p_(int* gpid, int* nj, int* ni, int* pidx, double (*pos)[3],
                                 double (*vel)[3], double* eps2, double* h2, double (*acc)[3],
                                 double (*jrk)[3], double* pot, int* nnb, float* rnnb) {
  return g6calc_lasthalf2p(*gpid, *nj, *ni, pidx, pos, vel, *eps2, h2, acc, jrk, pot, nnb, rnnb);
}

PGAMD_API int g6calc_lasthalfn(int gpid, int nj, int ni, int* pidx, double (*pos)[3],
                               double (*vel)[3], double eps2, double* h2, double (*acc)[3],
                               double (*jrk)[3], double* pot) {
  (void)gpid;
  return lasthalf_impl("g6calc_lasthalfn", nj, ni, pidx, pos, vel, eps2, h2, acc, jrk, pot,
                       nullptr, nullptr, false, true);
}
PGAMD_API int g6calc_lasthalfn_(int* gpid, int* nj, int* ni, int* pidx, double (*pos)[3],
                                double (*vel)[3], double* eps2, double* h2, double (*acc)[3],
                                double (*jrk)[3], double* pot) {
  return g6calc_lasthalfn(*gpid, *nj, *ni, pidx, pos, vel, *eps2, h2, acc, jrk, pot);
}

PGAMD_API int g6calc_lasthalf2n(int gpid, int nj, int ni, int* pidx, double (*pos)[3],
                                double (*vel)[3], double eps2, double* h2, double (*acc)[3],
                                double (*jrk)[3], double* pot, int* nnb) {
  (void)gpid;
  return lasthalf_impl("g6calc_lasthalf2n", nj, ni, pidx, pos, vel, eps2, h2, acc, jrk, pot,
                       nnb, nullptr, true, true);
}
PGAMD_API int g6calc_lasthalf2n_(int* gpid, int* nj, int* ni, int* pidx, double (*pos)[3],
                                 double (*vel)[3], double* eps2, double* h2, double (*acc)[3],
                                 double (*jrk)[3], double* pot, int* nnb) {
  return g6calc_lasthalf2n(*gpid, *nj, *ni, pidx, pos, vel, *eps2, h2, acc, jrk, pot, nnb);
}

PGAMD_API int g6calc_lasthalf2np(int gpid, int nj, int ni, int* pidx, double (*pos)[3],
                                 double (*vel)[3], double eps2, double* h2, double (*acc)[3],
                                 double (*jrk)[3], double* pot, int* nnb, float* rnnb) {
  (void)gpid;
  return lasthalf_impl("g6calc_lasthalf2np", nj, ni, pidx, pos, vel, eps2, h2, acc, jrk, pot,
                       nnb, rnnb, true, true);
}
PGAMD_API int g6calc_lasthalf2np_(int* gpid, int* nj, int* ni, int* pidx, double (*pos)[3],
                                  double (*vel)[3], double* eps2, double* h2, double (*acc)[3],
                                  double (*jrk)[3], double* pot, int* nnb, float* rnnb) {
  return g6calc_lasthalf2np(*gpid, *nj, *ni, pidx, pos, vel, *eps2, h2, acc, jrk, pot, nnb,
                            rnnb);
}

// 0 = every list of the last lasthalf*n call fitted; 1 = one overflowed
// (PG6/libpg6/phantom_g6.c:482-486, gravity.c:905-906).
PGAMD_API int g6_read_neighbour_list(int gpid) {
  (void)gpid;
  std::lock_guard<std::mutex> lk(g_mutex);
  return g_board.nblerror;
}
PGAMD_API int g6_read_neighbour_list_(int* gpid) { return g6_read_neighbour_list(*gpid); }

// Copies the list of i-particle `ipipe` of the last lasthalf*n call.
// Returns 1 if it is longer than maxlen (nothing copied), else 0
// (PG6/libpg6/gravity.c:135-147).
PGAMD_API int g6_get_neighbour_list(int gpid, int ipipe, int maxlen, int* nblen, int* nbl) {
  (void)gpid;
  std::lock_guard<std::mutex> lk(g_mutex);
  Board& b = g_board;
  if (ipipe < 0 || ipipe >= b.nb_ni) {
    api_error(3, "g6_get_neighbour_list: ipipe=%d outside the last call's %d i-particles",
              ipipe, b.nb_ni);
    return 1;
  }
  const int len = b.nblen[ipipe];
  if (len > maxlen || len > kMaxLen) return 1;
  *nblen = len;
  std::memcpy(nbl, b.nbl.data() + (size_t)ipipe * kMaxLen, sizeof(int) * (size_t)len);
  return 0;
}
PGAMD_API int g6_get_neighbour_list_(int* gpid, int* ipipe, int* maxlen, int* nblen, int* nbl) {
  return g6_get_neighbour_list(*gpid, *ipipe, *maxlen, nblen, nbl);
}

// Debug dumps with the reference's file names and formats
// (PG6/libpg6/phantom_g6.c:508-561).
PGAMD_API void g6_debugfunc(int tmp) {
  FILE* fp = std::fopen("dump.dat", "a");
  if (!fp) return;
  std::fprintf(fp, "%d\n", tmp);
  std::fclose(fp);
}
PGAMD_API void g6_debugfunc_(int* tmp) { g6_debugfunc(*tmp); }
PGAMD_API void g6_debugfunc_double(double tmp) {
  FILE* fp = std::fopen("dump.dat", "a");
  if (!fp) return;
  std::fprintf(fp, "%+e\n", tmp);
  std::fclose(fp);
}
PGAMD_API void g6_debugfunc_double_(double* tmp) { g6_debugfunc_double(*tmp); }
PGAMD_API void g6_dump(double tim, int n, double* m, double (*x)[3], double (*v)[3]) {
  char out[64];
  std::snprintf(out, sizeof(out), "dump.dat.%05d", (int)tim);
  FILE* fp = std::fopen(out, "w");
  if (!fp) return;
  for (int i = 0; i < n; i++) {
    std::fprintf(fp, "%+.16e", m[i]);
    std::fprintf(fp, " %+.16e %+.16e %+.16e", x[i][0], x[i][1], x[i][2]);
    std::fprintf(fp, " %+.16e %+.16e %+.16e", v[i][0], v[i][1], v[i][2]);
    std::fprintf(fp, "\n");
  }
  std::fclose(fp);
}
PGAMD_API void g6_dump_(double* tim, int* n, double* m, double (*x)[3], double (*v)[3]) {
  g6_dump(*tim, *n, m, x, v);
}

// ------------------------------------------------- device-pointer entry points

// Bulk load of j-particles 0..nj-1 from host arrays (same layout and
// conventions as nj calls of g6_set_j_particle: acc2 = a/2, jrk6 = j/6).
PGAMD_API int pg6_dev_set_j(int nj, const double* pos, const double* vel, const double* acc2,
                            const double* jrk6, const double* mass, const double* tj,
                            const int* idx) {
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!board_check_open("pg6_dev_set_j")) return -1;
  if (nj < 0) {
    api_error(3, "pg6_dev_set_j: negative nj %d", nj);
    return -1;
  }
  if (!board_reserve(g_board, nj)) return -1;
  for (int j = 0; j < nj; ++j)
    board_queue_j(g_board, j, idx[j], tj[j], mass[j], acc2 + 3 * (size_t)j,
                  jrk6 + 3 * (size_t)j, vel + 3 * (size_t)j, pos + 3 * (size_t)j);
  board_flush_j(g_board);
  return 0;
}

// Predicts j-particles 0..nj-1 to time ti on `stream` (nullptr = library stream).
PGAMD_API int pg6_dev_predict(int nj, double ti, void* stream) {
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!board_check_open("pg6_dev_predict")) return -1;
  if (nj < 0 || nj > g_board.cap) {
    api_error(2, "pg6_dev_predict: nj=%d outside the j-memory (capacity %d)", nj, g_board.cap);
    return -1;
  }
  // Lazy: the first force after this (or a flush of new j-records) predicts;
  // pg6_force_f does it inside its own launch.  `stream` orders nothing here.
  (void)stream;
  g_board.pred_virtual = false;  // superseded: pending records enter the new one
  board_flush_j(g_board);
  g_board.ti = ti;
  g_board.predict_flag = 1;
  g_board.pred_armed = true;
  g_board.pred_nj = nj;
  return 0;
}

// Force on ni i-particles held in device memory against the first nj
// predicted j-particles.  All pointers are device pointers:
//   pos[ni][3], vel[ni][3] (FP64), id[ni]; acc[ni][3], jrk[ni][3], pot[ni]
//   (FP64); nnb[ni], rnnb[ni] may be null (then no nearest-neighbour search).
// Asynchronous on `stream` (nullptr = library stream).  Calls that overlap in
// time on different streams share the partial-sum workspace and race.
PGAMD_API int pg6_dev_force(int nj, int ni, const double* pos, const double* vel, const int* id,
                            double eps2, double* acc, double* jrk, double* pot, int* nnb,
                            float* rnnb, void* stream) {
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!board_check_open("pg6_dev_force")) return -1;
  if (ni < 0 || nj < 0 || nj > g_board.cap) {
    api_error(2, "pg6_dev_force: ni=%d nj=%d (capacity %d)", ni, nj, g_board.cap);
    return -1;
  }
  Outputs o{acc, jrk, pot, nnb, rnnb};
  if (nnb && !rnnb) {
    api_error(3, "pg6_dev_force: nnb given without rnnb");
    return -1;
  }
  board_force(g_board, ni, nj, pos, vel, id, eps2, o,
              stream ? (hipStream_t)stream : g_board.stream);
  return 0;
}

PGAMD_API int pg6_dev_sync(void) {
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!board_check_open("pg6_dev_sync")) return -1;
  board_use_stream(g_board, g_board.stream);
  PGAMD_HIP_CHECK(hipStreamSynchronize(g_board.stream));
  return 0;
}

PGAMD_API void* pg6_dev_stream(void) { return g_board.open ? (void*)g_board.stream : nullptr; }

// Launch-shape overrides: bi in {0 (auto), 64, 128, 256}; blocks_per_cu in
// [1, 256] sets the j-split target (0 = auto); min_jper > 0 the smallest
// j-range (0 = 128).
PGAMD_API int pg6_dev_set_launch(int bi, int blocks_per_cu, int min_jper) {
  std::lock_guard<std::mutex> lk(g_mutex);
  if (bi != 0 && bi != 64 && bi != 128 && bi != 256) {
    api_error(3, "pg6_dev_set_launch: bi=%d not in {0,64,128,256}", bi);
    return -1;
  }
  if (blocks_per_cu < 0 || blocks_per_cu > 256) {
    api_error(3, "pg6_dev_set_launch: blocks_per_cu=%d not in [0, 256]", blocks_per_cu);
    return -1;
  }
  g_board.force_bi = bi;
  g_board.force_bpc = blocks_per_cu;
  g_board.min_jper = min_jper > 0 ? min_jper : 128;
  return 0;
}

// The launch shape the library would use for (ni, nj):
// {variant, bi, blocks_per_cu, nsplit, jper}.
PGAMD_API int pg6_dev_plan(int ni, int nj, int* out5) {
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!board_check_open("pg6_dev_plan")) return -1;
  LaunchPlan p = plan_launch(g_board, ni > 0 ? ni : 1, nj > 0 ? nj : 1);
  out5[0] = p.variant;
  out5[1] = p.bi;
  out5[2] = p.bpc;
  out5[3] = p.nsplit;
  out5[4] = p.jper;
  return 0;
}

// Copies the predicted j-particles 0..nj-1 to host memory: pos[nj][3] FP64,
// vel[nj][3] FP32 (as the force kernel reads them), mass[nj] FP32 (may be
// null).  Synchronous.  For checking the predictor.
PGAMD_API int pg6_dev_read_predicted(int nj, double* pos, float* vel, float* mass) {
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!board_check_open("pg6_dev_read_predicted")) return -1;
  Board& b = g_board;
  if (nj < 0 || nj > b.cap) {
    api_error(2, "pg6_dev_read_predicted: nj=%d outside the j-memory (capacity %d)", nj, b.cap);
    return -1;
  }
  if (nj == 0) return 0;
  std::vector<double> x((size_t)nj), y((size_t)nj), z((size_t)nj);
  std::vector<float4> vm((size_t)nj);
  board_flush_j(b);
  if (board_needs_pred(b, nj)) board_predict(b, nj, b.stream);
  board_use_stream(b, b.stream);
  PGAMD_HIP_CHECK(hipStreamSynchronize(b.stream));
  PGAMD_HIP_CHECK(hipMemcpy(x.data(), b.jp.x, sizeof(double) * nj, hipMemcpyDeviceToHost));
  PGAMD_HIP_CHECK(hipMemcpy(y.data(), b.jp.y, sizeof(double) * nj, hipMemcpyDeviceToHost));
  PGAMD_HIP_CHECK(hipMemcpy(z.data(), b.jp.z, sizeof(double) * nj, hipMemcpyDeviceToHost));
  PGAMD_HIP_CHECK(hipMemcpy(vm.data(), b.jp.vm, sizeof(float4) * nj, hipMemcpyDeviceToHost));
  for (int j = 0; j < nj; ++j) {
    pos[3 * j] = x[j];
    pos[3 * j + 1] = y[j];
    pos[3 * j + 2] = z[j];
    vel[3 * j] = vm[j].x;
    vel[3 * j + 1] = vm[j].y;
    vel[3 * j + 2] = vm[j].z;
    if (mass) mass[j] = vm[j].w;
  }
  return 0;
}

// Kernel timing: pg6_dev_timing_start(n) records HIP events around each of the
// next n force-kernel launches (on their launch stream);
// pg6_dev_timing_read(&ms, &count) waits for them and returns the summed
// kernel time in milliseconds and the number of launches timed, then stops.
PGAMD_API int pg6_dev_timing_start(int max_launches) {
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!board_check_open("pg6_dev_timing_start")) return -1;
  if (max_launches <= 0 || max_launches > (1 << 16)) {
    api_error(3, "pg6_dev_timing_start: %d launches not in [1, 65536]", max_launches);
    return -1;
  }
  g_timer.start(max_launches);
  return 0;
}

PGAMD_API int pg6_dev_timing_read(double* ms, int* count) {
  std::lock_guard<std::mutex> lk(g_mutex);
  double tot = 0.0;
  for (int k = 0; k < g_timer.used; ++k) {
    PGAMD_HIP_CHECK(hipEventSynchronize(g_timer.ev[2 * k + 1]));
    float t = 0.f;
    PGAMD_HIP_CHECK(hipEventElapsedTime(&t, g_timer.ev[2 * k], g_timer.ev[2 * k + 1]));
    tot += t;
  }
  if (ms) *ms = tot;
  if (count) *count = g_timer.used;
  g_timer.on = false;
  g_timer.used = 0;
  return 0;
}

// ------------------------------------------------------ batched many clusters

namespace {
int g_cl_bpc = 0;   // persistent blocks per CU (0 = default 8)
int g_cl_tile = 0;  // j-particles per LDS tile (0 = default 256)
constexpr int kClTiles[3] = {128, 256, 512};
DevBuf g_cl_counter;

struct ClArgs {
  int nitems;
  const int2* items;
  const int* coff;
  const double *pos, *vel, *mass;
  const float* eps2;
  double *acc, *jrk, *pot;
};

template <int TILE>
void launch_cl_t(const ClArgs& a, int* counter, int blocks, hipStream_t st) {
  hipLaunchKernelGGL(pgc::cluster_force<TILE>, dim3(blocks), dim3(pgc::kThreads), 0, st,
                     a.nitems, a.items, a.coff, a.pos, a.vel, a.mass, a.eps2, a.acc, a.jrk, a.pot,
                     counter);
}

// Counter reset + persistent kernel on st; tile in kClTiles.
void launch_cl(const ClArgs& a, int* counter, int bpc, int tile, hipStream_t st) {
  hipLaunchKernelGGL(pgc::zero_counter, dim3(1), dim3(64), 0, st, counter);
  PGAMD_HIP_CHECK(hipGetLastError());
  if (a.nitems <= 0) return;
  const int blocks = std::max(1, std::min(a.nitems, device_cu_count() * bpc));
  switch (tile) {
    case 128: launch_cl_t<128>(a, counter, blocks, st); break;
    case 256: launch_cl_t<256>(a, counter, blocks, st); break;
    case 512: launch_cl_t<512>(a, counter, blocks, st); break;
    default:
      api_error(3, "cluster batch: tile %d not in {128, 256, 512}", tile);
      return;
  }
  PGAMD_HIP_CHECK(hipGetLastError());
}

// 8 per CU: 5.48 ms on the 4096-cluster workload against 5.54-5.60 at 4
int cl_bpc() { return g_cl_bpc > 0 ? g_cl_bpc : 8; }
int cl_tile() { return g_cl_tile > 0 ? g_cl_tile : pgc::kTile; }

struct ClGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  DevBuf counter;
};
}  // namespace

// Work items for pg6_cluster_force from host cluster offsets coff[ncl+1]:
// one item (cluster, first i) per chunk of up to 256 i-particles, ordered by
// decreasing cost (chunk size x cluster size) so the largest start first.
// Writes 2*count ints to items (if non-null, at most maxitems items);
// returns the number of items.
PGAMD_API int pg6_cluster_items(int ncl, const int* coff, int* items, int maxitems) {
  std::vector<std::pair<long long, std::pair<int, int>>> v;
  for (int c = 0; c < ncl; ++c) {
    const int n = coff[c + 1] - coff[c];
    for (int i0 = 0; i0 < n; i0 += pgc::kThreads)
      v.push_back({-(long long)std::min(pgc::kThreads, n - i0) * n, {c, i0}});
  }
  std::stable_sort(v.begin(), v.end(),
                   [](const auto& a, const auto& b) { return a.first < b.first; });
  if (items) {
    if ((int)v.size() > maxitems) {
      api_error(3, "pg6_cluster_items: %d items exceed maxitems=%d", (int)v.size(), maxitems);
      return -1;
    }
    for (size_t k = 0; k < v.size(); ++k) {
      items[2 * k] = v[k].second.first;
      items[2 * k + 1] = v[k].second.second;
    }
  }
  return (int)v.size();
}

// Acc, jerk and pot of every particle of many independent clusters in one
// persistent launch (cluster_batch.h).  Device arrays: coff[ncl+1],
// items[2*nitems] from pg6_cluster_items, pos/vel [n][3], mass[n] (FP64),
// eps2[ncl] (FP32); outputs acc/jrk [n][3], pot[n].  Asynchronous on stream.
PGAMD_API int pg6_cluster_force(int nitems, const int* items, const int* coff, const double* pos,
                                const double* vel, const double* mass, const float* eps2,
                                double* acc, double* jrk, double* pot, void* stream) {
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!board_check_open("pg6_cluster_force")) return -1;
  if (nitems < 0) {
    api_error(3, "pg6_cluster_force: nitems=%d", nitems);
    return -1;
  }
  g_cl_counter.reserve(256);
  launch_cl(ClArgs{nitems, reinterpret_cast<const int2*>(items), coff, pos, vel, mass, eps2, acc,
                   jrk, pot},
            g_cl_counter.as<int>(), cl_bpc(), cl_tile(),
            stream ? (hipStream_t)stream : g_board.stream);
  return 0;
}

PGAMD_API void* pg6_cluster_graph(int nitems, const int* items, const int* coff,
                                  const double* pos, const double* vel, const double* mass,
                                  const float* eps2, double* acc, double* jrk, double* pot) {
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!board_check_open("pg6_cluster_graph")) return nullptr;
  ClGraph* g = new ClGraph();
  g->counter.reserve(256);
  hipStream_t cs;
  PGAMD_HIP_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  PGAMD_HIP_CHECK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
  launch_cl(ClArgs{nitems, reinterpret_cast<const int2*>(items), coff, pos, vel, mass, eps2, acc,
                   jrk, pot},
            g->counter.as<int>(), cl_bpc(), cl_tile(), cs);
  PGAMD_HIP_CHECK(hipStreamEndCapture(cs, &g->graph));
  PGAMD_HIP_CHECK(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
  PGAMD_HIP_CHECK(hipStreamDestroy(cs));
  return g;
}

PGAMD_API int pg6_cluster_graph_launch(void* h, void* stream) {
  if (!h) {
    api_error(3, "pg6_cluster_graph_launch: null graph");
    return -1;
  }
  PGAMD_HIP_CHECK(hipGraphLaunch(static_cast<ClGraph*>(h)->exec,
                                 stream ? (hipStream_t)stream : g_board.stream));
  return 0;
}

PGAMD_API int pg6_cluster_graph_destroy(void* h) {
  if (!h) return 0;
  ClGraph* g = static_cast<ClGraph*>(h);
  PGAMD_HIP_CHECK(hipDeviceSynchronize());
  PGAMD_HIP_CHECK(hipGraphExecDestroy(g->exec));
  PGAMD_HIP_CHECK(hipGraphDestroy(g->graph));
  g->counter.release();
  delete g;
  return 0;
}

// Autotuner over j-tiles of 128, 256 and 512 particles and {1, 2, 3, 4, 6,
// 8, 12, 16} persistent blocks per CU on these inputs (one warm-up, mean of
// reps): keeps the fastest pair and returns its blocks per CU through
// *best; ms[8] (nullable) receives the timings at the tile kept, which
// pg6_cluster_launch_shape reports.  Results depend on the tile only through
// the order of the FP32 partial sums.
PGAMD_API int pg6_cluster_autotune(int nitems, const int* items, const int* coff,
                                   const double* pos, const double* vel, const double* mass,
                                   const float* eps2, double* acc, double* jrk, double* pot,
                                   int reps, int* best, float* ms) {
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!board_check_open("pg6_cluster_autotune")) return -1;
  if (nitems < 0) {
    api_error(3, "pg6_cluster_autotune: nitems=%d", nitems);
    return -1;
  }
  const ClArgs a{nitems, reinterpret_cast<const int2*>(items), coff, pos, vel, mass, eps2, acc,
                 jrk, pot};
  g_cl_counter.reserve(256);
  const int cand[8] = {1, 2, 3, 4, 6, 8, 12, 16};
  hipStream_t st = g_board.stream;
  hipEvent_t e0, e1;
  PGAMD_HIP_CHECK(hipEventCreate(&e0));
  PGAMD_HIP_CHECK(hipEventCreate(&e1));
  float bestt = 1e30f, t_all[3][8];
  int bestc = cl_bpc(), bestkt = 1;
  while (bestkt < 2 && kClTiles[bestkt] != cl_tile()) ++bestkt;
  reps = std::max(1, reps);
  for (int kt = 0; kt < 3; ++kt)
    for (int k = 0; k < 8; ++k) {
      launch_cl(a, g_cl_counter.as<int>(), cand[k], kClTiles[kt], st);
      PGAMD_HIP_CHECK(hipEventRecord(e0, st));
      for (int r = 0; r < reps; ++r)
        launch_cl(a, g_cl_counter.as<int>(), cand[k], kClTiles[kt], st);
      PGAMD_HIP_CHECK(hipEventRecord(e1, st));
      PGAMD_HIP_CHECK(hipEventSynchronize(e1));
      float t = 0.f;
      PGAMD_HIP_CHECK(hipEventElapsedTime(&t, e0, e1));
      t /= reps;
      t_all[kt][k] = t;
      if (t < bestt) {
        bestt = t;
        bestc = cand[k];
        bestkt = kt;
      }
    }
  PGAMD_HIP_CHECK(hipEventDestroy(e0));
  PGAMD_HIP_CHECK(hipEventDestroy(e1));
  g_cl_bpc = bestc;
  g_cl_tile = kClTiles[bestkt];
  if (ms)
    for (int k = 0; k < 8; ++k) ms[k] = t_all[bestkt][k];
  if (best) *best = bestc;
  return 0;
}

// Launch shape of the cluster batch: tile = j-particles per LDS tile (128,
// 256 or 512; 0 = default 256), blocks_per_cu (0 = default 8); a negative
// value leaves that part as it is.
PGAMD_API int pg6_cluster_set_launch(int tile, int blocks_per_cu) {
  std::lock_guard<std::mutex> lk(g_mutex);
  if (tile > 0 && tile != 128 && tile != 256 && tile != 512) {
    api_error(3, "pg6_cluster_set_launch: tile %d not in {0, 128, 256, 512}", tile);
    return -1;
  }
  if (blocks_per_cu > 64) {
    api_error(3, "pg6_cluster_set_launch: blocks_per_cu %d > 64", blocks_per_cu);
    return -1;
  }
  if (tile >= 0) g_cl_tile = tile;
  if (blocks_per_cu >= 0) g_cl_bpc = blocks_per_cu;
  return 0;
}

// Launch shape the cluster batch uses now: out2 = {j-tile, persistent blocks
// per CU}.
PGAMD_API int pg6_cluster_launch_shape(int* out2) {
  if (!out2) {
    api_error(3, "pg6_cluster_launch_shape: null output");
    return -1;
  }
  std::lock_guard<std::mutex> lk(g_mutex);
  out2[0] = cl_tile();
  out2[1] = cl_bpc();
  return 0;
}

// Force kernel choice: 0 = j-particles staged in LDS tiles (pg6_force),
// 1 = j-particles broadcast through scalar registers (pg6_force_s), 2 = as 1
// with two-float positions and the chunked self-pair filter (pg6_force_h),
// 3 = prediction, force and j-range reduction in one launch (pg6_force_f, up
// to kFusedMaxNi i; more fall back to 2), 4 = j-particles in Morton order as
// FP32 offsets from their chunk's origin (pg6_force_o; its dx error scales
// with the chunk extent, so the automatic choice takes it only from
// kSortedMinNj j up; set here it runs at any nj), 5 = as 4 with 64 i per
// block and the block's j-range split over its bi / 64 wavefronts (fewer
// partial sums; pg6_force_ow), -1 = the autotuned or default choice for each
// i-count.  0 and 1 compute the same arithmetic; 2 and 3 form dx in FP32
// (within about one FP32 ulp of the others).
PGAMD_API int pg6_dev_set_kernel(int variant) {
  std::lock_guard<std::mutex> lk(g_mutex);
  if (variant < -1 || variant > 5) {
    api_error(3, "pg6_dev_set_kernel: variant %d not in {-1, 0, 1, 2, 3, 4, 5}", variant);
    return -1;
  }
  g_board.force_variant = variant;
  return 0;
}

namespace {
// Gathers i-particle arrays (AoS FP64 position and velocity, id) from the
// first n predicted j-particles, for autotuning on the caller's own data.
__global__ void pg6_gather_i(int n, JPred jp, double* pos, double* vel, int* id) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  pos[3 * k + 0] = jp.x[k];
  pos[3 * k + 1] = jp.y[k];
  pos[3 * k + 2] = jp.z[k];
  const float4 vm = jp.vm[k];
  vel[3 * k + 0] = vm.x;
  vel[3 * k + 1] = vm.y;
  vel[3 * k + 2] = vm.z;
  id[k] = jp.id[k];
}
}  // namespace

// Autotuner for the direct force: for each i-count bucket ni = 64, 128, ...,
// min(65536, nj), times every (kernel, bi, blocks_per_cu) candidate on the
// first nj j-particles (i-set = the first ni of them; prediction at the
// board's time, force, jerk, potential and nearest neighbour, j-range
// reduction included) and keeps the
// fastest for that bucket (the default shape unless one is more than 1 %
// faster); plan_launch uses it for every ni in the bucket.
// Call after pg6_dev_predict (or g6calc_firsthalf) so the predicted set is
// representative.  out[4 * k .. 4k+3] = {variant, bi, bpc, microseconds} per
// bucket k (may be NULL); returns the number of buckets tuned.  Explicit
// overrides (pg6_dev_set_kernel, pg6_dev_set_launch) still take precedence.
PGAMD_API int pg6_dev_autotune(int nj, double eps2, int reps, int* out) {
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!board_check_open("pg6_dev_autotune")) return -1;
  Board& b = g_board;
  if (nj < 64 || nj > b.cap) {
    api_error(3, "pg6_dev_autotune: nj=%d must be in [64, %d]", nj, b.cap);
    return -1;
  }
  board_flush_j(b);
  if (board_needs_pred(b, nj)) board_predict(b, nj, b.stream);
  if (reps <= 0) reps = 3;
  const int nimax = std::min(nj, 65536);
  DevBuf buf;
  // per i: pos[3], vel[3], acc[3], jrk[3], pot (FP64); id, nnb (int); rnnb
  const size_t per_i = 13 * sizeof(double) + 2 * sizeof(int) + sizeof(float);
  buf.reserve((size_t)nimax * per_i);
  double* ipos = buf.as<double>();
  double* ivel = ipos + 3 * (size_t)nimax;
  double* acc = ivel + 3 * (size_t)nimax;
  double* jrk = acc + 3 * (size_t)nimax;
  double* pot = jrk + 3 * (size_t)nimax;
  int* iid = reinterpret_cast<int*>(pot + (size_t)nimax);
  int* nnb = iid + nimax;
  float* rnnb = reinterpret_cast<float*>(nnb + nimax);
  hipStream_t st = b.stream;
  hipLaunchKernelGGL(pg6_gather_i, dim3(ceil_div(nimax, 256)), dim3(256), 0, st, nimax, b.jp, ipos,
                     ivel, iid);
  PGAMD_HIP_CHECK(hipGetLastError());
  const Outputs o{acc, jrk, pot, nnb, rnnb};
  hipEvent_t e0, e1;
  PGAMD_HIP_CHECK(hipEventCreate(&e0));
  PGAMD_HIP_CHECK(hipEventCreate(&e1));
  const bool timer_was_on = g_timer.on;
  g_timer.on = false;
  const int save_bi = b.force_bi, save_bpc = b.force_bpc, save_var = b.force_variant;
  static const int kBpc[] = {1, 2, 4, 8, 16, 32, 64};
  static const int kBpcF[] = {1, 2, 4, 8, 16};  // variant 3: 64 .. 1024 blocks
  int tuned = 0;
  for (int k = 0; k < kBuckets && (64 << k) <= nimax; ++k) {
    const int ni = 64 << k;
    float best = 1e30f;
    const LaunchChoice dc = default_choice(ni);
    LaunchChoice bc = dc;
    float t_def = 1e30f;  // the default's time, if it is a candidate here
    // Every timed call predicts the j-set first, as a block step does: the
    // kernels differ in what a prediction costs them (the sorted layout's
    // records, the one-launch kernel's in-kernel prediction).
    for (int var = 0; var <= 5; ++var)
      for (int bi : {64, 128, 256})
        for (int bk = 0; bk < 7; ++bk) {
          if (var == 3 && (bi != 64 || ni > kFusedMaxNi || bk >= 5)) continue;
          const int bpc = var == 3 ? kBpcF[bk] : kBpc[bk];
          if ((var == 4 || var == 5) && nj < kSortedMinNj) continue;
          b.force_variant = var;
          b.force_bi = bi;
          b.force_bpc = bpc;
          // the positions do not change here: no rebuild of the sorted
          // layout inside a timed run
          b.sort_age = 0;
          b.predict_flag = 1;
          board_force(b, ni, nj, ipos, ivel, iid, eps2, o, st);  // warm-up
          PGAMD_HIP_CHECK(hipEventRecord(e0, st));
          for (int r = 0; r < reps; ++r) {
            b.predict_flag = 1;
            board_force(b, ni, nj, ipos, ivel, iid, eps2, o, st);
          }
          PGAMD_HIP_CHECK(hipEventRecord(e1, st));
          PGAMD_HIP_CHECK(hipEventSynchronize(e1));
          float ms = 0.f;
          PGAMD_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
          if (var == dc.variant && bi == dc.bi && bpc == dc.bpc) t_def = ms;
          if (ms < best) {
            best = ms;
            bc = {var, bi, bpc};
          }
        }
    // keep the default unless a candidate beats it by more than 1 % (run-to-run
    // noise would otherwise pick a different shape on every call)
    if (t_def <= 1.01f * best) {
      bc = dc;
      best = t_def;
    }
    b.tuned[k] = bc;
    b.tuned_ok[k] = true;
    // the largest bucket also serves every larger ni
    if (out) {
      out[4 * k + 0] = bc.variant;
      out[4 * k + 1] = bc.bi;
      out[4 * k + 2] = bc.bpc;
      out[4 * k + 3] = (int)(1000.f * best / reps);
    }
    ++tuned;
  }
  for (int k = tuned; k < kBuckets; ++k) {
    b.tuned[k] = b.tuned[tuned - 1];
    b.tuned_ok[k] = true;
  }
  b.force_bi = save_bi;
  b.force_bpc = save_bpc;
  b.force_variant = save_var;
  // the timed calls predicted the same j-state to the same ti again (some
  // into the predicted set, some inside their launch): the set predicted
  // before them still stands
  b.predict_flag = 0;
  b.pred_armed = false;
  b.pred_virtual = false;
  b.nj_pred = std::max(b.nj_pred, nj);
  g_timer.on = timer_was_on;
  PGAMD_HIP_CHECK(hipEventDestroy(e0));
  PGAMD_HIP_CHECK(hipEventDestroy(e1));
  buf.release();
  return tuned;
}

// Forgets the autotuned choices (back to the defaults).
PGAMD_API int pg6_dev_autotune_reset(void) {
  std::lock_guard<std::mutex> lk(g_mutex);
  for (int k = 0; k < kBuckets; ++k) g_board.tuned_ok[k] = false;
  return 0;
}

// ======================================================================
// Block-time-step Hermite integration on the GPU (the reference's sample
// integrator PG6/banana, banana01.c:57-166, on the device).  It owns the
// board: its particles are the j-memory (address = id = index).
namespace {

struct Hermite {
  bool on = false;
  int n = 0;
  double eps2 = 0.0;
  pg6h::Params P{};
  DevBuf state, work;
  pg6h::State S{};
  pg6h::Ctl* ctl = nullptr;  // device
  // work arrays: active list and its predicted i data, force partials
  int* act = nullptr;
  double *ipos = nullptr, *ivel = nullptr, *acc1 = nullptr, *jrk1 = nullptr, *pot1 = nullptr;
  int* iid = nullptr;
  double* epart = nullptr;  // energy partial sums, 2 per block
  Partials part{};
  int target = 0, min_jper = 16;  // force split (pg6h_plan)
  int min_blocks = 0, force_blocks = 0, finish_blocks = 0;
  HostBuf host;  // pinned: a Ctl image, then t_end
  // a captured batch of kGraphSteps steps, and the board arrays it was captured with
  hipGraphExec_t graph = nullptr;
  const void* graph_key[2] = {nullptr, nullptr};
};

constexpr int kGraphSteps = 32;

Hermite g_h;

constexpr int kHermBlocks = 256;  // blocks of the energy reduction

void hermite_release(Hermite& h) {
  if (h.graph) PGAMD_HIP_CHECK(hipGraphExecDestroy(h.graph));
  h.state.release();
  h.work.release();
  h.host.release();
  h = Hermite();
}

// Runs the force on the first `ni` entries of ipos/ivel/iid against all n
// j-particles (predicted to `ti` beforehand) into acc1/jrk1/pot1.
void hermite_force(Board& b, Hermite& h, int ni) {
  const Outputs o{h.acc1, h.jrk1, h.pot1, nullptr, nullptr};
  board_force(b, ni, h.n, h.ipos, h.ivel, h.iid, h.eps2, o, b.stream);
}

// Queues one block step (three kernels, no host synchronisation).
void hermite_enqueue(Board& b, Hermite& h, hipStream_t st) {
  const int n = h.n;
  board_use_stream(b, st);
  hipLaunchKernelGGL(pg6h::pg6h_select, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, h.S, h.P,
                     h.ctl, b.js, b.jp, h.act, h.ipos, h.ivel, h.iid);
  hipLaunchKernelGGL(pg6h::pg6h_force, dim3(h.force_blocks), dim3(64 * pg6h::kForceWaves), 0, st,
                     n, (float)h.eps2,
                     h.P, h.ctl, h.target, h.min_jper, h.ipos, h.ivel, h.iid, b.jp, h.part);
  hipLaunchKernelGGL(pg6h::pg6h_finish, dim3(h.finish_blocks), dim3(256), 0, st, n, h.P, h.ctl,
                     h.target, h.min_jper, h.act, h.S, h.ipos, h.ivel, h.part, b.js);
  PGAMD_HIP_CHECK(hipGetLastError());
  b.nj_pred = -1;  // the board's predictions now belong to the integrator
  b.sort_age = kSortAge;  // its corrector moves every particle: rebuild the order
  b.predict_flag = 1;
  b.pred_virtual = false;
}

// Queues kGraphSteps block steps as one captured graph (captured on first
// use, again when the board's arrays have moved).  The steps are the same
// kernels with the same arguments as hermite_enqueue's; none clears memory.
void hermite_enqueue_batch(Board& b, Hermite& h, hipStream_t st) {
  if (h.graph && (h.graph_key[0] != b.js.x || h.graph_key[1] != b.jp.x)) {
    PGAMD_HIP_CHECK(hipGraphExecDestroy(h.graph));
    h.graph = nullptr;
  }
  board_use_stream(b, st);  // before the capture: inside it the stream is already current
  if (!h.graph) {
    hipGraph_t g = nullptr;
    PGAMD_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
    for (int k = 0; k < kGraphSteps; ++k) hermite_enqueue(b, h, st);
    PGAMD_HIP_CHECK(hipStreamEndCapture(st, &g));
    PGAMD_HIP_CHECK(hipGraphInstantiate(&h.graph, g, nullptr, nullptr, 0));
    PGAMD_HIP_CHECK(hipGraphDestroy(g));
    h.graph_key[0] = b.js.x;
    h.graph_key[1] = b.jp.x;
  }
  PGAMD_HIP_CHECK(hipGraphLaunch(h.graph, st));
  b.nj_pred = -1;
  b.sort_age = kSortAge;
  b.predict_flag = 1;
  b.pred_virtual = false;
}

// Sets Ctl::t_end (through the pinned buffer; the caller synchronises before
// the buffer is written again).
void hermite_set_end(Hermite& h, double t_end, hipStream_t st) {
  double* p = reinterpret_cast<double*>(h.host.as<char>() + 256);
  *p = t_end;
  PGAMD_HIP_CHECK(hipMemcpyAsync(&h.ctl->t_end, p, sizeof(double), hipMemcpyHostToDevice, st));
}

// Copies Ctl to the host (synchronous).
pg6h::Ctl hermite_read(Hermite& h, hipStream_t st) {
  pg6h::Ctl* c = h.host.as<pg6h::Ctl>();
  PGAMD_HIP_CHECK(hipMemcpyAsync(c, h.ctl, sizeof(pg6h::Ctl), hipMemcpyDeviceToHost, st));
  PGAMD_HIP_CHECK(hipStreamSynchronize(st));
  return *c;
}

double hermite_time(const Hermite& h, const pg6h::Ctl& c) {
  return (double)c.t_mod * h.P.dtmax + c.t_gl;
}

}  // namespace

// Starts an integration of n particles (host arrays x[n][3], v[n][3], m[n],
// FP64) at t = 0: uploads them as the j-particles, computes every force, jerk
// and potential, and the first block time steps (dt_startup with eta_s;
// timestep.c:24-48).  eps2 = softening^2 (0 allowed), eta = accuracy of the
// corrector's step criterion (timestep.c:50-79), dtmax = largest step (also
// the window after which times restart at 0).  Returns 0, or -1 on bad
// arguments or when a first step would fall below dtmax / 2^51.
PGAMD_API int pg6_hermite_init(int n, const double* x, const double* v, const double* m,
                               double eps2, double eta_s, double eta, double dtmax) {
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!board_check_open("pg6_hermite_init")) return -1;
  if (n < 2 || !x || !v || !m || !(eps2 >= 0.0) || !(eta_s > 0.0) || !(eta > 0.0) ||
      !(dtmax > 0.0)) {
    api_error(3, "pg6_hermite_init: n=%d eps2=%g eta_s=%g eta=%g dtmax=%g", n, eps2, eta_s, eta,
              dtmax);
    return -1;
  }
  Board& b = g_board;
  if (!board_reserve(b, n)) return -1;
  hermite_release(g_h);
  Hermite& h = g_h;
  h.n = n;
  h.eps2 = eps2;
  h.P.eta_s = eta_s;
  h.P.eta = eta;
  h.P.dtmax = dtmax;
  h.P.dtmin = dtmax / std::pow(2.0, 51.0);  // initialize_timestep (timestep.c:18-22)
  const int cus = device_cu_count();
  h.target = 16 * cus;  // wavefront items per force launch: 4 per SIMD
  h.min_blocks = std::min(ceil_div(n, 256), 4 * cus);
  h.force_blocks = 4 * cus * 4 / pg6h::kForceWaves;  // 4 wavefronts per SIMD
  // blocks of 256 threads: a wavefront per active particle of a typical block
  // step (a few hundred), and few blocks to count in Ctl::done
  h.finish_blocks = std::max(1, std::min(ceil_div(n, 256), cus));
  // largest force split: ns * ntiles <= target / kForceWaves + ntiles (pg6h_plan)
  const size_t nsplit_i = ((size_t)h.target / pg6h::kForceWaves + ceil_div(n, 64)) * 64;
  // state: x, v, a, j (3n each), pot, t, dt, m (n each)
  h.state.reserve(sizeof(double) * (size_t)n * 16);
  double* d = h.state.as<double>();
  h.S.x = d;
  h.S.v = d + 3 * (size_t)n;
  h.S.a = d + 6 * (size_t)n;
  h.S.j = d + 9 * (size_t)n;
  h.S.pot = d + 12 * (size_t)n;
  h.S.t = d + 13 * (size_t)n;
  h.S.dt = d + 14 * (size_t)n;
  h.S.m = d + 15 * (size_t)n;
  // work: ipos, ivel, acc1, jrk1 (3n each), pot1 (n), epart (2 * blocks), partials (7 *
  // nsplit_i), Ctl; act, iid (n each)
  const size_t dbytes =
      sizeof(double) * ((size_t)n * 13 + 2 * kHermBlocks + 7 * nsplit_i) + 256;
  const size_t ibytes = sizeof(int) * ((size_t)n * 2);
  h.work.reserve(dbytes + ibytes);
  double* w = h.work.as<double>();
  h.ipos = w;
  h.ivel = w + 3 * (size_t)n;
  h.acc1 = w + 6 * (size_t)n;
  h.jrk1 = w + 9 * (size_t)n;
  h.pot1 = w + 12 * (size_t)n;
  h.epart = w + 13 * (size_t)n;
  double* pp = h.epart + 2 * kHermBlocks;
  double** pf[7] = {&h.part.ax, &h.part.ay, &h.part.az, &h.part.pot,
                    &h.part.jx, &h.part.jy, &h.part.jz};
  for (int k = 0; k < 7; ++k) *pf[k] = pp + (size_t)k * nsplit_i;
  h.part.rmax = nullptr;
  h.part.nn = nullptr;
  static_assert(sizeof(pg6h::Ctl) <= 256, "Ctl fits its slot");
  h.ctl = reinterpret_cast<pg6h::Ctl*>(pp + 7 * nsplit_i);
  int* wi = reinterpret_cast<int*>(reinterpret_cast<char*>(h.ctl) + 256);
  h.act = wi;
  h.iid = wi + n;
  h.host.reserve(512);
  hipStream_t st = b.stream;
  // upload x, v, m; t = 0, a = j = 0
  PGAMD_HIP_CHECK(hipMemcpyAsync(h.S.x, x, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st));
  PGAMD_HIP_CHECK(hipMemcpyAsync(h.S.v, v, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st));
  PGAMD_HIP_CHECK(hipMemcpyAsync(h.S.m, m, sizeof(double) * n, hipMemcpyHostToDevice, st));
  PGAMD_HIP_CHECK(hipMemsetAsync(h.S.a, 0, sizeof(double) * 6 * (size_t)n, st));
  PGAMD_HIP_CHECK(hipMemsetAsync(h.S.t, 0, sizeof(double) * (size_t)n, st));
  PGAMD_HIP_CHECK(hipMemsetAsync(h.S.dt, 0, sizeof(double) * (size_t)n, st));
  // j-memory = the particles at t = 0 (set_particle_on_grape, banana01.c:206-236):
  // the start-up kernel below writes it; the i-set is every particle as is.
  PGAMD_HIP_CHECK(hipMemcpyAsync(h.ipos, x, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st));
  PGAMD_HIP_CHECK(hipMemcpyAsync(h.ivel, v, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st));
  std::vector<int> ids(n);
  for (int i = 0; i < n; ++i) ids[i] = i;
  PGAMD_HIP_CHECK(hipMemcpyAsync(h.iid, ids.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
  // control block: t = 0, n particle steps and 1 block step (the start-up)
  pg6h::Ctl* c = h.host.as<pg6h::Ctl>();
  *c = pg6h::Ctl{};
  c->tmin = pg6h::kInfBits;
  c->tmin_next = pg6h::kInfBits;
  c->t_end = HUGE_VAL;
  c->nsteps = n;
  c->nblocks = 1;
  PGAMD_HIP_CHECK(hipMemcpyAsync(h.ctl, c, sizeof(pg6h::Ctl), hipMemcpyHostToDevice, st));
  // j-state of the board from the particles (a = j = 0, t = 0), predicted to 0
  b.npend = 0;
  std::fill(b.slot.begin(), b.slot.end(), -1);
  hipLaunchKernelGGL(pg6h::pg6h_store_all, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, h.S,
                     b.js);
  PGAMD_HIP_CHECK(hipGetLastError());
  b.ti = 0.0;
  board_predict(b, n, st);
  // start-up forces of every particle, then the first steps (startup_sgmotion)
  hermite_force(b, h, n);
  hipLaunchKernelGGL(pg6h::pg6h_startup, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, h.S, h.P,
                     h.acc1, h.jrk1, h.pot1, b.js, &h.ctl->bad);
  hipLaunchKernelGGL(pg6h::pg6h_min_time, dim3(h.min_blocks), dim3(256), 0, st, n, h.S, h.ctl);
  PGAMD_HIP_CHECK(hipGetLastError());
  const pg6h::Ctl r = hermite_read(h, st);
  b.predict_flag = 1;
  b.pred_virtual = false;
  b.nj_pred = -1;
  h.on = true;
  if (r.bad) {
    api_error(6, "pg6_hermite_init: %d particles have no admissible first step", r.bad);
    return -1;
  }
  return 0;
}

// One block step (global_time + integrate_exmotion + correct_sgmotion of
// banana01.c:120-166): the particles with the smallest t + dt advance to it.
// After a step that ends a dtmax window every t restarts at 0 (reset_time).
// *t_out = t_mod * dtmax + t_gl after the step, *nact = active particles.
PGAMD_API int pg6_hermite_step(double* t_out, int* nact_out) {
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!board_check_open("pg6_hermite_step")) return -1;
  Hermite& h = g_h;
  if (!h.on) {
    api_error(1, "pg6_hermite_step: no integration started (pg6_hermite_init)");
    return -1;
  }
  Board& b = g_board;
  hipStream_t st = b.stream;
  hermite_set_end(h, HUGE_VAL, st);
  hermite_enqueue(b, h, st);
  const pg6h::Ctl c = hermite_read(h, st);
  if (c.last_nact <= 0 || c.last_nact > h.n) {
    api_error(6, "pg6_hermite_step: no active particle");
    return -1;
  }
  if (t_out) *t_out = hermite_time(h, c);
  if (nact_out) *nact_out = c.last_nact;
  return 0;
}

// Block steps until t = t_mod * dtmax + t_gl reaches t_end (finite).  Steps
// are queued in growing batches without waiting for each other; those queued
// past t_end find it reached on the device and do nothing.  Returns the
// number of block steps taken, or -1.
PGAMD_API long long pg6_hermite_run(double t_end) {
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!board_check_open("pg6_hermite_run")) return -1;
  Hermite& h = g_h;
  if (!h.on) {
    api_error(1, "pg6_hermite_run: no integration started (pg6_hermite_init)");
    return -1;
  }
  if (!std::isfinite(t_end)) {
    api_error(3, "pg6_hermite_run: t_end = %g is not finite", t_end);
    return -1;
  }
  Board& b = g_board;
  hipStream_t st = b.stream;
  pg6h::Ctl c = hermite_read(h, st);
  const long long blocks0 = c.nblocks;
  hermite_set_end(h, t_end, st);
  int batch = 1;  // graphs of kGraphSteps steps
  while (hermite_time(h, c) < t_end) {
    for (int k = 0; k < batch; ++k) hermite_enqueue_batch(b, h, st);
    c = hermite_read(h, st);
    batch = std::min(2 * batch, 8);
  }
  return c.nblocks - blocks0;
}

// Copies the integrator state to host arrays (each may be null): x, v, a, j
// [n][3]; pot, t, dt [n] (t inside the current window); info[4] =
// {t_mod, n_steps (particle steps), n_blocks, n}.  Synchronous.
PGAMD_API int pg6_hermite_state(double* x, double* v, double* a, double* j, double* pot,
                                double* t, double* dt, long long* info) {
  std::lock_guard<std::mutex> lk(g_mutex);
  Hermite& h = g_h;
  if (!h.on) {
    api_error(1, "pg6_hermite_state: no integration started (pg6_hermite_init)");
    return -1;
  }
  hipStream_t st = g_board.stream;
  const size_t n = (size_t)h.n;
  auto get = [&](double* out, const double* src, size_t cnt) {
    if (out)
      PGAMD_HIP_CHECK(hipMemcpyAsync(out, src, sizeof(double) * cnt, hipMemcpyDeviceToHost, st));
  };
  get(x, h.S.x, 3 * n);
  get(v, h.S.v, 3 * n);
  get(a, h.S.a, 3 * n);
  get(j, h.S.j, 3 * n);
  get(pot, h.S.pot, n);
  get(t, h.S.t, n);
  get(dt, h.S.dt, n);
  const pg6h::Ctl c = hermite_read(h, st);
  if (info) {
    info[0] = c.t_mod;
    info[1] = c.nsteps;
    info[2] = c.nblocks;
    info[3] = h.n;
  }
  return 0;
}

// Kinetic and potential energy, 0.5 sum m v^2 and 0.5 sum m pot (the
// reference's calc_energy.c:8-33; meaningful when every particle is
// synchronised, i.e. at t_gl = 0).
PGAMD_API int pg6_hermite_energy(double* ek, double* ep) {
  std::lock_guard<std::mutex> lk(g_mutex);
  Hermite& h = g_h;
  if (!h.on) {
    api_error(1, "pg6_hermite_energy: no integration started (pg6_hermite_init)");
    return -1;
  }
  hipStream_t st = g_board.stream;
  hipLaunchKernelGGL(pg6h::pg6h_energy_part, dim3(kHermBlocks), dim3(256), 0, st, h.n, h.S,
                     h.epart);
  PGAMD_HIP_CHECK(hipGetLastError());
  std::vector<double> p(2 * kHermBlocks);
  PGAMD_HIP_CHECK(hipMemcpyAsync(p.data(), h.epart, sizeof(double) * p.size(),
                                 hipMemcpyDeviceToHost, st));
  PGAMD_HIP_CHECK(hipStreamSynchronize(st));
  double k = 0.0, w = 0.0;
  for (int b = 0; b < kHermBlocks; ++b) {
    k += p[2 * b];
    w += p[2 * b + 1];
  }
  if (ek) *ek = 0.5 * k;
  if (ep) *ep = 0.5 * w;
  return 0;
}

PGAMD_API int pg6_hermite_end(void) {
  std::lock_guard<std::mutex> lk(g_mutex);
  hermite_release(g_h);
  return 0;
}
