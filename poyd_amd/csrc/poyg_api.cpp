// C ABI of the engine: contexts, plans and the per-call entry points.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <numeric>
#include <unordered_map>
#include <vector>

#include "poyg_band.h"
#include "poyg_ukk.h"
#include "poyg_internal.h"

namespace poyg {
hipError_t launch_reset(int *counters, int n, hipStream_t s);
hipError_t launch_plan_start(int *ctr, int nctr, int *status, int nstatus, TraceOut *tro, uint8_t *out,
                             int32_t *out_len, hipStream_t s);
hipError_t launch_collect_results(const uint32_t *small, int64_t small_words, const int32_t *status,
                                  const uint32_t *out, int64_t out_words, uint32_t *host, hipStream_t s);
hipError_t launch_prep(bool affine, int grid, const PairMeta *meta, int n, const uint8_t *seq, const DevTcm *tcm,
                       int *aux, hipStream_t s);
hipError_t launch_aff_fill(bool bt, bool glb, int grid, const PairMeta *meta, const int32_t *order, int n,
                           int *counter, const uint8_t *seq, const DevTcm *tcm, uint8_t *dm, int *aux,
                           int32_t *cost, int *rowscratch, int row_width, hipStream_t s);
int aff_diag_blocks_per_cu(bool bt, int sets);
int aff_quad_blocks_per_cu(bool bt);
hipError_t launch_pair_pure(int grid, const PairMeta *meta, const int32_t *order, int n, const uint8_t *seq,
                            uint8_t *pure, hipStream_t s);
hipError_t launch_aff_quad(bool bt, int grid, const PairMeta *meta, const int32_t *order, int n, int *counter,
                           const uint8_t *seq, const DevTcm *tcm, uint8_t *dm, int32_t *cost, uint32_t *sink,
                           const uint8_t *pure, const TraceOut *tro, const int *n_dev, hipStream_t s);
hipError_t launch_aff_pk(bool bt, int grid, const PairMeta *meta, const int32_t *order, int n, int *counter,
                         const uint8_t *seq, const DevTcm *tcm, uint8_t *dm, int32_t *cost, uint32_t *sink,
                         const uint8_t *pure, int32_t *redo, int *redo_n, int32_t *tlist, int *tlist_n,
                         hipStream_t s);
int aff_pk_blocks_per_cu(bool bt);
int aff_trace_list_blocks_per_cu();
hipError_t launch_aff_trace_list(int grid, const PairMeta *meta, const int32_t *list, const int *list_n, int *counter,
                                 const uint8_t *seq, const DevTcm *tcm, const uint8_t *dm, const TraceOut *tro,
                                 hipStream_t s);
hipError_t launch_aff_diag(bool bt, int sets, int grid, const PairMeta *meta, const int32_t *order, int n,
                           int *counter, const uint8_t *seq, const DevTcm *tcm, uint8_t *dm, int32_t *cost,
                           uint32_t *sink, const TraceOut *tro, hipStream_t s);
hipError_t launch_aff_trace(const PairMeta *meta, const int32_t *order, int n, const uint8_t *seq,
                            const DevTcm *tcm, const uint8_t *dm, uint8_t *out, int32_t *out_len, int32_t *status,
                            hipStream_t s);
hipError_t launch_lin_sys(bool with_dm, int grid, const PairMeta *meta, const int32_t *order, int n, int *counter,
                          const uint8_t *seq, const DevTcm *tcm, uint8_t *dm, const int *aux, int32_t *cost,
                          int32_t *status, int spin_limit, hipStream_t s);
int lin_sys_blocks_per_cu(bool with_dm);
int lin_cost16_blocks_per_cu(bool with_dm, int n_max);   // resident waves per CU
int lin_c16mw_blocks_per_cu(bool with_dm, int n_max);
hipError_t launch_lin_c16mw(bool with_dm, int grid, const PairMeta *meta, const int32_t *order, int n, int *counter,
                            const uint8_t *seq, const DevTcm *tcm, const int *aux, int32_t *cost, uint8_t *dm,
                            int n_max, hipStream_t s);
hipError_t launch_lin_cost16(bool with_dm, int grid_waves, const PairMeta *meta, const int32_t *order, int n,
                             int *counter, const uint8_t *seq, const DevTcm *tcm, const int *aux, int32_t *cost,
                             uint8_t *dm, int n_max, hipStream_t s);
hipError_t launch_lin_fill(bool with_dm, bool glb, int grid, const PairMeta *meta, const int32_t *order, int n,
                           int *counter, const uint8_t *seq, const DevTcm *tcm, uint8_t *dm, int *aux,
                           int32_t *cost, int *rowscratch, int row_width, hipStream_t s);
int lin_trace_blocks_per_cu();
hipError_t launch_lin_trace(int grid, const PairMeta *meta, const int32_t *order, int n, int *counter,
                            const uint8_t *seq, const DevTcm *tcm, const uint8_t *dm, uint8_t *out, int32_t *out_len,
                            int32_t *status, hipStream_t s);
hipError_t launch_fitch_down(int width, uint8_t *chars, int len, int n, const int32_t *ia, const int32_t *ib,
                             const int32_t *ires, int64_t *added, hipStream_t s);
hipError_t launch_fitch_up(int width, uint8_t *chars, int len, int n, const int32_t *iA, const int32_t *iN,
                           const int32_t *iD1, const int32_t *iD2, const int32_t *iF, hipStream_t s);
}  // namespace poyg

extern "C" {
int poyg_triple_plan_create(poyg_ctx *ctx, const poyg_cm3 *cm3, int32_t n, const int64_t *off_1,
                            const int32_t *len_1, const int64_t *off_2, const int32_t *len_2,
                            const int64_t *off_3, const int32_t *len_3, int device, int cus, void *stream,
                            void *side, void *side2, int64_t dm_budget, int waves_per_cu, void **out);
void poyg_triple_plan_destroy(void *p);
int64_t poyg_triple_plan_out_bytes(void *p);
int64_t poyg_triple_plan_cells(void *p);
int poyg_triple_plan_out_offsets(void *p, int64_t *o);
int poyg_triple_plan_run(void *pp, const uint8_t *d_seq, int32_t *d_cost, uint8_t *d_out, int32_t *d_out_len,
                         void *stream, float *fill_ms, float *trace_ms);
int poyg_triple_plan_status(void *pp, int32_t *bad);
}

using namespace poyg;

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) {                                                        \
      set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
      return POYG_EDEVICE;                                                         \
    }                                                                              \
  } while (0)

// Launch-shape knobs a tuner may set per context (-1: the built-in choice).
// lin_cost_kernel: 1 = the 32-bit systolic fill for every linear pair that
// fits it; the packed 16-bit fill wherever a pair allows it with 2 = one
// pair per wave for large batches and four waves per pair for small ones
// (built-in), 3 = one pair per wave, 4 = four waves per pair.
// aff_narrow_kernel: 1 = the 32-bit four-pair fill for the narrow affine
// bands, 2 = the packed eight-pair fill (built-in).
enum TuneKnob : int {
  kTuneQuadMaxLanes, kTuneSideDiv, kTuneSysBlocks, kTuneTraceBlocks, kTuneCubeWaves, kTuneLinCost, kTuneAffNarrow,
  kTuneKnobs
};
static const char *const kTuneNames[kTuneKnobs] = {"aff_quad_max_lanes", "aff_side_div", "lin_sys_blocks_per_cu",
                                                   "lin_trace_blocks_per_cu", "cube_waves_per_cu",
                                                   "lin_cost_kernel", "aff_narrow_kernel"};

struct poyg_ctx {
  int device = 0;
  int cus = 256;
  int tune[kTuneKnobs] = {-1, -1, -1, -1, -1, -1, -1};
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;    // second stream: fills beside the main one
  hipStream_t side2 = nullptr;   // third stream: the widest bands (lane sets, row kernels)
  hipStream_t capture = nullptr;   // the stream of a graph capture (created at the first one)
  // pinned host staging of the per-call entry points (one call at a time per
  // context: every call ends with a synchronization of `stream`): the plan's
  // descriptor block, and the call's sequences and results
  uint8_t *pin_plan = nullptr, *pin_io = nullptr;
  size_t pin_plan_cap = 0, pin_io_cap = 0;
  void *pin_io_dev = nullptr;   // pin_io's address for kernels (they write the results there)
};

// Grows a pinned staging buffer of the context (its contents are not kept);
// dev, if given, receives the buffer's address for kernels.
static uint8_t *ctx_pinned(uint8_t *&buf, size_t &cap, size_t bytes, void **dev = nullptr) {
  if (bytes <= cap && buf) return buf;
  if (buf) (void)hipHostFree(buf);
  buf = nullptr;
  cap = 0;
  size_t want = 1 << 16;
  while (want < bytes) want *= 2;
  if (hipHostMalloc((void **)&buf, want, hipHostMallocDefault) != hipSuccess) {
    buf = nullptr;
    return nullptr;
  }
  if (dev && hipHostGetDevicePointer(dev, buf, 0) != hipSuccess) {
    (void)hipHostFree(buf);
    buf = nullptr;
    return nullptr;
  }
  cap = want;
  return buf;
}

namespace poyg {
int ctx_cus(poyg_ctx *c) { return c ? c->cus : 1; }
}  // namespace poyg

// Device memory cache, one per device.  A tree pass builds and drops a plan
// with a dozen buffers per tree level; hipMalloc and hipFree calls cost more
// than the level's kernels.  Blocks of up to kPoolMax bytes are rounded to a
// power of two and kept: a released block waits in `pending` until a
// device-wide synchronization (poyg_ctx_sync, or the cache's own once the
// pending bytes pass kPoolPendingMax), after which no queued work can still
// use it, and is then handed out again.  Larger blocks (a batch's
// direction band or prefix sums: gigabytes for a TBR round's median plans,
// where a fresh hipMalloc / hipFree pair took longer than the plan's
// kernels) are rounded up to a size class and kept the same way, up to
// kBigCachedShare of the device's memory (a round's plans of tens of
// gigabytes each fit; with a smaller cap every plan paid a hipMalloc and a
// hipFree of its band again); a request takes the smallest kept block that
// holds it, and a failed hipMalloc first returns the kept blocks.
namespace {
constexpr int kPoolMinLog = 8, kPoolMaxLog = 26;   // 256 B .. 64 MiB
constexpr size_t kPoolMax = (size_t)1 << kPoolMaxLog;
constexpr size_t kPoolPendingMax = (size_t)256 << 20;
constexpr size_t kPoolCachedMax = (size_t)1 << 30;
constexpr double kBigCachedShare = 0.6;
constexpr int kPoolDevices = 64;
constexpr int kBigClass = -2;
struct DevPool {
  std::mutex mu;
  std::vector<void *> free_[kPoolMaxLog + 1];
  std::vector<std::pair<void *, int>> pending;
  std::unordered_map<void *, int> cls;   // every block handed out: its class (kBigClass: a large block)
  std::unordered_map<void *, size_t> big_size;          // large blocks handed out
  std::vector<std::pair<void *, size_t>> big_pending, big_free;
  size_t pending_bytes = 0, free_bytes = 0, big_free_bytes = 0;
  size_t big_cap = 0;   // kBigCachedShare of the device's memory, set at first use
};
DevPool g_pool[kPoolDevices];

int pool_class(size_t n) {
  if (n > kPoolMax) return -1;
  int c = kPoolMinLog;
  while (((size_t)1 << c) < n) ++c;
  return c;
}

DevPool *pool_of_current() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kPoolDevices) return nullptr;
  return &g_pool[dev];
}

// the large-block cap of the current device; caller holds the lock
size_t pool_big_cap(DevPool &P) {
  if (!P.big_cap) {
    size_t fr = 0, total = 0;
    P.big_cap = hipMemGetInfo(&fr, &total) == hipSuccess && total ? (size_t)(kBigCachedShare * (double)total)
                                                                   : (size_t)48 << 30;
  }
  return P.big_cap;
}

// pending -> free once the device is idle; caller holds the lock
void pool_promote(DevPool &P) {
  for (auto &b : P.pending) {
    const size_t sz = (size_t)1 << b.second;
    if (P.free_bytes + sz > kPoolCachedMax) {
      (void)hipFree(b.first);
      continue;
    }
    P.free_[b.second].push_back(b.first);
    P.free_bytes += sz;
  }
  P.pending.clear();
  for (auto &b : P.big_pending) {
    if (P.big_free_bytes + b.second > pool_big_cap(P)) {
      (void)hipFree(b.first);
      continue;
    }
    P.big_free.push_back(b);
    P.big_free_bytes += b.second;
  }
  P.big_pending.clear();
  P.pending_bytes = 0;
}

hipError_t pool_alloc(void **p, size_t n) {
  DevPool *P = pool_of_current();
  const int c = pool_class(n);
  if (P && c < 0) {
    // size classes a quarter of a power of two apart (at most 25% unused),
    // so that the similar batches of a round share blocks
    size_t top = kPoolMax;
    while (top * 2 <= n) top *= 2;
    const size_t grain = std::max(kPoolMax, top / 4);
    const size_t want = (n + grain - 1) / grain * grain;
    std::lock_guard<std::mutex> g(P->mu);
    int best = -1;
    for (int k = 0; k < (int)P->big_free.size(); ++k)
      if (P->big_free[k].second >= want && (best < 0 || P->big_free[k].second < P->big_free[best].second)) best = k;
    if (best >= 0) {
      *p = P->big_free[best].first;
      P->big_size[*p] = P->big_free[best].second;
      P->big_free_bytes -= P->big_free[best].second;
      P->big_free.erase(P->big_free.begin() + best);
      P->cls[*p] = kBigClass;
      return hipSuccess;
    }
    hipError_t e = hipMalloc(p, want);
    if (e != hipSuccess && !P->big_free.empty()) {
      // return the kept blocks to the device and try once more
      for (auto &b : P->big_free) (void)hipFree(b.first);
      P->big_free.clear();
      P->big_free_bytes = 0;
      (void)hipGetLastError();
      e = hipMalloc(p, want);
    }
    if (e != hipSuccess) return e;
    P->cls[*p] = kBigClass;
    P->big_size[*p] = want;
    return hipSuccess;
  }
  if (P && c >= 0) {
    std::lock_guard<std::mutex> g(P->mu);
    if (!P->free_[c].empty()) {
      *p = P->free_[c].back();
      P->free_[c].pop_back();
      P->free_bytes -= (size_t)1 << c;
      P->cls[*p] = c;
      return hipSuccess;
    }
  }
  const hipError_t e = hipMalloc(p, c >= 0 ? (size_t)1 << c : n);
  if (e != hipSuccess) return e;
  if (P) {
    std::lock_guard<std::mutex> g(P->mu);
    P->cls[*p] = c;
  }
  return hipSuccess;
}

hipError_t pool_release(void *p) {
  if (!p) return hipSuccess;
  DevPool *P = pool_of_current();
  if (!P) return hipFree(p);
  std::lock_guard<std::mutex> g(P->mu);
  auto it = P->cls.find(p);
  if (it == P->cls.end()) return hipErrorInvalidValue;
  const int c = it->second;
  P->cls.erase(it);
  if (c == kBigClass) {
    const size_t sz = P->big_size[p];
    P->big_size.erase(p);
    P->big_pending.emplace_back(p, sz);
    P->pending_bytes += sz;
  } else if (c < 0) {
    return hipFree(p);
  } else {
    P->pending.emplace_back(p, c);
    P->pending_bytes += (size_t)1 << c;
  }
  if (P->pending_bytes > kPoolPendingMax) {
    const hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return e;
    pool_promote(*P);
  }
  return hipSuccess;
}

// A block whose every use was on a stream that the caller has synchronized
// since: no queued work can touch it, so it is handed out again at once.
hipError_t pool_release_idle(void *p) {
  if (!p) return hipSuccess;
  DevPool *P = pool_of_current();
  if (!P) return hipFree(p);
  std::lock_guard<std::mutex> g(P->mu);
  auto it = P->cls.find(p);
  if (it == P->cls.end()) return hipErrorInvalidValue;
  const int c = it->second;
  P->cls.erase(it);
  if (c == kBigClass) {
    const size_t sz = P->big_size[p];
    P->big_size.erase(p);
    if (P->big_free_bytes + sz > pool_big_cap(*P)) return hipFree(p);
    P->big_free.emplace_back(p, sz);
    P->big_free_bytes += sz;
  } else if (c < 0) {
    return hipFree(p);
  } else {
    const size_t sz = (size_t)1 << c;
    if (P->free_bytes + sz > kPoolCachedMax) return hipFree(p);
    P->free_[c].push_back(p);
    P->free_bytes += sz;
  }
  return hipSuccess;
}

// after a device-wide synchronization of the current device
void pool_synced() {
  DevPool *P = pool_of_current();
  if (!P) return;
  std::lock_guard<std::mutex> g(P->mu);
  pool_promote(*P);
}
}  // namespace

// A device buffer owned by a plan, or a view into another one.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  bool own = true;
  void view(const DevBuf &base, size_t off, size_t n) {
    p = (char *)base.p + off;
    bytes = n;
    own = false;
  }
  int alloc(size_t n) {
    if (n == 0) n = 16;
    if (pool_alloc(&p, n) != hipSuccess) {
      p = nullptr;
      set_error("hipMalloc of %zu bytes failed", n);
      return POYG_ENOMEM;
    }
    bytes = n;
    return 0;
  }
  // idle: every use of the block has completed (see pool_release_idle)
  void release(bool idle = false) {
    if (p && own) (void)(idle ? pool_release_idle(p) : pool_release(p));
    p = nullptr;
    bytes = 0;
    own = true;
  }
};

struct poyg_plan {
  poyg_ctx *ctx = nullptr;
  int device = 0;           // the context's device (the plan may outlive the context object)
  void *triple = nullptr;   // three-sequence plan (mode POYG_MODE_TRIPLE)
  int mode = 0;
  int n = 0;
  int slots = 0;
  int64_t cells = 0;
  int64_t out_bytes = 0;
  std::vector<int64_t> out_off;
  // device
  // work lists (longest first): row kernels (LDS / global rows), the
  // four-pairs-per-wave kernel, the anti-diagonal kernel per lane-set count
  // (1, 2, 4; these two trace their own pairs), and the row kernels'
  // backtraces; pure: per pair, one sequence without gap-bearing characters
  // (written by a scan before the fills)
  // (linear pairs: the systolic kernel takes every pair whose column
  // sequence fits its LDS row, the row kernels the rest)
  DevBuf tcm, meta, order_all, order_lds, order_glb, order_sys, order_c16, order_quad, order_diag[3], order_rows,
      dm, aux, rows, counters, status, sink, pure, trout, blob, redo, tlist;
  int n_lds = 0, n_glb = 0, n_sys = 0, n_quad = 0, n_diag[3] = {0, 0, 0}, n_rows = 0;
  int n_c16 = 0, grid_c16 = 0, c16_nmax = 0;   // linear pairs of the packed 16-bit fill
  bool c16_mw = false;                          // ... four waves per pair (grid_c16 counts workgroups)
  // the narrow affine bands (order_quad) go to the packed eight-pair fill;
  // its pairs whose cost saturates 16 bits are listed in `redo` for the
  // 32-bit four-pair fill (grid_redo workgroups), launched after it
  // (alignment modes: the packed fill lists its other pairs in `tlist` for
  // the backtrace kernel, grid_tlist workgroups)
  bool aff_pk = false;
  int grid_redo = 0, grid_tlist = 0;
  int grid_lds = 0, grid_glb = 0, grid_sys = 0, grid_quad = 0, grid_diag[3] = {0, 0, 0},
      row_width = 0;
  // the fill with the most cells runs on the main stream ("main": 1 the
  // four-pair kernel, 2 the one-set anti-diagonal kernel, 0 none); every
  // other fill runs on the side stream, followed by a second launch of the
  // main kernel (grid_help workgroups) that pulls from the same work list
  int main_fill = 0, grid_help = 0, grid_help2 = 0;
  int trace_bpc = 1;   // resident workgroups per CU of the linear backtrace
  bool timing = false;
  float fill_ms = 0.f, trace_ms = 0.f;
  hipStream_t last_stream = nullptr;   // the stream of the last run (its status is read there)
  bool idle = false;   // set by an owner that synchronized every run: buffers go straight back to the cache
  bool preset = false;   // the upload initialised the first run's counters, status and backtrace pointers
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  hipEvent_t fork = nullptr, join = nullptr, join2 = nullptr;   // hand-off to and from the side streams
  ~poyg_plan() {
    if (triple) poyg_triple_plan_destroy(triple);
    for (DevBuf *b : {&tcm, &meta, &order_all, &order_lds, &order_glb, &order_sys, &order_c16, &order_quad, &order_diag[0],
                      &order_diag[1], &order_diag[2], &order_rows, &dm, &aux, &rows, &counters, &status,
                      &sink, &pure, &trout, &blob, &redo, &tlist})
      b->release(idle);
    for (auto &e : ev)
      if (e) (void)hipEventDestroy(e);
    if (fork) (void)hipEventDestroy(fork);
    if (join) (void)hipEventDestroy(join);
    if (join2) (void)hipEventDestroy(join2);
  }
};

static int64_t round16(int64_t v) { return (v + 15) & ~(int64_t)15; }
static constexpr int kPlanCounters = 16;   // work-list counters of a pair plan's run
static constexpr int kQuadH = 3;   // diagonal pairs per lane of the four-pair kernel

extern "C" {

int poyg_ctx_create(int device, poyg_ctx **out) {
  if (!out) return POYG_EINVAL;
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) {
    set_error("poyg_ctx_create: device %d not present (%d devices)", device, ndev);
    return POYG_EINVAL;
  }
  HIP_TRY(hipSetDevice(device));
  poyg_ctx *c = new poyg_ctx;
  c->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->cus = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&c->side2, hipStreamNonBlocking) != hipSuccess) {
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->side) (void)hipStreamDestroy(c->side);
    delete c;
    set_error("poyg_ctx_create: stream creation failed");
    return POYG_EDEVICE;
  }
  *out = c;
  return POYG_OK;
}

void poyg_ctx_destroy(poyg_ctx *ctx) {
  if (!ctx) return;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->pin_plan) (void)hipHostFree(ctx->pin_plan);
  if (ctx->pin_io) (void)hipHostFree(ctx->pin_io);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  if (ctx->side) (void)hipStreamDestroy(ctx->side);
  if (ctx->side2) (void)hipStreamDestroy(ctx->side2);
  if (ctx->capture) (void)hipStreamDestroy(ctx->capture);
  delete ctx;
}

int poyg_ctx_sync(poyg_ctx *ctx) {
  if (!ctx) return POYG_EINVAL;
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipDeviceSynchronize());
  pool_synced();
  return POYG_OK;
}

int poyg_dev_alloc(poyg_ctx *ctx, int64_t bytes, void **out) {
  if (!ctx || !out || bytes < 0) return POYG_EINVAL;
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(pool_alloc(out, (size_t)(bytes ? bytes : 16)));
  return POYG_OK;
}
int poyg_dev_free(poyg_ctx *ctx, void *p) {
  if (!ctx) return POYG_EINVAL;
  HIP_TRY(hipSetDevice(ctx->device));
  if (p) HIP_TRY(pool_release(p));
  return POYG_OK;
}
int poyg_dev_upload(poyg_ctx *ctx, void *dst, const void *src, int64_t bytes) {
  if (!ctx || bytes < 0) return POYG_EINVAL;
  if (bytes == 0) return POYG_OK;
  HIP_TRY(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return POYG_OK;
}
int poyg_dev_download(poyg_ctx *ctx, void *dst, const void *src, int64_t bytes) {
  if (!ctx || bytes < 0) return POYG_EINVAL;
  if (bytes == 0) return POYG_OK;
  HIP_TRY(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return POYG_OK;
}
int poyg_dev_zero(poyg_ctx *ctx, void *p, int64_t bytes, void *stream) {
  if (!ctx || bytes < 0 || (bytes & 3) || ((uintptr_t)p & 3) || (bytes && !p) || bytes / 4 > INT32_MAX) {
    set_error("poyg_dev_zero: bad arguments");
    return POYG_EINVAL;
  }
  if (bytes == 0) return POYG_OK;
  HIP_TRY(launch_reset((int *)p, (int)(bytes / 4), stream ? (hipStream_t)stream : ctx->stream));
  return POYG_OK;
}

struct poyg_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t stream = nullptr;   // the context's stream (launches with a NULL stream)
};

int poyg_graph_begin(poyg_ctx *ctx, void **capture_stream) {
  if (!ctx || !capture_stream) return POYG_EINVAL;
  HIP_TRY(hipSetDevice(ctx->device));
  if (!ctx->capture) HIP_TRY(hipStreamCreateWithFlags(&ctx->capture, hipStreamNonBlocking));
  HIP_TRY(hipStreamBeginCapture(ctx->capture, hipStreamCaptureModeThreadLocal));
  *capture_stream = ctx->capture;
  return POYG_OK;
}

int poyg_graph_end(poyg_ctx *ctx, poyg_graph **out) {
  if (!ctx || !out || !ctx->capture) return POYG_EINVAL;
  hipGraph_t g = nullptr;
  HIP_TRY(hipStreamEndCapture(ctx->capture, &g));
  hipGraphExec_t ex = nullptr;
  const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    set_error("poyg_graph_end: %s", hipGetErrorString(e));
    return POYG_EDEVICE;
  }
  poyg_graph *r = new poyg_graph;
  r->graph = g;
  r->exec = ex;
  r->stream = ctx->stream;
  *out = r;
  return POYG_OK;
}

int poyg_graph_launch(poyg_graph *g, void *stream) {
  if (!g) return POYG_EINVAL;
  HIP_TRY(hipGraphLaunch(g->exec, stream ? (hipStream_t)stream : g->stream));
  return POYG_OK;
}

void poyg_graph_destroy(poyg_graph *g) {
  if (!g) return;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
}

void *poyg_ctx_stream(poyg_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }
int poyg_ctx_device(poyg_ctx *ctx) { return ctx ? ctx->device : -1; }

static int tune_index(const char *knob) {
  if (!knob) return -1;
  for (int k = 0; k < kTuneKnobs; ++k)
    if (!strcmp(knob, kTuneNames[k])) return k;
  return -1;
}

int poyg_ctx_set_tuning(poyg_ctx *ctx, const char *knob, int value) {
  const int k = tune_index(knob);
  if (!ctx || k < 0) {
    set_error("poyg_ctx_set_tuning: unknown knob %s", knob ? knob : "(null)");
    return POYG_EINVAL;
  }
  static const int lo[kTuneKnobs] = {41, 2, 1, 1, 1, 1, 1}, hi[kTuneKnobs] = {48, 64, 64, 64, 64, 4, 2};
  if (value >= 0 && (value < lo[k] || value > hi[k])) {
    set_error("poyg_ctx_set_tuning: %s = %d outside %d..%d", knob, value, lo[k], hi[k]);
    return POYG_EINVAL;
  }
  ctx->tune[k] = value < 0 ? -1 : value;
  return POYG_OK;
}

int poyg_ctx_device_name(poyg_ctx *ctx, char *buf, int cap) {
  if (!ctx || !buf || cap < 1) return POYG_EINVAL;
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, ctx->device));
  snprintf(buf, (size_t)cap, "%s:%s:%dcu", prop.name, prop.gcnArchName, ctx->cus);
  return POYG_OK;
}

int poyg_ctx_get_tuning(poyg_ctx *ctx, const char *knob) {
  const int k = tune_index(knob);
  if (!ctx || k < 0) return POYG_EINVAL;
  return ctx->tune[k];
}

void poyg_plan_destroy(poyg_plan *plan) {
  if (!plan) return;
  // the plan's blocks return to its device's cache; the device was recorded
  // when the plan was made, so a plan released after its context is still
  // released correctly
  if (hipSetDevice(plan->device) != hipSuccess) (void)hipGetLastError();
  delete plan;
}
int poyg_plan_slots(const poyg_plan *p) { return p->slots; }
int64_t poyg_plan_out_bytes(const poyg_plan *p) {
  return p->triple ? poyg_triple_plan_out_bytes(p->triple) : p->out_bytes;
}
int64_t poyg_plan_cells(const poyg_plan *p) { return p->triple ? poyg_triple_plan_cells(p->triple) : p->cells; }
int poyg_plan_out_offsets(const poyg_plan *p, int64_t *o) {
  if (!o) return POYG_EINVAL;
  if (p->triple) return poyg_triple_plan_out_offsets(p->triple, o);
  memcpy(o, p->out_off.data(), sizeof(int64_t) * p->out_off.size());
  return POYG_OK;
}

int poyg_plan_triples(poyg_ctx *ctx, const poyg_cm3 *cm3, int32_t n, const int64_t *off_1, const int32_t *len_1,
                      const int64_t *off_2, const int32_t *len_2, const int64_t *off_3, const int32_t *len_3,
                      poyg_plan **out) {
  if (!ctx || !out) return POYG_EINVAL;
  // direction cubes of one batch: POYG_CUBE_DM_BUDGET_MB, else a little over
  // half of the free device memory (at least 1 GiB)
  const char *b = getenv("POYG_CUBE_DM_BUDGET_MB");
  int64_t budget = int64_t(1) << 30;
  if (b) {
    budget = atoll(b) << 20;
  } else {
    size_t free_b = 0, total_b = 0;
    HIP_TRY(hipSetDevice(ctx->device));
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) budget = std::max(budget, (int64_t)(free_b / 16 * 9));
  }
  void *tp = nullptr;
  int rc = poyg_triple_plan_create(ctx, cm3, n, off_1, len_1, off_2, len_2, off_3, len_3, ctx->device, ctx->cus,
                                   ctx->stream, ctx->side, ctx->side2, budget, ctx->tune[kTuneCubeWaves], &tp);
  if (rc) return rc;
  poyg_plan *P = new poyg_plan;
  P->ctx = ctx;
  P->device = ctx->device;
  P->triple = tp;
  P->mode = POYG_MODE_TRIPLE;
  P->n = n;
  P->slots = 4;
  *out = P;
  return POYG_OK;
}
int poyg_plan_set_timing(poyg_plan *p, int enable) {
  p->timing = enable != 0;
  if (p->timing && !p->ev[0])
    for (auto &e : p->ev) HIP_TRY(hipEventCreate(&e));
  return POYG_OK;
}
int poyg_plan_last_times(const poyg_plan *p, float *fill_ms, float *trace_ms) {
  if (fill_ms) *fill_ms = p->fill_ms;
  if (trace_ms) *trace_ms = p->trace_ms;
  return POYG_OK;
}

}  // extern "C"

static_assert(kBandFull == kLinFull && kBandCase2 == kLinCase2 && kBandCase3 == kLinCase3,
              "the band cases of the host geometry are the kernels' PairMeta::p0 values");

// Polls before a wait of the systolic linear fill for the previous strip of
// its pair gives up (the kernel's default when negative).  POYG_SYS_SPIN_LIMIT
// overrides it, so that a test can make the waits give up and check that the
// plan reports the pairs as failed.
static int sys_spin_limit() {
  const char *e = getenv("POYG_SYS_SPIN_LIMIT");
  return e && *e ? atoi(e) : -1;
}

static constexpr int kAffRBHost = 256;   // must match kAffRB in the kernels
static constexpr int kLinRBHost = 512;   // must match kLinRB
static constexpr int kLinStageHost = 2048;   // must match kLinStage

// The host-sequence route: the sequences travel in the plan's descriptor
// block (one upload from the context's pinned staging, with no
// synchronization): `a` at offset 0 and, if given, `b` at off_b.  The caller
// runs the plan on ctx->stream with d_seq = dev and synchronizes that stream
// before the next such plan reuses the staging.
struct PerCallSeqs {
  const uint8_t *a, *b;
  int64_t la, lb;
  int64_t off_b;                  // b's offset from a (16-byte aligned)
  const uint8_t *dev = nullptr;   // out: a's copy on the device
  // out: the run's results in the plan's block (costs, then output
  // lengths: small_words words; the output sequences)
  int32_t *small_dev = nullptr;
  uint8_t *out_dev = nullptr;
  int64_t small_words = 0;
};

static int plan_pairs_impl(poyg_ctx *ctx, const poyg_cm *cm, int mode, int32_t n, const int64_t *off_a,
                           const int32_t *len_a, const int64_t *off_b, const int32_t *len_b,
                           const int32_t *deltaw, poyg_plan **out, PerCallSeqs *pc);

extern "C" int poyg_plan_pairs(poyg_ctx *ctx, const poyg_cm *cm, int mode, int32_t n, const int64_t *off_a,
                               const int32_t *len_a, const int64_t *off_b, const int32_t *len_b,
                               const int32_t *deltaw, poyg_plan **out) {
  return plan_pairs_impl(ctx, cm, mode, n, off_a, len_a, off_b, len_b, deltaw, out, nullptr);
}

static int plan_pairs_impl(poyg_ctx *ctx, const poyg_cm *cm, int mode, int32_t n, const int64_t *off_a,
                           const int32_t *len_a, const int64_t *off_b, const int32_t *len_b,
                           const int32_t *deltaw, poyg_plan **out, PerCallSeqs *pc) {
  const bool per_call = pc != nullptr;
  if (!ctx || !cm || !out || n < 0 || (n > 0 && (!off_a || !len_a || !off_b || !len_b))) {
    set_error("poyg_plan_pairs: bad arguments");
    return POYG_EINVAL;
  }
  const bool affine = mode == POYG_MODE_AFFINE_COST || mode == POYG_MODE_AFFINE_ALIGN;
  const bool align = mode == POYG_MODE_AFFINE_ALIGN || mode == POYG_MODE_LINEAR_ALIGN;
  if (mode < 0 || mode > 3) return POYG_EINVAL;
  if (affine != (cm->cost_model == 1)) {
    set_error("poyg_plan_pairs: mode %d does not match the cost model %d of the matrix", mode, cm->cost_model);
    return POYG_EINVAL;
  }
  if (!cm->combinations || cm->lcm != 5 || cm->gap != kGap) {
    set_error("poyg_plan_pairs: only 5-bit DNA combination matrices are supported");
    return POYG_EUNSUPPORTED;
  }
  if (!affine && !deltaw) {
    set_error("poyg_plan_pairs: linear modes need the deltawh of every pair");
    return POYG_EINVAL;
  }
  HIP_TRY(hipSetDevice(ctx->device));
  poyg_plan *P = new poyg_plan;
  P->ctx = ctx;
  P->device = ctx->device;
  P->mode = mode;
  P->n = n;
  P->slots = align ? (affine ? 4 : 3) : 0;
  P->trace_bpc = ctx->tune[kTuneTraceBlocks] > 0 ? std::min(ctx->tune[kTuneTraceBlocks], lin_trace_blocks_per_cu())
                                                 : lin_trace_blocks_per_cu();
  std::vector<PairMeta> meta(n);
  std::vector<int64_t> cells(n);
  std::vector<char> glb(n, 0), diag(n, 0), quad(n, 0), sys(n, 0), c16(n, 0);
  DevTcm dt;
  to_dev_tcm(cm, &dt);
  // the packed 16-bit fill: linear plans whose tail costs equal the
  // deletion costs (its last and first columns need no rule of their own
  // then); per pair, every cost below 0xffff
  int64_t c16_cmax = 0;
  bool c16_matrix = !affine && ctx->tune[kTuneLinCost] != 1;
  for (int x = 0; x < 32; ++x) {
    c16_matrix = c16_matrix && dt.tail[x] == dt.cost[x * 32 + kGap] && dt.prepend[x] >= 0 && dt.tail[x] >= 0;
    c16_cmax = std::max<int64_t>(c16_cmax, std::max(dt.prepend[x], dt.tail[x]));
  }
  for (int x = 0; x < 32 * 32; ++x) {
    c16_matrix = c16_matrix && dt.cost[x] >= 0;
    c16_cmax = std::max<int64_t>(c16_cmax, dt.cost[x]);
  }
  // tuned launch shapes (see poyg_ctx_set_tuning)
  // (a per-call pair runs alone: the one-pair-per-wave kernel steps one
  // diagonal pair per lane instead of three, a shorter latency)
  const int quad_max =
      per_call ? 0 : (ctx->tune[kTuneQuadMaxLanes] > 0 ? ctx->tune[kTuneQuadMaxLanes] : 16 * kQuadH);
  const int side_div = ctx->tune[kTuneSideDiv] > 0 ? ctx->tune[kTuneSideDiv] : 4;
  const int sys_bpc = ctx->tune[kTuneSysBlocks] > 0 ? std::min(ctx->tune[kTuneSysBlocks], lin_sys_blocks_per_cu(align))
                                                    : lin_sys_blocks_per_cu(align);
  int64_t dm_bytes = 0, aux_ints = 0, out_bytes = 0;
  int row_width = 0;
  P->out_off.resize(n);
  for (int k = 0; k < n; ++k) {
    PairMeta &m = meta[k];
    const int la = len_a[k], lb = len_b[k];
    if (la < 1 || lb < 1) {
      delete P;
      set_error("poyg_plan_pairs: pair %d has an empty sequence (the leading gap is required)", k);
      return POYG_EINVAL;
    }
    memset(&m, 0, sizeof m);
    m.cap = (int32_t)round16((int64_t)la + lb + 2);
    m.out_off = out_bytes;
    P->out_off[k] = out_bytes;
    out_bytes += (int64_t)P->slots * m.cap;
    if (affine) {
      // algn_CAML_align_affine_3: the shorter sequence indexes the rows
      m.swap = la > lb;
      m.off_r = m.swap ? off_b[k] : off_a[k];
      m.off_c = m.swap ? off_a[k] : off_b[k];
      m.lr = m.swap ? lb : la;
      m.lc = m.swap ? la : lb;
      const int ra = m.lr - 1, cb = m.lc - 1;
      const int e0 = std::min(cb, std::max(40, cb - ra + 8));
      m.p0 = e0;
      const int lanes = ((e0 + 40) >> 1) + 1;   // band diagonal pairs
      if (lanes <= 4 * kWave) {
        // anti-diagonal wavefront (1, 2 or 4 lane sets): one dword per
        // diagonal pair and 4 steps
        diag[k] = lanes <= kWave ? 1 : (lanes <= 2 * kWave ? 2 : 4);
        // four pairs per wave: 16 lanes of H = 3 diagonal pairs (row 0
        // alone, ra == 0, stays with the one-set kernel).  Wider bands belong
        // to long pairs, whose latency the one-pair-per-wave kernel keeps
        // lower (a lane of it steps one diagonal pair instead of three).
        if (ra > 0 && lanes <= quad_max) quad[k] = 3;
        m.p1 = lanes;
        m.p2 = 1;
        if (align && ra > 0) {
          m.dm_off = dm_bytes;
          dm_bytes += (((int64_t)ra + cb + 4) >> 2) * 4 * lanes;
        }
      } else {
        m.stride = (int32_t)((std::min(cb + 1, e0 + 41) + 3) & ~3);
        m.aux_off = aux_ints;
        aux_ints += (2 * (int64_t)(cb + 1) + 63) & ~(int64_t)63;   // 256-byte aligned regions
        if (align) {
          m.dm_off = dm_bytes;
          dm_bytes += (int64_t)ra * m.stride;
        }
        if (m.stride > kAffRBHost) {
          glb[k] = 1;
          row_width = std::max(row_width, (cb + 2 + 63) & ~63);
        }
      }
      int64_t c = 0;
      for (int i = 1; i <= ra; ++i) {
        const int sp = 1 + (i > 40 ? i - 40 : 0), ep = std::min(cb, e0 + i - 1);
        c += ep - sp + 1;
      }
      cells[k] = c;
    } else {
      // Sequence.Align.cost_2 / create_edited_2: the longer sequence goes first
      m.swap = la < lb;
      m.off_r = m.swap ? off_b[k] : off_a[k];
      m.off_c = m.swap ? off_a[k] : off_b[k];
      m.lr = m.swap ? lb : la;
      m.lc = m.swap ? la : lb;
      const LinGeom g = lin_geom(m.lr, m.lc, deltaw[k]);
      m.p0 = g.cs;
      m.p1 = g.w;
      m.p2 = g.h;
      int maxw = 0;
      int64_t c = 0;
      lin_band_extent(m.lr, m.lc, g, maxw, c);
      cells[k] = c;
      m.stride = lin_dm_stride(maxw);   // room for a row start phase of up to 3 bytes (systolic fill)
      if (align && !lin_dm_fits(m.lr, maxw)) {
        delete P;
        set_error("poyg_plan_pairs: pair %d (%d x %d) has a direction band of %lld bytes, beyond the linear "
                  "kernels' 32-bit band offsets", k, m.lr, m.lc, (long long)m.lr * m.stride);
        return POYG_EUNSUPPORTED;
      }
      m.aux_off = aux_ints;
      aux_ints += (2 * (int64_t)m.lc + 63) & ~(int64_t)63;
      if (align) {
        m.dm_off = dm_bytes;
        dm_bytes += (int64_t)m.lr * m.stride;
      }
      if (c16_matrix && m.lc >= 2 && m.lc <= kLinStageHost &&
          ((int64_t)m.lr + m.lc + 2) * std::max<int64_t>(c16_cmax, 1) < 0xfff0) {
        c16[k] = 1;
        m.p3 = 1;
      } else if (m.lc <= kLinStageHost) {
        sys[k] = 1;
      } else if (maxw > kLinRBHost) {
        glb[k] = 1;
        row_width = std::max(row_width, (m.lc + 1 + 63) & ~63);
      }
    }
    P->cells += cells[k];
  }
  P->out_bytes = out_bytes;
  // longest-first work lists
  std::vector<int32_t> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cells[a] > cells[b]; });
  std::vector<int32_t> o_lds, o_glb, o_sys, o_c16, o_quad, o_diag[3], o_rows;
  int64_t c_quad3 = 0, c_diag1 = 0;
  for (int k : order) {
    if (quad[k]) {
      o_quad.push_back(k);
      c_quad3 += cells[k];
    } else if (diag[k]) {
      o_diag[diag[k] == 1 ? 0 : (diag[k] == 2 ? 1 : 2)].push_back(k);
      if (diag[k] == 1) c_diag1 += cells[k];
    } else if (c16[k]) {
      o_c16.push_back(k);
      P->c16_nmax = std::max(P->c16_nmax, (int)meta[k].lc);
    } else {
      (sys[k] ? o_sys : glb[k] ? o_glb : o_lds).push_back(k);
      o_rows.push_back(k);
    }
  }
  P->n_lds = (int)o_lds.size();
  P->n_glb = (int)o_glb.size();
  P->n_sys = (int)o_sys.size();
  P->n_rows = (int)o_rows.size();
  // persistent grids: exactly the workgroups that can be resident at once
  for (int q = 0; q < 3; ++q) {
    P->n_diag[q] = (int)o_diag[q].size();
    if (P->n_diag[q])
      P->grid_diag[q] = std::max(1, std::min((P->n_diag[q] + 3) / 4,
                                             ctx->cus * aff_diag_blocks_per_cu(align, q == 0 ? 1 : 2 * q)));
  }
  P->n_quad = (int)o_quad.size();
  P->aff_pk = P->n_quad > 0 && ctx->tune[kTuneAffNarrow] != 1;
  // workgroups of the narrow-band fill per CU (eight or four pairs per wave)
  auto narrow_bpc = [&]() { return P->aff_pk ? aff_pk_blocks_per_cu(align) : aff_quad_blocks_per_cu(align); };
  if (P->n_quad)   // four waves per workgroup
    P->grid_quad = std::max(1, std::min((P->n_quad + (P->aff_pk ? 31 : 15)) / (P->aff_pk ? 32 : 16),
                                        ctx->cus * narrow_bpc()));
  if (P->aff_pk) {
    P->grid_redo = std::max(1, std::min((P->n_quad + 15) / 16, ctx->cus));
    if (align) P->grid_tlist = std::max(1, std::min((P->n_quad + 3) / 4, ctx->cus * aff_trace_list_blocks_per_cu()));
  }
  P->row_width = row_width;
  const int per_cu = affine ? 4 : 8;
  P->grid_lds = std::max(1, std::min((P->n_lds + 3) / 4, ctx->cus * per_cu));
  P->grid_glb = std::max(1, std::min(P->n_glb, 256));
  if (P->n_sys) P->grid_sys = std::max(1, std::min(P->n_sys, ctx->cus * sys_bpc));
  P->n_c16 = (int)o_c16.size();
  if (P->n_c16) {
    // a batch that one pair per wave leaves far from filling the chip (tree
    // levels, per-call entries) gets four waves per pair: each pair's
    // latency is about a third
    const int lk = ctx->tune[kTuneLinCost];
    P->c16_mw = lk == 4 || (lk != 3 && P->n_c16 <= 2 * ctx->cus);
    if (P->c16_mw)   // grid_c16 counts workgroups
      P->grid_c16 = std::max(1, std::min(P->n_c16, ctx->cus * lin_c16mw_blocks_per_cu(align, P->c16_nmax)));
    else   // one pair per wave; grid_c16 counts waves
      P->grid_c16 = std::max(1, std::min(P->n_c16, ctx->cus * lin_cost16_blocks_per_cu(align, P->c16_nmax)));
  }
  if (affine) {
    // The fill with the most cells is the main kernel.  The side stream's
    // kernels are persistent too; the main kernel leaves them whole CUs, or
    // whichever kernel the hardware dispatches first would hold every slot
    // and the other would wait for it to drain.  Once the side work is done,
    // a second launch of the main kernel on the side stream pulls from the
    // same work list.
    P->main_fill = P->n_quad && c_quad3 >= c_diag1 ? 1 : (P->n_diag[0] ? 2 : (P->n_quad ? 1 : 0));
    auto cus_of = [](int grid, int bpc) { return (grid + bpc - 1) / std::max(1, bpc); };
    int side = 0, side2 = 0;   // CUs of the side stream and of the widest-band stream
    for (int q = 1; q < 3; ++q)
      if (P->n_diag[q]) side2 += cus_of(P->grid_diag[q], aff_diag_blocks_per_cu(align, 2 * q));
    if (P->n_lds) side2 += cus_of(P->grid_lds, per_cu);
    if (P->n_glb) side2 += cus_of(P->grid_glb, 8);
    if (P->n_quad && P->main_fill != 1) side += cus_of(P->grid_quad, narrow_bpc());
    if (P->n_diag[0] && P->main_fill != 2) side += cus_of(P->grid_diag[0], aff_diag_blocks_per_cu(align, 1));
    if ((side || side2) && P->main_fill) {
      side = std::min(side, ctx->cus / side_div);
      side2 = std::min(side2, ctx->cus / side_div);
      const int bpc = P->main_fill == 1 ? narrow_bpc() : aff_diag_blocks_per_cu(align, 1);
      int &grid_main = P->main_fill == 1 ? P->grid_quad : P->grid_diag[0];
      grid_main = std::max(1, std::min(grid_main, (ctx->cus - side - side2) * bpc));
      // a second (third) launch of the main kernel follows the side work on
      // each side stream and pulls from the same work list
      P->grid_help = std::max(1, side * bpc);
      P->grid_help2 = side2 ? side2 * bpc : 0;
    }
  }
  if (!affine) {
    // The linear kernels mark cells outside the band with 2^30 - 1 and add
    // costs to it: every entry must be non-negative and the most costly
    // alignment (every column at the largest entry) must stay below 2^30,
    // so that no sum overflows and the marker stays above every cell value.
    int64_t cmax = 0, longest = 0;
    bool negative = false;
    for (int x = 0; x < 32 * 32; ++x) {
      negative |= dt.cost[x] < 0;
      cmax = std::max<int64_t>(cmax, dt.cost[x]);
    }
    for (int x = 0; x < 32; ++x) {
      negative |= dt.prepend[x] < 0 || dt.tail[x] < 0;
      cmax = std::max<int64_t>(cmax, std::max(dt.prepend[x], dt.tail[x]));
    }
    for (int k = 0; k < n; ++k) longest = std::max<int64_t>(longest, (int64_t)len_a[k] + len_b[k] + 2);
    if (negative || cmax * longest >= ((int64_t)1 << 30)) {
      delete P;
      set_error("poyg_plan_pairs: linear costs must be non-negative and small enough that an alignment costs "
                "less than 2^30 (largest entry %lld, longest pair %lld)", (long long)cmax, (long long)longest);
      return POYG_EINVAL;
    }
  }
  if (affine) {
    // The affine kernels rank equal minima by packing (value << 2 | rank);
    // every DP value is non-negative and at most 3 kHigh plus the cost of
    // the whole alignment, which must stay below 2^29.
    int64_t cmax = std::max(0, dt.gap_open), longest = 0;
    bool negative = dt.gap_open < 0;
    for (int x = 0; x < 32 * 32; ++x) {
      negative |= dt.cost[x] < 0;
      cmax = std::max<int64_t>(cmax, dt.cost[x]);
    }
    for (int x = 0; x < 32; ++x) {
      negative |= dt.prepend[x] < 0;
      cmax = std::max<int64_t>(cmax, dt.prepend[x]);
    }
    for (int k = 0; k < n; ++k) longest = std::max<int64_t>(longest, (int64_t)len_a[k] + len_b[k] + 2);
    if (negative || 3 * (int64_t)kHighNum + 2 * cmax * longest >= ((int64_t)1 << 29)) {
      delete P;
      set_error("poyg_plan_pairs: affine costs must be non-negative and small enough that an alignment costs "
                "less than 2^28 (largest entry %lld, longest pair %lld)", (long long)cmax, (long long)longest);
      return POYG_EINVAL;
    }
  }
  // the cost matrix, the pair descriptors and the work lists travel as one
  // upload: sections of one block at 256-byte offsets
  std::vector<std::pair<DevBuf *, std::pair<const void *, size_t>>> sections = {
      {&P->tcm, {&dt, sizeof dt}},
      {&P->meta, {meta.data(), sizeof(PairMeta) * (size_t)n}},
      {&P->order_all, {order.data(), sizeof(int32_t) * (size_t)n}},
      {&P->order_lds, {o_lds.data(), sizeof(int32_t) * o_lds.size()}},
      {&P->order_glb, {o_glb.data(), sizeof(int32_t) * o_glb.size()}},
      {&P->order_sys, {o_sys.data(), sizeof(int32_t) * o_sys.size()}},
      {&P->order_c16, {o_c16.data(), sizeof(int32_t) * o_c16.size()}},
      {&P->order_quad, {o_quad.data(), sizeof(int32_t) * o_quad.size()}},
      {&P->order_diag[0], {o_diag[0].data(), sizeof(int32_t) * o_diag[0].size()}},
      {&P->order_diag[1], {o_diag[1].data(), sizeof(int32_t) * o_diag[1].size()}},
      {&P->order_diag[2], {o_diag[2].data(), sizeof(int32_t) * o_diag[2].size()}},
      {&P->order_rows, {o_rows.data(), sizeof(int32_t) * o_rows.size()}}};
  std::vector<size_t> sec_off;
  size_t blob_bytes = 0;
  for (auto &sec : sections) {
    sec_off.push_back(blob_bytes);
    blob_bytes += (std::max<size_t>(sec.second.second, 16) + 255) & ~(size_t)255;
  }
  const size_t seq_at = blob_bytes;   // the per-call route's sequences
  if (per_call)   // a at 0, b (if any) at off_b
    blob_bytes += (size_t)(std::max(pc->la, pc->b ? pc->off_b + pc->lb : 0) + 255) & ~(size_t)255;
  // the host-sequence route runs its plan once: the run's work-list
  // counters, status block and backtrace pointers travel initialised with
  // the descriptors (no start kernel), and its results (costs, output
  // lengths, output sequences) sit behind them in the same block, not
  // uploaded
  size_t ctr_at = 0, st_at = 0, tro_at = 0, small_at = 0, res_at = 0;
  int64_t small_words = 0;
  auto r256 = [](size_t x) { return (std::max<size_t>(x, 16) + 255) & ~(size_t)255; };
  if (per_call) {
    ctr_at = blob_bytes;
    blob_bytes += r256(sizeof(int) * kPlanCounters);
    st_at = blob_bytes;
    blob_bytes += r256(sizeof(int32_t) * ((size_t)kStatusFlags + n));
    tro_at = blob_bytes;
    blob_bytes += r256(sizeof(TraceOut));
  }
  const size_t upload_bytes = blob_bytes;
  if (per_call) {
    const int64_t n_cost = ((int64_t)n + 3) & ~(int64_t)3;
    small_words = n_cost + (((int64_t)n * P->slots + 3) & ~(int64_t)3);
    small_at = blob_bytes;
    blob_bytes += r256(4 * (size_t)small_words);
    res_at = blob_bytes;
    blob_bytes += r256((size_t)out_bytes);
  }
  // per-call plans also carve their scratch (purity flags, direction band,
  // prefix sums, rows, sink) from the block: one pool block per call
  size_t pure_at = 0, dm_at = 0, aux_at = 0, rows_at = 0, sink_at = 0;
  const size_t rows_bytes = P->n_glb ? sizeof(int) * (size_t)P->grid_glb * (affine ? 8 : 2) * row_width : 16;
  if (per_call) {
    pure_at = blob_bytes;
    blob_bytes += r256((size_t)n);
    dm_at = blob_bytes;
    blob_bytes += r256((size_t)dm_bytes);
    aux_at = blob_bytes;
    blob_bytes += r256(sizeof(int) * (size_t)aux_ints);
    rows_at = blob_bytes;
    blob_bytes += r256(rows_bytes);
    sink_at = blob_bytes;
    blob_bytes += r256(sizeof(uint32_t) * kWave);
  }
  std::vector<uint8_t> staging_v;
  uint8_t *staging = per_call ? ctx_pinned(ctx->pin_plan, ctx->pin_plan_cap, upload_bytes) : nullptr;
  const bool sync_upload = staging == nullptr;
  if (!staging) {
    staging_v.assign(upload_bytes, 0);
    staging = staging_v.data();
  } else {
    memset(staging, 0, seq_at);
    memset(staging + ctr_at, 0, upload_bytes - ctr_at);
  }
  for (size_t k = 0; k < sections.size(); ++k)
    if (sections[k].second.second) memcpy(staging + sec_off[k], sections[k].second.first, sections[k].second.second);
  if (per_call) {
    if (pc->la) memcpy(staging + seq_at, pc->a, (size_t)pc->la);
    if (pc->b && pc->lb) memcpy(staging + seq_at + pc->off_b, pc->b, (size_t)pc->lb);
  }
  int rc = 0;
  rc |= P->blob.alloc(blob_bytes);
  if (!rc)
    for (size_t k = 0; k < sections.size(); ++k)
      sections[k].first->view(P->blob, sec_off[k], std::max<size_t>(sections[k].second.second, 16));
  if (per_call && !rc) {
    P->pure.view(P->blob, pure_at, (size_t)n);
    P->dm.view(P->blob, dm_at, (size_t)dm_bytes);
    P->aux.view(P->blob, aux_at, sizeof(int) * (size_t)aux_ints);
    P->rows.view(P->blob, rows_at, rows_bytes);
    P->sink.view(P->blob, sink_at, sizeof(uint32_t) * kWave);
  } else if (!per_call) {
    rc |= P->pure.alloc((size_t)n);
    rc |= P->dm.alloc((size_t)dm_bytes);
    rc |= P->aux.alloc(sizeof(int) * (size_t)aux_ints);
    rc |= P->rows.alloc(rows_bytes);
  }
  if (per_call && !rc) {
    P->counters.view(P->blob, ctr_at, sizeof(int) * kPlanCounters);
    P->status.view(P->blob, st_at, sizeof(int32_t) * ((size_t)kStatusFlags + n));
    P->trout.view(P->blob, tro_at, sizeof(TraceOut));
    uint8_t *base = (uint8_t *)P->blob.p;
    pc->small_dev = (int32_t *)(base + small_at);
    pc->out_dev = base + res_at;
    pc->small_words = small_words;
    const int64_t n_cost = ((int64_t)n + 3) & ~(int64_t)3;
    const TraceOut t{pc->out_dev, P->slots ? pc->small_dev + n_cost : nullptr, (int32_t *)(base + st_at)};
    memcpy(staging + tro_at, &t, sizeof t);
    P->preset = true;
  } else if (!per_call) {
    rc |= P->counters.alloc(sizeof(int) * kPlanCounters);
    rc |= P->status.alloc(sizeof(int32_t) * ((size_t)kStatusFlags + n));
    rc |= P->trout.alloc(sizeof(TraceOut));
  }
  if (!per_call) rc |= P->sink.alloc(sizeof(uint32_t) * kWave);
  if (P->aff_pk) rc |= P->redo.alloc(sizeof(int32_t) * (size_t)P->n_quad);
  if (P->aff_pk && align) rc |= P->tlist.alloc(sizeof(int32_t) * (size_t)P->n_quad);
  if (!rc && affine && P->grid_help &&
      (hipEventCreateWithFlags(&P->fork, hipEventDisableTiming) != hipSuccess ||
       hipEventCreateWithFlags(&P->join, hipEventDisableTiming) != hipSuccess ||
       hipEventCreateWithFlags(&P->join2, hipEventDisableTiming) != hipSuccess))
    rc = POYG_EDEVICE;
  if (rc) {
    delete P;
    return rc;
  }
  hipStream_t s = ctx->stream;
  hipError_t e = hipSuccess;
  e = hipMemcpyAsync(P->blob.p, staging, upload_bytes, hipMemcpyHostToDevice, s);
  if (e == hipSuccess && sync_upload) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    delete P;
    set_error("poyg_plan_pairs: upload failed: %s", hipGetErrorString(e));
    return POYG_EDEVICE;
  }
  if (per_call) pc->dev = (const uint8_t *)P->blob.p + seq_at;
  *out = P;
  return POYG_OK;
}

extern "C" int poyg_plan_run(poyg_plan *P, const uint8_t *d_seq, int32_t *d_cost, uint8_t *d_out,
                             int32_t *d_out_len, void *stream) {
  if (!P || !d_seq || !d_cost) return POYG_EINVAL;
  if (P->n == 0) return POYG_OK;
  if (P->triple) {
    float *f = P->timing ? &P->fill_ms : nullptr, *t = P->timing ? &P->trace_ms : nullptr;
    return poyg_triple_plan_run(P->triple, d_seq, d_cost, d_out, d_out_len,
                                stream ? stream : (void *)P->ctx->stream, f, t);
  }
  const bool affine = P->mode >= POYG_MODE_AFFINE_COST;
  const bool align = P->mode == POYG_MODE_AFFINE_ALIGN || P->mode == POYG_MODE_LINEAR_ALIGN;
  if (align && (!d_out || !d_out_len)) {
    set_error("poyg_plan_run: alignment modes need output buffers");
    return POYG_EINVAL;
  }
  if (align && ((uintptr_t)d_out & 15) != 0) {
    set_error("poyg_plan_run: the output buffer must be 16-byte aligned");
    return POYG_EINVAL;
  }
  hipStream_t s = stream ? (hipStream_t)stream : P->ctx->stream;
  // the launches below report hipGetLastError; an error an earlier,
  // unrelated call left behind is not this run's
  (void)hipGetLastError();
  int *ctr = (int *)P->counters.p;
  const PairMeta *meta = (const PairMeta *)P->meta.p;
  const DevTcm *tcm = (const DevTcm *)P->tcm.p;
  if (P->timing) HIP_TRY(hipEventRecord(P->ev[0], s));
  if (P->preset)   // the first run of a host-route plan: the upload set the counters, status and pointers
    P->preset = false;
  else
    HIP_TRY(launch_plan_start(ctr, kPlanCounters, (int *)P->status.p, kStatusFlags + P->n,
                              affine && align ? (TraceOut *)P->trout.p : nullptr, d_out, d_out_len, s));
  P->last_stream = s;
  if (!affine || P->n_lds + P->n_glb)
    HIP_TRY(launch_prep(affine, std::max(1, std::min((P->n + 3) / 4, P->ctx->cus * 8)), meta, P->n, d_seq, tcm,
                        (int *)P->aux.p, s));
  if (affine) {
    // The main kernel runs on the main stream; the other one-set and
    // four-pair fills on the side stream, followed by a second launch of the
    // main kernel there; the widest bands (lane sets, row kernels: few, long
    // pairs) on a third stream, so that they start at once.
    // purity of the four-pair kernel's pairs, before any fill is released
    if (P->n_quad)
      HIP_TRY(launch_pair_pure(std::max(1, std::min((P->n_quad + 3) / 4, P->ctx->cus * 8)), meta,
                               (const int32_t *)P->order_quad.p, P->n_quad, d_seq, (uint8_t *)P->pure.p, s));
    const bool fork = P->fork != nullptr;
    hipStream_t rs = s, rs2 = s;
    if (fork) {
      rs = P->ctx->side;
      rs2 = P->ctx->side2;
      HIP_TRY(hipEventRecord(P->fork, s));
      HIP_TRY(hipStreamWaitEvent(rs, P->fork, 0));
      HIP_TRY(hipStreamWaitEvent(rs2, P->fork, 0));
    }
    uint8_t *dm = (uint8_t *)P->dm.p;
    uint32_t *sink = (uint32_t *)P->sink.p;
    // the narrow bands: the packed fill queues the pairs whose cost saturates
    // 16 bits (count in ctr[5]) for the 32-bit fill (work counter ctr[7]),
    // and lists the others for the backtrace (count ctr[8], work counter
    // ctr[9])
    auto quad = [&](int grid, hipStream_t st) {
      if (P->aff_pk)
        return launch_aff_pk(align, grid, meta, (const int32_t *)P->order_quad.p, P->n_quad, ctr + 6, d_seq, tcm, dm,
                             d_cost, sink, (const uint8_t *)P->pure.p, (int32_t *)P->redo.p, ctr + 5,
                             (int32_t *)P->tlist.p, ctr + 8, st);
      return launch_aff_quad(align, grid, meta, (const int32_t *)P->order_quad.p, P->n_quad, ctr + 6, d_seq, tcm, dm,
                             d_cost, sink, (const uint8_t *)P->pure.p, (const TraceOut *)P->trout.p, nullptr, st);
    };
    auto diag = [&](int q, int grid, hipStream_t st) {
      return launch_aff_diag(align, q == 0 ? 1 : 2 * q, grid, meta, (const int32_t *)P->order_diag[q].p,
                             P->n_diag[q], ctr + 2 + q, d_seq, tcm, dm, d_cost, sink, (const TraceOut *)P->trout.p,
                             st);
    };
    for (int q = 2; q >= 1; --q)   // the widest bands: 4 and 2 lane sets
      if (P->n_diag[q]) HIP_TRY(diag(q, P->grid_diag[q], rs2));
    if (P->n_lds)
      HIP_TRY(launch_aff_fill(align, false, P->grid_lds, meta, (const int32_t *)P->order_lds.p, P->n_lds, ctr,
                              d_seq, tcm, dm, (int *)P->aux.p, d_cost, nullptr, 0, rs2));
    if (P->n_glb)
      HIP_TRY(launch_aff_fill(align, true, P->grid_glb, meta, (const int32_t *)P->order_glb.p, P->n_glb, ctr + 1,
                              d_seq, tcm, dm, (int *)P->aux.p, d_cost, (int *)P->rows.p, P->row_width, rs2));
    if (P->n_quad && P->main_fill != 1) HIP_TRY(quad(P->grid_quad, rs));
    if (P->n_diag[0] && P->main_fill != 2) HIP_TRY(diag(0, P->grid_diag[0], rs));
    if (fork && P->grid_help) {
      if (P->main_fill == 1) HIP_TRY(quad(P->grid_help, rs));
      else HIP_TRY(diag(0, P->grid_help, rs));
    }
    if (fork && P->grid_help2) {
      if (P->main_fill == 1) HIP_TRY(quad(P->grid_help2, rs2));
      else HIP_TRY(diag(0, P->grid_help2, rs2));
    }
    if (P->main_fill == 1) HIP_TRY(quad(P->grid_quad, s));
    else if (P->main_fill == 2) HIP_TRY(diag(0, P->grid_diag[0], s));
    if (fork) {
      HIP_TRY(hipEventRecord(P->join, rs));
      HIP_TRY(hipEventRecord(P->join2, rs2));
      HIP_TRY(hipStreamWaitEvent(s, P->join, 0));
      HIP_TRY(hipStreamWaitEvent(s, P->join2, 0));
    }
    if (P->aff_pk)   // every launch of the packed fill is done: its redo list is complete
      HIP_TRY(launch_aff_quad(align, P->grid_redo, meta, (const int32_t *)P->redo.p, P->n_quad, ctr + 7, d_seq, tcm,
                              dm, d_cost, sink, (const uint8_t *)P->pure.p, (const TraceOut *)P->trout.p, ctr + 5,
                              s));
  } else {
    if (P->n_c16 && P->c16_mw)
      HIP_TRY(launch_lin_c16mw(align, P->grid_c16, meta, (const int32_t *)P->order_c16.p, P->n_c16, ctr + 4, d_seq,
                               tcm, (const int *)P->aux.p, d_cost, (uint8_t *)P->dm.p, P->c16_nmax, s));
    else if (P->n_c16)
      HIP_TRY(launch_lin_cost16(align, P->grid_c16, meta, (const int32_t *)P->order_c16.p, P->n_c16, ctr + 4, d_seq,
                                tcm, (const int *)P->aux.p, d_cost, (uint8_t *)P->dm.p, P->c16_nmax, s));
    if (P->n_sys)
      HIP_TRY(launch_lin_sys(align, P->grid_sys, meta, (const int32_t *)P->order_sys.p, P->n_sys, ctr + 2, d_seq,
                             tcm, (uint8_t *)P->dm.p, (const int *)P->aux.p, d_cost, (int32_t *)P->status.p,
                             sys_spin_limit(), s));
    if (P->n_lds)
      HIP_TRY(launch_lin_fill(align, false, P->grid_lds, meta, (const int32_t *)P->order_lds.p, P->n_lds, ctr,
                              d_seq, tcm, (uint8_t *)P->dm.p, (int *)P->aux.p, d_cost, nullptr, 0, s));
    if (P->n_glb)
      HIP_TRY(launch_lin_fill(align, true, P->grid_glb, meta, (const int32_t *)P->order_glb.p, P->n_glb, ctr + 1,
                              d_seq, tcm, (uint8_t *)P->dm.p, (int *)P->aux.p, d_cost, (int *)P->rows.p,
                              P->row_width, s));
  }
  if (P->timing) HIP_TRY(hipEventRecord(P->ev[1], s));
  if (align) {
    if (affine) {
      // the four-pair and anti-diagonal fills have traced their own pairs;
      // the packed fill listed its pairs for the backtrace kernel
      if (P->aff_pk)
        HIP_TRY(launch_aff_trace_list(P->grid_tlist, meta, (const int32_t *)P->tlist.p, ctr + 8, ctr + 9, d_seq, tcm,
                                      (const uint8_t *)P->dm.p, (const TraceOut *)P->trout.p, s));
      if (P->n_rows)
        HIP_TRY(launch_aff_trace(meta, (const int32_t *)P->order_rows.p, P->n_rows, d_seq, tcm,
                                 (const uint8_t *)P->dm.p, d_out, d_out_len, (int32_t *)P->status.p, s));
    }
    else
      HIP_TRY(launch_lin_trace(std::max(1, std::min((P->n + 3) / 4, P->ctx->cus * P->trace_bpc)), meta,
                               (const int32_t *)P->order_all.p, P->n, ctr + 3, d_seq, tcm, (const uint8_t *)P->dm.p,
                               d_out, d_out_len, (int32_t *)P->status.p, s));
  }
  if (P->timing) {
    HIP_TRY(hipEventRecord(P->ev[2], s));
    HIP_TRY(hipEventSynchronize(P->ev[2]));
    HIP_TRY(hipEventElapsedTime(&P->fill_ms, P->ev[0], P->ev[1]));
    HIP_TRY(hipEventElapsedTime(&P->trace_ms, P->ev[1], P->ev[2]));
  }
  return POYG_OK;
}

// Number of pairs of the last run that failed (0 when sane), read on the
// stream that run was queued on, after its kernels.
extern "C" int poyg_plan_status(poyg_plan *P, int32_t *bad) {
  if (!P || !bad) return POYG_EINVAL;
  if (P->triple) return poyg_triple_plan_status(P->triple, bad);
  *bad = 0;
  if (!P->last_stream) return POYG_OK;   // never run (or nothing to run)
  HIP_TRY(hipMemcpyAsync(bad, P->status.p, sizeof(int32_t), hipMemcpyDeviceToHost, P->last_stream));
  HIP_TRY(hipStreamSynchronize(P->last_stream));
  return POYG_OK;
}

// ---------------------------------------------------------------------------
// Per-call entry points: a one-pair plan over host data.
// ---------------------------------------------------------------------------
namespace {

// The host-sequence route: one upload (descriptor block, sequences and the
// run's initialised counters from pinned staging), the plan's kernels (no
// start kernel), one kernel that copies every result into the context's
// pinned result buffer, one synchronization.  The results
// (and the plan, for readers of its direction band) stay valid until the next
// call on the context.
struct HostRun {
  poyg_plan *plan = nullptr;
  bool idle = false;   // every queued use of the plan's buffers has completed
  const int32_t *h_cost = nullptr, *h_len = nullptr;
  const uint8_t *h_out = nullptr;
  int32_t failed = 0;
  ~HostRun() {
    if (plan) plan->idle = idle;
    poyg_plan_destroy(plan);
  }
  int run(poyg_ctx *ctx, const poyg_cm *cm, int mode, int32_t n, PerCallSeqs &in, const int64_t *off_a,
          const int32_t *len_a, const int64_t *off_b, const int32_t *len_b, const int32_t *dw) {
    hipStream_t s = ctx->stream;
    HIP_TRY(hipStreamSynchronize(s));   // the staging of an earlier call is free
    int rc = plan_pairs_impl(ctx, cm, mode, n, off_a, len_a, off_b, len_b, dw, &plan, &in);
    if (rc) return rc;
    const int slots = plan->slots;
    // device results (in the plan's block): costs (n, rounded to 4 words),
    // then output lengths, then the output sequences
    const int64_t n_cost = ((int64_t)n + 3) & ~(int64_t)3;
    const int64_t small_words = in.small_words;
    const int64_t out_bytes = slots ? plan->out_bytes : 0;   // a multiple of 16
    const size_t res_bytes = 64 + 4 * (size_t)small_words + (size_t)out_bytes;
    uint8_t *pin = ctx_pinned(ctx->pin_io, ctx->pin_io_cap, res_bytes, &ctx->pin_io_dev);
    if (!pin) {
      set_error("pinned staging of %zu bytes failed", res_bytes);
      return POYG_ENOMEM;
    }
    int32_t *d_small = in.small_dev;
    rc = poyg_plan_run(plan, in.dev, d_small, in.out_dev, slots ? d_small + n_cost : nullptr, s);
    if (rc) return rc;
    // pinned: word 0 the failed-pair count, from word 16 the device results,
    // then the output sequences
    HIP_TRY(launch_collect_results((const uint32_t *)d_small, small_words, (const int32_t *)plan->status.p,
                                   (const uint32_t *)in.out_dev, out_bytes / 4, (uint32_t *)ctx->pin_io_dev, s));
    HIP_TRY(hipStreamSynchronize(s));
    idle = true;
    const int32_t *h = (const int32_t *)pin;
    failed = h[0];
    h_cost = h + 16;
    h_len = h + 16 + n_cost;
    h_out = pin + 64 + 4 * small_words;
    return POYG_OK;
  }
};

// A one-pair run of the host-sequence route (the per-call entry points).
struct OnePair : HostRun {
  int32_t h_cost1 = 0;
  int run(poyg_ctx *ctx, const poyg_cm *cm, int mode, const uint8_t *a, int la, const uint8_t *b, int lb,
          int dw) {
    if (!a || !b || la < 1 || lb < 1) {
      set_error("sequence arguments must be non-empty (leading gap included)");
      return POYG_EINVAL;
    }
    int64_t offa = 0, offb = round16(la);
    PerCallSeqs in{a, b, la, lb, offb};
    int rc = HostRun::run(ctx, cm, mode, 1, in, &offa, &la, &offb, &lb, &dw);
    if (rc) return rc;
    h_cost1 = h_cost[0];
    if (failed) {
      set_error("backtrace left the filled band");
      return POYG_EBAND;
    }
    return POYG_OK;
  }
  // copy slot k right-aligned into dst (capacity cap)
  int take(int k, uint8_t *dst, int cap, int *olen) {
    const int c = (int)plan->out_bytes / plan->slots;
    const int l = h_len[k];
    if (l > cap) {
      set_error("output capacity %d is smaller than the result (%d)", cap, l);
      return POYG_EINVAL;
    }
    memcpy(dst + cap - l, h_out + (size_t)k * c + (c - l), l);
    if (olen) *olen = l;
    return POYG_OK;
  }
};

}  // namespace

extern "C" {

int poyg_run_pairs_host(poyg_ctx *ctx, const poyg_cm *cm, int mode, int32_t n, const uint8_t *seq,
                        int64_t seq_bytes, const int64_t *off_a, const int32_t *len_a, const int64_t *off_b,
                        const int32_t *len_b, const int32_t *deltaw, int32_t *cost, uint8_t *out, int64_t out_cap,
                        int32_t *out_len, int64_t *cells, int32_t *failed) {
  if (!ctx || !cm || n < 0 || seq_bytes < 0 || (n > 0 && (!seq || !off_a || !len_a || !off_b || !len_b || !cost)) ||
      mode < POYG_MODE_LINEAR_COST || mode > POYG_MODE_AFFINE_ALIGN) {
    set_error("poyg_run_pairs_host: bad arguments");
    return POYG_EINVAL;
  }
  if (cells) *cells = 0;
  if (failed) *failed = 0;
  if (n == 0) return POYG_OK;
  for (int32_t k = 0; k < n; ++k)
    if (off_a[k] < 0 || off_b[k] < 0 || len_a[k] < 1 || len_b[k] < 1 || off_a[k] + len_a[k] > seq_bytes ||
        off_b[k] + len_b[k] > seq_bytes) {
      set_error("poyg_run_pairs_host: pair %d lies outside the %lld sequence bytes", k, (long long)seq_bytes);
      return POYG_EINVAL;
    }
  // linear modes without deltaw: the default band of Sequence.Align.cost_2
  // (src/sequence.ml:691-713, ?deltaw absent): the larger count of
  // gap-bearing elements, plus half of 10% of the longer length when the
  // lengths differ by less than that, else 2
  std::vector<int32_t> dw_default;
  const bool linear = mode == POYG_MODE_LINEAR_COST || mode == POYG_MODE_LINEAR_ALIGN;
  if (linear && !deltaw) {
    dw_default.resize(n);
    auto gaps = [&](int64_t off, int32_t len) {
      int64_t g = 0;
      for (int32_t x = 0; x < len; ++x) g += (seq[off + x] & cm->gap) != 0;
      return g;
    };
    for (int32_t k = 0; k < n; ++k) {
      const int64_t longer = std::max(len_a[k], len_b[k]), dif = std::abs((int64_t)len_a[k] - len_b[k]);
      const int64_t lower = (int64_t)((double)longer * 0.10);
      dw_default[k] = (int32_t)(std::max(gaps(off_a[k], len_a[k]), gaps(off_b[k], len_b[k])) +
                                (dif < lower ? lower / 2 : 2));
    }
    deltaw = dw_default.data();
  }
  HostRun hr;
  PerCallSeqs in{seq, nullptr, seq_bytes, 0, 0};
  int rc = hr.run(ctx, cm, mode, n, in, off_a, len_a, off_b, len_b, deltaw);
  if (rc) return rc;
  const int slots = hr.plan->slots;
  if (slots) {
    if (!out || !out_len || out_cap < hr.plan->out_bytes) {
      set_error("poyg_run_pairs_host: the outputs need %lld bytes", (long long)hr.plan->out_bytes);
      return POYG_EINVAL;
    }
    memcpy(out, hr.h_out, (size_t)hr.plan->out_bytes);
    memcpy(out_len, hr.h_len, sizeof(int32_t) * (size_t)n * slots);
  }
  memcpy(cost, hr.h_cost, sizeof(int32_t) * (size_t)n);
  if (cells) *cells = hr.plan->cells;
  if (failed) *failed = hr.failed;
  return POYG_OK;
}

int poyg_algn_simple_2(poyg_ctx *ctx, const poyg_cm *cm, const uint8_t *s1, int l1, const uint8_t *s2, int l2,
                       int deltawh, int *cost) {
  if (l1 < l2) {
    set_error("poyg_algn_simple_2: s1 must be at least as long as s2");
    return POYG_EINVAL;
  }
  OnePair op;
  int rc = op.run(ctx, cm, POYG_MODE_LINEAR_COST, s1, l1, s2, l2, deltawh);
  if (!rc && cost) *cost = op.h_cost1;
  return rc;
}

// The direction matrix of algn_fill_plane_2 as Matrix.default holds it
// after cost_2 (the reader stubs mat_CAML_get_value, src/matrices.c:236-249,
// and algn_CAML_create_backtrack, src/algn.c:4301-4328, index it as
// row * l2 + column): the band that the pair's fill wrote, un-phased into
// full rows; 0 outside the band (the reference leaves those cells as the
// previous fill left them).
int poyg_algn_direction_matrix(poyg_ctx *ctx, const poyg_cm *cm, const uint8_t *s1, int l1, const uint8_t *s2,
                               int l2, int deltawh, uint16_t *dm, int64_t cap, int *cost) {
  if (l1 < l2 || l2 < 1 || !dm || cap < (int64_t)l1 * l2) {
    set_error("poyg_algn_direction_matrix: s1 must be at least as long as s2, dm must hold l1 * l2 entries");
    return POYG_EINVAL;
  }
  OnePair op;
  int rc = op.run(ctx, cm, POYG_MODE_LINEAR_ALIGN, s1, l1, s2, l2, deltawh);
  if (rc) return rc;
  poyg_plan *P = op.plan;
  const LinGeom g = lin_geom(l1, l2, deltawh);
  int maxw = 0;
  int64_t cells = 0;
  lin_band_extent(l1, l2, g, maxw, cells);
  const int S = lin_dm_stride(maxw);
  std::vector<uint8_t> band((size_t)l1 * S);
  HIP_TRY(hipMemcpy(band.data(), P->dm.p, band.size(), hipMemcpyDeviceToHost));
  // rows of the systolic fill start at phase (L + lo - lo_first) mod 4,
  // L the row's place in its strip (64 rows; 128 in the packed fill); the
  // row kernels (s2 longer than 2,048) at phase 0
  const int strip = P->n_c16 ? 128 : 64;
  memset(dm, 0, sizeof(uint16_t) * (size_t)l1 * l2);
  for (int i = 0; i < l1; ++i) {
    int lo, hi;
    lin_row_host(i, l1, l2, g, lo, hi);
    int phase = 0;
    if (i >= 1 && l2 <= kLinStageHost) {
      const int L = (i - 1) % strip;
      int slo, shi;
      lin_row_host(i - L, l1, l2, g, slo, shi);
      phase = (L + lo - slo) & 3;
    }
    for (int j = lo; j <= hi; ++j) dm[(size_t)i * l2 + j] = band[(size_t)i * S + phase + (j - lo)];
  }
  if (cost) *cost = op.h_cost1;
  return POYG_OK;
}

int poyg_algn_align_2d(poyg_ctx *ctx, const poyg_cm *cm, const uint8_t *s1, int l1, const uint8_t *s2, int l2,
                       int deltawh, int swaped, uint8_t *r1, uint8_t *r2, int cap, int *rlen, int *cost) {
  // A pair plan puts the longer sequence on the rows and backtracks with
  // swaped = (rows sequence is the caller's first), as
  // Sequence.Align.create_edited_2 does.  The stub's (s1 longer, swaped)
  // maps onto a plan of (s1, s2) when swaped, else of (s2, s1).
  if (l1 < l2) {
    set_error("poyg_algn_align_2d: s1 must be at least as long as s2");
    return POYG_EINVAL;
  }
  if (!swaped && l1 == l2) {
    set_error("poyg_algn_align_2d: swaped = false on equal lengths is not supported");
    return POYG_EUNSUPPORTED;
  }
  OnePair op;
  int rc = swaped ? op.run(ctx, cm, POYG_MODE_LINEAR_ALIGN, s1, l1, s2, l2, deltawh)
                  : op.run(ctx, cm, POYG_MODE_LINEAR_ALIGN, s2, l2, s1, l1, deltawh);
  if (rc) return rc;
  if (cost) *cost = op.h_cost1;
  rc = op.take(swaped ? 0 : 1, r1, cap, rlen);
  if (!rc) rc = op.take(swaped ? 1 : 0, r2, cap, nullptr);
  return rc;
}

int poyg_algn_align_affine_3(poyg_ctx *ctx, const poyg_cm *cm, const uint8_t *si, int li, const uint8_t *sj,
                             int lj, uint8_t *median, uint8_t *medianwg, uint8_t *resi, uint8_t *resj, int cap,
                             int *lens, int *cost) {
  OnePair op;
  int rc = op.run(ctx, cm, POYG_MODE_AFFINE_ALIGN, si, li, sj, lj, 0);
  if (rc) return rc;
  if (cost) *cost = op.h_cost1;
  uint8_t *dst[4] = {median, medianwg, resi, resj};
  for (int k = 0; k < 4 && !rc; ++k) rc = op.take(k, dst[k], cap, lens ? lens + k : nullptr);
  return rc;
}

int poyg_algn_cost_affine_3(poyg_ctx *ctx, const poyg_cm *cm, const uint8_t *si, int li, const uint8_t *sj,
                            int lj, int *cost) {
  OnePair op;
  int rc = op.run(ctx, cm, POYG_MODE_AFFINE_COST, si, li, sj, lj, 0);
  if (!rc && cost) *cost = op.h_cost1;
  return rc;
}

int poyg_algn_align_3d(poyg_ctx *ctx, const poyg_cm3 *cm3, const uint8_t *s1, int l1, const uint8_t *s2, int l2,
                       const uint8_t *s3, int l3, uint8_t *r1, uint8_t *r2, uint8_t *r3, int cap, int *rlen,
                       int *cost) {
  if (!ctx || !cm3 || !s1 || !s2 || !s3 || l1 < 1 || l2 < 1 || l3 < 1) return POYG_EINVAL;
  int64_t o1 = 0, o2 = round16(l1), o3 = o2 + round16(l2);
  poyg_plan *P = nullptr;
  int rc = poyg_plan_triples(ctx, cm3, 1, &o1, &l1, &o2, &l2, &o3, &l3, &P);
  if (rc) return rc;
  DevBuf seq, dcost, out, len;
  const int64_t ob = poyg_plan_out_bytes(P);
  if (seq.alloc((size_t)(o3 + l3)) || dcost.alloc(4) || out.alloc((size_t)ob) || len.alloc(16)) {
    poyg_plan_destroy(P);
    return POYG_ENOMEM;
  }
  hipStream_t s = ctx->stream;
  std::vector<uint8_t> h_out(ob);
  int32_t h_len[4] = {0, 0, 0, 0}, h_cost = 0, bad = 0;
  hipError_t e = hipMemcpyAsync((uint8_t *)seq.p + o1, s1, l1, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync((uint8_t *)seq.p + o2, s2, l2, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync((uint8_t *)seq.p + o3, s3, l3, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) {
    rc = poyg_plan_run(P, (const uint8_t *)seq.p, (int32_t *)dcost.p, (uint8_t *)out.p, (int32_t *)len.p, s);
    if (rc == POYG_OK) {
      e = hipMemcpyAsync(&h_cost, dcost.p, 4, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipMemcpyAsync(h_out.data(), out.p, ob, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipMemcpyAsync(h_len, len.p, 16, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      if (e == hipSuccess) rc = poyg_plan_status(P, &bad);
    }
  }
  const int c4 = (int)(ob / 4);
  if (e != hipSuccess) {
    set_error("align_3d: %s", hipGetErrorString(e));
    rc = POYG_EDEVICE;
  } else if (rc == POYG_OK && bad) {
    set_error("backtrack_3d left the cube");
    rc = POYG_EBAND;
  } else if (rc == POYG_OK && h_len[0] > cap) {
    rc = POYG_EINVAL;
  } else if (rc == POYG_OK) {
    uint8_t *dst[3] = {r1, r2, r3};
    for (int q = 0; q < 3; ++q) memcpy(dst[q] + cap - h_len[q], h_out.data() + (size_t)q * c4 + (c4 - h_len[q]), h_len[q]);
    if (rlen) *rlen = h_len[0];
    if (cost) *cost = h_cost;
  }
  seq.release(); dcost.release(); out.release(); len.release();
  poyg_plan_destroy(P);
  return rc;
}

// ---- powell_3D_align (src/ukkCommon.c:110-151) ---------------------------
extern "C++" {
namespace poyg {
void uk_model(int mm, int go, int ge, UkModel *M);
hipError_t launch_ukk_geom(const UkRange &r, int ge, int c0, int c1, UkLevel *lev, UkWalk *w, int reset,
                           hipStream_t s);
hipError_t launch_ukk_level(const UkModel *M, const UkRange &r, int c, int64_t diags, const UkLevel *lev, uint32_t *U,
                            const char *A, const char *B, const char *C, int target_state, int target_dist,
                            UkWalk *w, hipStream_t s);
hipError_t launch_ukk_esweep(const UkModel *M, const UkRange &r, int c, int final_cost, int64_t diags,
                             const UkLevel *lev, const uint32_t *U, int *E, int64_t ring_words, int ring_n, int half,
                             UkWalk *w, hipStream_t s);
hipError_t launch_ukk_walk(const UkModel *M, const UkRange &r, int cpcost, int cpw, const UkLevel *lev,
                           const uint32_t *U, int first_pass, UkWalk *w, hipStream_t s);
hipError_t launch_ukk_trace(const UkModel *M, const UkRange &r, const UkLevel *lev, const uint32_t *U, const char *A,
                            const char *B, const char *C, char *cols, int cap, UkWalk *w, hipStream_t s);
}  // namespace poyg
}

namespace {

// The level store of one search range on the device: the level boxes
// (computed on the device by ukk_geom_kernel, mirrored here to size the
// store) and 16 words per diagonal and level.
struct UkStore {
  DevBuf lev, U;
  size_t lev_cap = 0, u_words = 0;
  std::vector<UkLevel> box;   // host mirror of the current range's boxes
  void release() {
    lev.release();
    U.release();
    lev_cap = u_words = 0;
  }
};

constexpr int64_t kUkStoreLimit = int64_t(24) << 30;   // bytes of cost levels one call may hold

// Boxes of levels [c0, c1] of range r appended to the host mirror; returns
// the store's words after level c1.
int64_t uk_boxes(std::vector<UkLevel> &box, const UkRange &r, int ge, int c0, int c1) {
  int64_t base = 0;
  if (c0 > r.scost) base = box[c0 - 1 - r.scost].base + (int64_t)box[c0 - 1 - r.scost].nab * box[c0 - 1 - r.scost].nac * kUkStates;
  box.resize((size_t)(c1 - r.scost + 1));
  for (int c = c0; c <= c1; ++c) {
    box[c - r.scost] = uk_level_box(r, ge, c, base);
    base += (int64_t)box[c - r.scost].nab * box[c - r.scost].nac * kUkStates;
  }
  return base;
}

// Room for `words` level words (keep: copy what the store holds).
int uk_reserve(UkStore &st, size_t levels, int64_t words, bool keep, hipStream_t s) {
  if (words * 4 > kUkStoreLimit) {
    set_error("powell_3d: the search needs %lld bytes of cost levels, above the %lld-byte limit", (long long)words * 4,
              (long long)kUkStoreLimit);
    return POYG_ENOMEM;
  }
  if (levels > st.lev_cap) {
    DevBuf nl;
    const size_t cap = std::max(levels, 2 * st.lev_cap);
    if (nl.alloc(cap * sizeof(UkLevel))) return POYG_ENOMEM;
    if (keep && st.lev_cap && hipMemcpyAsync(nl.p, st.lev.p, st.lev_cap * sizeof(UkLevel), hipMemcpyDeviceToDevice, s) != hipSuccess)
      return POYG_EDEVICE;
    st.lev.release();
    st.lev = nl;
    st.lev_cap = cap;
  }
  if ((size_t)words > st.u_words) {
    DevBuf nu;
    const size_t cap = std::max((size_t)words, 2 * st.u_words);
    if (nu.alloc(cap * 4)) return POYG_ENOMEM;
    if (keep && st.u_words && hipMemcpyAsync(nu.p, st.U.p, st.u_words * 4, hipMemcpyDeviceToDevice, s) != hipSuccess)
      return POYG_EDEVICE;
    st.U.release();
    st.U = nu;
    st.u_words = cap;
  }
  return POYG_OK;
}

}  // namespace

// The check-point Ukkonen search of powell_3D_align (doUkk,
// src/ukk.checkp.c:113-214): the first pass over costs 1, 2, ... on the
// device in chunks of levels until the end diagonal's best state reaches
// |A|; the check-point cost from the evaluated cells (ukk_esweep_kernel);
// then doUkkInLimits ranges (src/ukk.checkp.c:216-322) from a stack in the
// reference's order (a split range's later half first), each its levels,
// then its traceback (base case, at most CPwidth costs) or its split; the
// columns come back in reverse and go out after the common prefix
// (printTraceBack, src/ukk.checkp.c:456-500).
int poyg_powell_3d(poyg_ctx *ctx, const uint8_t *a, int la, const uint8_t *b, int lb, const uint8_t *c, int lc,
                   int mm, int go, int ge, uint8_t *ra, uint8_t *rb, uint8_t *rc, int cap, int *rlen, int *cost) {
  if (!ctx || !a || !b || !c || !ra || !rb || !rc || la < 1 || lb < 1 || lc < 1) return POYG_EINVAL;
  // the reference's preconditions (src/ukkCommon.c:132)
  if (mm <= 0 || go < 0 || ge <= 0) {
    set_error("powell_3d: needs mm > 0, go >= 0, ge > 0");
    return POYG_EINVAL;
  }
  if (cap < la + lb + lc) {
    set_error("powell_3d: output capacity %d below %d", cap, la + lb + lc);
    return POYG_EINVAL;
  }
  // the first set base bit of every mask after the leading gap
  // (copySequence, src/ukkCommon.c:88-108), each string followed by a 0
  const int n[3] = {la - 1, lb - 1, lc - 1};
  const uint8_t *src[3] = {a, b, c};
  std::vector<char> x[3];
  for (int q = 0; q < 3; ++q) {
    if (n[q] > kUkMaxLen) {
      set_error("powell_3d: sequence %d has %d positions, above %d", q + 1, n[q], kUkMaxLen);
      return POYG_EUNSUPPORTED;
    }
    x[q].assign((size_t)n[q] + 8, 0);
    for (int i = 0; i < n[q]; ++i) {
      const uint8_t m = src[q][i + 1];
      const char ch = (m & 1) ? 'A' : (m & 2) ? 'C' : (m & 4) ? 'G' : (m & 8) ? 'T' : 0;
      if (!ch) {
        set_error("powell_3d: position %d of sequence %d holds no base (the reference fails: This is impossible!)",
                  i + 1, q + 1);
        return POYG_EINVAL;
      }
      x[q][i] = ch;
    }
  }
  const int LA = n[0], LB = n[1], LC = n[2];
  HIP_TRY(hipSetDevice(ctx->device));   // (a caller's worker thread may start on another device)
  UkModel M;
  uk_model(mm, go, ge, &M);
  const int cpw = M.max_step;
  int start = 0;   // the common prefix (src/ukk.checkp.c:139-147)
  while (start < LA && x[0][start] == x[1][start] && x[0][start] == x[2][start]) ++start;
  const int col_cap = LA + LB + LC + 2;
  // an upper bound of the least cost: every column costs at most
  // maxSingleStep + 2 mismatches
  const int cost_bound = (LA + LB + LC + 1) * (M.max_step + 2 * mm) + 8;

  hipStream_t s = ctx->stream;
  DevBuf dx[3], dM, dW, dCols, dE;
  UkStore st;
  auto cleanup = [&]() {
    for (auto &d : dx) d.release();
    dM.release(); dW.release(); dCols.release(); dE.release();
    st.release();
  };
  for (int q = 0; q < 3; ++q)
    if (dx[q].alloc(x[q].size())) { cleanup(); return POYG_ENOMEM; }
  if (dM.alloc(sizeof(UkModel)) || dW.alloc(sizeof(UkWalk)) || dCols.alloc((size_t)3 * col_cap)) {
    cleanup();
    return POYG_ENOMEM;
  }
  hipError_t e = hipSuccess;
  for (int q = 0; q < 3 && e == hipSuccess; ++q)
    e = hipMemcpyAsync(dx[q].p, x[q].data(), x[q].size(), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(dM.p, &M, sizeof M, hipMemcpyHostToDevice, s);
  const char *A = (const char *)dx[0].p, *B = (const char *)dx[1].p, *C = (const char *)dx[2].p;
  const UkModel *dm = (const UkModel *)dM.p;
  UkWalk *dw = (UkWalk *)dW.p;
  UkWalk hw;
  int rc_ = POYG_OK;
  auto fail = [&](int code, const char *what) {
    if (code == POYG_EDEVICE) set_error("powell_3d: %s: %s", what, hipGetErrorString(e));
    else if (code == POYG_EBAND) set_error("powell_3d: %s", what);
    rc_ = code;
    return code;
  };
  auto read_walk = [&]() -> bool {
    if (e == hipSuccess) e = hipMemcpyAsync(&hw, dw, sizeof hw, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e == hipSuccess;
  };
  // ---- first pass: levels 0, 1, ... until the end's best state reaches |A|
  UkRange r0{0, 0, 0, 0, start, LA - LB, LA - LC, 0, 0, LA, LA, LB, LC};
  int final_cost = -1;
  {
    const int chunk = 32;
    st.box.clear();
    for (int c0 = 0; e == hipSuccess && final_cost < 0; c0 += chunk) {
      if (c0 > cost_bound) { cleanup(); return fail(POYG_EBAND, "the first pass did not reach the end"); }
      const int c1 = c0 + chunk - 1;
      const int64_t words = uk_boxes(st.box, r0, ge, c0, c1);
      const int rr = uk_reserve(st, (size_t)(c1 + 1), words, true, s);
      if (rr) { cleanup(); return rr; }
      if (e == hipSuccess) e = launch_ukk_geom(r0, ge, c0, c1, (UkLevel *)st.lev.p, dw, c0 == 0 ? 1 : -1, s);
      for (int cc = c0; cc <= c1 && e == hipSuccess; ++cc)
        e = launch_ukk_level(dm, r0, cc, (int64_t)st.box[cc].nab * st.box[cc].nac, (const UkLevel *)st.lev.p,
                             (uint32_t *)st.U.p, A, B, C, -1, LA, dw, s);
      if (!read_walk()) break;
      if (hw.reached <= c1) final_cost = hw.reached;
    }
  }
  if (e != hipSuccess) { cleanup(); return fail(POYG_EDEVICE, "first pass"); }
  // ---- the check-point cost: cells evaluated by the top-level calls
  r0.fcost = final_cost;
  {
    const int ring_n = M.max_step + 1;
    const int64_t ring_words = (int64_t)st.box[final_cost].nab * st.box[final_cost].nac * kUkStates;
    if (dE.alloc((size_t)(ring_words * ring_n * 4))) { cleanup(); return POYG_ENOMEM; }
    for (int cc = final_cost; cc >= 0 && e == hipSuccess; --cc)
      e = launch_ukk_esweep(dm, r0, cc, final_cost, (int64_t)st.box[cc].nab * st.box[cc].nac,
                            (const UkLevel *)st.lev.p, (const uint32_t *)st.U.p, (int *)dE.p, ring_words, ring_n,
                            LA / 2, dw, s);
    if (e == hipSuccess)
      e = launch_ukk_walk(dm, r0, 0, cpw, (const UkLevel *)st.lev.p, (const uint32_t *)st.U.p, 1, dw, s);
    if (!read_walk()) { cleanup(); return fail(POYG_EDEVICE, "check-point walk"); }
    dE.release();
    if (hw.error) { cleanup(); return fail(POYG_EBAND, "the first pass's walk found no end state"); }
  }
  // ---- doUkkInLimits ranges, later halves first
  std::vector<UkRange> stack;
  r0.fstate = hw.fstate;
  if (hw.f_cost < 0) {
    stack.push_back(r0);   // checked too late: the whole range again
  } else {
    const UkRange first{0, 0, 0, 0, start, hw.f_ab, hw.f_ac, hw.f_cost, hw.f_state, hw.f_dist, hw.f_dist,
                        hw.f_dist - hw.f_ab, hw.f_dist - hw.f_ac};
    const UkRange later{hw.f_ab, hw.f_ac, hw.f_cost, hw.f_state, hw.f_dist, r0.fab, r0.fac, final_cost, r0.fstate,
                        LA, LA, LB, LC};
    stack.push_back(first);
    stack.push_back(later);
  }
  size_t ranges = 0;
  while (!stack.empty() && e == hipSuccess) {
    const UkRange r = stack.back();
    stack.pop_back();
    if (++ranges > (size_t)8 * col_cap + 64) { cleanup(); return fail(POYG_EBAND, "the check-point recursion did not close"); }
    st.box.clear();
    const int64_t words = uk_boxes(st.box, r, ge, r.scost, r.fcost);
    const int rr = uk_reserve(st, (size_t)(r.fcost - r.scost + 1), words, false, s);
    if (rr) { cleanup(); return rr; }
    e = launch_ukk_geom(r, ge, r.scost, r.fcost, (UkLevel *)st.lev.p, dw, 0, s);
    for (int cc = r.scost; cc <= r.fcost && e == hipSuccess; ++cc)
      e = launch_ukk_level(dm, r, cc, (int64_t)st.box[cc - r.scost].nab * st.box[cc - r.scost].nac,
                           (const UkLevel *)st.lev.p, (uint32_t *)st.U.p, A, B, C, r.fstate, r.fdist, dw, s);
    if (e != hipSuccess) break;
    if (r.fcost - r.scost <= cpw) {
      e = launch_ukk_trace(dm, r, (const UkLevel *)st.lev.p, (const uint32_t *)st.U.p, A, B, C, (char *)dCols.p,
                           col_cap, dw, s);
      continue;
    }
    const int cpcost = (r.fcost + r.scost - cpw + 1) / 2;
    e = launch_ukk_walk(dm, r, cpcost, cpw, (const UkLevel *)st.lev.p, (const uint32_t *)st.U.p, 0, dw, s);
    if (!read_walk()) break;
    if (hw.error || hw.f_cost < 0) { cleanup(); return fail(POYG_EBAND, "a range's end has no check point"); }
    const UkRange first{r.sab, r.sac, r.scost, r.sstate, r.sdist, hw.f_ab, hw.f_ac, hw.f_cost, hw.f_state, hw.f_dist,
                        hw.f_dist, hw.f_dist - hw.f_ab, hw.f_dist - hw.f_ac};
    const UkRange later{hw.f_ab, hw.f_ac, hw.f_cost, hw.f_state, hw.f_dist, r.fab, r.fac, hw.reached, r.fstate,
                        r.fdist, r.endA, r.endB, r.endC};
    stack.push_back(first);
    stack.push_back(later);
  }
  std::vector<char> cols((size_t)3 * col_cap);
  if (read_walk() && hw.ncols > 0)
    e = hipMemcpy(cols.data(), dCols.p, (size_t)3 * hw.ncols, hipMemcpyDeviceToHost);
  if (e != hipSuccess) { cleanup(); return fail(POYG_EDEVICE, "search"); }
  cleanup();
  if (hw.error) return fail(POYG_EBAND, "a traceback left its range");
  const int len = 1 + start + hw.ncols;
  if (len > cap) {
    set_error("powell_3d: alignment of %d columns does not fit", len);
    return POYG_EINVAL;
  }
  auto mask = [](char ch) -> uint8_t {
    return (uint8_t)(ch == 'A' ? 1 : ch == 'C' ? 2 : ch == 'G' ? 4 : ch == 'T' ? 8 : 16);
  };
  uint8_t *dst[3] = {ra, rb, rc};
  for (int q = 0; q < 3; ++q) {
    uint8_t *o = dst[q] + cap - len;
    o[0] = 16;
    for (int i = 0; i < start; ++i) o[1 + i] = mask(x[0][i]);
    for (int i = 0; i < hw.ncols; ++i) o[1 + start + i] = mask(cols[(size_t)3 * (hw.ncols - 1 - i) + q]);
  }
  if (rlen) *rlen = len;
  if (cost) *cost = final_cost;
  return rc_;
}

// ---- host-side column rules (O(L) per call) ------------------------------
static int put_rev(uint8_t *out, int cap, int &n, int v) {
  if (n >= cap) return POYG_EINVAL;
  out[cap - ++n] = (uint8_t)v;
  return 0;
}

int poyg_algn_median_2_no_gaps(const poyg_cm *cm, const uint8_t *a, const uint8_t *b, int len, uint8_t *out,
                               int cap, int *olen) {
  int n = 0;
  for (int i = len - 1; i >= 0; --i) {
    const int m = cm->median[(a[i] << cm->lcm) + b[i]];
    if (m != cm->gap && put_rev(out, cap, n, m)) return POYG_EINVAL;
  }
  if (put_rev(out, cap, n, cm->gap)) return POYG_EINVAL;
  *olen = n;
  return POYG_OK;
}

int poyg_algn_median_2_with_gaps(const poyg_cm *cm, const uint8_t *a, const uint8_t *b, int len, uint8_t *out,
                                 int cap, int *olen) {
  int n = 0;
  for (int i = len - 1; i >= 0; --i)
    if (put_rev(out, cap, n, cm->median[(a[i] << cm->lcm) + b[i]])) return POYG_EINVAL;
  *olen = n;
  return POYG_OK;
}

int poyg_algn_union(const uint8_t *a, const uint8_t *b, int len, uint8_t *out, int cap, int *olen) {
  int n = 0;
  for (int i = len - 1; i >= 0; --i)
    if (put_rev(out, cap, n, a[i] | b[i])) return POYG_EINVAL;
  *olen = n;
  return POYG_OK;
}

// algn_ancestor_2 (src/algn.c:4126-4147) with algn_correct_blocks_affine
// (:4080-4119) and algn_remove_gaps (:4058-4077).
int poyg_algn_ancestor_2(const poyg_cm *cm, const uint8_t *a, const uint8_t *b, int len, uint8_t *out, int cap,
                         int *olen) {
  const int gap = cm->gap;
  const bool keep_all = cm->combinations && cm->cost_model == 1;
  int n = 0;
  for (int i = len - 1; i >= 0; --i) {
    const int m = cm->median[(a[i] << cm->lcm) + b[i]];
    if ((keep_all || m != gap) && put_rev(out, cap, n, m)) return POYG_EINVAL;
  }
  if (!keep_all) {
    if (n == 0 || out[cap - n] != gap)
      if (put_rev(out, cap, n, gap)) return POYG_EINVAL;
    *olen = n;
    return POYG_OK;
  }
  uint8_t *s = out + cap - n;
  int extending = 0, inside = 0, prev_block = 0;
  for (int i = 0; i < n; ++i) {
    const int ab = a[i], bb = b[i];
    int sb = s[i];
    if (!inside && (!(ab & gap) || !(bb & gap))) inside = 0;
    else if (inside && (!(ab & gap) || !(bb & gap))) inside = 0;
    else if (((ab & gap) || (bb & gap)) && ((ab != gap) || (bb != gap))) inside = 1;
    else inside = 0;
    if (((gap & ab) || (gap & bb)) && !(sb & gap) && !extending) {
      prev_block = inside;
      extending = 1;
    } else if ((gap & ab) && (gap & bb) && (sb & gap) && (sb != gap) && extending && inside && !prev_block) {
      sb = (~gap) & sb;
      prev_block = 0;
    } else if ((gap & ab) && (gap & bb) && extending == 1) {
      prev_block = inside;
      extending = 0;
    }
    s[i] = (uint8_t)sb;
  }
  int src = cap - 1, dst = cap - 1, kept = 0;
  for (int i = n - 1; i >= 0; --i, --src)
    if (out[src] != gap) { out[dst--] = out[src]; ++kept; }
  if (put_rev(out, cap, kept, gap)) return POYG_EINVAL;
  *olen = kept;
  return POYG_OK;
}

int poyg_algn_median_3(const poyg_cm3 *cm3, const uint8_t *a, const uint8_t *b, const uint8_t *c, int len,
                       uint8_t *out, int cap, int *olen) {
  // algn_get_median_3d (src/algn.c:4160-4173) never advances its end
  // pointers: every position gets the median of the last column.
  if (len > cap) return POYG_EINVAL;
  const int m = len > 0 ? cm3->median[(((a[len - 1] << cm3->lcm) + b[len - 1]) << cm3->lcm) + c[len - 1]] : 0;
  for (int k = 1; k <= len; ++k) out[cap - k] = (uint8_t)m;
  *olen = len;
  return POYG_OK;
}

static int calc_aligned(const poyg_cm *c, const uint8_t *s1, const uint8_t *s2, int len, const int *tab) {
  // algn_calculate_from_2_aligned, src/algn.c:3316-3370
  int res = 0, gap_row = 0;
  const int gap = c->gap, go = c->gap_open;
  const bool comb = c->combinations != 0;
  if (len == 0) return 0;
  int i = ((comb && (gap & s1[0]) && (gap & s2[0])) || (!comb && gap == s1[0] && gap == s2[0])) ? 1 : 0;
  for (; i < len; ++i) {
    const int x = s1[i], y = s2[i];
    if (gap_row == 0) {
      if ((comb && (x & gap) && !(y & gap)) || (!comb && x == gap)) { res += go; gap_row = 1; }
      else if ((comb && (y & gap) && !(x & gap)) || (!comb && y == gap)) { res += go; gap_row = 2; }
    } else if (gap_row == 1) {
      if ((comb && !(x & gap)) || (!comb && x != gap)) {
        if ((comb && (y & gap) && !(x & gap)) || (!comb && y == gap)) { res += go; gap_row = 2; }
        else gap_row = 0;
      }
    } else {
      if ((comb && !(y & gap)) || (!comb && y != gap)) {
        if ((comb && (x & gap)) || (!comb && x == gap)) { res += go; gap_row = 1; }
        else gap_row = 0;
      }
    }
    res += tab[(x << c->lcm) + y];
  }
  return res;
}

int poyg_algn_worst_2(const poyg_cm *cm, const uint8_t *a, const uint8_t *b, int len, int *cost) {
  *cost = calc_aligned(cm, a, b, len, cm->worst);
  return POYG_OK;
}
int poyg_algn_verify_2(const poyg_cm *cm, const uint8_t *a, const uint8_t *b, int len, int *cost) {
  *cost = calc_aligned(cm, a, b, len, cm->cost);
  return POYG_OK;
}

// ---- Fitch ---------------------------------------------------------------
static bool fitch_width_ok(int w) { return w == 1 || w == 2 || w == 4; }

int poyg_fitch_down_w(poyg_ctx *ctx, int width, uint8_t *d_chars, int32_t len, int32_t n, const int32_t *d_ia,
                      const int32_t *d_ib, const int32_t *d_ires, int64_t *d_added, void *stream) {
  if (!ctx || n < 0 || len < 0 || !fitch_width_ok(width)) {
    set_error("poyg_fitch_down_w: bad arguments (width %d)", width);
    return POYG_EINVAL;
  }
  if (n == 0) return POYG_OK;
  hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
  HIP_TRY(launch_fitch_down(width, d_chars, len, n, d_ia, d_ib, d_ires, d_added, s));
  return POYG_OK;
}

int poyg_fitch_up_w(poyg_ctx *ctx, int width, uint8_t *d_chars, int32_t len, int32_t n, const int32_t *d_iA,
                    const int32_t *d_iN, const int32_t *d_iD1, const int32_t *d_iD2, const int32_t *d_iF,
                    void *stream) {
  if (!ctx || n < 0 || len < 0 || !fitch_width_ok(width)) {
    set_error("poyg_fitch_up_w: bad arguments (width %d)", width);
    return POYG_EINVAL;
  }
  if (n == 0) return POYG_OK;
  hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
  HIP_TRY(launch_fitch_up(width, d_chars, len, n, d_iA, d_iN, d_iD1, d_iD2, d_iF, s));
  return POYG_OK;
}

int poyg_fitch_down(poyg_ctx *ctx, uint8_t *d_chars, int32_t len, int32_t n, const int32_t *d_ia,
                    const int32_t *d_ib, const int32_t *d_ires, int64_t *d_added, void *stream) {
  return poyg_fitch_down_w(ctx, 1, d_chars, len, n, d_ia, d_ib, d_ires, d_added, stream);
}

int poyg_fitch_up(poyg_ctx *ctx, uint8_t *d_chars, int32_t len, int32_t n, const int32_t *d_iA,
                  const int32_t *d_iN, const int32_t *d_iD1, const int32_t *d_iD2, const int32_t *d_iF,
                  void *stream) {
  return poyg_fitch_up_w(ctx, 1, d_chars, len, n, d_iA, d_iN, d_iD1, d_iD2, d_iF, stream);
}

// Fitch down-pass medians of n row pairs on host arrays (one byte per
// character): one upload of both row sets from the context's pinned staging
// (with the zeroed change counts and the row index lists), the kernel, one
// download of the medians and counts into pinned memory, one
// synchronization.
int poyg_fitch_pairs_host(poyg_ctx *ctx, const uint8_t *a, const uint8_t *b, int32_t n, int32_t len, uint8_t *res,
                          int64_t *added) {
  if (!ctx || n < 0 || len < 0 || (n && (!a || !b || !res || !added))) {
    set_error("poyg_fitch_pairs_host: bad arguments");
    return POYG_EINVAL;
  }
  if (n == 0) return POYG_OK;
  const size_t nm = (size_t)n * len;
  const size_t add_off = (3 * nm + 255) & ~(size_t)255;   // the counts right after the medians
  const size_t idx_off = (add_off + 8 * (size_t)n + 255) & ~(size_t)255;
  const size_t bytes = idx_off + 12 * (size_t)n;
  uint8_t *stage = ctx_pinned(ctx->pin_plan, ctx->pin_plan_cap, bytes);
  uint8_t *back = ctx_pinned(ctx->pin_io, ctx->pin_io_cap, add_off + 8 * (size_t)n - 2 * nm, &ctx->pin_io_dev);
  if (!stage || !back) return POYG_ENOMEM;
  hipStream_t s = ctx->stream;
  // the staging buffers are free once the context's stream is idle
  HIP_TRY(hipStreamSynchronize(s));
  memcpy(stage, a, nm);
  memcpy(stage + nm, b, nm);
  memset(stage + 2 * nm, 0, idx_off - 2 * nm);
  int32_t *idx = (int32_t *)(stage + idx_off);
  for (int32_t k = 0; k < 3 * n; ++k) idx[k] = k;
  DevBuf d;
  if (d.alloc(bytes)) return POYG_ENOMEM;
  uint8_t *dp = (uint8_t *)d.p;
  hipError_t e = hipMemcpyAsync(dp, stage, bytes, hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = launch_fitch_down(1, dp, len, n, (const int32_t *)(dp + idx_off), (const int32_t *)(dp + idx_off) + n,
                          (const int32_t *)(dp + idx_off) + 2 * n, (int64_t *)(dp + add_off), s);
  if (e == hipSuccess) e = hipMemcpyAsync(back, dp + 2 * nm, add_off + 8 * (size_t)n - 2 * nm, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  d.release(e == hipSuccess);
  if (e != hipSuccess) {
    set_error("poyg_fitch_pairs_host: %s", hipGetErrorString(e));
    return POYG_EDEVICE;
  }
  memcpy(res, back, nm);
  memcpy(added, back + (add_off - 2 * nm), 8 * (size_t)n);
  return POYG_OK;
}

// One node on host arrays of `len` characters of `width` bytes: two inputs
// (down pass) or four (up pass).
static int fitch_host(poyg_ctx *ctx, int width, const void *const *in, int nin, void *res, int len, int64_t *added) {
  if (!ctx || len < 0 || !fitch_width_ok(width)) {
    set_error("nonadd median: bad arguments (width %d)", width);
    return POYG_EINVAL;
  }
  DevBuf chars, idx, acc;
  const int rows = nin + 1;
  const size_t row = (size_t)len * width;
  if (chars.alloc((size_t)rows * row) || idx.alloc(sizeof(int32_t) * 8) || acc.alloc(sizeof(int64_t))) {
    chars.release(); idx.release(); acc.release();
    return POYG_ENOMEM;
  }
  hipStream_t s = ctx->stream;
  int rc = POYG_OK;
  int32_t h_idx[8] = {0, 1, 2, 3, 4, 0, 0, 0};
  int64_t zero = 0;
  hipError_t e = hipSuccess;
  for (int k = 0; k < nin && e == hipSuccess; ++k)
    e = hipMemcpyAsync((uint8_t *)chars.p + (size_t)k * row, in[k], row, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(idx.p, h_idx, sizeof h_idx, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(acc.p, &zero, sizeof zero, hipMemcpyHostToDevice, s);
  const int32_t *I = (const int32_t *)idx.p;
  if (e == hipSuccess) {
    if (nin == 2) e = launch_fitch_down(width, (uint8_t *)chars.p, len, 1, I, I + 1, I + 2, (int64_t *)acc.p, s);
    else e = launch_fitch_up(width, (uint8_t *)chars.p, len, 1, I, I + 1, I + 2, I + 3, I + 4, s);
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(res, (uint8_t *)chars.p + (size_t)nin * row, row, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess && added) e = hipMemcpyAsync(added, acc.p, sizeof(int64_t), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    set_error("fitch: %s", hipGetErrorString(e));
    rc = POYG_EDEVICE;
  }
  chars.release(); idx.release(); acc.release();
  return rc;
}

int poyg_nonadd_median_w(poyg_ctx *ctx, int width, const void *a, const void *b, void *res, int len,
                         int64_t *added) {
  const void *in[2] = {a, b};
  return fitch_host(ctx, width, in, 2, res, len, added);
}

int poyg_nonadd_median_3_w(poyg_ctx *ctx, int width, const void *A, const void *N, const void *D1, const void *D2,
                           void *F, int len) {
  const void *in[4] = {A, N, D1, D2};
  return fitch_host(ctx, width, in, 4, F, len, nullptr);
}

int poyg_nonadd_median(poyg_ctx *ctx, const uint8_t *a, const uint8_t *b, uint8_t *res, int len, int64_t *added) {
  return poyg_nonadd_median_w(ctx, 1, a, b, res, len, added);
}

int poyg_nonadd_median_3(poyg_ctx *ctx, const uint8_t *A, const uint8_t *N, const uint8_t *D1, const uint8_t *D2,
                         uint8_t *F, int len) {
  return poyg_nonadd_median_3_w(ctx, 1, A, N, D1, D2, F, len);
}

}  // extern "C"
