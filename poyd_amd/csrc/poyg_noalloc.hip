// Batched linear alignment costs that allocate no device memory: for
// callers that cannot call hipMalloc (code under HIP graph capture, tight
// serving loops, harnesses that forbid it).  The same kernels as a
// POYG_MODE_LINEAR_COST plan (lin_prep_kernel, lin_sys_kernel), with every
// buffer such a plan allocates -- the cost matrix copy, the pair
// descriptors, the work list, the prefix sums, the counters and the status
// -- carved out of one file-scope device array instead.  Every call on a
// device shares that device's array: calls that overlap on one device race
// (the entry point waits for its own work before it returns).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <mutex>

#include "poyg_band.h"
#include "poyg_internal.h"

namespace poyg {

hipError_t launch_reset(int *counters, int n, hipStream_t s);
hipError_t launch_prep(bool affine, int grid, const PairMeta *meta, int n, const uint8_t *seq, const DevTcm *tcm,
                       int *aux, hipStream_t s);
hipError_t launch_lin_sys(bool with_dm, int grid, const PairMeta *meta, const int32_t *order, int n, int *counter,
                          const uint8_t *seq, const DevTcm *tcm, uint8_t *dm, const int *aux, int32_t *cost,
                          int32_t *status, int spin_limit, hipStream_t s);
int lin_sys_blocks_per_cu(bool with_dm);

// The scratch array: 32 MiB per device, held until the process ends.
constexpr int64_t kScratchBytes = (int64_t)32 << 20;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

constexpr int kNoAllocStage = 2048;   // the systolic fill's longest column sequence (kLinStage)
constexpr int kMaxDevices = 64;

namespace {
std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once
// per device.
hipError_t scratch_of(int device, uint8_t **out) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> g(g_mu);
  if (!g_scratch[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[device] = p;
  }
  *out = (uint8_t *)g_scratch[device];
  return hipSuccess;
}

int64_t up256(int64_t v) { return (v + 255) & ~(int64_t)255; }
}  // namespace

// Layout of the scratch for n pairs (byte offsets).
struct NoAllocLayout {
  int64_t tcm, ctr, status, order, meta, in_off, in_len, in_dw, aux, end;
  NoAllocLayout(int n, int64_t aux_ints) {
    tcm = 0;
    ctr = up256(tcm + (int64_t)sizeof(DevTcm));
    status = up256(ctr + 8 * 4);
    order = up256(status + 4 * ((int64_t)kStatusFlags + n));
    meta = up256(order + 4 * (int64_t)n);
    in_off = up256(meta + (int64_t)sizeof(PairMeta) * n);
    in_len = up256(in_off + 16 * (int64_t)n);
    in_dw = up256(in_len + 8 * (int64_t)n);
    aux = up256(in_dw + 4 * (int64_t)n);
    end = aux + 4 * aux_ints;
  }
};

// Pair descriptors on the device (one workgroup): the longer sequence
// indexes the rows (Sequence.Align.cost_2), the band geometry of
// algn_fill_plane_2 (src/algn.c:872-968), prefix-sum regions of
// round64(2 lc) ints in pair order (a block scan), the identity work list.
__global__ __launch_bounds__(1024) void noalloc_plan_kernel(int n, const int64_t *__restrict__ off,
                                                            const int32_t *__restrict__ len,
                                                            const int32_t *__restrict__ dw, PairMeta *meta,
                                                            int32_t *order) {
  __shared__ int64_t part[1024];
  const int t = threadIdx.x, T = blockDim.x;
  const int per = (n + T - 1) / T, k0 = t * per, k1 = min(n, k0 + per);
  int64_t sum = 0;
  for (int k = k0; k < k1; ++k) {
    const int lc = min(len[k], len[n + k]);
    sum += (2 * (int64_t)lc + 63) & ~(int64_t)63;
  }
  part[t] = sum;
  __syncthreads();
  for (int d = 1; d < T; d <<= 1) {   // inclusive scan of the per-thread sums
    const int64_t v = t >= d ? part[t - d] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int64_t at = part[t] - sum;
  for (int k = k0; k < k1; ++k) {
    PairMeta m;
    memset(&m, 0, sizeof m);
    const int la = len[k], lb = len[n + k];
    m.swap = la < lb;
    m.off_r = m.swap ? off[n + k] : off[k];
    m.off_c = m.swap ? off[k] : off[n + k];
    m.lr = m.swap ? lb : la;
    m.lc = m.swap ? la : lb;
    const LinGeom g = lin_geom(m.lr, m.lc, dw[k]);
    m.p0 = g.cs;
    m.p1 = g.w;
    m.p2 = g.h;
    m.aux_off = at;
    at += (2 * (int64_t)m.lc + 63) & ~(int64_t)63;
    meta[k] = m;
    order[k] = k;
  }
}

}  // namespace poyg

using namespace poyg;

#define HIP_TRYN(expr)                                                                           \
  do {                                                                                           \
    hipError_t e_ = (expr);                                                                      \
    if (e_ != hipSuccess) {                                                                      \
      set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__);      \
      return POYG_EDEVICE;                                                                       \
    }                                                                                            \
  } while (0)

extern "C" void *poyg_ctx_stream(poyg_ctx *ctx);
extern "C" int poyg_ctx_device(poyg_ctx *ctx);

extern "C" int64_t poyg_noalloc_scratch_bytes(void) { return kScratchBytes; }

// See include/poyg.h.
extern "C" int poyg_pairs_cost_linear_noalloc(poyg_ctx *ctx, const poyg_cm *cm, int32_t n, const int64_t *off_a,
                                              const int32_t *len_a, const int64_t *off_b, const int32_t *len_b,
                                              const int32_t *deltaw, const uint8_t *d_seq, int32_t *d_cost,
                                              void *stream) {
  if (!ctx || !cm || n < 0 || (n > 0 && (!off_a || !len_a || !off_b || !len_b || !deltaw || !d_seq || !d_cost))) {
    set_error("poyg_pairs_cost_linear_noalloc: bad arguments");
    return POYG_EINVAL;
  }
  if (cm->cost_model == 1 || !cm->combinations || cm->lcm != 5 || cm->gap != kGap) {
    set_error("poyg_pairs_cost_linear_noalloc: a linear 5-bit DNA combination matrix is required");
    return POYG_EUNSUPPORTED;
  }
  if (n == 0) return POYG_OK;
  // sizes first: every pair's shorter sequence fits the systolic fill, and
  // the descriptors and prefix sums fit the scratch
  int64_t aux_ints = 0, longest = 0;
  for (int k = 0; k < n; ++k) {
    const int la = len_a[k], lb = len_b[k];
    if (la < 1 || lb < 1) {
      set_error("poyg_pairs_cost_linear_noalloc: pair %d has an empty sequence (the leading gap is required)", k);
      return POYG_EINVAL;
    }
    const int lc = std::min(la, lb);
    if (lc > kNoAllocStage) {
      set_error("poyg_pairs_cost_linear_noalloc: pair %d: the shorter sequence has %d positions, more than the "
                "%d this entry point takes", k, lc, kNoAllocStage);
      return POYG_EUNSUPPORTED;
    }
    aux_ints += (2 * (int64_t)lc + 63) & ~(int64_t)63;
    longest = std::max<int64_t>(longest, (int64_t)la + lb + 2);
  }
  const NoAllocLayout lay(n, aux_ints);
  if (lay.end > kScratchBytes) {
    set_error("poyg_pairs_cost_linear_noalloc: %d pairs need %lld bytes of scratch, more than its %lld", n,
              (long long)lay.end, (long long)kScratchBytes);
    return POYG_EUNSUPPORTED;
  }
  DevTcm dt;
  to_dev_tcm(cm, &dt);
  int64_t cmax = 0;
  bool negative = false;
  for (int x = 0; x < 32 * 32; ++x) {
    negative |= dt.cost[x] < 0;
    cmax = std::max<int64_t>(cmax, dt.cost[x]);
  }
  for (int x = 0; x < 32; ++x) {
    negative |= dt.prepend[x] < 0 || dt.tail[x] < 0;
    cmax = std::max<int64_t>(cmax, std::max(dt.prepend[x], dt.tail[x]));
  }
  if (negative || cmax * longest >= ((int64_t)1 << 30)) {
    set_error("poyg_pairs_cost_linear_noalloc: costs must be non-negative and an alignment must cost less than 2^30");
    return POYG_EINVAL;
  }
  const int device = poyg_ctx_device(ctx);
  HIP_TRYN(hipSetDevice(device));
  uint8_t *S = nullptr;
  HIP_TRYN(scratch_of(device, &S));
  hipStream_t s = stream ? (hipStream_t)stream : (hipStream_t)poyg_ctx_stream(ctx);
  // inputs into the scratch (the host arrays and the table stay valid until
  // the final synchronisation below)
  HIP_TRYN(hipMemcpyAsync(S + lay.tcm, &dt, sizeof dt, hipMemcpyHostToDevice, s));
  HIP_TRYN(hipMemcpyAsync(S + lay.in_off, off_a, 8 * (size_t)n, hipMemcpyHostToDevice, s));
  HIP_TRYN(hipMemcpyAsync(S + lay.in_off + 8 * (int64_t)n, off_b, 8 * (size_t)n, hipMemcpyHostToDevice, s));
  HIP_TRYN(hipMemcpyAsync(S + lay.in_len, len_a, 4 * (size_t)n, hipMemcpyHostToDevice, s));
  HIP_TRYN(hipMemcpyAsync(S + lay.in_len + 4 * (int64_t)n, len_b, 4 * (size_t)n, hipMemcpyHostToDevice, s));
  HIP_TRYN(hipMemcpyAsync(S + lay.in_dw, deltaw, 4 * (size_t)n, hipMemcpyHostToDevice, s));
  // counters and status start at zero (fill kernels, not memsets)
  HIP_TRYN(launch_reset((int *)(S + lay.ctr), 8, s));
  HIP_TRYN(launch_reset((int *)(S + lay.status), kStatusFlags + n, s));
  hipLaunchKernelGGL(noalloc_plan_kernel, dim3(1), dim3(1024), 0, s, n, (const int64_t *)(S + lay.in_off),
                     (const int32_t *)(S + lay.in_len), (const int32_t *)(S + lay.in_dw), (PairMeta *)(S + lay.meta),
                     (int32_t *)(S + lay.order));
  HIP_TRYN(hipGetLastError());
  const PairMeta *meta = (const PairMeta *)(S + lay.meta);
  const DevTcm *tcm = (const DevTcm *)(S + lay.tcm);
  const int cus = ctx_cus(ctx);
  HIP_TRYN(launch_prep(false, std::max(1, std::min((n + 3) / 4, cus * 8)), meta, n, d_seq, tcm, (int *)(S + lay.aux),
                       s));
  HIP_TRYN(launch_lin_sys(false, std::max(1, std::min(n, cus * lin_sys_blocks_per_cu(false))), meta,
                          (const int32_t *)(S + lay.order), n, (int *)(S + lay.ctr), d_seq, tcm, S,
                          (const int *)(S + lay.aux), d_cost, (int32_t *)(S + lay.status), -1, s));
  int32_t bad = 0;
  HIP_TRYN(hipMemcpyAsync(&bad, S + lay.status, 4, hipMemcpyDeviceToHost, s));
  HIP_TRYN(hipStreamSynchronize(s));
  if (bad) {
    set_error("poyg_pairs_cost_linear_noalloc: %d pairs failed (a fill's wait for the previous strip gave up)", bad);
    return POYG_EDEVICE;
  }
  return POYG_OK;
}
