// Three-sequence alignment on gfx950: algn_fill_cube (src/algn.c:2885-3055)
// and backtrack_3d (:3829-3905).
//
// One wavefront fills one cube, row by row (a row = the n3 cells of one
// (plane, line)), lanes across the s3 axis and the self-gap chain along s3
// evaluated as a prefix minimum.  The reference walks three row pointers
// that advance n2-1 rows per plane while the output advances n2 rows, so
// plane p >= 1, line r >= 1 reads its "diagonal", "upper" and "previous"
// rows at flat row indices out-n2-p, out-n2-p+1 and out-p-1 (and the first
// line of plane p reads out-n2-p+1).  Those rows are kept in a per-wave ring
// of n1+n2 rows in global memory.  The direction cube (one code byte per
// cell) is walked afterwards by one thread per triple.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "poyg_internal.h"

namespace poyg {

struct TripleMeta {
  int64_t off1, off2, off3, out_off, dm_off;
  int32_t l1, l2, l3, cap;
};

constexpr int kP1 = 1, kP2 = 2, kP3 = 4, kS1 = 8, kS2 = 16, kS3 = 32, kSS = 64;

__device__ __forceinline__ int c_lane() { return (int)__lane_id(); }

__device__ __forceinline__ int64_t c_uniform64(int64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// LDS ordering across lanes of one wave (see wave_sync_mem in poyg_kernels.hip).
__device__ __forceinline__ void c_lds_sync() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0xC07F);
  __builtin_amdgcn_wave_barrier();
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// DPP-only cross-lane helpers (see poyg_kernels.hip).
// Eight wait states and a scheduling barrier between any EXEC change and the
// DPP moves that follow (a DPP issued shortly after an EXEC restore has been
// seen to treat enabled source lanes as disabled).
__device__ __forceinline__ void c_dpp_guard() {
  asm volatile("s_nop 7");
  __builtin_amdgcn_sched_barrier(0);
}

__device__ __forceinline__ int c_prefix_min(int v) {
  c_dpp_guard();
  constexpr int I = 0x7fffffff;
  v = min(v, __builtin_amdgcn_update_dpp(I, v, 0x111, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(I, v, 0x112, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(I, v, 0x114, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(I, v, 0x118, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(I, v, 0x142, 0xa, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(I, v, 0x143, 0xc, 0xf, false));
  return v;
}
__device__ __forceinline__ int c_prefix_sum(int v) {
  c_dpp_guard();
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);
  return v;
}

// Self-gap chain along s3 with the reference's strict-< rule: the value of
// a cell becomes mm[k-1] + ggs3[k] only when that is strictly smaller.
// v: value after the six plane/line moves; G3: prefix sums of ggs3.
struct ChainOut {
  int mm;
  bool ss;
};
__device__ __forceinline__ ChainOut c_chain(int v, bool act, int G3k, int g, int &carry_mm, int &carry_g3,
                                           bool k0) {
  const int big = 0x3fffffff;
  int pm = c_prefix_min(act ? v - G3k : big);
  pm = min(pm, carry_mm - carry_g3);
  const int mm = G3k + pm;
  c_dpp_guard();
  const int left = __builtin_amdgcn_update_dpp(carry_mm, mm, 0x138 /* wave_shr:1 */, 0xf, 0xf, false);
  ChainOut o;
  o.ss = act && !k0 && (left + g < v);
  o.mm = o.ss ? left + g : v;
  carry_mm = __builtin_amdgcn_readlane(mm, 63);
  carry_g3 = __builtin_amdgcn_readlane(G3k, 63);
  return o;
}

__global__ __launch_bounds__(64) void cube_fill_kernel(const TripleMeta *__restrict__ meta,
                                                       const int32_t *__restrict__ order, int n, int *counter,
                                                       const uint8_t *__restrict__ seq,
                                                       const DevTcm3 *__restrict__ tcm, uint8_t *dm_base,
                                                       int *scratch, int64_t scratch_per_wave, int32_t *cost_out) {
  __shared__ int s_r[4][32];   // ggs3, s1gs3, gs2s3, s1s2s3 rows indexed by the s3 symbol
  const int lane = c_lane();
  int *ring = scratch + (size_t)blockIdx.x * scratch_per_wave;
  const int gap = tcm->gap;
  if (lane < 32) s_r[0][lane] = tcm->cost[(gap * 32 + gap) * 32 + lane];
  c_lds_sync();
  for (;;) {
    int kk = 0;
    if (lane == 0) kk = atomicAdd(counter, 1);
    kk = __builtin_amdgcn_readfirstlane(kk);
    if (kk >= n) break;
    // the descriptor is wave-uniform: keep it in scalar registers so that the
    // loops over it compile to scalar branches
    const int t = __builtin_amdgcn_readfirstlane(order[kk]);
    TripleMeta m = meta[t];
    m.off1 = c_uniform64(m.off1);
    m.off2 = c_uniform64(m.off2);
    m.off3 = c_uniform64(m.off3);
    m.out_off = c_uniform64(m.out_off);
    m.dm_off = c_uniform64(m.dm_off);
    m.l1 = __builtin_amdgcn_readfirstlane(m.l1);
    m.l2 = __builtin_amdgcn_readfirstlane(m.l2);
    m.l3 = __builtin_amdgcn_readfirstlane(m.l3);
    m.cap = __builtin_amdgcn_readfirstlane(m.cap);
    const uint8_t *s1 = seq + m.off1, *s2 = seq + m.off2, *s3 = seq + m.off3;
    const int n1 = m.l1, n2 = m.l2, n3 = m.l3, W = n1 + n2;
    int *G3 = ring + (size_t)W * n3;
    uint8_t *dm = dm_base + m.dm_off;
    const int s30 = s3[0] & 31;
    // prefix sums of the (gap, gap, s3[k]) row
    {
      int carry = 0;
      for (int k0 = 0; k0 < n3; k0 += 64) {
        const int k = k0 + lane;
        const int g = (k >= 1 && k < n3) ? s_r[0][s3[k] & 31] : 0;
        const int ps = c_prefix_sum(g) + carry;
        if (k < n3) G3[k] = ps;
        carry = __builtin_amdgcn_readlane(ps, 63);
      }
    }
    __threadfence();
#define ROW(R) (ring + (size_t)((R) % W) * n3)
    // plane 0, line 0: 0, then the self-gap chain
    for (int k0 = 0; k0 < n3; k0 += 64) {
      const int k = k0 + lane;
      if (k < n3) {
        ROW(0)[k] = G3[k];
        dm[k] = k == 0 ? kS2 : kSS;
      }
    }
    __threadfence();
    int fenced = 0;   // rows [0, fenced] are visible to loads
    // plane 0, lines r >= 1
    for (int r = 1; r < n2; ++r) {
      const int x = s2[r] & 31;
      c_lds_sync();   // previous row's reads of the tables are done
      if (lane < 32) s_r[2][lane] = tcm->cost[(gap * 32 + x) * 32 + lane];
      c_lds_sync();
      const int *pv = ROW(r - 1);
      int *cur = ROW(r);
      uint8_t *dc = dm + (size_t)r * n3;
      const int gs2g = s_r[2][s30];
      int carry_mm = 0x3fffffff, carry_g3 = 0;
      for (int k0 = 0; k0 < n3; k0 += 64) {
        const int k = k0 + lane;
        const bool act = k < n3;
        int v = 0, code = kP1, G3k = 0, g = 0;
        if (act) {
          const int y = s3[k];
          G3k = G3[k];
          g = s_r[0][y];
          v = pv[k] + gs2g;
          if (k > 0) {
            const int tt = pv[k - 1] + s_r[2][y];
            if (tt < v) { v = tt; code = kS1; }
          }
        }
        const ChainOut o = c_chain(v, act, G3k, g, carry_mm, carry_g3, k == 0);
        if (act) {
          cur[k] = o.mm;
          dc[k] = (uint8_t)(o.ss ? kSS : code);
        }
      }
      __threadfence();
      fenced = r;
    }
    // planes p >= 1
    for (int p = 1; p < n1; ++p) {
      const int a = s1[p] & 31;
      c_lds_sync();
      if (lane < 32) s_r[1][lane] = tcm->cost[(a * 32 + gap) * 32 + lane];
      c_lds_sync();
      const int s1gg = s_r[1][s30];
      // line 0: reads the row the "diagonal" pointer holds, out-(n2+p-1)
      {
        const int out = p * n2;
        if (out - (n2 + p - 1) > fenced) { __threadfence(); fenced = out - 1; }
        const int *dg = ROW(out - (n2 + p - 1));
        int *cur = ROW(out);
        uint8_t *dc = dm + (size_t)out * n3;
        int carry_mm = 0x3fffffff, carry_g3 = 0;
        for (int k0 = 0; k0 < n3; k0 += 64) {
          const int k = k0 + lane;
          const bool act = k < n3;
          int v = 0, code = kP3, G3k = 0, g = 0;
          if (act) {
            const int y = s3[k] & 31;
            G3k = G3[k];
            g = s_r[0][y];
            v = dg[k] + s1gg;
            if (k > 0) {
              const int tt = dg[k - 1] + s_r[1][y];
              if (tt < v) { v = tt; code = kS3; }
            }
          }
          const ChainOut o = c_chain(v, act, G3k, g, carry_mm, carry_g3, k == 0);
          if (act) {
            cur[k] = o.mm;
            dc[k] = (uint8_t)(o.ss ? kSS : code);
          }
        }
      }
      for (int r = 1; r < n2; ++r) {
        const int out = p * n2 + r;
        const int b = s2[r] & 31;
        c_lds_sync();
        if (lane < 32) {
          s_r[2][lane] = tcm->cost[(gap * 32 + b) * 32 + lane];
          s_r[3][lane] = tcm->cost[(a * 32 + b) * 32 + lane];
        }
        c_lds_sync();
        if (out - p - 1 > fenced) { __threadfence(); fenced = out - 1; }
        const int *dg = ROW(out - n2 - p), *up = ROW(out - n2 - p + 1), *pv = ROW(out - p - 1);
        int *cur = ROW(out);
        uint8_t *dc = dm + (size_t)out * n3;
        const int gs2g = s_r[2][s30], s1s2g = s_r[3][s30];
        int carry_mm = 0x3fffffff, carry_g3 = 0;
        for (int k0 = 0; k0 < n3; k0 += 64) {
          const int k = k0 + lane;
          const bool act = k < n3;
          int v = 0, code = kP3, G3k = 0, g = 0;
          if (act) {
            const int y = s3[k] & 31;
            G3k = G3[k];
            g = s_r[0][y];
            // fill_parallel (:2840-2858) then fill_moved (:2812-2833), strict <
            v = up[k] + s1gg;
            int tt = pv[k] + gs2g;
            if (tt < v) { v = tt; code = kP1; }
            tt = dg[k] + s1s2g;
            if (tt < v) { v = tt; code = kP2; }
            if (k > 0) {
              tt = up[k - 1] + s_r[1][y];
              if (tt < v) { v = tt; code = kS3; }
              tt = pv[k - 1] + s_r[2][y];
              if (tt < v) { v = tt; code = kS1; }
              tt = dg[k - 1] + s_r[3][y];
              if (tt < v) { v = tt; code = kS2; }
            }
          }
          const ChainOut o = c_chain(v, act, G3k, g, carry_mm, carry_g3, k == 0);
          if (act) {
            cur[k] = o.mm;
            dc[k] = (uint8_t)(o.ss ? kSS : code);
          }
        }
      }
    }
    // the reference returns the last cell of the last row written; every
    // lane stores the same value (no lane-divergent code at the end of the
    // item loop body, which keeps the loop scalar)
    __threadfence();
    cost_out[t] = ROW(n1 * n2 - 1)[n3 - 1];
    __threadfence();
#undef ROW
  }
}

__device__ __forceinline__ void rev_put(uint8_t *base, int &pos, uint32_t &acc, int v) {
  --pos;
  acc |= (uint32_t)(v & 0xff) << (8 * (pos & 3));
  if ((pos & 3) == 0) {
    *(uint32_t *)(base + pos) = acc;
    acc = 0;
  }
}
__device__ __forceinline__ void rev_flush(uint8_t *base, int pos, uint32_t acc) {
  if (pos & 3) *(uint32_t *)(base + (pos & ~3)) = acc;
}

// backtrack_3d + algn_get_median_3d; slots: aligned s1, s2, s3, median.
__global__ void cube_trace_kernel(const TripleMeta *__restrict__ meta, const int32_t *__restrict__ order, int n,
                                  const uint8_t *__restrict__ seq, const DevTcm3 *__restrict__ tcm,
                                  const uint8_t *__restrict__ dm_base, uint8_t *out, int32_t *out_len,
                                  int32_t *status) {
  const int tt = blockIdx.x * blockDim.x + threadIdx.x;
  if (tt >= n) return;
  const int t = order[tt];
  const TripleMeta m = meta[t];
  const uint8_t *s1 = seq + m.off1, *s2 = seq + m.off2, *s3 = seq + m.off3;
  const int l1 = m.l1, l2 = m.l2, l3 = m.l3, gap = tcm->gap;
  const uint8_t *dm = dm_base + m.dm_off;
  uint8_t *o = out + m.out_off;
  int pos[3] = {m.cap, m.cap, m.cap};
  uint32_t acc[3] = {0, 0, 0};
  const int64_t a_plane = (int64_t)l2 * l3, a_line = l3;
  int64_t e = (int64_t)l1 * l2 * l3 - 1;
  int i1 = l1 - 1, i2 = l2 - 1, i3 = l3 - 1, len = 0, bad = 0;
  int last[3] = {gap, gap, gap};
  while (e > 0) {
    const int d = dm[e];
    int x1 = gap, x2 = gap, x3 = gap;
    const bool m1 = d & (kS2 | kS3 | kP3 | kP2), m2 = d & (kS2 | kS1 | kP1 | kP2), m3 = d & (kS2 | kS3 | kS1 | kSS);
    if ((m1 && i1 < 0) || (m2 && i2 < 0) || (m3 && i3 < 0)) { bad = 1; break; }
    if (d & kS2) { x1 = s1[i1--]; x2 = s2[i2--]; x3 = s3[i3--]; e -= a_plane + a_line + 1; }
    else if (d & kS3) { x1 = s1[i1--]; x3 = s3[i3--]; e -= a_plane + 1; }
    else if (d & kS1) { x2 = s2[i2--]; x3 = s3[i3--]; e -= a_line + 1; }
    else if (d & kP3) { x1 = s1[i1--]; e -= a_plane; }
    else if (d & kSS) { x3 = s3[i3--]; e -= 1; }
    else if (d & kP1) { x2 = s2[i2--]; e -= a_line; }
    else if (d & kP2) { x1 = s1[i1--]; x2 = s2[i2--]; e -= a_plane + a_line; }
    else { bad = 1; break; }
    if (len == 0) { last[0] = x1; last[1] = x2; last[2] = x3; }
    rev_put(o, pos[0], acc[0], x1);
    rev_put(o + m.cap, pos[1], acc[1], x2);
    rev_put(o + 2 * m.cap, pos[2], acc[2], x3);
    ++len;
    if (i1 < -1 || i2 < -1 || i3 < -1 || len > m.cap) { bad = 1; break; }
  }
  for (int q = 0; q < 3; ++q) rev_flush(o + q * m.cap, pos[q], acc[q]);
  // median_3 over the aligned triple: a constant sequence (the reference
  // never moves its end pointers), one symbol per column
  const int med = len > 0 ? tcm->median[((last[0] & 31) * 32 + (last[1] & 31)) * 32 + (last[2] & 31)] : 0;
  int mp = m.cap;
  uint32_t ma = 0;
  for (int q = 0; q < len; ++q) rev_put(o + 3 * m.cap, mp, ma, med);
  rev_flush(o + 3 * m.cap, mp, ma);
  for (int q = 0; q < 3; ++q) out_len[4 * t + q] = len;
  out_len[4 * t + 3] = len;
  if (bad) atomicAdd(status, 1);
}

hipError_t launch_reset(int *counters, int n, hipStream_t s);

}  // namespace poyg

using namespace poyg;

#define HIP_TRY3(expr)                                                               \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess) {                                                          \
      set_error("%s failed: %s", #expr, hipGetErrorString(e_));                      \
      return POYG_EDEVICE;                                                           \
    }                                                                                \
  } while (0)

// Triple plans are kept apart from pair plans; the C ABI shares poyg_plan
// through a tagged wrapper defined here.
struct TriplePlan {
  poyg_ctx *ctx;
  int n;
  std::vector<int32_t> order;
  std::vector<TripleMeta> meta;
  std::vector<std::pair<int, int>> batches;   // [begin, end) in `order`
  std::vector<int64_t> out_off;
  int64_t out_bytes = 0, cells = 0, dm_bytes = 0;
  int64_t scratch_per_wave = 0;
  int grid = 1;
  void *d_tcm = nullptr, *d_meta = nullptr, *d_order = nullptr, *d_dm = nullptr, *d_scratch = nullptr,
       *d_ctr = nullptr, *d_status = nullptr;
  hipStream_t stream = nullptr;
  ~TriplePlan() {
    for (void *p : {d_tcm, d_meta, d_order, d_dm, d_scratch, d_ctr, d_status})
      if (p) (void)hipFree(p);
  }
};

extern "C" {

int poyg_triple_plan_create(poyg_ctx *ctx, const poyg_cm3 *cm3, int32_t n, const int64_t *off_1,
                            const int32_t *len_1, const int64_t *off_2, const int32_t *len_2,
                            const int64_t *off_3, const int32_t *len_3, int device, int cus, void *stream,
                            int64_t dm_budget, void **out) {
  if (!cm3 || n < 0 || !out) return POYG_EINVAL;
  if (cm3->lcm != 5 || cm3->gap != kGap) {
    set_error("poyg_plan_triples: 5-bit DNA matrices only");
    return POYG_EUNSUPPORTED;
  }
  HIP_TRY3(hipSetDevice(device));
  TriplePlan *P = new TriplePlan;
  P->ctx = ctx;
  P->n = n;
  P->stream = (hipStream_t)stream;
  P->meta.resize(n);
  P->out_off.resize(n);
  std::vector<int64_t> cells(n);
  int maxw = 0, maxn3 = 0;
  for (int k = 0; k < n; ++k) {
    TripleMeta &m = P->meta[k];
    if (len_1[k] < 1 || len_2[k] < 1 || len_3[k] < 1) {
      delete P;
      set_error("poyg_plan_triples: empty sequence in triple %d", k);
      return POYG_EINVAL;
    }
    m.off1 = off_1[k]; m.off2 = off_2[k]; m.off3 = off_3[k];
    m.l1 = len_1[k]; m.l2 = len_2[k]; m.l3 = len_3[k];
    m.cap = (int32_t)(((int64_t)m.l1 + m.l2 + m.l3 + 1 + 15) & ~15);
    m.out_off = P->out_bytes;
    P->out_off[k] = P->out_bytes;
    P->out_bytes += 4 * (int64_t)m.cap;
    cells[k] = (int64_t)m.l1 * m.l2 * m.l3;
    P->cells += cells[k];
    maxw = std::max(maxw, m.l1 + m.l2);
    maxn3 = std::max(maxn3, m.l3);
  }
  P->order.resize(n);
  std::iota(P->order.begin(), P->order.end(), 0);
  std::stable_sort(P->order.begin(), P->order.end(), [&](int a, int b) { return cells[a] > cells[b]; });
  // split into batches whose direction cubes fit the budget
  int64_t cur = 0;
  int begin = 0;
  for (int q = 0; q < n; ++q) {
    const int k = P->order[q];
    if (cur > 0 && cur + cells[k] > dm_budget) {
      P->batches.push_back({begin, q});
      begin = q;
      cur = 0;
    }
    P->meta[k].dm_off = cur;
    cur += (cells[k] + 15) & ~(int64_t)15;
    P->dm_bytes = std::max(P->dm_bytes, cur);
  }
  if (n) P->batches.push_back({begin, n});
  P->scratch_per_wave = ((int64_t)maxw * maxn3 + maxn3 + 64 + 63) & ~(int64_t)63;   // 256-byte aligned
  P->grid = std::max(1, std::min(n, cus * 4));
  DevTcm3 *dt = new DevTcm3;
  to_dev_tcm3(cm3, dt);
  hipError_t e = hipSuccess;
  auto A = [&](void **p, size_t b) { if (e == hipSuccess) e = hipMalloc(p, b ? b : 16); };
  A(&P->d_tcm, sizeof(DevTcm3));
  A(&P->d_meta, sizeof(TripleMeta) * (size_t)n);
  A(&P->d_order, sizeof(int32_t) * (size_t)n);
  A(&P->d_dm, (size_t)P->dm_bytes);
  A(&P->d_scratch, sizeof(int) * (size_t)P->scratch_per_wave * P->grid);
  A(&P->d_ctr, 64);
  A(&P->d_status, 64);
  if (e == hipSuccess) e = hipMemcpy(P->d_tcm, dt, sizeof(DevTcm3), hipMemcpyHostToDevice);
  if (e == hipSuccess && n) e = hipMemcpy(P->d_meta, P->meta.data(), sizeof(TripleMeta) * n, hipMemcpyHostToDevice);
  if (e == hipSuccess && n) e = hipMemcpy(P->d_order, P->order.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice);
  delete dt;
  if (e != hipSuccess) {
    set_error("poyg_plan_triples: device setup failed: %s", hipGetErrorString(e));
    delete P;
    return POYG_ENOMEM;
  }
  *out = P;
  return POYG_OK;
}

void poyg_triple_plan_destroy(void *p) { delete (TriplePlan *)p; }
int64_t poyg_triple_plan_out_bytes(void *p) { return ((TriplePlan *)p)->out_bytes; }
int64_t poyg_triple_plan_cells(void *p) { return ((TriplePlan *)p)->cells; }
int poyg_triple_plan_out_offsets(void *p, int64_t *o) {
  TriplePlan *P = (TriplePlan *)p;
  std::copy(P->out_off.begin(), P->out_off.end(), o);
  return POYG_OK;
}

int poyg_triple_plan_run(void *pp, const uint8_t *d_seq, int32_t *d_cost, uint8_t *d_out, int32_t *d_out_len,
                         void *stream, float *fill_ms, float *trace_ms) {
  TriplePlan *P = (TriplePlan *)pp;
  if (!P || !d_seq || !d_cost || !d_out || !d_out_len) return POYG_EINVAL;
  hipStream_t s = stream ? (hipStream_t)stream : P->stream;
  hipEvent_t ev[3];
  const bool timing = fill_ms || trace_ms;
  if (timing)
    for (auto &x : ev) HIP_TRY3(hipEventCreate(&x));
  float f_tot = 0, t_tot = 0;
  for (auto &b : P->batches) {
    const int nb = b.second - b.first;
    const int32_t *ord = (const int32_t *)P->d_order + b.first;
    if (timing) HIP_TRY3(hipEventRecord(ev[0], s));
    HIP_TRY3(launch_reset((int *)P->d_ctr, 4, s));
    HIP_TRY3(launch_reset((int *)P->d_status, 4, s));
    const int grid = std::max(1, std::min(P->grid, nb));
    hipLaunchKernelGGL(cube_fill_kernel, dim3(grid), dim3(64), 0, s, (const TripleMeta *)P->d_meta, ord, nb,
                       (int *)P->d_ctr, d_seq, (const DevTcm3 *)P->d_tcm, (uint8_t *)P->d_dm, (int *)P->d_scratch,
                       P->scratch_per_wave, d_cost);
    HIP_TRY3(hipGetLastError());
    if (timing) HIP_TRY3(hipEventRecord(ev[1], s));
    hipLaunchKernelGGL(cube_trace_kernel, dim3((nb + 63) / 64), dim3(64), 0, s, (const TripleMeta *)P->d_meta, ord,
                       nb, d_seq, (const DevTcm3 *)P->d_tcm, (const uint8_t *)P->d_dm, d_out, d_out_len,
                       (int32_t *)P->d_status);
    HIP_TRY3(hipGetLastError());
    if (timing) {
      HIP_TRY3(hipEventRecord(ev[2], s));
      HIP_TRY3(hipEventSynchronize(ev[2]));
      float a = 0, c = 0;
      HIP_TRY3(hipEventElapsedTime(&a, ev[0], ev[1]));
      HIP_TRY3(hipEventElapsedTime(&c, ev[1], ev[2]));
      f_tot += a;
      t_tot += c;
    }
  }
  if (timing) {
    for (auto &x : ev) (void)hipEventDestroy(x);
    if (fill_ms) *fill_ms = f_tot;
    if (trace_ms) *trace_ms = t_tot;
  }
  return POYG_OK;
}

int poyg_triple_plan_status(void *pp, int32_t *bad) {
  TriplePlan *P = (TriplePlan *)pp;
  HIP_TRY3(hipMemcpy(bad, P->d_status, sizeof(int32_t), hipMemcpyDeviceToHost));
  return POYG_OK;
}

}  // extern "C"
