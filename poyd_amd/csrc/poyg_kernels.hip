// HIP kernels of the direct-optimization engine for gfx950 (MI355X).
//
// Pairwise band DP: one 64-lane wavefront aligns one pair at a time, taking
// pairs from a length-sorted work list through an atomic counter
// (persistent waves).  A row of the band is processed in chunks of 64
// columns, one column per lane.  Everything a cell needs from the previous
// row lives in two LDS row buffers (ring-indexed by column); the only
// dependency inside a row -- the horizontal gap chain -- is a min-plus
// recurrence x[j] = min(x[j-1] + a[j], b[j]) that the wave evaluates as a
// prefix minimum over b[j] - A[j], with A the per-pair prefix sum of a,
// precomputed once per pair.  All arithmetic is int32, exactly the
// comparisons of the reference, so costs and tie masks are bit-identical.
//
// Direction bands are stored one byte per cell, row by row, starting at the
// first filled column of the row; the backtraces run one thread per pair.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <utility>

#include <type_traits>

#include "poyg_band.h"
#include "poyg_internal.h"

namespace poyg {

constexpr int kHigh = kHighNum;
constexpr int kBig = 0x3fffffff;

// ---------------------------------------------------------------------------
// wave helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Separates DPP moves from any EXEC change before them: eight wait states,
// and the scheduler may move nothing across the barrier.  A DPP issued a few
// instructions after an EXEC restore has been seen to treat enabled source
// lanes as disabled (returning the fill value), depending on how waves
// interleave; every DPP sequence below starts with this guard.
__device__ __forceinline__ void dpp_exec_guard() {
  asm volatile("s_nop 7");
  __builtin_amdgcn_sched_barrier(0);
}

// Two unsigned 16-bit halves in one register (v_pk_* operations).
typedef unsigned short c16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ c16x2 c16v(uint32_t x) { return __builtin_bit_cast(c16x2, x); }
__device__ __forceinline__ uint32_t c16u(c16x2 x) { return __builtin_bit_cast(uint32_t, x); }
__device__ __forceinline__ uint32_t c16_add_sat(uint32_t a, uint32_t b) {
  return c16u(__builtin_elementwise_add_sat(c16v(a), c16v(b)));
}
__device__ __forceinline__ uint32_t c16_sub_sat(uint32_t a, uint32_t b) {
  return c16u(__builtin_elementwise_sub_sat(c16v(a), c16v(b)));
}
__device__ __forceinline__ uint32_t c16_min(uint32_t a, uint32_t b) {
  return c16u(__builtin_elementwise_min(c16v(a), c16v(b)));
}
__device__ __forceinline__ uint32_t c16_add(uint32_t a, uint32_t b) { return c16u(c16v(a) + c16v(b)); }
// per half: b where g == 0, 0xffff where g > 0 (g * 0xffff + b, saturated)
__device__ __forceinline__ uint32_t c16_gate(uint32_t g, uint32_t b) {
  uint32_t r;
  asm("v_pk_mad_u16 %0, %1, -1, %2 clamp" : "=v"(r) : "v"(g), "v"(b));
  return r;
}


// Cross-lane operations use DPP (VALU) and v_readlane only: lane i receives
// lane i-1's value (lane 0 receives `fill`).
__device__ __forceinline__ int shfl_up1(int v, int fill) {
  dpp_exec_guard();
  return __builtin_amdgcn_update_dpp(fill, v, 0x138 /* wave_shr:1 */, 0xf, 0xf, false);
}

// Inclusive prefix minimum over the 64 lanes: Hillis-Steele inside 16-lane
// rows (row_shr 1, 2, 4, 8), then row_bcast:15 and row_bcast:31 carry the
// row totals upward.  Disabled DPP lanes return the identity.
__device__ __forceinline__ int wave_prefix_min(int v) {
  constexpr int I = 0x7fffffff;
  dpp_exec_guard();
  v = min(v, __builtin_amdgcn_update_dpp(I, v, 0x111, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(I, v, 0x112, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(I, v, 0x114, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(I, v, 0x118, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(I, v, 0x142, 0xa, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(I, v, 0x143, 0xc, 0xf, false));
  return v;
}

__device__ __forceinline__ int wave_prefix_sum(int v) {
  dpp_exec_guard();
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);
  return v;
}

__device__ __forceinline__ int lane63(int v) { return __builtin_amdgcn_readlane(v, kWave - 1); }

// Index of this wave inside its workgroup as a wave-uniform (scalar) value.
// Per-wave LDS regions are addressed from it, so that their base lives in
// scalar registers: a per-lane copy of the base has been seen to go stale in
// some lanes across the divergent code of the fill loops.
__device__ __forceinline__ int wave_in_group() {
  return __builtin_amdgcn_readfirstlane((int)threadIdx.x / kWave);
}

// Orders this wave's LDS traffic across lanes: every LDS access issued
// before the call has completed (s_waitcnt lgkmcnt(0)) before any access
// after it, and the compiler may not move memory operations across it.
// Used between DP rows, where lane j reads what lane j-1 wrote.
__device__ __forceinline__ void wave_sync_mem() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0xC07F);   // gfx9 encoding of lgkmcnt(0), other counters unchanged
  __builtin_amdgcn_wave_barrier();
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Ordering of a wave's global row stores before its loads of them on the
// next row, for kernels with one wave per workgroup: workgroup scope needs
// no L2 write-back or invalidate (an agent-scope fence would do both).
__device__ __forceinline__ void wave_fence_global() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup"); }

// Wave-uniform copies of loaded values.  A work item's descriptor is the
// same in every lane; reading it through v_readfirstlane lets the compiler
// keep it in scalar registers and turn every branch on it into a scalar
// branch, instead of EXEC-masked code for control flow that never diverges.
// Together with the rule that no early exit (return, break) of a per-item
// body sits inside lane-divergent code, this keeps the persistent item loops
// scalar.  When that rule was broken (`if (lane == 0) store; return;`), the
// compiler built the item loop as a divergent loop that refreshed VGPR copies
// of the item's geometry under a partial EXEC mask, so some lanes computed
// with the previous item's geometry.
__device__ __forceinline__ int uniform32(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int64_t uniform64(int64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ PairMeta uniform_meta(const PairMeta *meta, int p) {
  PairMeta m = meta[p];
  m.off_r = uniform64(m.off_r);
  m.off_c = uniform64(m.off_c);
  m.out_off = uniform64(m.out_off);
  m.dm_off = uniform64(m.dm_off);
  m.aux_off = uniform64(m.aux_off);
  m.lr = uniform32(m.lr);
  m.lc = uniform32(m.lc);
  m.stride = uniform32(m.stride);
  m.cap = uniform32(m.cap);
  m.swap = uniform32(m.swap);
  m.p0 = uniform32(m.p0);
  m.p1 = uniform32(m.p1);
  m.p2 = uniform32(m.p2);
  m.p3 = uniform32(m.p3);
  return m;
}

// The one-pair-per-wave fills take the longest pairs, whose run time is
// a chain of dependent steps; when they share SIMDs with the four-pair
// kernel, their waves issue first.
__device__ __forceinline__ void wave_priority_long_pairs() { __builtin_amdgcn_s_setprio(2); }

__device__ __forceinline__ int next_item(int *counter) {
  int v = 0;
  if (lane_id() == 0) v = atomicAdd(counter, 1);
  return __builtin_amdgcn_readfirstlane(v);
}

__global__ void reset_counters_kernel(int *c, int n) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) c[k] = 0;
}

// The start of a pair plan's run in one launch: its work-list counters and
// status block cleared, and (alignment modes of the affine plans) the
// backtrace's output pointers set.
__global__ void plan_start_kernel(int *ctr, int nctr, int *status, int nstatus, TraceOut *tro, uint8_t *out,
                                  int32_t *out_len) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
  for (int k = t; k < nctr; k += stride) ctr[k] = 0;
  for (int k = t; k < nstatus; k += stride) status[k] = 0;
  if (tro && t == 0) *tro = TraceOut{out, out_len, (int32_t *)status};
}

// The host-sequence route's results, written into pinned host memory: word
// 0 the failed-pair count (status word 0), from word 16 the run's small
// results (costs, output lengths), then the output sequences; the host
// reads them once the stream is synchronized.
__global__ void collect_results_kernel(const uint32_t *small, int64_t small_words, const int32_t *status,
                                       const uint32_t *out, int64_t out_words, uint32_t *host) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
  if (t == 0) host[0] = (uint32_t)status[0];
  for (int64_t w = t; w < small_words; w += stride) host[16 + w] = small[w];
  for (int64_t w = t; w < out_words; w += stride) host[16 + small_words + w] = out[w];
  // (no fence: the kernel's completion, which the stream synchronization
  // waits for, makes its writes visible to the host; a system-scope fence
  // in every wave cost an L2 write-back per wave)
}

// Marks pair p of the run as failed: its flag (word kStatusFlags + p of the
// plan's status block) is set once, and the count in word 0 grows only for
// the first mark, so that a pair that fails in several waves or kernels
// counts once.  Safe to call from every lane of a wave.
__device__ __forceinline__ void mark_failed(int32_t *status, int p) {
  const int old = atomicOr(&status[kStatusFlags + p], 1);
  atomicAdd(&status[0], old == 0 ? 1 : 0);
}

// Right-to-left byte writer that stores whole dwords (slot bases and
// capacities are multiples of 16 bytes, so a final partial dword only
// touches unused capacity of the same slot).
struct RevWriter {
  uint8_t *base;
  int pos;
  uint32_t acc;
  __device__ void init(uint8_t *b, int cap) { base = b; pos = cap; acc = 0; }
  __device__ __forceinline__ void put(int v) {
    --pos;
    acc |= (uint32_t)(v & 0xff) << (8 * (pos & 3));
    if ((pos & 3) == 0) {
      *(uint32_t *)(base + pos) = acc;
      acc = 0;
    }
  }
  __device__ void flush() {
    if (pos & 3) *(uint32_t *)(base + (pos & ~3)) = acc;
  }
};

// ---------------------------------------------------------------------------
// Affine band (algn_fill_plane_3_aff / _nobt, src/algn.c:2287-2548)
// ---------------------------------------------------------------------------
// Direction byte of the affine band, the fields backtrace_affine reads:
//   bits 0-1: state entered from "todo" (0 H, 1 A, 2 V, 3 D) -- DO_* priority
//   bit 2 END_VERTICAL, bit 3 END_HORIZONTAL, bit 4 END_BLOCK
//   bits 5-6: state after an aligned column (0 stay, 1 H, 2 D, 3 V)
constexpr uint8_t kAffLeft = 2 | 4;   // DO_VERTICAL | END_VERTICAL
constexpr uint8_t kAffRight = 0 | 8;  // DO_HORIZONTAL | END_HORIZONTAL
constexpr int kAffRB = 256;           // LDS ring length (columns) of a row buffer
constexpr int kAffWaves = 4;          // waves per workgroup

struct RowRef {
  int *p;
  int mask, bstride, sstride;
  __device__ __forceinline__ int &at(int buf, int st, int j) const {
    return p[buf * bstride + st * sstride + (j & mask)];
  }
};
enum { ST_EH = 0, ST_EV = 1, ST_EB = 2, ST_CB = 3 };

// Per-pair prefix sums over the columns sequence B (length lb+1):
//   HX[j] = sum_{k<=j} sj_horizontal_extension[k]   (algn.c:2452-2457)
//   G[j]  = sum_{k<=j} prepend[B[k]]                (row 0, algn.c:2201-2211)
__device__ void aff_prefix(const uint8_t *B, int lb, const int *s_prep, int go, int *HX, int *G) {
  const int lane = lane_id();
  int carry_h = 0, carry_g = 0;
  for (int j0 = 0; j0 <= lb; j0 += kWave) {
    const int j = j0 + lane;
    int hx = 0, gr = 0;
    if (j >= 1 && j <= lb) {
      const int jc = B[j] & 31, jp = B[j - 1] & 31;
      gr = s_prep[jc];
      const int sj_go = (!(kGap & jp) && (kGap & jc)) ? 0 : go;
      hx = ((jp & kGap) && !(jc & kGap)) ? sj_go + gr : gr;
      if (j == 1) hx = gr;
    }
    const int ph = wave_prefix_sum(hx) + carry_h;
    const int pg = wave_prefix_sum(gr) + carry_g;
    if (j <= lb) {
      HX[j] = ph;
      G[j] = pg;
    }
    carry_h = lane63(ph);
    carry_g = lane63(pg);
  }
}

// One wave per pair computes the prefix sums the fill reads (a kernel of its
// own, so that the fill only reads them).
__global__ __launch_bounds__(kWave * 4) void aff_prep_kernel(const PairMeta *__restrict__ meta, int n,
                                                           const uint8_t *__restrict__ seq,
                                                           const DevTcm *__restrict__ tcm, int *aux) {
  __shared__ int s_prep[32];
  if (threadIdx.x < 32) s_prep[threadIdx.x] = tcm->prepend[threadIdx.x];
  __syncthreads();
  const int waves = gridDim.x * (blockDim.x / kWave);
  for (int p = blockIdx.x * (blockDim.x / kWave) + wave_in_group(); p < n; p += waves) {
    const PairMeta m = uniform_meta(meta, p);
    if (m.p2) continue;   // anti-diagonal pairs need no prefix sums
    int *HX = aux + m.aux_off;
    aff_prefix(seq + m.off_c, m.lc - 1, s_prep, tcm->gap_open, HX, HX + m.lc);
  }
}

template <bool kBt, bool kGlobalRows>
__device__ __forceinline__ void aff_fill_pair(const PairMeta &m, const uint8_t *__restrict__ seq, const int *s_cost,
                                              const int *s_prep, int go, uint8_t *__restrict__ dm_base, int *aux,
                                              int32_t *__restrict__ cost_out, int pidx, RowRef R) {
  const int lane = lane_id();
  // the pair's geometry is wave-uniform (scalar registers, scalar branches)
  const uint8_t *A = seq + uniform64(m.off_r);
  const uint8_t *B = seq + uniform64(m.off_c);
  const int la = uniform32(m.lr) - 1, lb = uniform32(m.lc) - 1, E0 = uniform32(m.p0), S = uniform32(m.stride);
  pidx = uniform32(pidx);
  go = uniform32(go);
  int *HX = aux + uniform64(m.aux_off);
  int *G = HX + (lb + 1);
  uint8_t *dm = dm_base + uniform64(m.dm_off);

  // row 0 (initialize_matrices_affine, src/algn.c:2183-2248)
  for (int j0 = 0; j0 <= E0; j0 += kWave) {
    const int j = j0 + lane;
    if (j <= E0) {
      const int eh = j == 0 ? go : go + G[j];
      R.at(0, ST_EH, j) = eh;
      R.at(0, ST_EV, j) = j == 0 ? go : kHigh;
      R.at(0, ST_EB, j) = j == 0 ? 0 : kHigh;
      R.at(0, ST_CB, j) = j == 0 ? 0 : eh;
    }
  }
  if (la == 0) {
    // every lane stores the same value: an early exit must not sit inside
    // lane-divergent code (see uniform32)
    cost_out[pidx] = lb == 0 ? 0 : go + G[lb];
    return;
  }
  if (kGlobalRows) wave_fence_global();
  else wave_sync_mem();

  int ic = A[0] & 31;
  for (int i = 1; i <= la; ++i) {
    const int pb = (i - 1) & 1, cbuf = i & 1;
    const int sp = 1 + (i > 40 ? i - 40 : 0);
    const int ep = min(lb, E0 + i - 1);
    const int ip = ic;
    ic = A[i] & 31;
    const int si_ge = s_cost[ic * 32 + kGap];
    const int si_go = (!(kGap & ip) && (kGap & ic)) ? 0 : go;
    const int si_ng = ic & 15;
    const int sive = (i > 1 && (ip & kGap) && !(ic & kGap)) ? si_go + si_ge : si_ge;
    const int ic_gap = ic & kGap;
    const int r = R.at(pb, ST_EV, sp - 1) + sive;
    uint8_t *drow = dm + (size_t)(i - 1) * S - (sp - 1);   // indexed by column j
    if (lane == 0) {
      R.at(cbuf, ST_EH, sp - 1) = kHigh;
      R.at(cbuf, ST_EV, sp - 1) = r;
      R.at(cbuf, ST_EB, sp - 1) = kHigh;
      R.at(cbuf, ST_CB, sp - 1) = kHigh;
      if (kBt) drow[sp - 1] = kAffLeft;
    }
    int carry_eh = kHigh, carry_cb = kHigh;
    int carry_hx = HX[sp - 1];
    const int *crow = s_cost + si_ng * 32;
    for (int j0 = sp; j0 <= ep; j0 += kWave) {
      const int j = j0 + lane;
      const bool act = j <= ep;
      int jc = 0, jp = 0, HXj = 0, pev_j = 0, pcb_j = 0, pev_l = 0, pcb_l = 0, peb_l = 0, peh_l = 0;
      if (act) {
        jc = B[j] & 31;
        jp = B[j - 1] & 31;
        HXj = HX[j];
        pev_j = R.at(pb, ST_EV, j);
        pcb_j = R.at(pb, ST_CB, j);
        pev_l = R.at(pb, ST_EV, j - 1);
        pcb_l = R.at(pb, ST_CB, j - 1);
        peb_l = R.at(pb, ST_EB, j - 1);
        peh_l = R.at(pb, ST_EH, j - 1);
      }
      const int gap_row = s_prep[jc];
      const int sj_go = (!(kGap & jp) && (kGap & jc)) ? 0 : go;
      int hx = ((jp & kGap) && !(jc & kGap)) ? sj_go + gap_row : gap_row;
      if (j == 1) hx = gap_row;
      // FILL_EXTEND_VERTICAL (algn.c:1812-1830)
      int ext = pev_j + sive, open = pcb_j + si_go + si_ge;
      const bool endv = !(ext < open);
      const int ev = endv ? open : ext;
      // FILL_EXTEND_BLOCK_DIAGONAL (bt :1861-1882, nobt :1837-1857)
      const bool both_gap = ic_gap && (jc & kGap);
      const int diag = both_gap ? 0 : kHigh;
      ext = peb_l + diag;
      if (kBt) {
        open = pcb_l + diag;
      } else {
        const bool flag2 = !(kGap & ip) && !(kGap & jc);
        open = pcb_l + (both_gap ? (flag2 ? 0 : 2 * go) : kHigh);
      }
      const bool endb = !(ext < open);
      const int eb = endb ? open : ext;
      // FILL_CLOSE_BLOCK_DIAGONAL (algn.c:1929-1977)
      const int dg = crow[jc & 15];
      const int xgo = (sj_go < si_go) ? si_go : sj_go;
      const int fv = pev_l + dg + (ic_gap ? sj_go : 0);
      const int fh = peh_l + dg + ((jc & kGap) ? si_go : 0);
      const int fd = peb_l + dg + xgo;
      int cb = pcb_l + dg, anext = 0;   // 0 = stay aligned
      bool a2v = false, a2h = false, a2d = false;
      if (cb >= fv) { if (cb > fv) { cb = fv; a2v = true; } else a2v = true; }
      if (cb >= fh) { if (cb > fh) { cb = fh; a2v = false; a2h = true; } else a2h = true; }
      if (cb >= fd) { if (cb > fd) { cb = fd; a2v = false; a2h = false; a2d = true; } else a2d = true; }
      // a strict improvement replaces the whole mask, including ALIGN_TO_ALIGN
      anext = a2h ? 1 : (a2d ? 2 : (a2v ? 3 : 0));
      // FILL_EXTEND_HORIZONTAL (algn.c:1761-1787) as a prefix minimum
      const int cb_left = shfl_up1(cb, carry_cb);
      const int bval = cb_left + sj_go + gap_row;
      int pm = wave_prefix_min(act ? bval - HXj : kBig);
      pm = min(pm, carry_eh - carry_hx);
      const int eh = HXj + pm;
      const int eh_left = shfl_up1(eh, carry_eh);
      const bool endh = !(eh_left + hx < bval);
      // ASSIGN_MINIMUM (algn.c:2250-2280)
      int f = eh;
      bool dh = true, dv = false, dd = false, da = false;
      if (f >= ev) { if (f > ev) { f = ev; dh = false; } dv = true; }
      if (f >= eb) { if (f > eb) { f = eb; dh = dv = false; } dd = true; }
      if (f >= cb) { if (f > cb) { f = cb; dh = dv = dd = false; } da = true; }
      const int todo = dh ? 0 : (da ? 1 : (dv ? 2 : 3));
      if (act) {
        R.at(cbuf, ST_EH, j) = eh;
        R.at(cbuf, ST_EV, j) = ev;
        R.at(cbuf, ST_EB, j) = eb;
        R.at(cbuf, ST_CB, j) = cb;
        if (kBt) drow[j] = (uint8_t)(todo | (endv << 2) | (endh << 3) | (endb << 4) | (anext << 5));
        if (i == la && j == lb) cost_out[pidx] = f;
      }
      carry_eh = lane63(eh);
      carry_cb = lane63(cb);
      carry_hx = lane63(HXj);
    }
    if (ep < lb && lane == 0) {
      R.at(cbuf, ST_EH, ep + 1) = kHigh;
      R.at(cbuf, ST_EV, ep + 1) = kHigh;
      R.at(cbuf, ST_EB, ep + 1) = kHigh;
      R.at(cbuf, ST_CB, ep + 1) = kHigh;
      if (kBt) drow[ep + 1] = kAffRight;
    }
    if (kGlobalRows) wave_fence_global();   // row buffers in global memory
    else wave_sync_mem();
  }
}

// The ALIGN_TO mask logic above must equal the reference's: start with
// ALIGN_TO_ALIGN; each candidate that is <= the running value either
// replaces the whole mask (strictly smaller) or ORs its flag (equal).
// backtrace_affine then tests ALIGN_TO_HORIZONTAL, ALIGN_TO_DIAGONAL,
// ALIGN_TO_VERTICAL in that order.

__device__ void load_tables(const DevTcm *tcm, int *s_cost, int *s_prep, uint8_t *s_med) {
  for (int k = threadIdx.x; k < 32 * 32; k += blockDim.x) {
    s_cost[k] = tcm->cost[k];
    if (s_med) s_med[k] = tcm->median[k];
  }
  for (int k = threadIdx.x; k < 32; k += blockDim.x) s_prep[k] = tcm->prepend[k];
  __syncthreads();
}

template <bool kBt>
__global__ __launch_bounds__(kWave * kAffWaves) void aff_fill_lds_kernel(
    const PairMeta *__restrict__ meta, const int32_t *__restrict__ order, int n, int *counter,
    const uint8_t *__restrict__ seq, const DevTcm *__restrict__ tcm, uint8_t *dm, int *aux,
    int32_t *cost_out) {
  __shared__ int s_cost[32 * 32];
  __shared__ int s_prep[32];
  __shared__ int s_rows[kAffWaves][2][4][kAffRB];
  load_tables(tcm, s_cost, s_prep, nullptr);
  wave_priority_long_pairs();
  const int w = wave_in_group();
  const int go = tcm->gap_open;
  RowRef R{&s_rows[w][0][0][0], kAffRB - 1, 4 * kAffRB, kAffRB};
  for (;;) {
    const int k = next_item(counter);
    if (k >= n) break;
    const int p = uniform32(order[k]);
    const PairMeta m = uniform_meta(meta, p);
    aff_fill_pair<kBt, false>(m, seq, s_cost, s_prep, go, dm, aux, cost_out, p, R);
  }
}

// Same fill with the two row buffers in global memory, for bands wider than
// the LDS ring (very unequal lengths).  One wave per workgroup.
template <bool kBt>
__global__ __launch_bounds__(kWave) void aff_fill_glb_kernel(
    const PairMeta *__restrict__ meta, const int32_t *__restrict__ order, int n, int *counter,
    const uint8_t *__restrict__ seq, const DevTcm *__restrict__ tcm, uint8_t *dm, int *aux,
    int32_t *cost_out, int *rowscratch, int row_width) {
  __shared__ int s_cost[32 * 32];
  __shared__ int s_prep[32];
  load_tables(tcm, s_cost, s_prep, nullptr);
  wave_priority_long_pairs();
  const int go = tcm->gap_open;
  int *mine = rowscratch + (size_t)blockIdx.x * 8 * row_width;
  RowRef R{mine, -1, 4 * row_width, row_width};
  for (;;) {
    const int k = next_item(counter);
    if (k >= n) break;
    const int p = uniform32(order[k]);
    const PairMeta m = uniform_meta(meta, p);
    aff_fill_pair<kBt, true>(m, seq, s_cost, s_prep, go, dm, aux, cost_out, p, R);
  }
}

// ---------------------------------------------------------------------------
// Affine band as an anti-diagonal wavefront (bands of at most 64 diagonal
// pairs, i.e. E0 <= 87).
// ---------------------------------------------------------------------------
// Lane l owns the band diagonals d = j - i = 2l - 40 and 2l - 39.  Step t
// computes the cells with i + j = t: the even diagonal on even steps, the odd
// one on odd steps, so a lane's cell is (i, j) = (s + 20 - l, s - 20 + l + p)
// at step t = 2s + p.  Of a cell's neighbours, (i-1, j-1) was computed two
// steps earlier by the same lane; (i, j-1) and (i-1, j) one step earlier, by
// the same lane or by lane l-1 (even steps, left) / lane l+1 (odd steps, up),
// which one DPP move fetches.  The horizontal gap chain of a row therefore
// advances in time inside one lane instead of across lanes.  Row characters
// move one lane up every two steps and column characters one lane down; the
// character entering at lane 0 / lane 63 is read from a 64-entry window held
// across the wave.  Direction bytes are stored per step and lane: dword
// (t >> 2) * W + l of the pair's band, byte t & 3 (W = lanes in use).
__device__ __forceinline__ int dpp_from_below(int v, int fill) {   // lane l <- lane l-1
  return __builtin_amdgcn_update_dpp(fill, v, 0x138 /* wave_shr:1 */, 0xf, 0xf, false);
}
__device__ __forceinline__ int dpp_from_above(int v, int fill) {   // lane l <- lane l+1
  return __builtin_amdgcn_update_dpp(fill, v, 0x130 /* wave_shl:1 */, 0xf, 0xf, false);
}

__device__ __forceinline__ int min3_i32(int a, int b, int c) { return min(min(a, b), c); }

// Branch-free bounded read: the index is clamped, the value then masked, so
// no lane is switched off (see aff_diag_kernel).
__device__ __forceinline__ int load_or0(const uint8_t *s, int idx, int len) {
  const int v = s[min(max(idx, 0), len)] & 31;
  return (idx >= 0 && idx <= len) ? v : 0;
}

struct AffDiag {
  int eh1, ev1, eb1, cb1;   // this lane's cell of the previous step
  int eh2, ev2, eb2, cb2;   // ... and of the step before
};

// One cell of the band (FILL_* of src/algn.c:1740-1977 and ASSIGN_MINIMUM
// :2250-2280 with the band rules of :2287-2548).  Returns the direction byte.
// (eh_l, cb_l) are the states of the left neighbour (i, j-1), (ev_u, cb_u)
// those of the upper one (i-1, j); the diagonal one is S's step-before-last.
// kSteady: every cell of the step has 41 <= i <= la and j <= lb, so the band
// edges are the fixed lanes d == -40 (left) and d == E0 (right), given as
// flags; no row 0 and no final cell occur.
template <bool kBt, bool kSteady, bool kPure = false, bool kGapFree = false>
__device__ __forceinline__ uint32_t aff_diag_cell(AffDiag &S, int eh_l, int cb_l, int ev_u, int cb_u, int i, int j,
                                                  int la, int lb, int E0, int go, const int *s_cost, int ic, int ip,
                                                  int si_ge, int si_go, int sive, int jc, int gap_row, int sj_go,
                                                  int hx, int &fcost, int open_mask, int left_mask) {
  // Every decision below is a select on int values: the cell code must not
  // branch (no lane may be switched off between the DPP moves of a step).
  const int d = j - i;
  // kGapFree (with kPure): neither character of any cell carries the gap
  // bit, so every gap-opening term is go and every extension the plain
  // indel cost (sive = si_ge, hx = gap_row)
  if (kGapFree) {
    si_go = go;
    sj_go = go;
    sive = si_ge;
    hx = gap_row;
  }
  // FILL_EXTEND_VERTICAL
  const int ext_v = ev_u + sive, open_v = cb_u + si_go + si_ge;
  const int endv = ext_v >= open_v;
  int ev = min(ext_v, open_v);
  const int ic_gap = kGapFree ? 0 : ic & kGap, jc_gap = kGapFree ? 0 : jc & kGap;
  const int dg = s_cost[(ic & 15) * 32 + (jc & 15)];
  const int fv = S.ev2 + dg + (ic_gap ? sj_go : 0);
  const int fh = S.eh2 + dg + (jc_gap ? si_go : 0);
  // The reference takes each minimum step by step (strict and non-strict
  // comparisons); a state is flagged exactly when it equals the final
  // minimum, with the precedence of its tests.  All values are non-negative
  // and below 2^29, so the minimum of (value << 2 | rank) gives both the
  // minimum and the first state reaching it.
  int eb, endb, cb, anext;
  if (kPure) {
    // No cell of the group pairs two gap-bearing characters (one sequence
    // has none past its leading gap), so the block-diagonal state is at
    // least kHigh everywhere past (0, 0): it is never a minimum, its
    // extension test always holds, and kHigh stands in for its value.
    eb = kHigh;
    endb = 1;
    const int kc = min3_i32((fh << 2) | 0, (fv << 2) | 2, ((S.cb2 + dg) << 2) | 3);
    cb = kc >> 2;
    anext = (kc + 1) & 3;
  } else {
    // FILL_EXTEND_BLOCK_DIAGONAL (bt / nobt)
    const int both_gap = (ic & jc) & kGap;
    const int diag = both_gap ? 0 : kHigh;
    const int ext_b = S.eb2 + diag;
    int open_b;
    if (kBt) {
      open_b = S.cb2 + diag;
    } else {
      const int flag2 = !(kGap & ip) && !jc_gap;
      open_b = S.cb2 + (both_gap ? (flag2 ? 0 : 2 * go) : kHigh);
    }
    endb = ext_b >= open_b;
    eb = min(ext_b, open_b);
    // FILL_CLOSE_BLOCK_DIAGONAL: the aligned state and where it came from
    // (H, then D, then V)
    const int fd = S.eb2 + dg + max(si_go, sj_go);
    const int kc = min(min((fh << 2) | 0, (fd << 2) | 1), min((fv << 2) | 2, ((S.cb2 + dg) << 2) | 3));
    cb = kc >> 2;
    anext = (kc + 1) & 3;   // rank 0 H -> 1, 1 D -> 2, 2 V -> 3, 3 none -> 0
  }
  // FILL_EXTEND_HORIZONTAL
  const int bval = cb_l + sj_go + gap_row, ext_h = eh_l + hx;
  const int endh = ext_h >= bval;
  int eh = min(ext_h, bval);
  // ASSIGN_MINIMUM: likewise, ties prefer H, then A, then V, then D
  const int kf = kPure ? min3_i32((eh << 2) | 0, (cb << 2) | 1, (ev << 2) | 2)
                       : min(min((eh << 2) | 0, (cb << 2) | 1), min((ev << 2) | 2, (eb << 2) | 3));
  const int f = kf >> 2;
  const int todo = kf & 3;
  uint32_t byte = (uint32_t)(todo | (endv << 2) | (endh << 3) | (endb << 4) | (anext << 5));
  // assemble the byte now: left to the scheduler, the comparison results of
  // all cells of a step stay alive as lane masks until the stores
  if (kBt) asm volatile("" : "+v"(byte));
  // band edges: the left boundary column carries only the vertical state,
  // the cell right of the band is closed (initialize_matrices_affine and the
  // per-row boundary writes of algn_fill_plane_3_aff)
  if (kSteady) {
    // open_mask: all ones for a cell inside the band; left_mask: all ones on
    // the left boundary column (bit selects, no lane masks)
    const int ev_closed = (left_mask & ext_v) | (~left_mask & kHigh);
    ev = (open_mask & ev) | (~open_mask & ev_closed);
    eh = (open_mask & eh) | (~open_mask & kHigh);
    eb = (open_mask & eb) | (~open_mask & kHigh);
    cb = (open_mask & cb) | (~open_mask & kHigh);
    const uint32_t byte_closed = ((uint32_t)left_mask & kAffLeft) | (~(uint32_t)left_mask & kAffRight);
    byte = ((uint32_t)open_mask & byte) | (~(uint32_t)open_mask & byte_closed);
  } else {
    (void)open_mask;
    (void)left_mask;
    const int valid = i >= 0 && i <= la && j >= 0 && j <= lb && d <= E0;
    const int lbnd = valid && i >= 1 && (j == 0 || d == -40);
    const int rbnd = valid && i >= 1 && d == E0;
    const int closed = lbnd || rbnd || !valid;
    const int row0 = valid && i == 0, j0 = j == 0;   // row 0
    // one code per cell (0 open, 1 left boundary, 2 closed, 3 row 0), made
    // opaque so that the selects below share it instead of each keeping its
    // own condition alive
    int code = row0 ? 3 : (closed ? (lbnd ? 1 : 2) : 0);
    asm volatile("" : "+v"(code));
    const int r0eh = j0 ? go : eh_l + gap_row;
    const int open = code == 0;
    ev = open ? ev : (code == 1 ? ext_v : (code == 3 && j0 ? go : kHigh));
    eh = open ? eh : (code == 3 ? r0eh : kHigh);
    eb = open ? eb : (code == 3 && j0 ? 0 : kHigh);
    cb = open ? cb : (code == 3 ? (j0 ? 0 : r0eh) : kHigh);
    byte = (code == 1 || code == 2) ? (code == 1 ? kAffLeft : kAffRight) : byte;
    fcost = (valid && i == la && j == lb) ? f : fcost;
  }
  S.eh2 = S.eh1; S.ev2 = S.ev1; S.eb2 = S.eb1; S.cb2 = S.cb1;
  S.eh1 = eh; S.ev1 = ev; S.eb1 = eb; S.cb1 = cb;
  return byte;
}

// Every half-step starts with dpp_exec_guard (see there), and the loop avoids
// divergence where it can: edges are selects, out-of-band lanes store their
// direction dword to `sink`, the cost is taken with v_readlane after the
// loop, and character reads clamp their index.
//
// R lane sets: for bands of up to 64 R diagonal pairs every lane holds R
// cells per step, lane l of set r owning the diagonal pair 64 r + l.  Values
// cross from set to set at lanes 63 / 0 through v_readlane.
template <int R>
struct AffDiagLanes {
  AffDiag S[R];
  int ic[R], ip[R], jc[R], jp[R];
  int si_ge[R], si_go[R], sive[R];
  int gap_row[R], sj_go[R], hx[R];
  int fcost[R];
};


template <int K>
__device__ __forceinline__ void aff_trace_pair(const PairMeta &m, int p, const uint8_t *__restrict__ seq,
                                               const uint8_t *__restrict__ dm_base, const uint8_t *s_med,
                                               uint32_t *tile, uint8_t *out, int32_t *out_len, int32_t *status);

// The pair's backtrace by the calling wave.  K: tile groups, 16 for bands
// of at most 64 diagonal pairs (the four-pair kernel's and the one-set
// kernel's), 8 for wider ones (see aff_trace_pair).
template <int K>
__device__ __forceinline__ void aff_trace_own(const PairMeta &m, int p, const uint8_t *__restrict__ seq,
                                              const uint8_t *__restrict__ dm_base, const uint8_t *s_med,
                                              uint32_t *tile, const TraceOut &tro) {
  aff_trace_pair<K>(m, p, seq, dm_base, s_med, tile, tro.out, tro.out_len, tro.status);
}

template <bool kBt, int R>
__global__ __launch_bounds__(kWave * 4) void aff_diag_kernel(
    const PairMeta *__restrict__ meta, const int32_t *__restrict__ order, int n, int *counter,
    const uint8_t *__restrict__ seq, const DevTcm *__restrict__ tcm, uint8_t *dm_base, int32_t *cost_out,
    uint32_t *sink, const TraceOut *trop) {
  __shared__ int s_cost[32 * 32];
  __shared__ int s_prep[32];
  __shared__ uint8_t s_med[kBt ? 32 * 32 : 4];
  __shared__ uint32_t s_tile[4][kBt ? kWave * 16 : 1];   // the backtrace's tile, per wave
  load_tables(tcm, s_cost, s_prep, kBt ? s_med : nullptr);
  wave_priority_long_pairs();
  const int go = tcm->gap_open;
  const int lane = lane_id();
  for (;;) {
    const int k = next_item(counter);
    if (k >= n) break;
    const int p = uniform32(order[k]);
    const PairMeta m = uniform_meta(meta, p);
    const uint8_t *A = seq + m.off_r;
    const uint8_t *B = seq + m.off_c;
    const int la = m.lr - 1, lb = m.lc - 1, E0 = m.p0, W = m.p1;
    if (la == 0) {   // only row 0: go + prefix of prepend costs
      int sum = 0;
      for (int j0 = 1; j0 <= lb; j0 += kWave) {
        const int j = j0 + lane;
        const int v = s_prep[B[min(j, lb)] & 31];
        sum += j <= lb ? v : 0;
      }
      sum = lane63(wave_prefix_sum(sum));
      cost_out[p] = lb == 0 ? 0 : go + sum;   // every lane stores the same value
      if (kBt) aff_trace_own<R == 1 ? 16 : 8>(m, p, seq, dm_base, s_med, s_tile[wave_in_group()], *trop);
      continue;
    }
    uint8_t *dm = dm_base + m.dm_off;
    AffDiagLanes<R> L;
    int i0[R], j0[R];
    uint32_t *dmw[R];
    int dmw_step[R];
    bool v0[R], v1[R], r0[R], r1[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int lr = lane + 64 * r;   // this lane's diagonal pair in set r
      i0[r] = 20 - lr;
      j0[r] = lr - 20;
      L.ic[r] = load_or0(A, 19 - lr, la);
      L.ip[r] = 0;
      L.jc[r] = load_or0(B, lr - 20, lb);
      L.jp[r] = load_or0(B, lr - 21, lb);
      L.si_ge[r] = L.si_go[r] = L.sive[r] = 0;
      L.gap_row[r] = s_prep[L.jc[r]];
      L.sj_go[r] = (!(kGap & L.jp[r]) && (kGap & L.jc[r])) ? 0 : go;
      L.hx[r] = ((L.jp[r] & kGap) && !(L.jc[r] & kGap)) ? L.sj_go[r] + L.gap_row[r] : L.gap_row[r];
      if (j0[r] == 1) L.hx[r] = L.gap_row[r];
      L.S[r] = AffDiag{kHigh, kHigh, kHigh, kHigh, kHigh, kHigh, kHigh, kHigh};
      L.fcost[r] = 0;
      dmw[r] = lr < W ? (uint32_t *)dm + lr : sink + lane;
      dmw_step[r] = lr < W ? W : 0;
      const int d0 = 2 * lr - 40;   // even diagonal of this lane in set r; odd is d0 + 1
      v0[r] = d0 <= E0;
      v1[r] = d0 + 1 <= E0;
      r0[r] = d0 == E0;
      r1[r] = d0 + 1 == E0;
    }
    const bool l0 = lane == 0;
    // characters entering at set 0 lane 0 (rows) and set R-1 lane 63 (columns)
    int awin = load_or0(A, 20 + lane, la), anext = load_or0(A, 84 + lane, la);
    const int bbase = 44 + 64 * (R - 1);
    int bwin = load_or0(B, bbase + lane, lb), bnext = load_or0(B, bbase + 64 + lane, lb);
    const int groups = (la + lb + 4) >> 2;   // ceil((la + lb + 1) / 4)
    // steady steps: t in [E0 + 82, min(2 lb - E0, 2 la - 40)]
    const int t_lo = E0 + 82, t_hi = min(2 * lb - E0, 2 * la - 40);
    int g_lo = (t_lo + 3) >> 2, g_hi = t_hi >= 3 ? (t_hi - 3) / 4 + 1 : 0;
    g_lo = min(g_lo, groups);
    g_hi = max(g_lo, min(g_hi, groups));
    auto group = [&](int g, auto steady_tag) {
      constexpr bool kSt = decltype(steady_tag)::value;
      if (g && (g & 31) == 0) {
        const int base = 2 * g;   // first pair-step of the new window
        awin = anext;
        bwin = bnext;
        anext = load_or0(A, 20 + base + 64 + lane, la);
        bnext = load_or0(B, bbase + base + 64 + lane, lb);
      }
      uint32_t acc[R];
#pragma unroll
      for (int r = 0; r < R; ++r) acc[r] = 0;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int s = 2 * g + h;
        {   // even step: next row character; left neighbours come from lane - 1
          // all DPP moves of the step first, pinned right after the guard
          // (no EXEC change of the cell code may come between them)
          dpp_exec_guard();
          const int fresh = __builtin_amdgcn_readlane(awin, s & 63);
          int top_ic[R], top_eh[R], top_cb[R], eh_ls[R], cb_ls[R];
#pragma unroll
          for (int r = 0; r < R; ++r) {
            top_ic[r] = __builtin_amdgcn_readlane(L.ic[r], 63);
            top_eh[r] = __builtin_amdgcn_readlane(L.S[r].eh1, 63);
            top_cb[r] = __builtin_amdgcn_readlane(L.S[r].cb1, 63);
          }
#pragma unroll
          for (int r = 0; r < R; ++r) {
            L.ip[r] = L.ic[r];
            L.ic[r] = dpp_from_below(L.ic[r], r == 0 ? fresh : top_ic[r > 0 ? r - 1 : 0]);
            eh_ls[r] = dpp_from_below(L.S[r].eh1, r == 0 ? kHigh : top_eh[r > 0 ? r - 1 : 0]);
            cb_ls[r] = dpp_from_below(L.S[r].cb1, r == 0 ? kHigh : top_cb[r > 0 ? r - 1 : 0]);
            asm volatile("" : "+v"(L.ic[r]), "+v"(eh_ls[r]), "+v"(cb_ls[r]));
          }
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const int i = i0[r] + s;
            const int ic = L.ic[r], ip = L.ip[r];
            L.si_ge[r] = s_cost[ic * 32 + kGap];
            L.si_go[r] = (!(kGap & ip) && (kGap & ic)) ? 0 : go;
            L.sive[r] = (i > 1 && (ip & kGap) && !(ic & kGap)) ? L.si_go[r] + L.si_ge[r] : L.si_ge[r];
            const int eh_l = eh_ls[r], cb_l = cb_ls[r];
            acc[r] |= aff_diag_cell<kBt, kSt>(L.S[r], eh_l, cb_l, L.S[r].ev1, L.S[r].cb1, i, j0[r] + s, la, lb, E0,
                                              go, s_cost, ic, ip, L.si_ge[r], L.si_go[r], L.sive[r], L.jc[r],
                                              L.gap_row[r], L.sj_go[r], L.hx[r], L.fcost[r],
                                              (r == 0 && l0) || r0[r] || !v0[r] ? 0 : -1, (r == 0 && l0) ? -1 : 0)
                      << (16 * h);
          }
        }
        {   // odd step: next column character; upper neighbours come from lane + 1
          dpp_exec_guard();
          const int fresh = __builtin_amdgcn_readlane(bwin, s & 63);
          int bot_jc[R], bot_ev[R], bot_cb[R], ev_us[R], cb_us[R];
#pragma unroll
          for (int r = 0; r < R; ++r) {
            bot_jc[r] = __builtin_amdgcn_readlane(L.jc[r], 0);
            bot_ev[r] = __builtin_amdgcn_readlane(L.S[r].ev1, 0);
            bot_cb[r] = __builtin_amdgcn_readlane(L.S[r].cb1, 0);
          }
#pragma unroll
          for (int r = 0; r < R; ++r) {
            L.jp[r] = L.jc[r];
            L.jc[r] = dpp_from_above(L.jc[r], r == R - 1 ? fresh : bot_jc[r < R - 1 ? r + 1 : 0]);
            ev_us[r] = dpp_from_above(L.S[r].ev1, r == R - 1 ? kHigh : bot_ev[r < R - 1 ? r + 1 : 0]);
            cb_us[r] = dpp_from_above(L.S[r].cb1, r == R - 1 ? kHigh : bot_cb[r < R - 1 ? r + 1 : 0]);
            asm volatile("" : "+v"(L.jc[r]), "+v"(ev_us[r]), "+v"(cb_us[r]));
          }
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const int i = i0[r] + s, j = j0[r] + s + 1;
            const int jc = L.jc[r], jp = L.jp[r];
            L.gap_row[r] = s_prep[jc];
            L.sj_go[r] = (!(kGap & jp) && (kGap & jc)) ? 0 : go;
            L.hx[r] = ((jp & kGap) && !(jc & kGap)) ? L.sj_go[r] + L.gap_row[r] : L.gap_row[r];
            if (j == 1) L.hx[r] = L.gap_row[r];
            const int ev_u = ev_us[r], cb_u = cb_us[r];
            acc[r] |= aff_diag_cell<kBt, kSt>(L.S[r], L.S[r].eh1, L.S[r].cb1, ev_u, cb_u, i, j, la, lb, E0, go,
                                              s_cost, L.ic[r], L.ip[r], L.si_ge[r], L.si_go[r], L.sive[r], jc,
                                              L.gap_row[r], L.sj_go[r], L.hx[r], L.fcost[r],
                                              (r1[r] || !v1[r]) ? 0 : -1, 0)
                      << (16 * h + 8);
          }
        }
      }
      if (kBt) {
#pragma unroll
        for (int r = 0; r < R; ++r) dmw[r][(size_t)g * dmw_step[r]] = acc[r];
      }
    };
    int g = 0;
    for (; g < g_lo; ++g) group(g, std::false_type{});
    for (; g < g_hi; ++g) group(g, std::true_type{});
    for (; g < groups; ++g) group(g, std::false_type{});
    // cell (la, lb) sits on diagonal lb - la, i.e. diagonal pair (lb - la + 40) / 2
    const int q = (lb - la + 40) >> 1, qr = q >> 6;
    int fc = L.fcost[0];
#pragma unroll
    for (int r = 1; r < R; ++r) fc = qr == r ? L.fcost[r] : fc;
    cost_out[p] = __builtin_amdgcn_readlane(fc, q & 63);
    if (kBt) aff_trace_own<R == 1 ? 16 : 8>(m, p, seq, dm_base, s_med, s_tile[wave_in_group()], *trop);
  }
}

// backtrace_affine over the anti-diagonal band layout, one wave per pair.
//
// The walk (src/algn.c:2380-2470) is a state machine over the cells of the
// band: in state A it moves diagonally until a cell whose bits 5-6 name
// another state; V, H and D move one cell and may fall back to "todo",
// which re-reads the same cell's bits 0-1.  Almost all of an alignment is
// long runs of A, so the walk is split in two phases:
//
// * The walker (wave-uniform, scalar registers) runs over tiles of K
//   groups (4K anti-diagonals).  A tile is held by the 64 lanes as K dwords
//   of the 64 band lanes around the walk; it is copied to LDS (for the
//   single-cell reads of V, H, D and todo) and reduced per lane to masks of
//   the cells that end an A run, ordered along each of the lane's two
//   diagonals.  A whole A run inside a tile costs one v_readlane and one
//   find-last-bit.  The walker records one 2-bit operation per step
//   (A 0, D 1, V 2, H 3) in an op log held one 16-step word per lane.
// * Every 1024 steps (and at the end) the lanes expand their 16 steps of
//   the log into the four outputs: positions follow from prefix sums of the
//   per-lane i and j decrements, characters are read from the sequences,
//   the aligned sequences and the median with gaps are stored as one
//   16-byte store per lane and output, and the gap-free median is
//   compacted with a prefix sum of the per-lane emit counts.
//
// Window (K = 8, bands wider than 64 diagonal pairs): the next tile is
// prefetched when the walk enters the current one, at most 63
// anti-diagonals before the walk leaves the next tile; the walk then stays
// within lanes ((d + 1) >> 1) - 32 .. +31 of its band lane (d = j - i + 40),
// which is where the prefetched window starts.  Bands of at most 64
// diagonal pairs fit the window whole, and use K = 16.
constexpr int kTrStep = 16;    // op-log steps per lane
constexpr int kTrBatch = kTrStep * kWave;

__device__ __forceinline__ uint32_t tr_spread(uint32_t x) {
  x = (x | (x << 4)) & 0x0F0F0F0Fu;
  x = (x | (x << 2)) & 0x33333333u;
  x = (x | (x << 1)) & 0x55555555u;
  return x;
}

// One pair's backtrace by the calling wave (every lane).  s_med: the median
// table in LDS; tile: this wave's LDS tile (64 K dwords).  K: groups per
// tile, 8 in general, 16 when the band has at most 64 diagonal pairs (the
// window then always holds the whole band, and twice the anti-diagonals per
// tile halve the tile changes and the A runs they cut).
template <int K>
__device__ __forceinline__ void aff_trace_pair(const PairMeta &m, int p, const uint8_t *__restrict__ seq,
                                               const uint8_t *__restrict__ dm_base, const uint8_t *s_med,
                                               uint32_t *tile, uint8_t *out, int32_t *out_len, int32_t *status) {
  static_assert(K == 8 || K == 16, "tiles of 8 or 16 groups");
  const int lane = lane_id();
  const uint8_t *tile8 = (const uint8_t *)tile;
  const uint8_t *A = seq + m.off_r;
  const uint8_t *B = seq + m.off_c;
  const int la = m.lr - 1, lb = m.lc - 1, E0 = m.p0, W = m.p1, cap = m.cap;
  const int G = (la + lb + 4) >> 2;   // groups of the band
  const uint32_t *dmw = (const uint32_t *)(dm_base + m.dm_off);
  uint8_t *o = out + m.out_off;
  uint8_t *om = o;
  uint8_t *ow = o + cap;
  uint8_t *oa = o + (m.swap ? 3 : 2) * cap;   // the rows sequence's slot
  uint8_t *ob = o + (m.swap ? 2 : 3) * cap;

  // ---- op log and the emit phase
  uint32_t oplog = 0;   // lane l: steps 16l .. 16l+15 of the batch
  uint32_t acc = 0;     // the word being filled (scalar)
  int s = 0;            // steps in the batch
  int S0 = 0, ib = la, jb = lb, wmpos = cap, nm = 0, last = -1, final_step = -1;

  auto expand = [&](int nst) {
    const uint32_t c = oplog;
    const int v = min(max(nst - kTrStep * lane, 0), kTrStep);
    const uint32_t vm = v >= kTrStep ? 0xffffffffu : ((1u << (2 * v)) - 1u);
    const uint32_t lo = c & 0x55555555u & vm, hi = (c >> 1) & 0x55555555u & vm;
    const int nI = v - __popc(lo & hi), nJ = v - __popc(hi & ~lo);
    const int cI = wave_prefix_sum(nI);
    const int cJ = wave_prefix_sum(nJ);
    uint32_t wa[4] = {0u, 0u, 0u, 0u}, wb[4] = {0u, 0u, 0u, 0u}, ww[4] = {0u, 0u, 0u, 0u};
    uint32_t em = 0u;
    int lastv = 0;
    const int gbase = S0 + kTrStep * lane;
    int ii = ib - (cI - nI), jj = jb - (cJ - nJ);
    // two halves of 8 steps keep the loaded characters of one half live
#pragma unroll
    for (int h = 0; h < kTrStep; h += 8) {
      int av[8], bv[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const uint32_t code = (c >> (2 * (h + q))) & 3u;
        av[q] = A[min(max(ii, 0), la)];
        bv[q] = B[min(max(jj, 0), lb)];
        ii -= code != 3u;
        jj -= code != 2u;
      }
#pragma unroll
      for (int q0 = 0; q0 < 8; ++q0) {
        const int q = h + q0;
        const uint32_t code = (c >> (2 * q)) & 3u;
        const int a = av[q0], b = bv[q0];
        const bool isA = code == 0u, isD = code == 1u, isV = code == 2u;
        const int md = s_med[(a & 15) * 32 + (b & 15)];
        const int gv = (a & kGap) ? kGap : (a | kGap);
        const int gh = (b & kGap) ? kGap : (b | kGap);
        int va = code == 3u ? kGap : a;
        int vb = isV ? kGap : b;
        int vw = isA ? md : (isD ? kGap : (isV ? gv : gh));
        bool e = isA || (isV && !(a & kGap)) || (code == 3u && !(b & kGap));
        if (gbase + q == final_step) {   // the closing all-gap column
          va = kGap;
          vb = kGap;
          vw = kGap;
        }
        if (q >= v) {
          va = vb = vw = 0;
          e = false;
        }
        const int byte = kTrStep - 1 - q;   // steps are written right to left
        wa[byte >> 2] |= (uint32_t)va << (8 * (byte & 3));
        wb[byte >> 2] |= (uint32_t)vb << (8 * (byte & 3));
        ww[byte >> 2] |= (uint32_t)vw << (8 * (byte & 3));
        em |= (uint32_t)e << q;
        lastv = e ? vw : lastv;
      }
    }
    const int P = cap - kTrStep - gbase;   // 16-byte aligned: cap and S0 are multiples of 16
    if (v > 0 && P >= 0) {
      *(uint4 *)(oa + P) = make_uint4(wa[0], wa[1], wa[2], wa[3]);
      *(uint4 *)(ob + P) = make_uint4(wb[0], wb[1], wb[2], wb[3]);
      *(uint4 *)(ow + P) = make_uint4(ww[0], ww[1], ww[2], ww[3]);
    }
    const int ne = __popc(em);
    const int cE = wave_prefix_sum(ne);
    const int wtop = wmpos - 1 - (cE - ne);
    int rank = 0;
#pragma unroll
    for (int q = 0; q < kTrStep; ++q) {
      const int byte = kTrStep - 1 - q;
      if ((em >> q) & 1u) om[wtop - rank] = (uint8_t)(ww[byte >> 2] >> (8 * (byte & 3)));
      rank += (em >> q) & 1u;
    }
    const int tot = lane63(cE);
    if (tot) {
      const uint64_t bal = __ballot(ne > 0);
      const int ll = 63 - __clzll((long long)bal);
      last = __builtin_amdgcn_readlane(lastv, ll);
    }
    wmpos -= tot;
    nm += tot;
    ib -= lane63(cI);
    jb -= lane63(cJ);
    S0 += nst;
    oplog = 0u;
  };
  // ---- the walker: each pass yields one run of an operation (0 A, 1 D,
  // 2 V, 3 H), then the run is appended to the op log (the only place the
  // log is expanded, so the emit phase is instantiated once)
  int i = la, j = lb, mode = 0, bad = 0;   // mode: 0 todo, 1 V, 2 H, 3 D, 4 A
  int stage = 0;                            // 0 walk, 1 tail V, 2 tail H, 3 closing column, 4 done
  uint32_t pre[K];
  uint32_t maskA = 0u, maskB = 0u;   // A-run ends (layout below)
  int g0 = 0, Lb = 0, g0n = 0, Lbn = 0;
  auto lbase = [&](int d) { return min(max(((d + 1) >> 1) - kWave / 2, 0), max(0, W - kWave)); };
  // Lanes past the band load (and then ignore) copies of its last lane;
  // groups outside the band are clamped to its ends: the walker never reads
  // them (anti-diagonals above the start or below row / column 0).
  auto fetch = [&](int gs, int lb0) {
    const int Lc = min(lb0 + lane, W - 1);
    if (gs >= 0 && gs + K <= G) {
      const uint32_t *src = dmw + (size_t)gs * W + Lc;
#pragma unroll
      for (int u = 0; u < K; ++u) pre[u] = src[(size_t)u * W];
    } else {
#pragma unroll
      for (int u = 0; u < K; ++u) pre[u] = dmw[(size_t)min(max(gs + u, 0), G - 1) * W + Lc];
    }
  };
  auto commit = [&]() {
    wave_sync_mem();   // the walker's reads of the previous tile are done
    uint4 *dst = (uint4 *)(tile + lane * K);
#pragma unroll
    for (int v = 0; v < K / 4; ++v) dst[v] = make_uint4(pre[4 * v], pre[4 * v + 1], pre[4 * v + 2], pre[4 * v + 3]);
    // bit 8b + u of T[h]: the cell of group 8h + u, byte b ends an A run
    // (bits 5-6 set)
    uint32_t T[K / 8];
#pragma unroll
    for (int h = 0; h < K / 8; ++h) {
      T[h] = 0u;
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const uint32_t x = pre[8 * h + u];
        const uint32_t y = (x | (x << 1)) & 0x40404040u;
        T[h] |= u <= 6 ? (y >> (6 - u)) : (y << (u - 6));
      }
    }
    if (K == 8) {
      // interleaved to bit 16 (b & 1) + 2u + (b >> 1): the cells of each of
      // the lane's two diagonals in anti-diagonal order
      const uint32_t z0 = (T[0] & 0xFFu) | ((T[0] & 0xFF00u) << 8);
      const uint32_t z1 = ((T[0] >> 16) & 0xFFu) | ((T[0] >> 8) & 0xFF0000u);
      maskA = tr_spread(z0) | (tr_spread(z1) << 1);
    } else {
      // one word per diagonal: bit 2u + (b >> 1) of mask (b & 1)
      auto row = [&](int b) { return ((T[0] >> (8 * b)) & 0xFFu) | (((T[K / 8 - 1] >> (8 * b)) & 0xFFu) << 8); };
      auto spread16 = [](uint32_t x) {
        x = (x | (x << 8)) & 0x00FF00FFu;
        return tr_spread(x);
      };
      maskA = spread16(row(0)) | (spread16(row(2)) << 1);
      maskB = spread16(row(1)) | (spread16(row(3)) << 1);
    }
    wave_sync_mem();
  };
  if (i > 0 && j > 0) {
    g0 = max(0, ((i + j) >> 2) - K + 1);
    Lb = lbase(j - i + 40);
    fetch(g0, Lb);
    commit();
    g0n = g0 - K;
    Lbn = Lb;
    if (g0n + K > 0) fetch(g0n, Lbn);
  }
  for (;;) {
    int code = 0, r = 0;
    if (stage == 0) {
      if (i > 0 && j > 0) {
        const int t = i + j, d = j - i + 40;
        if ((unsigned)d > (unsigned)(E0 + 40)) {   // left the band
          bad = 1;
          stage = 4;
        } else {
          if (t < 4 * g0) {
            commit();
            g0 = g0n;
            Lb = Lbn;
            g0n = g0 - K;
            Lbn = lbase(d);
            if (g0n + K > 0) fetch(g0n, Lbn);
          }
          const int kk = (d >> 1) - Lb;   // window lane, 0..63
          const int pos = t - 4 * g0;      // anti-diagonal in the tile, 0 .. 4K-1
          int f = 0;
          if (mode != 4) f = uniform32(tile8[kk * (4 * K) + pos]);
          if (mode == 0) mode = (0x3142 >> ((f & 3) * 4)) & 15;   // todo: 0 H, 1 A, 2 V, 3 D
          if (mode == 4) {
            const int par = pos & 1, key = pos >> 1;
            uint32_t mk;
            if (K == 8) {
              mk = (uint32_t)__builtin_amdgcn_readlane((int)maskA, kk) >> (16 * par);
            } else {
              mk = par ? (uint32_t)__builtin_amdgcn_readlane((int)maskB, kk)
                       : (uint32_t)__builtin_amdgcn_readlane((int)maskA, kk);
            }
            const uint32_t cand = mk & ((2u << key) - 1u);   // key <= 31: 2u << 31 wraps to 0
            int ks = -1;
            r = key + 1;
            if (cand) {
              ks = 31 - __clz((int)cand);
              r = key - ks + 1;
            }
            const int avail = min(i, j);
            if (r > avail) {
              r = avail;
              ks = -1;
            }
            code = 0;
            i -= r;
            j -= r;
            if (ks >= 0) {   // the run's last cell names the next state: 1 H, 2 D, 3 V
              const int f2 = uniform32(tile8[kk * (4 * K) + 2 * ks + par]);
              mode = (0x1324 >> (((f2 >> 5) & 3) * 4)) & 15;
            }
          } else if (mode == 1) {
            code = 2;
            r = 1;
            if (f & 4) mode = 0;
            --i;
          } else if (mode == 2) {
            code = 3;
            r = 1;
            if (f & 8) mode = 0;
            --j;
          } else {
            code = 1;
            r = 1;
            if (f & 16) mode = 0;
            --i;
            --j;
          }
        }
      } else {
        stage = 1;
      }
    } else if (stage == 1) {
      code = 2;
      r = i;
      i = 0;
      stage = 2;
    } else if (stage == 2) {
      code = 3;
      r = j;
      j = 0;
      stage = 3;
    } else if (stage == 3) {
      final_step = S0 + s;
      code = 1;
      r = 1;
      stage = 4;
    }
    const bool finish = stage == 4;
    if (!finish && r > 0 && s + r < kTrBatch) {   // the walk's runs: no batch ends inside
      if (code == 0) {
        if (((s + r) >> 4) != (s >> 4)) {   // A codes are zero: store the word it leaves
          oplog = lane == (s >> 4) ? acc : oplog;
          acc = 0u;
        }
        s += r;
        continue;
      }
      if (r == 1) {
        acc |= (uint32_t)code << (2 * (s & 15));
        ++s;
        if ((s & 15) == 0) {
          oplog = lane == ((s - 1) >> 4) ? acc : oplog;
          acc = 0u;
        }
        continue;
      }
    }
    const uint32_t pat = code == 0 ? 0u : (code == 1 ? 0x55555555u : (code == 2 ? 0xAAAAAAAAu : 0xFFFFFFFFu));
    while (r > 0 || (finish && s > 0)) {
      if (r > 0) {
        int take;
        if (code == 0) {   // A codes are zero: only completed words are stored
          take = min(r, kTrBatch - s);
          if (((s + take) >> 4) != (s >> 4)) {
            oplog = lane == (s >> 4) ? acc : oplog;
            acc = 0u;
          }
        } else {
          const int wq = s & 15;
          take = min(r, 16 - wq);
          const uint32_t msk = take >= 16 ? 0xffffffffu : ((1u << (2 * take)) - 1u);
          acc |= (pat & msk) << (2 * wq);
          if (((s + take) & 15) == 0) {
            oplog = lane == (s >> 4) ? acc : oplog;
            acc = 0u;
          }
        }
        s += take;
        r -= take;
      }
      if (s == kTrBatch || (finish && r == 0 && s > 0)) {
        if (s & 15) {
          oplog = lane == (s >> 4) ? acc : oplog;
          acc = 0u;
        }
        expand(s);
        s = 0;
      }
    }
    if (finish) break;
  }
  if (!bad && (nm == 0 || last != kGap)) {
    --wmpos;
    om[wmpos] = (uint8_t)kGap;   // every lane stores the same byte
  }
  out_len[4 * p + 0] = cap - wmpos;
  out_len[4 * p + 1] = S0;
  out_len[4 * p + 2] = S0;
  out_len[4 * p + 3] = S0;
  if (bad) mark_failed(status, p);
}

// ---------------------------------------------------------------------------
// Affine band, four pairs per wave (bands of at most 16 H diagonal pairs).
// ---------------------------------------------------------------------------
// The same anti-diagonal recurrence as aff_diag_kernel, with each lane
// holding H consecutive diagonal pairs ("virtual lanes" L = H q + k) and each
// 16-lane row of the wave working on its own pair.  A band of E0 <= 55 fills
// 41..48 of the 48 virtual lanes of H = 3 (against 41..64 of the 64 lanes in
// aff_diag_kernel), and the per-step overhead (DPP moves, incoming
// characters, stores) is shared by H cells per lane.  Between the virtual
// lanes of one lane values move by register; across lanes by DPP row_shr:1
// (from the lane below) and row_shl:1 (from the lane above), which stop at
// the 16-lane row boundaries and there return the band-edge value (kHigh) or
// the incoming character, exactly what lanes 0 / 63 receive in
// aff_diag_kernel.  Incoming characters come from a 16-entry window per row
// that rotates by one lane per step (row_ror:15), so the current entry is
// always at lane 0 (rows) / lane 15 (columns).  Row constants (gap costs of a
// row) move with their row character; only the row entering a lane is
// evaluated.
__device__ __forceinline__ int dpp_row_from_below(int v, int fill) {   // lane q <- q-1 in the row
  return __builtin_amdgcn_update_dpp(fill, v, 0x111 /* row_shr:1 */, 0xf, 0xf, false);
}
__device__ __forceinline__ int dpp_row_from_above(int v, int fill) {   // lane q <- q+1 in the row
  return __builtin_amdgcn_update_dpp(fill, v, 0x101 /* row_shl:1 */, 0xf, 0xf, false);
}
// lane q <- lane (q + 1) mod 16 of its row
__device__ __forceinline__ int dpp_row_rotate_down(int v) {
  return __builtin_amdgcn_update_dpp(0, v, 0x12f /* row_ror:15 */, 0xf, 0xf, false);
}

// Per pair: bit 0 / bit 1 set if the rows / columns sequence has no
// gap-bearing character past the leading gap.  One bit: no DP cell pairs two
// of them (the block-diagonal state never matters); both bits: no cell holds
// one (see aff_diag_cell).
__global__ __launch_bounds__(kWave * 4) void pair_pure_kernel(const PairMeta *__restrict__ meta,
                                                            const int32_t *__restrict__ order, int n,
                                                            const uint8_t *__restrict__ seq, uint8_t *pure) {
  const int lane = lane_id();
  const int waves = gridDim.x * (blockDim.x / kWave);
  for (int w = blockIdx.x * (blockDim.x / kWave) + wave_in_group(); w < n; w += waves) {
    const int p = uniform32(order[w]);
    const PairMeta m = uniform_meta(meta, p);
    int ga = 0, gb = 0;
    for (int x = 1 + lane; x < m.lr; x += kWave) ga |= seq[m.off_r + x];
    for (int x = 1 + lane; x < m.lc; x += kWave) gb |= seq[m.off_c + x];
    const bool any_a = __ballot(ga & kGap) != 0, any_b = __ballot(gb & kGap) != 0;
    pure[p] = (any_a ? 0 : 1) | (any_b ? 0 : 2);   // every lane stores the same value
  }
}

template <int H>
struct QuadLanes {
  AffDiag S[H];
  int ic[H], ip[H], si_ge[H], si_go[H], sive[H];   // row i = s + 20 - L and its gap costs
  int jc[H], jp[H], gap_row[H], sj_go[H], hx[H];   // column and its gap costs
  int fcost[H];
};

template <bool kBt, int H>
__global__ __launch_bounds__(kWave * 4, 3) void aff_quad_kernel(
    const PairMeta *__restrict__ meta, const int32_t *__restrict__ order, int n, int *counter,
    const uint8_t *__restrict__ seq, const DevTcm *__restrict__ tcm, uint8_t *dm_base, int32_t *cost_out,
    uint32_t *sink, const uint8_t *__restrict__ pure, const TraceOut *trop, const int *n_dev) {
  __shared__ int s_cost[32 * 32];
  __shared__ int s_prep[32];
  __shared__ uint8_t s_med[kBt ? 32 * 32 : 4];
  __shared__ uint32_t s_tile[4][kBt ? kWave * 16 : 1];   // the backtrace's tile, per wave
  // n_dev: the list's length is on the device (the packed kernel's redo list)
  if (n_dev) {
    n = uniform32(*n_dev);
    if (n == 0) return;   // uniform over the workgroup
  }
  load_tables(tcm, s_cost, s_prep, kBt ? s_med : nullptr);
  const int go = tcm->gap_open;
  const int lane = lane_id();
  const int q = lane & 15;
  for (;;) {
    int base = 0;
    if (lane == 0) base = atomicAdd(counter, 4);
    base = uniform32(base);
    if (base >= n) break;
    // row r of the wave takes item base + r; rows past the list repeat the
    // last item and store nothing
    const int item = base + (lane >> 4);
    const bool active = item < n;
    const int p = order[min(item, n - 1)];
    const PairMeta m = meta[p];
    const uint8_t *A = seq + m.off_r;
    const uint8_t *B = seq + m.off_c;
    const int la = m.lr - 1, lb = m.lc - 1, E0 = m.p0, W = m.p1;
    uint32_t *dmw = (uint32_t *)(dm_base + m.dm_off);
    QuadLanes<H> Q;
    // per diagonal pair: all ones if the even / odd cell is inside the band
    // (steady groups); lmask: all ones on lane 0 (the left boundary column)
    int Lk[H], open0[H], open1[H];
    const int lmask = q == 0 ? -1 : 0;
    auto row_consts = [&](int ic, int ip, int i, int &ge, int &gop, int &ve) {
      ge = s_cost[ic * 32 + kGap];
      gop = (!(kGap & ip) && (kGap & ic)) ? 0 : go;
      ve = (i > 1 && (ip & kGap) && !(ic & kGap)) ? gop + ge : ge;
    };
    auto col_consts = [&](int jc, int jp, int j, int &gr, int &gop, int &x) {
      gr = s_prep[jc];
      gop = (!(kGap & jp) && (kGap & jc)) ? 0 : go;
      x = ((jp & kGap) && !(jc & kGap)) ? gop + gr : gr;
      x = j == 1 ? gr : x;
    };
#pragma unroll
    for (int k = 0; k < H; ++k) {
      const int L = H * q + k;
      Lk[k] = L;
      // the state before step 0: row 19 - L, column L - 20
      Q.ic[k] = load_or0(A, 19 - L, la);
      Q.ip[k] = load_or0(A, 18 - L, la);
      row_consts(Q.ic[k], Q.ip[k], 19 - L, Q.si_ge[k], Q.si_go[k], Q.sive[k]);
      Q.jc[k] = load_or0(B, L - 20, lb);
      Q.jp[k] = load_or0(B, L - 21, lb);
      col_consts(Q.jc[k], Q.jp[k], L - 20, Q.gap_row[k], Q.sj_go[k], Q.hx[k]);
      Q.S[k] = AffDiag{kHigh, kHigh, kHigh, kHigh, kHigh, kHigh, kHigh, kHigh};
      Q.fcost[k] = 0;
      const int d0 = 2 * L - 40;
      open0[k] = (L == 0 || d0 >= E0) ? 0 : -1;
      open1[k] = d0 + 1 >= E0 ? 0 : -1;
    }
    const int Gs = (la + lb + 4) >> 2;   // groups of 4 steps of this row's pair
    const int pf = active ? pure[p] : 3;
    // steady groups of this row's pair (all cells inside the sequences)
    const int t_lo = E0 + 82, t_hi = min(2 * lb - E0, 2 * la - 40);
    const int glo = active ? (t_lo + 3) >> 2 : 0;
    const int ghi = active ? (t_hi >= 3 ? (t_hi - 3) / 4 + 1 : 0) : 0x3fffffff;
    int groups = 0, g_lo = 0, g_hi = 0x3fffffff, all_pure = 1, all_free = 1;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int f = __builtin_amdgcn_readlane(pf, 16 * r);
      all_pure &= f != 0;
      all_free &= f == 3;
      groups = max(groups, __builtin_amdgcn_readlane(Gs, 16 * r));
      g_lo = max(g_lo, __builtin_amdgcn_readlane(glo, 16 * r));
      g_hi = min(g_hi, __builtin_amdgcn_readlane(ghi, 16 * r));
    }
    // characters entering the row at pair-step s: A[s + 20] at lane 0 and
    // B[s - 20 + 16 H] at lane 15.  Windows of 16 steps: before the
    // rotations of a window's steps, lane q holds step 16 w + q (rows) and
    // step 16 w + ((q + 1) & 15) (columns)
    int awin = load_or0(A, 20 + q, la);
    int bwin = load_or0(B, 16 * H - 20 + ((q + 1) & 15), lb);
    int anext = 0, bnext = 0;
    // variant: 0 edge groups, 1 steady groups, 2 steady groups of four pure
    // pairs, 3 steady groups of four gap-free pairs (see aff_diag_cell)
    auto group = [&](int g, auto variant_tag) {
      constexpr int kVar = decltype(variant_tag)::value;
      constexpr bool kSt = kVar >= 1, kPu = kVar >= 2, kGF = kVar == 3;
      if ((g & 7) == 0) {   // a new 16-step window; prefetch the one after it
        if (g) {
          awin = anext;
          bwin = bnext;
        }
        anext = load_or0(A, 2 * g + 16 + 20 + q, la);
        bnext = load_or0(B, 2 * g + 16 + 16 * H - 20 + ((q + 1) & 15), lb);
      }
      {
        // per-pair values pass through an opaque move at every group, so
        // that the compiler cannot hoist the band-edge comparisons of all
        // cells out of the loop (as lane masks, that exhausts the scalar
        // registers and spills)
        int lag = la, lbg = lb, E0g = E0, Lg[H], e0g[H], e1g[H], lmg = lmask;
        asm volatile("" : "+v"(lag), "+v"(lbg), "+v"(E0g), "+v"(lmg));
#pragma unroll
        for (int k = 0; k < H; ++k) {
          Lg[k] = Lk[k];
          e0g[k] = open0[k];
          e1g[k] = open1[k];
          asm volatile("" : "+v"(Lg[k]), "+v"(e0g[k]), "+v"(e1g[k]));
        }
        uint32_t acc[H];
#pragma unroll
        for (int k = 0; k < H; ++k) acc[k] = 0u;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int s = 2 * g + h;
          {   // even step: next row character into every lane's virtual lane 0
            dpp_exec_guard();
            int nic = dpp_row_from_below(Q.ic[H - 1], awin);
            int eh_l0 = dpp_row_from_below(Q.S[H - 1].eh1, kHigh);
            int cb_l0 = dpp_row_from_below(Q.S[H - 1].cb1, kHigh);
            awin = dpp_row_rotate_down(awin);   // the next step's entry to lane 0
            asm volatile("" : "+v"(nic), "+v"(eh_l0), "+v"(cb_l0), "+v"(awin));
            const int old0 = Q.ic[0];
#pragma unroll
            for (int k = H - 1; k > 0; --k) {
              Q.ip[k] = Q.ic[k];
              Q.ic[k] = Q.ic[k - 1];
              Q.si_ge[k] = Q.si_ge[k - 1];
              if (!kGF) {
                Q.si_go[k] = Q.si_go[k - 1];
                Q.sive[k] = Q.sive[k - 1];
              }
            }
            Q.ip[0] = old0;
            Q.ic[0] = nic;
            if (kGF) Q.si_ge[0] = s_cost[nic * 32 + kGap];
            else row_consts(nic, old0, s + 20 - Lg[0], Q.si_ge[0], Q.si_go[0], Q.sive[0]);
            // descending k: virtual lane k reads k-1's odd cell before it changes
#pragma unroll
            for (int k = H - 1; k >= 0; --k) {
              const int eh_l = k == 0 ? eh_l0 : Q.S[k > 0 ? k - 1 : 0].eh1;
              const int cb_l = k == 0 ? cb_l0 : Q.S[k > 0 ? k - 1 : 0].cb1;
              acc[k] |= aff_diag_cell<kBt, kSt, kPu, kGF>(Q.S[k], eh_l, cb_l, Q.S[k].ev1, Q.S[k].cb1, s + 20 - Lg[k],
                                                s - 20 + Lg[k], lag, lbg, E0g, go, s_cost, Q.ic[k], Q.ip[k],
                                                Q.si_ge[k], Q.si_go[k], Q.sive[k], Q.jc[k], Q.gap_row[k], Q.sj_go[k],
                                                Q.hx[k], Q.fcost[k], e0g[k], k == 0 ? lmg : 0)
                        << (16 * h);
            }
          }
          {   // odd step: next column character into every lane's virtual lane H-1
            dpp_exec_guard();
            int njc = dpp_row_from_above(Q.jc[0], bwin);
            int ev_u0 = dpp_row_from_above(Q.S[0].ev1, kHigh);
            int cb_u0 = dpp_row_from_above(Q.S[0].cb1, kHigh);
            bwin = dpp_row_rotate_down(bwin);   // the next step's entry to lane 15
            asm volatile("" : "+v"(njc), "+v"(ev_u0), "+v"(cb_u0), "+v"(bwin));
            const int oldt = Q.jc[H - 1];
#pragma unroll
            for (int k = 0; k < H - 1; ++k) {
              Q.jp[k] = Q.jc[k];
              Q.jc[k] = Q.jc[k + 1];
              Q.gap_row[k] = Q.gap_row[k + 1];
              if (!kGF) {
                Q.sj_go[k] = Q.sj_go[k + 1];
                Q.hx[k] = Q.hx[k + 1];
              }
            }
            Q.jp[H - 1] = oldt;
            Q.jc[H - 1] = njc;
            if (kGF) Q.gap_row[H - 1] = s_prep[njc];
            else col_consts(njc, oldt, s - 19 + Lg[H - 1], Q.gap_row[H - 1], Q.sj_go[H - 1], Q.hx[H - 1]);
            // ascending k: virtual lane k reads k+1's even cell before it changes
#pragma unroll
            for (int k = 0; k < H; ++k) {
              const int ev_u = k == H - 1 ? ev_u0 : Q.S[k < H - 1 ? k + 1 : 0].ev1;
              const int cb_u = k == H - 1 ? cb_u0 : Q.S[k < H - 1 ? k + 1 : 0].cb1;
              acc[k] |= aff_diag_cell<kBt, kSt, kPu, kGF>(Q.S[k], Q.S[k].eh1, Q.S[k].cb1, ev_u, cb_u, s + 20 - Lg[k],
                                                s - 19 + Lg[k], lag, lbg, E0g, go, s_cost, Q.ic[k], Q.ip[k],
                                                Q.si_ge[k], Q.si_go[k], Q.sive[k], Q.jc[k], Q.gap_row[k], Q.sj_go[k],
                                                Q.hx[k], Q.fcost[k], e1g[k], 0)
                        << (16 * h + 8);
            }
          }
        }
        if (kBt) {
          const bool gin = active && g < Gs;
#pragma unroll
          for (int k = 0; k < H; ++k) {
            uint32_t *dst = (gin && Lk[k] < W) ? dmw + (size_t)g * W + Lk[k] : sink + lane;
            *dst = acc[k];
          }
        }
      }
    };
    // The longest items set the end of the kernel: their waves issue first
    // (the work list is longest-first, so they start first too).
    if (groups >= 1536) __builtin_amdgcn_s_setprio(2);
    else if (groups >= 768) __builtin_amdgcn_s_setprio(1);
    else __builtin_amdgcn_s_setprio(0);
    // three loops (edge groups, steady groups, edge groups), so that no code
    // of a group is shared with -- and hoisted above -- a branch between the
    // two variants
    const int gs_lo = min(g_lo, groups), gs_hi = max(gs_lo, min(g_hi, groups));
    int g = 0;
    for (; g < gs_lo; ++g) group(g, std::integral_constant<int, 0>{});
    if (all_free) {
      for (; g < gs_hi; ++g) group(g, std::integral_constant<int, 3>{});
      // the gap-free groups kept only si_ge and gap_row; for rows and columns
      // without gap-bearing characters the other costs follow from them
#pragma unroll
      for (int k = 0; k < H; ++k) {
        Q.si_go[k] = go;
        Q.sive[k] = Q.si_ge[k];
        Q.sj_go[k] = go;
        Q.hx[k] = Q.gap_row[k];
      }
    } else if (all_pure) {
      for (; g < gs_hi; ++g) group(g, std::integral_constant<int, 2>{});
    } else {
      for (; g < gs_hi; ++g) group(g, std::integral_constant<int, 1>{});
    }
    for (; g < groups; ++g) group(g, std::integral_constant<int, 0>{});
    // cell (la, lb): diagonal pair (lb - la + 40) / 2 = H qs + ks
    const int Ls = (lb - la + 40) >> 1;
    const int qs = Ls / H, ks = Ls - H * (Ls / H);
    int fc = Q.fcost[0];
#pragma unroll
    for (int k = 1; k < H; ++k) fc = ks == k ? Q.fcost[k] : fc;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int src = 16 * r + __builtin_amdgcn_readlane(qs, 16 * r);
      const int c = __builtin_amdgcn_readlane(fc, src);
      const int pr = __builtin_amdgcn_readlane(p, 16 * r);
      if (__builtin_amdgcn_readlane((int)active, 16 * r)) cost_out[pr] = c;   // every lane stores the same value
    }
    if (kBt) {
      // read after the fill, so that nothing of the backtrace is held
      // across the fill loop
      const TraceOut tro = *trop;
      uint32_t *tile = s_tile[wave_in_group()];
      __builtin_amdgcn_s_setprio(0);
      for (int r = 0; r < 4; ++r) {
        if (!__builtin_amdgcn_readlane((int)active, 16 * r)) continue;
        const int pr = __builtin_amdgcn_readlane(p, 16 * r);
        aff_trace_own<16>(uniform_meta(meta, pr), pr, seq, dm_base, s_med, tile, tro);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Affine band, eight pairs per wave in 16-bit halves.
// ---------------------------------------------------------------------------
// aff_quad_kernel's layout (16-lane rows, H = 3 diagonal pairs per lane) with
// two pairs per row: every register holds one quantity of two pairs, pair 2r
// of the wave's items in the low halves of row r's lanes and pair 2r + 1 in
// the high halves, so one packed instruction (v_pk_add_u16 with clamp,
// v_pk_min_u16, v_pk_sub_u16, bit selects) does the work of two.
//
// A value v is held as min(4 v, 0xffff): the two low bits are free for the
// rank of a ranked minimum (value << 2 | rank in aff_diag_cell), and sums and
// minima of saturated values are the saturated sums and minima of the true
// values, so every value below 0xfffc is exact.  The sentinel is 0xffff in
// both halves (kHigh saturates), which makes "closed" a bitwise OR with the
// inverted band mask.  The backtrace follows states whose values are at most
// the pair's cost, and a comparison of two values is exact whenever the
// smaller one is, so a pair whose cost is below 0x3fff gets exactly
// aff_quad_kernel's directions and cost.  A pair whose final value saturated
// is queued (redo list) for aff_quad_kernel, launched after this kernel.
constexpr uint32_t kPkHigh = 0xffffffffu;
constexpr uint32_t kPkRank1 = 0x00010001u, kPkRank2 = 0x00020002u, kPkRank3 = 0x00030003u;

__device__ __forceinline__ uint32_t c16_max(uint32_t a, uint32_t b) {
  return c16u(__builtin_elementwise_max(c16v(a), c16v(b)));
}
// per half: 0xffff where the bit `bit` of the half is set
template <int kBit>
__device__ __forceinline__ uint32_t c16_bitmask(uint32_t x) {
  typedef short s16x2 __attribute__((ext_vector_type(2)));
  const c16x2 up = c16v(x) << (c16x2){15 - kBit, 15 - kBit};
  const s16x2 m = __builtin_bit_cast(s16x2, up) >> (s16x2){15, 15};
  return __builtin_bit_cast(uint32_t, m);
}
// per half: 1 where a < b, else 0 (as two packed operations: written out,
// the compiler turns the saturated difference's minimum with 1 into a
// compare and a select per half)
__device__ __forceinline__ uint32_t c16_lt01(uint32_t a, uint32_t b) {
  uint32_t r;
  asm("v_pk_min_u16 %0, %1, %2" : "=v"(r) : "v"(c16_sub_sat(b, a)), "s"(kPkRank1));
  return r;
}
// two 16-bit table entries at the byte offsets held in the halves of a
__device__ __forceinline__ uint32_t lds_pair(const uint16_t *t, uint32_t a) {
  const uint32_t lo = *(const uint16_t *)((const char *)t + (a & 0xffffu));
  const uint32_t hi = *(const uint16_t *)((const char *)t + (a >> 16));
  return lo | (hi << 16);
}
// 0xffff in the low half if a, in the high half if b
__device__ __forceinline__ uint32_t halves(bool a, bool b) { return (a ? 0xffffu : 0u) | (b ? 0xffff0000u : 0u); }

// Characters are held scaled for the table reads: a rows character x as
// 64 x (the byte offset of its row of the 32 x 32 table of 16-bit entries),
// a columns character as 2 x; the gap bit is then bit 10 (rows) / bit 5
// (columns).
constexpr uint32_t kPkRowGap = 16u * 64u, kPkColGap = 16u * 2u;

// One cell of two pairs: aff_diag_cell's recurrence in packed halves.  The
// edge arguments (kSteady false) are per pair; i and j are the same for
// both pairs of a register (same step, same virtual lane).
template <bool kBt, bool kSteady, bool kPure, bool kGapFree>
__device__ __forceinline__ uint32_t aff_pk_cell(AffDiag &S, uint32_t eh_l, uint32_t cb_l, uint32_t ev_u,
                                                uint32_t cb_u, int i, int j, int laA, int laB, int lbA, int lbB,
                                                int E0A, int E0B, uint32_t go4, const uint16_t *s_c4, uint32_t ic,
                                                uint32_t ip, uint32_t si_ge, uint32_t si_go, uint32_t sive,
                                                uint32_t jc, uint32_t gap_row, uint32_t sj_go, uint32_t hx,
                                                uint32_t &fcost, uint32_t open_mask, uint32_t left_mask) {
  if (kGapFree) {
    si_go = go4;
    sj_go = go4;
    sive = si_ge;
    hx = gap_row;
  }
  uint32_t ev, endv_n, ext_v;
  if (kGapFree) {
    // both candidates add si_ge: min(ev_u, cb_u + go) + si_ge
    const uint32_t t_v = c16_add_sat(cb_u, go4);
    endv_n = c16_lt01(ev_u, t_v);   // 1 where the open is the smaller
    ev = c16_add_sat(c16_min(ev_u, t_v), si_ge);
    ext_v = c16_add_sat(ev_u, si_ge);
  } else {
    ext_v = c16_add_sat(ev_u, sive);
    const uint32_t open_v = c16_add_sat(c16_add_sat(cb_u, si_go), si_ge);
    endv_n = c16_lt01(ext_v, open_v);
    ev = c16_min(ext_v, open_v);
  }
  uint32_t ic_g = 0, jc_g = 0;
  if (!kGapFree) {
    ic_g = c16_bitmask<10>(ic);
    jc_g = c16_bitmask<5>(jc);
  }
  const uint32_t dg = kGapFree ? lds_pair(s_c4, ic + jc) : lds_pair(s_c4, (ic & 0x03c003c0u) + (jc & 0x001e001eu));
  const uint32_t fv = kGapFree ? c16_add_sat(S.ev2, dg) : c16_add_sat(c16_add_sat(S.ev2, dg), sj_go & ic_g);
  const uint32_t fh = kGapFree ? c16_add_sat(S.eh2, dg) : c16_add_sat(c16_add_sat(S.eh2, dg), si_go & jc_g);
  const uint32_t fc = c16_add_sat(S.cb2, dg);
  uint32_t eb, endb_n, kc;
  if (kPure) {
    eb = kPkHigh;
    endb_n = 0;
    kc = c16_min(c16_min(fh, fv | kPkRank2), fc | kPkRank3);
  } else {
    const uint32_t both_g = ic_g & jc_g;
    const uint32_t diag = ~both_g;   // 0 where both characters carry the gap, else the sentinel
    const uint32_t ext_b = c16_add_sat(S.eb2, diag);
    uint32_t open_b;
    if (kBt) {
      open_b = c16_add_sat(S.cb2, diag);
    } else {
      const uint32_t ip_g = c16_bitmask<10>(ip);
      const uint32_t flag2 = ~ip_g & ~jc_g;
      const uint32_t add = (both_g & ~flag2 & c16_add_sat(go4, go4)) | ~both_g;
      open_b = c16_add_sat(S.cb2, add);
    }
    endb_n = c16_lt01(ext_b, open_b);
    eb = c16_min(ext_b, open_b);
    const uint32_t fd = c16_add_sat(c16_add_sat(S.eb2, dg), c16_max(si_go, sj_go));
    kc = c16_min(c16_min(fh, fd | kPkRank1), c16_min(fv | kPkRank2, fc | kPkRank3));
  }
  uint32_t cb = kc & 0xfffcfffcu;
  const uint32_t anext = c16_add(kc, kPkRank1) & kPkRank3;
  uint32_t eh, endh_n;
  if (kGapFree) {
    const uint32_t t_h = c16_add_sat(cb_l, go4);
    endh_n = c16_lt01(eh_l, t_h);
    eh = c16_add_sat(c16_min(eh_l, t_h), gap_row);
  } else {
    const uint32_t bval = c16_add_sat(c16_add_sat(cb_l, sj_go), gap_row);
    const uint32_t ext_h = c16_add_sat(eh_l, hx);
    endh_n = c16_lt01(ext_h, bval);
    eh = c16_min(ext_h, bval);
  }
  const uint32_t kf = kPure ? c16_min(c16_min(eh, cb | kPkRank1), ev | kPkRank2)
                            : c16_min(c16_min(eh, cb | kPkRank1), c16_min(ev | kPkRank2, eb | kPkRank3));
  // todo | END_V << 2 | END_H << 3 | END_B << 4 | anext << 5, from the
  // inverted flags (one XOR turns them back)
  uint32_t byte = kf & kPkRank3;
  byte |= endv_n << 2;
  byte |= endh_n << 3;
  byte |= endb_n << 4;
  byte |= anext << 5;
  byte ^= 0x001c001cu;
  if (kBt) asm volatile("" : "+v"(byte));
  if (kSteady) {
    const uint32_t ev_closed = ext_v | ~left_mask;
    ev = (open_mask & ev) | (~open_mask & ev_closed);
    eh |= ~open_mask;
    eb |= ~open_mask;
    cb |= ~open_mask;
    const uint32_t byte_closed = (left_mask & (kAffLeft * 0x00010001u)) | (~left_mask & (kAffRight * 0x00010001u));
    byte = (open_mask & byte) | (~open_mask & byte_closed);
  } else {
    (void)open_mask;
    (void)left_mask;
    const int d = j - i;
    const bool vA = i >= 0 && i <= laA && j >= 0 && j <= lbA && d <= E0A;
    const bool vB = i >= 0 && i <= laB && j >= 0 && j <= lbB && d <= E0B;
    const bool lside = i >= 1 && (j == 0 || d == -40);
    const bool lbA_ = vA && lside, lbB_ = vB && lside;
    const bool rbA = vA && i >= 1 && d == E0A, rbB = vB && i >= 1 && d == E0B;
    const bool r0A = vA && i == 0, r0B = vB && i == 0;
    const bool j0 = j == 0;
    uint32_t m_open = halves(vA && !lbA_ && !rbA && !r0A, vB && !lbB_ && !rbB && !r0B);
    uint32_t m_lb = halves(lbA_, lbB_);
    uint32_t m_r0 = halves(r0A, r0B);
    uint32_t m_fin = halves(vA && i == laA && j == lbA, vB && i == laB && j == lbB);
    asm volatile("" : "+v"(m_open), "+v"(m_lb), "+v"(m_r0), "+v"(m_fin));
    const uint32_t r0eh = j0 ? go4 : c16_add_sat(eh_l, gap_row);
    const uint32_t alt_ev = (m_lb & ext_v) | (~m_lb & ((m_r0 & (j0 ? go4 : kPkHigh)) | ~m_r0));
    ev = (m_open & ev) | (~m_open & alt_ev);
    eh = (m_open & eh) | (~m_open & ((m_r0 & r0eh) | ~m_r0));
    eb = (m_open & eb) | (~m_open & ((m_r0 & (j0 ? 0u : kPkHigh)) | ~m_r0));
    cb = (m_open & cb) | (~m_open & ((m_r0 & (j0 ? 0u : r0eh)) | ~m_r0));
    const uint32_t keep = m_open | m_r0;
    const uint32_t edge_byte = (m_lb & (kAffLeft * 0x00010001u)) | (~m_lb & (kAffRight * 0x00010001u));
    byte = (keep & byte) | (~keep & edge_byte);
    fcost = (m_fin & kf) | (~m_fin & fcost);
  }
  S.eh2 = S.eh1; S.ev2 = S.ev1; S.eb2 = S.eb1; S.cb2 = S.cb1;
  S.eh1 = (int)eh; S.ev1 = (int)ev; S.eb1 = (int)eb; S.cb1 = (int)cb;
  return byte;
}

template <int H>
struct PkLanes {
  AffDiag S[H];
  uint32_t ic[H], ip[H], si_ge[H], si_go[H], sive[H];   // row i = s + 20 - L (scaled characters) and its gap costs
  uint32_t jc[H], jp[H], gap_row[H], sj_go[H], hx[H];   // column and its gap costs
  uint32_t fcost[H];
};

// A rows character (scaled, see kPkRowGap) of both pairs at position idx
// of their rows sequences (0 outside 0..la).
__device__ __forceinline__ uint32_t pk_row_char(const uint8_t *A0, const uint8_t *A1, int idx, int la0, int la1) {
  return ((uint32_t)load_or0(A0, idx, la0) << 6) | ((uint32_t)load_or0(A1, idx, la1) << 22);
}
__device__ __forceinline__ uint32_t pk_col_char(const uint8_t *B0, const uint8_t *B1, int idx, int lb0, int lb1) {
  return ((uint32_t)load_or0(B0, idx, lb0) << 1) | ((uint32_t)load_or0(B1, idx, lb1) << 17);
}

template <bool kBt, int H>
__global__ __launch_bounds__(kWave * 4, 2) void aff_pk_kernel(
    const PairMeta *__restrict__ meta, const int32_t *__restrict__ order, int n, int *counter,
    const uint8_t *__restrict__ seq, const DevTcm *__restrict__ tcm, uint8_t *dm_base, int32_t *cost_out,
    uint32_t *sink, const uint8_t *__restrict__ pure, int32_t *redo, int *redo_n, int32_t *tlist, int *tlist_n) {
  __shared__ uint16_t s_c4[32 * 32];   // min(4 cost, 0xffff)
  __shared__ uint16_t s_p4[32];        // min(4 prepend, 0xffff)
  for (int k = threadIdx.x; k < 32 * 32; k += blockDim.x) s_c4[k] = (uint16_t)min(4ll * tcm->cost[k], 0xffffll);
  for (int k = threadIdx.x; k < 32; k += blockDim.x) s_p4[k] = (uint16_t)min(4ll * tcm->prepend[k], 0xffffll);
  __syncthreads();
  const uint32_t go4 = (uint32_t)min(4ll * tcm->gap_open, 0xffffll) * 0x00010001u;
  const int lane = lane_id();
  const int q = lane & 15;
  for (;;) {
    int base = 0;
    if (lane == 0) base = atomicAdd(counter, 8);
    base = uniform32(base);
    if (base >= n) break;
    // row r of the wave takes items base + 2r (low halves) and base + 2r + 1
    // (high halves); slots past the list repeat the last item and store
    // nothing
    const int itA = base + 2 * (lane >> 4), itB = itA + 1;
    const bool actA = itA < n, actB = itB < n;
    const int pA = order[min(itA, n - 1)], pB = order[min(itB, n - 1)];
    const PairMeta mA = meta[pA], mB = meta[pB];
    const uint8_t *AA = seq + mA.off_r, *AB = seq + mB.off_r;
    const uint8_t *BA = seq + mA.off_c, *BB = seq + mB.off_c;
    const int laA = mA.lr - 1, lbA = mA.lc - 1, E0A = mA.p0, WA = mA.p1;
    const int laB = mB.lr - 1, lbB = mB.lc - 1, E0B = mB.p0, WB = mB.p1;
    uint32_t *dmwA = (uint32_t *)(dm_base + mA.dm_off);
    uint32_t *dmwB = (uint32_t *)(dm_base + mB.dm_off);
    PkLanes<H> Q;
    int Lk[H];
    uint32_t open0[H], open1[H];
    const uint32_t lmask = q == 0 ? kPkHigh : 0u;
    // row constants of both pairs (scaled characters)
    auto row_consts = [&](uint32_t ic, uint32_t ip, int i, uint32_t &ge, uint32_t &gop, uint32_t &ve) {
      ge = lds_pair(s_c4, ic + 2u * kGap * 0x00010001u);
      const uint32_t ic_g = c16_bitmask<10>(ic), ip_g = c16_bitmask<10>(ip);
      gop = go4 & ~(~ip_g & ic_g);
      const uint32_t m = i > 1 ? (ip_g & ~ic_g) : 0u;
      ve = (m & c16_add_sat(gop, ge)) | (~m & ge);
    };
    auto col_consts = [&](uint32_t jc, uint32_t jp, int j, uint32_t &gr, uint32_t &gop, uint32_t &x) {
      gr = lds_pair(s_p4, jc);
      const uint32_t jc_g = c16_bitmask<5>(jc), jp_g = c16_bitmask<5>(jp);
      gop = go4 & ~(~jp_g & jc_g);
      const uint32_t m = j == 1 ? 0u : (jp_g & ~jc_g);
      x = (m & c16_add_sat(gop, gr)) | (~m & gr);
    };
#pragma unroll
    for (int k = 0; k < H; ++k) {
      const int L = H * q + k;
      Lk[k] = L;
      // the state before step 0: row 19 - L, column L - 20
      Q.ic[k] = pk_row_char(AA, AB, 19 - L, laA, laB);
      Q.ip[k] = pk_row_char(AA, AB, 18 - L, laA, laB);
      row_consts(Q.ic[k], Q.ip[k], 19 - L, Q.si_ge[k], Q.si_go[k], Q.sive[k]);
      Q.jc[k] = pk_col_char(BA, BB, L - 20, lbA, lbB);
      Q.jp[k] = pk_col_char(BA, BB, L - 21, lbA, lbB);
      col_consts(Q.jc[k], Q.jp[k], L - 20, Q.gap_row[k], Q.sj_go[k], Q.hx[k]);
      Q.S[k] = AffDiag{(int)kPkHigh, (int)kPkHigh, (int)kPkHigh, (int)kPkHigh,
                       (int)kPkHigh, (int)kPkHigh, (int)kPkHigh, (int)kPkHigh};
      Q.fcost[k] = kPkHigh;
      const int d0 = 2 * L - 40;
      open0[k] = halves(!(L == 0 || d0 >= E0A), !(L == 0 || d0 >= E0B));
      open1[k] = halves(!(d0 + 1 >= E0A), !(d0 + 1 >= E0B));
    }
    const int GsA = (laA + lbA + 4) >> 2, GsB = (laB + lbB + 4) >> 2;
    const int pfA = actA ? pure[pA] : 3, pfB = actB ? pure[pB] : 3;
    // steady groups of each pair (all cells inside the sequences)
    auto steady = [&](bool act, int la, int lb, int E0, int &glo, int &ghi) {
      const int t_lo = E0 + 82, t_hi = min(2 * lb - E0, 2 * la - 40);
      glo = act ? (t_lo + 3) >> 2 : 0;
      ghi = act ? (t_hi >= 3 ? (t_hi - 3) / 4 + 1 : 0) : 0x3fffffff;
    };
    int gloA, ghiA, gloB, ghiB;
    steady(actA, laA, lbA, E0A, gloA, ghiA);
    steady(actB, laB, lbB, E0B, gloB, ghiB);
    int groups = 0, g_lo = 0, g_hi = 0x3fffffff, all_pure = 1, all_free = 1;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int fA = __builtin_amdgcn_readlane(pfA, 16 * r), fB = __builtin_amdgcn_readlane(pfB, 16 * r);
      all_pure &= fA != 0 && fB != 0;
      all_free &= fA == 3 && fB == 3;
      groups = max(groups, max(__builtin_amdgcn_readlane(actA ? GsA : 0, 16 * r),
                               __builtin_amdgcn_readlane(actB ? GsB : 0, 16 * r)));
      g_lo = max(g_lo, max(__builtin_amdgcn_readlane(gloA, 16 * r), __builtin_amdgcn_readlane(gloB, 16 * r)));
      g_hi = min(g_hi, min(__builtin_amdgcn_readlane(ghiA, 16 * r), __builtin_amdgcn_readlane(ghiB, 16 * r)));
    }
    // characters entering the row at pair-step s: A[s + 20] at lane 0 and
    // B[s - 20 + 16 H] at lane 15, from 16-step windows as in aff_quad_kernel
    uint32_t awin = pk_row_char(AA, AB, 20 + q, laA, laB);
    uint32_t bwin = pk_col_char(BA, BB, 16 * H - 20 + ((q + 1) & 15), lbA, lbB);
    uint32_t anext = 0, bnext = 0;
    auto group = [&](int g, auto variant_tag) {
      constexpr int kVar = decltype(variant_tag)::value;
      constexpr bool kSt = kVar >= 1, kPu = kVar >= 2, kGF = kVar == 3;
      if ((g & 7) == 0) {
        if (g) {
          awin = anext;
          bwin = bnext;
        }
        anext = pk_row_char(AA, AB, 2 * g + 16 + 20 + q, laA, laB);
        bnext = pk_col_char(BA, BB, 2 * g + 16 + 16 * H - 20 + ((q + 1) & 15), lbA, lbB);
      }
      {
        int laAg = laA, laBg = laB, lbAg = lbA, lbBg = lbB, E0Ag = E0A, E0Bg = E0B, Lg[H];
        uint32_t e0g[H], e1g[H], lmg = lmask;
        asm volatile("" : "+v"(laAg), "+v"(laBg), "+v"(lbAg), "+v"(lbBg), "+v"(E0Ag), "+v"(E0Bg), "+v"(lmg));
#pragma unroll
        for (int k = 0; k < H; ++k) {
          Lg[k] = Lk[k];
          e0g[k] = open0[k];
          e1g[k] = open1[k];
          asm volatile("" : "+v"(Lg[k]), "+v"(e0g[k]), "+v"(e1g[k]));
        }
        uint32_t by[H][4];   // direction bytes: even/odd step of the group's two pair-steps
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int s = 2 * g + h;
          {   // even step: next row character into every lane's virtual lane 0
            dpp_exec_guard();
            uint32_t nic = (uint32_t)dpp_row_from_below((int)Q.ic[H - 1], (int)awin);
            uint32_t eh_l0 = (uint32_t)dpp_row_from_below(Q.S[H - 1].eh1, (int)kPkHigh);
            uint32_t cb_l0 = (uint32_t)dpp_row_from_below(Q.S[H - 1].cb1, (int)kPkHigh);
            awin = (uint32_t)dpp_row_rotate_down((int)awin);
            asm volatile("" : "+v"(nic), "+v"(eh_l0), "+v"(cb_l0), "+v"(awin));
            const uint32_t old0 = Q.ic[0];
#pragma unroll
            for (int k = H - 1; k > 0; --k) {
              Q.ip[k] = Q.ic[k];
              Q.ic[k] = Q.ic[k - 1];
              Q.si_ge[k] = Q.si_ge[k - 1];
              if (!kGF) {
                Q.si_go[k] = Q.si_go[k - 1];
                Q.sive[k] = Q.sive[k - 1];
              }
            }
            Q.ip[0] = old0;
            Q.ic[0] = nic;
            if (kGF) Q.si_ge[0] = lds_pair(s_c4, nic + 2u * kGap * 0x00010001u);
            else row_consts(nic, old0, s + 20 - Lg[0], Q.si_ge[0], Q.si_go[0], Q.sive[0]);
#pragma unroll
            for (int k = H - 1; k >= 0; --k) {
              const uint32_t eh_l = k == 0 ? eh_l0 : (uint32_t)Q.S[k > 0 ? k - 1 : 0].eh1;
              const uint32_t cb_l = k == 0 ? cb_l0 : (uint32_t)Q.S[k > 0 ? k - 1 : 0].cb1;
              by[k][2 * h] = aff_pk_cell<kBt, kSt, kPu, kGF>(
                  Q.S[k], eh_l, cb_l, (uint32_t)Q.S[k].ev1, (uint32_t)Q.S[k].cb1, s + 20 - Lg[k], s - 20 + Lg[k],
                  laAg, laBg, lbAg, lbBg, E0Ag, E0Bg, go4, s_c4, Q.ic[k], Q.ip[k], Q.si_ge[k], Q.si_go[k],
                  Q.sive[k], Q.jc[k], Q.gap_row[k], Q.sj_go[k], Q.hx[k], Q.fcost[k], e0g[k], k == 0 ? lmg : 0u);
            }
          }
          {   // odd step: next column character into every lane's virtual lane H-1
            dpp_exec_guard();
            uint32_t njc = (uint32_t)dpp_row_from_above((int)Q.jc[0], (int)bwin);
            uint32_t ev_u0 = (uint32_t)dpp_row_from_above(Q.S[0].ev1, (int)kPkHigh);
            uint32_t cb_u0 = (uint32_t)dpp_row_from_above(Q.S[0].cb1, (int)kPkHigh);
            bwin = (uint32_t)dpp_row_rotate_down((int)bwin);
            asm volatile("" : "+v"(njc), "+v"(ev_u0), "+v"(cb_u0), "+v"(bwin));
            const uint32_t oldt = Q.jc[H - 1];
#pragma unroll
            for (int k = 0; k < H - 1; ++k) {
              Q.jp[k] = Q.jc[k];
              Q.jc[k] = Q.jc[k + 1];
              Q.gap_row[k] = Q.gap_row[k + 1];
              if (!kGF) {
                Q.sj_go[k] = Q.sj_go[k + 1];
                Q.hx[k] = Q.hx[k + 1];
              }
            }
            Q.jp[H - 1] = oldt;
            Q.jc[H - 1] = njc;
            if (kGF) Q.gap_row[H - 1] = lds_pair(s_p4, njc);
            else col_consts(njc, oldt, s - 19 + Lg[H - 1], Q.gap_row[H - 1], Q.sj_go[H - 1], Q.hx[H - 1]);
#pragma unroll
            for (int k = 0; k < H; ++k) {
              const uint32_t ev_u = k == H - 1 ? ev_u0 : (uint32_t)Q.S[k < H - 1 ? k + 1 : 0].ev1;
              const uint32_t cb_u = k == H - 1 ? cb_u0 : (uint32_t)Q.S[k < H - 1 ? k + 1 : 0].cb1;
              by[k][2 * h + 1] = aff_pk_cell<kBt, kSt, kPu, kGF>(
                  Q.S[k], (uint32_t)Q.S[k].eh1, (uint32_t)Q.S[k].cb1, ev_u, cb_u, s + 20 - Lg[k], s - 19 + Lg[k],
                  laAg, laBg, lbAg, lbBg, E0Ag, E0Bg, go4, s_c4, Q.ic[k], Q.ip[k], Q.si_ge[k], Q.si_go[k],
                  Q.sive[k], Q.jc[k], Q.gap_row[k], Q.sj_go[k], Q.hx[k], Q.fcost[k], e1g[k], 0u);
            }
          }
        }
        if (kBt) {
          const bool ginA = actA && g < GsA, ginB = actB && g < GsB;
#pragma unroll
          for (int k = 0; k < H; ++k) {
            // [even h0, odd h0, even h1, odd h1] of each pair, as
            // aff_quad_kernel stores them
            const uint32_t t0 = __builtin_amdgcn_perm(by[k][1], by[k][0], 0x06020400u);
            const uint32_t t1 = __builtin_amdgcn_perm(by[k][3], by[k][2], 0x06020400u);
            const uint32_t accA = __builtin_amdgcn_perm(t1, t0, 0x05040100u);
            const uint32_t accB = __builtin_amdgcn_perm(t1, t0, 0x07060302u);
            uint32_t *dA = (ginA && Lk[k] < WA) ? dmwA + (size_t)g * WA + Lk[k] : sink + lane;
            uint32_t *dB = (ginB && Lk[k] < WB) ? dmwB + (size_t)g * WB + Lk[k] : sink + lane;
            *dA = accA;
            *dB = accB;
          }
        }
      }
    };
    if (groups >= 1536) __builtin_amdgcn_s_setprio(2);
    else if (groups >= 768) __builtin_amdgcn_s_setprio(1);
    else __builtin_amdgcn_s_setprio(0);
    const int gs_lo = min(g_lo, groups), gs_hi = max(gs_lo, min(g_hi, groups));
    int g = 0;
    for (; g < gs_lo; ++g) group(g, std::integral_constant<int, 0>{});
    if (all_free) {
      for (; g < gs_hi; ++g) group(g, std::integral_constant<int, 3>{});
      // the gap-free groups kept only si_ge and gap_row; for rows and
      // columns without gap-bearing characters the other costs follow
#pragma unroll
      for (int k = 0; k < H; ++k) {
        Q.si_go[k] = go4;
        Q.sive[k] = Q.si_ge[k];
        Q.sj_go[k] = go4;
        Q.hx[k] = Q.gap_row[k];
      }
    } else if (all_pure) {
      for (; g < gs_hi; ++g) group(g, std::integral_constant<int, 2>{});
    } else {
      for (; g < gs_hi; ++g) group(g, std::integral_constant<int, 1>{});
    }
    for (; g < groups; ++g) group(g, std::integral_constant<int, 0>{});
    // cell (la, lb) of each pair: diagonal pair (lb - la + 40) / 2 = H qs + ks
    auto final_of = [&](int la, int lb, uint32_t shift, int &qs, uint32_t &fc) {
      const int Ls = (lb - la + 40) >> 1;
      qs = Ls / H;
      const int ks = Ls - H * (Ls / H);
      fc = Q.fcost[0];
#pragma unroll
      for (int k = 1; k < H; ++k) fc = ks == k ? Q.fcost[k] : fc;
      fc = (fc >> shift) & 0xffffu;
    };
    int qsA, qsB;
    uint32_t fcA, fcB;
    final_of(laA, lbA, 0, qsA, fcA);
    final_of(laB, lbB, 16, qsB, fcB);
    // per pair: its cost, or -1 when its final value saturated (redone by
    // the 32-bit kernel); lane 16 r holds row r's low pair, lane 16 r + 1 its
    // high pair
    int mine = -2;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int cA = __builtin_amdgcn_readlane((int)fcA, 16 * r + __builtin_amdgcn_readlane(qsA, 16 * r));
      const int cB = __builtin_amdgcn_readlane((int)fcB, 16 * r + __builtin_amdgcn_readlane(qsB, 16 * r));
      mine = lane == 16 * r ? cA : mine;
      mine = lane == 16 * r + 1 ? cB : mine;
    }
    const bool slot = (q < 2) && ((q == 0) ? actA : actB);
    const int pmine = q == 0 ? pA : pB;
    const bool sat = (uint32_t)(mine & 0xfffc) >= 0xfffcu;
    // the backtraces run in aff_trace_list_kernel, after every fill: a wave
    // that traced its eight pairs here would hold its SIMD's issue slots
    // through eight serial walks (two waves per SIMD fit this kernel)
    if (slot) {
      if (sat) {
        redo[atomicAdd(redo_n, 1)] = pmine;
      } else {
        cost_out[pmine] = (mine & 0xfffc) >> 2;
        if (kBt) tlist[atomicAdd(tlist_n, 1)] = pmine;
      }
    }
  }
}

// The backtraces of the pairs that the packed fill listed, one wave per pair
// (aff_trace_pair over tiles of 16 groups).  Launched after every fill of
// the run (the bands are complete); the list's length is on the device.
__global__ __launch_bounds__(kWave * 4, 5) void aff_trace_list_kernel(
    const PairMeta *__restrict__ meta, const int32_t *__restrict__ list, const int *list_n, int *counter,
    const uint8_t *__restrict__ seq, const DevTcm *__restrict__ tcm, const uint8_t *dm_base, const TraceOut *trop) {
  __shared__ uint8_t s_med[32 * 32];
  __shared__ uint32_t s_tile[4][kWave * 16];
  const int n = uniform32(*list_n);
  if (n == 0) return;   // uniform over the workgroup
  for (int k = threadIdx.x; k < 32 * 32; k += blockDim.x) s_med[k] = tcm->median[k];
  __syncthreads();
  const TraceOut tro = *trop;
  uint32_t *tile = s_tile[wave_in_group()];
  for (;;) {
    const int k = next_item(counter);
    if (k >= n) break;
    const int p = uniform32(list[k]);
    aff_trace_own<16>(uniform_meta(meta, p), p, seq, dm_base, s_med, tile, tro);
  }
}

// backtrace_affine (src/algn.c:1982-2097), one thread per pair.  Outputs per
// pair: median, medianwg, res of the caller's first and second sequence.
__global__ void aff_trace_kernel(const PairMeta *__restrict__ meta, const int32_t *__restrict__ order,
                                 int n, const uint8_t *__restrict__ seq, const DevTcm *__restrict__ tcm,
                                 const uint8_t *__restrict__ dm_base, uint8_t *out, int32_t *out_len,
                                 int32_t *status) {
  __shared__ uint8_t s_med[32 * 32];
  for (int k = threadIdx.x; k < 32 * 32; k += blockDim.x) s_med[k] = tcm->median[k];
  __syncthreads();
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int p = order[t];
  const PairMeta m = meta[p];
  const uint8_t *A = seq + m.off_r;
  const uint8_t *B = seq + m.off_c;
  const int la = m.lr - 1, lb = m.lc - 1, E0 = m.p0, S = m.stride;
  const uint8_t *dm = dm_base + m.dm_off;
  uint8_t *o = out + m.out_off;
  RevWriter wm, ww, wa, wb;
  wm.init(o, m.cap);
  ww.init(o + m.cap, m.cap);
  // resa belongs to the rows sequence: slot 2 if it is the caller's first
  wa.init(o + (m.swap ? 3 : 2) * m.cap, m.cap);
  wb.init(o + (m.swap ? 2 : 3) * m.cap, m.cap);
  int i = la, j = lb, mode = 0;   // 0 todo, 1 V, 2 H, 3 D, 4 A
  int ic = A[i], jc = B[j];
  int nm = 0, bad = 0;
  int last_med = -1;
  while (i != 0 && j != 0) {
    const int sp = 1 + (i > 40 ? i - 40 : 0);
    const int ep = min(lb, E0 + i - 1);
    const int hi = ep < lb ? ep + 1 : ep;
    if (j < sp - 1 || j > hi) { bad = 1; break; }
    // band layouts: rows (stride S from the first filled column) or the
    // anti-diagonal steps of aff_diag_kernel (W lanes of 4 steps per dword)
    const int f = m.p2 ? dm[(size_t)((i + j) >> 2) * (4 * m.p1) + (((j - i + 40) >> 1) << 2) + ((i + j) & 3)]
                       : dm[(size_t)(i - 1) * S + (j - (sp - 1))];
    if (mode == 0) {
      const int td = f & 3;
      mode = td == 0 ? 2 : (td == 1 ? 4 : (td == 2 ? 1 : 3));
    } else if (mode == 1) {
      if (f & 4) mode = 0;
      if (!(ic & kGap)) { wm.put(ic | kGap); ww.put(ic | kGap); ++nm; last_med = ic | kGap; }
      else ww.put(kGap);
      wa.put(ic); wb.put(kGap);
      --i; ic = A[i];
    } else if (mode == 2) {
      if (f & 8) mode = 0;
      if (!(jc & kGap)) { wm.put(jc | kGap); ww.put(jc | kGap); ++nm; last_med = jc | kGap; }
      else ww.put(kGap);
      wa.put(kGap); wb.put(jc);
      --j; jc = B[j];
    } else if (mode == 3) {
      if (f & 16) mode = 0;
      wa.put(ic); wb.put(jc); ww.put(kGap);
      --i; --j; jc = B[j]; ic = A[i];
    } else {
      const int an = (f >> 5) & 3;
      if (an == 1) mode = 2; else if (an == 2) mode = 3; else if (an == 3) mode = 1;
      const int prep = s_med[(ic & 15) * 32 + (jc & 15)];
      wm.put(prep); ww.put(prep); ++nm; last_med = prep;
      wa.put(ic); wb.put(jc);
      --i; --j; jc = B[j]; ic = A[i];
    }
  }
  if (!bad) {
    while (i != 0) {
      if (!(ic & kGap)) { wm.put(ic | kGap); ww.put(ic | kGap); ++nm; last_med = ic | kGap; }
      else ww.put(kGap);
      wa.put(ic); wb.put(kGap);
      --i; ic = A[i];
    }
    while (j != 0) {
      if (!(jc & kGap)) { wm.put(jc | kGap); ww.put(jc | kGap); ++nm; last_med = jc | kGap; }
      else ww.put(kGap);
      wa.put(kGap); wb.put(jc);
      --j; jc = B[j];
    }
    wa.put(kGap); wb.put(kGap); ww.put(kGap);
    if (nm == 0 || last_med != kGap) { wm.put(kGap); ++nm; }
  }
  wm.flush(); ww.flush(); wa.flush(); wb.flush();
  out_len[4 * p + 0] = m.cap - wm.pos;
  out_len[4 * p + 1] = m.cap - ww.pos;
  out_len[4 * p + (m.swap ? 3 : 2)] = m.cap - wa.pos;
  out_len[4 * p + (m.swap ? 2 : 3)] = m.cap - wb.pos;
  if (bad) mark_failed(status, p);
}


// ---------------------------------------------------------------------------
// Linear band (algn_fill_plane_2 + algn_fill_row family, src/algn.c:87-968)
// ---------------------------------------------------------------------------
constexpr int kLinRB = 512;
constexpr int kLinWaves = 4;
constexpr int kLinStage = 2048;   // longest s2 whose column data is staged in LDS
// LDS index of column j in the R-cells-per-lane rows: one pad word after every
// 32, so that lanes 4 or 8 words apart fall in distinct banks
__device__ __forceinline__ int lin_pad(int j) { return j + (j >> 5); }
constexpr int kLinRBP = kLinRB + kLinRB / 32;         // padded ring row
constexpr int kLinStageP = kLinStage + kLinStage / 32;
constexpr int kLinWideRB = 4096;  // widest row the wide-band kernel keeps in LDS
enum { LK_FIRST_TAIL = 0, LK_FIRST_COST = 1, LK_LEFT = 2 };
enum { RK_RIGHT = 0, RK_LAST = 1 };

// Row layout of algn_fill_plane_2; m = |s1| >= n = |s2| (with the leading
// gaps), w and h the clamped width/height, cs the case.
__device__ __forceinline__ void lin_row(int i, int m, int n, int cs, int w, int h, int &lo, int &hi,
                                        int &lk, int &rk) {
  if (cs == kLinFull) {
    lo = 0; hi = n - 1; lk = LK_FIRST_COST; rk = RK_LAST;
    return;
  }
  if (cs == kLinCase2) {
    if (i < h) { lo = 0; hi = w + i - 1; lk = LK_FIRST_TAIL; rk = RK_RIGHT; }
    else if (i < m - (h - 1)) { lo = 1 + (i - h); hi = lo + w + h - 2; lk = LK_LEFT; rk = RK_RIGHT; }
    else { lo = n - (w + h - 2) + (i - (m - (h - 1))); hi = n - 1; lk = LK_LEFT; rk = RK_LAST; }
    return;
  }
  if (i < n - w + 1) { lo = 0; hi = w + i - 1; lk = LK_FIRST_TAIL; rk = RK_RIGHT; }
  else if (i < m - (n - w) + 1) { lo = 0; hi = n - 1; lk = LK_FIRST_TAIL; rk = RK_LAST; }
  else { lo = 1 + (i - (m - (n - w) + 1)); hi = n - 1; lk = LK_LEFT; rk = RK_LAST; }
}

__device__ __forceinline__ int lin_row0_hi(int n, int cs, int w) { return cs == kLinFull ? n - 1 : w - 1; }

// Per-pair prefix sums over s2: GI[j] = sum cost(gap, s2[k]), k=1..j (the
// INSERT chain, precalc row `gap`), GP[j] = sum prepend[s2[k]] (row 0).
__device__ void lin_prefix(const uint8_t *s2, int n, const int *s_cost, const int *s_prep, int *GI,
                           int *GP) {
  const int lane = lane_id();
  int ci = 0, cp = 0;
  for (int j0 = 0; j0 < n; j0 += kWave) {
    const int j = j0 + lane;
    int gi = 0, gp = 0;
    if (j >= 1 && j < n) {
      const int x = s2[j] & 31;
      gi = s_cost[kGap * 32 + x];
      gp = s_prep[x];
    }
    const int a = wave_prefix_sum(gi) + ci;
    const int b = wave_prefix_sum(gp) + cp;
    if (j < n) { GI[j] = a; GP[j] = b; }
    ci = lane63(a);
    cp = lane63(b);
  }
}

__global__ __launch_bounds__(kWave * 4) void lin_prep_kernel(const PairMeta *__restrict__ meta, int n,
                                                           const uint8_t *__restrict__ seq,
                                                           const DevTcm *__restrict__ tcm, int *aux) {
  __shared__ int s_cost[32 * 32];
  __shared__ int s_prep[32];
  for (int k = threadIdx.x; k < 32 * 32; k += blockDim.x) s_cost[k] = tcm->cost[k];
  if (threadIdx.x < 32) s_prep[threadIdx.x] = tcm->prepend[threadIdx.x];
  __syncthreads();
  const int waves = gridDim.x * (blockDim.x / kWave);
  for (int p = blockIdx.x * (blockDim.x / kWave) + wave_in_group(); p < n; p += waves) {
    const PairMeta m = uniform_meta(meta, p);
    int *GI = aux + m.aux_off;
    lin_prefix(seq + m.off_c, m.lc, s_cost, s_prep, GI, GI + m.lc);
  }
}

// kStaged: the columns' insert-chain prefix sums and s2 symbols are copied
// to LDS (stage_gi, stage_x; n2 <= kLinStage) once per pair, so that a row
// reads them from LDS rather than from global memory.
template <bool kDm, bool kGlobalRows, bool kStaged = false>
__device__ __forceinline__ void lin_fill_pair(const PairMeta &m, const uint8_t *__restrict__ seq, const int *s_cost,
                                              const int *s_prep, const int *s_tail, uint8_t *__restrict__ dm_base,
                                              int *aux, int32_t *__restrict__ cost_out, int pidx, int *rows,
                                              int mask, int bstride, int *stage_gi = nullptr,
                                              uint8_t *stage_x = nullptr) {
  const int lane = lane_id();
  // the pair's geometry is wave-uniform (scalar registers, scalar branches)
  const uint8_t *s1 = seq + uniform64(m.off_r);   // the longer sequence (rows)
  const uint8_t *s2 = seq + uniform64(m.off_c);
  const int M = uniform32(m.lr), N = uniform32(m.lc), cs = uniform32(m.p0), w = uniform32(m.p1),
            h = uniform32(m.p2), S = uniform32(m.stride);
  pidx = uniform32(pidx);
  int *GI = aux + uniform64(m.aux_off);
  int *GP = GI + N;
  uint8_t *dm = dm_base + uniform64(m.dm_off);
  const int *gi_rd = GI;
  const uint8_t *x_rd = s2;
  if (kStaged) {
    for (int j = lane; j < N; j += kWave) {
      stage_gi[j] = GI[j];
      stage_x[j] = s2[j];
    }
    wave_sync_mem();
    gi_rd = stage_gi;
    x_rd = stage_x;
  }
  // row 0 (algn_fill_first_row, :584-607, or the first row of algn_fill_plane)
  const int hi0 = lin_row0_hi(N, cs, w);
  for (int j0 = 0; j0 <= hi0; j0 += kWave) {
    const int j = j0 + lane;
    if (j <= hi0) {
      rows[j & mask] = GP[j];
      if (kDm) dm[j] = j == 0 ? 1 : 2;   // ALIGN, then INSERT
    }
  }
  if (M == 1) {
    // every lane stores the same value: an early exit must not sit inside
    // lane-divergent code (see uniform32)
    cost_out[pidx] = GP[N - 1];
    return;
  }
  if (kGlobalRows) wave_fence_global();
  else wave_sync_mem();
  // the rows' symbols are loaded 64 at a time (lane k holds s1[i + k]) and
  // read with readlane: a load consumed on every row would wait for the
  // previous row's direction stores as well (one in-order counter)
  int sblk = 0;
  for (int i = 1; i < M; ++i) {
    int lo, hi, lk, rk;
    lin_row(i, M, N, cs, w, h, lo, hi, lk, rk);
    lo = uniform32(lo);
    hi = uniform32(hi);
    lk = uniform32(lk);
    rk = uniform32(rk);
    const int *prv = rows + ((i - 1) & 1) * bstride;
    int *cur = rows + (i & 1) * bstride;
    if (((i - 1) & (kWave - 1)) == 0) {
      const int ii = i + lane;
      sblk = s1[min(ii, M - 1)];
      __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0) here, once per 64 rows, not on every row
    }
    const int si = __builtin_amdgcn_readlane(sblk, (i - 1) & (kWave - 1)) & 31;
    const int c_del = s_cost[si * 32 + kGap];
    const int tail = s_tail[si];
    const int *crow = s_cost + si * 32;
    uint8_t *drow = dm + (size_t)i * S - lo;
    int carry_mm = kBig, carry_gi = 0;   // chain value before lo: none
    for (int j0 = lo; j0 <= hi; j0 += kWave) {
      // Every lane loads and computes (clamped column jj); only the stores
      // are conditional.  Keeping the body free of divergent branches keeps
      // the per-lane state of inactive lanes out of reach of the branches.
      const int j = j0 + lane;
      const bool act = j <= hi;
      const int jj = min(j, hi);
      const bool first = (j == lo) && lk != LK_LEFT;
      const bool has_ins = act && !(j == lo);
      const bool right = j == hi && rk == RK_RIGHT;
      const bool last = j == hi && rk == RK_LAST && j > 0;
      const int GIraw = gi_rd[jj];
      const int GIj = act ? GIraw : 0;
      const int x = x_rd[jj] & 31;
      const int gj = s_cost[kGap * 32 + x];
      const int p_left = prv[max(jj - 1, 0) & mask];   // unused (first) when jj == lo
      const int p_up = prv[jj & mask];
      const int p_first = prv[0 & mask];
      const int t3 = first ? kBig : p_left + crow[x];
      const int t1 = (first || right) ? kBig : p_up + c_del;
      const int cst = (!first && last) ? tail + p_up : kBig;
      int T = first ? p_first + (lk == LK_FIRST_TAIL ? tail : c_del) : min(min(t1, t3), cst);
      T = act ? T : kBig;
      int pm = wave_prefix_min(act ? T - GIj : kBig);
      pm = min(pm, carry_mm - carry_gi);
      const int mm = GIj + pm;
      const int mm_left = shfl_up1(mm, carry_mm);
      const int t2 = has_ins ? mm_left + gj : kBig;
      int best = min(min(t1, t2), t3);
      int msk = (t3 == best ? 1 : 0) | (has_ins && t2 == best ? 2 : 0) | (t1 == best ? 4 : 0);
      const bool cst_lt = cst < best, cst_eq = cst == best && cst != kBig;
      best = cst_lt ? cst : best;
      msk = cst_lt ? 4 : (cst_eq ? (msk | 4) : msk);
      best = first ? T : best;
      msk = first ? 4 : msk;
      if (act) {
        cur[j & mask] = best;
        if (kDm) drow[j] = (uint8_t)msk;
        if (i == M - 1 && j == N - 1) cost_out[pidx] = best;
      }
      carry_mm = lane63(mm);
      carry_gi = lane63(GIj);
    }
    if (kGlobalRows) wave_fence_global();
    else wave_sync_mem();
  }
}

// The row formulation with R consecutive cells per lane: cell j = lo + L*R + r
// of lane L (bands of up to 64R columns), so a row is one pass with one
// prefix minimum for the insert chain instead of one per 64 columns.  Rows
// are LDS rings of kLinRB columns; the columns' prefix sums and symbols are
// staged in LDS.  Same cell rules as lin_fill_pair (src/algn.c:375-620).
template <bool kDm, int R>
__device__ __forceinline__ void lin_fill_pair_r(const PairMeta &m, const uint8_t *__restrict__ seq,
                                                const int *s_cost, const int *s_prep, const int *s_tail,
                                                uint8_t *__restrict__ dm_base, int *aux,
                                                int32_t *__restrict__ cost_out, int pidx, int *rows,
                                                int *stage_gi, uint8_t *stage_x) {
  static_assert(R % 4 == 0 && 64 * R <= kLinRB, "a row pass covers at most one ring");
  constexpr int mask = kLinRB - 1;
  const int lane = lane_id();
  const uint8_t *s1 = seq + uniform64(m.off_r);
  const uint8_t *s2 = seq + uniform64(m.off_c);
  const int M = uniform32(m.lr), N = uniform32(m.lc), cs = uniform32(m.p0), w = uniform32(m.p1),
            h = uniform32(m.p2), S = uniform32(m.stride);
  pidx = uniform32(pidx);
  int *GI = aux + uniform64(m.aux_off);
  int *GP = GI + N;
  uint8_t *dm = dm_base + uniform64(m.dm_off);
  for (int j = lane; j < N; j += kWave) {
    stage_gi[lin_pad(j)] = GI[j];
    stage_x[j] = s2[j];
  }
  const int hi0 = lin_row0_hi(N, cs, w);
  for (int j0 = 0; j0 <= hi0; j0 += kWave) {
    const int j = j0 + lane;
    if (j <= hi0) {
      rows[lin_pad(j & mask)] = GP[j];
      if (kDm) dm[j] = j == 0 ? 1 : 2;   // ALIGN, then INSERT
    }
  }
  if (M == 1) {
    cost_out[pidx] = GP[N - 1];
    return;
  }
  wave_sync_mem();
  int sblk = 0;
  for (int i = 1; i < M; ++i) {
    int lo, hi, lk, rk;
    lin_row(i, M, N, cs, w, h, lo, hi, lk, rk);
    lo = uniform32(lo);
    hi = uniform32(hi);
    lk = uniform32(lk);
    rk = uniform32(rk);
    const int *prv = rows + ((i - 1) & 1) * kLinRBP;
    int *cur = rows + (i & 1) * kLinRBP;
    if (((i - 1) & (kWave - 1)) == 0) {
      sblk = s1[min(i + lane, M - 1)];
      __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0) here, once per 64 rows
    }
    const int si = __builtin_amdgcn_readlane(sblk, (i - 1) & (kWave - 1)) & 31;
    const int c_del = s_cost[si * 32 + kGap];
    const int tail = s_tail[si];
    const int *crow = s_cost + si * 32;
    const int p_first = prv[0];
    const int j0 = lo + lane * R;
    // per cell: T (DELETE / ALIGN / last-column tail, or the first cell's
    // value), t1, t3, cst and the chain key T - GI
    int T[R], t1v[R], t3v[R], cstv[R], gjv[R], giv[R], wk[R];
    int up_prev = prv[lin_pad(max(j0 - 1, 0) & mask)];
    int run = kBig;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int j = j0 + r;
      const bool act = j <= hi;
      const int jj = min(j, hi);
      const bool first = (j == lo) && lk != LK_LEFT;
      const bool right = j == hi && rk == RK_RIGHT;
      const bool last = j == hi && rk == RK_LAST && j > 0;
      const int GIj = act ? stage_gi[lin_pad(jj)] : 0;
      const int x = stage_x[jj] & 31;
      gjv[r] = s_cost[kGap * 32 + x];
      const int p_up = prv[lin_pad(jj & mask)];
      const int p_left = up_prev;
      up_prev = p_up;
      const int t3 = first ? kBig : p_left + crow[x];
      const int t1 = (first || right) ? kBig : p_up + c_del;
      const int cst = (!first && last) ? tail + p_up : kBig;
      int Tv = first ? p_first + (lk == LK_FIRST_TAIL ? tail : c_del) : min(min(t1, t3), cst);
      Tv = act ? Tv : kBig;
      T[r] = Tv;
      t1v[r] = t1;
      t3v[r] = t3;
      cstv[r] = cst;
      giv[r] = GIj;
      wk[r] = act ? Tv - GIj : kBig;
      run = min(run, wk[r]);
    }
    // insert chain: mm[j] = GI[j] + min over k <= j of (T[k] - GI[k])
    const int incl = wave_prefix_min(run);
    const int excl = shfl_up1(incl, kBig);
    int pm = excl;
    int mm[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      pm = min(pm, wk[r]);
      mm[r] = giv[r] + pm;
    }
    const int mm_in = shfl_up1(mm[R - 1], kBig);   // cell j0 - 1
    uint32_t codes[R / 4];
#pragma unroll
    for (int q = 0; q < R / 4; ++q) codes[q] = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int j = j0 + r;
      const bool act = j <= hi;
      const bool first = (j == lo) && lk != LK_LEFT;
      const bool has_ins = act && j != lo;
      const int mm_left = r ? mm[r - 1] : mm_in;
      const int t1 = t1v[r], t3 = t3v[r], cst = cstv[r];
      const int t2 = has_ins ? mm_left + gjv[r] : kBig;
      int best = min(min(t1, t2), t3);
      int msk = (t3 == best ? 1 : 0) | (has_ins && t2 == best ? 2 : 0) | (t1 == best ? 4 : 0);
      const bool cst_lt = cst < best, cst_eq = cst == best && cst != kBig;
      best = cst_lt ? cst : best;
      msk = cst_lt ? 4 : (cst_eq ? (msk | 4) : msk);
      best = first ? T[r] : best;
      msk = first ? 4 : msk;
      codes[r >> 2] |= (uint32_t)(msk & 0xff) << (8 * (r & 3));
      if (act) {
        cur[lin_pad(j & mask)] = best;
        if (i == M - 1 && j == N - 1) cost_out[pidx] = best;
      }
    }
    if (kDm) {
      // whole dwords of the row's stride (S covers hi - lo + 1 rounded to 4)
      uint32_t *drow = (uint32_t *)(dm + (size_t)i * S);
#pragma unroll
      for (int q = 0; q < R / 4; ++q)
        if (lane * R + 4 * q <= hi - lo) drow[(lane * R) / 4 + q] = codes[q];
    }
    wave_sync_mem();
  }
}

// Systolic form of the banded linear fill, for s2 of at most kLinStage
// columns.  The rows are cut into strips of 64: lane L of a strip holds row
// i0 + L and reaches column j = c0 + t - L at step t (c0 = the strip's first
// column), so the cells of one step lie on an anti-diagonal and depend only
// on the two steps before it: (i-1, j) is lane L-1's result of step t-1 (one
// DPP shift), (i-1, j-1) what this lane received from it a step earlier,
// (i, j-1) the lane's own previous result.  Lane 0 reads row i0-1 from the
// LDS row that lane 63 of the previous strip wrote.  Each lane's band
// (lin_row) decides where it holds cells; the band's lower and upper edges
// never decrease from row to row, so every lane's band lies inside the
// strip's steps.  Out-of-band results are never read (a cell reads (i-1, j)
// and (i-1, j-1) only where they lie in row i-1's band).  The step loop
// changes no EXEC mask (its branches are uniform): stores that only some
// lanes make go to per-lane LDS slots or through an out-of-range buffer
// offset, which the hardware drops.
// Same cell rules as lin_fill_pair (src/algn.c:375-620).
//
// The four waves of a workgroup share one pair: wave k takes strips k,
// k+4, ...  A strip runs behind its predecessor: every kSysSync steps it
// waits until the predecessor's progress counter (strip * kSysStripTag +
// last column its lane 63 has written) covers the columns its lane 0 reads
// next, and publishes its own.  Strip s writes hand-off row s % 4; strip
// s+4 overwrites it only after strips s+1 .. s+3 have each run 63 columns
// ahead of their successors, by which time strip s+1 has read it.
constexpr int kSysRow = kWave + 1 + kLinStage;   // 64 dummy slots, column -1, columns 0..kLinStage-1
constexpr int kSysWaves = 4;
constexpr int kSysDrop = 0x7ffffff0;             // buffer offset past any pair's direction band
constexpr int kSysSync = 8;                      // steps between hand-off checks
constexpr int kSysStripTag = 1 << 16;            // progress = strip * kSysStripTag + column
constexpr int kSysSpin = 1 << 22;                // polls before a hand-off wait gives up (default)

struct SysLds {
  int cost[32 * 32];
  int prep[32];
  int tail[32];
  int row[kSysWaves][kSysRow];   // hand-off rows; row[kSysWaves - 1] holds row 0 first
  int prog[kSysWaves + kWave];   // progress per hand-off row, then one dummy slot per lane
  uint8_t x[kWave + kLinStage + kWave + 2];   // s2, masked to 5 bits, at kWave; zero pads around it
  int item;
};

template <bool kDm>
__device__ __forceinline__ bool lin_sys_strip(int s, int nstrips, int M, int N, int cs, int w, int h, int S,
                                              const uint8_t *__restrict__ s1, SysLds &L,
                                              __amdgpu_buffer_rsrc_t drsrc, int32_t *__restrict__ cost_out,
                                              int pidx, int spin_limit) {
  const int lane = lane_id();
  const int i0 = 1 + kWave * s;
  const int *prow = &L.row[(s + kSysWaves - 1) % kSysWaves][kWave + 1];   // prow[-1] is a slot
  int *mine = &L.row[s % kSysWaves][0];   // mine[kWave + 1 + j] = column j, mine[lane] = dummy
  const int pslot = (s + kSysWaves - 1) % kSysWaves, mslot = s % kSysWaves;
  const uint8_t *sx = L.x + kWave;   // columns -64 .. N + 65 are readable
  const int *gcost = L.cost + kGap * 32;
  const int i = min(i0 + lane, M - 1);
  const bool valid = i0 + lane < M;
  int lo, hi, lk, rk;
  lin_row(i, M, N, cs, w, h, lo, hi, lk, rk);
  lo = valid ? lo : 0x3fffffff;   // rows past the last one hold no cell
  hi = valid ? hi : -1;
  const int wid = valid ? hi - lo : 0;   // a cell where 0 <= j - lo <= wid
  const int c0 = uniform32(lo);
  const int last_lane = min(kWave - 1, M - 1 - i0);
  const int t_end = __builtin_amdgcn_readlane(hi, last_lane) - c0 + last_lane;
  // the predecessor's last row (row 0 for the first strip) ends at hi_prev
  int plo, hi_prev, plk, prk;
  if (i0 - 1 == 0) hi_prev = lin_row0_hi(N, cs, w);
  else lin_row(i0 - 1, M, N, cs, w, h, plo, hi_prev, plk, prk);
  hi_prev = uniform32(hi_prev);
  bool ok = true;
  // waits until the predecessor's last row holds every column up to col
  auto wait_for = [&](int col) {
    if (s == 0 || !ok) return;
    const int want = (s - 1) * kSysStripTag + min(col, hi_prev);
    for (int k = 0;; ++k) {
      const int v = uniform32(__hip_atomic_load(&L.prog[pslot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
      if (v >= want) break;
      if (k >= spin_limit) {
        ok = false;   // not reached with the default limit: every strip runs to its end
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
  };
  // lane 0 publishes (the other lanes store to their dummy slots), after
  // this wave's earlier LDS writes have completed
  auto publish = [&](int col) {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0)
    __hip_atomic_store(&L.prog[lane == 0 ? mslot : kSysWaves + lane], s * kSysStripTag + col, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_WORKGROUP);
  };
  const int si = s1[i] & 31;
  const int c_del = L.cost[si * 32 + kGap];
  const int tail = L.tail[si];
  const int *crow = L.cost + si * 32;
  const int first_add = lk == LK_FIRST_TAIL ? tail : c_del;
  const int lo_first = lk != LK_LEFT ? lo : -2;   // column of the first-column cell, if any
  const int hi_right = rk == RK_RIGHT ? hi : -2;  // column without DELETE
  const int hi_last = rk == RK_LAST && hi > 0 ? hi : -2;   // column re-checked with tail
  // row i's bytes start at phase (lane + lo - c0) mod 4 of its stride, so that
  // byte (j - lo) of every lane lies at t mod 4 in its dword: the lanes fill
  // their dwords in step and store them together every fourth step
  const int drow = i * S + ((lane + lo - c0) & 3);
  wait_for(c0 + kSysSync);
  int res = kBig, acc = 0, fin = 0;
  int up_prev = prow[c0 - 1];
  int j = c0 - lane;
  // look-ahead: symbols two steps ahead, cost-table reads one step ahead
  int x1 = sx[j + 1];
  const int x0 = sx[j];
  int cr = crow[x0], gj = gcost[x0];
  int fill = prow[c0];
  dpp_exec_guard();
  for (int t = 0; t <= t_end; ++t, ++j) {
    if (t != 0 && (t & (kSysSync - 1)) == 0) {
      publish(c0 + t - kWave);   // lane 63 wrote columns up to c0 + t - 64
      wait_for(c0 + t + kSysSync);
    }
    const int x2 = sx[j + 2];   // lanes outside their band read the pads
    const int cr_n = crow[x1], gj_n = gcost[x1];
    const int fill_n = prow[min(c0 + t + 1, N - 1)];
    const int up = __builtin_amdgcn_update_dpp(fill, res, 0x138 /* wave_shr:1 */, 0xf, 0xf, false);
    const int diag = up_prev;
    const int t3 = diag + cr;
    const int t1 = j == hi_right ? kBig : up + c_del;
    const int b = j - lo;
    const bool has_ins = b != 0;
    const int t2 = has_ins ? res + gj : kBig;
    int best = min(min(t1, t2), t3);
    int msk = (int)(t3 == best) | ((int)(has_ins && t2 == best) << 1) | ((int)(t1 == best) << 2);
    // last column: DELETE with tail replaces a larger minimum and joins an
    // equal one (bit arithmetic, so that no branch splits the step)
    const bool is_last = j == hi_last;
    const int cst = tail + up;
    const int lt = (int)(is_last && cst < best), le = (int)(is_last && cst <= best);
    best = min(best, is_last ? cst : kBig);
    msk = (msk & (lt - 1)) | (le << 2);
    const bool first = j == lo_first;
    best = first ? up + first_add : best;
    msk = first ? 4 : msk;
    const bool act = (unsigned)b <= (unsigned)wid;
    res = act ? best : kBig;
    fin = act && j == N - 1 ? best : fin;
    // lane 63 hands its row to the next strip
    mine[act && lane == kWave - 1 ? kWave + 1 + j : lane] = best;
    if (kDm) {
      acc |= act ? msk << (8 * (t & 3)) : 0;   // every cell's code is non-zero
      if ((t & 3) == 3 || t == t_end) {          // uniform
        // the dword holding this step's byte position; lanes without a cell
        // in it store nothing
        __builtin_amdgcn_raw_buffer_store_b32((uint32_t)acc, drsrc,
                                              acc ? ((drow + b) & ~3) : kSysDrop, 0, 0);
        acc = 0;
      }
    }
    up_prev = up;
    fill = fill_n;
    cr = cr_n;
    gj = gj_n;
    x1 = x2;
  }
  publish(kSysStripTag - 1);   // every column
  if (s == nstrips - 1 && lane == M - 1 - i0) cost_out[pidx] = fin;
  return ok;
}

// One pair per workgroup (kSysWaves waves), pairs from the work list.
template <bool kDm>
__global__ __launch_bounds__(kWave * kSysWaves) void lin_sys_kernel(
    const PairMeta *__restrict__ meta, const int32_t *__restrict__ order, int n, int *counter,
    const uint8_t *__restrict__ seq, const DevTcm *__restrict__ tcm, uint8_t *dm_base, const int *aux,
    int32_t *cost_out, int32_t *status, int spin_limit) {
  __shared__ SysLds L;
  load_tables(tcm, L.cost, L.prep, nullptr);
  for (int k = threadIdx.x; k < 32; k += blockDim.x) L.tail[k] = tcm->tail[k];
  const int wv = wave_in_group();
  for (;;) {
    __syncthreads();   // the previous pair is done in every wave
    if (threadIdx.x == 0) L.item = atomicAdd(counter, 1);
    __syncthreads();
    const int k = uniform32(L.item);
    if (k >= n) break;
    const int p = uniform32(order[k]);
    const PairMeta m = uniform_meta(meta, p);
    const uint8_t *s1 = seq + m.off_r;   // the longer sequence (rows)
    const uint8_t *s2 = seq + m.off_c;
    const int M = m.lr, N = m.lc, cs = m.p0, w = m.p1, S = m.stride;
    const int *GP = aux + m.aux_off + N;
    uint8_t *dm = dm_base + m.dm_off;
    // stage s2 and row 0 (algn_fill_first_row, :584-607, or the first row of
    // algn_fill_plane), reset the progress counters
    const int hi0 = lin_row0_hi(N, cs, w);
    for (int j = threadIdx.x; j < N; j += blockDim.x) L.x[kWave + j] = s2[j] & 31;
    for (int j = threadIdx.x; j < 2 * kWave + 2; j += blockDim.x)
      L.x[j < kWave ? j : N + j] = 0;   // pads: columns -64..-1 and N..N+65
    for (int j = threadIdx.x; j <= hi0; j += blockDim.x) {
      L.row[kSysWaves - 1][kWave + 1 + j] = GP[j];
      if (kDm) dm[j] = j == 0 ? 1 : 2;   // ALIGN, then INSERT
    }
    if (threadIdx.x < kSysWaves) L.prog[threadIdx.x] = -1;
    __syncthreads();
    if (M == 1) {
      if (threadIdx.x == 0) cost_out[p] = GP[N - 1];
      continue;
    }
    const __amdgpu_buffer_rsrc_t drsrc = __builtin_amdgcn_make_buffer_rsrc(dm, 0, kDm ? M * S : 0, 0x00020000);
    const int nstrips = (M - 1 + kWave - 1) / kWave;
    bool ok = true;
    for (int s = wv; s < nstrips; s += kSysWaves)
      ok = lin_sys_strip<kDm>(s, nstrips, M, N, cs, w, m.p2, S, s1, L, drsrc, cost_out, p, spin_limit) && ok;
    if (!ok) mark_failed(status, p);
  }
}

// ---------------------------------------------------------------------------
// Linear band, 16-bit (algn_fill_plane_2, src/algn.c:872-968: the cost of
// Sequence.Align.cost_2, and with kDm the direction band that
// backtrack_2d reads)
// ---------------------------------------------------------------------------
// One wave per pair, two rows per lane in the two 16-bit halves of a
// register, so that every packed add / min / saturation handles two cells.
// A strip of 128 rows runs as an anti-diagonal wavefront: at step t lane L
// holds row i0 + 2L at column c = c0 + t - 2L (low half) and row i0 + 2L + 1
// at column c - 1 (high half).  A cell's neighbours are then all of the
// previous two steps:
//   up    low: lane L-1's high half (DPP wave_shr:1; lane 0 reads the
//              previous strip's last row from LDS); high: own low half
//   left  both halves of the own previous value
//   diag  the previous step's `up`
// Cells outside a row's band hold 0xffff, and every candidate is a
// saturating add, so an out-of-band neighbour never wins: that stands in
// for the reference's explicit band-edge cases (no INSERT at a row's first
// cell, no DELETE at its last cell of a right-open row, the first-column
// cells).  The last column's tail rule and the first column's tail rule
// reduce to the plain DELETE when tail[x] == cost[x][gap] for every x: the
// plan sends a pair here only then, only when every cost of the pair stays
// below 0xffff, and only when its shorter sequence has 2 .. kLinStage
// positions.  The lane (and half) holding a strip's last row stores it in
// LDS for the next strip; the pair's cost is that row's entry at column
// n - 1 after the last strip.
//
// Direction codes (kDm): a candidate's bit (ALIGN 1, INSERT 2, DELETE 4) is
// set when it equals the cell's minimum -- with the 0xffff neighbours the
// excluded moves never do, so the codes equal the reference's, the first
// column's forced DELETE and the last column's tail rule included (tail ==
// deletion cost).  Row r's bytes start at phase (r - i0 + lo_r - lo_i0)
// mod 4 of its stride (i0 the first row of r's 128-row strip), so that
// byte (c - lo_r) lies at step t mod 4 of its dword: every lane fills two
// dwords (its two rows) in step and stores them every fourth step.
// PairMeta::p3 = 1 tells the backtrace this phase.
constexpr int kC16Waves = 4;    // waves per workgroup, one pair each
constexpr int kC16Pad = 192;    // LDS columns kept before column 0 and after the last one
constexpr int kC16U = 8;        // steps per unrolled group
constexpr int kC16Rows = 2 * kWave;

// LDS words per wave for pairs whose shorter sequence has at most n_max
// positions: column words, the hand-off row, 64 lanes x kC16U dummy slots.
__host__ __device__ __forceinline__ int c16_cols(int n_max) { return (n_max + 2 * kC16Pad + 63) & ~63; }
__host__ __device__ __forceinline__ int c16_wave_words(int n_max) { return 2 * c16_cols(n_max) + kWave * kC16U; }

// One 128-row strip of the packed 16-bit fill on one wave: the state a
// strip carries from step to step.  begin() sets up strip i0's rows, reading
// the hand-off row of the strip above from Rin and writing this strip's last
// row into Rout (the same buffer when one wave runs the strips in turn: the
// writes trail the reads); run16() advances it 16 steps.
template <bool kDm>
struct C16Strip {
  const uint16_t *ta, *tb;
  uint32_t cdel, kv, wid, V, upp, wprev, acc_a, acc_b;
  uint32_t *wp;
  const uint32_t *wc, *rr;
  int winc, oa, ob, steps, t, wh;
  __amdgpu_buffer_rsrc_t drsrc;
  uint32_t wA[kC16U], fA[kC16U], wB[kC16U], fB[kC16U];

  __device__ __forceinline__ void load_wf(uint32_t (&w)[kC16U], uint32_t (&f)[kC16U], int off) {
#pragma unroll
    for (int u = 0; u < kC16U; ++u) {
      w[u] = wc[off + u];
      f[u] = rr[off + u];
    }
  }

  __device__ __forceinline__ void begin(int i0, int M, int N, const LinGeom &g, const uint8_t *s1, const uint16_t *T,
                                        const uint32_t *Wc, const uint32_t *Rin, uint32_t *Rout, uint32_t *D, int S,
                                        __amdgpu_buffer_rsrc_t rsrc) {
    const int lane = lane_id();
    drsrc = rsrc;
    int c0, hi_first, lo_l, hi_l;
    lin_row_host(i0, M, N, g, c0, hi_first);
    const int r_last = min(i0 + kC16Rows - 1, M - 1);
    lin_row_host(r_last, M, N, g, lo_l, hi_l);
    const int wl = (r_last - i0) >> 1;   // lane and half of the last row
    wh = (r_last - i0) & 1;
    const int t_end = hi_l - c0 + 2 * wl + wh;
    // one step more (the cell right of the band), rounded to two groups
    steps = (t_end + 2 + 2 * kC16U - 1) / (2 * kC16U) * (2 * kC16U);
    t = 0;
    // this lane's rows
    const int ra = i0 + 2 * lane, rb = ra + 1;
    const bool va = ra < M, vb = rb < M;
    int lo_a, hi_a, lo_b, hi_b;
    lin_row_host(va ? ra : M - 1, M, N, g, lo_a, hi_a);
    lin_row_host(vb ? rb : M - 1, M, N, g, lo_b, hi_b);
    const int sa = va ? (s1[ra] & 31) : 0, sb = vb ? (s1[rb] & 31) : 0;
    cdel = (uint32_t)T[sa * 32 + kGap] | ((uint32_t)T[sb * 32 + kGap] << 16);
    ta = T + sa * 32;
    tb = T + sb * 32;
    // column - lo of the step's cell (rows past the last one: never inside)
    kv = (va ? (uint32_t)(c0 - 2 * lane - lo_a) & 0xffffu : 0x8000u) |
         ((vb ? (uint32_t)(c0 - 2 * lane - 1 - lo_b) & 0xffffu : 0x8000u) << 16);
    wid = (va ? (uint32_t)(hi_a - lo_a) : 0u) | ((vb ? (uint32_t)(hi_b - lo_b) : 0u) << 16);
    const bool writer = lane == wl;
    wp = writer ? Rout + (c0 - 2 * wl - wh) : D;
    winc = writer ? kC16U : 0;
    wc = Wc + (c0 - 2 * lane);
    rr = Rin + c0;
    // byte offsets of the dwords that steps 0..3 fill (rows past the last
    // one store nothing: no byte of theirs is ever inside)
    oa = ob = 0;
    if (kDm) {
      oa = ra * S + ((lo_a - c0 + 2 * lane) & 3) + (c0 - 2 * lane - lo_a);
      ob = rb * S + ((lo_b - c0 + 2 * lane + 1) & 3) + (c0 - 2 * lane - 1 - lo_b);
    }
    acc_a = acc_b = 0;
    V = 0xffffffffu;
    upp = (Rin[c0 - 1] >> 16) | 0xffff0000u;   // lane 0's first diagonal; the rest lie outside
    wprev = wc[-1];
    // A group of kC16U steps reads its substitution costs at its start;
    // column words and lane 0's hand-off words are read one group ahead,
    // into two register sets that alternate (no copies between groups).
    load_wf(wA, fA, 0);
  }

  __device__ __forceinline__ void group(const uint32_t (&w)[kC16U], const uint32_t (&f)[kC16U]) {
    uint32_t cr[kC16U];
#pragma unroll
    for (int u = 0; u < kC16U; ++u) {
      const uint32_t wq = u ? w[u - 1] : wprev;
      cr[u] = (uint32_t)*(const uint16_t *)((const char *)ta + (w[u] >> 16)) |
              ((uint32_t)*(const uint16_t *)((const char *)tb + (wq >> 16)) << 16);
    }
#pragma unroll
    for (int u = 0; u < kC16U; ++u) {
      const uint32_t wq = u ? w[u - 1] : wprev;
      const uint32_t sh =
          (uint32_t)__builtin_amdgcn_update_dpp((int)f[u], (int)V, 0x138 /* wave_shr:1 */, 0xf, 0xf, false);
      const uint32_t up = __builtin_amdgcn_alignbit(V, sh, 16);
      const uint32_t gj = __builtin_amdgcn_perm(wq, w[u], 0x05040100u);
      const uint32_t t1 = c16_add_sat(up, cdel);    // DELETE
      const uint32_t t2 = c16_add_sat(V, gj);       // INSERT
      const uint32_t t3 = c16_add_sat(upp, cr[u]);  // ALIGN
      const uint32_t best = c16_min(c16_min(t1, t2), t3);
      const uint32_t gt = c16_sub_sat(kv, wid);
      V = c16_gate(gt, best);
      kv = c16_add(kv, 0x00010001u);
      wp[u] = V;
      upp = up;
      if (kDm) {
        // per half: 1 for a candidate equal to the minimum inside the band
        const uint32_t ib = c16_sub_sat(0x00010001u, gt);
        const uint32_t q1 = c16_sub_sat(ib, c16_sub_sat(t1, best));
        const uint32_t q2 = c16_sub_sat(ib, c16_sub_sat(t2, best));
        const uint32_t q3 = c16_sub_sat(ib, c16_sub_sat(t3, best));
        const uint32_t msk = c16_add(c16_add(q3, c16_add(q2, q2)), c16u(c16v(q1) << (unsigned short)2));
        // byte u mod 4 of the two dwords (the low half's code is byte 0
        // of msk, the high half's byte 2); position 0 starts a dword
        constexpr uint32_t sel_a[4] = {0x0c0c0c00u, 0x0c0c0004u, 0x0c000504u, 0x00060504u};
        constexpr uint32_t sel_b[4] = {0x0c0c0c02u, 0x0c0c0204u, 0x0c020504u, 0x02060504u};
        acc_a = __builtin_amdgcn_perm(acc_a, msk, sel_a[u & 3]);
        acc_b = __builtin_amdgcn_perm(acc_b, msk, sel_b[u & 3]);
        if ((u & 3) == 3) {
          __builtin_amdgcn_raw_buffer_store_b32(acc_a, drsrc, acc_a ? oa : kSysDrop, 0, 0);
          __builtin_amdgcn_raw_buffer_store_b32(acc_b, drsrc, acc_b ? ob : kSysDrop, 0, 0);
          oa += 4;
          ob += 4;
        }
      }
    }
    wprev = w[kC16U - 1];
    wp += winc;
  }

  __device__ __forceinline__ void run16() {
    load_wf(wB, fB, kC16U);
    group(wA, fA);
    load_wf(wA, fA, 2 * kC16U);
    group(wB, fB);
    wc += 2 * kC16U;
    rr += 2 * kC16U;
    t += 2 * kC16U;
  }
};

// Column words of the pair (gap cost of s2[c] | 2 s2[c] << 16, 0xffff past
// the ends) into Wc and row 0 into the hand-off row R0 (0xffff outside the
// band), by `count` threads from `first`.
__device__ __forceinline__ void c16_pair_rows(const uint8_t *s2, int N, const int *GP, int hi0, const uint16_t *T,
                                              uint32_t *Wc, uint32_t *R0, int first, int count) {
  for (int c = first - kC16Pad; c < N + kC16Pad; c += count) {
    const int x = (c >= 0 && c < N) ? (s2[c] & 31) : 0;
    Wc[c] = (uint32_t)T[kGap * 32 + x] | ((uint32_t)(2 * x) << 16);
    R0[c] = (c >= 0 && c <= hi0) ? ((uint32_t)GP[c] << 16) | 0xffffu : 0xffffffffu;
  }
}

template <bool kDm>
__global__ __launch_bounds__(kWave * kC16Waves) void lin_cost16_kernel(
    const PairMeta *__restrict__ meta, const int32_t *__restrict__ order, int n, int *counter,
    const uint8_t *__restrict__ seq, const DevTcm *__restrict__ tcm, const int *__restrict__ aux,
    int32_t *__restrict__ cost_out, uint8_t *dm_base, int n_max) {
  __shared__ uint16_t T[32 * 32];
  extern __shared__ uint32_t c16_dyn[];
  for (int k = threadIdx.x; k < 32 * 32; k += blockDim.x) T[k] = (uint16_t)min(tcm->cost[k], 0xffff);
  __syncthreads();
  const int lane = lane_id();
  const int cols = c16_cols(n_max);
  uint32_t *const base = c16_dyn + (size_t)wave_in_group() * c16_wave_words(n_max);
  uint32_t *const Wc = base + kC16Pad;          // column c: gcost[x] | 2x << 16 (x = s2[c] & 31)
  uint32_t *const R = base + cols + kC16Pad;    // hand-off row: value of column c in the high half
  uint32_t *const D = base + 2 * cols + lane * kC16U;   // this lane's dummy slots
  for (;;) {
    const int k = next_item(counter);
    if (k >= n) break;
    const int p = uniform32(order[k]);
    const PairMeta m = uniform_meta(meta, p);
    const uint8_t *s1 = seq + m.off_r;   // the longer sequence: rows
    const uint8_t *s2 = seq + m.off_c;
    const int M = m.lr, N = m.lc;
    const LinGeom g{m.p0, m.p1, m.p2};
    const int *GP = aux + m.aux_off + N;
    int lo0, hi0;
    lin_row_host(0, M, N, g, lo0, hi0);
    // column words and row 0 (algn_fill_first_row, :584-607, or the first
    // row of algn_fill_plane)
    c16_pair_rows(s2, N, GP, hi0, T, Wc, R, lane, kWave);
    uint8_t *dm = dm_base + m.dm_off;
    const int S = m.stride;
    if (kDm)
      for (int j = lane; j <= hi0; j += kWave) dm[j] = j == 0 ? 1 : 2;   // row 0: ALIGN, then INSERT
    const __amdgpu_buffer_rsrc_t drsrc = __builtin_amdgcn_make_buffer_rsrc(dm, 0, kDm ? M * S : 0, 0x00020000);
    wave_sync_mem();
    int last_half = 1;
    const int nstrips = (M - 1 + kC16Rows - 1) / kC16Rows;
    for (int s = 0; s < nstrips; ++s) {
      C16Strip<kDm> st;
      st.begin(1 + kC16Rows * s, M, N, g, s1, T, Wc, R, R, D, S, drsrc);
      dpp_exec_guard();
      while (st.t < st.steps) st.run16();
      last_half = st.wh;
      wave_sync_mem();
    }
    if (lane == 0) {
      const uint32_t e = R[N - 1];
      cost_out[p] = M == 1 ? GP[N - 1] : (int32_t)(last_half ? e >> 16 : e & 0xffffu);
    }
    wave_sync_mem();
  }
}

// The packed 16-bit fill with the four waves of a workgroup on one pair, for
// batches too small to fill the chip with one pair per wave (tree levels,
// per-call entries): wave w runs strips w, w + 4, ..., the workgroup
// advancing in chunks of 32 steps with a barrier between chunks.  Strip s
// reads the hand-off row that strip s - 1 writes into buffer (s - 1) mod 4;
// it starts enough chunks after strip s - 1 that every column it reads
// (lane 0's row, read up to 23 steps ahead) was written by strip s - 1's
// last row in an earlier chunk: strip s - 1 writes column c at its step
// c - c0(s - 1) + 127 at the latest, strip s reads it at its step
// c - c0(s) - 23 at the earliest, so the lag is c0(s) - c0(s - 1) + 150 steps
// plus one chunk, rounded up to chunks.  A wave takes its next strip when
// the previous one is done.  The schedule is the same arithmetic in every
// wave (no waits: only the barrier orders the buffers).
constexpr int kC16Chunk = 32;   // steps between the workgroup's barriers
constexpr int kC16MwStrips = 512;   // strips the multi-wave fill schedules (rows up to 65,537)
__host__ __device__ __forceinline__ int c16mw_words(int n_max) { return 5 * c16_cols(n_max) + kC16Waves * kWave * kC16U; }

template <bool kDm>
__global__ __launch_bounds__(kWave * kC16Waves) void lin_c16mw_kernel(
    const PairMeta *__restrict__ meta, const int32_t *__restrict__ order, int n, int *counter,
    const uint8_t *__restrict__ seq, const DevTcm *__restrict__ tcm, const int *__restrict__ aux,
    int32_t *__restrict__ cost_out, uint8_t *dm_base, int n_max) {
  __shared__ uint16_t T[32 * 32];
  __shared__ int s_start[kC16MwStrips];   // first chunk of each strip
  __shared__ int s_item, s_chunks;
  extern __shared__ uint32_t c16_dyn[];
  for (int k = threadIdx.x; k < 32 * 32; k += blockDim.x) T[k] = (uint16_t)min(tcm->cost[k], 0xffff);
  const int lane = lane_id(), wv = wave_in_group();
  const int cols = c16_cols(n_max);
  uint32_t *const Wc = c16_dyn + kC16Pad;
  uint32_t *const Rb = c16_dyn + cols + kC16Pad;   // hand-off buffer r at Rb + r * cols
  uint32_t *const D = c16_dyn + 5 * cols + (wv * kWave + lane) * kC16U;
  for (;;) {
    __syncthreads();   // the previous pair is done in every wave
    if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
    __syncthreads();
    const int k = uniform32(s_item);
    if (k >= n) break;
    const int p = uniform32(order[k]);
    const PairMeta m = uniform_meta(meta, p);
    const uint8_t *s1 = seq + m.off_r;
    const uint8_t *s2 = seq + m.off_c;
    const int M = m.lr, N = m.lc;
    const LinGeom g{m.p0, m.p1, m.p2};
    const int *GP = aux + m.aux_off + N;
    int lo0, hi0;
    lin_row_host(0, M, N, g, lo0, hi0);
    const int nstrips = (M - 1 + kC16Rows - 1) / kC16Rows;
    // row 0 is the hand-off row of strip -1 (buffer 3)
    c16_pair_rows(s2, N, GP, hi0, T, Wc, Rb + 3 * cols, threadIdx.x, blockDim.x);
    uint8_t *dm = dm_base + m.dm_off;
    const int S = m.stride;
    if (kDm)
      for (int j = threadIdx.x; j <= hi0; j += blockDim.x) dm[j] = j == 0 ? 1 : 2;   // row 0: ALIGN, then INSERT
    if (threadIdx.x == 0) {
      int prev_c0 = 0, end = 0, ends[kC16Waves] = {0, 0, 0, 0};
      for (int s = 0; s < nstrips; ++s) {
        const int i0 = 1 + kC16Rows * s;
        int c0, hi_first, lo_l, hi_l;
        lin_row_host(i0, M, N, g, c0, hi_first);
        const int r_last = min(i0 + kC16Rows - 1, M - 1);
        lin_row_host(r_last, M, N, g, lo_l, hi_l);
        const int t_end = hi_l - c0 + 2 * ((r_last - i0) >> 1) + ((r_last - i0) & 1);
        const int steps = (t_end + 2 + 2 * kC16U - 1) / (2 * kC16U) * (2 * kC16U);
        int st = s == 0 ? 0 : s_start[s - 1] + (c0 - prev_c0 + 150 + 2 * kC16Chunk - 1) / kC16Chunk;
        st = max(st, ends[s % kC16Waves]);   // this wave's previous strip is done
        s_start[s] = st;
        ends[s % kC16Waves] = st + (steps + kC16Chunk - 1) / kC16Chunk;
        end = max(end, ends[s % kC16Waves]);
        prev_c0 = c0;
      }
      s_chunks = end;
    }
    __syncthreads();
    const __amdgpu_buffer_rsrc_t drsrc = __builtin_amdgcn_make_buffer_rsrc(dm, 0, kDm ? M * S : 0, 0x00020000);
    const int chunks = uniform32(s_chunks);
    C16Strip<kDm> st;
    int cur = wv;   // this wave's strip
    bool on = false;
    for (int ch = 0; ch < chunks; ++ch) {
      if (ch) __syncthreads();
      if (cur < nstrips && !on && ch == s_start[cur]) {
        st.begin(1 + kC16Rows * cur, M, N, g, s1, T, Wc, Rb + ((cur + 3) & 3) * cols, Rb + (cur & 3) * cols, D, S,
                 drsrc);
        on = true;
      }
      if (on) {
        dpp_exec_guard();
#pragma unroll 1
        for (int q = 0; q < kC16Chunk / (2 * kC16U) && st.t < st.steps; ++q) st.run16();
        if (st.t >= st.steps) {
          on = false;
          cur += kC16Waves;
        }
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int last_half = 1;
      if (nstrips) last_half = (M - 1 - (1 + kC16Rows * (nstrips - 1))) & 1;
      const uint32_t e = Rb[((nstrips - 1) & 3) * cols + N - 1];
      cost_out[p] = M == 1 ? GP[N - 1] : (int32_t)(last_half ? e >> 16 : e & 0xffffu);
    }
  }
}

template <bool kDm>
__global__ __launch_bounds__(kWave * kLinWaves) void lin_fill_lds_kernel(
    const PairMeta *__restrict__ meta, const int32_t *__restrict__ order, int n, int *counter,
    const uint8_t *__restrict__ seq, const DevTcm *__restrict__ tcm, uint8_t *dm, int *aux,
    int32_t *cost_out) {
  __shared__ int s_cost[32 * 32];
  __shared__ int s_prep[32];
  __shared__ int s_tail[32];
  __shared__ int s_rows[kLinWaves][2][kLinRBP];
  __shared__ int s_gi[kLinWaves][kLinStageP];
  __shared__ uint8_t s_x[kLinWaves][kLinStage];
  load_tables(tcm, s_cost, s_prep, nullptr);
  for (int k = threadIdx.x; k < 32; k += blockDim.x) s_tail[k] = tcm->tail[k];
  __syncthreads();
  const int wv = wave_in_group();
  for (;;) {
    const int k = next_item(counter);
    if (k >= n) break;
    const int p = uniform32(order[k]);
    const PairMeta m = uniform_meta(meta, p);
    // R cells per lane where they fill most lanes (the widest row in
    // (48R, 64R]); 64-column chunks otherwise
    const int wide = uniform32(m.stride);
    if (uniform32(m.lc) <= kLinStage && wide > 192 && wide <= 256)
      lin_fill_pair_r<kDm, 4>(m, seq, s_cost, s_prep, s_tail, dm, aux, cost_out, p, &s_rows[wv][0][0],
                              &s_gi[wv][0], &s_x[wv][0]);
    else if (uniform32(m.lc) <= kLinStage && wide > 384)
      lin_fill_pair_r<kDm, 8>(m, seq, s_cost, s_prep, s_tail, dm, aux, cost_out, p, &s_rows[wv][0][0],
                              &s_gi[wv][0], &s_x[wv][0]);
    else if (uniform32(m.lc) <= kLinStage)
      lin_fill_pair<kDm, false, true>(m, seq, s_cost, s_prep, s_tail, dm, aux, cost_out, p, &s_rows[wv][0][0],
                                      kLinRB - 1, kLinRBP, &s_gi[wv][0], &s_x[wv][0]);
    else
      lin_fill_pair<kDm, false>(m, seq, s_cost, s_prep, s_tail, dm, aux, cost_out, p, &s_rows[wv][0][0],
                                kLinRB - 1, kLinRBP);
  }
}

// Bands wider than kLinRB: one wave per workgroup.  When a row fits
// kLinWideRB columns, both rows and the staged column data live in dynamic
// LDS (8 + 5 bytes per column); wider rows use global memory.
template <bool kDm>
__global__ __launch_bounds__(kWave) void lin_fill_glb_kernel(
    const PairMeta *__restrict__ meta, const int32_t *__restrict__ order, int n, int *counter,
    const uint8_t *__restrict__ seq, const DevTcm *__restrict__ tcm, uint8_t *dm, int *aux,
    int32_t *cost_out, int *rowscratch, int row_width) {
  __shared__ int s_cost[32 * 32];
  __shared__ int s_prep[32];
  __shared__ int s_tail[32];
  extern __shared__ int s_dyn[];
  load_tables(tcm, s_cost, s_prep, nullptr);
  for (int k = threadIdx.x; k < 32; k += blockDim.x) s_tail[k] = tcm->tail[k];
  __syncthreads();
  const bool lds_rows = row_width <= kLinWideRB;
  int *mine = rowscratch + (size_t)blockIdx.x * 2 * row_width;
  for (;;) {
    const int k = next_item(counter);
    if (k >= n) break;
    const int p = uniform32(order[k]);
    const PairMeta m = uniform_meta(meta, p);
    if (lds_rows)
      lin_fill_pair<kDm, false, true>(m, seq, s_cost, s_prep, s_tail, dm, aux, cost_out, p, s_dyn, -1, row_width,
                                      s_dyn + 2 * row_width, (uint8_t *)(s_dyn + 3 * row_width));
    else
      lin_fill_pair<kDm, true>(m, seq, s_cost, s_prep, s_tail, dm, aux, cost_out, p, mine, -1, row_width);
  }
}

// backtrack_2d (src/algn.c:3551-3666) + median_2 (:4042-4055), one wave per
// pair.  Slots: aligned s1 (caller's first), aligned s2, median_2.
//
// The walk itself is scalar.  It reads direction bytes from a 32 x 32 tile
// whose bottom-right corner is the cell where the tile was loaded: lane L
// holds 16 bytes of tile row L/2 (columns 16(L%2) .. +15) in four registers,
// so one round of loads serves at least 32 steps (each step leaves the row
// or the column of the previous one, never both tile edges at once).  One
// more register carries the symbols of the tile's rows (lanes 0-31) and
// columns (lanes 32-63).  A step's two output symbols go to lane (step % 64)
// of two registers; every 64 steps the lanes store them (consecutive bytes,
// right to left) and look up their medians, which a ballot compacts.
//
// Runs of ALIGN, most of an alignment, go in one step: per tile, lane L
// also holds a 32-bit mask of the ALIGN bits of diagonal L (cells (t, L-t)
// of the tile, built through an LDS copy of the tile), so the run from a
// cell is the trailing ones of one readlane, and its output columns are
// gathered from the symbol register by two ds_bpermutes.  Bytes of cells
// outside their row's band are zeroed when the tile is loaded, so a run
// never leaves the band.
constexpr int kTrTile = 32;
constexpr int kTrDrop = 0x7ffffff0;   // buffer offset past any pair's direction band

__global__ __launch_bounds__(kWave * 4) void lin_trace_wave_kernel(
    const PairMeta *__restrict__ meta, const int32_t *__restrict__ order, int n, int *counter,
    const uint8_t *__restrict__ seq, const DevTcm *__restrict__ tcm, const uint8_t *__restrict__ dm_base,
    uint8_t *out, int32_t *out_len, int32_t *status) {
  __shared__ uint8_t s_med[32 * 32];
  for (int k = threadIdx.x; k < 32 * 32; k += blockDim.x) s_med[k] = tcm->median[k];
  __syncthreads();
  const int lane = lane_id();
  __shared__ uint4 s_tt[4][kTrTile * kTrTile / 16];   // per wave: the tile, row-major
  uint4 *tt = s_tt[wave_in_group()];
  for (;;) {
    const int item = next_item(counter);
    if (item >= n) break;
    const int p = uniform32(order[item]);
    const PairMeta m = uniform_meta(meta, p);
    const uint8_t *s1 = seq + m.off_r;
    const uint8_t *s2 = seq + m.off_c;
    const int M = m.lr, N = m.lc, cs = m.p0, w = m.p1, h = m.p2, S = m.stride, cap = m.cap;
    const __amdgpu_buffer_rsrc_t drsrc =
        __builtin_amdgcn_make_buffer_rsrc((void *)(dm_base + m.dm_off), 0, M * S, 0x00020000);
    uint8_t *o = out + m.out_off;
    // swap == 0: the caller's first sequence is s1 (the longer one), and the
    // backtrack runs with swaped = true (Sequence.Align.create_edited_2)
    const bool swaped = m.swap == 0;
    uint8_t *o1 = o + (swaped ? 0 : 1) * cap;   // aligned s1
    uint8_t *o2 = o + (swaped ? 1 : 0) * cap;   // aligned s2
    uint8_t *om = o + 2 * cap;
    int i = M - 1, j = N - 1, len = 0, bad = 0;
    int I = 0x3fffffff, J = 0x3fffffff;   // tile corner; the first step loads
    uint32_t t0 = 0, t1 = 0, t2 = 0, t3 = 0, sym = 0, va = 0, vb = 0, dmk = 0;
    int tlo = 0, thi = 0;   // lanes 0-31: band of tile row lane (row I - lane)
    int outpos = cap, medpos = cap;
    // stores the batch's lanes 0..cnt-1 and their non-gap medians
    auto flush = [&](int cnt) {
      const int a = (int)(va & 0xff), b = (int)(vb & 0xff);
      const bool live = lane < cnt;
      const int med = swaped ? s_med[(a & 31) * 32 + (b & 31)] : s_med[(b & 31) * 32 + (a & 31)];
      const bool keep = live && med != kGap;
      const uint64_t km = __builtin_amdgcn_ballot_w64(keep);
      const int rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(km >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)km, 0));
      if (live) {
        o1[outpos - 1 - lane] = (uint8_t)a;
        o2[outpos - 1 - lane] = (uint8_t)b;
      }
      if (keep) om[medpos - 1 - rank] = (uint8_t)med;
      outpos -= cnt;
      medpos -= __builtin_popcountll(km);
    };
    while (i >= 0 && j >= 0) {
      if (i < I - (kTrTile - 1) || j < J - (kTrTile - 1)) {
        I = i;
        J = j;
        const int r = I - (lane >> 1);
        int rlo = 0, rhi = lin_row0_hi(N, cs, w);
        {
          int lk, rk, lo, hi;
          lin_row(max(r, 1), M, N, cs, w, h, lo, hi, lk, rk);
          rlo = r >= 1 ? lo : rlo;
          rhi = r >= 1 ? hi : rhi;
        }
        {
          const int rr = max(I - (lane & (kTrTile - 1)), 0);
          int lo, hi, lk, rk;
          lin_row(max(rr, 1), M, N, cs, w, h, lo, hi, lk, rk);
          tlo = rr == 0 ? 0 : lo;
          thi = rr == 0 ? lin_row0_hi(N, cs, w) : hi;
        }
        const int c0 = J - (kTrTile - 1) + 16 * (lane & 1);
        // rows the systolic fill wrote (s2 up to kLinStage) start at the
        // phase (L + lo - c0) mod 4 of their stride, L being the row's lane in
        // its strip and c0 the strip's first column
        int phase = 0;
        if (N <= kLinStage) {
          // strips of 64 rows (systolic fill) or 128 (packed fill, p3 = 1)
          const int L = (max(r, 1) - 1) & (m.p3 ? 2 * kWave - 1 : kWave - 1);
          int lk, rk, slo, shi;
          lin_row(max(r, 1) - L, M, N, cs, w, h, slo, shi, lk, rk);
          phase = r >= 1 ? (L + rlo - slo) & 3 : 0;
        }
        const int roff = r * S - rlo + phase;
        uint32_t wv[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          wv[q] = 0;
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const int c = c0 + 4 * q + k;
            // outside the row's band (or the matrix): no byte, the load reads 0
            const int off = r >= 0 && c >= rlo && c <= rhi ? roff + c : kTrDrop;
            wv[q] |= (uint32_t)__builtin_amdgcn_raw_buffer_load_b8(drsrc, off, 0, 0) << (8 * k);
          }
        }
        t0 = wv[0];
        t1 = wv[1];
        t2 = wv[2];
        t3 = wv[3];
        // the ALIGN bits of diagonal lane: cells (t, lane - t)
        tt[lane] = make_uint4(t0, t1, t2, t3);
        wave_sync_mem();
        const uint8_t *tb = (const uint8_t *)tt;
        dmk = 0;
#pragma unroll
        for (int t = 0; t < kTrTile; ++t) {
          const int c = lane - t;
          const int v = tb[t * kTrTile + min(max(c, 0), kTrTile - 1)];
          dmk |= (uint32_t)((c >= 0 && c < kTrTile) ? (v & 1) : 0) << t;
        }
        wave_sync_mem();   // the next tile's copy overwrites this one
        sym = lane < kTrTile ? s1[max(I - lane, 0)] : s2[max(J - (kTrTile - 1) + (lane - kTrTile), 0)];
        // wait here, once per tile, so that no step waits on the counter
        // that the output stores also hold
        __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0)
      }
      const int ri = I - i, cc = j - (J - (kTrTile - 1));
      {
        // a run of ALIGN along the walk's diagonal (lane cc + ri), inside the tile
        const uint32_t dm = (uint32_t)__builtin_amdgcn_readlane((int)dmk, cc + ri) >> ri;
        const int ones = ~dm ? __builtin_ctz(~dm) : kTrTile;
        const int run = min(ones, min(kTrTile - ri, cc + 1));
        if (run >= 2) {
          int left = run, r0 = ri, c0 = cc;
          while (left > 0) {
            const int x0 = len & (kWave - 1);
            const int chunk = min(left, kWave - x0);
            const int k = lane - x0;
            const bool in = k >= 0 && k < chunk;
            const int kk = in ? k : 0;
            const int sa = __builtin_amdgcn_ds_bpermute(4 * (r0 + kk), (int)sym);
            const int sb = __builtin_amdgcn_ds_bpermute(4 * (kTrTile + c0 - kk), (int)sym);
            va = in ? (uint32_t)sa : va;
            vb = in ? (uint32_t)sb : vb;
            len += chunk;
            r0 += chunk;
            c0 -= chunk;
            left -= chunk;
            if ((len & (kWave - 1)) == 0) flush(kWave);
          }
          i -= run;
          j -= run;
          if (j < 0 && i >= 0) {
            bad = 1;
            break;
          }
          continue;
        }
      }
      const int lo = __builtin_amdgcn_readlane(tlo, ri), hi = __builtin_amdgcn_readlane(thi, ri);
      if (j < lo || j > hi) {
        bad = 1;
        break;
      }
      const int ln = ri * 2 + (cc >> 4), q = (cc >> 2) & 3;
      // all four words, then a select: no branch per step
      const uint32_t w0 = (uint32_t)__builtin_amdgcn_readlane((int)t0, ln);
      const uint32_t w1 = (uint32_t)__builtin_amdgcn_readlane((int)t1, ln);
      const uint32_t w2 = (uint32_t)__builtin_amdgcn_readlane((int)t2, ln);
      const uint32_t w3 = (uint32_t)__builtin_amdgcn_readlane((int)t3, ln);
      const uint32_t word = (w0 & (0u - (uint32_t)(q == 0))) | (w1 & (0u - (uint32_t)(q == 1))) |
                            (w2 & (0u - (uint32_t)(q == 2))) | (w3 & (0u - (uint32_t)(q == 3)));
      const int d = (int)((word >> (8 * (cc & 3))) & 0xff);
      const int sa = __builtin_amdgcn_readlane((int)sym, ri), sb = __builtin_amdgcn_readlane((int)sym, kTrTile + cc);
      // ALIGN first, then INSERT before DELETE when swaped, else DELETE first
      const bool al = d & 1;
      const bool ins = !al && (swaped ? (d & 2) != 0 : (d & 4) == 0);
      const int a = ins ? kGap : sa, b = !al && !ins ? kGap : sb;
      i -= !ins;
      j -= al || ins;
      const bool mine = lane == (len & (kWave - 1));
      va = mine ? (uint32_t)a : va;
      vb = mine ? (uint32_t)b : vb;
      ++len;
      if ((len & (kWave - 1)) == 0) flush(kWave);
      if (j < 0 && i >= 0) {
        bad = 1;
        break;
      }
    }
    if (len & (kWave - 1)) flush(len & (kWave - 1));
    if (lane == 0) om[medpos - 1] = kGap;   // the leading gap of median_2
    out_len[3 * p + 0] = len;
    out_len[3 * p + 1] = len;
    out_len[3 * p + 2] = cap - (medpos - 1);
    if (bad) mark_failed(status, p);
  }
}

// ---------------------------------------------------------------------------
// Fitch over packed nonadditive characters (src/nonaddCSc.c:422-544), for
// the reference's three character widths (nonaddCSc.c compiled with CHARSIZE
// 8, 16 and 32, src/myocamlbuild.ml.in:255-269): W bytes per character.
// HBM-bound bit operations, 16 bytes per lane per step.
// ---------------------------------------------------------------------------
// All ones in every W-byte lane of v that is zero.
template <int W>
__device__ __forceinline__ uint32_t empty_lanes(uint32_t v) {
  if (W == 1) {
    uint32_t t = ((v & 0x7f7f7f7fu) + 0x7f7f7f7fu) | v;
    t = ~(t | 0x7f7f7f7fu);   // 0x80 where the byte is zero
    return (t >> 7) * 0xffu;
  } else if (W == 2) {
    uint32_t t = ((v & 0x7fff7fffu) + 0x7fff7fffu) | v;
    t = ~(t | 0x7fff7fffu);   // 0x8000 where the half is zero
    return (t >> 15) * 0xffffu;
  }
  return v ? 0u : 0xffffffffu;
}
template <int W>
using fitch_elt_t = typename std::conditional<W == 1, uint8_t, typename std::conditional<W == 2, uint16_t, uint32_t>::type>::type;

template <int W>
__global__ void fitch_down_kernel(uint8_t *chars, int len, int n, const int32_t *ia, const int32_t *ib,
                                  const int32_t *ires, int64_t *added) {
  using T = fitch_elt_t<W>;
  const int node = blockIdx.y;
  if (node >= n) return;
  const size_t row = (size_t)len * W;
  const uint8_t *a = chars + (size_t)ia[node] * row;
  const uint8_t *b = chars + (size_t)ib[node] * row;
  uint8_t *r = chars + (size_t)ires[node] * row;
  long long cnt = 0;
  const int nvec = (row % 16 == 0) ? (int)(row / 16) : 0;
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nvec; v += gridDim.x * blockDim.x) {
    const uint4 x = ((const uint4 *)a)[v];
    const uint4 y = ((const uint4 *)b)[v];
    uint32_t xs[4] = {x.x, x.y, x.z, x.w}, ys[4] = {y.x, y.y, y.z, y.w}, rs[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t in = xs[q] & ys[q], un = xs[q] | ys[q];
      const uint32_t emp = empty_lanes<W>(in);
      cnt += __popc(emp) / (8 * W);
      rs[q] = (emp & un) | in;
    }
    ((uint4 *)r)[v] = make_uint4(rs[0], rs[1], rs[2], rs[3]);
  }
  if (blockIdx.x == 0 && nvec == 0) {
    const T *ta = (const T *)a, *tb = (const T *)b;
    T *tr = (T *)r;
    for (int c = threadIdx.x; c < len; c += blockDim.x) {
      const T in = ta[c] & tb[c];
      if (!in) ++cnt;
      tr[c] = in ? in : (T)(ta[c] | tb[c]);
    }
  }
  // block reduction through LDS, one atomic per block
  __shared__ unsigned long long s_red[256];
  s_red[threadIdx.x] = (unsigned long long)cnt;
  __syncthreads();
  for (int d = blockDim.x / 2; d > 0; d >>= 1) {
    if ((int)threadIdx.x < d) s_red[threadIdx.x] += s_red[threadIdx.x + d];
    __syncthreads();
  }
  if (threadIdx.x == 0 && s_red[0]) atomicAdd((unsigned long long *)&added[node], s_red[0]);
}

template <int W>
__global__ void fitch_up_kernel(uint8_t *chars, int len, int n, const int32_t *iA, const int32_t *iN,
                                const int32_t *iD1, const int32_t *iD2, const int32_t *iF) {
  using T = fitch_elt_t<W>;
  const int node = blockIdx.y;
  if (node >= n) return;
  const size_t row = (size_t)len * W;
  const uint8_t *A = chars + (size_t)iA[node] * row;
  const uint8_t *N = chars + (size_t)iN[node] * row;
  const uint8_t *D1 = chars + (size_t)iD1[node] * row;
  const uint8_t *D2 = chars + (size_t)iD2[node] * row;
  uint8_t *F = chars + (size_t)iF[node] * row;
  const int nvec = (row % 16 == 0) ? (int)(row / 16) : 0;
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nvec; v += gridDim.x * blockDim.x) {
    const uint4 a4 = ((const uint4 *)A)[v], n4 = ((const uint4 *)N)[v];
    const uint4 d14 = ((const uint4 *)D1)[v], d24 = ((const uint4 *)D2)[v];
    uint32_t as[4] = {a4.x, a4.y, a4.z, a4.w}, ns[4] = {n4.x, n4.y, n4.z, n4.w};
    uint32_t d1[4] = {d14.x, d14.y, d14.z, d14.w}, d2[4] = {d24.x, d24.y, d24.z, d24.w}, fs[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      uint32_t inter = ns[q] | (as[q] & (d1[q] | d2[q]));
      uint32_t temp = inter;
      const uint32_t was_union = empty_lanes<W>(d1[q] & d2[q]);
      inter = (inter | as[q]) & was_union;
      temp |= inter;
      const uint32_t not_contained = ~empty_lanes<W>((as[q] & ns[q]) ^ as[q]);   // ones where A is not within N
      fs[q] = (~not_contained & as[q]) | (not_contained & temp);
    }
    ((uint4 *)F)[v] = make_uint4(fs[0], fs[1], fs[2], fs[3]);
  }
  if (blockIdx.x == 0 && nvec == 0) {
    const T *tA = (const T *)A, *tN = (const T *)N, *t1 = (const T *)D1, *t2 = (const T *)D2;
    T *tF = (T *)F;
    for (int c = threadIdx.x; c < len; c += blockDim.x) {
      const T a = tA[c], nn = tN[c], x1 = t1[c], x2 = t2[c];
      T f;
      if ((T)(a & nn) == a) f = a;
      else if ((T)(x1 & x2) == 0) f = (T)(nn | a);
      else f = (T)(nn | (a & (x1 | x2)));
      tF[c] = f;
    }
  }
}

}  // namespace poyg

// ---------------------------------------------------------------------------
// launch wrappers used by the host side
// ---------------------------------------------------------------------------
namespace poyg {

hipError_t launch_reset(int *counters, int n, hipStream_t s) {
  const int grid = std::max(1, std::min((n + 255) / 256, 1024));
  hipLaunchKernelGGL(reset_counters_kernel, dim3(grid), dim3(256), 0, s, counters, n);
  return hipGetLastError();
}

hipError_t launch_plan_start(int *ctr, int nctr, int *status, int nstatus, TraceOut *tro, uint8_t *out,
                             int32_t *out_len, hipStream_t s) {
  const int grid = std::max(1, std::min((std::max(nctr, nstatus) + 255) / 256, 1024));
  hipLaunchKernelGGL(plan_start_kernel, dim3(grid), dim3(256), 0, s, ctr, nctr, status, nstatus, tro, out, out_len);
  return hipGetLastError();
}

hipError_t launch_collect_results(const uint32_t *small, int64_t small_words, const int32_t *status,
                                  const uint32_t *out, int64_t out_words, uint32_t *host, hipStream_t s) {
  const int64_t words = std::max(small_words, out_words);
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((words + 255) / 256, 1024));
  hipLaunchKernelGGL(collect_results_kernel, dim3(grid), dim3(256), 0, s, small, small_words, status, out, out_words,
                     host);
  return hipGetLastError();
}

hipError_t launch_prep(bool affine, int grid, const PairMeta *meta, int n, const uint8_t *seq, const DevTcm *tcm,
                       int *aux, hipStream_t s) {
  if (affine)
    hipLaunchKernelGGL(aff_prep_kernel, dim3(grid), dim3(kWave * 4), 0, s, meta, n, seq, tcm, aux);
  else
    hipLaunchKernelGGL(lin_prep_kernel, dim3(grid), dim3(kWave * 4), 0, s, meta, n, seq, tcm, aux);
  return hipGetLastError();
}

hipError_t launch_aff_fill(bool bt, bool glb, int grid, const PairMeta *meta, const int32_t *order, int n,
                           int *counter, const uint8_t *seq, const DevTcm *tcm, uint8_t *dm, int *aux,
                           int32_t *cost, int *rowscratch, int row_width, hipStream_t s) {
  if (!glb) {
    if (bt)
      hipLaunchKernelGGL(aff_fill_lds_kernel<true>, dim3(grid), dim3(kWave * kAffWaves), 0, s, meta, order, n,
                         counter, seq, tcm, dm, aux, cost);
    else
      hipLaunchKernelGGL(aff_fill_lds_kernel<false>, dim3(grid), dim3(kWave * kAffWaves), 0, s, meta, order, n,
                         counter, seq, tcm, dm, aux, cost);
  } else {
    if (bt)
      hipLaunchKernelGGL(aff_fill_glb_kernel<true>, dim3(grid), dim3(kWave), 0, s, meta, order, n, counter, seq,
                         tcm, dm, aux, cost, rowscratch, row_width);
    else
      hipLaunchKernelGGL(aff_fill_glb_kernel<false>, dim3(grid), dim3(kWave), 0, s, meta, order, n, counter, seq,
                         tcm, dm, aux, cost, rowscratch, row_width);
  }
  return hipGetLastError();
}

// Occupancy answers per (device, kernel, launch shape), kept for the
// process: they depend only on the kernel's resources and the device, and a
// plan of one pair would otherwise spend more time asking than running.
static std::mutex g_occ_mu;
static std::map<std::pair<int, int64_t>, int> g_occ;
template <typename F>
static int occupancy_cached(int64_t key, F query) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return query();
  {
    std::lock_guard<std::mutex> g(g_occ_mu);
    const auto it = g_occ.find({dev, key});
    if (it != g_occ.end()) return it->second;
  }
  const int v = query();
  std::lock_guard<std::mutex> g(g_occ_mu);
  g_occ[{dev, key}] = v;
  return v;
}

// Workgroups of aff_diag_kernel<bt, R> one CU holds at once (register and
// LDS limits), for a persistent grid that is exactly resident.
template <bool kBt, int R>
static int diag_occupancy() {
  return occupancy_cached(0x100 + 16 * kBt + R, [] {
    int nb = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, aff_diag_kernel<kBt, R>, kWave * 4, 0);
    return (e == hipSuccess && nb > 0) ? nb : 2;
  });
}
int aff_diag_blocks_per_cu(bool bt, int sets) {
  if (sets == 1) return bt ? diag_occupancy<true, 1>() : diag_occupancy<false, 1>();
  if (sets == 2) return bt ? diag_occupancy<true, 2>() : diag_occupancy<false, 2>();
  return bt ? diag_occupancy<true, 4>() : diag_occupancy<false, 4>();
}

template <int R>
static void launch_diag_r(bool bt, int grid, const PairMeta *meta, const int32_t *order, int n, int *counter,
                          const uint8_t *seq, const DevTcm *tcm, uint8_t *dm, int32_t *cost, uint32_t *sink,
                          const TraceOut *tro, hipStream_t s) {
  if (bt)
    hipLaunchKernelGGL((aff_diag_kernel<true, R>), dim3(grid), dim3(kWave * 4), 0, s, meta, order, n, counter, seq,
                       tcm, dm, cost, sink, tro);
  else
    hipLaunchKernelGGL((aff_diag_kernel<false, R>), dim3(grid), dim3(kWave * 4), 0, s, meta, order, n, counter, seq,
                       tcm, dm, cost, sink, tro);
}

hipError_t launch_aff_diag(bool bt, int sets, int grid, const PairMeta *meta, const int32_t *order, int n,
                           int *counter, const uint8_t *seq, const DevTcm *tcm, uint8_t *dm, int32_t *cost,
                           uint32_t *sink, const TraceOut *tro, hipStream_t s) {
  if (sets == 1) launch_diag_r<1>(bt, grid, meta, order, n, counter, seq, tcm, dm, cost, sink, tro, s);
  else if (sets == 2) launch_diag_r<2>(bt, grid, meta, order, n, counter, seq, tcm, dm, cost, sink, tro, s);
  else launch_diag_r<4>(bt, grid, meta, order, n, counter, seq, tcm, dm, cost, sink, tro, s);
  return hipGetLastError();
}

template <bool kBt>
static int pk_occupancy() {
  return occupancy_cached(0x600 + kBt, [] {
    int nb = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, aff_pk_kernel<kBt, 3>, kWave * 4, 0);
    return (e == hipSuccess && nb > 0) ? nb : 1;
  });
}
int aff_pk_blocks_per_cu(bool bt) { return bt ? pk_occupancy<true>() : pk_occupancy<false>(); }

template <bool kBt>
static int quad_occupancy() {
  return occupancy_cached(0x200 + kBt, [] {
    int nb = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, aff_quad_kernel<kBt, 3>, kWave * 4, 0);
    return (e == hipSuccess && nb > 0) ? nb : 1;
  });
}
int aff_quad_blocks_per_cu(bool bt) { return bt ? quad_occupancy<true>() : quad_occupancy<false>(); }

hipError_t launch_pair_pure(int grid, const PairMeta *meta, const int32_t *order, int n, const uint8_t *seq,
                            uint8_t *pure, hipStream_t s) {
  hipLaunchKernelGGL(pair_pure_kernel, dim3(grid), dim3(kWave * 4), 0, s, meta, order, n, seq, pure);
  return hipGetLastError();
}

hipError_t launch_aff_quad(bool bt, int grid, const PairMeta *meta, const int32_t *order, int n, int *counter,
                           const uint8_t *seq, const DevTcm *tcm, uint8_t *dm, int32_t *cost, uint32_t *sink,
                           const uint8_t *pure, const TraceOut *tro, const int *n_dev, hipStream_t s) {
  if (bt)
    hipLaunchKernelGGL((aff_quad_kernel<true, 3>), dim3(grid), dim3(kWave * 4), 0, s, meta, order, n, counter, seq,
                       tcm, dm, cost, sink, pure, tro, n_dev);
  else
    hipLaunchKernelGGL((aff_quad_kernel<false, 3>), dim3(grid), dim3(kWave * 4), 0, s, meta, order, n, counter, seq,
                       tcm, dm, cost, sink, pure, tro, n_dev);
  return hipGetLastError();
}

hipError_t launch_aff_pk(bool bt, int grid, const PairMeta *meta, const int32_t *order, int n, int *counter,
                         const uint8_t *seq, const DevTcm *tcm, uint8_t *dm, int32_t *cost, uint32_t *sink,
                         const uint8_t *pure, int32_t *redo, int *redo_n, int32_t *tlist, int *tlist_n,
                         hipStream_t s) {
  if (bt)
    hipLaunchKernelGGL((aff_pk_kernel<true, 3>), dim3(grid), dim3(kWave * 4), 0, s, meta, order, n, counter, seq,
                       tcm, dm, cost, sink, pure, redo, redo_n, tlist, tlist_n);
  else
    hipLaunchKernelGGL((aff_pk_kernel<false, 3>), dim3(grid), dim3(kWave * 4), 0, s, meta, order, n, counter, seq,
                       tcm, dm, cost, sink, pure, redo, redo_n, tlist, tlist_n);
  return hipGetLastError();
}

int aff_trace_list_blocks_per_cu() {
  return occupancy_cached(0x700, [] {
    int nb = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, aff_trace_list_kernel, kWave * 4, 0);
    return (e == hipSuccess && nb > 0) ? nb : 2;
  });
}

hipError_t launch_aff_trace_list(int grid, const PairMeta *meta, const int32_t *list, const int *list_n, int *counter,
                                 const uint8_t *seq, const DevTcm *tcm, const uint8_t *dm, const TraceOut *tro,
                                 hipStream_t s) {
  hipLaunchKernelGGL(aff_trace_list_kernel, dim3(grid), dim3(kWave * 4), 0, s, meta, list, list_n, counter, seq, tcm,
                     dm, tro);
  return hipGetLastError();
}

hipError_t launch_aff_trace(const PairMeta *meta, const int32_t *order, int n, const uint8_t *seq,
                            const DevTcm *tcm, const uint8_t *dm, uint8_t *out, int32_t *out_len, int32_t *status,
                            hipStream_t s) {
  const int bs = 64;
  hipLaunchKernelGGL(aff_trace_kernel, dim3((n + bs - 1) / bs), dim3(bs), 0, s, meta, order, n, seq, tcm, dm, out,
                     out_len, status);
  return hipGetLastError();
}

template <bool kDm>
static int sys_occupancy() {
  return occupancy_cached(0x300 + kDm, [] {
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, lin_sys_kernel<kDm>, kWave * kSysWaves, 0) != hipSuccess)
      b = 1;
    return std::max(1, b);
  });
}
int lin_sys_blocks_per_cu(bool with_dm) { return with_dm ? sys_occupancy<true>() : sys_occupancy<false>(); }

hipError_t launch_lin_sys(bool with_dm, int grid, const PairMeta *meta, const int32_t *order, int n, int *counter,
                          const uint8_t *seq, const DevTcm *tcm, uint8_t *dm, const int *aux, int32_t *cost,
                          int32_t *status, int spin_limit, hipStream_t s) {
  if (spin_limit < 0) spin_limit = kSysSpin;
  if (with_dm)
    hipLaunchKernelGGL(lin_sys_kernel<true>, dim3(grid), dim3(kWave * kSysWaves), 0, s, meta, order, n, counter, seq,
                       tcm, dm, aux, cost, status, spin_limit);
  else
    hipLaunchKernelGGL(lin_sys_kernel<false>, dim3(grid), dim3(kWave * kSysWaves), 0, s, meta, order, n, counter,
                       seq, tcm, dm, aux, cost, status, spin_limit);
  return hipGetLastError();
}

// Waves per workgroup of the packed cost fill: as many as the per-workgroup
// LDS limit allows (at most kC16Waves), so that long columns still launch.
static int c16_waves(int n_max) {
  static int limit = -1;
  if (limit < 0) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
      v = 65536;
    limit = v;
  }
  const int per_wave = 4 * c16_wave_words(n_max);
  const int avail = limit - (int)(sizeof(uint16_t) * 32 * 32);
  return std::max(1, std::min(kC16Waves, avail / per_wave));
}

int lin_cost16_blocks_per_cu(bool with_dm, int n_max) {
  const int waves = c16_waves(n_max);
  const int lds = 4 * waves * c16_wave_words(n_max);
  // the launch shape depends on n_max only through waves and lds
  return occupancy_cached(((int64_t)0x400 + with_dm) << 32 | (int64_t)lds << 4 | waves, [&] {
    int b = 0;
    const hipError_t e =
        with_dm ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, lin_cost16_kernel<true>, kWave * waves, lds)
                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, lin_cost16_kernel<false>, kWave * waves, lds);
    if (e != hipSuccess) b = 1;
    return std::max(1, b) * waves;   // in waves
  });
}

hipError_t launch_lin_cost16(bool with_dm, int grid_waves, const PairMeta *meta, const int32_t *order, int n,
                             int *counter, const uint8_t *seq, const DevTcm *tcm, const int *aux, int32_t *cost,
                             uint8_t *dm, int n_max, hipStream_t s) {
  const int waves = c16_waves(n_max);
  const int grid = std::max(1, (grid_waves + waves - 1) / waves);
  const int lds = 4 * waves * c16_wave_words(n_max);
  if (with_dm)
    hipLaunchKernelGGL(lin_cost16_kernel<true>, dim3(grid), dim3(kWave * waves), lds, s, meta, order, n, counter, seq,
                       tcm, aux, cost, dm, n_max);
  else
    hipLaunchKernelGGL(lin_cost16_kernel<false>, dim3(grid), dim3(kWave * waves), lds, s, meta, order, n, counter,
                       seq, tcm, aux, cost, dm, n_max);
  return hipGetLastError();
}

int lin_c16mw_blocks_per_cu(bool with_dm, int n_max) {
  const int lds = 4 * c16mw_words(n_max);
  return occupancy_cached(((int64_t)0x800 + with_dm) << 32 | (int64_t)lds, [&] {
    int b = 0;
    const hipError_t e =
        with_dm ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, lin_c16mw_kernel<true>, kWave * kC16Waves, lds)
                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, lin_c16mw_kernel<false>, kWave * kC16Waves, lds);
    return (e == hipSuccess && b > 0) ? b : 1;
  });
}

hipError_t launch_lin_c16mw(bool with_dm, int grid, const PairMeta *meta, const int32_t *order, int n, int *counter,
                            const uint8_t *seq, const DevTcm *tcm, const int *aux, int32_t *cost, uint8_t *dm,
                            int n_max, hipStream_t s) {
  const int lds = 4 * c16mw_words(n_max);
  if (with_dm)
    hipLaunchKernelGGL(lin_c16mw_kernel<true>, dim3(grid), dim3(kWave * kC16Waves), lds, s, meta, order, n, counter,
                       seq, tcm, aux, cost, dm, n_max);
  else
    hipLaunchKernelGGL(lin_c16mw_kernel<false>, dim3(grid), dim3(kWave * kC16Waves), lds, s, meta, order, n, counter,
                       seq, tcm, aux, cost, dm, n_max);
  return hipGetLastError();
}

hipError_t launch_lin_fill(bool with_dm, bool glb, int grid, const PairMeta *meta, const int32_t *order, int n,
                           int *counter, const uint8_t *seq, const DevTcm *tcm, uint8_t *dm, int *aux,
                           int32_t *cost, int *rowscratch, int row_width, hipStream_t s) {
  if (!glb) {
    if (with_dm)
      hipLaunchKernelGGL(lin_fill_lds_kernel<true>, dim3(grid), dim3(kWave * kLinWaves), 0, s, meta, order, n,
                         counter, seq, tcm, dm, aux, cost);
    else
      hipLaunchKernelGGL(lin_fill_lds_kernel<false>, dim3(grid), dim3(kWave * kLinWaves), 0, s, meta, order, n,
                         counter, seq, tcm, dm, aux, cost);
  } else {
    // rows (two of row_width ints), the prefix sums (row_width ints) and the
    // symbols (row_width bytes); s2 is shorter than row_width
    const size_t dyn = row_width <= kLinWideRB ? (size_t)row_width * 13 : 0;
    if (with_dm)
      hipLaunchKernelGGL(lin_fill_glb_kernel<true>, dim3(grid), dim3(kWave), dyn, s, meta, order, n, counter, seq,
                         tcm, dm, aux, cost, rowscratch, row_width);
    else
      hipLaunchKernelGGL(lin_fill_glb_kernel<false>, dim3(grid), dim3(kWave), dyn, s, meta, order, n, counter,
                         seq, tcm, dm, aux, cost, rowscratch, row_width);
  }
  return hipGetLastError();
}

int lin_trace_blocks_per_cu() {
  return occupancy_cached(0x500, [] {
    int nb = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, lin_trace_wave_kernel, kWave * 4, 0);
    return (e == hipSuccess && nb > 0) ? nb : 2;
  });
}

hipError_t launch_lin_trace(int grid, const PairMeta *meta, const int32_t *order, int n, int *counter,
                            const uint8_t *seq, const DevTcm *tcm, const uint8_t *dm, uint8_t *out, int32_t *out_len,
                            int32_t *status, hipStream_t s) {
  hipLaunchKernelGGL(lin_trace_wave_kernel, dim3(grid), dim3(kWave * 4), 0, s, meta, order, n, counter, seq, tcm, dm,
                     out, out_len, status);
  return hipGetLastError();
}

// Blocks per node along the characters: enough for the whole launch to
// cover the chip (about 4,096 blocks) when there are few nodes, at most one
// block per 256 vectors of 16 characters.
static int fitch_blocks(int64_t row_bytes, int n) {
  const int bs = 256;
  const int64_t nvec = row_bytes / 16;
  int64_t gx = (nvec + bs - 1) / bs;
  const int want = std::max(64, 4096 / std::max(n, 1));
  if (gx > want) gx = want;
  return gx < 1 ? 1 : (int)gx;
}

hipError_t launch_fitch_down(int width, uint8_t *chars, int len, int n, const int32_t *ia, const int32_t *ib,
                             const int32_t *ires, int64_t *added, hipStream_t s) {
  const int bs = 256;
  const int gx = fitch_blocks((int64_t)len * width, n);
  if (width == 1)
    hipLaunchKernelGGL(fitch_down_kernel<1>, dim3(gx, n), dim3(bs), 0, s, chars, len, n, ia, ib, ires, added);
  else if (width == 2)
    hipLaunchKernelGGL(fitch_down_kernel<2>, dim3(gx, n), dim3(bs), 0, s, chars, len, n, ia, ib, ires, added);
  else
    hipLaunchKernelGGL(fitch_down_kernel<4>, dim3(gx, n), dim3(bs), 0, s, chars, len, n, ia, ib, ires, added);
  return hipGetLastError();
}

hipError_t launch_fitch_up(int width, uint8_t *chars, int len, int n, const int32_t *iA, const int32_t *iN,
                           const int32_t *iD1, const int32_t *iD2, const int32_t *iF, hipStream_t s) {
  const int bs = 256;
  const int gx = fitch_blocks((int64_t)len * width, n);
  if (width == 1)
    hipLaunchKernelGGL(fitch_up_kernel<1>, dim3(gx, n), dim3(bs), 0, s, chars, len, n, iA, iN, iD1, iD2, iF);
  else if (width == 2)
    hipLaunchKernelGGL(fitch_up_kernel<2>, dim3(gx, n), dim3(bs), 0, s, chars, len, n, iA, iN, iD1, iD2, iF);
  else
    hipLaunchKernelGGL(fitch_up_kernel<4>, dim3(gx, n), dim3(bs), 0, s, chars, len, n, iA, iN, iD1, iD2, iF);
  return hipGetLastError();
}

}  // namespace poyg
