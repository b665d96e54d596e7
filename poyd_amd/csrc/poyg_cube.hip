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
  int32_t dstride;   // bytes per row of the direction cube
  int32_t flags;     // bit 0: two 4-bit codes per byte (else one per byte); bits 1-2: wave priority
};

// Direction codes, one byte per cell: the index of the winning move in the
// reference's evaluation order (fill_parallel P3, P1, P2, then fill_moved S3,
// S1, S2, then the self pass SS; src/algn.c:2812-2858, 3041-3048).  With
// strict-< first-wins updates the winner is the first minimum in this order.
constexpr int kP3 = 0, kP1 = 1, kP2 = 2, kS3 = 3, kS1 = 4, kS2 = 5, kSS = 6;

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

// Ordering of one wave's own global stores before its later loads of the
// same rows (every cube kernel runs one wave per workgroup, so workgroup
// scope suffices: no L2 write-back or invalidate is needed).
__device__ __forceinline__ void c_wave_fence() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup"); }

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
    c_wave_fence();
#define ROW(R) (ring + (size_t)((R) % W) * n3)
    // plane 0, line 0: 0, then the self-gap chain
    for (int k0 = 0; k0 < n3; k0 += 64) {
      const int k = k0 + lane;
      if (k < n3) {
        ROW(0)[k] = G3[k];
        dm[k] = k == 0 ? kS2 : kSS;
      }
    }
    c_wave_fence();
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
      c_wave_fence();
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
        if (out - (n2 + p - 1) > fenced) { c_wave_fence(); fenced = out - 1; }
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
        if (out - p - 1 > fenced) { c_wave_fence(); fenced = out - 1; }
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
    c_wave_fence();
    cost_out[t] = ROW(n1 * n2 - 1)[n3 - 1];
    c_wave_fence();
#undef ROW
  }
}

// ---------------------------------------------------------------------------
// Row kernel: lane L owns the R consecutive cells k = L*R .. L*R+R-1 of every
// row (n3 <= 64*R), so a row is one pass: six candidate moves per cell as
// ranked keys (value*8 + move index; the minimum key is the reference's
// first minimum), then the self pass along s3 as one prefix minimum
// (mm[k] = min(v[k], mm[k-1] + g[k]) equals G3[k] + min_{j<=k}(v[j] - G3[j])
// with G3 the prefix sums of g; the self move wins exactly where
// mm[k-1] + g[k] < v[k]).
//
// The reference's row pointers lag (see the top of this file): row `out` of
// plane p >= 1 reads rows out-(n2+p-1) ("upper", or "diagonal" on line 0),
// out-(n2+p) ("diagonal") and out-(p+1) ("previous").  Every input of row
// out+1 is therefore complete once row out-1 is stored, so the loads of the
// next row are issued before the current row is computed, and the
// "diagonal" row of line r is the "upper" row of line r-1 (line 1 reuses the
// row of line 0), already in registers.  Rows live in a per-wave ring of
// n1+n2 rows; T is int16_t when every cell value is provably below 2^15
// (checked on the host), else int32_t.
constexpr int kBig3 = 1 << 28;
constexpr int kTop3 = 0x7fffffff;
constexpr int kAhead = 4;   // rows whose inputs are in flight

__device__ __forceinline__ int c_shr1(int v, int fill) {
  return __builtin_amdgcn_update_dpp(fill, v, 0x138 /* wave_shr:1 */, 0xf, 0xf, false);
}

template <typename T, int R>
struct RingRow {
  static constexpr int kWords = R * (int)sizeof(T) / 4;
  uint32_t w[kWords];
  __device__ __forceinline__ void load(const T *p) {
    static_assert(kWords % 2 == 0, "rows are loaded in 8-byte pieces");
    const uint2 *q = (const uint2 *)p;
#pragma unroll
    for (int i = 0; i < kWords / 2; ++i) {
      const uint2 v = q[i];
      w[2 * i] = v.x;
      w[2 * i + 1] = v.y;
    }
  }
  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int i = 0; i < kWords; ++i) w[i] = 0;
  }
  // cell j as a non-negative int (inactive cells may hold anything)
  __device__ __forceinline__ int get(int j) const {
    if constexpr (sizeof(T) == 2) return (int)((w[j >> 1] >> (16 * (j & 1))) & 0xffffu);
    else return (int)w[j];
  }
};

template <typename T, int R>
__device__ __forceinline__ void store_ring(T *p, const int (&mm)[R]) {
  constexpr int kWords = R * (int)sizeof(T) / 4;
  uint32_t w[kWords];
#pragma unroll
  for (int i = 0; i < kWords; ++i) {
    if constexpr (sizeof(T) == 2)
      w[i] = __builtin_amdgcn_perm((uint32_t)mm[2 * i + 1], (uint32_t)mm[2 * i], 0x05040100u);
    else
      w[i] = (uint32_t)mm[i];
  }
  uint2 *q = (uint2 *)p;
#pragma unroll
  for (int i = 0; i < kWords / 2; ++i) q[i] = make_uint2(w[2 * i], w[2 * i + 1]);
}

// R 4-bit codes (R/2 bytes at p, aligned to R/2 bytes)
template <int R>
__device__ __forceinline__ void store_codes(uint8_t *p, const uint32_t (&c)[(R + 7) / 8]) {
  if constexpr (R == 4) {
    *(uint16_t *)p = (uint16_t)c[0];
  } else if constexpr (R == 8) {
    *(uint32_t *)p = c[0];
  } else if constexpr (R == 12) {
    uint16_t *q = (uint16_t *)p;
    q[0] = (uint16_t)c[0];
    q[1] = (uint16_t)(c[0] >> 16);
    q[2] = (uint16_t)c[1];
  } else {
    static_assert(R == 16, "row widths are 4, 8, 12 or 16 cells per lane");
    *(uint2 *)p = make_uint2(c[0], c[1]);
  }
}

// One row.  kKind 0: plane 0, line >= 1 (previous-row moves only); 1: line 0
// of a plane >= 1 (upper moves only, the upper row being the reference's
// diagonal pointer there); 2: every other row.  U, D, P: raw values of the
// upper, diagonal and previous rows; tab: the LDS cost table row of s2[r].
// Out: mm (raw values) and the packed move codes.
template <int R, int kKind>
__device__ __forceinline__ void cube_row(const int (&U)[R], const int (&D)[R], const int (&P)[R],
                                         const int (&G3s)[R], const int (&yo)[R], const int4 *tab, int cP3,
                                         int cP1, int cP2, int (&mm)[R], uint32_t (&codes)[(R + 7) / 8]) {
  // cell k-1 of lane 0's first cell does not exist: the fill value makes its
  // moves lose to every real key
  constexpr int kNone = kBig3 >> 3;
  int Ul = 0, Dl = 0, Pl = 0;
  c_dpp_guard();
  if (kKind != 0) Ul = c_shr1(U[R - 1], kNone);
  if (kKind == 2) Dl = c_shr1(D[R - 1], kNone);
  if (kKind != 1) Pl = c_shr1(P[R - 1], kNone);
  int key[R];
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const int4 tb = tab[yo[j]];
    const int ul = j ? U[j - 1] : Ul, dl = j ? D[j - 1] : Dl, pl = j ? P[j - 1] : Pl;
    int k;
    if (kKind == 0) {
      k = min((P[j] << 3) + cP1, (pl << 3) + tb.x);
    } else if (kKind == 1) {
      k = min((U[j] << 3) + cP3, (ul << 3) + tb.z);
    } else {
      k = min(min((U[j] << 3) + cP3, (P[j] << 3) + cP1), (D[j] << 3) + cP2);
      k = min(min(k, (ul << 3) + tb.z), min((pl << 3) + tb.x, (dl << 3) + tb.y));
    }
    key[j] = k;
  }
  // self pass along s3
  int wv[R], lp[R];
  int run = kTop3;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    wv[j] = (key[j] & ~7) - G3s[j];
    run = min(run, wv[j]);
    lp[j] = run;
  }
  const int incl = c_prefix_min(run);
  c_dpp_guard();
  const int excl = c_shr1(incl, kTop3);
  int prev = excl;
#pragma unroll
  for (int j = 0; j < (R + 7) / 8; ++j) codes[j] = 0;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const int pm = min(excl, lp[j]);
    const bool ss = prev < wv[j];
    mm[j] = (G3s[j] + pm) >> 3;
    const uint32_t c = ss ? (uint32_t)kSS : (uint32_t)(key[j] & 7);
    codes[j >> 3] |= c << (4 * (j & 7));
    prev = pm;
  }
}

// Inclusive prefix minimum over the wave without the EXEC guard (for use
// inside a block whose EXEC is unchanged since its one guard).
__device__ __forceinline__ int c_prefix_min_ng(int v) {
  constexpr int I = 0x7fffffff;
  v = min(v, __builtin_amdgcn_update_dpp(I, v, 0x111, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(I, v, 0x112, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(I, v, 0x114, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(I, v, 0x118, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(I, v, 0x142, 0xa, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(I, v, 0x143, 0xc, 0xf, false));
  return v;
}

// Two consecutive rows of one plane (lines r and r+1 >= 1): independent of
// each other (the lag puts every input of either at least two rows back),
// computed in one block with one EXEC guard so that their dependent chains
// interleave.  The second row's diagonal row is the first row's upper row.
template <int R>
__device__ __forceinline__ void cube_row_pair(const int (&U0)[R], const int (&D0)[R], const int (&P0)[R],
                                              const int4 *tab0, int cP1a, int cP2a, const int (&U1)[R],
                                              const int (&P1)[R], const int4 *tab1, int cP1b, int cP2b,
                                              const int (&G3s)[R], const int (&yo)[R], int cP3,
                                              int (&mm0)[R], uint32_t (&codes0)[(R + 7) / 8], int (&mm1)[R],
                                              uint32_t (&codes1)[(R + 7) / 8]) {
  constexpr int kNone = kBig3 >> 3;
  c_dpp_guard();
  const int Ul0 = c_shr1(U0[R - 1], kNone), Dl0 = c_shr1(D0[R - 1], kNone), Pl0 = c_shr1(P0[R - 1], kNone);
  const int Ul1 = c_shr1(U1[R - 1], kNone), Pl1 = c_shr1(P1[R - 1], kNone);
  int key0[R], key1[R];
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const int4 ta = tab0[yo[j]], tb = tab1[yo[j]];
    const int ul0 = j ? U0[j - 1] : Ul0, dl0 = j ? D0[j - 1] : Dl0, pl0 = j ? P0[j - 1] : Pl0;
    const int ul1 = j ? U1[j - 1] : Ul1, dl1 = ul0, pl1 = j ? P1[j - 1] : Pl1;
    int k0 = min(min((U0[j] << 3) + cP3, (P0[j] << 3) + cP1a), (D0[j] << 3) + cP2a);
    k0 = min(min(k0, (ul0 << 3) + ta.z), min((pl0 << 3) + ta.x, (dl0 << 3) + ta.y));
    int k1 = min(min((U1[j] << 3) + cP3, (P1[j] << 3) + cP1b), (U0[j] << 3) + cP2b);
    k1 = min(min(k1, (ul1 << 3) + tb.z), min((pl1 << 3) + tb.x, (dl1 << 3) + tb.y));
    key0[j] = k0;
    key1[j] = k1;
  }
  int wv0[R], lp0[R], wv1[R], lp1[R];
  int run0 = kTop3, run1 = kTop3;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    wv0[j] = (key0[j] & ~7) - G3s[j];
    wv1[j] = (key1[j] & ~7) - G3s[j];
    run0 = min(run0, wv0[j]);
    run1 = min(run1, wv1[j]);
    lp0[j] = run0;
    lp1[j] = run1;
  }
  const int excl0 = c_shr1(c_prefix_min_ng(run0), kTop3);
  const int excl1 = c_shr1(c_prefix_min_ng(run1), kTop3);
  int prev0 = excl0, prev1 = excl1;
#pragma unroll
  for (int j = 0; j < (R + 7) / 8; ++j) codes0[j] = codes1[j] = 0;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const int pm0 = min(excl0, lp0[j]), pm1 = min(excl1, lp1[j]);
    const bool ss0 = prev0 < wv0[j], ss1 = prev1 < wv1[j];
    mm0[j] = (G3s[j] + pm0) >> 3;
    mm1[j] = (G3s[j] + pm1) >> 3;
    codes0[j >> 3] |= (ss0 ? (uint32_t)kSS : (uint32_t)(key0[j] & 7)) << (4 * (j & 7));
    codes1[j >> 3] |= (ss1 ? (uint32_t)kSS : (uint32_t)(key1[j] & 7)) << (4 * (j & 7));
    prev0 = pm0;
    prev1 = pm1;
  }
}

template <typename T, int R>
__global__ __launch_bounds__(64) void cube_rows_kernel(const TripleMeta *__restrict__ meta,
                                                       const int32_t *__restrict__ order, int n, int *counter,
                                                       const uint8_t *__restrict__ seq,
                                                       const DevTcm3 *__restrict__ tcm, uint8_t *dm_base,
                                                       T *scratch, int64_t scratch_per_wave,
                                                       int32_t *cost_out) {
  // [b][y]: x = cost(gap, b, y)*8 + S1, y = cost(a, b, y)*8 + S2, z = cost(a, gap, y)*8 + S3 for the
  // plane's s1 symbol a
  __shared__ int4 s_tab[32 * 32];
  constexpr int RS = 64 * R;   // ring row stride (elements)
  const int lane = c_lane();
  T *ring = scratch + (size_t)blockIdx.x * scratch_per_wave;
  const int gap = tcm->gap;
  for (int e = lane; e < 1024; e += 64) s_tab[e].x = (tcm->cost[gap * 1024 + e] << 3) | kS1;
  c_lds_sync();
  for (;;) {
    int kk = 0;
    if (lane == 0) kk = atomicAdd(counter, 1);
    kk = __builtin_amdgcn_readfirstlane(kk);
    if (kk >= n) break;
    const int t = __builtin_amdgcn_readfirstlane(order[kk]);
    const int64_t off1 = c_uniform64(meta[t].off1), off2 = c_uniform64(meta[t].off2),
                  off3 = c_uniform64(meta[t].off3), dm_off = c_uniform64(meta[t].dm_off);
    const int n1 = __builtin_amdgcn_readfirstlane(meta[t].l1), n2 = __builtin_amdgcn_readfirstlane(meta[t].l2),
              n3 = __builtin_amdgcn_readfirstlane(meta[t].l3),
              dstride = __builtin_amdgcn_readfirstlane(meta[t].dstride);
    // the triples with the most rows bound the launch: their waves issue first
    const int prio = (__builtin_amdgcn_readfirstlane(meta[t].flags) >> 1) & 3;
    if (prio >= 2) __builtin_amdgcn_s_setprio(2);
    else if (prio == 1) __builtin_amdgcn_s_setprio(1);
    else __builtin_amdgcn_s_setprio(0);
    const uint8_t *s1 = seq + off1, *s2 = seq + off2, *s3 = seq + off3;
    uint8_t *dm = dm_base + dm_off;
    const int W = n1 + n2;
    const int s30 = __builtin_amdgcn_readfirstlane((int)(s3[0] & 31));
    const bool act = lane * R < n3;   // lane holds at least one cell
    // per-cell symbols of s3 (inactive cells: the gap) and prefix sums of
    // cost(gap, gap, s3[k]) for k >= 1
    int yo[R], G3s[R];
    {
      int g[R], sum = 0;
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const int k = lane * R + j;
        const int y = k < n3 ? (int)(s3[k] & 31) : gap;
        yo[j] = y;
        g[j] = (k >= 1 && k < n3) ? (s_tab[gap * 32 + y].x >> 3) : 0;
        sum += g[j];
        G3s[j] = sum;
      }
      const int ex = c_prefix_sum(sum) - sum;
#pragma unroll
      for (int j = 0; j < R; ++j) G3s[j] = (G3s[j] + ex) << 3;
    }
    int mm[R];
    uint32_t codes[(R + 7) / 8];
    // plane 0, line 0: the self pass from 0
#pragma unroll
    for (int j = 0; j < R; ++j) mm[j] = G3s[j] >> 3;
#pragma unroll
    for (int j = 0; j < (R + 7) / 8; ++j) codes[j] = kSS * 0x11111111u;
    if (lane == 0) codes[0] = (codes[0] & ~0xfu) | kS2;   // the origin's code is never read
    if (act) {
      store_ring<T, R>(ring + lane * R, mm);
      store_codes<R>(dm + lane * (R / 2), codes);
    }
    int cur = 0;   // ring row of the row just written
    // plane 0, lines r >= 1 (the previous row is the one just computed)
    for (int r = 1; r < n2; ++r) {
      const int b = s2[r] & 31;
      const int4 *tab = s_tab + b * 32;
      const int cP1 = __builtin_amdgcn_readfirstlane(tab[s30].x) - (kS1 - kP1);
      int P[R];
#pragma unroll
      for (int j = 0; j < R; ++j) P[j] = mm[j];
      cube_row<R, 0>(P, P, P, G3s, yo, tab, 0, cP1, 0, mm, codes);
      cur = cur + 1 == W ? 0 : cur + 1;
      if (act) {
        store_ring<T, R>(ring + (size_t)cur * RS + lane * R, mm);
        store_codes<R>(dm + (size_t)r * dstride + lane * (R / 2), codes);
      }
    }
    int last = 0;   // cost: cell n3-1 of the last row
    {
      const int jl = (n3 - 1) % R;
      int v = 0;
#pragma unroll
      for (int j = 0; j < R; ++j) v = j == jl ? mm[j] : v;
      last = __builtin_amdgcn_readlane(v, (n3 - 1) / R);
    }
    if (n1 > 1) {
      // planes p >= 1 as one stream of rows f = p*n2 + r.  The inputs of row
      // g are loaded kAhead rows early (register slot (g - n2) % kAhead) when
      // every row they read is stored by then even if two rows are computed
      // together, i.e. g - (p+1) and g - (n2+p-1) are at most g - kAhead - 2;
      // other rows (the first planes) load their inputs when they are
      // computed.  A wave's own stores are seen by its later loads in program
      // order.
      RingRow<T, R> xs[kAhead], ps[kAhead];
      int bs[kAhead];
      bool pre[kAhead];
      int U[R], D[R];
#pragma unroll
      for (int j = 0; j < R; ++j) U[j] = D[j] = 0;
#pragma unroll
      for (int i = 0; i < kAhead; ++i) {
        xs[i].zero();
        ps[i].zero();
        bs[i] = 0;
        pre[i] = false;
      }
      const int total = n1 * n2;
      // row g's ring index and (plane, line), kAhead rows ahead of f
      int f = n2, p = 1, r = 0, cur_ring = n2;
      int g = f, gp = 1, gr = 0, g_ring = n2;
      auto issue = [&](int i) {
        // loads of row g into slot i when allowed, then advance g
        pre[i] = false;
        bs[i] = s2[gr];   // gr < n2 on every call: loaded unconditionally, used kAhead rows later
        if (g < total && gp >= kAhead + 1 && n2 + gp >= kAhead + 3) {
          int xi = g_ring - (n2 + gp - 1);
          xi += xi < 0 ? W : 0;
          if (act) xs[i].load(ring + (size_t)xi * RS + lane * R);
          if (gr >= 1) {
            int pi = g_ring - (gp + 1);
            pi += pi < 0 ? W : 0;
            if (act) ps[i].load(ring + (size_t)pi * RS + lane * R);
          }
          pre[i] = true;
        }
        ++g;
        g_ring = g_ring + 1 == W ? 0 : g_ring + 1;
        if (++gr == n2) { gr = 0; ++gp; }
      };
#pragma unroll
      for (int i = 0; i < kAhead; ++i) issue(i);
      int cP3 = 0;
      // the contents of slot i; true when they hold the row's inputs
      auto inputs = [&](int i, RingRow<T, R> &xc, RingRow<T, R> &pc, int &b) {
        xc = xs[i];
        pc = ps[i];
        b = bs[i];
        return pre[i];
      };
      auto load_now = [&](int pp, int rr, int rg, RingRow<T, R> &xc, RingRow<T, R> &pc, int &b) {
        int xi = rg - (n2 + pp - 1);
        xi += xi < 0 ? W : 0;
        if (act) xc.load(ring + (size_t)xi * RS + lane * R);
        if (rr >= 1) {
          int pi = rg - (pp + 1);
          pi += pi < 0 ? W : 0;
          if (act) pc.load(ring + (size_t)pi * RS + lane * R);
        }
        b = s2[rr];
      };
      auto store_row = [&](int rg, int ff) {
        if (act) {
          store_ring<T, R>(ring + (size_t)rg * RS + lane * R, mm);
          store_codes<R>(dm + (size_t)ff * dstride + lane * (R / 2), codes);
        }
      };
      auto advance = [&]() {
        ++f;
        cur_ring = cur_ring + 1 == W ? 0 : cur_ring + 1;
        if (++r == n2) { r = 0; ++p; }
      };
      auto single = [&](int i) {
        RingRow<T, R> xc, pc;
        int b;
        const bool had = inputs(i, xc, pc, b);
        issue(i);
        if (!had) {
          load_now(p, r, cur_ring, xc, pc, b);
          // complete these loads here, so that the rows computed from
          // prefetched slots do not wait for the loads just issued
          __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0)
        }
        b = __builtin_amdgcn_readfirstlane(b) & 31;
#pragma unroll
        for (int j = 0; j < R; ++j) {
          D[j] = U[j];
          U[j] = xc.get(j);
        }
        if (r == 0) {
          // plane table: cost(a, b, y) and cost(a, gap, y); all loads first
          const int a = __builtin_amdgcn_readfirstlane((int)(s1[p] & 31));
          int v[16];
#pragma unroll
          for (int q = 0; q < 16; ++q) v[q] = tcm->cost[a * 1024 + q * 64 + lane];
          const int z = (tcm->cost[(a * 32 + gap) * 32 + (lane & 31)] << 3) | kS3;
#pragma unroll
          for (int q = 0; q < 16; ++q) {
            s_tab[q * 64 + lane].y = (v[q] << 3) | kS2;
            s_tab[q * 64 + lane].z = z;
          }
          c_lds_sync();
          cP3 = __builtin_amdgcn_readfirstlane(s_tab[gap * 32 + s30].z) - (kS3 - kP3);
          cube_row<R, 1>(U, U, U, G3s, yo, s_tab + gap * 32, cP3, 0, 0, mm, codes);
        } else {
          const int4 *tab = s_tab + b * 32;
          const int4 c0 = tab[s30];
          const int cP1 = __builtin_amdgcn_readfirstlane(c0.x) - (kS1 - kP1);
          const int cP2 = __builtin_amdgcn_readfirstlane(c0.y) - (kS2 - kP2);
          int P[R];
#pragma unroll
          for (int j = 0; j < R; ++j) P[j] = pc.get(j);
          cube_row<R, 2>(U, D, P, G3s, yo, tab, cP3, cP1, cP2, mm, codes);
        }
        store_row(cur_ring, f);
        advance();
      };
      while (f < total) {
#pragma unroll
        for (int i = 0; i < kAhead; i += 2) {
          if (f + 1 < total && r >= 1 && r + 1 < n2) {
            // rows f and f+1: lines r and r+1 of plane p
            RingRow<T, R> xa, pa, xb, pb;
            int ba, bb;
            const int ring_b = cur_ring + 1 == W ? 0 : cur_ring + 1;
            const bool ha = inputs(i, xa, pa, ba);
            const bool hb = inputs(i + 1, xb, pb, bb);
            issue(i);
            issue(i + 1);
            if (!ha) load_now(p, r, cur_ring, xa, pa, ba);
            if (!hb) load_now(p, r + 1, ring_b, xb, pb, bb);
            if (!ha || !hb) __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0)
            ba = __builtin_amdgcn_readfirstlane(ba) & 31;
            bb = __builtin_amdgcn_readfirstlane(bb) & 31;
            int U1[R], P0[R], P1[R];
#pragma unroll
            for (int j = 0; j < R; ++j) {
              D[j] = U[j];
              U[j] = xa.get(j);
              P0[j] = pa.get(j);
              U1[j] = xb.get(j);
              P1[j] = pb.get(j);
            }
            const int4 *tab0 = s_tab + ba * 32, *tab1 = s_tab + bb * 32;
            const int4 c0 = tab0[s30], c1 = tab1[s30];
            const int cP1a = __builtin_amdgcn_readfirstlane(c0.x) - (kS1 - kP1);
            const int cP2a = __builtin_amdgcn_readfirstlane(c0.y) - (kS2 - kP2);
            const int cP1b = __builtin_amdgcn_readfirstlane(c1.x) - (kS1 - kP1);
            const int cP2b = __builtin_amdgcn_readfirstlane(c1.y) - (kS2 - kP2);
            int mm1[R];
            uint32_t codes1[(R + 7) / 8];
            cube_row_pair<R>(U, D, P0, tab0, cP1a, cP2a, U1, P1, tab1, cP1b, cP2b, G3s, yo, cP3, mm, codes, mm1,
                             codes1);
            store_row(cur_ring, f);
            advance();
#pragma unroll
            for (int j = 0; j < R; ++j) {
              D[j] = U[j];
              U[j] = U1[j];
              mm[j] = mm1[j];
            }
#pragma unroll
            for (int q = 0; q < (R + 7) / 8; ++q) codes[q] = codes1[q];
            store_row(cur_ring, f);
            advance();
          } else {
            if (f < total) single(i);
            if (f < total) single(i + 1);
          }
        }
      }
      const int jl = (n3 - 1) % R;
      int v = 0;
#pragma unroll
      for (int j = 0; j < R; ++j) v = j == jl ? mm[j] : v;
      last = __builtin_amdgcn_readlane(v, (n3 - 1) / R);
    }
    // every lane stores the same value (keeps the item loop scalar)
    cost_out[t] = last;
    c_wave_fence();
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
  const bool nib = m.flags & 1;   // e counts cells; a row holds dstride bytes
  const int64_t a_line = nib ? 2 * (int64_t)m.dstride : m.dstride, a_plane = (int64_t)l2 * a_line;
  int64_t e = ((int64_t)l1 * l2 - 1) * a_line + (l3 - 1);
  int i1 = l1 - 1, i2 = l2 - 1, i3 = l3 - 1, len = 0, bad = 0;
  int last[3] = {gap, gap, gap};
  // per move index: which sequences advance (bits 0-2) -- P3, P1, P2, S3, S1, S2, SS
  constexpr uint32_t kUse = 1u | (2u << 3) | (3u << 6) | (5u << 9) | (6u << 12) | (7u << 15) | (4u << 18);
  while (e > 0) {
    const int d = nib ? (dm[e >> 1] >> (4 * (e & 1))) & 15 : dm[e];
    if (d > kSS) { bad = 1; break; }
    const uint32_t u = (kUse >> (3 * d)) & 7u;
    const bool m1 = u & 1u, m2 = u & 2u, m3 = u & 4u;
    if ((m1 && i1 < 0) || (m2 && i2 < 0) || (m3 && i3 < 0)) { bad = 1; break; }
    const int x1 = m1 ? (int)s1[i1--] : gap, x2 = m2 ? (int)s2[i2--] : gap, x3 = m3 ? (int)s3[i3--] : gap;
    e -= (m1 ? a_plane : 0) + (m2 ? a_line : 0) + (m3 ? 1 : 0);
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
//
// Row classes: a triple with n3 <= 64*R (R = 4, 8, 12, 16) runs in the row
// kernel of that R; longer s3 fall back to cube_fill_kernel (64-cell chunks).
constexpr int kCubeClasses = 5;   // 0: chunked fallback, c >= 1: row kernel with R = 4c
struct TriplePlan {
  poyg_ctx *ctx;
  int n;
  std::vector<int32_t> order;
  std::vector<TripleMeta> meta;
  std::vector<std::pair<int, int>> batches;   // [begin, end) in `order`
  std::vector<int64_t> out_off;
  int64_t out_bytes = 0, cells = 0, dm_bytes = 0;
  int64_t scratch_bytes_per_wave = 0;
  int grid = 1;
  bool narrow = false;   // int16 ring rows
  std::vector<int> cls;  // row class per triple
  void *d_tcm = nullptr, *d_meta = nullptr, *d_order = nullptr, *d_dm = nullptr, *d_scratch = nullptr,
       *d_ctr = nullptr, *d_status = nullptr;
  hipStream_t stream = nullptr, side = nullptr, side2 = nullptr;
  hipStream_t last_stream = nullptr;   // the stream of the last run (its status is read there)
  hipEvent_t fork = nullptr, join = nullptr, join2 = nullptr;   // row classes run on three streams
  ~TriplePlan() {
    for (void *p : {d_tcm, d_meta, d_order, d_dm, d_scratch, d_ctr, d_status})
      if (p) (void)hipFree(p);
    for (hipEvent_t e : {fork, join, join2})
      if (e) (void)hipEventDestroy(e);
  }
};

static int cube_class(int l3) { return l3 > 1024 ? 0 : (l3 + 255) / 256; }

// Launch ri of a batch uses scratch region ri (launches run side by side).
template <typename T, int R>
static hipError_t launch_rows(const TriplePlan *P, int ri, const int32_t *ord, int nb, int *ctr,
                              const uint8_t *d_seq, int32_t *d_cost, hipStream_t s) {
  const int grid = std::max(1, std::min(P->grid, nb));
  T *scratch = (T *)((uint8_t *)P->d_scratch + (size_t)ri * P->grid * P->scratch_bytes_per_wave);
  hipLaunchKernelGGL((cube_rows_kernel<T, R>), dim3(grid), dim3(64), 0, s, (const TripleMeta *)P->d_meta, ord, nb,
                     ctr, d_seq, (const DevTcm3 *)P->d_tcm, (uint8_t *)P->d_dm, scratch,
                     (int64_t)(P->scratch_bytes_per_wave / (int64_t)sizeof(T)), d_cost);
  return hipGetLastError();
}

template <typename T>
static hipError_t launch_class(const TriplePlan *P, int c, int ri, const int32_t *ord, int nb, int *ctr,
                               const uint8_t *d_seq, int32_t *d_cost, hipStream_t s) {
  switch (c) {
    case 1: return launch_rows<T, 4>(P, ri, ord, nb, ctr, d_seq, d_cost, s);
    case 2: return launch_rows<T, 8>(P, ri, ord, nb, ctr, d_seq, d_cost, s);
    case 3: return launch_rows<T, 12>(P, ri, ord, nb, ctr, d_seq, d_cost, s);
    default: return launch_rows<T, 16>(P, ri, ord, nb, ctr, d_seq, d_cost, s);
  }
}

extern "C" {

int poyg_triple_plan_create(poyg_ctx *ctx, const poyg_cm3 *cm3, int32_t n, const int64_t *off_1,
                            const int32_t *len_1, const int64_t *off_2, const int32_t *len_2,
                            const int64_t *off_3, const int32_t *len_3, int device, int cus, void *stream,
                            void *side, void *side2, int64_t dm_budget, int waves_per_cu, void **out) {
  if (!cm3 || n < 0 || !out) return POYG_EINVAL;
  if (cm3->lcm != 5 || cm3->gap != kGap) {
    set_error("poyg_plan_triples: 5-bit DNA matrices only");
    return POYG_EUNSUPPORTED;
  }
  HIP_TRY3(hipSetDevice(device));
  DevTcm3 *dt = new DevTcm3;
  to_dev_tcm3(cm3, dt);
  // cost bounds of the masks a sequence can hold (1..31)
  int64_t maxc = 0, minc = 0;
  for (int a = 1; a < 32; ++a)
    for (int b = 1; b < 32; ++b)
      for (int c = 1; c < 32; ++c) {
        const int64_t v = dt->cost[(a * 32 + b) * 32 + c];
        maxc = std::max(maxc, v);
        minc = std::min(minc, v);
      }
  TriplePlan *P = new TriplePlan;
  P->ctx = ctx;
  P->n = n;
  P->stream = (hipStream_t)stream;
  P->side = (hipStream_t)side;
  P->side2 = (hipStream_t)side2;
  P->meta.resize(n);
  P->out_off.resize(n);
  P->cls.resize(n);
  std::vector<int64_t> cells(n), dmb(n);
  int64_t max_sum = 0;
  for (int k = 0; k < n; ++k) {
    TripleMeta &m = P->meta[k];
    if (len_1[k] < 1 || len_2[k] < 1 || len_3[k] < 1) {
      delete P;
      delete dt;
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
    max_sum = std::max(max_sum, (int64_t)m.l1 + m.l2 + m.l3);
  }
  // every cell value is at most (l1 + l2 + l3) * maxc (a path of single
  // moves), so the row kernel's keys (value * 8 + move) stay below 2^28 and,
  // for non-negative costs below 2^15, the ring rows fit int16
  const bool rows_ok = minc >= 0 && max_sum * std::max<int64_t>(maxc, 1) < (int64_t(1) << 24);
  P->narrow = rows_ok && max_sum * maxc < 32767;
  int64_t need = 0;
  for (int k = 0; k < n; ++k) {
    TripleMeta &m = P->meta[k];
    const int c = rows_ok ? cube_class(m.l3) : 0;
    P->cls[k] = c;
    // row kernel: 4-bit codes, rows padded to whole lanes' cells; fallback: bytes
    m.dstride = c ? (m.l3 + 4 * c - 1) / (4 * c) * (2 * c) : m.l3;
    m.flags = c ? 1 : 0;
    dmb[k] = (int64_t)m.l1 * m.l2 * m.dstride;
    const int64_t w = (int64_t)m.l1 + m.l2;
    const int64_t bytes = c ? w * 64 * 4 * c * (P->narrow ? 2 : 4)
                            : 4 * (w * m.l3 + m.l3 + 64);
    need = std::max(need, bytes);
  }
  P->scratch_bytes_per_wave = (need + 255) & ~(int64_t)255;
  // wave priority: the triples with the most rows (each is one wave's serial work)
  {
    int64_t max_rows = 1;
    for (int k = 0; k < n; ++k) max_rows = std::max(max_rows, (int64_t)P->meta[k].l1 * P->meta[k].l2);
    for (int k = 0; k < n; ++k) {
      const int64_t rows = (int64_t)P->meta[k].l1 * P->meta[k].l2;
      const int prio = rows * 2 >= max_rows ? 2 : rows * 4 >= max_rows ? 1 : 0;
      P->meta[k].flags |= prio << 1;
    }
  }
  P->order.resize(n);
  std::iota(P->order.begin(), P->order.end(), 0);
  std::stable_sort(P->order.begin(), P->order.end(), [&](int a, int b) {
    return P->cls[a] != P->cls[b] ? P->cls[a] > P->cls[b] : cells[a] > cells[b];
  });
  // split into batches whose direction cubes fit the budget
  int64_t cur = 0;
  int begin = 0;
  for (int q = 0; q < n; ++q) {
    const int k = P->order[q];
    if (cur > 0 && cur + dmb[k] > dm_budget) {
      P->batches.push_back({begin, q});
      begin = q;
      cur = 0;
    }
    P->meta[k].dm_off = cur;
    cur += (dmb[k] + 15) & ~(int64_t)15;
    P->dm_bytes = std::max(P->dm_bytes, cur);
  }
  if (n) P->batches.push_back({begin, n});
  const char *wv = getenv("POYG_CUBE_WAVES_PER_CU");
  const int waves = waves_per_cu > 0 ? waves_per_cu : (wv ? std::max(1, atoi(wv)) : 8);
  P->grid = std::max(1, std::min(n, cus * waves));
  hipError_t e = hipSuccess;
  auto A = [&](void **p, size_t b) { if (e == hipSuccess) e = hipMalloc(p, b ? b : 16); };
  A(&P->d_tcm, sizeof(DevTcm3));
  A(&P->d_meta, sizeof(TripleMeta) * (size_t)n);
  A(&P->d_order, sizeof(int32_t) * (size_t)n);
  A(&P->d_dm, (size_t)P->dm_bytes);
  // one scratch region per row class (a batch runs its classes side by side)
  int n_cls = 0;
  {
    bool seen[kCubeClasses] = {};
    for (int k = 0; k < n; ++k) seen[P->cls[k]] = true;
    for (bool x : seen) n_cls += x;
  }
  A(&P->d_scratch, (size_t)P->scratch_bytes_per_wave * P->grid * std::max(1, n_cls));
  A(&P->d_ctr, 64);
  A(&P->d_status, 64);
  for (hipEvent_t *ev : {&P->fork, &P->join, &P->join2})
    if (e == hipSuccess) e = hipEventCreateWithFlags(ev, hipEventDisableTiming);
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
  // the status counts the failed triples of the whole run (every batch)
  HIP_TRY3(launch_reset((int *)P->d_status, 4, s));
  P->last_stream = s;
  for (auto &b : P->batches) {
    const int nb = b.second - b.first;
    const int32_t *ord = (const int32_t *)P->d_order + b.first;
    if (timing) HIP_TRY3(hipEventRecord(ev[0], s));
    HIP_TRY3(launch_reset((int *)P->d_ctr, 16, s));
    // one launch per row class present in the batch (the order is sorted by
    // class), the classes side by side: the first on the third stream, the
    // last on the run stream, any others on the second
    std::vector<std::pair<int, int>> runs;
    for (int q = b.first; q < b.second;) {
      int q2 = q;
      while (q2 < b.second && P->cls[P->order[q2]] == P->cls[P->order[q]]) ++q2;
      runs.push_back({q, q2});
      q = q2;
    }
    const bool fan = runs.size() > 1 && P->side && P->side2;
    bool used_side = false, used_side2 = false;
    if (fan) {
      HIP_TRY3(hipEventRecord(P->fork, s));
      HIP_TRY3(hipStreamWaitEvent(P->side, P->fork, 0));
      HIP_TRY3(hipStreamWaitEvent(P->side2, P->fork, 0));
    }
    for (size_t ri = 0; ri < runs.size(); ++ri) {
      const int q = runs[ri].first, q2 = runs[ri].second;
      const int c = P->cls[P->order[q]];
      hipStream_t ls = s;
      if (fan && ri + 1 < runs.size()) {
        ls = ri == 0 ? P->side2 : P->side;
        (ri == 0 ? used_side2 : used_side) = true;
      }
      const int32_t *o2 = (const int32_t *)P->d_order + q;
      int *ctr = (int *)P->d_ctr + ri;
      if (c == 0) {
        const int grid = std::max(1, std::min(P->grid, q2 - q));
        hipLaunchKernelGGL(cube_fill_kernel, dim3(grid), dim3(64), 0, ls, (const TripleMeta *)P->d_meta, o2,
                           q2 - q, ctr, d_seq, (const DevTcm3 *)P->d_tcm, (uint8_t *)P->d_dm,
                           (int *)P->d_scratch + (size_t)ri * P->grid * (P->scratch_bytes_per_wave / 4),
                           P->scratch_bytes_per_wave / 4, d_cost);
        HIP_TRY3(hipGetLastError());
      } else if (P->narrow) {
        HIP_TRY3(launch_class<int16_t>(P, c, ri, o2, q2 - q, ctr, d_seq, d_cost, ls));
      } else {
        HIP_TRY3(launch_class<int32_t>(P, c, ri, o2, q2 - q, ctr, d_seq, d_cost, ls));
      }
    }
    if (used_side) {
      HIP_TRY3(hipEventRecord(P->join, P->side));
      HIP_TRY3(hipStreamWaitEvent(s, P->join, 0));
    }
    if (used_side2) {
      HIP_TRY3(hipEventRecord(P->join2, P->side2));
      HIP_TRY3(hipStreamWaitEvent(s, P->join2, 0));
    }
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

// Failed triples of the last run, read on the stream that run was queued on.
int poyg_triple_plan_status(void *pp, int32_t *bad) {
  TriplePlan *P = (TriplePlan *)pp;
  *bad = 0;
  if (!P->last_stream) return POYG_OK;
  HIP_TRY3(hipMemcpyAsync(bad, P->d_status, sizeof(int32_t), hipMemcpyDeviceToHost, P->last_stream));
  HIP_TRY3(hipStreamSynchronize(P->last_stream));
  return POYG_OK;
}

}  // extern "C"
