// Powell's three-way alignment with affine gaps (powell_3D_align,
// src/ukkCommon.c:110-151): the reference's check-point Ukkonen search
// (src/ukk.checkp.c) as level-synchronous device passes.
//
// A level is one cost d of one search range (doUkkInLimits): one thread per
// (diagonal, state) over the diagonals within the range's reach at cost d
// (16 consecutive lanes per diagonal).  A cell's value is calcUkk's
// (src/ukk.checkp.c:608-781) over the levels already stored: the state's
// neighbour diagonal at every previous state (the cost without the second
// mismatch when whichCharCost says 1, else one mismatch lower for states
// with a second cost), the same state one cost lower, and for the all-match
// state the furthest state of this level extended along equal characters;
// a candidate wins only when strictly further, in that order.  The winner's
// code is stored beside the dist, so the "from" records the reference keeps
// (coordinates in a base case; the check-point cell above the band) are
// walks over the stored levels.
//
// The first pass also finds the reference's check-point cost: the first
// cost d after whose top-level call the cells evaluated so far reach |A| / 2
// (src/ukk.checkp.c:152-170).  The cells Ukk evaluates up to that call are
// those some chain of calls reaches from the end diagonal's all-match cell
// at a cost of at most d: ukk_esweep_kernel gives every cell the least such
// d (its callers are the cell one cost higher, the all-match cell of its
// level, and the cells whose transition lowers the cost onto it), from the
// top level down.
#include <hip/hip_runtime.h>

#include <climits>
#include <mutex>

#include "poyg_internal.h"
#include "poyg_ukk.h"

namespace poyg {

// The model tables of setup (src/ukkCommon.c:247-330).
void uk_model(int mm, int go, int ge, UkModel *M) {
  int ns = 0;
  int st[kUkStates][3];
  for (int s = 0; s < 27; ++s) {
    const int t[3] = {s % 3, (s / 3) % 3, (s / 9) % 3};   // match 0, del 1, ins 2
    int nm = 0, ni = 0;
    for (int q = 0; q < 3; ++q) {
      nm += t[q] == 0;
      ni += t[q] == 2;
    }
    if (nm == 0 || ni > 1) continue;   // at least one match, at most one insert
    for (int q = 0; q < 3; ++q) st[ns][q] = t[q];
    ++ns;
  }
  M->ns = ns;
  M->mis = mm;
  M->go = go;
  M->ge = ge;
  for (int s = 0; s < ns; ++s) {
    int nm = 0, nd = 0, ni = 0;
    for (int q = 0; q < 3; ++q) {
      nm += st[s][q] == 0;
      nd += st[s][q] == 1;
      ni += st[s][q] == 2;
    }
    // no insert: the matching sequences advance; one insert: only it does
    int nb = 0;
    for (int q = 0; q < 3; ++q) nb |= (ni ? st[s][q] == 2 : st[s][q] == 0) << q;
    M->nb[s] = nb;
    if (ni) {
      M->cont[s] = ge;
      M->second[s] = 0;
    } else if (nm == 3) {
      M->cont[s] = mm;
      M->second[s] = 1;
    } else if (nd == 1) {
      M->cont[s] = ge;
      M->second[s] = 1;
    } else {
      M->cont[s] = 2 * ge;
      M->second[s] = 0;
    }
  }
  int max_step = 0;
  for (int a = 0; a < ns; ++a)
    for (int b = 0; b < ns; ++b) {
      int c = 0, nm = 0;
      for (int q = 0; q < 3; ++q) {
        if (st[b][q] != 0 && st[b][q] != st[a][q]) c += go;
        nm += st[b][q] == 0;
      }
      M->trans[a][b] = c;
      max_step = std::max(max_step, c + M->cont[b] + mm * (nm - 1));
    }
  M->max_step = max_step;
}

namespace {

__device__ __forceinline__ bool uk_ok(int a, int da, int end) {   // okIndex
  if (a < 0) return false;
  return da ? a < end : a <= end;
}

__device__ __forceinline__ int uk_chars(int a, int b, int c) {   // whichCharCost
  if (a == b && a == c) return 0;
  return (a == b || a == c || b == c) ? 1 : 2;
}

// withinMatrix: the range's start can reach diagonal (ab, ac) at cost c.
__device__ __forceinline__ bool uk_within(const UkRange &r, const UkModel &M, int ab, int ac, int c) {
  if (c < 0) return false;
  int x = abs(r.sab - ab), y = abs(r.sac - ac), z = abs((r.sac - r.sab) - (ac - ab));
  // the two smallest of three
  const int g = min(x, min(y, z));
  const int h = max(min(x, y), min(max(x, y), z));
  const int cheapest = r.sstate == 0 ? (g ? M.go + g * M.ge : 0) + (h ? M.go + h * M.ge : 0) : (g + h) * M.ge;
  return cheapest + r.scost <= c;
}

// A stored cell's dist; -INFINITY outside the stored levels and boxes.
__device__ __forceinline__ int uk_read(const UkLevel *lev, const uint32_t *U, const UkRange &r, int ab, int ac,
                                       int c, int s) {
  if (c < r.scost) return -kUkInf;
  const UkLevel L = lev[c - r.scost];
  const int i = ab - L.ab_lo, j = ac - L.ac_lo;
  if (i < 0 || i >= L.nab || j < 0 || j >= L.nac) return -kUkInf;
  const uint32_t w = U[L.base + ((int64_t)i * L.nac + j) * kUkStates + s] & 0xffffu;
  return w ? (int)w - 1 : -kUkInf;
}

__device__ __forceinline__ int64_t uk_index(const UkLevel &L, int ab, int ac, int s) {
  return L.base + ((int64_t)(ab - L.ab_lo) * L.nac + (ac - L.ac_lo)) * kUkStates + s;
}

__device__ __forceinline__ bool uk_in_box(const UkLevel &L, int ab, int ac) {
  return ab >= L.ab_lo && ab < L.ab_lo + L.nab && ac >= L.ac_lo && ac < L.ac_lo + L.nac;
}

__device__ __forceinline__ void uk_load_model(UkModel &sm, const UkModel *M) {
  for (int e = threadIdx.x; e < (int)(sizeof(UkModel) / 4); e += blockDim.x) ((int *)&sm)[e] = ((const int *)M)[e];
  __syncthreads();
}

}  // namespace

// Level boxes of levels [c0, c1] of a range (one thread; bases continue the
// previous level's); reset >= 0 first resets the walk record for the range
// (1: for a whole search).
__device__ void ukk_reset(UkWalk *w, int first);

__global__ void ukk_geom_kernel(UkRange r, int ge, int c0, int c1, UkLevel *lev, UkWalk *w, int reset) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (reset >= 0) ukk_reset(w, reset);
  int64_t base = 0;
  if (c0 > r.scost) {
    const UkLevel &p = lev[c0 - 1 - r.scost];
    base = p.base + (int64_t)p.nab * p.nac * kUkStates;
  }
  for (int c = c0; c <= c1; ++c) {
    lev[c - r.scost] = uk_level_box(r, ge, c, base);
    base += (int64_t)lev[c - r.scost].nab * lev[c - r.scost].nac * kUkStates;
  }
}

// One cost level of a range.  target_state >= 0: the level at which the end
// cell's state reaches target_dist goes to w->reached (the loops of
// doUkkInLimits); target_state < 0: the best state does (doUkk's loop).
__global__ __launch_bounds__(256) void ukk_level_kernel(const UkModel *__restrict__ Mg, UkRange r, int c,
                                                        const UkLevel *__restrict__ lev, uint32_t *U,
                                                        const char *__restrict__ A, const char *__restrict__ B,
                                                        const char *__restrict__ C, int target_state,
                                                        int target_dist, UkWalk *w) {
  __shared__ UkModel M;
  __shared__ int s_dist[256];
  uk_load_model(M, Mg);
  const UkLevel L = lev[c - r.scost];
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t diag = g >> 4;
  const int s = (int)(g & 15);
  const bool act = diag < (int64_t)L.nab * L.nac;
  int ab = 0, ac = 0;
  bool in = false;
  if (act) {
    ab = L.ab_lo + (int)(diag / L.nac);
    ac = L.ac_lo + (int)(diag % L.nac);
    in = uk_within(r, M, ab, ac, c);
  }
  const bool preset = in && c == r.scost && ab == r.sab && ac == r.sac && s == r.sstate;
  int best = -kUkInf, code = kUkNone;
  if (preset) {
    best = r.sdist;
    code = kUkStart;
  } else if (in) {
    const int nb = M.nb[s];
    const int da = nb & 1, db = (nb >> 1) & 1, dc = (nb >> 2) & 1;
    const int ab1 = ab - da + db, ac1 = ac - da + dc;
    if (ab1 >= -r.endB && ab1 <= r.endA && ac1 >= -r.endC && ac1 <= r.endA) {
      for (int fs = 0; fs < M.ns; ++fs) {
        const int cost = c - M.trans[fs][s] - M.cont[s];
        int dist = -kUkInf, v = 0;
        const int a1 = uk_read(lev, U, r, ab1, ac1, cost, fs);
        if (uk_ok(a1, da, r.endA) && uk_ok(a1 - ab1, db, r.endB) && uk_ok(a1 - ac1, dc, r.endC) &&
            uk_chars(da ? A[a1] : '-', db ? B[a1 - ab1] : '-', dc ? C[a1 - ac1] : '-') == 1) {
          dist = a1 + da;
        } else {
          if (!M.second[s]) continue;
          const int a2 = uk_read(lev, U, r, ab1, ac1, cost - M.mis, fs);
          if (uk_ok(a2, da, r.endA) && uk_ok(a2 - ab1, db, r.endB) && uk_ok(a2 - ac1, dc, r.endC)) {
            dist = a2 + da;
            v = 1;
          }
        }
        if (best < dist) {
          best = dist;
          code = 2 * fs + v;
        }
      }
    }
    const int d1 = uk_read(lev, U, r, ab, ac, c - 1, s);
    if (uk_ok(d1, 0, r.endA) && uk_ok(d1 - ab, 0, r.endB) && uk_ok(d1 - ac, 0, r.endC) && best < d1) {
      best = d1;
      code = kUkLower;
    }
  }
  s_dist[threadIdx.x] = best;
  __syncthreads();
  if (in && s == 0) {
    int top = -kUkInf;
    if (!preset) {   // extend the furthest state of this level along a run of equal characters
      int dist = -kUkInf, from = -1;
      for (int s2 = 0; s2 < M.ns; ++s2) {
        const int x = s2 == 0 ? best : s_dist[threadIdx.x + s2];
        if (x > dist) {
          dist = x;
          from = s2;
        }
      }
      while (uk_ok(dist, 1, r.endA) && uk_ok(dist - ab, 1, r.endB) && uk_ok(dist - ac, 1, r.endC) &&
             A[dist] == B[dist - ab] && A[dist] == C[dist - ac])
        ++dist;
      if (dist > best) {
        best = dist;
        if (from != 0) code = kUkExt + from;
      }
    }
    if (target_state < 0 && c > r.scost && ab == r.fab && ac == r.fac) {   // doUkk's loop starts at cost 1
      top = best;
      for (int s2 = 1; s2 < M.ns; ++s2) top = max(top, s_dist[threadIdx.x + s2]);
      if (top >= target_dist) atomicMin(&w->reached, c);
    }
  }
  if (in && target_state >= 0 && s == target_state && ab == r.fab && ac == r.fac && best >= target_dist)
    atomicMin(&w->reached, c);
  if (act) U[L.base + diag * kUkStates + s] = ((uint32_t)code << 16) | (best == -kUkInf ? 0u : (uint32_t)(best + 1));
}

// First pass, one level from the top down: every cell's first top-level
// cost at which the recursion evaluates it (E), and the least such cost over
// the cells that reach |A| / 2 (w->dstar).  E keeps ring_n levels.
__global__ __launch_bounds__(256) void ukk_esweep_kernel(const UkModel *__restrict__ Mg, UkRange r, int c,
                                                         int final_cost, const UkLevel *__restrict__ lev,
                                                         const uint32_t *__restrict__ U, int *E, int64_t ring_words,
                                                         int ring_n, int half, UkWalk *w) {
  __shared__ UkModel M;
  __shared__ int s_e[256];
  uk_load_model(M, Mg);
  const UkLevel L = lev[c - r.scost];
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t diag = g >> 4;
  const int s = (int)(g & 15);
  const bool act = diag < (int64_t)L.nab * L.nac;
  int ab = 0, ac = 0;
  bool in = false;
  if (act) {
    ab = L.ab_lo + (int)(diag / L.nac);
    ac = L.ac_lo + (int)(diag % L.nac);
    in = uk_within(r, M, ab, ac, c);
  }
  auto eread = [&](int xab, int xac, int xc, int xs) -> int {
    const UkLevel &X = lev[xc - r.scost];
    return E[(int64_t)(xc % ring_n) * ring_words + (uk_index(X, xab, xac, xs) - X.base)];
  };
  // the preset start cell is never evaluated and calls nothing
  const bool at_start = c == r.scost && ab == r.sab && ac == r.sac;
  int e = INT_MAX;
  if (in) {
    if (ab == r.fab && ac == r.fac && s == 0 && c >= 1) e = c;   // doUkk's call at cost c
    if (c + 1 <= final_cost) e = min(e, eread(ab, ac, c + 1, s));
    for (int ts = 0; ts < M.ns; ++ts) {   // callers whose transition lowers the cost onto this cell
      const int nb = M.nb[ts];
      const int da = nb & 1, db = (nb >> 1) & 1, dc = (nb >> 2) & 1;
      const int yab = ab + da - db, yac = ac + da - dc;
      const int c2 = c + M.trans[s][ts] + M.cont[ts];
      if (c2 > final_cost) continue;
      if (!uk_in_box(lev[c2 - r.scost], yab, yac) || !uk_within(r, M, yab, yac, c2)) continue;
      if (c2 == r.scost && yab == r.sab && yac == r.sac && ts == r.sstate) continue;
      e = min(e, eread(yab, yac, c2, ts));
    }
  }
  s_e[threadIdx.x] = e;
  __syncthreads();
  // the all-match cell of this level evaluates every other state of it
  if (in && s != 0 && !(at_start && r.sstate == 0)) e = min(e, s_e[threadIdx.x - s]);
  if (act) E[(int64_t)(c % ring_n) * ring_words + diag * kUkStates + s] = e;
  if (in && e != INT_MAX && !(at_start && s == r.sstate)) {
    const uint32_t x = U[L.base + diag * kUkStates + s] & 0xffffu;
    if (x && (int)x - 1 >= half) atomicMin(&w->dstar, e);
  }
}

namespace {

// The candidate a cell's code names.
__device__ __forceinline__ void uk_pred(const UkModel &M, int code, int &ab, int &ac, int &c, int &s) {
  if (code < kUkLower) {
    const int fs = code >> 1, v = code & 1;
    const int nb = M.nb[s];
    const int da = nb & 1, db = (nb >> 1) & 1, dc = (nb >> 2) & 1;
    c = c - M.trans[fs][s] - M.cont[s] - v * M.mis;
    ab = ab - da + db;
    ac = ac - da + dc;
    s = fs;
  } else if (code == kUkLower) {
    c -= 1;
  } else {
    s = code - kUkExt;
  }
}

__device__ __forceinline__ uint32_t uk_word(const UkLevel *lev, const uint32_t *U, const UkRange &r, int ab, int ac,
                                            int c, int s) {
  if (c < r.scost) return (uint32_t)kUkNone << 16;
  const UkLevel L = lev[c - r.scost];
  if (!uk_in_box(L, ab, ac)) return (uint32_t)kUkNone << 16;
  return U[uk_index(L, ab, ac, s)];
}

}  // namespace

// The check-point split of a range (getSplitRecurse, src/ukk.checkp.c:
// 324-368): from the end cell at the level its loop stopped at, back along
// the winning candidates to the first cell inside [cpcost, cpcost + cpw)
// -- the end cell's "from" record.  first_pass: the end state is the first
// best one and cpcost = dstar + 1 (src/ukk.checkp.c:160-187).  One thread.
__global__ void ukk_walk_kernel(const UkModel *__restrict__ M, UkRange r, int cpcost, int cpw,
                                const UkLevel *__restrict__ lev, const uint32_t *__restrict__ U, int first_pass,
                                UkWalk *w) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  w->f_cost = -1;
  const int end = w->reached;
  if (end == INT_MAX || end < r.scost) {
    w->error = 1;
    return;
  }
  int st = r.fstate;
  if (first_pass) {
    int best = -kUkInf;
    st = -1;
    for (int s = 0; s < M->ns; ++s) {
      const uint32_t x = uk_word(lev, U, r, r.fab, r.fac, end, s) & 0xffffu;
      const int d = x ? (int)x - 1 : -kUkInf;
      if (d > best) {
        best = d;
        st = s;
      }
    }
    w->fstate = st;
    if (w->dstar == INT_MAX) {
      w->error = 2;
      return;
    }
    cpcost = w->dstar + 1;
    if (st < 0) {
      w->error = 3;
      return;
    }
  }
  int ab = r.fab, ac = r.fac, c = end, s = st;
  for (int steps = 0; steps < (1 << 26); ++steps) {
    if (c >= cpcost && c < cpcost + cpw) {
      const uint32_t x = uk_word(lev, U, r, ab, ac, c, s);
      w->f_ab = ab;
      w->f_ac = ac;
      w->f_cost = c;
      w->f_state = s;
      w->f_dist = (int)(x & 0xffffu) - 1;
      return;
    }
    if (c < cpcost) return;   // no check point: the range is searched whole
    const int code = (int)(uk_word(lev, U, r, ab, ac, c, s) >> 16);
    if (code == kUkNone || code == kUkStart) return;
    uk_pred(*M, code, ab, ac, c, s);
  }
  w->error = 4;
}

// The traceback of a base-case range (traceBack, src/ukk.checkp.c:370-454):
// from the end cell back to the start cell, a run of all-match columns down
// to each candidate's dist, then one column of the cell's state; columns are
// appended (last column first) at w->ncols.  One thread.
__global__ void ukk_trace_kernel(const UkModel *__restrict__ M, UkRange r, const UkLevel *__restrict__ lev,
                                 const uint32_t *__restrict__ U, const char *__restrict__ A,
                                 const char *__restrict__ B, const char *__restrict__ C, char *cols, int cap,
                                 UkWalk *w) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int end = w->reached;
  if (end == INT_MAX || end < r.scost) {
    w->error = 5;
    return;
  }
  int n = w->ncols;
  int ab = r.fab, ac = r.fac, c = end, s = r.fstate;
  auto emit = [&](char x, char y, char z) {
    if (n >= cap) return false;
    cols[3 * (int64_t)n] = x;
    cols[3 * (int64_t)n + 1] = y;
    cols[3 * (int64_t)n + 2] = z;
    ++n;
    return true;
  };
  while (ab != r.sab || ac != r.sac || c != r.scost || s != r.sstate) {
    const uint32_t x = uk_word(lev, U, r, ab, ac, c, s);
    const int code = (int)(x >> 16);
    if (code == kUkNone || code == kUkStart || !(x & 0xffffu)) {
      w->error = 6;
      w->ncols = n;
      return;
    }
    int a = (int)(x & 0xffffu) - 1;
    int nab = ab, nac = ac, nc = c, ns = s;
    uk_pred(*M, code, nab, nac, nc, ns);
    const uint32_t y = uk_word(lev, U, r, nab, nac, nc, ns) & 0xffffu;
    if (!y) {
      w->error = 7;
      w->ncols = n;
      return;
    }
    const int a1 = (int)y - 1;
    int b = a - ab, cc = a - ac;
    const int b1 = a1 - nab, c1 = a1 - nac;
    bool ok = true;
    while (ok && a > a1 && b > b1 && cc > c1) {
      --a;
      --b;
      --cc;
      ok = emit(A[a], B[b], C[cc]);
    }
    if (ok && (a != a1 || b != b1 || cc != c1)) {
      const char p = a > a1 ? A[--a] : '-';
      const char q = b > b1 ? B[--b] : '-';
      const char t = cc > c1 ? C[--cc] : '-';
      ok = emit(p, q, t);
    }
    if (!ok || a != a1 || b != b1 || cc != c1) {
      w->error = 8;
      w->ncols = n;
      return;
    }
    ab = nab;
    ac = nac;
    c = nc;
    s = ns;
  }
  w->ncols = n;
}

// A range's walk record before its levels run.
__device__ void ukk_reset(UkWalk *w, int first) {
  w->reached = INT_MAX;
  w->f_cost = -1;
  if (first) {
    w->dstar = INT_MAX;
    w->ncols = 0;
    w->error = 0;
    w->fstate = -1;
  }
}


// Searches may run on several host threads at once (a context each); their
// launches go through one lock, so that the launch path sees one thread at a
// time (the waits, where the time goes, stay concurrent).
static std::mutex g_ukk_launch;


hipError_t launch_ukk_geom(const UkRange &r, int ge, int c0, int c1, UkLevel *lev, UkWalk *w, int reset,
                           hipStream_t s) {
  std::lock_guard<std::mutex> lock(g_ukk_launch);
  hipLaunchKernelGGL(ukk_geom_kernel, dim3(1), dim3(64), 0, s, r, ge, c0, c1, lev, w, reset);
  return hipGetLastError();
}

hipError_t launch_ukk_level(const UkModel *M, const UkRange &r, int c, int64_t diags, const UkLevel *lev, uint32_t *U,
                            const char *A, const char *B, const char *C, int target_state, int target_dist,
                            UkWalk *w, hipStream_t s) {
  std::lock_guard<std::mutex> lock(g_ukk_launch);
  if (diags <= 0) return hipSuccess;
  const int64_t blocks = (diags * kUkStates + 255) / 256;
  hipLaunchKernelGGL(ukk_level_kernel, dim3((unsigned)blocks), dim3(256), 0, s, M, r, c, lev, U, A, B, C,
                     target_state, target_dist, w);
  return hipGetLastError();
}

hipError_t launch_ukk_esweep(const UkModel *M, const UkRange &r, int c, int final_cost, int64_t diags,
                             const UkLevel *lev, const uint32_t *U, int *E, int64_t ring_words, int ring_n, int half,
                             UkWalk *w, hipStream_t s) {
  std::lock_guard<std::mutex> lock(g_ukk_launch);
  if (diags <= 0) return hipSuccess;
  const int64_t blocks = (diags * kUkStates + 255) / 256;
  hipLaunchKernelGGL(ukk_esweep_kernel, dim3((unsigned)blocks), dim3(256), 0, s, M, r, c, final_cost, lev, U, E,
                     ring_words, ring_n, half, w);
  return hipGetLastError();
}

hipError_t launch_ukk_walk(const UkModel *M, const UkRange &r, int cpcost, int cpw, const UkLevel *lev,
                           const uint32_t *U, int first_pass, UkWalk *w, hipStream_t s) {
  std::lock_guard<std::mutex> lock(g_ukk_launch);
  hipLaunchKernelGGL(ukk_walk_kernel, dim3(1), dim3(64), 0, s, M, r, cpcost, cpw, lev, U, first_pass, w);
  return hipGetLastError();
}

hipError_t launch_ukk_trace(const UkModel *M, const UkRange &r, const UkLevel *lev, const uint32_t *U, const char *A,
                            const char *B, const char *C, char *cols, int cap, UkWalk *w, hipStream_t s) {
  std::lock_guard<std::mutex> lock(g_ukk_launch);
  hipLaunchKernelGGL(ukk_trace_kernel, dim3(1), dim3(64), 0, s, M, r, lev, U, A, B, C, cols, cap, w);
  return hipGetLastError();
}

}  // namespace poyg
