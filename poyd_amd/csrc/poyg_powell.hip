// Powell's three-way alignment with affine gaps (powell_3D_align,
// src/ukkCommon.c:110-151; the model of src/ukkCommon.c:157-330, 380-445).
//
// The reference searches the alignment lattice Ukkonen-style for the least
// cost of any alignment under its model.  Here the whole lattice is solved
// exactly: a cell (i, j, k) holds one value per model state, and every
// state's predecessor is the cell its move left, on the anti-diagonal plane
// i+j+k = t-1, t-2 or t-3.  One launch computes one plane (a thread per
// cell, four rotating plane buffers), recording per cell and state the
// first least previous state; one thread then walks back from the first
// least final state.
#include <hip/hip_runtime.h>

#include "poyg_internal.h"

namespace poyg {

constexpr int kPwStates = 16;
constexpr int kPwInf = 0x3fffffff;

// Model states (src/ukkCommon.c:252-330): per sequence match 0 / del 1 /
// ins 2, enumerated as s = t0 + 3 t1 + 9 t2 and kept when at least one
// sequence matches and at most one inserts.
struct PwTables {
  int ns;
  int st[kPwStates][3];
  int adv[kPwStates][3];   // sequences the state consumes a character of
  int nins[kPwStates], ndel[kPwStates];
  int trans[kPwStates][kPwStates];   // entering state b from a: go per new del/ins
};

void pw_tables(int go, PwTables *T) {
  int ns = 0;
  for (int s = 0; s < 27; ++s) {
    const int st[3] = {s % 3, (s / 3) % 3, (s / 9) % 3};
    int nm = 0, nd = 0, ni = 0;
    for (int q = 0; q < 3; ++q) {
      nm += st[q] == 0;
      nd += st[q] == 1;
      ni += st[q] == 2;
    }
    if (nm == 0 || ni > 1) continue;
    for (int q = 0; q < 3; ++q) {
      T->st[ns][q] = st[q];
      // no insert: the matching sequences advance; one insert: only it does
      T->adv[ns][q] = ni == 0 ? st[q] == 0 : st[q] == 2;
    }
    T->nins[ns] = ni;
    T->ndel[ns] = nd;
    ++ns;
  }
  T->ns = ns;
  for (int a = 0; a < ns; ++a)
    for (int b = 0; b < ns; ++b) {
      int c = 0;
      for (int q = 0; q < 3; ++q)
        if (T->st[b][q] != 0 && T->st[b][q] != T->st[a][q]) c += go;
      T->trans[a][b] = c;
    }
}

// Column cost of state t consuming up to (i, j, k) (alignmentCost,
// src/ukkCommon.c:407-440): ge for an insert column, else ge per del plus mm
// per mismatch among the matched characters (whichCharCost for three).
__device__ __forceinline__ int pw_step(const PwTables &T, int t, int c0, int c1, int c2, int mm, int ge) {
  if (T.nins[t]) return ge;
  int cost = ge * T.ndel[t];
  int ch[3], n = 0;
  if (T.st[t][0] == 0) ch[n++] = c0;
  if (T.st[t][1] == 0) ch[n++] = c1;
  if (T.st[t][2] == 0) ch[n++] = c2;
  if (n == 3) {
    const int diff = (ch[0] == ch[1] && ch[0] == ch[2]) ? 0
                     : (ch[0] == ch[1] || ch[0] == ch[2] || ch[1] == ch[2]) ? 1 : 2;
    cost += mm * diff;
  } else if (n == 2) {
    cost += mm * (ch[0] != ch[1]);
  }
  return cost;
}

__global__ __launch_bounds__(256) void powell_plane_kernel(int t, int n1, int n2, int n3,
                                                           const uint8_t *__restrict__ x1,
                                                           const uint8_t *__restrict__ x2,
                                                           const uint8_t *__restrict__ x3, int *V,
                                                           uint8_t *from, const PwTables *__restrict__ tab,
                                                           int mm, int ge) {
  __shared__ PwTables T;
  for (int e = threadIdx.x; e < (int)(sizeof(PwTables) / 4); e += blockDim.x)
    ((int *)&T)[e] = ((const int *)tab)[e];
  __syncthreads();
  const int d2 = n2 + 1, d3 = n3 + 1;
  const int64_t plane_cells = (int64_t)(n1 + 1) * d2;
  const int64_t cell = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= plane_cells) return;
  const int i = (int)(cell / d2), j = (int)(cell % d2), k = t - i - j;
  if (k < 0 || k > n3) return;
  const int c0 = i ? x1[i - 1] : 0, c1 = j ? x2[j - 1] : 0, c2 = k ? x3[k - 1] : 0;
  int *Vt = V + ((int64_t)(t & 3) * plane_cells + cell) * kPwStates;
  uint8_t *ft = from + (cell * d3 + k) * kPwStates;
  for (int s2 = 0; s2 < T.ns; ++s2) {
    int best = kPwInf, arg = 0;
    if (t == 0) {
      best = s2 == 0 ? 0 : kPwInf;   // start in the all-match state
    } else {
      const int a0 = T.adv[s2][0], a1 = T.adv[s2][1], a2 = T.adv[s2][2];
      const int pi = i - a0, pj = j - a1, pk = k - a2;
      if (pi >= 0 && pj >= 0 && pk >= 0) {
        const int dt = a0 + a1 + a2;
        const int *Vp = V + ((int64_t)((t - dt) & 3) * plane_cells + (int64_t)pi * d2 + pj) * kPwStates;
        for (int s1 = 0; s1 < T.ns; ++s1) {
          const int v = Vp[s1];
          if (v >= kPwInf) continue;
          const int w = v + T.trans[s1][s2];
          if (w < best) {
            best = w;
            arg = s1;
          }
        }
        if (best < kPwInf) best += pw_step(T, s2, c0, c1, c2, mm, ge);
      }
    }
    Vt[s2] = best;
    ft[s2] = (uint8_t)arg;
  }
}

// One thread: the first least final state, then the recorded previous
// states back to the origin.  Columns are written right-aligned into three
// rows of cap bytes (masks, gap 16, the leading gap first); res = {length
// with the leading gap, cost}.
__global__ void powell_trace_kernel(int n1, int n2, int n3, const int *V, const uint8_t *from,
                                    const PwTables *__restrict__ tab, const uint8_t *x1, const uint8_t *x2,
                                    const uint8_t *x3, uint8_t *out, int cap, int *res) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const PwTables &T = *tab;
  const int d2 = n2 + 1, d3 = n3 + 1;
  const int64_t plane_cells = (int64_t)(n1 + 1) * d2;
  const int tEnd = n1 + n2 + n3;
  const int *Ve = V + ((int64_t)(tEnd & 3) * plane_cells + (int64_t)n1 * d2 + n2) * kPwStates;
  int cost = kPwInf, st = 0;
  for (int s = 0; s < T.ns; ++s)
    if (Ve[s] < cost) {
      cost = Ve[s];
      st = s;
    }
  int i = n1, j = n2, k = n3, pos = cap;
  const uint8_t *x[3] = {x1, x2, x3};
  while ((i > 0 || j > 0 || k > 0) && pos > 1) {
    if (T.adv[st][0] > i || T.adv[st][1] > j || T.adv[st][2] > k) {
      cost = -1;   // not reached: every recorded state of a finite path moves inside the lattice
      break;
    }
    --pos;
    const int at[3] = {i, j, k};
    for (int q = 0; q < 3; ++q) {
      const int ch = T.adv[st][q] ? x[q][at[q] - 1] : 0;
      out[q * cap + pos] = (uint8_t)(ch == 'A' ? 1 : ch == 'C' ? 2 : ch == 'G' ? 4 : ch == 'T' ? 8 : 16);
    }
    const int prev = from[(((int64_t)i * d2 + j) * d3 + k) * kPwStates + st];
    i -= T.adv[st][0];
    j -= T.adv[st][1];
    k -= T.adv[st][2];
    st = prev;
  }
  --pos;
  for (int q = 0; q < 3; ++q) out[q * cap + pos] = 16;
  res[0] = cap - pos;
  res[1] = cost;
}

hipError_t launch_powell(int n1, int n2, int n3, const uint8_t *x1, const uint8_t *x2, const uint8_t *x3, int *V,
                         uint8_t *from, const PwTables *tab, int mm, int ge, uint8_t *out, int cap, int *res,
                         hipStream_t s) {
  const int64_t plane_cells = (int64_t)(n1 + 1) * (n2 + 1);
  const int grid = (int)((plane_cells + 255) / 256);
  for (int t = 0; t <= n1 + n2 + n3; ++t) {
    hipLaunchKernelGGL(powell_plane_kernel, dim3(grid), dim3(256), 0, s, t, n1, n2, n3, x1, x2, x3, V, from, tab,
                       mm, ge);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(powell_trace_kernel, dim3(1), dim3(64), 0, s, n1, n2, n3, (const int *)V,
                     (const uint8_t *)from, tab, x1, x2, x3, out, cap, res);
  return hipGetLastError();
}

size_t pw_tables_size() { return sizeof(PwTables); }

}  // namespace poyg
