// Host-side geometry of the linear band (algn_fill_plane_2,
// src/algn.c:872-968): the case, the clamped width and height, each row's
// first and last column, and the band's widest row and cell count.  Plain
// C++ (no HIP), so that a CPU test can check it.
#pragma once

#include <stdint.h>

#include <algorithm>

#if defined(__HIPCC__)
#define POYG_HD __host__ __device__
#else
#define POYG_HD
#endif

namespace poyg {

enum LinBandCase : int { kBandFull = 0, kBandCase2 = 1, kBandCase3 = 2 };

// Linear band geometry of algn_fill_plane_2 (src/algn.c:872-968).
struct LinGeom {
  int cs, w, h;
};
POYG_HD inline LinGeom lin_geom(int m, int n, int dw) {
  LinGeom g;
  g.w = n < 50 + dw ? n : 50 + dw;
  g.h = m < (m - n) + 50 + dw ? m : (m - n) + 50 + dw;
  if ((float)m >= (float)((float)((float)3 / (float)2) * (float)n)) g.cs = kBandFull;
  else if (2 * g.h < m) g.cs = kBandCase2;
  else if (8 >= m - g.h) g.cs = kBandFull;
  else g.cs = kBandCase3;
  return g;
}
POYG_HD inline void lin_row_host(int i, int m, int n, const LinGeom &g, int &lo, int &hi) {
  if (g.cs == kBandFull) { lo = 0; hi = n - 1; return; }
  if (i == 0) { lo = 0; hi = g.w - 1; return; }
  const int w = g.w, h = g.h;
  if (g.cs == kBandCase2) {
    if (i < h) { lo = 0; hi = w + i - 1; }
    else if (i < m - (h - 1)) { lo = 1 + (i - h); hi = lo + w + h - 2; }
    else { lo = n - (w + h - 2) + (i - (m - (h - 1))); hi = n - 1; }
    return;
  }
  if (i < n - w + 1) { lo = 0; hi = w + i - 1; }
  else if (i < m - (n - w) + 1) { lo = 0; hi = n - 1; }
  else { lo = 1 + (i - (m - (n - w) + 1)); hi = n - 1; }
}

// The linear kernels address a pair's direction band with 32-bit offsets
// and use 0x7ffffff0 as their "out of range" offset: a pair whose band
// (rows x stride bytes) reaches it cannot be aligned by them.
constexpr int64_t kLinDmLimit = 0x7ffffff0;
// Bytes per direction row of a linear band whose widest row is maxw: room
// for a start phase of up to 3 bytes, rounded to 4.
POYG_HD inline int lin_dm_stride(int maxw) { return (maxw + 3 + 3) & ~3; }
POYG_HD inline bool lin_dm_fits(int m, int maxw) { return (int64_t)m * lin_dm_stride(maxw) < kLinDmLimit; }

// Widest row and number of cells of a linear band.  Within each of the three
// sections of lin_row_host the row width is affine in the row (growing by
// one, constant, shrinking by one), so a section's cells are an arithmetic
// series of its first and last widths and its widest row is one of the two.
POYG_HD inline void lin_band_extent(int m, int n, const LinGeom &g, int &maxw, int64_t &cells) {
  auto width = [&](int i) {
    int lo, hi;
    lin_row_host(i, m, n, g, lo, hi);
    return hi - lo + 1;
  };
  maxw = width(0);
  cells = maxw;
  if (g.cs == kBandFull) {
    cells = (int64_t)m * n;
    return;
  }
  const int t1 = g.cs == kBandCase2 ? g.h : n - g.w + 1;             // first row of the middle section
  const int t2 = g.cs == kBandCase2 ? m - (g.h - 1) : m - (n - g.w) + 1;   // first row of the last one
  const int sec[3][2] = {{1, std::min(t1, m)},
                         {std::max(1, t1), std::min(m, t2)},
                         {std::max(std::max(1, t1), t2), m}};
  for (const auto &r : sec) {
    if (r[0] >= r[1]) continue;
    const int a = width(r[0]), b = width(r[1] - 1);
    maxw = std::max(maxw, std::max(a, b));
    cells += (int64_t)(a + b) * (r[1] - r[0]) / 2;
  }
}

}  // namespace poyg
