// Sequence unions: union_merge (src/union.c:102-164) over the structure
// Sequence.Unions.u (src/sequence.ml:1810-1895).
//
// A union of n items is four arrays, left-aligned here (the reference keeps
// them right-aligned in capacity-sized bigarrays):
//   seq[k]  the item: the OR of the characters it stands for;
//   end[k]  0 on a homologous item; on a non-homologous one its 1-based
//           place in the run of non-homologous items it belongs to, counted
//           from the run's left end, so that the rightmost item of a run
//           holds the run's length;
//   ca[k], cb[k]  positions in the first / second child's union that a
//           homologous item came from (0 where none).
// A leaf's union (Unions.leaf, :1870-1873) is its sequence with end = 0 and
// ca = cb = 0, 1, ..., n - 1.
//
// The merge walks the aligned children (a, b) and their median with gaps
// from the last column to the first, consuming the children's union items
// from the right.  Before every column it copies the run of non-homologous
// items that ends at each child's current item (union_copy_non_homologous,
// :61-69), OR-ed with the gap.  A column whose median is the gap adds one
// non-homologous item (the gap OR the child items of the column); column 0
// closes the union with the gap and returns.  A column with a median
// character adds a homologous item (the median OR the child items) and
// writes the offsets of the non-homologous items prepended since the
// previous one (union_prepend_and_fill_items, :71-91).  The offsets advance
// separately from the items, as in the reference, and meet at every
// homologous item.
//
// Plain C++ with POYG_HD, so that the host entry point and the batched
// device kernel run the same code.
#pragma once

#include <stdint.h>

#include "poyg_band.h"

namespace poyg {

struct UnionIn {
  const uint8_t *seq;
  const int32_t *end;
  int n;
};

struct UnionOut {
  uint8_t *seq;
  int32_t *end, *ca, *cb;
  int cap;
};

enum : int { kUnionBadChild = -1, kUnionFull = -2, kUnionSkew = -3 };

// Returns the union's length (its items fill out.*[cap - n .. cap - 1]), or
// a negative kUnion* code when the children's unions do not match the
// alignment (an item is needed past a child's first item), the capacity is
// too small, or the item and offset arrays end up misaligned.  med(i) gives
// the median of column i (with gaps).
template <class Med>
POYG_HD inline int union_walk(const uint8_t *a, const uint8_t *b, Med med, int len, int gap, UnionIn A, UnionIn B,
                              UnionOut C) {
  int pa = A.n - 1, pb = B.n - 1;   // the children's current (rightmost unconsumed) items
  int cs = C.cap, ce = C.cap;       // next item slot cs - 1, next offset slot ce - 1
  int pending = 0;                  // items prepended since the last homologous one
  bool bad = false;
  auto put = [&](int v) {
    if (cs == 0) {
      bad = true;
      return;
    }
    C.seq[--cs] = (uint8_t)v;
  };
  auto child = [&](const UnionIn &U, int p) -> int {
    if (p < 0) {
      bad = true;
      return 0;
    }
    return U.seq[p];
  };
  // the run of non-homologous items ending at the child's current item
  auto copy_run = [&](const UnionIn &U, int &p) {
    int k = U.end[p];
    pending += k;
    for (; k > 0 && !bad; --k) {
      put(gap | child(U, p));
      --p;
    }
  };
  // offsets of the pending items (k, k-1, ..., 1 leftwards), then the
  // homologous item v with its child positions
  auto fill = [&](int v, int apos, int bpos) {
    for (int k = pending; k > 0; --k) {
      if (ce == 0) {
        bad = true;
        return;
      }
      --ce;
      C.end[ce] = k;
      C.ca[ce] = 0;
      C.cb[ce] = 0;
    }
    if (ce == 0) {
      bad = true;
      return;
    }
    --ce;
    C.ca[ce] = apos;
    C.cb[ce] = bpos;
    C.end[ce] = 0;
    put(v);
  };
  int i = len - 1;
  while ((pa >= 0 || pb >= 0) && !bad) {
    if (pa >= 0) copy_run(A, pa);
    if (pb >= 0 && !bad) copy_run(B, pb);
    if (bad) break;
    if (i < 0) {
      // no column is left: the reference keeps copying runs; a run of
      // length 0 would never end the loop
      if ((pa >= 0 && A.end[pa] == 0) || (pb >= 0 && B.end[pb] == 0)) bad = true;
      continue;
    }
    int x = med(i);
    if (x == gap) {
      if (i == 0) {
        fill(gap, 0, 0);
        if (bad) break;
        return ce == cs ? C.cap - cs : kUnionSkew;
      }
      if (a[i] == gap) {
        put(gap | child(B, pb));
        --pb;
      } else if (b[i] == gap) {
        put(gap | child(A, pa));
        --pa;
      } else {
        put(gap | child(A, pa) | child(B, pb));
        --pa;
        --pb;
      }
      ++pending;
    } else {
      int apos = 0, bpos = 0;
      if (a[i] == gap) {
        bpos = pb;
        x |= child(B, pb);
        --pb;
      } else if (b[i] == gap) {
        apos = pa;
        x |= child(A, pa);
        --pa;
      } else {
        bpos = pb;
        apos = pa;
        x |= child(A, pa);
        x |= child(B, pb);
        --pa;
        --pb;
      }
      fill(x, apos, bpos);
      pending = 0;
    }
    --i;
  }
  if (bad) return cs == 0 || ce == 0 ? kUnionFull : kUnionBadChild;
  // both children consumed before column 0 (src/union.c:155-160)
  if (cs == C.cap || C.seq[cs] != gap) {
    if (ce == 0 || cs == 0) return kUnionFull;
    --ce;
    C.ca[ce] = 0;
    C.cb[ce] = 0;
    C.end[ce] = 0;
    put(gap);
  }
  return ce == cs ? C.cap - cs : kUnionSkew;
}

}  // namespace poyg
