// Powell's three-way alignment search (powell_3D_align, src/ukkCommon.c:
// 110-151; src/ukk.checkp.c) as level-synchronous device passes: the types
// shared by the kernels (poyg_powell.hip) and the host driver that runs the
// check-point recursion (poyg_api.cpp).
//
// The reference computes U(ab, ac, d, s) -- the furthest index into A
// reached on diagonal (ab = a - b, ac = a - c) at cost at most d ending in
// state s -- by memoised recursion from the end diagonal, cost 1, 2, ...
// Every value is a function of (ab, ac, d, s) and the search range alone, so
// here one launch computes a whole cost level: every diagonal within the
// range's reach at that cost, 16 states each, from the levels below (a
// transition lowers the cost by at most maxSingleStep) and, for the
// all-match state, the other states of the same level.  Each stored cell
// keeps its dist and which candidate won it (the reference's "from" record
// follows from that and the check-point band), so the check-point split, the
// base-case traceback and the reference's choice of check-point cost (the
// first cost at which the cells its recursion has evaluated reach |A| / 2,
// computed from which cells call which) are walks over the stored levels.
#pragma once
#include <stdint.h>

namespace poyg {

constexpr int kUkStates = 16;
constexpr int kUkInf = 5000;        // the reference's INFINITY (MAXINT / 2)
constexpr int kUkMaxLen = 65534;    // a dist is stored in 16 bits as dist + 1 (0: -INFINITY)

// Candidate codes (the high half of a stored cell):
//   2 fs + v   the transition from state fs (v = 1: at one mismatch less)
//   32         the same state one cost lower
//   33 + s     the all-match state extended from state s of the same cost
//   62         the preset start cell of the range;  63  no candidate
constexpr int kUkLower = 32, kUkExt = 33, kUkStart = 62, kUkNone = 63;

// The model tables of setup (src/ukkCommon.c:247-330), states in the
// reference's enumeration order (state 0 = all match).
struct UkModel {
  int ns, mis, go, ge, max_step;
  int nb[kUkStates];       // neighbour step: bit 0 A, bit 1 B, bit 2 C advance
  int cont[kUkStates];     // contCost
  int second[kUkStates];   // secondCost
  int trans[kUkStates][kUkStates];
};

// One doUkkInLimits range (src/ukk.checkp.c:216-322): the start cell (preset
// to sdist), the end cell, and the ends every index is checked against.
struct UkRange {
  int sab, sac, scost, sstate, sdist;
  int fab, fac, fcost, fstate, fdist;
  int endA, endB, endC;
};

// Where level c (scost <= c) of a range sits in the level store: the box of
// diagonals within reach of the start at that cost, clipped to the ends.
struct UkLevel {
  int ab_lo, ac_lo, nab, nac;
  int64_t base;   // first word of the level (16 words per diagonal)
};

// Diagonals within the reach of cost c from the start: every diagonal whose
// cheapest cost (withinMatrix, src/ukk.checkp.c:545-576) is at most c has
// both |ab - sab| and |ac - sac| at most (c - scost) / ge.
__host__ __device__ inline UkLevel uk_level_box(const UkRange &r, int ge, int c, int64_t base) {
  UkLevel L;
  const int w = (c - r.scost) / ge;
  const int ab_lo = r.sab - w > -r.endB ? r.sab - w : -r.endB;
  const int ab_hi = r.sab + w < r.endA ? r.sab + w : r.endA;
  const int ac_lo = r.sac - w > -r.endC ? r.sac - w : -r.endC;
  const int ac_hi = r.sac + w < r.endA ? r.sac + w : r.endA;
  L.ab_lo = ab_lo;
  L.ac_lo = ac_lo;
  L.nab = ab_hi >= ab_lo ? ab_hi - ab_lo + 1 : 0;
  L.nac = ac_hi >= ac_lo ? ac_hi - ac_lo + 1 : 0;
  L.base = base;
  return L;
}

// Results of a walk (check-point split or first pass), read by the host.
struct UkWalk {
  int reached;            // first level at which the range's end condition held (INT_MAX: none)
  int dstar;              // first pass: the cost whose evaluated cells first reach |A| / 2
  int f_ab, f_ac, f_cost, f_state, f_dist;   // the check-point cell (f_cost -1: none)
  int fstate;             // first pass: the end cell's first best state
  int ncols;              // columns written so far (traceback, reverse order)
  int error;              // nonzero: a broken chain or an overflow
};

}  // namespace poyg
