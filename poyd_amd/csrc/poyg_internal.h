// Internal structures shared by the host side and the kernels of the engine.
#pragma once

#include <stdint.h>

#include "../../include/poyg.h"

namespace poyg {

constexpr int kHighNum = 1000000;     // sentinel of the affine DP, src/algn.c:38
constexpr int kOcamlMaxInt = 1073741823;
constexpr int kGap = 16;              // gap bit of the 5-bit DNA masks
constexpr int kCmSlots = 2048;        // 2 * 2^5 * 2^5, src/cm.c:344
constexpr int kCm3Slots = 64 * 64 * 64;
constexpr int kWave = 64;

// Device copy of a two-dimensional cost matrix: the 32 x 32 block of a
// 5-bit alphabet, with prepend/tail rows and the scalar parameters.
struct DevTcm {
  int cost[32 * 32];
  int prepend[32];
  int tail[32];
  uint8_t median[32 * 32];
  int gap, lcm, gap_open, cost_model, combinations, pad[3];
};

// Three-dimensional cost matrix block for 5-bit alphabets.
struct DevTcm3 {
  int cost[32 * 32 * 32];
  uint8_t median[32 * 32 * 32];
  int gap, lcm, a_sz, pad;
};

// Per-pair descriptor of a pair plan (device resident).
struct PairMeta {
  int64_t off_r;     // rows sequence: the shorter one for affine, s2-role for linear
  int64_t off_c;     // columns sequence
  int64_t out_off;   // byte offset of the pair's output region
  int64_t dm_off;    // byte offset of its direction band
  int64_t aux_off;   // int32 offset of its prefix-sum arrays
  int32_t lr, lc;    // lengths including the leading gap
  int32_t stride;    // bytes per direction-band row
  int32_t cap;       // capacity of one output sequence
  int32_t swap;      // 1 when the rows sequence is the caller's second one
  int32_t p0, p1, p2;  // mode specific band parameters
  int32_t p3;          // linear: 1 when the packed 16-bit fill (128-row strips) fills the pair
};

// Outputs of the backtrace that the affine fill kernels run on each pair
// right after its fill (the same wave: its own band stores are visible to
// its later loads, and the band is still in L2).  A plan keeps one in
// device memory, written at the start of each run.
struct TraceOut {
  uint8_t *out;
  int32_t *out_len;
  int32_t *status;
};

// A plan's status block: word 0 counts the pairs of the last run that
// failed, words kStatusFlags + p hold one flag per pair (so that a pair that
// fails in several waves or kernels counts once).
constexpr int kStatusFlags = 4;

// Linear band row descriptor kinds (src/algn.c:872-968).
enum LinearCase : int32_t { kLinFull = 0, kLinCase2 = 1, kLinCase3 = 2 };

}  // namespace poyg

struct poyg_cm {
  int a_sz, lcm, gap, cost_model, combinations, gap_open, is_metric, all_elements;
  int cost[poyg::kCmSlots];
  int worst[poyg::kCmSlots];
  int prepend[poyg::kCmSlots];
  int tail[poyg::kCmSlots];
  uint8_t median[poyg::kCmSlots];
};

struct poyg_cm3 {
  int a_sz, lcm, gap, cost_model, combinations, gap_open, all_elements;
  int cost[poyg::kCm3Slots];
  uint8_t median[poyg::kCm3Slots];
};

namespace poyg {
int ctx_cus(poyg_ctx *ctx);   // compute units of the context's device
void set_error(const char *fmt, ...);
void to_dev_tcm(const poyg_cm *cm, DevTcm *out);
void to_dev_tcm3(const poyg_cm3 *cm, DevTcm3 *out);
}  // namespace poyg
