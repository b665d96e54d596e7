// Sequence unions (union_CAML_make, src/union.c:209-226) on the host for
// one merge and on the device for a batch of merges, both through
// union_walk.  A merge is a sequential walk over its columns whose every
// step depends on the children's offsets read by the step before, so the
// batched kernel gives each merge one thread; a tree level's merges run
// side by side.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

#include "poyg_internal.h"
#include "poyg_union.h"

extern "C" {
#include "../../include/poyg.h"
}

namespace poyg {

// Column medians of a linear matrix, (a << 5) | b, passed by value.
struct UnionMedTab {
  uint8_t m[32 * 32];
};

__global__ __launch_bounds__(64) void union_merge_kernel(const poyg_union_job *__restrict__ jobs, int n, int gap,
                                                          UnionMedTab tab, int32_t *__restrict__ out_n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const poyg_union_job J = jobs[k];
  const uint8_t *a = J.a, *b = J.b, *med = J.median;
  auto column_median = [&](int i) -> int { return med ? med[i] : tab.m[((a[i] & 31) << 5) | (b[i] & 31)]; };
  const int r = union_walk(a, b, column_median, J.len, gap, UnionIn{J.ua_seq, J.ua_end, J.na},
                           UnionIn{J.ub_seq, J.ub_end, J.nb},
                           UnionOut{J.out_seq, J.out_end, J.out_ca, J.out_cb, J.cap});
  out_n[k] = r;
}

}  // namespace poyg

using namespace poyg;

extern "C" void *poyg_ctx_stream(poyg_ctx *ctx);

static bool fill_med_tab(const poyg_cm *cm, UnionMedTab &t) {
  if (cm->lcm > 5) return false;
  for (int x = 0; x < 32; ++x)
    for (int y = 0; y < 32; ++y)
      t.m[(x << 5) | y] = (x < (1 << cm->lcm) && y < (1 << cm->lcm)) ? cm->median[(x << cm->lcm) + y] : 0;
  return true;
}

extern "C" int poyg_union_make(const poyg_cm *cm, const uint8_t *a, const uint8_t *b, const uint8_t *median, int len,
                               const uint8_t *ua_seq, const int32_t *ua_end, int na, const uint8_t *ub_seq,
                               const int32_t *ub_end, int nb, uint8_t *out_seq, int32_t *out_end, int32_t *out_ca,
                               int32_t *out_cb, int cap, int *out_n) {
  if (!cm || len < 0 || na < 0 || nb < 0 || cap < 0 || (len > 0 && (!a || !b)) || (na > 0 && (!ua_seq || !ua_end)) ||
      (nb > 0 && (!ub_seq || !ub_end)) || (cap > 0 && (!out_seq || !out_end || !out_ca || !out_cb)) || !out_n) {
    set_error("poyg_union_make: bad arguments");
    return POYG_EINVAL;
  }
  const int lcm = cm->lcm;
  auto column_median = [&](int i) -> int { return median ? median[i] : cm->median[(a[i] << lcm) + b[i]]; };
  const int r = union_walk(a, b, column_median, len, cm->gap, UnionIn{ua_seq, ua_end, na}, UnionIn{ub_seq, ub_end, nb},
                           UnionOut{out_seq, out_end, out_ca, out_cb, cap});
  if (r < 0) {
    set_error(r == kUnionFull ? "poyg_union_make: the output capacity is too small"
                              : "poyg_union_make: the children's unions do not match the alignment");
    return POYG_EINVAL;
  }
  *out_n = r;
  return POYG_OK;
}

extern "C" int poyg_union_merge(poyg_ctx *ctx, const poyg_cm *cm, int32_t n, const poyg_union_job *d_jobs,
                                int32_t *d_out_n, void *stream) {
  UnionMedTab tab;
  if (!ctx || !cm || n < 0 || (n > 0 && (!d_jobs || !d_out_n)) || !fill_med_tab(cm, tab)) {
    set_error("poyg_union_merge: bad arguments (or an alphabet wider than 5 bits)");
    return POYG_EINVAL;
  }
  if (n == 0) return POYG_OK;
  hipStream_t s = stream ? (hipStream_t)stream : (hipStream_t)poyg_ctx_stream(ctx);
  hipLaunchKernelGGL(union_merge_kernel, dim3((n + 63) / 64), dim3(64), 0, s, d_jobs, n, cm->gap, tab, d_out_n);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("poyg_union_merge: %s", hipGetErrorString(e));
    return POYG_EDEVICE;
  }
  return POYG_OK;
}
