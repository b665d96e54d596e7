// Statistics of device-resident sequences that a tree pass needs before it
// plans the next batch of alignments: the band parameter of
// Sequence.Align.cost_2 counts the positions that carry the gap bit
// (seq_CAML_count, src/seq.c:571-581), and DOS.median / DOS.distance treat a
// sequence of gaps only as empty (Sequence.is_empty, src/sequence.ml:442-450).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "poyg_internal.h"

namespace poyg {

// One wave per sequence: stats[2s] = positions with the gap bit,
// stats[2s + 1] = 1 when every position is exactly the gap.
__global__ __launch_bounds__(256) void seq_stats_kernel(const uint8_t *__restrict__ seq,
                                                         const int64_t *__restrict__ off,
                                                         const int32_t *__restrict__ len, int n, int gap,
                                                         int32_t *__restrict__ stats) {
  const int lane = threadIdx.x & 63;
  const int waves = gridDim.x * (blockDim.x >> 6);
  for (int s = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); s < n; s += waves) {
    const uint8_t *p = seq + off[s];
    const int L = len[s];
    int g = 0, other = 0;
    for (int i = lane; i < L; i += 64) {
      const int x = p[i];
      g += (x & gap) ? 1 : 0;
      other += x != gap ? 1 : 0;
    }
    for (int d = 32; d >= 1; d >>= 1) {
      g += __shfl_xor(g, d);
      other += __shfl_xor(other, d);
    }
    if (lane == 0) {
      stats[2 * (int64_t)s] = g;
      stats[2 * (int64_t)s + 1] = other == 0 ? 1 : 0;
    }
  }
}

}  // namespace poyg

using namespace poyg;

extern "C" void *poyg_ctx_stream(poyg_ctx *ctx);

extern "C" int poyg_seq_stats(poyg_ctx *ctx, const uint8_t *d_seq, const int64_t *d_off, const int32_t *d_len,
                              int32_t n, int32_t *d_stats, void *stream) {
  if (!ctx || n < 0 || (n > 0 && (!d_seq || !d_off || !d_len || !d_stats))) {
    set_error("poyg_seq_stats: bad arguments");
    return POYG_EINVAL;
  }
  if (n == 0) return POYG_OK;
  hipStream_t s = stream ? (hipStream_t)stream : (hipStream_t)poyg_ctx_stream(ctx);
  hipLaunchKernelGGL(seq_stats_kernel, dim3(std::min((n + 3) / 4, 16384)), dim3(256), 0, s, d_seq, d_off, d_len, n,
                     kGap, d_stats);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("poyg_seq_stats: %s", hipGetErrorString(e));
    return POYG_EDEVICE;
  }
  return POYG_OK;
}
