// Learned join scorer: 4-mer profiles of sequences, pair features, and a
// 256 -> 256 -> 256 -> 1 ReLU network with bf16 weights and f32
// accumulation, trained and evaluated by the kernels below (bf16 matrix
// products on the MFMA units, 16x16x32 per instruction).  The network only
// orders candidate joins; every join it keeps is scored again by the exact
// alignment DP, so nothing here changes a reported cost.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "poyg_internal.h"

namespace poyg {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kH = 256;   // input width = hidden width

// float -> bf16 with round-to-nearest-even (inputs are finite).
__host__ __device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ __forceinline__ float bf2f(uint16_t h) {
  const uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

// ---------------------------------------------------------------------------
// 4-mer profiles.  A position's base is the lowest base bit of its 5-bit
// mask (A=1, C=2, G=4, T=8); a position with no base bit (the gap) ends
// every window that holds it.  Position 0 (the leading gap) is skipped.
// counts[s][b0*64 + b1*16 + b2*4 + b3] = windows of sequence s.
// One workgroup per sequence, an LDS histogram.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int base_of(uint8_t x) {
  const int b = x & 15;
  return b ? __builtin_ctz(b) : -1;
}

__global__ __launch_bounds__(256) void kmer_count_kernel(const uint8_t *__restrict__ seq,
                                                          const int64_t *__restrict__ off,
                                                          const int32_t *__restrict__ len, int n,
                                                          float *__restrict__ counts) {
  __shared__ int hist[256];
  for (int s = blockIdx.x; s < n; s += gridDim.x) {
    hist[threadIdx.x] = 0;
    __syncthreads();
    const uint8_t *p = seq + off[s];
    const int L = len[s];
    for (int i = 1 + (int)threadIdx.x; i + 3 < L; i += blockDim.x) {
      const int a = base_of(p[i]), b = base_of(p[i + 1]), c = base_of(p[i + 2]), d = base_of(p[i + 3]);
      if ((a | b | c | d) >= 0) atomicAdd(&hist[(a << 6) | (b << 4) | (c << 2) | d], 1);
    }
    __syncthreads();
    counts[(int64_t)s * 256 + threadIdx.x] = (float)hist[threadIdx.x];
    __syncthreads();
  }
}

// x[p][k] = bf16(|counts[ia[p]][k] - counts[ib[p]][k]| * scale)
__global__ __launch_bounds__(256) void pair_feature_kernel(const float *__restrict__ counts,
                                                            const int32_t *__restrict__ ia,
                                                            const int32_t *__restrict__ ib, int64_t n,
                                                            float scale, uint16_t *__restrict__ x) {
  for (int64_t p = blockIdx.x; p < n; p += gridDim.x) {
    const float a = counts[(int64_t)ia[p] * 256 + threadIdx.x];
    const float b = counts[(int64_t)ib[p] * 256 + threadIdx.x];
    x[p * 256 + threadIdx.x] = f2bf(fabsf(a - b) * scale);
  }
}

// ---------------------------------------------------------------------------
// C[i][j] = sum_k A(i, k) B(k, j) over bf16 operands with f32 accumulation,
// A(i, k) = A[i*a_rs + k*a_cs], B(k, j) = B[k*b_rs + j*b_cs].  A workgroup
// (four waves) computes a 64 x 64 tile of C; each wave a 32 x 32 quadrant
// as 2 x 2 blocks of v_mfma_f32_16x16x32_bf16.  Operand tiles of 32 k pass
// through LDS in k-contiguous rows (A as [i][k], B as [j][k]), so that the
// MFMA operands are one 16-byte LDS read per lane: lane l holds
// A[l%16][8*(l/16) .. +7] and B[8*(l/16) .. +7][l%16], and receives
// C[4*(l/16) + r][l%16], r = 0..3.  Rows i >= M, columns j >= N and k >= K
// read as zero and are not stored.
// ---------------------------------------------------------------------------
enum GemmEpi : int {
  kEpiBiasReluBf16 = 0,   // out bf16 = relu(acc + bias[j])
  kEpiMaskBf16 = 1,       // out bf16 = acc where mask(i, j) > 0, else 0
  kEpiF32 = 2,            // out f32 = acc
};

constexpr int kTM = 64, kTN = 64, kTK = 32, kLdsK = kTK + 8;   // row pad: 16 bytes

__device__ __forceinline__ void load_tile(uint16_t (*dst)[kLdsK], const uint16_t *__restrict__ src, int64_t rs,
                                          int64_t cs, int r0, int rmax, int k0, int kmax) {
  const int t = threadIdx.x;
  if (cs == 1 && (rs & 7) == 0 && k0 + kTK <= kmax) {
    // 16-byte loads: thread t takes row t / 4, k chunk (t % 4) * 8
    const int r = t >> 2, kc = (t & 3) * 8;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (r0 + r < rmax) v = *(const uint4 *)(src + (int64_t)(r0 + r) * rs + k0 + kc);
    *(uint4 *)&dst[r][kc] = v;
  } else {
    // element loads; consecutive threads take consecutive rows
    for (int e = t; e < kTM * kTK; e += 256) {
      const int r = e & (kTM - 1), k = e >> 6;
      uint16_t v = 0;
      if (r0 + r < rmax && k0 + k < kmax) v = src[(int64_t)(r0 + r) * rs + (int64_t)(k0 + k) * cs];
      dst[r][k] = v;
    }
  }
}

template <int kEpi>
__global__ __launch_bounds__(256) void gemm_bf16_kernel(int M, int N, int K, const uint16_t *__restrict__ A,
                                                        int64_t a_rs, int64_t a_cs, const uint16_t *__restrict__ B,
                                                        int64_t b_rs, int64_t b_cs, const float *__restrict__ bias,
                                                        const uint16_t *__restrict__ mask, int64_t ldm,
                                                        void *__restrict__ out, int64_t ldc, int kchunk,
                                                        int64_t split_stride) {
  __shared__ __attribute__((aligned(16))) uint16_t As[kTM][kLdsK];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[kTN][kLdsK];
  const int i0 = blockIdx.x * kTM, j0 = blockIdx.y * kTN;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int qi = (wv >> 1) * 32, qj = (wv & 1) * 32;
  // split K: workgroup z takes k in [z kchunk, (z + 1) kchunk) and writes
  // its partial product split_stride elements further (f32 output only)
  const int kz0 = blockIdx.z * kchunk, kz1 = min(K, kz0 + kchunk);
  if (kEpi == kEpiF32) out = (float *)out + blockIdx.z * split_stride;
  f32x4 acc[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = kz0; k0 < kz1; k0 += kTK) {
    __syncthreads();
    load_tile(As, A, a_rs, a_cs, i0, M, k0, kz1);
    // B(k, j): rows of the LDS tile are columns j, k contiguous
    load_tile(Bs, B, b_cs, b_rs, j0, N, k0, kz1);
    __syncthreads();
    const int kk = 8 * (lane >> 4);
    bf16x8 av[2], bv[2];
    for (int a = 0; a < 2; ++a) av[a] = *(const bf16x8 *)&As[qi + 16 * a + (lane & 15)][kk];
    for (int b = 0; b < 2; ++b) bv[b] = *(const bf16x8 *)&Bs[qj + 16 * b + (lane & 15)][kk];
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a], bv[b], acc[a][b], 0, 0, 0);
  }
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + qi + 16 * a + 4 * (lane >> 4) + r;
        const int j = j0 + qj + 16 * b + (lane & 15);
        if (i >= M || j >= N) continue;
        const float v = acc[a][b][r];
        if (kEpi == kEpiBiasReluBf16) {
          ((uint16_t *)out)[(int64_t)i * ldc + j] = f2bf(fmaxf(v + bias[j], 0.f));
        } else if (kEpi == kEpiMaskBf16) {
          const bool on = bf2f(mask[(int64_t)i * ldm + j]) > 0.f;
          ((uint16_t *)out)[(int64_t)i * ldc + j] = on ? f2bf(v) : (uint16_t)0;
        } else {
          ((float *)out)[(int64_t)i * ldc + j] = v;
        }
      }
}

// Output layer and loss.  y[m] = sum_k H2[m][k] w3[k] + b3; with targets,
// dy[m] = 2 (y[m] - t[m]) / n (mean squared error), dZ2[m][k] =
// bf16(dy[m] w3[k]) where H2[m][k] > 0, and loss[m] = (y - t)^2 / n.
// One workgroup of 256 threads per row.
__global__ __launch_bounds__(256) void head_kernel(int n, const uint16_t *__restrict__ H2,
                                                   const uint16_t *__restrict__ w3, const float *__restrict__ b3,
                                                   const float *__restrict__ t, float *__restrict__ y,
                                                   float *__restrict__ dy, float *__restrict__ loss,
                                                   uint16_t *__restrict__ dZ2) {
  __shared__ float part[4];
  const int k = threadIdx.x, lane = k & 63, wv = k >> 6;
  const float wk = bf2f(w3[k]);
  for (int m = blockIdx.x; m < n; m += gridDim.x) {
    const float h = bf2f(H2[(int64_t)m * kH + k]);
    float s = h * wk;
    for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d);
    if (lane == 0) part[wv] = s;
    __syncthreads();
    const float ym = ((part[0] + part[1]) + (part[2] + part[3])) + b3[0];
    __syncthreads();
    if (k == 0 && y) y[m] = ym;
    if (t) {
      const float e = ym - t[m];
      const float g = 2.f * e / (float)n;
      if (k == 0) {
        dy[m] = g;
        loss[m] = e * e / (float)n;
      }
      dZ2[(int64_t)m * kH + k] = h > 0.f ? f2bf(g * wk) : (uint16_t)0;
    }
  }
}

// Column sums over the batch: out[j] = sum_m X[m][j] (X bf16), or
// sum_m X[m][j] * w[m] when w is given.  Two passes with a fixed order, so
// that the sums are deterministic: workgroup b sums rows
// [b * rows, (b + 1) * rows) of every column into part[b][j] (one thread
// per column, four row streams interleaved to keep loads in flight); then
// one thread per column adds the kColParts partials in order.
constexpr int kColParts = 64;
__global__ __launch_bounds__(256) void colsum_part_kernel(int n, const uint16_t *__restrict__ X,
                                                          const float *__restrict__ w, float *__restrict__ part) {
  const int j = threadIdx.x;   // kH == 256 columns
  const int rows = (n + kColParts - 1) / kColParts;
  const int m0 = blockIdx.x * rows, m1 = min(n, m0 + rows);
  float s[4] = {0.f, 0.f, 0.f, 0.f};
  int m = m0;
  for (; m + 3 < m1; m += 4) {
#pragma unroll
    for (int r = 0; r < 4; ++r) s[r] += bf2f(X[(int64_t)(m + r) * kH + j]) * (w ? w[m + r] : 1.f);
  }
  for (; m < m1; ++m) s[0] += bf2f(X[(int64_t)m * kH + j]) * (w ? w[m] : 1.f);
  part[blockIdx.x * kH + j] = (s[0] + s[1]) + (s[2] + s[3]);
}

__global__ __launch_bounds__(256) void colsum_fin_kernel(const float *__restrict__ part, float *__restrict__ out) {
  const int j = threadIdx.x;
  float s = 0.f;
  for (int b = 0; b < kColParts; ++b) s += part[b * kH + j];
  out[j] = s;
}

// sum of a vector (one workgroup, deterministic order per thread then a
// fixed tree)
__global__ __launch_bounds__(256) void vsum_kernel(int n, const float *__restrict__ v, float *__restrict__ out) {
  __shared__ float part[256];
  float s = 0.f;
  for (int m = threadIdx.x; m < n; m += 256) s += v[m];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int d = 128; d >= 1; d >>= 1) {
    if ((int)threadIdx.x < d) part[threadIdx.x] += part[threadIdx.x + d];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = part[0];
}

// dst row r = src row idx[r] (rows of `words` 32-bit words): batch
// shuffles of training data on the device.
__global__ __launch_bounds__(256) void gather_rows_kernel(const uint32_t *__restrict__ src, int words,
                                                          const int32_t *__restrict__ idx, int64_t n,
                                                          uint32_t *__restrict__ dst) {
  for (int64_t r = blockIdx.x; r < n; r += gridDim.x) {
    const uint32_t *a = src + (int64_t)idx[r] * words;
    uint32_t *b = dst + r * words;
    for (int w = threadIdx.x; w < words; w += blockDim.x) b[w] = a[w];
  }
}

// Adam over the flat parameter vector, then the bf16 copies the products
// read: W1t[n][k] = W1[k][n], W2t[n][k] = W2[k][n], W2b = W2, w3b = w3.
struct MlpLayout {
  static constexpr int64_t W1 = 0, B1 = W1 + kH * kH, W2 = B1 + kH, B2 = W2 + kH * kH, W3 = B2 + kH,
                           B3 = W3 + kH, N = B3 + 1;
};

__global__ __launch_bounds__(256) void adam_kernel(int64_t n, float *__restrict__ p, const float *__restrict__ g,
                                                   float *__restrict__ m1, float *__restrict__ m2, float lr,
                                                   float b1, float b2, float eps, float c1, float c2) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    p[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
  }
}

__global__ __launch_bounds__(256) void cast_params_kernel(const float *__restrict__ p, uint16_t *__restrict__ w1t,
                                                          uint16_t *__restrict__ w2t, uint16_t *__restrict__ w2b,
                                                          uint16_t *__restrict__ w3b) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e < kH * kH) {
    const int k = (int)(e / kH), j = (int)(e % kH);
    w1t[(int64_t)j * kH + k] = f2bf(p[MlpLayout::W1 + e]);
    const uint16_t v = f2bf(p[MlpLayout::W2 + e]);
    w2t[(int64_t)j * kH + k] = v;
    w2b[e] = v;
  }
  if (e < kH) w3b[e] = f2bf(p[MlpLayout::W3 + e]);
}

}  // namespace poyg

using namespace poyg;

#define HIP_TRYM(expr)                                                                           \
  do {                                                                                           \
    hipError_t e_ = (expr);                                                                      \
    if (e_ != hipSuccess) {                                                                      \
      set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__);      \
      return POYG_EDEVICE;                                                                       \
    }                                                                                            \
  } while (0)

struct poyg_mlp {
  int device = 0;
  int batch_max = 0;
  hipStream_t stream = nullptr;
  int step = 0;   // Adam steps taken
  // f32: params, grads, Adam moments; bf16 copies; batch scratch
  float *p = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
  uint16_t *w1t = nullptr, *w2t = nullptr, *w2b = nullptr, *w3b = nullptr;
  uint16_t *h1 = nullptr, *h2 = nullptr, *dz2 = nullptr, *dz1 = nullptr;
  float *y = nullptr, *dy = nullptr, *loss = nullptr, *lsum = nullptr;
  float *cpart = nullptr;   // column-sum partials, kColParts x kH per sum
  float *wpart = nullptr;   // split-K partials of a weight gradient
  ~poyg_mlp() {
    for (void *q : {(void *)p, (void *)g, (void *)m1, (void *)m2, (void *)w1t, (void *)w2t, (void *)w2b,
                    (void *)w3b, (void *)h1, (void *)h2, (void *)dz2, (void *)dz1, (void *)y, (void *)dy,
                    (void *)loss, (void *)lsum, (void *)cpart, (void *)wpart})
      if (q) (void)hipFree(q);
  }
};

extern "C" void *poyg_ctx_stream(poyg_ctx *ctx);
extern "C" int poyg_ctx_device(poyg_ctx *ctx);

static hipError_t gemm(int epi, int M, int N, int K, const uint16_t *A, int64_t a_rs, int64_t a_cs,
                       const uint16_t *B, int64_t b_rs, int64_t b_cs, const float *bias, const uint16_t *mask,
                       int64_t ldm, void *C, int64_t ldc, hipStream_t s) {
  if (M <= 0 || N <= 0) return hipSuccess;
  dim3 grid((M + kTM - 1) / kTM, (N + kTN - 1) / kTN);
  const int kall = std::max(K, 1);
  if (epi == kEpiBiasReluBf16)
    hipLaunchKernelGGL(gemm_bf16_kernel<kEpiBiasReluBf16>, grid, dim3(256), 0, s, M, N, K, A, a_rs, a_cs, B, b_rs,
                       b_cs, bias, mask, ldm, C, ldc, kall, (int64_t)0);
  else if (epi == kEpiMaskBf16)
    hipLaunchKernelGGL(gemm_bf16_kernel<kEpiMaskBf16>, grid, dim3(256), 0, s, M, N, K, A, a_rs, a_cs, B, b_rs,
                       b_cs, bias, mask, ldm, C, ldc, kall, (int64_t)0);
  else
    hipLaunchKernelGGL(gemm_bf16_kernel<kEpiF32>, grid, dim3(256), 0, s, M, N, K, A, a_rs, a_cs, B, b_rs, b_cs,
                       bias, mask, ldm, C, ldc, kall, (int64_t)0);
  return hipGetLastError();
}

// C (f32, M x N, ldc) = sum over z of part[z] (M x N each), in order.
__global__ __launch_bounds__(256) void split_sum_kernel(int M, int N, int splits, const float *__restrict__ part,
                                                        float *__restrict__ C, int64_t ldc) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)M * N) return;
  float s = 0.f;
  for (int z = 0; z < splits; ++z) s += part[z * (int64_t)M * N + e];
  C[(e / N) * ldc + e % N] = s;
}

// Weight gradients (M = N = kH, K = the batch): K split into kWgSplits
// chunks so that 16 output tiles become 16 * kWgSplits workgroups; the
// partial products go to `part` (kWgSplits x M x N floats) and are added
// in a fixed order, so the result does not depend on scheduling.
constexpr int kWgSplits = 32;
static hipError_t gemm_wgrad(int M, int N, int K, const uint16_t *A, int64_t a_rs, int64_t a_cs, const uint16_t *B,
                             int64_t b_rs, int64_t b_cs, float *part, float *C, int64_t ldc, hipStream_t s) {
  if (M <= 0 || N <= 0) return hipSuccess;
  const int chunk = std::max(kTK, (K + kWgSplits - 1) / kWgSplits + kTK - 1) / kTK * kTK;
  const int splits = std::max(1, (K + chunk - 1) / chunk);
  dim3 grid((M + kTM - 1) / kTM, (N + kTN - 1) / kTN, splits);
  hipLaunchKernelGGL(gemm_bf16_kernel<kEpiF32>, grid, dim3(256), 0, s, M, N, K, A, a_rs, a_cs, B, b_rs, b_cs,
                     (const float *)nullptr, (const uint16_t *)nullptr, (int64_t)0, (void *)part, (int64_t)N, chunk,
                     (int64_t)M * N);
  const int64_t tot = (int64_t)M * N;
  hipLaunchKernelGGL(split_sum_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, M, N, splits, part, C,
                     ldc);
  return hipGetLastError();
}

// forward of n <= batch_max rows into m->h1, m->h2 and (optionally) y
static int mlp_forward_batch(poyg_mlp *m, const uint16_t *x, int n, float *y, const float *t, hipStream_t s) {
  HIP_TRYM(gemm(kEpiBiasReluBf16, n, kH, kH, x, kH, 1, m->w1t, 1, kH, m->p + MlpLayout::B1, nullptr, 0, m->h1, kH,
                s));
  HIP_TRYM(gemm(kEpiBiasReluBf16, n, kH, kH, m->h1, kH, 1, m->w2t, 1, kH, m->p + MlpLayout::B2, nullptr, 0, m->h2,
                kH, s));
  hipLaunchKernelGGL(head_kernel, dim3(std::min(n, 65536)), dim3(256), 0, s, n, m->h2, m->w3b, m->p + MlpLayout::B3,
                     t, y, m->dy, m->loss, m->dz2);
  HIP_TRYM(hipGetLastError());
  return POYG_OK;
}

extern "C" {

int poyg_mlp_set_params(poyg_mlp *m, const float *host, int64_t n);

int64_t poyg_mlp_param_count(void) { return MlpLayout::N; }

int poyg_mlp_create(poyg_ctx *ctx, int batch_max, uint64_t seed, poyg_mlp **out) {
  if (!ctx || !out || batch_max < 1) {
    set_error("poyg_mlp_create: bad arguments");
    return POYG_EINVAL;
  }
  poyg_mlp *m = new poyg_mlp;
  m->device = poyg_ctx_device(ctx);
  m->stream = (hipStream_t)poyg_ctx_stream(ctx);
  m->batch_max = batch_max;
  hipError_t e = hipSetDevice(m->device);
  const size_t np = (size_t)MlpLayout::N, nb = (size_t)batch_max;
  auto A = [&](void **q, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(q, std::max<size_t>(bytes, 16));
  };
  A((void **)&m->p, 4 * np);
  A((void **)&m->g, 4 * np);
  A((void **)&m->m1, 4 * np);
  A((void **)&m->m2, 4 * np);
  A((void **)&m->w1t, 2 * kH * kH);
  A((void **)&m->w2t, 2 * kH * kH);
  A((void **)&m->w2b, 2 * kH * kH);
  A((void **)&m->w3b, 2 * kH);
  A((void **)&m->h1, 2 * nb * kH);
  A((void **)&m->h2, 2 * nb * kH);
  A((void **)&m->dz2, 2 * nb * kH);
  A((void **)&m->dz1, 2 * nb * kH);
  A((void **)&m->y, 4 * nb);
  A((void **)&m->dy, 4 * nb);
  A((void **)&m->loss, 4 * nb);
  A((void **)&m->lsum, 16);
  A((void **)&m->cpart, 4 * 3 * kColParts * kH);
  A((void **)&m->wpart, 4 * (size_t)kWgSplits * kH * kH);
  if (e != hipSuccess) {
    delete m;
    set_error("poyg_mlp_create: hipMalloc failed: %s", hipGetErrorString(e));
    return POYG_ENOMEM;
  }
  // He-uniform weights (bound sqrt(6 / fan_in)), zero biases, from a
  // splitmix64 stream of the seed
  std::vector<float> h(np, 0.f);
  uint64_t z = seed;
  auto next = [&]() {
    z += 0x9e3779b97f4a7c15ull;
    uint64_t x = z;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    x ^= x >> 31;
    return (double)(x >> 11) * (1.0 / 9007199254740992.0);
  };
  const float bound = sqrtf(6.f / kH);
  for (int64_t i = 0; i < kH * kH; ++i) h[MlpLayout::W1 + i] = (float)((2 * next() - 1) * bound);
  for (int64_t i = 0; i < kH * kH; ++i) h[MlpLayout::W2 + i] = (float)((2 * next() - 1) * bound);
  for (int64_t i = 0; i < kH; ++i) h[MlpLayout::W3 + i] = (float)((2 * next() - 1) * bound);
  const int rc = poyg_mlp_set_params(m, h.data(), (int64_t)np);
  if (rc) {
    delete m;
    return rc;
  }
  *out = m;
  return POYG_OK;
}

void poyg_mlp_destroy(poyg_mlp *m) { delete m; }

int poyg_mlp_set_params(poyg_mlp *m, const float *host, int64_t n) {
  if (!m || !host || n != MlpLayout::N) return POYG_EINVAL;
  HIP_TRYM(hipSetDevice(m->device));
  HIP_TRYM(hipMemcpyAsync(m->p, host, 4 * (size_t)n, hipMemcpyHostToDevice, m->stream));
  HIP_TRYM(hipMemsetAsync(m->m1, 0, 4 * (size_t)n, m->stream));
  HIP_TRYM(hipMemsetAsync(m->m2, 0, 4 * (size_t)n, m->stream));
  hipLaunchKernelGGL(cast_params_kernel, dim3((kH * kH + 255) / 256), dim3(256), 0, m->stream, m->p, m->w1t, m->w2t,
                     m->w2b, m->w3b);
  HIP_TRYM(hipGetLastError());
  HIP_TRYM(hipStreamSynchronize(m->stream));
  m->step = 0;
  return POYG_OK;
}

int poyg_mlp_get_params(poyg_mlp *m, float *host, int64_t n) {
  if (!m || !host || n != MlpLayout::N) return POYG_EINVAL;
  HIP_TRYM(hipMemcpyAsync(host, m->p, 4 * (size_t)n, hipMemcpyDeviceToHost, m->stream));
  HIP_TRYM(hipStreamSynchronize(m->stream));
  return POYG_OK;
}

int poyg_mlp_get_grads(poyg_mlp *m, float *host, int64_t n) {
  if (!m || !host || n != MlpLayout::N) return POYG_EINVAL;
  HIP_TRYM(hipMemcpyAsync(host, m->g, 4 * (size_t)n, hipMemcpyDeviceToHost, m->stream));
  HIP_TRYM(hipStreamSynchronize(m->stream));
  return POYG_OK;
}

// y[r] for r < n of bf16 feature rows d_x[r][256] (device), in batches.
int poyg_mlp_forward(poyg_mlp *m, const uint16_t *d_x, int64_t n, float *d_y, void *stream) {
  if (!m || (n > 0 && (!d_x || !d_y)) || n < 0) return POYG_EINVAL;
  hipStream_t s = stream ? (hipStream_t)stream : m->stream;
  for (int64_t r = 0; r < n; r += m->batch_max) {
    const int nb = (int)std::min<int64_t>(m->batch_max, n - r);
    const int rc = mlp_forward_batch(m, d_x + r * kH, nb, d_y + r, nullptr, s);
    if (rc) return rc;
  }
  return POYG_OK;
}

// Gradient of the mean squared error over one batch (n <= batch_max rows
// of features d_x and targets d_t), kept for poyg_mlp_adam; the batch's
// loss (mean) goes to *loss when given (a synchronisation).
int poyg_mlp_grad(poyg_mlp *m, const uint16_t *d_x, const float *d_t, int n, float *loss) {
  if (!m || !d_x || !d_t || n < 1 || n > m->batch_max) {
    set_error("poyg_mlp_grad: bad arguments (batch of %d rows, at most %d)", n, m ? m->batch_max : 0);
    return POYG_EINVAL;
  }
  hipStream_t s = m->stream;
  int rc = mlp_forward_batch(m, d_x, n, m->y, d_t, s);
  if (rc) return rc;
  float *g = m->g;
  // dZ1 = (dZ2 W2^T) where H1 > 0: B(k', j) = W2[j][k'] (W2b is [in][out])
  HIP_TRYM(gemm(kEpiMaskBf16, n, kH, kH, m->dz2, kH, 1, m->w2b, 1, kH, nullptr, m->h1, kH, m->dz1, kH, s));
  // dW2[k][j] = sum_m H1[m][k] dZ2[m][j]; dW1[k][j] = sum_m X[m][k] dZ1[m][j]
  HIP_TRYM(gemm_wgrad(kH, kH, n, m->h1, 1, kH, m->dz2, kH, 1, m->wpart, g + MlpLayout::W2, kH, s));
  HIP_TRYM(gemm_wgrad(kH, kH, n, d_x, 1, kH, m->dz1, kH, 1, m->wpart, g + MlpLayout::W1, kH, s));
  float *cp = m->cpart;
  hipLaunchKernelGGL(colsum_part_kernel, dim3(kColParts), dim3(kH), 0, s, n, m->dz2, nullptr, cp);
  hipLaunchKernelGGL(colsum_part_kernel, dim3(kColParts), dim3(kH), 0, s, n, m->dz1, nullptr, cp + kColParts * kH);
  hipLaunchKernelGGL(colsum_part_kernel, dim3(kColParts), dim3(kH), 0, s, n, m->h2, m->dy, cp + 2 * kColParts * kH);
  hipLaunchKernelGGL(colsum_fin_kernel, dim3(1), dim3(kH), 0, s, cp, g + MlpLayout::B2);
  hipLaunchKernelGGL(colsum_fin_kernel, dim3(1), dim3(kH), 0, s, cp + kColParts * kH, g + MlpLayout::B1);
  hipLaunchKernelGGL(colsum_fin_kernel, dim3(1), dim3(kH), 0, s, cp + 2 * kColParts * kH, g + MlpLayout::W3);
  hipLaunchKernelGGL(vsum_kernel, dim3(1), dim3(256), 0, s, n, m->dy, g + MlpLayout::B3);
  HIP_TRYM(hipGetLastError());
  if (loss) {
    hipLaunchKernelGGL(vsum_kernel, dim3(1), dim3(256), 0, s, n, m->loss, m->lsum);
    HIP_TRYM(hipGetLastError());
    HIP_TRYM(hipMemcpyAsync(loss, m->lsum, 4, hipMemcpyDeviceToHost, s));
    HIP_TRYM(hipStreamSynchronize(s));
  }
  return POYG_OK;
}

// One Adam step on the gradient the last poyg_mlp_grad left.
int poyg_mlp_adam(poyg_mlp *m, float lr, float beta1, float beta2, float eps) {
  if (!m) return POYG_EINVAL;
  m->step += 1;
  const float c1 = 1.f - powf(beta1, (float)m->step), c2 = 1.f - powf(beta2, (float)m->step);
  hipLaunchKernelGGL(adam_kernel, dim3(256), dim3(256), 0, m->stream, MlpLayout::N, m->p, m->g, m->m1, m->m2, lr,
                     beta1, beta2, eps, c1, c2);
  hipLaunchKernelGGL(cast_params_kernel, dim3((kH * kH + 255) / 256), dim3(256), 0, m->stream, m->p, m->w1t,
                     m->w2t, m->w2b, m->w3b);
  HIP_TRYM(hipGetLastError());
  return POYG_OK;
}

// 4-mer profiles (float counts, 256 per sequence) of n device sequences.
int poyg_kmer_counts(poyg_ctx *ctx, const uint8_t *d_seq, const int64_t *d_off, const int32_t *d_len, int32_t n,
                     float *d_counts, void *stream) {
  if (!ctx || n < 0 || (n > 0 && (!d_seq || !d_off || !d_len || !d_counts))) return POYG_EINVAL;
  if (n == 0) return POYG_OK;
  hipStream_t s = stream ? (hipStream_t)stream : (hipStream_t)poyg_ctx_stream(ctx);
  hipLaunchKernelGGL(kmer_count_kernel, dim3(std::min(n, 65536)), dim3(256), 0, s, d_seq, d_off, d_len, n,
                     d_counts);
  HIP_TRYM(hipGetLastError());
  return POYG_OK;
}

// Pair features x[p] = bf16(|counts[ia[p]] - counts[ib[p]]| * scale).
int poyg_pair_features(poyg_ctx *ctx, const float *d_counts, const int32_t *d_ia, const int32_t *d_ib, int64_t n,
                       float scale, uint16_t *d_x, void *stream) {
  if (!ctx || n < 0 || (n > 0 && (!d_counts || !d_ia || !d_ib || !d_x))) return POYG_EINVAL;
  if (n == 0) return POYG_OK;
  hipStream_t s = stream ? (hipStream_t)stream : (hipStream_t)poyg_ctx_stream(ctx);
  hipLaunchKernelGGL(pair_feature_kernel, dim3((unsigned)std::min<int64_t>(n, 1 << 20)), dim3(256), 0, s, d_counts,
                     d_ia, d_ib, n, scale, d_x);
  HIP_TRYM(hipGetLastError());
  return POYG_OK;
}

// dst[r] = src[idx[r]] for rows of row_bytes (a multiple of 4) bytes.
int poyg_gather_rows(poyg_ctx *ctx, const void *d_src, int64_t row_bytes, const int32_t *d_idx, int64_t n,
                     void *d_dst, void *stream) {
  if (!ctx || n < 0 || row_bytes <= 0 || (row_bytes & 3) || (n > 0 && (!d_src || !d_idx || !d_dst))) {
    set_error("poyg_gather_rows: bad arguments");
    return POYG_EINVAL;
  }
  if (n == 0) return POYG_OK;
  hipStream_t s = stream ? (hipStream_t)stream : (hipStream_t)poyg_ctx_stream(ctx);
  hipLaunchKernelGGL(gather_rows_kernel, dim3((unsigned)std::min<int64_t>(n, 1 << 20)), dim3(256), 0, s,
                     (const uint32_t *)d_src, (int)(row_bytes / 4), d_idx, n, (uint32_t *)d_dst);
  HIP_TRYM(hipGetLastError());
  return POYG_OK;
}

// A plain bf16 product for tests of the MFMA tiles: C (f32, ldc) = A B with
// the strides above.
int poyg_gemm_bf16_f32(poyg_ctx *ctx, int M, int N, int K, const uint16_t *d_a, int64_t a_rs, int64_t a_cs,
                       const uint16_t *d_b, int64_t b_rs, int64_t b_cs, float *d_c, int64_t ldc, void *stream) {
  if (!ctx || M < 0 || N < 0 || K < 0) return POYG_EINVAL;
  hipStream_t s = stream ? (hipStream_t)stream : (hipStream_t)poyg_ctx_stream(ctx);
  HIP_TRYM(gemm(kEpiF32, M, N, K, d_a, a_rs, a_cs, d_b, b_rs, b_cs, nullptr, nullptr, 0, d_c, ldc, s));
  return POYG_OK;
}

}  // extern "C"
