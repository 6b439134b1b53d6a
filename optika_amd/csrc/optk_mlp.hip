// optk_mlp.hip -- kernels of the learned surrogate (a 5 -> 4 x 256 SiLU -> 3
// MLP): a bf16 MFMA GEMM with float32 accumulation (gfx950
// v_mfma_f32_16x16x32_bf16), the elementwise and small-layer kernels of the
// forward and backward passes, Adam, and the fused inference kernel.
//
// The surrogate has no counterpart in the reference (it replaces traces of
// the spectrograph by a learned map); the float64 trace generates its
// training data and judges its error.

#include <hip/hip_runtime.h>

#include <cstring>
#include <string>

#include "optk_internal.h"

using namespace optk;

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// float -> bfloat16 bits, round to nearest even (gfx950 v_cvt_pk_bf16_f32)
__device__ __forceinline__ uint16_t to_bf16(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }

// two floats -> two bfloat16 in one 32-bit word (a in the low half)
__device__ __forceinline__ uint32_t pack_bf16(float a, float b) {
  return (uint32_t)to_bf16(a) | ((uint32_t)to_bf16(b) << 16);
}

__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// sigmoid through the hardware exp2 and reciprocal (1 ulp each): a plain
// division would expand to a range-scaled sequence several times longer.
// exp(-x) = inf (x < -88) gives 0, the limit.
__device__ __forceinline__ float sigmoid(float x) { return __builtin_amdgcn_rcpf(1.0f + __expf(-x)); }

__device__ __forceinline__ float silu(float x) { return x * sigmoid(x); }

// d silu / dx = s (1 + x (1 - s)), s = sigmoid(x)
__device__ __forceinline__ float silu_grad(float x) {
  float s = sigmoid(x);
  return s * (1.0f + x * (1.0f - s));
}

// ---------------------------------------------------------------------------
// GEMM: C[m][n] = sum_k A(m, k) B(k, n), bf16 operands, float32 result.
// A(m, k) = a[m * lda + k] (TA = 0) or a[k * lda + m] (TA = 1);
// B(k, n) = b[k * ldb + n] (TB = 0) or b[n * ldb + k] (TB = 1).
// Block tile 64 x 64 x 32, 4 waves each owning a 32 x 32 quarter as 2 x 2
// MFMA 16x16x32 tiles.  Split-K: block z covers k in [z kc, (z + 1) kc) and
// writes C + z * m * n (a partial sum, reduced by gemm_reduce).
// Requires m, n multiples of 64 and k a multiple of 32 (checked on the host).
// ---------------------------------------------------------------------------

constexpr int BM = 64, BN = 64, BK = 32;

template <int TA, int TB>
__global__ __launch_bounds__(256) void gemm_bf16(const uint16_t *__restrict__ a, int64_t lda,
                                                 const uint16_t *__restrict__ b, int64_t ldb,
                                                 float *__restrict__ c, int64_t ldc, int64_t m,
                                                 int64_t n, int64_t kc) {
  // LDS tiles with k contiguous: As[row][k], Bs[col][k] (+8 pad against bank conflicts)
  __shared__ __attribute__((aligned(16))) uint16_t As[BM][BK + 8];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[BN][BK + 8];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int64_t m0 = (int64_t)blockIdx.y * BM, n0 = (int64_t)blockIdx.x * BN;
  const int64_t k_begin = (int64_t)blockIdx.z * kc;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  floatx4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};

  for (int64_t k0 = k_begin; k0 < k_begin + kc; k0 += BK) {
    // stage: thread t covers row t / 4, k columns (t % 4) * 8 .. + 8
    {
      const int r = tid >> 2, kk = (tid & 3) * 8;
      if (TA == 0) {
        const uint4 v = *reinterpret_cast<const uint4 *>(a + (m0 + r) * lda + k0 + kk);
        *reinterpret_cast<uint4 *>(&As[r][kk]) = v;
      } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) As[r][kk + j] = a[(k0 + kk + j) * lda + m0 + r];
      }
      if (TB == 1) {
        const uint4 v = *reinterpret_cast<const uint4 *>(b + (n0 + r) * ldb + k0 + kk);
        *reinterpret_cast<uint4 *>(&Bs[r][kk]) = v;
      } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) Bs[r][kk + j] = b[(k0 + kk + j) * ldb + n0 + r];
      }
    }
    __syncthreads();
    // A operand: lane holds A[row = lane % 16][k = 8 (lane / 16) .. + 8];
    // B operand: lane holds B[k = 8 (lane / 16) .. + 8][col = lane % 16]
    const int fr = lane & 15, fk = (lane >> 4) * 8;
    bf16x8 af[2], bfr[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      uint4 va = *reinterpret_cast<const uint4 *>(&As[wm + 16 * i + fr][fk]);
      af[i] = __builtin_bit_cast(bf16x8, va);
      uint4 vb = *reinterpret_cast<const uint4 *>(&Bs[wn + 16 * i + fr][fk]);
      bfr[i] = __builtin_bit_cast(bf16x8, vb);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  // result: lane holds C[row = 4 (lane / 16) + r][col = lane % 16] of each tile
  float *cz = c + (int64_t)blockIdx.z * m * n;
  const int64_t ld = gridDim.z > 1 ? n : ldc;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
        const int64_t col = n0 + wn + 16 * j + (lane & 15);
        cz[row * ld + col] = acc[i][j][r];
      }
}

// ---------------------------------------------------------------------------
// The training GEMMs of the surrogate (65536 x 256 x 256 and 256 x 256 x
// 65536 with bf16 operands) are bandwidth-bound on MI355X: 128 to 256 flops
// per byte against a ridge near 300.  Block tile 128 x 128 x 32, 256 threads,
// each wave a 64 x 64 quarter as 4 x 4 MFMA 16x16x32 tiles (16 MFMAs per
// k step).  Operands are staged in LDS with k contiguous, double-buffered:
// the next k step's 16-byte global loads are issued before this step's
// MFMAs and land in the other buffer after them, one barrier per step.  An
// operand stored with k as its row index (TA = 1, TB = 0) is loaded 16 bytes
// along m (or n) and transposed on its way into LDS.  The epilogue can fuse
// the layer's elementwise step (bias + SiLU forward, SiLU derivative
// backward), so the pre-activation is not re-read by a separate kernel.
// Requires m, n multiples of 128 and k a multiple of 32 * split.
// ---------------------------------------------------------------------------

constexpr int GM = 128, GN = 128, GK = 32, GPAD = 8;

enum { EPI_STORE = 0, EPI_BIAS_SILU = 1, EPI_SILU_BACK = 2 };

struct GemmEpilogue {
  float *c;          // EPI_STORE: C (or split partials); BIAS_SILU: pre out; SILU_BACK: dpre32 out
  int64_t ldc;
  const float *bias; // BIAS_SILU
  const float *pre;  // SILU_BACK: pre-activation in, [m][ldc]
  uint16_t *out16;   // BIAS_SILU: bf16 activations; SILU_BACK: bf16 dpre
};

// one 128 x 32 operand tile into LDS ([row][k], k contiguous); KC: stored
// with k contiguous (x[row * ld + k]), else k as the row index (x[k * ld + row])
template <bool KC> struct TileLoad {
  uint4 v[2];
  __device__ __forceinline__ void load(const uint16_t *__restrict__ x, int64_t ld, int64_t row0,
                                       int64_t k0, int tid) {
    if (KC) {
      const int r = tid >> 2, kk = (tid & 3) * 8;
#pragma unroll
      for (int h = 0; h < 2; ++h)
        v[h] = *reinterpret_cast<const uint4 *>(x + (row0 + r + 64 * h) * ld + k0 + kk);
    } else {
      const int kr = tid >> 4, rr = (tid & 15) * 8;
#pragma unroll
      for (int h = 0; h < 2; ++h)
        v[h] = *reinterpret_cast<const uint4 *>(x + (k0 + kr + 16 * h) * ld + row0 + rr);
    }
  }
  __device__ __forceinline__ void store(uint16_t (*t)[GK + GPAD], int tid) const {
    if (KC) {
      const int r = tid >> 2, kk = (tid & 3) * 8;
#pragma unroll
      for (int h = 0; h < 2; ++h) *reinterpret_cast<uint4 *>(&t[r + 64 * h][kk]) = v[h];
    } else {
      const int kr = tid >> 4, rr = (tid & 15) * 8;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const uint32_t w[4] = {v[h].x, v[h].y, v[h].z, v[h].w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          t[rr + 2 * j][kr + 16 * h] = (uint16_t)(w[j] & 0xffffu);
          t[rr + 2 * j + 1][kr + 16 * h] = (uint16_t)(w[j] >> 16);
        }
      }
    }
  }
};

template <int TA, int TB, int EPI>
__global__ __launch_bounds__(256) void gemm_bf16_128(const uint16_t *__restrict__ a, int64_t lda,
                                                     const uint16_t *__restrict__ b, int64_t ldb,
                                                     const GemmEpilogue E, int64_t m, int64_t n,
                                                     int64_t kc) {
  __shared__ __attribute__((aligned(16))) uint16_t As[2][GM][GK + GPAD];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[2][GN][GK + GPAD];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int64_t m0 = (int64_t)blockIdx.y * GM, n0 = (int64_t)blockIdx.x * GN;
  const int64_t k_begin = (int64_t)blockIdx.z * kc;
  const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
  const int fr = lane & 15, fk = (lane >> 4) * 8;
  floatx4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
  TileLoad<TA == 0> la;
  TileLoad<TB == 1> lb;
  la.load(a, lda, m0, k_begin, tid);
  lb.load(b, ldb, n0, k_begin, tid);
  la.store(As[0], tid);
  lb.store(Bs[0], tid);
  __syncthreads();
  const int steps = (int)(kc / GK);
  for (int step = 0; step < steps; ++step) {
    const int buf = step & 1;
    const bool more = step + 1 < steps;
    if (more) {
      la.load(a, lda, m0, k_begin + (int64_t)(step + 1) * GK, tid);
      lb.load(b, ldb, n0, k_begin + (int64_t)(step + 1) * GK, tid);
    }
    bf16x8 af[4], bfr[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      af[i] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(&As[buf][wm + 16 * i + fr][fk]));
      bfr[i] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(&Bs[buf][wn + 16 * i + fr][fk]));
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    if (more) {
      la.store(As[buf ^ 1], tid);
      lb.store(Bs[buf ^ 1], tid);
    }
    __syncthreads();
  }
  // lane holds C[row = 4 (lane / 16) + r][col = lane % 16] of each 16 x 16 tile
  float *cz = E.c + (int64_t)blockIdx.z * m * n;
  const int64_t ld = gridDim.z > 1 ? n : E.ldc;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t col = n0 + wn + 16 * j + (lane & 15);
      float bias = 0.0f;
      if (EPI == EPI_BIAS_SILU) bias = E.bias[col];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
        const int64_t at = row * ld + col;
        float v = acc[i][j][r];
        if (EPI == EPI_BIAS_SILU) {
          v += bias;
          cz[at] = v;
          E.out16[at] = to_bf16(silu(v));
        } else if (EPI == EPI_SILU_BACK) {
          const float g = v * silu_grad(E.pre[at]);
          cz[at] = g;
          E.out16[at] = to_bf16(g);
        } else {
          cz[at] = v;
        }
      }
    }
}

// c[i] = sum over splits of partial[s][i] (m * n elements)
__global__ __launch_bounds__(256) void gemm_reduce(const float *__restrict__ partial, int splits,
                                                   int64_t count, float *__restrict__ c) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  float s = 0.0f;
  for (int z = 0; z < splits; ++z) s += partial[(int64_t)z * count + i];
  c[i] = s;
}

// ---------------------------------------------------------------------------
// elementwise and small-layer kernels
// ---------------------------------------------------------------------------

// pre[r][o] += bias[o]; act[r][o] = bf16(silu(pre[r][o]))
__global__ __launch_bounds__(256) void bias_silu(float *__restrict__ pre, const float *__restrict__ bias,
                                                 uint16_t *__restrict__ act, int64_t rows, int cols) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * cols) return;
  float v = pre[i] + bias[i % cols];
  pre[i] = v;
  act[i] = to_bf16(silu(v));
}

// first layer: pre[r][o] = sum_i x[r][i] w[o][i] + b[o] (float32, in < 8 inputs)
__global__ __launch_bounds__(256) void dense_in(const float *__restrict__ x, int in,
                                                const float *__restrict__ w, const float *__restrict__ bias,
                                                float *__restrict__ pre, uint16_t *__restrict__ act,
                                                int64_t rows, int cols) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * cols) return;
  const int64_t r = i / cols;
  const int o = (int)(i % cols);
  float v = bias[o];
  for (int k = 0; k < in; ++k) v += x[r * in + k] * w[o * in + k];
  if (pre) pre[i] = v;
  act[i] = to_bf16(silu(v));
}

// last layer: y[r][j] = sum_i act[r][i] w[j][i] + b[j]; one thread per row
__global__ __launch_bounds__(256) void dense_out(const uint16_t *__restrict__ act, int hidden,
                                                 const float *__restrict__ w, const float *__restrict__ bias,
                                                 float *__restrict__ y, int out, int64_t rows) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  float acc[8];
  for (int j = 0; j < out; ++j) acc[j] = bias[j];
  for (int i = 0; i < hidden; ++i) {
    float h = from_bf16(act[r * hidden + i]);
    for (int j = 0; j < out; ++j) acc[j] += h * w[j * hidden + i];
  }
  for (int j = 0; j < out; ++j) y[r * out + j] = acc[j];
}

// weighted mean-square loss gradient: dy = 2 w_j (y - t) / rows; loss
// partial sums per block into loss_partial[blockIdx.x]
__global__ __launch_bounds__(256) void mse_grad(const float *__restrict__ y, const float *__restrict__ t,
                                                const float *__restrict__ wj, float *__restrict__ dy,
                                                int out, int64_t rows, float *__restrict__ loss_partial) {
  __shared__ float red[256];
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.0f;
  if (i < rows * out) {
    int j = (int)(i % out);
    float d = y[i] - t[i];
    dy[i] = 2.0f * wj[j] * d / (float)rows;
    l = wj[j] * d * d / (float)rows;
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_partial[blockIdx.x] = red[0];
}

// last layer backward, activations side:
// dpre[r][i] = (sum_j dy[r][j] w[j][i]) * silu'(pre[r][i]) -> bf16
__global__ __launch_bounds__(256) void dense_out_back(const float *__restrict__ dy, int out,
                                                     const float *__restrict__ w,
                                                     const float *__restrict__ pre,
                                                     uint16_t *__restrict__ dpre, float *__restrict__ dpre32,
                                                     int64_t rows, int hidden) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * hidden) return;
  const int64_t r = i / hidden;
  const int h = (int)(i % hidden);
  float s = 0.0f;
  for (int j = 0; j < out; ++j) s += dy[r * out + j] * w[j * hidden + h];
  float g = s * silu_grad(pre[i]);
  dpre[i] = to_bf16(g);
  dpre32[i] = g;
}

// dpre = dact * silu'(pre) (float32 in, bf16 and float32 out)
__global__ __launch_bounds__(256) void silu_back(const float *__restrict__ dact,
                                                 const float *__restrict__ pre,
                                                 uint16_t *__restrict__ dpre, float *__restrict__ dpre32,
                                                 int64_t count) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  float g = dact[i] * silu_grad(pre[i]);
  dpre[i] = to_bf16(g);
  dpre32[i] = g;
}

// Partial outer products for the small layers' weight gradients: block b
// sums rows [b R, (b + 1) R) of x[r][c] * g[r][j] for c < cols (one thread
// per column, cols <= 256) and j < inner (<= 8); the partials are reduced
// over blocks by gemm_reduce.  x is float32 (x32) or bf16 (x16).  Layout of
// a partial: [j][c] (transposed = 0) or [c][j] (transposed = 1).
__global__ __launch_bounds__(256) void outer_grad(const float *__restrict__ x32,
                                                  const uint16_t *__restrict__ x16, int cols,
                                                  const float *__restrict__ g, int inner,
                                                  int64_t rows, int64_t rows_per_block,
                                                  int transposed, float *__restrict__ partial) {
  const int c = threadIdx.x;
  if (c >= cols) return;
  float acc[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  const int64_t r0 = (int64_t)blockIdx.x * rows_per_block;
  int64_t r1 = r0 + rows_per_block;
  if (r1 > rows) r1 = rows;
  for (int64_t r = r0; r < r1; ++r) {
    float xv = x32 ? x32[r * cols + c] : from_bf16(x16[r * cols + c]);
    for (int j = 0; j < inner; ++j) acc[j] += xv * g[r * inner + j];
  }
  float *out = partial + (int64_t)blockIdx.x * cols * inner;
  for (int j = 0; j < inner; ++j) {
    if (transposed)
      out[(int64_t)c * inner + j] = acc[j];
    else
      out[(int64_t)j * cols + c] = acc[j];
  }
}

// column sums: gb[o] = sum_r g[r][o]
__global__ __launch_bounds__(256) void column_sum(const float *__restrict__ g, int64_t rows, int cols,
                                                  float *__restrict__ gb) {
  __shared__ float red[256];
  const int o = blockIdx.x;
  float s = 0.0f;
  for (int64_t r = threadIdx.x; r < rows; r += blockDim.x) s += g[r * cols + o];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) gb[o] = red[0];
}

// Adam (Kingma & Ba 2015) with bias correction; also refreshes the bf16 copy
__global__ __launch_bounds__(256) void adam(float *__restrict__ p, const float *__restrict__ g,
                                            float *__restrict__ m1, float *__restrict__ m2,
                                            uint16_t *__restrict__ p16, int64_t count, float lr,
                                            float beta1, float beta2, float eps, float c1, float c2) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  float gi = g[i];
  float a = beta1 * m1[i] + (1.0f - beta1) * gi;
  float b = beta2 * m2[i] + (1.0f - beta2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  float v = p[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
  p[i] = v;
  if (p16) p16[i] = to_bf16(v);
}

__global__ __launch_bounds__(256) void to_bf16_kernel(const float *__restrict__ x,
                                                      uint16_t *__restrict__ y, int64_t count) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) y[i] = to_bf16(x[i]);
}

unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }



// ---------------------------------------------------------------------------
// polynomial baseline in float64: normal equations and evaluation
// ---------------------------------------------------------------------------

constexpr int POLY_N = OPTK_SURROGATE_POLY;
constexpr int POLY_TILE = 64;                                   // rays per LDS tile
constexpr int POLY_COLS = POLY_N + OPTK_SURROGATE_MAX_TARGETS;  // F | targets
constexpr int POLY_PER_THREAD = (POLY_N * POLY_COLS + 255) / 256;
constexpr int POLY_BLOCKS = 512;                                // partial sums

struct PolyExp {
  uint32_t e[POLY_N];
};

// powers 1, u, u^2, u^3 of every input, and one monomial: the product of the
// powers in input order (the same rounding as a column-by-column product)
__device__ inline void poly_powers(const double *u, double pw[5][4]) {
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    pw[k][0] = 1.0;
    pw[k][1] = u[k];
    pw[k][2] = u[k] * u[k];
    pw[k][3] = pw[k][2] * u[k];
  }
}

__device__ inline double poly_monomial(const double pw[5][4], uint32_t e) {
  double c = 1.0;
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    const uint32_t ek = (e >> (2 * k)) & 3u;
    if (ek) c = c * pw[k][ek];
  }
  return c;
}

// one partial sum of the normal equations per block over a stride of
// 64-ray tiles; thread t owns entries t, t + 256, ...
__global__ __launch_bounds__(256) void poly_normal_partial(const double *__restrict__ u,
                                                           const double *__restrict__ tg,
                                                           const uint8_t *__restrict__ valid,
                                                           int64_t n, int nt, PolyExp X,
                                                           double *__restrict__ partial) {
  __shared__ double G[POLY_TILE][POLY_COLS + 1];
  const int t = threadIdx.x;
  const int cols = POLY_N + nt;
  const int entries = POLY_N * cols;
  double acc[POLY_PER_THREAD];
#pragma unroll
  for (int q = 0; q < POLY_PER_THREAD; ++q) acc[q] = 0.0;
  const int64_t tiles = (n + POLY_TILE - 1) / POLY_TILE;
  for (int64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    __syncthreads();
    {
      // thread t fills ray t % 64: monomials [14 (t / 64), +14), and the
      // targets in the last quarter
      const int r = t & (POLY_TILE - 1), part = t >> 6;
      const int64_t i = tile * POLY_TILE + r;
      const bool ok = i < n && valid[i];
      double ui[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
      if (ok)
        for (int k = 0; k < 5; ++k) ui[k] = u[i * 5 + k];
      double pw[5][4];
      poly_powers(ui, pw);
      for (int m = part * (POLY_N / 4); m < (part + 1) * (POLY_N / 4); ++m)
        G[r][m] = ok ? poly_monomial(pw, X.e[m]) : 0.0;
      if (part == 3)
        for (int j = 0; j < nt; ++j) G[r][POLY_N + j] = ok ? tg[i * nt + j] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < POLY_PER_THREAD; ++q) {
      const int p = t + 256 * q;
      if (p < entries) {
        const int a = p / cols, b = p - a * cols;
        double s = acc[q];
        for (int r = 0; r < POLY_TILE; ++r) s = fma(G[r][a], G[r][b], s);
        acc[q] = s;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < POLY_PER_THREAD; ++q) {
    const int p = t + 256 * q;
    if (p < entries) partial[(int64_t)blockIdx.x * entries + p] = acc[q];
  }
}

__global__ __launch_bounds__(256) void poly_normal_reduce(const double *__restrict__ partial,
                                                          int blocks, int entries,
                                                          double *__restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= entries) return;
  double s = 0.0;
  for (int b = 0; b < blocks; ++b) s += partial[(int64_t)b * entries + p];
  out[p] = s;
}

__global__ __launch_bounds__(256) void poly_eval_kernel(const double *__restrict__ u,
                                                        const double *__restrict__ coef, int64_t n,
                                                        int nt, PolyExp X, double *__restrict__ out) {
  __shared__ double C[POLY_N * OPTK_SURROGATE_MAX_TARGETS];
  for (int k = threadIdx.x; k < POLY_N * nt; k += blockDim.x) C[k] = coef[k];
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double ui[5], pw[5][4];
  for (int k = 0; k < 5; ++k) ui[k] = u[i * 5 + k];
  poly_powers(ui, pw);
  double s[OPTK_SURROGATE_MAX_TARGETS];
#pragma unroll
  for (int j = 0; j < OPTK_SURROGATE_MAX_TARGETS; ++j) s[j] = 0.0;
  for (int m = 0; m < POLY_N; ++m) {
    const double f = poly_monomial(pw, X.e[m]);
#pragma unroll
    for (int j = 0; j < OPTK_SURROGATE_MAX_TARGETS; ++j)
      if (j < nt) s[j] = fma(f, C[m * nt + j], s[j]);
  }
  for (int j = 0; j < nt; ++j) out[i * nt + j] = s[j];
}

// ---------------------------------------------------------------------------
// fused inference: input normalisation, the five layers, the polynomial,
// aperture-edge distances, the error-map lookup and the compaction of the
// rays that fall back to the trace -- one kernel, activations in LDS.
//
// Block: 256 threads (4 waves), a tile of 128 rays (64: four blocks); two
// blocks share a CU, so one block's MFMA phase overlaps the other's VALU phase (SiLU, the
// polynomial).  Every layer runs on MFMA 16x16x32 bf16 with float32
// accumulation.  Layer 1 (5 -> 256): the normalised inputs and the weights
// are split into bf16 high and low parts, k row [hi | lo | hi] against
// weight row [w_hi | w_hi | w_lo], so the one 32-wide k step carries the
// float32 product to ~2^-16.  Hidden layers: a 128 x 256 x 256 product
// each.  Wave w computes all 128 rays x features [64 w, 64 w + 64) as
// 8 x 4 tiles, A operands (activations act[ray][feature], bf16) from LDS,
// B operands (weights, row-major [out][in], k contiguous) straight from
// global memory (L2-resident), accumulators started at the bias.  The B
// operand of n-tile t and lane column c is weight row 64 w + 4 c + t, so
// each lane's four accumulator tiles hold four consecutive features: the
// epilogue (SiLU, bf16) writes them as one 8-byte store per row.  The
// output layer (3 rows, high and low bf16 parts) is one more MFMA tile per
// 16 rays; the polynomial and edge tests run one thread per ray.
// ---------------------------------------------------------------------------

#ifndef OPTK_INF_RAYS
#define OPTK_INF_RAYS 128
#endif
constexpr int INF_RAYS = OPTK_INF_RAYS;  // rays per tile: 64 or 128
constexpr int INF_MT = INF_RAYS / 16;    // 16-ray MFMA row tiles per tile
constexpr int INF_OQ = INF_MT / 4;       // output-layer row tiles per wave
constexpr int INF_BLOCKS_PER_CU = INF_RAYS == 64 ? 4 : 2;
static_assert(INF_RAYS == 64 || INF_RAYS == 128, "tile of 64 or 128 rays");
constexpr int INF_HIDDEN = 256;
constexpr int INF_PAD = 8;

__global__ __launch_bounds__(256, INF_BLOCKS_PER_CU) void surrogate_infer(const optk_surrogate_desc D,
                                                          const optk_surrogate_io IO) {
  __shared__ __attribute__((aligned(16))) uint16_t act[INF_RAYS][INF_HIDDEN + INF_PAD];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int64_t n = IO.n;
  const int64_t tiles = (n + INF_RAYS - 1) / INF_RAYS;
  const float *in[5] = {IO.field_x, IO.field_y, IO.pupil_x, IO.pupil_y, IO.wavelength};
  for (int64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    const int64_t base = tile * INF_RAYS;
    // ---- inputs: thread t < 128 -> ray t, normalised and split into bf16
    // high and low parts, as the 32-wide k row [hi | lo | hi | 0] that meets
    // the first layer's weight rows [w_hi | w_hi | w_lo | 0] ---------------
    if (tid < INF_RAYS) {
      const int64_t i = base + tid;
      float hi[5], lo[5];
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        const float v = i < n ? in[k][i] : 0.0f;
        const float u = (v - D.in_mid[k]) * D.in_inv_half[k];
        hi[k] = from_bf16(to_bf16(u));
        lo[k] = u - hi[k];
      }
      uint4 row0, row1;
      row0.x = pack_bf16(hi[0], hi[1]);
      row0.y = pack_bf16(hi[2], hi[3]);
      row0.z = pack_bf16(hi[4], lo[0]);
      row0.w = pack_bf16(lo[1], lo[2]);
      row1.x = pack_bf16(lo[3], lo[4]);
      row1.y = pack_bf16(hi[0], hi[1]);
      row1.z = pack_bf16(hi[2], hi[3]);
      row1.w = pack_bf16(hi[4], 0.0f);
      *reinterpret_cast<uint4 *>(&act[tid][0]) = row0;
      *reinterpret_cast<uint4 *>(&act[tid][8]) = row1;
      *reinterpret_cast<uint4 *>(&act[tid][16]) = uint4{0u, 0u, 0u, 0u};
      *reinterpret_cast<uint4 *>(&act[tid][24]) = uint4{0u, 0u, 0u, 0u};
    }
    __syncthreads();
    // ---- four layers on MFMA: wave w -> features [64 w, 64 w + 64) of all
    // 128 rays; the first layer's k extent is one 32-wide step ------------
    const int obase = wave * 64;
    const int fr = lane & 15, fk = (lane >> 4) * 8;
    for (int layer = 0; layer < 4; ++layer) {
      const uint16_t *W = layer == 0 ? D.w1_16 : D.w16[layer - 1];
      const float *bias = layer == 0 ? D.b1 : D.b[layer - 1];
      const int K = layer == 0 ? 32 : INF_HIDDEN;
      // lane (row group g, column c) accumulates features obase + 4 c + nt,
      // started at the bias
      const int col = obase + 4 * fr;
      floatx4 acc[INF_MT][4];
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        const float bn = bias[col + nt];
#pragma unroll
        for (int mt = 0; mt < INF_MT; ++mt) acc[mt][nt] = floatx4{bn, bn, bn, bn};
      }
      // software pipeline: the next k-chunk's weights are requested before
      // this chunk's 32 MFMAs, so the L2 latency hides behind them
      const uint16_t *wrow[4];
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) wrow[nt] = W + (int64_t)(col + nt) * K + fk;
      uint4 bnext[4];
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) bnext[nt] = *reinterpret_cast<const uint4 *>(wrow[nt]);
#pragma unroll
      for (int kc = 0; kc < INF_HIDDEN; kc += 32) {
        if (kc < K) {
          bf16x8 bw[4];
#pragma unroll
          for (int nt = 0; nt < 4; ++nt) bw[nt] = __builtin_bit_cast(bf16x8, bnext[nt]);
          if (kc + 32 < K) {
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) bnext[nt] = *reinterpret_cast<const uint4 *>(wrow[nt] + kc + 32);
          }
#pragma unroll
          for (int mt = 0; mt < INF_MT; ++mt) {
            uint4 v = *reinterpret_cast<const uint4 *>(&act[16 * mt + fr][kc + fk]);
            bf16x8 a = __builtin_bit_cast(bf16x8, v);
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
              acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw[nt], acc[mt][nt], 0, 0, 0);
          }
        }
      }
      __syncthreads();  // every wave has read this layer's inputs
#pragma unroll
      for (int mt = 0; mt < INF_MT; ++mt)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int row = 16 * mt + 4 * (lane >> 4) + q;
          uint2 v;
          v.x = pack_bf16(silu(acc[mt][0][q]), silu(acc[mt][1][q]));
          v.y = pack_bf16(silu(acc[mt][2][q]), silu(acc[mt][3][q]));
          *reinterpret_cast<uint2 *>(&act[row][col]) = v;
        }
      __syncthreads();
    }
    // ---- output layer on MFMA: wave w -> rays [16 INF_OQ w, +16 INF_OQ).
    // The three outputs' weights are split into bf16 high and low parts
    // (rows 0..2 and 3..5 of one 16-row tile, the rest zero), so the
    // products carry the float32 weights to ~2^-17 at no extra MFMA --------
    __shared__ float yout[INF_RAYS][8];
    {
      floatx4 oacc[INF_OQ];
#pragma unroll
      for (int q = 0; q < INF_OQ; ++q) oacc[q] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int kc = 0; kc < INF_HIDDEN; kc += 32) {
        uint4 vb = *reinterpret_cast<const uint4 *>(D.w5_16 + fr * INF_HIDDEN + kc + fk);
        bf16x8 b5 = __builtin_bit_cast(bf16x8, vb);
#pragma unroll
        for (int q = 0; q < INF_OQ; ++q) {
          uint4 va = *reinterpret_cast<const uint4 *>(&act[16 * (INF_OQ * wave + q) + fr][kc + fk]);
          oacc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, va), b5, oacc[q], 0, 0, 0);
        }
      }
      if (fr < 6) {
#pragma unroll
        for (int q = 0; q < INF_OQ; ++q)
#pragma unroll
          for (int r = 0; r < 4; ++r) yout[16 * (INF_OQ * wave + q) + 4 * (lane >> 4) + r][fr] = oacc[q][r];
      }
    }
    __syncthreads();
    // ---- polynomial, apertures, fallback: one thread per ray ------------
    bool fallback = false;
    const int64_t i = base + tid;
    const bool live = tid < INF_RAYS && i < n;
    if (live) {
      const float y0 = yout[tid][0] + yout[tid][3] + D.b5[0];
      const float y1 = yout[tid][1] + yout[tid][4] + D.b5[1];
      const float y2 = yout[tid][2] + yout[tid][5] + D.b5[2];
      float u[5];
#pragma unroll
      for (int k = 0; k < 5; ++k) u[k] = (in[k][i] - D.in_mid[k]) * D.in_inv_half[k];
      // cubic monomials, in the order of optk_surrogate_desc.poly_exp
      float tgt[OPTK_SURROGATE_MAX_TARGETS];
#pragma unroll
      for (int j = 0; j < OPTK_SURROGATE_MAX_TARGETS; ++j) tgt[j] = 0.0f;
      float pw[5][4];
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        pw[k][0] = 1.0f;
        pw[k][1] = u[k];
        pw[k][2] = u[k] * u[k];
        pw[k][3] = pw[k][2] * u[k];
      }
      const int nt = D.num_targets;
      for (int t = 0; t < OPTK_SURROGATE_POLY; ++t) {
        const uint32_t e = D.poly_exp[t];  // 2 bits per variable
        float mono = 1.0f;
#pragma unroll
        for (int k = 0; k < 5; ++k) {
          const uint32_t ek = (e >> (2 * k)) & 3u;
          mono *= ek == 0 ? 1.0f : (ek == 1 ? pw[k][1] : (ek == 2 ? pw[k][2] : pw[k][3]));
        }
        const float *c = D.poly + t * OPTK_SURROGATE_MAX_TARGETS;
#pragma unroll
        for (int j = 0; j < OPTK_SURROGATE_MAX_TARGETS; ++j)
          if (j < nt) tgt[j] += mono * c[j];
      }
      const float x = tgt[0] + D.scale[0] * y0;
      const float yy = tgt[1] + D.scale[1] * y1;
      const float li = tgt[2] + D.scale[2] * y2;
      bool inside = true, near = false;
#pragma unroll
      for (int a = 0; a < OPTK_SURROGATE_MAX_APERTURES; ++a) {
        if (a >= D.num_apertures) break;
        const optk_surrogate_aperture &ap = D.apertures[a];
        const float ax = tgt[3 + 2 * a], ay = tgt[4 + 2 * a];
        float d;
        if (ap.kind == OPTK_APERTURE_CIRCLE) {
          d = ap.p[0] - sqrtf(ax * ax + ay * ay);
        } else if (ap.kind == OPTK_APERTURE_RECTANGLE) {
          d = fminf(ap.p[0] - fabsf(ax), ap.p[1] - fabsf(ay));
        } else if (ap.kind == OPTK_APERTURE_POLYGON) {
          float dmin = INFINITY;
          bool in_poly = false;
          for (int v = 0; v < ap.num_vertices; ++v) {
            const int w = v == 0 ? ap.num_vertices - 1 : v - 1;
            const float x1 = ap.vx[w], y1 = ap.vy[w], x2 = ap.vx[v], y2 = ap.vy[v];
            const float ex = x2 - x1, ey = y2 - y1;
            float s = ((ax - x1) * ex + (ay - y1) * ey) / (ex * ex + ey * ey);
            s = fminf(fmaxf(s, 0.0f), 1.0f);
            const float dx = ax - (x1 + s * ex), dy = ay - (y1 + s * ey);
            dmin = fminf(dmin, sqrtf(dx * dx + dy * dy));
            if (((y1 > ay) != (y2 > ay)) && (ax < (x2 - x1) * (ay - y1) / (y2 - y1) + x1)) in_poly = !in_poly;
          }
          d = in_poly ? dmin : -dmin;
        } else {
          d = 0.0f;  // no distance model: always near
        }
        if (ap.inverted) d = -d;
        inside = inside && d > 0.0f;
        near = near || fabsf(d) < D.edge_margin;
      }
      // error-map cell: 16 wavelength x 8 field x x 8 field y bins of [-1, 1]
      int cw = (int)((u[4] + 1.0f) * 8.0f), cx = (int)((u[0] + 1.0f) * 4.0f), cy = (int)((u[1] + 1.0f) * 4.0f);
      cw = cw < 0 ? 0 : (cw > 15 ? 15 : cw);
      cx = cx < 0 ? 0 : (cx > 7 ? 7 : cx);
      cy = cy < 0 ? 0 : (cy > 7 ? 7 : cy);
      const bool flagged = D.error_map ? D.error_map[(cw * 8 + cx) * 8 + cy] != 0 : false;
      fallback = near || flagged || !(fabsf(x) < INFINITY && fabsf(yy) < INFINITY && fabsf(li) < INFINITY);
      IO.x[i] = x;
      IO.y[i] = yy;
      IO.intensity[i] = __expf(li);
      IO.unvignetted[i] = inside ? 1 : 0;
      IO.fallback[i] = fallback ? 1 : 0;
    }
    // ---- compaction of the fallback rays: one atomic per wave ------------
    if (IO.fallback_list && wave < INF_RAYS / 64) {
      const uint64_t mask = __ballot(live && fallback);
      if (mask) {
        int32_t wbase = 0;
        const int leader = __ffsll((long long)mask) - 1;
        if (lane == leader) wbase = atomicAdd(IO.fallback_count, (int32_t)__popcll(mask));
        wbase = __shfl(wbase, leader, 64);
        if (live && fallback) {
          const int rank = __popcll(mask & ((1ull << lane) - 1ull));
          const int32_t slot = wbase + rank;
          if (slot < IO.fallback_capacity) IO.fallback_list[slot] = (int32_t)i;
        }
      }
    }
    __syncthreads();  // act is rewritten by the next tile's first layer
  }
}

}  // namespace

extern "C" {

int32_t optk_gemm_bf16(int32_t trans_a, int32_t trans_b, int64_t m, int64_t n, int64_t k,
                       const void *a, int64_t lda, const void *b, int64_t ldb, float *c,
                       int64_t ldc, int32_t split_k, float *workspace, void *hip_stream) {
  if (!a || !b || !c) return fail(OPTK_EINVAL, "gemm: null operand");
  if (m <= 0 || n <= 0 || k <= 0 || m % BM || n % BN) return fail(OPTK_EINVAL, "gemm: m and n must be positive multiples of 64");
  if (split_k < 1) return fail(OPTK_EINVAL, "gemm: split_k must be >= 1");
  if (k % ((int64_t)BK * split_k)) return fail(OPTK_EINVAL, "gemm: k must be a multiple of 32 * split_k");
  if (split_k > 1 && !workspace) return fail(OPTK_EINVAL, "gemm: split_k > 1 needs a workspace of split_k * m * n floats");
  if ((trans_a ? lda < m : lda < k) || (trans_b ? ldb < k : ldb < n) || ldc < n)
    return fail(OPTK_EINVAL, "gemm: leading dimension too small");
  if ((!trans_a && lda % 8) || (trans_b && ldb % 8)) return fail(OPTK_EINVAL, "gemm: contiguous operands need leading dimensions that are multiples of 8");
  if ((!trans_a && ((uintptr_t)a % 16)) || (trans_b && ((uintptr_t)b % 16)))
    return fail(OPTK_EINVAL, "gemm: operands must be 16-byte aligned");
  hipStream_t stream = (hipStream_t)hip_stream;
  float *target = split_k > 1 ? workspace : c;
  const int64_t kc = k / split_k;
  const uint16_t *A = (const uint16_t *)a, *B = (const uint16_t *)b;
  if (m % GM == 0 && n % GN == 0) {
    // 128 x 128 tiles; the transposed operands also need 16-byte aligned rows
    const bool rows_ok = (trans_a ? (lda % 8 == 0 && (uintptr_t)a % 16 == 0) : true) &&
                         (!trans_b ? (ldb % 8 == 0 && (uintptr_t)b % 16 == 0) : true);
    if (rows_ok) {
      GemmEpilogue E;
      std::memset(&E, 0, sizeof(E));
      E.c = target;
      E.ldc = ldc;
      dim3 grid((unsigned)(n / GN), (unsigned)(m / GM), (unsigned)split_k);
      if (!trans_a && !trans_b)
        hipLaunchKernelGGL((gemm_bf16_128<0, 0, EPI_STORE>), grid, dim3(256), 0, stream, A, lda, B, ldb, E, m, n, kc);
      else if (!trans_a && trans_b)
        hipLaunchKernelGGL((gemm_bf16_128<0, 1, EPI_STORE>), grid, dim3(256), 0, stream, A, lda, B, ldb, E, m, n, kc);
      else if (trans_a && !trans_b)
        hipLaunchKernelGGL((gemm_bf16_128<1, 0, EPI_STORE>), grid, dim3(256), 0, stream, A, lda, B, ldb, E, m, n, kc);
      else
        hipLaunchKernelGGL((gemm_bf16_128<1, 1, EPI_STORE>), grid, dim3(256), 0, stream, A, lda, B, ldb, E, m, n, kc);
      hipError_t e = hipGetLastError();
      if (e == hipSuccess && split_k > 1) {
        if (ldc != n) return fail(OPTK_EINVAL, "gemm: split_k > 1 needs ldc == n");
        hipLaunchKernelGGL(gemm_reduce, dim3(blocks_for(m * n)), dim3(256), 0, stream, workspace,
                           split_k, m * n, c);
        e = hipGetLastError();
      }
      if (e != hipSuccess) return fail(OPTK_EDEVICE, std::string("gemm launch: ") + hipGetErrorString(e));
      return OPTK_OK;
    }
  }
  dim3 grid((unsigned)(n / BN), (unsigned)(m / BM), (unsigned)split_k);
  if (!trans_a && !trans_b)
    hipLaunchKernelGGL((gemm_bf16<0, 0>), grid, dim3(256), 0, stream, A, lda, B, ldb, target, ldc, m, n, kc);
  else if (!trans_a && trans_b)
    hipLaunchKernelGGL((gemm_bf16<0, 1>), grid, dim3(256), 0, stream, A, lda, B, ldb, target, ldc, m, n, kc);
  else if (trans_a && !trans_b)
    hipLaunchKernelGGL((gemm_bf16<1, 0>), grid, dim3(256), 0, stream, A, lda, B, ldb, target, ldc, m, n, kc);
  else
    hipLaunchKernelGGL((gemm_bf16<1, 1>), grid, dim3(256), 0, stream, A, lda, B, ldb, target, ldc, m, n, kc);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && split_k > 1) {
    if (ldc != n) return fail(OPTK_EINVAL, "gemm: split_k > 1 needs ldc == n");
    hipLaunchKernelGGL(gemm_reduce, dim3(blocks_for(m * n)), dim3(256), 0, stream, workspace,
                       split_k, m * n, c);
    e = hipGetLastError();
  }
  if (e != hipSuccess) return fail(OPTK_EDEVICE, std::string("gemm launch: ") + hipGetErrorString(e));
  return OPTK_OK;
}

int32_t optk_gemm_bf16_fused(int32_t trans_a, int32_t trans_b, int64_t m, int64_t n, int64_t k,
                             const void *a, int64_t lda, const void *b, int64_t ldb, float *c,
                             int64_t ldc, int32_t epilogue, const float *bias, const float *pre,
                             void *out16, void *hip_stream) {
  if (!a || !b || !c || !out16) return fail(OPTK_EINVAL, "gemm_fused: null operand");
  if (epilogue != OPTK_GEMM_BIAS_SILU && epilogue != OPTK_GEMM_SILU_BACK)
    return fail(OPTK_EINVAL, "gemm_fused: unknown epilogue");
  if (epilogue == OPTK_GEMM_BIAS_SILU && !bias) return fail(OPTK_EINVAL, "gemm_fused: bias missing");
  if (epilogue == OPTK_GEMM_SILU_BACK && !pre) return fail(OPTK_EINVAL, "gemm_fused: pre missing");
  if (m <= 0 || n <= 0 || k <= 0 || m % GM || n % GN || k % GK)
    return fail(OPTK_EINVAL, "gemm_fused: m and n must be positive multiples of 128, k of 32");
  if ((trans_a ? lda < m : lda < k) || (trans_b ? ldb < k : ldb < n) || ldc < n)
    return fail(OPTK_EINVAL, "gemm_fused: leading dimension too small");
  if (lda % 8 || ldb % 8 || (uintptr_t)a % 16 || (uintptr_t)b % 16)
    return fail(OPTK_EINVAL, "gemm_fused: operands need 16-byte aligned rows (leading dimensions multiples of 8)");
  GemmEpilogue E;
  E.c = c;
  E.ldc = ldc;
  E.bias = bias;
  E.pre = pre;
  E.out16 = (uint16_t *)out16;
  const uint16_t *A = (const uint16_t *)a, *B = (const uint16_t *)b;
  hipStream_t stream = (hipStream_t)hip_stream;
  dim3 grid((unsigned)(n / GN), (unsigned)(m / GM), 1);
#define OPTK_GEMM_FUSED(TA, TB)                                                                    \
  if (epilogue == OPTK_GEMM_BIAS_SILU)                                                             \
    hipLaunchKernelGGL((gemm_bf16_128<TA, TB, EPI_BIAS_SILU>), grid, dim3(256), 0, stream, A, lda, B, ldb, E, m, n, k); \
  else                                                                                             \
    hipLaunchKernelGGL((gemm_bf16_128<TA, TB, EPI_SILU_BACK>), grid, dim3(256), 0, stream, A, lda, B, ldb, E, m, n, k);
  if (!trans_a && !trans_b) {
    OPTK_GEMM_FUSED(0, 0)
  } else if (!trans_a && trans_b) {
    OPTK_GEMM_FUSED(0, 1)
  } else if (trans_a && !trans_b) {
    OPTK_GEMM_FUSED(1, 0)
  } else {
    OPTK_GEMM_FUSED(1, 1)
  }
#undef OPTK_GEMM_FUSED
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(OPTK_EDEVICE, std::string("gemm_fused launch: ") + hipGetErrorString(e));
  return OPTK_OK;
}

int32_t optk_mlp_op(int32_t op, const optk_mlp_args *A, void *hip_stream) {
  if (!A) return fail(OPTK_EINVAL, "null argument");
  hipStream_t s = (hipStream_t)hip_stream;
  const int64_t rows = A->rows;
  if (rows < 0 || A->cols < 0 || A->inner < 0 ||
      (A->inner > 8 && (op == OPTK_MLP_DENSE_IN || op == OPTK_MLP_DENSE_OUT || op == OPTK_MLP_DENSE_OUT_BACK)))
    return fail(OPTK_EINVAL, "mlp op: bad sizes");
  const int64_t count = rows * A->cols;
  if (count == 0 && op != OPTK_MLP_ADAM) return OPTK_OK;
#define NEED(cond, msg)                     \
  do {                                      \
    if (!(cond)) return fail(OPTK_EINVAL, msg); \
  } while (0)
  switch (op) {
  case OPTK_MLP_BIAS_SILU:
    NEED(A->f32[0] && A->f32[1] && A->u16[0], "bias_silu: pre, bias, act");
    hipLaunchKernelGGL(bias_silu, dim3(blocks_for(count)), dim3(256), 0, s, A->f32[0], A->f32[1],
                       A->u16[0], rows, A->cols);
    break;
  case OPTK_MLP_DENSE_IN:
    NEED(A->f32[0] && A->f32[1] && A->f32[2] && A->u16[0], "dense_in: x, w, b, act");
    hipLaunchKernelGGL(dense_in, dim3(blocks_for(count)), dim3(256), 0, s, A->f32[0], A->inner,
                       A->f32[1], A->f32[2], A->f32[3], A->u16[0], rows, A->cols);
    break;
  case OPTK_MLP_DENSE_OUT:
    NEED(A->u16[0] && A->f32[0] && A->f32[1] && A->f32[2], "dense_out: act, w, b, y");
    hipLaunchKernelGGL(dense_out, dim3(blocks_for(rows)), dim3(256), 0, s, A->u16[0], A->cols,
                       A->f32[0], A->f32[1], A->f32[2], A->inner, rows);
    break;
  case OPTK_MLP_MSE_GRAD:
    NEED(A->f32[0] && A->f32[1] && A->f32[2] && A->f32[3] && A->f32[4], "mse: y, t, w, dy, loss");
    hipLaunchKernelGGL(mse_grad, dim3(blocks_for(count)), dim3(256), 0, s, A->f32[0], A->f32[1],
                       A->f32[2], A->f32[3], A->cols, rows, A->f32[4]);
    break;
  case OPTK_MLP_DENSE_OUT_BACK:
    NEED(A->f32[0] && A->f32[1] && A->f32[2] && A->u16[0] && A->f32[3], "dense_out_back: dy, w, pre, dpre, dpre32");
    hipLaunchKernelGGL(dense_out_back, dim3(blocks_for(count)), dim3(256), 0, s, A->f32[0], A->inner,
                       A->f32[1], A->f32[2], A->u16[0], A->f32[3], rows, A->cols);
    break;
  case OPTK_MLP_SILU_BACK:
    NEED(A->f32[0] && A->f32[1] && A->u16[0] && A->f32[2], "silu_back: dact, pre, dpre, dpre32");
    hipLaunchKernelGGL(silu_back, dim3(blocks_for(count)), dim3(256), 0, s, A->f32[0], A->f32[1],
                       A->u16[0], A->f32[2], count);
    break;
  case OPTK_MLP_OUTER_GRAD: {
    // f32[0] or u16[0]: x [rows][cols]; f32[1]: g [rows][inner];
    // f32[2]: out (cols * inner); f32[3]: workspace (blocks * cols * inner)
    NEED((A->f32[0] || A->u16[0]) && A->f32[1] && A->f32[2] && A->f32[3], "outer_grad: x, g, out, workspace");
    NEED(A->cols <= 256 && A->inner >= 1 && A->inner <= 8, "outer_grad: cols <= 256, 1 <= inner <= 8");
    // 32 rows per block: enough blocks (2048 at 65536 rows) to keep every CU
    // busy; each thread's loads of its column are independent
    const int64_t rpb = 32;
    const int64_t nblk = (rows + rpb - 1) / rpb;
    NEED(A->count >= nblk * A->cols * A->inner, "outer_grad: workspace too small");
    hipLaunchKernelGGL(outer_grad, dim3((unsigned)nblk), dim3(256), 0, s, A->f32[0], A->u16[0],
                       A->cols, A->f32[1], A->inner, rows, rpb, A->flag, A->f32[3]);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(OPTK_EDEVICE, std::string("outer_grad: ") + hipGetErrorString(e));
    const int64_t cnt = (int64_t)A->cols * A->inner;
    hipLaunchKernelGGL(gemm_reduce, dim3(blocks_for(cnt)), dim3(256), 0, s, A->f32[3], (int)nblk,
                       cnt, A->f32[2]);
    break;
  }
  case OPTK_MLP_COLUMN_SUM:
    NEED(A->f32[0] && A->f32[1], "column_sum: g, gb");
    hipLaunchKernelGGL(column_sum, dim3((unsigned)A->cols), dim3(256), 0, s, A->f32[0], rows,
                       A->cols, A->f32[1]);
    break;
  case OPTK_MLP_ADAM: {
    NEED(A->f32[0] && A->f32[1] && A->f32[2] && A->f32[3], "adam: p, g, m, v");
    const int64_t nparam = A->count;
    if (nparam <= 0) return OPTK_OK;
    hipLaunchKernelGGL(adam, dim3(blocks_for(nparam)), dim3(256), 0, s, A->f32[0], A->f32[1],
                       A->f32[2], A->f32[3], A->u16[0], nparam, A->scalars[0], A->scalars[1],
                       A->scalars[2], A->scalars[3], A->scalars[4], A->scalars[5]);
    break;
  }
  case OPTK_MLP_TO_BF16:
    NEED(A->f32[0] && A->u16[0], "to_bf16: x, y");
    hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(count)), dim3(256), 0, s, A->f32[0],
                       A->u16[0], count);
    break;
  default:
    return fail(OPTK_EINVAL, "unknown mlp op");
  }
#undef NEED
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(OPTK_EDEVICE, std::string("mlp op launch: ") + hipGetErrorString(e));
  return OPTK_OK;
}

}  // extern "C"

extern "C" int32_t optk_surrogate_infer(const optk_surrogate_desc *desc, const optk_surrogate_io *io,
                                        void *hip_stream) {
  if (!desc || !io) return fail(OPTK_EINVAL, "null argument");
  const optk_surrogate_desc &D = *desc;
  const optk_surrogate_io &IO = *io;
  if (IO.n < 0) return fail(OPTK_EINVAL, "negative count");
  if (IO.n == 0) return OPTK_OK;
  if (IO.n > INT32_MAX) return fail(OPTK_EINVAL, "at most 2^31 - 1 rays per call");
  if (!IO.field_x || !IO.field_y || !IO.pupil_x || !IO.pupil_y || !IO.wavelength || !IO.x || !IO.y ||
      !IO.intensity || !IO.unvignetted || !IO.fallback)
    return fail(OPTK_EINVAL, "missing ray array");
  if (IO.fallback_list && (!IO.fallback_count || IO.fallback_capacity < 0))
    return fail(OPTK_EINVAL, "fallback list needs a counter and a capacity");
  if (!D.w1_16 || !D.b1 || !D.w5 || !D.b5 || !D.w5_16) return fail(OPTK_EINVAL, "missing weights");
  if ((uintptr_t)D.w1_16 % 16) return fail(OPTK_EINVAL, "bf16 weights must be 16-byte aligned");
  if ((uintptr_t)D.w5_16 % 16) return fail(OPTK_EINVAL, "bf16 weights must be 16-byte aligned");
  for (int l = 0; l < 3; ++l) {
    if (!D.w16[l] || !D.b[l]) return fail(OPTK_EINVAL, "missing weights");
    if ((uintptr_t)D.w16[l] % 16) return fail(OPTK_EINVAL, "bf16 weights must be 16-byte aligned");
  }
  if (D.num_targets != 3 + 2 * D.num_apertures || D.num_apertures < 0 ||
      D.num_apertures > OPTK_SURROGATE_MAX_APERTURES)
    return fail(OPTK_EINVAL, "num_targets must be 3 + 2 * num_apertures");
  for (int a = 0; a < D.num_apertures; ++a)
    if (D.apertures[a].num_vertices < 0 || D.apertures[a].num_vertices > OPTK_SURROGATE_MAX_VERTICES)
      return fail(OPTK_EINVAL, "aperture vertex count out of range");
  int dev = 0, cus = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return fail(OPTK_EDEVICE, std::string("device query: ") + hipGetErrorString(e));
  const int64_t tiles = (IO.n + INF_RAYS - 1) / INF_RAYS;
  int64_t blocks = cus > 0 ? (int64_t)cus * INF_BLOCKS_PER_CU : 1024;  // all resident
  if (blocks > tiles) blocks = tiles;
  hipLaunchKernelGGL(surrogate_infer, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)hip_stream, D, IO);
  e = hipGetLastError();
  if (e != hipSuccess) return fail(OPTK_EDEVICE, std::string("surrogate launch: ") + hipGetErrorString(e));
  return OPTK_OK;
}

extern "C" size_t optk_poly_workspace(int64_t n) {
  (void)n;
  return sizeof(double) * (size_t)POLY_BLOCKS * POLY_N * POLY_COLS;
}

static int32_t poly_check(const double *u, int64_t n, int32_t nt, const uint32_t *poly_exp,
                          const double *out, PolyExp &X) {
  if (!u || !out || !poly_exp) return fail(OPTK_EINVAL, "null argument");
  if (n < 0) return fail(OPTK_EINVAL, "negative count");
  if (nt < 1 || nt > OPTK_SURROGATE_MAX_TARGETS) return fail(OPTK_EINVAL, "num_targets out of range");
  for (int m = 0; m < POLY_N; ++m) {
    uint32_t e = poly_exp[m], deg = 0;
    if (e >> 10) return fail(OPTK_EINVAL, "poly_exp: 2 bits for each of 5 inputs");
    for (int k = 0; k < 5; ++k) deg += (e >> (2 * k)) & 3u;
    if (deg > 3) return fail(OPTK_EINVAL, "poly_exp: total degree above 3");
    X.e[m] = e;
  }
  return OPTK_OK;
}

extern "C" int32_t optk_poly_normal(const double *u, const double *targets, const uint8_t *valid,
                                    int64_t n, int32_t num_targets, const uint32_t *poly_exp,
                                    double *out, void *workspace, size_t workspace_bytes,
                                    void *hip_stream) {
  PolyExp X;
  int32_t rc = poly_check(u, n, num_targets, poly_exp, out, X);
  if (rc) return rc;
  if (!targets || !valid || !workspace) return fail(OPTK_EINVAL, "null argument");
  if (workspace_bytes < optk_poly_workspace(n)) return fail(OPTK_EINVAL, "workspace too small");
  const int entries = POLY_N * (POLY_N + num_targets);
  const int64_t tiles = (n + POLY_TILE - 1) / POLY_TILE;
  const int blocks = (int)(tiles < POLY_BLOCKS ? (tiles > 0 ? tiles : 1) : POLY_BLOCKS);
  hipStream_t s = (hipStream_t)hip_stream;
  double *partial = (double *)workspace;
  hipLaunchKernelGGL(poly_normal_partial, dim3(blocks), dim3(256), 0, s, u, targets, valid, n,
                     (int)num_targets, X, partial);
  hipLaunchKernelGGL(poly_normal_reduce, dim3((entries + 255) / 256), dim3(256), 0, s, partial,
                     blocks, entries, out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(OPTK_EDEVICE, std::string("poly normal: ") + hipGetErrorString(e));
  return OPTK_OK;
}

extern "C" int32_t optk_poly_eval(const double *u, const double *coef, int64_t n, int32_t num_targets,
                                  const uint32_t *poly_exp, double *out, void *hip_stream) {
  PolyExp X;
  int32_t rc = poly_check(u, n, num_targets, poly_exp, out, X);
  if (rc) return rc;
  if (!coef) return fail(OPTK_EINVAL, "null argument");
  if (n == 0) return OPTK_OK;
  hipLaunchKernelGGL(poly_eval_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)hip_stream, u, coef, n, (int)num_targets, X, out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(OPTK_EDEVICE, std::string("poly eval: ") + hipGetErrorString(e));
  return OPTK_OK;
}
