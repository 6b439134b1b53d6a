// optk_mlp.hip -- kernels of the learned surrogate (a 5 -> 4 x 256 SiLU -> 3
// MLP): a bf16 MFMA GEMM with float32 accumulation (gfx950
// v_mfma_f32_16x16x32_bf16), the elementwise and small-layer kernels of the
// forward and backward passes, Adam, and the fused inference kernel.
//
// The surrogate has no counterpart in the reference (it replaces traces of
// the spectrograph by a learned map); the float64 trace generates its
// training data and judges its error.

#include <hip/hip_runtime.h>

#include <algorithm>
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
// writes C + z * m * n (a partial sum, reduced by reduce_partials).
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
// k step).  Operands are staged in LDS with k contiguous, double-buffered,
// and requested from global memory GPF k steps ahead (16-byte loads into
// registers), one barrier per step.  An
// operand stored with k as its row index (TA = 1, TB = 0) is loaded 16 bytes
// along m (or n) and transposed on its way into LDS.  The epilogue can fuse
// the layer's elementwise step (bias + SiLU forward, SiLU derivative
// backward), so the pre-activation is not re-read by a separate kernel.
// Requires m, n multiples of 128 and k a multiple of 32 * split.
// ---------------------------------------------------------------------------

constexpr int GM = 128, GN = 128, GK = 32, GPAD = 8, GPF = 4;

enum { EPI_STORE = 0, EPI_BIAS_SILU = 1, EPI_SILU_BACK = 2 };

struct GemmEpilogue {
  float *c;          // EPI_STORE: C (or split partials); BIAS_SILU: pre out; SILU_BACK: dpre32 out
  int64_t ldc;
  const float *bias; // BIAS_SILU
  const float *pre;  // SILU_BACK: pre-activation in, [m][ldc]
  uint16_t *out16;   // BIAS_SILU: bf16 activations; SILU_BACK: bf16 dpre
  float *sums = nullptr;  // row-panel SILU_BACK: per-tile column sums of dpre ([tile][256]), or null
};

// 16 bytes as a vector register value (a plain vector type, which the
// compiler keeps in registers where an array of structs may go to scratch)
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// one 128 x 32 operand tile into LDS ([row][k], k contiguous); KC: stored
// with k contiguous (x[row * ld + k]), else k as the row index (x[k * ld + row])
template <bool KC> struct TileLoad {
  static __device__ __forceinline__ void load(u32x4 (&v)[2], const uint16_t *__restrict__ x, int64_t ld,
                                              int64_t row0, int64_t k0, int tid) {
    if (KC) {
      const int r = tid >> 2, kk = (tid & 3) * 8;
#pragma unroll
      for (int h = 0; h < 2; ++h)
        v[h] = *reinterpret_cast<const u32x4 *>(x + (row0 + r + 64 * h) * ld + k0 + kk);
    } else {
      // k rows 2 kp and 2 kp + 1, tile rows 8 r .. 8 r + 7
      const int kp = tid & 15, r = tid >> 4;
#pragma unroll
      for (int h = 0; h < 2; ++h)
        v[h] = *reinterpret_cast<const u32x4 *>(x + (k0 + 2 * kp + h) * ld + row0 + 8 * r);
    }
  }
  static __device__ __forceinline__ void store(const u32x4 (&v)[2], uint16_t (*t)[GK + GPAD], int tid) {
    if (KC) {
      const int r = tid >> 2, kk = (tid & 3) * 8;
#pragma unroll
      for (int h = 0; h < 2; ++h) *reinterpret_cast<u32x4 *>(&t[r + 64 * h][kk]) = v[h];
    } else {
      // tile row 8 r + j gets the word (k = 2 kp, k = 2 kp + 1): one 32-bit
      // store per row.  Within a wave kp runs over 16 and r over 4 values,
      // so a store's 64 lanes hit at most two lanes per bank (rows of 20
      // words, 8 rows apart: 32 banks apart)
      const int kp = tid & 15, r = tid >> 4;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const uint32_t lo = v[0][q], hi = v[1][q];
        *reinterpret_cast<uint32_t *>(&t[8 * r + 2 * q][2 * kp]) = (lo & 0xffffu) | (hi << 16);
        *reinterpret_cast<uint32_t *>(&t[8 * r + 2 * q + 1][2 * kp]) = (lo >> 16) | (hi & 0xffff0000u);
      }
    }
  }
};

template <int TA, int TB, int EPI>
__global__ __launch_bounds__(256) void gemm_bf16_128(const uint16_t *__restrict__ a, int64_t lda,
                                                     const uint16_t *__restrict__ b, int64_t ldb,
                                                     const GemmEpilogue E, int64_t m, int64_t n,
                                                     int64_t kc) {
  // operand tiles during the k loop; the epilogue reuses the same memory
  // to stage half of the C tile (64 x 128 floats)
  __shared__ __attribute__((aligned(16))) uint16_t smem[2 * 2 * GM * (GK + GPAD)];
  auto As = reinterpret_cast<uint16_t(*)[GM][GK + GPAD]>(smem);
  auto Bs = reinterpret_cast<uint16_t(*)[GN][GK + GPAD]>(smem + 2 * GM * (GK + GPAD));
  auto Cs = reinterpret_cast<float(*)[GN + 4]>(smem);
  static_assert(64 * (GN + 4) * sizeof(float) <= sizeof(smem), "C half tile fits the operand tiles");
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
  // Operand tiles are requested GPF k steps ahead into registers (slot j
  // holds the steps s with s % GPF == j) and moved into the LDS double
  // buffer one step ahead of their use: a block keeps GPF x 16 KB of loads
  // in flight, which HBM's latency needs at these tile sizes.
  typedef TileLoad<TA == 0> LA;
  typedef TileLoad<TB == 1> LB;
  u32x4 ra[GPF][2], rb[GPF][2];
  const int steps = (int)(kc / GK);
#pragma unroll
  for (int u = 0; u < GPF; ++u)
    if (u < steps) {
      LA::load(ra[u], a, lda, m0, k_begin + (int64_t)u * GK, tid);
      LB::load(rb[u], b, ldb, n0, k_begin + (int64_t)u * GK, tid);
    }
  LA::store(ra[0], As[0], tid);
  LB::store(rb[0], Bs[0], tid);
  if (GPF < steps) {
    LA::load(ra[0], a, lda, m0, k_begin + (int64_t)GPF * GK, tid);
    LB::load(rb[0], b, ldb, n0, k_begin + (int64_t)GPF * GK, tid);
  }
  __syncthreads();
  for (int s0 = 0; s0 < steps; s0 += GPF) {
#pragma unroll
    for (int u = 0; u < GPF; ++u) {
      const int step = s0 + u;
      if (step < steps) {
        const int buf = u & 1;  // GPF is even: step & 1
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
        // step + 1 (slot (u + 1) % GPF) into the other buffer, whose last
        // reads (step - 1) ended at the previous barrier; the slot then
        // requests step + 1 + GPF
        const int nxt = step + 1, slot = (u + 1) % GPF;
        if (nxt < steps) {
          LA::store(ra[slot], As[buf ^ 1], tid);
          LB::store(rb[slot], Bs[buf ^ 1], tid);
          if (nxt + GPF < steps) {
            LA::load(ra[slot], a, lda, m0, k_begin + (int64_t)(nxt + GPF) * GK, tid);
            LB::load(rb[slot], b, ldb, n0, k_begin + (int64_t)(nxt + GPF) * GK, tid);
          }
        }
        __syncthreads();
      }
    }
  }
  // lane holds C[row = 4 (lane / 16) + r][col = lane % 16] of each 16 x 16
  // tile.  The tile leaves through LDS, 64 rows at a time, so that every
  // global access of the epilogue is a 16-byte (float4) or 8-byte (four
  // bf16) access along a row.  The k loop ended on a barrier: the operand
  // tiles are free.
  float *cz = E.c ? E.c + (int64_t)blockIdx.z * m * n : nullptr;
  const int64_t ld = gridDim.z > 1 ? n : E.ldc;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    if ((wave >> 1) == h) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) Cs[16 * i + 4 * (lane >> 4) + r][wn + 16 * j + (lane & 15)] = acc[i][j][r];
    }
    __syncthreads();
    const int cl = 4 * (tid & 31);
    const int64_t col = n0 + cl;
    float4 bias = float4{0.0f, 0.0f, 0.0f, 0.0f};
    if (EPI == EPI_BIAS_SILU) bias = *reinterpret_cast<const float4 *>(E.bias + col);
#pragma unroll 2
    for (int pass = 0; pass < 8; ++pass) {
      const int rl = 8 * pass + (tid >> 5);
      float4 v = *reinterpret_cast<const float4 *>(&Cs[rl][cl]);
      const int64_t at = (m0 + 64 * h + rl) * ld + col;
      if (EPI == EPI_BIAS_SILU) {
        v.x += bias.x;
        v.y += bias.y;
        v.z += bias.z;
        v.w += bias.w;
        *reinterpret_cast<float4 *>(cz + at) = v;
        *reinterpret_cast<uint2 *>(E.out16 + at) =
            uint2{pack_bf16(silu(v.x), silu(v.y)), pack_bf16(silu(v.z), silu(v.w))};
      } else if (EPI == EPI_SILU_BACK) {
        const float4 p = *reinterpret_cast<const float4 *>(E.pre + at);
        v.x *= silu_grad(p.x);
        v.y *= silu_grad(p.y);
        v.z *= silu_grad(p.z);
        v.w *= silu_grad(p.w);
        if (cz) *reinterpret_cast<float4 *>(cz + at) = v;
        *reinterpret_cast<uint2 *>(E.out16 + at) = uint2{pack_bf16(v.x, v.y), pack_bf16(v.z, v.w)};
      } else {
        *reinterpret_cast<float4 *>(cz + at) = v;
      }
    }
    __syncthreads();
  }
}

// out[j] = sum over s < splits of partial[s * count + j].  A block owns 64
// outputs (one per lane); its 16 waves each sum every 16th split with eight
// loads in flight, and the 16 wave sums are added in a fixed order, so the
// result does not depend on timing.  Grid: ceil(count / 64) blocks.
constexpr int RED_WAVES = 16;

__global__ __launch_bounds__(64 * RED_WAVES) void reduce_partials(const float *__restrict__ partial,
                                                                  int splits, int64_t count,
                                                                  float *__restrict__ out) {
  __shared__ float red[RED_WAVES][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t j = (int64_t)blockIdx.x * 64 + lane;
  float s = 0.0f;
  if (j < count) {
    const float *p = partial + j;
    int z = w;
    for (; z + 7 * RED_WAVES < splits; z += 8 * RED_WAVES) {
      float v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = p[(int64_t)(z + u * RED_WAVES) * count];
      s += ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
    }
    for (; z < splits; z += RED_WAVES) s += p[(int64_t)z * count];
  }
  red[w][lane] = s;
  __syncthreads();
  if (w == 0 && j < count) {
    float t = 0.0f;
#pragma unroll
    for (int u = 0; u < RED_WAVES; ++u) t += red[u][lane];
    out[j] = t;
  }
}

inline void launch_reduce(const float *partial, int splits, int64_t count, float *out, hipStream_t s) {
  hipLaunchKernelGGL(reduce_partials, dim3((unsigned)((count + 63) / 64)), dim3(64 * RED_WAVES), 0, s,
                     partial, splits, count, out);
}

// ---------------------------------------------------------------------------
// Row-panel GEMM of the network's 256-wide layers: out[m][0, 256) =
// A[m][0, 256) . Bt^T, row j of Bt holding output column j's 256 weights
// (k contiguous), with the epilogues of gemm_bf16_128.  At k = n = 256 a
// product moves more bytes than the matrix cores need time for (per row:
// 512 B of A in, 1.5 KB of pre-activation and activation out, against
// 131,072 flops), so the kernel is laid out for HBM traffic:
//  - a block owns a 64-row tile and all 256 output columns, so every A row
//    is read from HBM once, into LDS (rows padded by 16 halves);
//  - each of the four waves owns 64 columns; its weights (L2-resident)
//    stream through a register ring two 32-wide k chunks ahead;
//  - the epilogue stores straight from the accumulators: the B operand's
//    columns are permuted so that lane (fr, g) holds columns col .. col + 3
//    (col = 64 wave + 4 fr) of rows 4 g .. 4 g + 3 of each 16-row tile, and
//    one store instruction writes four 256-byte (float) or 128-byte (bf16)
//    row segments;
//  - the SiLU-derivative epilogue requests its pre-activations before the k
//    loop, so they arrive while the matrix cores run, and can sum its
//    float32 results over the tile's rows (the previous layer's bias
//    gradient, one partial per tile, added up by reduce_partials).
// Every accumulator sums the same 32-wide k chunks in the same order with
// the same MFMA as gemm_bf16_128, and the epilogue arithmetic is the same:
// the results are bit-identical.  Requires m a multiple of 64, k = n = 256.
// ---------------------------------------------------------------------------

constexpr int PMT = 4, PM = 16 * PMT, PK = 256, PPAD = 16;  // PMT 16-row tiles per block
typedef uint16_t PanelRow[PK + PPAD];

template <int EPI>
__global__ __launch_bounds__(256, 2) void gemm_panel(const uint16_t *__restrict__ a, int64_t lda,
                                                  const uint16_t *__restrict__ bt, int64_t ldb,
                                                  const GemmEpilogue E) {
  __shared__ __attribute__((aligned(16))) PanelRow As[PM];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int fr = lane & 15, fk = (lane >> 4) * 8;
  const int col = wave * 64 + 4 * fr;
  const int64_t row0 = (int64_t)blockIdx.x * PM;
  // the tile's A rows: thread t moves 16-byte segments t + 256 j, j < 2 PMT
  // (row (t + 256 j) / 32, segment (t + 256 j) % 32 of it)
  uint4 st[2 * PMT];
#pragma unroll
  for (int j = 0; j < 2 * PMT; ++j) {
    const int seg = tid + 256 * j, r = seg >> 5, s = seg & 31;
    st[j] = *reinterpret_cast<const uint4 *>(a + (row0 + r) * lda + 8 * s);
  }
  // the lane's B operands of chunk c: rows col + nt of Bt, k = 32 c + fk ..
  const uint16_t *brow = bt + (int64_t)col * ldb + fk;
  uint4 ring[3][4];
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) ring[c][nt] = *reinterpret_cast<const uint4 *>(brow + nt * ldb + 32 * c);
#pragma unroll
  for (int j = 0; j < 2 * PMT; ++j) {
    const int seg = tid + 256 * j, r = seg >> 5, s = seg & 31;
    *reinterpret_cast<uint4 *>(&As[r][8 * s]) = st[j];
  }
  float4 bias = float4{0.0f, 0.0f, 0.0f, 0.0f};
  if (EPI == EPI_BIAS_SILU) bias = *reinterpret_cast<const float4 *>(E.bias + col);
  __syncthreads();
  // row of element q of row tile mt in this lane
  auto row_of = [&](int mt, int q) { return row0 + 16 * mt + 4 * (lane >> 4) + q; };
  float4 pre[PMT][4];
  if (EPI == EPI_SILU_BACK) {
#pragma unroll
    for (int mt = 0; mt < PMT; ++mt)
#pragma unroll
      for (int q = 0; q < 4; ++q) pre[mt][q] = *reinterpret_cast<const float4 *>(E.pre + row_of(mt, q) * E.ldc + col);
  }
  floatx4 acc[PMT][4];
  float4 csum = float4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int c = 0; c < PK / 32; ++c) {
    if (c + 2 < PK / 32) {
#pragma unroll
      for (int nt = 0; nt < 4; ++nt)
        ring[(c + 2) % 3][nt] = *reinterpret_cast<const uint4 *>(brow + nt * ldb + 32 * (c + 2));
    }
    bf16x8 b[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) b[nt] = __builtin_bit_cast(bf16x8, ring[c % 3][nt]);
#pragma unroll
    for (int mt = 0; mt < PMT; ++mt) {
      const bf16x8 av = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(&As[16 * mt + fr][32 * c + fk]));
#pragma unroll
      for (int nt = 0; nt < 4; ++nt)
        acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, b[nt], c == 0 ? floatx4{0.0f, 0.0f, 0.0f, 0.0f} : acc[mt][nt], 0, 0, 0);
    }
  }
#pragma unroll
  for (int mt = 0; mt < PMT; ++mt)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int64_t at = row_of(mt, q) * E.ldc + col;
      float4 v = float4{acc[mt][0][q], acc[mt][1][q], acc[mt][2][q], acc[mt][3][q]};
      if (EPI == EPI_BIAS_SILU) {
        v.x += bias.x;
        v.y += bias.y;
        v.z += bias.z;
        v.w += bias.w;
        *reinterpret_cast<float4 *>(E.c + at) = v;
        *reinterpret_cast<uint2 *>(E.out16 + at) =
            uint2{pack_bf16(silu(v.x), silu(v.y)), pack_bf16(silu(v.z), silu(v.w))};
      } else {
        const float4 p = pre[mt][q];
        v.x *= silu_grad(p.x);
        v.y *= silu_grad(p.y);
        v.z *= silu_grad(p.z);
        v.w *= silu_grad(p.w);
        if (E.c) *reinterpret_cast<float4 *>(E.c + at) = v;
        *reinterpret_cast<uint2 *>(E.out16 + at) = uint2{pack_bf16(v.x, v.y), pack_bf16(v.z, v.w)};
        csum.x += v.x;
        csum.y += v.y;
        csum.z += v.z;
        csum.w += v.w;
      }
    }
  if (EPI == EPI_SILU_BACK && E.sums) {
    // the lane's 4 PMT rows, then the four row groups (lanes fr, fr + 16,
    // fr + 32, fr + 48) in a fixed order
#pragma unroll
    for (int o = 16; o <= 32; o *= 2) {
      csum.x += __shfl_xor(csum.x, o, 64);
      csum.y += __shfl_xor(csum.y, o, 64);
      csum.z += __shfl_xor(csum.z, o, 64);
      csum.w += __shfl_xor(csum.w, o, 64);
    }
    if (lane < 16) *reinterpret_cast<float4 *>(E.sums + (int64_t)blockIdx.x * PK + col) = csum;
  }
}

// y[c][r] = x[r][c] for a rows x cols array of 16-bit values: 32 x 32
// tiles through LDS (row padded by one to spread the banks)
__global__ __launch_bounds__(256) void transpose16(const uint16_t *__restrict__ x, uint16_t *__restrict__ y,
                                                   int64_t rows, int cols) {
  __shared__ uint16_t t[32][33];
  const int64_t r0 = (int64_t)blockIdx.y * 32;
  const int c0 = blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
#pragma unroll
  for (int j = 0; j < 32; j += 8) {
    const int64_t r = r0 + ty + j;
    const int c = c0 + tx;
    if (r < rows && c < cols) t[ty + j][tx] = x[r * cols + c];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 32; j += 8) {
    const int c = c0 + ty + j;
    const int64_t r = r0 + tx;
    if (r < rows && c < cols) y[(int64_t)c * rows + r] = t[tx][ty + j];
  }
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
// over blocks by reduce_partials.  x is float32 (x32) or bf16 (x16).  Layout of
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
#pragma unroll 4
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

// ---- vectorised forms of the small layers (four consecutive features per
// thread, 16-byte float and 8-byte bf16 accesses); the same arithmetic
// order as the scalar kernels above, which serve the shapes these do not

// dense_in for IN inputs, cols a multiple of 4
template <int IN>
__global__ __launch_bounds__(256) void dense_in4(const float *__restrict__ x, const float *__restrict__ w,
                                                 const float *__restrict__ bias, float *__restrict__ pre,
                                                 uint16_t *__restrict__ act, int64_t rows, int cols) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int cq = cols / 4;
  if (q >= rows * cq) return;
  const int64_t r = q / cq;
  const int o = (int)(q % cq) * 4;
  float xv[IN];
#pragma unroll
  for (int k = 0; k < IN; ++k) xv[k] = x[r * IN + k];
  float v[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    v[u] = bias[o + u];
#pragma unroll
    for (int k = 0; k < IN; ++k) v[u] += xv[k] * w[(o + u) * IN + k];
  }
  const int64_t at = r * cols + o;
  if (pre) *reinterpret_cast<float4 *>(pre + at) = float4{v[0], v[1], v[2], v[3]};
  *reinterpret_cast<uint2 *>(act + at) =
      uint2{pack_bf16(silu(v[0]), silu(v[1])), pack_bf16(silu(v[2]), silu(v[3]))};
}

// dense_out for hidden == 256 and OUT outputs: one wave per row, lane l
// holding features 4 l .. 4 l + 3; the 64 lane sums meet in a butterfly
template <int OUT>
__global__ __launch_bounds__(256) void dense_out_wave(const uint16_t *__restrict__ act,
                                                      const float *__restrict__ w,
                                                      const float *__restrict__ bias,
                                                      float *__restrict__ y, int64_t rows) {
  const int lane = threadIdx.x & 63;
  const int64_t first = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t stride = ((int64_t)gridDim.x * blockDim.x) >> 6;
  float wl[OUT][4];
#pragma unroll
  for (int j = 0; j < OUT; ++j)
#pragma unroll
    for (int u = 0; u < 4; ++u) wl[j][u] = w[j * 256 + 4 * lane + u];
  for (int64_t r = first; r < rows; r += stride) {
    const uint2 a = *reinterpret_cast<const uint2 *>(act + r * 256 + 4 * lane);
    const float h[4] = {__uint_as_float(a.x << 16), __uint_as_float(a.x & 0xffff0000u),
                        __uint_as_float(a.y << 16), __uint_as_float(a.y & 0xffff0000u)};
    float s[OUT];
#pragma unroll
    for (int j = 0; j < OUT; ++j) {
      s[j] = (h[0] * wl[j][0] + h[1] * wl[j][1]) + (h[2] * wl[j][2] + h[3] * wl[j][3]);
#pragma unroll
      for (int m = 32; m > 0; m >>= 1) s[j] += __shfl_xor(s[j], m, 64);
    }
    if (lane == 0) {
#pragma unroll
      for (int j = 0; j < OUT; ++j) y[r * OUT + j] = s[j] + bias[j];
    }
  }
}

// dense_out_back for OUT outputs and 256 hidden features: a wave takes rows
// in turn (lane l: features 4 l .. 4 l + 3), writes the bf16 dpre and, if
// asked, the float32 one, and sums its rows' float32 dpre per feature (the
// layer's bias gradient); the block's four waves add theirs in a fixed
// order into partial[block][256]
constexpr int DOB_BLOCKS = 1024;

template <int OUT>
__global__ __launch_bounds__(256) void dense_out_back4(const float *__restrict__ dy,
                                                       const float *__restrict__ w,
                                                       const float *__restrict__ pre,
                                                       uint16_t *__restrict__ dpre,
                                                       float *__restrict__ dpre32, int64_t rows,
                                                       float *__restrict__ partial) {
  __shared__ float4 red[4][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int h = 4 * lane;
  float wl[OUT][4];
#pragma unroll
  for (int j = 0; j < OUT; ++j)
#pragma unroll
    for (int u = 0; u < 4; ++u) wl[j][u] = w[j * 256 + h + u];
  float4 cs = float4{0.0f, 0.0f, 0.0f, 0.0f};
  const int64_t nw = (int64_t)gridDim.x * 4;
  for (int64_t r = (int64_t)blockIdx.x * 4 + wave; r < rows; r += nw) {
    float d[OUT];
#pragma unroll
    for (int j = 0; j < OUT; ++j) d[j] = dy[r * OUT + j];
    const int64_t at = r * 256 + h;
    const float4 p = *reinterpret_cast<const float4 *>(pre + at);
    const float pv[4] = {p.x, p.y, p.z, p.w};
    float g[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      float sum = 0.0f;
#pragma unroll
      for (int j = 0; j < OUT; ++j) sum += d[j] * wl[j][u];
      g[u] = sum * silu_grad(pv[u]);
    }
    *reinterpret_cast<uint2 *>(dpre + at) = uint2{pack_bf16(g[0], g[1]), pack_bf16(g[2], g[3])};
    if (dpre32) *reinterpret_cast<float4 *>(dpre32 + at) = float4{g[0], g[1], g[2], g[3]};
    cs.x += g[0];
    cs.y += g[1];
    cs.z += g[2];
    cs.w += g[3];
  }
  if (!partial) return;
  red[wave][lane] = cs;
  __syncthreads();
  if (wave == 0) {
    float4 t = red[0][lane];
#pragma unroll
    for (int v = 1; v < 4; ++v) {
      t.x += red[v][lane].x;
      t.y += red[v][lane].y;
      t.z += red[v][lane].z;
      t.w += red[v][lane].w;
    }
    *reinterpret_cast<float4 *>(partial + (int64_t)blockIdx.x * 256 + h) = t;
  }
}

// column sums of a narrow float32 array (COLS <= 8 columns) times one
// column of g: out[c] = sum_r x[r][c] g[r], per block over its rows, the
// block's waves added in a fixed order into partial[block][COLS]
template <int COLS>
__global__ __launch_bounds__(256) void narrow_outer1(const float *__restrict__ x, const float *__restrict__ g,
                                                     int64_t rows, int64_t rows_per_block,
                                                     float *__restrict__ partial) {
  __shared__ float red[4][COLS];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  float acc[COLS];
#pragma unroll
  for (int c = 0; c < COLS; ++c) acc[c] = 0.0f;
  const int64_t r0 = (int64_t)blockIdx.x * rows_per_block;
  const int64_t r1 = r0 + rows_per_block < rows ? r0 + rows_per_block : rows;
  for (int64_t r = r0 + threadIdx.x; r < r1; r += 256) {
    const float gv = g[r];
#pragma unroll
    for (int c = 0; c < COLS; ++c) acc[c] += x[r * COLS + c] * gv;
  }
#pragma unroll
  for (int c = 0; c < COLS; ++c)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc[c] += __shfl_xor(acc[c], o, 64);
  if (lane == 0)
#pragma unroll
    for (int c = 0; c < COLS; ++c) red[wave][c] = acc[c];
  __syncthreads();
  if (threadIdx.x < COLS)
    partial[(int64_t)blockIdx.x * COLS + threadIdx.x] =
        (red[0][threadIdx.x] + red[1][threadIdx.x]) + (red[2][threadIdx.x] + red[3][threadIdx.x]);
}

// outer_grad for cols a multiple of 4 (<= 256) and INNER columns of g:
// thread (wave p, lane c) sums rows r0 + p, r0 + p + 4, ... of columns
// 4 c .. 4 c + 3; the four row phases are added in LDS in a fixed order
template <int INNER, bool X16>
__global__ __launch_bounds__(256) void outer_grad4(const float *__restrict__ x32,
                                                   const uint16_t *__restrict__ x16, int cols,
                                                   const float *__restrict__ g, int64_t rows,
                                                   int64_t rows_per_block, int transposed,
                                                   float *__restrict__ partial) {
  __shared__ float red[4][256 * INNER];
  const int c4 = 4 * (threadIdx.x & 63), phase = threadIdx.x >> 6;
  float acc[4][INNER];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int j = 0; j < INNER; ++j) acc[u][j] = 0.0f;
  const int64_t r0 = (int64_t)blockIdx.x * rows_per_block;
  const int64_t r1 = r0 + rows_per_block < rows ? r0 + rows_per_block : rows;
  if (c4 < cols) {
#pragma unroll 2
    for (int64_t r = r0 + phase; r < r1; r += 4) {
      float xv[4];
      if (X16) {
        const uint2 a = *reinterpret_cast<const uint2 *>(x16 + r * cols + c4);
        xv[0] = __uint_as_float(a.x << 16);
        xv[1] = __uint_as_float(a.x & 0xffff0000u);
        xv[2] = __uint_as_float(a.y << 16);
        xv[3] = __uint_as_float(a.y & 0xffff0000u);
      } else {
        const float4 a = *reinterpret_cast<const float4 *>(x32 + r * cols + c4);
        xv[0] = a.x;
        xv[1] = a.y;
        xv[2] = a.z;
        xv[3] = a.w;
      }
#pragma unroll
      for (int j = 0; j < INNER; ++j) {
        const float gj = g[r * INNER + j];
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[u][j] += xv[u] * gj;
      }
    }
  }
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int j = 0; j < INNER; ++j) red[phase][(c4 + u) * INNER + j] = acc[u][j];
  __syncthreads();
  float *out = partial + (int64_t)blockIdx.x * cols * INNER;
  for (int idx = threadIdx.x; idx < cols * INNER; idx += blockDim.x) {
    const float v = (red[0][idx] + red[1][idx]) + (red[2][idx] + red[3][idx]);
    const int c = idx / INNER, j = idx % INNER;
    out[transposed ? idx : j * cols + c] = v;
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
// Block: 256 threads (4 waves), one per CU, a persistent loop over pairs
// of 128-ray tiles (A, B), each with its own LDS activation buffer.  The
// two tiles' layers are staggered by half a layer: while a wave issues
// tile A's MFMAs of layer L it applies bias + SiLU to tile B's layer L - 1
// accumulators (row tile c during k chunk c), and the other way round in
// the next stage, so each wave keeps its matrix core and its VALU busy at
// once; one barrier separates the stages.  The two accumulator sets take
// 256 registers, which the one wave per SIMD leaves room for.
//
// Every layer runs on MFMA 16x16x32 bf16 with float32 accumulation.
// Layer 1 (5 -> 256): the normalised inputs and the weights are split into
// bf16 high and low parts, k row [hi | lo | hi] against weight row
// [w_hi | w_hi | w_lo], so the one 32-wide k step carries the float32
// product to ~2^-16.  Hidden layers: a 128 x 256 x 256 product per tile.
// Wave w computes all 128 rays x features [64 w, 64 w + 64) as 8 x 4
// tiles, A operands (activations act[ray][feature], bf16) from LDS, B
// operands (weights, row-major [out][in], k contiguous) straight from
// global memory (L2-resident), streamed through three register buffers
// two 32-wide k chunks ahead of their MFMAs (across stages too).  The B operand
// of n-tile t and lane column c is weight row 64 w + 4 c + t, so each
// lane's four accumulator tiles hold four consecutive features: the SiLU
// writes them as one 8-byte store per row.  The output layer (3 rows, high
// and low bf16 parts) is one more MFMA tile per 16 rays.  The polynomial,
// the edge tests, the error map and the stores run one thread per ray.
// ---------------------------------------------------------------------------

constexpr int INF_RAYS = 128;          // rays per tile; a block holds two tiles
constexpr int INF_MT = INF_RAYS / 16;  // 16-ray MFMA row tiles per tile
constexpr int INF_OQ = INF_MT / 4;     // output-layer row tiles per wave
constexpr int INF_BLOCKS_PER_CU = 1;
constexpr int INF_HIDDEN = 256;
constexpr int INF_PAD = 16;
constexpr int INF_TARGETS = OPTK_SURROGATE_MAX_TARGETS;

// The kernel's order of the 56 cubic monomials: depth first over
// nondecreasing variable indices (1, u0, u0^2, u0^3, u0^2 u1, ..., u4^3).
// Monomial t is u[a] u[b] u[c] with u[5] = 1; make_poly_factors().f[t]
// packs 4 a, 4 b and 4 c in its low three bytes.  The launch maps every
// kernel-order monomial to its row of the caller's poly.
struct PolyOrder {
  uint8_t row[OPTK_SURROGATE_POLY];
};
struct PolyFactors {
  uint32_t f[OPTK_SURROGATE_POLY];
};

constexpr uint32_t poly_factor(int a, int b, int c) {
  return (uint32_t)(4 * a) | ((uint32_t)(4 * b) << 8) | ((uint32_t)(4 * c) << 16);
}

constexpr PolyFactors make_poly_factors() {
  PolyFactors p{};
  int t = 0;
  p.f[t++] = poly_factor(5, 5, 5);
  for (int a = 0; a < 5; ++a) {
    p.f[t++] = poly_factor(a, 5, 5);
    for (int b = a; b < 5; ++b) {
      p.f[t++] = poly_factor(a, b, 5);
      for (int c = b; c < 5; ++c) p.f[t++] = poly_factor(a, b, c);
    }
  }
  return p;
}

// exponent codes (2 bits per input) of the kernel-order monomials
static void kernel_poly_codes(uint32_t (&codes)[OPTK_SURROGATE_POLY]) {
  constexpr PolyFactors p = make_poly_factors();
  for (int t = 0; t < OPTK_SURROGATE_POLY; ++t) {
    uint32_t code = 0;
    for (int k = 0; k < 3; ++k) {
      const int v = (int)((p.f[t] >> (8 * k)) & 255u) / 4;
      if (v < 5) code += 1u << (2 * v);
    }
    codes[t] = code;
  }
}

// The polynomial of a tile pair runs on the float32 matrix cores
// (v_mfma_f32_16x16x4_f32): each thread writes its ray's 56 monomials to
// LDS (rows of POLY_MONO_STRIDE floats, bank-conflict free for the A
// operand reads), each wave multiplies its 64 rays' rows by the [56][16]
// coefficient matrix (targets padded to 16 columns with zeros), and the
// products go back to LDS, a row of 16 per ray.  A wave only reads the
// rows of its own rays.
constexpr int INF_COEF_STRIDE = 16;   // coefficient row: the targets, zero padded
constexpr int POLY_MONO_STRIDE = 60;  // floats per monomial row (56 used)

typedef float PolyRow[POLY_MONO_STRIDE];

// the ray's monomials in kernel order into its LDS row
__device__ __forceinline__ void poly_monomials(const float (&u)[5], float *row) {
  constexpr PolyFactors PF = make_poly_factors();
  const float U[6] = {u[0], u[1], u[2], u[3], u[4], 1.0f};
  float m[OPTK_SURROGATE_POLY];
#pragma unroll
  for (int t = 0; t < OPTK_SURROGATE_POLY; ++t) {
    const uint32_t f = PF.f[t];
    m[t] = (U[(f & 255u) / 4] * U[((f >> 8) & 255u) / 4]) * U[(f >> 16) / 4];
  }
#pragma unroll
  for (int q = 0; q < OPTK_SURROGATE_POLY / 4; ++q)
    *reinterpret_cast<float4 *>(row + 4 * q) = float4{m[4 * q], m[4 * q + 1], m[4 * q + 2], m[4 * q + 3]};
}

// targets of rays [row0, row0 + 64) (four 16-row tiles): out[ray][j] =
// sum_t mono[ray][t] coef[t][j]
__device__ __forceinline__ void poly_mfma(const PolyRow *mono, int row0, const float (*coef)[INF_COEF_STRIDE],
                                          float (*out)[INF_COEF_STRIDE], int lane) {
  const int fr = lane & 15, fk = lane >> 4;
  floatx4 acc[4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) acc[mt] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int ks = 0; ks < OPTK_SURROGATE_POLY / 4; ++ks) {
    const float b = coef[4 * ks + fk][fr];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
      acc[mt] = __builtin_amdgcn_mfma_f32_16x16x4f32(mono[row0 + 16 * mt + fr][4 * ks + fk], b, acc[mt], 0, 0, 0);
  }
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int r = 0; r < 4; ++r) out[row0 + 16 * mt + 4 * fk + r][fr] = acc[mt][r];
}

typedef uint16_t ActRow[INF_HIDDEN + INF_PAD];
typedef floatx4 Acc[INF_MT][4];

// two floats -> two bfloat16 (round to nearest even) in one word, a low
__device__ __forceinline__ uint32_t pack2_bf16(float a, float b) {
  typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
  typedef float float2_t __attribute__((ext_vector_type(2)));
  return __builtin_bit_cast(uint32_t, __builtin_convertvector((float2_t{a, b}), bf16x2_t));
}

// this lane's B operands of one 32-wide k chunk: rows col .. col + 3 of W
// ([out][K] bf16), k values k0 + fk .. k0 + fk + 7.  voff = (col K + fk) * 2
// bytes: one 32-bit offset against the uniform base, so that no per-lane
// 64-bit row addresses stay live across the tile loop
template <int K>
__device__ __forceinline__ void wload(uint4 (&w)[4], const uint16_t *__restrict__ W, uint32_t voff, int k0) {
  const char *base = reinterpret_cast<const char *>(W) + 2 * k0;
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) w[nt] = *reinterpret_cast<const uint4 *>(base + voff + 2 * K * nt);
}

// acc (+)= act[:, k0 .. k0 + 32) x the chunk's weights (FIRST: acc = c0 +
// product, c0[nt] the scaled bias of the lane's column in n-tile nt)
template <bool FIRST, typename Row>
__device__ __forceinline__ void mfma_chunk(Acc &acc, const uint4 (&w)[4], const Row *act, int k0,
                                           int fr, int fk, const floatx4 (&c0)[4]) {
  bf16x8 b[4];
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) b[nt] = __builtin_bit_cast(bf16x8, w[nt]);
#pragma unroll
  for (int mt = 0; mt < INF_MT; ++mt) {
    const bf16x8 a = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(&act[16 * mt + fr][k0 + fk]));
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
      acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[nt], FIRST ? c0[nt] : acc[mt][nt], 0, 0, 0);
  }
}

// The kernel carries every pre-activation and activation scaled by
// log2(e): v' = log2(e) v, s = log2(e) silu(v) = v' / (1 + 2^-v').  The
// scale enters through the inputs (u' = log2(e) u) and the biases, passes
// through the hidden layers unchanged (their inputs carry it), and leaves
// at the output layer (times ln 2); the SiLU then needs no multiply before
// its exponential and no bias add (the bias starts the accumulator).
constexpr float INF_LOG2E = 1.44269504088896341f, INF_LN2 = 0.693147180559945309f;

__device__ __forceinline__ float silu_scaled(float v) {
  return v * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-v));
}

// act[row][col .. col + 3] = bf16(silu_scaled(acc)) for the lane's rows of
// row tile mt
__device__ __forceinline__ void silu_rows(const Acc &acc, ActRow *act, int col, int lane, int mt) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int row = 16 * mt + 4 * (lane >> 4) + q;
    uint2 v;
    v.x = pack2_bf16(silu_scaled(acc[mt][0][q]), silu_scaled(acc[mt][1][q]));
    v.y = pack2_bf16(silu_scaled(acc[mt][2][q]), silu_scaled(acc[mt][3][q]));
    *reinterpret_cast<uint2 *>(&act[row][col]) = v;
  }
}

// the accumulator start of a layer: log2(e) times the bias of the lane's
// four columns, one per n-tile
__device__ __forceinline__ void bias_init(floatx4 (&c0)[4], const float *__restrict__ bias, uint32_t col) {
  const char *base = reinterpret_cast<const char *>(bias);
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    const float b = INF_LOG2E * *reinterpret_cast<const float *>(base + 4 * col + 4 * nt);
    c0[nt] = floatx4{b, b, b, b};
  }
}

// one row (q) of row tile mt of acc through the SiLU into act (four
// features, one 8-byte store)
__device__ __forceinline__ void silu_row(const Acc &acc, ActRow *act, int col, int lane, int mt, int q) {
  const int row = 16 * mt + 4 * (lane >> 4) + q;
  uint2 v;
  v.x = pack2_bf16(silu_scaled(acc[mt][0][q]), silu_scaled(acc[mt][1][q]));
  v.y = pack2_bf16(silu_scaled(acc[mt][2][q]), silu_scaled(acc[mt][3][q]));
  *reinterpret_cast<uint2 *>(&act[row][col]) = v;
}

// one stage: accx = c0 + actx . W (K = 256, eight 32-wide k chunks), and,
// while the MFMAs run, row tile c of accy through the SiLU into acty.
// Weights stream through a ring of three register buffers: chunk c sits in
// slot (S + c) % 3 and requests chunk c + 2 before its MFMAs; past the end
// the requests fetch the next stage's first two chunks from Wn, which
// therefore starts at slot (S + 2) % 3.  Wn must be a valid weight matrix
// (the last stage passes its own: those requests go unused).  The A
// operands (activation rows) are read one chunk ahead: row tile mt of
// chunk c + 1 replaces that of chunk c as soon as its four MFMAs are
// issued.  Each chunk is cut into four scheduling regions of two row tiles
// (eight MFMAs) and one SiLU row (four values): a small region lets the
// scheduler put the SiLU's VALU work between the MFMAs, which execute
// meanwhile.
template <int S>
__device__ __forceinline__ void stage(Acc &accx, uint4 (&ring)[3][4], const uint16_t *__restrict__ W,
                                      const uint16_t *__restrict__ Wn, const ActRow *actx, uint32_t voff,
                                      const floatx4 (&c0)[4], const Acc &accy, ActRow *acty, int col,
                                      int lane, int fr, int fk) {
  uint4 a[INF_MT];
#pragma unroll
  for (int mt = 0; mt < INF_MT; ++mt) a[mt] = *reinterpret_cast<const uint4 *>(&actx[16 * mt + fr][fk]);
  __builtin_amdgcn_sched_barrier(0);  // chunk 0's operands first, in a region of their own
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    if (c + 2 < 8)
      wload<INF_HIDDEN>(ring[(S + c + 2) % 3], W, voff, 32 * (c + 2));
    else
      wload<INF_HIDDEN>(ring[(S + c + 2) % 3], Wn, voff, 32 * (c + 2 - 8));
    bf16x8 b[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) b[nt] = __builtin_bit_cast(bf16x8, ring[(S + c) % 3][nt]);
#pragma unroll
    for (int mt = 0; mt < INF_MT; ++mt) {
      const bf16x8 am = __builtin_bit_cast(bf16x8, a[mt]);
#pragma unroll
      for (int nt = 0; nt < 4; ++nt)
        accx[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, b[nt], c == 0 ? c0[nt] : accx[mt][nt], 0, 0, 0);
      if (c + 1 < 8) a[mt] = *reinterpret_cast<const uint4 *>(&actx[16 * mt + fr][32 * (c + 1) + fk]);
      if (mt & 1) {
        silu_row(accy, acty, col, lane, c, mt >> 1);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
}

// the output layer of one tile: wave w -> rays [16 INF_OQ w, +16 INF_OQ).
// The three outputs' weights are split into bf16 high and low parts (rows
// 0..2 and 3..5 of one 16-row tile, the rest zero), so the products carry
// the float32 weights to ~2^-17 at no extra MFMA
__device__ __forceinline__ void output_layer(const uint4 (&w5)[INF_HIDDEN / 32], const ActRow *act,
                                             float (*yout)[8], int wave, int lane, int fr, int fk) {
  floatx4 oacc[INF_OQ];
#pragma unroll
  for (int q = 0; q < INF_OQ; ++q) oacc[q] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int kc = 0; kc < INF_HIDDEN / 32; ++kc) {
    const bf16x8 b5 = __builtin_bit_cast(bf16x8, w5[kc]);
#pragma unroll
    for (int q = 0; q < INF_OQ; ++q) {
      const uint4 va = *reinterpret_cast<const uint4 *>(&act[16 * (INF_OQ * wave + q) + fr][32 * kc + fk]);
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

__global__ __launch_bounds__(256, INF_BLOCKS_PER_CU) void surrogate_infer(const optk_surrogate_desc D,
                                                                          const optk_surrogate_io IO,
                                                                          const PolyOrder P) {
  __shared__ __attribute__((aligned(16))) ActRow act[2][INF_RAYS];
  __shared__ __attribute__((aligned(16))) float yout[2][INF_RAYS][8];
  // tile A's layer-1 input rows live in yout's memory (free until the
  // output layer), so that tile A's layer-1 SiLU may write act[0] while
  // other waves still read those rows
  typedef uint16_t InRow[32];
  auto ain = reinterpret_cast<InRow *>(&yout[0][0][0]);
  static_assert(INF_RAYS * sizeof(InRow) <= sizeof(yout), "tile A's input rows fit yout");
  __shared__ float us[2][INF_RAYS][5];  // normalised inputs
  __shared__ __attribute__((aligned(16))) float coef[OPTK_SURROGATE_POLY][INF_COEF_STRIDE];  // kernel order
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int fr = lane & 15, fk = (lane >> 4) * 8;
  const int col = wave * 64 + 4 * fr;  // the lane's four output features
  const uint32_t voff_h = 2u * (uint32_t)(col * INF_HIDDEN + fk), voff1_h = 2u * (uint32_t)(col * 32 + fk);
  const uint32_t voff5_h = 2u * (uint32_t)(fr * INF_HIDDEN + fk), bcol_h = (uint32_t)col;
  const int64_t n = IO.n;
  const int64_t pairs = (n + 2 * INF_RAYS - 1) / (2 * INF_RAYS);
  const float *in[5] = {IO.field_x, IO.field_y, IO.pupil_x, IO.pupil_y, IO.wavelength};
  for (int idx = tid; idx < OPTK_SURROGATE_POLY * INF_COEF_STRIDE; idx += blockDim.x) {
    const int t = idx / INF_COEF_STRIDE, j = idx % INF_COEF_STRIDE;
    coef[t][j] = j < INF_TARGETS ? D.poly[P.row[t] * INF_TARGETS + j] : 0.0f;
  }
  // the error map and, per polygon edge from vertex v - 1 to v, its start,
  // direction, 1 / |direction|^2 and inverse slope dx / dy (0 for a
  // horizontal edge, which the crossing rule never uses)
  __shared__ uint8_t emap[16 * 8 * 8];
  __shared__ float edge[OPTK_SURROGATE_MAX_APERTURES][OPTK_SURROGATE_MAX_VERTICES][6];
  for (int idx = tid; idx < 16 * 8 * 8; idx += blockDim.x) emap[idx] = D.error_map ? D.error_map[idx] : 0;
  if (tid < OPTK_SURROGATE_MAX_APERTURES * OPTK_SURROGATE_MAX_VERTICES) {
    const int ap = tid / OPTK_SURROGATE_MAX_VERTICES, v = tid % OPTK_SURROGATE_MAX_VERTICES;
    const int nv = D.apertures[ap].num_vertices;
    if (ap < D.num_apertures && v < nv) {
      const int w = v == 0 ? nv - 1 : v - 1;
      const float x1 = D.apertures[ap].vx[w], y1 = D.apertures[ap].vy[w];
      const float ex = D.apertures[ap].vx[v] - x1, ey = D.apertures[ap].vy[v] - y1;
      edge[ap][v][0] = x1;
      edge[ap][v][1] = y1;
      edge[ap][v][2] = ex;
      edge[ap][v][3] = ey;
      edge[ap][v][4] = 1.0f / (ex * ex + ey * ey);
      edge[ap][v][5] = ey != 0.0f ? ex / ey : 0.0f;
    }
  }
  // the first pair's inputs; each pair requests the next one's in its tail
  float nxt[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    const int64_t i0 = (int64_t)blockIdx.x * 2 * INF_RAYS + tid;
    nxt[k] = i0 < n ? in[k][i0] : 0.0f;
  }
  for (int64_t pair = blockIdx.x; pair < pairs; pair += gridDim.x) {
    // rays [base, base + 128) form tile A, [base + 128, base + 256) tile B;
    // thread t stages, and at the end finishes, ray base + t
    const int64_t base = pair * 2 * INF_RAYS;
    const int X = tid >> 7, r = tid & (INF_RAYS - 1);
    const int64_t i = base + tid;
    // opaque per pair: the weight and bias addresses are formed here, not
    // hoisted out of the loop as dozens of live 64-bit pointers
    uint32_t voff = voff_h, voff1 = voff1_h, voff5 = voff5_h, bcol = bcol_h;
    asm volatile("" : "+v"(voff), "+v"(voff1), "+v"(voff5), "+v"(bcol));
    uint4 w1[4], ring[3][4];
    wload<32>(w1, D.w1_16, voff1, 0);
    wload<INF_HIDDEN>(ring[0], D.w16[0], voff, 0);
    wload<INF_HIDDEN>(ring[1], D.w16[0], voff, 32);
    floatx4 c0[4];
    bias_init(c0, D.b1, bcol);
    // ---- inputs, normalised and split into bf16 high and low parts, as
    // the 32-wide k row [hi | lo | hi | 0] that meets the first layer's
    // weight rows [w_hi | w_hi | w_lo | 0] ---------------------------------
    {
      float hi[5], lo[5];
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        const float u = (nxt[k] - D.in_mid[k]) * D.in_inv_half[k];
        us[X][r][k] = u;
        const float us2 = INF_LOG2E * u;
        hi[k] = from_bf16(to_bf16(us2));
        lo[k] = us2 - hi[k];
      }
      uint4 row0, row1;
      row0.x = pack2_bf16(hi[0], hi[1]);
      row0.y = pack2_bf16(hi[2], hi[3]);
      row0.z = pack2_bf16(hi[4], lo[0]);
      row0.w = pack2_bf16(lo[1], lo[2]);
      row1.x = pack2_bf16(lo[3], lo[4]);
      row1.y = pack2_bf16(hi[0], hi[1]);
      row1.z = pack2_bf16(hi[2], hi[3]);
      row1.w = pack2_bf16(hi[4], 0.0f);
      uint16_t *dst = X == 0 ? &ain[r][0] : &act[1][r][0];
      *reinterpret_cast<uint4 *>(dst) = row0;
      *reinterpret_cast<uint4 *>(dst + 8) = row1;
      *reinterpret_cast<uint4 *>(dst + 16) = uint4{0u, 0u, 0u, 0u};
      *reinterpret_cast<uint4 *>(dst + 24) = uint4{0u, 0u, 0u, 0u};
    }
    __syncthreads();
    Acc accA, accB;
    // ---- the staggered layers ---------------------------------------------
    mfma_chunk<true>(accA, w1, ain, 0, fr, fk, c0);     // A layer 1
    mfma_chunk<true>(accB, w1, act[1], 0, fr, fk, c0);  // B layer 1 | A SiLU 1
#pragma unroll
    for (int mt = 0; mt < INF_MT; ++mt) silu_rows(accA, act[0], bcol, lane, mt);
    __syncthreads();
    bias_init(c0, D.b[0], bcol);
    stage<0>(accA, ring, D.w16[0], D.w16[0], act[0], voff, c0, accB, act[1], bcol, lane, fr, fk);  // A 2 | B 1
    __syncthreads();
    stage<2>(accB, ring, D.w16[0], D.w16[1], act[1], voff, c0, accA, act[0], bcol, lane, fr, fk);  // B 2 | A 2
    __syncthreads();
    bias_init(c0, D.b[1], bcol);
    stage<1>(accA, ring, D.w16[1], D.w16[1], act[0], voff, c0, accB, act[1], bcol, lane, fr, fk);  // A 3 | B 2
    __syncthreads();
    stage<0>(accB, ring, D.w16[1], D.w16[2], act[1], voff, c0, accA, act[0], bcol, lane, fr, fk);  // B 3 | A 3
    __syncthreads();
    bias_init(c0, D.b[2], bcol);
    stage<2>(accA, ring, D.w16[2], D.w16[2], act[0], voff, c0, accB, act[1], bcol, lane, fr, fk);  // A 4 | B 3
    __syncthreads();
    stage<1>(accB, ring, D.w16[2], D.w16[2], act[1], voff, c0, accA, act[0], bcol, lane, fr, fk);  // B 4 | A 4
    uint4 w5[INF_HIDDEN / 32];
#pragma unroll
    for (int kc = 0; kc < INF_HIDDEN / 32; ++kc)
      w5[kc] = *reinterpret_cast<const uint4 *>(reinterpret_cast<const char *>(D.w5_16) + 64 * kc + voff5);
    __syncthreads();
    output_layer(w5, act[0], yout[0], wave, lane, fr, fk);  // A output | B SiLU 4
#pragma unroll
    for (int mt = 0; mt < INF_MT; ++mt) silu_rows(accB, act[1], bcol, lane, mt);
    __syncthreads();
    output_layer(w5, act[1], yout[1], wave, lane, fr, fk);  // B output
    // ---- the next pair's inputs, requested now: they arrive during the tail
    {
      const int64_t inext = (pair + gridDim.x) * 2 * INF_RAYS + tid;
#pragma unroll
      for (int k = 0; k < 5; ++k) nxt[k] = inext < n ? in[k][inext] : 0.0f;
    }
    // ---- polynomial on the matrix cores (activations are dead: the
    // monomials go to act[0]'s memory, the targets to act[1]'s) -------------
    auto mono = reinterpret_cast<PolyRow *>(&act[0][0][0]);
    auto ptgt = reinterpret_cast<float(*)[INF_COEF_STRIDE]>(&act[1][0][0]);
    static_assert(2 * INF_RAYS * POLY_MONO_STRIDE * sizeof(float) <= sizeof(act[0]), "monomials fit");
    static_assert(2 * INF_RAYS * INF_COEF_STRIDE * sizeof(float) <= sizeof(act[1]), "targets fit");
    {
      float u[5];
#pragma unroll
      for (int k = 0; k < 5; ++k) u[k] = us[X][r][k];
      poly_monomials(u, mono[tid]);  // act[0] is free: tile A's output layer is done
    }
    __syncthreads();  // every wave has read act[1] (tile B's output layer)
    // a wave reads only its own rays' monomials and targets, and a wave's
    // LDS accesses take effect in order: no barrier around the product
    poly_mfma(mono, 64 * wave, coef, ptgt, lane);
    // ---- apertures, error map, stores: one thread per ray ------------------
    bool fallback = false;
    const bool live = i < n;
    float tgt[INF_TARGETS];
#pragma unroll
    for (int j = 0; j < INF_TARGETS; ++j) tgt[j] = ptgt[tid][j];
    if (live) {
      // the activations carry log2(e): the output layer's sums times ln 2
      const float y0 = (yout[X][r][0] + yout[X][r][3]) * INF_LN2 + D.b5[0];
      const float y1 = (yout[X][r][1] + yout[X][r][4]) * INF_LN2 + D.b5[1];
      const float y2 = (yout[X][r][2] + yout[X][r][5]) * INF_LN2 + D.b5[2];
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
          // distance to the nearest edge (the minimum of the squares, one
          // square root) and the crossing rule, from the edge table
          float dmin2 = INFINITY;
          bool in_poly = false;
          for (int v = 0; v < ap.num_vertices; ++v) {
            const float *e = edge[a][v];
            const float x1 = e[0], y1 = e[1], ex = e[2], ey = e[3];
            float s = ((ax - x1) * ex + (ay - y1) * ey) * e[4];
            s = fminf(fmaxf(s, 0.0f), 1.0f);
            const float dx = ax - (x1 + s * ex), dy = ay - (y1 + s * ey);
            dmin2 = fminf(dmin2, dx * dx + dy * dy);
            if (((y1 > ay) != (y1 + ey > ay)) && (ax < e[5] * (ay - y1) + x1)) in_poly = !in_poly;
          }
          d = in_poly ? sqrtf(dmin2) : -sqrtf(dmin2);
        } else {
          d = 0.0f;  // no distance model: always near
        }
        if (ap.inverted) d = -d;
        inside = inside && d > 0.0f;
        near = near || fabsf(d) < D.edge_margin;
      }
      // error-map cell: 16 wavelength x 8 field x x 8 field y bins of [-1, 1]
      int cw = (int)((us[X][r][4] + 1.0f) * 8.0f), cx = (int)((us[X][r][0] + 1.0f) * 4.0f),
          cy = (int)((us[X][r][1] + 1.0f) * 4.0f);
      cw = cw < 0 ? 0 : (cw > 15 ? 15 : cw);
      cx = cx < 0 ? 0 : (cx > 7 ? 7 : cx);
      cy = cy < 0 ? 0 : (cy > 7 ? 7 : cy);
      const bool flagged = emap[(cw * 8 + cx) * 8 + cy] != 0;
      fallback = near || flagged || !(fabsf(x) < INFINITY && fabsf(yy) < INFINITY && fabsf(li) < INFINITY);
      IO.x[i] = x;
      IO.y[i] = yy;
      IO.intensity[i] = __expf(li);
      IO.unvignetted[i] = inside ? 1 : 0;
      IO.fallback[i] = fallback ? 1 : 0;
    }
    // ---- compaction of the fallback rays: one atomic per wave ------------
    if (IO.fallback_list) {
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
    __syncthreads();  // act, us and yout are rewritten by the next pair
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
    // (the epilogue stores float4 along rows of C)
    const bool rows_ok = (trans_a ? (lda % 8 == 0 && (uintptr_t)a % 16 == 0) : true) &&
                         (!trans_b ? (ldb % 8 == 0 && (uintptr_t)b % 16 == 0) : true) &&
                         (split_k > 1 || (ldc % 4 == 0 && (uintptr_t)c % 16 == 0)) &&
                         (split_k == 1 || (uintptr_t)workspace % 16 == 0);
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
        launch_reduce(workspace, split_k, m * n, c, stream);
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
    launch_reduce(workspace, split_k, m * n, c, stream);
    e = hipGetLastError();
  }
  if (e != hipSuccess) return fail(OPTK_EDEVICE, std::string("gemm launch: ") + hipGetErrorString(e));
  return OPTK_OK;
}

int32_t optk_gemm_bf16_fused(int32_t trans_a, int32_t trans_b, int64_t m, int64_t n, int64_t k,
                             const void *a, int64_t lda, const void *b, int64_t ldb, float *c,
                             int64_t ldc, int32_t epilogue, const float *bias, const float *pre,
                             void *out16, void *hip_stream) {
  if (!a || !b || !out16) return fail(OPTK_EINVAL, "gemm_fused: null operand");
  if (!c && epilogue != OPTK_GEMM_SILU_BACK) return fail(OPTK_EINVAL, "gemm_fused: c missing");
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
  if (ldc % 4 || (uintptr_t)c % 16 || (uintptr_t)out16 % 8 || (uintptr_t)bias % 16 || (uintptr_t)pre % 16)
    return fail(OPTK_EINVAL, "gemm_fused: c, pre and bias need 16-byte, out16 8-byte alignment and ldc a multiple of 4");
  GemmEpilogue E;
  E.c = c;
  E.ldc = ldc;
  E.bias = bias;
  E.pre = pre;
  E.out16 = (uint16_t *)out16;
  const uint16_t *A = (const uint16_t *)a, *B = (const uint16_t *)b;
  hipStream_t stream = (hipStream_t)hip_stream;
  if (!trans_a && trans_b && n == PK && k == PK && m % PM == 0) {
    // the network's 256-wide layers: row panels (every A row read once)
    if (epilogue == OPTK_GEMM_BIAS_SILU)
      hipLaunchKernelGGL(gemm_panel<EPI_BIAS_SILU>, dim3((unsigned)(m / PM)), dim3(256), 0, stream, A, lda, B, ldb, E);
    else
      hipLaunchKernelGGL(gemm_panel<EPI_SILU_BACK>, dim3((unsigned)(m / PM)), dim3(256), 0, stream, A, lda, B, ldb, E);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(OPTK_EDEVICE, std::string("gemm_fused launch: ") + hipGetErrorString(e));
    return OPTK_OK;
  }
  dim3 grid((unsigned)(n / GN), (unsigned)(m / GM), 1);
#define OPTK_GEMM_FUSED(TA, TB)                                                                  \
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

int32_t optk_gemm_bf16_silu_back_sums(int64_t m, const void *a, int64_t lda, const void *b, int64_t ldb,
                                      const float *pre, int64_t ldc, float *c, void *out16, float *sums,
                                      float *workspace, void *hip_stream) {
  if (!a || !b || !pre || !out16 || !sums || !workspace)
    return fail(OPTK_EINVAL, "gemm_silu_back_sums: null operand");
  if (m <= 0 || m % PM) return fail(OPTK_EINVAL, "gemm_silu_back_sums: m must be a positive multiple of 64");
  if (lda < PK || ldb < PK || ldc < PK || lda % 8 || ldb % 8 || ldc % 4)
    return fail(OPTK_EINVAL, "gemm_silu_back_sums: leading dimensions below 256 or misaligned");
  if ((uintptr_t)a % 16 || (uintptr_t)b % 16 || (uintptr_t)pre % 16 || (uintptr_t)c % 16 ||
      (uintptr_t)out16 % 8 || (uintptr_t)workspace % 16)
    return fail(OPTK_EINVAL, "gemm_silu_back_sums: a, b, pre, c and workspace need 16-byte, out16 8-byte alignment");
  GemmEpilogue E;
  E.c = c;
  E.ldc = ldc;
  E.bias = nullptr;
  E.pre = pre;
  E.out16 = (uint16_t *)out16;
  E.sums = workspace;
  hipStream_t stream = (hipStream_t)hip_stream;
  hipLaunchKernelGGL(gemm_panel<EPI_SILU_BACK>, dim3((unsigned)(m / PM)), dim3(256), 0, stream,
                     (const uint16_t *)a, lda, (const uint16_t *)b, ldb, E);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) {
    launch_reduce(workspace, (int)(m / PM), PK, sums, stream);
    e = hipGetLastError();
  }
  if (e != hipSuccess) return fail(OPTK_EDEVICE, std::string("gemm_silu_back_sums launch: ") + hipGetErrorString(e));
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
    if (A->inner == 5 && A->cols % 4 == 0 && (uintptr_t)A->f32[3] % 16 == 0 && (uintptr_t)A->u16[0] % 8 == 0)
      hipLaunchKernelGGL(dense_in4<5>, dim3(blocks_for(count / 4)), dim3(256), 0, s, A->f32[0], A->f32[1],
                         A->f32[2], A->f32[3], A->u16[0], rows, A->cols);
    else
      hipLaunchKernelGGL(dense_in, dim3(blocks_for(count)), dim3(256), 0, s, A->f32[0], A->inner,
                         A->f32[1], A->f32[2], A->f32[3], A->u16[0], rows, A->cols);
    break;
  case OPTK_MLP_DENSE_OUT:
    NEED(A->u16[0] && A->f32[0] && A->f32[1] && A->f32[2], "dense_out: act, w, b, y");
    if (A->cols == 256 && A->inner == 3 && (uintptr_t)A->u16[0] % 8 == 0) {
      // four rows per wave
      const int64_t nb = std::min<int64_t>((rows + 15) / 16, 8192);
      hipLaunchKernelGGL(dense_out_wave<3>, dim3((unsigned)nb), dim3(256), 0, s, A->u16[0], A->f32[0],
                         A->f32[1], A->f32[2], rows);
    } else {
      hipLaunchKernelGGL(dense_out, dim3(blocks_for(rows)), dim3(256), 0, s, A->u16[0], A->cols,
                         A->f32[0], A->f32[1], A->f32[2], A->inner, rows);
    }
    break;
  case OPTK_MLP_MSE_GRAD:
    NEED(A->f32[0] && A->f32[1] && A->f32[2] && A->f32[3] && A->f32[4], "mse: y, t, w, dy, loss");
    hipLaunchKernelGGL(mse_grad, dim3(blocks_for(count)), dim3(256), 0, s, A->f32[0], A->f32[1],
                       A->f32[2], A->f32[3], A->cols, rows, A->f32[4]);
    break;
  case OPTK_MLP_DENSE_OUT_BACK:
    NEED(A->f32[0] && A->f32[1] && A->f32[2] && A->u16[0] && (A->f32[3] || A->f32[4]),
         "dense_out_back: dy, w, pre, dpre, and dpre32 or sums");
    NEED(!A->f32[4] || A->f32[5], "dense_out_back: sums need a workspace");
    if (A->inner == 3 && A->cols == 256 && (uintptr_t)A->f32[2] % 16 == 0 && (uintptr_t)A->f32[3] % 16 == 0 &&
        (uintptr_t)A->u16[0] % 8 == 0 && (uintptr_t)A->f32[5] % 16 == 0) {
      const int64_t nblk = std::min<int64_t>(DOB_BLOCKS, (rows + 3) / 4);
      NEED(!A->f32[4] || A->count >= nblk * 256, "dense_out_back: workspace too small");
      hipLaunchKernelGGL(dense_out_back4<3>, dim3((unsigned)nblk), dim3(256), 0, s, A->f32[0], A->f32[1],
                         A->f32[2], A->u16[0], A->f32[3], rows, A->f32[4] ? A->f32[5] : nullptr);
      if (A->f32[4] && hipGetLastError() == hipSuccess) launch_reduce(A->f32[5], (int)nblk, 256, A->f32[4], s);
    } else {
      NEED(A->f32[3] && !A->f32[4], "dense_out_back: this shape needs dpre32 and no sums");
      hipLaunchKernelGGL(dense_out_back, dim3(blocks_for(count)), dim3(256), 0, s, A->f32[0], A->inner,
                         A->f32[1], A->f32[2], A->u16[0], A->f32[3], rows, A->cols);
    }
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
    // about 512 blocks (two per CU), at least 32 rows each: the partials
    // stay few enough for one short reduction
    const int64_t rpb = std::max<int64_t>(32, (rows + 511) / 512);
    const int64_t nblk = (rows + rpb - 1) / rpb;
    NEED(A->count >= nblk * A->cols * A->inner, "outer_grad: workspace too small");
    const bool x16 = A->f32[0] == nullptr;
    const bool vec = A->cols % 4 == 0 &&
                     (x16 ? (uintptr_t)A->u16[0] % 8 == 0 : (uintptr_t)A->f32[0] % 16 == 0);
#define OPTK_OUTER4(IN)                                                                               \
  if (x16)                                                                                         \
    hipLaunchKernelGGL((outer_grad4<IN, true>), dim3((unsigned)nblk), dim3(256), 0, s, nullptr, A->u16[0], \
                       A->cols, A->f32[1], rows, rpb, A->flag, A->f32[3]);                        \
  else                                                                                             \
    hipLaunchKernelGGL((outer_grad4<IN, false>), dim3((unsigned)nblk), dim3(256), 0, s, A->f32[0], nullptr, \
                       A->cols, A->f32[1], rows, rpb, A->flag, A->f32[3]);
    if (!x16 && A->cols == 3 && A->inner == 1) {
      hipLaunchKernelGGL(narrow_outer1<3>, dim3((unsigned)nblk), dim3(256), 0, s, A->f32[0], A->f32[1], rows,
                         rpb, A->f32[3]);
    } else if (vec && A->inner == 1) {
      OPTK_OUTER4(1)
    } else if (vec && A->inner == 3) {
      OPTK_OUTER4(3)
    } else if (vec && A->inner == 5) {
      OPTK_OUTER4(5)
    } else {
      hipLaunchKernelGGL(outer_grad, dim3((unsigned)nblk), dim3(256), 0, s, A->f32[0], A->u16[0],
                         A->cols, A->f32[1], A->inner, rows, rpb, A->flag, A->f32[3]);
    }
#undef OPTK_OUTER4
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(OPTK_EDEVICE, std::string("outer_grad: ") + hipGetErrorString(e));
    const int64_t cnt = (int64_t)A->cols * A->inner;
    launch_reduce(A->f32[3], (int)nblk, cnt, A->f32[2], s);
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
  case OPTK_MLP_TRANSPOSE16:
    NEED(A->u16[0] && A->u16[1] && A->u16[0] != A->u16[1], "transpose16: x, y (distinct)");
    hipLaunchKernelGGL(transpose16, dim3((unsigned)((A->cols + 31) / 32), (unsigned)((rows + 31) / 32)), dim3(256),
                       0, s, A->u16[0], A->u16[1], rows, A->cols);
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
  if (!D.w1_16 || !D.b1 || !D.w5 || !D.b5 || !D.w5_16 || !D.poly) return fail(OPTK_EINVAL, "missing weights");
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
  const int64_t pairs = (IO.n + 2 * INF_RAYS - 1) / (2 * INF_RAYS);
  int64_t blocks = cus > 0 ? (int64_t)cus * INF_BLOCKS_PER_CU : 256;  // all resident
  if (blocks > pairs) blocks = pairs;
  // the caller's row of every monomial in the kernel's order
  uint32_t codes[OPTK_SURROGATE_POLY];
  kernel_poly_codes(codes);
  PolyOrder P;
  for (int t = 0; t < OPTK_SURROGATE_POLY; ++t) {
    int row = -1;
    for (int r = 0; r < OPTK_SURROGATE_POLY; ++r)
      if (D.poly_exp[r] == codes[t]) {
        if (row >= 0) return fail(OPTK_EINVAL, "poly_exp lists a monomial twice");
        row = r;
      }
    if (row < 0) return fail(OPTK_EINVAL, "poly_exp must list every monomial of degree <= 3 in the five inputs");
    P.row[t] = (uint8_t)row;
  }
  hipLaunchKernelGGL(surrogate_infer, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)hip_stream, D, IO, P);
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
