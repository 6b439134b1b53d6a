// optk_batch.h -- the persistent work loop of the batched trace, shared by
// the float64 and float32 forms (device code; include after
// optk_internal.h).
//
// A launch holds as many blocks as the device keeps resident.  Each block
// takes work tiles (block * rays_per_thread rays of one instance) from a
// global counter until the queue is empty, so instances with 1k rays and
// instances with 64k rays load-balance.  Per tile, the block stages that
// instance's records into LDS (only when the instance changes), traces its
// rays and reduces per-path statistics of the final positions with wave
// shuffles and one LDS pass: one float64 atomic per statistic per tile.
#pragma once

#include "optk_internal.h"

namespace optk {

#define OPTK_BATCH_MAX_WAVES 8  // 512 threads per block at most

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// sums acc[0..5] over the block and adds them to stats[0..5]
__device__ __forceinline__ void block_accumulate(double acc[6], double *stats,
                                                 double (*s_red)[6]) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int waves = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int k = 0; k < 6; ++k) acc[k] = wave_sum(acc[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < 6; ++k) s_red[wave][k] = acc[k];
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    double s = 0.0;
    for (int w = 0; w < waves; ++w) s += s_red[w][threadIdx.x];
    if (s != 0.0) unsafeAtomicAdd(&stats[threadIdx.x], s);
  }
  __syncthreads();
}

// Tracer: stage(instance) copies that instance's records into LDS (all
// threads), trace(i, path, x, y, intensity, unvignetted) traces ray i.
template <class Tracer>
__device__ __forceinline__ void batch_loop(const BatchArgs &A, Tracer &tr) {
  __shared__ int s_tile;
  __shared__ double s_red[OPTK_BATCH_MAX_WAVES][6];
  int current = -1;
  const int tid = threadIdx.x;
  const int block = blockDim.x;
  for (;;) {
    if (tid == 0) s_tile = atomicAdd(A.counter, 1);
    __syncthreads();
    const int64_t t = s_tile;
    if (t >= A.num_tiles) break;  // uniform: every thread read the same s_tile
    const BatchTile tile = A.tiles[t];
    if (tile.instance != current) {
      tr.stage(tile.instance);
      __syncthreads();
      current = tile.instance;
    }
    for (int path = 0; path < tr.num_paths(); ++path) {
      double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      for (int j = 0; j < A.rays_per_thread; ++j) {
        const int local = tid + j * block;
        if (local < tile.count) {
          double x, y, inten;
          bool unv;
          tr.trace(tile.start + local, path, x, y, inten, unv);
          if (unv) {
            acc[0] += 1.0;
            acc[1] += inten;
            acc[2] += x;
            acc[3] += y;
            acc[4] += x * x;
            acc[5] += y * y;
          }
        }
      }
      block_accumulate(acc, A.stats + ((int64_t)tile.instance * A.num_paths + path) * OPTK_BATCH_STATS,
                       s_red);
    }
    // every thread is past its reads of s_tile and of the staged records
    __syncthreads();
  }
}

__device__ __forceinline__ void stage_words(void *dst, const void *src, size_t bytes) {
  uint64_t *d = reinterpret_cast<uint64_t *>(dst);
  const uint64_t *s = reinterpret_cast<const uint64_t *>(src);
  const int words = (int)(bytes / 8);
  for (int w = threadIdx.x; w < words; w += blockDim.x) d[w] = s[w];
}

}  // namespace optk
