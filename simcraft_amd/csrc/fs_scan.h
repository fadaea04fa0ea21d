// fs_scan.h - exclusive prefix sum of int counts on the device, in three passes: per-tile sums
// (2048 elements per 256-thread block), one block scanning the tile sums, and every tile adding
// its offset to its own exclusive scan. Deterministic; no atomics and no host synchronisation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fsscan {

constexpr int TPB = 256;
constexpr int ITEMS = 8;               // elements per thread
constexpr int TILE = TPB * ITEMS;      // elements per block

inline int64_t tiles_for(int64_t n) { return (n + TILE - 1) / TILE; }

// exclusive scan of one value per thread over the block (Hillis-Steele in LDS); *total = sum
__device__ inline int block_exclusive(int v, int* sm, int* total) {
  sm[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < TPB; off <<= 1) {
    const int t = threadIdx.x >= (unsigned)off ? sm[threadIdx.x - off] : 0;
    __syncthreads();
    sm[threadIdx.x] += t;
    __syncthreads();
  }
  *total = sm[TPB - 1];
  const int ex = sm[threadIdx.x] - v;
  __syncthreads();
  return ex;
}

static __global__ __launch_bounds__(TPB) void k_tile_sums(int64_t n, const int* __restrict__ in,
                                                          int* __restrict__ tile_sum) {
  __shared__ int sm[TPB];
  const int64_t b = (int64_t)blockIdx.x * TILE + (int64_t)threadIdx.x * ITEMS;
  int s = 0;
#pragma unroll
  for (int k = 0; k < ITEMS; k++)
    if (b + k < n) s += in[b + k];
  int total;
  block_exclusive(s, sm, &total);
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = total;
}

// one block: exclusive scan of the tile sums in place, 256 at a time; the grand total after them
static __global__ __launch_bounds__(TPB) void k_scan_tile_sums(int ntiles, int* __restrict__ tile_sum) {
  __shared__ int sm[TPB];
  int carry = 0;
  for (int base = 0; base < ntiles; base += TPB) {
    const int i = base + threadIdx.x;
    const int v = i < ntiles ? tile_sum[i] : 0;
    int t;
    const int ex = block_exclusive(v, sm, &t);
    if (i < ntiles) tile_sum[i] = carry + ex;
    carry += t;
  }
  if (threadIdx.x == 0) tile_sum[ntiles] = carry;
}

static __global__ __launch_bounds__(TPB) void k_apply(int64_t n, const int* __restrict__ in,
                                                      const int* __restrict__ tile_off, int* __restrict__ out) {
  __shared__ int sm[TPB];
  const int64_t b = (int64_t)blockIdx.x * TILE + (int64_t)threadIdx.x * ITEMS;
  int v[ITEMS], s = 0;
#pragma unroll
  for (int k = 0; k < ITEMS; k++) {
    v[k] = b + k < n ? in[b + k] : 0;
    s += v[k];
  }
  int total;
  int run = tile_off[blockIdx.x] + block_exclusive(s, sm, &total);
#pragma unroll
  for (int k = 0; k < ITEMS; k++) {
    if (b + k < n) out[b + k] = run;
    run += v[k];
  }
}

// out[i] = in[0] + ... + in[i-1] for i < n; tiles must hold tiles_for(n) + 1 ints, and
// tiles[tiles_for(n)] receives the total. in and out may not alias.
inline hipError_t exclusive_scan(hipStream_t st, int64_t n, const int* in, int* out, int* tiles) {
  const int64_t nt = tiles_for(n);
  if (nt == 0) return hipSuccess;
  k_tile_sums<<<(int)nt, TPB, 0, st>>>(n, in, tiles);
  k_scan_tile_sums<<<1, TPB, 0, st>>>((int)nt, tiles);
  k_apply<<<(int)nt, TPB, 0, st>>>(n, in, tiles, out);
  return hipGetLastError();
}

}  // namespace fsscan
