// Device helpers for gfx950 (wave64): block ids remapped per XCD, wave
// reductions, block scans, sorted-list search and MurmurHash2.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#define FSIG_BLOCK 256

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X, each
// with its own L2.  Remapping makes XCD x own the contiguous logical range
// [x*n/8, (x+1)*n/8), so neighbouring rows (which share parents, hence
// candidate lists) meet in one L2.
__device__ __forceinline__ uint32_t xcd_block() {
  const uint32_t b = blockIdx.x, n = gridDim.x;
  if ((n & 7u) != 0) return b;
  return (b & 7u) * (n >> 3) + (b >> 3);
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

// Exclusive scan of one u64 per thread over a 256-thread block; `total`
// receives the block sum.  lds must hold 4 words.  All threads must call it.
__device__ __forceinline__ uint64_t block_exclusive_scan(uint64_t v, uint64_t* lds, uint64_t& total) {
  const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
  uint64_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint64_t y = __shfl_up(x, o, 64);
    if (lane >= (uint32_t)o) x += y;
  }
  if (lane == 63) lds[wid] = x;
  __syncthreads();
  uint64_t pre = 0, tot = 0;
#pragma unroll
  for (uint32_t w = 0; w < FSIG_BLOCK / 64; w++) {
    uint64_t s = lds[w];
    if (w < wid) pre += s;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return pre + x - v;
}

__device__ __forceinline__ uint64_t block_sum(uint64_t v, uint64_t* lds) {
  uint64_t t;
  (void)block_exclusive_scan(v, lds, t);
  return t;
}

// Membership of x in the ascending list a[0..n).
__device__ __forceinline__ bool sorted_contains(const uint32_t* __restrict__ a, uint32_t n, uint32_t x) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo < n && a[lo] == x;
}

// Index of x in the ascending list a[0..n) (x must be present).
__device__ __forceinline__ uint32_t sorted_index(const uint32_t* __restrict__ a, uint32_t n, uint32_t x) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// MurmurHash2 of one 4-byte key (the published algorithm for len == 4; the
// reference's device copy is match/GSI_kernels.cuh:61-93).
__device__ __forceinline__ uint32_t murmur2_u32(uint32_t key, uint32_t seed) {
  const uint32_t mul = 0x5bd1e995u;
  uint32_t h = seed ^ 4u;
  uint32_t w = key * mul;
  w ^= w >> 24;
  w *= mul;
  h = (h * mul) ^ w;
  h ^= h >> 13;
  h *= mul;
  h ^= h >> 15;
  return h;
}

// Order-independent digest of an embedding set.  A row (data vertex ids in
// query-vertex order) hashes to x by chaining the splitmix64 finaliser
// (Steele, Lea and Flood; public domain) over its words; the set's digest is
// the pair (sum of x, sum of mix64(x ^ kDigestSalt)) modulo 2^64, which does
// not depend on the order rows are produced in.
constexpr uint64_t kDigestSalt = 0xd6e8feb86659fd93ull;

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

__host__ __device__ __forceinline__ uint64_t row_hash_start(uint32_t k) {
  return 0x9e3779b97f4a7c15ull * (uint64_t)(k + 1);
}

__host__ __device__ __forceinline__ uint64_t row_hash_step(uint64_t x, uint32_t i, uint32_t v) {
  return mix64(x + (uint64_t)v + ((uint64_t)i << 32));
}
