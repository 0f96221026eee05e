// Kernels shared by the alg 1 and alg 2 pipelines: fills, segmented
// compaction, output conversion and the host-copy sink.  Everything has
// internal linkage; each translation unit gets its own copy.
#pragma once

#include <algorithm>
#include <cstring>

#include "device.h"
#include "internal.h"
#include "join.h"

namespace {

constexpr uint32_t kTile = FSIG_BLOCK * 4;

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_iota(uint32_t* __restrict__ p, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = (uint32_t)i;
}

// ---------------------------------------------------------------------------
// Order-preserving segmented compaction (one segment per class).  PRED 0
// keeps src[i] when bit `bit` of mask[src[i]] is set, PRED 1 when
// keep[keep_base + i] != 0.  Replaces the reference's flag array + CUB scan
// + scatter (match/MMatch.cu:677-702, match/MMatch_kernels.cuh:18-26, 242-254).
struct CompactSeg {
  const uint32_t* src;
  uint32_t* dst;
  uint32_t len;
  uint32_t bit;
  uint32_t tile0;
  uint32_t keep_base;
};
struct CompactTable {
  uint32_t n;
  uint32_t ntiles;
  CompactSeg s[MAXQ];
};

template <int PRED>
__device__ __forceinline__ bool compact_pred(const CompactSeg& sg, uint32_t i, const uint32_t* mask,
                                             const uint8_t* keep) {
  if (PRED == 0) return (mask[sg.src[i]] >> sg.bit) & 1u;
  return keep[sg.keep_base + i] != 0;
}

__device__ __forceinline__ uint32_t find_seg(const CompactTable& T, uint32_t tile) {
  uint32_t s = 0;
  while (s + 1 < T.n && T.s[s + 1].tile0 <= tile) s++;
  return s;
}

template <int PRED>
__global__ void __launch_bounds__(FSIG_BLOCK) k_compact_count(CompactTable T, const uint32_t* __restrict__ mask,
                                                              const uint8_t* __restrict__ keep,
                                                              uint32_t* __restrict__ tilecnt) {
  __shared__ uint64_t lds[4];
  const uint32_t s = find_seg(T, blockIdx.x);
  const CompactSeg& sg = T.s[s];
  const uint32_t base = (blockIdx.x - sg.tile0) * kTile;
  uint64_t c = 0;
  for (uint32_t j = 0; j < 4; j++) {
    const uint32_t i = base + j * FSIG_BLOCK + threadIdx.x;
    if (i < sg.len && compact_pred<PRED>(sg, i, mask, keep)) c++;
  }
  c = block_sum(c, lds);
  if (threadIdx.x == 0) tilecnt[blockIdx.x] = (uint32_t)c;
}

__global__ void __launch_bounds__(FSIG_BLOCK) k_compact_offsets(CompactTable T, uint32_t* __restrict__ tilecnt,
                                                                uint64_t* __restrict__ sizes) {
  __shared__ uint64_t lds[4];
  const CompactSeg& sg = T.s[blockIdx.x];
  const uint32_t nt = (sg.len + kTile - 1) / kTile;
  uint64_t carry = 0;
  for (uint32_t b = 0; b < nt; b += FSIG_BLOCK) {
    const uint32_t i = b + threadIdx.x;
    const uint64_t v = i < nt ? tilecnt[sg.tile0 + i] : 0;
    uint64_t tot;
    const uint64_t ex = block_exclusive_scan(v, lds, tot);
    if (i < nt) tilecnt[sg.tile0 + i] = (uint32_t)(carry + ex);
    carry += tot;
  }
  if (threadIdx.x == 0) sizes[blockIdx.x] = carry;
}

template <int PRED>
__global__ void __launch_bounds__(FSIG_BLOCK) k_compact_scatter(CompactTable T, const uint32_t* __restrict__ mask,
                                                                const uint8_t* __restrict__ keep,
                                                                const uint32_t* __restrict__ tileoff) {
  __shared__ uint64_t lds[4];
  const uint32_t s = find_seg(T, blockIdx.x);
  const CompactSeg& sg = T.s[s];
  const uint32_t base = (blockIdx.x - sg.tile0) * kTile + threadIdx.x * 4;
  bool f[4];
  uint64_t c = 0;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const uint32_t i = base + j;
    f[j] = i < sg.len && compact_pred<PRED>(sg, i, mask, keep);
    c += f[j];
  }
  uint64_t tot;
  uint64_t at = block_exclusive_scan(c, lds, tot) + tileoff[blockIdx.x];
#pragma unroll
  for (int j = 0; j < 4; j++)
    if (f[j]) sg.dst[at++] = sg.src[base + j];
}

// Level-k rows (candidate indices in join order) -> data vertex ids in
// query-vertex order.
struct EmitMap {
  uint32_t k;
  uint32_t qv[MAXQ];              // query vertex of join position c
  const uint32_t* cand[MAXQ];     // candidate array of join position c (null: word is a vertex id)
};

__global__ void k_emit(EmitMap M, const uint32_t* __restrict__ rows, uint64_t n, uint32_t* __restrict__ out) {
  const uint64_t words = n * M.k;
  for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < words; w += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t r = w / M.k;
    const uint32_t c = (uint32_t)(w - r * M.k);
    const uint32_t x = rows[w];
    out[r * M.k + M.qv[c]] = M.cand[c] ? M.cand[c][x] : x;
  }
}

// Digest of level-k rows: one row per thread, words taken in query-vertex
// order (join position M.pos[i] holds query vertex i); per-wave sums are
// added to acc[0..1].
struct DigestMap {
  uint32_t k;
  uint32_t pos[MAXQ];             // join position of query vertex i
  const uint32_t* cand[MAXQ];     // candidate array of join position c (null: word is a vertex id)
};

__global__ void __launch_bounds__(FSIG_BLOCK) k_digest(DigestMap M, const uint32_t* __restrict__ rows, uint64_t n,
                                                       unsigned long long* __restrict__ acc) {
  unsigned long long s0 = 0, s1 = 0;
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t* row = rows + r * M.k;
    uint64_t x = row_hash_start(M.k);
    for (uint32_t i = 0; i < M.k; i++) {
      const uint32_t c = M.pos[i];
      const uint32_t w = row[c];
      x = row_hash_step(x, i, M.cand[c] ? M.cand[c][w] : w);
    }
    s0 += x;
    s1 += mix64(x ^ kDigestSalt);
  }
  s0 = wave_sum_u64(s0);
  s1 = wave_sum_u64(s1);
  if ((threadIdx.x & 63u) == 0) {
    atomicAdd(&acc[0], s0);
    atomicAdd(&acc[1], s1);
  }
}

uint32_t grid_for(const fsig_ctx* c, uint64_t items, uint32_t per_block) {
  uint64_t b = (items + per_block - 1) / per_block;
  b = std::min<uint64_t>(b, (uint64_t)c->num_cus * 16);
  return (uint32_t)std::max<uint64_t>(b, 1);
}

// Copies level-k tables to the caller's host buffer in query-vertex order.
struct HostSink : LeafSink {
  fsig_ctx* c;
  EmitMap map;
  uint32_t* host;
  uint64_t cap;      // rows the host buffer holds
  uint64_t written = 0;
  int take(const uint32_t* rows, uint64_t n) override {
    if (!host || written >= cap) return FSIG_OK;
    const uint64_t m = std::min<uint64_t>(n, cap - written);
    const size_t mk = c->arena.mark();
    uint32_t* buf = (uint32_t*)c->arena.alloc(4ull * map.k * m);
    if (!buf) return set_err(c, FSIG_ENOMEM, "emit: arena exhausted");
    k_emit<<<grid_for(c, m * map.k, FSIG_BLOCK), FSIG_BLOCK, 0, c->stream>>>(map, rows, m, buf);
    FSIG_HIP(c, hipGetLastError());
    FSIG_HIP(c, hipMemcpyAsync(host + written * map.k, buf, 4ull * map.k * m, hipMemcpyDeviceToHost, c->stream));
    FSIG_HIP(c, hipStreamSynchronize(c->stream));
    c->arena.release(mk);
    written += m;
    return FSIG_OK;
  }
};

// Folds level-k tables into the set digest (acc: two device words, zeroed
// by the caller before the match).
struct DigestSink : LeafSink {
  fsig_ctx* c;
  DigestMap map;
  unsigned long long* acc;
  int take(const uint32_t* rows, uint64_t n) override {
    if (n == 0) return FSIG_OK;
    k_digest<<<grid_for(c, n, FSIG_BLOCK), FSIG_BLOCK, 0, c->stream>>>(map, rows, n, acc);
    FSIG_HIP(c, hipGetLastError());
    return FSIG_OK;
  }
};

// The sink a match hands its level-k tables to: the digest sink when the
// caller asked for a digest, else the host copy when it gave a buffer.
struct LeafSinks {
  HostSink host;
  DigestSink digest;
  // pos_of_qv[i]: join position of query vertex i; cand_of_pos[c]: candidate
  // array of position c (null when rows hold vertex ids)
  LeafSink* setup(fsig_ctx* c, uint32_t k, const uint32_t* order, const uint32_t* const* cand_of_pos,
                  uint32_t* emb_out, uint64_t emb_cap, unsigned long long* digest_acc) {
    host.c = c;
    std::memset(&host.map, 0, sizeof(host.map));
    host.map.k = k;
    digest.c = c;
    std::memset(&digest.map, 0, sizeof(digest.map));
    digest.map.k = k;
    digest.acc = digest_acc;
    for (uint32_t p = 0; p < k; p++) {
      host.map.qv[p] = order[p];
      host.map.cand[p] = cand_of_pos ? cand_of_pos[p] : nullptr;
      digest.map.pos[order[p]] = p;
      digest.map.cand[p] = cand_of_pos ? cand_of_pos[p] : nullptr;
    }
    host.host = emb_out;
    host.cap = emb_out ? emb_cap : 0;
    if (digest_acc) return &digest;
    return emb_out ? &host : nullptr;
  }
};

// Host copy of the candidate sets of a match (fsig_set_keep_candidates):
// stage 0 after the filter, stage 1 after refinement, per class.
inline int keep_candidate_sets(fsig_ctx* c, int stage, const QueryPlan& P, uint32_t* const* cand,
                               const uint32_t* size) {
  if (!c->keep_cand) return FSIG_OK;
  std::vector<std::vector<uint32_t>> per_class(P.nclass);
  for (uint32_t a = 0; a < P.nclass; a++) {
    per_class[a].resize(size[a]);
    if (size[a])
      FSIG_HIP(c, hipMemcpyAsync(per_class[a].data(), cand[a], 4ull * size[a], hipMemcpyDeviceToHost, c->stream));
  }
  FSIG_HIP(c, hipStreamSynchronize(c->stream));
  auto& dst = c->cand_sets[stage];
  dst.assign(P.k, {});
  for (uint32_t u = 0; u < P.k; u++) dst[u] = per_class[P.cls[u]];
  return FSIG_OK;
}

}  // namespace
