// Candidate-edge rows of high-degree candidates, cut into fixed segments.
//
// A candidate-edge row (pair a -> b, candidate v of a) is N(v) filtered by
// b's candidate mask.  Rows of at most kHubDeg entries are swept by one lane
// group each; the adjacency of a hub (a power-law graph's high-degree
// candidates, tens of thousands of entries) is cut into segments of
// kSegEdges entries dealt one per wavefront over the whole device, so that a
// few hubs no longer leave one workgroup sweeping while the rest idle.
//
// Pair tables (PT) carry: acand, asize (rows of the pair), bcand, bsize,
// bpos (vertex -> candidate index of b, or null: binary search), bits (b's
// candidate bitmap, n bits, or null), mask (one class-bit word per data
// vertex, read when bits is null), bbit (b's bit), off_out (the pair's list
// offsets, asize + 1 words) and flat (the pair's first row in the flat row
// space shared by every pair).  A bitmap is n/8 bytes against the mask's 4n:
// on large graphs it stays in L2 while the entries of one pair are tested.
#pragma once

#include <algorithm>

#include "device.h"
#include "internal.h"

namespace {

constexpr uint32_t kHubDeg = 512;
constexpr uint32_t kSegEdges = 512;

// Whether data vertex x is a candidate of the pair's target class.
template <class PT>
__device__ __forceinline__ bool cedge_keep(const PT& pd, uint32_t x) {
  return pd.bits ? (pd.bits[x >> 5] >> (x & 31u)) & 1u : (pd.mask[x] >> pd.bbit) & 1u;
}

// Pair of flat row f: the last pair whose first row is <= f.
template <class PT>
__device__ __forceinline__ uint32_t cedge_pair_index(const PT* __restrict__ P, uint32_t np, uint64_t f) {
  uint32_t lo = 0, hi = np;
  while (hi - lo > 1) {
    const uint32_t m = (lo + hi) >> 1;
    if (P[m].flat <= f) lo = m; else hi = m;
  }
  return lo;
}

template <class PT>
__device__ __forceinline__ const PT& cedge_pair(const PT* __restrict__ P, uint32_t np, uint64_t f) {
  return P[cedge_pair_index(P, np, f)];
}

__global__ void k_seg_zero(unsigned long long* p) {
  if (threadIdx.x == 0) *p = 0;
}

// Segments per flat row (0 for rows of at most kHubDeg entries); a hub
// row's count starts at zero for k_seg_cedge's additions.  *entries +=
// the adjacency entries of every row (the build's algorithmic volume).
template <class PT>
__global__ void __launch_bounds__(FSIG_BLOCK) k_seg_rows(const PT* __restrict__ P, uint32_t np, uint64_t rows,
                                                         const uint32_t* __restrict__ off, uint32_t* __restrict__ nseg,
                                                         uint32_t* __restrict__ cnt,
                                                         unsigned long long* __restrict__ entries) {
  unsigned long long mine = 0;
  for (uint64_t f = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; f < rows;
       f += (uint64_t)gridDim.x * blockDim.x) {
    const PT& pd = cedge_pair(P, np, f);
    const uint32_t v = pd.acand[f - pd.flat];
    const uint32_t deg = off[v + 1] - off[v];
    const bool hub = deg > kHubDeg;
    nseg[f] = hub ? (deg + kSegEdges - 1) / kSegEdges : 0u;
    if (hub) cnt[f] = 0;
    mine += deg;
  }
  mine = wave_sum_u64(mine);
  if ((threadIdx.x & 63u) == 0 && mine) atomicAdd(entries, mine);
}

// Row, pair and adjacency range of every segment (rows < 2^32), so that a
// segment's wavefront reaches its adjacency after one load instead of a
// chain of row, pair, candidate and offset lookups.
template <class PT>
__global__ void __launch_bounds__(FSIG_BLOCK) k_seg_expand(const PT* __restrict__ P, uint32_t np, uint64_t rows,
                                                           const uint32_t* __restrict__ off,
                                                           const uint32_t* __restrict__ nseg,
                                                           const uint64_t* __restrict__ segbase,
                                                           uint32_t* __restrict__ seg_row,
                                                           uint32_t* __restrict__ seg_pair,
                                                           uint2* __restrict__ seg_range) {
  for (uint64_t f = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; f < rows;
       f += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t m = nseg[f];
    if (!m) continue;
    const uint32_t pi = cedge_pair_index(P, np, f);
    const uint32_t v = P[pi].acand[f - P[pi].flat];
    const uint32_t lo = off[v], hi = off[v + 1];
    const uint64_t b = segbase[f];
    for (uint32_t j = 0; j < m; j++) {
      seg_row[b + j] = (uint32_t)f;
      seg_pair[b + j] = pi;
      const uint32_t s0 = lo + j * kSegEdges;
      seg_range[b + j] = make_uint2(s0, min(hi, s0 + kSegEdges));
    }
  }
}

// Work items [0, n) of a grid whose blocks each run `per_block` workers
// (lane groups or wavefronts): each XCD walks its own contiguous eighth, so
// the items one pair owns, and that pair's target bitmap, stay in one L2 at
// a time; a grid that is no multiple of 8 strides over all items.
struct XcdRange {
  uint64_t lo, hi, first, stride;
};
__device__ __forceinline__ XcdRange xcd_range(uint64_t n, uint32_t per_block, uint32_t worker) {
  XcdRange r{0, n, (uint64_t)blockIdx.x * per_block + worker, (uint64_t)gridDim.x * per_block};
  if ((gridDim.x & 7u) == 0) {
    const uint32_t xcd = blockIdx.x & 7u;
    r.lo = n * xcd / 8;
    r.hi = n * (xcd + 1) / 8;
    r.first = r.lo + (uint64_t)(blockIdx.x >> 3) * per_block + worker;
    r.stride = (uint64_t)(gridDim.x >> 3) * per_block;
  }
  return r;
}

// Rows of at most kHubDeg entries, G lanes each; U strides of G entries are
// loaded before any is tested, so a group keeps U gathers in flight.
constexpr int kRowUnroll = 4;

template <class PT, int G>
__global__ void __launch_bounds__(FSIG_BLOCK) k_row_count(const PT* __restrict__ P, uint32_t np, uint64_t rows,
                                                          const uint32_t* __restrict__ off,
                                                          const uint32_t* __restrict__ adj,
                                                          uint32_t* __restrict__ cnt) {
  constexpr int GPB = FSIG_BLOCK / G, U = kRowUnroll;
  const uint32_t lane = threadIdx.x & 63u, gl = lane & (G - 1), gw = threadIdx.x / G;
  const uint64_t gmask = (G == 64) ? ~0ull : (((1ull << G) - 1) << (lane - gl));
  const XcdRange xr = xcd_range(rows, GPB, gw);
  for (uint64_t f = xr.first; f < xr.hi; f += xr.stride) {
    const PT& pd = cedge_pair(P, np, f);
    const uint32_t v = pd.acand[f - pd.flat];
    const uint32_t s = off[v], e = off[v + 1];
    if (e - s > kHubDeg) continue;  // k_seg_cedge
    uint32_t got = 0;
    for (uint32_t j = s; j < e; j += U * G) {
      uint32_t x[U];
      bool ok[U];
#pragma unroll
      for (int t = 0; t < U; t++) {
        const uint32_t i = j + t * G + gl;
        ok[t] = i < e;
        x[t] = ok[t] ? adj[i] : 0u;
      }
#pragma unroll
      for (int t = 0; t < U; t++) ok[t] = ok[t] && cedge_keep(pd, x[t]);
#pragma unroll
      for (int t = 0; t < U; t++) got += (uint32_t)__popcll(__ballot(ok[t]) & gmask);
    }
    if (gl == 0) cnt[f] = got;
  }
}

template <class PT, int G>
__global__ void __launch_bounds__(FSIG_BLOCK) k_row_fill(const PT* __restrict__ P, uint32_t np, uint64_t rows,
                                                         const uint32_t* __restrict__ off,
                                                         const uint32_t* __restrict__ adj,
                                                         const uint64_t* __restrict__ gscan,
                                                         uint32_t* __restrict__ lst_all) {
  constexpr int GPB = FSIG_BLOCK / G, U = kRowUnroll;
  const uint32_t lane = threadIdx.x & 63u, gl = lane & (G - 1), gw = threadIdx.x / G;
  const uint64_t gmask = (G == 64) ? ~0ull : (((1ull << G) - 1) << (lane - gl));
  const uint64_t below = (1ull << lane) - 1;
  const XcdRange xr = xcd_range(rows, GPB, gw);
  for (uint64_t f = xr.first; f < xr.hi; f += xr.stride) {
    const PT& pd = cedge_pair(P, np, f);
    const uint32_t i0 = (uint32_t)(f - pd.flat);
    const uint32_t v = pd.acand[i0];
    const uint32_t s = off[v], e = off[v + 1];
    if (e - s > kHubDeg) continue;  // k_seg_cedge
    const uint64_t pbase = gscan[pd.flat];
    uint64_t at = gscan[f];
    if (gl == 0) {
      pd.off_out[i0] = (uint32_t)(at - pbase);
      if (i0 + 1 == pd.asize) pd.off_out[pd.asize] = (uint32_t)(gscan[f + 1] - pbase);
    }
    for (uint32_t j = s; j < e; j += U * G) {
      uint32_t x[U];
      bool ok[U];
#pragma unroll
      for (int t = 0; t < U; t++) {
        const uint32_t i = j + t * G + gl;
        ok[t] = i < e;
        x[t] = ok[t] ? adj[i] : 0u;
      }
#pragma unroll
      for (int t = 0; t < U; t++) ok[t] = ok[t] && cedge_keep(pd, x[t]);
#pragma unroll
      for (int t = 0; t < U; t++) {
        const uint64_t bal = __ballot(ok[t]) & gmask;
        if (ok[t]) lst_all[at + __popcll(bal & below)] = pd.bpos ? pd.bpos[x[t]] : sorted_index(pd.bcand, pd.bsize, x[t]);
        at += __popcll(bal);
      }
    }
  }
}

// One segment of kSegEdges = 8 x 64 entries per wavefront, for the
// segments [s_lo, s_hi); segment s belongs to row seg_row[s] of pair
// seg_pair[s], and segbase[f] (exclusive scan of the segments per row) is
// row f's first segment.  FILL == false: the segment's entries are tested (all eight
// loads in flight at once), the kept ones recorded as eight 64-bit words
// segbits[8s..8s+7], their number as segcnt[s] and added to the row's count
// cnt[f].  FILL == true: the kept entries (from segbits, no second test) are
// written after those of the row's earlier segments (their segcnt sum), so
// every list stays ascending; a row's first segment writes its offsets.
static_assert(kSegEdges == 8 * 64, "a segment is eight wavefront strides");

template <class PT, bool FILL>
__global__ void __launch_bounds__(FSIG_BLOCK) k_seg_cedge(const PT* __restrict__ P, uint64_t rows,
                                                          const uint64_t* __restrict__ segbase,
                                                          const uint32_t* __restrict__ seg_row,
                                                          const uint32_t* __restrict__ seg_pair,
                                                          const uint2* __restrict__ seg_range,
                                                          const uint32_t* __restrict__ adj,
                                                          uint32_t* __restrict__ segcnt,
                                                          unsigned long long* __restrict__ segbits,
                                                          uint32_t* __restrict__ cnt,
                                                          const uint64_t* __restrict__ gscan,
                                                          uint32_t* __restrict__ lst_all, uint64_t s_lo,
                                                          uint64_t s_hi) {
  constexpr uint32_t WPB = FSIG_BLOCK / 64;
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t below = (1ull << lane) - 1;
  const XcdRange xr = xcd_range(s_hi - s_lo, WPB, threadIdx.x >> 6);
  for (uint64_t s = s_lo + xr.first; s < s_lo + xr.hi; s += xr.stride) {
    const uint64_t f = seg_row[s];
    const PT& pd = P[seg_pair[s]];
    const uint2 range = seg_range[s];
    const uint32_t s0 = range.x, s1 = range.y;
    if (!FILL) {
      uint32_t x[8];
      bool ok[8];
#pragma unroll
      for (int t = 0; t < 8; t++) {
        const uint32_t i = s0 + t * 64 + lane;
        ok[t] = i < s1;
        x[t] = ok[t] ? adj[i] : 0u;
      }
#pragma unroll
      for (int t = 0; t < 8; t++) ok[t] = ok[t] && cedge_keep(pd, x[t]);
      uint32_t got = 0;
      unsigned long long mine = 0;
#pragma unroll
      for (int t = 0; t < 8; t++) {
        const unsigned long long bal = __ballot(ok[t]);
        got += (uint32_t)__popcll(bal);
        if (lane == (uint32_t)t) mine = bal;
      }
      if (lane < 8) segbits[s * 8 + lane] = mine;
      if (lane == 0) {
        segcnt[s] = got;
        if (got) atomicAdd(&cnt[f], got);
      }
    } else {
      // entries of the row's earlier segments
      uint64_t pre = 0;
      for (uint64_t t = segbase[f] + lane; t < s; t += 64) pre += segcnt[t];
      pre = wave_sum_u64(pre);
      const uint32_t i0 = (uint32_t)(f - pd.flat);
      const uint64_t pbase = gscan[pd.flat];
      uint64_t at = gscan[f] + pre;
      if (lane == 0 && s == segbase[f]) {
        pd.off_out[i0] = (uint32_t)(gscan[f] - pbase);
        if (i0 + 1 == pd.asize) pd.off_out[pd.asize] = (uint32_t)(gscan[f + 1] - pbase);
      }
      const unsigned long long w = segbits[s * 8 + (lane & 7u)];
#pragma unroll
      for (int t = 0; t < 8; t++) {
        const unsigned long long bal = __shfl(w, t, 64);
        if ((bal >> lane) & 1ull) {
          const uint32_t x = adj[s0 + t * 64 + lane];
          lst_all[at + __popcll(bal & below)] = pd.bpos ? pd.bpos[x] : sorted_index(pd.bcand, pd.bsize, x);
        }
        at += __popcll(bal);
      }
    }
  }
}

// Host side of the segmented hub rows: seg_prepare (segments per row, their
// scan and the hub rows' counts zeroed) before seg_count, seg_fill once the
// row counts are scanned into gscan.
struct SegRows {
  uint64_t* segbase = nullptr;  // rows + 1 entries, then the entry count
  uint32_t* segcnt = nullptr;
  unsigned long long* segbits = nullptr;  // 8 words per segment
  uint32_t* seg_row = nullptr;
  uint32_t* seg_pair = nullptr;
  uint2* seg_range = nullptr;  // adjacency entries [x, y) of the segment
  uint64_t nsegs = 0;
  uint64_t entries = 0;  // adjacency entries of all rows
};

template <class PT>
int seg_prepare(fsig_ctx* c, const PT* dP, uint32_t np, uint64_t rows, const uint32_t* off, uint32_t* cnt,
                SegRows& sr) {
  if (rows == 0) return FSIG_OK;
  if (rows >= 0xffffffffull) return set_err(c, FSIG_EUNSUPPORTED, "cedge: too many candidate rows");
  uint32_t* nseg = (uint32_t*)c->arena.alloc(4ull * rows);
  sr.segbase = (uint64_t*)c->arena.alloc(8ull * (rows + 2));
  if (!nseg || !sr.segbase) return set_err(c, FSIG_ENOMEM, "cedge: arena exhausted");
  unsigned long long* entries = (unsigned long long*)(sr.segbase + rows + 1);
  k_seg_zero<<<1, 64, 0, c->stream>>>(entries);
  const uint64_t blocks = std::min<uint64_t>((rows + FSIG_BLOCK - 1) / FSIG_BLOCK, (uint64_t)c->num_cus * 16);
  k_seg_rows<PT><<<(uint32_t)blocks, FSIG_BLOCK, 0, c->stream>>>(dP, np, rows, off, nseg, cnt, entries);
  FSIG_HIP(c, hipGetLastError());
  int rc;
  if ((rc = scan_u32_to_u64(c, nseg, sr.segbase, rows))) return rc;
  uint64_t tail[2];
  if ((rc = read_u64(c, sr.segbase + rows, tail, 2))) return rc;
  sr.nsegs = tail[0];
  sr.entries = tail[1];
  if (sr.nsegs) {
    sr.segcnt = (uint32_t*)c->arena.alloc(4ull * sr.nsegs);
    sr.segbits = (unsigned long long*)c->arena.alloc(64ull * sr.nsegs);
    sr.seg_row = (uint32_t*)c->arena.alloc(4ull * sr.nsegs);
    sr.seg_pair = (uint32_t*)c->arena.alloc(4ull * sr.nsegs);
    sr.seg_range = (uint2*)c->arena.alloc(8ull * sr.nsegs);
    if (!sr.segcnt || !sr.segbits || !sr.seg_row || !sr.seg_pair || !sr.seg_range)
      return set_err(c, FSIG_ENOMEM, "cedge: arena exhausted");
    k_seg_expand<PT><<<(uint32_t)blocks, FSIG_BLOCK, 0, c->stream>>>(dP, np, rows, off, nseg, sr.segbase, sr.seg_row,
                                                                     sr.seg_pair, sr.seg_range);
    FSIG_HIP(c, hipGetLastError());
  }
  return FSIG_OK;
}

inline uint32_t seg_grid(const fsig_ctx* c, uint64_t nsegs) {
  uint64_t b = (nsegs + FSIG_BLOCK / 64 - 1) / (FSIG_BLOCK / 64);
  b = std::min<uint64_t>(b, (uint64_t)c->num_cus * 16);
  if (b >= 64) b &= ~7ull;
  return (uint32_t)std::max<uint64_t>(b, 1);
}

// s_hi == ~0: every segment.
template <class PT>
int seg_count(fsig_ctx* c, const PT* dP, uint32_t np, uint64_t rows, const uint32_t* off, const uint32_t* adj,
              uint32_t* cnt, const SegRows& sr, uint64_t s_lo = 0, uint64_t s_hi = ~0ull) {
  s_hi = std::min(s_hi, sr.nsegs);
  if (s_lo >= s_hi) return FSIG_OK;
  k_seg_cedge<PT, false><<<seg_grid(c, s_hi - s_lo), FSIG_BLOCK, 0, c->stream>>>(
      dP, rows, sr.segbase, sr.seg_row, sr.seg_pair, sr.seg_range, adj, sr.segcnt, sr.segbits, cnt, nullptr, nullptr,
      s_lo, s_hi);
  FSIG_HIP(c, hipGetLastError());
  return FSIG_OK;
}

template <class PT>
int seg_fill(fsig_ctx* c, const PT* dP, uint32_t np, uint64_t rows, const uint32_t* off, const uint32_t* adj,
             const uint64_t* gscan, uint32_t* lst_all, const SegRows& sr, uint64_t s_lo = 0, uint64_t s_hi = ~0ull) {
  s_hi = std::min(s_hi, sr.nsegs);
  if (s_lo >= s_hi) return FSIG_OK;
  k_seg_cedge<PT, true><<<seg_grid(c, s_hi - s_lo), FSIG_BLOCK, 0, c->stream>>>(
      dP, rows, sr.segbase, sr.seg_row, sr.seg_pair, sr.seg_range, adj, sr.segcnt, sr.segbits, nullptr, gscan,
      lst_all, s_lo, s_hi);
  FSIG_HIP(c, hipGetLastError());
  return FSIG_OK;
}

// Per-row passes of the rows of at most kHubDeg entries, G lanes per row.
template <class PT>
int row_count(fsig_ctx* c, uint32_t G, const PT* dP, uint32_t np, uint64_t rows, const uint32_t* off,
              const uint32_t* adj, uint32_t* cnt) {
  if (!rows) return FSIG_OK;
  auto grid = [&](uint32_t g) {
    const uint64_t b = std::min<uint64_t>((rows + FSIG_BLOCK / g - 1) / (FSIG_BLOCK / g), (uint64_t)c->num_cus * 16);
    return (uint32_t)std::max<uint64_t>(b, 1);
  };
  switch (G) {
    case 8: k_row_count<PT, 8><<<grid(8), FSIG_BLOCK, 0, c->stream>>>(dP, np, rows, off, adj, cnt); break;
    case 16: k_row_count<PT, 16><<<grid(16), FSIG_BLOCK, 0, c->stream>>>(dP, np, rows, off, adj, cnt); break;
    default: k_row_count<PT, 64><<<grid(64), FSIG_BLOCK, 0, c->stream>>>(dP, np, rows, off, adj, cnt); break;
  }
  FSIG_HIP(c, hipGetLastError());
  return FSIG_OK;
}

template <class PT>
int row_fill(fsig_ctx* c, uint32_t G, const PT* dP, uint32_t np, uint64_t rows, const uint32_t* off,
             const uint32_t* adj, const uint64_t* gscan, uint32_t* lst_all) {
  if (!rows) return FSIG_OK;
  auto grid = [&](uint32_t g) {
    const uint64_t b = std::min<uint64_t>((rows + FSIG_BLOCK / g - 1) / (FSIG_BLOCK / g), (uint64_t)c->num_cus * 16);
    return (uint32_t)std::max<uint64_t>(b, 1);
  };
  switch (G) {
    case 8: k_row_fill<PT, 8><<<grid(8), FSIG_BLOCK, 0, c->stream>>>(dP, np, rows, off, adj, gscan, lst_all); break;
    case 16: k_row_fill<PT, 16><<<grid(16), FSIG_BLOCK, 0, c->stream>>>(dP, np, rows, off, adj, gscan, lst_all); break;
    default: k_row_fill<PT, 64><<<grid(64), FSIG_BLOCK, 0, c->stream>>>(dP, np, rows, off, adj, gscan, lst_all); break;
  }
  FSIG_HIP(c, hipGetLastError());
  return FSIG_OK;
}

// Candidate bitmaps (n bits, zeroed first) of the listed classes.
__global__ void k_set_bits(const uint32_t* __restrict__ cand, uint32_t size, uint32_t* __restrict__ bits) {
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < size; j += gridDim.x * blockDim.x) {
    const uint32_t v = cand[j];
    atomicOr(&bits[v >> 5], 1u << (v & 31u));
  }
}

}  // namespace
