// Batched engine: many queries against one resident graph.
//
// Setup (filter, refinement, candidate edges) runs for all queries at once:
// every kernel walks a device table of (query, class) or (query, class
// pair) descriptors, so the whole batch costs a fixed number of launches and
// host round trips instead of a dozen per query.
//
// Two join engines:
//  * lockstep: level p of every query in one count launch, one scan and one
//    write launch over all queries' rows (the row's query is looked up by
//    binary search over the row bases); the last two levels of a query are
//    counted by the fused tail in its count launch.  The host syncs once per
//    level of the batch instead of once per level of every query.
//  * persistent: one kernel fed by a device work queue of (query, root
//    range) tasks, launched through a captured hipGraph.  Each group of G
//    lanes pops a task and walks its roots depth-first, keeping one
//    candidate list per level on a private stack in HBM, and donates
//    unexplored ranges to idle groups; the last level is counted, never
//    stored.  It needs no level tables, so it also runs batches whose
//    levels do not fit the arena.
//
// The autotuner times the lockstep shapes on the whole batch and the
// persistent walk (lanes per group, LDS staging) on a sample, and runs the
// batch with the fastest; choices are cached per batch shape.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>

#include "common.h"
#include "device.h"
#include "internal.h"
#include "join.h"
#include "cedge.h"

namespace {

// ---------------------------------------------------------------------------
// Batched label/degree filter: one segment per (query, query label).
struct BSeg {
  const uint32_t* src;   // inverted-index segment of the label
  uint32_t* mask;        // the query's mask[n]
  uint32_t len;
  uint32_t cmask;        // classes of this label
  uint32_t cls_base;     // global index of the query's class 0
  uint32_t pad;
  uint64_t flat;         // first element in the flat numbering
};

template <class T>
__device__ __forceinline__ uint32_t find_flat(const T* __restrict__ d, uint32_t n, uint64_t f) {
  uint32_t lo = 0, hi = n;  // last entry with flat <= f
  while (hi - lo > 1) {
    const uint32_t m = (lo + hi) >> 1;
    if (d[m].flat <= f) lo = m; else hi = m;
  }
  return lo;
}

// With nlf non-null, global class gc also needs, for each entry e of
// req[gc * MAXQ .. + nreq[gc]), nlf[v * L + (e >> 8)] >= (e & 255).
__global__ void __launch_bounds__(FSIG_BLOCK) k_bldf(const BSeg* __restrict__ S, uint32_t ns, uint64_t total,
                                                     const uint32_t* __restrict__ cdeg,
                                                     const uint32_t* __restrict__ off, const uint8_t* __restrict__ nlf,
                                                     uint32_t L, const uint32_t* __restrict__ req,
                                                     const uint32_t* __restrict__ nreq) {
  for (uint64_t f = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; f < total; f += (uint64_t)gridDim.x * blockDim.x) {
    const BSeg& s = S[find_flat(S, ns, f)];
    const uint32_t v = s.src[f - s.flat];
    const uint32_t d = off[v + 1] - off[v];
    uint32_t m = 0, cm = s.cmask;
    while (cm) {
      const uint32_t c = __builtin_ctz(cm);
      cm &= cm - 1;
      const uint32_t gc = s.cls_base + c;
      if (d < cdeg[gc]) continue;
      bool ok = true;
      if (nlf) {
        const uint8_t* row = nlf + (uint64_t)v * L;
        const uint32_t* rq = req + (uint64_t)gc * MAXQ;
        for (uint32_t i = 0; i < nreq[gc] && ok; i++) ok = row[rq[i] >> 8] >= (rq[i] & 255u);
      }
      if (ok) m |= 1u << c;
    }
    s.mask[v] = m;
  }
}

// ---------------------------------------------------------------------------
// Batched order-preserving compaction, one segment per global class.
struct BComp {
  const uint32_t* src;
  uint32_t* dst;
  const uint32_t* mask;
  uint32_t len;
  uint32_t bit;
  uint64_t keep_base;
  uint64_t flat;  // first tile of the segment (find_flat over tiles)
};

template <int PRED>
__device__ __forceinline__ bool bpred(const BComp& s, uint32_t i, const uint8_t* keep) {
  if (PRED == 0) return (s.mask[s.src[i]] >> s.bit) & 1u;
  return keep[s.keep_base + i] != 0;
}

template <int PRED>
__global__ void __launch_bounds__(FSIG_BLOCK) k_bcompact_count(const BComp* __restrict__ S, uint32_t ns,
                                                               const uint8_t* __restrict__ keep,
                                                               uint32_t* __restrict__ tilecnt) {
  __shared__ uint64_t lds[4];
  const BComp& s = S[find_flat(S, ns, blockIdx.x)];
  const uint32_t base = (uint32_t)(blockIdx.x - s.flat) * kTile;
  uint64_t c = 0;
  for (uint32_t j = 0; j < 4; j++) {
    const uint32_t i = base + j * FSIG_BLOCK + threadIdx.x;
    if (i < s.len && bpred<PRED>(s, i, keep)) c++;
  }
  c = block_sum(c, lds);
  if (threadIdx.x == 0) tilecnt[blockIdx.x] = (uint32_t)c;
}

__global__ void __launch_bounds__(FSIG_BLOCK) k_bcompact_offsets(const BComp* __restrict__ S,
                                                                 uint32_t* __restrict__ tilecnt,
                                                                 uint64_t* __restrict__ sizes) {
  __shared__ uint64_t lds[4];
  const BComp& s = S[blockIdx.x];
  const uint32_t nt = (s.len + kTile - 1) / kTile;
  uint64_t carry = 0;
  for (uint32_t b = 0; b < nt; b += FSIG_BLOCK) {
    const uint32_t i = b + threadIdx.x;
    const uint64_t v = i < nt ? tilecnt[s.flat + i] : 0;
    uint64_t tot;
    const uint64_t ex = block_exclusive_scan(v, lds, tot);
    if (i < nt) tilecnt[s.flat + i] = (uint32_t)(carry + ex);
    carry += tot;
  }
  if (threadIdx.x == 0) sizes[blockIdx.x] = carry;
}

template <int PRED>
__global__ void __launch_bounds__(FSIG_BLOCK) k_bcompact_scatter(const BComp* __restrict__ S, uint32_t ns,
                                                                 const uint8_t* __restrict__ keep,
                                                                 const uint32_t* __restrict__ tileoff) {
  __shared__ uint64_t lds[4];
  const BComp& s = S[find_flat(S, ns, blockIdx.x)];
  const uint32_t base = (uint32_t)(blockIdx.x - s.flat) * kTile + threadIdx.x * 4;
  bool f[4];
  uint64_t c = 0;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    f[j] = base + j < s.len && bpred<PRED>(s, base + j, keep);
    c += f[j];
  }
  uint64_t tot;
  uint64_t at = block_exclusive_scan(c, lds, tot) + tileoff[blockIdx.x];
#pragma unroll
  for (int j = 0; j < 4; j++)
    if (f[j]) s.dst[at++] = s.src[base + j];
}

// ---------------------------------------------------------------------------
// Batched refinement pass (same rule as k_refine, per-query mask).
struct BClass {
  const uint32_t* cand;
  uint32_t* mask;
  uint32_t size;
  uint32_t bit;
  uint32_t tmask;
  uint32_t thr_row;  // row of the u8 threshold table
  uint64_t flat;
};

template <int G>
__global__ void __launch_bounds__(FSIG_BLOCK) k_brefine(const BClass* __restrict__ C, uint32_t nc, uint64_t total,
                                                        const uint8_t* __restrict__ thr,
                                                        const uint32_t* __restrict__ off,
                                                        const uint32_t* __restrict__ adj,
                                                        uint8_t* __restrict__ keep) {
  constexpr int GPB = FSIG_BLOCK / G;
  __shared__ uint8_t rem[GPB][MAXQ];
  const uint32_t lane = threadIdx.x & 63u, gl = lane & (G - 1), gw = threadIdx.x / G;
  const uint64_t gmask = (G == 64) ? ~0ull : (((1ull << G) - 1) << (lane - gl));
  const uint64_t ngroups = (uint64_t)gridDim.x * GPB;
  for (uint64_t f = (uint64_t)xcd_block() * GPB + gw; f < total; f += ngroups) {
    const BClass& cl = C[find_flat(C, nc, f)];
    const uint32_t v = cl.cand[f - cl.flat];
    const uint32_t s = off[v], e = off[v + 1];
    const uint32_t* mask = cl.mask;
    const uint8_t* need = thr + (uint64_t)cl.thr_row * MAXQ;
    uint32_t open = cl.tmask;
    for (uint32_t tm = open; tm; tm &= tm - 1) rem[gw][__builtin_ctz(tm)] = need[__builtin_ctz(tm)];
    for (uint32_t j = s; j < e && open; j += G) {
      const uint32_t i = j + gl;
      const uint32_t m = i < e ? (mask[adj[i]] & open) : 0u;
      for (uint32_t tm = open; tm; tm &= tm - 1) {
        const uint32_t b = __builtin_ctz(tm);
        const uint32_t got = __popcll(__ballot((m >> b) & 1u) & gmask);
        const uint32_t r = rem[gw][b];
        const uint32_t nr = r > got ? r - got : 0;
        rem[gw][b] = (uint8_t)nr;
        if (nr == 0) open &= ~(1u << b);
      }
    }
    if (gl == 0) keep[f] = open == 0 ? 1 : 0;
  }
}

__global__ void __launch_bounds__(FSIG_BLOCK) k_brefine_apply(const BClass* __restrict__ C, uint32_t nc,
                                                              uint64_t total, const uint8_t* __restrict__ keep) {
  for (uint64_t f = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; f < total; f += (uint64_t)gridDim.x * blockDim.x) {
    if (keep[f]) continue;
    const BClass& cl = C[find_flat(C, nc, f)];
    atomicAnd(&cl.mask[cl.cand[f - cl.flat]], ~(1u << cl.bit));
  }
}

// ---------------------------------------------------------------------------
// Batched candidate-edge CSR, with the longest list of every pair.
struct BPair {
  const uint32_t* acand;
  const uint32_t* bcand;
  const uint32_t* mask;
  const uint32_t* bits;  // candidate bitmap of b (null: the mask)
  uint32_t* off_out;
  const uint32_t* bpos;  // vertex -> candidate index of b (null: binary search)
  uint32_t asize, bsize, bbit, pad;
  uint64_t flat;
};

__global__ void k_bscatter_pos(const uint32_t* __restrict__ cand, uint32_t n, uint32_t* __restrict__ pos) {
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) pos[cand[j]] = j;
}

// out[j] = src[idx[j]] for j < m.
__global__ void k_gather_at(const uint64_t* __restrict__ src, const uint64_t* __restrict__ idx, uint32_t m,
                            uint64_t* __restrict__ out) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < m) out[j] = src[idx[j]];
}

__global__ void k_bpair_base(const uint64_t* __restrict__ gscan, const BPair* __restrict__ P, uint32_t np,
                             uint64_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < np) out[i] = gscan[P[i].flat];
}

// One (query, class a) with every target class b of a pair a -> b: the
// candidate-edge lists of all of a's pairs come from one sweep over each
// candidate's adjacency.
struct BSrc {
  const uint32_t* acand;
  const uint32_t* mask;        // the query's class mask (null: bitmaps)
  uint32_t asize;
  uint32_t tmask;
  uint64_t flat;
  uint32_t pidx[MAXQ];         // pair index of target b
  const uint32_t* bits[MAXQ];  // candidate bitmap (n bits) of target class b
};

// Target classes among tm that vertex w is a candidate of: one bit test per
// target in n-bit class bitmaps, which stay in L2 while the 4-byte-per-vertex
// masks of a whole batch do not, or (no room for bitmaps) one gather of the
// query's class-mask word.
__device__ __forceinline__ uint32_t target_bits(const BSrc& sc, uint32_t w, uint32_t tm) {
  if (sc.mask) return sc.mask[w] & tm;
  uint32_t m = 0;
  for (uint32_t t = tm; t; t &= t - 1) {
    const uint32_t b = __builtin_ctz(t);
    m |= ((sc.bits[b][w >> 5] >> (w & 31u)) & 1u) << b;
  }
  return m;
}

// Candidate bitmaps of the target classes: class j's candidates set bits
// in its n-bit map (the maps are zeroed before).
struct BBits {
  const uint32_t* cand;
  uint32_t* bits;
  uint32_t* pos;  // vertex -> candidate index map (n words) or null
  uint32_t size;
  uint32_t pad;
  uint64_t flat;
};

// Bits of every class's candidates; with_pos: also pos[C[j]] = j.
__global__ void k_bset_bits(const BBits* __restrict__ T, uint32_t nt, uint64_t total, bool with_pos) {
  for (uint64_t f = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; f < total;
       f += (uint64_t)gridDim.x * blockDim.x) {
    const BBits& t = T[find_flat(T, nt, f)];
    const uint32_t v = t.cand[f - t.flat];
    atomicOr(&t.bits[v >> 5], 1u << (v & 31u));
    if (with_pos && t.pos) t.pos[v] = (uint32_t)(f - t.flat);
  }
}

// Candidates of more than kHubDeg neighbours are cut into segments
// (k_seg_cedge, per pair) instead of one lane group.  A group loads
// kRowUnroll strides of a row before it tests any, so that many entry
// gathers, then that many rounds of bitmap tests, are in flight at once.

template <int G>
__global__ void __launch_bounds__(FSIG_BLOCK) k_bcedge_count_multi(const BSrc* __restrict__ S, uint32_t ns,
                                                                   uint64_t total, const BPair* __restrict__ P,
                                                                   const uint32_t* __restrict__ off,
                                                                   const uint32_t* __restrict__ adj,
                                                                   uint32_t* __restrict__ cnt, uint64_t f_lo) {
  constexpr int GPB = FSIG_BLOCK / G;
  __shared__ uint32_t acc[GPB][MAXQ];
  const uint32_t lane = threadIdx.x & 63u, gl = lane & (G - 1), gw = threadIdx.x / G;
  const uint64_t gmask = (G == 64) ? ~0ull : (((1ull << G) - 1) << (lane - gl));
  const XcdRange xr = xcd_range(total - f_lo, GPB, gw);  // source rows [f_lo, total)
  for (uint64_t f = f_lo + xr.first; f < f_lo + xr.hi; f += xr.stride) {
    const BSrc& sc = S[find_flat(S, ns, f)];
    const uint32_t i0 = (uint32_t)(f - sc.flat);
    const uint32_t v = sc.acand[i0];
    const uint32_t tm = sc.tmask;
    const uint32_t s = off[v], e = off[v + 1];
    if (e - s > kHubDeg) continue;  // k_seg_cedge
    if (gl == 0)
      for (uint32_t t = tm; t; t &= t - 1) acc[gw][__builtin_ctz(t)] = 0;
    for (uint32_t j = s; j < e; j += kRowUnroll * G) {
      uint32_t x[kRowUnroll], m[kRowUnroll];
#pragma unroll
      for (int u = 0; u < kRowUnroll; u++) {
        const uint32_t i = j + u * G + gl;
        x[u] = i < e ? adj[i] : 0u;
      }
#pragma unroll
      for (int u = 0; u < kRowUnroll; u++) m[u] = j + u * G + gl < e ? target_bits(sc, x[u], tm) : 0u;
      for (uint32_t t = tm; t; t &= t - 1) {
        const uint32_t b = __builtin_ctz(t);
        uint32_t got = 0;
#pragma unroll
        for (int u = 0; u < kRowUnroll; u++) got += (uint32_t)__popcll(__ballot((m[u] >> b) & 1u) & gmask);
        if (gl == 0) acc[gw][b] += got;
      }
    }
    if (gl == 0)
      for (uint32_t t = tm; t; t &= t - 1) {
        const uint32_t b = __builtin_ctz(t);
        cnt[P[sc.pidx[b]].flat + i0] = acc[gw][b];
      }
  }
}

template <int G>
__global__ void __launch_bounds__(FSIG_BLOCK) k_bcedge_fill_multi(const BSrc* __restrict__ S, uint32_t ns,
                                                                  uint64_t total, const BPair* __restrict__ P,
                                                                  const uint32_t* __restrict__ off,
                                                                  const uint32_t* __restrict__ adj,
                                                                  const uint64_t* __restrict__ gscan,
                                                                  uint32_t* __restrict__ lst_all, uint64_t f_lo) {
  constexpr int GPB = FSIG_BLOCK / G;
  __shared__ uint64_t at[GPB][MAXQ];
  const uint32_t lane = threadIdx.x & 63u, gl = lane & (G - 1), gw = threadIdx.x / G;
  const uint64_t gmask = (G == 64) ? ~0ull : (((1ull << G) - 1) << (lane - gl));
  const uint64_t below = (1ull << lane) - 1;
  const XcdRange xr = xcd_range(total - f_lo, GPB, gw);  // source rows [f_lo, total)
  for (uint64_t f = f_lo + xr.first; f < f_lo + xr.hi; f += xr.stride) {
    const BSrc& sc = S[find_flat(S, ns, f)];
    const uint32_t i0 = (uint32_t)(f - sc.flat);
    const uint32_t v = sc.acand[i0];
    const uint32_t tm = sc.tmask;
    const uint32_t s = off[v], e = off[v + 1];
    if (e - s > kHubDeg) continue;  // k_seg_cedge
    if (gl == 0)
      for (uint32_t t = tm; t; t &= t - 1) {
        const uint32_t b = __builtin_ctz(t);
        const BPair& pd = P[sc.pidx[b]];
        const uint64_t pbase = gscan[pd.flat];
        const uint64_t a0 = gscan[pd.flat + i0];
        at[gw][b] = a0;
        pd.off_out[i0] = (uint32_t)(a0 - pbase);
        if (i0 + 1 == pd.asize) pd.off_out[pd.asize] = (uint32_t)(gscan[pd.flat + i0 + 1] - pbase);
      }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (uint32_t j = s; j < e; j += kRowUnroll * G) {
      uint32_t x[kRowUnroll], m[kRowUnroll];
#pragma unroll
      for (int u = 0; u < kRowUnroll; u++) {
        const uint32_t i = j + u * G + gl;
        x[u] = i < e ? adj[i] : 0u;
      }
#pragma unroll
      for (int u = 0; u < kRowUnroll; u++) m[u] = j + u * G + gl < e ? target_bits(sc, x[u], tm) : 0u;
      // strides in order, each stride's targets in order: every list ascends
#pragma unroll
      for (int u = 0; u < kRowUnroll; u++) {
        for (uint32_t t = tm; t; t &= t - 1) {
          const uint32_t b = __builtin_ctz(t);
          const bool hit = (m[u] >> b) & 1u;
          const uint64_t bal = __ballot(hit) & gmask;
          if (bal == 0) continue;
          const uint64_t a0 = at[gw][b];
          if (hit) {
            const BPair& pd = P[sc.pidx[b]];
            lst_all[a0 + __popcll(bal & below)] = pd.bpos ? pd.bpos[x[u]] : sorted_index(pd.bcand, pd.bsize, x[u]);
          }
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
          if (gl == 0) at[gw][b] = a0 + __popcll(bal);
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
      }
    }
  }
}

// Longest candidate-edge list of each pair (one block per pair), from the
// list offsets.
__global__ void __launch_bounds__(FSIG_BLOCK) k_bpair_maxlen(const BPair* __restrict__ P,
                                                             const uint64_t* __restrict__ gscan,
                                                             unsigned int* __restrict__ maxlen) {
  __shared__ uint32_t red[FSIG_BLOCK];
  const BPair& pd = P[blockIdx.x];
  uint32_t mx = 0;
  for (uint32_t i = threadIdx.x; i < pd.asize; i += FSIG_BLOCK)
    mx = max(mx, (uint32_t)(gscan[pd.flat + i + 1] - gscan[pd.flat + i]));
  red[threadIdx.x] = mx;
  __syncthreads();
  for (uint32_t o = FSIG_BLOCK / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) maxlen[blockIdx.x] = red[0];
}

// ---------------------------------------------------------------------------
// Persistent depth-first join.
struct BLevel {
  uint32_t nb, ncf, cfsame, any_cross;
  uint32_t wsize;
  uint32_t stack_off;   // words of the group stack before this level's list
  uint32_t bcol[MAXQ];
  uint32_t cfcol[MAXQ];
  const uint32_t* boff[MAXQ];
  const uint32_t* blst[MAXQ];
  const uint32_t* cfcand[MAXQ];
  const uint32_t* wcand;
};

struct BQuery {
  uint32_t k;
  uint32_t level_base;  // index of this query's level 0 in the level table
  // Edge tasks: when join position 1 has position 0 as its only backward
  // neighbour, a task is a range of entries of that candidate-edge list,
  // i.e. of (root, first neighbour) pairs, so a hub root's subtree is spread
  // over many tasks.  n1 = roots, off1/lst1 = the list.
  uint32_t edge_tasks;
  uint32_t n1;
  const uint32_t* off1;
  const uint32_t* lst1;
};

struct Task {
  uint32_t q;
  uint32_t r0, r1;  // roots, or entries of the first list for edge tasks
};

// The candidates of level p (the join position p) for the path in `path`:
// intersection of the backward lists minus earlier vertices of the same
// label; written to `dst` when STORE, counted otherwise.
template <int G, bool STORE, bool LDS_STAGE>
__device__ __forceinline__ uint32_t level_candidates(const BLevel& L, const uint32_t* path, uint32_t* sh_s,
                                                     uint32_t* sh_n, uint32_t* sh_cf, uint32_t* tile,
                                                     uint32_t tile_words, uint32_t gl, uint64_t gmask,
                                                     uint64_t below, uint32_t* dst, uint64_t& bytes) {
  uint32_t dlen = L.wsize, d = 0;
  uint64_t el = 0;
  for (uint32_t i = 0; i < L.nb; i++) {
    const uint32_t x = path[L.bcol[i]];
    const uint32_t* o = L.boff[i];
    const uint32_t s = o[x], n = o[x + 1] - s;
    sh_s[i] = s;
    sh_n[i] = n;
    el += n;
    if (i == 0 || n < dlen) { dlen = n; d = i; }
  }
  for (uint32_t j = 0; j < L.ncf; j++) {
    uint32_t x = path[L.cfcol[j]];
    if (!((L.cfsame >> j) & 1u)) x = L.cfcand[j][x];
    sh_cf[j] = x;
  }
  bytes += 8ull * L.nb + 4ull * el;
  // LDS staging of the second list (the common two-neighbour case) when it
  // fits the group's tile: the probes then stay in LDS.
  bool staged = false;
  uint32_t sidx = 0;
  if (LDS_STAGE && L.nb == 2) {
    sidx = d ^ 1u;
    if (sh_n[sidx] <= tile_words) {
      const uint32_t* src = L.blst[sidx] + sh_s[sidx];
      for (uint32_t t = gl; t < sh_n[sidx]; t += G) tile[t] = src[t];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      staged = true;
    }
  }
  const uint32_t* dl = L.nb ? L.blst[d] + sh_s[d] : nullptr;
  uint32_t base = 0;
  if (!STORE && L.nb == 1) {
    // one backward list: the count is its length minus the earlier vertices
    // of the same label found in it (binary search, one lane per conflict)
    uint32_t hit = 0;
    for (uint32_t j0 = 0; j0 < L.ncf; j0 += G) {
      const uint32_t j = j0 + gl;
      bool in = false;
      if (j < L.ncf) {
        uint32_t target = sh_cf[j];
        bool ok = true;
        if (!((L.cfsame >> j) & 1u)) {
          const uint32_t at = sorted_index(L.wcand, L.wsize, target);
          ok = at < L.wsize && L.wcand[at] == target;
          target = at;
        }
        in = ok && sorted_contains(dl, dlen, target);
      }
      hit += __popcll(__ballot(in) & gmask);
    }
    bytes -= 4ull * el;  // the list itself is not read
    bytes += 4ull * L.ncf * (32 - __builtin_clz(dlen | 1));
    if (staged) __builtin_amdgcn_wave_barrier();
    return dlen - hit;
  }
  for (uint32_t t0 = 0; t0 < dlen; t0 += G) {
    const uint32_t t = t0 + gl;
    bool ok = t < dlen;
    uint32_t x = 0;
    if (ok) x = L.nb ? dl[t] : t;
    if (staged) {
      if (ok) ok = sorted_contains(tile, sh_n[sidx], x);
    } else {
      for (uint32_t i = 0; i < L.nb && ok; i++) {
        if (i == d) continue;
        ok = sorted_contains(L.blst[i] + sh_s[i], sh_n[i], x);
      }
    }
    if (ok && L.ncf) {
      const uint32_t xid = L.any_cross ? L.wcand[x] : x;
      for (uint32_t j = 0; j < L.ncf && ok; j++) ok = (((L.cfsame >> j) & 1u) ? x : xid) != sh_cf[j];
    }
    const uint64_t bal = __ballot(ok) & gmask;
    if (STORE && ok) dst[base + __popcll(bal & below)] = x;
    base += __popcll(bal);
  }
  if (staged) __builtin_amdgcn_wave_barrier();
  if (STORE) bytes += 4ull * base;
  return base;
}

// Work queue of the persistent join.  Seed tasks (ranges of roots or of
// first-list entries, built on the host) are taken first; donated tasks are
// (query, level p, path prefix, candidate range at p) records pushed by busy
// groups while other groups wait, each taken by one ticket.  `outstanding`
// counts seed chunks plus donated tasks not finished yet; a waiting group
// leaves when it reaches 0 (no task exists and none can be created).
struct DTask {
  uint32_t q, p, a, b;
  uint32_t path[MAXQ];
};

struct Queue {
  unsigned int seed_head;
  unsigned int don_head;     // tickets taken
  unsigned int don_reserve;  // slots reserved by donors
  unsigned int outstanding;
  unsigned int error;        // a wait ran past its bound
  unsigned int pad[3];
};

constexpr uint64_t kSpinLimit = 1ull << 27;
constexpr uint32_t kDonateMin = 4;

// The persistent join is a state machine: every iteration of one convergent
// loop advances each lane group by one bounded step (take a task, one DFS
// step, or one poll of the queue).  Groups sharing a wavefront therefore
// interleave; a group waiting for donated work never holds up the wave-mate
// that would donate it.
enum : uint32_t { kNeed = 0, kSeedItem = 1, kWalk = 2, kWait = 3, kDone = 4 };

// Count of the last level for one lane's own candidate x of level p (the
// path is path[0..p-1] followed by x): each lane works alone, so a group
// counts G leaf parents per step.  Returns false when the level has more
// backward neighbours than the lane keeps in registers (the caller then uses
// the group-cooperative count).
constexpr uint32_t kLaneNb = 6;
constexpr uint32_t kLaneCf = 8;

__device__ __forceinline__ bool lane_count(const BLevel& L, const uint32_t* path, uint32_t p, uint32_t x,
                                           uint32_t& out, uint64_t& bytes) {
  if (L.nb > kLaneNb || L.ncf > kLaneCf) return false;
  uint32_t s[kLaneNb], n[kLaneNb];
  const uint32_t* l[kLaneNb];
  uint32_t d = 0, dn = 0xffffffffu;
#pragma unroll
  for (uint32_t i = 0; i < kLaneNb; i++) {
    if (i < L.nb) {
      const uint32_t c = L.bcol[i];
      const uint32_t v = c == p ? x : path[c];
      const uint32_t* o = L.boff[i];
      s[i] = o[v];
      n[i] = o[v + 1] - s[i];
      l[i] = L.blst[i] + s[i];
      if (n[i] < dn) { dn = n[i]; d = i; }
    }
  }
  bytes += 8ull * L.nb;
  if (L.nb == 0) { dn = L.wsize; }
  // conflicts in the index space of w (data vertex ids mapped by search)
  uint32_t cnt = 0;
  if (L.nb == 1) {
    cnt = dn;
    for (uint32_t j = 0; j < L.ncf; j++) {
      const uint32_t c = L.cfcol[j];
      uint32_t target = c == p ? x : path[c];
      bool ok = true;
      if (!((L.cfsame >> j) & 1u)) {
        const uint32_t id = L.cfcand[j][target];
        const uint32_t at = sorted_index(L.wcand, L.wsize, id);
        ok = at < L.wsize && L.wcand[at] == id;
        target = at;
      }
      if (ok && sorted_contains(l[0], dn, target)) cnt--;
    }
    bytes += 4ull * L.ncf * (32 - __builtin_clz(dn | 1));
  } else {
    uint32_t cf[kLaneCf];
#pragma unroll
    for (uint32_t j = 0; j < kLaneCf; j++) {
      if (j < L.ncf) {
        const uint32_t c = L.cfcol[j];
        uint32_t v = c == p ? x : path[c];
        if (!((L.cfsame >> j) & 1u)) v = L.cfcand[j][v];
        cf[j] = v;
      }
    }
    for (uint32_t t = 0; t < dn; t++) {
      const uint32_t y = L.nb ? l[d][t] : t;
      bool ok = true;
#pragma unroll
      for (uint32_t i = 0; i < kLaneNb; i++)
        if (i < L.nb && i != d && ok) ok = sorted_contains(l[i], n[i], y);
      if (ok && L.ncf) {
        const uint32_t yid = L.any_cross ? L.wcand[y] : y;
#pragma unroll
        for (uint32_t j = 0; j < kLaneCf; j++)
          if (j < L.ncf) ok &= (((L.cfsame >> j) & 1u) ? y : yid) != cf[j];
      }
      cnt += ok;
    }
    uint64_t el = 0;
#pragma unroll
    for (uint32_t i = 0; i < kLaneNb; i++)
      if (i < L.nb) el += n[i];
    bytes += 4ull * el;
  }
  out = cnt;
  return true;
}

template <int G, bool LDS_STAGE>
__global__ void __launch_bounds__(FSIG_BLOCK) k_dfs(const BQuery* __restrict__ Q, const BLevel* __restrict__ LV,
                                                    const Task* __restrict__ tasks, uint32_t ntasks,
                                                    Queue* __restrict__ qu, DTask* __restrict__ dtasks,
                                                    unsigned int* __restrict__ ready, uint32_t dcap,
                                                    uint32_t* __restrict__ stacks, uint32_t stack_words,
                                                    uint32_t tile_words, unsigned long long* __restrict__ counts,
                                                    unsigned long long* __restrict__ stats) {
  constexpr int GPB = FSIG_BLOCK / G;
  __shared__ uint32_t sh_path[GPB][MAXQ];
  __shared__ uint32_t sh_len[GPB][MAXQ];
  __shared__ uint32_t sh_pos[GPB][MAXQ];
  __shared__ uint32_t sh_s[GPB][MAXQ];
  __shared__ uint32_t sh_n[GPB][MAXQ];
  __shared__ uint32_t sh_cf[GPB][MAXQ];
  extern __shared__ uint32_t sh_tile[];
  const uint32_t lane = threadIdx.x & 63u, gl = lane & (G - 1), gw = threadIdx.x / G;
  const uint64_t gmask = (G == 64) ? ~0ull : (((1ull << G) - 1) << (lane - gl));
  const uint64_t below = (1ull << lane) - 1;
  const uint64_t gid = (uint64_t)blockIdx.x * GPB + gw;
  uint32_t* stk = stacks + gid * stack_words;
  uint32_t* tile = sh_tile + gw * tile_words;
  uint32_t* path = sh_path[gw];
  uint32_t* len = sh_len[gw];
  uint32_t* pos = sh_pos[gw];
  uint64_t bytes = 0, nodes = 0, spins = 0;
  // group state (uniform within the group)
  uint32_t st = kNeed;
  bool seeds_left = true;
  bool from_seed = false;
  uint32_t qi = 0, k = 0, p0 = 0, p = 0, item = 0, item_end = 0, ticket = 0, since_check = 0;
  const BLevel* lv = LV;
  unsigned long long cnt = 0;

  // Starts the walk below path[0..s0-1] over candidates [a, b) of level s0.
  auto start_walk = [&](uint32_t s0, uint32_t a, uint32_t b) {
    if (s0 + 1 == k) {  // level s0 is the last one: count it
      cnt += level_candidates<G, false, LDS_STAGE>(lv[s0], path, sh_s[gw], sh_n[gw], sh_cf[gw], tile, tile_words, gl,
                                                   gmask, below, nullptr, bytes);
      return false;
    }
    const uint32_t full = level_candidates<G, true, LDS_STAGE>(lv[s0], path, sh_s[gw], sh_n[gw], sh_cf[gw], tile,
                                                               tile_words, gl, gmask, below, stk + lv[s0].stack_off,
                                                               bytes);
    len[s0] = b < full ? b : full;
    pos[s0] = a;
    p0 = s0;
    p = s0;
    nodes++;
    return true;
  };
  auto finish_task = [&]() {
    if (gl == 0) {
      atomicAdd(&counts[qi], cnt);
      atomicSub(&qu->outstanding, 1u);
    }
    cnt = 0;
    st = kNeed;
  };

  while (st != kDone) {
    if (st == kNeed) {
      if (seeds_left) {
        uint32_t ti = 0;
        if (gl == 0) ti = atomicAdd(&qu->seed_head, 1u);
        ti = __shfl(ti, 0, G);
        if (ti < ntasks) {
          const Task T = tasks[ti];
          qi = T.q;
          const BQuery q = Q[qi];
          k = q.k;
          lv = LV + q.level_base;
          item = T.r0;
          item_end = T.r1;
          cnt = 0;
          st = kSeedItem;
        } else {
          seeds_left = false;
        }
      } else {
        uint32_t t = 0;
        if (gl == 0) t = atomicAdd(&qu->don_head, 1u);
        t = __shfl(t, 0, G);
        if (t >= dcap) {
          st = kDone;  // no task can ever land in this slot
        } else {
          ticket = t;
          spins = 0;
          st = kWait;
        }
      }
    } else if (st == kSeedItem) {
      if (item >= item_end) {
        finish_task();
      } else {
        const BQuery q = Q[qi];
        uint32_t s0;
        bool skip = false;
        if (q.edge_tasks) {
          uint32_t lo = 0, hi = q.n1;  // root: last r with off1[r] <= item
          while (hi - lo > 1) {
            const uint32_t m = (lo + hi) >> 1;
            if (q.off1[m] <= item) lo = m; else hi = m;
          }
          path[0] = lo;
          path[1] = q.lst1[item];
          s0 = 2;
          if (k == 2) { cnt++; skip = true; }
        } else {
          path[0] = item;
          s0 = 1;
          if (k == 1) { cnt++; skip = true; }
        }
        item++;
        if (!skip && start_walk(s0, 0, 0xffffffffu)) {
          from_seed = true;
          since_check = 0;
          st = kWalk;
        }
      }
    } else if (st == kWait) {
      // Relaxed polls: an agent-scope acquire invalidates the L2 of the XCD,
      // so it is issued once, after the flag is seen.
      unsigned int state = 0;  // 1 ready, 2 leave
      if (gl == 0) {
        if (__hip_atomic_load(&ready[ticket], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) state = 1;
        else if (__hip_atomic_load(&qu->outstanding, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) state = 2;
        else if (++spins > kSpinLimit) { atomicExch(&qu->error, 1u); state = 2; }
      }
      state = __shfl(state, 0, G);
      if (state == 2) {
        st = kDone;
      } else if (state == 1) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        const DTask* dt = dtasks + ticket;
        qi = dt->q;
        const uint32_t tp = dt->p, a = dt->a, b = dt->b;
        for (uint32_t c = gl; c < tp; c += G) path[c] = dt->path[c];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const BQuery q = Q[qi];
        k = q.k;
        lv = LV + q.level_base;
        cnt = 0;
        from_seed = false;
        since_check = 0;
        if (start_walk(tp, a, b)) st = kWalk;
        else finish_task();
      } else {
        __builtin_amdgcn_s_sleep(1);
      }
    } else {  // kWalk: one step of the depth-first walk
      if (pos[p] >= len[p]) {
        if (p == p0) {
          if (from_seed) st = kSeedItem;
          else finish_task();
        } else {
          p--;
        }
        continue;
      }
      // Donation: while groups wait for work, hand half of the shallowest
      // level with at least kDonateMin candidates left to the queue (rare and
      // coarse: each donation costs one agent-scope release).
      if (++since_check >= 32) {
        since_check = 0;
        unsigned int hungry = 0;
        if (gl == 0)
          hungry = __hip_atomic_load(&qu->don_head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >
                   __hip_atomic_load(&qu->don_reserve, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        hungry = __shfl(hungry, 0, G);
        if (hungry) {
          uint32_t dp = p0;
          while (dp <= p && len[dp] - pos[dp] < kDonateMin) dp++;
          if (dp <= p) {
            unsigned int slot = 0;
            if (gl == 0) slot = atomicAdd(&qu->don_reserve, 1u);
            slot = __shfl(slot, 0, G);
            if (slot < dcap) {
              const uint32_t mid = pos[dp] + (len[dp] - pos[dp]) / 2;
              DTask* t = dtasks + slot;
              for (uint32_t c = gl; c < dp; c += G) t->path[c] = path[c];
              if (gl == 0) {
                t->q = qi;
                t->p = dp;
                t->a = mid;
                t->b = len[dp];
              }
              len[dp] = mid;
              if (gl == 0) atomicAdd(&qu->outstanding, 1u);
              __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
              if (gl == 0) __hip_atomic_store(&ready[slot], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
          }
        }
      }
      if (p + 2 == k && lv[p + 1].nb <= kLaneNb && lv[p + 1].ncf <= kLaneCf) {
        // leaf parents: each lane counts the last level for its own candidate
        const uint32_t take = len[p] - pos[p] < G ? len[p] - pos[p] : G;
        uint32_t mine = 0;
        if (gl < take) {
          const uint32_t x = stk[lv[p].stack_off + pos[p] + gl];
          lane_count(lv[p + 1], path, p, x, mine, bytes);
        }
        unsigned long long sum = mine;
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) sum += __shfl_xor(sum, o, G);
        cnt += sum;
        pos[p] = pos[p] + take;
        nodes += take;
        continue;
      }
      const uint32_t x = stk[lv[p].stack_off + pos[p]];
      pos[p] = pos[p] + 1;
      path[p] = x;
      nodes++;
      if (p + 2 == k) {
        cnt += level_candidates<G, false, LDS_STAGE>(lv[p + 1], path, sh_s[gw], sh_n[gw], sh_cf[gw], tile,
                                                     tile_words, gl, gmask, below, nullptr, bytes);
      } else {
        len[p + 1] = level_candidates<G, true, LDS_STAGE>(lv[p + 1], path, sh_s[gw], sh_n[gw], sh_cf[gw], tile,
                                                          tile_words, gl, gmask, below, stk + lv[p + 1].stack_off,
                                                          bytes);
        pos[p + 1] = 0;
        p++;
      }
    }
  }
  if (stats && gl == 0) {
    atomicAdd(&stats[0], (unsigned long long)bytes);
    atomicAdd(&stats[1], (unsigned long long)nodes);
  }
}

// ---------------------------------------------------------------------------
// Lockstep batched join: level p of every query that still expands is one
// table (rows of p candidate indices, queries back to back); one count
// launch, one scan and one write launch serve the whole batch per level.
// The per-row work is the plain join's (k_join); the level description is
// looked up per row from the query the row belongs to.
struct MLevel {
  JoinLevel L;
  CandEdgeSrc S;
  uint32_t last;  // level p+1 is the query's last: count only
  uint32_t qid;   // batch index of the query
  uint32_t tail;  // level p+2 is the last: both counted in one pass (L2, S2)
  JoinLevel L2;
  CandEdgeSrc S2;
};

template <int G>
struct PlainShared {
  static constexpr int GPB = FSIG_BLOCK / G;
  uint32_t s1[GPB][MAXQ], n1[GPB][MAXQ], cf1[GPB][MAXQ];
};

template <int G, bool WRITE>
__global__ void __launch_bounds__(FSIG_BLOCK) k_join_multi(const MLevel* __restrict__ ML,
                                                           const uint64_t* __restrict__ qbase, uint32_t nlive,
                                                           const uint32_t* __restrict__ rows, uint64_t row_lo,
                                                           uint64_t row_hi,
                                                           uint32_t p, uint32_t* __restrict__ cnt,
                                                           const uint64_t* __restrict__ off,
                                                           uint32_t* __restrict__ out,
                                                           unsigned long long* __restrict__ qcounts,
                                                           unsigned long long* __restrict__ stats,
                                                           uint32_t tile_words) {
  constexpr int GPB = FSIG_BLOCK / G;
  constexpr int NR = (MAXQ + G - 1) / G;
  // the count pass of G >= 8 also runs fused tails (their LDS tiles)
  constexpr bool TAIL = !WRITE && G >= 8;
  using Sh = typename std::conditional<TAIL, TailShared<G>, PlainShared<G>>::type;
  __shared__ Sh shm;
  extern __shared__ uint32_t tail_tile[];
  auto& sh_s = shm.s1;
  auto& sh_n = shm.n1;
  auto& sh_cf = shm.cf1;
  const uint32_t lane = threadIdx.x & 63u, gl = lane & (G - 1), gw = threadIdx.x / G;
  const uint64_t gmask = (G == 64) ? ~0ull : (((1ull << G) - 1) << (lane - gl));
  const uint64_t below = (1ull << lane) - 1;
  const uint64_t ngroups = (uint64_t)gridDim.x * GPB;
  unsigned long long my_bytes = 0, my_units = 0;
  uint32_t cur_q = 0xffffffffu;
  unsigned long long acc = 0;
  for (uint64_t r = row_lo + (uint64_t)xcd_block() * GPB + gw; r < row_hi; r += ngroups) {
    uint32_t lo = 0, hi = nlive;  // query of the row: last j with qbase[j] <= r
    while (hi - lo > 1) {
      const uint32_t m = (lo + hi) >> 1;
      if (qbase[m] <= r) lo = m; else hi = m;
    }
    const MLevel& M = ML[lo];
    const JoinLevel& L = M.L;
    const bool last = M.last != 0;
    if (WRITE && (last || off[r + 1] == off[r])) continue;  // counted in the count pass, or no children
    uint32_t rv[NR];
    const uint32_t* row = rows + r * p;
#pragma unroll
    for (int j = 0; j < NR; j++) {
      const uint32_t col = j * G + gl;
      rv[j] = col < p ? row[col] : 0u;
    }
    if constexpr (TAIL) {
      if (M.tail) {
        uint64_t kids = 0;
        bool defer = false;
        const uint64_t leaves = group_sum_u64<G>(tail_row<G>(L, M.S, M.L2, M.S2, rv, shm.at(gw, tail_tile, tile_words), gl, gmask, below,
                                                             kids, defer, my_bytes));
        if (gl == 0) {
          my_units += 1;
          if (M.qid != cur_q) {
            if (cur_q != 0xffffffffu && acc) atomicAdd(&qcounts[cur_q], acc);
            cur_q = M.qid;
            acc = 0;
          }
          acc += leaves;
          cnt[r] = defer ? (uint32_t)kids : 0u;  // a deferred row's children are written
        }
        continue;
      }
    }
    uint32_t dlen = L.wsize, d = 0;
    uint64_t elems = 0;
    const uint32_t nb = L.nb, ncf = L.ncf;
    for (uint32_t i = 0; i < nb; i++) {
      const uint32_t x = group_col<G>(rv, L.bcol[i]);
      const uint32_t* o = M.S.off[i];
      const uint32_t s0 = o[x], n0 = o[x + 1] - s0;
      sh_s[gw][i] = s0;
      sh_n[gw][i] = n0;
      elems += n0;
      if (i == 0 || n0 < dlen) { dlen = n0; d = i; }
    }
    for (uint32_t j = 0; j < ncf; j++) {
      uint32_t x = group_col<G>(rv, L.cfcol[j]);
      if (!((L.cfsame >> j) & 1u)) x = L.cfcand[j][x];
      sh_cf[gw][j] = x;
    }
    const uint32_t* dl = nb ? M.S.lst[d] + sh_s[gw][d] : nullptr;
    uint64_t base = 0;
    const bool direct = !WRITE && nb == 1;
    if (direct) {
      uint32_t hit = 0;
      for (uint32_t j0 = 0; j0 < ncf; j0 += G) {
        const uint32_t j = j0 + gl;
        bool in = false;
        if (j < ncf) {
          uint32_t target = sh_cf[gw][j];
          bool ok = true;
          if (!((L.cfsame >> j) & 1u)) {
            const uint32_t at = sorted_index(L.wcand, L.wsize, target);
            ok = at < L.wsize && L.wcand[at] == target;
            target = at;
          }
          in = ok && sorted_contains(dl, dlen, target);
        }
        hit += __popcll(__ballot(in) & gmask);
      }
      base = dlen - hit;
      if (gl == 0) my_bytes += 4ull * ncf * (32 - __builtin_clz(dlen | 1));
      elems = 0;
    }
    uint64_t child0 = WRITE ? off[r] : 0;
    for (uint32_t t0 = 0; !direct && t0 < dlen; t0 += G) {
      const uint32_t t = t0 + gl;
      bool ok = t < dlen;
      uint32_t x = 0;
      if (ok) x = nb ? dl[t] : t;
      for (uint32_t i = 0; i < nb && ok; i++) {
        if (i == d) continue;
        ok = sorted_contains(M.S.lst[i] + sh_s[gw][i], sh_n[gw][i], x);
      }
      if (ok && ncf) {
        const uint32_t xid = L.any_cross ? L.wcand[x] : x;
        for (uint32_t j = 0; j < ncf && ok; j++) ok = (((L.cfsame >> j) & 1u) ? x : xid) != sh_cf[gw][j];
      }
      const uint64_t bal = __ballot(ok) & gmask;
      if (WRITE) {
        uint32_t* dst = out + (child0 + base + __popcll(bal & below)) * (p + 1);
        for (uint32_t c = 0; c < p; c++) {
          const uint32_t v = group_col<G>(rv, c);
          if (ok) dst[c] = v;
        }
        if (ok) dst[p] = x;
      }
      base += __popcll(bal);
    }
    if (gl == 0) {
      my_bytes += 4ull * p + 8ull * nb + 4ull * elems + (WRITE ? 8 + base * 4ull * (p + 1) : 4);
      my_units += WRITE ? base : 1;
      if (!WRITE) {
        if (last) {
          if (M.qid != cur_q) {
            if (cur_q != 0xffffffffu && acc) atomicAdd(&qcounts[cur_q], acc);
            cur_q = M.qid;
            acc = 0;
          }
          acc += base;
          cnt[r] = 0;
        } else {
          cnt[r] = (uint32_t)base;
        }
      }
    }
  }
  if (!WRITE && gl == 0 && cur_q != 0xffffffffu && acc) atomicAdd(&qcounts[cur_q], acc);
  my_bytes = wave_sum_u64(my_bytes);
  my_units = wave_sum_u64(my_units);
  if (stats && lane == 0) {
    atomicAdd(&stats[0], my_bytes);
    atomicAdd(&stats[1], my_units);
  }
}

__global__ void k_iota_multi(const uint64_t* __restrict__ qbase, uint32_t nlive, uint32_t* __restrict__ rows,
                             uint64_t nrows) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrows; r += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t lo = 0, hi = nlive;
    while (hi - lo > 1) {
      const uint32_t m = (lo + hi) >> 1;
      if (qbase[m] <= r) lo = m; else hi = m;
    }
    rows[r] = (uint32_t)(r - qbase[lo]);
  }
}

// off values at the query boundaries of a level: the next level's row bases.
__global__ void k_gather_bases(const uint64_t* __restrict__ off, const uint64_t* __restrict__ qbase, uint32_t n,
                               uint64_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= n) out[i] = off[qbase[i]];
}


// Resets the work queue for a run over `seeds` seed chunks (a kernel, so the
// reset can live inside the captured graph).
__global__ void k_queue_reset(Queue* qu, unsigned int* ready, uint32_t dcap, uint32_t seeds) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < dcap; i += gridDim.x * blockDim.x) ready[i] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    qu->seed_head = 0;
    qu->don_head = 0;
    qu->don_reserve = 0;
    qu->outstanding = seeds;
    qu->error = 0;
  }
}

// ---------------------------------------------------------------------------
struct TuneChoice {
  uint32_t G;
  bool lds;
  uint32_t tile;
  bool lockstep;  // lockstep batched BFS instead of the persistent walk
};

struct DfsArgs {
  const BQuery* Q;
  const BLevel* LV;
  const Task* T;
  uint32_t nt;
  Queue* qu;
  DTask* dtasks;
  unsigned int* ready;
  uint32_t dcap;
  uint32_t* stacks;
  uint32_t stack_words;
  unsigned long long* counts;
  unsigned long long* stats;
};

template <int G, bool LDS>
int launch_dfs_t(fsig_ctx* c, uint32_t grid, hipStream_t st, const DfsArgs& a, uint32_t tile) {
  const size_t dyn = LDS ? (size_t)(FSIG_BLOCK / G) * tile * 4 : 0;
  k_dfs<G, LDS><<<grid, FSIG_BLOCK, dyn, st>>>(a.Q, a.LV, a.T, a.nt, a.qu, a.dtasks, a.ready, a.dcap, a.stacks,
                                               a.stack_words, tile, a.counts, a.stats);
  FSIG_HIP(c, hipGetLastError());
  return FSIG_OK;
}

// Queue reset + persistent join for a run over the first `nt` seed chunks.
int launch_dfs(fsig_ctx* c, const TuneChoice& tc, uint32_t grid, hipStream_t st, const DfsArgs& a) {
  k_queue_reset<<<(a.dcap + 1023) / 1024, 1024, 0, st>>>(a.qu, a.ready, a.dcap, a.nt);
  FSIG_HIP(c, hipGetLastError());
  const uint32_t tile = tc.lds ? tc.tile : 0;
#define FSIG_DFS_CASE(GG) \
  case GG:                \
    return tc.lds ? launch_dfs_t<GG, true>(c, grid, st, a, tile) : launch_dfs_t<GG, false>(c, grid, st, a, tile);
  switch (tc.G) {
    FSIG_DFS_CASE(4)
    FSIG_DFS_CASE(8)
    FSIG_DFS_CASE(16)
    FSIG_DFS_CASE(32)
    FSIG_DFS_CASE(64)
  }
#undef FSIG_DFS_CASE
  return set_err(c, FSIG_EINVAL, "dfs: bad lanes per group");
}

uint32_t bgrid(const fsig_ctx* c, uint64_t items, uint32_t per_block) {
  uint64_t b = (items + per_block - 1) / per_block;
  b = std::min<uint64_t>(b, (uint64_t)c->num_cus * 16);
  return (uint32_t)std::max<uint64_t>(b, 1);
}

template <class T>
T* upload_table(fsig_ctx* c, const std::vector<T>& h) {
  T* d = (T*)c->arena.alloc(sizeof(T) * std::max<size_t>(h.size(), 1));
  if (!d) return nullptr;
  if (!h.empty() && hipMemcpyAsync(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, c->stream) != hipSuccess)
    return nullptr;
  return d;
}

}  // namespace

// Stage timing on stderr when FSIG_DEBUG_TIMING is set (host clock, after a
// stream synchronisation).
struct StageClock {
  bool on;
  fsig_ctx* c;
  std::chrono::steady_clock::time_point t0;
  explicit StageClock(fsig_ctx* ctx) : on(getenv("FSIG_DEBUG_TIMING") != nullptr), c(ctx),
                                       t0(std::chrono::steady_clock::now()) {}
  void mark(const char* what) {
    if (!on) return;
    (void)hipStreamSynchronize(c->stream);
    const auto t = std::chrono::steady_clock::now();
    fprintf(stderr, "[fsig batch] %-24s %9.3f ms\n", what, std::chrono::duration<double, std::milli>(t - t0).count());
    t0 = t;
  }
};

// Beam of a bounded query in the lockstep join: the children of its first
// parent rows up to this many (fsig_tune.beam_log2 overrides).  Queries with
// more embeddings than their bound usually find them inside the beam, all in
// the batch's shared launches.
constexpr uint32_t kBeamLog2 = 16;

// Rows of a lockstep level whose first child index reaches their query's cut
// get no children (cnt = 0); first[j] receives query j's first such row (a
// query's rows and their child offsets ascend, so its cut rows are a suffix
// of its range; first[j] keeps its initial value when nothing is cut).
__global__ void k_beam_cut(const uint64_t* __restrict__ off, const uint64_t* __restrict__ qbase, uint32_t nlive,
                           const uint64_t* __restrict__ cut, uint64_t nrows, uint32_t* __restrict__ cnt,
                           uint64_t* __restrict__ first) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrows; r += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t lo = 0, hi = nlive;  // query of the row: last j with qbase[j] <= r
    while (hi - lo > 1) {
      const uint32_t m = (lo + hi) >> 1;
      if (qbase[m] <= r) lo = m; else hi = m;
    }
    if (off[r] >= cut[lo]) {
      cnt[r] = 0;
      if (r == qbase[lo] || off[r - 1] < cut[lo]) first[lo] = r;
    }
  }
}

// Level descriptions of one query of a batch for the single-query driver.
struct BatchProvider {
  using SrcType = CandEdgeSrc;
  std::function<void(uint32_t, JoinLevel&, CandEdgeSrc&)> fill;
  JoinLevel level(uint32_t p) const {
    JoinLevel L;
    CandEdgeSrc S;
    std::memset(&L, 0, sizeof(L));
    std::memset(&S, 0, sizeof(S));
    fill(p, L, S);
    return L;
  }
  CandEdgeSrc src(uint32_t p) const {
    JoinLevel L;
    CandEdgeSrc S;
    std::memset(&L, 0, sizeof(L));
    std::memset(&S, 0, sizeof(S));
    fill(p, L, S);
    return S;
  }
  bool chain_tail(uint32_t, ChainTail&) const { return false; }
};

// bounds (may be null): per query, 0 = exact count, else the count is
// min(count, bounds[i]) and the search of query i stops once that many
// embeddings are found (fsig_match's max_embeddings).
static int batch_impl(fsig_ctx* c, const fsig_graph* g, uint32_t nq, const uint32_t* ks,
                      const int32_t* const* qlabels, const uint32_t* const* qoff, const uint32_t* const* qadj,
                      const uint64_t* bounds, uint64_t* counts_out, int32_t* status_out, double* join_ms) {
  if (!c) return FSIG_EINVAL;
  if (!g || !ks || !qlabels || !qoff || !qadj || !counts_out) return set_err(c, FSIG_EINVAL, "batch: null argument");
  if (join_ms) *join_ms = 0;
  if (nq == 0) return FSIG_OK;
  FSIG_HIP(c, hipSetDevice(c->device));
  std::vector<uint64_t> bound(nq, 0);
  bool any_bound = false;
  for (uint32_t i = 0; bounds && i < nq; i++) {
    bound[i] = bounds[i];
    any_bound |= bounds[i] != 0;
  }
  // results of bounded queries finished outside the lockstep join
  std::vector<uint8_t> settled(nq, 0);
  std::vector<uint64_t> settled_count(nq, 0);
  std::vector<int> settled_rc(nq, FSIG_OK);
  std::vector<uint8_t> beamed(nq, 0);  // lockstep kept only part of its rows: its count is a lower bound
  const uint32_t n = g->n;
  Arena& A = c->arena;
  const size_t mark = A.mark();
  A.peak = mark;
  struct Guard {
    Arena& a;
    size_t m;
    ~Guard() { a.release(m); }
  } guard{A, mark};
  int rc;
  StageClock clk(c);
  // ---- plans
  std::vector<QueryPlan> plans(nq);
  std::vector<uint8_t> active(nq, 1);
  for (uint32_t i = 0; i < nq; i++) {
    counts_out[i] = 0;
    rc = plan_classes(plans[i], ks[i], qlabels[i], qoff[i], qadj[i], nullptr, FSIG_ALG_MMATCH, c->err);
    if (rc) {
      c->err = "batch query " + std::to_string(i) + ": " + c->err;
      return rc;
    }
    for (uint32_t u = 0; u < ks[i]; u++)
      if ((uint32_t)qlabels[i][u] >= g->num_labels) active[i] = 0;
  }
  // global classes
  std::vector<uint32_t> cls_base(nq + 1, 0);
  for (uint32_t i = 0; i < nq; i++) cls_base[i + 1] = cls_base[i] + plans[i].nclass;
  const uint32_t ncls = cls_base[nq];
  KTimer tf(c, FSIG_K_FILTER);
  uint32_t* masks = (uint32_t*)A.alloc(4ull * std::max<uint32_t>(n, 1) * nq);
  if (!masks) return set_err(c, FSIG_ENOMEM, "batch: arena exhausted for candidate masks");
  k_fill_u32<<<bgrid(c, (uint64_t)n * nq, FSIG_BLOCK), FSIG_BLOCK, 0, c->stream>>>(masks, (uint64_t)n * nq, 0);
  // LDF segments and compaction segments
  std::vector<BSeg> segs;
  std::vector<uint32_t> cdeg(std::max<uint32_t>(ncls, 1), 0);
  std::vector<BComp> comp(ncls);
  std::vector<uint32_t> cap(ncls, 0);
  uint64_t flat = 0, tiles = 0;
  for (uint32_t i = 0; i < nq; i++) {
    const QueryPlan& P = plans[i];
    uint32_t* mask = masks + (uint64_t)i * n;
    for (uint32_t a = 0; a < P.nclass; a++) {
      cdeg[cls_base[i] + a] = P.class_deg[a];
      if (!active[i]) continue;
      const int32_t lab = P.class_label[a];
      bool found = false;
      for (size_t s = segs.size(); s-- > 0;) {
        if (segs[s].mask != mask) break;
        if (segs[s].src == g->d_inv_list + g->h_inv_index[lab]) {
          segs[s].cmask |= 1u << a;
          found = true;
          break;
        }
      }
      if (!found) {
        BSeg s;
        s.src = g->d_inv_list + g->h_inv_index[lab];
        s.mask = mask;
        s.len = g->h_inv_index[lab + 1] - g->h_inv_index[lab];
        s.cmask = 1u << a;
        s.cls_base = cls_base[i];
        s.pad = 0;
        s.flat = flat;
        flat += s.len;
        segs.push_back(s);
      }
    }
  }
  // empty segments would break the flat search; drop them
  std::vector<BSeg> segs_nz;
  for (auto& s : segs)
    if (s.len) segs_nz.push_back(s);
  const uint32_t* dcdeg = upload_table(c, cdeg);
  const BSeg* dsegs = upload_table(c, segs_nz);
  if (!dcdeg || !dsegs) return set_err(c, FSIG_ENOMEM, "batch: arena exhausted for tables");
  const bool use_nlf = g->d_nlf && c->tune.nlf_filter == 0;
  const uint32_t* dreq = nullptr;
  const uint32_t* dnreq = nullptr;
  if (use_nlf) {
    std::vector<uint32_t> req((size_t)std::max<uint32_t>(ncls, 1) * MAXQ, 0), nreq(std::max<uint32_t>(ncls, 1), 0);
    for (uint32_t i = 0; i < nq; i++)
      for (uint32_t a = 0; a < plans[i].nclass; a++) {
        const uint32_t gc = cls_base[i] + a;
        nreq[gc] = plans[i].nlf_n[a];
        for (uint32_t j = 0; j < plans[i].nlf_n[a]; j++) req[(size_t)gc * MAXQ + j] = plans[i].nlf_req[a][j];
      }
    dreq = upload_table(c, req);
    dnreq = upload_table(c, nreq);
    if (!dreq || !dnreq) return set_err(c, FSIG_ENOMEM, "batch: arena exhausted for tables");
  }
  if (flat)
    k_bldf<<<bgrid(c, flat, FSIG_BLOCK), FSIG_BLOCK, 0, c->stream>>>(dsegs, (uint32_t)segs_nz.size(), flat, dcdeg,
                                                                     g->d_off, use_nlf ? g->d_nlf : nullptr,
                                                                     g->num_labels, dreq, dnreq);
  std::vector<uint32_t*> cand_a(ncls, nullptr), cand_b(ncls, nullptr);
  std::vector<uint32_t> size(ncls, 0);
  for (uint32_t i = 0; i < nq; i++) {
    const QueryPlan& P = plans[i];
    for (uint32_t a = 0; a < P.nclass; a++) {
      const uint32_t gc = cls_base[i] + a;
      // an inactive query may carry a label the graph does not have: its
      // segments are empty and the label never indexes the inverted index
      const int32_t lab = P.class_label[a];
      const uint32_t len = active[i] ? g->h_inv_index[lab + 1] - g->h_inv_index[lab] : 0;
      const uint32_t* src = active[i] ? g->d_inv_list + g->h_inv_index[lab] : g->d_inv_list;
      cand_a[gc] = (uint32_t*)A.alloc(4ull * std::max<uint32_t>(len, 1));
      if (!cand_a[gc]) return set_err(c, FSIG_ENOMEM, "batch: arena exhausted for candidates");
      comp[gc] = BComp{src, cand_a[gc], masks + (uint64_t)i * n, len, a, 0, tiles};
      tiles += (len + kTile - 1) / kTile;
    }
  }
  uint32_t* tilecnt = (uint32_t*)A.alloc(4ull * std::max<uint64_t>(tiles, 1));
  uint64_t* dsizes = (uint64_t*)A.alloc(8ull * std::max<uint32_t>(ncls, 1));
  if (!tilecnt || !dsizes) return set_err(c, FSIG_ENOMEM, "batch: arena exhausted");
  auto run_compact = [&](const std::vector<BComp>& tab, uint64_t ntiles, int pred, const uint8_t* keep) -> int {
    // segments without tiles are kept in the table (offsets kernel needs one
    // block per class) but the tile search only sees the ones with tiles
    const BComp* d = upload_table(c, tab);
    if (!d) return set_err(c, FSIG_ENOMEM, "batch: arena exhausted");
    std::vector<BComp> nz;
    for (auto& s : tab)
      if (s.len) nz.push_back(s);
    const BComp* dnz = upload_table(c, nz);
    if (!dnz) return set_err(c, FSIG_ENOMEM, "batch: arena exhausted");
    if (ntiles) {
      if (pred == 0) k_bcompact_count<0><<<(uint32_t)ntiles, FSIG_BLOCK, 0, c->stream>>>(dnz, (uint32_t)nz.size(), keep, tilecnt);
      else k_bcompact_count<1><<<(uint32_t)ntiles, FSIG_BLOCK, 0, c->stream>>>(dnz, (uint32_t)nz.size(), keep, tilecnt);
    }
    k_bcompact_offsets<<<(uint32_t)tab.size(), FSIG_BLOCK, 0, c->stream>>>(d, tilecnt, dsizes);
    if (ntiles) {
      if (pred == 0) k_bcompact_scatter<0><<<(uint32_t)ntiles, FSIG_BLOCK, 0, c->stream>>>(dnz, (uint32_t)nz.size(), keep, tilecnt);
      else k_bcompact_scatter<1><<<(uint32_t)ntiles, FSIG_BLOCK, 0, c->stream>>>(dnz, (uint32_t)nz.size(), keep, tilecnt);
    }
    FSIG_HIP(c, hipGetLastError());
    std::vector<uint64_t> hs(tab.size());
    for (size_t i = 0; i < tab.size(); i += kPinnedWords)
      if ((rc = read_u64(c, dsizes + i, hs.data() + i, std::min(kPinnedWords, tab.size() - i)))) return rc;
    for (size_t i = 0; i < tab.size(); i++) size[i] = (uint32_t)hs[i];
    return FSIG_OK;
  };
  clk.mark("plans + LDF");
  if (ncls && (rc = run_compact(comp, tiles, 0, nullptr))) return rc;
  tf.done(4ull * n * nq + 16ull * flat, flat);
  clk.mark("LDF compaction");

  // ---- refinement passes for every query
  std::vector<uint8_t> thr((size_t)std::max<uint32_t>(ncls, 1) * MAXQ, 0);
  std::vector<uint32_t> tmask(ncls, 0);
  bool any_target = false;
  for (uint32_t i = 0; i < nq; i++)
    for (uint32_t a = 0; a < plans[i].nclass; a++)
      for (uint32_t b = 0; b < plans[i].nclass; b++)
        if (plans[i].thr[a][b]) {
          thr[(size_t)(cls_base[i] + a) * MAXQ + b] = (uint8_t)plans[i].thr[a][b];
          tmask[cls_base[i] + a] |= 1u << b;
          any_target = true;
        }
  const uint8_t* dthr = upload_table(c, thr);
  if (!dthr) return set_err(c, FSIG_ENOMEM, "batch: arena exhausted");
  uint64_t cand_total = 0;
  for (uint32_t gc = 0; gc < ncls; gc++) {
    cand_b[gc] = (uint32_t*)A.alloc(4ull * std::max<uint32_t>(size[gc], 1));
    if (!cand_b[gc]) return set_err(c, FSIG_ENOMEM, "batch: arena exhausted for candidates");
    cand_total += size[gc];
  }
  uint8_t* keep = (uint8_t*)A.alloc(std::max<uint64_t>(cand_total, 1));
  if (!keep) return set_err(c, FSIG_ENOMEM, "batch: arena exhausted");
  std::vector<uint32_t*>* cur = &cand_a;
  std::vector<uint32_t*>* nxt = &cand_b;
  // a query is dead once it has an empty class (its count is 0)
  auto kill_empty = [&]() {
    for (uint32_t i = 0; i < nq; i++)
      for (uint32_t a = 0; a < plans[i].nclass && active[i]; a++)
        if (size[cls_base[i] + a] == 0) active[i] = 0;
  };
  kill_empty();
  const double avg_deg = n ? (double)g->nnz / n : 0;
  const uint32_t RG = avg_deg <= 12 ? 8 : (avg_deg <= 40 ? 16 : 64);
  const uint32_t bpasses = refine_passes(c);
  for (uint32_t pass = 0; any_target && pass < bpasses; pass++) {
    KTimer tr(c, FSIG_K_REFINE);
    std::vector<BClass> cl;
    std::vector<BComp> rc_tab(ncls);
    uint64_t f = 0, rt = 0;
    for (uint32_t i = 0; i < nq; i++)
      for (uint32_t a = 0; a < plans[i].nclass; a++) {
        const uint32_t gc = cls_base[i] + a;
        const uint32_t sz = active[i] ? size[gc] : 0;
        if (sz) cl.push_back(BClass{(*cur)[gc], masks + (uint64_t)i * n, sz, a, tmask[gc], gc, f});
        rc_tab[gc] = BComp{(*cur)[gc], (*nxt)[gc], masks + (uint64_t)i * n, sz, a, f, rt};
        f += sz;
        rt += (sz + kTile - 1) / kTile;
      }
    if (f == 0) break;
    const BClass* dcl = upload_table(c, cl);
    if (!dcl) return set_err(c, FSIG_ENOMEM, "batch: arena exhausted");
    switch (RG) {
      case 8: k_brefine<8><<<bgrid(c, f, FSIG_BLOCK / 8), FSIG_BLOCK, 0, c->stream>>>(dcl, (uint32_t)cl.size(), f, dthr, g->d_off, g->d_adj, keep); break;
      case 16: k_brefine<16><<<bgrid(c, f, FSIG_BLOCK / 16), FSIG_BLOCK, 0, c->stream>>>(dcl, (uint32_t)cl.size(), f, dthr, g->d_off, g->d_adj, keep); break;
      default: k_brefine<64><<<bgrid(c, f, FSIG_BLOCK / 64), FSIG_BLOCK, 0, c->stream>>>(dcl, (uint32_t)cl.size(), f, dthr, g->d_off, g->d_adj, keep); break;
    }
    k_brefine_apply<<<bgrid(c, f, FSIG_BLOCK), FSIG_BLOCK, 0, c->stream>>>(dcl, (uint32_t)cl.size(), f, keep);
    const std::vector<uint32_t> before = size;
    if ((rc = run_compact(rc_tab, rt, 1, keep))) return rc;
    bool changed = false;
    for (uint32_t gc = 0; gc < ncls; gc++) changed |= size[gc] != before[gc];
    std::swap(cur, nxt);
    kill_empty();
    tr.done(f * 18, f);
    if (!changed) break;
  }
  const std::vector<uint32_t*>& cand = *cur;
  clk.mark("refinement");

  // ---- orders and the class pairs each query's join needs
  std::vector<BPair> pairs;
  std::vector<std::vector<int>> pair_of(nq);  // [i][a*MAXQ+b] -> pair index
  std::vector<uint64_t> pair_row_lo(nq + 1, 0);  // query i's pair rows start here
  uint64_t pflat = 0;
  for (uint32_t i = 0; i < nq; i++) {
    pair_row_lo[i] = pflat;
    if (!active[i]) continue;
    QueryPlan& P = plans[i];
    plan_order_mmatch(P, &size[cls_base[i]]);
    pair_of[i].assign(MAXQ * MAXQ, -1);
    for (uint32_t p = 1; p < P.k; p++) {
      const uint32_t w = P.order[p];
      for (uint32_t cp = 0; cp < p; cp++) {
        const uint32_t u = P.order[cp];
        if (!((P.adjm[w] >> u) & 1u)) continue;
        const uint32_t a = P.cls[u], b = P.cls[w];
        if (pair_of[i][a * MAXQ + b] >= 0) continue;
        pair_of[i][a * MAXQ + b] = (int)pairs.size();
        const uint32_t ga = cls_base[i] + a, gb = cls_base[i] + b;
        BPair pd{cand[ga], cand[gb], masks + (uint64_t)i * n, nullptr, nullptr, nullptr, size[ga], size[gb], b, 0, pflat};
        pd.off_out = (uint32_t*)A.alloc(4ull * (size[ga] + 1));
        if (!pd.off_out) return set_err(c, FSIG_ENOMEM, "batch: arena exhausted for candidate edges");
        pflat += size[ga];
        pairs.push_back(pd);
      }
    }
  }
  pair_row_lo[nq] = pflat;
  const uint32_t np = (uint32_t)pairs.size();
  std::vector<uint64_t> pbase_h(np), maxlen_h(np), ptotal_h(np);
  uint32_t* lst_all = nullptr;
  if (np) {
    KTimer tc(c, FSIG_K_CEDGE);
    // Candidate bitmaps of the target classes, n bits each: a bitmap stays
    // in L2 while one pair's entries are tested, a query's 4n-byte class
    // mask does not on large graphs.  All of a batch's bitmaps are built at
    // once while they fit the last-level cache (kBitmapCache); beyond that
    // the build runs query by query, each query's bitmaps rebuilt in one
    // reused region right before its rows and segments are swept (bitmaps
    // built long before their sweep are read back from HBM).  No room in the
    // arena (a quarter of what is free): the class masks.
    const uint64_t nwords = (n + 31ull) / 32;
    std::vector<uint32_t*> tbits(std::max<uint32_t>(ncls, 1), nullptr);
    std::vector<std::vector<uint32_t>> qtgt(nq);
    size_t ntgt = 0, max_t = 0;
    {
      std::vector<uint8_t> seen(std::max<uint32_t>(ncls, 1), 0);
      for (uint32_t i = 0; i < nq; i++) {
        for (uint32_t x = 0; x < MAXQ * MAXQ && active[i]; x++)
          if (pair_of[i][x] >= 0) {
            const uint32_t gb = cls_base[i] + x % MAXQ;
            if (!seen[gb]) { seen[gb] = 1; qtgt[i].push_back(gb); }
          }
        ntgt += qtgt[i].size();
        max_t = std::max(max_t, qtgt[i].size());
      }
    }
    constexpr uint64_t kBitmapCache = 192ull << 20;
    const bool per_query = 4ull * nwords * ntgt > kBitmapCache;
    const uint64_t region_words = nwords * (per_query ? max_t : ntgt);
    const bool use_bits = ntgt && 4ull * region_words <= A.avail() / 4;
    if (clk.on)
      fprintf(stderr, "[fsig batch] %zu target classes: %s\n", ntgt,
              !use_bits ? "class masks (no room for bitmaps)"
                        : per_query ? "candidate bitmaps, query by query" : "candidate bitmaps, all at once");
    uint32_t* region = nullptr;
    // Vertex -> candidate index maps of the target classes (n words each),
    // read by the fills for the entries they keep.  Query by query: one
    // reused region holds a query's maps, scattered with its bitmaps right
    // before its fill, when the largest query's fit a quarter of the free
    // arena.  Otherwise every target class has its own map while a quarter
    // of the free arena holds them, and the classes beyond search their
    // candidate list instead.
    uint32_t* posregion = nullptr;
    std::vector<uint32_t*> tpos(std::max<uint32_t>(ncls, 1), nullptr);
    // chunks of the build: every query alone (per_query) or the whole batch;
    // chunk j: bitmap table entries [bt_lo[j], bt_lo[j+1]) (flats local to
    // the chunk), source rows [src_lo[j], src_hi[j]), pair rows [row_lo[j],
    // row_hi[j])
    std::vector<BBits> bt;
    std::vector<uint32_t> chunk_q;
    std::vector<size_t> bt_lo;
    std::vector<uint64_t> bt_total, chunk_words;
    if (use_bits) {
      region = (uint32_t*)A.alloc(4ull * region_words);
      if (!region) return set_err(c, FSIG_ENOMEM, "batch: arena exhausted for candidate bitmaps");
      if (per_query && 4ull * n * max_t <= A.avail() / 4) posregion = (uint32_t*)A.alloc(4ull * n * max_t);
      size_t slot = 0;
      uint64_t bflat = 0;
      for (uint32_t i = 0; i < nq; i++) {
        if (qtgt[i].empty()) continue;
        if (per_query || chunk_q.empty()) {  // a chunk starts
          chunk_q.push_back(i);
          bt_lo.push_back(bt.size());
          bflat = 0;
          slot = 0;
        }
        for (uint32_t gb : qtgt[i]) {
          tbits[gb] = region + slot * nwords;
          if (posregion) tpos[gb] = posregion + slot * (uint64_t)n;
          slot++;
          if (size[gb]) {
            bt.push_back(BBits{cand[gb], tbits[gb], tpos[gb], size[gb], 0, bflat});
            bflat += size[gb];
          }
        }
        if (per_query) {
          bt_total.push_back(bflat);
          chunk_words.push_back(slot * nwords);
        }
      }
      if (!per_query && !chunk_q.empty()) {
        bt_total.push_back(bflat);
        chunk_words.push_back(slot * nwords);
      }
      bt_lo.push_back(bt.size());
      for (uint32_t i = 0; i < nq; i++)
        for (uint32_t x = 0; x < MAXQ * MAXQ && active[i]; x++)
          if (pair_of[i][x] >= 0) {
            pairs[pair_of[i][x]].bits = tbits[cls_base[i] + x % MAXQ];
            pairs[pair_of[i][x]].bpos = tpos[cls_base[i] + x % MAXQ];
          }
    }
    if (!posregion) {
      std::map<const uint32_t*, uint32_t*> made;
      const uint64_t budget = A.avail() / 4;
      uint64_t used = 0;
      for (auto& pd : pairs) {
        auto it = made.find(pd.bcand);
        if (it == made.end()) {
          uint32_t* m = nullptr;
          if (used + 4ull * n <= budget && (m = (uint32_t*)A.alloc(4ull * n))) {
            used += 4ull * n;
            k_bscatter_pos<<<bgrid(c, pd.bsize, FSIG_BLOCK), FSIG_BLOCK, 0, c->stream>>>(pd.bcand, pd.bsize, m);
          }
          it = made.emplace(pd.bcand, m).first;
        }
        pd.bpos = it->second;
      }
    }
    // sources: every (query, class a) with its targets, one sweep each
    std::vector<BSrc> srcs;
    std::vector<uint64_t> qsrc_lo(nq + 1, 0), qrow_lo(nq + 1, 0);
    uint64_t sflat = 0;
    for (uint32_t i = 0; i < nq; i++) {
      qsrc_lo[i] = sflat;
      if (!active[i]) continue;
      for (uint32_t a = 0; a < plans[i].nclass; a++) {
        BSrc sc;
        std::memset(&sc, 0, sizeof(sc));
        for (uint32_t b = 0; b < plans[i].nclass; b++) {
          const int pi = pair_of[i][a * MAXQ + b];
          if (pi < 0) continue;
          sc.tmask |= 1u << b;
          sc.pidx[b] = (uint32_t)pi;
          sc.bits[b] = tbits[cls_base[i] + b];
        }
        if (!sc.tmask) continue;
        const uint32_t ga = cls_base[i] + a;
        sc.acand = cand[ga];
        sc.mask = use_bits ? nullptr : masks + (uint64_t)i * n;
        sc.asize = size[ga];
        sc.flat = sflat;
        sflat += size[ga];
        if (size[ga]) srcs.push_back(sc);
      }
    }
    qsrc_lo[nq] = sflat;
    for (uint32_t i = 0; i <= nq; i++) qrow_lo[i] = pair_row_lo[i];
    const BPair* dP = upload_table(c, pairs);
    const BSrc* dS = upload_table(c, srcs);
    const BBits* dbt = upload_table(c, bt);
    const uint32_t nsrc = (uint32_t)srcs.size();
    uint32_t* cnt = (uint32_t*)A.alloc(4ull * std::max<uint64_t>(pflat, 1));
    uint64_t* gscan = (uint64_t*)A.alloc(8ull * (pflat + 1));
    unsigned int* dmax = (unsigned int*)A.alloc(4ull * np);
    uint64_t* pb = (uint64_t*)A.alloc(8ull * np);
    if (!dP || !dS || (!bt.empty() && !dbt) || !cnt || !gscan || !dmax || !pb)
      return set_err(c, FSIG_ENOMEM, "batch: arena exhausted");
    SegRows segs;  // hub rows of every pair, one wavefront per segment
    if ((rc = seg_prepare(c, dP, np, pflat, g->d_off, cnt, segs))) return rc;
    // the chunks' source, pair-row and segment ranges
    const size_t nch = use_bits ? chunk_q.size() : 1;
    std::vector<uint64_t> ch_src(nch + 1), ch_row(nch + 1), ch_seg(nch + 1, 0);
    for (size_t j = 0; j < nch; j++) {
      const uint32_t i = use_bits ? chunk_q[j] : 0;
      ch_src[j] = use_bits ? qsrc_lo[i] : 0;
      ch_row[j] = use_bits ? qrow_lo[i] : 0;
    }
    ch_src[nch] = sflat;
    ch_row[nch] = pflat;
    if (!per_query || !use_bits) {
      ch_src[0] = 0;
      ch_row[0] = 0;
    }
    if (segs.nsegs) {
      uint64_t* didx = (uint64_t*)upload_table(c, ch_row);
      uint64_t* dseg = (uint64_t*)A.alloc(8ull * (nch + 1));
      if (!didx || !dseg) return set_err(c, FSIG_ENOMEM, "batch: arena exhausted");
      k_gather_at<<<(uint32_t)((nch + 1 + 255) / 256), 256, 0, c->stream>>>(segs.segbase, didx, (uint32_t)(nch + 1), dseg);
      FSIG_HIP(c, hipGetLastError());
      if ((rc = read_u64(c, dseg, ch_seg.data(), nch + 1))) return rc;
    }
    auto build_bits = [&](size_t j, bool with_pos) -> int {
      if (!use_bits) return FSIG_OK;
      k_fill_u32<<<bgrid(c, chunk_words[j], FSIG_BLOCK), FSIG_BLOCK, 0, c->stream>>>(region, chunk_words[j], 0);
      const size_t nt = bt_lo[j + 1] - bt_lo[j];
      if (nt && bt_total[j])
        k_bset_bits<<<bgrid(c, bt_total[j], FSIG_BLOCK), FSIG_BLOCK, 0, c->stream>>>(dbt + bt_lo[j], (uint32_t)nt,
                                                                                     bt_total[j], with_pos);
      FSIG_HIP(c, hipGetLastError());
      return FSIG_OK;
    };
    for (size_t j = 0; j < nch; j++) {
      if ((rc = build_bits(j, false))) return rc;
      const uint64_t lo = ch_src[j], hi = ch_src[j + 1];
      if (hi > lo) {
        switch (RG) {
          case 8: k_bcedge_count_multi<8><<<bgrid(c, hi - lo, FSIG_BLOCK / 8), FSIG_BLOCK, 0, c->stream>>>(dS, nsrc, hi, dP, g->d_off, g->d_adj, cnt, lo); break;
          case 16: k_bcedge_count_multi<16><<<bgrid(c, hi - lo, FSIG_BLOCK / 16), FSIG_BLOCK, 0, c->stream>>>(dS, nsrc, hi, dP, g->d_off, g->d_adj, cnt, lo); break;
          default: k_bcedge_count_multi<64><<<bgrid(c, hi - lo, FSIG_BLOCK / 64), FSIG_BLOCK, 0, c->stream>>>(dS, nsrc, hi, dP, g->d_off, g->d_adj, cnt, lo); break;
        }
      }
      if ((rc = seg_count(c, dP, np, pflat, g->d_off, g->d_adj, cnt, segs, ch_seg[j], ch_seg[j + 1]))) return rc;
    }
    FSIG_HIP(c, hipGetLastError());
    if ((rc = scan_u32_to_u64(c, cnt, gscan, pflat))) return rc;
    uint64_t total_edges;
    if ((rc = read_u64(c, gscan + pflat, &total_edges, 1))) return rc;
    lst_all = (uint32_t*)A.alloc(4ull * std::max<uint64_t>(total_edges, 1));
    if (!lst_all) return set_err(c, FSIG_ENOMEM, "batch: arena exhausted for candidate lists");
    for (size_t j = 0; j < nch; j++) {
      // one chunk: its bitmaps are still built (and its maps, if any, are the
      // classes' own)
      if (nch > 1 && (rc = build_bits(j, posregion != nullptr))) return rc;
      const uint64_t lo = ch_src[j], hi = ch_src[j + 1];
      if (hi > lo) {
        switch (RG) {
          case 8: k_bcedge_fill_multi<8><<<bgrid(c, hi - lo, FSIG_BLOCK / 8), FSIG_BLOCK, 0, c->stream>>>(dS, nsrc, hi, dP, g->d_off, g->d_adj, gscan, lst_all, lo); break;
          case 16: k_bcedge_fill_multi<16><<<bgrid(c, hi - lo, FSIG_BLOCK / 16), FSIG_BLOCK, 0, c->stream>>>(dS, nsrc, hi, dP, g->d_off, g->d_adj, gscan, lst_all, lo); break;
          default: k_bcedge_fill_multi<64><<<bgrid(c, hi - lo, FSIG_BLOCK / 64), FSIG_BLOCK, 0, c->stream>>>(dS, nsrc, hi, dP, g->d_off, g->d_adj, gscan, lst_all, lo); break;
        }
      }
      if ((rc = seg_fill(c, dP, np, pflat, g->d_off, g->d_adj, gscan, lst_all, segs, ch_seg[j], ch_seg[j + 1]))) return rc;
    }
    k_bpair_maxlen<<<np, FSIG_BLOCK, 0, c->stream>>>(dP, gscan, dmax);
    k_bpair_base<<<(np + 255) / 256, 256, 0, c->stream>>>(gscan, dP, np, pb);
    FSIG_HIP(c, hipGetLastError());
    for (uint32_t i = 0; i < np; i += kPinnedWords)
      if ((rc = read_u64(c, pb + i, pbase_h.data() + i, std::min<size_t>(kPinnedWords, np - i)))) return rc;
    std::vector<unsigned int> mx(np);
    FSIG_HIP(c, hipMemcpyAsync(mx.data(), dmax, 4ull * np, hipMemcpyDeviceToHost, c->stream));
    FSIG_HIP(c, hipStreamSynchronize(c->stream));
    for (uint32_t i = 0; i < np; i++) maxlen_h[i] = mx[i];
    for (uint32_t i = 0; i < np; i++) ptotal_h[i] = (i + 1 < np ? pbase_h[i + 1] : total_edges) - pbase_h[i];
    // bytes as the per-query build's, with sources' rows read once for all
    // their targets
    tc.done(28ull * pflat + 16ull * segs.entries + 8ull * total_edges + 4ull * (pflat + np), segs.entries);
  }

  clk.mark("candidate edges");
  // ---- level tables, stacks and tasks
  std::vector<BQuery> bq(nq);
  std::vector<BLevel> lv;
  std::vector<Task> tasks;
  uint32_t stack_words = 1;
  for (uint32_t i = 0; i < nq; i++) {
    bq[i].k = ks[i];
    bq[i].level_base = (uint32_t)lv.size();
    bq[i].edge_tasks = 0;
    bq[i].n1 = 0;
    bq[i].off1 = nullptr;
    bq[i].lst1 = nullptr;
    if (!active[i]) continue;
    const QueryPlan& P = plans[i];
    uint32_t soff = 0;
    for (uint32_t p = 0; p < P.k; p++) {
      BLevel L;
      std::memset(&L, 0, sizeof(L));
      const uint32_t w = P.order[p];
      const uint32_t gw = cls_base[i] + P.cls[w];
      L.wcand = cand[gw];
      L.wsize = size[gw];
      uint64_t cap_p = p == 0 ? 0 : size[gw];
      for (uint32_t cp = 0; p > 0 && cp < p; cp++) {
        const uint32_t u = P.order[cp];
        if ((P.adjm[w] >> u) & 1u) {
          const int pi = pair_of[i][P.cls[u] * MAXQ + P.cls[w]];
          L.bcol[L.nb] = cp;
          L.boff[L.nb] = pairs[pi].off_out;
          L.blst[L.nb] = lst_all + pbase_h[pi];
          cap_p = std::min<uint64_t>(cap_p, maxlen_h[pi]);
          L.nb++;
        } else if (P.label[u] == P.label[w]) {
          const uint32_t j = L.ncf++;
          L.cfcol[j] = cp;
          if (P.cls[u] == P.cls[w]) L.cfsame |= 1u << j;
          else { L.cfcand[j] = cand[cls_base[i] + P.cls[u]]; L.any_cross = 1; }
        }
      }
      L.stack_off = soff;
      // levels 1..k-2 keep their candidate list on the stack
      if (p >= 1 && p + 1 < P.k) soff += (uint32_t)std::max<uint64_t>(cap_p, 1);
      lv.push_back(L);
    }
    stack_words = std::max(stack_words, soff);
    // tasks: ranges of (root, first neighbour) list entries when position 1
    // hangs off position 0 alone, else ranges of roots; smaller ranges for
    // larger queries
    const uint32_t n1 = size[cls_base[i] + P.cls[P.order[0]]];
    const uint32_t chunk = P.k <= 4 ? 64 : (P.k <= 8 ? 16 : 4);
    const BLevel& l1 = lv[bq[i].level_base + 1 < lv.size() ? bq[i].level_base + 1 : bq[i].level_base];
    if (P.k >= 2 && l1.nb == 1 && l1.bcol[0] == 0 && l1.ncf == 0) {
      const int pi = pair_of[i][P.cls[P.order[0]] * MAXQ + P.cls[P.order[1]]];
      const uint64_t e1 = ptotal_h[pi];
      if (e1 >= 0xffffffffull) return set_err(c, FSIG_EUNSUPPORTED, "batch: first list too long");
      bq[i].edge_tasks = 1;
      bq[i].n1 = n1;
      bq[i].off1 = pairs[pi].off_out;
      bq[i].lst1 = lst_all + pbase_h[pi];
      for (uint64_t r = 0; r < e1; r += chunk) tasks.push_back(Task{i, (uint32_t)r, (uint32_t)std::min<uint64_t>(e1, r + chunk)});
    } else {
      for (uint32_t r = 0; r < n1; r += chunk) tasks.push_back(Task{i, r, std::min(n1, r + chunk)});
    }
  }
  if (tasks.empty()) {
    for (uint32_t i = 0; i < nq; i++) counts_out[i] = 0;
    if (join_ms) *join_ms = 0;
    return FSIG_OK;
  }
  // interleave tasks of different queries so a burst of heavy roots spreads
  // over the machine
  {
    std::vector<Task> mixed;
    mixed.reserve(tasks.size());
    std::vector<std::vector<Task>> perq(nq);
    for (auto& t : tasks) perq[t.q].push_back(t);
    size_t left = tasks.size();
    std::vector<size_t> at(nq, 0);
    while (left) {
      for (uint32_t i = 0; i < nq; i++)
        if (at[i] < perq[i].size()) { mixed.push_back(perq[i][at[i]++]); left--; }
    }
    tasks.swap(mixed);
  }
  const BQuery* dQ = upload_table(c, bq);
  const BLevel* dL = upload_table(c, lv);
  const Task* dT = upload_table(c, tasks);
  unsigned long long* dcounts = (unsigned long long*)A.alloc(8ull * nq);
  Queue* qu = (Queue*)A.alloc(sizeof(Queue));
  const uint32_t dcap = 1u << 20;
  DTask* dtasks = (DTask*)A.alloc(sizeof(DTask) * (size_t)dcap);
  unsigned int* ready = (unsigned int*)A.alloc(4ull * dcap);
  if (!dQ || !dL || !dT || !dcounts || !qu || !dtasks || !ready)
    return set_err(c, FSIG_ENOMEM, "batch: arena exhausted");

  clk.mark("level tables + tasks");
  // ---- join engines
  // Lockstep batched BFS over a set of live queries; counts land in dcounts.
  auto fill_level = [&](uint32_t i, uint32_t p, JoinLevel& L, CandEdgeSrc& S) {
    const QueryPlan& P = plans[i];
    const uint32_t w = P.order[p];
    L.p = p;
    L.wcand = cand[cls_base[i] + P.cls[w]];
    L.wsize = size[cls_base[i] + P.cls[w]];
    for (uint32_t cp = 0; cp < p; cp++) {
      const uint32_t u = P.order[cp];
      if ((P.adjm[w] >> u) & 1u) {
        const int pi = pair_of[i][P.cls[u] * MAXQ + P.cls[w]];
        L.bcol[L.nb] = cp;
        S.off[L.nb] = pairs[pi].off_out;
        S.lst[L.nb] = lst_all + pbase_h[pi];
        L.nb++;
      } else if (P.label[u] == P.label[w]) {
        const uint32_t j = L.ncf++;
        L.cfcol[j] = cp;
        if (P.cls[u] == P.cls[w]) L.cfsame |= 1u << j;
        else { L.cfcand[j] = cand[cls_base[i] + P.cls[u]]; L.any_cross = 1; }
      }
    }
  };
  // Level p of query i; with `tail` the query's last two levels (p + 2 == k)
  // are counted in one pass when the last vertex has a backward neighbour.
  auto make_level = [&](uint32_t i, uint32_t p, bool tail) {
    MLevel M;
    std::memset(&M, 0, sizeof(M));
    M.qid = i;
    const QueryPlan& P = plans[i];
    if (p >= P.k) { M.last = 1; return M; }
    fill_level(i, p, M.L, M.S);
    M.last = p + 1 == P.k;
    if (tail && p + 2 == P.k) {
      fill_level(i, p + 1, M.L2, M.S2);
      M.tail = tail_applies(M.L2) ? 1 : 0;
    }
    return M;
  };
  unsigned long long* stats_ls = family_counters(c, FSIG_K_BATCH);
  // Lanes of the single-query driver for query i: from the mean length of
  // its candidate-edge lists, as fsig_match picks them.
  auto lanes_for = [&](uint32_t i) {
    uint64_t edges = 0, rows = 0;
    for (uint32_t x = 0; x < MAXQ * MAXQ; x++) {
      const int pi = pair_of[i].empty() ? -1 : pair_of[i][x];
      if (pi < 0) continue;
      edges += ptotal_h[pi];
      rows += pairs[pi].asize;
    }
    return pick_lanes(c, rows ? (double)edges / (double)rows : 1.0);
  };
  // Bounded queries at a lockstep level, after its count pass and scan.  A
  // query whose count already reached its bound is settled at the bound (its
  // rows get no children).  A query whose next level would exceed the beam
  // keeps only its first parent rows, those whose children start within the
  // beam (the later rows get none): from then on its count is a lower bound
  // of its count, which settles it once it reaches the bound.  The cut rows
  // (a suffix of the query's rows at that level) are recorded; a beamed
  // query that ends below its bound has the subtrees of its cut rows counted
  // afterwards by the bounded depth-first driver, from those rows, while the
  // level tables are still held.  The level is then rescanned.
  const uint64_t beam = 1ull << (c->tune.beam_log2 ? c->tune.beam_log2 : kBeamLog2);
  struct CutRows {
    uint32_t p;            // level of the rows (p words each)
    const uint32_t* rows;  // in that level's table
    uint64_t n;
  };
  std::vector<std::vector<CutRows>> cut_rows(nq);
  auto settle_bounded = [&](uint32_t p, const uint32_t* rows, uint32_t* cnt, uint64_t* off, uint64_t R,
                            const uint64_t* dqb, uint64_t* dqb2, const std::vector<uint64_t>& qb,
                            const std::vector<uint32_t>& live, uint64_t& T) -> int {
    const uint32_t nl = (uint32_t)live.size();
    k_gather_bases<<<(nl + 256) / 256, 256, 0, c->stream>>>(off, dqb, nl, dqb2);
    FSIG_HIP(c, hipGetLastError());
    std::vector<uint64_t> nb(nl + 1), hc(nq);
    int rc2;
    for (uint32_t j = 0; j <= nl; j += kPinnedWords)
      if ((rc2 = read_u64(c, dqb2 + j, nb.data() + j, std::min<size_t>(kPinnedWords, nl + 1 - j)))) return rc2;
    for (uint32_t i = 0; i < nq; i += kPinnedWords)
      if ((rc2 = read_u64(c, dcounts + i, hc.data() + i, std::min<size_t>(kPinnedWords, nq - i)))) return rc2;
    // per live query: rows whose first child index is at or past cut[j] get
    // no children
    std::vector<uint64_t> cut(nl, ~0ull);
    bool changed = false;
    for (uint32_t j = 0; j < nl; j++) {
      const uint32_t i = live[j];
      if (!bound[i] || settled[i] || qb[j + 1] == qb[j]) continue;
      if (hc[i] >= bound[i]) {
        settled[i] = 1;
        settled_count[i] = bound[i];
        cut[j] = 0;
        changed = true;
      } else if (nb[j + 1] - nb[j] > beam) {
        beamed[i] = 1;
        cut[j] = nb[j] + beam;
        changed = true;
      }
    }
    if (!changed) return FSIG_OK;
    const size_t mk2 = A.mark();
    std::vector<uint64_t> first(qb.begin() + 1, qb.end());  // nothing cut: the query's end
    const uint64_t* dcut = upload_table(c, cut);
    uint64_t* dfirst = (uint64_t*)upload_table(c, first);
    if (!dcut || !dfirst) return set_err(c, FSIG_ENOMEM, "batch: arena exhausted");
    k_beam_cut<<<bgrid(c, R, FSIG_BLOCK), FSIG_BLOCK, 0, c->stream>>>(off, dqb, nl, dcut, R, cnt, dfirst);
    FSIG_HIP(c, hipGetLastError());
    for (uint32_t j = 0; j < nl; j += kPinnedWords)
      if ((rc2 = read_u64(c, dfirst + j, first.data() + j, std::min<size_t>(kPinnedWords, nl - j)))) return rc2;
    A.release(mk2);  // both tables are read before any later allocation is written
    for (uint32_t j = 0; j < nl; j++)
      if (cut[j] != ~0ull && cut[j] != 0 && first[j] < qb[j + 1])
        cut_rows[live[j]].push_back(CutRows{p, rows + first[j] * p, qb[j + 1] - first[j]});
    if ((rc2 = scan_u32_to_u64(c, cnt, off, R))) return rc2;
    return read_u64(c, off + R, &T, 1);
  };
  // The subtrees of the cut rows of every beamed query that ended below its
  // bound, counted by the bounded depth-first driver (bound: what the beam
  // left); settles those queries.
  auto resume_beamed = [&]() -> int {
    std::vector<uint64_t> hc(nq);
    int rc2;
    for (uint32_t i = 0; i < nq; i += kPinnedWords)
      if ((rc2 = read_u64(c, dcounts + i, hc.data() + i, std::min<size_t>(kPinnedWords, nq - i)))) return rc2;
    for (uint32_t i = 0; i < nq; i++) {
      if (!beamed[i] || settled[i] || hc[i] >= bound[i]) continue;
      BatchProvider pv;
      pv.fill = [&fill_level, i](uint32_t pp, JoinLevel& L, CandEdgeSrc& S) { fill_level(i, pp, L, S); };
      JoinDriver<BatchProvider> drv(c, pv, ks[i], lanes_for(i), FSIG_MODE_COUNT, bound[i] - hc[i], nullptr);
      rc2 = FSIG_OK;
      for (const CutRows& cr : cut_rows[i]) {
        if (drv.total() >= bound[i] - hc[i]) break;
        if ((rc2 = drv.run_at(cr.p, cr.rows, cr.n))) break;
      }
      if (rc2 && rc2 != FSIG_ELIMIT) return rc2;
      settled[i] = 1;
      settled_count[i] = rc2 ? 0 : hc[i] + drv.total();
      settled_rc[i] = rc2;  // over the context's work limit: refused
      if (clk.on) {
        fprintf(stderr, "[fsig batch]   resumed query %u (k %u): beam count %llu, %zu cut levels, status %d\n", i,
                ks[i], (unsigned long long)hc[i], cut_rows[i].size(), rc2);
        clk.mark("  resume");
      }
    }
    return FSIG_OK;
  };
  // Lockstep batched BFS.  G = 0: lanes by role at each level: 64 for the
  // queries whose last two levels are counted here (long shared
  // intersections, a 1024-word LDS tile), 16 for the others.  The live
  // queries are ordered by size, so each role owns a contiguous row range of
  // every level (one launch per role).
  auto lockstep = [&](uint32_t G, const std::vector<uint32_t>& qs, float& ms) -> int {
    const size_t mk = A.mark();
    std::vector<uint32_t> live;
    uint32_t maxk = 0;
    for (uint32_t i : qs)
      if (active[i] && ks[i] >= 2) { live.push_back(i); maxk = std::max(maxk, ks[i]); }
    std::stable_sort(live.begin(), live.end(), [&](uint32_t x, uint32_t y) { return ks[x] < ks[y]; });
    const uint32_t nl = (uint32_t)live.size();
    auto lanes_of = [&](uint32_t i, uint32_t p) { return G ? G : (ks[i] == p + 2 ? 64u : 16u); };
    // lane groups of level p: [gstart[j], gstart[j+1]) of live with glanes[j]
    std::vector<uint32_t> gstart, glanes;
    auto make_groups = [&](uint32_t p) {
      gstart.clear();
      glanes.clear();
      for (uint32_t j = 0; j < nl; j++)
        if (j == 0 || lanes_of(live[j], p) != lanes_of(live[j - 1], p)) {
          gstart.push_back(j);
          glanes.push_back(lanes_of(live[j], p));
        }
      gstart.push_back(nl);
    };
    FSIG_HIP(c, hipEventRecord(c->ev_a, c->stream));
    if (nl) {
      std::vector<uint64_t> qb(nl + 1, 0);
      for (uint32_t j = 0; j < nl; j++)
        qb[j + 1] = qb[j] + size[cls_base[live[j]] + plans[live[j]].cls[plans[live[j]].order[0]]];
      uint64_t* dqb = (uint64_t*)A.alloc(8ull * (nl + 1));
      uint64_t* dqb2 = (uint64_t*)A.alloc(8ull * (nl + 1));
      MLevel* dml = (MLevel*)A.alloc(sizeof(MLevel) * nl);
      uint32_t* rows = (uint32_t*)A.alloc(4ull * std::max<uint64_t>(qb[nl], 1));
      if (!dqb || !dqb2 || !dml || !rows) { A.release(mk); return set_err(c, FSIG_ENOMEM, "lockstep: arena exhausted"); }
      FSIG_HIP(c, hipMemcpyAsync(dqb, qb.data(), 8ull * (nl + 1), hipMemcpyHostToDevice, c->stream));
      uint64_t R = qb[nl];
      k_iota_multi<<<bgrid(c, R, FSIG_BLOCK), FSIG_BLOCK, 0, c->stream>>>(dqb, nl, rows, R);
      std::vector<MLevel> ml(nl);
      uint64_t prev_bytes = 0;  // debug timing: device byte counter before the level
      if (clk.on && stats_ls) (void)read_u64(c, stats_ls, &prev_bytes, 1);
      // one launch per lane-count group over its row range of the level
      auto launch = [&](bool write, uint32_t p, uint32_t* cnt, uint64_t* off, uint32_t* out) -> int {
        for (size_t gi = 0; gi + 1 < gstart.size(); gi++) {
          const uint64_t lo = qb[gstart[gi]], hi = qb[gstart[gi + 1]];
          if (hi <= lo) continue;
          const uint32_t GG = glanes[gi];
          const uint32_t grid = bgrid(c, hi - lo, FSIG_BLOCK / GG);
          // tile of the fused tails among this group's queries
          uint32_t tw = 0;
          if (!write && GG >= 8)
            for (uint32_t j = gstart[gi]; j < gstart[gi + 1]; j++)
              if (ml[j].tail) tw = std::max(tw, tail_tile_words(GG, ml[j].L2, 0));
          const size_t lds = (size_t)tw * (FSIG_BLOCK / GG) * sizeof(uint32_t);
#define FSIG_LS(GV, W) k_join_multi<GV, W><<<grid, FSIG_BLOCK, lds, c->stream>>>(dml, dqb, nl, rows, lo, hi, p, cnt, off, out, dcounts, stats_ls, tw)
          if (!write) {
            switch (GG) { case 4: FSIG_LS(4, false); break; case 8: FSIG_LS(8, false); break;
                          case 16: FSIG_LS(16, false); break; case 32: FSIG_LS(32, false); break;
                          default: FSIG_LS(64, false); break; }
          } else {
            switch (GG) { case 4: FSIG_LS(4, true); break; case 8: FSIG_LS(8, true); break;
                          case 16: FSIG_LS(16, true); break; case 32: FSIG_LS(32, true); break;
                          default: FSIG_LS(64, true); break; }
          }
#undef FSIG_LS
          FSIG_HIP(c, hipGetLastError());
        }
        return FSIG_OK;
      };
      for (uint32_t p = 1; p < maxk && R > 0; p++) {
        if (clk.on) {
          clk.mark("lockstep level start");
          if (stats_ls) (void)read_u64(c, stats_ls, &prev_bytes, 1);
        }
        bool any_more = false;
        make_groups(p);
        for (uint32_t j = 0; j < nl; j++) {
          ml[j] = make_level(live[j], p, lanes_of(live[j], p) >= 8 && c->tune.fused_tail == 0);
          any_more |= ml[j].last == 0;  // tail rows may defer children to the last level
        }
        FSIG_HIP(c, hipMemcpyAsync(dml, ml.data(), sizeof(MLevel) * nl, hipMemcpyHostToDevice, c->stream));
        uint32_t* cnt = (uint32_t*)A.alloc(4ull * R);
        uint64_t* off = (uint64_t*)A.alloc(8ull * (R + 1));
        if (!cnt || !off) { A.release(mk); return set_err(c, FSIG_ENOMEM, "lockstep: arena exhausted"); }
        int rc2 = launch(false, p, cnt, off, nullptr);
        if (rc2) { A.release(mk); return rc2; }
        if (!any_more) break;
        if ((rc2 = scan_u32_to_u64(c, cnt, off, R))) { A.release(mk); return rc2; }
        uint64_t T;
        if ((rc2 = read_u64(c, off + R, &T, 1))) { A.release(mk); return rc2; }
        if (any_bound && (rc2 = settle_bounded(p, rows, cnt, off, R, dqb, dqb2, qb, live, T))) {
          A.release(mk);
          return rc2;
        }
        if (clk.on) {
          uint64_t st2[2] = {0, 0};
          if (stats_ls && read_u64(c, stats_ls, st2, 2) == FSIG_OK) {
            fprintf(stderr, "[fsig batch]   level %u: %llu rows -> %llu, count-pass bytes/row %.1f\n", p,
                    (unsigned long long)R, (unsigned long long)T, R ? (double)(st2[0] - prev_bytes) / R : 0.0);
          }
          clk.mark("  count + scan");
        }
        if (T == 0) break;
        uint32_t* out = (uint32_t*)A.alloc(4ull * (p + 1) * T);
        if (!out) { A.release(mk); return set_err(c, FSIG_ENOMEM, "lockstep: a level does not fit the arena"); }
        if ((rc2 = launch(true, p, cnt, off, out))) { A.release(mk); return rc2; }
        k_gather_bases<<<(nl + 256) / 256, 256, 0, c->stream>>>(off, dqb, nl, dqb2);
        FSIG_HIP(c, hipGetLastError());
        std::swap(dqb, dqb2);
        // row bases of the next level, for the per-group launch ranges
        for (uint32_t j = 0; j <= nl; j += kPinnedWords)
          if ((rc2 = read_u64(c, dqb + j, qb.data() + j, std::min<size_t>(kPinnedWords, nl + 1 - j)))) {
            A.release(mk);
            return rc2;
          }
        rows = out;
        R = T;
      }
    }
    FSIG_HIP(c, hipEventRecord(c->ev_b, c->stream));
    FSIG_HIP(c, hipEventSynchronize(c->ev_b));
    FSIG_HIP(c, hipEventElapsedTime(&ms, c->ev_a, c->ev_b));
    clk.mark("lockstep levels");
    const int rc2 = any_bound ? resume_beamed() : FSIG_OK;
    A.release(mk);
    return rc2;
  };
  // Persistent work-queue walk over the seed tasks of a set of queries.
  auto groups_for = [&](uint32_t G, uint32_t grid) { return (uint64_t)grid * (FSIG_BLOCK / G); };
  auto grid_for_G = [&](uint32_t G) {
    uint64_t grid = (uint64_t)c->num_cus * 8;
    const uint64_t avail = A.avail() / 2;
    while (grid > 64 && groups_for(G, (uint32_t)grid) * stack_words * 4ull > avail) grid >>= 1;
    return (uint32_t)grid;
  };
  auto persistent = [&](const TuneChoice& tc, const std::vector<uint32_t>& qs, bool capture, float& ms,
                        unsigned long long* stats) -> int {
    const size_t mk = A.mark();
    std::vector<uint8_t> in(nq, 0);
    for (uint32_t i : qs) in[i] = 1;
    std::vector<Task> sub;
    for (const Task& t : tasks)
      if (in[t.q]) sub.push_back(t);
    const Task* dsub = upload_table(c, sub);
    const uint32_t grid = grid_for_G(tc.G);
    uint32_t* stacks = (uint32_t*)A.alloc(groups_for(tc.G, grid) * stack_words * 4ull);
    if (!dsub || !stacks) { A.release(mk); return set_err(c, FSIG_ENOMEM, "batch: arena exhausted for DFS stacks"); }
    const DfsArgs da{dQ, dL, dsub, (uint32_t)sub.size(), qu, dtasks, ready, dcap, stacks, stack_words, dcounts, stats};
    int rc2;
    if (capture) {
      // queue reset + persistent kernel as one graph.  The instantiated graph
      // of this engine shape is kept by the context and updated from each
      // new capture (its kernel arguments change with the batch); a capture
      // whose topology no longer matches is instantiated afresh.
      hipGraph_t graph;
      FSIG_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      rc2 = launch_dfs(c, tc, grid, c->stream, da);
      hipError_t ce = hipStreamEndCapture(c->stream, &graph);
      if (rc2) { A.release(mk); return rc2; }
      if (ce != hipSuccess) { A.release(mk); return set_err(c, FSIG_EHIP, std::string("batch: capture: ") + hipGetErrorString(ce)); }
      const uint64_t ekey = ((uint64_t)tc.G << 48) ^ ((uint64_t)tc.lds << 47) ^ ((uint64_t)tc.tile << 32) ^ grid;
      auto it = c->batch_exec.find(ekey);
      bool ready = false;
      if (it != c->batch_exec.end()) {
        hipGraphExecUpdateResult ur;
        hipGraphNode_t bad = nullptr;
        ready = hipGraphExecUpdate(it->second, graph, &bad, &ur) == hipSuccess && ur == hipGraphExecUpdateSuccess;
        if (!ready) {
          (void)hipGetLastError();
          (void)hipGraphExecDestroy(it->second);
          c->batch_exec.erase(it);
        }
      }
      if (!ready) {
        hipGraphExec_t exec;
        const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (ie != hipSuccess) {
          (void)hipGraphDestroy(graph);
          A.release(mk);
          return set_err(c, FSIG_EHIP, std::string("batch: instantiate: ") + hipGetErrorString(ie));
        }
        it = c->batch_exec.emplace(ekey, exec).first;
      }
      (void)hipGraphDestroy(graph);
      FSIG_HIP(c, hipEventRecord(c->ev_a, c->stream));
      FSIG_HIP(c, hipGraphLaunch(it->second, c->stream));
      FSIG_HIP(c, hipEventRecord(c->ev_b, c->stream));
      FSIG_HIP(c, hipEventSynchronize(c->ev_b));
    } else {
      FSIG_HIP(c, hipEventRecord(c->ev_a, c->stream));
      if ((rc2 = launch_dfs(c, tc, grid, c->stream, da))) { A.release(mk); return rc2; }
      FSIG_HIP(c, hipEventRecord(c->ev_b, c->stream));
      FSIG_HIP(c, hipEventSynchronize(c->ev_b));
    }
    FSIG_HIP(c, hipEventElapsedTime(&ms, c->ev_a, c->ev_b));
    uint32_t qerr = 0;
    FSIG_HIP(c, hipMemcpyAsync(c->h_pinned, &qu->error, 4, hipMemcpyDeviceToHost, c->stream));
    FSIG_HIP(c, hipStreamSynchronize(c->stream));
    qerr = *(const uint32_t*)c->h_pinned;
    A.release(mk);
    if (qerr) return set_err(c, FSIG_EHIP, "batch: a work-queue wait ran past its bound");
    return FSIG_OK;
  };
  std::vector<uint32_t> all_q;
  for (uint32_t i = 0; i < nq; i++)
    if (active[i]) all_q.push_back(i);
  // batch shape of the tuning cache (with the graph's pointer and size, so
  // a graph uploaded later at a freed graph's address does not inherit it)
  const uint64_t shape = ((uint64_t)nq << 40) ^ ((uint64_t)g->n << 8) ^ (uint64_t)stack_words ^ (g->nnz << 20);
  TuneChoice best{16, false, 0, true};
  if (any_bound || c->tune.strategy == 3) {
    // bounded batches settle queries inside the lockstep join, which timing
    // runs of the autotuner would repeat
    best = TuneChoice{0, false, 0, true};  // lockstep, lanes per query
  } else if (c->tune.lanes_per_row) {
    best = TuneChoice{c->tune.lanes_per_row, c->tune.strategy == 2, c->tune.lds_tile ? c->tune.lds_tile : 256,
                      c->tune.strategy != 1 && c->tune.strategy != 2};
  } else {
    bool cached = false;
    {
      auto it = c->batch_tuned.find({g, shape});
      if (it != c->batch_tuned.end()) {
        best = TuneChoice{it->second[0], it->second[1] != 0, it->second[2], it->second[3] != 0};
        cached = true;
      }
    }
    if (!cached) {
      // The lockstep shapes are timed on the whole batch (a few heavy
      // queries can dominate it and a sample may miss them).  The persistent
      // walk is timed on every 8th query and scaled by the best lockstep
      // shape's whole-batch / sample ratio.
      std::vector<uint32_t> sample;
      for (size_t j = 0; j < all_q.size(); j += 8) sample.push_back(all_q[j]);
      const TuneChoice ls_cands[] = {{0, false, 0, true}, {8, false, 0, true}, {16, false, 0, true},
                                     {64, false, 0, true}};
      const TuneChoice ps_cands[] = {{16, false, 0, false}, {64, false, 0, false}, {16, true, 256, false}};
      float best_ms = 1e30f;
      for (const TuneChoice& tc : ls_cands) {
        float ms = 0;
        FSIG_HIP(c, hipMemsetAsync(dcounts, 0, 8ull * nq, c->stream));
        rc = lockstep(tc.G, all_q, ms);
        if (rc == FSIG_ENOMEM) continue;  // a level does not fit: the walk needs no tables
        if (rc) return rc;
        if (ms < best_ms) { best_ms = ms; best = tc; }
      }
      float ls_sample = 0;
      if (best_ms < 1e30f) {
        FSIG_HIP(c, hipMemsetAsync(dcounts, 0, 8ull * nq, c->stream));
        if ((rc = lockstep(best.G, sample, ls_sample)) && rc != FSIG_ENOMEM) return rc;
      }
      const double scale = (best_ms < 1e30f && ls_sample > 0) ? (double)best_ms / ls_sample : 8.0;
      for (const TuneChoice& tc : ps_cands) {
        float ms = 0;
        FSIG_HIP(c, hipMemsetAsync(dcounts, 0, 8ull * nq, c->stream));
        rc = persistent(tc, sample, false, ms, nullptr);
        if (rc == FSIG_ENOMEM) continue;
        if (rc) return rc;
        if (ms * scale < best_ms) { best_ms = (float)(ms * scale); best = tc; }
      }
      c->batch_tuned[{g, shape}] = {best.G, best.lds ? 1u : 0u, best.tile, best.lockstep ? 1u : 0u};
    }
  }
  clk.mark("autotune");
  FSIG_HIP(c, hipMemsetAsync(dcounts, 0, 8ull * nq, c->stream));
  float ms = 0;
  {
    KTimer tj(c, FSIG_K_BATCH);
    rc = best.lockstep ? lockstep(best.G, all_q, ms) : persistent(best, all_q, true, ms, family_counters(c, FSIG_K_BATCH));
    tj.done(0, 0);
  }
  if (rc == FSIG_ENOMEM && best.lockstep && !any_bound) {  // a level did not fit: the walk needs no tables
    FSIG_HIP(c, hipMemsetAsync(dcounts, 0, 8ull * nq, c->stream));
    best = TuneChoice{64, false, 0, false};
    rc = persistent(best, all_q, true, ms, family_counters(c, FSIG_K_BATCH));
  }
  if (rc) return rc;
  std::vector<uint64_t> hc(nq);
  for (uint32_t i = 0; i < nq; i += kPinnedWords)
    if ((rc = read_u64(c, dcounts + i, hc.data() + i, std::min<size_t>(kPinnedWords, nq - i)))) return rc;
  clk.mark("lockstep join");
  if (clk.on) {
    uint32_t nbeamed = 0;
    for (uint32_t i = 0; i < nq; i++) nbeamed += beamed[i];
    fprintf(stderr, "[fsig batch] beamed %u queries\n", nbeamed);
  }
  for (uint32_t i = 0; i < nq; i++) {
    counts_out[i] = active[i] ? hc[i] : 0;
    if (active[i] && ks[i] == 1) counts_out[i] = size[cls_base[i] + plans[i].cls[0]];
    if (settled[i]) counts_out[i] = settled_count[i];
    if (bound[i]) counts_out[i] = std::min(counts_out[i], bound[i]);
    if (status_out) status_out[i] = settled_rc[i];
  }
  bool refused = false;
  for (uint32_t i = 0; i < nq; i++) refused |= settled_rc[i] != FSIG_OK;
  if (refused && !status_out) return set_err(c, FSIG_ELIMIT, "batch: a query exceeded the work limit");
  clk.mark("join");
  if (join_ms) *join_ms = ms;
  c->last_tune_G = best.G;
  c->last_tune_lds = best.lds ? best.tile : 0;
  c->last_tune_engine = best.lockstep ? 1 : 2;
  c->last_peak = A.peak - mark;
  FSIG_HIP(c, hipStreamSynchronize(c->stream));
  return resolve_timings(c);
}

extern "C" int fsig_match_batch(fsig_ctx* c, const fsig_graph* g, uint32_t nq, const uint32_t* ks,
                                const int32_t* const* qlabels, const uint32_t* const* qoff,
                                const uint32_t* const* qadj, uint64_t* counts_out, double* join_ms) {
  return batch_impl(c, g, nq, ks, qlabels, qoff, qadj, nullptr, counts_out, nullptr, join_ms);
}

extern "C" int fsig_match_batch_bounded(fsig_ctx* c, const fsig_graph* g, uint32_t nq, const uint32_t* ks,
                                        const int32_t* const* qlabels, const uint32_t* const* qoff,
                                        const uint32_t* const* qadj, const uint64_t* max_embeddings,
                                        uint64_t* counts_out, int32_t* status_out, double* join_ms) {
  if (c && !max_embeddings) return set_err(c, FSIG_EINVAL, "batch: null bounds");
  return batch_impl(c, g, nq, ks, qlabels, qoff, qadj, max_embeddings, counts_out, status_out, join_ms);
}
