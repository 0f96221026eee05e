// Batched engine: many queries against one resident graph.
//
// Setup (filter, refinement, candidate edges) runs for all queries at once:
// every kernel walks a device table of (query, class) or (query, class
// pair) descriptors, so the whole batch costs a fixed number of launches and
// host round trips instead of a dozen per query.
//
// The join is one persistent kernel fed by a device work queue of
// (query, root range) tasks, launched through a captured hipGraph.  Each
// group of G lanes pops a task and walks its roots depth-first, keeping one
// candidate list per level on a private stack in HBM; the last level is
// counted, never stored.  No host synchronisation happens between levels or
// queries, which is where the per-query plain path spends most of its time
// on small queries (it syncs per level to size the next table).
//
// The autotuner times the persistent kernel over a sample of the task queue
// for each candidate (lanes per group, intersection strategy, LDS tile) and
// runs the batch with the fastest; choices are cached per batch shape.
#include <algorithm>
#include <cstring>
#include <map>

#include "common.h"
#include "device.h"
#include "internal.h"
#include "join.h"

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

__global__ void __launch_bounds__(FSIG_BLOCK) k_bldf(const BSeg* __restrict__ S, uint32_t ns, uint64_t total,
                                                     const uint32_t* __restrict__ cdeg,
                                                     const uint32_t* __restrict__ off) {
  for (uint64_t f = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; f < total; f += (uint64_t)gridDim.x * blockDim.x) {
    const BSeg& s = S[find_flat(S, ns, f)];
    const uint32_t v = s.src[f - s.flat];
    const uint32_t d = off[v + 1] - off[v];
    uint32_t m = 0, cm = s.cmask;
    while (cm) {
      const uint32_t c = __builtin_ctz(cm);
      cm &= cm - 1;
      if (d >= cdeg[s.cls_base + c]) m |= 1u << c;
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
  const uint32_t lane = threadIdx.x & 63u, gl = lane & (G - 1), gw = threadIdx.x / G;
  const uint64_t gmask = (G == 64) ? ~0ull : (((1ull << G) - 1) << (lane - gl));
  const uint64_t ngroups = (uint64_t)gridDim.x * GPB;
  for (uint64_t f = (uint64_t)xcd_block() * GPB + gw; f < total; f += ngroups) {
    const BClass& cl = C[find_flat(C, nc, f)];
    const uint32_t v = cl.cand[f - cl.flat];
    const uint32_t s = off[v], e = off[v + 1];
    const uint32_t* mask = cl.mask;
    bool ok = true;
    uint32_t tm = cl.tmask;
    while (tm && ok) {
      const uint32_t b = __builtin_ctz(tm);
      tm &= tm - 1;
      const uint32_t need = thr[(uint64_t)cl.thr_row * MAXQ + b];
      uint32_t got = 0;
      for (uint32_t j = s; j < e && got < need; j += G) {
        const uint32_t i = j + gl;
        const bool hit = i < e && ((mask[adj[i]] >> b) & 1u);
        got += __popcll(__ballot(hit) & gmask);
      }
      ok = got >= need;
    }
    if (gl == 0) keep[f] = ok ? 1 : 0;
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
  uint32_t* off_out;
  uint32_t asize, bsize, bbit, pad;
  uint64_t flat;
};

template <int G>
__global__ void __launch_bounds__(FSIG_BLOCK) k_bcedge_count(const BPair* __restrict__ P, uint32_t np, uint64_t total,
                                                             const uint32_t* __restrict__ off,
                                                             const uint32_t* __restrict__ adj,
                                                             uint32_t* __restrict__ cnt,
                                                             unsigned int* __restrict__ maxlen) {
  constexpr int GPB = FSIG_BLOCK / G;
  const uint32_t lane = threadIdx.x & 63u, gl = lane & (G - 1), gw = threadIdx.x / G;
  const uint64_t gmask = (G == 64) ? ~0ull : (((1ull << G) - 1) << (lane - gl));
  const uint64_t ngroups = (uint64_t)gridDim.x * GPB;
  for (uint64_t f = (uint64_t)xcd_block() * GPB + gw; f < total; f += ngroups) {
    const uint32_t pi = find_flat(P, np, f);
    const BPair& pd = P[pi];
    const uint32_t v = pd.acand[f - pd.flat];
    const uint32_t s = off[v], e = off[v + 1];
    uint32_t got = 0;
    for (uint32_t j = s; j < e; j += G) {
      const uint32_t i = j + gl;
      const bool hit = i < e && ((pd.mask[adj[i]] >> pd.bbit) & 1u);
      got += __popcll(__ballot(hit) & gmask);
    }
    if (gl == 0) {
      cnt[f] = got;
      atomicMax(&maxlen[pi], got);
    }
  }
}

template <int G>
__global__ void __launch_bounds__(FSIG_BLOCK) k_bcedge_fill(const BPair* __restrict__ P, uint32_t np, uint64_t total,
                                                            const uint32_t* __restrict__ off,
                                                            const uint32_t* __restrict__ adj,
                                                            const uint64_t* __restrict__ gscan,
                                                            uint32_t* __restrict__ lst_all) {
  constexpr int GPB = FSIG_BLOCK / G;
  const uint32_t lane = threadIdx.x & 63u, gl = lane & (G - 1), gw = threadIdx.x / G;
  const uint64_t gmask = (G == 64) ? ~0ull : (((1ull << G) - 1) << (lane - gl));
  const uint64_t below = (1ull << lane) - 1;
  const uint64_t ngroups = (uint64_t)gridDim.x * GPB;
  for (uint64_t f = (uint64_t)xcd_block() * GPB + gw; f < total; f += ngroups) {
    const BPair& pd = P[find_flat(P, np, f)];
    const uint32_t i0 = (uint32_t)(f - pd.flat);
    const uint32_t v = pd.acand[i0];
    const uint64_t pbase = gscan[pd.flat];
    uint64_t at = gscan[f];
    if (gl == 0) {
      pd.off_out[i0] = (uint32_t)(at - pbase);
      if (i0 + 1 == pd.asize) pd.off_out[pd.asize] = (uint32_t)(gscan[f + 1] - pbase);
    }
    const uint32_t s = off[v], e = off[v + 1];
    for (uint32_t j = s; j < e; j += G) {
      const uint32_t i = j + gl;
      uint32_t x = 0;
      bool hit = false;
      if (i < e) {
        x = adj[i];
        hit = (pd.mask[x] >> pd.bbit) & 1u;
      }
      const uint64_t bal = __ballot(hit) & gmask;
      if (hit) lst_all[at + __popcll(bal & below)] = sorted_index(pd.bcand, pd.bsize, x);
      at += __popcll(bal);
    }
  }
}

__global__ void k_bpair_base(const uint64_t* __restrict__ gscan, const BPair* __restrict__ P, uint32_t np,
                             uint64_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < np) out[i] = gscan[P[i].flat];
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
};

struct Task {
  uint32_t q;
  uint32_t r0, r1;  // roots (candidate indices of the first join vertex)
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

template <int G, bool LDS_STAGE>
__global__ void __launch_bounds__(FSIG_BLOCK) k_dfs(const BQuery* __restrict__ Q, const BLevel* __restrict__ LV,
                                                    const Task* __restrict__ tasks, uint32_t ntasks,
                                                    unsigned int* __restrict__ head, uint32_t* __restrict__ stacks,
                                                    uint32_t stack_words, uint32_t tile_words,
                                                    unsigned long long* __restrict__ counts,
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
  uint64_t bytes = 0, nodes = 0;
  for (;;) {
    uint32_t ti = 0;
    if (gl == 0) ti = atomicAdd(head, 1u);
    ti = __shfl(ti, 0, G);
    if (ti >= ntasks) break;  // every group leaves once the queue is drained
    const Task T = tasks[ti];
    const BQuery q = Q[T.q];
    const BLevel* lv = LV + q.level_base;
    unsigned long long cnt = 0;
    for (uint32_t root = T.r0; root < T.r1; root++) {
      path[0] = root;
      if (q.k == 1) { cnt++; continue; }
      if (q.k == 2) {
        cnt += level_candidates<G, false, LDS_STAGE>(lv[1], path, sh_s[gw], sh_n[gw], sh_cf[gw], tile, tile_words,
                                                     gl, gmask, below, nullptr, bytes);
        nodes++;
        continue;
      }
      len[1] = level_candidates<G, true, LDS_STAGE>(lv[1], path, sh_s[gw], sh_n[gw], sh_cf[gw], tile, tile_words,
                                                    gl, gmask, below, stk + lv[1].stack_off, bytes);
      pos[1] = 0;
      nodes++;
      uint32_t p = 1;
      for (;;) {
        if (pos[p] == len[p]) {
          if (p == 1) break;
          p--;
          continue;
        }
        const uint32_t x = stk[lv[p].stack_off + pos[p]];
        pos[p] = pos[p] + 1;
        path[p] = x;
        nodes++;
        if (p + 2 == q.k) {
          cnt += level_candidates<G, false, LDS_STAGE>(lv[p + 1], path, sh_s[gw], sh_n[gw], sh_cf[gw], tile,
                                                       tile_words, gl, gmask, below, nullptr, bytes);
        } else {
          len[p + 1] = level_candidates<G, true, LDS_STAGE>(lv[p + 1], path, sh_s[gw], sh_n[gw], sh_cf[gw], tile,
                                                            tile_words, gl, gmask, below,
                                                            stk + lv[p + 1].stack_off, bytes);
          pos[p + 1] = 0;
          p++;
        }
      }
    }
    if (gl == 0) atomicAdd(&counts[T.q], cnt);
  }
  if (stats && gl == 0) {
    atomicAdd(&stats[0], (unsigned long long)bytes);
    atomicAdd(&stats[1], (unsigned long long)nodes);
  }
}

__global__ void k_zero_u32(unsigned int* p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------------------
struct TuneChoice {
  uint32_t G;
  bool lds;
  uint32_t tile;
};

template <int G, bool LDS>
int launch_dfs_t(fsig_ctx* c, uint32_t grid, hipStream_t st, const BQuery* Q, const BLevel* LV, const Task* T,
                 uint32_t nt, unsigned int* head, uint32_t* stacks, uint32_t sw, uint32_t tile,
                 unsigned long long* counts, unsigned long long* stats) {
  const size_t dyn = LDS ? (size_t)(FSIG_BLOCK / G) * tile * 4 : 0;
  k_dfs<G, LDS><<<grid, FSIG_BLOCK, dyn, st>>>(Q, LV, T, nt, head, stacks, sw, tile, counts, stats);
  FSIG_HIP(c, hipGetLastError());
  return FSIG_OK;
}

int launch_dfs(fsig_ctx* c, const TuneChoice& tc, uint32_t grid, hipStream_t st, const BQuery* Q, const BLevel* LV,
               const Task* T, uint32_t nt, unsigned int* head, uint32_t* stacks, uint32_t sw,
               unsigned long long* counts, unsigned long long* stats) {
  const uint32_t tile = tc.lds ? tc.tile : 0;
#define FSIG_DFS_CASE(GG)                                                                                   \
  case GG:                                                                                                   \
    return tc.lds ? launch_dfs_t<GG, true>(c, grid, st, Q, LV, T, nt, head, stacks, sw, tile, counts, stats) \
                  : launch_dfs_t<GG, false>(c, grid, st, Q, LV, T, nt, head, stacks, sw, tile, counts, stats);
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

// Per-context cache of tuning decisions, keyed by a coarse batch shape.
static std::map<std::pair<const fsig_ctx*, uint64_t>, TuneChoice> g_tuned;

extern "C" int fsig_match_batch(fsig_ctx* c, const fsig_graph* g, uint32_t nq, const uint32_t* ks,
                                const int32_t* const* qlabels, const uint32_t* const* qoff,
                                const uint32_t* const* qadj, uint64_t* counts_out, double* join_ms) {
  if (!c) return FSIG_EINVAL;
  if (!g || !ks || !qlabels || !qoff || !qadj || !counts_out) return set_err(c, FSIG_EINVAL, "batch: null argument");
  FSIG_HIP(c, hipSetDevice(c->device));
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
  if (flat)
    k_bldf<<<bgrid(c, flat, FSIG_BLOCK), FSIG_BLOCK, 0, c->stream>>>(dsegs, (uint32_t)segs_nz.size(), flat, dcdeg,
                                                                     g->d_off);
  std::vector<uint32_t*> cand_a(ncls, nullptr), cand_b(ncls, nullptr);
  std::vector<uint32_t> size(ncls, 0);
  for (uint32_t i = 0; i < nq; i++) {
    const QueryPlan& P = plans[i];
    for (uint32_t a = 0; a < P.nclass; a++) {
      const uint32_t gc = cls_base[i] + a;
      const int32_t lab = P.class_label[a];
      const uint32_t len = active[i] ? g->h_inv_index[lab + 1] - g->h_inv_index[lab] : 0;
      cand_a[gc] = (uint32_t*)A.alloc(4ull * std::max<uint32_t>(len, 1));
      if (!cand_a[gc]) return set_err(c, FSIG_ENOMEM, "batch: arena exhausted for candidates");
      comp[gc] = BComp{g->d_inv_list + g->h_inv_index[lab], cand_a[gc], masks + (uint64_t)i * n, len, a, 0, tiles};
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
  if (ncls && (rc = run_compact(comp, tiles, 0, nullptr))) return rc;
  tf.done(4ull * n * nq + 16ull * flat, flat);

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
  for (uint32_t pass = 0; any_target && pass < c->tune.refine_passes; pass++) {
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

  // ---- orders and the class pairs each query's join needs
  std::vector<BPair> pairs;
  std::vector<std::vector<int>> pair_of(nq);  // [i][a*MAXQ+b] -> pair index
  uint64_t pflat = 0;
  for (uint32_t i = 0; i < nq; i++) {
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
        BPair pd{cand[ga], cand[gb], masks + (uint64_t)i * n, nullptr, size[ga], size[gb], b, 0, pflat};
        pd.off_out = (uint32_t*)A.alloc(4ull * (size[ga] + 1));
        if (!pd.off_out) return set_err(c, FSIG_ENOMEM, "batch: arena exhausted for candidate edges");
        pflat += size[ga];
        pairs.push_back(pd);
      }
    }
  }
  const uint32_t np = (uint32_t)pairs.size();
  std::vector<uint64_t> pbase_h(np), maxlen_h(np);
  uint32_t* lst_all = nullptr;
  if (np) {
    KTimer tc(c, FSIG_K_CEDGE);
    const BPair* dP = upload_table(c, pairs);
    uint32_t* cnt = (uint32_t*)A.alloc(4ull * std::max<uint64_t>(pflat, 1));
    uint64_t* gscan = (uint64_t*)A.alloc(8ull * (pflat + 1));
    unsigned int* dmax = (unsigned int*)A.alloc(4ull * np);
    uint64_t* pb = (uint64_t*)A.alloc(8ull * np);
    if (!dP || !cnt || !gscan || !dmax || !pb) return set_err(c, FSIG_ENOMEM, "batch: arena exhausted");
    k_zero_u32<<<(np + 255) / 256, 256, 0, c->stream>>>(dmax, np);
    switch (RG) {
      case 8: k_bcedge_count<8><<<bgrid(c, pflat, FSIG_BLOCK / 8), FSIG_BLOCK, 0, c->stream>>>(dP, np, pflat, g->d_off, g->d_adj, cnt, dmax); break;
      case 16: k_bcedge_count<16><<<bgrid(c, pflat, FSIG_BLOCK / 16), FSIG_BLOCK, 0, c->stream>>>(dP, np, pflat, g->d_off, g->d_adj, cnt, dmax); break;
      default: k_bcedge_count<64><<<bgrid(c, pflat, FSIG_BLOCK / 64), FSIG_BLOCK, 0, c->stream>>>(dP, np, pflat, g->d_off, g->d_adj, cnt, dmax); break;
    }
    FSIG_HIP(c, hipGetLastError());
    if ((rc = scan_u32_to_u64(c, cnt, gscan, pflat))) return rc;
    uint64_t total_edges;
    if ((rc = read_u64(c, gscan + pflat, &total_edges, 1))) return rc;
    lst_all = (uint32_t*)A.alloc(4ull * std::max<uint64_t>(total_edges, 1));
    if (!lst_all) return set_err(c, FSIG_ENOMEM, "batch: arena exhausted for candidate lists");
    switch (RG) {
      case 8: k_bcedge_fill<8><<<bgrid(c, pflat, FSIG_BLOCK / 8), FSIG_BLOCK, 0, c->stream>>>(dP, np, pflat, g->d_off, g->d_adj, gscan, lst_all); break;
      case 16: k_bcedge_fill<16><<<bgrid(c, pflat, FSIG_BLOCK / 16), FSIG_BLOCK, 0, c->stream>>>(dP, np, pflat, g->d_off, g->d_adj, gscan, lst_all); break;
      default: k_bcedge_fill<64><<<bgrid(c, pflat, FSIG_BLOCK / 64), FSIG_BLOCK, 0, c->stream>>>(dP, np, pflat, g->d_off, g->d_adj, gscan, lst_all); break;
    }
    k_bpair_base<<<(np + 255) / 256, 256, 0, c->stream>>>(gscan, dP, np, pb);
    FSIG_HIP(c, hipGetLastError());
    for (uint32_t i = 0; i < np; i += kPinnedWords)
      if ((rc = read_u64(c, pb + i, pbase_h.data() + i, std::min<size_t>(kPinnedWords, np - i)))) return rc;
    std::vector<unsigned int> mx(np);
    FSIG_HIP(c, hipMemcpyAsync(mx.data(), dmax, 4ull * np, hipMemcpyDeviceToHost, c->stream));
    FSIG_HIP(c, hipStreamSynchronize(c->stream));
    for (uint32_t i = 0; i < np; i++) maxlen_h[i] = mx[i];
    tc.done(2ull * pflat * 12 + 12ull * total_edges, pflat);
  }

  // ---- level tables, stacks and tasks
  std::vector<BQuery> bq(nq);
  std::vector<BLevel> lv;
  std::vector<Task> tasks;
  uint32_t stack_words = 1;
  for (uint32_t i = 0; i < nq; i++) {
    bq[i].k = ks[i];
    bq[i].level_base = (uint32_t)lv.size();
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
    // tasks: roots in chunks; smaller chunks for larger queries
    const uint32_t n1 = size[cls_base[i] + P.cls[P.order[0]]];
    const uint32_t chunk = P.k <= 4 ? 64 : (P.k <= 8 ? 16 : 4);
    for (uint32_t r = 0; r < n1; r += chunk) tasks.push_back(Task{i, r, std::min(n1, r + chunk)});
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
  unsigned int* head = (unsigned int*)A.alloc(64);
  if (!dQ || !dL || !dT || !dcounts || !head) return set_err(c, FSIG_ENOMEM, "batch: arena exhausted");
  const uint32_t nt = (uint32_t)tasks.size();

  // ---- autotune over a sample of the queue, then the captured full run
  auto groups_for = [&](uint32_t G, uint32_t grid) { return (uint64_t)grid * (FSIG_BLOCK / G); };
  auto grid_for_G = [&](uint32_t G) {
    uint64_t grid = (uint64_t)c->num_cus * 8;
    const uint64_t avail = A.avail() / 2;
    while (grid > 64 && groups_for(G, (uint32_t)grid) * stack_words * 4ull > avail) grid >>= 1;
    return (uint32_t)grid;
  };
  const uint32_t max_grid = grid_for_G(4);
  uint32_t* stacks = (uint32_t*)A.alloc(groups_for(4, max_grid) * stack_words * 4ull);
  if (!stacks) return set_err(c, FSIG_ENOMEM, "batch: arena exhausted for DFS stacks");
  const uint64_t shape = ((uint64_t)nq << 40) ^ ((uint64_t)g->n << 8) ^ (uint64_t)stack_words;
  TuneChoice best{16, false, 0};
  if (c->tune.lanes_per_row) {
    best = TuneChoice{c->tune.lanes_per_row, c->tune.strategy == 2, c->tune.lds_tile ? c->tune.lds_tile : 256};
  } else {
    auto it = g_tuned.find({c, shape});
    if (it != g_tuned.end()) {
      best = it->second;
    } else {
      // sample: a prefix of the interleaved queue (every query represented)
      const uint32_t sample = std::min<uint32_t>(nt, std::max<uint32_t>(nq * 2, nt / 16));
      const TuneChoice cands[] = {{4, false, 0}, {8, false, 0}, {16, false, 0}, {32, false, 0},
                                  {8, true, 128}, {16, true, 256}, {32, true, 512}};
      float best_ms = 1e30f;
      for (const TuneChoice& tc : cands) {
        const uint32_t grid = std::min(max_grid, grid_for_G(tc.G));
        k_zero_u32<<<1, 64, 0, c->stream>>>(head, 1);
        FSIG_HIP(c, hipMemsetAsync(dcounts, 0, 8ull * nq, c->stream));
        FSIG_HIP(c, hipEventRecord(c->ev_a, c->stream));
        if ((rc = launch_dfs(c, tc, grid, c->stream, dQ, dL, dT, sample, head, stacks, stack_words, dcounts, nullptr)))
          return rc;
        FSIG_HIP(c, hipEventRecord(c->ev_b, c->stream));
        FSIG_HIP(c, hipEventSynchronize(c->ev_b));
        float ms = 0;
        FSIG_HIP(c, hipEventElapsedTime(&ms, c->ev_a, c->ev_b));
        if (ms < best_ms) { best_ms = ms; best = tc; }
      }
      g_tuned[{c, shape}] = best;
    }
  }
  const uint32_t grid = std::min(max_grid, grid_for_G(best.G));
  // the full run: queue reset + persistent kernel, captured once as a graph
  FSIG_HIP(c, hipMemsetAsync(dcounts, 0, 8ull * nq, c->stream));
  unsigned long long* stats = family_counters(c, FSIG_K_BATCH);
  hipGraph_t graph;
  hipGraphExec_t exec;
  FSIG_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  k_zero_u32<<<1, 64, 0, c->stream>>>(head, 1);
  rc = launch_dfs(c, best, grid, c->stream, dQ, dL, dT, nt, head, stacks, stack_words, dcounts, stats);
  hipError_t ce = hipStreamEndCapture(c->stream, &graph);
  if (rc) return rc;
  if (ce != hipSuccess) return set_err(c, FSIG_EHIP, std::string("batch: capture: ") + hipGetErrorString(ce));
  FSIG_HIP(c, hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  KTimer tj(c, FSIG_K_BATCH);
  FSIG_HIP(c, hipEventRecord(c->ev_a, c->stream));
  FSIG_HIP(c, hipGraphLaunch(exec, c->stream));
  FSIG_HIP(c, hipEventRecord(c->ev_b, c->stream));
  tj.done(0, 0);
  std::vector<uint64_t> hc(nq);
  for (uint32_t i = 0; i < nq; i += kPinnedWords)
    if ((rc = read_u64(c, dcounts + i, hc.data() + i, std::min<size_t>(kPinnedWords, nq - i)))) return rc;
  float ms = 0;
  FSIG_HIP(c, hipEventElapsedTime(&ms, c->ev_a, c->ev_b));
  (void)hipGraphExecDestroy(exec);
  (void)hipGraphDestroy(graph);
  for (uint32_t i = 0; i < nq; i++) {
    counts_out[i] = active[i] ? hc[i] : 0;
    if (active[i] && ks[i] == 1) counts_out[i] = size[cls_base[i] + plans[i].cls[0]];
  }
  if (join_ms) *join_ms = ms;
  c->last_tune_G = best.G;
  c->last_tune_lds = best.lds ? best.tile : 0;
  c->last_peak = A.peak - mark;
  FSIG_HIP(c, hipStreamSynchronize(c->stream));
  return resolve_timings(c);
}
