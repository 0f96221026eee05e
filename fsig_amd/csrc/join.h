// Join (enumeration) kernels and the memory-blocked depth-first driver,
// shared by alg 1 (candidate-edge CSR) and alg 2 (PCSR).
//
// A table at level p holds partial embeddings of the first p join-order
// vertices, row-major, p u32 words per row.  Extending row r by the vertex w
// = order[p] takes, for every backward neighbour b of w (query neighbour
// earlier in the order), the sorted list of admissible data vertices for w
// given phi(b); the children of r are the elements common to all those lists
// that differ from every earlier data vertex of the same label.  This is the
// per-row computation of the reference's precompute/compute/extend kernels
// (match/MMatch_kernels.cuh:374-792) and of GSI's load/intersection/link
// kernels (match/GSI_kernels.cuh:95-560).
//
// Work mapping (gfx950, wave64): G lanes of a wavefront share one row
// (G in {4,8,16,32,64}).  The shortest list drives; each lane takes one
// driver element per step and probes the other lists by binary search; a
// wave ballot counts or places the survivors.  Lists never leave their
// sorted order, so children come out sorted within a row.
//
// Levels are processed depth-first over chunks of children whose size is
// bounded by the free arena (the reference's form_blocks,
// memory_manager/memory_manager.h:220-242, and extend_embeddings_recursively,
// match/MMatch.cu:1834-1985), but children are windows of a global child
// numbering, so one huge row can also be split, and nothing spills to host.
#pragma once

#include <algorithm>
#include <type_traits>

#include "device.h"
#include "internal.h"

struct JoinLevel {
  uint32_t p;                 // columns of the parent rows
  uint32_t nb;                // backward neighbours
  uint32_t ncf;               // conflict columns (same label, not adjacent)
  uint32_t cfsame;            // bit j: conflict j compares row words directly
  uint32_t any_cross;         // some conflict compares data vertex ids
  uint32_t wsize;             // |C(w)|; the driver when nb == 0
  uint32_t range_ids;         // nb == 0: driver element t is wcand[t], else t
  uint32_t bcol[MAXQ];
  uint32_t cfcol[MAXQ];
  const uint32_t* cfcand[MAXQ];  // row word -> data vertex for cross conflicts
  const uint32_t* wcand;         // candidate index -> data vertex of w
  // data vertex -> candidate index of w (n words; meaningful where
  // wcand[wpos[v]] == v), or null: cand_index then binary-searches wcand
  const uint32_t* wpos;
  // symmetry breaking (count mode): a child x must exceed the row word of
  // every column in gtmask and stay below that of every column in ltmask
  // (columns of w's class: the words compare as the data vertex ids do)
  uint32_t gtmask, ltmask;
};

constexpr uint32_t kNone = 0xffffffffu;

// First index of the ascending a[0..n) whose element is >= x.
__device__ __forceinline__ uint32_t lower_bound_u32(const uint32_t* __restrict__ a, uint32_t n, uint32_t x) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Bounds [lo, hi) that the symmetry conditions of a level leave for the
// child of a row (rv: the row, G lanes; uniform over the group).
template <int G>
__device__ __forceinline__ uint32_t group_col(const uint32_t (&rv)[(MAXQ + G - 1) / G], uint32_t c);

template <int G>
__device__ __forceinline__ void sym_bounds(uint32_t gtmask, uint32_t ltmask, const uint32_t (&rv)[(MAXQ + G - 1) / G],
                                           uint32_t& lo, uint32_t& hi) {
  lo = 0;
  hi = kNone;
  for (uint32_t m = gtmask; m; m &= m - 1) lo = max(lo, group_col<G>(rv, __builtin_ctz(m)) + 1);
  for (uint32_t m = ltmask; m; m &= m - 1) hi = min(hi, group_col<G>(rv, __builtin_ctz(m)));
}

// Driver positions [ta, tb) whose elements fall in [lo, hi): the driver is
// the ascending list dl, or (dl null) the identity over [0, dlen), or wcand.
__device__ __forceinline__ void driver_window(const uint32_t* dl, const uint32_t* wcand, bool ids, uint32_t dlen,
                                              uint32_t lo, uint32_t hi, uint32_t& ta, uint32_t& tb) {
  if (dl) {
    ta = lo ? lower_bound_u32(dl, dlen, lo) : 0;
    tb = hi != kNone ? lower_bound_u32(dl, dlen, hi) : dlen;
  } else if (ids) {
    ta = lo ? lower_bound_u32(wcand, dlen, lo) : 0;
    tb = hi != kNone ? lower_bound_u32(wcand, dlen, hi) : dlen;
  } else {
    ta = min(lo, dlen);
    tb = min(hi, dlen);
  }
  if (tb < ta) tb = ta;
}

// Candidate index of data vertex id in w's class, or kNone when id is no
// candidate of w: one or two gathers through wpos, else a binary search.
__device__ __forceinline__ uint32_t cand_index(const JoinLevel& L, uint32_t id) {
  const uint32_t at = L.wpos ? L.wpos[id] : sorted_index(L.wcand, L.wsize, id);
  return (at < L.wsize && L.wcand[at] == id) ? at : kNone;
}

struct JoinIO {
  const uint32_t* rows;
  uint64_t row_lo, row_hi;    // rows [row_lo, row_hi) are processed
  uint32_t* cnt;              // count pass: children per row (absolute row index), may be null
  const uint64_t* off;        // write pass: first child index of each row (absolute)
  uint32_t* out;              // write pass: child table, p+1 words per child
  uint64_t out_lo, out_hi;    // write pass: children [out_lo, out_hi) are written, at out_lo -> out[0]
  unsigned long long* total;  // count pass: += children (may be null)
  unsigned long long* stats;  // [0] += algorithmic bytes, [1] += units (rows or children)
};

// Candidate-edge CSR lists (alg 1): list i of row value x (a candidate index of
// the backward neighbour's class) is lst[i][off[i][x] .. off[i][x+1]), holding
// candidate indices of w's class.
struct CandEdgeSrc {
  static constexpr bool kAllMembers = true;  // lists hold only candidates of w
  const uint32_t* off[MAXQ];
  const uint32_t* lst[MAXQ];
  __device__ __forceinline__ void range(uint32_t i, uint32_t x, uint32_t& s, uint32_t& n) const {
    const uint32_t* o = off[i];
    s = o[x];
    n = o[x + 1] - s;
  }
  __device__ __forceinline__ const uint32_t* list(uint32_t i) const { return lst[i]; }
  __device__ __forceinline__ bool member(uint32_t) const { return true; }
};

// PCSR lists (alg 2): list i of a data vertex x is x's neighbour list in the
// unit of the label pair (L(w), L(b_i)), found by MurmurHash2 bucket probing
// (the lookup of match/GSI_kernels.cuh:95-133 with a correct chain walk).
// Elements are data vertex ids; membership in C(w) is a bit of `mask`.
struct PcsrSrc {
  static constexpr bool kAllMembers = false;  // lists hold every neighbour of the partner label
  const uint32_t* groups;
  const uint32_t* edges;
  const uint32_t* mask;
  uint32_t wbit;
  UnitMeta unit[MAXQ];
  __device__ __forceinline__ void range(uint32_t i, uint32_t x, uint32_t& s, uint32_t& n) const {
    const UnitMeta u = unit[i];
    s = 0;
    n = 0;
    if (u.keys == 0) return;
    uint32_t grp = murmur2_u32(x, 17u) % u.keys;
    for (uint32_t hops = 0; hops <= u.keys && grp != FSIG_INVALID; hops++) {
      const uint32_t* gw = groups + u.group_base + 32ull * grp;
      for (uint32_t slot = 0; slot < 15; slot++) {
        const uint32_t key = gw[2 * slot];
        if (key == FSIG_INVALID) return;  // slots fill from 0; a hole ends the bucket
        if (key == x) {  // list is edges[edge_base + start, edge_base + end)
          s = gw[2 * slot + 1];
          n = gw[2 * slot + 3] - s;
          return;
        }
      }
      grp = gw[30];
    }
  }
  __device__ __forceinline__ const uint32_t* list(uint32_t i) const { return edges + unit[i].edge_base; }
  __device__ __forceinline__ bool member(uint32_t x) const { return (mask[x] >> wbit) & 1u; }
};

template <int G>
__device__ __forceinline__ uint32_t group_col(const uint32_t (&rv)[(MAXQ + G - 1) / G], uint32_t c) {
  constexpr int NR = (MAXQ + G - 1) / G;
  const uint32_t j = c / G;
  uint32_t v = rv[0];
#pragma unroll
  for (int jj = 1; jj < NR; jj++)
    if ((uint32_t)jj == j) v = rv[jj];
  return __shfl(v, (int)(c % G), G);
}

// WRITE == false: count pass (children per row into io.cnt, total into
// counters[0]).  WRITE == true: children of rows [row_lo, row_hi) whose global
// index falls in [out_lo, out_hi) are written as p+1 word rows.
template <int G, bool WRITE, class Src>
__global__ void __launch_bounds__(FSIG_BLOCK) k_join(JoinLevel L, Src S, JoinIO io) {
  constexpr int GPB = FSIG_BLOCK / G;  // groups per block
  constexpr int NR = (MAXQ + G - 1) / G;
  __shared__ uint32_t sh_s[GPB][MAXQ];
  __shared__ uint32_t sh_n[GPB][MAXQ];
  __shared__ uint32_t sh_cf[GPB][MAXQ];
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t gl = lane & (G - 1);
  const uint32_t gw = threadIdx.x / G;
  const uint64_t gmask = (G == 64) ? ~0ull : (((1ull << G) - 1) << (lane - gl));
  const uint64_t below = (1ull << lane) - 1;
  const uint64_t ngroups = (uint64_t)gridDim.x * GPB;
  unsigned long long my_children = 0, my_bytes = 0, my_rows = 0;

  for (uint64_t r = io.row_lo + (uint64_t)xcd_block() * GPB + gw; r < io.row_hi; r += ngroups) {
    uint32_t rv[NR];
    const uint32_t* row = io.rows + r * L.p;
#pragma unroll
    for (int j = 0; j < NR; j++) {
      const uint32_t col = j * G + gl;
      rv[j] = col < L.p ? row[col] : 0u;
    }
    // list ranges of the backward neighbours; the shortest one drives
    uint32_t dlen = L.wsize, d = 0;
    uint64_t elems = 0;
    for (uint32_t i = 0; i < L.nb; i++) {
      const uint32_t x = group_col<G>(rv, L.bcol[i]);
      uint32_t s, n;
      S.range(i, x, s, n);
      sh_s[gw][i] = s;
      sh_n[gw][i] = n;
      elems += n;
      if (i == 0 || n < dlen) { dlen = n; d = i; }
    }
    for (uint32_t j = 0; j < L.ncf; j++) {
      uint32_t x = group_col<G>(rv, L.cfcol[j]);
      if (!((L.cfsame >> j) & 1u)) x = L.cfcand[j][x];
      sh_cf[gw][j] = x;
    }
    uint64_t base = 0;
    uint64_t child0 = 0;
    if (WRITE) {
      child0 = io.off[r];
      if (io.off[r + 1] == child0) continue;  // no children (or counted by a fused tail)
    }
    const uint32_t* dl = L.nb ? S.list(d) + sh_s[gw][d] : nullptr;
    // symmetry conditions: the children lie in [lo, hi), driver positions
    // [ta, tb)
    uint32_t lo = 0, hi = kNone, ta = 0, tb = dlen;
    if (L.gtmask | L.ltmask) {
      sym_bounds<G>(L.gtmask, L.ltmask, rv, lo, hi);
      driver_window(dl, L.wcand, L.range_ids != 0, dlen, lo, hi, ta, tb);
    }
    // Count pass with one backward list of candidates only: the children are
    // the list minus the earlier vertices of the same label that occur in
    // it, found by binary search (one lane per conflict).
    const bool direct = !WRITE && Src::kAllMembers && L.nb == 1;
    if (direct) {
      uint32_t hit = 0;
      for (uint32_t j0 = 0; j0 < L.ncf; j0 += G) {
        const uint32_t j = j0 + gl;
        bool in = false;
        if (j < L.ncf) {
          uint32_t target = sh_cf[gw][j];
          if (!((L.cfsame >> j) & 1u)) target = cand_index(L, target);  // data vertex id -> candidate index of w
          in = target != kNone && target >= lo && target < hi && sorted_contains(dl, dlen, target);
        }
        hit += __popcll(__ballot(in) & gmask);
      }
      base = (tb - ta) - hit;
      elems = 0;  // the list itself is not read
    }
    for (uint32_t t0 = ta; !direct && t0 < tb; t0 += G) {
      const uint32_t t = t0 + gl;
      bool ok = t < tb;
      uint32_t x = 0;
      if (ok) x = L.nb ? dl[t] : (L.range_ids ? L.wcand[t] : t);
      for (uint32_t i = 0; i < L.nb && ok; i++) {
        if (i == d) continue;
        ok = sorted_contains(S.list(i) + sh_s[gw][i], sh_n[gw][i], x);
      }
      if (ok) ok = S.member(x);
      if (ok && L.ncf) {
        const uint32_t xid = L.any_cross ? L.wcand[x] : x;
        for (uint32_t j = 0; j < L.ncf && ok; j++) {
          const uint32_t cmp = ((L.cfsame >> j) & 1u) ? x : xid;
          ok = cmp != sh_cf[gw][j];
        }
      }
      const uint64_t bal = __ballot(ok) & gmask;
      if (WRITE) {
        const uint64_t gidx = child0 + base + __popcll(bal & below);
        const bool put = ok && gidx >= io.out_lo && gidx < io.out_hi;
        uint32_t* dst = io.out + (gidx - io.out_lo) * (L.p + 1);
        for (uint32_t c = 0; c < L.p; c++) {
          const uint32_t v = group_col<G>(rv, c);
          if (put) dst[c] = v;
        }
        if (put) dst[L.p] = x;
      }
      base += __popcll(bal);
    }
    if (!WRITE && io.cnt && gl == 0) io.cnt[r] = (uint32_t)base;
    if (gl == 0) {
      // algorithmic bytes: row words, an offset pair per backward neighbour,
      // each list element read once (none on the direct count path, where
      // only the probes of the binary searches are charged); then a u32
      // count (count pass) or a u64 offset and 4(p+1) bytes per child
      // written (write pass)
      uint64_t b = 4ull * L.p + 8ull * L.nb + 4ull * elems;
      if (direct) b += 4ull * L.ncf * (32 - __builtin_clz(dlen | 1));
      if (WRITE) {
        const uint64_t lo = child0 > io.out_lo ? child0 : io.out_lo;
        const uint64_t hi = (child0 + base) < io.out_hi ? child0 + base : io.out_hi;
        const uint64_t put = hi > lo ? hi - lo : 0;
        b += 8 + put * 4ull * (L.p + 1);
        my_rows += put;
      } else {
        b += 4;
        my_rows += 1;
      }
      my_children += base;
      my_bytes += b;
    }
  }
  my_children = wave_sum_u64(my_children);
  my_bytes = wave_sum_u64(my_bytes);
  my_rows = wave_sum_u64(my_rows);
  if (lane == 0) {
    if (!WRITE && io.total) atomicAdd(io.total, my_children);
    if (io.stats) {
      atomicAdd(&io.stats[0], my_bytes);
      atomicAdd(&io.stats[1], my_rows);
    }
  }
}

// ---------------------------------------------------------------------------
// Fused count of the last two levels (count mode, candidate-edge lists).
//
// For a parent row at level p (p columns) the children are the candidates x
// of join position p (level L1) and each child's leaves are the candidates y
// of the last position (level L2).  L2's backward lists split into the ones
// of parent columns, which all children share, and at most one of column p
// (the child's own).  The shared intersection I (minus the leaves ruled out
// by parent-column conflicts) is built once per parent in LDS; a child's leaf
// count is then |I ∩ own(x)| (or |I| minus x's own image when w is not
// adjacent to position p).  When I does not fit the tile, every child is
// counted by the group against all lists; with no shared list the count is
// |own(x)| minus the parent-column conflicts found in it.  Level p+1 is never
// written.
template <int G>
__device__ __forceinline__ uint64_t group_sum_u64(uint64_t v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
  return v;
}

struct TailLds {
  uint32_t* s1;
  uint32_t* n1;
  uint32_t* cf1;
  uint32_t* s2;
  uint32_t* n2;
  uint32_t* cf2;
  uint32_t* slot;  // G words: children counted by the group
  uint32_t* tile;  // tw words
  uint32_t tw;
};

__device__ __forceinline__ void lds_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Leaves of one parent row; returns this lane's share of the leaf count
// (sum over the group's lanes = the row's leaves) and sets `children` (group
// uniform) to the number of children x.
template <int G>
__device__ __forceinline__ uint64_t tail_row(const JoinLevel& L1, const CandEdgeSrc& S1, const JoinLevel& L2,
                                             const CandEdgeSrc& S2, const uint32_t (&rv)[(MAXQ + G - 1) / G],
                                             const TailLds& sh, uint32_t gl, uint64_t gmask, uint64_t below,
                                             uint64_t& children, bool& defer, unsigned long long& bytes) {
  const uint32_t p = L1.p;
  defer = false;
  uint32_t dlen = L1.wsize, d = 0;
  uint64_t el = 0;
  for (uint32_t i = 0; i < L1.nb; i++) {
    const uint32_t x = group_col<G>(rv, L1.bcol[i]);
    const uint32_t s = S1.off[i][x], n = S1.off[i][x + 1] - s;
    sh.s1[i] = s;
    sh.n1[i] = n;
    el += n;
    if (i == 0 || n < dlen) { dlen = n; d = i; }
  }
  for (uint32_t j = 0; j < L1.ncf; j++) {
    uint32_t x = group_col<G>(rv, L1.cfcol[j]);
    if (!((L1.cfsame >> j) & 1u)) x = L1.cfcand[j][x];
    sh.cf1[j] = x;
  }
  // leaf lists of parent columns (shared) and of column p (own)
  uint32_t ns = 0, own = kNone, ds = 0, dslen = kNone;
  for (uint32_t i = 0; i < L2.nb; i++) {
    const uint32_t c = L2.bcol[i];
    if (c == p) { own = i; continue; }
    const uint32_t x = group_col<G>(rv, c);
    const uint32_t s = S2.off[i][x], n = S2.off[i][x + 1] - s;
    sh.s2[i] = s;
    sh.n2[i] = n;
    el += n;
    if (n < dslen) { dslen = n; ds = i; }
    ns++;
  }
  // leaf conflicts with parent columns, as candidate indices of the leaf
  // class (kNone: that vertex is no candidate, nothing to exclude); ccf is the
  // conflict with column p, resolved per child
  uint32_t ccf = kNone;
  for (uint32_t j = 0; j < L2.ncf; j++) {
    const uint32_t c = L2.cfcol[j];
    if (c == p) { ccf = j; sh.cf2[j] = kNone; continue; }
    uint32_t t = group_col<G>(rv, c);
    if (!((L2.cfsame >> j) & 1u)) t = cand_index(L2, L2.cfcand[j][t]);
    sh.cf2[j] = t;
  }
  uint64_t lbytes = gl == 0 ? 4ull * p + 8ull * (L1.nb + ns) + 4ull * el : 0;
  const uint32_t* dl = L1.nb ? S1.lst[d] + sh.s1[d] : nullptr;
  // symmetry conditions: children x in [lo1, hi1) (driver positions [ta1,
  // tb1)); leaves y in [lo2, hi2), and above (g2p) or below (l2p) their own
  // x
  const uint32_t pbit = 1u << p;
  const bool g2p = (L2.gtmask & pbit) != 0, l2p = (L2.ltmask & pbit) != 0;
  uint32_t lo1 = 0, hi1 = kNone, lo2 = 0, hi2 = kNone, ta1 = 0, tb1 = dlen;
  if (L1.gtmask | L1.ltmask) {
    sym_bounds<G>(L1.gtmask, L1.ltmask, rv, lo1, hi1);
    driver_window(dl, nullptr, false, dlen, lo1, hi1, ta1, tb1);
  }
  if ((L2.gtmask | L2.ltmask) & ~pbit) sym_bounds<G>(L2.gtmask & ~pbit, L2.ltmask & ~pbit, rv, lo2, hi2);
  const bool leaf_bounds = lo2 != 0 || hi2 != kNone;
  if (own == kNone) {
    // No own list: every child has the same leaves I except its own image.
    // |I| once per parent (one shared list: its length minus the conflicts
    // found in it; more: a counting walk), then one membership test per child.
    uint64_t isz = 0;
    const uint32_t* dl2 = S2.lst[ds] + sh.s2[ds];
    uint32_t a2 = 0, b2 = dslen;
    if (leaf_bounds) driver_window(dl2, nullptr, false, dslen, lo2, hi2, a2, b2);
    if (ns == 1) {
      uint32_t hit = 0;
      for (uint32_t j0 = 0; j0 < L2.ncf; j0 += G) {
        const uint32_t j = j0 + gl;
        const bool in = j < L2.ncf && sh.cf2[j] != kNone && sh.cf2[j] >= lo2 && sh.cf2[j] < hi2 &&
                        sorted_contains(dl2, dslen, sh.cf2[j]);
        hit += (uint32_t)__popcll(__ballot(in) & gmask);
      }
      isz = (b2 - a2) - hit;
      if (gl == 0) lbytes -= 4ull * dslen;  // the list itself is not read
    } else {
      for (uint32_t t0 = a2; t0 < b2; t0 += G) {
        const uint32_t t = t0 + gl;
        bool ok = t < b2;
        const uint32_t y = ok ? dl2[t] : 0u;
        for (uint32_t i = 0; i < L2.nb && ok; i++) {
          if (i == ds) continue;
          ok = sorted_contains(S2.lst[i] + sh.s2[i], sh.n2[i], y);
        }
        for (uint32_t j = 0; j < L2.ncf && ok; j++) ok = y != sh.cf2[j];
        isz += (uint32_t)__popcll(__ballot(ok) & gmask);
      }
    }
    if (ccf == kNone && L1.nb == 1) {
      // Every child has exactly |I| leaves and the children are the one
      // list minus the parent-column conflicts found in it: count those
      // (lanes take one conflict each) instead of walking the children.
      uint32_t hit = 0;
      for (uint32_t j0 = 0; j0 < L1.ncf; j0 += G) {
        const uint32_t j = j0 + gl;
        bool in = false;
        if (j < L1.ncf) {
          uint32_t x = sh.cf1[j];
          if (!((L1.cfsame >> j) & 1u)) x = cand_index(L1, x);  // a data vertex id
          in = x != kNone && x >= lo1 && x < hi1 && sorted_contains(dl, dlen, x);
        }
        hit += (uint32_t)__popcll(__ballot(in) & gmask);
      }
      children = (tb1 - ta1) - hit;
      if (gl == 0) lbytes += 4ull * L1.ncf * (2 + (32 - __builtin_clz(dlen | 1))) - 4ull * dlen;  // list not read
      bytes += lbytes;
      return gl == 0 ? children * isz : 0;
    }
    uint64_t kids = 0, leaves = 0;
    for (uint32_t t0 = ta1; t0 < tb1; t0 += G) {
      const uint32_t t = t0 + gl;
      bool ok = t < tb1;
      uint32_t x = 0;
      if (ok) x = L1.nb ? dl[t] : t;
      for (uint32_t i = 0; i < L1.nb && ok; i++) {
        if (i == d) continue;
        ok = sorted_contains(S1.lst[i] + sh.s1[i], sh.n1[i], x);
      }
      if (ok && L1.ncf) {
        const uint32_t xid = L1.any_cross ? L1.wcand[x] : x;
        for (uint32_t j = 0; j < L1.ncf && ok; j++) ok = (((L1.cfsame >> j) & 1u) ? x : xid) != sh.cf1[j];
      }
      kids += __popcll(__ballot(ok) & gmask);
      if (ok) {
        uint64_t cnt = isz;
        if (ccf != kNone) {  // x's image is a leaf candidate unless some shared list lacks it
          uint32_t cx = x;
          if (!((L2.cfsame >> ccf) & 1u)) cx = cand_index(L2, L2.cfcand[ccf][x]);
          bool in = cx != kNone && cx >= lo2 && cx < hi2;
          for (uint32_t i = 0; i < L2.nb && in; i++) in = sorted_contains(S2.lst[i] + sh.s2[i], sh.n2[i], cx);
          if (in) cnt--;
          lbytes += 4ull * ns * (32 - __builtin_clz(dslen | 1));
        }
        leaves += cnt;
      }
    }
    children = kids;
    bytes += lbytes;
    return leaves;
  }
  // A row whose shared intersection does not fit the tile, or that may have
  // very many children (> 16G) while I is large (> 4G), is deferred: its
  // children are counted here and written by the caller, which counts their
  // leaves one row per group.  Other rows stage I in LDS once and count each
  // child's leaves cooperatively: lanes split the shorter of I and the
  // child's own list and search the other.
  bool defer_row = false;
  uint32_t isz = 0;
  if (ns > 0) {
    const uint32_t* dl2 = S2.lst[ds] + sh.s2[ds];
    uint32_t a2 = 0, b2 = dslen;
    if (leaf_bounds) driver_window(dl2, nullptr, false, dslen, lo2, hi2, a2, b2);
    for (uint32_t t0 = a2; t0 < b2; t0 += G) {
      const uint32_t t = t0 + gl;
      bool ok = t < b2;
      const uint32_t y = ok ? dl2[t] : 0u;
      for (uint32_t i = 0; i < L2.nb && ok; i++) {
        if (i == ds || i == own) continue;
        ok = sorted_contains(S2.lst[i] + sh.s2[i], sh.n2[i], y);
      }
      for (uint32_t j = 0; j < L2.ncf && ok; j++) ok = y != sh.cf2[j];
      const uint64_t bal = __ballot(ok) & gmask;
      const uint32_t nb = (uint32_t)__popcll(bal);
      if (isz + nb > sh.tw) { defer_row = true; break; }  // I does not fit
      if (ok) sh.tile[isz + __popcll(bal & below)] = y;
      isz += nb;
    }
    lds_wave_sync();
    // many children, each with a large I to intersect: spread them instead
    defer_row = defer_row || (tb1 - ta1 > 16u * G && isz > 4u * G);
  }
  uint64_t kids = 0, leaves = 0;
  for (uint32_t t0 = ta1; t0 < tb1; t0 += G) {
    const uint32_t t = t0 + gl;
    bool ok = t < tb1;
    uint32_t x = 0;
    if (ok) x = L1.nb ? dl[t] : t;
    for (uint32_t i = 0; i < L1.nb && ok; i++) {
      if (i == d) continue;
      ok = sorted_contains(S1.lst[i] + sh.s1[i], sh.n1[i], x);
    }
    if (ok && L1.ncf) {
      const uint32_t xid = L1.any_cross ? L1.wcand[x] : x;
      for (uint32_t j = 0; j < L1.ncf && ok; j++) ok = (((L1.cfsame >> j) & 1u) ? x : xid) != sh.cf1[j];
    }
    const uint64_t bal = __ballot(ok) & gmask;
    kids += __popcll(bal);
    if (defer_row) continue;
    if (ns == 0) {
      // own list only: its length minus the parent-column conflicts in it
      if (ok) {
        const uint32_t s = S2.off[own][x], n = S2.off[own][x + 1] - s;
        const uint32_t* l = S2.lst[own] + s;
        uint32_t lox = lo2, hix = hi2, a = 0, b = n;
        if (g2p) lox = max(lox, x + 1);
        if (l2p) hix = min(hix, x);
        if (lox != 0 || hix != kNone) driver_window(l, nullptr, false, n, lox, hix, a, b);
        uint64_t cnt = b - a;
        for (uint32_t j = 0; j < L2.ncf; j++)
          if (sh.cf2[j] != kNone && sh.cf2[j] >= lox && sh.cf2[j] < hix && sorted_contains(l, n, sh.cf2[j])) cnt--;
        leaves += cnt;
        lbytes += 8 + 4ull * L2.ncf * (32 - __builtin_clz(n | 1));
      }
      continue;
    }
    if (ok) sh.slot[__popcll(bal & below)] = x;
    lds_wave_sync();
    const uint32_t nsurv = (uint32_t)__popcll(bal);
    for (uint32_t q = 0; q < nsurv; q++) {
      const uint32_t xx = sh.slot[q];
      const uint32_t s = S2.off[own][xx], n = S2.off[own][xx + 1] - s;
      const uint32_t* l = S2.lst[own] + s;
      // leaves above (g2p) or below (l2p) the child itself
      const uint32_t lox = g2p ? xx + 1 : 0, hix = l2p ? xx : kNone;
      uint32_t got = 0;
      if (isz <= n) {
        for (uint32_t e0 = 0; e0 < isz; e0 += G) {
          const uint32_t e = e0 + gl;
          const bool in = e < isz && sh.tile[e] >= lox && sh.tile[e] < hix && sorted_contains(l, n, sh.tile[e]);
          got += (uint32_t)__popcll(__ballot(in) & gmask);
        }
        if (gl == 0) lbytes += 8 + 4ull * isz * (32 - __builtin_clz(n | 1));
      } else {
        for (uint32_t e0 = 0; e0 < n; e0 += G) {
          const uint32_t e = e0 + gl;
          const bool in = e < n && l[e] >= lox && l[e] < hix && sorted_contains(sh.tile, isz, l[e]);
          got += (uint32_t)__popcll(__ballot(in) & gmask);
        }
        if (gl == 0) lbytes += 8 + 4ull * n;
      }
      if (gl == 0) leaves += got;
    }
    lds_wave_sync();
  }
  defer = defer_row;
  children = kids;
  bytes += lbytes;
  return leaves;
}

// Per-group LDS of the fused tail: list ranges and conflicts of both
// levels (at most kTailMax entries each) and the shared-intersection tile,
// which lives in dynamic LDS sized per launch (0 when no row of the launch
// needs it).
constexpr uint32_t kTailMax = 16;
// Levels whose lists and conflicts all fit this many entries take a tail
// with arrays of that size: a 64-group block then needs 13 KiB of LDS
// instead of 25 KiB, so LDS no longer caps the waves per CU.
constexpr uint32_t kTailSmall = 8;

// N1: entries of the level-p arrays (MAXQ when a lockstep launch also keeps
// its plain rows there; the plain tail's levels are limited to kTailMax
// backward lists and conflicts: less LDS per group, more groups per CU);
// N2: entries of the leaf-level arrays.
template <int G, int N1 = MAXQ, int N2 = kTailMax>
struct TailShared {
  static constexpr int GPB = FSIG_BLOCK / G;
  static constexpr uint32_t kTileWords = 4096 / GPB;  // at most 16 KiB of tiles per block
  uint32_t s1[GPB][N1], n1[GPB][N1], cf1[GPB][N1];
  uint32_t s2[GPB][N2], n2[GPB][N2], cf2[GPB][N2];
  uint32_t slot[GPB][G];
  __device__ TailLds at(uint32_t gw, uint32_t* tile, uint32_t tw) {
    return TailLds{s1[gw], n1[gw], cf1[gw], s2[gw], n2[gw], cf2[gw], slot[gw], tile + gw * tw, tw};
  }
};

struct TailIO {
  const uint32_t* rows;
  uint64_t nrows;
  uint32_t* cnt;              // per row: children of a deferred row, else 0
  unsigned long long* res;    // [0] += leaves, [1] += children of counted rows
  unsigned long long* stats;  // [0] += algorithmic bytes, [1] += parent rows
  uint32_t tile_words;        // shared-intersection tile per group (dynamic LDS)
};

// N: entries of the per-group list and conflict arrays of both levels.  At
// most 72 VGPRs, so seven waves per SIMD fit where LDS allows (the tail is
// bound by dependent gathers; eight waves need spills and run slower); 64
// four-lane groups with 16-entry arrays fit six per SIMD by their LDS.
template <int G, int N>
__global__ void __launch_bounds__(FSIG_BLOCK)
    __attribute__((amdgpu_waves_per_eu(G == 4 && N > (int)kTailSmall ? 6 : 7, 8)))
k_join_tail(JoinLevel L1, CandEdgeSrc S1, JoinLevel L2, CandEdgeSrc S2, TailIO io) {
  constexpr int GPB = FSIG_BLOCK / G;
  constexpr int NR = (MAXQ + G - 1) / G;
  __shared__ TailShared<G, N, N> sh;
  extern __shared__ uint32_t tail_tile[];
  const uint32_t lane = threadIdx.x & 63u, gl = lane & (G - 1), gw = threadIdx.x / G;
  const uint64_t gmask = (G == 64) ? ~0ull : (((1ull << G) - 1) << (lane - gl));
  const uint64_t below = (1ull << lane) - 1;
  const uint64_t ngroups = (uint64_t)gridDim.x * GPB;
  const TailLds t = sh.at(gw, tail_tile, io.tile_words);
  unsigned long long my_leaves = 0, my_kids = 0, my_bytes = 0, my_rows = 0;
  for (uint64_t r = (uint64_t)xcd_block() * GPB + gw; r < io.nrows; r += ngroups) {
    uint32_t rv[NR];
    const uint32_t* row = io.rows + r * L1.p;
#pragma unroll
    for (int j = 0; j < NR; j++) {
      const uint32_t col = j * G + gl;
      rv[j] = col < L1.p ? row[col] : 0u;
    }
    uint64_t kids = 0;
    bool defer = false;
    my_leaves += tail_row<G>(L1, S1, L2, S2, rv, t, gl, gmask, below, kids, defer, my_bytes);
    if (gl == 0) {
      io.cnt[r] = defer ? (uint32_t)kids : 0u;
      if (!defer) my_kids += kids;
      my_rows += 1;
    }
  }
  my_leaves = wave_sum_u64(my_leaves);
  my_kids = wave_sum_u64(my_kids);
  my_bytes = wave_sum_u64(my_bytes);
  my_rows = wave_sum_u64(my_rows);
  if (lane == 0) {
    atomicAdd(&io.res[0], my_leaves);
    atomicAdd(&io.res[1], my_kids);
    if (io.stats) {
      atomicAdd(&io.stats[0], my_bytes);
      atomicAdd(&io.stats[1], my_rows);
    }
  }
}

// Whether the fused tail applies to leaf level L2: it needs a backward list,
// its lists and conflicts must fit the group arrays, and a symmetry
// condition between leaf and child needs the leaf adjacent to the child.
inline bool tail_applies(const JoinLevel& L2) {
  if (!(L2.nb >= 1 && L2.nb <= kTailMax && L2.ncf <= kTailMax)) return false;
  // a symmetry condition between the leaf and the child needs the child's
  // own list to bound
  const uint32_t pbit = 1u << (L2.p - 1);
  if (!((L2.gtmask | L2.ltmask) & pbit)) return true;
  for (uint32_t i = 0; i < L2.nb; i++)
    if (L2.bcol[i] == L2.p - 1) return true;
  return false;
}

// Tile words per group the tail of L2 needs: only a leaf with both shared
// lists (parent columns) and an own list (column p) intersects in LDS.
inline uint32_t tail_tile_words(uint32_t G, const JoinLevel& L2, uint32_t cap) {
  bool shared = false, own = false;
  for (uint32_t i = 0; i < L2.nb; i++) (L2.bcol[i] == L2.p - 1 ? own : shared) = true;
  if (!shared || !own) return 0;
  const uint32_t w = 4096 / (FSIG_BLOCK / G);
  return cap && cap < w ? cap : w;
}

inline int launch_join_tail(fsig_ctx* c, uint32_t G, const JoinLevel& L1, const CandEdgeSrc& S1,
                            const JoinLevel& L2, const CandEdgeSrc& S2, const TailIO& io) {
  if (io.nrows == 0) return FSIG_OK;
  const uint64_t gpb = FSIG_BLOCK / G;
  uint64_t blocks = std::min<uint64_t>((io.nrows + gpb - 1) / gpb, (uint64_t)c->num_cus * 16);
  if (blocks >= 64) blocks &= ~7ull;
  const uint32_t grid = (uint32_t)std::max<uint64_t>(blocks, 1);
  const size_t lds = (size_t)io.tile_words * gpb * sizeof(uint32_t);
  const bool small = L1.nb <= kTailSmall && L1.ncf <= kTailSmall && L2.nb <= kTailSmall && L2.ncf <= kTailSmall;
#define FSIG_TAIL(GV)                                                                       \
  (small ? k_join_tail<GV, kTailSmall><<<grid, FSIG_BLOCK, lds, c->stream>>>(L1, S1, L2, S2, io) \
         : k_join_tail<GV, kTailMax><<<grid, FSIG_BLOCK, lds, c->stream>>>(L1, S1, L2, S2, io))
  switch (G) {
    case 4: FSIG_TAIL(4); break;
    case 8: FSIG_TAIL(8); break;
    case 16: FSIG_TAIL(16); break;
    case 32: FSIG_TAIL(32); break;
    default: FSIG_TAIL(64); break;
  }
#undef FSIG_TAIL
  FSIG_HIP(c, hipGetLastError());
  return FSIG_OK;
}

// ---------------------------------------------------------------------------
// Chain tail (count mode).  When the leaf w (position p+1) has one backward
// neighbour, the previous position p, and position p has one backward list
// (of parent column c1), a parent row's children are C1 = list(phi(c1))
// minus X1 (the parent columns of p's label that occur in the list), and a
// child x has |own(x)| leaves minus the parent columns of w's label in
// own(x) (the conflicts y_j).  Summed over the children:
//
//   leaves = S[phi(c1)] - sum_{x in X1} (|own(x)| - #{j : y_j in own(x)})
//                       - sum_j |list(phi(c1)) ∩ rev(y_j)|
//
// where S[v] = sum of |own(x)| over list(v) is computed once per match
// (k_chain_sum) and rev(y) is the list of p's candidates adjacent to leaf
// candidate y (the candidate-edge list of the reverse class pair): a row
// costs a few gathers and one intersection per conflict instead of a walk
// over its children.  Paths and other tree-shaped queries end this way.
struct ChainTail {
  const unsigned long long* sum;  // S[v] for every candidate v of c1's class
  const uint32_t* rev_off;        // reverse pair (w's class -> p's class)
  const uint32_t* rev_lst;
};

// S[v] = sum over x in list(v) (pair c1 -> p) of |own(x)| (pair p -> w); G
// lanes per candidate v.
template <int G>
__global__ void __launch_bounds__(FSIG_BLOCK) k_chain_sum(const uint32_t* __restrict__ off1,
                                                          const uint32_t* __restrict__ lst1,
                                                          const uint32_t* __restrict__ own_off, uint32_t nv,
                                                          unsigned long long* __restrict__ sum) {
  constexpr int GPB = FSIG_BLOCK / G;
  const uint32_t gl = threadIdx.x & (G - 1), gw = threadIdx.x / G;
  for (uint64_t v = (uint64_t)xcd_block() * GPB + gw; v < nv; v += (uint64_t)gridDim.x * GPB) {
    unsigned long long s = 0;
    for (uint32_t t = off1[v] + gl; t < off1[v + 1]; t += G) {
      const uint32_t x = lst1[t];
      s += own_off[x + 1] - own_off[x];
    }
    s = group_sum_u64<G>(s);
    if (gl == 0) sum[v] = s;
  }
}

template <int G>
__global__ void __launch_bounds__(FSIG_BLOCK) k_join_tail_chain(JoinLevel L1, CandEdgeSrc S1, JoinLevel L2,
                                                                CandEdgeSrc S2, ChainTail T, TailIO io) {
  constexpr int GPB = FSIG_BLOCK / G;
  constexpr int NR = (MAXQ + G - 1) / G;
  const uint32_t lane = threadIdx.x & 63u, gl = lane & (G - 1), gw = threadIdx.x / G;
  const uint64_t gmask = (G == 64) ? ~0ull : (((1ull << G) - 1) << (lane - gl));
  const uint64_t ngroups = (uint64_t)gridDim.x * GPB;
  const uint32_t p = L1.p;
  const uint32_t* own_off = S2.off[0];
  const uint32_t* own_lst = S2.lst[0];
  __shared__ uint32_t sh_y[GPB][kTailMax];
  __shared__ uint32_t sh_x1[GPB][kTailMax];
  uint32_t* y = sh_y[gw];
  uint32_t* x1 = sh_x1[gw];
  unsigned long long my_leaves = 0, my_kids = 0, my_bytes = 0, my_rows = 0;
  for (uint64_t r = (uint64_t)xcd_block() * GPB + gw; r < io.nrows; r += ngroups) {
    uint32_t rv[NR];
    const uint32_t* row = io.rows + r * p;
#pragma unroll
    for (int j = 0; j < NR; j++) {
      const uint32_t col = j * G + gl;
      rv[j] = col < p ? row[col] : 0u;
    }
    const uint32_t v = group_col<G>(rv, L1.bcol[0]);
    const uint32_t ls = S1.off[0][v], ln = S1.off[0][v + 1] - ls;
    const uint32_t* lst = S1.lst[0] + ls;
    int64_t leaves = (int64_t)T.sum[v];
    uint64_t kids = ln;
    uint64_t bytes = 4ull * p + 16;
    // conflicts of the leaf: parent columns of w's label, as candidate
    // indices of w's class (kNone: not a candidate, nothing to exclude);
    // every lane of the group computes the same values
    lds_wave_sync();  // the previous row's reads of y are done
    for (uint32_t j = 0; j < L2.ncf; j++) {
      uint32_t t = group_col<G>(rv, L2.cfcol[j]);
      if (!((L2.cfsame >> j) & 1u)) {
        t = cand_index(L2, L2.cfcand[j][t]);
        bytes += 12;
      }
      if (gl == 0) y[j] = t;
    }
    // X1 candidates: parent columns of p's label as candidate indices of
    // p's class (a cross-lane read of the row, so in a loop every lane runs)
    for (uint32_t j = 0; j < L1.ncf; j++) {
      uint32_t x = group_col<G>(rv, L1.cfcol[j]);
      if (!((L1.cfsame >> j) & 1u)) x = cand_index(L1, L1.cfcand[j][x]);  // a data vertex id
      if (gl == 0) x1[j] = x;
    }
    lds_wave_sync();
    // X1: parent columns of p's label found in the list; each removes a
    // child and its leaves (lanes take one conflict each)
    for (uint32_t j0 = 0; j0 < L1.ncf; j0 += G) {
      const uint32_t j = j0 + gl;
      int64_t sub = 0;
      bool in = false;
      uint32_t x = kNone;
      if (j < L1.ncf) {
        x = x1[j];
        in = x != kNone && sorted_contains(lst, ln, x);
        if (in) {
          const uint32_t os = own_off[x], on = own_off[x + 1] - os;
          sub = on;
          for (uint32_t k2 = 0; k2 < L2.ncf; k2++)
            if (y[k2] != kNone && sorted_contains(own_lst + os, on, y[k2])) sub--;
        }
      }
      kids -= (uint64_t)__popcll(__ballot(in) & gmask);
      leaves -= (int64_t)group_sum_u64<G>((uint64_t)sub);
      if (gl == 0) bytes += 4ull * (G < L1.ncf - j0 ? G : L1.ncf - j0) * (2 + (32 - __builtin_clz(ln | 1)));
    }
    // children that hold a conflict in their own list: |list ∩ rev(y_j)|,
    // the shorter list driving and the other searched
    for (uint32_t j = 0; j < L2.ncf; j++) {
      if (y[j] == kNone) continue;
      const uint32_t rs = T.rev_off[y[j]], rn = T.rev_off[y[j] + 1] - rs;
      const uint32_t* rl = T.rev_lst + rs;
      const bool drive_list = ln <= rn;
      const uint32_t* a = drive_list ? lst : rl;
      const uint32_t* b = drive_list ? rl : lst;
      const uint32_t an = drive_list ? ln : rn, bn = drive_list ? rn : ln;
      uint32_t hit = 0;
      for (uint32_t t0 = 0; t0 < an; t0 += G) {
        const uint32_t t = t0 + gl;
        const bool in = t < an && sorted_contains(b, bn, a[t]);
        hit += (uint32_t)__popcll(__ballot(in) & gmask);
      }
      leaves -= hit;
      bytes += 8ull + 4ull * an * (1 + (32 - __builtin_clz(bn | 1)));
    }
    if (gl == 0) {
      io.cnt[r] = 0;
      my_leaves += (unsigned long long)leaves;
      my_kids += kids;
      my_bytes += bytes;
      my_rows += 1;
    }
  }
  my_leaves = wave_sum_u64(my_leaves);
  my_kids = wave_sum_u64(my_kids);
  my_bytes = wave_sum_u64(my_bytes);
  my_rows = wave_sum_u64(my_rows);
  if (lane == 0) {
    atomicAdd(&io.res[0], my_leaves);
    atomicAdd(&io.res[1], my_kids);
    if (io.stats) {
      atomicAdd(&io.stats[0], my_bytes);
      atomicAdd(&io.stats[1], my_rows);
    }
  }
}

// Whether the chain tail applies to levels p (L1) and p+1 (L2).
inline bool chain_tail_applies(const JoinLevel& L1, const JoinLevel& L2) {
  return L2.nb == 1 && L2.bcol[0] == L1.p && L1.nb == 1 && L1.ncf <= kTailMax && L2.ncf <= kTailMax &&
         !(L1.gtmask | L1.ltmask | L2.gtmask | L2.ltmask);  // its sums cover every child and leaf
}

inline int launch_join_tail_chain(fsig_ctx* c, uint32_t G, const JoinLevel& L1, const CandEdgeSrc& S1,
                                  const JoinLevel& L2, const CandEdgeSrc& S2, const ChainTail& T,
                                  const TailIO& io) {
  if (io.nrows == 0) return FSIG_OK;
  const uint64_t gpb = FSIG_BLOCK / G;
  uint64_t blocks = std::min<uint64_t>((io.nrows + gpb - 1) / gpb, (uint64_t)c->num_cus * 16);
  if (blocks >= 64) blocks &= ~7ull;
  const uint32_t grid = (uint32_t)std::max<uint64_t>(blocks, 1);
  switch (G) {
    case 4: k_join_tail_chain<4><<<grid, FSIG_BLOCK, 0, c->stream>>>(L1, S1, L2, S2, T, io); break;
    case 8: k_join_tail_chain<8><<<grid, FSIG_BLOCK, 0, c->stream>>>(L1, S1, L2, S2, T, io); break;
    case 16: k_join_tail_chain<16><<<grid, FSIG_BLOCK, 0, c->stream>>>(L1, S1, L2, S2, T, io); break;
    default: k_join_tail_chain<32><<<grid, FSIG_BLOCK, 0, c->stream>>>(L1, S1, L2, S2, T, io); break;
  }
  FSIG_HIP(c, hipGetLastError());
  return FSIG_OK;
}

// Parent rows of the child window [lo, hi): out[0] = last r with off[r] <= lo,
// out[1] = first r with off[r] >= hi (off has nrows+1 entries).
static __global__ void k_window_rows(const uint64_t* __restrict__ off, uint64_t nrows, uint64_t lo, uint64_t hi,
                                     unsigned long long* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint64_t a = 0, b = nrows + 1;
  while (a < b) {
    const uint64_t m = (a + b) >> 1;
    if (off[m] <= lo) a = m + 1; else b = m;
  }
  out[0] = a - 1;
  a = 0;
  b = nrows + 1;
  while (a < b) {
    const uint64_t m = (a + b) >> 1;
    if (off[m] < hi) a = m + 1; else b = m;
  }
  out[1] = a > nrows ? nrows : a;
}

// Host launcher for the join kernel, dispatching on G.
template <bool WRITE, class Src>
int launch_join(fsig_ctx* c, uint32_t G, const JoinLevel& L, const Src& S, const JoinIO& io) {
  const uint64_t nrows = io.row_hi - io.row_lo;
  if (nrows == 0) return FSIG_OK;
  const uint64_t gpb = FSIG_BLOCK / G;
  uint64_t blocks = (nrows + gpb - 1) / gpb;
  const uint64_t max_blocks = (uint64_t)c->num_cus * 16;  // grid-stride beyond this
  blocks = std::min(blocks, max_blocks);
  if (blocks >= 64) blocks &= ~7ull;  // whole multiples of 8 XCDs
  const uint32_t grid = (uint32_t)std::max<uint64_t>(blocks, 1);
  switch (G) {
    case 4: k_join<4, WRITE, Src><<<grid, FSIG_BLOCK, 0, c->stream>>>(L, S, io); break;
    case 8: k_join<8, WRITE, Src><<<grid, FSIG_BLOCK, 0, c->stream>>>(L, S, io); break;
    case 16: k_join<16, WRITE, Src><<<grid, FSIG_BLOCK, 0, c->stream>>>(L, S, io); break;
    case 32: k_join<32, WRITE, Src><<<grid, FSIG_BLOCK, 0, c->stream>>>(L, S, io); break;
    default: k_join<64, WRITE, Src><<<grid, FSIG_BLOCK, 0, c->stream>>>(L, S, io); break;
  }
  FSIG_HIP(c, hipGetLastError());
  return FSIG_OK;
}

// Receives completed embeddings (level k tables, join-order words).
struct LeafSink {
  virtual ~LeafSink() = default;
  virtual int take(const uint32_t* rows, uint64_t n) = 0;
};

// Lanes per row of a launch over R parent rows: G, doubled while the launch
// would leave most of the device's lanes idle (few rows, each a long walk
// whose children and intersections the lanes split), up to Gmax.  A tuned
// lane count is kept as it is.
inline uint32_t widen_lanes(const fsig_ctx* c, uint32_t G, uint64_t R, uint32_t Gmax) {
  if (c->tune.lanes_per_row) return G;
  const uint64_t lanes = (uint64_t)c->num_cus * 32 * 64;  // 32 resident waves of 64 per CU
  while (G < Gmax && R * G * 2 <= lanes) G <<= 1;
  return G;
}

// Provider: supplies the level description and list source for each level:
//   using SrcType = ...; JoinLevel level(uint32_t p); SrcType src(uint32_t p);
template <class Provider>
class JoinDriver {
 public:
  using Src = typename Provider::SrcType;
  JoinDriver(fsig_ctx* c, Provider& pv, uint32_t k, uint32_t G, int mode, uint64_t max_emb, LeafSink* sink)
      : c_(c), pv_(pv), k_(k), G_(G), mode_(mode), max_emb_(max_emb), sink_(sink) {
    for (uint32_t p = 0; p <= MAXQ; p++) chunk_[p] = window_[p] = kGrow0;
  }

  uint64_t total() const { return total_; }

  int run(const uint32_t* level1, uint64_t n1) { return expand(1, level1, n1); }
  // Continues from R level-p rows (p words each) produced elsewhere.
  int run_at(uint32_t p, const uint32_t* rows, uint64_t R) { return expand(p, rows, R); }

 private:
  bool bounded() const { return max_emb_ != 0; }
  uint64_t remaining() const { return bounded() ? max_emb_ - total_ : ~0ull; }

  // A bounded match takes the parent rows of a level in chunks, and the
  // children of a chunk in windows, whose sizes start at kGrow0 rows and
  // double each time the level takes one (for the whole match, so a level
  // reached from many windows above does not start small every time); it
  // stops as soon as the bound is reached, so most children of a huge query
  // are never counted or written.  An unbounded match takes every row at once.
  static constexpr uint64_t kGrow0 = 1ull << 9;
  static constexpr uint64_t kGrowMax = 1ull << 28;
  uint64_t chunk_[MAXQ + 1];
  uint64_t window_[MAXQ + 1];
  // Embeddings per level-p row over the last chunk of that level that found
  // any (0: none seen yet).  Near the bound, a chunk or window takes 1.25
  // times the rows the remaining bound needs at that yield, not the doubled
  // size: the rows beyond would be counted or written only to be cut off.
  double yield_[MAXQ + 1] = {};

  uint64_t need_rows(uint32_t p) const {
    if (yield_[p] <= 0) return ~0ull;
    const double r = 1.25 * (double)remaining() / yield_[p];
    return r >= (double)kGrowMax ? kGrowMax : std::max<uint64_t>(kGrow0, (uint64_t)r + 1);
  }

  int expand(uint32_t p, const uint32_t* rows, uint64_t R) {
    if (R == 0) return FSIG_OK;
    if (p == k_ || !bounded()) return expand_rows(p, rows, R);
    for (uint64_t r0 = 0; r0 < R && remaining() > 0;) {
      const uint64_t n = std::min(std::min(chunk_[p], R - r0), need_rows(p));
      chunk_[p] = std::min(chunk_[p] * 2, kGrowMax);
      const uint64_t before = total_;
      const int rc = expand_rows(p, rows + r0 * p, n);
      if (rc) return rc;
      if (total_ > before) yield_[p] = (double)(total_ - before) / (double)n;
      r0 += n;
    }
    return FSIG_OK;
  }

  int expand_rows(uint32_t p, const uint32_t* rows, uint64_t R) {
    if (R == 0) return FSIG_OK;
    if (p == k_) {
      uint64_t n = std::min<uint64_t>(R, remaining());
      if (sink_ && n) {
        int rc = sink_->take(rows, n);
        if (rc) return rc;
      }
      total_ += n;
      return FSIG_OK;
    }
    const size_t mk = c_->arena.mark();
    const JoinLevel L = pv_.level(p);
    const Src S = pv_.src(p);
    const bool last = p + 1 == k_;
    unsigned long long* res = c_->d_counters;  // [0] children of this launch
    int rc;
    bool fused = false;
    if constexpr (std::is_same<Src, CandEdgeSrc>::value)
      // (under a work limit the tail's children count after its pass)
      fused = mode_ == FSIG_MODE_COUNT && p + 2 == k_ && c_->tune.fused_tail == 0 && tail_applies(pv_.level(p + 1)) &&
              L.nb <= kTailMax && L.ncf <= kTailMax;
    if (last && mode_ == FSIG_MODE_COUNT) {
      FSIG_HIP(c_, hipMemsetAsync(res, 0, sizeof(unsigned long long), c_->stream));
      JoinIO io{rows, 0, R, nullptr, nullptr, nullptr, 0, 0, res, family_counters(c_, FSIG_K_JOIN_COUNT)};
      KTimer t(c_, FSIG_K_JOIN_COUNT);
      if ((rc = launch_join<false>(c_, widen_lanes(c_, G_, R, 64), L, S, io))) return rc;
      t.done(0, 0);
      uint64_t h;
      if ((rc = read_u64(c_, res, &h, 1))) return rc;
      total_ += std::min(h, remaining());
      return FSIG_OK;
    }
    uint32_t* cnt = (uint32_t*)c_->arena.alloc(4 * R);
    uint64_t* off = (uint64_t*)c_->arena.alloc(8 * (R + 1));
    if (!cnt || !off) return set_err(c_, FSIG_ENOMEM, "join: arena exhausted for row counts");
    if constexpr (std::is_same<Src, CandEdgeSrc>::value) {
      if (fused) {
        // rows whose leaves the tail counts get 0 children; deferred rows
        // keep theirs and go through the write pass and the last level
        if ((rc = count_tail(rows, R, cnt, L, S, pv_.level(p + 1), pv_.src(p + 1)))) return rc;
      }
    }
    if (!fused) {
      JoinIO io{rows, 0, R, cnt, nullptr, nullptr, 0, 0, nullptr, family_counters(c_, FSIG_K_JOIN_COUNT)};
      KTimer t(c_, FSIG_K_JOIN_COUNT);
      if ((rc = launch_join<false>(c_, widen_lanes(c_, G_, R, 64), L, S, io))) return rc;
      t.done(0, 0);
    }
    if ((rc = scan_u32_to_u64(c_, cnt, off, R))) return rc;
    uint64_t T;
    if ((rc = read_u64(c_, off + R, &T, 1))) return rc;
    if (T == 0) { c_->arena.release(mk); return FSIG_OK; }
    if (last) T = std::min(T, remaining());
    const uint64_t limit = c_->work_limit;
    // Children per window: at most half of the free arena, each child
    // costing its row plus the count/offset words it needs when expanded in
    // turn.  Unbounded: balanced windows, all children counted as produced
    // now.  Bounded: windows growing per level from kGrow0, each counted as
    // it is produced.
    const uint64_t per_child = 4ull * (p + 1) + 16;
    uint64_t cap = c_->arena.avail() / 2 / per_child;
    if (cap == 0) return set_err(c_, FSIG_ENOMEM, "join: arena exhausted for child rows");
    const bool grow = bounded();
    uint64_t wsize;
    if (grow) {
      wsize = std::min(std::min(cap, window_[p]), need_rows(p + 1));
    } else {
      if (limit && !last && rows_ + T > limit) return set_err(c_, FSIG_ELIMIT, "join: work limit exceeded");
      rows_ += last ? 0 : T;
      const uint64_t nwin = (T + cap - 1) / cap;
      wsize = (T + nwin - 1) / nwin;  // balanced windows
    }
    const bool split = T > wsize;
    unsigned long long* dwin = split ? (unsigned long long*)c_->arena.alloc(16) : nullptr;
    if (split && !dwin) return set_err(c_, FSIG_ENOMEM, "join: arena exhausted for windows");
    for (uint64_t lo = 0; lo < T && remaining() > 0;) {
      uint64_t hi = std::min(T, lo + wsize);
      if (grow) {
        // the limit is exact: a window is clipped to the rows the limit
        // leaves (the bound may be reached inside it), and the match stops
        // only when it needs a row beyond the limit
        if (limit && !last) {
          if (rows_ >= limit) return set_err(c_, FSIG_ELIMIT, "join: work limit exceeded");
          hi = std::min(hi, lo + (limit - rows_));
        }
        rows_ += last ? 0 : hi - lo;
        window_[p] = std::min(window_[p] * 2, kGrowMax);
        wsize = std::min(std::min(cap, window_[p]), need_rows(p + 1));
      }
      uint64_t win[2] = {0, R};
      if (split) {  // parent rows of this window, found on the device
        k_window_rows<<<1, 64, 0, c_->stream>>>(off, R, lo, hi, dwin);
        FSIG_HIP(c_, hipGetLastError());
        if ((rc = read_u64(c_, dwin, win, 2))) return rc;
      }
      const size_t m2 = c_->arena.mark();
      uint32_t* child = (uint32_t*)c_->arena.alloc(4ull * (p + 1) * (hi - lo));
      if (!child) return set_err(c_, FSIG_ENOMEM, "join: arena exhausted for a window");
      JoinIO io{rows, win[0], win[1], nullptr, off, child, lo, hi, nullptr,
                family_counters(c_, FSIG_K_JOIN_WRITE)};
      {
        KTimer t(c_, FSIG_K_JOIN_WRITE);
        if ((rc = launch_join<true>(c_, widen_lanes(c_, G_, win[1] - win[0], 64), L, S, io))) return rc;
        t.done(0, 0);
      }
      if ((rc = expand(p + 1, child, hi - lo))) return rc;
      c_->arena.release(m2);
      lo = hi;
    }
    c_->arena.release(mk);
    return FSIG_OK;
  }

  // The last two levels in one pass (count mode): leaves of the counted rows
  // are added to the total and their children count towards the work limit;
  // deferred rows keep their children in cnt.
  int count_tail(const uint32_t* rows, uint64_t R, uint32_t* cnt, const JoinLevel& L1, const CandEdgeSrc& S1,
                 const JoinLevel& L2, const CandEdgeSrc& S2) {
    unsigned long long* res = c_->d_counters;
    FSIG_HIP(c_, hipMemsetAsync(res, 0, 2 * sizeof(unsigned long long), c_->stream));
    ChainTail chain;
    const bool use_chain = chain_tail_applies(L1, L2) && pv_.chain_tail(L1.p, chain);
    // lanes per parent row: at least 8 where the shared intersection goes
    // through the LDS tile, else few (rows are mostly a handful of dependent
    // gathers: more of them in flight)
    const bool needs_tile = tail_tile_words(8, L2, 0) != 0;
    const uint32_t G = widen_lanes(c_, use_chain || !needs_tile ? std::min<uint32_t>(std::max<uint32_t>(G_, 4), 8)
                                                                : std::max<uint32_t>(G_, 8),
                                   R, use_chain ? 32 : 64);
    const TailIO io{rows, R, cnt, res, family_counters(c_, FSIG_K_JOIN_TAIL),
                    use_chain ? 0u : tail_tile_words(G, L2, c_->tune.lds_tile)};
    KTimer t(c_, FSIG_K_JOIN_TAIL);
    int rc = use_chain ? launch_join_tail_chain(c_, G, L1, S1, L2, S2, chain, io)
                       : launch_join_tail(c_, G, L1, S1, L2, S2, io);
    if (rc) return rc;
    t.done(0, 0);
    uint64_t h[2];
    if ((rc = read_u64(c_, res, h, 2))) return rc;
    rows_ += h[1];
    const uint64_t limit = c_->work_limit;
    if (limit && rows_ > limit) return set_err(c_, FSIG_ELIMIT, "join: work limit exceeded");
    total_ += std::min(h[0], remaining());
    return FSIG_OK;
  }

  fsig_ctx* c_;
  Provider& pv_;
  uint32_t k_, G_;
  int mode_;
  uint64_t max_emb_;
  LeafSink* sink_;
  uint64_t total_ = 0;
  uint64_t rows_ = 0;  // intermediate rows produced so far
};

// Lanes per row: the tuned value, or the power of two at or above the
// expected driver-list length, clamped to [4, 64].
inline uint32_t pick_lanes(const fsig_ctx* c, double expected_list) {
  if (c->tune.lanes_per_row) return c->tune.lanes_per_row;
  uint32_t g = 4;
  while (g < 64 && g < expected_list) g <<= 1;
  return g;
}
