// Alg 1 (the MMatch path): label/degree filter over the inverted index,
// count-based refinement, candidate-edge CSR and the join.
//
// Data layout in HBM for one query (all from the context arena):
//   mask[n]        u32, bit a set iff data vertex v is a candidate of class a
//                  (one word per vertex serves all <= 32 classes: a refinement
//                  probe of a neighbour is one gather for every class)
//   cand[a]        ascending candidate vertices of class a (ping-pong pair)
//   keep[]         u8 per candidate, refinement verdicts
//   cedge (a->b)   off[|C(a)|+1] u32 and a list of candidate indices of b per
//                  candidate of a: N(C(a)[i]) intersected with C(b)
//   tables         join levels, p u32 candidate indices per row
#include <algorithm>
#include <cstring>

#include "device.h"
#include "internal.h"
#include "join.h"
#include "common.h"
#include "cedge.h"

namespace {

// ---------------------------------------------------------------------------
// Label/degree filter (LDF) over the inverted index: for every vertex of a
// query label, one word of class bits, set where the degree suffices.
// Reference: initialise_kernel (match/MMatch_kernels.cuh:6-16), which scans
// all n vertices once per class; here only |V_L| vertices per query label are
// touched and all classes of that label are decided in one pass.
struct LabelSegs {
  uint32_t n;
  uint32_t start[MAXQ];
  uint32_t flat[MAXQ + 1];
  uint32_t cmask[MAXQ];
  uint32_t deg[MAXQ];  // per class
  // neighbour-label filter (nlf null: off): class c needs, for each of its
  // nreq[c] entries e = req[c * MAXQ + i], nlf[v * L + (e >> 8)] >= (e & 255)
  uint32_t nreq[MAXQ];
  const uint32_t* req;
  const uint8_t* nlf;
  uint32_t L;
};

__device__ __forceinline__ bool nlf_ok(const uint8_t* __restrict__ nlf, uint32_t L, uint32_t v,
                                       const uint32_t* __restrict__ req, uint32_t nreq) {
  const uint8_t* row = nlf + (uint64_t)v * L;
  for (uint32_t i = 0; i < nreq; i++) {
    const uint32_t e = req[i];
    if (row[e >> 8] < (e & 255u)) return false;
  }
  return true;
}

__global__ void __launch_bounds__(FSIG_BLOCK) k_ldf(LabelSegs S, const uint32_t* __restrict__ inv_list,
                                                    const uint32_t* __restrict__ off, uint32_t* __restrict__ mask) {
  const uint32_t total = S.flat[S.n];
  for (uint32_t f = blockIdx.x * blockDim.x + threadIdx.x; f < total; f += gridDim.x * blockDim.x) {
    uint32_t s = 0;
    while (s + 1 < S.n && S.flat[s + 1] <= f) s++;
    const uint32_t v = inv_list[S.start[s] + (f - S.flat[s])];
    const uint32_t d = off[v + 1] - off[v];
    uint32_t m = 0, cm = S.cmask[s];
    while (cm) {
      const uint32_t c = __builtin_ctz(cm);
      cm &= cm - 1;
      if (d >= S.deg[c] && (!S.nlf || nlf_ok(S.nlf, S.L, v, S.req + c * MAXQ, S.nreq[c]))) m |= 1u << c;
    }
    mask[v] = m;
  }
}

// ---------------------------------------------------------------------------
// Count-based refinement, one Jacobi pass: candidate v of class a survives iff
// for every class b, |N(v) ∩ C(b)| >= thr[a][b] (the number of neighbours a
// member of a has in b).  Reference rule: count_neighbors_kernel +
// check_re_kernel (match/MMatch_kernels.cuh:28-117, 190-214) driven by
// kernel_filter_forest / kernel_filter_core (match/MMatch.cu:810-885,
// 1127-1195).  G lanes share one candidate; each class bit of the gathered
// neighbour words is counted with one wave ballot, and a target stops as soon
// as its threshold is reached.
struct RefineTable {
  uint32_t n;
  uint32_t base[MAXQ + 1];
  uint32_t tmask[MAXQ];
  const uint32_t* cand[MAXQ];
  uint8_t thr[MAXQ][MAXQ];
};

template <int G>
__global__ void __launch_bounds__(FSIG_BLOCK) k_refine(RefineTable T, const uint32_t* __restrict__ off,
                                                       const uint32_t* __restrict__ adj,
                                                       const uint32_t* __restrict__ mask,
                                                       uint8_t* __restrict__ keep) {
  constexpr int GPB = FSIG_BLOCK / G;
  __shared__ uint8_t rem[GPB][MAXQ];  // outstanding need per target class
  const uint32_t lane = threadIdx.x & 63u, gl = lane & (G - 1), gw = threadIdx.x / G;
  const uint64_t gmask = (G == 64) ? ~0ull : (((1ull << G) - 1) << (lane - gl));
  const uint32_t total = T.base[T.n];
  const uint64_t ngroups = (uint64_t)gridDim.x * GPB;
  for (uint64_t f = (uint64_t)xcd_block() * GPB + gw; f < total; f += ngroups) {
    uint32_t a = 0;
    while (a + 1 < T.n && T.base[a + 1] <= f) a++;
    const uint32_t v = T.cand[a][f - T.base[a]];
    const uint32_t s = off[v], e = off[v + 1];
    // one pass over the adjacency: each gathered mask word serves every
    // target class; stop once all thresholds are met
    uint32_t open = T.tmask[a];
    for (uint32_t tm = open; tm; tm &= tm - 1) rem[gw][__builtin_ctz(tm)] = T.thr[a][__builtin_ctz(tm)];
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

// Clears the class bit of every rejected candidate (after the pass has read
// all bits, so the pass is a Jacobi step and its result deterministic).
__global__ void __launch_bounds__(FSIG_BLOCK) k_refine_apply(RefineTable T, const uint8_t* __restrict__ keep,
                                                             uint32_t* __restrict__ mask) {
  const uint32_t total = T.base[T.n];
  for (uint32_t f = blockIdx.x * blockDim.x + threadIdx.x; f < total; f += gridDim.x * blockDim.x) {
    if (keep[f]) continue;
    uint32_t a = 0;
    while (a + 1 < T.n && T.base[a + 1] <= f) a++;
    atomicAnd(&mask[T.cand[a][f - T.base[a]]], ~(1u << a));
  }
}

// ---------------------------------------------------------------------------
// Candidate-edge CSR for every class pair (a -> b) the join uses.
// Reference: collect_forest_edges / collect_core_edges with
// re_examine_kernel (match/MMatch.cu:1229-1423, match/MMatch_kernels.cuh:
// 302-372), whose offsets are indexed by candidate position but whose
// lookups scan the candidate array linearly (match/MMatch_kernels.cuh:
// 394-415).  Here list entries are candidate indices of b, so a row stores
// indices and every lookup is direct.
struct PairDesc {
  const uint32_t* acand;
  const uint32_t* bcand;
  const uint32_t* mask;  // the query's class mask
  const uint32_t* bits;  // candidate bitmap of b (null: the mask)
  uint32_t* off_out;
  const uint32_t* bpos;  // vertex -> candidate index of b (null: binary search)
  uint32_t asize, bsize, bbit, pad;
  uint64_t flat;  // first row of the pair in the flat row space
};

// bpos[C[j]] = j: the vertex -> candidate index map of one class.
__global__ void k_scatter_pos(const uint32_t* __restrict__ cand, uint32_t n, uint32_t* __restrict__ pos) {
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) pos[cand[j]] = j;
}

__global__ void k_gather_u64(const uint64_t* __restrict__ src, const PairDesc* __restrict__ P, uint32_t np,
                             uint64_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < np) out[i] = src[P[i].flat];
}

}  // namespace

struct MMatchProvider {
  using SrcType = CandEdgeSrc;
  const QueryPlan* P;
  const uint32_t* cand[MAXQ];  // per class
  uint32_t csize[MAXQ];
  // pair lists: index by (class a, class b)
  const uint32_t* poff[MAXQ][MAXQ];
  const uint32_t* plst[MAXQ][MAXQ];
  // data vertex -> candidate index maps of the classes that have one (null:
  // binary search)
  const uint32_t* pos[MAXQ];

  JoinLevel level(uint32_t p) const {
    JoinLevel L;
    std::memset(&L, 0, sizeof(L));
    const uint32_t w = P->order[p];
    L.p = p;
    L.wcand = cand[P->cls[w]];
    L.wsize = csize[P->cls[w]];
    L.wpos = pos[P->cls[w]];
    for (uint32_t c = 0; c < p; c++) {
      const uint32_t u = P->order[c];
      if ((P->adjm[w] >> u) & 1u) {
        L.bcol[L.nb++] = c;
      } else if (P->label[u] == P->label[w]) {
        const uint32_t j = L.ncf++;
        L.cfcol[j] = c;
        if (P->cls[u] == P->cls[w]) {
          L.cfsame |= 1u << j;
        } else {
          L.cfcand[j] = cand[P->cls[u]];
          L.any_cross = 1;
        }
      }
    }
    for (uint32_t c = 0; c < p; c++) {  // symmetry conditions against earlier columns
      const uint32_t u = P->order[c];
      if ((P->sym_lt[u] >> w) & 1u) L.gtmask |= 1u << c;  // phi(u) < phi(w)
      if ((P->sym_lt[w] >> u) & 1u) L.ltmask |= 1u << c;  // phi(w) < phi(u)
    }
    return L;
  }
  CandEdgeSrc src(uint32_t p) const {
    CandEdgeSrc S;
    std::memset(&S, 0, sizeof(S));
    const uint32_t w = P->order[p];
    uint32_t i = 0;
    for (uint32_t c = 0; c < p; c++) {
      const uint32_t u = P->order[c];
      if ((P->adjm[w] >> u) & 1u) {
        S.off[i] = poff[P->cls[u]][P->cls[w]];
        S.lst[i] = plst[P->cls[u]][P->cls[w]];
        i++;
      }
    }
    return S;
  }
  // chain tail of the last two levels (set up by run_mmatch when the query
  // ends that way; chain_p is its parent level)
  ChainTail chain;
  uint32_t chain_p = 0xffffffffu;
  bool chain_tail(uint32_t p, ChainTail& t) const {
    if (p != chain_p) return false;
    t = chain;
    return true;
  }
};

static uint32_t refine_lanes(const fsig_graph* g) {
  const double avg = g->n ? (double)g->nnz / g->n : 0;
  return avg <= 12 ? 8 : (avg <= 40 ? 16 : 64);
}

int run_mmatch(fsig_ctx* c, const fsig_graph* g, QueryPlan& P, int mode, uint64_t max_emb, uint64_t* count_out,
               uint32_t* emb_out, uint64_t emb_cap, unsigned long long* digest_acc) {
  const uint32_t nc = P.nclass, k = P.k;
  Arena& A = c->arena;
  int rc;
  // ---- label/degree filter over the inverted index
  KTimer tf(c, FSIG_K_FILTER);
  uint32_t* mask = (uint32_t*)A.alloc(4ull * std::max<uint32_t>(g->n, 1));
  if (!mask) return set_err(c, FSIG_ENOMEM, "filter: arena exhausted");
  k_fill_u32<<<grid_for(c, g->n, FSIG_BLOCK), FSIG_BLOCK, 0, c->stream>>>(mask, g->n, 0);
  LabelSegs LS;
  std::memset(&LS, 0, sizeof(LS));
  uint32_t seg_of_class[MAXQ];
  int32_t seg_label[MAXQ];
  for (uint32_t a = 0; a < nc; a++) {
    const int32_t lab = P.class_label[a];
    uint32_t s = 0;
    while (s < LS.n && seg_label[s] != lab) s++;
    if (s == LS.n) {
      seg_label[s] = lab;
      LS.start[s] = g->h_inv_index[lab];
      LS.flat[s + 1] = LS.flat[s] + (g->h_inv_index[lab + 1] - g->h_inv_index[lab]);
      LS.n++;
    }
    LS.cmask[s] |= 1u << a;
    LS.deg[a] = P.class_deg[a];
    seg_of_class[a] = s;
  }
  if (g->d_nlf && c->tune.nlf_filter == 0) {
    uint32_t* dreq = (uint32_t*)A.alloc(4ull * MAXQ * MAXQ);
    if (!dreq) return set_err(c, FSIG_ENOMEM, "filter: arena exhausted");
    FSIG_HIP(c, hipMemcpyAsync(dreq, P.nlf_req, 4ull * MAXQ * MAXQ, hipMemcpyHostToDevice, c->stream));
    for (uint32_t a = 0; a < nc; a++) LS.nreq[a] = P.nlf_n[a];
    LS.req = dreq;
    LS.nlf = g->d_nlf;
    LS.L = g->num_labels;
  }
  const uint64_t ldf_items = LS.flat[LS.n];
  if (ldf_items)
    k_ldf<<<grid_for(c, ldf_items, FSIG_BLOCK), FSIG_BLOCK, 0, c->stream>>>(LS, g->d_inv_list, g->d_off, mask);
  // compaction per class over its label segment
  uint32_t* cand_a[MAXQ];
  uint32_t* cand_b[MAXQ];
  uint32_t size[MAXQ];
  CompactTable CT;
  std::memset(&CT, 0, sizeof(CT));
  CT.n = nc;
  uint32_t tiles = 0;
  for (uint32_t a = 0; a < nc; a++) {
    const uint32_t s = seg_of_class[a];
    const uint32_t len = LS.flat[s + 1] - LS.flat[s];
    cand_a[a] = (uint32_t*)A.alloc(4ull * std::max<uint32_t>(len, 1));
    if (!cand_a[a]) return set_err(c, FSIG_ENOMEM, "filter: arena exhausted");
    CT.s[a] = CompactSeg{g->d_inv_list + LS.start[s], cand_a[a], len, a, tiles, 0};
    tiles += (len + kTile - 1) / kTile;
  }
  CT.ntiles = tiles;
  uint32_t* tilecnt = (uint32_t*)A.alloc(4ull * std::max<uint32_t>(tiles, 1));
  uint64_t* dsizes = (uint64_t*)A.alloc(8ull * MAXQ);
  if (!tilecnt || !dsizes) return set_err(c, FSIG_ENOMEM, "filter: arena exhausted");
  if (tiles) {
    k_compact_count<0><<<tiles, FSIG_BLOCK, 0, c->stream>>>(CT, mask, nullptr, tilecnt);
    k_compact_offsets<<<nc, FSIG_BLOCK, 0, c->stream>>>(CT, tilecnt, dsizes);
    k_compact_scatter<0><<<tiles, FSIG_BLOCK, 0, c->stream>>>(CT, mask, nullptr, tilecnt);
  } else {
    FSIG_HIP(c, hipMemsetAsync(dsizes, 0, 8ull * nc, c->stream));
  }
  FSIG_HIP(c, hipGetLastError());
  uint64_t hs[MAXQ];
  if ((rc = read_u64(c, dsizes, hs, nc))) return rc;
  uint64_t cand_total = 0;
  for (uint32_t a = 0; a < nc; a++) { size[a] = (uint32_t)hs[a]; cand_total += size[a]; }
  if ((rc = keep_candidate_sets(c, 0, P, cand_a, size))) return rc;
  tf.done(4ull * g->n + 12ull * ldf_items + 4ull * ldf_items * nc + 4ull * cand_total, ldf_items);

  // ---- refinement passes
  bool any_target = false;
  RefineTable RT;
  std::memset(&RT, 0, sizeof(RT));
  RT.n = nc;
  for (uint32_t a = 0; a < nc; a++)
    for (uint32_t b = 0; b < nc; b++)
      if (P.thr[a][b]) {
        RT.tmask[a] |= 1u << b;
        RT.thr[a][b] = (uint8_t)P.thr[a][b];
        any_target = true;
      }
  for (uint32_t a = 0; a < nc; a++) {
    cand_b[a] = (uint32_t*)A.alloc(4ull * std::max<uint32_t>(size[a], 1));
    if (!cand_b[a]) return set_err(c, FSIG_ENOMEM, "refine: arena exhausted");
  }
  uint8_t* keep = (uint8_t*)A.alloc(std::max<uint64_t>(cand_total, 1));
  if (!keep) return set_err(c, FSIG_ENOMEM, "refine: arena exhausted");
  uint32_t** cur = cand_a;
  uint32_t** nxt = cand_b;
  const uint32_t RG = refine_lanes(g);
  bool empty = false;
  for (uint32_t a = 0; a < nc; a++) empty |= size[a] == 0;
  for (uint32_t pass = 0; any_target && !empty && pass < refine_passes(c); pass++) {
    KTimer tr(c, FSIG_K_REFINE);
    uint64_t flat = 0, deg_sum = 0;
    for (uint32_t a = 0; a < nc; a++) {
      RT.base[a] = (uint32_t)flat;
      RT.cand[a] = cur[a];
      flat += size[a];
    }
    RT.base[nc] = (uint32_t)flat;
    (void)deg_sum;
    switch (RG) {
      case 8: k_refine<8><<<grid_for(c, flat, FSIG_BLOCK / 8), FSIG_BLOCK, 0, c->stream>>>(RT, g->d_off, g->d_adj, mask, keep); break;
      case 16: k_refine<16><<<grid_for(c, flat, FSIG_BLOCK / 16), FSIG_BLOCK, 0, c->stream>>>(RT, g->d_off, g->d_adj, mask, keep); break;
      default: k_refine<64><<<grid_for(c, flat, FSIG_BLOCK / 64), FSIG_BLOCK, 0, c->stream>>>(RT, g->d_off, g->d_adj, mask, keep); break;
    }
    k_refine_apply<<<grid_for(c, flat, FSIG_BLOCK), FSIG_BLOCK, 0, c->stream>>>(RT, keep, mask);
    CompactTable RC;
    std::memset(&RC, 0, sizeof(RC));
    RC.n = nc;
    uint32_t rt = 0;
    for (uint32_t a = 0; a < nc; a++) {
      RC.s[a] = CompactSeg{cur[a], nxt[a], size[a], a, rt, RT.base[a]};
      rt += (size[a] + kTile - 1) / kTile;
    }
    RC.ntiles = rt;
    if (rt) {
      k_compact_count<1><<<rt, FSIG_BLOCK, 0, c->stream>>>(RC, mask, keep, tilecnt);
      k_compact_offsets<<<nc, FSIG_BLOCK, 0, c->stream>>>(RC, tilecnt, dsizes);
      k_compact_scatter<1><<<rt, FSIG_BLOCK, 0, c->stream>>>(RC, mask, keep, tilecnt);
    }
    FSIG_HIP(c, hipGetLastError());
    if ((rc = read_u64(c, dsizes, hs, nc))) return rc;
    bool changed = false;
    uint64_t new_total = 0;
    for (uint32_t a = 0; a < nc; a++) {
      changed |= hs[a] != size[a];
      size[a] = (uint32_t)hs[a];
      new_total += size[a];
      empty |= size[a] == 0;
    }
    std::swap(cur, nxt);
    // bytes: candidate ids, offset pairs, adjacency, one mask word per probe
    // (upper bound: full lists), keep flags and the compaction
    tr.done(flat * (4 + 8 + 1 + 1 + 4) + 8ull * flat * 0 + new_total * 4, flat);
    if (!changed) break;
  }
  c->last_cand.assign(k, 0);
  for (uint32_t u = 0; u < k; u++) c->last_cand[u] = size[P.cls[u]];
  if ((rc = keep_candidate_sets(c, 1, P, cur, size))) return rc;
  if (empty) { *count_out = 0; return FSIG_OK; }

  // ---- join order and the class pairs it needs
  plan_order_mmatch(P, size);
  std::vector<PairDesc> pairs;
  int pair_id[MAXQ][MAXQ];
  for (auto& row : pair_id) for (int& x : row) x = -1;
  for (uint32_t p = 1; p < k; p++) {
    const uint32_t w = P.order[p];
    for (uint32_t cpos = 0; cpos < p; cpos++) {
      const uint32_t u = P.order[cpos];
      if (!((P.adjm[w] >> u) & 1u)) continue;
      const uint32_t a = P.cls[u], b = P.cls[w];
      if (pair_id[a][b] >= 0) continue;
      pair_id[a][b] = (int)pairs.size();
      PairDesc pd;
      pd.bpos = nullptr;
      pd.acand = cur[a];
      pd.bcand = cur[b];
      pd.asize = size[a];
      pd.bsize = size[b];
      pd.bbit = b;
      pd.mask = mask;
      pd.bits = nullptr;
      pd.off_out = nullptr;
      pd.pad = 0;
      pd.flat = 0;
      pairs.push_back(pd);
    }
  }
  MMatchProvider pv;
  std::memset(&pv, 0, sizeof(pv));
  pv.P = &P;
  pv.chain_p = 0xffffffffu;
  for (uint32_t a = 0; a < nc; a++) { pv.cand[a] = cur[a]; pv.csize[a] = size[a]; }
  // A count whose last two levels form a chain (see ChainTail) also needs
  // the reverse pair (last class -> previous class).
  const bool chain = mode == FSIG_MODE_COUNT && k >= 3 && c->tune.fused_tail == 0 &&
                     chain_tail_applies(pv.level(k - 2), pv.level(k - 1));
  const uint32_t cls_w = P.cls[P.order[k - 1]], cls_p = P.cls[P.order[k >= 2 ? k - 2 : 0]];
  if (chain && pair_id[cls_w][cls_p] < 0) {
    pair_id[cls_w][cls_p] = (int)pairs.size();
    PairDesc pd;
    pd.bpos = nullptr;
    pd.acand = cur[cls_w];
    pd.bcand = cur[cls_p];
    pd.asize = size[cls_w];
    pd.bsize = size[cls_p];
    pd.bbit = cls_p;
    pd.mask = mask;
    pd.bits = nullptr;
    pd.off_out = nullptr;
    pd.pad = 0;
    pd.flat = 0;
    pairs.push_back(pd);
  }
  double avg_list = 1;
  if (!pairs.empty()) {
    KTimer tc(c, FSIG_K_CEDGE);
    const uint32_t np = (uint32_t)pairs.size();
    uint64_t flat = 0;
    for (auto& pd : pairs) {
      pd.flat = flat;
      flat += pd.asize;
      pd.off_out = (uint32_t*)A.alloc(4ull * (pd.asize + 1));
      if (!pd.off_out) return set_err(c, FSIG_ENOMEM, "cedge: arena exhausted");
    }
    if (flat >= 0xffffffffull) return set_err(c, FSIG_EUNSUPPORTED, "cedge: too many candidate rows");
    // vertex -> candidate index maps of the target classes (n words each),
    // when the arena holds them; otherwise the fill binary-searches
    {
      uint32_t* bpos_of[MAXQ] = {nullptr};
      bool fit = true;
      for (auto& pd : pairs) {
        const uint32_t b = pd.bbit;
        if (!bpos_of[b] && fit) {
          bpos_of[b] = (uint32_t*)A.alloc(4ull * g->n);
          if (!bpos_of[b]) { fit = false; continue; }
          k_scatter_pos<<<grid_for(c, pd.bsize, FSIG_BLOCK), FSIG_BLOCK, 0, c->stream>>>(pd.bcand, pd.bsize, bpos_of[b]);
        }
        pd.bpos = fit ? bpos_of[b] : nullptr;
      }
      if (!fit)
        for (auto& pd : pairs) pd.bpos = nullptr;
      for (uint32_t b = 0; b < nc && fit; b++) pv.pos[b] = bpos_of[b];  // the join's conflict lookups
    }
    // candidate bitmaps of the target classes on graphs whose class mask
    // outgrows L2 (n/8 bytes each instead of the mask's 4n)
    if (g->n >= (1u << 20)) {
      const uint64_t words = (g->n + 31ull) / 32;
      uint32_t* bits_of[MAXQ] = {nullptr};
      for (auto& pd : pairs) {
        const uint32_t b = pd.bbit;
        if (!bits_of[b]) {
          bits_of[b] = (uint32_t*)A.alloc(4ull * words);
          if (!bits_of[b]) return set_err(c, FSIG_ENOMEM, "cedge: arena exhausted");
          k_fill_u32<<<grid_for(c, words, FSIG_BLOCK), FSIG_BLOCK, 0, c->stream>>>(bits_of[b], words, 0);
          k_set_bits<<<grid_for(c, pd.bsize, FSIG_BLOCK), FSIG_BLOCK, 0, c->stream>>>(pd.bcand, pd.bsize, bits_of[b]);
        }
        pd.bits = bits_of[b];
      }
    }
    PairDesc* dP = (PairDesc*)A.alloc(sizeof(PairDesc) * np);
    uint32_t* cnt = (uint32_t*)A.alloc(4ull * flat);
    uint64_t* gscan = (uint64_t*)A.alloc(8ull * (flat + 1));
    uint64_t* pbase = (uint64_t*)A.alloc(8ull * (np + 1));
    if (!dP || !cnt || !gscan || !pbase) return set_err(c, FSIG_ENOMEM, "cedge: arena exhausted");
    FSIG_HIP(c, hipMemcpyAsync(dP, pairs.data(), sizeof(PairDesc) * np, hipMemcpyHostToDevice, c->stream));
    if ((rc = row_count(c, RG, dP, np, flat, g->d_off, g->d_adj, cnt))) return rc;
    SegRows segs;  // hub rows, one wavefront per segment
    if ((rc = seg_prepare(c, dP, np, flat, g->d_off, cnt, segs))) return rc;
    if ((rc = seg_count(c, dP, np, flat, g->d_off, g->d_adj, cnt, segs))) return rc;
    FSIG_HIP(c, hipGetLastError());
    if ((rc = scan_u32_to_u64(c, cnt, gscan, flat))) return rc;
    uint64_t total_edges;
    if ((rc = read_u64(c, gscan + flat, &total_edges, 1))) return rc;
    uint32_t* lst_all = (uint32_t*)A.alloc(4ull * std::max<uint64_t>(total_edges, 1));
    if (!lst_all) return set_err(c, FSIG_ENOMEM, "cedge: arena exhausted for lists");
    if ((rc = row_fill(c, RG, dP, np, flat, g->d_off, g->d_adj, gscan, lst_all))) return rc;
    if ((rc = seg_fill(c, dP, np, flat, g->d_off, g->d_adj, gscan, lst_all, segs))) return rc;
    k_gather_u64<<<(np + 255) / 256, 256, 0, c->stream>>>(gscan, dP, np, pbase);
    FSIG_HIP(c, hipGetLastError());
    std::vector<uint64_t> hb(np);
    for (uint32_t i = 0; i < np; i += 4096)
      if ((rc = read_u64(c, pbase + i, hb.data() + i, std::min<uint32_t>(4096, np - i)))) return rc;
    for (uint32_t a = 0; a < nc; a++)
      for (uint32_t b = 0; b < nc; b++)
        if (pair_id[a][b] >= 0) {
          const PairDesc& pd = pairs[pair_id[a][b]];
          pv.poff[a][b] = pd.off_out;
          pv.plst[a][b] = lst_all + hb[pair_id[a][b]];
        }
    avg_list = flat ? (double)total_edges / (double)flat : 1;
    // bytes: both passes read each row's candidate and offsets and, per
    // adjacency entry, the entry and its mask word; the fill reads the
    // candidate index of each kept entry and writes it; counts, offsets
    tc.done(28ull * flat + 16ull * segs.entries + 8ull * total_edges + 4ull * (flat + np), segs.entries);
  }
  if (chain) {
    // S[v] over the candidates v of the previous position's backward
    // neighbour: the leaves of a row before conflicts are removed
    const JoinLevel L1 = pv.level(k - 2);
    const uint32_t cls_c1 = P.cls[P.order[L1.bcol[0]]];
    const uint32_t nv = size[cls_c1];
    unsigned long long* sum = (unsigned long long*)A.alloc(8ull * std::max<uint32_t>(nv, 1));
    if (!sum) return set_err(c, FSIG_ENOMEM, "join: arena exhausted for the chain tail");
    KTimer tt(c, FSIG_K_JOIN_TAIL);
    k_chain_sum<16><<<grid_for(c, nv, FSIG_BLOCK / 16), FSIG_BLOCK, 0, c->stream>>>(
        pv.poff[cls_c1][cls_p], pv.plst[cls_c1][cls_p], pv.poff[cls_p][cls_w], nv, sum);
    FSIG_HIP(c, hipGetLastError());
    tt.done(16ull * nv, 0);  // offsets of each list; its entries and own offsets are charged by the tail
    pv.chain = ChainTail{sum, pv.poff[cls_w][cls_p], pv.plst[cls_w][cls_p]};
    pv.chain_p = k - 2;
  }

  // ---- join
  const uint32_t G = pick_lanes(c, avg_list);
  const uint32_t n1 = size[P.cls[P.order[0]]];
  uint32_t* level1 = (uint32_t*)A.alloc(4ull * std::max<uint32_t>(n1, 1));
  if (!level1) return set_err(c, FSIG_ENOMEM, "join: arena exhausted");
  k_iota<<<grid_for(c, n1, FSIG_BLOCK), FSIG_BLOCK, 0, c->stream>>>(level1, n1);
  const uint32_t* cand_of_pos[MAXQ];
  for (uint32_t pos = 0; pos < k; pos++) cand_of_pos[pos] = cur[P.cls[P.order[pos]]];
  LeafSinks sinks;
  LeafSink* sink = sinks.setup(c, k, P.order, cand_of_pos, emb_out, emb_cap, digest_acc);
  JoinDriver<MMatchProvider> drv(c, pv, k, G, mode, max_emb, mode == FSIG_MODE_ENUMERATE ? sink : nullptr);
  if ((rc = drv.run(level1, n1))) return rc;
  *count_out = drv.total();
  return FSIG_OK;
}
