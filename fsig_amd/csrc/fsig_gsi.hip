// Alg 2 (the GSI path): neighbour-label signature filter over the inverted
// index and the join over the resident PCSR.
//
// Reference: GSI::signature_filter (match/GSI.cu:34-103) with filter_kernel
// (match/GSI_kernels.cuh:8-24) runs one N-wide pass per query vertex; here one
// pass over the inverted-index segment of each query label decides every
// query vertex of that label.  The join (match/GSI.cu:162-327) uploads one
// PCSR unit per level and neighbour and materialises every level in full;
// here all units are resident, the lists come from the shared join kernel
// (PcsrSrc) and levels are processed depth-first in memory-bounded windows.
// The candidate bitmap is per level (the reference ORs every level's
// candidates into one bitmap that is never cleared, match/GSI.cu:172-192).
#include <algorithm>
#include <cstring>

#include "common.h"
#include "device.h"
#include "internal.h"
#include "join.h"

namespace {

struct SigSegs {
  uint32_t n;
  uint32_t start[MAXQ];
  uint32_t flat[MAXQ + 1];
  uint32_t cmask[MAXQ];
  uint32_t qsig[MAXQ][16];  // per class; word 0 (the label) is implied by the segment
};

// C(u) = { v : label(v) = label(u) and (q_i & d_i) == q_i for words 1..15 }
// (match/GSI_kernels.cuh:8-24; word 0 of a signature is the label).
__global__ void __launch_bounds__(FSIG_BLOCK) k_sig_filter(SigSegs S, const uint32_t* __restrict__ inv_list,
                                                           const uint32_t* __restrict__ sig, uint32_t n,
                                                           uint32_t* __restrict__ mask) {
  const uint32_t total = S.flat[S.n];
  for (uint32_t f = blockIdx.x * blockDim.x + threadIdx.x; f < total; f += gridDim.x * blockDim.x) {
    uint32_t s = 0;
    while (s + 1 < S.n && S.flat[s + 1] <= f) s++;
    const uint32_t v = inv_list[S.start[s] + (f - S.flat[s])];
    uint32_t d[16];
#pragma unroll
    for (int w = 1; w < 16; w++) d[w] = sig[(uint64_t)w * n + v];
    uint32_t m = 0, cm = S.cmask[s];
    while (cm) {
      const uint32_t c = __builtin_ctz(cm);
      cm &= cm - 1;
      bool ok = true;
#pragma unroll
      for (int w = 1; w < 16; w++) ok &= (S.qsig[c][w] & d[w]) == S.qsig[c][w];
      if (ok) m |= 1u << c;
    }
    mask[v] = m;
  }
}

struct GsiProvider {
  using SrcType = PcsrSrc;
  const QueryPlan* P;
  const fsig_graph* g;
  const uint32_t* mask;
  const uint32_t* cand[MAXQ];
  uint32_t csize[MAXQ];

  uint32_t unit_of(int32_t a, int32_t b) const {
    const int32_t hi = std::max(a, b), lo = std::min(a, b);
    if (g->ref_unit_ids) return (uint32_t)(hi == 0 ? lo : (hi - 1) * hi / 2 + lo);
    return (uint32_t)(hi * (hi + 1) / 2 + lo);
  }
  JoinLevel level(uint32_t p) const {
    JoinLevel L;
    std::memset(&L, 0, sizeof(L));
    const uint32_t w = P->order[p];
    L.p = p;
    L.wcand = cand[P->cls[w]];
    L.wsize = csize[P->cls[w]];
    L.range_ids = 1;
    for (uint32_t c = 0; c < p; c++) {
      const uint32_t u = P->order[c];
      if ((P->adjm[w] >> u) & 1u) {
        L.bcol[L.nb++] = c;
      } else if (P->label[u] == P->label[w]) {
        L.cfcol[L.ncf] = c;
        L.cfsame |= 1u << L.ncf;  // rows hold data vertex ids
        L.ncf++;
      }
      if ((P->sym_lt[u] >> w) & 1u) L.gtmask |= 1u << c;  // symmetry: phi(u) < phi(w)
      if ((P->sym_lt[w] >> u) & 1u) L.ltmask |= 1u << c;  // phi(w) < phi(u)
    }
    return L;
  }
  bool chain_tail(uint32_t, ChainTail&) const { return false; }
  PcsrSrc src(uint32_t p) const {
    PcsrSrc S;
    std::memset(&S, 0, sizeof(S));
    S.groups = g->d_groups;
    S.edges = g->d_pedges;
    S.mask = mask;
    const uint32_t w = P->order[p];
    S.wbit = P->cls[w];
    uint32_t i = 0;
    for (uint32_t c = 0; c < p; c++) {
      const uint32_t u = P->order[c];
      if ((P->adjm[w] >> u) & 1u) S.unit[i++] = g->h_units[unit_of(P->label[w], P->label[u])];
    }
    return S;
  }
};

}  // namespace

int run_gsi(fsig_ctx* c, const fsig_graph* g, QueryPlan& P, int mode, uint64_t max_emb, uint64_t* count_out,
            uint32_t* emb_out, uint64_t emb_cap, unsigned long long* digest_acc) {
  const uint32_t nc = P.nclass, k = P.k;
  Arena& A = c->arena;
  int rc;
  KTimer tf(c, FSIG_K_FILTER);
  uint32_t* mask = (uint32_t*)A.alloc(4ull * std::max<uint32_t>(g->n, 1));
  if (!mask) return set_err(c, FSIG_ENOMEM, "filter: arena exhausted");
  k_fill_u32<<<grid_for(c, g->n, FSIG_BLOCK), FSIG_BLOCK, 0, c->stream>>>(mask, g->n, 0);
  SigSegs SS;
  std::memset(&SS, 0, sizeof(SS));
  uint32_t seg_of_class[MAXQ];
  int32_t seg_label[MAXQ];
  for (uint32_t a = 0; a < nc; a++) {
    const int32_t lab = P.class_label[a];
    uint32_t s = 0;
    while (s < SS.n && seg_label[s] != lab) s++;
    if (s == SS.n) {
      seg_label[s] = lab;
      SS.start[s] = g->h_inv_index[lab];
      SS.flat[s + 1] = SS.flat[s] + (g->h_inv_index[lab + 1] - g->h_inv_index[lab]);
      SS.n++;
    }
    SS.cmask[s] |= 1u << a;
    std::memcpy(SS.qsig[a], P.class_sig[a], 64);
    seg_of_class[a] = s;
  }
  const uint64_t items = SS.flat[SS.n];
  if (items)
    k_sig_filter<<<grid_for(c, items, FSIG_BLOCK), FSIG_BLOCK, 0, c->stream>>>(SS, g->d_inv_list, g->d_sig, g->n, mask);
  uint32_t* cand[MAXQ];
  uint32_t size[MAXQ];
  CompactTable CT;
  std::memset(&CT, 0, sizeof(CT));
  CT.n = nc;
  uint32_t tiles = 0;
  for (uint32_t a = 0; a < nc; a++) {
    const uint32_t s = seg_of_class[a];
    const uint32_t len = SS.flat[s + 1] - SS.flat[s];
    cand[a] = (uint32_t*)A.alloc(4ull * std::max<uint32_t>(len, 1));
    if (!cand[a]) return set_err(c, FSIG_ENOMEM, "filter: arena exhausted");
    CT.s[a] = CompactSeg{g->d_inv_list + SS.start[s], cand[a], len, a, tiles, 0};
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
  bool empty = false;
  for (uint32_t a = 0; a < nc; a++) {
    size[a] = (uint32_t)hs[a];
    cand_total += size[a];
    empty |= size[a] == 0;
  }
  // bytes: zero fill, per segment vertex its id, 15 signature words and the
  // mask word, compaction re-reads id + mask per class and writes candidates
  tf.done(4ull * g->n + items * (4 + 60 + 4) + 8ull * items * nc + 4ull * cand_total, items);
  c->last_cand.assign(k, 0);
  for (uint32_t u = 0; u < k; u++) c->last_cand[u] = size[P.cls[u]];
  // the signature filter is alg 2's only filter: both stages hold its sets
  if ((rc = keep_candidate_sets(c, 0, P, cand, size))) return rc;
  if ((rc = keep_candidate_sets(c, 1, P, cand, size))) return rc;
  if (empty) { *count_out = 0; return FSIG_OK; }

  plan_order_gsi(P, size);
  GsiProvider pv;
  pv.P = &P;
  pv.g = g;
  pv.mask = mask;
  for (uint32_t a = 0; a < nc; a++) { pv.cand[a] = cand[a]; pv.csize[a] = size[a]; }
  const double avg = g->n ? (double)g->nnz / g->n / std::max<uint32_t>(g->num_labels, 1) : 1;
  const uint32_t G = pick_lanes(c, avg);
  LeafSinks sinks;  // rows hold data vertex ids
  LeafSink* sink = sinks.setup(c, k, P.order, nullptr, emb_out, emb_cap, digest_acc);
  JoinDriver<GsiProvider> drv(c, pv, k, G, mode, max_emb, mode == FSIG_MODE_ENUMERATE ? sink : nullptr);
  // level-1 rows are the candidate vertices of the first join vertex
  if ((rc = drv.run(cand[P.cls[P.order[0]]], size[P.cls[P.order[0]]]))) return rc;
  *count_out = drv.total();
  return FSIG_OK;
}
