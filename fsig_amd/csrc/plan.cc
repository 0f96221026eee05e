// Query planning on the host: equivalence classes, refinement thresholds and
// join orders.
//
// Classes.  The reference groups query vertices by (degree, label) and then
// splits groups whose neighbourhoods cannot be mapped onto each other, trying
// permutations of the neighbour lists (match/MMatch.cu:44-173, factorial in
// the degree).  Here the classes are the stable colouring of colour
// refinement started from (label, degree): two vertices share a class iff
// they have the same label, degree and multiset of neighbour classes.  All
// members of a class then have identical candidate sets under the
// count-based refinement, which is what the reference's classes are for, and
// the cost is polynomial in k.
//
// Thresholds.  thr[a][b] = number of neighbours a member of class a has in
// class b: the "count" of the reference's BasicOperator (match/MMatch.h:31-35,
// built at match/MMatch.cu:334-341).
#include <algorithm>
#include <cstring>
#include <map>

#include "internal.h"

static uint32_t popc(uint32_t x) { return (uint32_t)__builtin_popcount(x); }

int plan_classes(QueryPlan& P, uint32_t k, const int32_t* qlabels, const uint32_t* qoff,
                 const uint32_t* qadj, const uint32_t* qsig, int alg, std::string& err) {
  if (k == 0) { err = "query has no vertices"; return FSIG_EINVAL; }
  if (k > MAXQ) { err = "query has more than 32 vertices"; return FSIG_EUNSUPPORTED; }
  if (alg == FSIG_ALG_GSI && !qsig) { err = "alg 2 needs the query signature table"; return FSIG_EINVAL; }
  P = QueryPlan();
  P.k = k;
  for (uint32_t u = 0; u < k; u++) {
    P.label[u] = qlabels[u];
    if (qlabels[u] < 0) { err = "negative query label"; return FSIG_EINVAL; }
    P.adjm[u] = 0;
    for (uint32_t i = qoff[u]; i < qoff[u + 1]; i++) {
      uint32_t w = qadj[i];
      if (w >= k || w == u) { err = "query adjacency out of range or self-loop"; return FSIG_EINVAL; }
      P.adjm[u] |= 1u << w;
    }
  }
  for (uint32_t u = 0; u < k; u++) {
    P.deg[u] = popc(P.adjm[u]);
    for (uint32_t w = 0; w < k; w++)
      if (((P.adjm[u] >> w) & 1) != ((P.adjm[w] >> u) & 1)) {
        err = "query adjacency is not symmetric";
        return FSIG_EINVAL;
      }
  }
  std::vector<uint32_t> color(k);
  if (alg == FSIG_ALG_GSI) {
    // GSI keeps one candidate set per query vertex, filtered by signature
    // (match/GSI.cu:34-103); vertices with equal signatures share it.
    std::map<std::vector<uint32_t>, uint32_t> ids;
    for (uint32_t u = 0; u < k; u++) {
      std::vector<uint32_t> s(qsig + (size_t)u * 16, qsig + (size_t)u * 16 + 16);
      auto it = ids.find(s);
      if (it == ids.end()) it = ids.emplace(s, (uint32_t)ids.size()).first;
      color[u] = it->second;
    }
  } else {
    // colour refinement from (label, degree)
    std::map<std::vector<uint32_t>, uint32_t> ids;
    for (uint32_t u = 0; u < k; u++) {
      std::vector<uint32_t> key{(uint32_t)P.label[u], P.deg[u]};
      auto it = ids.emplace(key, 0).first;
      (void)it;
    }
    uint32_t r = 0;
    for (auto& kv : ids) kv.second = r++;
    for (uint32_t u = 0; u < k; u++) color[u] = ids[{(uint32_t)P.label[u], P.deg[u]}];
    uint32_t ncol = (uint32_t)ids.size();
    while (true) {
      std::map<std::vector<uint32_t>, uint32_t> sig;
      std::vector<std::vector<uint32_t>> keys(k);
      for (uint32_t u = 0; u < k; u++) {
        std::vector<uint32_t> nb;
        for (uint32_t w = 0; w < k; w++)
          if ((P.adjm[u] >> w) & 1) nb.push_back(color[w]);
        std::sort(nb.begin(), nb.end());
        keys[u].push_back(color[u]);
        keys[u].insert(keys[u].end(), nb.begin(), nb.end());
        sig.emplace(keys[u], 0);
      }
      uint32_t rr = 0;
      for (auto& kv : sig) kv.second = rr++;
      for (uint32_t u = 0; u < k; u++) color[u] = sig[keys[u]];
      if (rr == ncol) break;
      ncol = rr;
    }
  }
  // number classes by first occurrence
  std::vector<uint32_t> remap(k, FSIG_INVALID);
  P.nclass = 0;
  for (uint32_t u = 0; u < k; u++) {
    if (remap[color[u]] == FSIG_INVALID) {
      remap[color[u]] = P.nclass;
      P.rep[P.nclass] = u;
      P.class_label[P.nclass] = P.label[u];
      P.class_deg[P.nclass] = P.deg[u];
      if (qsig) std::memcpy(P.class_sig[P.nclass], qsig + (size_t)u * 16, 64);
      P.nclass++;
    }
    P.cls[u] = remap[color[u]];
  }
  std::memset(P.thr, 0, sizeof(P.thr));
  for (uint32_t a = 0; a < P.nclass; a++) {
    uint32_t u = P.rep[a];
    for (uint32_t w = 0; w < k; w++)
      if ((P.adjm[u] >> w) & 1) P.thr[a][P.cls[w]]++;
  }
  // neighbour-label frequency: an injective image of u needs, for every
  // label l, at least as many label-l neighbours as u has
  for (uint32_t a = 0; a < P.nclass; a++) {
    const uint32_t u = P.rep[a];
    uint32_t lab[MAXQ], cnt[MAXQ], m = 0;
    for (uint32_t w = 0; w < k; w++) {
      if (!((P.adjm[u] >> w) & 1)) continue;
      uint32_t j = 0;
      while (j < m && lab[j] != (uint32_t)P.label[w]) j++;
      if (j == m) { lab[m] = (uint32_t)P.label[w]; cnt[m] = 0; m++; }
      cnt[j]++;
    }
    P.nlf_n[a] = m;
    for (uint32_t j = 0; j < m; j++) P.nlf_req[a][j] = (lab[j] << 8) | cnt[j];
  }
  return FSIG_OK;
}

// Join order of alg 1.  The reference starts from the smallest candidate set
// (core vertices first) and then repeatedly takes the frontier vertex with
// the smallest score (match/MMatch.cu:1575-1641).  Here the frontier vertex
// with the most already-ordered neighbours is taken first (every one of them
// becomes an intersection that prunes rows), ties broken by candidate-set
// size, degree and id.  Any order gives the same embedding set.
void plan_order_mmatch(QueryPlan& P, const uint32_t* class_size) {
  const uint32_t k = P.k;
  uint32_t used = 0;
  for (uint32_t i = 0; i < k; i++) {
    uint32_t best = FSIG_INVALID;
    uint64_t best_key = ~0ull;
    for (uint32_t u = 0; u < k; u++) {
      if ((used >> u) & 1) continue;
      uint32_t back = popc(P.adjm[u] & used);
      bool frontier = i == 0 || back > 0;
      // key: frontier first, more back neighbours, smaller set, larger degree, id
      uint64_t key = ((uint64_t)(frontier ? 0 : 1) << 62) | ((uint64_t)(MAXQ - back) << 56) |
                     ((uint64_t)std::min<uint32_t>(class_size[P.cls[u]], (1u << 31) - 1) << 24) |
                     ((uint64_t)(MAXQ - P.deg[u]) << 8) | u;
      if (key < best_key) { best_key = key; best = u; }
    }
    P.order[i] = best;
    P.pos[best] = i;
    used |= 1u << best;
  }
}

// Join order of alg 2, after GSI::generate_join_order (match/GSI.cu:105-149):
// smallest candidate set first, then the frontier vertex of smallest score,
// where matching a vertex multiplies its neighbours' scores by 0.9.  Ties go
// to the smaller id (the reference leaves them to its heap).
void plan_order_gsi(QueryPlan& P, const uint32_t* class_size) {
  const uint32_t k = P.k;
  float score[MAXQ];
  for (uint32_t u = 0; u < k; u++) score[u] = (float)class_size[P.cls[u]];
  uint32_t used = 0, frontier = 0;
  for (uint32_t i = 0; i < k; i++) {
    uint32_t pool = i == 0 ? ((k == 32) ? 0xffffffffu : ((1u << k) - 1)) : (frontier & ~used);
    if (pool == 0) pool = ((k == 32) ? 0xffffffffu : ((1u << k) - 1)) & ~used;  // disconnected
    uint32_t best = FSIG_INVALID;
    for (uint32_t u = 0; u < k; u++)
      if ((pool >> u) & 1)
        if (best == FSIG_INVALID || score[u] < score[best]) best = u;
    P.order[i] = best;
    P.pos[best] = i;
    used |= 1u << best;
    frontier |= P.adjm[best];
    for (uint32_t w = 0; w < k; w++)
      if ((P.adjm[best] >> w) & 1) score[w] *= 0.9f;
  }
}

// ---------------------------------------------------------------------------
// Symmetry breaking for counts (Grochow and Kellis, "Network motif discovery
// using subgraph enumeration and symmetry-breaking", RECOMB 2007).  An
// automorphism s of the query (a permutation of its vertices that keeps
// labels and edges) turns every embedding phi into the embedding phi o s,
// and distinct automorphisms give distinct embeddings, so the embeddings
// fall into orbits of exactly |Aut| each.  Walking down a stabiliser chain
// (take a vertex v of largest orbit under the automorphisms that fix the
// vertices taken so far, ask phi(v) < phi(w) for the other members w of that
// orbit, then fix v) yields conditions that exactly one embedding of every
// orbit satisfies, and |Aut| is the product of the orbit sizes taken.
// Automorphic vertices share a refinement class, hence a candidate array,
// so a condition compares two row words of the same class.

namespace {

struct AutSearch {
  const QueryPlan& P;
  uint32_t img[MAXQ];
  uint64_t steps = 0;
  static constexpr uint64_t kBudget = 1u << 22;
  explicit AutSearch(const QueryPlan& p) : P(p) {}

  // Extends the partial permutation img[0..u) (images in `used`) to a full
  // automorphism; pre[u] != FSIG_INVALID prescribes u's image.
  bool extend(uint32_t u, uint32_t used, const uint32_t* pre) {
    if (u == P.k) return true;
    if (++steps > kBudget) return false;
    for (uint32_t t = 0; t < P.k; t++) {
      if ((used >> t) & 1u) continue;
      if (pre[u] != FSIG_INVALID && t != pre[u]) continue;
      if (P.cls[t] != P.cls[u]) continue;  // classes are kept by automorphisms
      bool ok = true;
      for (uint32_t a = 0; a < u && ok; a++)
        ok = (((P.adjm[u] >> a) & 1u) == ((P.adjm[t] >> img[a]) & 1u));
      if (!ok) continue;
      img[u] = t;
      if (extend(u + 1, used | (1u << t), pre)) return true;
      if (steps > kBudget) return false;
    }
    return false;
  }
};

}  // namespace

bool plan_symmetry(QueryPlan& P) {
  const uint32_t k = P.k;
  P.aut = 1;
  for (uint32_t u = 0; u < MAXQ; u++) P.sym_lt[u] = 0;
  AutSearch S(P);
  uint32_t fixed = 0;
  uint32_t lt[MAXQ] = {0};
  uint64_t aut = 1;
  for (;;) {
    // orbits under the automorphisms fixing `fixed` pointwise
    uint32_t orbit[MAXQ];
    for (uint32_t v = 0; v < k; v++) orbit[v] = 1u << v;
    uint32_t pre[MAXQ];
    for (uint32_t v = 0; v < k; v++) {
      if ((fixed >> v) & 1u) continue;
      for (uint32_t w = v + 1; w < k; w++) {
        if (((fixed >> w) & 1u) || P.cls[w] != P.cls[v] || ((orbit[v] >> w) & 1u)) continue;
        for (uint32_t x = 0; x < k; x++) pre[x] = ((fixed >> x) & 1u) ? x : FSIG_INVALID;
        pre[v] = w;
        const bool found = S.extend(0, 0, pre);
        if (S.steps > AutSearch::kBudget) return false;
        if (!found) continue;
        const uint32_t merged = orbit[v] | orbit[w];
        for (uint32_t x = 0; x < k; x++)
          if ((merged >> x) & 1u) orbit[x] = merged;
      }
    }
    uint32_t best = FSIG_INVALID, best_n = 1;
    for (uint32_t v = 0; v < k; v++) {
      if ((fixed >> v) & 1u) continue;
      const uint32_t n = (uint32_t)__builtin_popcount(orbit[v]);
      if (n > best_n) { best_n = n; best = v; }
    }
    if (best == FSIG_INVALID) break;
    lt[best] |= orbit[best] & ~(1u << best);
    aut *= best_n;
    fixed |= 1u << best;
  }
  P.aut = aut;
  for (uint32_t u = 0; u < k; u++) P.sym_lt[u] = lt[u];
  return true;
}
