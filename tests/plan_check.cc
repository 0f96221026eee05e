// Randomised check of the host query planner (fsig_amd/csrc/plan.cc), built
// with g++ and AddressSanitizer / UndefinedBehaviorSanitizer by
// tests/test_plan_host.py.  Random queries (k 1..32, edge density 0..1,
// 1..5 labels, connected or not) must give: alg 1 classes whose thresholds
// thr[a][b] equal every member's neighbour count in class b and whose
// neighbour-label requirements every member meets exactly; alg 2 classes of
// equal signatures; join orders (both algorithms) that are permutations with
// pos the inverse of order; symmetry-breaking conditions (plan_symmetry)
// whose |Aut| equals brute force (and the known orders of named graphs) and
// that exactly one automorphic image of every injective assignment
// satisfies.  Usage: plan_check <iterations> <seed>.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>
#include "internal.h"

// Every automorphism of a small query (label- and edge-preserving
// permutations), by backtracking.
static void all_auts(uint32_t k, const std::vector<uint32_t>& adjm, const std::vector<int32_t>& lab,
                     std::vector<uint32_t>& img, uint32_t u, uint32_t used, std::vector<std::vector<uint32_t>>& out) {
  if (u == k) { out.push_back(img); return; }
  for (uint32_t t = 0; t < k; t++) {
    if (((used >> t) & 1u) || lab[t] != lab[u]) continue;
    bool ok = true;
    for (uint32_t a = 0; a < u && ok; a++) ok = ((adjm[u] >> a) & 1u) == ((adjm[t] >> img[a]) & 1u);
    if (!ok) continue;
    img[u] = t;
    all_auts(k, adjm, lab, img, u + 1, used | (1u << t), out);
  }
}

// The symmetry conditions of plan_symmetry against brute force: |Aut| is the
// group's order, and for distinct values on the query vertices exactly one
// automorphism s makes val o s satisfy every condition phi(a) < phi(b).
static bool check_symmetry(uint32_t k, const std::vector<uint32_t>& adjm, const std::vector<int32_t>& lab,
                           std::mt19937_64& rng, uint64_t want_aut, const char* what) {
  std::vector<uint32_t> off{0}, adj;
  for (uint32_t a = 0; a < k; a++) {
    for (uint32_t b = 0; b < k; b++) if ((adjm[a] >> b) & 1u) adj.push_back(b);
    off.push_back((uint32_t)adj.size());
  }
  if (adj.empty()) adj.push_back(0);
  QueryPlan P;
  std::string err;
  if (plan_classes(P, k, lab.data(), off.data(), adj.data(), nullptr, FSIG_ALG_MMATCH, err)) {
    printf("%s: plan failed %s\n", what, err.c_str());
    return false;
  }
  if (!plan_symmetry(P)) { printf("%s: automorphism search over budget\n", what); return false; }
  std::vector<std::vector<uint32_t>> auts;
  std::vector<uint32_t> img(k);
  all_auts(k, adjm, lab, img, 0, 0, auts);
  if (P.aut != auts.size() || (want_aut && P.aut != want_aut)) {
    printf("%s: |Aut| %llu, brute force %zu, expected %llu\n", what, (unsigned long long)P.aut, auts.size(),
           (unsigned long long)want_aut);
    return false;
  }
  for (uint32_t a = 0; a < k; a++)
    for (uint32_t b = 0; b < k; b++)
      if (((P.sym_lt[a] >> b) & 1u) && (a == b || P.cls[a] != P.cls[b])) {
        printf("%s: condition %u < %u across classes\n", what, a, b);
        return false;
      }
  for (int trial = 0; trial < 4; trial++) {
    std::vector<uint64_t> val(k);
    for (auto& v : val) v = rng();
    size_t pass = 0;
    for (const auto& s : auts) {
      bool ok = true;
      for (uint32_t a = 0; a < k && ok; a++)
        for (uint32_t b = 0; b < k && ok; b++)
          if ((P.sym_lt[a] >> b) & 1u) ok = val[s[a]] < val[s[b]];
      pass += ok;
    }
    if (pass != 1) { printf("%s: %zu automorphic images satisfy the conditions\n", what, pass); return false; }
  }
  return true;
}

static std::vector<uint32_t> graph_of(uint32_t k, std::initializer_list<std::pair<int, int>> edges) {
  std::vector<uint32_t> m(k, 0);
  for (auto e : edges) { m[e.first] |= 1u << e.second; m[e.second] |= 1u << e.first; }
  return m;
}

int main(int argc, char** argv) {
  const long iters = argc > 1 ? atol(argv[1]) : 20000;
  std::mt19937_64 rng(argc > 2 ? strtoull(argv[2], nullptr, 10) : 12345);
  long checked = 0;
  {
    // automorphism groups of known order
    struct Named { const char* what; uint32_t k; std::vector<uint32_t> adjm; uint64_t aut; };
    std::vector<Named> named = {
        {"triangle", 3, graph_of(3, {{0, 1}, {1, 2}, {2, 0}}), 6},
        {"4-cycle", 4, graph_of(4, {{0, 1}, {1, 2}, {2, 3}, {3, 0}}), 8},
        {"4-clique", 4, graph_of(4, {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}}), 24},
        {"house", 5, graph_of(5, {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {2, 4}, {3, 4}}), 2},
        {"path 4", 4, graph_of(4, {{0, 1}, {1, 2}, {2, 3}}), 2},
        {"star 1+5", 6, graph_of(6, {{0, 1}, {0, 2}, {0, 3}, {0, 4}, {0, 5}}), 120},
        {"K3,3", 6, graph_of(6, {{0, 3}, {0, 4}, {0, 5}, {1, 3}, {1, 4}, {1, 5}, {2, 3}, {2, 4}, {2, 5}}), 72},
        {"cube", 8, graph_of(8, {{0, 1}, {1, 3}, {3, 2}, {2, 0}, {4, 5}, {5, 7}, {7, 6}, {6, 4},
                                 {0, 4}, {1, 5}, {2, 6}, {3, 7}}), 48},
        {"Petersen", 10, graph_of(10, {{0, 1}, {1, 2}, {2, 3}, {3, 4}, {4, 0}, {0, 5}, {1, 6}, {2, 7}, {3, 8},
                                       {4, 9}, {5, 7}, {7, 9}, {9, 6}, {6, 8}, {8, 5}}), 120},
    };
    for (const auto& g : named)
      if (!check_symmetry(g.k, g.adjm, std::vector<int32_t>(g.k, 0), rng, g.aut, g.what)) return 1;
    // labels split the 4-cycle's group: labels a b a b keep 4 of its 8
    if (!check_symmetry(4, named[1].adjm, {0, 1, 0, 1}, rng, 4, "4-cycle abab")) return 1;
  }
  for (long it = 0; it < iters; it++) {
    uint32_t k = 1 + rng() % 32;
    uint32_t L = 1 + rng() % 5;
    double p = (rng() % 100) / 100.0;
    std::vector<uint32_t> adjm(k, 0);
    for (uint32_t a = 0; a < k; a++)
      for (uint32_t b = a + 1; b < k; b++)
        if ((rng() % 1000) < p * 1000) { adjm[a] |= 1u << b; adjm[b] |= 1u << a; }
    std::vector<int32_t> lab(k);
    for (auto& l : lab) l = (int32_t)(rng() % L);
    std::vector<uint32_t> off{0}, adj;
    for (uint32_t a = 0; a < k; a++) {
      for (uint32_t b = 0; b < k; b++) if ((adjm[a] >> b) & 1u) adj.push_back(b);
      off.push_back((uint32_t)adj.size());
    }
    if (adj.empty()) adj.push_back(0);
    QueryPlan P;
    std::string err;
    const int alg = (it % 4 == 0) ? FSIG_ALG_GSI : FSIG_ALG_MMATCH;
    std::vector<uint32_t> sig(16 * k, 0);
    for (uint32_t u = 0; u < k; u++) sig[16 * u] = (uint32_t)lab[u];
    int rc = plan_classes(P, k, lab.data(), off.data(), adj.data(), sig.data(), alg, err);
    if (rc) { printf("rc %d %s at it %ld\n", rc, err.c_str(), it); return 1; }
    // partition + thresholds
    for (uint32_t u = 0; u < k; u++) {
      if (P.cls[u] >= P.nclass) { printf("bad class\n"); return 1; }
      if (P.label[u] != lab[u]) { printf("bad label\n"); return 1; }
      if (alg == FSIG_ALG_GSI) {
        for (int w = 0; w < 16; w++)
          if (P.class_sig[P.cls[u]][w] != sig[16 * u + w]) { printf("gsi class sig mismatch\n"); return 1; }
        continue;
      }
      // NLF: every member of the class meets the representative's requirements exactly
      for (uint32_t j = 0; j < P.nlf_n[P.cls[u]]; j++) {
        uint32_t l = P.nlf_req[P.cls[u]][j] >> 8, need = P.nlf_req[P.cls[u]][j] & 255u, have = 0;
        for (uint32_t w = 0; w < k; w++) if (((adjm[u] >> w) & 1u) && (uint32_t)lab[w] == l) have++;
        if (have != need) { printf("nlf mismatch it %ld\n", it); return 1; }
      }
      for (uint32_t b = 0; b < P.nclass; b++) {
        uint32_t c = 0;
        for (uint32_t w = 0; w < k; w++) if (((adjm[u] >> w) & 1u) && P.cls[w] == b) c++;
        if (c != P.thr[P.cls[u]][b]) {
          printf("thr mismatch it %ld alg %d k %u nclass %u u %u cls %u b %u count %u thr %u\n", it, alg, k, P.nclass, u, P.cls[u], b, c, P.thr[P.cls[u]][b]);
          for (uint32_t w = 0; w < k; w++) printf("v%u lab %d cls %u adj %x\n", w, lab[w], P.cls[w], adjm[w]);
          return 1;
        }
      }
      if (P.class_label[P.cls[u]] != lab[u]) { printf("class label\n"); return 1; }
    }
    std::vector<uint32_t> sizes(P.nclass);
    for (auto& s : sizes) s = rng() % 100000;
    for (int ord = 0; ord < 2; ord++) {
      if (ord == 0) plan_order_mmatch(P, sizes.data()); else plan_order_gsi(P, sizes.data());
      uint32_t seen = 0;
      for (uint32_t i = 0; i < k; i++) {
        uint32_t u = P.order[i];
        if (u >= k || ((seen >> u) & 1u) || P.pos[u] != i) { printf("bad order it %ld\n", it); return 1; }
        seen |= 1u << u;
      }
    }
    if (k <= 7 && alg == FSIG_ALG_MMATCH && !check_symmetry(k, adjm, lab, rng, 0, "random query")) {
      printf("at it %ld\n", it);
      return 1;
    }
    checked++;
  }
  printf("checked %ld queries\n", checked);
  return 0;
}
