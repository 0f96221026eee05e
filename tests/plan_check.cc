// Randomised check of the host query planner (fsig_amd/csrc/plan.cc), built
// with g++ and AddressSanitizer / UndefinedBehaviorSanitizer by
// tests/test_plan_host.py.  Random queries (k 1..32, edge density 0..1,
// 1..5 labels, connected or not) must give: alg 1 classes whose thresholds
// thr[a][b] equal every member's neighbour count in class b and whose
// neighbour-label requirements every member meets exactly; alg 2 classes of
// equal signatures; join orders (both algorithms) that are permutations with
// pos the inverse of order.  Usage: plan_check <iterations> <seed>.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>
#include "internal.h"
int main(int argc, char** argv) {
  const long iters = argc > 1 ? atol(argv[1]) : 20000;
  std::mt19937_64 rng(argc > 2 ? strtoull(argv[2], nullptr, 10) : 12345);
  long checked = 0;
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
    checked++;
  }
  printf("checked %ld queries\n", checked);
  return 0;
}
