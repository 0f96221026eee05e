// ref_dump: links the reference's own host sources (compiled where they lie
// under the reference tree: graph/graph.cc, IO/graph_reader.cc,
// graph_format_transformer/graph_format_transformer.cc) and prints what they
// build for a graph file, so tests can compare the product's reader, CSR,
// PCSR, signature and MurmurHash2 with the reference bit for bit.
// Test infrastructure only (built into oracle/_ref/, never shipped).
//
// Output: whitespace-separated records, each introduced by a tag line:
//   CSR n m L / offsets / edges / labels
//   INV index... / list...
//   PCSR units / then per unit: UNIT keys nedges / offset words / edges
//   SIGQ (row-major [n][16]) / SIGD (column-major [16][n])
//   MURMUR count / hashes of the 4-byte keys 0..count-1 with seed 17
// With "time" as the third argument it prints only the wall time of each
// stage the reference runs for alg 2 (`TIME <stage> <seconds>`: read, pcsr,
// csr, signature), the host-ingest baseline.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "IO/graph_reader.h"
#include "graph/graph.h"
#include "graph_format_transformer/graph_format_transformer.h"
#include "utility/inverted_index.h"

static void dump(const char* tag, const unsigned* a, size_t n) {
  printf("%s %zu\n", tag, n);
  for (size_t i = 0; i < n; i++) printf(i + 1 < n ? "%u " : "%u", a[i]);
  printf("\n");
}

static int time_stages(const char* path, bool labeled) {
  using clk = std::chrono::steady_clock;
  auto secs = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); };
  auto t0 = clk::now();
  UndirectedSurveyPlainReader reader;
  Graph* g = nullptr;
  reader.read(path, g, labeled);
  UndirectedPlainGraph* plain = dynamic_cast<UndirectedPlainGraph*>(g);
  auto t1 = clk::now();
  Graph* pg = nullptr;
  UndirectedPlainToPCSRTransformer pt;
  pt.transform(plain, pg);
  auto t2 = clk::now();
  Graph* cg = nullptr;
  UndirectedPlainToCSRTransformer ct;
  ct.transform(plain, cg);
  auto t3 = clk::now();
  Vertex* sd = nullptr;
  UndirectedPlainToPCSRTransformer::build_signature(true, sd, plain);
  auto t4 = clk::now();
  printf("TIME read %.6f\nTIME pcsr %.6f\nTIME csr %.6f\nTIME signature %.6f\n", secs(t0, t1), secs(t1, t2),
         secs(t2, t3), secs(t3, t4));
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: ref_dump <graph file> <labeled 0|1> [murmur_count | time]\n");
    return 2;
  }
  const bool labeled = atoi(argv[2]) != 0;
  if (argc > 3 && std::strcmp(argv[3], "time") == 0) return time_stages(argv[1], labeled);
  const unsigned murmur_count = argc > 3 ? (unsigned)atoi(argv[3]) : 64;
  UndirectedSurveyPlainReader reader;
  Graph* g = nullptr;
  reader.read(argv[1], g, labeled);
  UndirectedPlainGraph* plain = dynamic_cast<UndirectedPlainGraph*>(g);

  // PCSR first: like main() for alg 2, it sorts the plain adjacency in place.
  Graph* pg = nullptr;
  UndirectedPlainToPCSRTransformer pt;
  pt.transform(plain, pg);
  UndirectedPCSRGraph* pcsr = dynamic_cast<UndirectedPCSRGraph*>(pg);

  Graph* cg = nullptr;
  UndirectedPlainToCSRTransformer ct;
  ct.transform(plain, cg);
  UndirectedCSRGraph* csr = dynamic_cast<UndirectedCSRGraph*>(cg);
  const unsigned n = csr->num_vertices(), m = csr->num_edges();
  printf("CSR %u %u %d\n", n, m, (int)csr->num_labels());
  dump("OFF", csr->offset_, n + 1);
  dump("ADJ", csr->edges_, 2ull * m);
  dump("LAB", (const unsigned*)csr->labels_, n);

  InvertedIndex<Label, Vertex> inv(n, csr->num_labels(), csr->labels_);
  dump("INVIDX", inv.index_, (size_t)csr->num_labels() + 1);
  dump("INVLIST", inv.list_, n);

  printf("PCSR %u\n", pcsr->num_units_);
  for (unsigned u = 0; u < pcsr->num_units_; u++) {
    PCSRUnit* unit = pcsr->units_[u];
    printf("UNIT %u %u\n", unit->num_keys_, unit->num_edges_);
    dump("UOFF", unit->offset_, 32ull * unit->num_keys_);
    dump("UEDGE", unit->edges_, unit->num_edges_);
  }

  Vertex* sq = nullptr;
  Vertex* sd = nullptr;
  UndirectedPlainToPCSRTransformer::build_signature(false, sq, plain);
  UndirectedPlainToPCSRTransformer::build_signature(true, sd, plain);
  dump("SIGQ", sq, 16ull * n);
  dump("SIGD", sd, 16ull * n);

  printf("MURMUR %u\n", murmur_count);
  for (unsigned i = 0; i < murmur_count; i++) {
    unsigned key = i;
    printf(i + 1 < murmur_count ? "%u " : "%u", UndirectedPlainToPCSRTransformer::MurmurHash2(&key, 4, HASHSEED));
  }
  printf("\n");
  return 0;
}
