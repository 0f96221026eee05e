// C ABI of libfsig_host.so (include/fsig_host.h).
#include "../../include/fsig_host.h"

#include <algorithm>
#include <cstdio>
#include <memory>
#include <mutex>
#include <thread>

#include "gen.h"
#include "graph.h"
#include "io.h"
#include "transform.h"

using namespace fsig;

struct fsig_hgraph {
  uint32_t n = 0;
  int32_t num_labels = 1;
  std::vector<uint32_t> off{0};
  std::vector<uint32_t> adj;
  std::vector<int32_t> labels;
  // walk starts: the non-isolated vertices, built on the first walk query
  mutable std::vector<uint32_t> walk_starts;
  mutable bool walk_starts_ready = false;
  mutable std::mutex walk_mu;
};

struct fsig_hpcsr {
  std::vector<PCSRUnit> units;
  // graph header and labels when the units come from a cache file
  uint32_t n = 0, num_edges = 0;
  int32_t num_labels = 0;
  std::vector<int32_t> labels;
};

// Binary caches.  Every field is 4 bytes, native byte order, no padding.
//   PCSR (the reference's commented STORE / READ blocks,
//   graph_format_transformer/graph_format_transformer.cc:47-74, 197-213):
//     u32 num_vertices, u32 num_edges, i32 num_labels, i32 labels[num_vertices],
//     u32 num_units, then per unit: u32 num_keys, u32 num_edges,
//     u32 offset[32 * num_keys], u32 edges[num_edges]
//   CSR (the same header): u32 num_vertices, u32 num_edges (undirected),
//     i32 num_labels, i32 labels[num_vertices], u32 offset[num_vertices + 1],
//     u32 adjacency[2 * num_edges]
namespace {

struct BinFile {
  FILE* f = nullptr;
  explicit BinFile(const char* path, const char* mode) : f(std::fopen(path, mode)) {}
  ~BinFile() {
    if (f) std::fclose(f);
  }
  template <class T>
  bool put(const T* p, size_t n) { return n == 0 || std::fwrite(p, sizeof(T), n, f) == n; }
  template <class T>
  bool get(T* p, size_t n) { return n == 0 || std::fread(p, sizeof(T), n, f) == n; }
  bool at_end() { return std::fgetc(f) == EOF; }
  // true when at least `bytes` remain after the read position: a corrupt
  // header is rejected before it sizes a buffer
  bool has(uint64_t bytes) {
    const long at = std::ftell(f);
    if (at < 0 || std::fseek(f, 0, SEEK_END) != 0) return false;
    const long end = std::ftell(f);
    if (end < 0 || std::fseek(f, at, SEEK_SET) != 0) return false;
    return (uint64_t)(end - at) >= bytes;
  }
};

}  // namespace

// Frees the adjacency lists of a large plain graph on several threads (one
// free per vertex, millions of them at the largest sizes).
static void release_lists(UndirectedPlainGraph& g) {
  const size_t n = g.edges_.size();
  const unsigned T = n > (1u << 20) ? std::max(1u, std::min(16u, std::thread::hardware_concurrency())) : 1u;
  auto drop = [&](size_t a, size_t b) {
    for (size_t v = a; v < b; v++) std::vector<Vertex>().swap(g.edges_[v]);
  };
  if (T == 1) {
    drop(0, n);
  } else {
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T; t++) th.emplace_back(drop, n * t / T, n * (t + 1) / T);
    for (auto& x : th) x.join();
  }
  std::vector<std::vector<Vertex>>().swap(g.edges_);
}

static void fix_num_labels(fsig_hgraph* g) {
  int32_t mx = 0;
  for (int32_t l : g->labels) mx = std::max(mx, l);
  g->num_labels = mx + 1;
}

extern "C" {

fsig_hgraph* fsig_hg_read(const char* path, int is_labeled, int* status) {
  UndirectedPlainGraph plain;
  int rc = read_plain_graph(path, is_labeled != 0, plain);
  if (status) *status = rc;
  if (rc) return nullptr;
  Graph* out = nullptr;
  UndirectedPlainToCSRTransformer t(true);
  if (t.transform(&plain, out)) return nullptr;
  std::unique_ptr<UndirectedCSRGraph> csr(static_cast<UndirectedCSRGraph*>(out));
  release_lists(plain);
  auto* g = new fsig_hgraph();
  g->n = csr->num_vertices_;
  g->off = std::move(csr->offset_);
  g->adj = std::move(csr->edges_);
  g->labels = std::move(csr->labels_);
  g->num_labels = std::max<int32_t>(csr->num_labels_, 1);
  return g;
}

fsig_hgraph* fsig_hg_from_csr(uint32_t n, const uint32_t* off, const uint32_t* adj, const int32_t* labels) {
  auto* g = new fsig_hgraph();
  g->n = n;
  g->off.assign(off, off + n + 1);
  g->adj.assign(adj, adj + off[n]);
  g->labels.assign(labels, labels + n);
  fix_num_labels(g);
  return g;
}

fsig_hgraph* fsig_hg_from_edges(uint32_t n, uint64_t m, const uint32_t* pairs, const int32_t* labels) {
  EdgeList e(m);
  for (uint64_t i = 0; i < m; i++) e[i] = {pairs[2 * i], pairs[2 * i + 1]};
  for (auto& x : e)
    if (x.first >= n || x.second >= n) return nullptr;
  auto* g = new fsig_hgraph();
  g->n = n;
  csr_from_edge_list(n, e, g->off, g->adj);
  if (labels) g->labels.assign(labels, labels + n); else g->labels.assign(n, 0);
  fix_num_labels(g);
  return g;
}

fsig_hgraph* fsig_hg_chung_lu(uint32_t n, uint64_t m, double gamma, uint64_t seed) {
  EdgeList e = gen_chung_lu(n, m, gamma, seed);
  auto* g = new fsig_hgraph();
  g->n = n;
  csr_from_edge_list(n, e, g->off, g->adj);
  g->labels.assign(n, 0);
  return g;
}

fsig_hgraph* fsig_hg_rmat(uint32_t scale, uint64_t m, double a, double b, double c, uint64_t seed) {
  EdgeList e = gen_rmat(scale, m, a, b, c, seed);
  auto* g = new fsig_hgraph();
  g->n = 1u << scale;
  csr_from_edge_list(g->n, e, g->off, g->adj);
  g->labels.assign(g->n, 0);
  return g;
}

int fsig_hg_set_labels(fsig_hgraph* g, int kind, int32_t num_labels, double s, uint64_t seed) {
  if (!g || num_labels <= 0) return -1;
  g->labels = kind == 0 ? labels_uniform(g->n, num_labels, seed) : labels_zipf(g->n, num_labels, s, seed);
  g->num_labels = num_labels;
  return 0;
}

int fsig_hg_write(const fsig_hgraph* g, const char* path) {
  return write_graph(path, g->n, g->labels.data(), g->off.data(), g->adj.data());
}

void fsig_hg_info(const fsig_hgraph* g, uint32_t* n, uint64_t* nnz, int32_t* num_labels, uint32_t* max_degree) {
  if (n) *n = g->n;
  if (nnz) *nnz = g->off[g->n];
  if (num_labels) *num_labels = g->num_labels;
  if (max_degree) {
    uint32_t md = 0;
    for (uint32_t v = 0; v < g->n; v++) md = std::max(md, g->off[v + 1] - g->off[v]);
    *max_degree = md;
  }
}

void fsig_hg_export(const fsig_hgraph* g, uint32_t* off, uint32_t* adj, int32_t* labels) {
  if (off) std::copy(g->off.begin(), g->off.end(), off);
  if (adj) std::copy(g->adj.begin(), g->adj.end(), adj);
  if (labels) std::copy(g->labels.begin(), g->labels.end(), labels);
}

const uint32_t* fsig_hg_off(const fsig_hgraph* g) { return g->off.data(); }
const uint32_t* fsig_hg_adj(const fsig_hgraph* g) { return g->adj.data(); }
const int32_t* fsig_hg_labels(const fsig_hgraph* g) { return g->labels.data(); }

void fsig_hg_inverted_index(const fsig_hgraph* g, uint32_t* index, uint32_t* list) {
  InvertedIndex<int32_t, uint32_t> inv(g->n, g->num_labels, g->labels.data());
  std::copy(inv.index_.begin(), inv.index_.end(), index);
  std::copy(inv.list_.begin(), inv.list_.end(), list);
}

void fsig_hg_signature(const fsig_hgraph* g, int column_oriented, int x86_shift, uint32_t* out) {
  auto s = signature_from_csr(g->n, g->off.data(), g->adj.data(), g->labels.data(), column_oriented != 0,
                              x86_shift != 0);
  std::copy(s.begin(), s.end(), out);
}

void fsig_hg_free(fsig_hgraph* g) { delete g; }

int fsig_hg_walk_query(const fsig_hgraph* g, uint32_t k, uint32_t min_e, uint32_t max_e, uint64_t seed, int bias,
                       uint32_t max_attempts, int32_t* labels, uint32_t* edges, uint32_t* walk) {
  WalkQuery q;
  {
    std::lock_guard<std::mutex> lk(g->walk_mu);
    if (!g->walk_starts_ready) {
      for (uint32_t v = 0; v < g->n; v++)
        if (g->off[v + 1] > g->off[v]) g->walk_starts.push_back(v);
      g->walk_starts_ready = true;
    }
  }
  if (random_walk_query(g->n, g->off.data(), g->adj.data(), g->labels.data(), k, min_e, max_e, seed, bias,
                        max_attempts, q, &g->walk_starts))
    return -1;
  for (uint32_t i = 0; i < k; i++) {
    labels[i] = q.labels[i];
    walk[i] = q.data_vertices[i];
  }
  for (size_t i = 0; i < q.edges.size(); i++) {
    edges[2 * i] = q.edges[i].first;
    edges[2 * i + 1] = q.edges[i].second;
  }
  return (int)q.edges.size();
}

fsig_hpcsr* fsig_pcsr_build(const fsig_hgraph* g, int reference_unit_ids) {
  if (!g || g->num_labels > kPcsrMaxLabels) return nullptr;
  auto* p = new fsig_hpcsr();
  pcsr_from_csr(g->n, g->off.data(), g->adj.data(), g->labels.data(), g->num_labels, reference_unit_ids != 0,
                p->units);
  return p;
}

uint32_t fsig_pcsr_num_units(const fsig_hpcsr* p) { return (uint32_t)p->units.size(); }

void fsig_pcsr_unit_info(const fsig_hpcsr* p, uint32_t u, uint32_t* keys, uint32_t* nedges) {
  if (keys) *keys = p->units[u].num_keys_;
  if (nedges) *nedges = p->units[u].num_edges_;
}

const uint32_t* fsig_pcsr_unit_offset(const fsig_hpcsr* p, uint32_t u) { return p->units[u].offset_.data(); }
const uint32_t* fsig_pcsr_unit_edges(const fsig_hpcsr* p, uint32_t u) { return p->units[u].edges_.data(); }
void fsig_pcsr_free(fsig_hpcsr* p) { delete p; }

int fsig_hg_save_bin(const fsig_hgraph* g, const char* path) {
  if (!g || !path) return -1;
  BinFile b(path, "wb");
  if (!b.f) return -1;
  const uint32_t m = g->off[g->n] / 2;
  const bool ok = b.put(&g->n, 1) && b.put(&m, 1) && b.put(&g->num_labels, 1) && b.put(g->labels.data(), g->n) &&
                  b.put(g->off.data(), g->n + 1ull) && b.put(g->adj.data(), g->adj.size());
  return ok ? 0 : -1;
}

fsig_hgraph* fsig_hg_load_bin(const char* path, int* status) {
  auto fail = [&](int code) -> fsig_hgraph* {
    if (status) *status = code;
    return nullptr;
  };
  BinFile b(path, "rb");
  if (!b.f) return fail(kIoCannotOpen);
  auto g = std::make_unique<fsig_hgraph>();
  uint32_t m = 0;
  if (!b.get(&g->n, 1) || !b.get(&m, 1) || !b.get(&g->num_labels, 1)) return fail(kIoBadHeader);
  if (!b.has(4ull * g->n + 4ull * (g->n + 1ull) + 8ull * m)) return fail(kIoBadHeader);
  g->labels.resize(g->n);
  g->off.resize(g->n + 1ull);
  if (!b.get(g->labels.data(), g->n) || !b.get(g->off.data(), g->n + 1ull)) return fail(kIoBadRecord);
  if (g->off[0] != 0 || g->off[g->n] != 2ull * m) return fail(kIoBadRecord);
  g->adj.resize(2ull * m);
  if (!b.get(g->adj.data(), g->adj.size()) || !b.at_end()) return fail(kIoBadRecord);
  for (uint32_t v = 0; v < g->n; v++)
    if (g->off[v + 1] < g->off[v] || g->labels[v] < 0 || g->labels[v] >= g->num_labels) return fail(kIoBadRecord);
  for (uint32_t x : g->adj)
    if (x >= g->n) return fail(kIoEdgeOutOfRange);
  if (status) *status = 0;
  return g.release();
}

int fsig_pcsr_save_bin(const fsig_hpcsr* p, const fsig_hgraph* g, const char* path) {
  if (!p || !g || !path) return -1;
  BinFile b(path, "wb");
  if (!b.f) return -1;
  const uint32_t m = g->off[g->n] / 2, nu = (uint32_t)p->units.size();
  bool ok = b.put(&g->n, 1) && b.put(&m, 1) && b.put(&g->num_labels, 1) && b.put(g->labels.data(), g->n) &&
            b.put(&nu, 1);
  for (const PCSRUnit& u : p->units)
    ok = ok && b.put(&u.num_keys_, 1) && b.put(&u.num_edges_, 1) && b.put(u.offset_.data(), 32ull * u.num_keys_) &&
         b.put(u.edges_.data(), u.num_edges_);
  return ok ? 0 : -1;
}

fsig_hpcsr* fsig_pcsr_load_bin(const char* path, int* status) {
  auto fail = [&](int code) -> fsig_hpcsr* {
    if (status) *status = code;
    return nullptr;
  };
  BinFile b(path, "rb");
  if (!b.f) return fail(kIoCannotOpen);
  auto p = std::make_unique<fsig_hpcsr>();
  uint32_t nu = 0;
  if (!b.get(&p->n, 1) || !b.get(&p->num_edges, 1) || !b.get(&p->num_labels, 1)) return fail(kIoBadHeader);
  if (!b.has(4ull * p->n + 4)) return fail(kIoBadHeader);
  p->labels.resize(p->n);
  if (!b.get(p->labels.data(), p->n) || !b.get(&nu, 1)) return fail(kIoBadRecord);
  if (p->num_labels < 1 || nu != (uint64_t)p->num_labels * (p->num_labels + 1) / 2) return fail(kIoBadHeader);
  if (!b.has(8ull * nu)) return fail(kIoBadHeader);
  p->units.resize(nu);
  for (PCSRUnit& u : p->units) {
    if (!b.get(&u.num_keys_, 1) || !b.get(&u.num_edges_, 1)) return fail(kIoBadRecord);
    if (!b.has(128ull * u.num_keys_ + 4ull * u.num_edges_)) return fail(kIoBadRecord);
    u.offset_.resize(32ull * u.num_keys_);
    u.edges_.resize(u.num_edges_);
    if (!b.get(u.offset_.data(), u.offset_.size()) || !b.get(u.edges_.data(), u.edges_.size()))
      return fail(kIoBadRecord);
  }
  if (!b.at_end()) return fail(kIoBadRecord);
  if (status) *status = 0;
  return p.release();
}

void fsig_pcsr_graph_info(const fsig_hpcsr* p, uint32_t* n, uint32_t* num_edges, int32_t* num_labels) {
  if (n) *n = p->n;
  if (num_edges) *num_edges = p->num_edges;
  if (num_labels) *num_labels = p->num_labels;
}

const int32_t* fsig_pcsr_labels(const fsig_hpcsr* p) { return p->labels.data(); }

uint32_t fsig_murmur2(const void* key, int len, uint32_t seed) { return murmur_hash2(key, len, seed); }

int fsig_write_embeddings(const char* path, uint32_t k, uint64_t count, const uint32_t* rows, uint64_t nrows) {
  std::vector<uint32_t> r(rows, rows + nrows * k);
  return write_embeddings(path, k, count, r);
}

}  // extern "C"
