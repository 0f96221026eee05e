// Storage transforms (CSR, PCSR, signatures) and MurmurHash2.
#include "transform.h"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>

namespace fsig {

uint32_t murmur_hash2(const void* key, int len, uint32_t seed) {
  const uint32_t mul = 0x5bd1e995u;
  const unsigned char* bytes = static_cast<const unsigned char*>(key);
  uint32_t h = seed ^ (uint32_t)len;
  int rest = len;
  while (rest >= 4) {
    uint32_t w;
    std::memcpy(&w, bytes, 4);
    w *= mul;
    w ^= w >> 24;
    w *= mul;
    h = (h * mul) ^ w;
    bytes += 4;
    rest -= 4;
  }
  if (rest > 0) {
    if (rest >= 3) h ^= (uint32_t)bytes[2] << 16;
    if (rest >= 2) h ^= (uint32_t)bytes[1] << 8;
    h ^= bytes[0];
    h *= mul;
  }
  h ^= h >> 13;
  h *= mul;
  h ^= h >> 15;
  return h;
}

static void sort_ranges(uint32_t n, const std::vector<uint32_t>& off, std::vector<uint32_t>& adj,
                        int threads) {
  if (threads <= 0) threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 100000) threads = 1;
  auto work = [&](uint32_t lo, uint32_t hi) {
    for (uint32_t v = lo; v < hi; v++) std::sort(adj.begin() + off[v], adj.begin() + off[v + 1]);
  };
  if (threads == 1) { work(0, n); return; }
  std::vector<std::thread> pool;
  uint32_t step = (n + threads - 1) / threads;
  for (int t = 0; t < threads; t++) {
    uint32_t lo = std::min<uint64_t>((uint64_t)t * step, n), hi = std::min<uint64_t>((uint64_t)(t + 1) * step, n);
    pool.emplace_back(work, lo, hi);
  }
  for (auto& th : pool) th.join();
}

// Every list ends sorted and free of repeats, so the scatter may fill a list
// in any order.  Large edge lists are counted by threads over edge ranges
// (relaxed atomic degree counts) and scattered by threads that own vertex
// ranges; then each list is sorted, cleared of repeats in place and packed,
// by threads over vertex ranges.
void csr_from_edge_list(uint32_t n, const std::vector<std::pair<uint32_t, uint32_t>>& edges,
                        std::vector<uint32_t>& offset, std::vector<uint32_t>& adj, int threads) {
  if (threads <= 0) threads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const uint64_t ne = edges.size();
  const int T = ne >= (1u << 20) ? threads : 1;
  auto run = [&](auto&& fn) {
    if (T == 1) return fn(0);
    std::vector<std::thread> th;
    for (int t = 0; t < T; t++) th.emplace_back(fn, t);
    for (auto& x : th) x.join();
  };
  std::vector<uint32_t> deg((size_t)n + 1, 0);
  run([&](int t) {
    for (uint64_t i = ne * t / T; i < ne * (t + 1) / T; i++) {
      const auto& e = edges[i];
      if (e.first == e.second) continue;
      __atomic_fetch_add(&deg[e.first], 1u, __ATOMIC_RELAXED);
      __atomic_fetch_add(&deg[e.second], 1u, __ATOMIC_RELAXED);
    }
  });
  std::vector<uint32_t> raw((size_t)n + 1, 0);
  for (uint32_t v = 0; v < n; v++) raw[v + 1] = raw[v] + deg[v];
  std::vector<uint32_t> lists(raw[n]);
  std::vector<uint32_t> cur(raw.begin(), raw.end() - 1);
  // each thread owns a vertex range of about equal adjacency and writes only
  // its own vertices' entries (hub cursors are never shared)
  std::vector<uint32_t> own(T + 1, n);
  own[0] = 0;
  for (int t = 1; t < T; t++)
    own[t] = (uint32_t)(std::lower_bound(raw.begin(), raw.end(), (uint32_t)((uint64_t)raw[n] * t / T)) - raw.begin());
  for (int t = 1; t <= T; t++) own[t] = std::max(own[t], own[t - 1]);
  run([&](int t) {
    const uint32_t a0 = own[t], a1 = own[t + 1];
    for (const auto& e : edges) {
      if (e.first == e.second) continue;
      if (e.first >= a0 && e.first < a1) lists[cur[e.first]++] = e.second;
      if (e.second >= a0 && e.second < a1) lists[cur[e.second]++] = e.first;
    }
  });
  std::vector<uint32_t>().swap(cur);
  sort_ranges(n, raw, lists, threads);
  // repeats out, in place per list, then the kept prefixes packed
  auto vrange = [&](int t) {
    return std::make_pair((uint32_t)((uint64_t)n * t / T), (uint32_t)((uint64_t)n * (t + 1) / T));
  };
  run([&](int t) {
    const auto r = vrange(t);
    for (uint32_t v = r.first; v < r.second; v++)
      deg[v] = (uint32_t)(std::unique(lists.begin() + raw[v], lists.begin() + raw[v + 1]) - (lists.begin() + raw[v]));
  });
  offset.assign((size_t)n + 1, 0);
  for (uint32_t v = 0; v < n; v++) offset[v + 1] = offset[v] + deg[v];
  adj.resize(offset[n]);
  run([&](int t) {
    const auto r = vrange(t);
    for (uint32_t v = r.first; v < r.second; v++)
      std::copy(lists.begin() + raw[v], lists.begin() + raw[v] + deg[v], adj.begin() + offset[v]);
  });
}

int UndirectedPlainToCSRTransformer::transform(const Graph* input, Graph*& output) {
  const auto* g = dynamic_cast<const UndirectedPlainGraph*>(input);
  if (!g) return -1;
  auto* csr = new UndirectedCSRGraph();
  const Vertex n = g->num_vertices_;
  csr->num_vertices_ = n;
  csr->num_labels_ = g->num_labels_;
  csr->is_labeled_ = g->is_labeled_;
  csr->labels_.assign(g->labels_.begin(), g->labels_.begin() + n);
  csr->offset_.assign((size_t)n + 1, 0);
  // Each list is copied to its raw position, sorted and (with dedupe) cleared
  // of repeats and self-loops in place; the kept prefixes are then packed.
  // Both passes split the vertices into ranges of about equal adjacency.
  std::vector<uint64_t> raw((size_t)n + 1, 0);
  for (Vertex v = 0; v < n; v++) raw[v + 1] = raw[v] + g->edges_[v].size();
  const uint64_t total = raw[n];
  std::vector<Vertex> tmp(total);
  std::vector<Edge> len(n);
  unsigned T = total > (1u << 20) ? std::max(1u, std::min(16u, std::thread::hardware_concurrency())) : 1u;
  std::vector<Vertex> lo(T + 1, n);
  lo[0] = 0;
  for (unsigned t = 1; t < T; t++)
    lo[t] = (Vertex)(std::lower_bound(raw.begin(), raw.end(), total * t / T) - raw.begin());
  for (unsigned t = 1; t <= T; t++) lo[t] = std::max(lo[t], lo[t - 1]);
  auto run = [&](auto&& fn) {
    if (T == 1) return fn(0u, n);
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T; t++) th.emplace_back(fn, lo[t], lo[t + 1]);
    for (auto& x : th) x.join();
  };
  run([&](Vertex a, Vertex b) {
    for (Vertex v = a; v < b; v++) {
      Vertex* p = tmp.data() + raw[v];
      Vertex* e = std::copy(g->edges_[v].begin(), g->edges_[v].end(), p);
      std::sort(p, e);
      if (dedupe_) {
        e = std::unique(p, e);
        e = std::remove(p, e, v);
      }
      len[v] = (Edge)(e - p);
    }
  });
  for (Vertex v = 0; v < n; v++) csr->offset_[v + 1] = csr->offset_[v] + len[v];
  csr->edges_.resize(csr->offset_[n]);
  run([&](Vertex a, Vertex b) {
    for (Vertex v = a; v < b; v++)
      std::copy(tmp.data() + raw[v], tmp.data() + raw[v] + len[v], csr->edges_.data() + csr->offset_[v]);
  });
  csr->num_edges_ = dedupe_ ? csr->offset_[n] / 2 : g->num_edges_;
  Vertex md = n;
  for (Vertex v = 0; v < n; v++) md = std::min(md, csr->degree(v));
  csr->min_degree = md;
  output = csr;
  return 0;
}

void pcsr_from_csr(uint32_t n, const uint32_t* off, const uint32_t* adj, const int32_t* labels,
                   int32_t num_labels, bool ref_ids, std::vector<PCSRUnit>& units) {
  const uint32_t num_units = (uint32_t)((int64_t)num_labels * (num_labels + 1) / 2);
  auto unit_of = [&](Label a, Label b) {
    return (uint32_t)(ref_ids ? UndirectedPCSRGraph::reference_edge_label(a, b)
                              : UndirectedPCSRGraph::injective_edge_label(a, b));
  };
  // Pass 1, one sweep of the adjacency: per unit its keys in ascending vertex
  // order, each key's list length, and the lists themselves back to back in
  // key order (a key's list keeps its adjacency order, ascending).  Large
  // graphs sweep contiguous vertex ranges on threads; a unit's pieces are
  // joined in range order, which keeps the keys ascending.
  const uint64_t nnz = off[n];
  unsigned T = nnz > (1u << 20) ? std::max(1u, std::min(16u, std::thread::hardware_concurrency())) : 1u;
  T = std::min<unsigned>(T, std::max<uint32_t>(num_units, 1));
  const unsigned T1 = (uint64_t)T * num_units <= (1u << 16) ? T : 1u;
  struct Piece {
    std::vector<uint32_t> keys, key_len, list;
  };
  std::vector<std::vector<Piece>> piece(T1, std::vector<Piece>(num_units));
  auto sweep = [&](unsigned t) {
    const uint32_t v0 = (uint32_t)(std::lower_bound(off, off + n + 1, nnz * t / T1) - off);
    const uint32_t v1 = t + 1 == T1 ? n : (uint32_t)(std::lower_bound(off, off + n + 1, nnz * (t + 1) / T1) - off);
    std::vector<Piece>& P = piece[t];
    std::vector<uint32_t> last_key(num_units, kInvalid);
    for (uint32_t v = v0; v < v1; v++) {
      for (uint32_t i = off[v]; i < off[v + 1]; i++) {
        const uint32_t u = unit_of(labels[v], labels[adj[i]]);
        Piece& pc = P[u];
        if (last_key[u] != v) {
          last_key[u] = v;
          pc.keys.push_back(v);
          pc.key_len.push_back(0);
        }
        pc.key_len.back()++;
        pc.list.push_back(adj[i]);
      }
    }
  };
  if (T1 == 1) {
    sweep(0);
  } else {
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T1; t++) th.emplace_back(sweep, t);
    for (auto& x : th) x.join();
  }
  std::vector<std::vector<uint32_t>> keys(num_units), key_len(num_units), lists(num_units);
  std::vector<uint64_t> unit_edges(num_units, 0);
  for (uint32_t u = 0; u < num_units; u++)
    for (unsigned t = 0; t < T1; t++) unit_edges[u] += piece[t][u].list.size();
  units.assign(num_units, PCSRUnit());
  // Pass 2, units in parallel (largest first): hash the keys into K groups of
  // 15 slots (counting sort by bucket, insertion order kept within a bucket),
  // chain a full bucket's overflow through the empty groups in ascending
  // order, then copy each key's list into place group by group, slot by slot.
  std::vector<uint32_t> order(num_units);
  for (uint32_t u = 0; u < num_units; u++) order[u] = u;
  std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return unit_edges[a] > unit_edges[b]; });
  std::atomic<uint32_t> next{0};
  auto worker = [&]() {
    std::vector<uint32_t> hash, bstart, byb, slot_key, kstart;
    for (uint32_t t = next++; t < num_units; t = next++) {
      const uint32_t u = order[t];
      if (T1 == 1) {
        keys[u].swap(piece[0][u].keys);
        key_len[u].swap(piece[0][u].key_len);
        lists[u].swap(piece[0][u].list);
      } else {
        for (unsigned r = 0; r < T1; r++) {
          Piece& pc = piece[r][u];
          keys[u].insert(keys[u].end(), pc.keys.begin(), pc.keys.end());
          key_len[u].insert(key_len[u].end(), pc.key_len.begin(), pc.key_len.end());
          lists[u].insert(lists[u].end(), pc.list.begin(), pc.list.end());
          pc = Piece();
        }
      }
      PCSRUnit& unit = units[u];
      const uint32_t K = (uint32_t)keys[u].size();
      const uint64_t ne = lists[u].size();
      unit.num_keys_ = K;
      unit.num_edges_ = (uint32_t)ne;
      unit.offset_.assign((size_t)32 * K, 0);
      for (size_t w = 0; w < (size_t)32 * K; w += 2) unit.offset_[w] = kInvalid;
      unit.edges_.resize(ne);
      if (K == 0) continue;
      const uint32_t* ku = keys[u].data();
      hash.resize(K);
      bstart.assign((size_t)K + 1, 0);
      for (uint32_t j = 0; j < K; j++) {
        uint32_t key = ku[j];
        hash[j] = murmur_hash2(&key, 4, kHashSeed) % K;
        bstart[hash[j] + 1]++;
      }
      for (uint32_t b = 0; b < K; b++) bstart[b + 1] += bstart[b];
      byb.resize(K);
      {
        std::vector<uint32_t>& cur = kstart;  // reused as the fill cursor
        cur.assign(bstart.begin(), bstart.end() - 1);
        for (uint32_t j = 0; j < K; j++) byb[cur[hash[j]]++] = j;
      }
      // key index held by each (group, slot); kInvalid where empty
      slot_key.assign((size_t)15 * K, kInvalid);
      uint32_t empty_at = 0;  // FIFO over the empty buckets in ascending order
      auto next_empty = [&]() {
        while (bstart[empty_at + 1] != bstart[empty_at]) empty_at++;
        return empty_at++;
      };
      for (uint32_t b = 0; b < K; b++) {
        const uint32_t* bk = byb.data() + bstart[b];
        const uint32_t cnt = bstart[b + 1] - bstart[b];
        if (cnt == 0) continue;
        uint32_t group = b, placed = 0;
        for (uint32_t f = 0; f < cnt / 15; f++) {
          for (uint32_t s = 0; s < 15; s++) {
            slot_key[(size_t)15 * group + s] = bk[placed];
            unit.offset_[(size_t)32 * group + 2 * s] = ku[bk[placed++]];
          }
          const uint32_t nx = next_empty();
          unit.offset_[(size_t)32 * group + 30] = nx;
          group = nx;
        }
        for (uint32_t s = 0; placed < cnt; s++) {
          slot_key[(size_t)15 * group + s] = bk[placed];
          unit.offset_[(size_t)32 * group + 2 * s] = ku[bk[placed++]];
        }
      }
      kstart.resize((size_t)K + 1);
      kstart[0] = 0;
      for (uint32_t j = 0; j < K; j++) kstart[j + 1] = kstart[j] + key_len[u][j];
      const uint32_t* src = lists[u].data();
      uint32_t pos = 0;
      for (uint32_t grp = 0; grp < K; grp++) {
        for (uint32_t s = 0; s < 15; s++) {
          const uint32_t j = slot_key[(size_t)15 * grp + s];
          if (j == kInvalid) break;
          unit.offset_[(size_t)32 * grp + 2 * s + 1] = pos;
          const uint32_t len = key_len[u][j];
          std::memcpy(unit.edges_.data() + pos, src + kstart[j], 4ull * len);
          pos += len;
          unit.offset_[(size_t)32 * grp + 2 * (s + 1) + 1] = pos;
        }
      }
      std::vector<uint32_t>().swap(keys[u]);
      std::vector<uint32_t>().swap(key_len[u]);
      std::vector<uint32_t>().swap(lists[u]);
    }
  };
  if (T == 1) {
    worker();
  } else {
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T; t++) th.emplace_back(worker);
    for (auto& x : th) x.join();
  }
}

int UndirectedPlainToPCSRTransformer::transform(const Graph* input, Graph*& output) {
  const auto* g = dynamic_cast<const UndirectedPlainGraph*>(input);
  if (!g || g->num_labels_ > kPcsrMaxLabels) return -1;
  // Same adjacency the reference uses: each list sorted ascending
  // (graph_format_transformer.cc:79-83), duplicates kept.
  const Vertex n = g->num_vertices_;
  std::vector<uint32_t> off((size_t)n + 1, 0), adj;
  for (Vertex v = 0; v < n; v++) off[v + 1] = off[v] + (uint32_t)g->edges_[v].size();
  adj.resize(off[n]);
  for (Vertex v = 0; v < n; v++) {
    std::copy(g->edges_[v].begin(), g->edges_[v].end(), adj.begin() + off[v]);
    std::sort(adj.begin() + off[v], adj.begin() + off[v + 1]);
  }
  auto* p = new UndirectedPCSRGraph();
  p->num_vertices_ = n;
  p->num_edges_ = g->num_edges_;
  p->num_labels_ = g->num_labels_;
  p->is_labeled_ = g->is_labeled_;
  p->labels_.assign(g->labels_.begin(), g->labels_.begin() + n);
  p->reference_unit_ids_ = reference_unit_ids_;
  pcsr_from_csr(n, off.data(), adj.data(), p->labels_.data(), p->num_labels_, reference_unit_ids_, p->units_);
  p->num_units_ = (Vertex)p->units_.size();
  output = p;
  return 0;
}

std::vector<uint32_t> signature_from_csr(uint32_t n, const uint32_t* off, const uint32_t* adj,
                                         const int32_t* labels, bool column_oriented, bool x86_shift) {
  const uint32_t words = kSigWords;                // 16
  const uint32_t slots = (kSigBits - 32) / 2;      // 240 two-bit counters
  // slot of each neighbour label, hashed once per label when labels are small
  int32_t max_label = -1;
  bool nonneg = true;
  for (uint32_t v = 0; v < n; v++) {
    max_label = std::max(max_label, labels[v]);
    nonneg = nonneg && labels[v] >= 0;
  }
  std::vector<uint32_t> slot_of;
  if (nonneg && max_label >= 0 && max_label < (1 << 20)) {
    slot_of.resize((size_t)max_label + 1);
    for (int32_t l = 0; l <= max_label; l++) slot_of[l] = murmur_hash2(&l, 4, kHashSeed) % slots;
  }
  std::vector<uint32_t> row((size_t)words * n, 0);
  auto build = [&](uint32_t lo, uint32_t hi) {
    for (uint32_t v = lo; v < hi; v++) {
      uint32_t* sig = &row[(size_t)v * words];
      sig[0] = (uint32_t)labels[v];
      for (uint32_t i = off[v]; i < off[v + 1]; i++) {
        int32_t nl = labels[adj[i]];
        uint32_t h = slot_of.empty() ? murmur_hash2(&nl, 4, kHashSeed) % slots : slot_of[nl];
        uint32_t word = 1 + (2 * h) / 32;
        uint32_t bit = (2 * h) % 32;
        uint32_t shift = x86_shift ? ((2 * bit) & 31u) : bit;
        uint32_t cnt = (sig[word] >> shift) & 3u;
        uint32_t next = cnt == 0 ? 1u : 3u;
        sig[word] |= next << shift;
      }
    }
  };
  const unsigned T = off[n] > (1u << 20) ? std::max(1u, std::min(16u, std::thread::hardware_concurrency())) : 1u;
  auto parallel = [&](auto&& fn) {
    if (T == 1) return fn(0u, n);
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T; t++)
      th.emplace_back(fn, (uint32_t)((uint64_t)n * t / T), (uint32_t)((uint64_t)n * (t + 1) / T));
    for (auto& x : th) x.join();
  };
  parallel(build);
  if (!column_oriented) return row;
  std::vector<uint32_t> col((size_t)words * n);
  parallel([&](uint32_t lo, uint32_t hi) {
    for (uint32_t w = 0; w < words; w++)
      for (uint32_t v = lo; v < hi; v++) col[(size_t)w * n + v] = row[(size_t)v * words + w];
  });
  return col;
}

std::vector<uint32_t> UndirectedPlainToPCSRTransformer::build_signature(bool column_oriented,
                                                                        const UndirectedPlainGraph& g,
                                                                        bool x86_shift) {
  const Vertex n = g.num_vertices_;
  std::vector<uint32_t> off((size_t)n + 1, 0), adj;
  for (Vertex v = 0; v < n; v++) off[v + 1] = off[v] + (uint32_t)g.edges_[v].size();
  adj.resize(off[n]);
  for (Vertex v = 0; v < n; v++) std::copy(g.edges_[v].begin(), g.edges_[v].end(), adj.begin() + off[v]);
  return signature_from_csr(n, off.data(), adj.data(), g.labels_.data(), column_oriented, x86_shift);
}

}  // namespace fsig
