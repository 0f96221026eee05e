// Seeded synthetic graph and query generators.
#include "gen.h"

#include <algorithm>
#include <cmath>
#include <thread>

namespace fsig {

static uint64_t splitmix64(uint64_t& x) {
  uint64_t z = (x += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

Rng::Rng(uint64_t seed) {
  uint64_t x = seed;
  for (int i = 0; i < 4; i++) s[i] = splitmix64(x);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

uint64_t Rng::next() {
  uint64_t r = rotl(s[1] * 5, 7) * 9;
  uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl(s[3], 45);
  return r;
}

double Rng::uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }

uint32_t Rng::below(uint32_t bound) { return (uint32_t)(((next() >> 32) * (uint64_t)bound) >> 32); }

namespace {

// Open-addressing set of non-zero 64-bit keys, first-come insertion.
class EdgeSet {
 public:
  explicit EdgeSet(uint64_t expected) {
    uint64_t cap = 16;
    while (cap < expected * 2) cap <<= 1;
    slots_.assign(cap, 0);
    mask_ = cap - 1;
  }
  bool insert(uint64_t key) {  // key != 0
    if (2 * (count_ + 1) > slots_.size()) grow();  // stays at most half full
    uint64_t i = slot(key);
    while (true) {
      uint64_t cur = slots_[i];
      if (cur == key) return false;
      if (cur == 0) {
        slots_[i] = key;
        count_++;
        return true;
      }
      i = (i + 1) & mask_;
    }
  }
 private:
  uint64_t slot(uint64_t key) const { return ((key * 0x9e3779b97f4a7c15ull) >> 17) & mask_; }
  void grow() {
    std::vector<uint64_t> old(2 * slots_.size(), 0);
    old.swap(slots_);
    mask_ = slots_.size() - 1;
    for (uint64_t key : old) {
      if (!key) continue;
      uint64_t i = slot(key);
      while (slots_[i]) i = (i + 1) & mask_;
      slots_[i] = key;
    }
  }
  std::vector<uint64_t> slots_;
  uint64_t mask_;
  uint64_t count_ = 0;
};

inline uint64_t pair_key(uint32_t u, uint32_t v) {
  uint32_t lo = std::min(u, v), hi = std::max(u, v);
  return ((uint64_t)hi << 32 | lo) + 1;  // +1 keeps the key non-zero
}

// Vose alias table over weights.
struct Alias {
  std::vector<double> prob;
  std::vector<uint32_t> alias;
  explicit Alias(const std::vector<double>& w) {
    size_t n = w.size();
    prob.resize(n);
    alias.resize(n);
    double sum = 0;
    for (double x : w) sum += x;
    std::vector<double> scaled(n);
    std::vector<uint32_t> small, large;
    for (size_t i = 0; i < n; i++) {
      scaled[i] = w[i] * (double)n / sum;
      (scaled[i] < 1.0 ? small : large).push_back((uint32_t)i);
    }
    while (!small.empty() && !large.empty()) {
      uint32_t s = small.back(); small.pop_back();
      uint32_t l = large.back(); large.pop_back();
      prob[s] = scaled[s];
      alias[s] = l;
      scaled[l] = scaled[l] + scaled[s] - 1.0;
      (scaled[l] < 1.0 ? small : large).push_back(l);
    }
    for (uint32_t i : large) { prob[i] = 1.0; alias[i] = i; }
    for (uint32_t i : small) { prob[i] = 1.0; alias[i] = i; }
  }
  uint32_t draw(Rng& r) const {
    uint32_t i = r.below((uint32_t)prob.size());
    return r.uniform() < prob[i] ? i : alias[i];
  }
};

std::vector<uint32_t> shuffled_ids(uint32_t n, uint64_t seed) {
  std::vector<uint32_t> p(n);
  for (uint32_t i = 0; i < n; i++) p[i] = i;
  Rng r(seed ^ 0x5bd1e995ull);
  for (uint32_t i = n; i > 1; i--) std::swap(p[i - 1], p[r.below(i)]);
  return p;
}

}  // namespace

EdgeList gen_chung_lu(uint32_t n, uint64_t m, double gamma, uint64_t seed) {
  const double alpha = 1.0 / (gamma - 1.0);
  const double target = 2.0 * (double)m;
  const double cap = std::sqrt(target);
  std::vector<double> base(n), w(n);
  for (uint32_t i = 0; i < n; i++) base[i] = std::pow((double)i + 1.0, -alpha);
  double scale = target / [&] { double s = 0; for (double x : base) s += x; return s; }();
  for (int it = 0; it < 60; it++) {  // fixed point: capped weights sum to 2m
    double sum = 0;
    for (uint32_t i = 0; i < n; i++) sum += std::min(cap, scale * base[i]);
    scale *= target / sum;
  }
  for (uint32_t i = 0; i < n; i++) w[i] = std::min(cap, scale * base[i]);
  Alias table(w);
  Rng r(seed);
  EdgeSet seen(m);
  EdgeList edges;
  edges.reserve(m);
  while (edges.size() < m) {
    uint32_t u = table.draw(r), v = table.draw(r);
    if (u == v) continue;
    if (seen.insert(pair_key(u, v))) edges.emplace_back(std::min(u, v), std::max(u, v));
  }
  std::vector<uint32_t> perm = shuffled_ids(n, seed);
  for (auto& e : edges) e = {perm[e.first], perm[e.second]};
  ensure_last_vertex_has_edge(n, edges);
  return edges;
}

EdgeList gen_rmat(uint32_t scale, uint64_t m, double a, double b, double c, uint64_t seed, int threads) {
  const uint32_t n = 1u << scale;
  const uint64_t ta = (uint64_t)(a * 4294967296.0), tb = (uint64_t)((a + b) * 4294967296.0),
                 tc = (uint64_t)((a + b + c) * 4294967296.0);
  if (threads <= 0) threads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  // Candidates are drawn in fixed chunks, each from its own stream, and then
  // filtered in chunk order, so the result does not depend on `threads`.
  const uint64_t chunk = 1u << 20;
  // Large graphs keep their keys in 16 shards.  Every round but the last
  // inserts only keys it accepts, and the last inserts at most one round of
  // candidates, so m + threads * chunk keys bound what the shards ever hold;
  // each table has room for twice its share of that bound.  Small graphs keep
  // one set, filled only until m edges are accepted.
  const uint32_t nshards = m >= (1u << 22) ? 16u : 1u;
  std::vector<EdgeSet> shards;
  shards.reserve(nshards);
  if (nshards == 1) {
    shards.emplace_back(m + m / 4);
  } else {
    for (uint32_t sh = 0; sh < nshards; sh++) shards.emplace_back(16);
    std::vector<std::thread> al;
    for (uint32_t sh = 0; sh < nshards; sh++)
      al.emplace_back([&, sh] { shards[sh] = EdgeSet((m + (uint64_t)threads * chunk) / nshards + 1024); });
    for (auto& th : al) th.join();
  }
  auto shard_of = [&](uint64_t key) -> uint32_t {
    return nshards == 1 ? 0u : (uint32_t)((key * 0xd6e8feb86659fd93ull) >> 60);
  };
  std::vector<uint8_t> flags;
  // byshard[t][sh]: positions in chunk t of the non-loop candidates of shard sh, in order
  std::vector<std::vector<std::vector<uint32_t>>> byshard(
      nshards == 1 ? 0 : threads, std::vector<std::vector<uint32_t>>(nshards));
  EdgeList edges;
  edges.reserve(m);
  uint64_t next_chunk = 0;
  std::vector<std::vector<uint64_t>> bufs(threads);
  while (edges.size() < m) {
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++) {
      pool.emplace_back([&, t] {
        uint64_t cid = next_chunk + t;
        Rng r(seed * 0x100000001b3ull + cid + 1);
        auto& buf = bufs[t];
        buf.resize(chunk);
        for (uint64_t i = 0; i < chunk; i++) {
          uint32_t u = 0, v = 0;
          for (uint32_t lvl = 0; lvl < scale; lvl += 2) {
            uint64_t x = r.next();
            for (uint32_t h = 0; h < 2 && lvl + h < scale; h++) {
              uint64_t q = (x >> (32 * h)) & 0xffffffffull;
              uint32_t bu = 0, bv = 0;
              if (q < ta) {} else if (q < tb) { bv = 1; } else if (q < tc) { bu = 1; } else { bu = 1; bv = 1; }
              u = (u << 1) | bu;
              v = (v << 1) | bv;
            }
          }
          buf[i] = (uint64_t)u << 32 | v;
        }
        if (nshards > 1) {
          for (auto& l : byshard[t]) l.clear();
          for (uint64_t i = 0; i < chunk; i++) {
            const uint32_t u = (uint32_t)(buf[i] >> 32), v = (uint32_t)buf[i];
            if (u != v) byshard[t][shard_of(pair_key(u, v))].push_back((uint32_t)i);
          }
        }
      });
    }
    for (auto& th : pool) th.join();
    next_chunk += threads;
    // Flag each candidate that is the first of its key so far, in candidate
    // order.  Large graphs split the key set into shards by hash, one thread
    // per group of shards, each thread scanning the candidates in order and
    // handling its shards' keys: a key's candidates all meet in one shard in
    // the same order, so the flags are those of one set.
    if (nshards == 1) {
      for (int t = 0; t < threads && edges.size() < m; t++)
        for (uint64_t x : bufs[t]) {
          const uint32_t u = (uint32_t)(x >> 32), v = (uint32_t)x;
          if (u == v) continue;
          if (shards[0].insert(pair_key(u, v))) {
            edges.emplace_back(std::min(u, v), std::max(u, v));
            if (edges.size() == m) break;
          }
        }
      continue;
    }
    flags.assign((uint64_t)threads * chunk, 0);
    auto mark = [&](int t, int nt) {
      for (uint32_t sh = (uint32_t)t; sh < nshards; sh += (uint32_t)nt)
        for (int b = 0; b < threads; b++) {
          const uint64_t* buf = bufs[b].data();
          uint8_t* fl = flags.data() + (uint64_t)b * chunk;
          for (uint32_t i : byshard[b][sh]) {
            const uint32_t u = (uint32_t)(buf[i] >> 32), v = (uint32_t)buf[i];
            fl[i] = shards[sh].insert(pair_key(u, v)) ? 1 : 0;
          }
        }
    };
    {
      std::vector<std::thread> mk;
      const int nt = std::min<int>(threads, (int)nshards);
      for (int t = 0; t < nt; t++) mk.emplace_back(mark, t, nt);
      for (auto& th : mk) th.join();
    }
    for (int t = 0; t < threads && edges.size() < m; t++) {
      const uint8_t* fl = flags.data() + (uint64_t)t * chunk;
      for (uint64_t i = 0; i < chunk; i++) {
        if (!fl[i]) continue;
        const uint32_t u = (uint32_t)(bufs[t][i] >> 32), v = (uint32_t)bufs[t][i];
        edges.emplace_back(std::min(u, v), std::max(u, v));
        if (edges.size() == m) break;
      }
    }
  }
  std::vector<uint32_t> perm = shuffled_ids(n, seed);
  {
    std::vector<std::thread> pm;
    const uint64_t ne = edges.size();
    const int tp = ne >= (1u << 20) ? threads : 1;
    for (int t = 0; t < tp; t++)
      pm.emplace_back([&, t] {
        for (uint64_t i = ne * t / tp; i < ne * (t + 1) / tp; i++)
          edges[i] = {perm[edges[i].first], perm[edges[i].second]};
      });
    for (auto& th : pm) th.join();
  }
  ensure_last_vertex_has_edge(n, edges);
  return edges;
}

void ensure_last_vertex_has_edge(uint32_t n, EdgeList& edges) {
  if (n == 0 || edges.empty()) return;
  uint32_t hi = 0;
  for (auto& e : edges) hi = std::max(hi, std::max(e.first, e.second));
  if (hi == n - 1) return;
  const uint32_t last = n - 1;
  for (auto& e : edges) {  // swap the names `hi` and `last`
    if (e.first == hi) e.first = last; else if (e.first == last) e.first = hi;
    if (e.second == hi) e.second = last; else if (e.second == last) e.second = hi;
    if (e.first > e.second) std::swap(e.first, e.second);
  }
}

std::vector<int32_t> labels_uniform(uint32_t n, int32_t L, uint64_t seed) {
  std::vector<int32_t> out(n);
  Rng r(seed ^ 0x1abe1ull);
  for (uint32_t i = 0; i < n; i++) out[i] = (int32_t)r.below((uint32_t)L);
  return out;
}

std::vector<int32_t> labels_zipf(uint32_t n, int32_t L, double s, uint64_t seed) {
  std::vector<double> cdf(L);
  double acc = 0;
  for (int32_t i = 0; i < L; i++) { acc += 1.0 / std::pow((double)i + 1.0, s); cdf[i] = acc; }
  for (auto& x : cdf) x /= acc;
  std::vector<int32_t> out(n);
  Rng r(seed ^ 0x21bfull);
  for (uint32_t i = 0; i < n; i++) {
    double u = r.uniform();
    int32_t l = (int32_t)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
    out[i] = std::min(l, L - 1);
  }
  return out;
}

static bool has_edge(const uint32_t* off, const uint32_t* adj, uint32_t u, uint32_t v) {
  return std::binary_search(adj + off[u], adj + off[u + 1], v);
}

int random_walk_query(uint32_t n, const uint32_t* off, const uint32_t* adj, const int32_t* labels,
                      uint32_t k, uint32_t min_edges, uint32_t max_edges, uint64_t seed, int bias,
                      uint32_t max_attempts, WalkQuery& out, const std::vector<uint32_t>* given) {
  Rng r(seed ^ 0x3a1cull);
  std::vector<uint32_t> local;
  if (!given) {
    for (uint32_t v = 0; v < n; v++)
      if (off[v + 1] > off[v]) local.push_back(v);
  }
  const std::vector<uint32_t>& nonisolated = given ? *given : local;
  if (nonisolated.empty() || k == 0) return -1;
  std::vector<uint32_t> visit;
  for (uint32_t attempt = 0; attempt < max_attempts; attempt++) {
    visit.clear();
    uint32_t cur = nonisolated[r.below((uint32_t)nonisolated.size())];
    visit.push_back(cur);
    for (uint32_t step = 0; step < 64 * k && visit.size() < k; step++) {
      uint32_t d = off[cur + 1] - off[cur];
      uint32_t nxt = adj[off[cur] + r.below(d)];
      for (int b = 0; b < (bias < 0 ? -bias : bias); b++) {
        uint32_t alt = adj[off[cur] + r.below(d)];
        const uint32_t da = off[alt + 1] - off[alt], dn = off[nxt + 1] - off[nxt];
        if (bias > 0 ? da > dn : da < dn) nxt = alt;
      }
      cur = nxt;
      if (std::find(visit.begin(), visit.end(), cur) == visit.end()) visit.push_back(cur);
    }
    if (visit.size() < k) continue;
    std::vector<std::pair<uint32_t, uint32_t>> qe;
    for (uint32_t i = 0; i < k; i++)
      for (uint32_t j = i + 1; j < k; j++)
        if (has_edge(off, adj, visit[i], visit[j])) qe.emplace_back(i, j);
    if (qe.size() < min_edges || qe.size() > max_edges) continue;
    out.k = k;
    out.labels.resize(k);
    for (uint32_t i = 0; i < k; i++) out.labels[i] = labels[visit[i]];
    out.edges = qe;
    out.data_vertices = visit;
    return 0;
  }
  return -1;
}

}  // namespace fsig
